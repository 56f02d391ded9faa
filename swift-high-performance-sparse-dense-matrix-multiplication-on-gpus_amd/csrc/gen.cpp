// gen.cpp -- seeded synthetic matrices of the benchmark configurations (host).
//
// The reference runs on downloaded SuiteSparse matrices with values from
// srand(time(0)) (src/main.cu:125-137), so nothing of its inputs is
// reproducible; these generators give the configurations a fixed, documented
// shape.  PRNG: splitmix64-seeded xoshiro256**, one stream per row block so the
// output does not depend on thread count.
#include <new>
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "swift_spmm.h"

namespace {

struct Rng {
    uint64_t s[4];
    static uint64_t splitmix(uint64_t &x) {
        uint64_t z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    explicit Rng(uint64_t seed) {
        uint64_t x = seed;
        for (auto &v : s) v = splitmix(x);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }          // [0,1)
    uint64_t below(uint64_t n) { return n ? (uint64_t)(uniform() * (double)n) % n : 0; }
    double normal() {
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

// Truncated Zipf(a) on [lo, hi] by inverse CDF over a table.
struct Zipf {
    std::vector<double> cdf;
    int lo;
    Zipf(double a, int lo_, int hi) : lo(lo_) {
        cdf.resize((size_t)(hi - lo + 1));
        double acc = 0;
        for (int k = lo; k <= hi; k++) cdf[(size_t)(k - lo)] = (acc += std::pow((double)k, -a));
        for (auto &c : cdf) c /= acc;
    }
    int draw(Rng &r) const {
        double u = r.uniform();
        size_t i = (size_t)(std::lower_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
        if (i >= cdf.size()) i = cdf.size() - 1;
        return lo + (int)i;
    }
};

// k distinct values in [lo, hi) in ascending order.
void distinct_sorted(Rng &r, int lo, int hi, int k, std::vector<int> &out) {
    out.clear();
    int range = hi - lo;
    if (k >= range) {
        for (int v = lo; v < hi; v++) out.push_back(v);
        return;
    }
    if (k * 4 > range) {  // dense: selection sampling
        int need = k;
        for (int v = lo; v < hi && need > 0; v++) {
            if (r.below((uint64_t)(hi - v)) < (uint64_t)need) {
                out.push_back(v);
                need--;
            }
        }
        return;
    }
    while ((int)out.size() < k) {
        size_t before = out.size();
        for (int i = (int)before; i < k; i++) out.push_back(lo + (int)r.below((uint64_t)range));
        std::sort(out.begin(), out.end());
        out.erase(std::unique(out.begin(), out.end()), out.end());
    }
}

struct Csr {
    int m = 0, n = 0;
    std::vector<int> rowptr, colidx;
    std::vector<double> val;
};

int emit(const Csr &A, int *m, int *n, int *nnz, int **rowptr, int **colidx, double **val) {
    size_t z = A.colidx.size();
    if (z > 0x7FFFFFFFull) return SWIFT_ERR_SIZE;
    int *rp = (int *)malloc(sizeof(int) * ((size_t)A.m + 1));
    int *ci = (int *)malloc(sizeof(int) * (z ? z : 1));
    double *vv = (double *)malloc(sizeof(double) * (z ? z : 1));
    if (!rp || !ci || !vv) {
        free(rp); free(ci); free(vv);
        return SWIFT_ERR_ALLOC;
    }
    memcpy(rp, A.rowptr.data(), sizeof(int) * ((size_t)A.m + 1));
    if (z) {
        memcpy(ci, A.colidx.data(), sizeof(int) * z);
        memcpy(vv, A.val.data(), sizeof(double) * z);
    }
    *m = A.m; *n = A.n; *nnz = (int)z;
    *rowptr = rp; *colidx = ci; *val = vv;
    return SWIFT_OK;
}

// Row lengths -> CSR with per-row column sampler.
template <class ColFn, class ValFn>
void fill_rows(Csr &A, const std::vector<int> &len, uint64_t seed, ColFn cols, ValFn vals) {
    A.rowptr.assign((size_t)A.m + 1, 0);
    for (int i = 0; i < A.m; i++) A.rowptr[i + 1] = A.rowptr[i] + len[i];
    A.colidx.resize((size_t)A.rowptr[A.m]);
    A.val.resize((size_t)A.rowptr[A.m]);
    std::vector<int> tmp;
    for (int i = 0; i < A.m; i++) {
        Rng r(seed ^ (0xA5A5A5A5ull * (uint64_t)(i + 1)));
        cols(r, i, len[i], tmp);
        int base = A.rowptr[i];
        for (int q = 0; q < len[i]; q++) {
            A.colidx[base + q] = tmp[q];
            A.val[base + q] = vals(r);
        }
    }
}

// Config 1: 4096^2, Zipf(2.0) lengths on [1,512] scaled by a common factor to
// 80k nonzeros; 30% of rows are unions of contiguous 8-wide column runs.
void gen1(double scale, uint64_t seed, Csr &A) {
    A.m = A.n = std::max(32, (int)std::lround(4096 * scale));
    Rng r(seed);
    Zipf z(2.0, 1, 512);
    std::vector<int> len((size_t)A.m);
    std::vector<char> runs((size_t)A.m);
    double sum = 0;
    for (int i = 0; i < A.m; i++) {
        len[i] = z.draw(r);
        runs[i] = r.uniform() < 0.3;
        sum += len[i];
    }
    // common factor, refined for the clipping at 512, so that the total is ~80k
    const double target = 80000.0 * ((double)A.m / 4096.0);
    const int cap = std::min(512, A.n);
    double f = target / sum;
    for (int it = 0; it < 8; it++) {
        double tot = 0;
        for (int i = 0; i < A.m; i++) tot += std::min(cap, std::max(1, (int)std::lround(len[i] * f)));
        f *= target / tot;
    }
    for (int i = 0; i < A.m; i++) len[i] = std::min(cap, std::max(1, (int)std::lround(len[i] * f)));
    const int N = A.n;
    fill_rows(A, len, seed * 31 + 7,
              [&](Rng &rr, int i, int L, std::vector<int> &out) {
                  if (!runs[i] || N < 8) {
                      distinct_sorted(rr, 0, N, L, out);
                      return;
                  }
                  std::vector<char> mark((size_t)N, 0);
                  int have = 0;
                  while (have < L) {
                      int s = (int)rr.below((uint64_t)(N - 7));
                      for (int c = s; c < s + 8 && have < L; c++)
                          if (!mark[c]) { mark[c] = 1; have++; }
                  }
                  out.clear();
                  for (int c = 0; c < N; c++)
                      if (mark[c]) out.push_back(c);
              },
              [](Rng &rr) { return rr.uniform() * 2.0 - 1.0; });
}

// Config 2: even rows banded (|j - i| <= 32, 8..32 nnz), odd rows Zipf(1.8)
// on [1,20000] with uniform columns; the power-law lengths are scaled by one
// common factor so that the matrix holds 16e6 * scale nonzeros.
void gen2(double scale, uint64_t seed, Csr &A) {
    A.m = A.n = std::max(64, (int)std::lround(1000000 * scale));
    const int M = A.m, N = A.n;
    Rng r(seed);
    Zipf z(1.8, 1, std::min(20000, N));
    std::vector<int> len((size_t)M);
    long long banded = 0;
    double plsum = 0;
    for (int i = 0; i < M; i++) {
        if ((i & 1) == 0) {
            int lo = std::max(0, i - 32), hi = std::min(N, i + 33);
            len[i] = std::min(hi - lo, 8 + (int)r.below(25));
            banded += len[i];
        } else {
            len[i] = z.draw(r);
            plsum += len[i];
        }
    }
    double target = 16.0e6 * scale - (double)banded;
    double f = plsum > 0 ? std::max(0.05, target / plsum) : 1.0;
    for (int i = 1; i < M; i += 2)
        len[i] = std::min(N, std::max(1, (int)std::lround(len[i] * f)));
    fill_rows(A, len, seed * 131 + 3,
              [&](Rng &rr, int i, int L, std::vector<int> &out) {
                  if ((i & 1) == 0) distinct_sorted(rr, std::max(0, i - 32), std::min(N, i + 33), L, out);
                  else distinct_sorted(rr, 0, N, L, out);
              },
              [](Rng &rr) { return rr.normal(); });
}

// Config 4: Chung-Lu graph.  Weights w_i = c (i + i0)^(-1/(gamma-1)) with
// gamma = 2.1, capped at sqrt(sum w), scaled to mean degree 50; n*50/2 edges
// drawn with endpoints proportional to w; loops and duplicates dropped,
// symmetrised, identity added; values 1/sqrt(d_i d_j) with d counting the
// self loop.
void gen4(double scale, uint64_t seed, Csr &A) {
    const int n = std::max(64, (int)std::lround(200000 * scale));
    A.m = A.n = n;
    const double gamma = 2.1, mean_deg = 50.0;
    std::vector<double> w((size_t)n);
    const double i0 = 10.0;
    double sum = 0;
    for (int i = 0; i < n; i++) sum += (w[i] = std::pow(i + i0, -1.0 / (gamma - 1.0)));
    for (int it = 0; it < 4; it++) {
        double f = mean_deg * n / sum;
        double cap = std::sqrt(mean_deg * n);
        sum = 0;
        for (auto &x : w) sum += (x = std::min(cap, x * f));
    }
    std::vector<double> cdf((size_t)n);
    double acc = 0;
    for (int i = 0; i < n; i++) cdf[i] = (acc += w[i]);
    Rng r(seed);
    // random relabelling so that hubs are spread over the index range
    std::vector<int> label((size_t)n);
    for (int i = 0; i < n; i++) label[i] = i;
    for (int i = n - 1; i > 0; i--) std::swap(label[i], label[(size_t)r.below((uint64_t)i + 1)]);
    const long long E = (long long)(mean_deg * n / 2.0);
    std::vector<uint64_t> edges;
    edges.reserve((size_t)E * 2 + n);
    auto pick = [&]() {
        double u = r.uniform() * acc;
        int i = (int)(std::lower_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
        return label[std::min(i, n - 1)];
    };
    for (long long e = 0; e < E; e++) {
        int a = pick(), b = pick();
        if (a == b) continue;
        edges.push_back(((uint64_t)a << 32) | (uint32_t)b);
        edges.push_back(((uint64_t)b << 32) | (uint32_t)a);
    }
    for (int i = 0; i < n; i++) edges.push_back(((uint64_t)i << 32) | (uint32_t)i);
    std::sort(edges.begin(), edges.end());
    edges.erase(std::unique(edges.begin(), edges.end()), edges.end());
    A.rowptr.assign((size_t)n + 1, 0);
    for (uint64_t e : edges) A.rowptr[(e >> 32) + 1]++;
    for (int i = 0; i < n; i++) A.rowptr[i + 1] += A.rowptr[i];
    A.colidx.resize(edges.size());
    A.val.resize(edges.size());
    for (size_t k = 0; k < edges.size(); k++) A.colidx[k] = (int)(uint32_t)edges[k];
    for (int i = 0; i < n; i++) {
        for (int p = A.rowptr[i]; p < A.rowptr[i + 1]; p++) {
            int j = A.colidx[p];
            double di = A.rowptr[i + 1] - A.rowptr[i], dj = A.rowptr[j + 1] - A.rowptr[j];
            A.val[p] = 1.0 / std::sqrt(di * dj);
        }
    }
}

// Config 5 item: square, rows log-U[1k,64k], mean nnz/row log-U[4,128],
// Zipf(2.0) lengths on [1, n] scaled to that mean, uniform columns, N(0,1).
void gen5(int index, double scale, uint64_t seed, Csr &A, int *k) {
    Rng r(seed * 1000003ull + (uint64_t)index);
    double lr = std::log(1000.0) + r.uniform() * (std::log(64000.0) - std::log(1000.0));
    double la = std::log(4.0) + r.uniform() * (std::log(128.0) - std::log(4.0));
    static const int ks[3] = {32, 64, 128};
    int kk = ks[r.below(3)];
    if (k) *k = kk;
    A.m = A.n = std::max(64, (int)std::lround(std::exp(lr) * scale));
    double mean = std::exp(la);
    Zipf z(2.0, 1, std::min(A.n, 4096));
    std::vector<int> len((size_t)A.m);
    double sum = 0;
    for (int i = 0; i < A.m; i++) sum += (len[i] = z.draw(r));
    double f = mean * A.m / sum;
    for (auto &L : len) L = std::min(A.n, std::max(1, (int)std::lround(L * f)));
    const int N = A.n;
    fill_rows(A, len, seed * 7 + (uint64_t)index * 13 + 1,
              [&](Rng &rr, int, int L, std::vector<int> &out) { distinct_sorted(rr, 0, N, L, out); },
              [](Rng &rr) { return rr.normal(); });
}

}  // namespace

extern "C" int swift_generate(int which, int index, double scale, uint64_t seed, int *m, int *n,
                              int *nnz, int **rowptr, int **colidx, double **val, int *k_hint) {
    if (!m || !n || !nnz || !rowptr || !colidx || !val || !(scale > 0.0) || scale > 1.0)
        return SWIFT_ERR_ARG;
    try {  // host allocation failures must not escape through C
        Csr A;
        int k = 128;
        switch (which) {
            case 1: gen1(scale, seed, A); k = 32; break;
            case 2: gen2(scale, seed, A); break;
            case 4: gen4(scale, seed, A); break;
            case 5: gen5(index, scale, seed, A, &k); break;
            default: return SWIFT_ERR_ARG;
        }
        if (k_hint) *k_hint = k;
        return emit(A, m, n, nnz, rowptr, colidx, val);
    } catch (const std::bad_alloc &) {
        return SWIFT_ERR_ALLOC;
    }
}

extern "C" int swift_generate_dense(size_t count, int dist, double lo, double hi, uint64_t seed,
                                    double *out) {
    if (!out && count) return SWIFT_ERR_ARG;
    const size_t chunk = 1 << 16;
    for (size_t b = 0; b < count; b += chunk) {
        Rng r(seed ^ (0x9E37ull * (uint64_t)(b / chunk + 1)));
        size_t e = std::min(count, b + chunk);
        for (size_t i = b; i < e; i++)
            out[i] = dist == 1 ? r.normal() : lo + (hi - lo) * r.uniform();
    }
    return SWIFT_OK;
}
