// host_check.cpp -- the library's host code (Matrix Market loader with its
// threaded parser, seeded generators) and the CPU oracle run under the
// sanitizers: AddressSanitizer + UndefinedBehaviorSanitizer, or
// ThreadSanitizer for the threaded loader alone (built without -DWITH_ORACLE).
// Built with g++ from the sources by tests/test_host_sanitizers.py; prints
// "ok" and exits 0 when every comparison holds, exits 1 on a mismatch (a
// sanitizer report aborts on its own).
//
// usage: host_check <scratch dir> <file.mtx>...
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "swift_spmm.h"
#ifdef WITH_ORACLE
#include "swift_oracle.h"
#endif

namespace {

int failures = 0;

void fail(const std::string &what) {
    fprintf(stderr, "MISMATCH %s\n", what.c_str());
    failures++;
}

struct Loaded {
    int rc = 0, m = 0, n = 0, nnz = 0, sym = 0;
    std::vector<int> rp, ci;
    std::vector<double> v;
};

Loaded load(const std::string &path, swift_dtype dt, int threads) {
    Loaded L;
    int *rp = nullptr, *ci = nullptr;
    void *v = nullptr;
    L.rc = swift_mmio_allinone_ex(&L.m, &L.n, &L.nnz, &L.sym, &rp, &ci, &v, dt, path.c_str(),
                                  threads);
    if (L.rc != 0) return L;
    L.rp.assign(rp, rp + L.m + 1);
    L.ci.assign(ci, ci + L.nnz);
    L.v.resize((size_t)L.nnz);
    for (int i = 0; i < L.nnz; i++)
        L.v[i] = dt == SWIFT_F64 ? static_cast<double *>(v)[i] : static_cast<float *>(v)[i];
    swift_host_free(rp);
    swift_host_free(ci);
    swift_host_free(v);
    return L;
}

bool same(const Loaded &a, const Loaded &b) {
    if (a.rc != b.rc) return false;
    if (a.rc != 0) return true;
    return a.m == b.m && a.n == b.n && a.nnz == b.nnz && a.sym == b.sym && a.rp == b.rp &&
           a.ci == b.ci && std::memcmp(a.v.data(), b.v.data(), a.v.size() * sizeof(double)) == 0;
}

void check_threads(const std::string &path) {
    for (swift_dtype dt : {SWIFT_F64, SWIFT_F32}) {
        const Loaded ref = load(path, dt, 1);
        for (int t : {2, 8})
            if (!same(ref, load(path, dt, t))) fail(path + " threads " + std::to_string(t));
    }
}

// Generated matrix -> .mtx text -> loader, sequential and threaded.
void check_generated(const std::string &dir, int which, int index, double scale) {
    int m, n, nnz, k;
    int *rp, *ci;
    double *v;
    if (swift_generate(which, index, scale, 7, &m, &n, &nnz, &rp, &ci, &v, &k) != 0) {
        fail("generate " + std::to_string(which));
        return;
    }
    const std::string path = dir + "/gen" + std::to_string(which) + ".mtx";
    FILE *f = fopen(path.c_str(), "w");
    fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n%d %d %d\n", m, n, nnz);
    for (int i = 0; i < m; i++)
        for (int e = rp[i]; e < rp[i + 1]; e++) fprintf(f, "%d %d %.17g\n", i + 1, ci[e] + 1, v[e]);
    fclose(f);
    for (int t : {1, 8}) {
        const Loaded L = load(path, SWIFT_F64, t);
        const bool ok = L.rc == 0 && L.m == m && L.n == n && L.nnz == nnz &&
                        std::equal(L.rp.begin(), L.rp.end(), rp) &&
                        std::equal(L.ci.begin(), L.ci.end(), ci) &&
                        std::memcmp(L.v.data(), v, sizeof(double) * (size_t)nnz) == 0;
        if (!ok) fail(path + " round trip, threads " + std::to_string(t));
    }
#ifdef WITH_ORACLE
    // oracle: Swift build, row sort and Swift_cpu against the FP64 CSR product
    oracle_swift *s = oracle_build(m, n, nnz, rp, ci, v);
    oracle_row_idx_sort(s);
    const int K = 8;
    std::vector<double> B((size_t)n * K), sB((size_t)n * K), C((size_t)m * K, 0.0),
        C64((size_t)m * K), absC((size_t)m * K);
    for (size_t i = 0; i < B.size(); i++) B[i] = std::sin(0.37 * (double)i);
    for (int i = 0; i < n; i++)
        for (int kk = 0; kk < K; kk++) sB[(size_t)i * K + kk] = B[(size_t)s->perm[i] * K + kk];
    oracle_swift_cpu_f64(s, K, sB.data(), C.data());
    oracle_spmm_csr_f64(m, n, rp, ci, v, K, B.data(), C64.data(), absC.data());
    for (size_t i = 0; i < C.size(); i++)
        if (std::fabs(C[i] - C64[i]) > 1e-12 * (absC[i] + 1e-30)) {
            fail("oracle Swift_cpu vs CSR, generator " + std::to_string(which));
            break;
        }
    long long blocks, empty, dense, dense_nnz;
    int mx;
    oracle_block_stats(m, n, rp, ci, &blocks, &empty);
    oracle_block_density(m, n, rp, ci, 512, &dense, &dense_nnz, &mx);
    if (empty > blocks || dense > blocks - empty) fail("oracle block statistics");
    oracle_swift_free(s);
#endif
    swift_host_free(rp);
    swift_host_free(ci);
    swift_host_free(v);
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 2) {
        fprintf(stderr, "usage: %s <scratch dir> <file.mtx>...\n", argv[0]);
        return 2;
    }
    for (int i = 2; i < argc; i++) check_threads(argv[i]);
    check_generated(argv[1], 1, 0, 0.25);   // 1024 x 1024 CPU-path case
    check_generated(argv[1], 4, 0, 0.01);   // 2,000-node graph
    check_generated(argv[1], 5, 3, 1.0);    // one batch matrix
    if (failures) return 1;
    printf("ok\n");
    return 0;
}
