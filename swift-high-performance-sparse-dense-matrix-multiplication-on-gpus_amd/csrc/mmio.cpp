// mmio.cpp -- Matrix Market loader (host).
//
// Behaviour of mmio_allinone (src/auxiliary.h:167-331) with the banner and
// size-line rules of src/mmio.h:120-256: coordinate files only; real,
// integer, pattern (-> 1.0) and complex (real part) values; symmetric and
// hermitian off-diagonal entries mirrored right after the original entry;
// skew-symmetric read as stored; 1-based indices converted to 0-based; rows
// keep file order.  Unlike the reference, truncated or out-of-range entry
// lines return SWIFT_ERR_DATA instead of reading uninitialised indices, and
// FP32 values are parsed as double and rounded once (the reference's FP32
// fork scans %lg into a float, test/k-32/FP32/auxiliary.h:238-248).
//
// Large files are parsed by several threads (swift_mmio_allinone_ex).  The
// entry section is mapped and cut into chunks at whitespace; a first pass
// counts each chunk's tokens, so every chunk knows the index of its first
// token, and a second pass parses the entries whose first token lies in the
// chunk (an entry may continue into the next chunk: entries are counted in
// tokens, not lines, as fscanf reads them).  Tokens are converted with strtol
// / strtod, which accept what %d / %lg accept; a token that one of them does
// not consume whole sends the file to the sequential fscanf parser, so odd
// files keep the reference's reading.  The CSR build after parsing is shared,
// so both parsers give the same arrays.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cctype>
#include <cerrno>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <new>
#include <thread>
#include <vector>

#include "swift_spmm.h"

namespace {

struct FileCloser {
    FILE *f;
    ~FileCloser() { if (f) fclose(f); }
};

int parse_banner(FILE *f, char &field, char &symmetry) {
    char line[1025], banner[64], object[64], format[64], type[64], sym[64];
    if (!fgets(line, sizeof line, f)) return -1;
    if (sscanf(line, "%63s %63s %63s %63s %63s", banner, object, format, type, sym) != 5)
        return -1;
    for (char *s : {object, format, type, sym})
        for (char *p = s; *p; ++p) *p = (char)tolower((unsigned char)*p);
    if (strncmp(banner, "%%MatrixMarket", 14) != 0) return -1;
    if (strcmp(object, "matrix") != 0) return -1;
    // dense "array" files pass the reference's banner check but are then
    // misread as coordinates; they are rejected here.
    if (strcmp(format, "coordinate") != 0) return -1;
    if (!strcmp(type, "real")) field = 'R';
    else if (!strcmp(type, "complex")) field = 'C';
    else if (!strcmp(type, "pattern")) field = 'P';
    else if (!strcmp(type, "integer")) field = 'I';
    else return -1;
    if (!strcmp(sym, "general")) symmetry = 'G';
    else if (!strcmp(sym, "symmetric")) symmetry = 'S';
    else if (!strcmp(sym, "hermitian")) symmetry = 'H';
    else if (!strcmp(sym, "skew-symmetric")) symmetry = 'K';
    else return -1;
    return 0;
}

int parse_size(FILE *f, int &m, int &n, long long &nz) {
    char line[1025];
    m = n = 0;
    nz = 0;
    do {
        if (!fgets(line, sizeof line, f)) return -1;
    } while (line[0] == '%');
    if (sscanf(line, "%d %d %lld", &m, &n, &nz) == 3) return 0;
    int got;
    do {
        got = fscanf(f, "%d %d %lld", &m, &n, &nz);
        if (got == EOF) return -1;
    } while (got != 3);
    return 0;
}

// Entries in file order, 0-based: rows, columns, values.
struct Entries {
    std::vector<int> ri, ci;
    std::vector<double> vv;
};

// The reference's reading: fscanf per entry (src/auxiliary.h:238-276).
int parse_sequential(FILE *f, char field, bool mirror, int M, int N, long long NZ, Entries &E) {
    for (long long e = 0; e < NZ; e++) {
        int i, j, iv;
        double x = 1.0, xi;
        int got;
        switch (field) {
            case 'R': got = fscanf(f, "%d %d %lg", &i, &j, &x) == 3; break;
            case 'C': got = fscanf(f, "%d %d %lg %lg", &i, &j, &x, &xi) == 4; break;
            case 'I': got = fscanf(f, "%d %d %d", &i, &j, &iv) == 3; x = got ? iv : 0; break;
            default: got = fscanf(f, "%d %d", &i, &j) == 2; x = 1.0; break;
        }
        if (!got || i < 1 || i > M || j < 1 || j > N) return SWIFT_ERR_DATA;
        if (mirror && i != j && j > M) return SWIFT_ERR_DATA;
        E.ri[e] = i - 1;
        E.ci[e] = j - 1;
        E.vv[e] = x;
    }
    return SWIFT_OK;
}

inline bool is_ws(char c) { return c == ' ' || c == '\n' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; }

// Token [p, q) as a NUL-terminated copy (tokens longer than 63 bytes are not
// numbers either parser reads whole).
inline bool token_copy(const char *p, const char *q, char (&buf)[64]) {
    const size_t len = (size_t)(q - p);
    if (len == 0 || len >= sizeof buf) return false;
    memcpy(buf, p, len);
    buf[len] = 0;
    return true;
}

inline bool token_int(const char *p, const char *q, long &v) {
    char buf[64];
    if (!token_copy(p, q, buf)) return false;
    char *end;
    errno = 0;
    v = strtol(buf, &end, 10);
    return *end == 0 && errno == 0 && v >= INT_MIN && v <= INT_MAX;
}

inline bool token_double(const char *p, const char *q, double &v) {
    char buf[64];
    if (!token_copy(p, q, buf)) return false;
    char *end;
    v = strtod(buf, &end);
    return *end == 0;
}

// Runs fn(0) .. fn(n-1) on n threads.  If a thread cannot be started, the
// ones already running are joined before the error propagates (a joinable
// std::thread must not be destroyed).
template <class F>
void run_threads(int n, F fn) {
    std::vector<std::thread> th;
    th.reserve((size_t)n);
    try {
        for (int t = 0; t < n; t++) th.emplace_back(fn, t);
    } catch (...) {
        for (auto &x : th) x.join();
        throw;
    }
    for (auto &x : th) x.join();
}

// Outcome of the parallel parse: OK, a data error, or "use the sequential
// parser" (a token that strtol/strtod do not consume whole).
enum { PAR_OK = 0, PAR_DATA = 1, PAR_FALLBACK = 2 };

int parse_parallel(const char *d, const char *end, char field, bool mirror, int M, int N,
                   long long NZ, int nthreads, Entries &E) {
    const int tpe = field == 'P' ? 2 : field == 'C' ? 4 : 3;  // tokens per entry
    const size_t len = (size_t)(end - d);
    std::vector<const char *> cut((size_t)nthreads + 1);
    cut[0] = d;
    cut[nthreads] = end;
    for (int t = 1; t < nthreads; t++) {
        const char *p = d + len * t / nthreads;
        p = std::max(p, cut[t - 1]);
        while (p < end && !is_ws(*p)) p++;  // a chunk starts on whitespace
        cut[t] = p;
    }
    // pass 1: tokens per chunk
    std::vector<long long> ntok((size_t)nthreads + 1, 0);
    run_threads(nthreads, [&](int t) {
        long long c = 0;
        bool in = false;
        for (const char *p = cut[t]; p < cut[t + 1]; p++) {
            const bool w = is_ws(*p);
            if (!w && !in) c++;
            in = !w;
        }
        ntok[t + 1] = c;
    });
    for (int t = 0; t < nthreads; t++) ntok[t + 1] += ntok[t];
    if (ntok[nthreads] < NZ * tpe) return PAR_DATA;  // fscanf would meet the end
    // pass 2: entries whose first token is in the chunk
    std::atomic<int> status{PAR_OK};
    {
        run_threads(nthreads, [&](int t) {
                long long g = ntok[t];  // global index of the next token
                const char *p = cut[t];
                auto next = [&](const char *&a, const char *&b) -> bool {
                    while (p < end && is_ws(*p)) p++;
                    if (p >= end) return false;
                    a = p;
                    while (p < end && !is_ws(*p)) p++;
                    b = p;
                    g++;
                    return true;
                };
                const char *a, *b;
                // tokens of an entry that began in an earlier chunk
                while (g % tpe && g < ntok[t + 1])
                    if (!next(a, b)) return;
                while (g < ntok[t + 1] && g / tpe < NZ && status.load(std::memory_order_relaxed) == PAR_OK) {
                    const long long e = g / tpe;
                    const char *ta[4], *tb[4];
                    for (int k = 0; k < tpe; k++)
                        if (!next(ta[k], tb[k])) { status = PAR_DATA; return; }
                    long i, j, iv;
                    double x = 1.0, xi;
                    bool ok = token_int(ta[0], tb[0], i) && token_int(ta[1], tb[1], j);
                    if (field == 'R') ok = ok && token_double(ta[2], tb[2], x);
                    else if (field == 'C') ok = ok && token_double(ta[2], tb[2], x) && token_double(ta[3], tb[3], xi);
                    else if (field == 'I') { ok = ok && token_int(ta[2], tb[2], iv); x = (double)iv; }
                    if (!ok) { status = PAR_FALLBACK; return; }
                    if (i < 1 || i > M || j < 1 || j > N || (mirror && i != j && j > M)) {
                        int want = PAR_OK;
                        status.compare_exchange_strong(want, PAR_DATA);
                        return;
                    }
                    E.ri[e] = (int)i - 1;
                    E.ci[e] = (int)j - 1;
                    E.vv[e] = x;
                }
            });
    }
    return status.load();
}

// Files with at least this many entry bytes are parsed in parallel when the
// thread count is automatic.
constexpr size_t PARALLEL_BYTES = 1u << 20;

template <class T>
int load(int *m_out, int *n_out, int *nnz_out, int *sym_out, int **rowptr_out,
         int **colidx_out, T **val_out, const char *path, int nthreads) {
    if (!m_out || !n_out || !nnz_out || !sym_out || !rowptr_out || !colidx_out || !val_out ||
        !path || nthreads < 0)
        return SWIFT_ERR_ARG;
    FILE *f = fopen(path, "r");
    if (!f) return SWIFT_ERR_OPEN;
    FileCloser closer{f};
    char field = 'R', symmetry = 'G';
    if (parse_banner(f, field, symmetry) != 0) return SWIFT_ERR_BANNER;
    int M, N;
    long long NZ;
    if (parse_size(f, M, N, NZ) != 0) return SWIFT_ERR_SIZE;
    if (M < 0 || N < 0 || NZ < 0 || NZ > 0x7FFFFFFFLL) return SWIFT_ERR_SIZE;
    const bool mirror = symmetry == 'S' || symmetry == 'H';
    Entries E;
    E.ri.resize((size_t)NZ);
    E.ci.resize((size_t)NZ);
    E.vv.resize((size_t)NZ);
    // entry section: from the current position to the end of the file
    const long start = ftell(f);
    struct stat st;
    const bool sized = start >= 0 && fstat(fileno(f), &st) == 0 && st.st_size >= start;
    const size_t body = sized ? (size_t)(st.st_size - start) : 0;
    int threads = nthreads;
    if (threads == 0) {
        const unsigned hw = std::thread::hardware_concurrency();
        threads = body >= PARALLEL_BYTES ? (int)std::min(16u, std::max(1u, hw)) : 1;
    }
    int par = PAR_FALLBACK;
    if (threads > 1 && sized && body > 0 && NZ > 0) {
        void *map = mmap(nullptr, (size_t)st.st_size, PROT_READ, MAP_PRIVATE, fileno(f), 0);
        if (map != MAP_FAILED) {
            const char *base = static_cast<const char *>(map);
            try {
                par = parse_parallel(base + start, base + st.st_size, field, mirror, M, N, NZ,
                                     threads, E);
            } catch (const std::exception &) {  // threads could not be started
                par = PAR_FALLBACK;
            }
            munmap(map, (size_t)st.st_size);
        }
    }
    if (par == PAR_DATA) return SWIFT_ERR_DATA;
    if (par == PAR_FALLBACK) {
        if (fseek(f, start, SEEK_SET) != 0) return SWIFT_ERR_DATA;
        const int rc = parse_sequential(f, field, mirror, M, N, NZ, E);
        if (rc != SWIFT_OK) return rc;
    }
    const std::vector<int> &ri = E.ri, &ci = E.ci;
    const std::vector<double> &vv = E.vv;
    std::vector<long long> count((size_t)M + 1, 0);
    for (long long e = 0; e < NZ; e++) {
        count[ri[e]]++;
        if (mirror && ri[e] != ci[e]) count[ci[e]]++;
    }
    long long total = 0;
    std::vector<int> rowptr((size_t)M + 1);
    for (int r = 0; r < M; r++) {
        rowptr[r] = (int)total;
        total += count[r];
    }
    if (total > 0x7FFFFFFFLL) return SWIFT_ERR_SIZE;
    rowptr[M] = (int)total;
    int *rp = (int *)malloc(sizeof(int) * ((size_t)M + 1));
    int *cc = (int *)malloc(sizeof(int) * (size_t)(total ? total : 1));
    T *cv = (T *)malloc(sizeof(T) * (size_t)(total ? total : 1));
    if (!rp || !cc || !cv) {
        free(rp); free(cc); free(cv);
        return SWIFT_ERR_ALLOC;
    }
    memcpy(rp, rowptr.data(), sizeof(int) * ((size_t)M + 1));
    std::vector<int> fill((size_t)M, 0);
    for (long long e = 0; e < NZ; e++) {
        int r = ri[e], c = ci[e];
        int p = rp[r] + fill[r]++;
        cc[p] = c;
        cv[p] = (T)vv[e];
        if (mirror && r != c) {
            p = rp[c] + fill[c]++;
            cc[p] = r;
            cv[p] = (T)vv[e];
        }
    }
    *m_out = M;
    *n_out = N;
    *nnz_out = (int)total;
    *sym_out = mirror ? 1 : 0;
    *rowptr_out = rp;
    *colidx_out = cc;
    *val_out = cv;
    return SWIFT_OK;
}


// The C entry points: host allocation failures become SWIFT_ERR_ALLOC
// instead of escaping through C.
template <class T>
int load_c(int *m, int *n, int *nnz, int *sym, int **rp, int **ci, T **v, const char *path,
           int nthreads) {
    try {
        return load<T>(m, n, nnz, sym, rp, ci, v, path, nthreads);
    } catch (const std::bad_alloc &) {
        return SWIFT_ERR_ALLOC;
    }
}

}  // namespace

extern "C" int swift_mmio_allinone(int *m, int *n, int *nnz, int *isSymmetric, int **csrrowptr,
                                   int **csrcolidx, double **csrval, const char *filename) {
    return load_c<double>(m, n, nnz, isSymmetric, csrrowptr, csrcolidx, csrval, filename, 0);
}

extern "C" int swift_mmio_allinone_f32(int *m, int *n, int *nnz, int *isSymmetric,
                                       int **csrrowptr, int **csrcolidx, float **csrval,
                                       const char *filename) {
    return load_c<float>(m, n, nnz, isSymmetric, csrrowptr, csrcolidx, csrval, filename, 0);
}

extern "C" int swift_mmio_allinone_ex(int *m, int *n, int *nnz, int *isSymmetric,
                                      int **csrrowptr, int **csrcolidx, void **csrval,
                                      swift_dtype vtype, const char *filename, int nthreads) {
    if (vtype == SWIFT_F64)
        return load_c<double>(m, n, nnz, isSymmetric, csrrowptr, csrcolidx,
                              reinterpret_cast<double **>(csrval), filename, nthreads);
    if (vtype == SWIFT_F32)
        return load_c<float>(m, n, nnz, isSymmetric, csrrowptr, csrcolidx,
                             reinterpret_cast<float **>(csrval), filename, nthreads);
    return SWIFT_ERR_ARG;
}

extern "C" void swift_host_free(void *p) { free(p); }
