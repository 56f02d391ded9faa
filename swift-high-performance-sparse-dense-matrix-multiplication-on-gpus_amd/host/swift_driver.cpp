// swift_driver.cpp -- the C driver of the Swift SpMM path on one MI355X.
//
// Same flow and command line as the reference driver (src/main.cu:31-491):
//
//   swift_driver -d <device> <A.mtx> [K]
//
// load the Matrix Market file -> random dense B and values -> naive CSR golden
// on the host -> vendor SpMM (rocSPARSE here, cuSPARSE in the reference,
// src/gpu_cusparse_spmm.h) -> Swift preprocessing on the GPU -> every Swift
// kernel variant (Swift_gpu1 / Swift_gpu2 / Swift_GPU / Swift_GPU_3 /
// Swift_GPU_4 counterparts), each verified with the reference's ResultVerify
// rule (|C - golden| > 1e-2 counts as an error, src/auxiliary.h:156) -> the
// minimum time over passing variants, appended as one line in the
// reference's result format (src/main.cu:460).
//
// --mode selects one of the reference's study drivers instead; each appends
// its result line in that driver's format, so the reference's DataProcess
// parsers (which read fixed token positions) read ours:
//   main        the flow above (default)
//   ksweep      K sweep of test/irregular_K (script/Run_irregular_K.sh: K =
//               24 48 96 192 384 768, or --ks), one file per K:
//               "<name> m M n N width K nnz NNZ rocSPARSE T final T"
//   cma         row- vs column-major dense operand (test/k-32/CMA-compare):
//               "... withoutCMA T withCMA T"
//   irropt      irregular part per reside column vs per balance block
//               (test/k-32/irr_opt): "... non-opt T opt T"
//   preprocess  preprocessing cost (test/preprocess/main.cu:440-456):
//               "... sort T ratio R block T ratio R preprocess T final T overall_time T"
//   vendor      rocSPARSE SpMM alone, column- and row-major (the vendor
//               baseline of bench.py)
//   mtxinfo     matrix statistics (test/MTX_info/main.cu:484-493 and
//               test/irr_regular_ratio/main.cu:216): "... rocSPARSE: T
//               sparsity S emptyblkratio E final T" and "<name> nnz NNZ
//               regulatpart R f irregularpart I f" (nonzeros of Swift's two
//               parts here, tiles of its variant tiling there)
//
// Options: --fp32 (default is double, like src/), --seed N (default time(0)
// as in the reference), --reps N, --out FILE (result file; ksweep appends
// _K<k>.txt to it), --ks K1,K2,..., --gen W [--scale S] to use a seeded
// configuration matrix instead of a file.
#include <hip/hip_runtime.h>
#include <rocsparse/rocsparse.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <sys/stat.h>
#include <sys/time.h>
#include <type_traits>
#include <vector>

#include "swift_spmm.h"

#define HIP_OK(x)                                                                   \
    do {                                                                            \
        hipError_t e_ = (x);                                                        \
        if (e_ != hipSuccess) {                                                     \
            fprintf(stderr, "%s:%d %s: %s\n", __FILE__, __LINE__, #x,             \
                    hipGetErrorString(e_));                                         \
            exit(1);                                                                \
        }                                                                           \
    } while (0)
#define RS_OK(x)                                                                    \
    do {                                                                            \
        rocsparse_status s_ = (x);                                                  \
        if (s_ != rocsparse_status_success) {                                       \
            fprintf(stderr, "%s:%d %s: rocsparse status %d\n", __FILE__, __LINE__, #x, \
                    (int)s_);                                                       \
            return -1.0f;                                                           \
        }                                                                           \
    } while (0)

static double now_ms() {
    timeval t;
    gettimeofday(&t, nullptr);
    return t.tv_sec * 1000.0 + t.tv_usec / 1000.0;
}

template <class T> struct Traits;
template <> struct Traits<double> {
    static constexpr swift_dtype st = SWIFT_F64;
    static constexpr rocsparse_datatype rt = rocsparse_datatype_f64_r;
    static const char *name() { return "FP64"; }
};
template <> struct Traits<float> {
    static constexpr swift_dtype st = SWIFT_F32;
    static constexpr rocsparse_datatype rt = rocsparse_datatype_f32_r;
    static const char *name() { return "FP32"; }
};

// dense_mtx2dense_mtx_spmm (src/auxiliary.h:53): rows x cols row-major ->
// column-major.
template <class T>
static void to_col_major(int rows, int cols, const T *in, T *out) {
    size_t k = 0;
    for (int c = 0; c < cols; c++)
        for (int r = 0; r < rows; r++) out[k++] = in[(size_t)r * cols + c];
}

template <class T>
static int verify(const T *got, const T *golden, size_t n) {
    int bad = 0;
    for (size_t i = 0; i < n; i++)
        if (std::fabs((double)golden[i] - (double)got[i]) > 1e-2) bad++;
    return bad;
}

// rocSPARSE CSR SpMM with column-major B/C (the reference's cuSPARSE call,
// src/gpu_cusparse_spmm.h:75-95); returns the median kernel time in ms.
// One rocsparse_spmm configuration on device operands: descriptors, buffer,
// preprocess, then `reps` timed compute stages after 2 warm-ups; the median
// kernel time in ms, or -1 when rocSPARSE rejects the configuration (some
// algorithms take only one dense order).
template <class T>
static float vendor_run(rocsparse_handle h, int m, int n, int nnz, int K, int *drp, int *dci,
                        T *dv, T *dB, T *dC, bool row_major, rocsparse_spmm_alg alg, int reps) {
    rocsparse_spmat_descr A = nullptr;
    rocsparse_dnmat_descr Bd = nullptr, Cd = nullptr;
    void *buf = nullptr;
    float result = -1.0f;
    do {
        if (rocsparse_create_csr_descr(&A, m, n, nnz, drp, dci, dv, rocsparse_indextype_i32,
                                       rocsparse_indextype_i32, rocsparse_index_base_zero,
                                       Traits<T>::rt) != rocsparse_status_success)
            break;
        const rocsparse_order order = row_major ? rocsparse_order_row : rocsparse_order_column;
        if (rocsparse_create_dnmat_descr(&Bd, n, K, row_major ? K : n, dB, Traits<T>::rt, order) !=
                rocsparse_status_success ||
            rocsparse_create_dnmat_descr(&Cd, m, K, row_major ? K : m, dC, Traits<T>::rt, order) !=
                rocsparse_status_success)
            break;
        T alpha = 1, beta = 0;
        size_t bytes = 0;
        if (rocsparse_spmm(h, rocsparse_operation_none, rocsparse_operation_none, &alpha, A, Bd,
                           &beta, Cd, Traits<T>::rt, alg, rocsparse_spmm_stage_buffer_size,
                           &bytes, nullptr) != rocsparse_status_success)
            break;
        HIP_OK(hipMalloc(&buf, std::max<size_t>(bytes, 16)));
        if (rocsparse_spmm(h, rocsparse_operation_none, rocsparse_operation_none, &alpha, A, Bd,
                           &beta, Cd, Traits<T>::rt, alg, rocsparse_spmm_stage_preprocess, &bytes,
                           buf) != rocsparse_status_success)
            break;
        hipEvent_t e0, e1;
        HIP_OK(hipEventCreate(&e0));
        HIP_OK(hipEventCreate(&e1));
        std::vector<float> t;
        bool ok = true;
        for (int r = 0; r < reps + 2 && ok; r++) {
            HIP_OK(hipEventRecord(e0, nullptr));
            ok = rocsparse_spmm(h, rocsparse_operation_none, rocsparse_operation_none, &alpha, A,
                                Bd, &beta, Cd, Traits<T>::rt, alg, rocsparse_spmm_stage_compute,
                                &bytes, buf) == rocsparse_status_success;
            HIP_OK(hipEventRecord(e1, nullptr));
            HIP_OK(hipEventSynchronize(e1));
            float ms;
            HIP_OK(hipEventElapsedTime(&ms, e0, e1));
            if (r >= 2) t.push_back(ms);
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (!ok || t.empty()) break;
        std::sort(t.begin(), t.end());
        result = t[t.size() / 2];
    } while (false);
    if (A) rocsparse_destroy_spmat_descr(A);
    if (Bd) rocsparse_destroy_dnmat_descr(Bd);
    if (Cd) rocsparse_destroy_dnmat_descr(Cd);
    if (buf) (void)hipFree(buf);
    return result;
}

// rocSPARSE SpMM (the reference compares against cuSPARSE SpMM, CSR with
// column-major dense operands, src/gpu_cusparse_spmm.h:75-95) with algorithm
// `alg`; B and C are column-major, or row-major (rocsparse_order_row) when
// row_major.  Returns the median kernel time in ms (-1: not supported) and
// leaves the product in Ccol (host, in the same order as B).
template <class T>
static float vendor_spmm(int m, int n, int nnz, int K, const int *rp, const int *ci,
                         const T *val, const T *Bcol, T *Ccol, int reps, bool row_major = false,
                         rocsparse_spmm_alg alg = rocsparse_spmm_alg_default) {
    int *drp, *dci;
    T *dv, *dB, *dC;
    HIP_OK(hipMalloc(&drp, sizeof(int) * (m + 1)));
    HIP_OK(hipMalloc(&dci, sizeof(int) * std::max(nnz, 1)));
    HIP_OK(hipMalloc(&dv, sizeof(T) * std::max(nnz, 1)));
    HIP_OK(hipMalloc(&dB, sizeof(T) * (size_t)n * K));
    HIP_OK(hipMalloc(&dC, sizeof(T) * (size_t)m * K));
    HIP_OK(hipMemcpy(drp, rp, sizeof(int) * (m + 1), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(dci, ci, sizeof(int) * nnz, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(dv, val, sizeof(T) * nnz, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(dB, Bcol, sizeof(T) * (size_t)n * K, hipMemcpyHostToDevice));
    rocsparse_handle h;
    float t = -1.0f;
    if (rocsparse_create_handle(&h) == rocsparse_status_success) {
        t = vendor_run<T>(h, m, n, nnz, K, drp, dci, dv, dB, dC, row_major, alg, reps);
        rocsparse_destroy_handle(h);
    }
    if (t >= 0) HIP_OK(hipMemcpy(Ccol, dC, sizeof(T) * (size_t)m * K, hipMemcpyDeviceToHost));
    for (void *p : {(void *)drp, (void *)dci, (void *)dv, (void *)dB, (void *)dC})
        (void)hipFree(p);
    return t;
}

struct Args {
    int device = 0;
    std::string path;
    int K = 32;
    bool fp32 = false;
    long long seed = -1;
    int reps = 10;
    int gen = 0;
    double scale = 1.0;
    std::string out;
    std::string mode = "main";
    std::vector<int> ks = {24, 48, 96, 192, 384, 768};  // script/Run_irregular_K.sh
};

// One loaded problem: the pattern, random values (src/main.cu:125-137) and the
// Swift format built on the GPU.
template <class T>
struct Problem {
    std::string name;
    int m = 0, n = 0, nnz = 0;
    int *rp = nullptr, *ci = nullptr;
    std::vector<T> val;
    swift_matrix *M = nullptr;
    swift_info inf{};
    std::vector<int> perm;
    double t_build = 0, t_rowsort = 0;
    ~Problem() {
        if (M) swift_destroy(M);
        swift_host_free(rp);
        swift_host_free(ci);
    }
};

template <class T>
static int load(const Args &a, Problem<T> &P) {
    int sym = 0;
    double *v64 = nullptr;
    if (a.gen) {
        int rc = swift_generate(a.gen, 0, a.scale, 0, &P.m, &P.n, &P.nnz, &P.rp, &P.ci, &v64,
                                nullptr);
        if (rc) { fprintf(stderr, "swift_generate: %d\n", rc); return 1; }
        P.name = "config" + std::to_string(a.gen);
    } else {
        int rc = swift_mmio_allinone(&P.m, &P.n, &P.nnz, &sym, &P.rp, &P.ci, &v64, a.path.c_str());
        if (rc) {
            fprintf(stderr, "mmio_allinone(%s) failed: %d\n", a.path.c_str(), rc);
            return 1;
        }
        P.name = a.path.substr(a.path.find_last_of("/\\") + 1);
        P.name = P.name.substr(0, P.name.find_last_of('.'));
    }
    swift_host_free(v64);
    // random values (src/main.cu:125-137); the dense operands follow per K
    unsigned seed = a.seed >= 0 ? (unsigned)a.seed : (unsigned)time(0);
    srand(seed);
    P.val.resize(P.nnz);
    for (auto &x : P.val) x = T(rand() % 1000 - 500) / 1000;
    printf("read success,input matrix %s :(%i,%i) nnz =%i  %s\n", P.name.c_str(), P.m, P.n,
           P.nnz, Traits<T>::name());
    // Swift preprocessing on the GPU
    double t0 = now_ms();
    int rc = swift_build_host(&P.M, P.m, P.n, P.nnz, Traits<T>::st, P.rp, P.ci, P.val.data());
    P.t_build = now_ms() - t0;
    if (rc) {
        fprintf(stderr, "swift_build: %s\n", swift_last_error());
        return 1;
    }
    t0 = now_ms();
    rc = swift_row_idx_sort(P.M, nullptr);
    HIP_OK(hipDeviceSynchronize());
    P.t_rowsort = now_ms() - t0;
    if (rc) {
        fprintf(stderr, "swift_row_idx_sort: %s\n", swift_last_error());
        return 1;
    }
    swift_get_info(P.M, &P.inf);
    printf("Swift format: tilenum %d regular_nnz %d reside_col %d reside_nnz %d numblk %d "
           "(build %.2f ms: sort %.2f block %.2f schedules %.2f; row sort %.2f ms)\n",
           P.inf.tilenum, P.inf.regular_nnz, P.inf.reside_col, P.inf.reside_nnz, P.inf.numblk,
           P.t_build, P.inf.build_ms[0], P.inf.build_ms[1], P.inf.build_ms[2], P.t_rowsort);
    P.perm.resize(P.n);
    if (P.n) HIP_OK(hipMemcpy(P.perm.data(), P.inf.perm, sizeof(int) * P.n, hipMemcpyDeviceToHost));
    return 0;
}

// Dense operands of one K: random B (src/main.cu:125-130), its column-sorted
// copy (Pre4sortedx) in both layouts, and the naive CSR golden
// (src/main.cu:145-155) in both layouts.
template <class T>
struct Operands {
    int K = 0;
    std::vector<T> B, Bcol, sB, sBcol, golden, Gcol;
};

template <class T>
static void make_operands(const Problem<T> &P, int K, Operands<T> &o) {
    const int m = P.m, n = P.n;
    o.K = K;
    o.B.resize((size_t)n * K);
    for (auto &x : o.B) x = T(rand() % 100 - 50) / 100;
    o.golden.assign((size_t)m * K, T(0));
    for (int i = 0; i < m; i++)
        for (int e = P.rp[i]; e < P.rp[i + 1]; e++)
            for (int k = 0; k < K; k++)
                o.golden[(size_t)i * K + k] += P.val[e] * o.B[(size_t)P.ci[e] * K + k];
    o.Bcol.resize((size_t)n * K);
    o.Gcol.resize((size_t)m * K);
    to_col_major(n, K, o.B.data(), o.Bcol.data());
    to_col_major(m, K, o.golden.data(), o.Gcol.data());
    o.sB.resize((size_t)n * K);
    o.sBcol.resize((size_t)n * K);
    for (int i = 0; i < n; i++)
        std::copy(&o.B[(size_t)P.perm[i] * K], &o.B[(size_t)P.perm[i] * K] + K,
                  &o.sB[(size_t)i * K]);
    to_col_major(n, K, o.sB.data(), o.sBcol.data());
}

// rocSPARSE (col-major) with ResultVerify; -1111 when it disagrees.
template <class T>
static float vendor_time(const Problem<T> &P, const Operands<T> &o, int reps) {
    std::vector<T> Ccol((size_t)P.m * o.K);
    float t = vendor_spmm<T>(P.m, P.n, P.nnz, o.K, P.rp, P.ci, P.val.data(), o.Bcol.data(),
                             Ccol.data(), reps);
    if (t < 0) {
        printf("error rocSPARSE SpMM: not supported\n");
        return -1111;
    }
    int bad = verify(Ccol.data(), o.Gcol.data(), Ccol.size());
    if (bad) {
        printf("error rocSPARSE SpMM, error = %d\n", bad);
        return -1111;
    }
    printf("success rocSPARSE SpMM, process:%f ms\n", t);
    return t;
}

// rocSPARSE with algorithm alg in either dense order, verified against the
// host golden (ResultVerify's rule); -1 when rocSPARSE rejects the
// configuration or the result disagrees.
template <class T>
static float vendor_time_alg(const Problem<T> &P, const Operands<T> &o, int reps, bool row_major,
                             rocsparse_spmm_alg alg) {
    std::vector<T> C((size_t)P.m * o.K);
    const float t = vendor_spmm<T>(P.m, P.n, P.nnz, o.K, P.rp, P.ci, P.val.data(),
                                   row_major ? o.B.data() : o.Bcol.data(), C.data(), reps,
                                   row_major, alg);
    if (t < 0) return -1;
    const int bad = verify(C.data(), row_major ? o.golden.data() : o.Gcol.data(), C.size());
    return bad ? -1 : t;
}

// One Swift variant on host pointers (swift_spmm_host: kernel-only events,
// median, ResultVerify); -1 when it fails the check.
template <class T>
static float swift_time(const Problem<T> &P, const Operands<T> &o, swift_layout layout,
                        swift_kernel kernel, int reps, const char *label) {
    const bool row = layout == SWIFT_ROW_MAJOR;
    std::vector<T> C((size_t)P.m * o.K);
    float ms = -1;
    double gf = 0;
    int rc = swift_spmm_host(P.M, o.K, layout, Traits<T>::st, kernel,
                             row ? o.sB.data() : o.sBcol.data(), C.data(),
                             row ? o.golden.data() : o.Gcol.data(), 2, reps, &ms, &gf);
    if (rc == SWIFT_OK) {
        printf("success %-44s %10.4f ms %9.2f GFLOP/s\n", label, ms, gf);
        return ms;
    }
    printf("error %-46s %s\n", label, swift_last_error());
    return -1;
}

static float min_passing(std::initializer_list<float> ts) {
    float best = -1;
    for (float t : ts)
        if (t >= 0 && (best < 0 || t < best)) best = t;
    return best;
}

// Appends one line to the result file, creating its directory (the
// reference's drivers write into ./data/).
static void append(const std::string &path, const char *fmt, ...) {
    const size_t slash = path.find_last_of('/');
    if (slash != std::string::npos && slash > 0) {
        std::string dir = path.substr(0, slash);
        for (size_t i = 1; i <= dir.size(); i++)
            if (i == dir.size() || dir[i] == '/') mkdir(dir.substr(0, i).c_str(), 0755);
    }
    FILE *f = fopen(path.c_str(), "a");
    if (!f) {
        fprintf(stderr, "Writing results fails (%s).\n", path.c_str());
        return;
    }
    va_list ap;
    va_start(ap, fmt);
    vfprintf(f, fmt, ap);
    va_end(ap);
    fclose(f);
}

template <class T>
static std::string out_path(const Args &a, const std::string &stem) {
    return a.out.empty() ? "./data/" + stem + "_" + Traits<T>::name() + ".txt" : a.out;
}

// main: every variant, minimum over the passing ones (src/main.cu:443-461).
template <class T>
static int mode_main(const Args &a, Problem<T> &P) {
    Operands<T> o;
    make_operands(P, a.K, o);
    const float t_vendor = vendor_time(P, o, a.reps);
    struct Variant {
        const char *name;
        swift_layout layout;
        swift_kernel kernel;
    } variants[] = {
        {"Swift_gpu1 (col-major, unbalanced)", SWIFT_COL_MAJOR, SWIFT_KERNEL_UNBALANCED},
        {"Swift_gpu2 (row-major, balanced)", SWIFT_ROW_MAJOR, SWIFT_KERNEL_BALANCED},
        {"Swift_GPU (col-major, balanced)", SWIFT_COL_MAJOR, SWIFT_KERNEL_BALANCED},
        {"Swift_GPU_3 (col-major, K chunks)", SWIFT_COL_MAJOR, SWIFT_KERNEL_KCHUNK},
        {"Swift_GPU_4 (col-major, row-sorted groups)", SWIFT_COL_MAJOR, SWIFT_KERNEL_ROWSORT},
        {"Swift row-major default", SWIFT_ROW_MAJOR, SWIFT_KERNEL_DEFAULT},
    };
    float best = -1;
    for (auto &v : variants) best = min_passing({best, swift_time(P, o, v.layout, v.kernel, a.reps, v.name)});
    const std::string out = a.out.empty() ? "./data/results_Swift_" + std::string(Traits<T>::name()) +
                                                "_" + std::to_string(a.K) + ".txt"
                                          : a.out;
    if (best >= 0) {
        printf("Swift GPU SpMM time %f ms\n", best);
        append(out, "%s m %d n %d width %d nnz %d rocSPARSE: %f Swift %f \n", P.name.c_str(), P.m,
               P.n, a.K, P.nnz, t_vendor, best);
        return 0;
    }
    printf("Swift GPU SpMM Check NO PASS!\n");
    append(out, "erro Swift (dense mtx col-major)%s \n", P.name.c_str());
    return 2;
}

// ksweep: test/irregular_K, one result file per K (FP32 and FP64 label
// formats of test/irregular_K/FP32/main.cu:482 and FP64/main.cu:467).
template <class T>
static int mode_ksweep(const Args &a, Problem<T> &P) {
    int rc = 0;
    const std::string base = a.out.empty() ? "./data/Swift_irregular_K_" + std::string(Traits<T>::name())
                                           : a.out;
    for (int K : a.ks) {
        printf("---- K = %d\n", K);
        Operands<T> o;
        make_operands(P, K, o);
        const float tv = vendor_time(P, o, a.reps);
        const float tr = swift_time(P, o, SWIFT_ROW_MAJOR, SWIFT_KERNEL_DEFAULT, a.reps,
                                    "Swift row-major default");
        const float tc = swift_time(P, o, SWIFT_COL_MAJOR, SWIFT_KERNEL_DEFAULT, a.reps,
                                    "Swift col-major default");
        const float best = min_passing({tr, tc});
        const std::string out = base + "_K" + std::to_string(K) + ".txt";
        if (best >= 0) {
            if (std::is_same<T, float>::value)
                append(out, "%s m %d n %d width %d nnz %d rocSPARSE %f final %f \n", P.name.c_str(),
                       P.m, P.n, K, P.nnz, tv, best);
            else
                append(out, "%s m %d n %d width %d nnz %d rocSPARSE: %f final %f \n",
                       P.name.c_str(), P.m, P.n, K, P.nnz, tv, best);
        } else {
            append(out, "erro Swift (dense mtx col-major)%s \n", P.name.c_str());
            rc = 2;
        }
    }
    return rc;
}

// cma: the reference's Swift kernels with a row-major dense operand
// (Swift_gpu2 order, "withoutCMA") against the column-major one (Swift_GPU,
// "withCMA"), test/k-32/CMA-compare/main.cu:316-357; the tuned kernels of
// both layouts are printed beside them.
template <class T>
static int mode_cma(const Args &a, Problem<T> &P) {
    Operands<T> o;
    make_operands(P, a.K, o);
    const float t_row = swift_time(P, o, SWIFT_ROW_MAJOR, SWIFT_KERNEL_BALANCED, a.reps,
                                   "Swift_gpu2 (row-major, balanced)");
    const float t_col = swift_time(P, o, SWIFT_COL_MAJOR, SWIFT_KERNEL_BALANCED, a.reps,
                                   "Swift_GPU (col-major, balanced)");
    swift_time(P, o, SWIFT_ROW_MAJOR, SWIFT_KERNEL_DEFAULT, a.reps, "Swift row-major default");
    swift_time(P, o, SWIFT_COL_MAJOR, SWIFT_KERNEL_DEFAULT, a.reps, "Swift col-major default");
    append(out_path<T>(a, "Swift_CMA_compare"), "%s m %d n %d width %d nnz %d withoutCMA %f withCMA %f \n",
           P.name.c_str(), P.m, P.n, a.K, P.nnz, t_row, t_col);
    return (t_row >= 0 && t_col >= 0) ? 0 : 2;
}

// irropt: the irregular (reside) part per reside column (Swift_gpu1's
// unbalanced launch) against per balance block (Swift_GPU), kernel time of
// that part alone (swift_spmm_timed phase 2, median), each product checked
// with ResultVerify; test/k-32/irr_opt/main.cu:300-372.
template <class T>
static int mode_irropt(const Args &a, Problem<T> &P) {
    Operands<T> o;
    make_operands(P, a.K, o);
    const int K = a.K;
    T *dB, *dC;
    HIP_OK(hipMalloc(&dB, sizeof(T) * std::max<size_t>(1, (size_t)P.n * K)));
    HIP_OK(hipMalloc(&dC, sizeof(T) * std::max<size_t>(1, (size_t)P.m * K)));
    HIP_OK(hipMemcpy(dB, o.sBcol.data(), sizeof(T) * (size_t)P.n * K, hipMemcpyHostToDevice));
    std::vector<T> C((size_t)P.m * K);
    float t[2] = {-1, -1};
    const swift_kernel ks[2] = {SWIFT_KERNEL_UNBALANCED, SWIFT_KERNEL_BALANCED};
    for (int v = 0; v < 2; v++) {
        std::vector<float> r;
        int rc = SWIFT_OK;
        for (int it = 0; it < a.reps + 2 && rc == SWIFT_OK; it++) {
            float ms[3];
            rc = swift_spmm_timed(P.M, K, SWIFT_COL_MAJOR, Traits<T>::st, ks[v], SWIFT_B_SORTED,
                                  dB, dC, nullptr, ms);
            if (it >= 2) r.push_back(ms[2]);
        }
        if (rc != SWIFT_OK) {
            printf("error %s\n", swift_last_error());
            continue;
        }
        HIP_OK(hipMemcpy(C.data(), dC, sizeof(T) * C.size(), hipMemcpyDeviceToHost));
        const int bad = verify(C.data(), o.Gcol.data(), C.size());
        std::sort(r.begin(), r.end());
        t[v] = bad ? -1 : r[r.size() / 2];
        printf("%s Swift GPU SpMM irregular part %s: %f ms\n", bad ? "error" : "success",
               v ? "with optimization (balance blocks)" : "without optimization (per column)",
               t[v]);
    }
    (void)hipFree(dB);
    (void)hipFree(dC);
    if (t[0] >= 0 && t[1] >= 0) {
        append(out_path<T>(a, "Swift_irr_opt"), "%s m %d n %d width %d nnz %d non-opt %f opt %f\n",
               P.name.c_str(), P.m, P.n, K, P.nnz, t[0], t[1]);
        return 0;
    }
    append(out_path<T>(a, "Swift_irr_opt"), "erro Swift (dense mtx col-major)%s \n", P.name.c_str());
    return 2;
}

// preprocess: the build's phase times (column sort; blocking = staircase
// tiles, reside part, categorisation, balance and the kernel schedules)
// and the product, test/preprocess/main.cu:440-456.
template <class T>
static int mode_preprocess(const Args &a, Problem<T> &P) {
    Operands<T> o;
    make_operands(P, a.K, o);
    const float final_ms = swift_time(P, o, SWIFT_ROW_MAJOR, SWIFT_KERNEL_DEFAULT, a.reps,
                                      "Swift row-major default");
    const double sort = P.inf.build_ms[0], block = P.inf.build_ms[1] + P.inf.build_ms[2];
    const double pre = sort + block;
    printf("Sort %f ms, block %f ms\nPreprocess %f ms, Process %f ms\n", sort, block, pre, final_ms);
    if (final_ms < 0) {
        append(out_path<T>(a, "Swift_preprocess"), "erro Swift %s \n", P.name.c_str());
        return 2;
    }
    append(out_path<T>(a, "Swift_preprocess"),
           "%s m %d n %d nnz %d sort %f ratio %f block %f ratio %f preprocess %f final %f "
           "overall_time %f \n",
           P.name.c_str(), P.m, P.n, P.nnz, sort, pre > 0 ? sort / pre : 0.0, block,
           pre > 0 ? block / pre : 0.0, pre, final_ms, pre + final_ms);
    return 0;
}

// vendor: rocSPARSE SpMM alone, every CSR algorithm in both dense orders,
// each verified against the host golden; one line per algorithm, then the
// best per order:
// "vendor rocSPARSE m M n N width K nnz NNZ row-major T col-major T
//  row-alg NAME col-alg NAME" (T = -1: no algorithm passed).
template <class T>
static int mode_vendor(const Args &a, Problem<T> &P) {
    Operands<T> o;
    make_operands(P, a.K, o);
    const struct {
        const char *name;
        rocsparse_spmm_alg alg;
    } algs[] = {{"default", rocsparse_spmm_alg_default},
                {"csr", rocsparse_spmm_alg_csr},
                {"csr_row_split", rocsparse_spmm_alg_csr_row_split},
                {"csr_nnz_split", rocsparse_spmm_alg_csr_nnz_split},
                {"csr_merge_path", rocsparse_spmm_alg_csr_merge_path}};
    float best[2] = {-1, -1};
    const char *best_name[2] = {"none", "none"};
    for (const auto &g : algs) {
        float t[2];
        for (int r = 0; r < 2; r++) {
            t[r] = vendor_time_alg(P, o, a.reps, r == 0, g.alg);
            if (t[r] >= 0 && (best[r] < 0 || t[r] < best[r])) {
                best[r] = t[r];
                best_name[r] = g.name;
            }
        }
        printf("vendor-alg %s row-major %f col-major %f\n", g.name, t[0], t[1]);
    }
    printf("vendor rocSPARSE m %d n %d width %d nnz %d row-major %f col-major %f row-alg %s "
           "col-alg %s\n",
           P.m, P.n, a.K, P.nnz, best[0], best[1], best_name[0], best_name[1]);
    return (best[0] < 0 || best[1] < 0) ? 2 : 0;
}

// mtxinfo: sparsity and the share of empty 32x32 blocks with the product
// times (test/MTX_info/main.cu:484-493), and the sizes of Swift's regular
// and irregular parts (test/irr_regular_ratio/main.cu:216).
template <class T>
static int mode_mtxinfo(const Args &a, Problem<T> &P) {
    swift_stats st;
    if (swift_matrix_stats(P.M, &st, nullptr)) {
        fprintf(stderr, "swift_matrix_stats: %s\n", swift_last_error());
        return 1;
    }
    printf("sparsity %f empty block ratio %f (%lld of %lld blocks hold entries); regular %d "
           "(%.4f) irregular %d; tile formats %d/%d/%d; rows <= %d, columns <= %d\n",
           st.sparsity, st.empty_block_ratio, st.nonempty_blocks, st.blocks, st.regular_nnz,
           st.regular_ratio, st.reside_nnz, st.tile_format_count[0], st.tile_format_count[1],
           st.tile_format_count[2], st.max_row_nnz, st.max_col_nnz);
    printf("dense blocks (>= 512 of 1024 entries) %lld holding %lld entries (%.4f of nnz); "
           "fullest block %d entries\n",
           st.dense_blocks, st.dense_block_nnz,
           P.nnz ? (double)st.dense_block_nnz / P.nnz : 0.0, st.max_block_nnz);
    Operands<T> o;
    make_operands(P, a.K, o);
    const float tv = vendor_time(P, o, a.reps);
    const float best = min_passing({swift_time(P, o, SWIFT_ROW_MAJOR, SWIFT_KERNEL_DEFAULT, a.reps,
                                               "Swift row-major default"),
                                    swift_time(P, o, SWIFT_COL_MAJOR, SWIFT_KERNEL_DEFAULT, a.reps,
                                               "Swift col-major default")});
    const std::string out = out_path<T>(a, "Swift_MTX_info");
    if (best >= 0)
        append(out, "%s m %d n %d width %d nnz %d rocSPARSE: %f sparsity %f emptyblkratio %f "
                    "final %f \n",
               P.name.c_str(), P.m, P.n, a.K, P.nnz, tv, st.sparsity, st.empty_block_ratio, best);
    else
        append(out, "erro Swift (dense mtx col-major)%s \n", P.name.c_str());
    const int sum = st.regular_nnz + st.reside_nnz;
    std::string ratio = out.substr(0, out.find_last_of('.')) + "_irr_regular_ratio.txt";
    append(ratio, "%s nnz %d regulatpart %d %f irregularpart %d %f \n", P.name.c_str(), P.nnz,
           st.regular_nnz, sum ? (double)st.regular_nnz / sum : 0.0, st.reside_nnz,
           sum ? (double)st.reside_nnz / sum : 0.0);
    return best >= 0 ? 0 : 2;
}

template <class T>
static int run(const Args &a) {
    Problem<T> P;
    if (int rc = load(a, P)) return rc;
    if (a.mode == "main") return mode_main(a, P);
    if (a.mode == "ksweep") return mode_ksweep(a, P);
    if (a.mode == "cma") return mode_cma(a, P);
    if (a.mode == "irropt") return mode_irropt(a, P);
    if (a.mode == "preprocess") return mode_preprocess(a, P);
    if (a.mode == "mtxinfo") return mode_mtxinfo(a, P);
    if (a.mode == "vendor") return mode_vendor(a, P);
    fprintf(stderr, "unknown --mode %s\n", a.mode.c_str());
    return 2;
}

int main(int argc, char **argv) {
    Args a;
    std::vector<std::string> pos;
    for (int i = 1; i < argc; i++) {
        std::string s = argv[i];
        auto next = [&]() -> std::string {
            if (i + 1 >= argc) { fprintf(stderr, "missing value for %s\n", s.c_str()); exit(2); }
            return argv[++i];
        };
        if (s == "-d") a.device = atoi(next().c_str());
        else if (s == "--fp32") a.fp32 = true;
        else if (s == "--fp64") a.fp32 = false;
        else if (s == "--seed") a.seed = atoll(next().c_str());
        else if (s == "--reps") a.reps = std::max(1, atoi(next().c_str()));
        else if (s == "--out") a.out = next();
        else if (s == "--gen") a.gen = atoi(next().c_str());
        else if (s == "--scale") a.scale = atof(next().c_str());
        else if (s == "--mode") a.mode = next();
        else if (s == "--ks") {
            a.ks.clear();
            std::string v = next();
            for (size_t p = 0; p < v.size();) {
                size_t q = v.find(',', p);
                if (q == std::string::npos) q = v.size();
                const int k = atoi(v.substr(p, q - p).c_str());
                if (k <= 0) { fprintf(stderr, "--ks: bad K in %s\n", v.c_str()); return 2; }
                a.ks.push_back(k);
                p = q + 1;
            }
        } else pos.push_back(s);
    }
    if (!a.gen) {
        if (pos.empty()) {
            fprintf(stderr, "usage: %s -d <device> <A.mtx> [K] [--fp32] [--seed N] [--reps N] "
                            "[--out FILE] [--mode main|ksweep|cma|irropt|preprocess|mtxinfo] "
                            "[--ks K1,K2,..] | --gen W [--scale S] [K]\n", argv[0]);
            return 2;
        }
        a.path = pos[0];
        if (pos.size() > 1) a.K = atoi(pos[1].c_str());
    } else if (!pos.empty()) {
        a.K = atoi(pos[0].c_str());
    }
    if (a.K <= 0) { fprintf(stderr, "K must be positive\n"); return 2; }
    HIP_OK(hipSetDevice(a.device));
    hipDeviceProp_t p;
    HIP_OK(hipGetDeviceProperties(&p, a.device));
    printf("---------------------------------------------------------------------------------------------\n");
    printf("Device [ %i ] %s (%s) @ %4.2f MHz, %s, mode %s\n", a.device, p.name, p.gcnArchName,
           p.clockRate * 1e-3f, swift_version(), a.mode.c_str());
    return a.fp32 ? run<float>(a) : run<double>(a);
}
