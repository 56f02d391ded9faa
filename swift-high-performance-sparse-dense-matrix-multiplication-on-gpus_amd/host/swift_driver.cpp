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
// Extra options: --fp32 (default is double, like src/), --seed N (default
// time(0) as in the reference), --reps N, --out FILE, --gen W [--scale S] to
// use a seeded configuration matrix instead of a file.
#include <hip/hip_runtime.h>
#include <rocsparse/rocsparse.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <sys/time.h>
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
template <class T>
static float vendor_spmm(int m, int n, int nnz, int K, const int *rp, const int *ci,
                         const T *val, const T *Bcol, T *Ccol, int reps) {
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
    RS_OK(rocsparse_create_handle(&h));
    rocsparse_spmat_descr A;
    rocsparse_dnmat_descr Bd, Cd;
    RS_OK(rocsparse_create_csr_descr(&A, m, n, nnz, drp, dci, dv, rocsparse_indextype_i32,
                                     rocsparse_indextype_i32, rocsparse_index_base_zero,
                                     Traits<T>::rt));
    RS_OK(rocsparse_create_dnmat_descr(&Bd, n, K, n, dB, Traits<T>::rt, rocsparse_order_column));
    RS_OK(rocsparse_create_dnmat_descr(&Cd, m, K, m, dC, Traits<T>::rt, rocsparse_order_column));
    T alpha = 1, beta = 0;
    size_t bytes = 0;
    RS_OK(rocsparse_spmm(h, rocsparse_operation_none, rocsparse_operation_none, &alpha, A, Bd,
                         &beta, Cd, Traits<T>::rt, rocsparse_spmm_alg_default,
                         rocsparse_spmm_stage_buffer_size, &bytes, nullptr));
    void *buf = nullptr;
    HIP_OK(hipMalloc(&buf, std::max<size_t>(bytes, 16)));
    RS_OK(rocsparse_spmm(h, rocsparse_operation_none, rocsparse_operation_none, &alpha, A, Bd,
                         &beta, Cd, Traits<T>::rt, rocsparse_spmm_alg_default,
                         rocsparse_spmm_stage_preprocess, &bytes, buf));
    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    std::vector<float> t;
    for (int r = 0; r < reps + 2; r++) {
        HIP_OK(hipEventRecord(e0, nullptr));
        RS_OK(rocsparse_spmm(h, rocsparse_operation_none, rocsparse_operation_none, &alpha, A,
                             Bd, &beta, Cd, Traits<T>::rt, rocsparse_spmm_alg_default,
                             rocsparse_spmm_stage_compute, &bytes, buf));
        HIP_OK(hipEventRecord(e1, nullptr));
        HIP_OK(hipEventSynchronize(e1));
        float ms;
        HIP_OK(hipEventElapsedTime(&ms, e0, e1));
        if (r >= 2) t.push_back(ms);
    }
    HIP_OK(hipMemcpy(Ccol, dC, sizeof(T) * (size_t)m * K, hipMemcpyDeviceToHost));
    rocsparse_destroy_spmat_descr(A);
    rocsparse_destroy_dnmat_descr(Bd);
    rocsparse_destroy_dnmat_descr(Cd);
    rocsparse_destroy_handle(h);
    for (void *p : {buf, (void *)drp, (void *)dci, (void *)dv, (void *)dB, (void *)dC})
        (void)hipFree(p);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    std::sort(t.begin(), t.end());
    return t[t.size() / 2];
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
};

template <class T>
static int run(const Args &a) {
    int m, n, nnz, sym = 0;
    int *rp = nullptr, *ci = nullptr;
    double *v64 = nullptr;
    std::string name;
    if (a.gen) {
        int rc = swift_generate(a.gen, 0, a.scale, 0, &m, &n, &nnz, &rp, &ci, &v64, nullptr);
        if (rc) { fprintf(stderr, "swift_generate: %d\n", rc); return 1; }
        name = "config" + std::to_string(a.gen);
    } else {
        int rc = swift_mmio_allinone(&m, &n, &nnz, &sym, &rp, &ci, &v64, a.path.c_str());
        if (rc) {
            fprintf(stderr, "mmio_allinone(%s) failed: %d\n", a.path.c_str(), rc);
            return 1;
        }
        name = a.path.substr(a.path.find_last_of("/\\") + 1);
        name = name.substr(0, name.find_last_of('.'));
    }
    const int K = a.K;
    printf("read success,input matrix %s :(%i,%i) nnz =%i  width=%d  %s\n", name.c_str(), m, n,
           nnz, K, Traits<T>::name());
    // random dense matrix and values (src/main.cu:125-137)
    unsigned seed = a.seed >= 0 ? (unsigned)a.seed : (unsigned)time(0);
    srand(seed);
    std::vector<T> B((size_t)n * K), val(nnz);
    for (auto &x : B) x = T(rand() % 100 - 50) / 100;
    for (auto &x : val) x = T(rand() % 1000 - 500) / 1000;
    swift_host_free(v64);
    // golden (src/main.cu:145-155), value type T, row-major
    std::vector<T> golden((size_t)m * K, T(0));
    for (int i = 0; i < m; i++)
        for (int e = rp[i]; e < rp[i + 1]; e++)
            for (int k = 0; k < K; k++)
                golden[(size_t)i * K + k] += val[e] * B[(size_t)ci[e] * K + k];
    // vendor SpMM, column-major
    std::vector<T> Bcol((size_t)n * K), Ccol((size_t)m * K), Gcol((size_t)m * K);
    to_col_major(n, K, B.data(), Bcol.data());
    to_col_major(m, K, golden.data(), Gcol.data());
    float t_vendor = vendor_spmm<T>(m, n, nnz, K, rp, ci, val.data(), Bcol.data(), Ccol.data(),
                                    a.reps);
    int bad = verify(Ccol.data(), Gcol.data(), Ccol.size());
    if (bad) {
        printf("error rocSPARSE SpMM, error = %d\n", bad);
        t_vendor = -1111;
    } else {
        printf("success rocSPARSE SpMM, process:%f ms\n", t_vendor);
    }
    // Swift preprocessing on the GPU
    double t0 = now_ms();
    swift_matrix *M = nullptr;
    int rc = swift_build_host(&M, m, n, nnz, Traits<T>::st, rp, ci, val.data());
    double t_build = now_ms() - t0;
    if (rc) {
        fprintf(stderr, "swift_build: %s\n", swift_last_error());
        return 1;
    }
    t0 = now_ms();
    rc = swift_row_idx_sort(M, nullptr);
    HIP_OK(hipDeviceSynchronize());
    double t_rowsort = now_ms() - t0;
    if (rc) {
        fprintf(stderr, "swift_row_idx_sort: %s\n", swift_last_error());
        return 1;
    }
    swift_info inf;
    swift_get_info(M, &inf);
    printf("Swift format: tilenum %d regular_nnz %d reside_col %d reside_nnz %d numblk %d "
           "(build %.2f ms, row sort %.2f ms)\n", inf.tilenum, inf.regular_nnz, inf.reside_col,
           inf.reside_nnz, inf.numblk, t_build, t_rowsort);
    // sorted dense matrix (Pre4sortedx) in both layouts
    std::vector<int> perm(n);
    if (n) HIP_OK(hipMemcpy(perm.data(), inf.perm, sizeof(int) * n, hipMemcpyDeviceToHost));
    std::vector<T> sB((size_t)n * K), sBcol((size_t)n * K);
    for (int i = 0; i < n; i++)
        std::copy(&B[(size_t)perm[i] * K], &B[(size_t)perm[i] * K] + K, &sB[(size_t)i * K]);
    to_col_major(n, K, sB.data(), sBcol.data());
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
    std::vector<T> C((size_t)m * K);
    float best = -1;
    for (auto &v : variants) {
        bool row = v.layout == SWIFT_ROW_MAJOR;
        float ms = -1;
        double gf = 0;
        rc = swift_spmm_host(M, K, v.layout, Traits<T>::st, v.kernel,
                             row ? sB.data() : sBcol.data(), C.data(),
                             row ? golden.data() : Gcol.data(), 2, a.reps, &ms, &gf);
        if (rc == SWIFT_OK) {
            printf("success %-44s %10.4f ms %9.2f GFLOP/s\n", v.name, ms, gf);
            if (best < 0 || ms < best) best = ms;
        } else {
            printf("error %-46s %s\n", v.name, swift_last_error());
        }
    }
    swift_destroy(M);
    std::string out = a.out.empty() ? "./data/results_Swift_" + std::string(Traits<T>::name()) +
                                          "_" + std::to_string(K) + ".txt"
                                    : a.out;
    FILE *f = fopen(out.c_str(), "a");
    if (best >= 0) {
        printf("Swift GPU SpMM time %f ms\n", best);
        if (f) fprintf(f, "%s m %d n %d width %d nnz %d rocSPARSE: %f Swift %f \n", name.c_str(),
                       m, n, K, nnz, t_vendor, best);
    } else {
        printf("Swift GPU SpMM Check NO PASS!\n");
        if (f) fprintf(f, "erro Swift (dense mtx col-major)%s \n", name.c_str());
    }
    if (f) fclose(f);
    else fprintf(stderr, "Writing results fails (%s).\n", out.c_str());
    swift_host_free(rp);
    swift_host_free(ci);
    return best >= 0 ? 0 : 2;
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
        else pos.push_back(s);
    }
    if (!a.gen) {
        if (pos.empty()) {
            fprintf(stderr, "usage: %s -d <device> <A.mtx> [K] [--fp32] [--seed N] [--reps N] "
                            "[--out FILE] | --gen W [--scale S] [K]\n", argv[0]);
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
    printf("Device [ %i ] %s (%s) @ %4.2f MHz, %s\n", a.device, p.name, p.gcnArchName,
           p.clockRate * 1e-3f, swift_version());
    return a.fp32 ? run<float>(a) : run<double>(a);
}
