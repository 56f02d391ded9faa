// abi.hip -- the extern "C" entry points of include/swift_spmm.h.
//
// Argument checking, stream/device handling and the host-pointer launcher
// live here; the kernels are in preprocess.hip, spmm.hip, backward.hip and
// batched.hip.
#include "swift_internal.h"

#include <algorithm>
#include <cmath>
#include <new>
#include <vector>

namespace swift {
int build_impl(swift_matrix *M, const int *d_rowptr, const int *d_colidx, const void *d_val,
               hipStream_t s);
int row_sort_impl(swift_matrix *M, hipStream_t s);
int build_segments(swift_matrix *M, int W, hipStream_t s);
int permute_dense(const swift_matrix *M, int K, swift_layout layout, size_t es, const void *in,
                  void *out, hipStream_t s);
void destroy_matrix(swift_matrix *M);
int csr_to_csc_entry(int m, int n, int nnz, swift_dtype vtype, const void *d_val,
                     const int *d_rowptr, const int *d_colidx, void *d_cscval, int *d_cscrow,
                     int *d_cscptr, int *d_npc, hipStream_t s);
int col_sort_entry(int n, int nnz, int K, swift_dtype vtype, const int *d_npc,
                   const int *d_cscptr, const int *d_cscrow, const void *d_cscval, int *d_srow,
                   void *d_sval, int *d_sortnnz, int *d_perm, const void *d_B, void *d_sortB,
                   hipStream_t s);
int spmm_impl(const swift_matrix *M, int K, swift_layout layout, swift_dtype dtype,
              swift_kernel kernel, unsigned flags, const void *B, void *C, const float *kscale,
              hipStream_t s);
int spmm_timed_impl(const swift_matrix *M, int K, swift_layout layout, swift_dtype dtype,
                    swift_kernel kernel, unsigned flags, const void *B, void *C,
                    const float *kscale, hipStream_t s, float *ms3);
int spmm_noalloc_impl(const swift_matrix *M, int K, swift_layout layout, swift_dtype dtype,
                      swift_kernel kernel, unsigned flags, const void *B, void *C,
                      hipStream_t s);
int prepare_values_impl(swift_matrix *M, swift_dtype dtype, hipStream_t s);
int convert_dense_impl(const float *in, void *out, size_t rows, int K, swift_layout layout,
                       swift_dtype dtype, float *col_scale, hipStream_t s);
int build_transpose_impl(const swift_matrix *A, swift_matrix **out, hipStream_t s);
int backward_impl(const swift_matrix *AT, int K, swift_dtype dtype, const void *dC, const void *B,
                  void *dB, void *dvals, hipStream_t s);
int sddmm_impl(const swift_matrix *A, int K, swift_dtype dtype, const void *dC, const void *B,
               void *dvals, hipStream_t s);
int update_values_impl(swift_matrix *M, const void *csr, hipStream_t s);
int matrix_stats_impl(const swift_matrix *M, swift_stats *o, hipStream_t s);
}  // namespace swift

using namespace swift;

static inline hipStream_t S(swift_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

#define SWIFT_GUARD(...)                                                      \
    try {                                                                     \
        __VA_ARGS__                                                           \
    } catch (const std::bad_alloc &) {                                        \
        set_error("host allocation failed");                                  \
        return SWIFT_ERR_ALLOC;                                               \
    } catch (...) {                                                           \
        set_error("unexpected C++ exception");                                \
        return SWIFT_ERR_ARG;                                                 \
    }

static bool value_type_ok(swift_dtype t) { return t == SWIFT_F32 || t == SWIFT_F64; }

extern "C" int swift_csr_to_csc(int m, int n, int nnz, swift_dtype vtype, const void *d_csr_val,
                                const int *d_csr_row_ptr, const int *d_csr_col_ind,
                                void *d_csc_val, int *d_csc_row_ind, int *d_csc_col_ptr,
                                int *d_nnzpercol, swift_stream_t stream) {
    if (m < 0 || n < 0 || nnz < 0 || !value_type_ok(vtype) || !d_csr_row_ptr ||
        (nnz && (!d_csr_col_ind || !d_csc_row_ind)) || !d_csc_col_ptr || !d_nnzpercol ||
        (d_csc_val && !d_csr_val)) {
        set_error("swift_csr_to_csc: bad arguments");
        return SWIFT_ERR_ARG;
    }
    SWIFT_GUARD(return csr_to_csc_entry(m, n, nnz, vtype, d_csr_val, d_csr_row_ptr,
                                        d_csr_col_ind, d_csc_val, d_csc_row_ind, d_csc_col_ptr,
                                        d_nnzpercol, S(stream));)
}

extern "C" int swift_col_sort(int n, int nnz, int K, swift_dtype vtype, const int *d_nnzpercol,
                              const int *d_csc_col_ptr, const int *d_csc_row_ind,
                              const void *d_csc_val, int *d_sortrowidx_tmp, void *d_sortval_tmp,
                              int *d_sortnnz, int *d_perm, const void *d_B, void *d_sortB,
                              swift_stream_t stream) {
    if (n < 0 || nnz < 0 || !value_type_ok(vtype) || !d_nnzpercol || !d_csc_col_ptr ||
        (nnz && (!d_csc_row_ind || !d_sortrowidx_tmp)) || !d_sortnnz || !d_perm ||
        (d_sortval_tmp && !d_csc_val)) {
        set_error("swift_col_sort: bad arguments");
        return SWIFT_ERR_ARG;
    }
    SWIFT_GUARD(return col_sort_entry(n, nnz, K, vtype, d_nnzpercol, d_csc_col_ptr,
                                      d_csc_row_ind, d_csc_val, d_sortrowidx_tmp, d_sortval_tmp,
                                      d_sortnnz, d_perm, d_B, d_sortB, S(stream));)
}

extern "C" int swift_build(swift_matrix **out, int m, int n, int nnz, swift_dtype vtype,
                           const int *d_csr_row_ptr, const int *d_csr_col_ind,
                           const void *d_csr_val, swift_stream_t stream) {
    if (!out || m < 0 || n < 0 || nnz < 0 || !value_type_ok(vtype) || !d_csr_row_ptr ||
        (nnz && (!d_csr_col_ind || !d_csr_val))) {
        set_error("swift_build: bad arguments");
        return SWIFT_ERR_ARG;
    }
    *out = nullptr;
    SWIFT_GUARD({
        swift_matrix *M = new swift_matrix();
        M->m = m;
        M->n = n;
        M->nnz = nnz;
        M->vtype = vtype;
        (void)hipGetDevice(&M->device);
        int rc = build_impl(M, d_csr_row_ptr, d_csr_col_ind, d_csr_val, S(stream));
        if (rc != SWIFT_OK) {
            destroy_matrix(M);
            return rc;
        }
        *out = M;
        return SWIFT_OK;
    })
}

extern "C" int swift_build_host(swift_matrix **out, int m, int n, int nnz, swift_dtype vtype,
                                const int *csr_row_ptr, const int *csr_col_ind,
                                const void *csr_val) {
    if (!out || m < 0 || n < 0 || nnz < 0 || !value_type_ok(vtype) || !csr_row_ptr ||
        (nnz && (!csr_col_ind || !csr_val))) {
        set_error("swift_build_host: bad arguments");
        return SWIFT_ERR_ARG;
    }
    // the row pointer must describe nnz entries with valid columns; checked
    // here on the host before the copies (swift_build checks the device
    // arrays again), so the message can name the entry
    if (csr_row_ptr[0] != 0 || csr_row_ptr[m] != nnz) {
        set_error("swift_build_host: rowptr[0]=%d rowptr[m]=%d nnz=%d", csr_row_ptr[0],
                  csr_row_ptr[m], nnz);
        return SWIFT_ERR_ARG;
    }
    for (int i = 0; i < m; i++)
        if (csr_row_ptr[i + 1] < csr_row_ptr[i]) {
            set_error("swift_build_host: rowptr decreases at row %d", i);
            return SWIFT_ERR_ARG;
        }
    for (int e = 0; e < nnz; e++)
        if (csr_col_ind[e] < 0 || csr_col_ind[e] >= n) {
            set_error("swift_build_host: column %d out of range at entry %d", csr_col_ind[e], e);
            return SWIFT_ERR_ARG;
        }
    const size_t vs = vtype == SWIFT_F64 ? 8 : 4;
    Buffer rp, ci, vv;
    SWIFT_TRY(rp.alloc(sizeof(int) * ((size_t)m + 1)));
    SWIFT_TRY(ci.alloc(sizeof(int) * (size_t)nnz));
    SWIFT_TRY(vv.alloc(vs * (size_t)nnz));
    SWIFT_HIP(hipMemcpy(rp.ptr, csr_row_ptr, sizeof(int) * ((size_t)m + 1), hipMemcpyHostToDevice));
    if (nnz) {
        SWIFT_HIP(hipMemcpy(ci.ptr, csr_col_ind, sizeof(int) * (size_t)nnz, hipMemcpyHostToDevice));
        SWIFT_HIP(hipMemcpy(vv.ptr, csr_val, vs * (size_t)nnz, hipMemcpyHostToDevice));
    }
    int rc = swift_build(out, m, n, nnz, vtype, rp.as<int>(), ci.as<int>(), vv.ptr, nullptr);
    (void)hipDeviceSynchronize();
    return rc;
}

extern "C" void swift_destroy(swift_matrix *mat) {
    if (mat) {
        (void)hipDeviceSynchronize();
        destroy_matrix(mat);
    }
}

extern "C" int swift_row_idx_sort(swift_matrix *mat, swift_stream_t stream) {
    if (!mat) {
        set_error("swift_row_idx_sort: null matrix");
        return SWIFT_ERR_ARG;
    }
    SWIFT_GUARD(return row_sort_impl(mat, S(stream));)
}

extern "C" int swift_set_segment_items(swift_matrix *mat, int items, swift_stream_t stream) {
    if (!mat || items < 8 || items > (1 << 20)) {
        set_error("swift_set_segment_items: items must be in [8, 2^20]");
        return SWIFT_ERR_ARG;
    }
    SWIFT_GUARD(int rc = build_segments(mat, items, S(stream));
                if (rc == SWIFT_OK) mat->user_items = items;
                return rc;)
}

extern "C" int swift_matrix_stats(const swift_matrix *mat, swift_stats *stats,
                                  swift_stream_t stream) {
    if (!mat || !stats) {
        set_error("swift_matrix_stats: bad arguments");
        return SWIFT_ERR_ARG;
    }
    SWIFT_GUARD(return matrix_stats_impl(mat, stats, S(stream));)
}

extern "C" int swift_get_info(const swift_matrix *M, swift_info *o) {
    if (!M || !o) {
        set_error("swift_get_info: bad arguments");
        return SWIFT_ERR_ARG;
    }
    o->m = M->m; o->n = M->n; o->nnz = M->nnz; o->vtype = M->vtype;
    o->tilenum = M->tilenum; o->regular_nnz = M->regular_nnz;
    o->reside_col = M->reside_col; o->reside_nnz = M->reside_nnz;
    o->segsum = M->segsum; o->numblk = M->numblk; o->nsteps = M->nsteps;
    o->has_rowsort = M->has_rowsort;
    o->perm = M->perm; o->sortnnz = M->sortnnz; o->sortrowidx_tmp = M->sortrowidx_tmp;
    o->sortval_tmp = M->sortval_tmp;
    o->tile_ptr = M->tile_ptr; o->tile_len = M->tile_len; o->tile_colidx = M->tile_colidx;
    o->tile_tall = M->tile_tall; o->tile_format = M->tile_format;
    o->sortrowidx = M->sortrowidx; o->sortval = M->sortval;
    o->countrow = M->countrow; o->segmentoffset = M->segmentoffset;
    o->sortrowindex = M->sortrowindex;
    o->reside_cscptr = M->reside_cscptr; o->reside_cscrowidx = M->reside_cscrowidx;
    o->reside_val = M->reside_val;
    o->colorblk = M->colorblk; o->blkstart = M->blkstart; o->blkstop = M->blkstop;
    o->blkptr = M->blkptr;
    o->val_position = M->val_position; o->val_offset = M->val_offset;
    o->rowsorted = M->rowsorted;
    o->sortsrc = M->sortsrc; o->reside_src = M->reside_src;
    for (int i = 0; i < 4; i++) o->build_ms[i] = M->build_ms[i];
    o->rowview_nnz = M->rv_nnz;
    o->segment_items = M->user_items;
    o->tuned_K = M->tuned_K;
    o->tuned_slots = M->tuned_slots;
    o->tuned_kernel = M->tuned_kernel;
    o->tuned_cols = M->tuned_cols;
    return SWIFT_OK;
}

extern "C" int swift_permute_dense(const swift_matrix *mat, int K, swift_layout layout,
                                   swift_dtype dtype, const void *d_in, void *d_out,
                                   swift_stream_t stream) {
    if (!mat || K < 0 || !d_in || !d_out || d_in == d_out) {
        set_error("swift_permute_dense: bad arguments (in-place is not supported)");
        return SWIFT_ERR_ARG;
    }
    SWIFT_GUARD(return permute_dense(mat, K, layout, value_size(dtype), d_in, d_out, S(stream));)
}

extern "C" int swift_spmm(const swift_matrix *mat, int K, swift_layout layout, swift_dtype dtype,
                          swift_kernel kernel, unsigned flags, const void *d_B, void *d_C,
                          swift_stream_t stream) {
    SWIFT_GUARD(return spmm_impl(mat, K, layout, dtype, kernel, flags, d_B, d_C, nullptr,
                                 S(stream));)
}

extern "C" int swift_spmm_noalloc(const swift_matrix *mat, int K, swift_layout layout,
                                  swift_dtype dtype, swift_kernel kernel, unsigned flags,
                                  const void *d_B, void *d_C, swift_stream_t stream) {
    SWIFT_GUARD(return spmm_noalloc_impl(mat, K, layout, dtype, kernel, flags, d_B, d_C,
                                         S(stream));)
}

extern "C" int swift_spmm_timed(const swift_matrix *mat, int K, swift_layout layout,
                                swift_dtype dtype, swift_kernel kernel, unsigned flags,
                                const void *d_B, void *d_C, swift_stream_t stream,
                                float kernel_ms[3]) {
    if (!kernel_ms) {
        set_error("swift_spmm_timed: null kernel_ms");
        return SWIFT_ERR_ARG;
    }
    SWIFT_GUARD(return spmm_timed_impl(mat, K, layout, dtype, kernel, flags, d_B, d_C, nullptr,
                                       S(stream), kernel_ms);)
}

extern "C" int swift_spmm_scaled(const swift_matrix *mat, int K, swift_layout layout,
                                 swift_kernel kernel, unsigned flags, const void *d_B,
                                 const float *d_col_scale, float *d_C, swift_stream_t stream) {
    if (!d_col_scale) {
        set_error("swift_spmm_scaled: null scale");
        return SWIFT_ERR_ARG;
    }
    SWIFT_GUARD(return spmm_impl(mat, K, layout, SWIFT_F16, kernel, flags, d_B, d_C, d_col_scale,
                                 S(stream));)
}

extern "C" int swift_prepare_values(swift_matrix *mat, swift_dtype dtype, swift_stream_t stream) {
    if (!mat) {
        set_error("swift_prepare_values: null matrix");
        return SWIFT_ERR_ARG;
    }
    SWIFT_GUARD(return prepare_values_impl(mat, dtype, S(stream));)
}

extern "C" int swift_convert_dense(const float *d_in, void *d_out, size_t rows, int K,
                                   swift_layout layout, swift_dtype dtype, float *d_col_scale,
                                   swift_stream_t stream) {
    if (!d_in || !d_out || K <= 0) {
        set_error("swift_convert_dense: bad arguments");
        return SWIFT_ERR_ARG;
    }
    SWIFT_GUARD(return convert_dense_impl(d_in, d_out, rows, K, layout, dtype, d_col_scale,
                                          S(stream));)
}

// ----------------------------------------------------- host launcher --

template <class T>
static int verify_count(const T *C, const T *G, size_t n) {
    size_t bad = 0;
    for (size_t i = 0; i < n; i++)
        if (std::fabs((double)G[i] - (double)C[i]) > 1e-2) bad++;
    return (int)std::min<size_t>(bad, 0x7FFFFFFF);
}

extern "C" int swift_spmm_host(const swift_matrix *mat, int K, swift_layout layout,
                               swift_dtype dtype, swift_kernel kernel, const void *h_sortB,
                               void *h_C, const void *h_golden, int warmup, int reps, float *ms,
                               double *gflops) {
    if (!mat || K <= 0 || !h_sortB || !h_C || reps < 1 || warmup < 0 ||
        (dtype != SWIFT_F32 && dtype != SWIFT_F64)) {
        set_error("swift_spmm_host: bad arguments (F32/F64 only)");
        return SWIFT_ERR_ARG;
    }
    if (ms) *ms = -1.0f;
    if (gflops) *gflops = 0.0;
    SWIFT_GUARD({
        const size_t es = value_size(dtype);
        Buffer B, C;
        SWIFT_TRY(B.alloc(es * (size_t)mat->n * K));
        SWIFT_TRY(C.alloc(es * (size_t)mat->m * K));
        SWIFT_HIP(hipMemcpy(B.ptr, h_sortB, es * (size_t)mat->n * K, hipMemcpyHostToDevice));
        hipEvent_t e0, e1;
        SWIFT_HIP(hipEventCreate(&e0));
        SWIFT_HIP(hipEventCreate(&e1));
        std::vector<float> times;
        int rc = SWIFT_OK;
        for (int it = 0; it < warmup + reps && rc == SWIFT_OK; it++) {
            (void)hipEventRecord(e0, nullptr);
            rc = spmm_impl(mat, K, layout, dtype, kernel, SWIFT_B_SORTED, B.ptr, C.ptr, nullptr,
                           nullptr);
            (void)hipEventRecord(e1, nullptr);
            (void)hipEventSynchronize(e1);
            float t = 0;
            (void)hipEventElapsedTime(&t, e0, e1);
            if (it >= warmup) times.push_back(t);
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (rc != SWIFT_OK) return rc;
        SWIFT_HIP(hipMemcpy(h_C, C.ptr, es * (size_t)mat->m * K, hipMemcpyDeviceToHost));
        std::sort(times.begin(), times.end());
        float med = times[times.size() / 2];
        if (h_golden) {
            size_t n = (size_t)mat->m * K;
            int bad = dtype == SWIFT_F64
                          ? verify_count((const double *)h_C, (const double *)h_golden, n)
                          : verify_count((const float *)h_C, (const float *)h_golden, n);
            if (bad) {
                set_error("swift_spmm_host: %d elements differ by more than 1e-2", bad);
                return SWIFT_ERR_VERIFY;
            }
        }
        if (ms) *ms = med;
        if (gflops && med > 0) *gflops = 2.0 * mat->nnz * (double)K / (med * 1e6);
        return SWIFT_OK;
    })
}

// ------------------------------------------------------------ backward --

extern "C" int swift_build_transpose(const swift_matrix *A, swift_matrix **AT,
                                     swift_stream_t stream) {
    if (!A || !AT) {
        set_error("swift_build_transpose: bad arguments");
        return SWIFT_ERR_ARG;
    }
    *AT = nullptr;
    SWIFT_GUARD(return build_transpose_impl(A, AT, S(stream));)
}

extern "C" int swift_sddmm(const swift_matrix *A, int K, swift_dtype dtype, const void *d_dC,
                           const void *d_B, void *d_dvals, swift_stream_t stream) {
    SWIFT_GUARD(return sddmm_impl(A, K, dtype, d_dC, d_B, d_dvals, S(stream));)
}

extern "C" int swift_backward_fused(const swift_matrix *AT, int K, swift_dtype dtype,
                                    const void *d_dC, const void *d_B, void *d_dB, void *d_dvals,
                                    swift_stream_t stream) {
    SWIFT_GUARD(return backward_impl(AT, K, dtype, d_dC, d_B, d_dB, d_dvals, S(stream));)
}

extern "C" int swift_update_values(swift_matrix *mat, const void *d_csr_val,
                                   swift_stream_t stream) {
    if (!mat || (!d_csr_val && mat->nnz)) {
        set_error("swift_update_values: bad arguments");
        return SWIFT_ERR_ARG;
    }
    SWIFT_GUARD(return update_values_impl(mat, d_csr_val, S(stream));)
}
