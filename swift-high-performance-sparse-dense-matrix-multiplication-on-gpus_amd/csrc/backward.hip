// backward.hip -- gradients of C = A*B over the Swift layout (gfx950).
//
//   dB    = A^T * dC      (swift_spmm on the transposed Swift layout, or the
//                          fused kernel below)
//   dvals = SDDMM(dC, B)  dvals[e] = <dC[row_e,:], B[col_e,:]> on A's pattern
//
// The fused kernel walks A's own column tiles: the 32 columns of a work item
// are fixed across its depth tiles, so dB rows for those columns accumulate in
// registers and are flushed with one atomic per column per work item, while
// each entry's dC row is read once and reused for its dvals dot product.
#include "swift_internal.h"

#include <algorithm>

namespace swift {

int build_impl(swift_matrix *M, const int *d_rowptr, const int *d_colidx, const void *d_val,
               hipStream_t s);
void destroy_matrix(swift_matrix *M);

template <class T>
__device__ __forceinline__ void atomic_add_bw(T *p, T v) { unsafeAtomicAdd(p, v); }

// ---------------------------------------------------- value scatter/gather --

template <class T>
__global__ void k_scatter_vals(const int *__restrict__ src, const T *__restrict__ v, size_t n,
                               T *__restrict__ out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[src[i]] = v[i];
}

template <class T>
__global__ void k_gather_vals(const int *__restrict__ src, const T *__restrict__ csr, size_t n,
                              T *__restrict__ out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = csr[src[i]];
}

__global__ void k_compose(int *__restrict__ a, const int *__restrict__ map, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = map[a[i]];
}

static unsigned g1(size_t n) { return (unsigned)std::max<size_t>(1, (n + 255) / 256); }

template <class T>
static int csr_values_t(const swift_matrix *A, T *out, hipStream_t s) {
    if (A->regular_nnz)
        hipLaunchKernelGGL(k_scatter_vals<T>, dim3(g1(A->regular_nnz)), dim3(256), 0, s, A->sortsrc,
                           (const T *)A->sortval, (size_t)A->regular_nnz, out);
    if (A->reside_nnz)
        hipLaunchKernelGGL(k_scatter_vals<T>, dim3(g1(A->reside_nnz)), dim3(256), 0, s,
                           A->reside_src, (const T *)A->reside_val, (size_t)A->reside_nnz, out);
    SWIFT_CHECK_LAUNCH();
    return SWIFT_OK;
}

// A's values back in CSR order.
int csr_values(const swift_matrix *A, void *out, hipStream_t s) {
    return A->vtype == SWIFT_F64 ? csr_values_t<double>(A, (double *)out, s)
                                 : csr_values_t<float>(A, (float *)out, s);
}

template <class T>
static int update_values_t(swift_matrix *M, const T *csr, hipStream_t s) {
    if (M->regular_nnz)
        hipLaunchKernelGGL(k_gather_vals<T>, dim3(g1(M->regular_nnz)), dim3(256), 0, s, M->sortsrc,
                           csr, (size_t)M->regular_nnz, (T *)M->sortval);
    if (M->reside_nnz)
        hipLaunchKernelGGL(k_gather_vals<T>, dim3(g1(M->reside_nnz)), dim3(256), 0, s,
                           M->reside_src, csr, (size_t)M->reside_nnz, (T *)M->reside_val);
    if (M->nnz)
        hipLaunchKernelGGL(k_gather_vals<T>, dim3(g1(M->nnz)), dim3(256), 0, s, M->sortsrc_tmp, csr,
                           (size_t)M->nnz, (T *)M->sortval_tmp);
    if (M->regular_nnz)
        hipLaunchKernelGGL(k_gather_vals<T>, dim3(g1(M->regular_nnz)), dim3(256), 0, s, M->rv_src,
                           csr, (size_t)M->regular_nnz, (T *)M->rv_val);
    SWIFT_CHECK_LAUNCH();
    return SWIFT_OK;
}

int update_values_impl(swift_matrix *M, const void *csr, hipStream_t s) {
    SWIFT_TRY(M->vtype == SWIFT_F64 ? update_values_t<double>(M, (const double *)csr, s)
                                    : update_values_t<float>(M, (const float *)csr, s));
    if (M->sortval_lp) {
        extern int prepare_values_impl(swift_matrix *, swift_dtype, hipStream_t);
        SWIFT_TRY(prepare_values_impl(M, M->lp_type, s));
    }
    return SWIFT_OK;
}

// ------------------------------------------------------------ transpose --

int csr_to_csc_with_src(int m, int n, int nnz, swift_dtype vtype, const void *d_val,
                        const int *d_rowptr, const int *d_colidx, void *d_cscval, int *d_cscrow,
                        int *d_cscptr, int *d_npc, int *d_cscsrc, hipStream_t s);

int build_transpose_impl(const swift_matrix *A, swift_matrix **out, hipStream_t s) {
    const size_t vs = A->vtype == SWIFT_F64 ? 8 : 4;
    Buffer vals, cscptr, cscrow, cscval, cscsrc, npc;
    SWIFT_TRY(vals.alloc(vs * (size_t)A->nnz));
    SWIFT_TRY(csr_values(A, vals.ptr, s));
    SWIFT_TRY(cscptr.alloc(sizeof(int) * ((size_t)A->n + 1)));
    SWIFT_TRY(npc.alloc(sizeof(int) * ((size_t)A->n + 1)));
    SWIFT_TRY(cscrow.alloc(sizeof(int) * (size_t)A->nnz));
    SWIFT_TRY(cscsrc.alloc(sizeof(int) * (size_t)A->nnz));
    SWIFT_TRY(cscval.alloc(vs * (size_t)A->nnz));
    SWIFT_TRY(csr_to_csc_with_src(A->m, A->n, A->nnz, A->vtype, vals.ptr, A->csr_rowptr,
                                  A->csr_colidx, cscval.ptr, cscrow.as<int>(), cscptr.as<int>(),
                                  npc.as<int>(), cscsrc.as<int>(), s));
    swift_matrix *T = new swift_matrix();
    T->m = A->n;
    T->n = A->m;
    T->nnz = A->nnz;
    T->vtype = A->vtype;
    T->device = A->device;
    int rc = build_impl(T, cscptr.as<int>(), cscrow.as<int>(), cscval.ptr, s);
    if (rc != SWIFT_OK) {
        destroy_matrix(T);
        return rc;
    }
    // entry sources: positions in A^T's CSR (= A's CSC) -> positions in A's CSR
    if (T->regular_nnz)
        hipLaunchKernelGGL(k_compose, dim3(g1(T->regular_nnz)), dim3(256), 0, s, T->sortsrc,
                           cscsrc.as<int>(), (size_t)T->regular_nnz);
    if (T->reside_nnz)
        hipLaunchKernelGGL(k_compose, dim3(g1(T->reside_nnz)), dim3(256), 0, s, T->reside_src,
                           cscsrc.as<int>(), (size_t)T->reside_nnz);
    if (T->nnz)
        hipLaunchKernelGGL(k_compose, dim3(g1(T->nnz)), dim3(256), 0, s, T->sortsrc_tmp,
                           cscsrc.as<int>(), (size_t)T->nnz);
    SWIFT_CHECK_LAUNCH();
    SWIFT_HIP(hipStreamSynchronize(s));
    *out = T;
    return SWIFT_OK;
}

// ------------------------------------------------------- SDDMM / fused --

struct BwArgs {
    int nwork;
    const int2 *work;
    const int *tile_colidx, *srow, *ssrc;
    const void *sval;
    int nblk;
    const unsigned *colorblk;
    const int *blkstart, *blkstop, *rptr, *rrow, *rsrc;
    const void *rval;
    int n, K;
    const int *perm;
    const void *dC, *B;
    void *dB, *dvals;
};

template <class T> struct V16;
template <> struct V16<float> { using t = float4; static constexpr int N = 4; };
template <> struct V16<double> { using t = double2; static constexpr int N = 2; };

template <class T, int VEC>
__device__ __forceinline__ void ld(const T *p, T (&o)[VEC]) {
    if constexpr (VEC == 4) {
        float4 v = *reinterpret_cast<const float4 *>(p);
        o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
    } else if constexpr (VEC == 2) {
        double2 v = *reinterpret_cast<const double2 *>(p);
        o[0] = v.x; o[1] = v.y;
    } else {
        o[0] = p[0];
    }
}

// Sum over the LPE lanes of one entry group.
template <class T, int LPE>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
    for (int off = LPE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// One wave per work item; lanes along K (LPE lanes x VEC x Q elements).
// MODE bit 0: dvals, bit 1: dB.
template <class T, int VEC, int Q, int LPE, int MODE>
__global__ __launch_bounds__(256) void k_bw_regular(BwArgs a) {
    constexpr int EPS = 64 / LPE, STEPS = TILE / EPS;
    const int lane = threadIdx.x & 63;
    const int h = lane / LPE, sub = lane % LPE;
    const int w = (int)((blockIdx.x * (size_t)blockDim.x + threadIdx.x) >> 6);
    if (w >= a.nwork) return;
    const int2 wk = a.work[w];
    const int c0 = a.tile_colidx[wk.x];
    const int my = lane & 31;
    const int mycol = a.perm[c0 + my];
    const T *dC = static_cast<const T *>(a.dC);
    const T *B = static_cast<const T *>(a.B);
    const T *sval = static_cast<const T *>(a.sval);
    T db[STEPS][Q][VEC];
    if constexpr (MODE & 2) {
#pragma unroll
        for (int i = 0; i < STEPS; i++)
#pragma unroll
            for (int q = 0; q < Q; q++)
#pragma unroll
                for (int v = 0; v < VEC; v++) db[i][q][v] = (T)0;
    }
    for (int t = wk.x; t < wk.x + wk.y; t++) {
        const size_t e0 = (size_t)t * TILE;
        const int myrow = a.srow[e0 + my];
        const T myval = sval[e0 + my];
        const int mysrc = a.ssrc[e0 + my];
#pragma unroll
        for (int i = 0; i < STEPS; i++) {
            const int l = h + i * EPS;
            const int r = __shfl(myrow, l, 64);
            const T v = __shfl(myval, l, 64);
            const int bc = __shfl(mycol, l, 64);
            T dot = (T)0;
#pragma unroll
            for (int q = 0; q < Q; q++) {
                const int k0 = q * LPE * VEC + sub * VEC;
                T g[VEC];
                ld<T, VEC>(dC + (size_t)r * a.K + k0, g);
                if constexpr (MODE & 1) {
                    T b[VEC];
                    ld<T, VEC>(B + (size_t)bc * a.K + k0, b);
#pragma unroll
                    for (int x = 0; x < VEC; x++) dot += g[x] * b[x];
                }
                if constexpr (MODE & 2) {
#pragma unroll
                    for (int x = 0; x < VEC; x++) db[i][q][x] += v * g[x];
                }
            }
            if constexpr (MODE & 1) {
                dot = group_sum<T, LPE>(dot);
                const int src = __shfl(mysrc, l, 64);
                if (sub == 0) static_cast<T *>(a.dvals)[src] = dot;
            }
        }
    }
    if constexpr (MODE & 2) {
        T *dB = static_cast<T *>(a.dB);
#pragma unroll
        for (int i = 0; i < STEPS; i++) {
            const int bc = __shfl(mycol, h + i * EPS, 64);
#pragma unroll
            for (int q = 0; q < Q; q++)
#pragma unroll
                for (int x = 0; x < VEC; x++)
                    atomic_add_bw(dB + (size_t)bc * a.K + q * LPE * VEC + sub * VEC + x, db[i][q][x]);
        }
    }
}

template <class T, int VEC, int Q, int LPE, int MODE>
__global__ __launch_bounds__(256) void k_bw_reside(BwArgs a) {
    constexpr int EPS = 64 / LPE;
    const int lane = threadIdx.x & 63;
    const int h = lane / LPE, sub = lane % LPE;
    const int blk = (int)((blockIdx.x * (size_t)blockDim.x + threadIdx.x) >> 6);
    if (blk >= a.nblk) return;
    const unsigned u = a.colorblk[blk];
    const int r = (int)(u & 0x7FFFFFFFu);
    const bool split = (u >> 31) != 0;
    const int start = split ? a.blkstart[blk] : a.rptr[r];
    const int stop = split ? a.blkstop[blk] : a.rptr[r + 1];
    const int bc = a.perm[a.n - 1 - r];
    const T *dC = static_cast<const T *>(a.dC);
    const T *B = static_cast<const T *>(a.B);
    const T *rval = static_cast<const T *>(a.rval);
    T b[Q][VEC], db[Q][VEC];
#pragma unroll
    for (int q = 0; q < Q; q++) {
        ld<T, VEC>(B + (size_t)bc * a.K + q * LPE * VEC + sub * VEC, b[q]);
#pragma unroll
        for (int x = 0; x < VEC; x++) db[q][x] = (T)0;
    }
    for (int e0 = start; e0 < stop; e0 += EPS) {
        const int e = e0 + h;
        const bool ok = e < stop;
        const int row = ok ? a.rrow[e] : 0;
        const T v = ok ? rval[e] : (T)0;
        T dot = (T)0;
#pragma unroll
        for (int q = 0; q < Q; q++) {
            T g[VEC];
            if (ok) ld<T, VEC>(dC + (size_t)row * a.K + q * LPE * VEC + sub * VEC, g);
            else
#pragma unroll
                for (int x = 0; x < VEC; x++) g[x] = (T)0;
#pragma unroll
            for (int x = 0; x < VEC; x++) {
                dot += g[x] * b[q][x];
                db[q][x] += v * g[x];
            }
        }
        if constexpr (MODE & 1) {
            dot = group_sum<T, LPE>(dot);
            if (ok && sub == 0) static_cast<T *>(a.dvals)[a.rsrc[e]] = dot;
        }
    }
    if constexpr (MODE & 2) {
        // the EPS groups hold partial sums of the same B row
#pragma unroll
        for (int q = 0; q < Q; q++)
#pragma unroll
            for (int x = 0; x < VEC; x++) {
                T v = db[q][x];
#pragma unroll
                for (int off = LPE; off < 64; off <<= 1) v += __shfl_xor(v, off, 64);
                if (h == 0)
                    atomic_add_bw(static_cast<T *>(a.dB) + (size_t)bc * a.K + q * LPE * VEC +
                                      sub * VEC + x,
                                  v);
            }
    }
}

template <class T, int VEC, int Q, int LPE, int MODE>
static int launch_bw(const BwArgs &a, hipStream_t s) {
    if (a.nwork)
        hipLaunchKernelGGL((k_bw_regular<T, VEC, Q, LPE, MODE>),
                           dim3(ceil_div((size_t)a.nwork * 64, 256)), dim3(256), 0, s, a);
    if (a.nblk)
        hipLaunchKernelGGL((k_bw_reside<T, VEC, Q, LPE, MODE>),
                           dim3(ceil_div((size_t)a.nblk * 64, 256)), dim3(256), 0, s, a);
    SWIFT_CHECK_LAUNCH();
    return SWIFT_OK;
}

template <class T, int MODE>
static int dispatch_bw(const BwArgs &a, hipStream_t s) {
    constexpr int V = V16<T>::N;
    const int K = a.K;
    if (K % V != 0) {
        set_error("backward kernels need K %% %d == 0 (K=%d)", V, K);
        return SWIFT_ERR_UNSUPPORTED;
    }
    const int lanes = K / V;  // lanes x Q
    switch (lanes) {
        case 8: return launch_bw<T, V, 1, 8, MODE>(a, s);
        case 16: return launch_bw<T, V, 1, 16, MODE>(a, s);
        case 32: return launch_bw<T, V, 1, 32, MODE>(a, s);
        case 64: return launch_bw<T, V, 1, 64, MODE>(a, s);
        case 128: return launch_bw<T, V, 2, 64, MODE>(a, s);
        default: break;
    }
    set_error("backward kernels support K/%d in {8,16,32,64,128} (K=%d)", V, K);
    return SWIFT_ERR_UNSUPPORTED;
}

int backward_impl(const swift_matrix *A, int K, swift_dtype dtype, const void *dC, const void *B,
                  void *dB, void *dvals, hipStream_t s) {
    if (!A || K <= 0 || !dC || (!dB && !dvals) || (dvals && !B)) {
        set_error("swift_backward: bad arguments");
        return SWIFT_ERR_ARG;
    }
    if (dtype != A->vtype) {
        set_error("swift_backward: dtype must equal the build value type");
        return SWIFT_ERR_ARG;
    }
    const size_t es = dtype == SWIFT_F64 ? 8 : 4;
    if (dB) SWIFT_TRY(fill_bytes_zero(dB, (size_t)A->n * K * es, s));
    BwArgs a{};
    a.nwork = A->nwork;
    a.work = A->work;
    a.tile_colidx = A->tile_colidx;
    a.srow = A->sortrowidx;
    a.ssrc = A->sortsrc;
    a.sval = A->sortval;
    a.nblk = A->numblk;
    a.colorblk = A->colorblk;
    a.blkstart = A->blkstart;
    a.blkstop = A->blkstop;
    a.rptr = A->reside_cscptr;
    a.rrow = A->reside_cscrowidx;
    a.rsrc = A->reside_src;
    a.rval = A->reside_val;
    a.n = A->n;
    a.K = K;
    a.perm = A->perm;
    a.dC = dC;
    a.B = B;
    a.dB = dB;
    a.dvals = dvals;
    const int mode = (dvals ? 1 : 0) | (dB ? 2 : 0);
    if (dtype == SWIFT_F64) {
        if (mode == 1) return dispatch_bw<double, 1>(a, s);
        if (mode == 2) return dispatch_bw<double, 2>(a, s);
        return dispatch_bw<double, 3>(a, s);
    }
    if (mode == 1) return dispatch_bw<float, 1>(a, s);
    if (mode == 2) return dispatch_bw<float, 2>(a, s);
    return dispatch_bw<float, 3>(a, s);
}

}  // namespace swift
