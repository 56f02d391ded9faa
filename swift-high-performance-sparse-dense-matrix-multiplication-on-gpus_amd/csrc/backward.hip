// backward.hip -- gradients of C = A*B over the Swift layout (gfx950).
//
//   dB    = A^T * dC      (swift_spmm on the transposed Swift layout, or the
//                          fused kernel below)
//   dvals = SDDMM(dC, B)  dvals[e] = <dC[row_e,:], B[col_e,:]> on A's pattern
//
// The fused pass walks the transposed Swift layout (built once by
// swift_build_transpose): each dC row is gathered once per nonzero and feeds
// both dB and dvals, and dB rows are written once, not scattered.
#include "rowview.h"

#include <algorithm>
#include <type_traits>

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

// out[src[i]] = i: the inverse of a permutation of 0..n-1.
__global__ void k_invert(const unsigned *__restrict__ src, size_t n, int *__restrict__ out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[src[i]] = (int)i;
}

// keys[i] = src[i] >> lg (src: signed positions or unsigned walk indices)
template <class I>
__global__ void k_shift_keys(const I *__restrict__ src, size_t n, int lg,
                             unsigned *__restrict__ keys) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) keys[i] = (unsigned)src[i] >> lg;
}

// off[s] = CSR position, inside its bucket, of the entry in bucket-order slot s.
__global__ void k_bucket_offsets(const unsigned *__restrict__ walk, const int *__restrict__ src,
                                 size_t n, int lg, unsigned short *__restrict__ off) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) off[i] = (unsigned short)(src[walk[i]] & ((1 << lg) - 1));
}

// For position p of the (chunk, bucket, walk) order, whose entry is walk
// position e = order[p]: rank[e] = p inside the chunk, dst[p] = slot of e in
// bucket order.
__global__ void k_chunk_plan(const unsigned *__restrict__ order, const int *__restrict__ slot,
                             size_t n, int lg, unsigned short *__restrict__ rank,
                             int *__restrict__ dst) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned e = order[i];
    rank[e] = (unsigned short)(i & ((1u << lg) - 1));
    dst[i] = slot[e];
}

// Sizes (log2 of values) of the dvals reordering: a bucket (pass 2) holds
// 128 KiB of values, a chunk (pass 1) 64 KiB, so that two pass-1 workgroups
// share a CU; each is reordered in LDS by one workgroup.
template <class T> struct DvPlan;
template <> struct DvPlan<float> { static constexpr int BUCKET = 15, CHUNK = 13; };
template <> struct DvPlan<double> { static constexpr int BUCKET = 14, CHUNK = 13; };
static int dv_bucket_log(swift_dtype t) {
    return t == SWIFT_F64 ? DvPlan<double>::BUCKET : DvPlan<float>::BUCKET;
}
static int dv_chunk_log(swift_dtype t) {
    return t == SWIFT_F64 ? DvPlan<double>::CHUNK : DvPlan<float>::CHUNK;
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
    if (M->rv_nnz)
        hipLaunchKernelGGL(k_gather_vals<T>, dim3(g1(M->rv_nnz)), dim3(256), 0, s, M->rv_src,
                           csr, (size_t)M->rv_nnz, (T *)M->rv_val);
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

// The plan that moves the fused backward's dvals from walk order into the
// original matrix's CSR order (swift_matrix::dv_rank / dv_dst / dv_off).  The
// row view of a transposed layout holds every entry once, so rv_src is a
// permutation of the CSR positions.  Bucket order: walk positions stably
// sorted by bucket (CSR position >> dv_log); (chunk, bucket, walk) order:
// bucket order stably sorted by chunk (walk position >> dv_clog).
static int build_dv_plan(swift_matrix *T, hipStream_t s) {
    if (T->rv_nnz != T->nnz || !T->nnz) return SWIFT_OK;
    const size_t n = (size_t)T->nnz;
    const int lg = dv_bucket_log(T->vtype), lc = dv_chunk_log(T->vtype);
    Buffer keys, kout, border, corder, kalt, valt, slot, tmp;
    SWIFT_TRY(keys.alloc(4 * n));
    SWIFT_TRY(kout.alloc(4 * n));
    SWIFT_TRY(border.alloc(4 * n));
    SWIFT_TRY(corder.alloc(4 * n));
    SWIFT_TRY(kalt.alloc(4 * n));
    SWIFT_TRY(valt.alloc(4 * n));
    SWIFT_TRY(slot.alloc(4 * n));
    SWIFT_TRY(tmp.alloc(radix_temp_bytes(n)));
    hipLaunchKernelGGL(k_shift_keys<int>, dim3(g1(n)), dim3(256), 0, s, T->rv_src, n, lg,
                       keys.as<unsigned>());
    SWIFT_CHECK_LAUNCH();
    SWIFT_TRY(radix_sort_pairs(keys.as<unsigned>(), nullptr, kout.as<unsigned>(),
                               border.as<unsigned>(), kalt.as<unsigned>(), valt.as<unsigned>(),
                               n, bits_for((unsigned)((n - 1) >> lg)), tmp.ptr, s));
    hipLaunchKernelGGL(k_shift_keys<unsigned>, dim3(g1(n)), dim3(256), 0, s,
                       border.as<unsigned>(), n, lc, keys.as<unsigned>());
    SWIFT_CHECK_LAUNCH();
    SWIFT_TRY(radix_sort_pairs(keys.as<unsigned>(), border.as<unsigned>(), kout.as<unsigned>(),
                               corder.as<unsigned>(), kalt.as<unsigned>(), valt.as<unsigned>(),
                               n, bits_for((unsigned)((n - 1) >> lc)), tmp.ptr, s));
    if (hipMalloc((void **)&T->dv_rank, 2 * n) != hipSuccess ||
        hipMalloc((void **)&T->dv_off, 2 * n) != hipSuccess ||
        hipMalloc((void **)&T->dv_dst, 4 * n) != hipSuccess) {
        set_error("swift_build_transpose: hipMalloc(%zu) failed", 8 * n);
        return SWIFT_ERR_ALLOC;
    }
    T->dv_log = lg;
    T->dv_clog = lc;
    hipLaunchKernelGGL(k_invert, dim3(g1(n)), dim3(256), 0, s, border.as<unsigned>(), n,
                       slot.as<int>());
    hipLaunchKernelGGL(k_bucket_offsets, dim3(g1(n)), dim3(256), 0, s, border.as<unsigned>(),
                       T->rv_src, n, lg, T->dv_off);
    hipLaunchKernelGGL(k_chunk_plan, dim3(g1(n)), dim3(256), 0, s, corder.as<unsigned>(),
                       slot.as<int>(), n, lc, T->dv_rank, T->dv_dst);
    SWIFT_CHECK_LAUNCH();
    SWIFT_HIP(hipStreamSynchronize(s));  // the temporaries are freed on return
    return SWIFT_OK;
}

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
    T->rv_all = true;  // the backward walks cover the reside entries too
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
    if (T->rv_nnz)
        hipLaunchKernelGGL(k_compose, dim3(g1(T->rv_nnz)), dim3(256), 0, s, T->rv_src,
                           cscsrc.as<int>(), (size_t)T->rv_nnz);
    rc = build_dv_plan(T, s);
    if (rc == SWIFT_OK && hipGetLastError() != hipSuccess) {
        set_error("swift_build_transpose: a kernel launch failed");
        rc = SWIFT_ERR_HIP;
    }
    if (rc == SWIFT_OK && hipStreamSynchronize(s) != hipSuccess) rc = SWIFT_ERR_HIP;
    if (rc != SWIFT_OK) {
        destroy_matrix(T);
        return rc;
    }
    *out = T;
    return SWIFT_OK;
}

// ------------------------------------------------------- SDDMM / fused --
//
// Both passes walk a row view with merge-path segments (one wave per
// segment, EPS sub-groups of LPE lanes, lanes along K), like the forward
// row-view kernel:
//   SDDMM over A:    row i keeps H = dC[i] in registers; each entry (j, src)
//                    gathers G = B[j]; dvals[src] = <H, G>.
//   fused over A^T:  row j of A^T (= column j of A) keeps H = B[j]; each entry
//                    (i, v) at row-view position p gathers G = dC[i] once and
//                    uses it twice: dv[p] = <H, G> and dB[j] += v * G; dv is
//                    then reordered into A's CSR order (pass 1 in
//                    k_bw_tail, pass 2 in k_dv_bucket_order).
// The walked matrix's reside blocks (one column, <= 32 entries) follow with
// the roles swapped: the column operand is loaded once, dB takes atomics.

struct BwArgs {
    int m, nseg;               // walked matrix rows, merge-path segments
    const int2 *seg;
    const int *rv_ptr, *rv_col, *rv_src;
    const void *rv_val;
    int n, nblk;               // walked matrix columns, reside blocks
    const unsigned *colorblk;
    const int *blkstart, *blkstop, *rptr, *rrow, *rsrc;
    const void *rval;
    const int *perm;
    int K;                     // row length (elements) of G, H and dB
    int kbase, kcount;         // K slice handled by this launch
    int dv_add;                // dvals += this slice's partial dots (not the first slice)
    const void *G;             // operand gathered per entry (by column)
    const void *H;             // operand per walked row
    void *dB;                  // walked-row output (fused pass) or nullptr
    void *dvals;               // CSR-order output (walk order with WALK) or nullptr
    // dB rows cut by segment boundaries: a segment that starts inside a row
    // stores that row's piece into carry slot seg_slot[sid - 1] (K values per
    // slot); k_fix_rows adds a cut row's slots onto it after the walk
    void *carry;
    const int *seg_slot;
    const int4 *fix;
    int nfix, nlong;
    unsigned gbytes, gstride;  // buffer-resource path for G (0: plain loads)
    // dvals alone: the reside blocks run as blocks walk_blocks.. of the
    // walk's launch
    int walk_blocks;
};

// Elements per lane load of a stored operand: 16 bytes for FP32 / FP64, 8
// bytes (four values) for BF16 storage, as in the forward walk.
template <class T> struct V16;
template <> struct V16<float> { static constexpr int N = 4; };
template <> struct V16<double> { static constexpr int N = 2; };
template <> struct V16<__hip_bfloat16> { static constexpr int N = 4; };

template <class T, int VEC>
__device__ __forceinline__ void ldv(const T *p, T (&o)[VEC]) {
    if constexpr (std::is_same<T, float>::value && VEC == 4) {
        float4 v = *reinterpret_cast<const float4 *>(p);
        o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
    } else if constexpr (std::is_same<T, double>::value && VEC == 2) {
        double2 v = *reinterpret_cast<const double2 *>(p);
        o[0] = v.x; o[1] = v.y;
    } else {
#pragma unroll
        for (int i = 0; i < VEC; i++) o[i] = p[i];
    }
}

// The walked row's own operand row (H) is read once: non-temporal loads.
template <class T, int VEC>
__device__ __forceinline__ void ldv_stream(const T *p, T (&o)[VEC]) {
    if constexpr (std::is_same<T, float>::value && VEC == 4) {
        const f32x4 v = __builtin_nontemporal_load(reinterpret_cast<const f32x4 *>(p));
        o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
    } else if constexpr (std::is_same<T, double>::value && VEC == 2) {
        const f64x2 v = __builtin_nontemporal_load(reinterpret_cast<const f64x2 *>(p));
        o[0] = v.x; o[1] = v.y;
    } else {
#pragma unroll
        for (int i = 0; i < VEC; i++) o[i] = ld_stream(p + i);
    }
}

// VEC elements of a stored operand (TS) widened to the compute type T.
template <class TS, class T, int VEC>
__device__ __forceinline__ void ldw(const TS *p, T (&o)[VEC]) {
    if constexpr (std::is_same<TS, T>::value) ldv<T, VEC>(p, o);
    else load_b<TS, VEC, T>(p, o);
}
template <class TS, class T, int VEC>
__device__ __forceinline__ void ldw_stream(const TS *p, T (&o)[VEC]) {
    if constexpr (std::is_same<TS, T>::value) ldv_stream<T, VEC>(p, o);
    else load_b<TS, VEC, T>(p, o);
}

// dB rows are written once: non-temporal stores.
template <class T, int VEC>
__device__ __forceinline__ void stv(T *p, const T (&x)[VEC]) {
    if constexpr (std::is_same<T, float>::value && VEC == 4) {
        st_stream4(p, x[0], x[1], x[2], x[3]);
    } else if constexpr (std::is_same<T, double>::value && VEC == 2) {
        st_stream2(p, x[0], x[1]);
    } else {
#pragma unroll
        for (int i = 0; i < VEC; i++) p[i] = x[i];
    }
}

__device__ __forceinline__ unsigned xcd_block_bw(unsigned b, unsigned nblocks) {
    const unsigned per = nblocks / 8, rem = nblocks % 8, x = b % 8;
    return x * per + min(x, rem) + b / 8;
}

// Sum over the LPE lanes of one sub-group.  These sums exchange lanes with
// lane_xor (VALU lane swaps and DPP where the distance allows): on the
// 200k-node graph the fused backward went 0.623 -> 0.606-0.616 ms (K=128),
// 0.200 -> 0.197 ms (K=32), BF16 0.378 -> 0.374 ms.
template <class T, int LPE>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
    for (int off = LPE / 2; off > 0; off >>= 1) v += lane_xor(v, off);
    return v;
}

// N sums at once over the LPE lanes of a sub-group (N <= LPE, powers of two):
// each halving step exchanges half of the values with the partner lane, so N
// values cost N-1 + log2(LPE/N) shuffles instead of N*log2(LPE).  Lane sub
// ends up with the full sum of value sub / (LPE/N).
template <class T, int LPE, int N>
__device__ __forceinline__ T multi_group_sum(const T (&d)[N], int sub) {
    static_assert(N <= LPE && (N & (N - 1)) == 0, "N must be a power of two <= LPE");
    T x[N];
#pragma unroll
    for (int i = 0; i < N; i++) x[i] = d[i];
#pragma unroll
    for (int lvl = 0; (N >> lvl) > 1; lvl++) {
        const int off = LPE >> (lvl + 1);
        const int half = (N >> lvl) / 2;
        const bool up = (sub & off) != 0;
#pragma unroll
        for (int i = 0; i < half; i++) {
            const T send = up ? x[i] : x[i + half];
            const T keep = up ? x[i + half] : x[i];
            x[i] = keep + lane_xor(send, off);
        }
    }
    T r = x[0];
#pragma unroll
    for (int off = (LPE / N) / 2; off > 0; off >>= 1) r += lane_xor(r, off);
    return r;
}

// dvals of one entry: the first K slice stores its partial dot, later slices
// add theirs (one lane per entry and launch, launches in stream order).  The
// entry's slot is its CSR position, or with WALK its row-view position
// (consecutive entries, coalesced stores; the caller reorders afterwards).
template <class T>
__device__ __forceinline__ void put_dval(const BwArgs &a, int slot, T r) {
    T *dv = static_cast<T *>(a.dvals) + slot;
    *dv = a.dv_add ? *dv + r : r;
}

// N slot rows of one buffer-load step of the backward walk, without branches
// (only the last slot row can reach past the row's end; its unused slots
// read zeros through OOB_OFFSET), as in the forward walk's gather_body.
// DB: acc += v * G[c] for the row's dB.  DV: dvals[src] = <G[c], H[row]>, the
// slots' partial dots summed over the lane group together (lane sub then
// holds slot sub / (LPE / UNROLL)).
template <class TS, class T, int VEC, int Q, int LPE, bool MASK, bool DB, bool DV, bool WALK, int UNROLL, int N>
__device__ __forceinline__ void bw_body(const BwArgs &a, __amdgpu_buffer_rsrc_t rsrc, int base,
                                        int j, int nb, int h, int sub, int mc, int ms, T mv,
                                        const T (&hr)[Q][VEC], T (&acc)[Q][VEC], int pos) {
    constexpr int EPS = 64 / LPE;
    T g[N][Q][VEC];
    T vv[N];
    int src[N];
#pragma unroll
    for (int u = 0; u < N; u++) {
        // one permute base per step, G rows by 24-bit offsets
        const int addr = base + ((u * EPS) << 2);
        const int c = __builtin_amdgcn_ds_bpermute(addr, mc);
        src[u] = __builtin_amdgcn_ds_bpermute(addr, ms);
        const T vs = bperm_val(addr, mv);
        const bool ok = u < N - 1 || j + u * EPS + h < nb;
        vv[u] = ok ? vs : (T)0;
        const unsigned roff = __umul24((unsigned)c, a.gstride);
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const int k0 = q * LPE * VEC + sub * VEC;
            const bool in = ok && (!MASK || k0 < a.kcount);
            load_b_buf<TS, VEC, T>(rsrc, in ? roff + (unsigned)(k0 * (int)sizeof(TS)) : OOB_OFFSET,
                                  g[u][q]);
        }
    }
    if (DB) {
#pragma unroll
        for (int u = 0; u < N; u++)
#pragma unroll
            for (int q = 0; q < Q; q++)
#pragma unroll
                for (int v = 0; v < VEC; v++) acc[q][v] += vv[u] * g[u][q][v];
    }
    if (DV) {
        constexpr int LEAD = LPE / UNROLL;
        T d[UNROLL];
        int sv = 0;
        const int us = sub / LEAD;
#pragma unroll
        for (int u = 0; u < UNROLL; u++) d[u] = (T)0;
#pragma unroll
        for (int u = 0; u < N; u++) {
            T x = (T)0;
#pragma unroll
            for (int q = 0; q < Q; q++)
#pragma unroll
                for (int v = 0; v < VEC; v++) x += g[u][q][v] * hr[q][v];
            d[u] = x;
            if (us == u) sv = src[u];
        }
        const T r = multi_group_sum<T, LPE, UNROLL>(d, sub);
        // slots past the row's end (us >= N among them) write nothing
        if (sub % LEAD == 0 && j + us * EPS + h < nb) put_dval(a, WALK ? pos + us * EPS : sv, r);
    }
}

// The body for nu = min(UNROLL, slot rows left) slot rows (wave-uniform).
template <class TS, class T, int VEC, int Q, int LPE, bool MASK, bool DB, bool DV, bool WALK, int UNROLL, int N>
__device__ __forceinline__ void bw_pick(int nu, const BwArgs &a, __amdgpu_buffer_rsrc_t rsrc,
                                        int base, int j, int nb, int h, int sub, int mc, int ms,
                                        T mv, const T (&hr)[Q][VEC], T (&acc)[Q][VEC], int pos) {
    if constexpr (N == 1) {
        bw_body<TS, T, VEC, Q, LPE, MASK, DB, DV, WALK, UNROLL, 1>(a, rsrc, base, j, nb, h, sub, mc, ms, mv,
                                                         hr, acc, pos);
    } else {
        if (nu >= N)
            bw_body<TS, T, VEC, Q, LPE, MASK, DB, DV, WALK, UNROLL, N>(a, rsrc, base, j, nb, h, sub, mc, ms,
                                                             mv, hr, acc, pos);
        else
            bw_pick<TS, T, VEC, Q, LPE, MASK, DB, DV, WALK, UNROLL, N - 1>(nu, a, rsrc, base, j, nb, h, sub,
                                                                 mc, ms, mv, hr, acc, pos);
    }
}

// Reside blocks of the walked matrix: column operand G[c] loaded once, each
// entry gathers H[row]; dvals = <H[row], G[c]>, dB[row] += v * G[c].
template <class TS, class T, int VEC, int Q, int LPE, bool MASK, bool DB, bool DV>
__device__ __forceinline__ void bw_reside_block(const BwArgs &a, int blk) {
    constexpr int EPS = 64 / LPE;
    const int lane = threadIdx.x & 63;
    const int h = lane / LPE, sub = lane % LPE;
    const unsigned u = a.colorblk[blk];
    const int r = (int)(u & 0x7FFFFFFFu);
    const bool split = (u >> 31) != 0;
    const int start = split ? a.blkstart[blk] : a.rptr[r];
    const int stop = split ? a.blkstop[blk] : a.rptr[r + 1];
    const int c = a.perm[a.n - 1 - r];
    const TS *G = static_cast<const TS *>(a.G) + a.kbase;
    const TS *H = static_cast<const TS *>(a.H) + a.kbase;
    const TS *rval = static_cast<const TS *>(a.rval);
    // lane-strided layout k = q*LPE*VEC + x*LPE + sub: for a fixed (q, x) the
    // LPE lanes touch LPE consecutive elements, so each dB atomic instruction
    // covers whole cache lines
    T gc[Q][VEC];
#pragma unroll
    for (int q = 0; q < Q; q++)
#pragma unroll
        for (int x = 0; x < VEC; x++) {
            const int k = q * LPE * VEC + x * LPE + sub;
            gc[q][x] = (!MASK || k < a.kcount) ? to_acc(G[(size_t)c * a.K + k]) : (T)0;
        }
    for (int e0 = start; e0 < stop; e0 += EPS) {
        const int e = e0 + h;
        const bool ok = e < stop;
        const int row = ok ? a.rrow[e] : 0;
        if (DV) {
            T d = (T)0;
#pragma unroll
            for (int q = 0; q < Q; q++)
#pragma unroll
                for (int x = 0; x < VEC; x++) {
                    const int k = q * LPE * VEC + x * LPE + sub;
                    if (ok && (!MASK || k < a.kcount)) d += to_acc(H[(size_t)row * a.K + k]) * gc[q][x];
                }
            d = group_sum<T, LPE>(d);
            if (ok && sub == 0) put_dval(a, a.rsrc[e], d);
        }
        if (DB && ok) {
            const T v = to_acc(rval[e]);
            T *drow = static_cast<T *>(a.dB) + (size_t)row * a.K + a.kbase;
#pragma unroll
            for (int q = 0; q < Q; q++)
#pragma unroll
                for (int x = 0; x < VEC; x++) {
                    const int k = q * LPE * VEC + x * LPE + sub;
                    if (!MASK || k < a.kcount) atomic_add_bw(drow + k, v * gc[q][x]);
                }
        }
    }
}

// Slots per step of the backward walk.  8 for 32/64-lane groups (as in the
// forward walk) was measured slower on the 1M-row matrix (fused 1.06 -> 1.30
// ms at K=128), so every shape takes 4.
#ifndef SWIFT_BW_UNROLL_LONG
#define SWIFT_BW_UNROLL_LONG 4
#endif
template <class TS, class T, int VEC, int Q, int LPE, bool MASK, bool DB, bool DV, bool WALK, bool BUF>
__global__ __launch_bounds__(256) void k_bw_walk(BwArgs a) {
    if constexpr (!DB) {
        if (blockIdx.x >= (unsigned)a.walk_blocks) {
            const int blk =
                (int)(((blockIdx.x - a.walk_blocks) * (size_t)blockDim.x + threadIdx.x) >> 6);
            if (blk < a.nblk) bw_reside_block<TS, T, VEC, Q, LPE, MASK, DB, DV>(a, blk);
            return;
        }
    }
    constexpr int EPS = 64 / LPE;
    constexpr int UNROLL = (LPE >= 32 && Q <= 2) ? SWIFT_BW_UNROLL_LONG : 4;
    constexpr int STEP = EPS * UNROLL;
    const int lane = threadIdx.x & 63;
    const int h = lane / LPE, sub = lane % LPE;
    const unsigned lb = xcd_block_bw(blockIdx.x, (unsigned)a.walk_blocks);
    const int sid = __builtin_amdgcn_readfirstlane((int)((lb * (size_t)blockDim.x + threadIdx.x) >> 6));
    if (sid >= a.nseg) return;
    const int2 st = a.seg[sid], en = a.seg[sid + 1];
    const int r1 = en.x, e1 = en.y;
    int row = st.x, e = st.y;
    // requested at once, as in the forward walk: the carry slot of the
    // boundary before the segment and whether its end cuts the last row
    const int slot_before = DB && sid > 0 ? a.seg_slot[sid - 1] : 0;
    const bool carried_in = row < a.m && e > a.rv_ptr[row];
    const bool cut_end = r1 < a.m && e1 > a.rv_ptr[r1];
    // carry slot of the dB piece this segment starts with
    const int cslot = carried_in ? slot_before : 0;
    const TS *val = static_cast<const TS *>(a.rv_val);
    const TS *G = static_cast<const TS *>(a.G) + a.kbase;
    const TS *H = static_cast<const TS *>(a.H) + a.kbase;
    __amdgpu_buffer_rsrc_t rsrc;
    if constexpr (BUF) rsrc = make_rsrc(G, a.gbytes);
    // staged row ends and entries, as in the forward row-view kernel
    int rbase = row;
    int rend_l = (rbase + lane < a.m) ? ld_stream(a.rv_ptr + rbase + lane + 1) : e1;
    int ebase = e, mc = 0, ms = 0;
    T mv = (T)0;
    if (ebase + lane < e1) {
        mc = ld_stream(a.rv_col + ebase + lane);
        if (DV && !WALK) ms = ld_stream(a.rv_src + ebase + lane);
        if (DB) mv = to_acc(ld_stream(val + ebase + lane));  // dvals need no values
    }
    // the next batch, requested one batch early (as in the forward walk):
    // the 200k-node graph at K=128 0.633 -> 0.621 ms
    int nc = 0, ns = 0;
    T nv = (T)0;
    if (ebase + 64 < e1 && ebase + 64 + lane < e1) {
        nc = ld_stream(a.rv_col + ebase + 64 + lane);
        if (DV && !WALK) ns = ld_stream(a.rv_src + ebase + 64 + lane);
        if (DB) nv = to_acc(ld_stream(val + ebase + 64 + lane));
    }
    T acc[Q][VEC], hr[Q][VEC];
#pragma unroll
    for (int q = 0; q < Q; q++)
#pragma unroll
        for (int v = 0; v < VEC; v++) acc[q][v] = hr[q][v] = (T)0;
    while (true) {
        if (row - rbase >= 64) {
            rbase = row;
            rend_l = (rbase + lane < a.m) ? ld_stream(a.rv_ptr + rbase + lane + 1) : e1;
        }
        const int rend = __builtin_amdgcn_readlane(rend_l, row - rbase);
        const int row_end = row < r1 ? rend : e1;
        if (DV && e < row_end) {
#pragma unroll
            for (int q = 0; q < Q; q++) {
                const int k0 = q * LPE * VEC + sub * VEC;
                if (!MASK || k0 < a.kcount) ldw_stream<TS, T, VEC>(H + (size_t)row * a.K + k0, hr[q]);
            }
        }
        while (e < row_end) {
            if (e - ebase >= 64) {
                ebase = e;
                mc = nc;
                ms = ns;
                mv = nv;
                nc = ns = 0;
                nv = (T)0;
                if (ebase + 64 < e1 && ebase + 64 + lane < e1) {
                    nc = ld_stream(a.rv_col + ebase + 64 + lane);
                    if (DV && !WALK) ns = ld_stream(a.rv_src + ebase + 64 + lane);
                    if (DB) nv = to_acc(ld_stream(val + ebase + 64 + lane));
                }
            }
            const int off = e - ebase;
            const int nb = min(64 - off, row_end - e);
            for (int j = 0; j < nb; j += STEP) {
                T g[UNROLL][Q][VEC];
                T vv[UNROLL];
                int src[UNROLL];
                if constexpr (BUF) {
                    const int nu = min(UNROLL, (nb - j + EPS - 1) / EPS);
                    bw_pick<TS, T, VEC, Q, LPE, MASK, DB, DV, WALK, UNROLL, UNROLL>(
                        nu, a, rsrc, (off + j + h) << 2, j, nb, h, sub, mc, ms, mv, hr, acc,
                        e + j + h);
                    continue;
                }
#pragma unroll
                for (int u = 0; u < UNROLL; u++) {
                    const int idx = j + u * EPS + h;
                    // unconditional shuffles: a lane reading a masked-off source
                    // lane would get a stale value
                    const int c = __shfl(mc, (off + idx) & 63, 64);
                    src[u] = __shfl(ms, (off + idx) & 63, 64);
                    const T vs = __shfl(mv, (off + idx) & 63, 64);
                    vv[u] = idx < nb ? vs : (T)0;
#pragma unroll
                    for (int q = 0; q < Q; q++) {
                        const int k0 = q * LPE * VEC + sub * VEC;
                        if (idx < nb && (!MASK || k0 < a.kcount))
                            ldw<TS, T, VEC>(G + (size_t)c * a.K + k0, g[u][q]);
                        else
#pragma unroll
                            for (int v = 0; v < VEC; v++) g[u][q][v] = (T)0;
                    }
                }
                if (DB) {
#pragma unroll
                    for (int u = 0; u < UNROLL; u++) {
#pragma unroll
                        for (int q = 0; q < Q; q++)
#pragma unroll
                            for (int v = 0; v < VEC; v++) acc[q][v] += vv[u] * g[u][q][v];
                    }
                }
                if (DV) {
                    // partial dots of the UNROLL slots, summed over the lane
                    // group together; lane sub then holds slot sub / LEAD
                    constexpr int LEAD = LPE / UNROLL;
                    T d[UNROLL];
                    int sv = 0;
                    const int us = sub / LEAD;
#pragma unroll
                    for (int u = 0; u < UNROLL; u++) {
                        T x = (T)0;
#pragma unroll
                        for (int q = 0; q < Q; q++)
#pragma unroll
                            for (int v = 0; v < VEC; v++) x += g[u][q][v] * hr[q][v];
                        if (us == u) sv = src[u];
                        d[u] = x;
                    }
                    const T r = multi_group_sum<T, LPE, UNROLL>(d, sub);
                    if (sub % LEAD == 0 && j + us * EPS + h < nb)
                        put_dval(a, WALK ? e + j + us * EPS + h : sv, r);
                }
            }
            e += nb;
        }
        const bool last = row >= r1;
        if (!DB) {
            if (last) break;
            row++;
            continue;
        }
        if (last && !cut_end) break;
        subgroup_sum<T, VEC, Q, LPE>(acc);
        if (h == 0) {
            // the piece of a row this segment starts inside goes to its carry
            // slot; every other row (and the first piece of a cut row) to dB
            T *drow = (row == st.x && carried_in)
                          ? static_cast<T *>(a.carry) + (size_t)cslot * a.K + a.kbase
                          : static_cast<T *>(a.dB) + (size_t)row * a.K + a.kbase;
#pragma unroll
            for (int q = 0; q < Q; q++) {
                const int k0 = q * LPE * VEC + sub * VEC;
                if (MASK) {
#pragma unroll
                    for (int v = 0; v < VEC; v++)
                        if (k0 + v < a.kcount) drow[k0 + v] = acc[q][v];
                } else {
                    stv<T, VEC>(drow + k0, acc[q]);
                }
            }
        }
#pragma unroll
        for (int q = 0; q < Q; q++)
#pragma unroll
            for (int v = 0; v < VEC; v++) acc[q][v] = (T)0;
        if (last) break;
        row++;
    }
}

// dvals from walk order into CSR order in two passes through LDS, one
// workgroup per chunk of 2^LG values.  Pass 1: chunk c's values are put in
// LDS by rank (bucket, then walk order) and written to their bucket-order
// slots, so the chunk's values for one bucket go out as one run of adjacent
// slots (about 2^LG / buckets of them).  Chunk c + 1's run follows chunk c's
// in every bucket; neighbouring chunks run on one XCD, so the lines the two
// runs share are completed in one L2.
template <class T, int LG>
__device__ __forceinline__ void dv_chunk_to_buckets(const T *__restrict__ v,
                                                    const unsigned short *__restrict__ rank,
                                                    const int *__restrict__ dst, size_t n,
                                                    T *__restrict__ out, unsigned chunk) {
    constexpr int PER = (1 << LG) / 1024;
    __shared__ T lds[1 << LG];
    const size_t base = (size_t)chunk << LG;
    const int cnt = (int)min((size_t)1 << LG, n - base);
    unsigned short r[PER];
    T x[PER];
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const int t = threadIdx.x + i * 1024;
        r[i] = t < cnt ? rank[base + t] : (unsigned short)0;
        x[i] = t < cnt ? v[base + t] : (T)0;
    }
#pragma unroll
    for (int i = 0; i < PER; i++)
        if ((int)threadIdx.x + i * 1024 < cnt) lds[r[i]] = x[i];
    // the destinations are loaded before the barrier, so their latency
    // overlaps the wait for the other waves' LDS stores
    int d[PER];
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const int t = threadIdx.x + i * 1024;
        d[i] = t < cnt ? dst[base + t] : 0;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const int t = threadIdx.x + i * 1024;
        if (t < cnt) out[d[i]] = lds[t];
    }
}

// The fused backward's tail: the fix-up of the dB rows that segment
// boundaries cut and pass 1 of the dvals reorder, which both wait only on
// the walk, in one launch.  Workgroups [0, long_blocks) fix up the rows of
// more than FIX_LONG slots, four at a time (one per 256-thread group); then
// up to fix_blocks the other cut rows, one per wave; the rest are pass 1's
// chunks.  On the 200k-node graph the separate fix-up launch took 9 us; the
// fused backward went 0.617 -> 0.614 ms (K=128) and 0.205 -> 0.202 ms (K=32).
template <class T, int LG>
__global__ __launch_bounds__(1024) void k_bw_tail(RowArgs f, unsigned long_blocks,
                                                  unsigned fix_blocks, const T *__restrict__ v,
                                                  const unsigned short *__restrict__ rank,
                                                  const int *__restrict__ dst, size_t n,
                                                  T *__restrict__ out) {
    if (blockIdx.x < long_blocks) {  // uniform over the workgroup
        const int i = (int)(blockIdx.x * 4 + threadIdx.x / 256);
        split_fixup<T, 256, 4>(f, i, i < f.nlong);
        return;
    }
    if (blockIdx.x < fix_blocks) {
        const int w = (int)(((blockIdx.x - long_blocks) * (size_t)blockDim.x + threadIdx.x) >> 6);
        if (f.nlong + w < f.nfix) split_fixup<T, 64>(f, f.nlong + w);
        return;
    }
    dv_chunk_to_buckets<T, LG>(v, rank, dst, n, out,
                               xcd_block_bw(blockIdx.x - fix_blocks, gridDim.x - fix_blocks));
}

// Pass 2, in place: one workgroup per bucket of 2^LG CSR positions; its
// values (bucket-order slots) go to LDS at their CSR offsets and are written
// back in order.
template <class T, int LG>
__global__ __launch_bounds__(1024) void k_dv_bucket_order(T *__restrict__ d,
                                                          const unsigned short *__restrict__ off,
                                                          size_t n) {
    constexpr int PER = (1 << LG) / 1024;
    __shared__ T lds[1 << LG];
    const size_t base = (size_t)blockIdx.x << LG;
    const int cnt = (int)min((size_t)1 << LG, n - base);
    unsigned short o[PER];
    T x[PER];
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const int t = threadIdx.x + i * 1024;
        o[i] = t < cnt ? off[base + t] : (unsigned short)0;
        x[i] = t < cnt ? d[base + t] : (T)0;
    }
#pragma unroll
    for (int i = 0; i < PER; i++)
        if ((int)threadIdx.x + i * 1024 < cnt) lds[o[i]] = x[i];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const int t = threadIdx.x + i * 1024;
        if (t < cnt) d[base + t] = lds[t];
    }
}

template <class TS, class T, int VEC, int Q, int LPE, bool MASK, bool DB, bool DV>
__global__ __launch_bounds__(256) void k_bw_reside(BwArgs a) {
    const int blk = (int)((blockIdx.x * (size_t)blockDim.x + threadIdx.x) >> 6);
    if (blk < a.nblk) bw_reside_block<TS, T, VEC, Q, LPE, MASK, DB, DV>(a, blk);
}

// dvals alone (SDDMM): the reside blocks run as extra blocks of the walk's
// launch, since the two parts write different entries (C2 K=128: 0.947 ->
// 0.905 ms).  With dB the reside blocks add into rows the walk stores, so they
// run in a launch of their own after the walk.  (Flagging those rows for
// atomics, as the forward launch does, cost more than it saved: the
// transposed 1M-row matrix has many of them, fused backward 1.06 -> 1.13 ms,
// and the flag took the 8-lane walk from 6 to 5 waves per SIMD.)
template <class TS, class T, int VEC, int Q, int LPE, bool MASK, bool DB, bool DV, bool WALK>
static int launch_bw(const BwArgs &a0, hipStream_t s) {
    BwArgs a = a0;
    constexpr bool fuse = !DB;
    a.walk_blocks = (int)ceil_div((size_t)a.nseg * 64, 256);
    const size_t res_blocks = ceil_div((size_t)a.nblk * 64, 256);
    const dim3 grid((unsigned)(a.walk_blocks + (fuse ? res_blocks : 0)));
    if (grid.x) {
        if (a.gstride)
            hipLaunchKernelGGL((k_bw_walk<TS, T, VEC, Q, LPE, MASK, DB, DV, WALK, true>), grid, dim3(256), 0,
                               s, a);
        else
            hipLaunchKernelGGL((k_bw_walk<TS, T, VEC, Q, LPE, MASK, DB, DV, WALK, false>), grid, dim3(256),
                               0, s, a);
    }
    if (!fuse && res_blocks)
        hipLaunchKernelGGL((k_bw_reside<TS, T, VEC, Q, LPE, MASK, DB, DV>), dim3(res_blocks),
                           dim3(256), 0, s, a);
    SWIFT_CHECK_LAUNCH();
    return SWIFT_OK;
}

// After the walk's K slices: dB[row] += the carry slots of each row the
// segments cut (split_fixup over all K; rows with more than FIX_LONG slots by
// a workgroup each, the others by a wave each).
template <class T>
__global__ __launch_bounds__(256) void k_fix_rows(RowArgs a) {
    if (blockIdx.x < (unsigned)a.nlong) {
        split_fixup<T, 256>(a, (int)blockIdx.x);
        return;
    }
    const int w = (int)(((blockIdx.x - a.nlong) * (size_t)blockDim.x + threadIdx.x) >> 6);
    if (a.nlong + w < a.nfix) split_fixup<T, 64>(a, a.nlong + w);
}

// Any K: the slices of plan_row_slices (the forward walk's planner), each one
// launch with its own lane shape; the full-vector shapes need K % VEC == 0
// and 16-byte aligned operands, the rest of K takes masked scalar slices of
// up to 64.  Every slice stores its own columns of dB (the pieces of cut rows
// into carry slots, added by k_fix_rows after the last slice), and dvals
// takes the first slice's partial dots, then each later slice's added on.
// RowArgs of the fix-up of dB's cut rows (split_fixup).
static RowArgs fix_args(const BwArgs &a) {
    RowArgs f{};
    f.C = a.dB;
    f.ldc = a.K;
    f.ldk = a.K;
    f.carry = a.carry;
    f.fix = a.fix;
    f.nfix = a.nfix;
    f.nlong = a.nlong;
    return f;
}

// defer_fix: the caller runs the fix-up of dB's cut rows itself (k_bw_tail).
template <class TS, class T, bool DB, bool DV, bool WALK>
static int dispatch_bw(const BwArgs &a0, hipStream_t s, bool defer_fix = false) {
    constexpr int V = V16<TS>::N;
    const int K = a0.K;
    BwArgs a = a0;
    const bool aligned = K % V == 0 &&
                         (((uintptr_t)a.G | (uintptr_t)a.H | (uintptr_t)a.dB) % 16) == 0;
    const auto slices = plan_row_slices(K, V, aligned);
    for (size_t i = 0; i < slices.size(); i++) {
        const RowTarget &t = slices[i].second;
        a.kbase = slices[i].first;
        a.kcount = t.mask ? std::min(64, K - a.kbase) : t.vec * t.lpe * t.q;
        a.dv_add = i > 0 ? 1 : 0;
        if (a.gstride) a.gbytes = (unsigned)((a0.gbytes / sizeof(TS) - (size_t)a.kbase) * sizeof(TS));
        if (t.mask) {
            SWIFT_TRY((launch_bw<TS, T, 1, 1, 64, true, DB, DV, WALK>(a, s)));
            continue;
        }
#define SWIFT_BW_CASE(Q_, L_)                                                 \
        if (t.q == Q_ && t.lpe == L_) {                                      \
            SWIFT_TRY((launch_bw<TS, T, V, Q_, L_, false, DB, DV, WALK>(a, s)));       \
            continue;                                                        \
        }
        SWIFT_BW_CASE(4, 64)
        SWIFT_BW_CASE(2, 64)
        SWIFT_BW_CASE(1, 64)
        SWIFT_BW_CASE(1, 32)
        SWIFT_BW_CASE(1, 16)
        SWIFT_BW_CASE(1, 8)
#undef SWIFT_BW_CASE
        set_error("no backward kernel for vec=%d q=%d lpe=%d", t.vec, t.q, t.lpe);
        return SWIFT_ERR_UNSUPPORTED;
    }
    if (DB && a.nfix && !defer_fix) {
        hipLaunchKernelGGL(k_fix_rows<T>, dim3((unsigned)a.nlong + ceil_div((size_t)(a.nfix - a.nlong) * 64, 256)),
                           dim3(256), 0, s, fix_args(a));
        SWIFT_CHECK_LAUNCH();
    }
    return SWIFT_OK;
}

static BwArgs walk_args(const swift_matrix *W, int K) {
    BwArgs a{};
    a.m = W->m;
    a.nseg = W->prim.nseg;
    a.seg = W->prim.seg;
    a.rv_ptr = W->rv_ptr;
    a.rv_col = W->rv_col;
    a.rv_src = W->rv_src;
    a.rv_val = W->rv_val;
    a.n = W->n;
    a.nblk = W->rv_all ? 0 : W->numblk;  // rv_all: the walk holds the reside entries
    a.colorblk = W->colorblk;
    a.blkstart = W->blkstart;
    a.blkstop = W->blkstop;
    a.rptr = W->reside_cscptr;
    a.rrow = W->reside_cscrowidx;
    a.rsrc = W->reside_src;
    a.rval = W->reside_val;
    a.perm = W->perm;
    a.seg_slot = W->prim.seg_slot;
    a.fix = W->prim.fix;
    a.nfix = W->prim.nfix;
    a.nlong = W->prim.nlong;
    a.K = K;
    return a;
}

// G is gathered by the walked matrix's column index: W->n rows of K.
template <class T>
static void set_g_buffer(BwArgs &a, const swift_matrix *W, int K) {
    const bool ok = buffer_ok(W->n, K, sizeof(T));
    a.gstride = ok ? (unsigned)(K * sizeof(T)) : 0u;
    a.gbytes = ok ? (unsigned)((size_t)W->n * K * sizeof(T)) : 0u;
}

// dvals = SDDMM(dC, B) walking A: H = dC (per row), G = B (per entry).
// dtype BF16 (an FP32 build): dC and B in BF16, FP32 products, sums and
// dvals; the walk reads no values of A.
int sddmm_impl(const swift_matrix *A, int K, swift_dtype dtype, const void *dC, const void *B,
               void *dvals, hipStream_t s) {
    const bool bf16 = A && dtype == SWIFT_BF16 && A->vtype == SWIFT_F32;
    if (!A || K <= 0 || !dC || !B || (!dvals && A->nnz) || (dtype != A->vtype && !bf16)) {
        set_error("swift_sddmm: bad arguments (dtype must equal the build value type, or be "
                  "BF16 for an FP32 build)");
        return SWIFT_ERR_ARG;
    }
    if (!A->nnz) return SWIFT_OK;  // no entries, no dvals
    BwArgs a = walk_args(A, K);
    a.H = dC;
    a.G = B;
    a.dvals = dvals;
    if (bf16) {
        set_g_buffer<__hip_bfloat16>(a, A, K);
        return dispatch_bw<__hip_bfloat16, float, false, true, false>(a, s);
    }
    if (dtype == SWIFT_F64) set_g_buffer<double>(a, A, K);
    else set_g_buffer<float>(a, A, K);
    return dtype == SWIFT_F64 ? dispatch_bw<double, double, false, true, false>(a, s)
                              : dispatch_bw<float, float, false, true, false>(a, s);
}

// Fused backward walking A^T: H = B (per A^T row = A column), G = dC (per
// entry, gathered once), dB rows are A^T's rows.  dtype BF16: dC, B and A^T's
// values (swift_prepare_values on AT) in BF16, FP32 products, sums, dB and
// dvals; AT must be built from an FP32 matrix.
template <class TS, class T>
static int backward_t(const swift_matrix *AT, int K, const void *dC, const void *B, void *dB,
                      void *dvals, hipStream_t s) {
    BwArgs a = walk_args(AT, K);
    if (!std::is_same<TS, T>::value) a.rv_val = AT->rv_val_lp;
    a.H = B;
    a.G = dC;
    a.dB = dB;
    a.dvals = dvals;
    // dvals in walk order, then two passes through LDS into CSR order: a
    // transpose's entries sit far apart in the original CSR order, and one
    // 4-byte store per entry to scattered lines costs more than the reorder
    const bool walk = dvals && AT->nnz;
    if (walk) {
        if (!AT->dv_dst || a.nblk != 0 || AT->dv_log != DvPlan<T>::BUCKET ||
            AT->dv_clog != DvPlan<T>::CHUNK) {
            set_error("swift_backward_fused: the handle is not a transposed layout "
                      "(use swift_build_transpose)");
            return SWIFT_ERR_ARG;
        }
        SWIFT_TRY(AT->dv_ws.alloc(sizeof(T) * (size_t)AT->nnz));
        a.dvals = AT->dv_ws.ptr;
    }
    set_g_buffer<TS>(a, AT, K);
    if (dB && AT->prim.nsplit)
        SWIFT_TRY(carry_workspace(AT, (size_t)AT->prim.nsplit * K * sizeof(T), s, &a.carry));
    // with dB and walk-order dvals, the fix-up of dB's cut rows runs in the
    // launch of the reorder's first pass (k_bw_tail)
    const bool fused_tail = dB && walk;
    int rc;
    if (dB && dvals) rc = dispatch_bw<TS, T, true, true, true>(a, s, fused_tail);
    else if (dB) rc = dispatch_bw<TS, T, true, false, false>(a, s);
    else rc = dispatch_bw<TS, T, false, true, true>(a, s);
    if (rc != SWIFT_OK || !walk) return rc;
    const size_t n = (size_t)AT->nnz;
    using P = DvPlan<T>;
    const unsigned chunks = ceil_div(n, (size_t)1 << P::CHUNK);
    const unsigned nfix = fused_tail ? (unsigned)a.nfix : 0u;
    const unsigned nlong = fused_tail ? (unsigned)a.nlong : 0u;
    const unsigned long_blocks = ceil_div(nlong, 4);
    const unsigned fix_blocks = long_blocks + ceil_div((size_t)(nfix - nlong) * 64, 1024);
    hipLaunchKernelGGL((k_bw_tail<T, P::CHUNK>), dim3(fix_blocks + chunks), dim3(1024), 0, s,
                       fix_args(a), long_blocks, fix_blocks, (const T *)AT->dv_ws.ptr,
                       AT->dv_rank, AT->dv_dst, n, (T *)dvals);
    hipLaunchKernelGGL((k_dv_bucket_order<T, P::BUCKET>), dim3(ceil_div(n, (size_t)1 << P::BUCKET)),
                       dim3(1024), 0, s, (T *)dvals, AT->dv_off, n);
    SWIFT_CHECK_LAUNCH();
    return SWIFT_OK;
}

int backward_impl(const swift_matrix *AT, int K, swift_dtype dtype, const void *dC, const void *B,
                  void *dB, void *dvals, hipStream_t s) {
    const bool bf16 = AT && dtype == SWIFT_BF16 && AT->vtype == SWIFT_F32;
    if (!AT || K <= 0 || !dC || (!dB && !dvals) || (dvals && !B) ||
        (dtype != AT->vtype && !bf16)) {
        set_error("swift_backward_fused: bad arguments (dtype must equal the build value type, "
                  "or be BF16 for an FP32 build)");
        return SWIFT_ERR_ARG;
    }
    if (bf16) {
        if (!AT->rv_val_lp || AT->lp_type != SWIFT_BF16) {
            set_error("swift_backward_fused: BF16 needs swift_prepare_values(AT, BF16) first");
            return SWIFT_ERR_ARG;
        }
        if (!AT->rv_all) {  // the BF16 walk has no reside blocks of its own
            set_error("swift_backward_fused: BF16 needs a transposed layout "
                      "(swift_build_transpose)");
            return SWIFT_ERR_ARG;
        }
        return backward_t<__hip_bfloat16, float>(AT, K, dC, B, dB, dvals, s);
    }
    return dtype == SWIFT_F64 ? backward_t<double, double>(AT, K, dC, B, dB, dvals, s)
                              : backward_t<float, float>(AT, K, dC, B, dB, dvals, s);
}

}  // namespace swift
