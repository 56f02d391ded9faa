// preprocess.hip -- Swift format construction on the GPU (gfx950).
//
// Every output index array is bit-exact with the reference's CPU
// preprocessing (src/coo2csc.h, src/block_catgory.h, src/balance.h,
// src/back_process.h) under the stable column-sort tie rule of the
// reference's GPU ColSort (src/ColSort_cuda.h:108).  The reference's GPU
// kernels for this (Pre4SortIdxVal, prepare, auxiliaryTrans) are quadratic;
// here every stage is a scan, a stable radix sort or an embarrassingly
// parallel gather, so a 10^6-column matrix builds in milliseconds.
#include "swift_internal.h"

#include <algorithm>
#include <chrono>
#include <vector>

namespace swift {

// ------------------------------------------------------------ utilities --

// Largest i in [0, n) with a[i] <= x (a non-decreasing, a[0] <= x).
__device__ __forceinline__ int upper_index(const int *__restrict__ a, int n, int x) {
    int lo = 0, hi = n;  // answer in [lo, hi)
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (a[mid] <= x) lo = mid; else hi = mid;
    }
    return lo;
}

template <class T>
__device__ __forceinline__ void copy_val(void *dst, size_t di, const void *src, size_t si) {
    static_cast<T *>(dst)[di] = static_cast<const T *>(src)[si];
}

__device__ __forceinline__ void copy_value(int vs, void *dst, size_t di,
                                           const void *src, size_t si) {
    if (vs == 8) copy_val<double>(dst, di, src, si);
    else copy_val<float>(dst, di, src, si);
}

static int vsize_of(swift_dtype t) { return t == SWIFT_F64 ? 8 : 4; }

static unsigned grid1(size_t n, unsigned block = 256) {
    size_t g = (n + block - 1) / block;
    return (unsigned)(g ? g : 1);
}

// ------------------------------------------------------- input checking --

// One in-bounds read pass over a device CSR before anything is scattered by
// its indices.  Error bits: 1 rowptr[0] != 0, 2 rowptr[m] != nnz, 4 rowptr
// decreases somewhere, 8 a column index outside [0, n).  A monotone row
// pointer from 0 to nnz keeps every row range inside the column array.
__global__ void k_check_csr(const int *__restrict__ rowptr, int m,
                            const int *__restrict__ colidx, int nnz, int n,
                            int *__restrict__ flag) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    int bad = 0;
    if (t == 0) {
        if (rowptr[0] != 0) bad |= 1;
        if (rowptr[m] != nnz) bad |= 2;
    }
    for (size_t i = t; i < (size_t)m; i += stride)
        if (rowptr[i + 1] < rowptr[i]) bad |= 4;
    for (size_t e = t; e < (size_t)nnz; e += stride)
        if ((unsigned)colidx[e] >= (unsigned)n) bad |= 8;
    for (int o = 32; o > 0; o >>= 1) bad |= __shfl_xor(bad, o, 64);
    if ((threadIdx.x & 63) == 0 && bad) atomicOr(flag, bad);
}

// Validates a caller's device CSR (m rows, n columns, nnz entries) on stream
// s; returns SWIFT_ERR_ARG naming the first problem found.
static int check_csr(const char *who, int m, int n, int nnz, const int *d_rowptr,
                     const int *d_colidx, hipStream_t s) {
    Buffer flag;
    SWIFT_TRY(flag.alloc(sizeof(int)));
    SWIFT_TRY(fill_i32(flag.as<int>(), 0, 1, s));
    const size_t work = std::max<size_t>((size_t)m, (size_t)nnz);
    const unsigned grid = (unsigned)std::min<size_t>(std::max<size_t>(1, (work + 255) / 256), 4096);
    hipLaunchKernelGGL(k_check_csr, dim3(grid), dim3(256), 0, s, d_rowptr, m, d_colidx, nnz, n,
                       flag.as<int>());
    SWIFT_CHECK_LAUNCH();
    int bad = 0;
    SWIFT_HIP(hipMemcpyAsync(&bad, flag.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
    SWIFT_HIP(hipStreamSynchronize(s));
    if (bad & 1) { set_error("%s: rowptr[0] != 0", who); return SWIFT_ERR_ARG; }
    if (bad & 2) { set_error("%s: rowptr[m] != nnz (%d)", who, nnz); return SWIFT_ERR_ARG; }
    if (bad & 4) { set_error("%s: rowptr decreases", who); return SWIFT_ERR_ARG; }
    if (bad & 8) {
        set_error("%s: a column index is outside [0, %d)", who, n);
        return SWIFT_ERR_ARG;
    }
    return SWIFT_OK;
}

// ---------------------------------------------------------- csr -> csc --

__global__ void k_col_hist(const int *__restrict__ colidx, int nnz, int *__restrict__ npc) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < nnz) atomicAdd(&npc[colidx[e]], 1);
}

// p-th CSC slot <- CSR entry e = order[p]: row by binary search of rowptr.
__global__ void k_csc_gather(const unsigned *__restrict__ order, int nnz,
                             const int *__restrict__ rowptr, int m, int vs,
                             const void *__restrict__ csr_val, int *__restrict__ cscrow,
                             void *__restrict__ cscval, int *__restrict__ cscsrc) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nnz) return;
    int e = (int)order[p];
    cscrow[p] = upper_index(rowptr, m + 1, e);
    if (cscval) copy_value(vs, cscval, p, csr_val, e);
    if (cscsrc) cscsrc[p] = e;
}

struct CscTemp {
    Buffer keys, vals, kalt, valt, tmp;
};

// Stable counting transpose: radix sort of (col, e) pairs keeps e ascending
// inside each column, which is CSR order = rows ascending (src/coo2csc.h:44-56).
static int csr_to_csc_impl(int m, int n, int nnz, swift_dtype vtype,
                           const void *d_val, const int *d_rowptr,
                           const int *d_colidx, void *d_cscval, int *d_cscrow,
                           int *d_cscptr, int *d_npc, int *d_cscsrc, CscTemp &T,
                           hipStream_t s) {
    SWIFT_TRY(fill_i32(d_npc, 0, (size_t)n, s));
    if (nnz > 0) {
        hipLaunchKernelGGL(k_col_hist, dim3(grid1(nnz)), dim3(256), 0, s, d_colidx, nnz, d_npc);
        SWIFT_CHECK_LAUNCH();
    }
    SWIFT_TRY(T.tmp.alloc(std::max(scan_temp_bytes((size_t)n), radix_temp_bytes((size_t)nnz))));
    SWIFT_TRY(exclusive_scan(d_npc, d_cscptr, (size_t)n, d_cscptr + n, T.tmp.ptr, s));
    if (nnz == 0) return SWIFT_OK;
    size_t b = (size_t)nnz * sizeof(unsigned);
    SWIFT_TRY(T.keys.alloc(b)); SWIFT_TRY(T.vals.alloc(b));
    SWIFT_TRY(T.kalt.alloc(b)); SWIFT_TRY(T.valt.alloc(b));
    int bits = bits_for((unsigned)(n > 0 ? n - 1 : 0));
    SWIFT_TRY(radix_sort_pairs((const unsigned *)d_colidx, nullptr, T.keys.as<unsigned>(),
                               T.vals.as<unsigned>(), T.kalt.as<unsigned>(),
                               T.valt.as<unsigned>(), (size_t)nnz, bits, T.tmp.ptr, s));
    hipLaunchKernelGGL(k_csc_gather, dim3(grid1(nnz)), dim3(256), 0, s,
                       T.vals.as<unsigned>(), nnz, d_rowptr, m, vsize_of(vtype), d_val,
                       d_cscrow, d_cscval, d_cscsrc);
    SWIFT_CHECK_LAUNCH();
    return SWIFT_OK;
}

// ------------------------------------------------------------ col sort --

__global__ void k_max_reduce(const int *__restrict__ a, int n, int *__restrict__ out) {
    int v = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        v = max(v, a[i]);
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_down(v, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, v);
}

// sorted slot p <- CSC slot of column perm[i] at depth p - cptr[i]
__global__ void k_sorted_gather(int nnz, int n, const int *__restrict__ cptr,
                                const int *__restrict__ perm, const int *__restrict__ cscptr,
                                const int *__restrict__ cscrow, const void *__restrict__ cscval,
                                const int *__restrict__ cscsrc, int vs,
                                int *__restrict__ srow, void *__restrict__ sval,
                                int *__restrict__ ssrc) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nnz) return;
    int i = upper_index(cptr, n + 1, p);
    int src = cscptr[perm[i]] + (p - cptr[i]);
    srow[p] = cscrow[src];
    if (sval) copy_value(vs, sval, p, cscval, src);
    if (ssrc) ssrc[p] = cscsrc ? cscsrc[src] : src;
}

template <class T>
__global__ void k_permute_rows(const T *__restrict__ in, T *__restrict__ out,
                               const int *__restrict__ perm, int n, int K) {
    size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t total = (size_t)n * K;
    for (; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
        int i = (int)(idx / K), k = (int)(idx % K);
        out[idx] = in[(size_t)perm[i] * K + k];
    }
}

template <class T>
__global__ void k_permute_cols(const T *__restrict__ in, T *__restrict__ out,
                               const int *__restrict__ perm, int n, int K) {
    size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t total = (size_t)n * K;
    for (; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
        int k = (int)(idx / n), i = (int)(idx % n);
        out[idx] = in[(size_t)k * n + perm[i]];
    }
}

static int permute_dense_impl(const int *perm, int n, int K, swift_layout layout,
                              size_t es, const void *in, void *out, hipStream_t s) {
    size_t total = (size_t)n * K;
    if (total == 0) return SWIFT_OK;
    unsigned g = (unsigned)std::min<size_t>((total + 255) / 256, 65536);
    if (layout == SWIFT_ROW_MAJOR) {
        if (es == 8) hipLaunchKernelGGL(k_permute_rows<double>, dim3(g), dim3(256), 0, s, (const double *)in, (double *)out, perm, n, K);
        else if (es == 4) hipLaunchKernelGGL(k_permute_rows<float>, dim3(g), dim3(256), 0, s, (const float *)in, (float *)out, perm, n, K);
        else hipLaunchKernelGGL(k_permute_rows<unsigned short>, dim3(g), dim3(256), 0, s, (const unsigned short *)in, (unsigned short *)out, perm, n, K);
    } else {
        if (es == 8) hipLaunchKernelGGL(k_permute_cols<double>, dim3(g), dim3(256), 0, s, (const double *)in, (double *)out, perm, n, K);
        else if (es == 4) hipLaunchKernelGGL(k_permute_cols<float>, dim3(g), dim3(256), 0, s, (const float *)in, (float *)out, perm, n, K);
        else hipLaunchKernelGGL(k_permute_cols<unsigned short>, dim3(g), dim3(256), 0, s, (const unsigned short *)in, (unsigned short *)out, perm, n, K);
    }
    SWIFT_CHECK_LAUNCH();
    return SWIFT_OK;
}

struct SortTemp {
    Buffer keys, kalt, valt, tmp, maxv;
};

// Stable argsort of nnzpercol (radix sort of (nnz, j) keeps j ascending among
// equal counts), then the sorted-CSC gather.  cptr (n+1) is the exclusive scan
// of sortnnz, i.e. nnzcol_ptr of src/coo2csc.h:162.
static int col_sort_impl(int n, int nnz, swift_dtype vtype, const int *d_npc,
                         const int *d_cscptr, const int *d_cscrow, const void *d_cscval,
                         const int *d_cscsrc, int *d_srow, void *d_sval, int *d_ssrc,
                         int *d_sortnnz, int *d_perm, int *d_cptr, int *h_maxcount,
                         SortTemp &T, hipStream_t s) {
    if (n == 0) return SWIFT_OK;
    SWIFT_TRY(T.maxv.alloc(sizeof(int)));
    SWIFT_TRY(fill_i32(T.maxv.as<int>(), 0, 1, s));
    hipLaunchKernelGGL(k_max_reduce, dim3(std::min<unsigned>(grid1(n), 1024)), dim3(256), 0, s,
                       d_npc, n, T.maxv.as<int>());
    SWIFT_CHECK_LAUNCH();
    int maxc = 0;
    SWIFT_HIP(hipMemcpyAsync(&maxc, T.maxv.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
    SWIFT_HIP(hipStreamSynchronize(s));
    if (h_maxcount) *h_maxcount = maxc;
    size_t b = (size_t)n * sizeof(unsigned);
    SWIFT_TRY(T.kalt.alloc(b)); SWIFT_TRY(T.valt.alloc(b));
    SWIFT_TRY(T.tmp.alloc(std::max(radix_temp_bytes((size_t)n), scan_temp_bytes((size_t)n))));
    SWIFT_TRY(radix_sort_pairs((const unsigned *)d_npc, nullptr, (unsigned *)d_sortnnz,
                               (unsigned *)d_perm, T.kalt.as<unsigned>(), T.valt.as<unsigned>(),
                               (size_t)n, bits_for((unsigned)maxc), T.tmp.ptr, s));
    SWIFT_TRY(exclusive_scan(d_sortnnz, d_cptr, (size_t)n, d_cptr + n, T.tmp.ptr, s));
    if (nnz > 0) {
        hipLaunchKernelGGL(k_sorted_gather, dim3(grid1(nnz)), dim3(256), 0, s, nnz, n, d_cptr,
                           d_perm, d_cscptr, d_cscrow, d_cscval, d_cscsrc, vsize_of(vtype),
                           d_srow, d_sval, d_ssrc);
        SWIFT_CHECK_LAUNCH();
    }
    return SWIFT_OK;
}

// ---------------------------------------------------- staircase tiling --

__global__ void k_step_flags(const int *__restrict__ sn, int n, int *__restrict__ flag) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    int base = c ? sn[c - 1] : 0;
    flag[c] = (sn[c] - base) != 0;
}

// Step s at sorted column c: base depth, tall = new depths, its tile count
// floor((N-c)/32)*tall, and leftover width (N-c)%32 (src/coo2csc.h:165-197).
__global__ void k_step_compact(const int *__restrict__ sn, int n, const int *__restrict__ flag,
                               const int *__restrict__ idx, int *__restrict__ scol,
                               int *__restrict__ sbase, int *__restrict__ stall,
                               int *__restrict__ sntile, int *__restrict__ sw) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n || !flag[c]) return;
    int s = idx[c];
    int base = c ? sn[c - 1] : 0;
    int tall = sn[c] - base;
    scol[s] = c;
    sbase[s] = base;
    stall[s] = tall;
    sntile[s] = ((n - c) / TILE) * tall;
    sw[s] = (n - c) % TILE;
}

// One block: reside column r collects tall_s from every step with width > r
// (src/coo2csc.h:356-380); reside_col = max width.
__global__ void k_reside_counts(const int *__restrict__ sw, const int *__restrict__ stall,
                                int nsteps, int *__restrict__ rcount, int *__restrict__ rmax) {
    __shared__ int cnt[TILE];
    __shared__ int mx;
    if (threadIdx.x < TILE) cnt[threadIdx.x] = 0;
    if (threadIdx.x == 0) mx = 0;
    __syncthreads();
    int local[TILE];
#pragma unroll
    for (int r = 0; r < TILE; r++) local[r] = 0;
    int lmax = 0;
    for (int s = threadIdx.x; s < nsteps; s += blockDim.x) {
        int w = sw[s], t = stall[s];
        lmax = max(lmax, w);
#pragma unroll
        for (int r = 0; r < TILE; r++) local[r] += (r < w) ? t : 0;
    }
#pragma unroll
    for (int r = 0; r < TILE; r++)
        if (local[r]) atomicAdd(&cnt[r], local[r]);
    atomicMax(&mx, lmax);
    __syncthreads();
    if (threadIdx.x < TILE) rcount[threadIdx.x] = cnt[threadIdx.x];
    if (threadIdx.x == 0) *rmax = mx;
}

// Regular entry p = 32*t + l: tile t lies in step s; (j, jx) = block and
// depth inside the step; source = depth base+jx of sorted column c0+l
// (src/coo2csc.h:221-278).
__global__ void k_tile_fill(int regular_nnz, int nsteps, const int *__restrict__ tileoff,
                            const int *__restrict__ scol, const int *__restrict__ sbase,
                            const int *__restrict__ stall, const int *__restrict__ cptr,
                            const int *__restrict__ srow_tmp, const void *__restrict__ sval_tmp,
                            const int *__restrict__ ssrc_tmp, int vs,
                            int *__restrict__ tile_ptr, int *__restrict__ tile_len,
                            int *__restrict__ tile_colidx, int *__restrict__ tile_tall,
                            int *__restrict__ srow, void *__restrict__ sval,
                            int *__restrict__ ssrc) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= regular_nnz) return;
    int t = p >> 5, l = p & 31;
    int s = upper_index(tileoff, nsteps + 1, t);
    int rel = t - tileoff[s];
    int tall = stall[s];
    int j = rel / tall, jx = rel - j * tall;
    int c0 = scol[s] + TILE * j;
    int src = cptr[c0 + l] + sbase[s] + jx;
    srow[p] = srow_tmp[src];
    copy_value(vs, sval, p, sval_tmp, src);
    ssrc[p] = ssrc_tmp[src];
    if (l == 0) {
        tile_ptr[t] = TILE * t;
        tile_len[t] = TILE;
        tile_colidx[t] = c0;
        tile_tall[t] = tall;
    }
}

__global__ void k_set_int(int *p, int v) { *p = v; }

// One block per reside column r (sorted column jj = N-1-r).  Depth d of jj
// belongs to the step whose depth range holds d; it is reside iff that
// step's leftover width exceeds r.  Depth order = (step, depth) order of
// src/coo2csc.h:384-400.
__global__ void k_reside_fill(int n, int nsteps, const int *__restrict__ sbase,
                              const int *__restrict__ sw, const int *__restrict__ sn,
                              const int *__restrict__ cptr, const int *__restrict__ rptr,
                              const int *__restrict__ srow_tmp, const void *__restrict__ sval_tmp,
                              const int *__restrict__ ssrc_tmp, int vs,
                              int *__restrict__ rrow, void *__restrict__ rval,
                              int *__restrict__ rsrc) {
    const int r = blockIdx.x;
    const int jj = n - 1 - r;
    const int len = sn[jj];
    int out = rptr[r];
    __shared__ int wsum[4];
    for (int d0 = 0; d0 < len; d0 += blockDim.x) {
        int d = d0 + threadIdx.x;
        int take = 0;
        if (d < len) {
            int s = upper_index(sbase, nsteps, d);
            take = sw[s] > r;
        }
        // block exclusive scan of take (256 threads)
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        uint64_t bal = __ballot(take);
        int before_w = __popcll(bal & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))));
        if (lane == 0) wsum[w] = __popcll(bal);
        __syncthreads();
        int before = 0, all = 0;
        for (int q = 0; q < (int)(blockDim.x >> 6); q++) {
            before += q < w ? wsum[q] : 0;
            all += wsum[q];
        }
        if (take) {
            int pos = out + before + before_w;
            int src = cptr[jj] + d;
            rrow[pos] = srow_tmp[src];
            copy_value(vs, rval, pos, sval_tmp, src);
            rsrc[pos] = ssrc_tmp[src];
        }
        out += all;
        __syncthreads();
    }
}

// ------------------------------------------------------ categorisation --

// Two tiles per wave, one entry per lane.  Row changes between neighbouring
// lanes give the format (src/block_catgory.h:11-72) and run count.
__global__ void k_tile_cat(int tilenum, const int *__restrict__ srow,
                           int *__restrict__ fmt, int *__restrict__ countrow) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    int t = p >> 5, l = p & 31;
    bool valid = t < tilenum;
    int row = valid ? srow[p] : 0;
    int prev = __shfl_up(row, 1, 64);
    bool change = valid && l != 0 && row != prev;
    uint64_t bal = __ballot(change);
    const int lane = threadIdx.x & 63;
    unsigned half = (unsigned)(bal >> (lane & 32));
    if (valid && l == 0) {
        int changes = __popc(half);
        fmt[t] = changes == 0 ? 0 : (changes < 16 ? 2 : 1);
        countrow[t] = changes + 1;
    }
}

// Run heads: segmentoffset[run] = run length - 1 and sortrowindex[head] =
// same, other slots 0 (src/block_catgory.h:110-159).
__global__ void k_tile_seg(int tilenum, const int *__restrict__ srow,
                           const int *__restrict__ countrow, int *__restrict__ segoff,
                           int *__restrict__ sortrowindex) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    int t = p >> 5, l = p & 31;
    bool valid = t < tilenum;
    int row = valid ? srow[p] : 0;
    int prev = __shfl_up(row, 1, 64);
    bool head = valid && (l == 0 || row != prev);
    uint64_t bal = __ballot(head);
    const int lane = threadIdx.x & 63;
    unsigned heads = (unsigned)(bal >> (lane & 32));
    if (!valid) return;
    int v = 0;
    if (head) {
        unsigned after = (l == 31) ? 0u : (heads >> (l + 1));
        int next = after ? (l + 1 + __builtin_ctz(after)) : TILE;
        int run = __popc(heads & ((l == 31) ? 0xFFFFFFFFu : ((1u << (l + 1)) - 1u))) - 1;
        v = next - l - 1;
        segoff[countrow[t] + run] = v;
    }
    sortrowindex[p] = v;
}

// --------------------------------------------------------------- balance --

// Reside columns with > 32 entries split into ceil(n/32) blocks of length
// ceil(n/nb) (src/balance.h:1-69).  One thread per reside column, prefix in
// a single wave (reside_col <= 31).
__global__ void k_balance(int reside_col, const int *__restrict__ rptr,
                          unsigned *__restrict__ colorblk, int *__restrict__ blkstart,
                          int *__restrict__ blkstop, int *__restrict__ blkptr,
                          int *__restrict__ numblk_out) {
    int r = threadIdx.x;
    int len = 0, nb = 0;
    if (r < reside_col) {
        len = rptr[r + 1] - rptr[r];
        nb = len <= TILE ? 1 : (len + TILE - 1) / TILE;
    }
    int inc = nb;
    for (int off = 1; off < 64; off <<= 1) {
        int t = __shfl_up(inc, off, 64);
        if (r >= off) inc += t;
    }
    int first = inc - nb;
    if (r == 63) *numblk_out = inc;
    if (r >= reside_col) return;
    blkptr[r] = first;
    if (len <= TILE) {
        colorblk[first] = (unsigned)r;
        blkstart[first] = 0;
        blkstop[first] = 0;
    } else {
        int bl = (len + nb - 1) / nb;
        for (int j = 0; j < nb; j++) {
            colorblk[first + j] = (unsigned)r | 0x80000000u;
            blkstart[first + j] = rptr[r] + j * bl;
            blkstop[first + j] = (j == nb - 1) ? rptr[r] + len : rptr[r] + (j + 1) * bl;
        }
    }
}

// ----------------------------------------------------- work list build --

// Work item = up to `depth` consecutive depth tiles of one (step, column
// block): the tiles that read the same 32 B rows.
__global__ void k_work_count(int nsteps, int n, const int *__restrict__ scol,
                             const int *__restrict__ stall, int depth, int *__restrict__ cnt) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsteps) return;
    int nblk = (n - scol[s]) / TILE;
    cnt[s] = nblk * ((stall[s] + depth - 1) / depth);
}

__global__ void k_work_fill(int nwork, int nsteps, const int *__restrict__ woff,
                            const int *__restrict__ tileoff, const int *__restrict__ stall,
                            int depth, int2 *__restrict__ work) {
    int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nwork) return;
    int s = upper_index(woff, nsteps + 1, w);
    int rel = w - woff[s];
    int tall = stall[s];
    int per = (tall + depth - 1) / depth;
    int j = rel / per, c = rel - j * per;
    int start = tileoff[s] + j * tall + c * depth;
    int cnt = min(depth, tall - c * depth);
    work[w] = make_int2(start, cnt);
}

// ------------------------------------------------------------ row view --

__global__ void k_row_hist(const int *__restrict__ rows, int n, int *__restrict__ cnt) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) atomicAdd(&cnt[rows[p]], 1);
}

// q-th entry of the row view <- entry p = order[q].  p < R: regular entry p,
// whose B row is the original column of sorted column tile_colidx[p/32] +
// p%32.  p >= R (row views of every entry): reside entry p - R of reside
// column r (rptr[r] <= p - R < rptr[r+1], at most 32 columns), sorted column
// n-1-r.
__global__ void k_rv_fill(int R, int RV, const unsigned *__restrict__ order,
                          const int *__restrict__ tile_colidx, const int *__restrict__ perm,
                          const int *__restrict__ srow, const void *__restrict__ sval,
                          const int *__restrict__ ssrc, int vs, int n, int rcols,
                          const int *__restrict__ rptr, const int *__restrict__ rrow,
                          const void *__restrict__ rval, const int *__restrict__ rsrc,
                          int *__restrict__ rv_col, void *__restrict__ rv_val,
                          int *__restrict__ rv_src, int *__restrict__ rv_row) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= RV) return;
    int p = (int)order[q];
    if (p < R) {
        int c = tile_colidx[p >> 5] + (p & 31);
        rv_col[q] = perm[c];
        copy_value(vs, rv_val, q, sval, p);
        rv_src[q] = ssrc[p];
        rv_row[q] = srow[p];
    } else {
        const int e = p - R;
        int lo = 0, hi = rcols;  // largest r with rptr[r] <= e
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (rptr[mid] <= e) lo = mid;
            else hi = mid;
        }
        rv_col[q] = perm[n - 1 - lo];
        copy_value(vs, rv_val, q, rval, e);
        rv_src[q] = rsrc[e];
        rv_row[q] = rrow[e];
    }
}
// rows whose regular part is empty
__global__ void k_empty_flag(int m, const int *__restrict__ ptr, int *__restrict__ flag) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < m) flag[r] = ptr[r + 1] == ptr[r] ? 1 : 0;
}
__global__ void k_empty_compact(int m, const int *__restrict__ flag, const int *__restrict__ idx,
                                int *__restrict__ out) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < m && flag[r]) out[idx[r]] = r;
}

__global__ void k_mark_rows(const int *__restrict__ rows, int n, int *__restrict__ flag) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flag[rows[i]] = 1;  // equal values: concurrent writes agree
}
__global__ void k_flag_bytes(int m, const int *__restrict__ flag, unsigned char *__restrict__ out) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < m) out[r] = flag[r] ? 1 : 0;
}

__global__ void k_inv_perm(const int *__restrict__ perm, int n, int *__restrict__ inv) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) inv[perm[i]] = i;
}

// Merge path over rows and entries: the end of row i is item rv_ptr[i+1] + i,
// entry e of row r is item e + r.  Segment s nominally starts at item d = s*W
// (the rows finished before d are those with rv_ptr[i+1] + i < d); a start
// that falls inside row r is moved back to r's first entry when r began less
// than W items earlier, so that only rows longer than a segment are split.
// Starts stay non-decreasing: a moved start lies after the previous nominal
// start.
__global__ void k_mp_seg(int m, int R, int W, int nseg, const int *__restrict__ ptr,
                         int2 *__restrict__ seg) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s > nseg) return;
    long long d = min((long long)s * W, (long long)m + R);
    int lo = 0, hi = m;  // count of i in [0, m) with ptr[i+1] + i < d
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if ((long long)ptr[mid + 1] + mid < d) lo = mid + 1; else hi = mid;
    }
    int e = (int)(d - lo);
    if (s > 0 && s < nseg && lo < m && e > ptr[lo] && (long long)ptr[lo] + lo > d - W)
        e = ptr[lo];
    seg[s] = make_int2(lo, e);
}

// A segment ending inside row r (entries of r before and after its end) makes
// r a split row, walked in pieces by several segments.
__global__ void k_mp_split_flag(int m, int nseg, const int *__restrict__ ptr,
                                const int2 *__restrict__ seg, int *__restrict__ flag) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg) return;
    int2 end = seg[s + 1];
    flag[s] = (end.x < m && end.y > ptr[end.x]) ? 1 : 0;
}

__global__ void k_mp_split_compact(int nseg, const int2 *__restrict__ seg,
                                   const int *__restrict__ flag, const int *__restrict__ idx,
                                   int *__restrict__ rows) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nseg && flag[s]) rows[idx[s]] = seg[s + 1].x;
}

// Boundary b opens a run when it cuts another row than boundary b - 1.
__global__ void k_run_flag(int n, const int *__restrict__ rows, int *__restrict__ flag) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < n) flag[b] = (b == 0 || rows[b] != rows[b - 1]) ? 1 : 0;
}

// first[i] = first boundary of run i; first[nruns] = n.
__global__ void k_run_compact(int n, const int *__restrict__ flag, const int *__restrict__ idx,
                              int nruns, int *__restrict__ first) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < n && flag[b]) first[idx[b]] = b;
    if (b == n) first[nruns] = n;
}

// Sort key of run i: longer runs (more boundaries) first.
__global__ void k_run_key(int nruns, int n, const int *__restrict__ first,
                          unsigned *__restrict__ key) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nruns) key[i] = (unsigned)(n - (first[i + 1] - first[i]));
}

// fix[i] = (row, first boundary, end boundary, 0) of run order[i].
__global__ void k_run_fix(int nruns, const unsigned *__restrict__ order,
                          const int *__restrict__ first, const int *__restrict__ rows,
                          int4 *__restrict__ fix) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nruns) return;
    const int r = (int)order[i];
    const int b0 = first[r];
    fix[i] = make_int4(rows[b0], b0, first[r + 1], 0);
}

// ------------------------------------------------------------ row sort --

// Groups of 8 tiles = 256 entries (the last group holds tilenum%8 tiles).
// Bitonic sort in LDS of (row << 8 | slot) keys: unique, so the result is the
// stable sort of src/back_process.h:73 with ties by slot.
__global__ __launch_bounds__(256) void k_row_sort(int tilenum, const int *__restrict__ srow,
                                                  int *__restrict__ val_position,
                                                  int *__restrict__ val_offset,
                                                  int *__restrict__ rowsorted,
                                                  int *__restrict__ rowsort_src) {
    __shared__ unsigned long long key[256];
    __shared__ int sorted[256];
    const int g = blockIdx.x;
    const int t0 = g * 8;
    const int nt = min(8, tilenum - t0);
    const int len = nt * TILE;
    const size_t base = (size_t)t0 * TILE;
    const int i = threadIdx.x;
    key[i] = (i < len) ? (((unsigned long long)(unsigned)srow[base + i] << 8) | (unsigned)i)
                       : ~0ull;
    __syncthreads();
    for (int k = 2; k <= 256; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            int ixj = i ^ j;
            if (ixj > i) {
                unsigned long long a = key[i], b = key[ixj];
                bool up = (i & k) == 0;
                if ((a > b) == up) { key[i] = b; key[ixj] = a; }
            }
            __syncthreads();
        }
    }
    if (i < len) {
        unsigned long long kv = key[i];
        int row = (int)(kv >> 8);
        int slot = (int)(kv & 0xFF);
        sorted[i] = row;
        rowsorted[base + i] = row;
        val_position[base + slot] = i;
        rowsort_src[base + i] = slot;
    }
    __syncthreads();
    // val_offset (src/back_process.h:128-205): head = run length - 1,
    // member = 0, singleton = -1.  The remainder group's loop has no branch
    // for a final singleton that follows a multi-entry run (it stays 0).
    if (i < len) {
        int row = sorted[i];
        bool head = (i == 0) || sorted[i - 1] != row;
        int v = 0;
        if (head) {
            int e = i + 1;
            while (e < len && sorted[e] == row) e++;
            v = (e - i > 1) ? e - i - 1 : -1;
            // remainder group, last slot, singleton after a run of >= 2
            if (v == -1 && nt < 8 && i == len - 1 && i >= 2 && sorted[i - 2] == sorted[i - 1])
                v = 0;
        }
        val_offset[base + i] = v;
    }
}

// ------------------------------------------------------------ matrix build --

static void free_seg_table(swift_matrix::SegTable &t) {
    if (!t.shared)
        for (void *p : {(void *)t.seg, (void *)t.split_rows, (void *)t.seg_slot, (void *)t.fix})
            if (p) (void)hipFree(p);
    t = swift_matrix::SegTable();
}

// A table of W items that one of the matrix's default tables already holds
// is shared with it, not built again.  The default tables are built once,
// before the primary one, and freed with the matrix, so a shared table never
// outlives the arrays it points to.
static bool share_table(swift_matrix *M, int W, swift_matrix::SegTable &t) {
    for (const auto &o : M->cls) {
        if (&o == &t || !o.seg || o.shared || o.W != W) continue;
        free_seg_table(t);
        t = o;
        t.shared = true;
        return true;
    }
    return false;
}

static void free_matrix_arrays(swift_matrix *M) {
    void *ptrs[] = {M->perm, M->sortnnz, M->cptr, M->sortrowidx_tmp, M->sortval_tmp,
                    M->sortsrc_tmp, M->step_col, M->step_base, M->step_tall,
                    M->step_tileoff, M->tile_ptr, M->tile_len, M->tile_colidx,
                    M->tile_tall, M->tile_format, M->sortrowidx, M->sortval, M->sortsrc,
                    M->countrow, M->segmentoffset, M->sortrowindex, M->reside_cscptr,
                    M->reside_cscrowidx, M->reside_val, M->reside_src, M->colorblk,
                    M->blkstart, M->blkstop, M->blkptr, M->val_position, M->val_offset,
                    M->rowsorted, M->rowsort_src, M->csr_rowptr, M->csr_colidx,
                    M->sortval_lp, M->reside_val_lp, M->work, M->rv_ptr, M->rv_col,
                    M->rv_val, M->rv_src, M->rv_val_lp, M->inv_perm, M->rv_row, M->empty_rows, M->resrow_flag, M->dv_rank,
                    M->dv_off, M->dv_dst};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    free_seg_table(M->prim);
    for (auto &t : M->cls) free_seg_table(t);
    M->scratch.release();
    M->dv_ws.release();
    std::lock_guard<std::mutex> g(M->carry_mu);
    M->carry_ws.clear();
}

template <class T>
static int dalloc(T **p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)p, count * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        set_error("hipMalloc(%zu): %s", count * sizeof(T), hipGetErrorString(e));
        return SWIFT_ERR_ALLOC;
    }
    return SWIFT_OK;
}

static int dalloc_bytes(void **p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) bytes = 8;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        set_error("hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
        return SWIFT_ERR_ALLOC;
    }
    return SWIFT_OK;
}

int build_rowview(swift_matrix *M, hipStream_t s);

int build_work_list(swift_matrix *M, int depth, hipStream_t s) {
    if (M->work) { (void)hipFree(M->work); M->work = nullptr; }
    M->nwork = 0;
    M->work_depth = depth;
    if (M->tilenum == 0) return SWIFT_OK;
    Buffer cnt, tmp;
    SWIFT_TRY(cnt.alloc(sizeof(int) * ((size_t)M->nsteps + 1)));
    SWIFT_TRY(tmp.alloc(scan_temp_bytes((size_t)M->nsteps + 1)));
    hipLaunchKernelGGL(k_work_count, dim3(grid1(M->nsteps)), dim3(256), 0, s, M->nsteps, M->n,
                       M->step_col, M->step_tall, depth, cnt.as<int>());
    SWIFT_CHECK_LAUNCH();
    int *woff = cnt.as<int>();
    SWIFT_TRY(exclusive_scan(woff, woff, (size_t)M->nsteps, woff + M->nsteps, tmp.ptr, s));
    int nwork = 0;
    SWIFT_HIP(hipMemcpyAsync(&nwork, woff + M->nsteps, sizeof(int), hipMemcpyDeviceToHost, s));
    SWIFT_HIP(hipStreamSynchronize(s));
    SWIFT_TRY(dalloc(&M->work, (size_t)nwork));
    if (nwork > 0) {
        hipLaunchKernelGGL(k_work_fill, dim3(grid1(nwork)), dim3(256), 0, s, nwork, M->nsteps,
                           woff, M->step_tileoff, M->step_tall, depth, M->work);
        SWIFT_CHECK_LAUNCH();
    }
    M->nwork = nwork;
    SWIFT_HIP(hipStreamSynchronize(s));
    cnt.release();
    tmp.release();
    return SWIFT_OK;
}

// Host clock in ms; the build synchronises its stream at each phase end, so
// phase times taken on the host are the phases' device times plus launch and
// allocation overhead, the way the reference times its preprocessing
// (gettimeofday, test/preprocess/main.cu:440-456).
static double wall_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

int build_impl(swift_matrix *M, const int *d_rowptr, const int *d_colidx,
               const void *d_val, hipStream_t s) {
    const int m = M->m, n = M->n, nnz = M->nnz;
    const int vs = vsize_of(M->vtype);
    // the kernels below scatter by the caller's indices: check them first
    // (outside the phase times)
    SWIFT_TRY(check_csr("swift_build", m, n, nnz, d_rowptr, d_colidx, s));
    const double t0 = wall_ms();
    double t_sort = t0, t_block = t0;
    auto stamp = [&]() {
        const double t = wall_ms();
        M->build_ms[0] = (float)(t_sort - t0);
        M->build_ms[1] = (float)(t_block - t_sort);
        M->build_ms[2] = (float)(t - t_block);
        M->build_ms[3] = (float)(t - t0);
    };
    // keep the input pattern (transposed layout, SDDMM, value updates)
    SWIFT_TRY(dalloc(&M->csr_rowptr, (size_t)m + 1));
    SWIFT_TRY(dalloc(&M->csr_colidx, (size_t)nnz));
    SWIFT_HIP(hipMemcpyAsync(M->csr_rowptr, d_rowptr, sizeof(int) * ((size_t)m + 1),
                             hipMemcpyDeviceToDevice, s));
    if (nnz)
        SWIFT_HIP(hipMemcpyAsync(M->csr_colidx, d_colidx, sizeof(int) * (size_t)nnz,
                                 hipMemcpyDeviceToDevice, s));

    // 1. CSR -> CSC
    int *cscptr = nullptr, *cscrow = nullptr, *npc = nullptr, *cscsrc = nullptr;
    void *cscval = nullptr;
    SWIFT_TRY(dalloc(&cscptr, (size_t)n + 1));
    SWIFT_TRY(dalloc(&npc, (size_t)n));
    SWIFT_TRY(dalloc(&cscrow, (size_t)nnz));
    SWIFT_TRY(dalloc(&cscsrc, (size_t)nnz));
    SWIFT_TRY(dalloc_bytes(&cscval, (size_t)nnz * vs));
    struct Guard {
        void *p[5];
        ~Guard() { for (void *q : p) if (q) (void)hipFree(q); }
    } guard{{cscptr, npc, cscrow, cscsrc, cscval}};
    {
        CscTemp T;
        SWIFT_TRY(csr_to_csc_impl(m, n, nnz, M->vtype, d_val, d_rowptr, d_colidx, cscval, cscrow,
                                  cscptr, npc, cscsrc, T, s));
        SWIFT_HIP(hipStreamSynchronize(s));
    }
    // 2. column sort
    SWIFT_TRY(dalloc(&M->perm, (size_t)n));
    SWIFT_TRY(dalloc(&M->sortnnz, (size_t)n));
    SWIFT_TRY(dalloc(&M->cptr, (size_t)n + 1));
    SWIFT_TRY(dalloc(&M->sortrowidx_tmp, (size_t)nnz));
    SWIFT_TRY(dalloc(&M->sortsrc_tmp, (size_t)nnz));
    SWIFT_TRY(dalloc_bytes(&M->sortval_tmp, (size_t)nnz * vs));
    {
        SortTemp T;
        SWIFT_TRY(col_sort_impl(n, nnz, M->vtype, npc, cscptr, cscrow, cscval, cscsrc,
                                M->sortrowidx_tmp, M->sortval_tmp, M->sortsrc_tmp, M->sortnnz,
                                M->perm, M->cptr, &M->max_col_nnz, T, s));
        SWIFT_HIP(hipStreamSynchronize(s));
    }
    t_sort = t_block = wall_ms();
    if (n == 0) {
        SWIFT_TRY(build_rowview(M, s));
        stamp();
        return SWIFT_OK;
    }
    // 3. steps
    Buffer flag, idx, ntile, sw, tmp, small;
    SWIFT_TRY(flag.alloc(sizeof(int) * ((size_t)n + 1)));
    SWIFT_TRY(idx.alloc(sizeof(int) * ((size_t)n + 1)));
    SWIFT_TRY(tmp.alloc(scan_temp_bytes((size_t)n + 1)));
    SWIFT_TRY(small.alloc(sizeof(int) * 64));
    hipLaunchKernelGGL(k_step_flags, dim3(grid1(n)), dim3(256), 0, s, M->sortnnz, n,
                       flag.as<int>());
    SWIFT_CHECK_LAUNCH();
    SWIFT_TRY(exclusive_scan(flag.as<int>(), idx.as<int>(), (size_t)n, small.as<int>(), tmp.ptr, s));
    int nsteps = 0;
    SWIFT_HIP(hipMemcpyAsync(&nsteps, small.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
    SWIFT_HIP(hipStreamSynchronize(s));
    M->nsteps = nsteps;
    SWIFT_TRY(dalloc(&M->step_col, (size_t)nsteps));
    SWIFT_TRY(dalloc(&M->step_base, (size_t)nsteps));
    SWIFT_TRY(dalloc(&M->step_tall, (size_t)nsteps));
    SWIFT_TRY(dalloc(&M->step_tileoff, (size_t)nsteps + 1));
    SWIFT_TRY(ntile.alloc(sizeof(int) * ((size_t)nsteps + 1)));
    SWIFT_TRY(sw.alloc(sizeof(int) * ((size_t)nsteps + 1)));
    hipLaunchKernelGGL(k_step_compact, dim3(grid1(n)), dim3(256), 0, s, M->sortnnz, n,
                       flag.as<int>(), idx.as<int>(), M->step_col, M->step_base, M->step_tall,
                       ntile.as<int>(), sw.as<int>());
    SWIFT_CHECK_LAUNCH();
    int *rcount = small.as<int>() + 1;   // 32 counts
    int *rmax = small.as<int>() + 33;
    int *tilenum_d = small.as<int>() + 34;
    SWIFT_TRY(exclusive_scan(ntile.as<int>(), M->step_tileoff, (size_t)nsteps,
                             M->step_tileoff + nsteps, tmp.ptr, s));
    hipLaunchKernelGGL(k_reside_counts, dim3(1), dim3(256), 0, s, sw.as<int>(), M->step_tall,
                       nsteps, rcount, rmax);
    SWIFT_CHECK_LAUNCH();
    int hs[34];
    SWIFT_HIP(hipMemcpyAsync(hs, small.as<int>() + 1, sizeof(int) * 33, hipMemcpyDeviceToHost, s));
    SWIFT_HIP(hipMemcpyAsync(tilenum_d, M->step_tileoff + nsteps, sizeof(int),
                             hipMemcpyDeviceToDevice, s));
    int tilenum = 0;
    SWIFT_HIP(hipMemcpyAsync(&tilenum, M->step_tileoff + nsteps, sizeof(int),
                             hipMemcpyDeviceToHost, s));
    SWIFT_HIP(hipStreamSynchronize(s));
    int reside_col = hs[32];
    std::vector<int> rptr_h((size_t)reside_col + 1, 0);
    for (int r = 0; r < reside_col; r++) rptr_h[r + 1] = rptr_h[r] + hs[r];
    M->tilenum = tilenum;
    M->regular_nnz = tilenum * TILE;
    M->reside_col = reside_col;
    M->reside_nnz = rptr_h[reside_col];
    if ((long long)M->regular_nnz + M->reside_nnz != nnz) {
        set_error("swift_build: regular %d + reside %d != nnz %d", M->regular_nnz,
                  M->reside_nnz, nnz);
        return SWIFT_ERR_ARG;
    }
    // 4. tiles
    const int T = tilenum, R = M->regular_nnz;
    SWIFT_TRY(dalloc(&M->tile_ptr, (size_t)T + 1));
    SWIFT_TRY(dalloc(&M->tile_len, (size_t)T));
    SWIFT_TRY(dalloc(&M->tile_colidx, (size_t)T));
    SWIFT_TRY(dalloc(&M->tile_tall, (size_t)T));
    SWIFT_TRY(dalloc(&M->tile_format, (size_t)T));
    SWIFT_TRY(dalloc(&M->countrow, (size_t)T + 1));
    SWIFT_TRY(dalloc(&M->sortrowidx, (size_t)R));
    SWIFT_TRY(dalloc(&M->sortsrc, (size_t)R));
    SWIFT_TRY(dalloc(&M->sortrowindex, (size_t)R));
    SWIFT_TRY(dalloc_bytes(&M->sortval, (size_t)R * vs));
    if (R > 0) {
        hipLaunchKernelGGL(k_tile_fill, dim3(grid1(R)), dim3(256), 0, s, R, nsteps,
                           M->step_tileoff, M->step_col, M->step_base, M->step_tall, M->cptr,
                           M->sortrowidx_tmp, M->sortval_tmp, M->sortsrc_tmp, vs, M->tile_ptr,
                           M->tile_len, M->tile_colidx, M->tile_tall, M->sortrowidx,
                           M->sortval, M->sortsrc);
        SWIFT_CHECK_LAUNCH();
    }
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, s, M->tile_ptr + T, R);
    SWIFT_CHECK_LAUNCH();
    // 5. reside
    SWIFT_TRY(dalloc(&M->reside_cscptr, (size_t)reside_col + 1));
    SWIFT_TRY(dalloc(&M->reside_cscrowidx, (size_t)M->reside_nnz));
    SWIFT_TRY(dalloc(&M->reside_src, (size_t)M->reside_nnz));
    SWIFT_TRY(dalloc_bytes(&M->reside_val, (size_t)M->reside_nnz * vs));
    SWIFT_HIP(hipMemcpyAsync(M->reside_cscptr, rptr_h.data(), sizeof(int) * rptr_h.size(),
                             hipMemcpyHostToDevice, s));
    if (reside_col > 0) {
        hipLaunchKernelGGL(k_reside_fill, dim3(reside_col), dim3(256), 0, s, n, nsteps,
                           M->step_base, sw.as<int>(), M->sortnnz, M->cptr, M->reside_cscptr,
                           M->sortrowidx_tmp, M->sortval_tmp, M->sortsrc_tmp, vs,
                           M->reside_cscrowidx, M->reside_val, M->reside_src);
        SWIFT_CHECK_LAUNCH();
    }
    // 6. categorisation
    if (T > 0) {
        hipLaunchKernelGGL(k_tile_cat, dim3(grid1(R)), dim3(256), 0, s, T, M->sortrowidx,
                           M->tile_format, M->countrow);
        SWIFT_CHECK_LAUNCH();
    }
    SWIFT_TRY(tmp.alloc(scan_temp_bytes((size_t)T + 1)));
    SWIFT_TRY(exclusive_scan(M->countrow, M->countrow, (size_t)T, M->countrow + T, tmp.ptr, s));
    int segsum = 0;
    SWIFT_HIP(hipMemcpyAsync(&segsum, M->countrow + T, sizeof(int), hipMemcpyDeviceToHost, s));
    SWIFT_HIP(hipStreamSynchronize(s));
    M->segsum = segsum;
    SWIFT_TRY(dalloc(&M->segmentoffset, (size_t)segsum + 1));
    SWIFT_TRY(fill_i32(M->segmentoffset, 0, (size_t)segsum + 1, s));
    if (T > 0) {
        hipLaunchKernelGGL(k_tile_seg, dim3(grid1(R)), dim3(256), 0, s, T, M->sortrowidx,
                           M->countrow, M->segmentoffset, M->sortrowindex);
        SWIFT_CHECK_LAUNCH();
    }
    // 7. balance
    int nb_bound = reside_col + (M->reside_nnz + TILE - 1) / TILE + 1;
    SWIFT_TRY(dalloc(&M->colorblk, (size_t)nb_bound));
    SWIFT_TRY(dalloc(&M->blkstart, (size_t)nb_bound));
    SWIFT_TRY(dalloc(&M->blkstop, (size_t)nb_bound));
    SWIFT_TRY(dalloc(&M->blkptr, (size_t)nb_bound));
    SWIFT_TRY(fill_i32(M->blkptr, 0, (size_t)nb_bound, s));
    hipLaunchKernelGGL(k_balance, dim3(1), dim3(64), 0, s, reside_col, M->reside_cscptr,
                       M->colorblk, M->blkstart, M->blkstop, M->blkptr, small.as<int>());
    SWIFT_CHECK_LAUNCH();
    int numblk = 0;
    SWIFT_HIP(hipMemcpyAsync(&numblk, small.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
    SWIFT_HIP(hipStreamSynchronize(s));
    M->numblk = numblk;
    t_block = wall_ms();
    // 8. kernel schedules: column-tile work list and the row view
    SWIFT_TRY(build_work_list(M, 16, s));
    SWIFT_TRY(build_rowview(M, s));
    SWIFT_HIP(hipStreamSynchronize(s));
    stamp();
    return SWIFT_OK;
}

// Merge-path segment table of W items: (row, entry) start of every segment,
// the rows cut by a segment boundary, the carry slot of each segment end and
// the runs of boundaries that cut one row.
static int build_seg_table(swift_matrix *M, int W, hipStream_t s, swift_matrix::SegTable &t) {
    free_seg_table(t);
    t.W = W;
    const long long items = (long long)M->m + M->rv_nnz;
    const int nseg = (int)((items + W - 1) / W);
    t.nseg = nseg;
    SWIFT_TRY(dalloc(&t.seg, (size_t)nseg + 1));
    hipLaunchKernelGGL(k_mp_seg, dim3(grid1((size_t)nseg + 1)), dim3(256), 0, s, M->m,
                       M->rv_nnz, W, nseg, M->rv_ptr, t.seg);
    SWIFT_CHECK_LAUNCH();
    SWIFT_TRY(dalloc(&t.seg_slot, (size_t)nseg + 1));
    if (nseg == 0) {
        SWIFT_TRY(dalloc(&t.split_rows, 1));
        return dalloc(&t.fix, 1);
    }
    Buffer flag, tmp;
    SWIFT_TRY(flag.alloc(sizeof(int) * ((size_t)nseg + 1)));
    SWIFT_TRY(tmp.alloc(scan_temp_bytes((size_t)nseg + 1)));
    hipLaunchKernelGGL(k_mp_split_flag, dim3(grid1(nseg)), dim3(256), 0, s, M->m, nseg, M->rv_ptr,
                       t.seg, flag.as<int>());
    SWIFT_CHECK_LAUNCH();
    SWIFT_TRY(exclusive_scan(flag.as<int>(), t.seg_slot, (size_t)nseg, t.seg_slot + nseg,
                             tmp.ptr, s));
    int nsplit = 0;
    SWIFT_HIP(hipMemcpyAsync(&nsplit, t.seg_slot + nseg, sizeof(int), hipMemcpyDeviceToHost, s));
    SWIFT_HIP(hipStreamSynchronize(s));
    SWIFT_TRY(dalloc(&t.split_rows, (size_t)nsplit));
    if (nsplit) {
        hipLaunchKernelGGL(k_mp_split_compact, dim3(grid1(nseg)), dim3(256), 0, s, nseg, t.seg,
                           flag.as<int>(), t.seg_slot, t.split_rows);
        SWIFT_CHECK_LAUNCH();
        // runs of equal rows in split_rows (one per cut row)
        Buffer run, ridx;
        SWIFT_TRY(run.alloc(sizeof(int) * ((size_t)nsplit + 1)));
        SWIFT_TRY(ridx.alloc(sizeof(int) * ((size_t)nsplit + 1)));
        hipLaunchKernelGGL(k_run_flag, dim3(grid1(nsplit)), dim3(256), 0, s, nsplit, t.split_rows,
                           run.as<int>());
        SWIFT_CHECK_LAUNCH();
        SWIFT_TRY(exclusive_scan(run.as<int>(), ridx.as<int>(), (size_t)nsplit,
                                 ridx.as<int>() + nsplit, tmp.ptr, s));
        int nfix = 0;
        SWIFT_HIP(hipMemcpyAsync(&nfix, ridx.as<int>() + nsplit, sizeof(int),
                                 hipMemcpyDeviceToHost, s));
        SWIFT_HIP(hipStreamSynchronize(s));
        // run i starts at first[i]; the fix-up takes the runs longest first
        Buffer first, key, order, alt, rtmp;
        SWIFT_TRY(first.alloc(sizeof(int) * ((size_t)nfix + 1)));
        hipLaunchKernelGGL(k_run_compact, dim3(grid1((size_t)nsplit + 1)), dim3(256), 0, s, nsplit,
                           run.as<int>(), ridx.as<int>(), nfix, first.as<int>());
        SWIFT_CHECK_LAUNCH();
        SWIFT_TRY(key.alloc(sizeof(unsigned) * 2 * (size_t)nfix));
        SWIFT_TRY(order.alloc(sizeof(unsigned) * (size_t)nfix));
        SWIFT_TRY(alt.alloc(sizeof(unsigned) * 2 * (size_t)nfix));
        SWIFT_TRY(rtmp.alloc(radix_temp_bytes((size_t)nfix)));
        unsigned *k0 = key.as<unsigned>(), *k1 = k0 + nfix;
        unsigned *a0 = alt.as<unsigned>(), *a1 = a0 + nfix;
        hipLaunchKernelGGL(k_run_key, dim3(grid1(nfix)), dim3(256), 0, s, nfix, nsplit,
                           first.as<int>(), k0);
        SWIFT_CHECK_LAUNCH();
        SWIFT_TRY(radix_sort_pairs(k0, nullptr, k1, order.as<unsigned>(), a0, a1, (size_t)nfix,
                                   bits_for((unsigned)nsplit), rtmp.ptr, s));
        SWIFT_TRY(dalloc(&t.fix, (size_t)nfix));
        hipLaunchKernelGGL(k_run_fix, dim3(grid1(nfix)), dim3(256), 0, s, nfix,
                           order.as<unsigned>(), first.as<int>(), t.split_rows, t.fix);
        SWIFT_CHECK_LAUNCH();
        // runs of more than FIX_LONG slots: sorted keys below nsplit - FIX_LONG
        std::vector<unsigned> keys((size_t)nfix);
        SWIFT_HIP(hipMemcpyAsync(keys.data(), k1, sizeof(unsigned) * nfix, hipMemcpyDeviceToHost, s));
        SWIFT_HIP(hipStreamSynchronize(s));  // the temporaries are freed at the end of this block
        t.nlong = (int)(std::lower_bound(keys.begin(), keys.end(),
                                         (unsigned)std::max(0, nsplit - FIX_LONG)) - keys.begin());
        t.nfix = nfix;
    } else {
        SWIFT_TRY(dalloc(&t.fix, 1));
    }
    SWIFT_HIP(hipStreamSynchronize(s));
    t.nsplit = nsplit;
    return SWIFT_OK;
}

// The primary segment table (used by every row-view kernel unless the narrow
// row table applies).
int build_segments(swift_matrix *M, int W, hipStream_t s) {
    M->mp_items = W;
    if (share_table(M, W, M->prim)) return SWIFT_OK;
    return build_seg_table(M, W, s, M->prim);
}

// The default tables.  A wave's segment of W items should keep enough B
// bytes in flight and leave enough segments to fill the device: W is the
// size measured best on the 1M-row benchmark matrix for the class's B rows
// (48 items for >= 1 KiB rows, 64 for >= 512 B, 128 for 256..511 B of 16-bit
// elements and 96 for wider ones, 128 for narrower rows; FP32 K=128 64 items
// 0.719 ms against 96 items 0.744 ms, BF16 K=128 128 items 0.416 ms against
// 192 items 0.447 ms), raised to 3 mean rows (items per row), which a matrix
// of long rows (the 200k-node graph: 51 items per row) prefers.  A cut row
// costs one carry slot.  Classes that come to the same W share one table.
int build_class_segments(swift_matrix *M, hipStream_t s) {
    // measured on the 1M-row matrix with the carry slots and the batch
    // prefetch: FP32 K=32 128 items 0.258 ms against 256 items 0.270 ms, FP64
    // K=32 96 items 0.348 against 128 items 0.361 ms, FP32 K=64 96 items
    // 0.340 against 128 items 0.352 ms, BF16 K=32 128 items 0.264 against
    // 256 items 0.268 ms
    static const int by_bytes[SEG_CLASSES] = {48, 64, 128, 96, 128};
    const long long m = std::max(M->m, 1);
    const int mean = (int)((m + M->rv_nnz + m - 1) / m);
    for (int c = 0; c < SEG_CLASSES; c++) {
        const int W = std::min(1024, std::max(by_bytes[c], 3 * mean));
        if (!share_table(M, W, M->cls[c])) SWIFT_TRY(build_seg_table(M, W, s, M->cls[c]));
    }
    return SWIFT_OK;
}

// Row view of the regular part (with rv_all, of every entry): stable radix
// sort of the entries by row, so each row keeps its regular entries in tile
// order, followed by its reside entries.
int build_rowview(swift_matrix *M, hipStream_t s) {
    const int m = M->m, R0 = M->regular_nnz;
    // A small reside part (at most 1% of the entries) is walked with the
    // regular entries: its balance blocks would each load a B row for few
    // entries and add into finished rows with atomics, in a launch of their
    // own after the walk.
    if (!M->rv_all && M->reside_nnz > 0 && (long long)M->reside_nnz * 100 <= (long long)M->nnz)
        M->rv_all = true;
    const int NRES = M->rv_all ? M->reside_nnz : 0;
    const int R = R0 + NRES;
    M->rv_nnz = R;
    const int vs = vsize_of(M->vtype);
    SWIFT_TRY(dalloc(&M->rv_ptr, (size_t)m + 1));
    SWIFT_TRY(dalloc(&M->rv_col, (size_t)R));
    SWIFT_TRY(dalloc(&M->rv_src, (size_t)R));
    SWIFT_TRY(dalloc(&M->rv_row, (size_t)R));
    SWIFT_TRY(dalloc_bytes(&M->rv_val, (size_t)R * vs));
    SWIFT_TRY(dalloc(&M->inv_perm, (size_t)M->n));
    if (M->n) {
        hipLaunchKernelGGL(k_inv_perm, dim3(grid1(M->n)), dim3(256), 0, s, M->perm, M->n,
                           M->inv_perm);
        SWIFT_CHECK_LAUNCH();
    }
    Buffer cnt, tmp, keys, vals, kalt, valt;
    SWIFT_TRY(cnt.alloc(sizeof(int) * ((size_t)m + 1)));
    SWIFT_TRY(fill_i32(cnt.as<int>(), 0, (size_t)m + 1, s));
    if (R0) {
        hipLaunchKernelGGL(k_row_hist, dim3(grid1(R0)), dim3(256), 0, s, M->sortrowidx, R0,
                           cnt.as<int>());
        SWIFT_CHECK_LAUNCH();
    }
    if (NRES) {
        hipLaunchKernelGGL(k_row_hist, dim3(grid1(NRES)), dim3(256), 0, s, M->reside_cscrowidx,
                           NRES, cnt.as<int>());
        SWIFT_CHECK_LAUNCH();
    }
    SWIFT_TRY(tmp.alloc(std::max(scan_temp_bytes((size_t)m + 1), radix_temp_bytes((size_t)R))));
    SWIFT_TRY(exclusive_scan(cnt.as<int>(), M->rv_ptr, (size_t)m, M->rv_ptr + m, tmp.ptr, s));
    if (R) {
        size_t b = sizeof(unsigned) * (size_t)R;
        SWIFT_TRY(keys.alloc(b)); SWIFT_TRY(vals.alloc(b));
        SWIFT_TRY(kalt.alloc(b)); SWIFT_TRY(valt.alloc(b));
        // sort keys: the regular entries' rows, then the reside entries' rows
        Buffer rows;
        const unsigned *kin = (const unsigned *)M->sortrowidx;
        if (NRES) {
            SWIFT_TRY(rows.alloc(b));
            if (R0)
                SWIFT_HIP(hipMemcpyAsync(rows.ptr, M->sortrowidx, sizeof(int) * (size_t)R0,
                                         hipMemcpyDeviceToDevice, s));
            SWIFT_HIP(hipMemcpyAsync(rows.as<int>() + R0, M->reside_cscrowidx,
                                     sizeof(int) * (size_t)NRES, hipMemcpyDeviceToDevice, s));
            kin = rows.as<unsigned>();
        }
        SWIFT_TRY(radix_sort_pairs(kin, nullptr, keys.as<unsigned>(), vals.as<unsigned>(),
                                   kalt.as<unsigned>(), valt.as<unsigned>(), (size_t)R,
                                   bits_for((unsigned)(m > 0 ? m - 1 : 0)), tmp.ptr, s));
        hipLaunchKernelGGL(k_rv_fill, dim3(grid1(R)), dim3(256), 0, s, R0, R,
                           vals.as<unsigned>(), M->tile_colidx, M->perm, M->sortrowidx,
                           M->sortval, M->sortsrc, vs, M->n, M->reside_col, M->reside_cscptr,
                           M->reside_cscrowidx, M->reside_val, M->reside_src, M->rv_col,
                           M->rv_val, M->rv_src, M->rv_row);
        SWIFT_CHECK_LAUNCH();
        SWIFT_HIP(hipStreamSynchronize(s));  // rows is freed at the end of this block
    }
    // rows without regular entries
    if (m) {
        int *flag = cnt.as<int>();  // m+1 ints, free again
        Buffer idx;
        SWIFT_TRY(idx.alloc(sizeof(int) * ((size_t)m + 1)));
        hipLaunchKernelGGL(k_empty_flag, dim3(grid1(m)), dim3(256), 0, s, m, M->rv_ptr, flag);
        SWIFT_CHECK_LAUNCH();
        SWIFT_TRY(exclusive_scan(flag, idx.as<int>(), (size_t)m, idx.as<int>() + m, tmp.ptr, s));
        int ne = 0;
        SWIFT_HIP(hipMemcpyAsync(&ne, idx.as<int>() + m, sizeof(int), hipMemcpyDeviceToHost, s));
        SWIFT_HIP(hipStreamSynchronize(s));
        SWIFT_TRY(dalloc(&M->empty_rows, (size_t)ne));
        if (ne) {
            hipLaunchKernelGGL(k_empty_compact, dim3(grid1(m)), dim3(256), 0, s, m, flag,
                               idx.as<int>(), M->empty_rows);
            SWIFT_CHECK_LAUNCH();
        }
        M->nempty = ne;
        // rows that receive reside entries
        SWIFT_TRY(fill_i32(flag, 0, (size_t)m + 1, s));
        if (M->reside_nnz) {
            hipLaunchKernelGGL(k_mark_rows, dim3(grid1(M->reside_nnz)), dim3(256), 0, s,
                               M->reside_cscrowidx, M->reside_nnz, flag);
            SWIFT_CHECK_LAUNCH();
        }
        SWIFT_TRY(dalloc(&M->resrow_flag, (size_t)m));
        hipLaunchKernelGGL(k_flag_bytes, dim3(grid1(m)), dim3(256), 0, s, m, flag,
                           M->resrow_flag);
        SWIFT_CHECK_LAUNCH();
    }
    SWIFT_HIP(hipStreamSynchronize(s));
    // the default tables first: the primary table shares one of the same size
    SWIFT_TRY(build_class_segments(M, s));
    return build_segments(M, 128, s);
}

int row_sort_impl(swift_matrix *M, hipStream_t s) {
    const size_t R = (size_t)M->regular_nnz;
    if (!M->val_position) {
        SWIFT_TRY(dalloc(&M->val_position, R));
        SWIFT_TRY(dalloc(&M->val_offset, R));
        SWIFT_TRY(dalloc(&M->rowsorted, R));
        SWIFT_TRY(dalloc(&M->rowsort_src, R));
    }
    int groups = (M->tilenum + 7) / 8;
    if (groups > 0) {
        hipLaunchKernelGGL(k_row_sort, dim3(groups), dim3(256), 0, s, M->tilenum, M->sortrowidx,
                           M->val_position, M->val_offset, M->rowsorted, M->rowsort_src);
        SWIFT_CHECK_LAUNCH();
    }
    SWIFT_HIP(hipStreamSynchronize(s));
    M->has_rowsort = 1;
    return SWIFT_OK;
}

int permute_dense(const swift_matrix *M, int K, swift_layout layout, size_t es,
                  const void *in, void *out, hipStream_t s) {
    return permute_dense_impl(M->perm, M->n, K, layout, es, in, out, s);
}

void destroy_matrix(swift_matrix *M) {
    if (!M) return;
    free_matrix_arrays(M);
    delete M;
}

// ------------------------------------------------------ C entry points --

int csr_to_csc_entry(int m, int n, int nnz, swift_dtype vtype, const void *d_val,
                     const int *d_rowptr, const int *d_colidx, void *d_cscval, int *d_cscrow,
                     int *d_cscptr, int *d_npc, hipStream_t s) {
    SWIFT_TRY(check_csr("swift_csr_to_csc", m, n, nnz, d_rowptr, d_colidx, s));
    CscTemp T;
    SWIFT_TRY(csr_to_csc_impl(m, n, nnz, vtype, d_val, d_rowptr, d_colidx, d_cscval, d_cscrow,
                              d_cscptr, d_npc, nullptr, T, s));
    SWIFT_HIP(hipStreamSynchronize(s));
    return SWIFT_OK;
}

int csr_to_csc_with_src(int m, int n, int nnz, swift_dtype vtype, const void *d_val,
                        const int *d_rowptr, const int *d_colidx, void *d_cscval, int *d_cscrow,
                        int *d_cscptr, int *d_npc, int *d_cscsrc, hipStream_t s) {
    CscTemp T;
    SWIFT_TRY(csr_to_csc_impl(m, n, nnz, vtype, d_val, d_rowptr, d_colidx, d_cscval, d_cscrow,
                              d_cscptr, d_npc, d_cscsrc, T, s));
    SWIFT_HIP(hipStreamSynchronize(s));
    return SWIFT_OK;
}

int col_sort_entry(int n, int nnz, int K, swift_dtype vtype, const int *d_npc,
                   const int *d_cscptr, const int *d_cscrow, const void *d_cscval,
                   int *d_srow, void *d_sval, int *d_sortnnz, int *d_perm, const void *d_B,
                   void *d_sortB, hipStream_t s) {
    SortTemp T;
    Buffer cptr;
    SWIFT_TRY(cptr.alloc(sizeof(int) * ((size_t)n + 1)));
    SWIFT_TRY(col_sort_impl(n, nnz, vtype, d_npc, d_cscptr, d_cscrow, d_cscval, nullptr, d_srow,
                            d_sval, nullptr, d_sortnnz, d_perm, cptr.as<int>(), nullptr, T, s));
    if (d_B && d_sortB && K > 0)
        SWIFT_TRY(permute_dense_impl(d_perm, n, K, SWIFT_ROW_MAJOR, (size_t)vsize_of(vtype), d_B,
                                     d_sortB, s));
    SWIFT_HIP(hipStreamSynchronize(s));
    return SWIFT_OK;
}

}  // namespace swift
