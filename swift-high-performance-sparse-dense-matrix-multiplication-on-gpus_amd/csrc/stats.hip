// stats.hip -- statistics of a built Swift matrix, on the GPU (gfx950).
//
// The figures the reference's study drivers print per matrix: sparsity and
// the share of empty 32x32 blocks (test/MTX_info/main.cu:143-176, :484), and
// the sizes of Swift's regular and irregular parts (test/irr_regular_ratio),
// plus the tile formats of block_catgory and the row/column extremes that the
// batch autotuner keys on.
#include "swift_internal.h"

#include <algorithm>
#include <cstring>

namespace swift {

constexpr int SBLK = 32;             // block edge of the emptiness count
constexpr int SWORDS = 8192;         // bitmap words per pass: 262144 column blocks, 32 KiB

__device__ __forceinline__ unsigned long long block_sum_u64(unsigned long long v) {
    __shared__ unsigned long long part[4];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    unsigned long long t = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x >> 6); i++) t += part[i];
    return t;  // valid in thread 0
}

// One workgroup (256 threads) per band of 32 rows: the column blocks that its
// entries touch are marked in an LDS bitmap and counted, SWORDS*32 column
// blocks per pass.
__global__ __launch_bounds__(256) void k_nonempty_blocks(int m, int nbc,
                                                         const int *__restrict__ rowptr,
                                                         const int *__restrict__ colidx,
                                                         unsigned long long *__restrict__ count) {
    __shared__ unsigned bits[SWORDS];
    const int r0 = blockIdx.x * SBLK, r1 = min(m, r0 + SBLK);
    const int e0 = rowptr[r0], e1 = rowptr[r1];
    unsigned long long local = 0;
    for (int base = 0; base < nbc; base += SWORDS * 32) {
        const int words = min(SWORDS, (nbc - base + 31) / 32);
        for (int w = threadIdx.x; w < words; w += blockDim.x) bits[w] = 0u;
        __syncthreads();
        for (int e = e0 + (int)threadIdx.x; e < e1; e += blockDim.x) {
            const int cb = colidx[e] / SBLK - base;
            if (cb >= 0 && cb < words * 32) atomicOr(&bits[cb >> 5], 1u << (cb & 31));
        }
        __syncthreads();
        for (int w = threadIdx.x; w < words; w += blockDim.x) local += __popc(bits[w]);
        __syncthreads();
    }
    const unsigned long long t = block_sum_u64(local);
    if (threadIdx.x == 0 && t) atomicAdd(count, t);
}

// Entry count of every 32x32 block of a band of 32 rows, SCNT column blocks
// per pass in LDS counters: blocks with at least DENSE_BLOCK entries (half
// full), the entries they hold, and the largest count.  These are the density
// figures that decide whether dense matrix-core tiles could pay.
constexpr int SCNT = 8192;         // counters per pass, 32 KiB
constexpr int DENSE_BLOCK = 512;   // entries of a half-full 32x32 block

__global__ __launch_bounds__(256) void k_block_density(int m, int nbc,
                                                       const int *__restrict__ rowptr,
                                                       const int *__restrict__ colidx,
                                                       unsigned long long *__restrict__ out,
                                                       int *__restrict__ maxcnt) {
    __shared__ unsigned cnt[SCNT];
    const int r0 = blockIdx.x * SBLK, r1 = min(m, r0 + SBLK);
    const int e0 = rowptr[r0], e1 = rowptr[r1];
    if (e0 == e1) return;  // uniform over the workgroup: an empty band
    unsigned long long nd = 0, nz = 0;
    int mx = 0;
    for (int base = 0; base < nbc; base += SCNT) {
        const int words = min(SCNT, nbc - base);
        for (int w = threadIdx.x; w < words; w += blockDim.x) cnt[w] = 0u;
        __syncthreads();
        for (int e = e0 + (int)threadIdx.x; e < e1; e += blockDim.x) {
            const int cb = colidx[e] / SBLK - base;
            if (cb >= 0 && cb < words) atomicAdd(&cnt[cb], 1u);
        }
        __syncthreads();
        for (int w = threadIdx.x; w < words; w += blockDim.x) {
            const int c = (int)cnt[w];
            mx = max(mx, c);
            if (c >= DENSE_BLOCK) {
                nd++;
                nz += (unsigned long long)c;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = max(mx, __shfl_xor(mx, o, 64));
    if ((threadIdx.x & 63) == 0 && mx) atomicMax(maxcnt, mx);
    const unsigned long long tnd = block_sum_u64(nd);
    __syncthreads();  // block_sum_u64's partials are reused by the next sum
    const unsigned long long tnz = block_sum_u64(nz);
    if (threadIdx.x == 0 && tnd) {
        atomicAdd(out, tnd);
        atomicAdd(out + 1, tnz);
    }
}

// Longest row (entries per row from the kept CSR row pointer).
__global__ void k_max_row(int m, const int *__restrict__ rowptr, int *__restrict__ out) {
    int best = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x)
        best = max(best, rowptr[i + 1] - rowptr[i]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) best = max(best, __shfl_xor(best, o, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, best);
}

// Tiles per format (block_catgory's 0 / 1 / 2).
__global__ void k_tile_formats(int tilenum, const int *__restrict__ fmt, int *__restrict__ out) {
    int c[3] = {0, 0, 0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < tilenum; i += gridDim.x * blockDim.x) {
        const int f = fmt[i];
        if (f >= 0 && f < 3) c[f]++;
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        int v = c[k];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        if ((threadIdx.x & 63) == 0 && v) atomicAdd(out + k, v);
    }
}

int matrix_stats_impl(const swift_matrix *M, swift_stats *o, hipStream_t s) {
    *o = swift_stats();
    const long long m = M->m, n = M->n;
    o->sparsity = (m && n) ? (double)M->nnz / ((double)m * (double)n) : 0.0;
    const long long nbr = (m + SBLK - 1) / SBLK, nbc = (n + SBLK - 1) / SBLK;
    o->blocks = nbr * nbc;
    o->regular_nnz = M->regular_nnz;
    o->reside_nnz = M->reside_nnz;
    o->regular_ratio = M->nnz ? (double)M->regular_nnz / M->nnz : 0.0;
    o->tilenum = M->tilenum;
    o->max_col_nnz = M->max_col_nnz;
    o->mean_row_nnz = m ? (double)M->nnz / m : 0.0;
    Buffer buf;
    SWIFT_TRY(buf.alloc(64));
    SWIFT_TRY(fill_bytes_zero(buf.ptr, 64, s));
    // [0] non-empty blocks, [1] dense blocks, [2] their entries (u64); then
    // ints: [0] max row, [1..3] tile formats, [4] max block count
    unsigned long long *cnt = buf.as<unsigned long long>();
    int *ints = reinterpret_cast<int *>(cnt + 3);
    if (nbr && nbc && M->nnz) {
        hipLaunchKernelGGL(k_nonempty_blocks, dim3((unsigned)nbr), dim3(256), 0, s, M->m,
                           (int)nbc, M->csr_rowptr, M->csr_colidx, cnt);
        SWIFT_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_block_density, dim3((unsigned)nbr), dim3(256), 0, s, M->m,
                           (int)nbc, M->csr_rowptr, M->csr_colidx, cnt + 1, ints + 4);
        SWIFT_CHECK_LAUNCH();
    }
    if (m) {
        const unsigned g = (unsigned)std::min<long long>(1024, (m + 255) / 256);
        hipLaunchKernelGGL(k_max_row, dim3(g), dim3(256), 0, s, M->m, M->csr_rowptr, ints);
        SWIFT_CHECK_LAUNCH();
    }
    if (M->tilenum) {
        const unsigned g = (unsigned)std::min<long long>(1024, ((long long)M->tilenum + 255) / 256);
        hipLaunchKernelGGL(k_tile_formats, dim3(g), dim3(256), 0, s, M->tilenum, M->tile_format,
                           ints + 1);
        SWIFT_CHECK_LAUNCH();
    }
    unsigned char h[64];
    SWIFT_HIP(hipMemcpyAsync(h, buf.ptr, 64, hipMemcpyDeviceToHost, s));
    SWIFT_HIP(hipStreamSynchronize(s));
    unsigned long long ne[3];
    int hi[5];
    std::memcpy(ne, h, 24);
    std::memcpy(hi, h + 24, 20);
    o->nonempty_blocks = (long long)ne[0];
    o->dense_blocks = (long long)ne[1];
    o->dense_block_nnz = (long long)ne[2];
    o->max_block_nnz = hi[4];
    o->empty_block_ratio = o->blocks ? (double)(o->blocks - o->nonempty_blocks) / o->blocks : 0.0;
    o->max_row_nnz = hi[0];
    for (int k = 0; k < 3; k++) o->tile_format_count[k] = hi[1 + k];
    return SWIFT_OK;
}

}  // namespace swift
