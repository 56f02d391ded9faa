// sort_scan.hip -- device-wide exclusive scan and stable LSD radix sort.
//
// Both are written for 64-lane wavefronts: wave scans use DPP-free shuffles
// (preprocessing is a one-off cost next to the SpMM it enables), ranks inside
// a wave come from 8 ballots per 8-bit digit, and every block is 256 threads
// (4 waves).  The radix sort is stable, which is what makes the column sort
// bit-exact: equal keys keep input order.
#include "swift_internal.h"

#include <cstdarg>
#include <cstring>

namespace swift {

// ------------------------------------------------------------ error text --
static thread_local char g_err[512];

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

const char *last_error() { return g_err; }

int Buffer::alloc(size_t n) {
    if (n <= bytes && ptr) return SWIFT_OK;
    release();
    if (n == 0) n = 256;
    hipError_t e = hipMalloc(&ptr, n);
    if (e != hipSuccess) {
        ptr = nullptr;
        set_error("hipMalloc(%zu): %s", n, hipGetErrorString(e));
        return SWIFT_ERR_ALLOC;
    }
    bytes = n;
    return SWIFT_OK;
}

void Buffer::release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

int bits_for(unsigned v) {
    int b = 0;
    while (b < 32 && (v >> b) != 0u) b++;
    return b;
}

// ------------------------------------------------------------------ fills --
__global__ void k_fill_i32(int *p, int v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

__global__ void k_fill_zero16(uint4 *p, size_t n16) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n16; i += stride) p[i] = make_uint4(0, 0, 0, 0);
}

__global__ void k_fill_zero1(unsigned char *p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void k_iota(int *p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = (int)i;
}

static unsigned grid_for(size_t n, unsigned block = 256, unsigned cap = 8192) {
    size_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

int fill_i32(int *p, int v, size_t n, hipStream_t s) {
    if (n == 0) return SWIFT_OK;
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(n)), dim3(256), 0, s, p, v, n);
    SWIFT_CHECK_LAUNCH();
    return SWIFT_OK;
}

int fill_bytes_zero(void *p, size_t bytes, hipStream_t s) {
    if (bytes == 0) return SWIFT_OK;
    uintptr_t a = (uintptr_t)p;
    if ((a & 15) == 0) {
        size_t n16 = bytes / 16;
        if (n16) {
            hipLaunchKernelGGL(k_fill_zero16, dim3(grid_for(n16, 256, 16384)),
                               dim3(256), 0, s, (uint4 *)p, n16);
            SWIFT_CHECK_LAUNCH();
        }
        size_t tail = bytes - n16 * 16;
        if (tail) {
            hipLaunchKernelGGL(k_fill_zero1, dim3(1), dim3(256), 0, s,
                               (unsigned char *)p + n16 * 16, tail);
            SWIFT_CHECK_LAUNCH();
        }
        return SWIFT_OK;
    }
    hipLaunchKernelGGL(k_fill_zero1, dim3(ceil_div(bytes, 256)), dim3(256), 0, s,
                       (unsigned char *)p, bytes);
    SWIFT_CHECK_LAUNCH();
    return SWIFT_OK;
}

int iota_i32(int *p, size_t n, hipStream_t s) {
    if (n == 0) return SWIFT_OK;
    hipLaunchKernelGGL(k_iota, dim3(grid_for(n)), dim3(256), 0, s, p, n);
    SWIFT_CHECK_LAUNCH();
    return SWIFT_OK;
}

// ------------------------------------------------------------------- scan --
constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_CHUNK = SCAN_BLOCK * SCAN_ITEMS;  // 4096 ints per block

__device__ __forceinline__ int wave_inclusive_scan(int v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        int t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
    }
    return v;
}

// Block-wide exclusive scan of one int per thread (256 threads); returns the
// exclusive prefix and the block total in *total.
__device__ __forceinline__ int block_exclusive_scan(int v, int *total) {
    __shared__ int wsum[SCAN_BLOCK / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int inc = wave_inclusive_scan(v);
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    int before = 0, all = 0;
#pragma unroll
    for (int i = 0; i < SCAN_BLOCK / 64; i++) {
        int x = wsum[i];
        before += (i < w) ? x : 0;
        all += x;
    }
    __syncthreads();
    *total = all;
    return before + inc - v;
}

__global__ void k_scan_reduce(const int *__restrict__ in, size_t n,
                              int *__restrict__ block_sums) {
    size_t base = (size_t)blockIdx.x * SCAN_CHUNK;
    int acc = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++) {
        size_t idx = base + (size_t)i * SCAN_BLOCK + threadIdx.x;
        if (idx < n) acc += in[idx];
    }
    int total;
    block_exclusive_scan(acc, &total);
    if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

// One block scans the per-block sums in place (exclusive) and writes the
// grand total.
__global__ void k_scan_sums(int *sums, int nblocks, int *total_out) {
    int carry = 0;
    for (int base = 0; base < nblocks; base += SCAN_BLOCK) {
        int i = base + threadIdx.x;
        int v = i < nblocks ? sums[i] : 0;
        int tot;
        int ex = block_exclusive_scan(v, &tot);
        if (i < nblocks) sums[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0 && total_out) *total_out = carry;
}

// Each thread owns SCAN_ITEMS consecutive elements of the block's chunk.
__global__ void k_scan_down(const int *in, int *out, size_t n,
                            const int *__restrict__ block_sums) {
    size_t base = (size_t)blockIdx.x * SCAN_CHUNK + (size_t)threadIdx.x * SCAN_ITEMS;
    int v[SCAN_ITEMS];
    int acc = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++) {
        size_t idx = base + i;
        v[i] = idx < n ? in[idx] : 0;
        acc += v[i];
    }
    int tot;
    int run = block_exclusive_scan(acc, &tot) + block_sums[blockIdx.x];
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++) {
        size_t idx = base + i;
        if (idx < n) out[idx] = run;
        run += v[i];
    }
}

size_t scan_temp_bytes(size_t n) {
    size_t nb = (n + SCAN_CHUNK - 1) / SCAN_CHUNK;
    return (nb + 1) * sizeof(int) + 256;
}

int exclusive_scan(const int *in, int *out, size_t n, int *total, void *tmp,
                   hipStream_t s) {
    if (n == 0) {
        if (total) SWIFT_TRY(fill_i32(total, 0, 1, s));
        return SWIFT_OK;
    }
    int nb = (int)((n + SCAN_CHUNK - 1) / SCAN_CHUNK);
    int *sums = static_cast<int *>(tmp);
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(SCAN_BLOCK), 0, s, in, n, sums);
    SWIFT_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SCAN_BLOCK), 0, s, sums, nb, total);
    SWIFT_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(SCAN_BLOCK), 0, s, in, out, n, sums);
    SWIFT_CHECK_LAUNCH();
    return SWIFT_OK;
}

// ------------------------------------------------------------- radix sort --
constexpr int RS_BLOCK = 256;
constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = RS_BLOCK * RS_ITEMS;  // 4096 keys per block
constexpr int RS_BINS = 256;

// hist[d * ntiles + tile] = count of digit d in the tile
__global__ __launch_bounds__(RS_BLOCK) void k_radix_hist(
    const unsigned *__restrict__ keys, size_t n, int shift, int ntiles,
    int *__restrict__ hist) {
    __shared__ int cnt[RS_BINS];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    size_t base = (size_t)blockIdx.x * RS_TILE;
#pragma unroll 4
    for (int i = 0; i < RS_ITEMS; i++) {
        size_t idx = base + (size_t)i * RS_BLOCK + threadIdx.x;
        if (idx < n) atomicAdd(&cnt[(keys[idx] >> shift) & 0xFF], 1);
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * ntiles + blockIdx.x] = cnt[threadIdx.x];
}

// Lanes of the wave whose 8-bit digit equals this lane's.
__device__ __forceinline__ uint64_t match_digit(unsigned d) {
    uint64_t m = ~0ull;
#pragma unroll
    for (int b = 0; b < 8; b++) {
        uint64_t bal = __ballot((d >> b) & 1u);
        m &= ((d >> b) & 1u) ? bal : ~bal;
    }
    return m;
}

// Stable scatter: items are visited in index order (round-major, then
// thread), a running per-digit base in LDS carries order across rounds, and
// per-wave digit counts order the 4 waves of a round.
__global__ __launch_bounds__(RS_BLOCK) void k_radix_scatter(
    const unsigned *__restrict__ keys_in, const unsigned *__restrict__ vals_in,
    unsigned *__restrict__ keys_out, unsigned *__restrict__ vals_out, size_t n,
    int shift, int ntiles, const int *__restrict__ offs) {
    __shared__ int base_d[RS_BINS];
    __shared__ int wcnt[RS_BLOCK / 64][RS_BINS];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    base_d[t] = offs[(size_t)t * ntiles + blockIdx.x];
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    size_t tile = (size_t)blockIdx.x * RS_TILE;
    for (int r = 0; r < RS_ITEMS; r++) {
#pragma unroll
        for (int q = 0; q < RS_BLOCK / 64; q++) wcnt[q][t] = 0;
        __syncthreads();
        size_t idx = tile + (size_t)r * RS_BLOCK + t;
        bool valid = idx < n;
        unsigned key = valid ? keys_in[idx] : 0u;
        unsigned val = valid ? (vals_in ? vals_in[idx] : (unsigned)idx) : 0u;
        // lanes past n are excluded from every peer mask
        unsigned d = (key >> shift) & 0xFFu;
        uint64_t valid_mask = __ballot(valid);
        uint64_t peers = match_digit(d) & valid_mask;
        int rank = __popcll(peers & lt_mask);
        if (valid && rank == 0) wcnt[w][d] = __popcll(peers);
        __syncthreads();
        {   // prefix over waves for digit t, and advance the running base
            int run = base_d[t];
#pragma unroll
            for (int q = 0; q < RS_BLOCK / 64; q++) {
                int c = wcnt[q][t];
                wcnt[q][t] = run;
                run += c;
            }
            base_d[t] = run;
        }
        __syncthreads();
        if (valid) {
            int pos = wcnt[w][d] + rank;
            keys_out[pos] = key;
            vals_out[pos] = val;
        }
        __syncthreads();
    }
}

size_t radix_temp_bytes(size_t n) {
    size_t ntiles = (n + RS_TILE - 1) / RS_TILE;
    size_t h = ntiles * RS_BINS;
    return h * sizeof(int) + scan_temp_bytes(h) + 256;
}

int radix_sort_pairs(const unsigned *keys_in, const unsigned *vals_in,
                     unsigned *keys_out, unsigned *vals_out, unsigned *keys_alt,
                     unsigned *vals_alt, size_t n, int bits, void *tmp,
                     hipStream_t s) {
    if (n == 0) return SWIFT_OK;
    if (n > 0x7FFFFFFFull) {
        set_error("radix_sort_pairs: n=%zu exceeds int32 positions", n);
        return SWIFT_ERR_ARG;
    }
    if (bits < 1) bits = 1;
    int passes = (bits + 7) / 8;
    int ntiles = (int)((n + RS_TILE - 1) / RS_TILE);
    size_t h = (size_t)ntiles * RS_BINS;
    int *hist = static_cast<int *>(tmp);
    void *scan_tmp = hist + h;
    // Ping-pong so that the last pass lands in *_out.
    const unsigned *ki = keys_in, *vi = vals_in;
    for (int p = 0; p < passes; p++) {
        bool last = (p == passes - 1);
        bool to_out = ((passes - 1 - p) % 2) == 0;
        unsigned *ko = to_out ? keys_out : keys_alt;
        unsigned *vo = to_out ? vals_out : vals_alt;
        (void)last;
        hipLaunchKernelGGL(k_radix_hist, dim3(ntiles), dim3(RS_BLOCK), 0, s, ki, n,
                           8 * p, ntiles, hist);
        SWIFT_CHECK_LAUNCH();
        SWIFT_TRY(exclusive_scan(hist, hist, h, nullptr, scan_tmp, s));
        hipLaunchKernelGGL(k_radix_scatter, dim3(ntiles), dim3(RS_BLOCK), 0, s, ki,
                           vi, ko, vo, n, 8 * p, ntiles, hist);
        SWIFT_CHECK_LAUNCH();
        ki = ko;
        vi = vo;
    }
    return SWIFT_OK;
}

}  // namespace swift

extern "C" const char *swift_last_error(void) { return swift::last_error(); }
extern "C" const char *swift_version(void) { return "swift-mi355x 0.1 (gfx950)"; }
