// spmm.hip -- Swift SpMM kernels for gfx950 (CDNA4, 64-lane waves).
//
// SWIFT_KERNEL_DEFAULT, row-major B/C: the merge-path walk of the row view
// (k_rowview_row over rowview.h's rowview_walk).  Every wave owns a segment
// of W items (rows + entries), lanes run along K (LPE lanes x VEC elements
// per entry, EPS = 64/LPE entries side by side), B rows are gathered with
// buffer loads and each finished row is stored once.  A row cut by segment
// boundaries is stored in pieces: the piece holding its start into C, the
// others into carry slots, which one tail launch adds onto C in boundary
// order (k_row_tail) while its other workgroups run the reside blocks with
// atomics onto the finished rows.  The autotuner's schedule
// (segment size, slot rows per step, lane width, window walk) applies per K.
// Column-major B/C: B is transposed to row-major, the same walk runs, and C
// is transposed back.
//
// The reference's own orders, kept for variant parity (BALANCED, UNBALANCED,
// KCHUNK, ROWSORT): the regular part in 32-wide staircase tiles and the
// irregular part as reside CSC balance blocks.  Row-major: lanes along K, the
// wave walks one work item of up to 16 depth tiles that share 32 sorted
// columns and flushes a row's sum when the row changes.  Column-major (the
// reference's Swift_GPU layout): lanes along the 32 entries of a tile (two
// tiles per wave), each k reads 32 consecutive sorted columns of B, and
// equal-row runs are summed with a 5-step segmented shuffle reduction before
// one atomic per run head (format 0/2 tiles; format 1 tiles skip it).
#include "rowview.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <vector>

namespace swift {

// ---------------------------------------------------- row-major kernels --

// Tile order, one wave per work item (up to 16 depth tiles of one column
// block, whose B rows stay hot in L1/L2).  Register accumulators per group
// and C atomics only when the row changes.
template <class TV, class TB, class TC, int VEC, int Q, int LPE, bool MASK>
__global__ __launch_bounds__(256) void k_tiles_row(RowArgs a) {
    constexpr int EPS = 64 / LPE;       // entries in flight per wave
    constexpr int STEPS = TILE / EPS;   // steps per tile
    const int lane = threadIdx.x & 63;
    const int h = lane / LPE, sub = lane % LPE;
    const int w = (int)((blockIdx.x * (size_t)blockDim.x + threadIdx.x) >> 6);
    if (w >= a.nwork) return;
    const int2 wk = a.work[w];
    const int c0 = a.tile_colidx[wk.x];
    const int my = lane & 31;
    const int mycol = a.perm ? a.perm[c0 + my] : c0 + my;
    const TV *sval = static_cast<const TV *>(a.sval);
    TC acc[Q][VEC];
#pragma unroll
    for (int q = 0; q < Q; q++)
#pragma unroll
        for (int v = 0; v < VEC; v++) acc[q][v] = (TC)0;
    int cur = -1;
    for (int t = wk.x; t < wk.x + wk.y; t++) {
        const size_t e0 = (size_t)t * TILE;
        const int myrow = a.srow[e0 + my];
        const TC myval = to_acc(sval[e0 + my]);
#pragma unroll
        for (int i = 0; i < STEPS; i++) {
            const int l = h + i * EPS;
            const int r = __shfl(myrow, l, 64);
            const TC v = __shfl(myval, l, 64);
            const int bc = __shfl(mycol, l, 64);
            if (r != cur) {
                if (cur >= 0) flush_strided<TC, VEC, Q, LPE, MASK>(a, cur, sub, acc);
                cur = r;
            }
            TC b[Q][VEC];
            load_strided<TB, TC, VEC, Q, LPE, MASK>(a, bc, sub, b);
#pragma unroll
            for (int q = 0; q < Q; q++)
#pragma unroll
                for (int x = 0; x < VEC; x++) acc[q][x] += v * b[q][x];
        }
    }
    if (cur >= 0) flush_strided<TC, VEC, Q, LPE, MASK>(a, cur, sub, acc);
}

// Waves per SIMD the walk may reach.  The FP32 walk of 32-lane groups (K =
// 128) fits 64 VGPRs, i.e. 8 waves per SIMD, and runs faster at 7: on the
// 10^6-row matrix 0.829-0.831 ms at 8 against 0.821-0.823 ms at 7 (the
// 200k-node graph: equal).  Other shapes keep the limit of their register
// count.
template <class TB, int Q, int LPE>
constexpr int walk_max_waves() {
    return (sizeof(TB) == 4 && Q == 1 && LPE == 32) ? 7 : 8;
}

// Thin kernels over the device functions of rowview.h.  The walk stores the
// pieces that its segments start with into carry slots (k_row_tail adds them).
template <class TV, class TB, class TC, int VEC, int Q, int LPE, bool MASK, bool BUF,
          int SLOTS = 0>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, walk_max_waves<TB, Q, LPE>())))
void k_rowview_row(RowArgs a) {
    const unsigned lb = xcd_block(blockIdx.x, gridDim.x);
    // wave-uniform segment id, kept in a scalar register
    const int sid = __builtin_amdgcn_readfirstlane((int)((lb * (size_t)blockDim.x + threadIdx.x) >> 6));
    if (sid >= a.nseg) return;
    rowview_walk<TV, TB, TC, VEC, Q, LPE, MASK, BUF, false, SLOTS>(a, sid);
}

// The row-view walk with a.slots slot rows per step where the autotuner
// chose a count other than the rule that the shape can take (buffer loads
// only), else with the rule.
template <class TV, class TB, class TC, int VEC, int Q, int LPE, bool MASK, bool BUF>
static void launch_view(const RowArgs &a, dim3 grid, hipStream_t s) {
    if constexpr (BUF && !MASK) {
        constexpr int RULE = rule_slots(LPE, Q, (int)sizeof(TB));
#define SWIFT_SLOTS_CASE(N_)                                                              \
        if constexpr (slots_ok(LPE, Q, N_) && RULE != N_) {                              \
            if (a.slots == N_) {                                                         \
                hipLaunchKernelGGL((k_rowview_row<TV, TB, TC, VEC, Q, LPE, MASK, BUF, N_>), \
                                   grid, dim3(256), 0, s, a);                            \
                return;                                                                  \
            }                                                                            \
        }
        SWIFT_SLOTS_CASE(4)
        SWIFT_SLOTS_CASE(8)
        SWIFT_SLOTS_CASE(16)
#undef SWIFT_SLOTS_CASE
    }
    hipLaunchKernelGGL((k_rowview_row<TV, TB, TC, VEC, Q, LPE, MASK, BUF>), grid, dim3(256), 0, s,
                       a);
}

// Window walk.  One workgroup of WIN_WAVES waves takes WIN_WAVES consecutive
// segments (one per wave) and first picks the B rows its entries reuse most:
// WIN_CAND candidate windows of wcap rows (LDS bytes / B row bytes), centred
// on the columns of evenly spaced entries, are scored by how many of the
// group's entries they hold; the best is shrunk to the span of the columns it
// holds and, if every staged row is used WIN_REUSE times on average, copied
// into LDS once.  The waves then walk their segments as k_rowview_row does,
// reading a step's B rows from LDS when all of them are staged.  Rows of
// banded or locally clustered matrices thus cost one L2 read per workgroup
// instead of one per entry; scattered rows keep the plain walk's loads.
// Two workgroups (32 waves, the plain walk's occupancy) share a CU's 160 KB.
#ifndef SWIFT_WIN_WAVES
#define SWIFT_WIN_WAVES 16
#endif
#ifndef SWIFT_WIN_LDS_KB
#define SWIFT_WIN_LDS_KB 76
#endif
#ifndef SWIFT_WIN_MIN_ENTRIES
#define SWIFT_WIN_MIN_ENTRIES 256
#endif
constexpr int WIN_WAVES = SWIFT_WIN_WAVES;
constexpr int WIN_THREADS = WIN_WAVES * 64;
constexpr unsigned WIN_LDS_BYTES = SWIFT_WIN_LDS_KB * 1024u;
constexpr int WIN_CAND = 32;
constexpr int WIN_MIN_ENTRIES = SWIFT_WIN_MIN_ENTRIES;
constexpr int WIN_REUSE = 2;
constexpr int WIN_MIN_ROWS = 16;

template <class TV, class TB, class TC, int VEC, int Q, int LPE>
__global__ __launch_bounds__(WIN_THREADS, 8) void k_rowview_win(RowArgs a) {
    __shared__ __align__(16) unsigned char wbuf[WIN_LDS_BYTES];
    __shared__ int cand[WIN_CAND], cnt[WIN_CAND], sel[4];
    const unsigned g = xcd_block(blockIdx.x, gridDim.x);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int s0 = (int)g * WIN_WAVES;
    const int s1 = min(s0 + WIN_WAVES, a.nseg);
    const int E0 = a.seg[s0].y, E1 = a.seg[s1].y;
    const int n = E1 - E0;
    const unsigned rb = (unsigned)(a.kcount * (int)sizeof(TB));
    const int wcap = (int)(WIN_LDS_BYTES / rb);
    Window win{wbuf, 0, 0, rb};
    // every branch below that holds a barrier is uniform over the workgroup
    if (n >= WIN_MIN_ENTRIES) {
        auto col = [&](int e) {
            const int c = a.rv_col[e];
            return a.inv_perm ? a.inv_perm[c] : c;
        };
        if (tid < WIN_CAND) {
            const int c = col(E0 + (int)(((long long)tid * n) / WIN_CAND));
            cand[tid] = max(0, c - wcap / 2);
            cnt[tid] = 0;
        }
        if (tid == 0) {
            sel[2] = 0x7FFFFFFF;
            sel[3] = -1;
        }
        __syncthreads();
        const int iters = (n + WIN_THREADS - 1) / WIN_THREADS;
        int mine = 0;  // lane i: entries of this wave inside candidate i
        for (int it = 0; it < iters; it++) {
            const int e = E0 + it * WIN_THREADS + tid;
            const bool live = e < E1;
            const int c = live ? col(e) : 0;
#pragma unroll 8
            for (int i = 0; i < WIN_CAND; i++) {
                const bool in = live && (unsigned)(c - cand[i]) < (unsigned)wcap;
                const int pop = __popcll(__ballot(in));
                mine += lane == i ? pop : 0;
            }
        }
        if (lane < WIN_CAND && mine) atomicAdd(&cnt[lane], mine);
        __syncthreads();
        if (wave == 0) {
            int bc = lane < WIN_CAND ? cnt[lane] : -1, bi = lane;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                const int oc = __shfl_xor(bc, o, 64), oi = __shfl_xor(bi, o, 64);
                if (oc > bc || (oc == bc && oi < bi)) {
                    bc = oc;
                    bi = oi;
                }
            }
            if (lane == 0) {
                sel[0] = cand[bi];
                sel[1] = bc;
            }
        }
        __syncthreads();
        const int ws = __builtin_amdgcn_readfirstlane(sel[0]);
        int lo = 0x7FFFFFFF, hi = -1;
        for (int it = 0; it < iters; it++) {
            const int e = E0 + it * WIN_THREADS + tid;
            if (e < E1) {
                const int c = col(e);
                if ((unsigned)(c - ws) < (unsigned)wcap) {
                    lo = min(lo, c);
                    hi = max(hi, c);
                }
            }
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            lo = min(lo, __shfl_xor(lo, o, 64));
            hi = max(hi, __shfl_xor(hi, o, 64));
        }
        if (lane == 0 && hi >= 0) {
            atomicMin(&sel[2], lo);
            atomicMax(&sel[3], hi);
        }
        __syncthreads();
        const int w0 = __builtin_amdgcn_readfirstlane(sel[2]);
        const int wh = __builtin_amdgcn_readfirstlane(sel[3]);
        const int wn = __builtin_amdgcn_readfirstlane(sel[1]);
        const int wl = wh - w0 + 1;
        if (wh >= 0 && wn >= WIN_REUSE * wl) {
            // B rows [w0, w0 + wl) (wl <= wcap, so wl * rb <= WIN_LDS_BYTES)
            // in 16-byte pieces through the walk's buffer resource
            const __amdgpu_buffer_rsrc_t rsrc =
                make_rsrc(static_cast<const TB *>(a.B) + a.kbase, a.bbytes);
            const int cpr = (int)(rb >> 4);
            const int chunks = wl * cpr;
            for (int ch = tid; ch < chunks; ch += WIN_THREADS) {
                const int r = ch / cpr, p = ch - r * cpr;
                const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(
                    rsrc, (unsigned)(w0 + r) * a.bstride + (unsigned)(p << 4), 0, 0);
                *reinterpret_cast<u32x4 *>(wbuf + ((unsigned)ch << 4)) = v;
            }
            win.w0 = w0;
            win.wlen = wl;
            __syncthreads();
        }
    }
    // wave-uniform segment id, kept in a scalar register
    const int sid = __builtin_amdgcn_readfirstlane(s0 + wave);
    if (sid < a.nseg) rowview_walk<TV, TB, TC, VEC, Q, LPE, false, true, true>(a, sid, win);
}

template <class TV, class TB, class TC, int VEC, int Q, int LPE, bool MASK>
__global__ __launch_bounds__(256) void k_reside_row(RowArgs a) {
    const int blk = (int)((blockIdx.x * (size_t)blockDim.x + threadIdx.x) >> 6);
    if (blk >= a.nblk) return;
    reside_block<TV, TB, TC, VEC, Q, LPE, MASK>(a, blk);
}

// After the row-view walk, adds the a.nfix cut rows' carry slots onto C
// (split_fixup, all K columns): the a.nlong rows with the most slots by one
// workgroup each, then the others by one wave each.  The workgroups after
// them run the reside blocks of this K slice, one wave each, with atomics
// onto the finished rows (a.nblk = 0: none).
template <class TV, class TB, class TC, int VEC, int Q, int LPE, bool MASK>
__global__ __launch_bounds__(256) void k_row_tail(RowArgs a) {
    const unsigned long_blocks = (unsigned)a.nlong;
    const unsigned fix_blocks = long_blocks + (unsigned)(a.nfix - a.nlong + 3) / 4u;
    if (blockIdx.x < long_blocks) {
        split_fixup<TC, 256>(a, (int)blockIdx.x);
        return;
    }
    if (blockIdx.x < fix_blocks) {
        const int w = (int)(((blockIdx.x - long_blocks) * (size_t)blockDim.x + threadIdx.x) >> 6);
        if (a.nlong + w < a.nfix) split_fixup<TC, 64>(a, a.nlong + w);
        return;
    }
    const int w = (int)(((blockIdx.x - fix_blocks) * (size_t)blockDim.x + threadIdx.x) >> 6);
    if (w < a.nblk) reside_block<TV, TB, TC, VEC, Q, LPE, MASK>(a, w);
}

#ifndef SWIFT_FLAT_U
#define SWIFT_FLAT_U 4
#endif

template <class TV, class TB, class TC, int VEC, int Q, int LPE, bool MASK, int U>
__global__ __launch_bounds__(256) void k_flat_row(RowArgs a) {
    const unsigned lb = xcd_block(blockIdx.x, gridDim.x);
    const int t = (int)(lb * (size_t)blockDim.x + threadIdx.x);
    if ((t & ~63) / LPE >= a.nseg) return;  // every walker of this wave is past the end
    flat_walk<TV, TB, TC, VEC, Q, LPE, MASK, U>(a, t / LPE);
}

// Zeroes C[row, 0 : K) of each listed row; one wave per row.
template <class TC>
__global__ __launch_bounds__(256) void k_zero_rows(const int *__restrict__ rows, int nrows,
                                                   TC *__restrict__ C, int ldc, int K) {
    const int i = (int)((blockIdx.x * (size_t)blockDim.x + threadIdx.x) >> 6);
    if (i >= nrows) return;
    const int r = rows[i];
    TC *c = C + (size_t)r * ldc;
    for (int k = threadIdx.x & 63; k < K; k += 64) c[k] = (TC)0;
}

// ---------------------------------------------------- col-major kernels --

struct ColArgs {
    int tilenum;
    const int *tile_colidx, *tile_format;
    const int *srow;
    const void *sval;
    int nblk;
    const unsigned *colorblk;
    const int *blkstart, *blkstop, *rptr, *rrow;
    const void *rval;
    int m, n;
    const int *perm;
    const void *B;              // K x n, B[k*ldb + col]
    void *C;                    // K x m, C[k*ldc + row]
    int ldb, ldc;
    int kbase, kchunk, K;       // k range of blockIdx.y
    const float *kscale;
    float gscale;
    // row-sorted groups (Swift_GPU_4)
    const int *val_position, *rowsorted;
};

template <class TC>
__device__ __forceinline__ TC kscale_of(const ColArgs &a, int k) {
    return a.kscale ? (TC)(a.kscale[k] * a.gscale) : (TC)a.gscale;
}

// Two tiles per wave; blockIdx.y selects a K chunk.
template <class TV, class TB, class TC>
__global__ __launch_bounds__(256) void k_regular_col(ColArgs a) {
    const int lane = threadIdx.x & 63;
    const int l = lane & 31;
    const int t = (int)((blockIdx.x * (size_t)blockDim.x + threadIdx.x) >> 5);
    const bool valid = t < a.tilenum;
    const size_t e = (size_t)(valid ? t : 0) * TILE + l;
    const int row = valid ? a.srow[e] : 0;
    const TC val = valid ? to_acc(static_cast<const TV *>(a.sval)[e]) : (TC)0;
    const int c = valid ? a.tile_colidx[t] + l : 0;
    const int bcol = a.perm ? a.perm[c] : c;
    const int fmt = valid ? a.tile_format[t] : 1;
    // equal-row runs inside the half: heads and each lane's run end
    const int prev = __shfl_up(row, 1, 64);
    const bool head = (l == 0) || prev != row;
    const unsigned heads = (unsigned)(__ballot(head) >> (lane & 32));
    const unsigned after = (l == 31) ? 0u : (heads >> (l + 1));
    const int last = after ? l + __builtin_ctz(after) : 31;  // last lane of my run
    const bool reduce = fmt != 1;
    const int k0 = a.kbase + blockIdx.y * a.kchunk;
    const int k1 = min(a.K, k0 + a.kchunk);
    const TB *B = static_cast<const TB *>(a.B);
    TC *C = static_cast<TC *>(a.C);
    for (int k = k0; k < k1; k++) {
        TC p = val * to_acc(B[(size_t)k * a.ldb + bcol]) * kscale_of<TC>(a, k);
        if (reduce) {  // uniform per half-wave
#pragma unroll
            for (int off = 1; off < 32; off <<= 1) {
                TC q = __shfl_down(p, off, 32);
                if (l + off <= last) p += q;
            }
            if (valid && head) atomic_add(C + (size_t)k * a.ldc + row, p);
        } else if (valid) {
            atomic_add(C + (size_t)k * a.ldc + row, p);
        }
    }
}

// Balance block per half-wave; lanes along entries (<= 32 when balanced,
// a whole reside column when unbalanced).
template <class TV, class TB, class TC>
__global__ __launch_bounds__(256) void k_reside_col(ColArgs a) {
    const int lane = threadIdx.x & 63;
    const int l = lane & 31;
    const int blk = (int)((blockIdx.x * (size_t)blockDim.x + threadIdx.x) >> 5);
    if (blk >= a.nblk) return;
    int r, start, stop;
    if (a.colorblk) {
        const unsigned u = a.colorblk[blk];
        r = (int)(u & 0x7FFFFFFFu);
        const bool split = (u >> 31) != 0;
        start = split ? a.blkstart[blk] : a.rptr[r];
        stop = split ? a.blkstop[blk] : a.rptr[r + 1];
    } else {
        r = blk;
        start = a.rptr[r];
        stop = a.rptr[r + 1];
    }
    const int jj = a.n - 1 - r;
    const int bcol = a.perm ? a.perm[jj] : jj;
    const int k0 = a.kbase + blockIdx.y * a.kchunk;
    const int k1 = min(a.K, k0 + a.kchunk);
    const TB *B = static_cast<const TB *>(a.B);
    TC *C = static_cast<TC *>(a.C);
    const TV *rval = static_cast<const TV *>(a.rval);
    for (int e = start + l; e < stop; e += 32) {
        const int row = a.rrow[e];
        const TC val = to_acc(rval[e]);
        for (int k = k0; k < k1; k++)
            atomic_add(C + (size_t)k * a.ldc + row,
                       val * to_acc(B[(size_t)k * a.ldb + bcol]) * kscale_of<TC>(a, k));
    }
}

// Row-sorted groups of 8 tiles (256 entries, one per thread).  Products go to
// their sorted slot in LDS; each run head of equal rows sums its run and
// issues one atomic per (run, k).  Run bounds come from the sorted keys.
template <class TV, class TB, class TC>
__global__ __launch_bounds__(256) void k_regular_col_rowsort(ColArgs a) {
    __shared__ TC prod[256];
    const int g = blockIdx.x;
    const int t0 = g * 8;
    const int nt = min(8, a.tilenum - t0);
    const int len = nt * TILE;
    const int i = threadIdx.x;
    const size_t base = (size_t)t0 * TILE;
    const bool valid = i < len;
    TC val = (TC)0;
    int bcol = 0, pos = 0, srow = 0, runlen = 0;
    if (valid) {
        const int t = t0 + (i >> 5);
        val = to_acc(static_cast<const TV *>(a.sval)[base + i]);
        const int c = a.tile_colidx[t] + (i & 31);
        bcol = a.perm ? a.perm[c] : c;
        pos = a.val_position[base + i];
        srow = a.rowsorted[base + i];
        const bool head = i == 0 || a.rowsorted[base + i - 1] != srow;
        if (head) {
            int e = i + 1;
            while (e < len && a.rowsorted[base + e] == srow) e++;
            runlen = e - i;
        }
    }
    const int k0 = a.kbase + blockIdx.y * a.kchunk;
    const int k1 = min(a.K, k0 + a.kchunk);
    const TB *B = static_cast<const TB *>(a.B);
    TC *C = static_cast<TC *>(a.C);
    for (int k = k0; k < k1; k++) {
        if (valid) prod[pos] = val * to_acc(B[(size_t)k * a.ldb + bcol]);
        __syncthreads();
        if (runlen > 0) {
            TC s = (TC)0;
            for (int j = 0; j < runlen; j++) s += prod[i + j];
            atomic_add(C + (size_t)k * a.ldc + srow, s * kscale_of<TC>(a, k));
        }
        __syncthreads();
    }
}

// --------------------------------------------------------------- launch --

// Optional events between the phases of one call (zero | regular | reside).
struct PhaseEvents {
    hipEvent_t ev[4];
};
static thread_local PhaseEvents *g_phase = nullptr;

static inline void mark(int i, hipStream_t s) {
    if (g_phase) (void)hipEventRecord(g_phase->ev[i], s);
}


// Caller-provided workspace of the calling thread's current product (set by
// spmm_noalloc_impl around spmm_impl); nullptr: the matrix's own, grown on
// demand.  g_ws_off: bytes of it already taken by an enclosing call (the
// column-major transposes).
struct Workspace {
    void *ptr;
    size_t bytes;
};
static thread_local const Workspace *g_ws = nullptr;
static thread_local size_t g_ws_off = 0;

// Carry slots for a row-view product (or backward walk): from the caller's
// workspace, else the matrix's workspace for this stream (grown on demand).
int carry_workspace(const swift_matrix *M, size_t bytes, hipStream_t s, void **out) {
    if (g_ws) {
        const size_t off = (g_ws_off + 255) & ~(size_t)255;
        if (off + bytes > g_ws->bytes) {
            set_error("swift_spmm_noalloc: the product needs %zu bytes of workspace (%zu of "
                      "them for the carry slots of its split rows), more than the %zu of the "
                      "scratch array", off + bytes, bytes, g_ws->bytes);
            return SWIFT_ERR_ARG;
        }
        *out = static_cast<char *>(g_ws->ptr) + off;
        return SWIFT_OK;
    }
    std::lock_guard<std::mutex> g(M->carry_mu);
    auto &slot = M->carry_ws[s];
    if (!slot) slot.reset(new Buffer());
    SWIFT_TRY(slot->alloc(bytes));
    *out = slot->ptr;
    return SWIFT_OK;
}

enum RowKind { ROW_TILES = 0, ROW_RESIDE = 1, ROW_VIEW = 2, ROW_FLAT = 3, ROW_WIN = 4, ROW_TAIL = 5 };

template <class TV, class TB, class TC, int VEC, int Q, int LPE, bool MASK>
static void launch_row_shape(const RowArgs &a, RowKind kind, hipStream_t s) {
    if (kind == ROW_WIN) {
        // the window walk needs buffer loads, whole 8/16-byte lane loads, one
        // chunk per lane (K slices up to 64 lanes x one load: wider rows
        // leave too few rows in the window) and room for WIN_MIN_ROWS B rows;
        // otherwise the plain walk runs
        if constexpr (!MASK && Q == 1 && (sizeof(TB) * VEC == 16 || sizeof(TB) * VEC == 8)) {
            const unsigned rb = (unsigned)(a.kcount * (int)sizeof(TB));
            if (a.nseg && a.bstride && WIN_LDS_BYTES / rb >= (unsigned)WIN_MIN_ROWS) {
                hipLaunchKernelGGL((k_rowview_win<TV, TB, TC, VEC, Q, LPE>),
                                   dim3(ceil_div((size_t)a.nseg, WIN_WAVES)), dim3(WIN_THREADS),
                                   0, s, a);
                return;
            }
        }
        kind = ROW_VIEW;
    }
    if (kind == ROW_FLAT) {
        constexpr int U = SWIFT_FLAT_U < LPE ? SWIFT_FLAT_U : LPE;
        if (a.nseg)
            hipLaunchKernelGGL((k_flat_row<TV, TB, TC, VEC, Q, LPE, MASK, U>),
                               dim3(ceil_div((size_t)a.nseg * LPE, 256)), dim3(256), 0, s, a);
        return;
    }
    if (kind == ROW_VIEW) {
        if (!a.nseg) return;
        const dim3 grid(ceil_div((size_t)a.nseg * 64, 256));
        if (a.bstride)
            launch_view<TV, TB, TC, VEC, Q, LPE, MASK, true>(a, grid, s);
        else
            launch_view<TV, TB, TC, VEC, Q, LPE, MASK, false>(a, grid, s);
        return;
    }
    if (kind == ROW_TAIL) {
        const size_t blocks = (size_t)a.nlong + ceil_div((size_t)(a.nfix - a.nlong) * 64, 256) +
                              ceil_div((size_t)a.nblk * 64, 256);
        if (blocks)
            hipLaunchKernelGGL((k_row_tail<TV, TB, TC, VEC, Q, LPE, MASK>), dim3((unsigned)blocks),
                               dim3(256), 0, s, a);
        return;
    }
    const unsigned items = kind == ROW_TILES ? (unsigned)a.nwork : (unsigned)a.nblk;
    if (!items) return;
    dim3 grid(ceil_div((size_t)items * 64, 256)), block(256);
    if (kind == ROW_TILES)
        hipLaunchKernelGGL((k_tiles_row<TV, TB, TC, VEC, Q, LPE, MASK>), grid, block, 0, s, a);
    else
        hipLaunchKernelGGL((k_reside_row<TV, TB, TC, VEC, Q, LPE, MASK>), grid, block, 0, s, a);
}

template <class TV, class TB, class TC>
static int launch_row(const RowArgs &a, const RowTarget &t, RowKind kind, hipStream_t s) {
    constexpr int V = Vec16<TB>::N;
    if (t.mask) {
        launch_row_shape<TV, TB, TC, 1, 1, 64, true>(a, kind, s);
        SWIFT_CHECK_LAUNCH();
        return SWIFT_OK;
    }
#define SWIFT_ROW_CASE(Q_, L_)                                                \
    if (t.q == Q_ && t.lpe == L_) {                                           \
        launch_row_shape<TV, TB, TC, V, Q_, L_, false>(a, kind, s);           \
        SWIFT_CHECK_LAUNCH();                                                 \
        return SWIFT_OK;                                                      \
    }
    SWIFT_ROW_CASE(4, 64)
    SWIFT_ROW_CASE(2, 64)
    SWIFT_ROW_CASE(1, 64)
    SWIFT_ROW_CASE(1, 32)
    SWIFT_ROW_CASE(1, 16)
    SWIFT_ROW_CASE(1, 8)
    if constexpr (V == 8) { SWIFT_ROW_CASE(1, 4) }
#undef SWIFT_ROW_CASE
    set_error("no row kernel for vec=%d q=%d lpe=%d", t.vec, t.q, t.lpe);
    return SWIFT_ERR_UNSUPPORTED;
}

// Row-major C = A*B.  DEFAULT / ROWVIEW: row view (regular part, output
// stationary) then the balanced reside blocks.  BALANCED / UNBALANCED: the
// reference's tile order (zero C, tile kernel, reside kernel).
template <class TV, class TB, class TC>
static int spmm_row(const swift_matrix *M, int K, swift_kernel kernel, unsigned flags,
                    const void *B, void *C, const void *sval, const void *rval,
                    const void *rv_val, const float *kscale, float gscale, hipStream_t s) {
    RowArgs a{};
    a.nwork = M->nwork;
    a.work = M->work;
    a.tile_colidx = M->tile_colidx;
    a.srow = M->sortrowidx;
    a.sval = sval;
    const bool unbalanced = kernel == SWIFT_KERNEL_UNBALANCED;
    a.nblk = unbalanced ? M->reside_col : M->numblk;
    a.colorblk = unbalanced ? nullptr : M->colorblk;
    a.blkstart = M->blkstart;
    a.blkstop = M->blkstop;
    a.rptr = M->reside_cscptr;
    a.rrow = M->reside_cscrowidx;
    a.rval = rval;
    a.m = M->m;
    constexpr int V = Vec16<TB>::N;
    const bool aligned = (K % V) == 0 && ((uintptr_t)B % 16) == 0 && ((uintptr_t)C % 16) == 0;
    // the autotuner's choice for this K and element size (DEFAULT only)
    const bool tuned = kernel == SWIFT_KERNEL_DEFAULT && M->tuned_K == K &&
                       M->tuned_elem == (int)sizeof(TB);
    // the default product's slice width: the autotuner's cap, else the rule
    const int cap = tuned && M->tuned_cols ? M->tuned_cols
                    : kernel == SWIFT_KERNEL_DEFAULT ? default_slice_cols(K, sizeof(TB))
                                                     : 0;
    const auto slices = plan_row_slices(K, V, aligned, cap);
    // the default table of the bytes one entry gathers in the first slice,
    // unless the size was set
    const RowTarget &t0 = slices.front().second;
    const int cols0 = t0.mask ? std::min(64, K) : t0.vec * t0.lpe * t0.q;
    const swift_matrix::SegTable &tb = M->cls[seg_class((size_t)cols0 * sizeof(TB), sizeof(TB))];
    const swift_matrix::SegTable &st = !M->user_items && tb.seg ? tb : M->prim;
    a.nseg = st.nseg;
    a.seg = st.seg;
    a.rv_ptr = M->rv_ptr;
    a.rv_col = M->rv_col;
    a.rv_val = rv_val;
    a.rv_row = M->rv_row;
    const bool sorted = (flags & SWIFT_B_SORTED) != 0;
    a.inv_perm = sorted ? M->inv_perm : nullptr;
    a.accumulate = (flags & SWIFT_ACCUMULATE) ? 1 : 0;
    a.n = M->n;
    a.perm = sorted ? nullptr : M->perm;
    a.B = B;
    a.C = C;
    a.ldb = K;
    a.ldc = K;
    a.kscale = kscale;
    a.gscale = gscale;
    a.slots = tuned ? M->tuned_slots : 0;
    if (tuned && M->tuned_kernel == SWIFT_KERNEL_WINDOW) kernel = SWIFT_KERNEL_WINDOW;
    const bool flat = kernel == SWIFT_KERNEL_FLAT;
    const bool winw = kernel == SWIFT_KERNEL_WINDOW;
    const bool view = kernel == SWIFT_KERNEL_DEFAULT || kernel == SWIFT_KERNEL_ROWVIEW || flat ||
                      winw;
    // The row-view walks (not FLAT) store the pieces of rows cut by a segment
    // boundary into carry slots; after the walk one tail launch adds them
    // onto C (k_row_tail) and runs the reside blocks, which add onto finished
    // rows with atomics.  The flat walk zeroes its split and empty rows first
    // and stores split rows with atomics.
    const bool carry = view && !flat;
    a.fix = st.fix;
    a.seg_slot = st.seg_slot;
    a.nfix = carry ? st.nfix : 0;
    a.nlong = carry ? st.nlong : 0;
    a.ldk = K;
    if (carry && st.nsplit)
        SWIFT_TRY(carry_workspace(M, (size_t)st.nsplit * K * sizeof(TC), s, &a.carry));
    const bool reside = a.nblk > 0 && !(view && M->rv_all);  // rv_all: the walk holds them
    a.rflag = reside ? M->resrow_flag : nullptr;
    if (flat) {
        if (!a.accumulate && st.nsplit) {
            hipLaunchKernelGGL(k_zero_rows<TC>, dim3(ceil_div((size_t)st.nsplit * 64, 256)),
                               dim3(256), 0, s, st.split_rows, st.nsplit, (TC *)C, K, K);
            SWIFT_CHECK_LAUNCH();
        }
        if (!a.accumulate && M->nempty) {
            hipLaunchKernelGGL(k_zero_rows<TC>, dim3(ceil_div((size_t)M->nempty * 64, 256)),
                               dim3(256), 0, s, M->empty_rows, M->nempty, (TC *)C, K, K);
            SWIFT_CHECK_LAUNCH();
        }
    } else if (!view && !a.accumulate) {
        SWIFT_TRY(fill_bytes_zero(C, (size_t)M->m * K * sizeof(TC), s));
    }
    mark(1, s);
    const bool buf = buffer_ok(M->n, K, sizeof(TB));
    for (auto &sl : slices) {
        a.kbase = sl.first;
        const RowTarget &t = sl.second;
        a.kcount = t.mask ? std::min(64, K - sl.first) : t.vec * t.lpe * t.q;
        a.bstride = buf ? (unsigned)(K * sizeof(TB)) : 0u;
        a.bbytes = buf ? (unsigned)(((size_t)M->n * K - sl.first) * sizeof(TB)) : 0u;
        const RowKind kind = flat ? ROW_FLAT : winw ? ROW_WIN : view ? ROW_VIEW : ROW_TILES;
        SWIFT_TRY((launch_row<TV, TB, TC>(a, t, kind, s)));
    }
    mark(2, s);
    const int nblk = a.nblk;
    bool first = true;
    for (auto &sl : slices) {
        a.kbase = sl.first;
        const RowTarget &t = sl.second;
        a.kcount = t.mask ? std::min(64, K - sl.first) : t.vec * t.lpe * t.q;
        a.nblk = reside ? nblk : 0;
        if (carry && first) {
            // fix-up of the cut rows (all of K) with this slice's reside blocks
            SWIFT_TRY((launch_row<TV, TB, TC>(a, t, ROW_TAIL, s)));
        } else if (a.nblk) {
            SWIFT_TRY((launch_row<TV, TB, TC>(a, t, ROW_RESIDE, s)));
        }
        first = false;
    }
    mark(3, s);
    return SWIFT_OK;
}

// out[c * rows + r] (+)= in[r * cols + c]: 32x32 tiles staged through LDS so
// both the read and the write are coalesced; one block per tile (1-D grid).
template <class T>
__global__ __launch_bounds__(256) void k_transpose(const T *__restrict__ in, T *__restrict__ out,
                                                   int rows, int cols, int tiles_c, int accumulate) {
    __shared__ T tile[32][33];
    const int tcol = (int)(blockIdx.x % (unsigned)tiles_c), trow = (int)(blockIdx.x / (unsigned)tiles_c);
    const int c0 = tcol * 32, r0 = trow * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
#pragma unroll
    for (int i = 0; i < 32; i += 8) {
        const int r = r0 + ty + i, c = c0 + tx;
        if (r < rows && c < cols) tile[ty + i][tx] = in[(size_t)r * cols + c];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 32; i += 8) {
        const int c = c0 + ty + i, r = r0 + tx;
        if (r < rows && c < cols) {
            T *o = out + (size_t)c * rows + r;
            *o = accumulate ? (T)(*o + tile[tx][ty + i]) : tile[tx][ty + i];
        }
    }
}

// 16-byte form: tiles of 32 rows x 32*V columns (V elements per 16 bytes,
// 16 KiB per tile), read as 4 row vectors and written as 4 column vectors per
// thread, so each block keeps 16 KiB of loads in flight instead of 4 KiB.
// Needs rows and cols multiples of V and 16-byte aligned bases.
template <class T>
__global__ __launch_bounds__(256) void k_transpose16(const T *__restrict__ in, T *__restrict__ out,
                                                     int rows, int cols, int tiles_c,
                                                     int accumulate) {
    constexpr int V = 16 / (int)sizeof(T);
    constexpr int TC = 32 * V;           // tile columns
    constexpr int RG = 32 / V;           // row vectors per output row of the tile
    __shared__ T tile[32][TC + 1];
    const int tcol = (int)(blockIdx.x % (unsigned)tiles_c), trow = (int)(blockIdx.x / (unsigned)tiles_c);
    const int c0 = tcol * TC, r0 = trow * 32;
    const int t = threadIdx.x;
    // 32 rows x 32 vectors: vector t + 256*i is (row (t>>5) + 8i, column vector t & 31)
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int rr = (t >> 5) + 8 * i, cv = t & 31;
        const int r = r0 + rr, c = c0 + cv * V;
        if (r < rows && c < cols) {
            // read once and written once: non-temporal loads and stores (a
            // C2 column-major product 1.115 -> 1.110 ms)
            const u32x4 vv =
                __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(in + (size_t)r * cols + c));
            const uint4 v = make_uint4(vv.x, vv.y, vv.z, vv.w);
            const T *e = reinterpret_cast<const T *>(&v);
#pragma unroll
            for (int k = 0; k < V; k++) tile[rr][cv * V + k] = e[k];
        }
    }
    __syncthreads();
    // TC output rows x RG vectors: vector t + 256*i is (output row
    // (t + 256i) / RG, rows rg*V .. rg*V + V-1 of the tile)
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int q = t + 256 * i;
        const int cc = q / RG, rg = q % RG;
        const int c = c0 + cc, r = r0 + rg * V;
        if (c < cols && r < rows) {
            T x[V];
#pragma unroll
            for (int k = 0; k < V; k++) x[k] = tile[rg * V + k][cc];
            T *o = out + (size_t)c * rows + r;
            if (accumulate) {
                const uint4 old = *reinterpret_cast<const uint4 *>(o);
                const T *e = reinterpret_cast<const T *>(&old);
#pragma unroll
                for (int k = 0; k < V; k++) x[k] = (T)(x[k] + e[k]);
            }
            u32x4 w;
            __builtin_memcpy(&w, x, 16);
            __builtin_nontemporal_store(w, reinterpret_cast<u32x4 *>(o));
        }
    }
}

template <class T>
static int transpose(const T *in, T *out, int rows, int cols, bool accumulate, hipStream_t s) {
    if (rows == 0 || cols == 0) return SWIFT_OK;
    constexpr int V = 16 / (int)sizeof(T);
    if (rows % V == 0 && cols % V == 0 && ((uintptr_t)in % 16) == 0 && ((uintptr_t)out % 16) == 0) {
        const size_t tc = ceil_div(cols, 32 * V), tr = ceil_div(rows, 32);
        if (tc * tr <= 0x7FFFFFFFull) {
            hipLaunchKernelGGL(k_transpose16<T>, dim3((unsigned)(tc * tr)), dim3(256), 0, s, in,
                               out, rows, cols, (int)tc, accumulate ? 1 : 0);
            SWIFT_CHECK_LAUNCH();
            return SWIFT_OK;
        }
    }
    const size_t tc = ceil_div(cols, 32), tr = ceil_div(rows, 32);
    if (tc * tr > 0x7FFFFFFFull) {
        set_error("transpose: %d x %d is too large", rows, cols);
        return SWIFT_ERR_UNSUPPORTED;
    }
    hipLaunchKernelGGL(k_transpose<T>, dim3((unsigned)(tc * tr)), dim3(256), 0, s, in, out, rows,
                       cols, (int)tc, accumulate ? 1 : 0);
    SWIFT_CHECK_LAUNCH();
    return SWIFT_OK;
}

// ------------------------------------------------ allocation-free entry --
//
// swift_spmm_noalloc takes every workspace from `scratch`: one array per GPU,
// allocated with the code object and held until the process ends, so a call
// allocates nothing (for HIP graph capture and loops that must not allocate).
// The host reaches it only through the device address hipGetSymbolAddress
// returns for the current device (looked up once per device and kept):
// `scratch` itself names a host-side stand-in.  Every call on one GPU shares
// the array; its transposes overwrite all they later read, so nothing needs
// clearing, and calls that overlap in time on one GPU race.
static __device__ __align__(256) unsigned char scratch[SWIFT_NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_addr[MAX_DEVICES];  // device address of scratch, per device

static int scratch_address(void **out) {
    int dev = 0;
    SWIFT_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) {
        set_error("swift_spmm_noalloc: device %d beyond the %d supported", dev, MAX_DEVICES);
        return SWIFT_ERR_UNSUPPORTED;
    }
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_addr[dev]) {
        void *p = nullptr;
        SWIFT_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) {
            set_error("swift_spmm_noalloc: no device address for the scratch array");
            return SWIFT_ERR_HIP;
        }
        g_scratch_addr[dev] = p;
    }
    *out = g_scratch_addr[dev];
    return SWIFT_OK;
}


// Column-major C = A*B.  DEFAULT: B is transposed into the matrix's
// workspace, the row-view product runs row-major, and C is transposed back
// (3 HBM-bound passes instead of strided scatter atomics).  Other kernels:
// the reference's tile order with lanes along the 32 columns of a tile.
template <class TV, class TB, class TC>
static int spmm_col(const swift_matrix *M, int K, swift_kernel kernel, unsigned flags,
                    const void *B, void *C, const void *sval, const void *rval,
                    const void *rv_val, const float *kscale, float gscale, hipStream_t s) {
    const bool accumulate = (flags & SWIFT_ACCUMULATE) != 0;
    if (kernel == SWIFT_KERNEL_DEFAULT) {
        const size_t bB = ((size_t)M->n * K * sizeof(TB) + 255) & ~(size_t)255;
        const size_t bC = (size_t)M->m * K * sizeof(TC);
        void *ws;
        if (g_ws) {
            if (bB + bC + 256 > g_ws->bytes) {
                set_error("swift_spmm_noalloc: the column-major product needs %zu bytes of "
                          "workspace, more than the %zu of the scratch array",
                          bB + bC + 256, g_ws->bytes);
                return SWIFT_ERR_ARG;
            }
            ws = g_ws->ptr;
        } else {
            SWIFT_TRY(M->scratch.alloc(bB + bC + 256));
            ws = M->scratch.ptr;
        }
        TB *Bt = static_cast<TB *>(ws);
        TC *Ct = reinterpret_cast<TC *>(static_cast<char *>(ws) + bB);
        SWIFT_TRY(transpose<TB>(static_cast<const TB *>(B), Bt, K, M->n, false, s));
        // the row-major product's carry slots follow the transposes' workspace
        g_ws_off = bB + bC + 256;
        const int rc = spmm_row<TV, TB, TC>(M, K, SWIFT_KERNEL_DEFAULT, flags & ~SWIFT_ACCUMULATE,
                                            Bt, Ct, sval, rval, rv_val, kscale, gscale, s);
        g_ws_off = 0;
        SWIFT_TRY(rc);
        return transpose<TC>(Ct, static_cast<TC *>(C), M->m, K, accumulate, s);
    }
    if (!accumulate) SWIFT_TRY(fill_bytes_zero(C, (size_t)M->m * K * sizeof(TC), s));
    mark(1, s);
    ColArgs a{};
    a.tilenum = M->tilenum;
    a.tile_colidx = M->tile_colidx;
    a.tile_format = M->tile_format;
    a.srow = M->sortrowidx;
    a.sval = sval;
    const bool unbalanced = kernel == SWIFT_KERNEL_UNBALANCED;
    a.nblk = unbalanced ? M->reside_col : M->numblk;
    a.colorblk = unbalanced ? nullptr : M->colorblk;
    a.blkstart = M->blkstart;
    a.blkstop = M->blkstop;
    a.rptr = M->reside_cscptr;
    a.rrow = M->reside_cscrowidx;
    a.rval = rval;
    a.m = M->m;
    a.n = M->n;
    a.perm = (flags & SWIFT_B_SORTED) ? nullptr : M->perm;
    a.B = B;
    a.C = C;
    a.ldb = M->n;
    a.ldc = M->m;
    a.kbase = 0;
    a.K = K;
    a.kscale = kscale;
    a.gscale = gscale;
    a.val_position = M->val_position;
    a.rowsorted = M->rowsorted;
    // K chunk per blockIdx.y: the reference's Swift_GPU_3/4 use 32-wide chunks
    // (src/Swift_3.h:338); the others walk all of K in one wave.
    const bool chunked = kernel == SWIFT_KERNEL_KCHUNK || kernel == SWIFT_KERNEL_ROWSORT;
    a.kchunk = chunked ? std::min(K, 32) : K;
    const unsigned ky = (unsigned)((K + a.kchunk - 1) / a.kchunk);
    if (kernel == SWIFT_KERNEL_ROWSORT) {
        if (!M->has_rowsort) {
            set_error("SWIFT_KERNEL_ROWSORT needs swift_row_idx_sort first");
            return SWIFT_ERR_ARG;
        }
        unsigned groups = (unsigned)((M->tilenum + 7) / 8);
        if (groups) {
            hipLaunchKernelGGL((k_regular_col_rowsort<TV, TB, TC>), dim3(groups, ky), dim3(256), 0,
                               s, a);
            SWIFT_CHECK_LAUNCH();
        }
    } else if (M->tilenum) {
        hipLaunchKernelGGL((k_regular_col<TV, TB, TC>),
                           dim3(ceil_div((size_t)M->tilenum * 32, 256), ky), dim3(256), 0, s, a);
        SWIFT_CHECK_LAUNCH();
    }
    mark(2, s);
    if (a.nblk) {
        hipLaunchKernelGGL((k_reside_col<TV, TB, TC>),
                           dim3(ceil_div((size_t)a.nblk * 32, 256), ky), dim3(256), 0, s, a);
        SWIFT_CHECK_LAUNCH();
    }
    mark(3, s);
    return SWIFT_OK;
}

int spmm_impl(const swift_matrix *M, int K, swift_layout layout, swift_dtype dtype,
              swift_kernel kernel, unsigned flags, const void *B, void *C,
              const float *kscale, hipStream_t s);

int spmm_timed_impl(const swift_matrix *M, int K, swift_layout layout, swift_dtype dtype,
                    swift_kernel kernel, unsigned flags, const void *B, void *C,
                    const float *kscale, hipStream_t s, float *ms3) {
    PhaseEvents pe;
    for (auto &e : pe.ev) SWIFT_HIP(hipEventCreate(&e));
    g_phase = &pe;
    int rc = spmm_impl(M, K, layout, dtype, kernel, flags, B, C, kscale, s);
    g_phase = nullptr;
    if (rc == SWIFT_OK) {
        SWIFT_HIP(hipEventSynchronize(pe.ev[3]));
        for (int i = 0; i < 3; i++) SWIFT_HIP(hipEventElapsedTime(&ms3[i], pe.ev[i], pe.ev[i + 1]));
    }
    for (auto &e : pe.ev) (void)hipEventDestroy(e);
    return rc;
}

int spmm_noalloc_impl(const swift_matrix *M, int K, swift_layout layout, swift_dtype dtype,
                      swift_kernel kernel, unsigned flags, const void *B, void *C,
                      hipStream_t s) {
    void *p = nullptr;
    SWIFT_TRY(scratch_address(&p));
    const Workspace w{p, (size_t)SWIFT_NOALLOC_SCRATCH_BYTES};
    g_ws = &w;
    const int rc = spmm_impl(M, K, layout, dtype, kernel, flags, B, C, nullptr, s);
    g_ws = nullptr;
    return rc;
}

int spmm_impl(const swift_matrix *M, int K, swift_layout layout, swift_dtype dtype,
              swift_kernel kernel, unsigned flags, const void *B, void *C,
              const float *kscale, hipStream_t s) {
    if (!M || K <= 0 || !B || !C) {
        set_error("swift_spmm: bad arguments");
        return SWIFT_ERR_ARG;
    }
    if (layout != SWIFT_ROW_MAJOR && layout != SWIFT_COL_MAJOR) {
        set_error("swift_spmm: bad layout");
        return SWIFT_ERR_ARG;
    }
    const bool lowp = dtype == SWIFT_BF16 || dtype == SWIFT_F16;
    if (!lowp && dtype != M->vtype) {
        set_error("swift_spmm: dtype %d differs from the build value type %d", (int)dtype,
                  (int)M->vtype);
        return SWIFT_ERR_ARG;
    }
    if (lowp && (M->lp_type != dtype || !M->sortval_lp)) {
        set_error("swift_spmm: call swift_prepare_values(%d) first", (int)dtype);
        return SWIFT_ERR_ARG;
    }
    if (kernel == SWIFT_KERNEL_ROWSORT && layout == SWIFT_ROW_MAJOR) {
        set_error("SWIFT_KERNEL_ROWSORT is a column-major kernel");
        return SWIFT_ERR_UNSUPPORTED;
    }
    if ((kernel == SWIFT_KERNEL_ROWVIEW || kernel == SWIFT_KERNEL_FLAT ||
         kernel == SWIFT_KERNEL_WINDOW) &&
        layout == SWIFT_COL_MAJOR) {
        set_error("SWIFT_KERNEL_ROWVIEW / FLAT / WINDOW are row-major kernels");
        return SWIFT_ERR_UNSUPPORTED;
    }
    if ((int)kernel < 0 || (int)kernel > (int)SWIFT_KERNEL_WINDOW) {
        set_error("swift_spmm: unknown kernel %d", (int)kernel);
        return SWIFT_ERR_ARG;
    }
    mark(0, s);
    const void *sv = lowp ? M->sortval_lp : M->sortval;
    const void *rv = lowp ? M->reside_val_lp : M->reside_val;
    const void *vv = lowp ? M->rv_val_lp : M->rv_val;
    const float gs = M->lp_scale;
    if (layout == SWIFT_ROW_MAJOR) {
        switch (dtype) {
            case SWIFT_F32: return spmm_row<float, float, float>(M, K, kernel, flags, B, C, sv, rv, vv, nullptr, 1.f, s);
            case SWIFT_F64: return spmm_row<double, double, double>(M, K, kernel, flags, B, C, sv, rv, vv, nullptr, 1.f, s);
            case SWIFT_BF16: return spmm_row<__hip_bfloat16, __hip_bfloat16, float>(M, K, kernel, flags, B, C, sv, rv, vv, nullptr, 1.f, s);
            case SWIFT_F16: return spmm_row<__half, __half, float>(M, K, kernel, flags, B, C, sv, rv, vv, kscale, gs, s);
        }
    } else {
        switch (dtype) {
            case SWIFT_F32: return spmm_col<float, float, float>(M, K, kernel, flags, B, C, sv, rv, vv, nullptr, 1.f, s);
            case SWIFT_F64: return spmm_col<double, double, double>(M, K, kernel, flags, B, C, sv, rv, vv, nullptr, 1.f, s);
            case SWIFT_BF16: return spmm_col<__hip_bfloat16, __hip_bfloat16, float>(M, K, kernel, flags, B, C, sv, rv, vv, nullptr, 1.f, s);
            case SWIFT_F16: return spmm_col<__half, __half, float>(M, K, kernel, flags, B, C, sv, rv, vv, kscale, gs, s);
        }
    }
    set_error("swift_spmm: unsupported dtype");
    return SWIFT_ERR_UNSUPPORTED;
}

// ------------------------------------------------ low-precision storage --

template <class TO>
__device__ __forceinline__ TO from_f32(float x);
template <> __device__ __forceinline__ __hip_bfloat16 from_f32(float x) { return __float2bfloat16(x); }
template <> __device__ __forceinline__ __half from_f32(float x) { return __float2half(x); }

template <class TI, class TO>
__global__ void k_convert(const TI *__restrict__ in, TO *__restrict__ out, size_t n, float inv) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) out[i] = from_f32<TO>((float)in[i] * inv);
}

// Per-k power-of-two scale so that max |B[:,k]| / scale <= 2^14 (F16 max is
// 65504): one atomicMax on the float bits per block and column.
__global__ void k_colmax(const float *__restrict__ in, size_t rows, int K, int col_major,
                         unsigned *__restrict__ cmax) {
    size_t total = rows * (size_t)K;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < total; i += stride) {
        int k = col_major ? (int)(i / rows) : (int)(i % K);
        float v = fabsf(in[i]);
        atomicMax(&cmax[k], __float_as_uint(v));
    }
}

__global__ void k_scale_from_max(const unsigned *__restrict__ cmax, int K, float *__restrict__ sc) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    float mx = __uint_as_float(cmax[k]);
    float s = 1.0f;
    if (mx > 0.f) {
        int e;
        frexpf(mx / 16384.0f, &e);
        s = ldexpf(1.0f, e);
    }
    sc[k] = s;
}

template <class TO>
__global__ void k_convert_scaled(const float *__restrict__ in, TO *__restrict__ out, size_t rows,
                                 int K, int col_major, const float *__restrict__ sc) {
    size_t total = rows * (size_t)K;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < total; i += stride) {
        int k = col_major ? (int)(i / rows) : (int)(i % K);
        out[i] = from_f32<TO>(in[i] / sc[k]);
    }
}

static unsigned grid_cap(size_t n) {
    size_t g = (n + 255) / 256;
    return (unsigned)std::max<size_t>(1, std::min<size_t>(g, 65536));
}

int convert_dense_impl(const float *in, void *out, size_t rows, int K, swift_layout layout,
                       swift_dtype dtype, float *col_scale, hipStream_t s) {
    size_t n = rows * (size_t)K;
    if (n == 0) return SWIFT_OK;
    if (dtype == SWIFT_BF16) {
        hipLaunchKernelGGL((k_convert<float, __hip_bfloat16>), dim3(grid_cap(n)), dim3(256), 0, s,
                           in, (__hip_bfloat16 *)out, n, 1.0f);
    } else if (dtype == SWIFT_F16) {
        if (col_scale) {
            Buffer cmax;
            SWIFT_TRY(cmax.alloc(sizeof(unsigned) * (size_t)K));
            SWIFT_TRY(fill_i32(cmax.as<int>(), 0, (size_t)K, s));
            hipLaunchKernelGGL(k_colmax, dim3(grid_cap(n)), dim3(256), 0, s, in, rows, K,
                               layout == SWIFT_COL_MAJOR, cmax.as<unsigned>());
            hipLaunchKernelGGL(k_scale_from_max, dim3(ceil_div(K, 256)), dim3(256), 0, s,
                               cmax.as<unsigned>(), K, col_scale);
            hipLaunchKernelGGL((k_convert_scaled<__half>), dim3(grid_cap(n)), dim3(256), 0, s, in,
                               (__half *)out, rows, K, layout == SWIFT_COL_MAJOR, col_scale);
            SWIFT_CHECK_LAUNCH();
            SWIFT_HIP(hipStreamSynchronize(s));
            return SWIFT_OK;
        }
        hipLaunchKernelGGL((k_convert<float, __half>), dim3(grid_cap(n)), dim3(256), 0, s, in,
                           (__half *)out, n, 1.0f);
    } else {
        set_error("swift_convert_dense: dtype must be BF16 or F16");
        return SWIFT_ERR_ARG;
    }
    SWIFT_CHECK_LAUNCH();
    return SWIFT_OK;
}

__global__ void k_absmax_f32(const float *__restrict__ a, size_t n, unsigned *out) {
    unsigned v = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        v = max(v, __float_as_uint(fabsf(a[i])));
    atomicMax(out, v);
}

__global__ void k_absmax_f64(const double *__restrict__ a, size_t n, unsigned *out) {
    unsigned v = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        v = max(v, __float_as_uint(fabsf((float)a[i])));
    atomicMax(out, v);
}

// A's values to BF16, or to F16 divided by one power-of-two scale chosen so
// that max |a| / scale <= 2^14.
int prepare_values_impl(swift_matrix *M, swift_dtype dtype, hipStream_t s) {
    if (dtype != SWIFT_BF16 && dtype != SWIFT_F16) {
        set_error("swift_prepare_values: dtype must be BF16 or F16");
        return SWIFT_ERR_ARG;
    }
    for (void **p : {&M->sortval_lp, &M->reside_val_lp, &M->rv_val_lp})
        if (*p) { (void)hipFree(*p); *p = nullptr; }
    SWIFT_HIP(hipMalloc(&M->sortval_lp, std::max<size_t>(2, (size_t)M->regular_nnz * 2)));
    SWIFT_HIP(hipMalloc(&M->reside_val_lp, std::max<size_t>(2, (size_t)M->reside_nnz * 2)));
    SWIFT_HIP(hipMalloc(&M->rv_val_lp, std::max<size_t>(2, (size_t)M->rv_nnz * 2)));
    float scale = 1.0f;
    if (dtype == SWIFT_F16) {
        Buffer mx;
        SWIFT_TRY(mx.alloc(sizeof(unsigned)));
        SWIFT_TRY(fill_i32(mx.as<int>(), 0, 1, s));
        if (M->vtype == SWIFT_F64) {
            if (M->regular_nnz) hipLaunchKernelGGL(k_absmax_f64, dim3(grid_cap(M->regular_nnz)), dim3(256), 0, s, (const double *)M->sortval, (size_t)M->regular_nnz, mx.as<unsigned>());
            if (M->reside_nnz) hipLaunchKernelGGL(k_absmax_f64, dim3(grid_cap(M->reside_nnz)), dim3(256), 0, s, (const double *)M->reside_val, (size_t)M->reside_nnz, mx.as<unsigned>());
        } else {
            if (M->regular_nnz) hipLaunchKernelGGL(k_absmax_f32, dim3(grid_cap(M->regular_nnz)), dim3(256), 0, s, (const float *)M->sortval, (size_t)M->regular_nnz, mx.as<unsigned>());
            if (M->reside_nnz) hipLaunchKernelGGL(k_absmax_f32, dim3(grid_cap(M->reside_nnz)), dim3(256), 0, s, (const float *)M->reside_val, (size_t)M->reside_nnz, mx.as<unsigned>());
        }
        SWIFT_CHECK_LAUNCH();
        unsigned bits = 0;
        SWIFT_HIP(hipMemcpyAsync(&bits, mx.ptr, 4, hipMemcpyDeviceToHost, s));
        SWIFT_HIP(hipStreamSynchronize(s));
        float amax;
        std::memcpy(&amax, &bits, 4);
        if (amax > 0.f) {
            int e;
            std::frexp(amax / 16384.0f, &e);
            scale = std::ldexp(1.0f, e);
        }
    }
    const float inv = 1.0f / scale;
    auto conv = [&](const void *src, void *dst, size_t n) -> int {
        if (n == 0) return SWIFT_OK;
        if (M->vtype == SWIFT_F64) {
            if (dtype == SWIFT_BF16) hipLaunchKernelGGL((k_convert<double, __hip_bfloat16>), dim3(grid_cap(n)), dim3(256), 0, s, (const double *)src, (__hip_bfloat16 *)dst, n, inv);
            else hipLaunchKernelGGL((k_convert<double, __half>), dim3(grid_cap(n)), dim3(256), 0, s, (const double *)src, (__half *)dst, n, inv);
        } else {
            if (dtype == SWIFT_BF16) hipLaunchKernelGGL((k_convert<float, __hip_bfloat16>), dim3(grid_cap(n)), dim3(256), 0, s, (const float *)src, (__hip_bfloat16 *)dst, n, inv);
            else hipLaunchKernelGGL((k_convert<float, __half>), dim3(grid_cap(n)), dim3(256), 0, s, (const float *)src, (__half *)dst, n, inv);
        }
        SWIFT_CHECK_LAUNCH();
        return SWIFT_OK;
    };
    SWIFT_TRY(conv(M->sortval, M->sortval_lp, (size_t)M->regular_nnz));
    SWIFT_TRY(conv(M->reside_val, M->reside_val_lp, (size_t)M->reside_nnz));
    SWIFT_TRY(conv(M->rv_val, M->rv_val_lp, (size_t)M->rv_nnz));
    SWIFT_HIP(hipStreamSynchronize(s));
    M->lp_type = dtype;
    M->lp_scale = scale;
    return SWIFT_OK;
}

}  // namespace swift
