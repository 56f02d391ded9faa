// batched.hip -- many uneven Swift SpMMs in one persistent launch, and the
// autotuner that picks each matrix's row-view schedule.
//
// A batch is a list of built matrices (FP32, row-major, K in {32, 64, 128,
// 256} per matrix).  Their row-view segments are flattened into one work list
// of (matrix, segment) items, cut into 8 contiguous ranges of about equal
// work, one per XCD; a persistent kernel sized to the device (32 waves per
// CU) hands items to waves, two per claim, through one atomic counter per
// range (each counter on a cache line of its own), and each wave runs the
// same walk as the single-matrix row-view kernel with the matrix's own lane
// shape (chosen by a wave-uniform switch on K), storing the pieces of split
// rows into the batch's carry slots.  One more kernel adds every matrix's
// carry slots onto its split rows and runs the reside blocks of all matrices.
// swift_batch_graph_launch captures the three launches (counter reset,
// persistent walk, fix-up and reside blocks) into a HIP graph once and
// replays it; new B/C pointers only rewrite the descriptor table.
//
// Autotuner: matrices are keyed by (K, log2 mean nonzeros per row, reside
// share bucket, whether the row view holds the reside entries); the matrices
// of each key run as one batch under segment sizes {64, 128, 256, 512} x two
// slot-row counts, and at half the lane width (K in two halves, walked as two
// descriptors), and the fastest is applied to all of them.  Results are
// cached per process.  (Timing single products instead favours small
// segments, which fill the device for one small matrix but, pooled in a
// batch, only add items and split rows.)
#include "rowview.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

namespace swift {

int build_segments(swift_matrix *M, int W, hipStream_t s);
int spmm_impl(const swift_matrix *M, int K, swift_layout layout, swift_dtype dtype,
              swift_kernel kernel, unsigned flags, const void *B, void *C, const float *kscale,
              hipStream_t s);

// ------------------------------------------------------------ kernels --

constexpr int NXCD = 8;  // XCDs of an MI355X; blocks are dealt to them round-robin
// The claim counters sit one per 128-byte line (CPAD ints apart): atomics on
// one line serialise in one L2 channel, and with the 8 counters in one line
// the 10^6 claims of the 512-matrix batch took as long as the whole walk
// (13.9 ms against 6.6 ms with one line per counter).
constexpr int CPAD = 32;
// Items handed out per claim: 2 halves the atomics and keeps the tail short
// (512 matrices: 1 item 6.59 ms, 2 items 6.13-6.25 ms, 4 items 6.30 ms, 16
// items 7.79 ms).
constexpr int CLAIM = 2;

__global__ void k_batch_reset(int *counter) {
    if (threadIdx.x < NXCD) counter[threadIdx.x * CPAD] = 0;
}

template <int LPE, bool BUF, int SLOTS = 0>
__device__ __forceinline__ void batch_walk(const RowArgs &a, int sid) {
    rowview_walk<float, float, float, 4, 1, LPE, false, BUF, false, SLOTS>(a, sid);
}

// Slot-row count the batched walk can run besides the rule for a K: 8 for
// K = 32 and 64 (rule 4), 4 for K = 128 and 256 (rule 8).  Larger counts
// would raise the register count of the one kernel that holds every shape.
constexpr int batch_alt_slots(int K) { return K <= 64 ? 8 : 4; }

// Next CLAIM items of a wave: items range[x] .. range[x+1]-1 belong to XCD x
// and counter[x * CPAD] counts the ones handed out.  A wave drains its own
// XCD's range first (block b runs on XCD b % 8, so the waves of one XCD walk
// neighbouring segments of the same matrices and their B rows meet in that
// XCD's L2), then the other ranges in turn, so the tail stays balanced.
// Returns the first item of the claim, or -1 once all NXCD ranges are
// exhausted: counters only grow, so every wave reaches the exit after at most
// NXCD empty claims.
__device__ __forceinline__ int batch_claim(const int *__restrict__ range,
                                           int *__restrict__ counter, int &x, int &left) {
    while (left > 0) {
        int it = 0;
        if ((threadIdx.x & 63) == 0) it = atomicAdd(counter + x * CPAD, CLAIM);
        it = __builtin_amdgcn_readfirstlane(__shfl(it, 0, 64)) + range[x];
        if (it < range[x + 1]) return it;
        x = x + 1 == NXCD ? 0 : x + 1;
        left--;
    }
    return -1;
}

// Persistent: every wave claims CLAIM items at a time until every range is
// exhausted (an exit every wave reaches), and walks each with the lane shape
// of its matrix's K.  BUF: every matrix's B fits the buffer-resource path.
template <bool BUF>
__global__ __launch_bounds__(256) void k_batch_walk(const RowArgs *__restrict__ desc,
                                                    const int2 *__restrict__ items,
                                                    const int *__restrict__ range,
                                                    int *__restrict__ counter) {
    int x = (int)(blockIdx.x % NXCD), left = NXCD;
    for (;;) {
        const int first = batch_claim(range, counter, x, left);
        if (first < 0) break;
        const int end = min(first + CLAIM, range[x + 1]);
        for (int it = first; it < end; it++) {
            const int2 w = items[it];
            const int mi = __builtin_amdgcn_readfirstlane(w.x);
            const int sid = __builtin_amdgcn_readfirstlane(w.y);
            const RowArgs &a = desc[mi];
            // lane shape by the slice width (K, or K/2 for a matrix walked in
            // two halves), slot rows by the autotuner (0: the rule)
            switch (a.kcount | (a.slots << 12)) {
                case 32: batch_walk<8, BUF>(a, sid); break;
                case 32 | (batch_alt_slots(32) << 12): batch_walk<8, BUF, batch_alt_slots(32)>(a, sid); break;
                case 64: batch_walk<16, BUF>(a, sid); break;
                case 64 | (batch_alt_slots(64) << 12): batch_walk<16, BUF, batch_alt_slots(64)>(a, sid); break;
                case 128: batch_walk<32, BUF>(a, sid); break;
                case 128 | (batch_alt_slots(128) << 12): batch_walk<32, BUF, batch_alt_slots(128)>(a, sid); break;
                case 256 | (batch_alt_slots(256) << 12): batch_walk<64, BUF, batch_alt_slots(256)>(a, sid); break;
                // K = 256 at 8 slot rows: the rule's 16 for a single product
                // would raise the register count of the kernel that holds
                // every shape
                default: batch_walk<64, BUF, 8>(a, sid); break;  // 256
            }
        }
    }
}

// One wave per item: the first nfix items (matrix, cut row) add the row's
// carry slots onto it (split_fixup), the next nres (matrix, balance block)
// run that reside block with atomics onto the finished rows.
__global__ __launch_bounds__(256) void k_batch_tail(const RowArgs *__restrict__ desc,
                                                    const int2 *__restrict__ fix, int nfix,
                                                    const int2 *__restrict__ items, int n) {
    int i = (int)((blockIdx.x * (size_t)blockDim.x + threadIdx.x) >> 6);
    if (i < nfix) {
        const int2 f = fix[i];
        split_fixup<float, 64>(desc[f.x], f.y);
        return;
    }
    i -= nfix;
    if (i >= n) return;
    const int2 w = items[i];
    const int mi = __builtin_amdgcn_readfirstlane(w.x);
    const int blk = __builtin_amdgcn_readfirstlane(w.y);
    const RowArgs a = desc[mi];
    switch (a.ldb) {
        case 32: reside_block<float, float, float, 4, 1, 8, false>(a, blk); break;
        case 64: reside_block<float, float, float, 4, 1, 16, false>(a, blk); break;
        case 128: reside_block<float, float, float, 4, 1, 32, false>(a, blk); break;
        default: reside_block<float, float, float, 4, 1, 64, false>(a, blk); break;
    }
}

// ---------------------------------------------------------- autotuner --

// (K, log2 mean nonzeros per row, reside share bucket, row view holds the
// reside entries): matrices of one key walk the same entry sets with the
// same lane shape.
using TuneKey = std::tuple<int, int, int, int, int>;
struct TuneChoice {
    int W = 128, slots = 0;
    swift_kernel kernel = SWIFT_KERNEL_DEFAULT;
    int cols = 0;  // widest K slice of one launch (0: no cap)
};
static std::map<TuneKey, TuneChoice> g_tuned;
static std::mutex g_tune_mu;
// swift_batch_set_lane_width: which lane widths the batch autotuner times
static std::atomic<int> g_batch_width{SWIFT_BATCH_WIDTH_TUNED};

static TuneKey tune_key(const swift_matrix *M, int K) {
    const double mean = M->m ? (double)M->nnz / M->m : 0.0;
    const int dens = (int)std::floor(std::log2(mean + 1.0));
    const double share = M->nnz ? (double)M->reside_nnz / M->nnz : 0.0;
    const int res = share == 0.0 ? 0 : share < 0.01 ? 1 : share < 0.1 ? 2 : 3;
    return TuneKey(K, dens, res, M->rv_all ? 1 : 0, g_batch_width);
}

// Makes c the matrix's schedule for products of K columns with elem-byte B
// elements: its segment table (W = 0: the default tables by B-row bytes),
// slot rows per step and walk.
static int apply_choice(swift_matrix *M, int K, size_t elem, const TuneChoice &c, hipStream_t s) {
    if (c.W && (M->mp_items != c.W || !M->user_items)) SWIFT_TRY(build_segments(M, c.W, s));
    M->user_items = c.W;
    M->tuned_K = K;
    M->tuned_elem = (int)elem;
    M->tuned_slots = c.slots;
    M->tuned_kernel = c.kernel;
    M->tuned_cols = c.cols;
    return SWIFT_OK;
}

// Times the SWIFT_KERNEL_DEFAULT product of M on the given operands for every
// candidate schedule and keeps the fastest (median of 3 after one warm-up).
// Candidates: segment sizes {default, 64, 128, 256, 512} (default: the table
// of the B-row bytes, W = 0 in the choice) x the slot-row counts of the
// product's lane shape (slot_candidates), plus at each size the window walk
// (LDS-staged column windows, row-major only) and the other slice widths
// (width_candidates: half the width, in two slices of narrower lane groups,
// and full width where the default slices K).
int autotune_impl(swift_matrix *M, int K, swift_dtype dtype, swift_layout layout, const void *B,
                  void *C, hipStream_t s, TuneChoice *best_out) {
    static const int sizes[] = {0, 64, 128, 256, 512};
    const size_t elem = value_size(dtype);
    std::vector<TuneChoice> cands;
    const int vec = vec_of(elem);
    const bool aligned = K % vec == 0 && (((uintptr_t)B | (uintptr_t)C) % 16) == 0;
    const std::vector<int> slots = slot_candidates(K, elem, aligned);
    const std::vector<int> widths = width_candidates(K, elem, aligned);
    for (int W : sizes) {
        for (int n : slots) cands.push_back({W, n, SWIFT_KERNEL_DEFAULT, 0});
        for (int w : widths)
            if (w) cands.push_back({W, 0, SWIFT_KERNEL_DEFAULT, w});
        if (layout == SWIFT_ROW_MAJOR) cands.push_back({W, 0, SWIFT_KERNEL_WINDOW, 0});
    }
    hipEvent_t e0, e1;
    SWIFT_HIP(hipEventCreate(&e0));
    SWIFT_HIP(hipEventCreate(&e1));
    TuneChoice best;
    float best_ms = 1e30f;
    int rc = SWIFT_OK;
    for (const TuneChoice &c : cands) {
        if ((rc = apply_choice(M, K, elem, c, s)) != SWIFT_OK) break;
        std::vector<float> t;
        for (int r = 0; r < 4 && rc == SWIFT_OK; r++) {
            (void)hipEventRecord(e0, s);
            rc = spmm_impl(M, K, layout, dtype, SWIFT_KERNEL_DEFAULT, 0, B, C, nullptr, s);
            (void)hipEventRecord(e1, s);
            (void)hipEventSynchronize(e1);
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (r) t.push_back(ms);
        }
        if (rc != SWIFT_OK) break;
        std::sort(t.begin(), t.end());
        if (t[1] < best_ms) {
            best_ms = t[1];
            best = c;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != SWIFT_OK) return rc;
    SWIFT_TRY(apply_choice(M, K, elem, best, s));
    if (best_out) *best_out = best;
    return SWIFT_OK;
}

}  // namespace swift

using namespace swift;

struct swift_batch {
    std::vector<swift_matrix *> mats;
    std::vector<int> Ks, W, slots;  // per matrix: K, segment size, slot rows (0: rule)
    std::vector<int> cols;          // per matrix: K/2 when walked in two halves, else 0
    // descriptors: one per matrix (full K: the split-row fix-up, reside blocks and
    // whole-width walks), then two per matrix walked in halves (kbase 0 and
    // K/2); half_of[j] is the matrix of descriptor count + j
    std::vector<RowArgs> desc_h;
    std::vector<int> half_of;
    RowArgs *desc = nullptr;
    int2 *items = nullptr, *res_items = nullptr, *fix_items = nullptr;
    int nitems = 0, nres = 0, nfix = 0;
    void *carry = nullptr;  // carry slots of every matrix's split rows
    int *counter = nullptr;  // NXCD claim counters, CPAD ints apart
    int *range = nullptr;    // NXCD+1 item bounds of the per-XCD ranges
    int grid = 0;
    bool buf = true;  // all B operands fit the buffer-resource path
    std::vector<const float *> lastB;
    std::vector<float *> lastC;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t graph_stream = nullptr;
};

static void batch_free(swift_batch *b) {
    if (!b) return;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (void *p : {(void *)b->desc, (void *)b->items, (void *)b->res_items, (void *)b->fix_items,
                    (void *)b->carry, (void *)b->counter, (void *)b->range})
        if (p) (void)hipFree(p);
    delete b;
}

static RowArgs row_args_of(const swift_matrix *M, int K) {
    RowArgs a{};
    a.nblk = M->numblk;
    a.colorblk = M->colorblk;
    a.blkstart = M->blkstart;
    a.blkstop = M->blkstop;
    a.rptr = M->reside_cscptr;
    a.rrow = M->reside_cscrowidx;
    a.rval = M->reside_val;
    a.m = M->m;
    a.nseg = M->prim.nseg;
    a.seg = M->prim.seg;
    a.rv_ptr = M->rv_ptr;
    a.rv_col = M->rv_col;
    a.rv_val = M->rv_val;
    a.n = M->n;
    a.perm = M->perm;
    a.ldb = K;
    a.ldc = K;
    a.kbase = 0;
    a.kcount = K;
    a.gscale = 1.0f;
    if (buffer_ok(M->n, K, sizeof(float))) {
        a.bstride = (unsigned)(K * sizeof(float));
        a.bbytes = (unsigned)((size_t)M->n * K * sizeof(float));
    }
    a.seg_slot = M->prim.seg_slot;
    a.fix = M->prim.fix;
    a.nfix = M->prim.nfix;
    a.ldk = K;
    a.rflag = M->rv_all ? nullptr : M->resrow_flag;
    return a;
}

static int batch_launch(swift_batch *b, hipStream_t s);

// Builds a batch's work from b->mats / Ks / W / slots / cols: makes each
// matrix's schedule its own (apply_choice; a matrix listed more than once
// keeps its first segment size), then the flattened work lists, descriptors,
// carry slots, claim counters and per-XCD ranges on the device.
static int batch_build(swift_batch *b, hipStream_t s) {
    const int count = (int)b->mats.size();
    std::map<const swift_matrix *, int> first;
    for (int i = 0; i < count; i++) {
        auto it = first.emplace(b->mats[i], b->W[i]).first;
        b->W[i] = it->second;
    }
    for (int i = 0; i < count; i++)
        SWIFT_TRY(apply_choice(b->mats[i], b->Ks[i], sizeof(float),
                               TuneChoice{b->W[i], b->slots[i], SWIFT_KERNEL_DEFAULT, b->cols[i]}, s));
    std::vector<int2> items, res, fix;
    std::vector<size_t> carry_off;  // per matrix: its carry slots in b->carry
    size_t carry_bytes = 0;
    std::vector<double> work;  // prefix of segment items x K per item
    double total = 0.0;
    for (int i = 0; i < count; i++) {
        swift_matrix *M = b->mats[i];
        if (b->cols[i]) {
            // two halves of K, each one descriptor; every segment of the first
            // half, then of the second, so that a stretch of the XCD's work
            // gathers from one half of B
            const int d0 = count + (int)b->half_of.size();
            b->half_of.push_back(i);
            b->half_of.push_back(i);
            for (int h = 0; h < 2; h++)
                for (int q = 0; q < M->prim.nseg; q++) {
                    items.push_back(make_int2(d0 + h, q));
                    total += (double)b->W[i] * b->cols[i];
                    work.push_back(total);
                }
        } else {
            for (int q = 0; q < M->prim.nseg; q++) {
                items.push_back(make_int2(i, q));
                total += (double)b->W[i] * b->Ks[i];
                work.push_back(total);
            }
        }
        if (!M->rv_all)  // rv_all: the walk holds the reside entries
            for (int q = 0; q < M->numblk; q++) res.push_back(make_int2(i, q));
        for (int r = 0; r < M->prim.nfix; r++) fix.push_back(make_int2(i, r));
        carry_off.push_back(carry_bytes);
        carry_bytes += ((size_t)M->prim.nsplit * b->Ks[i] * sizeof(float) + 255) & ~(size_t)255;
        b->desc_h.push_back(row_args_of(M, b->Ks[i]));
        b->desc_h.back().slots = b->slots[i];
        if (!b->desc_h.back().bstride) b->buf = false;
    }
    if (hipMalloc(&b->carry, std::max<size_t>(carry_bytes, 256)) != hipSuccess) {
        set_error("swift_batch_create: hipMalloc(%zu) for the carry slots failed", carry_bytes);
        return SWIFT_ERR_ALLOC;
    }
    for (int i = 0; i < count; i++)
        b->desc_h[i].carry = static_cast<char *>(b->carry) + carry_off[i];
    for (size_t j = 0; j < b->half_of.size(); j++) {
        const int i = b->half_of[j];
        RowArgs h = b->desc_h[i];
        h.kcount = b->cols[i];
        h.kbase = (int)(j % 2) * b->cols[i];
        h.slots = 0;
        if (h.bstride) h.bbytes -= (unsigned)(h.kbase * sizeof(float));
        b->desc_h.push_back(h);
    }
    auto up = [&](void **dst, const void *src, size_t bytes) -> int {
        if (hipMalloc(dst, std::max<size_t>(bytes, 16)) != hipSuccess) {
            set_error("swift_batch_create: hipMalloc(%zu) failed", bytes);
            return SWIFT_ERR_ALLOC;
        }
        if (bytes && hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice) != hipSuccess) {
            set_error("swift_batch_create: upload failed");
            return SWIFT_ERR_HIP;
        }
        return SWIFT_OK;
    };
    SWIFT_TRY(up((void **)&b->items, items.data(), items.size() * sizeof(int2)));
    SWIFT_TRY(up((void **)&b->res_items, res.data(), res.size() * sizeof(int2)));
    SWIFT_TRY(up((void **)&b->fix_items, fix.data(), fix.size() * sizeof(int2)));
    SWIFT_TRY(up((void **)&b->desc, b->desc_h.data(), b->desc_h.size() * sizeof(RowArgs)));
    const int zeros[NXCD * CPAD] = {};
    SWIFT_TRY(up((void **)&b->counter, zeros, sizeof(zeros)));
    // range x ends at the first item whose work prefix reaches (x+1)/NXCD of
    // the total
    int bounds[NXCD + 1];
    bounds[0] = 0;
    for (int x = 1; x <= NXCD; x++) {
        const double goal = total * x / NXCD;
        bounds[x] = x == NXCD ? (int)items.size()
                              : (int)(std::lower_bound(work.begin(), work.end(), goal) -
                                      work.begin());
        bounds[x] = std::max(bounds[x], bounds[x - 1]);
    }
    SWIFT_TRY(up((void **)&b->range, bounds, sizeof(bounds)));
    b->nitems = (int)items.size();
    b->nres = (int)res.size();
    b->nfix = (int)fix.size();
    int dev = 0, cus = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    b->grid = std::max(1, cus) * 8;  // 8 blocks x 4 waves = 32 waves per CU
    b->lastB.assign(count, nullptr);
    b->lastC.assign(count, nullptr);
    return SWIFT_OK;
}

// Tunes the schedule of one key's matrices on the batched walk itself: every
// candidate (segment sizes {64, 128, 256, 512} x the rule's and the
// alternative slot rows, and half the lane width at each size) is applied to
// all of them, they run as one batch (3 timed launches after a warm-up, the
// median kept) and the fastest wins.  The operands are one zeroed B and one C
// shared by all the matrices (sized for the largest): only the time is used.
static int tune_batch_key(swift_matrix *const *mats, const int *Ks, const std::vector<int> &idx,
                          hipStream_t s, TuneChoice *best_out) {
    const int K = Ks[idx[0]];
    std::vector<int> slots{0, batch_alt_slots(K)};
    std::vector<int> widths = width_candidates(K, sizeof(float), true, false);
    if (g_batch_width == SWIFT_BATCH_WIDTH_FULL) {
        widths = {0};
    } else if (g_batch_width == SWIFT_BATCH_WIDTH_HALF && widths.size() > 1) {
        widths = {widths[1]};
        slots.clear();  // half width only, at the rule's slot rows
    }
    std::vector<TuneChoice> cands;
    for (int W : {64, 128, 256, 512}) {
        for (int n : slots) cands.push_back({W, n, SWIFT_KERNEL_DEFAULT, 0});
        for (int w : widths)
            if (w) cands.push_back({W, 0, SWIFT_KERNEL_DEFAULT, w});
    }
    size_t nb = 16, nc = 16;
    for (int i : idx) {
        nb = std::max(nb, (size_t)mats[i]->n * K);
        nc = std::max(nc, (size_t)mats[i]->m * K);
    }
    Buffer Bs, Cs;
    SWIFT_TRY(Bs.alloc(sizeof(float) * nb));
    SWIFT_TRY(Cs.alloc(sizeof(float) * nc));
    SWIFT_TRY(fill_bytes_zero(Bs.ptr, sizeof(float) * nb, s));
    hipEvent_t e0, e1;
    SWIFT_HIP(hipEventCreate(&e0));
    SWIFT_HIP(hipEventCreate(&e1));
    TuneChoice best = cands[0];
    float best_ms = 1e30f;
    int rc = SWIFT_OK;
    for (const TuneChoice &c : cands) {
        swift_batch *sub = new (std::nothrow) swift_batch();
        if (!sub) { rc = SWIFT_ERR_ALLOC; break; }
        for (int i : idx) {
            sub->mats.push_back(mats[i]);
            sub->Ks.push_back(Ks[i]);
            sub->W.push_back(c.W);
            sub->slots.push_back(c.slots);
            sub->cols.push_back(c.cols);
        }
        rc = batch_build(sub, s);
        if (rc == SWIFT_OK) {
            for (auto &d : sub->desc_h) {
                d.B = Bs.ptr;
                d.C = Cs.ptr;
            }
            if (hipMemcpy(sub->desc, sub->desc_h.data(), sub->desc_h.size() * sizeof(RowArgs),
                          hipMemcpyHostToDevice) != hipSuccess) {
                set_error("swift_batch_create: descriptor upload failed");
                rc = SWIFT_ERR_HIP;
            }
        }
        std::vector<float> t;
        for (int r = 0; r < 4 && rc == SWIFT_OK; r++) {
            (void)hipEventRecord(e0, s);
            rc = batch_launch(sub, s);
            (void)hipEventRecord(e1, s);
            (void)hipEventSynchronize(e1);
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (r) t.push_back(ms);
        }
        (void)hipStreamSynchronize(s);
        batch_free(sub);
        if (rc != SWIFT_OK) break;
        std::sort(t.begin(), t.end());
        if (t[1] < best_ms) {
            best_ms = t[1];
            best = c;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc == SWIFT_OK) *best_out = best;
    return rc;
}

extern "C" int swift_batch_create(swift_batch **out, swift_matrix *const *mats, const int *Ks,
                                  int count, swift_stream_t stream) {
    if (!out || !mats || !Ks || count <= 0) {
        set_error("swift_batch_create: bad arguments");
        return SWIFT_ERR_ARG;
    }
    *out = nullptr;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    for (int i = 0; i < count; i++) {
        const int K = Ks[i];
        if (!mats[i] || mats[i]->vtype != SWIFT_F32 || !(K == 32 || K == 64 || K == 128 || K == 256)) {
            set_error("swift_batch_create: matrix %d must be FP32 with K in {32,64,128,256}", i);
            return SWIFT_ERR_UNSUPPORTED;
        }
    }
    swift_batch *b = new (std::nothrow) swift_batch();
    if (!b) return SWIFT_ERR_ALLOC;
    int rc = SWIFT_OK;
    try {
        b->mats.assign(mats, mats + count);
        b->Ks.assign(Ks, Ks + count);
        b->W.assign(count, 128);
        b->slots.assign(count, 0);
        b->cols.assign(count, 0);
        // the matrices of each key, tuned together on the batched walk
        std::map<TuneKey, std::vector<int>> keys;
        for (int i = 0; i < count; i++) keys[tune_key(mats[i], Ks[i])].push_back(i);
        for (auto &kv : keys) {
            TuneChoice c;
            bool have = false;
            {
                std::lock_guard<std::mutex> g(g_tune_mu);
                auto it = g_tuned.find(kv.first);
                if (it != g_tuned.end()) {
                    c = it->second;
                    have = true;
                }
            }
            if (!have) {
                if ((rc = tune_batch_key(mats, Ks, kv.second, s, &c)) != SWIFT_OK) break;
                std::lock_guard<std::mutex> g(g_tune_mu);
                g_tuned[kv.first] = c;
            }
            for (int i : kv.second) {
                b->W[i] = c.W;
                b->slots[i] = c.slots;
                b->cols[i] = c.cols;
            }
        }
        if (rc == SWIFT_OK) rc = batch_build(b, s);
    } catch (...) {
        set_error("swift_batch_create: host allocation failed");
        rc = SWIFT_ERR_ALLOC;
    }
    if (rc != SWIFT_OK) {
        batch_free(b);
        return rc;
    }
    *out = b;
    return SWIFT_OK;
}

static int batch_set_operands(swift_batch *b, const float *const *dB, float *const *dC,
                              hipStream_t s, bool *changed) {
    if (!dB || !dC) {
        set_error("swift_batch_spmm: null operand arrays");
        return SWIFT_ERR_ARG;
    }
    bool ch = false;
    for (size_t i = 0; i < b->mats.size(); i++) {
        const swift_matrix *M = b->mats[i];
        if (M->prim.seg != b->desc_h[i].seg || M->prim.nseg != b->desc_h[i].nseg ||
            M->mp_items != b->W[i]) {
            set_error("swift_batch_spmm: matrix %zu was re-segmented after swift_batch_create", i);
            return SWIFT_ERR_ARG;
        }
        if (!dB[i] || !dC[i]) {
            set_error("swift_batch_spmm: null operand %zu", i);
            return SWIFT_ERR_ARG;
        }
        if (((uintptr_t)dB[i] | (uintptr_t)dC[i]) % 16) {
            set_error("swift_batch_spmm: operands of matrix %zu are not 16-byte aligned", i);
            return SWIFT_ERR_ARG;
        }
        if (dB[i] != b->lastB[i] || dC[i] != b->lastC[i]) {
            b->desc_h[i].B = dB[i];
            b->desc_h[i].C = dC[i];
            b->lastB[i] = dB[i];
            b->lastC[i] = dC[i];
            ch = true;
        }
    }
    const size_t count = b->mats.size();
    for (size_t j = 0; j < b->half_of.size(); j++) {
        b->desc_h[count + j].B = b->desc_h[b->half_of[j]].B;
        b->desc_h[count + j].C = b->desc_h[b->half_of[j]].C;
    }
    if (ch) {
        SWIFT_HIP(hipMemcpyAsync(b->desc, b->desc_h.data(), b->desc_h.size() * sizeof(RowArgs),
                                 hipMemcpyHostToDevice, s));
        SWIFT_HIP(hipStreamSynchronize(s));  // desc_h may change on the next call
    }
    if (changed) *changed = ch;
    return SWIFT_OK;
}

static int batch_launch(swift_batch *b, hipStream_t s) {
    hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, s, b->counter);
    if (b->nitems) {
        if (b->buf)
            hipLaunchKernelGGL(k_batch_walk<true>, dim3(b->grid), dim3(256), 0, s, b->desc,
                               b->items, b->range, b->counter);
        else
            hipLaunchKernelGGL(k_batch_walk<false>, dim3(b->grid), dim3(256), 0, s, b->desc,
                               b->items, b->range, b->counter);
    }
    if (b->nfix + b->nres)
        hipLaunchKernelGGL(k_batch_tail, dim3(ceil_div((size_t)(b->nfix + b->nres) * 64, 256)),
                           dim3(256), 0, s, b->desc, b->fix_items, b->nfix, b->res_items, b->nres);
    SWIFT_CHECK_LAUNCH();
    return SWIFT_OK;
}

extern "C" int swift_batch_spmm(swift_batch *b, const float *const *d_B, float *const *d_C,
                                swift_stream_t stream) {
    if (!b) {
        set_error("swift_batch_spmm: null batch");
        return SWIFT_ERR_ARG;
    }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    SWIFT_TRY(batch_set_operands(b, d_B, d_C, s, nullptr));
    return batch_launch(b, s);
}

extern "C" int swift_batch_graph_launch(swift_batch *b, const float *const *d_B,
                                        float *const *d_C, swift_stream_t stream) {
    if (!b) {
        set_error("swift_batch_graph_launch: null batch");
        return SWIFT_ERR_ARG;
    }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    SWIFT_TRY(batch_set_operands(b, d_B, d_C, s, nullptr));
    if (!b->exec) {
        // capture on a private stream (capture is not allowed on the legacy
        // null stream), then replay on the caller's stream
        hipStream_t cs;
        SWIFT_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
        int rc = e == hipSuccess ? batch_launch(b, cs) : SWIFT_ERR_HIP;
        hipGraph_t g = nullptr;
        hipError_t e2 = hipStreamEndCapture(cs, &g);
        (void)hipStreamDestroy(cs);
        if (e != hipSuccess || e2 != hipSuccess || rc != SWIFT_OK) {
            if (g) (void)hipGraphDestroy(g);
            set_error("swift_batch_graph_launch: capture failed (%s / %s)", hipGetErrorString(e),
                      hipGetErrorString(e2));
            return rc != SWIFT_OK ? rc : SWIFT_ERR_HIP;
        }
        b->graph = g;
        SWIFT_HIP(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
    }
    SWIFT_HIP(hipGraphLaunch(b->exec, s));
    return SWIFT_OK;
}

extern "C" int swift_batch_set_lane_width(int mode) {
    if (mode != SWIFT_BATCH_WIDTH_TUNED && mode != SWIFT_BATCH_WIDTH_FULL &&
        mode != SWIFT_BATCH_WIDTH_HALF) {
        set_error("swift_batch_set_lane_width: mode must be SWIFT_BATCH_WIDTH_TUNED, _FULL or "
                  "_HALF");
        return SWIFT_ERR_ARG;
    }
    g_batch_width = mode;
    return SWIFT_OK;
}

extern "C" int swift_batch_params(const swift_batch *b, int *work_units, int *depth,
                                  int *block_threads) {
    if (!b) {
        set_error("swift_batch_params: null batch");
        return SWIFT_ERR_ARG;
    }
    if (work_units) *work_units = b->nitems;
    if (depth) {  // most common tuned segment size
        std::map<int, int> cnt;
        for (int w : b->W) cnt[w]++;
        int best = 0, bc = -1;
        for (auto &kv : cnt)
            if (kv.second > bc) { bc = kv.second; best = kv.first; }
        *depth = best;
    }
    if (block_threads) *block_threads = 256;
    return SWIFT_OK;
}

extern "C" void swift_batch_destroy(swift_batch *b) {
    if (b) (void)hipDeviceSynchronize();
    batch_free(b);
}

extern "C" int swift_autotune(swift_matrix *mat, int K, swift_dtype dtype, swift_layout layout,
                              const void *d_B, void *d_C, swift_stream_t stream, int *best_items) {
    if (!mat || K <= 0 || !d_B || !d_C) {
        set_error("swift_autotune: bad arguments");
        return SWIFT_ERR_ARG;
    }
    try {
        TuneChoice c;
        const int rc = autotune_impl(mat, K, dtype, layout, d_B, d_C,
                                     reinterpret_cast<hipStream_t>(stream), &c);
        if (rc == SWIFT_OK && best_items) *best_items = c.W;
        return rc;
    } catch (...) {
        set_error("swift_autotune: host allocation failed");
        return SWIFT_ERR_ALLOC;
    }
}
