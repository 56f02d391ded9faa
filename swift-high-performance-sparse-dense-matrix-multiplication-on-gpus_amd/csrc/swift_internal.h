// swift_internal.h -- shared declarations of the Swift SpMM library (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <map>
#include <memory>
#include <mutex>
#include <string>

#include "swift_spmm.h"

namespace swift {

constexpr int TILE = 32;   // tile width (TILESIZE/slidesize of the reference)
constexpr int WAVE = 64;   // CDNA wavefront

// Last error text of the calling thread (swift_last_error()).
void set_error(const char *fmt, ...);

#define SWIFT_HIP(call)                                                       \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess) {                                               \
            ::swift::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call,     \
                               hipGetErrorString(e_));                        \
            return SWIFT_ERR_HIP;                                             \
        }                                                                     \
    } while (0)

#define SWIFT_TRY(call)                                                       \
    do {                                                                      \
        int s_ = (call);                                                      \
        if (s_ != SWIFT_OK) return s_;                                        \
    } while (0)

#define SWIFT_CHECK_LAUNCH() SWIFT_HIP(hipGetLastError())

inline size_t value_size(swift_dtype t) {
    switch (t) {
        case SWIFT_F32: return 4;
        case SWIFT_F64: return 8;
        case SWIFT_BF16: return 2;
        case SWIFT_F16: return 2;
    }
    return 0;
}

inline unsigned ceil_div(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

// Cut rows with more carry slots than this are fixed up by a 256-thread
// workgroup (8 groups of 32 threads at K = 128 FP32), the others by one wave.
constexpr int FIX_LONG = 32;

// Segment-table class of a product whose B rows are row_bytes long, of
// elements elem bytes wide.
constexpr int SEG_CLASSES = 5;
inline int seg_class(size_t row_bytes, size_t elem) {
    if (row_bytes >= 1024) return 0;
    if (row_bytes >= 512) return 1;
    if (row_bytes >= 256) return elem == 2 ? 2 : 3;
    return 4;
}

// Device workspace that grows on demand; owned by a matrix or a call site
// and released when it goes out of scope (hipFree synchronises the device,
// so work still queued on it finishes first).  Not copyable.
struct Buffer {
    void *ptr = nullptr;
    size_t bytes = 0;
    Buffer() = default;
    Buffer(const Buffer &) = delete;
    Buffer &operator=(const Buffer &) = delete;
    ~Buffer() { release(); }
    int alloc(size_t n);   // (re)allocates only when n > bytes
    void release();
    template <class T> T *as() const { return static_cast<T *>(ptr); }
};

// ---------------------------------------------------------------- scan/sort

// Exclusive scan of n ints; out may alias in.  If total != nullptr the sum is
// written there (device pointer).  tmp needs scan_temp_bytes(n).
size_t scan_temp_bytes(size_t n);
int exclusive_scan(const int *in, int *out, size_t n, int *total, void *tmp,
                   hipStream_t s);

// Stable LSD radix sort of (key, value) pairs on bits [0, bits).  If vals_in is
// null the values are the input positions 0..n-1.  Results land in
// keys_out/vals_out; the *_alt buffers (n each) are scratch.  tmp needs
// radix_temp_bytes(n).
size_t radix_temp_bytes(size_t n);
int radix_sort_pairs(const unsigned *keys_in, const unsigned *vals_in,
                     unsigned *keys_out, unsigned *vals_out, unsigned *keys_alt,
                     unsigned *vals_alt, size_t n, int bits, void *tmp,
                     hipStream_t s);

int bits_for(unsigned max_value);

// Carry slots of bytes for a product or backward walk on M issued on stream
// s: from swift_spmm_noalloc's scratch array when that call is in progress
// on this thread, else from M's workspace for s (grown on demand).
int carry_workspace(const swift_matrix *M, size_t bytes, hipStream_t s, void **out);

// Fill helpers (kernels, never hipMemsetAsync, so that graph capture of the
// SpMM stays valid).
int fill_i32(int *p, int v, size_t n, hipStream_t s);
int fill_bytes_zero(void *p, size_t bytes, hipStream_t s);
int iota_i32(int *p, size_t n, hipStream_t s);

}  // namespace swift

// ------------------------------------------------------------------ matrix
// Device-resident Swift format.  Field names follow the reference's
// slide_matrix (src/format.h:6-39) so every index array can be dumped and
// compared with the CPU restatement.
struct swift_matrix {
    int m = 0, n = 0, nnz = 0;
    swift_dtype vtype = SWIFT_F32;  // value type of the build (F32 or F64)
    int device = 0;

    // scalars
    int tilenum = 0, regular_nnz = 0, reside_col = 0, reside_nnz = 0;
    int segsum = 0, numblk = 0, nsteps = 0, max_col_nnz = 0;
    int has_rowsort = 0;

    // column sort (ColSort): perm[i] = original column at sorted position i
    int *perm = nullptr, *sortnnz = nullptr, *cptr = nullptr;  // cptr n+1
    int *sortrowidx_tmp = nullptr;
    void *sortval_tmp = nullptr;
    int *sortsrc_tmp = nullptr;  // CSR position of each sorted-CSC entry

    // staircase steps (our derived schedule data)
    int *step_col = nullptr, *step_base = nullptr, *step_tall = nullptr;
    int *step_tileoff = nullptr;  // nsteps+1

    // regular tiles
    int *tile_ptr = nullptr, *tile_len = nullptr, *tile_colidx = nullptr;
    int *tile_tall = nullptr, *tile_format = nullptr;
    int *sortrowidx = nullptr;
    void *sortval = nullptr;
    int *sortsrc = nullptr;          // CSR position of each regular entry

    // categorisation
    int *countrow = nullptr, *segmentoffset = nullptr, *sortrowindex = nullptr;

    // reside (irregular) CSC over reversed sorted columns
    int *reside_cscptr = nullptr, *reside_cscrowidx = nullptr;
    void *reside_val = nullptr;
    int *reside_src = nullptr;

    // balance
    unsigned *colorblk = nullptr;
    int *blkstart = nullptr, *blkstop = nullptr, *blkptr = nullptr;

    // row sort (Swift_GPU_4's row_idx_sort)
    int *val_position = nullptr, *val_offset = nullptr, *rowsorted = nullptr;
    int *rowsort_src = nullptr;      // position inside the group of sorted slot

    // CSR copy of the input pattern (kept for the transposed layout / SDDMM)
    int *csr_rowptr = nullptr, *csr_colidx = nullptr;

    // low-precision value copies (swift_prepare_values)
    void *sortval_lp = nullptr, *reside_val_lp = nullptr;
    swift_dtype lp_type = SWIFT_F32;
    float lp_scale = 1.0f;

    // column-tile work list for the tile-order (scatter) kernels
    int nwork = 0;
    int2 *work = nullptr;        // (first tile, tile count) per work item
    int work_depth = 0;          // max tiles per work item

    // Row view of the regular part: the regular tiles' entries regrouped by
    // output row (stable, so Swift order is kept inside a row), for the
    // output-stationary kernels.
    int *rv_ptr = nullptr;       // m+1
    int *rv_col = nullptr;       // B row (original column) of each entry
    void *rv_val = nullptr;      // value (vtype)
    int *rv_src = nullptr;       // CSR position of each entry
    int *rv_row = nullptr;       // output row of each entry
    // transposed layouts (swift_build_transpose, whose row view holds every
    // entry): the plan that moves dvals from walk order into the original
    // matrix's CSR order in two passes through LDS.  Chunks are ranges of
    // 2^dv_clog walk positions, buckets ranges of 2^dv_log CSR positions, and
    // bucket order lists the entries by bucket, in walk order inside one.
    //   dv_rank[e]: rank of walk position e inside its chunk, by (bucket, e)
    //   dv_dst[p]:  slot in bucket order of the chunk's p-th entry by rank
    //   dv_off[s]:  CSR position, inside its bucket, of bucket-order slot s
    // dv_ws holds the nnz walk-order values.
    unsigned short *dv_rank = nullptr, *dv_off = nullptr;
    int *dv_dst = nullptr;
    int dv_log = 0, dv_clog = 0;
    mutable swift::Buffer dv_ws;
    int *empty_rows = nullptr;   // rows with no regular entry (zeroed before a flat walk)
    int nempty = 0;
    void *rv_val_lp = nullptr;   // BF16/F16 copy (swift_prepare_values)
    int *inv_perm = nullptr;     // inv_perm[perm[i]] = i
    // The row view holds the regular entries, or with rv_all every entry
    // (regular, then reside, within each row).  rv_all is set for the
    // transposed layout of the backward pass: its walks then need no reside
    // launch, whose atomics into dB cost more than walking the few entries.
    // build_rowview also sets it for a matrix whose reside part is at most 1%
    // of its entries.
    bool rv_all = false;
    int rv_nnz = 0;
    // rows that receive reside entries: the fix-up of a cut row adds with
    // atomics where reside blocks of the same tail launch add into the row
    unsigned char *resrow_flag = nullptr;  // m bytes, 1 = row has reside entries
    // Merge-path split of (rows + regular entries) into segments of W items.
    // A boundary that falls inside a row cuts it into pieces: the piece
    // holding the row's start is stored into C like any row, every later
    // piece (it starts its segment) into the carry slot of the boundary it
    // starts at, and a fix-up pass adds a row's carry slots onto it in order.
    //   split_rows[b]: row cut by boundary b (boundaries in segment order, so
    //                  one row's boundaries are consecutive)
    //   seg_slot[s]:   boundary index of the end of segment s, where it cuts
    //                  a row (exclusive scan of the cut flags)
    //   fix[i]:        (row, first boundary, end boundary, 0) of the i-th cut
    //                  row, rows with the most boundaries first
    struct SegTable {
        int W = 0, nseg = 0, nsplit = 0, nfix = 0;
        int nlong = 0;               // cut rows of more than FIX_LONG slots (first in fix)
        bool shared = false;         // the arrays belong to a default table of the same W
        int2 *seg = nullptr;         // nseg+1 (row, entry) segment starts
        int *split_rows = nullptr;
        int *seg_slot = nullptr;
        int4 *fix = nullptr;
    };
    // the table of mp_items items (swift_set_segment_items, the autotuner)
    int mp_items = 0;
    SegTable prim;
    // Default tables by the bytes of one B row (K * element size) gathered
    // per entry (swift::seg_class) -- used unless the segment size was set
    // explicitly (user_items != 0).
    SegTable cls[swift::SEG_CLASSES];
    int user_items = 0;
    // Carry slots of the row-view walks (nsplit rows of K values), one
    // workspace per stream so that products on different streams do not
    // share one; grown on demand.
    mutable std::mutex carry_mu;
    mutable std::map<hipStream_t, std::unique_ptr<swift::Buffer>> carry_ws;
    // autotuner choice (swift_autotune, swift_batch_create) for products of
    // tuned_K columns with tuned_elem-byte B elements: the walk that
    // SWIFT_KERNEL_DEFAULT runs (the row-view walk or the window walk), the
    // row-view walk's slot rows per step (0: the rule) and the widest K slice
    // of one launch (0: no cap)
    int tuned_K = 0, tuned_elem = 0, tuned_slots = 0, tuned_cols = 0;
    swift_kernel tuned_kernel = SWIFT_KERNEL_DEFAULT;

    // build phase times (ms): column sort (csr->csc + ColSort), blocking
    // (staircase tiles, reside part, categorisation, balance), kernel
    // schedules (work list, row view, segment tables), total
    float build_ms[4] = {0, 0, 0, 0};

    mutable swift::Buffer scratch;  // reusable device workspace
};
