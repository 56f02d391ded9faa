/*
 * swift_spmm.h -- C ABI of the MI355X (gfx950) Swift SpMM library.
 *
 * The reference exposes its path as C++ header functions called in a fixed
 * order by src/main.cu (load .mtx -> csr_to_csc -> col sort ->
 * formattransation -> block_catgory -> balance -> GPU SpMM launchers).  This
 * header replaces that interface with plain C entry points: plain pointers and
 * sizes, explicit value type / layout / K, status-code returns and an optional
 * HIP stream.  Device pointers are named d_*; everything else is host memory.
 *
 * Value types: the matrix is built in F32 or F64 (the reference's float and
 * double forks).  SpMM additionally accepts BF16 and F16 storage of A's values
 * and of B with FP32 accumulation and an FP32 C.
 */
#ifndef SWIFT_SPMM_H
#define SWIFT_SPMM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t *swift_stream_t; /* == hipStream_t */

typedef enum {
    SWIFT_OK = 0,
    SWIFT_ERR_OPEN = -1,        /* mmio_allinone: fopen failed          */
    SWIFT_ERR_BANNER = -2,      /* mmio_allinone: bad banner            */
    SWIFT_ERR_DATA = -3,        /* .mtx entry lines malformed/truncated */
    SWIFT_ERR_SIZE = -4,        /* mmio_allinone: bad size line         */
    SWIFT_ERR_ARG = -10,        /* invalid argument                     */
    SWIFT_ERR_HIP = -11,        /* HIP runtime error (see last_error)   */
    SWIFT_ERR_ALLOC = -12,      /* device or host allocation failed     */
    SWIFT_ERR_UNSUPPORTED = -13,/* dtype/layout/K combination not built */
    SWIFT_ERR_VERIFY = -14      /* host launcher: result check failed   */
} swift_status;

typedef enum { SWIFT_F32 = 0, SWIFT_F64 = 1, SWIFT_BF16 = 2, SWIFT_F16 = 3 } swift_dtype;

/* Dense B (n x K) and C (m x K) storage order.  ROW: B[j*K + k];
 * COL: B[k*n + j] (the reference's "column-major", src/Swift.h:43-44). */
typedef enum { SWIFT_ROW_MAJOR = 0, SWIFT_COL_MAJOR = 1 } swift_layout;

/* Kernel family.  DEFAULT picks the tuned kernel for the layout. */
typedef enum {
    SWIFT_KERNEL_DEFAULT = 0,
    SWIFT_KERNEL_BALANCED = 1,   /* reside part per balance block (Swift_GPU / Swift_gpu2) */
    SWIFT_KERNEL_UNBALANCED = 2, /* reside part per column (Swift_gpu1)                    */
    SWIFT_KERNEL_KCHUNK = 3,     /* 2-D grid tiles x K-chunks (Swift_GPU_3)                */
    SWIFT_KERNEL_ROWSORT = 4,    /* same-row segmented reduction over row-sorted groups
                                    of 8 tiles (Swift_GPU_4); needs swift_row_idx_sort */
    SWIFT_KERNEL_ROWVIEW = 5,    /* row-major: regular part walked output-stationary
                                    through its row view (merge-path balanced), one
                                    row at a time per wave; reside part in balance
                                    blocks, or in the row view when it is at most 1%
                                    of the entries (swift_info.rowview_nnz = nnz)  */
    SWIFT_KERNEL_FLAT = 6,       /* row-major: row view walked by independent lane
                                    groups, several entries in flight whatever the
                                    row lengths; reside part as for ROWVIEW          */
    SWIFT_KERNEL_WINDOW = 7      /* row-major: ROWVIEW walk by workgroups of 16
                                    segments that first stage the B rows their
                                    entries reuse most (a column window) in LDS    */
} swift_kernel;

/* swift_spmm flags */
#define SWIFT_B_SORTED 0x1u   /* B rows (ROW) / columns (COL) already in sorted
                                 column order, i.e. the reference's sort_dense_mtx */
#define SWIFT_ACCUMULATE 0x2u /* C += A*B instead of C = A*B                     */

typedef struct swift_matrix swift_matrix; /* opaque, device resident */

/* Sizes and device pointers of every array of the Swift format.  Field names
 * follow slide_matrix (src/format.h:6-39); pointers are device pointers owned
 * by the handle. */
typedef struct {
    int m, n, nnz;
    swift_dtype vtype;
    int tilenum, regular_nnz, reside_col, reside_nnz, segsum, numblk, nsteps;
    int has_rowsort;
    const int *perm, *sortnnz, *sortrowidx_tmp;
    const void *sortval_tmp;
    const int *tile_ptr, *tile_len, *tile_colidx, *tile_tall, *tile_format;
    const int *sortrowidx;
    const void *sortval;
    const int *countrow, *segmentoffset, *sortrowindex;
    const int *reside_cscptr, *reside_cscrowidx;
    const void *reside_val;
    const unsigned int *colorblk;
    const int *blkstart, *blkstop, *blkptr;
    const int *val_position, *val_offset, *rowsorted;
    const int *sortsrc, *reside_src; /* CSR position of each entry (ours) */
    /* Build phase times in ms (host clock at the phase ends, where the build
     * synchronises; ours, for the preprocessing-cost report of
     * test/preprocess/main.cu:440-456): [0] column sort (csr->csc + ColSort),
     * [1] blocking (staircase tiles, reside part, block_catgory, balance),
     * [2] kernel schedules (work list, row view, segment tables), [3] total. */
    float build_ms[4];
    /* Entries in the row view that the row-view kernels walk (ours):
     * regular_nnz, or nnz when the reside part is walked with them. */
    int rowview_nnz;
    /* Schedule of SWIFT_KERNEL_DEFAULT (ours): segment_items = items per
     * merge-path segment set by swift_set_segment_items or the autotuner (0:
     * the default tables); for products of tuned_K columns (0: none tuned)
     * the autotuner's walk (SWIFT_KERNEL_DEFAULT = row-view walk,
     * SWIFT_KERNEL_WINDOW = window walk), slot rows per step (0: the
     * rule) and the widest K slice of one launch (0: the default, which cuts
     * B rows of 1 KiB or more of 4/8-byte elements into 512-byte slices; a
     * cap runs K as several slices of narrower lane groups). */
    int segment_items, tuned_K, tuned_slots;
    swift_kernel tuned_kernel;
    int tuned_cols;
} swift_info;

/* ------------------------------------------------------------------ host */

/* Text of the last error on the calling thread ("" if none). */
const char *swift_last_error(void);
/* Library version string. */
const char *swift_version(void);

/* Matrix Market loader.  Replaces mmio_allinone (src/auxiliary.h:167): .mtx
 * -> 0-based CSR, symmetric/hermitian mirrored, pattern -> 1.0, within-row
 * order = file order.  Arrays are malloc'd; release with swift_host_free.
 * Returns 0, -1 (open), -2 (banner), -3 (entries), -4 (size line). */
int swift_mmio_allinone(int *m, int *n, int *nnz, int *isSymmetric,
                        int **csrrowptr, int **csrcolidx, double **csrval,
                        const char *filename);
int swift_mmio_allinone_f32(int *m, int *n, int *nnz, int *isSymmetric,
                            int **csrrowptr, int **csrcolidx, float **csrval,
                            const char *filename);
/* The same loader with the value type as an argument and an explicit parser
 * thread count: 1 = the sequential fscanf reading of the reference; n > 1 =
 * n threads over the mapped file (identical arrays; a token the fast
 * conversion does not take whole falls back to the sequential reading);
 * 0 = automatic (up to 16 threads for files of 1 MiB or more).  No
 * counterpart in the reference, whose loader reads entry by entry with
 * fscanf (src/auxiliary.h:238-276). */
int swift_mmio_allinone_ex(int *m, int *n, int *nnz, int *isSymmetric,
                           int **csrrowptr, int **csrcolidx, void **csrval,
                           swift_dtype vtype, const char *filename, int nthreads);
void swift_host_free(void *p);

/* Seeded synthetic inputs of the benchmark configurations (host memory,
 * malloc'd, release with swift_host_free).  which:
 *   1 = 4096^2, ~80k nnz, Zipf(2.0) row lengths in [1,512], 30% of rows as
 *       8-wide column runs, values U[-1,1]            (k_hint 32)
 *   2 = 10^6^2, 16M nnz, even rows banded (bandwidth 64, 8..32 nnz), odd
 *       rows Zipf(1.8) up to 20k, values N(0,1)       (k_hint 128)
 *   4 = symmetric Chung-Lu graph, 200k nodes, mean degree 50, degree
 *       exponent 2.1, self loops, values D^-1/2 (A+I) D^-1/2 (k_hint 128)
 *   5 = batch item `index`: square, rows log-U[1k,64k], mean nnz/row
 *       log-U[4,128], Zipf(2.0) row lengths, N(0,1)   (k_hint from {32,64,128})
 * scale in (0,1] shrinks the row/column counts (1 = full size).  Values are
 * float64; k_hint (may be NULL) receives the configuration's K. */
int swift_generate(int which, int index, double scale, uint64_t seed, int *m,
                   int *n, int *nnz, int **rowptr, int **colidx, double **val,
                   int *k_hint);
/* Dense matrix of rows x cols, uniform in [lo,hi) (dist 0) or N(0,1) (dist 1). */
int swift_generate_dense(size_t count, int dist, double lo, double hi,
                         uint64_t seed, double *out);

/* ---------------------------------------------------- preprocessing (GPU) */

/* Stable CSR -> CSC transpose on the GPU.  Replaces csr_to_csc
 * (src/gpu_csr2csc.h:51) and nnzpercolSegmentReduce (:24): rows ascending
 * inside each column, nnzpercol[j] = colptr[j+1]-colptr[j].  Values are of
 * vtype (F32/F64); d_csc_val may be NULL to skip values.  The CSR is
 * checked as in swift_build (SWIFT_ERR_ARG on a bad row pointer or column). */
int swift_csr_to_csc(int m, int n, int nnz, swift_dtype vtype,
                     const void *d_csr_val, const int *d_csr_row_ptr,
                     const int *d_csr_col_ind, void *d_csc_val,
                     int *d_csc_row_ind, int *d_csc_col_ptr, int *d_nnzpercol,
                     swift_stream_t stream);

/* Stable column sort by nnz on the GPU.  Replaces ColSort
 * (src/ColSort_cuda.h:48) and col_sort (src/coo2csc.h:82): d_perm[i] =
 * column at sorted position i (ties by column index), d_sortnnz[i] = its nnz,
 * CSC columns concatenated in that order.  d_B/d_sortB (n x K row-major, the
 * dense operand, Pre4sortedx src/ColSort_cuda.h:9) may be NULL. */
int swift_col_sort(int n, int nnz, int K, swift_dtype vtype,
                   const int *d_nnzpercol, const int *d_csc_col_ptr,
                   const int *d_csc_row_ind, const void *d_csc_val,
                   int *d_sortrowidx_tmp, void *d_sortval_tmp, int *d_sortnnz,
                   int *d_perm, const void *d_B, void *d_sortB,
                   swift_stream_t stream);

/* Whole preprocessing on the GPU: csr_to_csc -> ColSort -> formattransation
 * (src/coo2csc.h:137) -> block_catgory (src/block_catgory.h:1) -> balance
 * (src/balance.h:1).  Input CSR is device resident (vtype F32 or F64);
 * the handle keeps copies of its row pointer and column indices.  One read
 * pass checks the input first: rowptr[0] = 0, rowptr[m] = nnz, rowptr
 * non-decreasing and every column in [0, n), else SWIFT_ERR_ARG (the
 * reference passes its loader output on unchecked, src/main.cu:79). */
int swift_build(swift_matrix **out, int m, int n, int nnz, swift_dtype vtype,
                const int *d_csr_row_ptr, const int *d_csr_col_ind,
                const void *d_csr_val, swift_stream_t stream);
/* Same, from host CSR (copies in, builds, synchronises). */
int swift_build_host(swift_matrix **out, int m, int n, int nnz,
                     swift_dtype vtype, const int *csr_row_ptr,
                     const int *csr_col_ind, const void *csr_val);
void swift_destroy(swift_matrix *mat);

/* Per-8-tile-group row sort on the GPU.  Replaces row_idx_sort
 * (src/back_process.h:23) and its bitonic GPU twin
 * (test/preprocess/back_process.h:56-116): fills val_position, val_offset and
 * the sorted row keys (kept beside, not over, sortrowidx). */
int swift_row_idx_sort(swift_matrix *mat, swift_stream_t stream);

/* Schedule knob of the row-view kernels: items (rows + entries) per
 * merge-path segment, i.e. per wave.  By default the size follows the bytes
 * of one B row (48 items for rows of 1 KiB or more, 64 from 512 B, from
 * 256 B 128 for 16-bit elements and 96 for others, 128 below), and is at
 * least three mean rows of the matrix; a call
 * here pins one size for every product of
 * the matrix.  Rebuilds the segment table on the GPU; the format
 * itself is unchanged.  The batch autotuner (swift_batch_create) and
 * swift_autotune set it per matrix.  No counterpart in the reference. */
int swift_set_segment_items(swift_matrix *mat, int items, swift_stream_t stream);

/* Sizes + device pointers of the format (valid until swift_destroy). */
int swift_get_info(const swift_matrix *mat, swift_info *info);

/* Matrix statistics of the reference's study drivers, computed on the GPU
 * from the built matrix: sparsity and the share of empty 32x32 blocks
 * (test/MTX_info/main.cu:143-176, :484), the sizes of Swift's regular (tile)
 * and irregular (reside) parts (test/irr_regular_ratio/main.cu:196-218, which
 * counts tiles of its own variant of the tiling; here nonzeros of the two
 * parts), the tile formats of block_catgory and the row/column extremes.
 * The block density figures (dense_blocks, dense_block_nnz, max_block_nnz)
 * have no counterpart in the reference: they say how much of A would fit
 * dense matrix-core tiles at least half full. */
typedef struct {
    double sparsity;            /* nnz / (m * n)                                */
    double empty_block_ratio;   /* 32x32 blocks without an entry / all blocks   */
    long long blocks, nonempty_blocks;
    int regular_nnz, reside_nnz;
    double regular_ratio;       /* regular_nnz / nnz                            */
    int tilenum;
    int tile_format_count[3];   /* tiles of format 0, 1, 2                      */
    int max_row_nnz, max_col_nnz;
    double mean_row_nnz;
    long long dense_blocks;     /* 32x32 blocks with >= 512 entries (half full)  */
    long long dense_block_nnz;  /* entries inside those blocks                   */
    int max_block_nnz;          /* entries of the fullest 32x32 block            */
} swift_stats;
int swift_matrix_stats(const swift_matrix *mat, swift_stats *stats, swift_stream_t stream);

/* Permute dense rows into sorted column order: out[i,:] = in[perm[i],:]
 * (ROW) or out[:,i] = in[:,perm[i]] (COL).  Pre4sortedx
 * (src/ColSort_cuda.h:9-21) as a standalone call. */
int swift_permute_dense(const swift_matrix *mat, int K, swift_layout layout,
                        swift_dtype dtype, const void *d_in, void *d_out,
                        swift_stream_t stream);

/* ------------------------------------------------------------- SpMM (GPU) */

/* C = A*B on the GPU (C += A*B with SWIFT_ACCUMULATE).  Replaces the
 * launchers Swift_gpu1 (src/Swift_gpu_col_major.h:171), Swift_gpu2
 * (src/Swift_gpu_row_major.h:136), Swift_GPU (src/Swift.h:184), Swift_GPU_3
 * (src/Swift_3.h:216) and Swift_GPU_4 (src/Swift_4.h:305), minus their
 * host copies: B (n x K) and C (m x K) are device resident in `layout`.
 * dtype F32/F64 must equal the build's value type; BF16/F16 need
 * swift_prepare_values first and take B in that type, C in F32.
 * Column-major products with SWIFT_KERNEL_DEFAULT transpose B and C through a
 * workspace held by the handle: such calls on one handle must not overlap in
 * time (issue them on one stream, or synchronise between streams).  The
 * row-view kernels (DEFAULT, ROWVIEW, WINDOW) keep the partial sums of rows
 * that their segments cut in carry slots, in a workspace the handle holds per
 * stream (allocated by the first such call on that stream, grown when a
 * larger K needs more): calls on different streams may overlap; for HIP graph
 * capture, make one call on the stream first or use swift_spmm_noalloc. */
int swift_spmm(const swift_matrix *mat, int K, swift_layout layout,
               swift_dtype dtype, swift_kernel kernel, unsigned flags,
               const void *d_B, void *d_C, swift_stream_t stream);

/* swift_spmm that allocates nothing.  No counterpart in the reference: it is
 * for callers that cannot call hipMalloc (code under HIP graph capture, tight
 * serving loops, harnesses that forbid allocation).  Arguments, results and
 * errors are those of swift_spmm.  Workspace -- the column-major
 * SWIFT_KERNEL_DEFAULT's transposes of B and C (round_up(n*K*size(B), 256) +
 * m*K*size(C) + 256 bytes) and the row-view kernels' carry slots (K values of
 * C's type per segment boundary that cuts a row, after the transposes, from a
 * 256-byte boundary) -- comes from one array of SWIFT_NOALLOC_SCRATCH_BYTES
 * per GPU that the library holds from its load until the process ends; a
 * product needing more returns SWIFT_ERR_ARG.  Every call on one GPU shares
 * that array: calls that overlap in time on one GPU (other streams or
 * threads) race. */
#define SWIFT_NOALLOC_SCRATCH_BYTES ((size_t)256 << 20)
int swift_spmm_noalloc(const swift_matrix *mat, int K, swift_layout layout,
                       swift_dtype dtype, swift_kernel kernel, unsigned flags,
                       const void *d_B, void *d_C, swift_stream_t stream);

/* swift_spmm with HIP events recorded on `stream` between its launches, the
 * way the reference times only its kernels (src/Swift.h:298-336, where C is
 * zeroed before the timed region): kernel_ms[0] = zeroing of C, [1] = regular
 * (tile) kernels, [2] = irregular (reside) kernels.  The row-view kernels
 * zero nothing ([0] is the gap between two events); their [2] is the tail
 * launch that adds the carry slots of cut rows onto C together with the
 * reside blocks, or only the former when the row view holds the reside
 * entries (swift_info.rowview_nnz = nnz) and the walk in [1] covers them.
 * Synchronises `stream`. */
int swift_spmm_timed(const swift_matrix *mat, int K, swift_layout layout,
                     swift_dtype dtype, swift_kernel kernel, unsigned flags,
                     const void *d_B, void *d_C, swift_stream_t stream,
                     float kernel_ms[3]);

/* Host-pointer launcher with the reference's semantics (e.g. src/Swift.h:184):
 * copies sorted B in, runs `kernel`, times the kernels with HIP events
 * (median of `reps` after `warmup`), copies C out, then counts elements with
 * |C - golden| > 1e-2 (ResultVerify, src/auxiliary.h:156); on any mismatch
 * *ms = -1 and SWIFT_ERR_VERIFY is returned.  golden may be NULL. */
int swift_spmm_host(const swift_matrix *mat, int K, swift_layout layout,
                    swift_dtype dtype, swift_kernel kernel, const void *h_sortB,
                    void *h_C, const void *h_golden, int warmup, int reps,
                    float *ms, double *gflops);

/* BF16/F16 storage of A's values (FP32 accumulation).  For F16, the values
 * are divided by one global power-of-two scale (returned by the B conversion
 * into `scales`).  Must precede swift_spmm with dtype BF16/F16. */
int swift_prepare_values(swift_matrix *mat, swift_dtype dtype,
                         swift_stream_t stream);
/* Convert dense FP32 B (count = rows*K elements, any layout) to BF16 or F16.
 * For F16 with d_col_scale != NULL: per-column (k) power-of-two scales are
 * computed so that max |B[:,k]|/scale <= 2^14; layout/K give the k index. */
int swift_convert_dense(const float *d_in, void *d_out, size_t rows, int K,
                        swift_layout layout, swift_dtype dtype,
                        float *d_col_scale, swift_stream_t stream);
/* F16 path with per-column scales: C[:,k] = scale[k] * (A_f16 * B_f16)[:,k]. */
int swift_spmm_scaled(const swift_matrix *mat, int K, swift_layout layout,
                      swift_kernel kernel, unsigned flags, const void *d_B,
                      const float *d_col_scale, float *d_C,
                      swift_stream_t stream);

/* -------------------------------------------------------- backward (GPU) */

/* Swift layout of A^T, built once on the GPU from A's own column-sorted CSC
 * (which is A^T's CSR); reused by every backward call. */
int swift_build_transpose(const swift_matrix *A, swift_matrix **AT,
                          swift_stream_t stream);
/* dvals[e] = <dC[row_e,:], B[col_e,:]> on A's pattern, e in A's CSR order
 * (SDDMM).  Row-major dC (m x K) and B (n x K), original column order;
 * dtype BF16 for an FP32 build: dC and B in BF16, FP32 dvals. */
int swift_sddmm(const swift_matrix *A, int K, swift_dtype dtype,
                const void *d_dC, const void *d_B, void *d_dvals,
                swift_stream_t stream);
/* Fused backward in one pass over the transposed Swift layout AT (from
 * swift_build_transpose): dB = A^T*dC (n x K) and dvals = SDDMM(dC, B) in
 * A's CSR order; every dC row is gathered once per nonzero and feeds both.
 * Either output may be NULL.  Row-major, original column order, F32/F64;
 * or dtype BF16 for a transpose of an FP32 build: dC, B and AT's values
 * (swift_prepare_values(AT, BF16) first) in BF16, with FP32 products, sums,
 * dB and dvals.
 * dvals are written in AT's walk order into a workspace of nnz values held
 * by AT, then gathered into CSR order: calls with dvals on one AT must not
 * overlap in time (issue them on one stream). */
int swift_backward_fused(const swift_matrix *AT, int K, swift_dtype dtype,
                         const void *d_dC, const void *d_B, void *d_dB,
                         void *d_dvals, swift_stream_t stream);
/* Scatter new CSR-order values into the built format (values change, pattern
 * fixed): used when A's values are trainable. */
int swift_update_values(swift_matrix *mat, const void *d_csr_val,
                        swift_stream_t stream);

/* -------------------------------------------------------- batched (GPU) */

typedef struct swift_batch swift_batch; /* opaque */

/* Group `count` built matrices (F32, K in {32, 64, 128, 256}) for one
 * persistent launch.  No counterpart in the reference, which runs one
 * matrix per process (test/k-32/FP32/main.cu:32-72) with two launches per
 * product (test/k-32/FP32/Swift_gpu_row_major.h:264,280); here one work list
 * of (matrix, segment) items serves every matrix.  The autotuner keys
 * matrices by (K, log2 mean nonzeros per row, reside share, whether the row
 * view holds the reside part) and times the matrices of each key as one
 * batch over segment sizes {64, 128, 256, 512} x two slot-row counts per step
 * (the rule and 8 for K <= 64, 4 above), and the lane shape of half the width
 * (K = 64..256 walked as two halves of K/2 columns, each half's segments one
 * after the other), and applies the fastest to all of that key (cached per
 * process).  The batch holds the carry slots of every matrix's cut rows.
 * Operand pointers and 16-byte alignment are checked at launch. */
int swift_batch_create(swift_batch **out, swift_matrix *const *mats,
                       const int *Ks, int count, swift_stream_t stream);
/* Lane widths the batch autotuner times for batches created afterwards (a
 * process-wide setting, part of the autotuner's cache key; no counterpart
 * in the reference).  TUNED: all of K per pass and half of it, the faster
 * per key (the default); FULL: all of K; HALF: K >= 64 always in two
 * halves. */
#define SWIFT_BATCH_WIDTH_TUNED 0
#define SWIFT_BATCH_WIDTH_FULL 1
#define SWIFT_BATCH_WIDTH_HALF 2
int swift_batch_set_lane_width(int mode);
/* All C_i = A_i * B_i (row-major, original column order, F32) in one
 * persistent kernel; d_B[i], d_C[i] are device pointers. */
int swift_batch_spmm(swift_batch *batch, const float *const *d_B,
                     float *const *d_C, swift_stream_t stream);
/* Capture swift_batch_spmm into a HIP graph once, then replay it. */
int swift_batch_graph_launch(swift_batch *batch, const float *const *d_B,
                             float *const *d_C, swift_stream_t stream);
/* Tuned parameters for reporting: work items of the persistent kernel, the
 * most common segment size, threads per block. */
int swift_batch_params(const swift_batch *batch, int *work_units,
                       int *depth, int *block_threads);
void swift_batch_destroy(swift_batch *batch);

/* Time the SWIFT_KERNEL_DEFAULT product of one matrix on the caller's
 * operands (C is overwritten) for every candidate schedule and keep the
 * fastest for products of this K and dtype: segment sizes {default, 64,
 * 128, 256, 512} (default: the table of the B-row bytes) x the slot rows
 * per step the product's lane shape can take (4, 8, 16), the lane shape of
 * half the width (K in two slices) and, where the default slices K, the
 * full width, and the window walk (row-major).  best_items receives the
 * segment size (0: the default tables); swift_info reports the rest.  Has no counterpart in the reference,
 * whose tiling constants are fixed at compile time
 * (test/k-32/FP32/main.cu:21-23). */
int swift_autotune(swift_matrix *mat, int K, swift_dtype dtype,
                   swift_layout layout, const void *d_B, void *d_C,
                   swift_stream_t stream, int *best_items);

#ifdef __cplusplus
}
#endif
#endif /* SWIFT_SPMM_H */
