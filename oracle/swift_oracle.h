/*
 * swift_oracle.h -- CPU oracle for the Swift SpMM path.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C restatement of the reference's
 * CPU algorithms (loader, CSR->CSC, column sort, staircase tiling, tile
 * categorisation, balance, row sort, Swift_cpu), which gives the expected
 * results of the tests and of the smoke entry point, and the CPU-baseline leg
 * of the benchmark. The product library never links, loads or calls it.
 *
 * Each function cites the reference file:line it restates.
 */
#ifndef SWIFT_ORACLE_H
#define SWIFT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* The full Swift format, host-resident, values held as double (float inputs
 * convert exactly).  Field names follow slide_matrix (src/format.h:6-39). */
typedef struct {
    int m, n, nnz;
    /* column sort */
    int *perm;            /* perm[i] = original column at sorted position i   */
    int *sortnnz;         /* sortnnz[i] = nnz of column perm[i]               */
    int *sortrowidx_tmp;  /* CSC concatenated in perm order                   */
    double *sortval_tmp;
    /* staircase tiling */
    int tilenum, regular_nnz, reside_col, reside_nnz;
    int *tile_ptr;        /* tilenum+1 */
    int *tile_len, *tile_colidx, *tile_tall;  /* tilenum */
    int *sortrowidx;      /* regular_nnz */
    double *sortval;
    int *reside_cscptr;   /* reside_col+1 */
    int *reside_cscrowidx;
    double *reside_val;
    /* categorisation */
    int *tile_format;     /* tilenum */
    int segsum;
    int *countrow;        /* tilenum+1 (scanned) */
    int *segmentoffset;   /* segsum+1 */
    int *sortrowindex;    /* regular_nnz */
    /* balance */
    int numblk;
    unsigned int *colorblk;
    int *blkstart, *blkstop, *blkptr;  /* numblk each */
    /* row sort (filled by oracle_row_idx_sort) */
    int *val_position, *val_offset, *rowsorted;  /* regular_nnz each */
} oracle_swift;

/* src/auxiliary.h:83-98: in-place exclusive scan of `length` ints. */
void oracle_exclusive_scan(int *a, int length);
/* src/auxiliary.h:100-113: in-place adjacent difference (a[0] kept). */
void oracle_exclusive_reduce(int *a, int length);

/* src/auxiliary.h:167-331 + src/mmio.h:90-256.  Returns 0, -1 (open),
 * -2 (banner), -4 (size line).  Arrays are malloc'd; free with oracle_free_ptr. */
int oracle_mmio_load(const char *path, int *m, int *n, int *nnz, int *is_sym,
                     int **rowptr, int **colidx, double **val);

/* src/coo2csc.h:11-79: stable counting transpose. */
void oracle_csr2csc(int m, int n, int nnz, const int *rowptr, const int *colidx,
                    const double *val, int *cscptr, int *cscrowidx,
                    double *cscval, int *nnzpercol);

/* Column sort (src/coo2csc.h:82-134), ties broken by original column index
 * ascending, which is the stable order of the reference's GPU ColSort
 * (thrust::sort_by_key, src/ColSort_cuda.h:108). */
void oracle_col_sort(int n, const int *nnzpercol, const int *cscptr,
                     const int *cscrowidx, const double *cscval, int *perm,
                     int *sortnnz, int *sortrowidx_tmp, double *sortval_tmp);

/* Whole build: csr2csc -> col_sort -> formattransation -> block_catgory ->
 * balance.  Returns a malloc'd struct (free with oracle_swift_free). */
oracle_swift *oracle_build(int m, int n, int nnz, const int *rowptr,
                           const int *colidx, const double *val);

/* Tiling + categorisation + balance from an already column-sorted CSC
 * (perm may be NULL).  Lets tests feed any column order. */
oracle_swift *oracle_build_from_sorted(int m, int n, int nnz, const int *perm,
                                       const int *sortnnz,
                                       const int *sortrowidx_tmp,
                                       const double *sortval_tmp);

/* src/back_process.h:23-229: per-8-tile group row sort, ties by original
 * lane position ascending. Fills val_position, val_offset, rowsorted. */
void oracle_row_idx_sort(oracle_swift *s);

void oracle_swift_free(oracle_swift *s);
void oracle_free_ptr(void *p);

/* src/Swift_cpu.h:2-78: C += regular tiles then reside columns, row-major
 * sorted B (n x K) and C (m x K), exact reference loop order. */
void oracle_swift_cpu_f64(const oracle_swift *s, int K, const double *sortB,
                          double *C);
void oracle_swift_cpu_f32(const oracle_swift *s, int K, const float *sortB,
                          float *C);
/* Same loops, with the reference timer (gettimeofday) around them; returns ms. */
double oracle_swift_cpu_f32_timed(const oracle_swift *s, int K,
                                  const float *sortB, float *C);

/* FP64 reference C = A*B over CSR, row-major B (n x K) / C (m x K), plus
 * absC = |A|*|B| used to normalise the per-element error. */
void oracle_spmm_csr_f64(int m, int n, const int *rowptr, const int *colidx,
                         const double *val, int K, const double *B, double *C,
                         double *absC);
/* FP64 dB = A^T * dC (n x K), dvals[e] = <dC[row_e,:], B[col_e,:]>. */
void oracle_spmm_t_csr_f64(int m, int n, const int *rowptr, const int *colidx,
                           const double *val, int K, const double *dC,
                           double *dB);
void oracle_sddmm_csr_f64(int m, int n, const int *rowptr, const int *colidx,
                          int K, const double *dC, const double *B,
                          double *dvals);

/* Empty 32x32 blocks of the CSR pattern, test/MTX_info/main.cu:143-176:
 * *blocks = ceil(m/32)*ceil(n/32), *empty = blocks holding no entry. */
void oracle_block_stats(int m, int n, const int *rowptr, const int *colidx,
                        long long *blocks, long long *empty);
void oracle_block_density(int m, int n, const int *rowptr, const int *colidx, int thr,
                          long long *dense, long long *dense_nnz, int *max_count);

#ifdef __cplusplus
}
#endif
#endif
