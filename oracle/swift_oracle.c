/*
 * swift_oracle.c -- CPU oracle for the Swift SpMM path.
 *
 * TEST INFRASTRUCTURE ONLY: the expected results of tests/ and of the smoke
 * entry point, and the CPU-baseline leg of bench.py.  The product library never links or calls
 * this file.  Every function restates a reference CPU routine and cites it.
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off, so float sums are not
 * contracted into FMAs and the Swift_cpu restatement stays bitwise equal to the
 * reference loops compiled the same way).
 */
#include "swift_oracle.h"

#include <ctype.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#define TILE 32 /* TILESIZE / slidesize, src/auxiliary.h:7, src/format.h:77 */

static void *xcalloc(size_t n, size_t sz) {
    void *p = calloc(n ? n : 1, sz);
    if (!p) {
        fprintf(stderr, "oracle: out of memory\n");
        abort();
    }
    return p;
}

void oracle_free_ptr(void *p) { free(p); }

/* src/auxiliary.h:83-98 */
void oracle_exclusive_scan(int *a, int length) {
    if (length <= 1) return; /* the reference leaves a 1-element array as is */
    int run = 0;
    for (int i = 0; i < length; i++) {
        int v = a[i];
        a[i] = run;
        run += v;
    }
}

/* src/auxiliary.h:100-113 */
void oracle_exclusive_reduce(int *a, int length) {
    if (length <= 1) return;
    int prev = a[0];
    for (int i = 1; i < length; i++) {
        int v = a[i];
        a[i] = v - prev;
        prev = v;
    }
}

/* ---------------------------------------------------------------- loader */

static void lower(char *s) {
    for (; *s; s++) *s = (char)tolower((unsigned char)*s);
}

/* Banner: src/mmio.h:178-256.  Returns 0 and fills the 4 type letters. */
static int read_banner(FILE *f, char code[4]) {
    char line[1025], banner[64], mtx[64], crd[64], dtype[64], storage[64];
    if (!fgets(line, sizeof line, f)) return -1;
    if (sscanf(line, "%63s %63s %63s %63s %63s", banner, mtx, crd, dtype,
               storage) != 5)
        return -1;
    lower(mtx); lower(crd); lower(dtype); lower(storage);
    if (strncmp(banner, "%%MatrixMarket", 14) != 0) return -1;
    if (strcmp(mtx, "matrix") != 0) return -1;
    if (strcmp(crd, "coordinate") == 0) code[1] = 'C';
    else if (strcmp(crd, "array") == 0) code[1] = 'A';
    else return -1;
    if (strcmp(dtype, "real") == 0) code[2] = 'R';
    else if (strcmp(dtype, "complex") == 0) code[2] = 'C';
    else if (strcmp(dtype, "pattern") == 0) code[2] = 'P';
    else if (strcmp(dtype, "integer") == 0) code[2] = 'I';
    else return -1;
    if (strcmp(storage, "general") == 0) code[3] = 'G';
    else if (strcmp(storage, "symmetric") == 0) code[3] = 'S';
    else if (strcmp(storage, "hermitian") == 0) code[3] = 'H';
    else if (strcmp(storage, "skew-symmetric") == 0) code[3] = 'K';
    else return -1;
    code[0] = 'M';
    return 0;
}

/* Size line: src/mmio.h:147-172 (skip '%' lines, then three ints). */
static int read_size(FILE *f, int *m, int *n, int *nz) {
    char line[1025];
    *m = *n = *nz = 0;
    do {
        if (!fgets(line, sizeof line, f)) return -1;
    } while (line[0] == '%');
    if (sscanf(line, "%d %d %d", m, n, nz) == 3) return 0;
    int got;
    do {
        got = fscanf(f, "%d %d %d", m, n, nz);
        if (got == EOF) return -1;
    } while (got != 3);
    return 0;
}

/* src/auxiliary.h:167-331: COO in file order -> CSR, within-row order = the
 * order entries reach the row; symmetric/hermitian off-diagonal entries are
 * mirrored immediately after the original; pattern -> 1.0; complex -> real. */
int oracle_mmio_load(const char *path, int *m, int *n, int *nnz, int *is_sym,
                     int **rowptr, int **colidx, double **val) {
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    char code[4] = {' ', ' ', ' ', 'G'};
    if (read_banner(f, code) != 0) { fclose(f); return -2; }
    int M, N, NZ;
    if (read_size(f, &M, &N, &NZ) != 0) { fclose(f); return -4; }
    int sym = (code[3] == 'S' || code[3] == 'H');
    int *ri = (int *)xcalloc((size_t)NZ, sizeof(int));
    int *ci = (int *)xcalloc((size_t)NZ, sizeof(int));
    double *vv = (double *)xcalloc((size_t)NZ, sizeof(double));
    int *cnt = (int *)xcalloc((size_t)M + 1, sizeof(int));
    for (int e = 0; e < NZ; e++) {
        int i = 0, j = 0, iv = 0;
        double x = 0.0, xi = 0.0;
        if (code[2] == 'R') { if (fscanf(f, "%d %d %lg\n", &i, &j, &x) != 3) break; }
        else if (code[2] == 'C') { if (fscanf(f, "%d %d %lg %lg\n", &i, &j, &x, &xi) != 4) break; }
        else if (code[2] == 'I') { if (fscanf(f, "%d %d %d\n", &i, &j, &iv) != 3) break; x = iv; }
        else { if (fscanf(f, "%d %d\n", &i, &j) != 2) break; x = 1.0; }
        ri[e] = i - 1;
        ci[e] = j - 1;
        vv[e] = x;
        cnt[i - 1]++;
    }
    fclose(f);
    if (sym)
        for (int e = 0; e < NZ; e++)
            if (ri[e] != ci[e]) cnt[ci[e]]++;
    oracle_exclusive_scan(cnt, M + 1);
    int total = cnt[M];
    int *rp = (int *)xcalloc((size_t)M + 1, sizeof(int));
    int *cc = (int *)xcalloc((size_t)total, sizeof(int));
    double *cv = (double *)xcalloc((size_t)total, sizeof(double));
    memcpy(rp, cnt, sizeof(int) * ((size_t)M + 1));
    int *fill = (int *)xcalloc((size_t)M + 1, sizeof(int));
    for (int e = 0; e < NZ; e++) {
        int r = ri[e], c = ci[e];
        int p = rp[r] + fill[r]++;
        cc[p] = c;
        cv[p] = vv[e];
        if (sym && r != c) {
            p = rp[c] + fill[c]++;
            cc[p] = r;
            cv[p] = vv[e];
        }
    }
    free(fill); free(cnt); free(ri); free(ci); free(vv);
    *m = M; *n = N; *nnz = total; *is_sym = sym;
    *rowptr = rp; *colidx = cc; *val = cv;
    return 0;
}

/* ------------------------------------------------------------- csr2csc */

/* src/coo2csc.h:11-79: count per column, scan, then sweep rows in order so
 * each column receives its rows ascending (CSR order inside a row). */
void oracle_csr2csc(int m, int n, int nnz, const int *rowptr, const int *colidx,
                    const double *val, int *cscptr, int *cscrowidx,
                    double *cscval, int *nnzpercol) {
    (void)nnz;
    memset(nnzpercol, 0, sizeof(int) * (size_t)n);
    for (int e = 0; e < rowptr[m]; e++) nnzpercol[colidx[e]]++;
    cscptr[0] = 0;
    for (int j = 0; j < n; j++) cscptr[j + 1] = cscptr[j] + nnzpercol[j];
    int *next = (int *)xcalloc((size_t)n, sizeof(int));
    memcpy(next, cscptr, sizeof(int) * (size_t)n);
    for (int i = 0; i < m; i++)
        for (int e = rowptr[i]; e < rowptr[i + 1]; e++) {
            int p = next[colidx[e]]++;
            cscrowidx[p] = i;
            cscval[p] = val[e];
        }
    free(next);
}

/* ------------------------------------------------------------ col sort */

/* src/coo2csc.h:82-134 with the stable tie rule of src/ColSort_cuda.h:108:
 * counting sort of columns by nnz, equal counts keep column order. */
void oracle_col_sort(int n, const int *nnzpercol, const int *cscptr,
                     const int *cscrowidx, const double *cscval, int *perm,
                     int *sortnnz, int *sortrowidx_tmp, double *sortval_tmp) {
    int maxc = 0;
    for (int j = 0; j < n; j++)
        if (nnzpercol[j] > maxc) maxc = nnzpercol[j];
    int *start = (int *)xcalloc((size_t)maxc + 2, sizeof(int));
    for (int j = 0; j < n; j++) start[nnzpercol[j] + 1]++;
    for (int v = 0; v <= maxc; v++) start[v + 1] += start[v];
    for (int j = 0; j < n; j++) perm[start[nnzpercol[j]]++] = j;
    free(start);
    int w = 0;
    for (int i = 0; i < n; i++) {
        int j = perm[i];
        sortnnz[i] = nnzpercol[j];
        for (int p = cscptr[j]; p < cscptr[j + 1]; p++, w++) {
            sortrowidx_tmp[w] = cscrowidx[p];
            sortval_tmp[w] = cscval[p];
        }
    }
}

/* -------------------------------------------------------- tiling etc. */

/* Staircase tiling, src/coo2csc.h:137-435.  A "step" is a sorted column c
 * whose nnz exceeds its left neighbour's (c = 0 counts if nnz > 0); it adds
 * tall = sortnnz[c] - sortnnz[c-1] new depths starting at base = sortnnz[c-1].
 * Each step emits floor((N-c)/32) column blocks x tall depths of 32-wide
 * tiles (order: step, block, depth), and its (N-c)%32 leftover columns at the
 * same depths go to the reversed-column "reside" CSC. */
static void formattransation(oracle_swift *s) {
    const int N = s->n;
    const int *sn = s->sortnnz;
    int *cptr = (int *)xcalloc((size_t)N + 1, sizeof(int));
    for (int c = 0; c < N; c++) cptr[c + 1] = cptr[c] + sn[c];
    int *scol = (int *)xcalloc((size_t)N, sizeof(int));
    int *sbase = (int *)xcalloc((size_t)N, sizeof(int));
    int *stall = (int *)xcalloc((size_t)N, sizeof(int));
    int nsteps = 0;
    for (int c = 0; c < N; c++) {
        int base = c ? sn[c - 1] : 0;
        int d = sn[c] - base;
        if (d != 0) {
            scol[nsteps] = c;
            sbase[nsteps] = base;
            stall[nsteps] = d;
            nsteps++;
        }
    }
    long long tilenum = 0;
    int reside_col = 0;
    long long reside_nnz = 0;
    for (int i = 0; i < nsteps; i++) {
        tilenum += (long long)((N - scol[i]) / TILE) * stall[i];
        int w = (N - scol[i]) % TILE;
        if (w > reside_col) reside_col = w;
        reside_nnz += (long long)w * stall[i];
    }
    s->tilenum = (int)tilenum;
    s->regular_nnz = (int)(tilenum * TILE);
    s->reside_col = reside_col;
    s->reside_nnz = (int)reside_nnz;
    s->tile_ptr = (int *)xcalloc((size_t)tilenum + 1, sizeof(int));
    s->tile_len = (int *)xcalloc((size_t)tilenum, sizeof(int));
    s->tile_colidx = (int *)xcalloc((size_t)tilenum, sizeof(int));
    s->tile_tall = (int *)xcalloc((size_t)tilenum, sizeof(int));
    s->sortrowidx = (int *)xcalloc((size_t)s->regular_nnz, sizeof(int));
    s->sortval = (double *)xcalloc((size_t)s->regular_nnz, sizeof(double));
    long long t = 0;
    for (int i = 0; i < nsteps; i++) {
        int c = scol[i];
        int nblk = (N - c) / TILE;
        for (int j = 0; j < nblk; j++)
            for (int jx = 0; jx < stall[i]; jx++, t++) {
                int c0 = c + TILE * j;
                s->tile_ptr[t] = (int)(TILE * t);
                s->tile_len[t] = TILE;
                s->tile_colidx[t] = c0;
                s->tile_tall[t] = stall[i];
                for (int l = 0; l < TILE; l++) {
                    int src = cptr[c0 + l] + sbase[i] + jx;
                    s->sortrowidx[TILE * t + l] = s->sortrowidx_tmp[src];
                    s->sortval[TILE * t + l] = s->sortval_tmp[src];
                }
            }
    }
    s->tile_ptr[tilenum] = (int)(TILE * tilenum);

    /* Reside: column r holds sorted column N-1-r; entries ordered by
     * (step, depth) ascending. */
    s->reside_cscptr = (int *)xcalloc((size_t)reside_col + 1, sizeof(int));
    s->reside_cscrowidx = (int *)xcalloc((size_t)reside_nnz, sizeof(int));
    s->reside_val = (double *)xcalloc((size_t)reside_nnz, sizeof(double));
    for (int i = 0; i < nsteps; i++) {
        int w = (N - scol[i]) % TILE;
        for (int r = 0; r < w; r++) s->reside_cscptr[r + 1] += stall[i];
    }
    for (int r = 0; r < reside_col; r++)
        s->reside_cscptr[r + 1] += s->reside_cscptr[r];
    int *fill = (int *)xcalloc((size_t)reside_col + 1, sizeof(int));
    for (int i = 0; i < nsteps; i++) {
        int w = (N - scol[i]) % TILE;
        for (int jx = 0; jx < stall[i]; jx++)
            for (int jj = N - w; jj < N; jj++) {
                int r = N - 1 - jj;
                int p = s->reside_cscptr[r] + fill[r]++;
                int src = cptr[jj] + sbase[i] + jx;
                s->reside_cscrowidx[p] = s->sortrowidx_tmp[src];
                s->reside_val[p] = s->sortval_tmp[src];
            }
    }
    free(fill); free(cptr); free(scol); free(sbase); free(stall);
}

/* src/block_catgory.h:1-160.  Format 0: all 32 rows equal; 2: between 1 and
 * 15 row changes; otherwise 1.  countrow = row-runs per tile (scanned),
 * segmentoffset[run] = run length - 1, sortrowindex[run start] = same. */
static void block_catgory(oracle_swift *s) {
    int T = s->tilenum;
    s->tile_format = (int *)xcalloc((size_t)T, sizeof(int));
    s->countrow = (int *)xcalloc((size_t)T + 1, sizeof(int));
    for (int t = 0; t < T; t++) {
        const int *r = s->sortrowidx + s->tile_ptr[t];
        int len = s->tile_ptr[t + 1] - s->tile_ptr[t];
        int changes = 0;
        for (int l = 1; l < len; l++) changes += (r[l] != r[l - 1]);
        s->tile_format[t] = changes == 0 ? 0 : (changes < 16 ? 2 : 1);
        s->countrow[t] = changes + 1;
    }
    oracle_exclusive_scan(s->countrow, T + 1);
    s->segsum = s->countrow[T];
    s->segmentoffset = (int *)xcalloc((size_t)s->segsum + 1, sizeof(int));
    s->sortrowindex = (int *)xcalloc((size_t)s->regular_nnz, sizeof(int));
    for (int t = 0; t < T; t++) {
        int b = s->tile_ptr[t], e = s->tile_ptr[t + 1];
        int run = s->countrow[t];
        int head = b;
        for (int p = b + 1; p <= e; p++) {
            if (p == e || s->sortrowidx[p] != s->sortrowidx[p - 1]) {
                s->segmentoffset[run] = p - head - 1;
                s->sortrowindex[head] = p - head - 1;
                run++;
                head = p;
            }
        }
    }
}

/* src/balance.h:1-69 over the reside CSC: columns with <= 32 entries are one
 * block (colorblk = r, blkstart/blkstop left 0); longer ones split into
 * nb = ceil(n/32) blocks of len = ceil(n/nb), flagged with bit 31. */
static void balance(oracle_swift *s) {
    int R = s->reside_col;
    const int *ptr = s->reside_cscptr;
    int nblk = 0;
    for (int r = 0; r < R; r++) {
        int len = ptr[r + 1] - ptr[r];
        nblk += len <= TILE ? 1 : (len + TILE - 1) / TILE;
    }
    s->numblk = nblk;
    s->colorblk = (unsigned int *)xcalloc((size_t)nblk, sizeof(unsigned int));
    s->blkstart = (int *)xcalloc((size_t)nblk, sizeof(int));
    s->blkstop = (int *)xcalloc((size_t)nblk, sizeof(int));
    s->blkptr = (int *)xcalloc((size_t)nblk, sizeof(int));
    int b = 0;
    for (int r = 0; r < R; r++) {
        int len = ptr[r + 1] - ptr[r];
        s->blkptr[r] = b;
        if (len <= TILE) {
            s->colorblk[b++] = (unsigned int)r;
        } else {
            int nb = (len + TILE - 1) / TILE;
            int bl = (len + nb - 1) / nb;
            for (int j = 0; j < nb; j++, b++) {
                s->colorblk[b] = (unsigned int)r | 0x80000000u;
                s->blkstart[b] = ptr[r] + j * bl;
                s->blkstop[b] = (j == nb - 1) ? ptr[r] + len : ptr[r] + (j + 1) * bl;
            }
        }
    }
}

oracle_swift *oracle_build_from_sorted(int m, int n, int nnz, const int *perm,
                                       const int *sortnnz,
                                       const int *sortrowidx_tmp,
                                       const double *sortval_tmp) {
    oracle_swift *s = (oracle_swift *)xcalloc(1, sizeof(oracle_swift));
    s->m = m; s->n = n; s->nnz = nnz;
    s->perm = (int *)xcalloc((size_t)n, sizeof(int));
    s->sortnnz = (int *)xcalloc((size_t)n, sizeof(int));
    s->sortrowidx_tmp = (int *)xcalloc((size_t)nnz, sizeof(int));
    s->sortval_tmp = (double *)xcalloc((size_t)nnz, sizeof(double));
    if (perm) memcpy(s->perm, perm, sizeof(int) * (size_t)n);
    memcpy(s->sortnnz, sortnnz, sizeof(int) * (size_t)n);
    memcpy(s->sortrowidx_tmp, sortrowidx_tmp, sizeof(int) * (size_t)nnz);
    memcpy(s->sortval_tmp, sortval_tmp, sizeof(double) * (size_t)nnz);
    formattransation(s);
    block_catgory(s);
    balance(s);
    return s;
}

oracle_swift *oracle_build(int m, int n, int nnz, const int *rowptr,
                           const int *colidx, const double *val) {
    int *cscptr = (int *)xcalloc((size_t)n + 1, sizeof(int));
    int *cscrow = (int *)xcalloc((size_t)nnz, sizeof(int));
    double *cscval = (double *)xcalloc((size_t)nnz, sizeof(double));
    int *npc = (int *)xcalloc((size_t)n, sizeof(int));
    oracle_csr2csc(m, n, nnz, rowptr, colidx, val, cscptr, cscrow, cscval, npc);
    int *perm = (int *)xcalloc((size_t)n, sizeof(int));
    int *sn = (int *)xcalloc((size_t)n, sizeof(int));
    int *sr = (int *)xcalloc((size_t)nnz, sizeof(int));
    double *sv = (double *)xcalloc((size_t)nnz, sizeof(double));
    oracle_col_sort(n, npc, cscptr, cscrow, cscval, perm, sn, sr, sv);
    oracle_swift *s = oracle_build_from_sorted(m, n, nnz, perm, sn, sr, sv);
    free(cscptr); free(cscrow); free(cscval); free(npc);
    free(perm); free(sn); free(sr); free(sv);
    return s;
}

/* ------------------------------------------------------------ row sort */

typedef struct { int key, pos; } kp;
static int kp_cmp(const void *a, const void *b) {
    const kp *x = (const kp *)a, *y = (const kp *)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->pos < y->pos ? -1 : (x->pos > y->pos);
}

/* val_offset marking, src/back_process.h:128-205: run head = run length - 1,
 * run member = 0, singleton = -1.  `full` selects the full-group loop; the
 * remainder-group loop (src/back_process.h:172-205) lacks the branch that
 * marks a final singleton after a multi-entry run, so that entry stays 0. */
static void mark_offsets(const int *k, int *off, int len, int full) {
    int count = 0, cmp = k[0];
    for (int j = 1; j < len; j++) {
        int v = k[j], last = (j == len - 1);
        if (v == cmp) {
            count++;
            off[j - count]++;
        } else if (count != 0 && !last) {
            cmp = v; count = 0;
        } else if (count == 0 && !last) {
            cmp = v; off[j - 1] = -1;
        } else if (count == 0 && last) {
            cmp = v; off[j - 1] = -1; off[j] = -1;
        } else if (full) { /* count != 0 && last */
            cmp = v; count = 0; off[j] = -1;
        }
    }
}

/* src/back_process.h:23-229.  Groups of 8 tiles (256 entries; the last group
 * holds the tilenum%8 leftover tiles); rows sorted ascending with ties by
 * lane position (the reference's std::sort leaves tie order unspecified). */
void oracle_row_idx_sort(oracle_swift *s) {
    int R = s->regular_nnz;
    free(s->val_position); free(s->val_offset); free(s->rowsorted);
    s->val_position = (int *)xcalloc((size_t)R, sizeof(int));
    s->val_offset = (int *)xcalloc((size_t)R, sizeof(int));
    s->rowsorted = (int *)xcalloc((size_t)R, sizeof(int));
    kp *tmp = (kp *)xcalloc(256, sizeof(kp));
    int T = s->tilenum;
    for (int g = 0; g * 8 < T; g++) {
        int t0 = g * 8, t1 = t0 + 8 < T ? t0 + 8 : T;
        int b = s->tile_ptr[t0], len = s->tile_ptr[t1] - b;
        for (int i = 0; i < len; i++) {
            tmp[i].key = s->sortrowidx[b + i];
            tmp[i].pos = i;
        }
        qsort(tmp, (size_t)len, sizeof(kp), kp_cmp);
        for (int i = 0; i < len; i++) {
            s->rowsorted[b + i] = tmp[i].key;
            s->val_position[b + tmp[i].pos] = i;
        }
        mark_offsets(s->rowsorted + b, s->val_offset + b, len, t1 - t0 == 8);
    }
    free(tmp);
}

void oracle_swift_free(oracle_swift *s) {
    if (!s) return;
    void *ptrs[] = {s->perm, s->sortnnz, s->sortrowidx_tmp, s->sortval_tmp,
                    s->tile_ptr, s->tile_len, s->tile_colidx, s->tile_tall,
                    s->sortrowidx, s->sortval, s->reside_cscptr,
                    s->reside_cscrowidx, s->reside_val, s->tile_format,
                    s->countrow, s->segmentoffset, s->sortrowindex, s->colorblk,
                    s->blkstart, s->blkstop, s->blkptr, s->val_position,
                    s->val_offset, s->rowsorted};
    for (size_t i = 0; i < sizeof ptrs / sizeof ptrs[0]; i++) free(ptrs[i]);
    free(s);
}

/* ------------------------------------------------------------ Swift_cpu */

/* src/Swift_cpu.h:31-65, value type double. */
void oracle_swift_cpu_f64(const oracle_swift *s, int K, const double *B,
                          double *C) {
    for (int t = 0; t < s->tilenum; t++) {
        int b = s->tile_ptr[t], e = s->tile_ptr[t + 1], len = s->tile_len[t];
        for (int p = b; p < e; p++) {
            int col = s->tile_colidx[t] + (p - b) % len;
            int row = s->sortrowidx[p];
            double v = s->sortval[p];
            for (int k = 0; k < K; k++)
                C[(size_t)row * K + k] += v * B[(size_t)col * K + k];
        }
    }
    for (int r = 0; r < s->reside_col; r++) {
        int col = s->n - r - 1;
        for (int p = s->reside_cscptr[r]; p < s->reside_cscptr[r + 1]; p++) {
            int row = s->reside_cscrowidx[p];
            double v = s->reside_val[p];
            for (int k = 0; k < K; k++)
                C[(size_t)row * K + k] += v * B[(size_t)col * K + k];
        }
    }
}

/* test/k-32/FP32/Swift_cpu.h: the same loops with value type float. */
void oracle_swift_cpu_f32(const oracle_swift *s, int K, const float *B,
                          float *C) {
    for (int t = 0; t < s->tilenum; t++) {
        int b = s->tile_ptr[t], e = s->tile_ptr[t + 1], len = s->tile_len[t];
        for (int p = b; p < e; p++) {
            int col = s->tile_colidx[t] + (p - b) % len;
            int row = s->sortrowidx[p];
            float v = (float)s->sortval[p];
            for (int k = 0; k < K; k++)
                C[(size_t)row * K + k] += v * B[(size_t)col * K + k];
        }
    }
    for (int r = 0; r < s->reside_col; r++) {
        int col = s->n - r - 1;
        for (int p = s->reside_cscptr[r]; p < s->reside_cscptr[r + 1]; p++) {
            int row = s->reside_cscrowidx[p];
            float v = (float)s->reside_val[p];
            for (int k = 0; k < K; k++)
                C[(size_t)row * K + k] += v * B[(size_t)col * K + k];
        }
    }
}

double oracle_swift_cpu_f32_timed(const oracle_swift *s, int K,
                                  const float *B, float *C) {
    struct timeval t1, t2;
    gettimeofday(&t1, NULL);
    oracle_swift_cpu_f32(s, K, B, C);
    gettimeofday(&t2, NULL);
    return (t2.tv_sec - t1.tv_sec) * 1000.0 + (t2.tv_usec - t1.tv_usec) / 1000.0;
}

/* ------------------------------------------------------ FP64 references */

/* The naive CSR golden of src/main.cu:145-155, in double, plus |A||B|.
 * Rows are independent and each row sums its entries in CSR order, so the
 * result is the same bits whatever the number of threads. */
void oracle_spmm_csr_f64(int m, int n, const int *rowptr, const int *colidx,
                         const double *val, int K, const double *B, double *C,
                         double *absC) {
    (void)n;
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < m; i++) {
        double *c = C + (size_t)i * K;
        double *a = absC ? absC + (size_t)i * K : NULL;
        for (int k = 0; k < K; k++) c[k] = 0.0;
        for (int e = rowptr[i]; e < rowptr[i + 1]; e++) {
            const double *b = B + (size_t)colidx[e] * K;
            const double v = val[e];
            for (int k = 0; k < K; k++) c[k] += v * b[k];
        }
        if (!a) continue;
        for (int k = 0; k < K; k++) a[k] = 0.0;
        for (int e = rowptr[i]; e < rowptr[i + 1]; e++) {
            const double *b = B + (size_t)colidx[e] * K;
            const double v = val[e] < 0 ? -val[e] : val[e];
            for (int k = 0; k < K; k++) a[k] += v * (b[k] < 0 ? -b[k] : b[k]);
        }
    }
}

void oracle_spmm_t_csr_f64(int m, int n, const int *rowptr, const int *colidx,
                           const double *val, int K, const double *dC,
                           double *dB) {
    memset(dB, 0, sizeof(double) * (size_t)n * K);
    for (int i = 0; i < m; i++)
        for (int e = rowptr[i]; e < rowptr[i + 1]; e++) {
            double *d = dB + (size_t)colidx[e] * K;
            const double *g = dC + (size_t)i * K;
            for (int k = 0; k < K; k++) d[k] += val[e] * g[k];
        }
}

void oracle_sddmm_csr_f64(int m, int n, const int *rowptr, const int *colidx,
                          int K, const double *dC, const double *B,
                          double *dvals) {
    (void)n;
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < m; i++)
        for (int e = rowptr[i]; e < rowptr[i + 1]; e++) {
            const double *g = dC + (size_t)i * K;
            const double *b = B + (size_t)colidx[e] * K;
            double acc = 0.0;
            for (int k = 0; k < K; k++) acc += g[k] * b[k];
            dvals[e] = acc;
        }
}

/* test/MTX_info/main.cu:143-176: an entry count per 32x32 block (row-major
 * over block rows), then the blocks whose count is zero. */
void oracle_block_stats(int m, int n, const int *rowptr, const int *colidx,
                        long long *blocks, long long *empty) {
    const long long rows = m % 32 == 0 ? m / 32 : m / 32 + 1;
    const long long cols = n % 32 == 0 ? n / 32 : n / 32 + 1;
    const long long total = rows * cols;
    int *count = (int *)calloc(total > 0 ? (size_t)total : 1, sizeof(int));
    for (int i = 0; i < m; i++)
        for (int j = rowptr[i]; j < rowptr[i + 1]; j++)
            count[(long long)(i / 32) * cols + colidx[j] / 32]++;
    long long e = 0;
    for (long long b = 0; b < total; b++)
        if (count[b] == 0) e++;
    free(count);
    *blocks = total;
    *empty = e;
}

/* Entry counts of the 32x32 blocks as above: blocks holding at least thr
 * entries, the entries they hold, and the largest count (the density figures
 * behind the choice of dense matrix-core tiles; no counterpart in the
 * reference). */
void oracle_block_density(int m, int n, const int *rowptr, const int *colidx, int thr,
                          long long *dense, long long *dense_nnz, int *max_count) {
    const long long rows = m % 32 == 0 ? m / 32 : m / 32 + 1;
    const long long cols = n % 32 == 0 ? n / 32 : n / 32 + 1;
    const long long total = rows * cols;
    int *count = (int *)calloc(total > 0 ? (size_t)total : 1, sizeof(int));
    for (int i = 0; i < m; i++)
        for (int j = rowptr[i]; j < rowptr[i + 1]; j++)
            count[(long long)(i / 32) * cols + colidx[j] / 32]++;
    long long d = 0, dz = 0;
    int mx = 0;
    for (long long b = 0; b < total; b++) {
        if (count[b] > mx) mx = count[b];
        if (count[b] >= thr) {
            d++;
            dz += count[b];
        }
    }
    free(count);
    *dense = d;
    *dense_nnz = dz;
    *max_count = mx;
}
