// ref_harness.cpp -- C-linkage shim over the reference's own CPU headers.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file against the
// reference sources where they lie (-I <reference>/src for double,
// -I <reference>/test/k-32/FP32 for float) into oracle/_ref/.  Nothing of the
// reference is copied; this file only calls its functions:
//   mmio_allinone      src/auxiliary.h:167
//   csr2csc            src/coo2csc.h:11
//   col_sort           src/coo2csc.h:82
//   formattransation   src/coo2csc.h:137
//   block_catgory      src/block_catgory.h:1
//   balance            src/balance.h:1
//   row_idx_sort       src/back_process.h:23
//   Swift_cpu          src/Swift_cpu.h:2
// The tests use it to pin the plain-C oracle, and the benchmark may time its
// Swift_cpu as the reference CPU baseline.
#include <sys/time.h>

#include "coo2csc.h"  // pulls auxiliary.h, mmio.h, format.h
#include "block_catgory.h"
#include "balance.h"
#include "back_process.h"
#include "Swift_cpu.h"

#ifdef REF_F32
typedef float VT;
#else
typedef double VT;
#endif

extern "C" {

struct ref_view {
    int tilenum, regular_nnz, reside_col, reside_nnz, segsum, numblk;
    VT *sortval;
    int *sortrowidx, *tile_ptr, *tile_len, *tile_colidx, *tile_tall;
    int *tile_format, *segmentoffset, *sortrowindex;
    int *reside_cscptr, *reside_cscrowidx;
    VT *reside_val;
    unsigned int *colorblk;
    int *blkstart, *blkstop, *blkptr;
    int *val_position, *val_offset;
};

int ref_value_size(void) { return (int)sizeof(VT); }

#ifndef REF_F32
int ref_mmio_load(const char *path, int *m, int *n, int *nnz, int *sym,
                  int **rowptr, int **colidx, double **val) {
    return mmio_allinone(m, n, nnz, sym, rowptr, colidx, val,
                         const_cast<char *>(path));
}
#endif

void ref_csr2csc(int nnz, int m, int n, int *colidx, int *rowptr, VT *val,
                 int **cscrow, int **cscptr, int **npc, VT **cscval) {
    csr2csc(nnz, m, n, colidx, rowptr, val, cscrow, cscptr, npc, cscval);
}

void ref_col_sort(int n, int K, int *npc, int *cscptr, int *cscrow, VT *cscval,
                  int *sortrow, VT *sortval, int *sortnnz, VT *B, VT *sortB) {
    col_sort(n, K, npc, cscptr, cscrow, cscval, sortrow, sortval, sortnnz, B,
             sortB);
}

// formattransation + block_catgory, then balance over the reside CSC, in the
// order of src/main.cu:279-291 and :367-370.
void *ref_format(int *sortrow, VT *sortval, int *sortnnz, int nnz, int m, int n) {
    slide_matrix *s = (slide_matrix *)calloc(1, sizeof(slide_matrix));
#ifdef REF_F32
    formattransation(0.0f, s, sortrow, sortval, sortnnz, nnz, m, n);
    block_catgory(s);
#else
    double ratio = 0, shuffle = 0;
    formattransation(0.0, s, sortrow, sortval, sortnnz, nnz, m, n, ratio);
    block_catgory(s, shuffle);
#endif
    balance(s, m, s->reside_col, s->reside_nnz, s->reside_cscptr,
            s->reside_cscrowidx);
    return s;
}

void ref_row_idx_sort(void *h, VT *sortB, int m, int n, int K) {
    slide_matrix *s = (slide_matrix *)h;
#ifdef REF_F32
    float ratio = 0;
    row_idx_sort(s, sortB, m, n, K, ratio);
#else
    row_idx_sort(s, sortB, m, n, K);
#endif
}

void ref_export(void *h, ref_view *v) {
    slide_matrix *s = (slide_matrix *)h;
    v->tilenum = s->tilenum;
    v->regular_nnz = s->regular_nnz;
    v->reside_col = s->reside_col;
    v->reside_nnz = s->reside_nnz;
    v->segsum = s->segsum;
    v->numblk = s->numblk;
    v->sortval = s->sortval;
    v->sortrowidx = s->sortrowidx;
    v->tile_ptr = s->tile_ptr;
    v->tile_len = s->tile_len;
    v->tile_colidx = s->tile_colidx;
    v->tile_tall = s->tile_tall;
    v->tile_format = s->tile_format;
    v->segmentoffset = s->segmentoffset;
    v->sortrowindex = s->sortrowindex;
    v->reside_cscptr = s->reside_cscptr;
    v->reside_cscrowidx = s->reside_cscrowidx;
    v->reside_val = s->reside_val;
    v->colorblk = s->colorblk;
    v->blkstart = s->blkstart;
    v->blkstop = s->blkstop;
    v->blkptr = s->blkptr;
    v->val_position = s->val_position;
    v->val_offset = s->val_offset;
}

// Swift_cpu over row-major sorted B / C; returns the reference's own
// gettimeofday measurement in ms.
double ref_swift_cpu(void *h, int nnz, int m, int n, int K, VT *sortB, VT *C) {
#ifdef REF_F32
    float ms = 0;
#else
    double ms = 0;
#endif
    Swift_cpu(ms, (slide_matrix *)h, nnz, m, n, K, sortB, C);
    return (double)ms;
}

// Only the buffers this harness can reach are released; the reference's
// Tile_destroy (src/format.h:48) frees fields that were never allocated.
void ref_free(void *h) {
    slide_matrix *s = (slide_matrix *)h;
    if (!s) return;
    void *p[] = {s->sortval, s->sortrowidx, s->tile_ptr, s->tile_len,
                 s->tile_colidx, s->tile_tall, s->tile_format,
                 s->segmentoffset, s->sortrowindex, s->reside_cscptr,
                 s->reside_cscrowidx, s->reside_val, s->colorblk, s->blkstart,
                 s->blkstop, s->blkptr, s->val_position, s->val_offset};
    for (unsigned i = 0; i < sizeof p / sizeof p[0]; i++) free(p[i]);
    free(s);
}

void ref_free_ptr(void *p) { free(p); }

}  // extern "C"
