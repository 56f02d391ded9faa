"""Writes the golden fixtures of tests/golden/ from the reference's own CPU code.

Runs in the build container only: it loads oracle/_ref/libswiftref_{f64,f32}.so,
the reference's CPU headers (src/ and test/k-32/FP32/) compiled in place by
oracle/Makefile, and records what they output for small seeded inputs:

  csr2csc      src/coo2csc.h:11        CSC of each matrix
  col_sort     src/coo2csc.h:82        sorted nnz (tie-independent) and, for a
                                       matrix with distinct column counts, the
                                       full sorted CSC
  formattransation + block_catgory + balance (src/coo2csc.h:137,
               src/block_catgory.h:1, src/balance.h:1) fed the column order of
               the stable tie rule
  row_idx_sort src/back_process.h:23   sorted keys and val_offset
  Swift_cpu    src/Swift_cpu.h:2       C in float and double
  mmio_allinone src/auxiliary.h:167    CSR of each .mtx fixture

Usage: python tests/golden/make_golden.py   (rewrites tests/golden/*.npz)
"""
import glob
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.dirname(HERE))

import matgen  # noqa: E402
import oracle as O  # noqa: E402

CASES = {
    "random_300x200": lambda: matgen.random_csr(300, 200, 0.05, 11),
    "powerlaw_500x333": lambda: matgen.powerlaw_csr(500, 333, 12),
    "densecol_400x100": lambda: matgen.dense_column_csr(400, 100, 13),
    "equalrows_100x70": lambda: matgen.equal_rows_csr(100, 70, 14),
    "unsorted_rows_64x96": lambda: matgen.random_csr(64, 96, 0.1, 15, shuffle_rows=True),
    "small_5x40": lambda: matgen.random_csr(5, 40, 0.3, 16),
}

FORMAT_KEYS = ("tilenum", "regular_nnz", "reside_col", "reside_nnz", "segsum", "numblk",
               "sortrowidx", "tile_ptr", "tile_len", "tile_colidx", "tile_tall",
               "tile_format", "segmentoffset", "sortrowindex", "reside_cscptr",
               "reside_cscrowidx", "colorblk", "blkstart", "blkstop", "blkptr")


def make_case(name, gen, r64, r32):
    m, n, rp, ci, v = gen()
    out = dict(m=m, n=n, rowptr=rp, colidx=ci, val=v)
    cptr, crow, cval, npc = r64.csr2csc(m, n, rp, ci, v)
    out.update(csc_ptr=cptr, csc_row=crow, csc_val=cval, nnzpercol=npc)
    # reference col_sort (its tie order is libstdc++'s): sortnnz only
    B1 = np.arange(n, dtype=np.float64).reshape(n, 1)
    _, _, sn_ref, sB = r64.col_sort(n, npc, cptr, crow, cval, B1)
    out["ref_sortnnz"] = sn_ref
    out["ref_colsort_perm"] = sB[:, 0].astype(np.int32)
    # stable tie order (the reference GPU ColSort), then the reference tiler
    perm, sn, sr, sv = O.col_sort(n, npc, cptr, crow, cval)
    out.update(perm=perm, sortnnz=sn, sortrowidx_tmp=sr, sortval_tmp=sv)
    h, f = r64.format(m, n, sr, sv, sn)
    for k in FORMAT_KEYS:
        out["fmt_" + k] = np.asarray(f[k])
    out["fmt_sortval"] = f["sortval"]
    out["fmt_reside_val"] = f["reside_val"]
    K = 8
    rng = np.random.default_rng(1000 + m)
    B = rng.uniform(-1, 1, (n, K))
    sB = B[perm]
    C64, _ = r64.swift_cpu(h, m, n, len(ci), sB)
    out.update(B=B, C64=C64)
    rs = r64.row_idx_sort(h, m, n)
    out["rowsorted"] = rs["sortrowidx"]
    out["val_offset"] = rs["val_offset"]
    r64.free(h)
    h32, _ = r32.format(m, n, sr, sv.astype(np.float32), sn)
    C32, _ = r32.swift_cpu(h32, m, n, len(ci), sB.astype(np.float32))
    out["C32"] = C32
    r32.free(h32)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print(name, m, n, len(ci), "tiles", f["tilenum"], "reside", f["reside_nnz"])


def make_mtx(r64):
    for path in sorted(glob.glob(os.path.join(HERE, "mtx", "*.mtx"))):
        rc, d = r64.mmio_load(path)
        base = os.path.splitext(os.path.basename(path))[0]
        out = dict(rc=rc)
        if rc == 0:
            out.update(d)
        np.savez_compressed(os.path.join(HERE, "mtx", base + ".npz"), **out)
        print(base, rc, "" if rc else (d["m"], d["n"], d["nnz"]))


def main():
    r64, r32 = O.Ref(np.float64), O.Ref(np.float32)
    for name, gen in CASES.items():
        make_case(name, gen, r64, r32)
    make_mtx(r64)


if __name__ == "__main__":
    main()
