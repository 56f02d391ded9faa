"""Shared helpers of the GPU parity tests."""
import numpy as np

import matgen


def cases():
    """name -> (m, n, rowptr, colidx, val) covering the edge cases of the path."""
    c = {
        "random_300x200": matgen.random_csr(300, 200, 0.05, 11),
        "powerlaw_500x333": matgen.powerlaw_csr(500, 333, 12),
        "densecol_400x100": matgen.dense_column_csr(400, 100, 13),
        "equalrows_100x70": matgen.equal_rows_csr(100, 70, 14),
        "unsorted_rows_64x96": matgen.random_csr(64, 96, 0.1, 15, shuffle_rows=True),
        "narrow_50x20": matgen.random_csr(50, 20, 0.3, 3),
        "tall_2000x64": matgen.random_csr(2000, 64, 0.05, 4),
        "wide_40x3000": matgen.random_csr(40, 3000, 0.02, 5),
        "powerlaw_3000x2500": matgen.powerlaw_csr(3000, 2500, 6, max_len=400),
        # the benchmark matrix's structure at 1/250 scale: even rows banded,
        # odd rows Zipf lengths with uniform columns
        "banded_mixed_4000": matgen.banded_mixed_csr(4000, 4000, 21),
        # every row banded (bandwidth 24): workgroups of the window walk stage
        # their B rows in LDS
        "banded_3000": matgen.banded_mixed_csr(3000, 3000, 22, bw=24, mixed=False),
    }
    # uniform columns, N % 32 == 0: no irregular part at all
    m = n = 64
    rp = np.arange(0, 3 * m + 1, 3, dtype=np.int32)
    ci = np.concatenate([np.sort([(i + d) % n for d in range(3)]) for i in range(m)])
    c["uniform_64x64"] = (m, n, rp, ci.astype(np.int32), np.linspace(-1, 1, 3 * m))
    # empty matrix and empty rows/cols
    c["empty_10x40"] = (10, 40, np.zeros(11, np.int32), np.zeros(0, np.int32), np.zeros(0))
    return c


def rel_err(C, C64, absC):
    """max |C - C64| / (|A||B| + 1e-30), the per-element error normalisation."""
    C = np.asarray(C, np.float64)
    return float(np.max(np.abs(C - C64) / (absC + 1e-30))) if C.size else 0.0


def row_max_rel(C, C64):
    """per-row max relative error against |C64| (reported for the 16-bit forms)."""
    C = np.asarray(C, np.float64)
    den = np.maximum(np.max(np.abs(C64), axis=1), 1e-30)
    return np.max(np.abs(C - C64), axis=1) / den
