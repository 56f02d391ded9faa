"""Fused backward on transposed rows that segment boundaries cut into many pieces.

The matrix has five dense columns (1,500-2,600 entries each) beside a few
uniform random entries per row.  In its transpose those columns are rows that
segments of 8 items cut into more than 32 pieces, so the fix-up of dB's cut
rows takes the workgroup path (four rows per workgroup, the last workgroup
with three groups idle), beside many rows of a few pieces; the default
segments of the same matrix cut the dense rows into fewer pieces.  dB and
dvals of the fused backward (FP32 and BF16, K = 32, 100 with masked slices,
and 128), dB alone and dvals alone, each against the FP64 oracle: FP32 within
1e-5 of the absolute sums, BF16 within 1e-5 on the BF16-rounded operands."""
import numpy as np
import pytest

import gpu_common

pytestmark = pytest.mark.gpu

_CASE = {}
DENSE = (7, 300, 1100, 1500, 1999)


def case():
    """m = 3000, n = 2000: 3-12 random columns per row, and each dense column
    in a random 50-87% of the rows."""
    if not _CASE:
        rng = np.random.default_rng(2024)
        m, n = 3000, 2000
        share = rng.uniform(0.5, 0.87, len(DENSE))
        rows = []
        for i in range(m):
            cols = set(rng.choice(n, int(rng.integers(3, 13)), replace=False).tolist())
            cols.update(c for c, p in zip(DENSE, share) if rng.random() < p)
            rows.append(sorted(cols))
        rp = np.zeros(m + 1, np.int32)
        rp[1:] = np.cumsum([len(r) for r in rows])
        ci = np.concatenate([np.asarray(r, np.int32) for r in rows])
        v = rng.uniform(-1, 1, len(ci)).astype(np.float32)
        _CASE.update(m=m, n=n, rp=rp, ci=ci, v=v)
    c = _CASE
    return c["m"], c["n"], c["rp"], c["ci"], c["v"]


def cuts_per_row(rowptr, W):
    """Segment boundaries inside each row, for segments of W items over
    (rows + entries): segment s starts at diagonal s*W, moved back to the
    start of the row it falls in when that row began less than W items
    earlier (the rule test_gpu_carry.py checks against the device)."""
    ptr = np.asarray(rowptr, np.int64)
    m, R = len(ptr) - 1, int(ptr[-1])
    nseg = -(-(m + R) // W)
    s = np.arange(1, nseg, dtype=np.int64)
    d = s * W
    row = np.searchsorted(ptr[1:] + np.arange(m), d, side="left")
    e = d - row
    start = ptr[row]
    e = np.where((e > start) & (start + row > d - W), start, e)
    cut = e > start
    return np.bincount(row[cut], minlength=m)


def test_dense_columns():
    """At 8 items per segment the transpose's five dense rows are cut into
    more than 32 pieces and no other row is; at the default size (128 items)
    none is."""
    m, n, rp, ci, v = case()
    tptr = np.zeros(n + 1, np.int64)
    tptr[1:] = np.cumsum(np.bincount(ci, minlength=n))
    long8 = np.flatnonzero(cuts_per_row(tptr, 8) > 32)
    assert long8.tolist() == list(DENSE), long8
    assert (cuts_per_row(tptr, 128) <= 32).all()


def _refs(oracle, m, n, rp, ci, vals, dC, B):
    dB64 = oracle.spmm_t_f64(m, n, rp, ci, vals, dC)
    absB = oracle.spmm_t_f64(m, n, rp, ci, np.abs(vals), np.abs(dC))
    dv64 = oracle.sddmm_f64(m, n, rp, ci, dC, B)
    absv = oracle.sddmm_f64(m, n, rp, ci, np.abs(dC), np.abs(B))
    return dB64, absB, dv64, absv


def _dv_err(dv, dv64, absv):
    return float(np.max(np.abs(np.asarray(dv, np.float64) - dv64) / (absv + 1e-30), initial=0))


@pytest.mark.parametrize("items", [8, 0])
@pytest.mark.parametrize("K", [32, 100, 128])
def test_fused_backward_fp32(swift, oracle, cuda, K, items):
    import torch
    m, n, rp, ci, v = case()
    print("case fp32 K", K, "items", items, flush=True)
    rng = np.random.default_rng(K)
    B = rng.standard_normal((n, K)).astype(np.float32)
    dC = rng.standard_normal((m, K)).astype(np.float32)
    M = swift.SwiftMatrix(m, n, rp, ci, v, dtype=torch.float32)
    if items:
        M.transpose().set_segment_items(items)
    dCt, Bt = torch.tensor(dC, device="cuda"), torch.tensor(B, device="cuda")
    dB, dv = M.backward_fused(dCt, Bt)
    dB_only, none_v = M.backward_fused(dCt, Bt, want_dvals=False)
    none_b, dv_only = M.backward_fused(dCt, Bt, want_dB=False)
    assert none_v is None and none_b is None
    dB64, absB, dv64, absv = _refs(oracle, m, n, rp, ci, v.astype(np.float64),
                                   dC.astype(np.float64), B.astype(np.float64))
    errs = (gpu_common.rel_err(dB.cpu().numpy(), dB64, absB),
            gpu_common.rel_err(dB_only.cpu().numpy(), dB64, absB),
            _dv_err(dv.cpu().numpy(), dv64, absv), _dv_err(dv_only.cpu().numpy(), dv64, absv))
    print("dB %.3g dB-only %.3g dvals %.3g dvals-only %.3g" % errs, flush=True)
    assert max(errs) <= 1e-5
    M.close()


@pytest.mark.parametrize("items", [8, 0])
@pytest.mark.parametrize("K", [32, 128])
def test_fused_backward_bf16(swift, oracle, cuda, K, items):
    import torch
    m, n, rp, ci, v = case()
    print("case bf16 K", K, "items", items, flush=True)
    rng = np.random.default_rng(K + 1)
    B = rng.standard_normal((n, K)).astype(np.float32)
    dC = rng.standard_normal((m, K)).astype(np.float32)
    rnd = lambda x: torch.tensor(x).to(torch.bfloat16).double().numpy()
    M = swift.SwiftMatrix(m, n, rp, ci, v, dtype=torch.float32)
    if items:
        M.transpose().set_segment_items(items)
    dB, dv = M.backward_fused(torch.tensor(dC, device="cuda").to(torch.bfloat16),
                              torch.tensor(B, device="cuda").to(torch.bfloat16))
    dB64, absB, dv64, absv = _refs(oracle, m, n, rp, ci, rnd(v), rnd(dC), rnd(B))
    e1 = gpu_common.rel_err(dB.cpu().numpy(), dB64, absB)
    e2 = _dv_err(dv.cpu().numpy(), dv64, absv)
    print("dB %.3g dvals %.3g" % (e1, e2), flush=True)
    assert e1 <= 1e-5 and e2 <= 1e-5
    M.close()
