"""Every product form on the benchmark matrix's structure at 1/20 scale.

50,000 rows, 804,644 nonzeros: even rows banded, odd rows with Zipf lengths
(153 rows of more than 256 entries, the longest 13,831), so the default
segment tables cut many rows (their pieces go through the carry slots and
the fix-up), and 420 reside entries (under 1%), which the row view holds.  Checked against the FP64 oracle: FP32
at K=32/128 (plain and accumulating into NaN-free and NaN-prefilled outputs),
BF16, F16 with column scales, column-major B, and SDDMM."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_MAT = {}


def matrix(swift, oracle):
    if not _MAT:
        d = swift.generate(2, scale=0.05, seed=0)
        f = oracle.build(d["m"], d["n"], d["rowptr"], d["colidx"], d["val"].astype(np.float64))
        rows = np.unique(np.asarray(f["reside_cscrowidx"]))
        _MAT.update(d=d, reside_rows=rows)
    return _MAT["d"], _MAT["reside_rows"]


def test_matrix_structure(swift, oracle, cuda):
    """Rows longer than the default segments (the row view holds 17 items per
    row on average and a default segment at least three such rows, 51 items)
    and a reside part of under 1% of the entries."""
    import torch
    d, rows = matrix(swift, oracle)
    lens = np.diff(d["rowptr"])
    assert lens.max() > 256 and (lens > 256).sum() > 100, (lens.max(), (lens > 256).sum())
    M = swift.SwiftMatrix(d["m"], d["n"], d["rowptr"], d["colidx"], d["val"].astype(np.float32),
                          dtype=torch.float32)
    inf = M.info()
    assert 0 < inf.reside_nnz * 100 <= d["nnz"] and inf.rowview_nnz == d["nnz"]
    M.close()


def _setup(swift, oracle, K, seed=0):
    import torch
    d, _ = matrix(swift, oracle)
    vals = d["val"].astype(np.float32)
    B = np.random.default_rng(seed).uniform(-1, 1, (d["n"], K)).astype(np.float32)
    C64, absC = oracle.spmm_f64(d["m"], d["n"], d["rowptr"], d["colidx"], vals.astype(np.float64),
                                B.astype(np.float64))
    M = swift.SwiftMatrix(d["m"], d["n"], d["rowptr"], d["colidx"], vals, dtype=torch.float32)
    return d, M, B, C64, absC


def _err(C, C64, absC):
    return float(np.max(np.abs(np.asarray(C, np.float64) - C64) / (absC + 1e-30)))


@pytest.mark.parametrize("K", [32, 128])
def test_structure_f32(swift, oracle, cuda, K):
    import torch
    d, M, B, C64, absC = _setup(swift, oracle, K)
    Bt = torch.tensor(B, device="cuda")
    C = torch.full((d["m"], K), float("nan"), device="cuda")
    M.spmm(Bt, out=C)
    err = _err(C.cpu().numpy(), C64, absC)
    print("bench structure f32 K", K, "err", err, flush=True)
    assert err <= 1e-5
    # C += A*B: split and reside rows take atomics onto the existing values
    C0 = np.random.default_rng(1).uniform(-1, 1, (d["m"], K)).astype(np.float32)
    C = torch.tensor(C0, device="cuda")
    M.spmm(Bt, out=C, flags=swift.ACCUMULATE)
    err = _err(C.cpu().numpy(), C64 + C0, absC + np.abs(C0))
    print("bench structure f32 accumulate K", K, "err", err, flush=True)
    assert err <= 1e-5
    # the separate-launch kernel gives the same product
    Cr = M.spmm(Bt, kernel=swift.KERNEL_ROWVIEW)
    assert _err(Cr.cpu().numpy(), C64, absC) <= 1e-5
    M.close()


def test_structure_lowp(swift, oracle, cuda):
    import torch
    K = 128
    d, M, B, C64, absC = _setup(swift, oracle, K)
    Bt = torch.tensor(B, device="cuda")
    M.prepare_values(swift.BF16)
    C = torch.full((d["m"], K), float("nan"), device="cuda")
    M.spmm(swift.convert_dense(Bt, swift.BF16), out=C, dtype=swift.BF16)
    err = _err(C.cpu().numpy(), C64, absC)
    print("bench structure bf16 err", err, flush=True)
    assert err <= 1e-2
    M.prepare_values(swift.F16)
    Bh, sc = swift.convert_dense(Bt, swift.F16, col_scale=True)
    C = torch.full((d["m"], K), float("nan"), device="cuda")
    M.spmm_scaled(Bh, sc, out=C)
    err = _err(C.cpu().numpy(), C64, absC)
    print("bench structure f16 err", err, flush=True)
    assert err <= 1e-2
    M.close()


def test_structure_col_major(swift, oracle, cuda):
    import torch
    K = 128
    d, M, B, C64, absC = _setup(swift, oracle, K)
    Bt = torch.tensor(B.T.copy(), device="cuda")
    C = M.spmm(Bt, layout=swift.COL_MAJOR).cpu().numpy().T
    err = _err(C, C64, absC)
    print("bench structure col-major err", err, flush=True)
    assert err <= 1e-5
    M.close()


def test_structure_sddmm(swift, oracle, cuda):
    import torch
    K = 128
    d, M, B, _, _ = _setup(swift, oracle, K)
    dC = np.random.default_rng(2).uniform(-1, 1, (d["m"], K)).astype(np.float32)
    ref = oracle.sddmm_f64(d["m"], d["n"], d["rowptr"], d["colidx"], dC.astype(np.float64),
                           B.astype(np.float64))
    den = oracle.sddmm_f64(d["m"], d["n"], d["rowptr"], d["colidx"],
                           np.abs(dC).astype(np.float64), np.abs(B).astype(np.float64))
    dv = M.sddmm(torch.tensor(dC, device="cuda"), torch.tensor(B, device="cuda")).cpu().numpy()
    err = float(np.max(np.abs(dv.astype(np.float64) - ref) / (den + 1e-30)))
    print("bench structure sddmm err", err, flush=True)
    assert err <= 1e-5
    M.close()
