"""Row views that hold the reside entries.

A matrix whose reside part is at most 1% of its entries has its reside
entries walked with the regular ones (regular, then reside, within each row),
with no reside blocks.  The Chung-Lu graph at 1/4 scale (50,000 nodes,
2,474,324 nonzeros, 22,484 reside entries) and the benchmark matrix at 1/20
scale (420 reside entries) qualify; tall_2000x64 (12% of its entries reside)
does not.  The test checks the choice from the oracle's format, then runs
every row-view product form through the graph against the FP64 oracle: FP32
under each row-view kernel (plain and accumulating), BF16, column-major B,
SDDMM and the fused backward."""
import numpy as np
import pytest

import gpu_common

pytestmark = pytest.mark.gpu

_MAT = {}


def graph(swift, oracle):
    if not _MAT:
        d = swift.generate(4, scale=0.25, seed=0)
        f = oracle.build(d["m"], d["n"], d["rowptr"], d["colidx"], d["val"].astype(np.float64))
        _MAT.update(d=d, reside=np.asarray(f["reside_cscrowidx"]))
    return _MAT["d"], _MAT["reside"]


def _matrix(swift, d):
    import torch
    return swift.SwiftMatrix(d["m"], d["n"], d["rowptr"], d["colidx"],
                             d["val"].astype(np.float32), dtype=torch.float32)


def test_choice_of_row_view(swift, oracle, cuda):
    import torch
    d, reside = graph(swift, oracle)
    assert 0 < len(reside) * 100 <= d["nnz"], len(reside)
    M = _matrix(swift, d)
    inf = M.info()
    assert inf.reside_nnz == len(reside)
    assert inf.rowview_nnz == d["nnz"] == inf.regular_nnz + inf.reside_nnz
    M.close()
    c = swift.generate(2, scale=0.05, seed=0)
    M = _matrix(swift, c)
    inf = M.info()
    assert 0 < inf.reside_nnz * 100 <= c["nnz"] and inf.rowview_nnz == c["nnz"]
    M.close()
    # a reside part over 1% keeps its balance blocks: the row view holds the
    # regular entries only
    m, n, rp, ci, v = gpu_common.cases()["tall_2000x64"]
    f = oracle.build(m, n, rp, ci, np.asarray(v, np.float64))
    assert len(f["reside_cscrowidx"]) * 100 > len(ci)
    M = swift.SwiftMatrix(m, n, rp, ci, np.asarray(v, np.float32), dtype=torch.float32)
    inf = M.info()
    assert inf.reside_nnz > 0 and inf.rowview_nnz == inf.regular_nnz
    M.close()


@pytest.mark.parametrize("K", [32, 128])
def test_forward_kernels(swift, oracle, cuda, K):
    import torch
    d, _ = graph(swift, oracle)
    M = _matrix(swift, d)
    vals = d["val"].astype(np.float32).astype(np.float64)
    rng = np.random.default_rng(K)
    B = rng.uniform(-1, 1, (d["n"], K)).astype(np.float32)
    C64, absC = oracle.spmm_f64(d["m"], d["n"], d["rowptr"], d["colidx"], vals,
                                B.astype(np.float64))
    Bt = torch.tensor(B, device="cuda")
    for name in ("KERNEL_DEFAULT", "KERNEL_ROWVIEW", "KERNEL_FLAT", "KERNEL_WINDOW"):
        C = torch.full((d["m"], K), float("nan"), device="cuda")
        M.spmm(Bt, out=C, kernel=getattr(swift, name))
        err = gpu_common.rel_err(C.cpu().numpy(), C64, absC)
        print(name, "K", K, "err", err, flush=True)
        assert err <= 1e-5, name
    C0 = rng.uniform(-1, 1, (d["m"], K)).astype(np.float32)
    C = torch.tensor(C0, device="cuda")
    M.spmm(Bt, out=C, flags=swift.ACCUMULATE)
    assert gpu_common.rel_err(C.cpu().numpy(), C64 + C0, absC + np.abs(C0)) <= 1e-5
    Cc = M.spmm(torch.tensor(B.T.copy(), device="cuda"), layout=swift.COL_MAJOR)
    assert gpu_common.rel_err(Cc.cpu().numpy().T, C64, absC) <= 1e-5
    M.prepare_values(swift.BF16)
    C = torch.full((d["m"], K), float("nan"), device="cuda")
    M.spmm(swift.convert_dense(Bt, swift.BF16), out=C, dtype=swift.BF16)
    assert gpu_common.rel_err(C.cpu().numpy(), C64, absC) <= 1e-2
    M.close()


def test_sddmm_and_backward(swift, oracle, cuda):
    import torch
    d, _ = graph(swift, oracle)
    K = 64
    M = _matrix(swift, d)
    rng = np.random.default_rng(5)
    B = rng.uniform(-1, 1, (d["n"], K)).astype(np.float32)
    dC = rng.uniform(-1, 1, (d["m"], K)).astype(np.float32)
    args = (d["m"], d["n"], d["rowptr"], d["colidx"])
    ref = oracle.sddmm_f64(*args, dC.astype(np.float64), B.astype(np.float64))
    den = oracle.sddmm_f64(*args, np.abs(dC).astype(np.float64), np.abs(B).astype(np.float64))
    dCt, Bt = torch.tensor(dC, device="cuda"), torch.tensor(B, device="cuda")
    dv = M.sddmm(dCt, Bt).cpu().numpy()
    print("sddmm done", flush=True)
    assert float(np.max(np.abs(dv - ref) / (den + 1e-30))) <= 1e-5
    # dB = A^T dC through the transposed layout, dvals through the same walk
    v64 = d["val"].astype(np.float32).astype(np.float64)
    dB64 = oracle.spmm_t_f64(*args, v64, dC.astype(np.float64))
    absB = oracle.spmm_t_f64(*args, np.abs(v64), np.abs(dC).astype(np.float64))
    dB, dv2 = M.backward_fused(dCt, Bt)
    assert gpu_common.rel_err(dB.cpu().numpy(), dB64, absB) <= 1e-5
    assert gpu_common.rel_err(M.transpose().spmm(dCt).cpu().numpy(), dB64, absB) <= 1e-5
    assert float(np.max(np.abs(dv2.cpu().numpy() - ref) / (den + 1e-30))) <= 1e-5
    M.close()
