"""16-bit storage forms and the backward kernels on the GPU.

BF16 / F16 (per-column scaled) storage of A's values and of B with FP32
accumulation: per-element error normalised by |A||B| <= 1e-2 against the FP64
oracle (per-row max relative error is printed).  Backward: dB = A^T dC through
the transposed Swift layout and through the fused kernel, dvals = SDDMM(dC, B),
both against FP64 closed forms; the autograd wrapper against FP64 finite
differences of the loss."""
import numpy as np
import pytest

import gpu_common
import matgen

pytestmark = pytest.mark.gpu

CASES = gpu_common.cases()


def setup(swift, oracle, name, K, seed=0):
    import torch
    m, n, rp, ci, v = CASES[name]
    vv = np.asarray(v, np.float32)
    B = np.random.default_rng(seed).uniform(-1, 1, (n, K)).astype(np.float32)
    C64, absC = oracle.spmm_f64(m, n, rp, ci, vv.astype(np.float64), B.astype(np.float64))
    M = swift.SwiftMatrix(m, n, rp, ci, vv, dtype=torch.float32)
    return M, B, C64, absC


ROW_KERNELS = {"ROW_FLAT": "KERNEL_FLAT", "ROW_WIN": "KERNEL_WINDOW"}


@pytest.mark.parametrize("name", ["powerlaw_3000x2500", "densecol_400x100", "equalrows_100x70",
                                  "narrow_50x20", "banded_3000"])
@pytest.mark.parametrize("K", [32, 100, 128, 256])
@pytest.mark.parametrize("layout", ["ROW", "COL", "ROW_FLAT", "ROW_WIN"])
def test_bf16_storage(swift, oracle, cuda, name, K, layout):
    import torch
    M, B, C64, absC = setup(swift, oracle, name, K)
    lay = swift.COL_MAJOR if layout == "COL" else swift.ROW_MAJOR
    kernel = getattr(swift, ROW_KERNELS.get(layout, "KERNEL_DEFAULT"))
    M.prepare_values(swift.BF16)
    Bt = torch.tensor(B if lay == swift.ROW_MAJOR else B.T.copy(), device="cuda")
    Bh = swift.convert_dense(Bt, swift.BF16, layout=lay)
    C = M.spmm(Bh, layout=lay, dtype=swift.BF16, kernel=kernel).cpu().numpy()
    if lay == swift.COL_MAJOR:
        C = C.T
    err = gpu_common.rel_err(C, C64, absC)
    print("bf16", name, K, layout, "err", err,
          "row_max_rel", float(np.max(gpu_common.row_max_rel(C, C64))), flush=True)
    assert err <= 1e-2
    M.close()


@pytest.mark.parametrize("name", ["powerlaw_3000x2500", "banded_3000", "equalrows_100x70"])
@pytest.mark.parametrize("K", [32, 100, 128])
@pytest.mark.parametrize("form", ["f32", "bf16"])
def test_nonfinite_propagation(swift, cuda, name, K, form):
    """inf / -inf / NaN in B reach exactly the outputs they reach in the plain sum: the
    BF16 walk multiplies two entries' values per instruction, and unused slots read
    zeros, so neither may leak a non-finite value into another output or lose one."""
    import torch
    import scipy.sparse as sp
    m, n, rp, ci, v = CASES[name]
    rng = np.random.default_rng(7)
    B = rng.uniform(-1, 1, (n, K)).astype(np.float32)
    cols = np.asarray(ci)
    for x in (np.inf, -np.inf, np.nan):
        for _ in range(3):  # B rows that some entry gathers
            B[cols[rng.integers(cols.size)], rng.integers(K)] = x
    vals = np.asarray(v, np.float32)
    M = swift.SwiftMatrix(m, n, rp, ci, vals, dtype=torch.float32)
    Bt = torch.tensor(B, device="cuda")
    if form == "bf16":
        M.prepare_values(swift.BF16)
        C = M.spmm(swift.convert_dense(Bt, swift.BF16), dtype=swift.BF16).cpu().numpy()
        # the reference rounds both operands the same way (to nearest even)
        vq = torch.tensor(vals).to(torch.bfloat16).double().numpy()
        Bq = Bt.cpu().to(torch.bfloat16).double().numpy()
    else:
        C = M.spmm(Bt).cpu().numpy()
        vq, Bq = vals.astype(np.float64), B.astype(np.float64)
    A = sp.csr_matrix((vq, np.asarray(ci), np.asarray(rp)), shape=(m, n))
    with np.errstate(invalid="ignore"):
        R = A @ Bq
        absR = abs(A) @ np.abs(np.nan_to_num(Bq, nan=0.0, posinf=0.0, neginf=0.0))
    assert np.array_equal(np.isnan(C), np.isnan(R))
    assert np.array_equal(np.isposinf(C), np.isposinf(R))
    assert np.array_equal(np.isneginf(C), np.isneginf(R))
    fin = np.isfinite(R)
    assert not fin.all()
    # both sides sum the same (rounded) operands: only the FP32 accumulation differs
    err = float(np.max(np.abs(C[fin] - R[fin]) / (absR[fin] + 1e-30)))
    assert err <= 1e-5, err
    M.close()


@pytest.mark.parametrize("name", ["powerlaw_3000x2500", "banded_3000"])
@pytest.mark.parametrize("K", [32, 128])
def test_f16_subnormal_products(swift, cuda, name, K):
    """F16 storage keeps B values that scale into the F16 subnormal range.
    Column k of B holds one value of 1 and otherwise values near 2^-30, which
    its per-column scale (max / scale <= 2^14) maps to about 2^-16, below the
    smallest normal F16 (2^-14).  A's values are multiples of 1/8 in [1/8, 1],
    exact in F16 at any power-of-two scale.  The product must equal an FP64 sum
    of the same F16-rounded operands to 1e-5 of (|A||B|)[i,k]: a walk that
    flushed subnormal products to zero would miss whole outputs."""
    import torch
    import scipy.sparse as sp
    m, n, rp, ci, _ = CASES[name]
    rng = np.random.default_rng(11)
    vals = (rng.integers(1, 9, len(ci)) / 8.0 * rng.choice([-1, 1], len(ci))).astype(np.float32)
    B = (rng.uniform(0.5, 1.0, (n, K)) * 2.0 ** -30).astype(np.float32)
    B *= rng.choice([-1, 1], (n, K)).astype(np.float32)
    B[rng.integers(n, size=K), np.arange(K)] = 1.0  # one large value per column
    M = swift.SwiftMatrix(m, n, rp, ci, vals, dtype=torch.float32)
    M.prepare_values(swift.F16)
    Bt = torch.tensor(B, device="cuda")
    Bh, sc = swift.convert_dense(Bt, swift.F16, col_scale=True)
    C = M.spmm_scaled(Bh, sc, layout=swift.ROW_MAJOR).cpu().numpy()
    scale = sc.cpu().numpy().astype(np.float64)
    Bq = (B / scale[None, :]).astype(np.float16).astype(np.float64) * scale[None, :]
    assert np.abs(Bh.cpu().numpy().astype(np.float64) * scale[None, :] - Bq).max() == 0.0
    assert (np.abs(Bq / scale[None, :]) < 2.0 ** -14).mean() > 0.9  # mostly subnormal
    A = sp.csr_matrix((vals.astype(np.float64), np.asarray(ci), np.asarray(rp)), shape=(m, n))
    R = A @ Bq
    absR = abs(A) @ np.abs(Bq)
    err = float(np.max(np.abs(C - R) / (absR + 1e-300)))
    print("f16 subnormal", name, K, "err", err, flush=True)
    assert err <= 1e-5
    M.close()


@pytest.mark.parametrize("name", ["powerlaw_3000x2500", "densecol_400x100", "banded_3000"])
@pytest.mark.parametrize("K", [32, 100, 128])
@pytest.mark.parametrize("layout", ["ROW", "COL", "ROW_FLAT", "ROW_WIN"])
def test_f16_scaled_storage(swift, oracle, cuda, name, K, layout):
    import torch
    M, B, C64, absC = setup(swift, oracle, name, K)
    lay = swift.COL_MAJOR if layout == "COL" else swift.ROW_MAJOR
    kernel = getattr(swift, ROW_KERNELS.get(layout, "KERNEL_DEFAULT"))
    # widen the dynamic range so that the per-column scales matter
    B = B * np.logspace(-3, 3, K, dtype=np.float32)[None, :]
    C64, absC = oracle.spmm_f64(*CASES[name][:4], np.asarray(CASES[name][4], np.float32)
                                .astype(np.float64), B.astype(np.float64))
    M.prepare_values(swift.F16)
    Bt = torch.tensor(B if lay == swift.ROW_MAJOR else B.T.copy(), device="cuda")
    Bh, sc = swift.convert_dense(Bt, swift.F16, layout=lay, col_scale=True)
    C = M.spmm_scaled(Bh, sc, layout=lay, kernel=kernel).cpu().numpy()
    if lay == swift.COL_MAJOR:
        C = C.T
    err = gpu_common.rel_err(C, C64, absC)
    print("f16", name, K, layout, "err", err, flush=True)
    assert err <= 1e-2
    M.close()


def _operand(torch, a, offset):
    """a on the GPU; offset > 0 places it `offset` elements into a larger
    buffer, so its address is not 16-byte aligned."""
    if not offset:
        return torch.tensor(a, device="cuda")
    flat = torch.zeros(a.size + offset, dtype=torch.float32 if a.dtype == np.float32
                       else torch.float64, device="cuda")
    flat[offset:] = torch.tensor(a.ravel(), device="cuda")
    return flat[offset:].view(a.shape)


def check_backward(swift, oracle, name, K, dt, offset=0):
    """Fused backward, the transposed-layout product and SDDMM against the
    FP64 oracle: dB = A^T dC within tol of (|A|^T|dC|), dvals within tol of
    <|dC[row]|, |B[col]|>."""
    import torch
    m, n, rp, ci, v = CASES[name]
    npdt = np.float32 if dt == "f32" else np.float64
    dtype = torch.float32 if dt == "f32" else torch.float64
    tol = 1e-5 if dt == "f32" else 1e-12
    vv = np.asarray(v, npdt)
    rng = np.random.default_rng(K)
    B = rng.standard_normal((n, K)).astype(npdt)
    dC = rng.standard_normal((m, K)).astype(npdt)
    dB64 = oracle.spmm_t_f64(m, n, rp, ci, vv.astype(np.float64), dC.astype(np.float64))
    absB = oracle.spmm_t_f64(m, n, rp, ci, np.abs(vv.astype(np.float64)),
                             np.abs(dC.astype(np.float64)))
    dv64 = oracle.sddmm_f64(m, n, rp, ci, dC.astype(np.float64), B.astype(np.float64))
    absv = oracle.sddmm_f64(m, n, rp, ci, np.abs(dC.astype(np.float64)),
                            np.abs(B.astype(np.float64)))
    M = swift.SwiftMatrix(m, n, rp, ci, vv, dtype=dtype)
    dCt = _operand(torch, dC, offset)
    Bt = _operand(torch, B, offset)
    dB_f, dv_f = M.backward_fused(dCt, Bt)
    dB_t = M.transpose().spmm(dCt)
    dv_s = M.sddmm(dCt, Bt)
    for dB in (dB_f, dB_t):
        assert gpu_common.rel_err(dB.cpu().numpy(), dB64, absB) <= tol
    for dv in (dv_f, dv_s):
        got = dv.cpu().numpy().astype(np.float64)
        assert float(np.max(np.abs(got - dv64) / (absv + 1e-30), initial=0)) <= tol
    M.close()


# banded_3000 and banded_mixed_4000 hold more entries than one chunk and one
# bucket of the dvals reordering (8,192 and 32,768 FP32 / 16,384 FP64 values)
@pytest.mark.parametrize("name", ["powerlaw_3000x2500", "densecol_400x100", "equalrows_100x70",
                                  "unsorted_rows_64x96", "narrow_50x20", "banded_3000",
                                  "banded_mixed_4000", "empty_10x40"])
@pytest.mark.parametrize("K", [32, 128])
@pytest.mark.parametrize("dt", ["f32", "f64"])
def test_backward(swift, oracle, cuda, name, K, dt):
    print("case", name, K, dt, flush=True)
    check_backward(swift, oracle, name, K, dt)


@pytest.mark.parametrize("K", [100, 128])
@pytest.mark.parametrize("dt", ["f32", "f64"])
def test_backward_reorder_many_buckets(swift, oracle, cuda, K, dt):
    """The benchmark matrix's structure at 1/20 scale (50,000 rows, ~800k
    entries: ~100 chunks and 25-50 buckets of the dvals reordering, the last
    ones partial); K = 100 adds later K slices onto the walk-order dvals
    before the reorder.  Fused dvals and dB against the FP64 oracle."""
    import torch
    d = swift.generate(2, scale=0.05, seed=3)
    m, n, rp, ci = d["m"], d["n"], d["rowptr"], d["colidx"]
    npdt = np.float32 if dt == "f32" else np.float64
    tol = 1e-5 if dt == "f32" else 1e-12
    v = d["val"].astype(npdt)
    print("case reorder", m, d["nnz"], K, dt, flush=True)
    rng = np.random.default_rng(K)
    B = rng.uniform(-1, 1, (n, K)).astype(npdt)
    dC = rng.uniform(-1, 1, (m, K)).astype(npdt)
    v64, B64, dC64 = v.astype(np.float64), B.astype(np.float64), dC.astype(np.float64)
    dv64 = oracle.sddmm_f64(m, n, rp, ci, dC64, B64)
    absv = oracle.sddmm_f64(m, n, rp, ci, np.abs(dC64), np.abs(B64))
    dB64 = oracle.spmm_t_f64(m, n, rp, ci, v64, dC64)
    absB = oracle.spmm_t_f64(m, n, rp, ci, np.abs(v64), np.abs(dC64))
    M = swift.SwiftMatrix(m, n, rp, ci, v, dtype=torch.float32 if dt == "f32" else torch.float64)
    dB, dv = M.backward_fused(torch.tensor(dC, device="cuda"), torch.tensor(B, device="cuda"))
    got = dv.cpu().numpy().astype(np.float64)
    err = float(np.max(np.abs(got - dv64) / (absv + 1e-30)))
    print("dvals err %.3g" % err, flush=True)
    assert err <= tol
    assert gpu_common.rel_err(dB.cpu().numpy(), dB64, absB) <= tol
    M.close()


# the K sweep of test/irregular_K (script/Run_irregular_K.sh) plus K = 100 and
# small K; each K runs as slices of one lane shape per launch, the part of K
# past the last full vector shape in masked scalar slices
@pytest.mark.parametrize("name", ["powerlaw_3000x2500", "densecol_400x100", "tall_2000x64"])
@pytest.mark.parametrize("K", [1, 3, 24, 48, 96, 100, 192, 384, 768])
@pytest.mark.parametrize("dt", ["f32", "f64"])
def test_backward_irregular_k(swift, oracle, cuda, name, K, dt):
    print("case", name, K, dt, flush=True)
    check_backward(swift, oracle, name, K, dt)


@pytest.mark.parametrize("name", ["powerlaw_3000x2500", "densecol_400x100", "tall_2000x64",
                                  "banded_mixed_4000"])
@pytest.mark.parametrize("K", [4, 32, 100, 128, 256])
def test_backward_bf16(swift, oracle, cuda, name, K):
    """BF16 storage of dC, B and A's values, FP32 accumulation (the fused
    backward, and SDDMM walking A, which reads no values): a product of
    two BF16 values is exact in FP32, so against the FP64 oracle on the
    BF16-rounded operands the error is that of FP32 sums (<= 1e-5 of the
    absolute sums); against the unrounded FP32 operands it is within the
    16-bit storage bound (<= 1e-2)."""
    import torch
    print("case bf16", name, K, flush=True)
    m, n, rp, ci, v = CASES[name]
    rng = np.random.default_rng(K + 7)
    vv = np.asarray(v, np.float32)
    B = rng.standard_normal((n, K)).astype(np.float32)
    dC = rng.standard_normal((m, K)).astype(np.float32)
    Bh = torch.tensor(B, device="cuda").to(torch.bfloat16)
    dCh = torch.tensor(dC, device="cuda").to(torch.bfloat16)
    rnd = lambda x: torch.tensor(x).to(torch.bfloat16).double().numpy()
    B_r, dC_r, v_r = rnd(B), rnd(dC), rnd(vv)
    M = swift.SwiftMatrix(m, n, rp, ci, vv, dtype=torch.float32)
    dB, dv = M.backward_fused(dCh, Bh)
    dv_s = M.sddmm(dCh, Bh)  # SDDMM walking A itself
    assert dB.dtype == torch.float32 and dv.dtype == torch.float32 and dv_s.dtype == torch.float32
    dB, dv = dB.cpu().numpy().astype(np.float64), dv.cpu().numpy().astype(np.float64)
    dv_s = dv_s.cpu().numpy().astype(np.float64)
    for (vals, bb, dd), tol in (((v_r, B_r, dC_r), 1e-5),
                                ((vv.astype(np.float64), B.astype(np.float64),
                                  dC.astype(np.float64)), 1e-2)):
        dB64 = oracle.spmm_t_f64(m, n, rp, ci, vals, dd)
        absB = oracle.spmm_t_f64(m, n, rp, ci, np.abs(vals), np.abs(dd))
        dv64 = oracle.sddmm_f64(m, n, rp, ci, dd, bb)
        absv = oracle.sddmm_f64(m, n, rp, ci, np.abs(dd), np.abs(bb))
        e1 = gpu_common.rel_err(dB, dB64, absB)
        e2 = float(np.max(np.abs(dv - dv64) / (absv + 1e-30), initial=0))
        e3 = float(np.max(np.abs(dv_s - dv64) / (absv + 1e-30), initial=0))
        print("tol %g: dB %.3g dvals %.3g sddmm %.3g" % (tol, e1, e2, e3), flush=True)
        assert e1 <= tol and e2 <= tol and e3 <= tol
    M.close()


@pytest.mark.parametrize("K", [32, 100])
@pytest.mark.parametrize("dt", ["f32", "f64"])
def test_backward_unaligned_operands(swift, oracle, cuda, K, dt):
    """dC and B one element off a 16-byte boundary: every slice is masked."""
    print("case unaligned", K, dt, flush=True)
    check_backward(swift, oracle, "powerlaw_3000x2500", K, dt, offset=1)


def test_transpose_layout_bit_exact(swift, oracle, cuda):
    """The transposed Swift layout equals the oracle's format of A^T."""
    import scipy.sparse as sp
    import torch
    m, n, rp, ci, v = CASES["powerlaw_3000x2500"]
    vv = np.asarray(v, np.float32)
    M = swift.SwiftMatrix(m, n, rp, ci, vv)
    T = M.transpose()
    At = sp.csr_matrix((vv.astype(np.float64), ci, rp), shape=(m, n)).T.tocsr()
    # CSR of A^T with rows in ascending column order = A's CSC (stable)
    ptr, row, val, _ = oracle.csr2csc(m, n, rp, ci, vv.astype(np.float64))
    ref = oracle.build(n, m, ptr, row, val)
    got = T.arrays()
    for k in ("perm", "sortnnz", "tile_colidx", "sortrowidx", "reside_cscptr",
              "reside_cscrowidx", "tile_format", "colorblk"):
        assert np.array_equal(got[k], np.asarray(ref[k]).astype(got[k].dtype)), k
    assert np.array_equal(got["sortval"], ref["sortval"].astype(np.float32))
    assert At.nnz == len(ci)
    M.close()


def test_update_values(swift, oracle, cuda):
    import torch
    m, n, rp, ci, v = CASES["densecol_400x100"]
    M = swift.SwiftMatrix(m, n, rp, ci, np.asarray(v, np.float32))
    v2 = np.random.default_rng(9).standard_normal(len(ci)).astype(np.float32)
    M.update_values(torch.tensor(v2, device="cuda"))
    B = np.random.default_rng(10).uniform(-1, 1, (n, 32)).astype(np.float32)
    C64, absC = oracle.spmm_f64(m, n, rp, ci, v2.astype(np.float64), B.astype(np.float64))
    C = M.spmm(torch.tensor(B, device="cuda")).cpu().numpy()
    assert gpu_common.rel_err(C, C64, absC) <= 1e-5
    M.close()


@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("K", [16, 24, 100])
def test_autograd_finite_differences(swift, cuda, fused, K):
    """d loss / d vals and d loss / d B of loss = <W, A(vals) B> against FP64
    central differences on a small power-law graph (induced subgraph size)."""
    import torch
    d = swift.generate(4, scale=0.01, seed=1)  # 2,000 nodes
    m, n, rp, ci, v = d["m"], d["n"], d["rowptr"], d["colidx"], d["val"]
    print("case fd", K, fused, flush=True)
    M = swift.SwiftMatrix(m, n, rp, ci, v, dtype=torch.float64)
    Fn = swift.spmm_function()
    rng = np.random.default_rng(0)
    vals = torch.tensor(v, device="cuda", requires_grad=True)
    B = torch.tensor(rng.standard_normal((n, K)), device="cuda", requires_grad=True)
    W = torch.tensor(rng.standard_normal((m, K)), device="cuda")
    loss = (Fn.apply(vals, B, M, fused) * W).sum()
    loss.backward()
    gv, gB = vals.grad.cpu().numpy(), B.grad.cpu().numpy()

    def f(vv, BB):
        M.update_values(torch.tensor(vv, device="cuda"))
        return float((M.spmm(torch.tensor(BB, device="cuda")) * W).sum())

    h = 1e-6
    Bn = B.detach().cpu().numpy()
    for e in rng.choice(len(v), 6, replace=False):
        vp, vm = v.copy(), v.copy()
        vp[e] += h
        vm[e] -= h
        fd = (f(vp, Bn) - f(vm, Bn)) / (2 * h)
        assert abs(fd - gv[e]) <= 1e-6 * max(1.0, abs(fd))
    for (i, k) in [(0, 0), (n // 2, 3), (n - 1, K - 1)]:
        Bp, Bm = Bn.copy(), Bn.copy()
        Bp[i, k] += h
        Bm[i, k] -= h
        fd = (f(v, Bp) - f(v, Bm)) / (2 * h)
        assert abs(fd - gB[i, k]) <= 1e-6 * max(1.0, abs(fd))
    M.close()


@pytest.mark.parametrize("K", [32, 128])
def test_autograd_bf16(swift, oracle, cuda, K):
    """The autograd wrapper with BF16 B (FP32 values): C from BF16 storage,
    gradients from the BF16 fused backward.  Against FP64 closed forms of
    loss = <W, A(vals) B>: d/dvals[e] = <W[row_e], B[col_e]> and d/dB = A^T W,
    with dC = W rounded to BF16 as the backward takes it: within 1e-2 of the
    absolute sums (16-bit storage of A's values, B and dC)."""
    import torch
    d = swift.generate(4, scale=0.01, seed=1)  # 2,000 nodes
    m, n, rp, ci, v = d["m"], d["n"], d["rowptr"], d["colidx"], d["val"].astype(np.float32)
    print("case autograd bf16", K, flush=True)
    M = swift.SwiftMatrix(m, n, rp, ci, v, dtype=torch.float32)
    Fn = swift.spmm_function()
    rng = np.random.default_rng(K)
    vals = torch.tensor(v, device="cuda", requires_grad=True)
    B = torch.tensor(rng.standard_normal((n, K)), device="cuda",
                     dtype=torch.float32).to(torch.bfloat16).requires_grad_(True)
    W = torch.tensor(rng.standard_normal((m, K)), device="cuda", dtype=torch.float32)
    C = Fn.apply(vals, B, M, True)
    assert C.dtype == torch.float32
    (C * W).sum().backward()
    assert B.grad.dtype == torch.bfloat16 and vals.grad.dtype == torch.float32
    v64, B64, W64 = v.astype(np.float64), B.detach().double().cpu().numpy(), \
        W.double().cpu().numpy()
    C64, absC = oracle.spmm_f64(m, n, rp, ci, v64, B64)
    assert gpu_common.rel_err(C.detach().cpu().numpy(), C64, absC) <= 1e-2
    gv64 = oracle.sddmm_f64(m, n, rp, ci, W64, B64)
    absv = oracle.sddmm_f64(m, n, rp, ci, np.abs(W64), np.abs(B64))
    gB64 = oracle.spmm_t_f64(m, n, rp, ci, v64, W64)
    absB = oracle.spmm_t_f64(m, n, rp, ci, np.abs(v64), np.abs(W64))
    ev = float(np.max(np.abs(vals.grad.double().cpu().numpy() - gv64) / (absv + 1e-30)))
    eB = gpu_common.rel_err(B.grad.double().cpu().numpy(), gB64, absB)
    print("dvals err %.3g dB err %.3g" % (ev, eB), flush=True)
    assert ev <= 1e-2 and eB <= 1e-2
    M.close()


def _gnn_graph(swift):
    d = swift.generate(4, scale=1.0, seed=0)  # 200,000 nodes, ~10.2M edges
    return d["m"], d["rowptr"].astype(np.int64), d["colidx"].astype(np.int64), d["val"]


@pytest.mark.parametrize("fused", [True, False])
def test_gnn_induced_subgraph_gradients(swift, cuda, fused):
    """The 2,000-node subgraph induced by the highest-degree nodes of the
    200k-node graph, with the full graph's D^-1/2 (A+I) D^-1/2 values: the
    autograd wrapper's gradients of loss = <W, A(vals) B> against dense FP64
    torch autograd on the CPU, then against FP64 central differences."""
    import torch
    n_all, rp, ci, v = _gnn_graph(swift)
    deg = np.diff(rp)
    nodes = np.sort(np.argsort(-deg, kind="stable")[:2000])
    new = np.full(n_all, -1, np.int64)
    new[nodes] = np.arange(len(nodes))
    row = np.repeat(np.arange(n_all), deg)
    keep = (new[row] >= 0) & (new[ci] >= 0)
    r2, c2, v2 = new[row[keep]], new[ci[keep]], v[keep]
    n = len(nodes)
    rp2 = np.zeros(n + 1, np.int64)
    np.cumsum(np.bincount(r2, minlength=n), out=rp2[1:])
    assert len(v2) > 20 * n  # the hubs are densely connected among themselves
    K = 16
    rng = np.random.default_rng(0)
    B0 = rng.standard_normal((n, K))
    W0 = rng.standard_normal((n, K))
    M = swift.SwiftMatrix(n, n, rp2.astype(np.int32), c2.astype(np.int32), v2,
                          dtype=torch.float64)
    Fn = swift.spmm_function()
    vals = torch.tensor(v2, device="cuda", requires_grad=True)
    B = torch.tensor(B0, device="cuda", requires_grad=True)
    W = torch.tensor(W0, device="cuda")
    (Fn.apply(vals, B, M, fused) * W).sum().backward()
    gv, gB = vals.grad.cpu().numpy(), B.grad.cpu().numpy()
    # dense reference
    rv = torch.tensor(v2, dtype=torch.float64, requires_grad=True)
    rB = torch.tensor(B0, requires_grad=True)
    A = torch.zeros((n, n), dtype=torch.float64).index_put(
        (torch.tensor(r2), torch.tensor(c2)), rv, accumulate=True)
    (torch.mm(A, rB) * torch.tensor(W0)).sum().backward()
    np.testing.assert_allclose(gv, rv.grad.numpy(), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(gB, rB.grad.numpy(), rtol=1e-12, atol=1e-12)

    def f(vv, BB):
        M.update_values(torch.tensor(vv, device="cuda"))
        return float((M.spmm(torch.tensor(BB, device="cuda")) * W).sum())

    h = 1e-6
    for e in rng.choice(len(v2), 6, replace=False):
        vp, vm = v2.copy(), v2.copy()
        vp[e] += h
        vm[e] -= h
        fd = (f(vp, B0) - f(vm, B0)) / (2 * h)
        assert abs(fd - gv[e]) <= 1e-6 * max(1.0, abs(fd))
    M.close()


@pytest.mark.slow
@pytest.mark.parametrize("fused", [True, False])
def test_gnn_full_adjoint_identity(swift, cuda, fused):
    """At the full 200k-node size (FP32): <dC, A B> = <A^T dC, B> =
    sum_e vals_e dvals_e, summed in FP64.  The three agree to ~1.4e-11 of
    S = <|dC|, |A| |B|> (~1e7 here) while the sums themselves are ~170 (random
    signs cancel); the bound 1e-9 S (~0.01) is tight enough that one dropped
    or doubled edge (~0.2) fails it."""
    import torch
    n, rp, ci, v = _gnn_graph(swift)
    K = 128
    M = swift.SwiftMatrix(n, n, rp.astype(np.int32), ci.astype(np.int32),
                          v.astype(np.float32), dtype=torch.float32)
    Mabs = swift.SwiftMatrix(n, n, rp.astype(np.int32), ci.astype(np.int32),
                             np.abs(v).astype(np.float32), dtype=torch.float32)
    g = torch.Generator(device="cuda").manual_seed(3)
    B = torch.randn((n, K), device="cuda", generator=g)
    dC = torch.randn((n, K), device="cuda", generator=g)
    C = M.spmm(B)
    if fused:
        dB, dv = M.backward_fused(dC, B)
    else:
        dB, dv = M.transpose().spmm(dC), M.sddmm(dC, B)
    s1 = float((dC.double() * C.double()).sum())
    s2 = float((dB.double() * B.double()).sum())
    s3 = float((torch.tensor(v, device="cuda") * dv.double()).sum())
    scale = float((dC.abs().double() * Mabs.spmm(B.abs()).double()).sum())
    print("adjoint", s1, s2, s3, "scale", scale, flush=True)
    assert abs(s1 - s2) <= 1e-9 * scale
    assert abs(s1 - s3) <= 1e-9 * scale
    Mabs.close()
    M.close()
