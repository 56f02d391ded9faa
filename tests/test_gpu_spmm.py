"""GPU SpMM parity: every layout / kernel / value type / K against the FP64
oracle, per element, with the error normalised by (|A||B|)[i,k]:
    FP32 <= 1e-5, FP64 <= 1e-12, BF16/F16 storage <= 1e-2,
and against the reference CPU path (Swift_cpu, FP32) to the same bound."""
import numpy as np
import pytest

import gpu_common

pytestmark = pytest.mark.gpu

CASES = gpu_common.cases()
TOL = {"f32": 1e-5, "f64": 1e-12}


def run_case(swift, oracle, name, K, dt, layout, kernel, flags=0, seed=0):
    import torch
    m, n, rp, ci, v = CASES[name]
    dtype = torch.float32 if dt == "f32" else torch.float64
    npdt = np.float32 if dt == "f32" else np.float64
    vv = np.asarray(v, npdt)
    rng = np.random.default_rng(seed + K)
    B = rng.uniform(-1, 1, (n, K)).astype(npdt)
    C64, absC = oracle.spmm_f64(m, n, rp, ci, vv.astype(np.float64), B.astype(np.float64))
    M = swift.SwiftMatrix(m, n, rp, ci, vv, dtype=dtype)
    Bt = torch.tensor(B, device="cuda")
    if flags & swift.B_SORTED:
        Bt = Bt[torch.tensor(M.arrays(["perm"])["perm"].astype(np.int64), device="cuda")]
    if layout == swift.COL_MAJOR:
        Bt = Bt.t().contiguous()
    C = M.spmm(Bt, layout=layout, kernel=kernel, flags=flags)
    torch.cuda.synchronize()
    C = C.cpu().numpy()
    if layout == swift.COL_MAJOR:
        C = C.T
    M.close()
    return C, C64, absC


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("K", [32, 128])
@pytest.mark.parametrize("dt", ["f32", "f64"])
def test_row_major(swift, oracle, cuda, name, K, dt):
    print("case", name, K, dt, flush=True)
    C, C64, absC = run_case(swift, oracle, name, K, dt, swift.ROW_MAJOR, swift.KERNEL_DEFAULT)
    assert gpu_common.rel_err(C, C64, absC) <= TOL[dt]


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("K", [32, 128])
@pytest.mark.parametrize("dt", ["f32", "f64"])
def test_col_major(swift, oracle, cuda, name, K, dt):
    print("case", name, K, dt, flush=True)
    C, C64, absC = run_case(swift, oracle, name, K, dt, swift.COL_MAJOR, swift.KERNEL_DEFAULT)
    assert gpu_common.rel_err(C, C64, absC) <= TOL[dt]


@pytest.mark.parametrize("kernel", ["BALANCED", "UNBALANCED", "KCHUNK", "ROWVIEW", "FLAT",
                                    "WINDOW"])
@pytest.mark.parametrize("layout", ["ROW", "COL"])
@pytest.mark.parametrize("name", ["densecol_400x100", "equalrows_100x70", "powerlaw_3000x2500",
                                  "empty_10x40", "uniform_64x64"])
def test_kernel_variants(swift, oracle, cuda, kernel, layout, name):
    print("case", name, kernel, layout, flush=True)
    if layout == "ROW" and kernel == "KCHUNK":
        pytest.skip("K chunks are a column-major grid dimension")
    if layout == "COL" and kernel in ("ROWVIEW", "FLAT", "WINDOW"):
        pytest.skip("the row-view walks are row-major kernels")
    k = getattr(swift, "KERNEL_" + kernel)
    lay = swift.ROW_MAJOR if layout == "ROW" else swift.COL_MAJOR
    C, C64, absC = run_case(swift, oracle, name, 64, "f32", lay, k)
    assert gpu_common.rel_err(C, C64, absC) <= 1e-5


@pytest.mark.parametrize("name", ["equalrows_100x70", "powerlaw_3000x2500", "random_300x200"])
@pytest.mark.parametrize("dt", ["f32", "f64"])
def test_rowsort_kernel(swift, oracle, cuda, name, dt):
    """Swift_GPU_4's segmented reduction over row-sorted groups of 8 tiles."""
    import torch
    print("case", name, dt, flush=True)
    m, n, rp, ci, v = CASES[name]
    dtype = torch.float32 if dt == "f32" else torch.float64
    npdt = np.float32 if dt == "f32" else np.float64
    vv = np.asarray(v, npdt)
    B = np.random.default_rng(7).uniform(-1, 1, (n, 96)).astype(npdt)
    C64, absC = oracle.spmm_f64(m, n, rp, ci, vv.astype(np.float64), B.astype(np.float64))
    M = swift.SwiftMatrix(m, n, rp, ci, vv, dtype=dtype)
    M.row_idx_sort()
    C = M.spmm(torch.tensor(B.T.copy(), device="cuda"), layout=swift.COL_MAJOR,
               kernel=swift.KERNEL_ROWSORT).cpu().numpy().T
    assert gpu_common.rel_err(C, C64, absC) <= TOL[dt]
    M.close()


@pytest.mark.parametrize("K", [1, 3, 8, 24, 48, 96, 100, 192, 256, 384, 512, 768])
@pytest.mark.parametrize("layout", ["ROW", "COL"])
def test_irregular_k(swift, oracle, cuda, K, layout):
    """K sweep of test/irregular_K (24..768) plus non-multiples of the vector width."""
    print("case K", K, layout, flush=True)
    lay = swift.ROW_MAJOR if layout == "ROW" else swift.COL_MAJOR
    C, C64, absC = run_case(swift, oracle, "powerlaw_500x333", K, "f32", lay,
                            swift.KERNEL_DEFAULT)
    assert gpu_common.rel_err(C, C64, absC) <= 1e-5


@pytest.mark.parametrize("layout", ["ROW", "COL"])
@pytest.mark.parametrize("dt", ["f32", "f64"])
def test_b_sorted_flag(swift, oracle, cuda, layout, dt):
    """B already in sorted column order (the reference's sort_dense_mtx)."""
    lay = swift.ROW_MAJOR if layout == "ROW" else swift.COL_MAJOR
    C, C64, absC = run_case(swift, oracle, "powerlaw_3000x2500", 32, dt, lay,
                            swift.KERNEL_DEFAULT, flags=swift.B_SORTED)
    assert gpu_common.rel_err(C, C64, absC) <= TOL[dt]


# Which entries the row view holds decides the launches of a row-view product:
# powerlaw_3000x2500 (reside part under 1% of its entries) walks every entry
# in the row view; tall_2000x64 (12% reside) runs its reside blocks after the
# walk (in the tail launch of the row-view walks, in a launch of their own
# after the flat walk).
ROWVIEW_HOLDS_ALL = {"powerlaw_3000x2500": True, "tall_2000x64": False}


def assert_rowview_path(M, name):
    inf = M.info()
    if ROWVIEW_HOLDS_ALL[name]:
        assert inf.rowview_nnz == inf.nnz and inf.reside_nnz > 0
    else:
        assert inf.rowview_nnz == inf.regular_nnz < inf.nnz


@pytest.mark.parametrize("kernel", ["DEFAULT", "BALANCED", "ROWVIEW", "FLAT", "WINDOW"])
@pytest.mark.parametrize("K", [32, 100])
@pytest.mark.parametrize("layout", ["ROW", "COL"])
@pytest.mark.parametrize("name", list(ROWVIEW_HOLDS_ALL))
def test_accumulate_flag(swift, oracle, cuda, kernel, K, layout, name):
    import torch
    if layout == "COL" and kernel in ("ROWVIEW", "FLAT", "WINDOW"):
        pytest.skip("the row-view walks are row-major kernels")
    print("case", name, kernel, K, layout, flush=True)
    m, n, rp, ci, v = CASES[name]
    vv = np.asarray(v, np.float32)
    B = np.random.default_rng(1).uniform(-1, 1, (n, K)).astype(np.float32)
    C0 = np.random.default_rng(2).uniform(-1, 1, (m, K)).astype(np.float32)
    C64, absC = oracle.spmm_f64(m, n, rp, ci, vv.astype(np.float64), B.astype(np.float64))
    M = swift.SwiftMatrix(m, n, rp, ci, vv)
    assert_rowview_path(M, name)
    col = layout == "COL"
    out = torch.tensor(C0.T.copy() if col else C0, device="cuda")
    M.spmm(torch.tensor(B.T.copy() if col else B, device="cuda"), out=out,
           flags=swift.ACCUMULATE, kernel=getattr(swift, "KERNEL_" + kernel),
           layout=swift.COL_MAJOR if col else swift.ROW_MAJOR)
    got = out.cpu().numpy().astype(np.float64)
    if col:
        got = got.T
    assert gpu_common.rel_err(got - C0, C64, absC + np.abs(C0)) <= 1e-5
    M.close()


@pytest.mark.parametrize("name", ["random_300x200", "densecol_400x100", "powerlaw_3000x2500"])
def test_matches_reference_cpu_path(swift, oracle, cuda, name):
    """Against Swift_cpu (src/Swift_cpu.h:31-65) in FP32 on the same inputs."""
    import torch
    m, n, rp, ci, v = CASES[name]
    vv = np.asarray(v, np.float32)
    f = oracle.build(m, n, rp, ci, vv.astype(np.float64))
    B = np.random.default_rng(3).uniform(-1, 1, (n, 32)).astype(np.float32)
    C_cpu = f.swift_cpu(32, B[f["perm"]], np.float32)
    _, absC = oracle.spmm_f64(m, n, rp, ci, vv.astype(np.float64), B.astype(np.float64))
    M = swift.SwiftMatrix(m, n, rp, ci, vv)
    C = M.spmm(torch.tensor(B, device="cuda")).cpu().numpy()
    assert gpu_common.rel_err(C, C_cpu.astype(np.float64), absC) <= 2e-5
    M.close()


def test_host_launcher(swift, oracle, cuda):
    """swift_spmm_host keeps the reference launcher contract: sorted host B in,
    host C out, ResultVerify at 1e-2, ms = -1 on a mismatch."""
    import ctypes as C
    m, n, rp, ci, v = CASES["powerlaw_3000x2500"]
    K = 32
    vv = np.asarray(v, np.float64)
    f = oracle.build(m, n, rp, ci, vv)
    B = np.random.default_rng(4).uniform(-1, 1, (n, K))
    golden, _ = oracle.spmm_f64(m, n, rp, ci, vv, B)
    M = swift.SwiftMatrix(m, n, rp, ci, vv, dtype=__import__("torch").float64)
    sB = np.ascontiguousarray(B[f["perm"]])
    out = np.zeros((m, K))
    ms, gf = C.c_float(), C.c_double()
    L = swift.lib()
    rc = L.swift_spmm_host(M.h, K, swift.ROW_MAJOR, swift.F64, swift.KERNEL_DEFAULT,
                           sB.ctypes.data, out.ctypes.data, golden.ctypes.data, 1, 3,
                           C.byref(ms), C.byref(gf))
    assert rc == 0 and ms.value > 0 and gf.value > 0
    np.testing.assert_allclose(out, golden, rtol=1e-10, atol=1e-10)
    bad = golden.copy()
    bad[0, 0] += 1.0
    rc = L.swift_spmm_host(M.h, K, swift.ROW_MAJOR, swift.F64, swift.KERNEL_DEFAULT,
                           sB.ctypes.data, out.ctypes.data, bad.ctypes.data, 0, 1,
                           C.byref(ms), C.byref(gf))
    assert rc == -14 and ms.value == -1.0
    M.close()


def test_rowview_col_major_rejected(swift, cuda):
    import torch
    m, n, rp, ci, v = CASES["random_300x200"]
    M = swift.SwiftMatrix(m, n, rp, ci, v, dtype=torch.float32)
    for k in (swift.KERNEL_ROWVIEW, swift.KERNEL_FLAT, swift.KERNEL_WINDOW):
        with pytest.raises(swift.SwiftError):
            M.spmm(torch.zeros((32, n), device="cuda"), layout=swift.COL_MAJOR, kernel=k)
    with pytest.raises(swift.SwiftError):
        M.spmm(torch.zeros((n, 32), device="cuda"), kernel=99)
    M.close()


@pytest.mark.parametrize("W", [8, 24, 128, 512])
@pytest.mark.parametrize("K", [32, 100, 128, 256])
@pytest.mark.parametrize("name", list(CASES))
def test_flat_walk(swift, oracle, cuda, name, K, W):
    """Lane-group walkers over segments of W items (W = 8 splits almost every
    row, leaves walkers without entries); C starts as NaN so a row that is not
    written (empty, split or stored) shows up."""
    import torch
    print("case", name, K, W, flush=True)
    m, n, rp, ci, v = CASES[name]
    vv = np.asarray(v, np.float32)
    B = np.random.default_rng(K + W).uniform(-1, 1, (n, K)).astype(np.float32)
    C64, absC = oracle.spmm_f64(m, n, rp, ci, vv.astype(np.float64), B.astype(np.float64))
    M = swift.SwiftMatrix(m, n, rp, ci, vv, dtype=torch.float32)
    M.set_segment_items(W)
    out = torch.full((m, K), float("nan"), device="cuda")
    M.spmm(torch.tensor(B, device="cuda"), out=out, kernel=swift.KERNEL_FLAT)
    C = out.cpu().numpy()
    assert not np.isnan(C).any()
    assert gpu_common.rel_err(C, C64, absC) <= 1e-5
    M.close()


@pytest.mark.parametrize("W", [24, 96, 512])
@pytest.mark.parametrize("K", [32, 100, 128, 256])
@pytest.mark.parametrize("name", list(CASES))
def test_window_walk(swift, oracle, cuda, name, K, W):
    """Workgroups of 16 segments of W items that stage their most reused B
    rows in LDS (K = 100 runs three K slices, two of them windowed, at
    nonzero K offsets); C starts as NaN so a row left unwritten shows up."""
    import torch
    print("case", name, K, W, flush=True)
    m, n, rp, ci, v = CASES[name]
    vv = np.asarray(v, np.float32)
    B = np.random.default_rng(K + W).uniform(-1, 1, (n, K)).astype(np.float32)
    C64, absC = oracle.spmm_f64(m, n, rp, ci, vv.astype(np.float64), B.astype(np.float64))
    M = swift.SwiftMatrix(m, n, rp, ci, vv, dtype=torch.float32)
    M.set_segment_items(W)
    out = torch.full((m, K), float("nan"), device="cuda")
    M.spmm(torch.tensor(B, device="cuda"), out=out, kernel=swift.KERNEL_WINDOW)
    C = out.cpu().numpy()
    assert not np.isnan(C).any()
    assert gpu_common.rel_err(C, C64, absC) <= 1e-5
    M.close()


@pytest.mark.parametrize("name", ["banded_3000", "banded_mixed_4000", "powerlaw_3000x2500"])
@pytest.mark.parametrize("K", [16, 64, 128])
def test_window_walk_f64_sorted(swift, oracle, cuda, name, K):
    """FP64 window walk, B in original and in sorted column order (the
    window is then chosen in sorted order)."""
    import torch
    print("case", name, K, flush=True)
    m, n, rp, ci, v = CASES[name]
    vv = np.asarray(v, np.float64)
    B = np.random.default_rng(K).uniform(-1, 1, (n, K))
    C64, absC = oracle.spmm_f64(m, n, rp, ci, vv, B)
    M = swift.SwiftMatrix(m, n, rp, ci, vv, dtype=torch.float64)
    perm = torch.tensor(M.arrays(["perm"])["perm"].astype(np.int64), device="cuda")
    Bt = torch.tensor(B, device="cuda")
    out = torch.full((m, K), float("nan"), device="cuda", dtype=torch.float64)
    M.spmm(Bt, out=out, kernel=swift.KERNEL_WINDOW)
    assert gpu_common.rel_err(out.cpu().numpy(), C64, absC) <= 1e-12
    out.fill_(float("nan"))
    M.spmm(Bt[perm].contiguous(), out=out, kernel=swift.KERNEL_WINDOW, flags=swift.B_SORTED)
    assert gpu_common.rel_err(out.cpu().numpy(), C64, absC) <= 1e-12
    M.close()


@pytest.mark.parametrize("name", ["powerlaw_3000x2500", "unsorted_rows_64x96", "empty_10x40"])
@pytest.mark.parametrize("K", [8, 64, 100])
def test_flat_walk_f64_sorted(swift, oracle, cuda, name, K):
    import torch
    m, n, rp, ci, v = CASES[name]
    vv = np.asarray(v, np.float64)
    B = np.random.default_rng(K).uniform(-1, 1, (n, K))
    C64, absC = oracle.spmm_f64(m, n, rp, ci, vv, B)
    M = swift.SwiftMatrix(m, n, rp, ci, vv, dtype=torch.float64)
    M.set_segment_items(24)
    perm = torch.tensor(M.arrays(["perm"])["perm"].astype(np.int64), device="cuda")
    Bt = torch.tensor(B, device="cuda")
    out = torch.full((m, K), float("nan"), device="cuda", dtype=torch.float64)
    M.spmm(Bt, out=out, kernel=swift.KERNEL_FLAT)
    assert gpu_common.rel_err(out.cpu().numpy(), C64, absC) <= 1e-12
    out.fill_(float("nan"))
    M.spmm(Bt[perm].contiguous(), out=out, kernel=swift.KERNEL_FLAT, flags=swift.B_SORTED)
    assert gpu_common.rel_err(out.cpu().numpy(), C64, absC) <= 1e-12
    M.close()


@pytest.mark.parametrize("name", list(ROWVIEW_HOLDS_ALL))
def test_output_has_no_stale_values(swift, oracle, cuda, name):
    """C = A*B overwrites every row (the row view stores rows, no zero pass):
    fill C with garbage first."""
    import torch
    m, n, rp, ci, v = CASES[name]
    vv = np.asarray(v, np.float32)
    B = np.random.default_rng(5).uniform(-1, 1, (n, 128)).astype(np.float32)
    C64, absC = oracle.spmm_f64(m, n, rp, ci, vv.astype(np.float64), B.astype(np.float64))
    M = swift.SwiftMatrix(m, n, rp, ci, vv)
    assert_rowview_path(M, name)
    out = torch.full((m, 128), 1e30, device="cuda")
    M.spmm(torch.tensor(B, device="cuda"), out=out)
    assert gpu_common.rel_err(out.cpu().numpy(), C64, absC) <= 1e-5
    M.close()


def test_dtype_mismatch_rejected(swift, cuda):
    import torch
    m, n, rp, ci, v = CASES["random_300x200"]
    M = swift.SwiftMatrix(m, n, rp, ci, v, dtype=torch.float32)
    with pytest.raises(swift.SwiftError):
        M.spmm(torch.zeros((n, 32), dtype=torch.float64, device="cuda"))
    with pytest.raises(swift.SwiftError):
        M.spmm(torch.zeros((n, 32), dtype=torch.bfloat16, device="cuda"))  # not prepared
    M.close()


@pytest.mark.slow
@pytest.mark.parametrize("layout", ["ROW", "COL"])
def test_config2_full(swift, oracle, cuda, layout):
    """10^6 x 10^6, 16M nonzeros: K=32 against the FP64 oracle element by
    element; K=128 through the column-sum checksum 1^T C = (1^T A) B and
    linearity A(B1 + B2) = A B1 + A B2."""
    import torch
    d = swift.generate(2, scale=1.0, seed=0)
    m, n, rp, ci = d["m"], d["n"], d["rowptr"], d["colidx"]
    vv = d["val"].astype(np.float32)
    print("config2 full", layout, d["nnz"], flush=True)
    lay = swift.ROW_MAJOR if layout == "ROW" else swift.COL_MAJOR
    M = swift.SwiftMatrix(m, n, rp, ci, vv)
    B = swift.generate_dense(n * 32, 0, -1.0, 1.0, seed=5).reshape(n, 32).astype(np.float32)
    C64, absC = oracle.spmm_f64(m, n, rp, ci, vv.astype(np.float64), B.astype(np.float64))
    Bt = torch.tensor(B if lay == swift.ROW_MAJOR else B.T.copy(), device="cuda")
    C = M.spmm(Bt, layout=lay).cpu().numpy()
    if lay == swift.COL_MAJOR:
        C = C.T
    assert gpu_common.rel_err(C, C64, absC) <= 1e-5
    del C64, absC, C
    K = 128
    g = torch.Generator(device="cuda").manual_seed(0)
    B1 = torch.rand((n, K), device="cuda", generator=g) * 2 - 1
    B2 = torch.rand((n, K), device="cuda", generator=g) * 2 - 1
    T = (lambda x: x) if lay == swift.ROW_MAJOR else (lambda x: x.t().contiguous())
    C1 = T(M.spmm(T(B1), layout=lay))
    C2 = T(M.spmm(T(B2), layout=lay))
    C12 = T(M.spmm(T(B1 + B2), layout=lay))
    colsum = np.bincount(ci, weights=vv.astype(np.float64), minlength=n)
    want = torch.tensor(colsum, device="cuda") @ B1.double()
    got = C1.double().sum(0)
    absA = np.bincount(ci, weights=np.abs(vv.astype(np.float64)), minlength=n)
    scale = torch.tensor(absA, device="cuda") @ B1.double().abs()
    assert float(((got - want).abs() / (scale + 1e-30)).max()) <= 1e-5
    Mabs = swift.SwiftMatrix(m, n, rp, ci, np.abs(vv))
    S = T(Mabs.spmm(T(B1.abs() + B2.abs()), layout=lay)).double()
    lin = (C12 - C1 - C2).abs().double() / (S + 1e-30)
    assert float(lin.max()) <= 3e-5
    Mabs.close()
    M.close()
