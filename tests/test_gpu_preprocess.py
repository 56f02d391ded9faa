"""GPU preprocessing is bit-exact with the CPU oracle (and through it with the
reference's CPU code): CSC, column sort, staircase tiles, reside CSC,
categorisation, balance and the row sort, on edge-case matrices and on the
full-size configurations."""
import numpy as np
import pytest

import gpu_common

pytestmark = pytest.mark.gpu

FIELDS = ("perm", "sortnnz", "sortrowidx_tmp", "tile_ptr", "tile_len", "tile_colidx",
          "tile_tall", "tile_format", "sortrowidx", "countrow", "segmentoffset",
          "sortrowindex", "reside_cscptr", "reside_cscrowidx", "colorblk", "blkstart",
          "blkstop", "blkptr")
SCALARS = ("tilenum", "regular_nnz", "reside_col", "reside_nnz", "segsum", "numblk")
CASES = gpu_common.cases()


def compare(swift, oracle, torch, m, n, rp, ci, v, dtype):
    npdt = np.float64 if dtype == torch.float64 else np.float32
    vv = np.asarray(v, npdt)
    ref = oracle.build(m, n, rp, ci, vv.astype(np.float64))
    M = swift.SwiftMatrix(m, n, rp, ci, vv, dtype=dtype)
    got = M.arrays()
    for k in SCALARS:
        assert got[k] == ref[k], k
    for k in FIELDS:
        assert np.array_equal(got[k], np.asarray(ref[k]).astype(got[k].dtype)), k
    for k, rk in (("sortval", "sortval"), ("reside_val", "reside_val"),
                  ("sortval_tmp", "sortval_tmp")):
        assert np.array_equal(got[k], np.asarray(ref[rk]).astype(npdt)), k
    # entry sources map every stored entry back to its CSR position
    assert np.array_equal(np.asarray(vv)[got["sortsrc"]], got["sortval"])
    assert np.array_equal(np.asarray(vv)[got["reside_src"]], got["reside_val"])
    return M, ref


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("dt", ["f32", "f64"])
def test_build_bit_exact(swift, oracle, cuda, name, dt):
    import torch
    print("case", name, dt, flush=True)
    m, n, rp, ci, v = CASES[name]
    dtype = torch.float32 if dt == "f32" else torch.float64
    M, ref = compare(swift, oracle, torch, m, n, rp, ci, v, dtype)
    M.close()


@pytest.mark.parametrize("name", list(CASES))
def test_row_sort_bit_exact(swift, oracle, cuda, name):
    import torch
    print("case", name, flush=True)
    m, n, rp, ci, v = CASES[name]
    M = swift.SwiftMatrix(m, n, rp, ci, v, dtype=torch.float32)
    M.row_idx_sort()
    ref = oracle.build(m, n, rp, ci, np.asarray(v, np.float32).astype(np.float64))
    ref.row_idx_sort()
    got = M.arrays()
    for k in ("val_position", "val_offset", "rowsorted"):
        assert np.array_equal(got[k], ref[k]), k
    M.close()


def test_stage_entry_points(swift, oracle, cuda):
    """swift_csr_to_csc and swift_col_sort as standalone calls."""
    import ctypes as C
    import torch
    m, n, rp, ci, v = CASES["densecol_400x100"]
    dev = torch.device("cuda")
    rp_t = torch.tensor(rp, device=dev)
    ci_t = torch.tensor(ci, device=dev)
    v_t = torch.tensor(v, device=dev, dtype=torch.float64)
    nnz = len(ci)
    cptr = torch.empty(n + 1, dtype=torch.int32, device=dev)
    crow = torch.empty(nnz, dtype=torch.int32, device=dev)
    cval = torch.empty(nnz, dtype=torch.float64, device=dev)
    npc = torch.empty(n, dtype=torch.int32, device=dev)
    L = swift.lib()
    s = swift._stream()
    swift.check("swift_csr_to_csc", L.swift_csr_to_csc(
        m, n, nnz, swift.F64, C.c_void_p(v_t.data_ptr()), C.c_void_p(rp_t.data_ptr()),
        C.c_void_p(ci_t.data_ptr()), C.c_void_p(cval.data_ptr()), C.c_void_p(crow.data_ptr()),
        C.c_void_p(cptr.data_ptr()), C.c_void_p(npc.data_ptr()), s))
    torch.cuda.synchronize()
    want = oracle.csr2csc(m, n, rp, ci, v)
    for got, w in zip((cptr, crow, cval, npc), want):
        assert np.array_equal(got.cpu().numpy(), w)
    K = 3
    B = torch.arange(n * K, dtype=torch.float64, device=dev).reshape(n, K)
    sB = torch.empty_like(B)
    sr = torch.empty(nnz, dtype=torch.int32, device=dev)
    sv = torch.empty(nnz, dtype=torch.float64, device=dev)
    sn = torch.empty(n, dtype=torch.int32, device=dev)
    perm = torch.empty(n, dtype=torch.int32, device=dev)
    swift.check("swift_col_sort", L.swift_col_sort(
        n, nnz, K, swift.F64, C.c_void_p(npc.data_ptr()), C.c_void_p(cptr.data_ptr()),
        C.c_void_p(crow.data_ptr()), C.c_void_p(cval.data_ptr()), C.c_void_p(sr.data_ptr()),
        C.c_void_p(sv.data_ptr()), C.c_void_p(sn.data_ptr()), C.c_void_p(perm.data_ptr()),
        C.c_void_p(B.data_ptr()), C.c_void_p(sB.data_ptr()), s))
    torch.cuda.synchronize()
    wperm, wsn, wsr, wsv = oracle.col_sort(n, *want[3:4], want[0], want[1], want[2])
    assert np.array_equal(perm.cpu().numpy(), wperm)
    assert np.array_equal(sn.cpu().numpy(), wsn)
    assert np.array_equal(sr.cpu().numpy(), wsr)
    assert np.array_equal(sv.cpu().numpy(), wsv)
    assert np.array_equal(sB.cpu().numpy(), B.cpu().numpy()[wperm])


@pytest.mark.parametrize("which", [1, 4])
def test_config_build_bit_exact(swift, oracle, cuda, which):
    """Configuration matrices (C1 full size, C4 at 1/10 scale)."""
    import torch
    d = swift.generate(which, scale=1.0 if which == 1 else 0.1, seed=0)
    print("config", which, d["m"], d["nnz"], flush=True)
    M, ref = compare(swift, oracle, torch, d["m"], d["n"], d["rowptr"], d["colidx"], d["val"],
                     torch.float32)
    M.close()


@pytest.mark.slow
def test_config2_full_build_bit_exact(swift, oracle, cuda):
    """The 10^6 x 10^6, 16M-nonzero matrix: every index array bit-exact."""
    import torch
    d = swift.generate(2, scale=1.0, seed=0)
    print("config 2 full", d["m"], d["nnz"], flush=True)
    M, ref = compare(swift, oracle, torch, d["m"], d["n"], d["rowptr"], d["colidx"], d["val"],
                     torch.float32)
    M.close()


@pytest.mark.parametrize("flaw", ["col_high", "col_negative", "rowptr_decreasing",
                                  "rowptr_end", "rowptr_start"])
def test_device_csr_is_checked(swift, cuda, flaw):
    """swift_build and swift_csr_to_csc take device CSR arrays and scatter by
    their indices: a bad column or row pointer must come back as
    SWIFT_ERR_ARG from the read-only check pass, before any scatter runs."""
    import ctypes as C
    import torch
    m, n, rp, ci, v = CASES["random_300x200"]
    rp, ci = rp.copy(), ci.copy()
    if flaw == "col_high":
        ci[len(ci) // 2] = n + 1000
    elif flaw == "col_negative":
        ci[3] = -1
    elif flaw == "rowptr_decreasing":
        rp[10] = rp[12] + 1
    elif flaw == "rowptr_end":
        rp[-1] -= 1
    else:
        rp[0] = 1
    print("case", flaw, flush=True)
    with pytest.raises(swift.SwiftError) as e:
        swift.SwiftMatrix(m, n, rp, ci, v, dtype=torch.float32)
    assert e.value.code == -10
    dev = torch.device("cuda")
    d_rp = torch.tensor(rp, device=dev)
    d_ci = torch.tensor(ci, device=dev)
    nnz = len(ci)
    cscptr = torch.empty(n + 1, dtype=torch.int32, device=dev)
    cscrow = torch.empty(max(nnz, 1), dtype=torch.int32, device=dev)
    npc = torch.empty(n, dtype=torch.int32, device=dev)
    P = lambda t: C.c_void_p(t.data_ptr())
    rc = swift.lib().swift_csr_to_csc(m, n, nnz, swift.F32, None, P(d_rp), P(d_ci), None,
                                      P(cscrow), P(cscptr), P(npc), swift._stream())
    assert rc == -10
    # the same arrays without the flaw still build
    m, n, rp, ci, v = CASES["random_300x200"]
    swift.SwiftMatrix(m, n, rp, ci, v, dtype=torch.float32).close()
