"""swift_spmm_noalloc: swift_spmm without device allocation.

The column-major default product transposes B and C through a workspace, and
the row-view walks keep the partial sums of rows that segment boundaries cut
in carry slots; swift_spmm grows these per matrix handle (the carry slots per
stream) with hipMalloc, swift_spmm_noalloc takes them from the library's
per-GPU scratch array.  Checked through the C ABI:
two calls in a row (the second on a smaller input, so stale scratch contents
from the first would show) against swift_spmm's output and against the FP64
oracle at the FP32 / FP64 bounds; row-major products; the size limit; and that
the call leaves the device's free memory unchanged where swift_spmm's first
column-major call on a fresh handle takes its workspace."""
import ctypes as C

import numpy as np
import pytest

import gpu_common

pytestmark = pytest.mark.gpu

LIMIT = 256 << 20  # SWIFT_NOALLOC_SCRATCH_BYTES


def _matrix(swift, oracle, scale, K, dtype, seed):
    import torch
    d = swift.generate(2, scale=scale, seed=seed)
    npdt = np.float32 if dtype == torch.float32 else np.float64
    v = d["val"].astype(npdt)
    B = np.random.default_rng(seed).uniform(-1, 1, (d["n"], K)).astype(npdt)
    C64, absC = oracle.spmm_f64(d["m"], d["n"], d["rowptr"], d["colidx"], v.astype(np.float64),
                                B.astype(np.float64))
    M = swift.SwiftMatrix(d["m"], d["n"], d["rowptr"], d["colidx"], v, dtype=dtype)
    return d, M, B, C64, absC


def _call(swift, fn, M, K, layout, code, B, out):
    return getattr(swift.lib(), fn)(M.h, K, layout, code, swift.KERNEL_DEFAULT, 0,
                                    C.c_void_p(B.data_ptr()), C.c_void_p(out.data_ptr()),
                                    swift._stream())


def _product(swift, fn, M, B, layout, code, m, K, dtype):
    """C (row-major host array) from entry point fn, C prefilled with NaN."""
    import torch
    Bt = torch.tensor(B if layout == swift.ROW_MAJOR else np.ascontiguousarray(B.T),
                      device="cuda")
    out = torch.full((m, K) if layout == swift.ROW_MAJOR else (K, m), float("nan"),
                     device="cuda", dtype=dtype)
    rc = getattr(swift.lib(), fn)(M.h, K, layout, code, swift.KERNEL_DEFAULT, 0,
                                  C.c_void_p(Bt.data_ptr()), C.c_void_p(out.data_ptr()),
                                  swift._stream())
    assert rc == 0, swift.lib().swift_last_error().decode()
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    return got if layout == swift.ROW_MAJOR else got.T


@pytest.mark.parametrize("layout", ["COL", "ROW"])
def test_noalloc_matches_swift_spmm(swift, oracle, cuda, layout):
    """Two calls in a row, the second smaller (and FP64): each against
    swift_spmm and the oracle."""
    import torch
    lay = swift.COL_MAJOR if layout == "COL" else swift.ROW_MAJOR
    for scale, K, dtype, code, tol in ((0.2, 64, torch.float32, swift.F32, 1e-5),
                                       (0.05, 32, torch.float64, swift.F64, 1e-12)):
        print("case", layout, scale, K, dtype, flush=True)
        d, M, B, C64, absC = _matrix(swift, oracle, scale, K, dtype, seed=int(scale * 100))
        got = _product(swift, "swift_spmm_noalloc", M, B, lay, code, d["m"], K, dtype)
        ref = _product(swift, "swift_spmm", M, B, lay, code, d["m"], K, dtype)
        assert not np.isnan(got).any()
        assert gpu_common.rel_err(got, C64, absC) <= tol
        assert gpu_common.rel_err(ref, C64, absC) <= tol
        # cut rows that also take reside entries add with atomics, in any order
        assert gpu_common.rel_err(got, ref.astype(np.float64), absC) <= 2 * tol
        M.close()


def test_noalloc_rejects_beyond_scratch(swift, oracle, cuda):
    import torch
    K = 256
    d = swift.generate(2, scale=0.2, seed=1)
    v = d["val"].astype(np.float32)
    need = -(-d["n"] * K * 4 // 256) * 256 + d["m"] * K * 4 + 256
    assert need > LIMIT
    M = swift.SwiftMatrix(d["m"], d["n"], d["rowptr"], d["colidx"], v)
    B = torch.zeros((K, d["n"]), device="cuda")
    out = torch.zeros((K, d["m"]), device="cuda")
    rc = _call(swift, "swift_spmm_noalloc", M, K, swift.COL_MAJOR, swift.F32, B, out)
    assert rc == -10
    assert "scratch" in swift.lib().swift_last_error().decode()
    # the same product fits swift_spmm, whose workspace grows
    assert _call(swift, "swift_spmm", M, K, swift.COL_MAJOR, swift.F32, B, out) == 0
    torch.cuda.synchronize()
    M.close()


def test_noalloc_leaves_free_memory(swift, oracle, cuda):
    """On fresh handles: swift_spmm's first column-major call takes its
    workspace (free memory falls by at least its size), swift_spmm_noalloc's
    takes none."""
    import torch
    K = 64
    d = swift.generate(2, scale=0.2, seed=3)
    v = d["val"].astype(np.float32)
    work = d["n"] * K * 4 + d["m"] * K * 4
    Bt = torch.rand((K, d["n"]), device="cuda")
    out = torch.empty((K, d["m"]), device="cuda")
    M0 = swift.SwiftMatrix(d["m"], d["n"], d["rowptr"], d["colidx"], v)
    # load every kernel of the path once
    assert _call(swift, "swift_spmm_noalloc", M0, K, swift.COL_MAJOR, swift.F32, Bt, out) == 0
    torch.cuda.synchronize()
    M1 = swift.SwiftMatrix(d["m"], d["n"], d["rowptr"], d["colidx"], v)
    M2 = swift.SwiftMatrix(d["m"], d["n"], d["rowptr"], d["colidx"], v)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    assert _call(swift, "swift_spmm_noalloc", M1, K, swift.COL_MAJOR, swift.F32, Bt, out) == 0
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    assert _call(swift, "swift_spmm", M2, K, swift.COL_MAJOR, swift.F32, Bt, out) == 0
    torch.cuda.synchronize()
    free2 = torch.cuda.mem_get_info()[0]
    print("free memory change: noalloc %d, swift_spmm %d, workspace %d"
          % (free1 - free0, free2 - free1, work), flush=True)
    assert free0 - free1 < (8 << 20)
    assert free1 - free2 >= work
    for M in (M0, M1, M2):
        M.close()
