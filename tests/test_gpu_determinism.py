"""Determinism: the same input built twice gives bit-identical Swift format
arrays (every array swift_get_info exposes), and the same product, repeated
and from the second build, agrees to 1e-6 of (|A||B|)[i,k] -- split rows and
reside blocks add with atomics, whose order may vary; everything else is
written once per element."""
import numpy as np
import pytest

import gpu_common

pytestmark = pytest.mark.gpu

CASES = gpu_common.cases()


def _matrices(swift):
    d = swift.generate(1, scale=1.0, seed=0)  # the 4096^2 CPU-path case
    out = {"config1": (d["m"], d["n"], d["rowptr"], d["colidx"], d["val"])}
    for name in ("powerlaw_3000x2500", "banded_mixed_4000", "densecol_400x100", "empty_10x40"):
        out[name] = CASES[name]
    return out


@pytest.mark.parametrize("name", ["config1", "powerlaw_3000x2500", "banded_mixed_4000",
                                  "densecol_400x100", "empty_10x40"])
def test_build_and_product_repeat(swift, cuda, name):
    import torch
    m, n, rp, ci, v = _matrices(swift)[name]
    vv = np.asarray(v, np.float32)
    A1 = swift.SwiftMatrix(m, n, rp, ci, vv)
    A2 = swift.SwiftMatrix(m, n, rp, ci, vv)
    a1, a2 = A1.arrays(), A2.arrays()
    assert sorted(a1) == sorted(a2)
    for k in a1:
        x, y = a1[k], a2[k]
        if isinstance(x, np.ndarray):
            assert x.tobytes() == y.tobytes(), k
        else:
            assert x == y, k
    K = 64
    B = torch.tensor(np.random.default_rng(1).uniform(-1, 1, (n, K)).astype(np.float32),
                     device="cuda")
    Babs = B.abs()
    Aabs = swift.SwiftMatrix(m, n, rp, ci, np.abs(vv))
    scale = Aabs.spmm(Babs).double() + 1e-30
    for kern in (swift.KERNEL_DEFAULT, swift.KERNEL_BALANCED, swift.KERNEL_FLAT):
        c1 = A1.spmm(B, kernel=kern).double()
        c2 = A1.spmm(B, kernel=kern).double()
        c3 = A2.spmm(B, kernel=kern).double()
        for c in (c2, c3):
            assert float(((c - c1).abs() / scale).max()) <= 1e-6
    Aabs.close()
    A1.close()
    A2.close()


def test_no_device_memory_growth(swift, cuda):
    """Build, transpose, statistics, 16-bit preparation and conversion and a
    product, four times over: the device memory the library takes for its
    temporaries comes back (free memory is measured by the HIP runtime, so
    allocations outside torch count)."""
    import torch
    d = swift.generate(2, scale=0.1, seed=0)  # 100k x 100k, ~1.6M nonzeros
    vals = d["val"].astype(np.float32)

    def cycle():
        M = swift.SwiftMatrix(d["m"], d["n"], d["rowptr"], d["colidx"], vals)
        M.transpose()
        M.stats()
        M.prepare_values(swift.F16)
        B = torch.rand((d["n"], 64), device="cuda")
        Bh, sc = swift.convert_dense(B, swift.F16, col_scale=True)
        M.spmm_scaled(Bh, sc)
        M.spmm(B)
        torch.cuda.synchronize()
        M.close()

    cycle()
    torch.cuda.empty_cache()
    free0 = torch.cuda.mem_get_info()[0]
    for _ in range(4):
        cycle()
    torch.cuda.empty_cache()
    free1 = torch.cuda.mem_get_info()[0]
    print("free before %d after %d (MiB)" % (free0 >> 20, free1 >> 20), flush=True)
    assert free0 - free1 < (64 << 20)
