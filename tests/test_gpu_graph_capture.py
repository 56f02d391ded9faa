"""Single products captured into a HIP graph (torch.cuda.CUDAGraph) and replayed.

swift_spmm keeps the carry slots of rows cut by segment boundaries in a
workspace per stream, and the column-major default's transposes in a
workspace per handle; the first call allocates them, so after one call on the
capture stream a captured product allocates nothing.  swift_spmm_noalloc needs
no such call.  On the benchmark matrix's structure at 1/20 scale (the default
segments cut rows, and the row view holds the reside entries, so no atomics
and every product is repeatable bit for bit): FP32 K=128 row-major and
column-major, BF16 and swift_spmm_noalloc are each captured once and replayed
three times, with new B values written and an eager product run between
replays.  Every replay equals the eager product bit for bit; the first replay
is also checked against the FP64 oracle."""
import ctypes as C

import numpy as np
import pytest

import gpu_common

pytestmark = pytest.mark.gpu

K = 128
_MAT = {}


def matrix(swift):
    if not _MAT:
        _MAT["d"] = swift.generate(2, scale=0.05, seed=0)
    return _MAT["d"]


@pytest.mark.parametrize("form", ["row", "col", "bf16", "noalloc"])
def test_captured_product_replays(swift, oracle, cuda, form):
    import torch
    d = matrix(swift)
    m, n = d["m"], d["n"]
    M = swift.SwiftMatrix(m, n, d["rowptr"], d["colidx"], d["val"].astype(np.float32),
                          dtype=torch.float32)
    inf = M.info()
    assert inf.rowview_nnz == d["nnz"]  # no reside blocks: no atomics in the product
    col = form == "col"
    layout = swift.COL_MAJOR if col else swift.ROW_MAJOR
    rng = np.random.default_rng(11)
    if form == "bf16":
        M.prepare_values(swift.BF16)
    shapeB, shapeC = ((K, n), (K, m)) if col else ((n, K), (m, K))
    Bst = torch.empty(shapeB, device="cuda",
                      dtype=torch.bfloat16 if form == "bf16" else torch.float32)
    Cst = torch.full(shapeC, float("nan"), device="cuda")

    def write_b(Bh):
        x = torch.tensor(Bh.T.copy() if col else Bh, device="cuda")
        Bst.copy_(swift.convert_dense(x, swift.BF16) if form == "bf16" else x)

    def product(out, stream):
        if form == "noalloc":
            rc = swift.lib().swift_spmm_noalloc(
                M.h, K, layout, swift.F32, swift.KERNEL_DEFAULT, 0, C.c_void_p(Bst.data_ptr()),
                C.c_void_p(out.data_ptr()), swift._stream(stream))
            assert rc == 0, swift.lib().swift_last_error().decode()
        else:
            M.spmm(Bst, out=out, layout=layout,
                   dtype=swift.BF16 if form == "bf16" else None, stream=stream)

    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    write_b(rng.uniform(-1, 1, (n, K)).astype(np.float32))
    torch.cuda.synchronize()
    if form != "noalloc":
        with torch.cuda.stream(s):
            product(Cst, s)  # allocates this stream's workspace
        s.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        product(Cst, s)
    print("captured", form, flush=True)
    vals = d["val"].astype(np.float32).astype(np.float64)
    for rep in range(3):
        Bh = rng.uniform(-1, 1, (n, K)).astype(np.float32)
        write_b(Bh)
        Cst.fill_(float("nan"))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        eager = torch.full(shapeC, float("nan"), device="cuda")
        product(eager, torch.cuda.current_stream())
        torch.cuda.synchronize()
        got = Cst.cpu().numpy()
        assert np.array_equal(got, eager.cpu().numpy()), (form, rep)
        if rep == 0:
            Bref = Bh.astype(np.float64)
            if form == "bf16":
                Bref = swift.convert_dense(torch.tensor(Bh, device="cuda"), swift.BF16) \
                    .float().cpu().numpy().astype(np.float64)
                vals_r = (torch.tensor(vals, dtype=torch.float32).bfloat16().double().numpy())
            else:
                vals_r = vals
            C64, absC = oracle.spmm_f64(m, n, d["rowptr"], d["colidx"], vals_r, Bref)
            err = gpu_common.rel_err(got.T if col else got, C64, absC)
            print(form, "replay err", err, flush=True)
            assert err <= 1e-5, (form, err)
    M.close()
