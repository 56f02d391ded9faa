"""Every benchmark configuration at its stated size, against the FP64 oracle.

- 4096^2 CPU-path matrix (~80k nonzeros), K = 32: FP32 and FP64, row- and
  column-major, element by element.
- 10^6 x 10^6 with 16M nonzeros: FP64 at K = 32 and K = 128 and FP32 at
  K = 128, row- and column-major, element by element (FP32 <= 1e-5, FP64 <=
  1e-12 of (|A||B|)[i,k]).
- The same matrix with BF16 storage of A's values and B, and F16 storage with
  per-column scales, at K = 128 and K = 256 (FP32 accumulation and C): within
  1e-2 of (|A||B|)[i,k].  Rounding both factors of a product to BF16 (8
  significant bits, unit roundoff u = 2^-8) moves it by at most (2u + u^2) =
  7.8e-3 of |a||b|, so the bound holds with room for the FP32 sums.
- 512 batched matrices (square, rows log-uniform in [1k, 64k]): one launch and
  two graph replays; every product checked through a row checksum
  C x = A (B x) and a column checksum 1^T C = (1^T A) B against FP64 sums, and
  every 8th product element by element.  An element error of at most
  1e-5 (|A||B|)[i,k] bounds the checksum errors by 1e-5 |A|(|B||x|) and
  1e-5 (1^T|A|)|B|, so those are the bounds used.
"""
import concurrent.futures as cf

import numpy as np
import pytest

import gpu_common

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

_C2 = {}


def config2(swift):
    if not _C2:
        _C2["d"] = swift.generate(2, scale=1.0, seed=0)
    return _C2["d"]


def _product(swift, M, B, layout, **kw):
    """C = A*B with B given row-major (n x K) in host memory; C returned row-major."""
    import torch
    Bt = torch.tensor(B if layout == swift.ROW_MAJOR else np.ascontiguousarray(B.T), device="cuda")
    C = M.spmm(Bt, layout=layout, **kw)
    torch.cuda.synchronize()
    out = C.cpu().numpy()
    del Bt, C
    return out if layout == swift.ROW_MAJOR else out.T


@pytest.mark.parametrize("dt", ["f32", "f64"])
def test_config1_full(swift, oracle, cuda, dt):
    import torch
    d = swift.generate(1, scale=1.0, seed=0)
    m, n = d["m"], d["n"]
    npdt = np.float32 if dt == "f32" else np.float64
    v = d["val"].astype(npdt)
    B = np.random.default_rng(0).uniform(-1, 1, (n, 32)).astype(npdt)
    C64, absC = oracle.spmm_f64(m, n, d["rowptr"], d["colidx"], v.astype(np.float64),
                                B.astype(np.float64))
    M = swift.SwiftMatrix(m, n, d["rowptr"], d["colidx"], v,
                          dtype=torch.float32 if dt == "f32" else torch.float64)
    for layout in (swift.ROW_MAJOR, swift.COL_MAJOR):
        err = gpu_common.rel_err(_product(swift, M, B, layout), C64, absC)
        print("config1", dt, "layout", layout, "err %.3g" % err, flush=True)
        assert err <= (1e-5 if dt == "f32" else 1e-12)
    M.close()


@pytest.mark.parametrize("dt,K", [("f64", 32), ("f64", 128), ("f32", 128)])
def test_config2_full_elementwise(swift, oracle, cuda, dt, K):
    """The 10^6-row matrix, both layouts, every element against the oracle."""
    import torch
    d = config2(swift)
    m, n = d["m"], d["n"]
    npdt = np.float32 if dt == "f32" else np.float64
    v = d["val"].astype(npdt)
    print("config2", dt, "K", K, "nnz", d["nnz"], flush=True)
    B = swift.generate_dense(n * K, 0, -1.0, 1.0, seed=11 + K).reshape(n, K).astype(npdt)
    C64, absC = oracle.spmm_f64(m, n, d["rowptr"], d["colidx"], v.astype(np.float64),
                                B.astype(np.float64))
    M = swift.SwiftMatrix(m, n, d["rowptr"], d["colidx"], v,
                          dtype=torch.float32 if dt == "f32" else torch.float64)
    for layout in (swift.ROW_MAJOR, swift.COL_MAJOR):
        C = _product(swift, M, B, layout)
        err = gpu_common.rel_err(C, C64, absC)
        print("config2", dt, "K", K, "layout", layout, "err %.3g" % err, flush=True)
        assert err <= (1e-5 if dt == "f32" else 1e-12)
        del C
    M.close()
    del C64, absC, B
    torch.cuda.empty_cache()


@pytest.mark.parametrize("K", [128, 256])
def test_config3_full(swift, oracle, cuda, K):
    """BF16 storage and F16 storage with per-column scales on the full
    matrix; the largest per-row max relative error against the row's max
    |C64| is printed beside the normalised error."""
    import torch
    d = config2(swift)
    m, n = d["m"], d["n"]
    v = d["val"].astype(np.float32)
    B = swift.generate_dense(n * K, 0, -1.0, 1.0, seed=3 + K).reshape(n, K).astype(np.float32)
    C64, absC = oracle.spmm_f64(m, n, d["rowptr"], d["colidx"], v.astype(np.float64),
                                B.astype(np.float64))
    M = swift.SwiftMatrix(m, n, d["rowptr"], d["colidx"], v, dtype=torch.float32)
    Bt = torch.tensor(B, device="cuda")
    del B
    M.prepare_values(swift.BF16)
    Bh = swift.convert_dense(Bt, swift.BF16)
    C = M.spmm(Bh, dtype=swift.BF16).cpu().numpy()
    del Bh
    err = gpu_common.rel_err(C, C64, absC)
    row = float(gpu_common.row_max_rel(C, C64).max())
    print("config3 bf16 K", K, "normalised %.3g row max rel %.3g" % (err, row), flush=True)
    assert err <= 1e-2
    del C
    M.prepare_values(swift.F16)
    Bf, sc = swift.convert_dense(Bt, swift.F16, col_scale=True)
    C = M.spmm_scaled(Bf, sc).cpu().numpy()
    del Bf, sc
    err = gpu_common.rel_err(C, C64, absC)
    row = float(gpu_common.row_max_rel(C, C64).max())
    print("config3 f16 scaled K", K, "normalised %.3g row max rel %.3g" % (err, row), flush=True)
    assert err <= 1e-2
    M.close()
    del C, C64, absC, Bt
    torch.cuda.empty_cache()


def _gen5(swift, i):
    return swift.generate(5, scale=1.0, seed=0, index=i)


def test_config5_full_batch(swift, oracle, cuda):
    """512 generated matrices in one persistent launch and in graph replays."""
    import torch
    count = 512
    with cf.ThreadPoolExecutor(8) as ex:  # the generator releases the GIL
        mats = list(ex.map(lambda i: _gen5(swift, i), range(count)))
    nnz = sum(d["nnz"] for d in mats)
    print("config5", count, "matrices", nnz, "nnz", flush=True)
    g = torch.Generator(device="cuda").manual_seed(7)
    Ms, Ks, Bs, Cs = [], [], [], []
    for d in mats:
        d["val32"] = d["val"].astype(np.float32)
        Ms.append(swift.SwiftMatrix(d["m"], d["n"], d["rowptr"], d["colidx"], d["val32"],
                                    dtype=torch.float32))
        Ks.append(d["k"])
        Bs.append(torch.rand((d["n"], d["k"]), device="cuda", generator=g) * 2 - 1)
        Cs.append(torch.empty((d["m"], d["k"]), device="cuda"))
    assert sorted(set(Ks)) == [32, 64, 128]
    batch = swift.Batch(Ms, Ks)
    # FP64 expectations, computed once: row checksums A (B x) with their bound
    # |A| (|B||x|), column checksums (1^T A) B with (1^T|A|)|B|, and every 8th
    # product in full
    xs = [torch.rand(K, device="cuda", dtype=torch.float64, generator=g) * 2 - 1 for K in Ks]
    expect = []
    for i, (d, b, x) in enumerate(zip(mats, Bs, xs)):
        m, n, rp, ci = d["m"], d["n"], d["rowptr"], d["colidx"]
        v64 = d["val32"].astype(np.float64)
        bx = (b.double() @ x).cpu().numpy().reshape(n, 1)
        abx = (b.double().abs() @ x.abs()).cpu().numpy().reshape(n, 1)
        row, _ = oracle.spmm_f64(m, n, rp, ci, v64, bx, want_abs=False)
        rbound, _ = oracle.spmm_f64(m, n, rp, ci, np.abs(v64), abx, want_abs=False)
        colA = torch.tensor(np.bincount(ci, weights=v64, minlength=n), device="cuda")
        absA = torch.tensor(np.bincount(ci, weights=np.abs(v64), minlength=n), device="cuda")
        e = {"row": torch.tensor(row[:, 0], device="cuda"),
             "rbound": torch.tensor(rbound[:, 0], device="cuda"),
             "col": colA @ b.double(), "cbound": absA @ b.double().abs()}
        if i % 8 == 0:
            e["C64"], e["absC"] = oracle.spmm_f64(m, n, rp, ci, v64,
                                                  b.cpu().numpy().astype(np.float64))
        expect.append(e)

    def check(tag):
        torch.cuda.synchronize()
        for i, (d, c, x, e) in enumerate(zip(mats, Cs, xs, expect)):
            cd = c.double()
            rerr = float(((cd @ x - e["row"]).abs() / (e["rbound"] + 1e-30)).max()) \
                if d["m"] else 0.0
            cerr = float(((cd.sum(0) - e["col"]).abs() / (e["cbound"] + 1e-30)).max())
            assert rerr <= 1e-5 and cerr <= 1e-5, "%s matrix %d (%dx%d, K=%d): row %.3g col %.3g" \
                % (tag, i, d["m"], d["n"], d["k"], rerr, cerr)
            if "C64" in e:
                err = gpu_common.rel_err(c.cpu().numpy(), e["C64"], e["absC"])
                assert err <= 1e-5, "%s matrix %d elementwise %.3g" % (tag, i, err)
        print("config5", tag, "checked", flush=True)

    for c in Cs:
        c.fill_(float("nan"))
    batch.spmm(Bs, Cs)
    check("launch")
    for rep in range(2):
        for c in Cs:
            c.fill_(float("nan"))
        batch.spmm(Bs, Cs, graph=True)
        check("graph replay %d" % rep)
    batch.close()
    for M in Ms:
        M.close()
    del Bs, Cs
    torch.cuda.empty_cache()
