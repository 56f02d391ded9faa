"""B operands past the buffer-resource limits of the walk.

The row-view walks read B rows through a buffer resource with 24-bit row
offsets while B has fewer than 2^24 rows and fewer than 2^31 bytes; past either
limit they take 64-bit global loads.  These tests run both limits at real
size: 4.4M B rows x K=128 FP32 (2.25 GB), and 17M rows (> 2^24) x K=32.  A has
a few thousand rows with entries spread over all of B (the last B row
included).  The FP64 reference sums the gathered B rows on the GPU."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _matrix(m, n, per_row, seed):
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, 2 * per_row + 1, m)
    lens[0] = max(lens[0], 1)
    rp = np.zeros(m + 1, np.int32)
    np.cumsum(lens, out=rp[1:])
    ci = rng.integers(0, n, int(rp[-1])).astype(np.int32)
    ci[0] = n - 1  # the last B row, at the largest offsets
    for i in range(m):  # sorted columns within a row, as a CSR usually has
        ci[rp[i]:rp[i + 1]].sort()
    v = rng.uniform(-1, 1, int(rp[-1])).astype(np.float32)
    return rp, ci, v


def _reference(torch, m, rp, ci, v, B):
    """C[i] = sum_e v_e * B[c_e] in FP64 and sum_e |v_e| |B[c_e]|."""
    cols = torch.tensor(ci.astype(np.int64), device="cuda")
    rows = torch.repeat_interleave(torch.arange(m, device="cuda"),
                                   torch.tensor(np.diff(rp).astype(np.int64), device="cuda"))
    vals = torch.tensor(v.astype(np.float64), device="cuda")[:, None]
    G = B[cols].double()
    C = torch.zeros((m, B.shape[1]), dtype=torch.float64, device="cuda")
    Cabs = torch.zeros_like(C)
    C.index_add_(0, rows, vals * G)
    Cabs.index_add_(0, rows, vals.abs() * G.abs())
    return C, Cabs


@pytest.mark.parametrize("n,K", [(4_400_000, 128), (17_000_000, 32)])
def test_large_b_forward(swift, cuda, n, K):
    import torch
    m = 3000
    rp, ci, v = _matrix(m, n, 8, seed=n % 1000)
    assert n * K * 4 >= 2 ** 31 or n >= 2 ** 24
    M = swift.SwiftMatrix(m, n, rp, ci, v, dtype=torch.float32)
    g = torch.Generator(device="cuda").manual_seed(1)
    B = torch.rand((n, K), device="cuda", generator=g) * 2 - 1
    R, Rabs = _reference(torch, m, rp, ci, v, B)
    C = M.spmm(B)
    err = float(((C.double() - R).abs() / (Rabs + 1e-30)).max())
    print("large f32", n, K, "err", err, flush=True)
    assert err <= 1e-5
    if K == 128:
        # column-major: B^T (K x n) is transposed into scratch, C back
        Ct = M.spmm(B.t().contiguous(), layout=swift.COL_MAJOR)
        err = float(((Ct.t().double() - R).abs() / (Rabs + 1e-30)).max())
        print("large f32 col-major", n, K, "err", err, flush=True)
        assert err <= 1e-5
    else:
        # BF16 storage: 1.09 GB of B, but more than 2^24 rows
        M.prepare_values(swift.BF16)
        Bh = swift.convert_dense(B, swift.BF16)
        C = M.spmm(Bh, dtype=swift.BF16)
        err = float(((C.double() - R).abs() / (Rabs + 1e-30)).max())
        print("large bf16", n, K, "err", err, flush=True)
        assert err <= 1e-2
        del Bh
    M.close()
    del B, C, R, Rabs
    torch.cuda.empty_cache()


def test_large_b_sddmm(swift, cuda):
    """dvals = SDDMM(dC, B) gathers B rows by A's columns: 4.4M x 128 FP32."""
    import torch
    m, n, K = 3000, 4_400_000, 128
    rp, ci, v = _matrix(m, n, 8, seed=5)
    M = swift.SwiftMatrix(m, n, rp, ci, v, dtype=torch.float32)
    g = torch.Generator(device="cuda").manual_seed(2)
    B = torch.rand((n, K), device="cuda", generator=g) * 2 - 1
    dC = torch.rand((m, K), device="cuda", generator=g) * 2 - 1
    dv = M.sddmm(dC, B)
    rows = torch.repeat_interleave(torch.arange(m, device="cuda"),
                                   torch.tensor(np.diff(rp).astype(np.int64), device="cuda"))
    cols = torch.tensor(ci.astype(np.int64), device="cuda")
    G, H = B[cols].double(), dC[rows].double()
    ref = (G * H).sum(1)
    den = (G.abs() * H.abs()).sum(1)
    err = float(((dv.double() - ref).abs() / (den + 1e-30)).max())
    print("large sddmm err", err, flush=True)
    assert err <= 1e-5
    M.close()
    del B, dC
    torch.cuda.empty_cache()


@pytest.mark.parametrize("graph", [False, True])
def test_large_b_batched(swift, cuda, graph):
    """A batch holding one 4.4M-row B (K=128, 2.25 GB) walks every matrix with
    64-bit loads; a small matrix beside it must come out the same."""
    import torch
    shapes = [(3000, 4_400_000, 128), (500, 700, 32)]
    mats, Bs, refs = [], [], []
    g = torch.Generator(device="cuda").manual_seed(3)
    for i, (m, n, K) in enumerate(shapes):
        rp, ci, v = _matrix(m, n, 8, seed=11 + i)
        mats.append(swift.SwiftMatrix(m, n, rp, ci, v, dtype=torch.float32))
        B = torch.rand((n, K), device="cuda", generator=g) * 2 - 1
        Bs.append(B)
        refs.append(_reference(torch, m, rp, ci, v, B))
    Cs = [torch.full((m, K), float("nan"), device="cuda") for m, _, K in shapes]
    batch = swift.Batch(mats, [K for _, _, K in shapes])
    for _ in range(2 if graph else 1):  # a graph is captured once, then replayed
        for c in Cs:
            c.fill_(float("nan"))
        batch.spmm(Bs, Cs, graph=graph)
        torch.cuda.synchronize()
        for (R, Rabs), C in zip(refs, Cs):
            err = float(((C.double() - R).abs() / (Rabs + 1e-30)).max())
            print("large batched graph=%s err %g" % (graph, err), flush=True)
            assert err <= 1e-5
    batch.close()
    for M in mats:
        M.close()
    del Bs, Cs
    torch.cuda.empty_cache()
