"""Rows cut by merge-path segment boundaries (carry slots and the fix-up).

A row-view walk stores the piece of a cut row that holds the row's start into
C and every later piece into a carry slot; one tail launch adds each cut
row's slots onto it in slot order (one wave per row, a workgroup for rows of
more than 32 slots) while its other workgroups run the reside blocks with
atomics.  Small segment sizes (W = 8 cuts almost every row) put every
matrix through that path, and a matrix with rows of 600-900 entries gives
rows of more than 32 slots; C starts as NaN so a row left unwritten shows
up.  Checked against
the FP64 oracle for FP32 (plain and accumulating), FP64 and BF16, for the
default, row-view and window walks, at K with one lane shape and with masked
slices (K = 100); the split count of the default table against a NumPy
restatement of the merge-path rule; and that a product whose rows take no
atomics is bitwise repeatable."""
import numpy as np
import pytest

import gpu_common

pytestmark = pytest.mark.gpu

CASES = gpu_common.cases()


def long_rows(m=600, n=1000, seed=17):
    """Ten rows of 600-900 entries among short random rows: at W = 8 the
    long rows are cut into 75-113 pieces."""
    rng = np.random.default_rng(seed)
    rows = []
    for i in range(m):
        k = int(rng.integers(600, 901)) if i % 60 == 7 else int(rng.integers(0, 12))
        rows.append(np.sort(rng.choice(n, size=k, replace=False)))
    rp = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
    ci = np.concatenate(rows).astype(np.int32)
    return m, n, rp, ci, rng.standard_normal(len(ci))


CASES["longrows_600x1000"] = long_rows()


def _product(swift, M, B, K, kernel, flags=0, dtype=None, fill=float("nan"), C0=None):
    import torch
    m = M.info().m
    out = (torch.tensor(C0, device="cuda") if C0 is not None else
           torch.full((m, K), fill, device="cuda", dtype=torch.float64 if B.dtype == torch.float64
                      else torch.float32))
    M.spmm(B, out=out, kernel=kernel, flags=flags, dtype=dtype)
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("kernel", ["DEFAULT", "ROWVIEW", "WINDOW"])
@pytest.mark.parametrize("K", [32, 100, 128])
@pytest.mark.parametrize("W", [8, 24])
@pytest.mark.parametrize("name", list(CASES))
def test_cut_rows_f32(swift, oracle, cuda, name, W, K, kernel):
    import torch
    print("case", name, W, K, kernel, flush=True)
    m, n, rp, ci, v = CASES[name]
    vv = np.asarray(v, np.float32)
    B = np.random.default_rng(W + K).uniform(-1, 1, (n, K)).astype(np.float32)
    C64, absC = oracle.spmm_f64(m, n, rp, ci, vv.astype(np.float64), B.astype(np.float64))
    M = swift.SwiftMatrix(m, n, rp, ci, vv, dtype=torch.float32)
    M.set_segment_items(W)
    C = _product(swift, M, torch.tensor(B, device="cuda"), K, getattr(swift, "KERNEL_" + kernel))
    assert not np.isnan(C).any()
    assert gpu_common.rel_err(C, C64, absC) <= 1e-5
    M.close()


@pytest.mark.parametrize("K", [32, 128])
@pytest.mark.parametrize("name", ["powerlaw_3000x2500", "densecol_400x100", "tall_2000x64",
                                  "banded_mixed_4000", "longrows_600x1000"])
def test_cut_rows_accumulate(swift, oracle, cuda, name, K):
    """C += A*B: the piece holding a row's start adds onto C, the fix-up adds
    the rest."""
    import torch
    m, n, rp, ci, v = CASES[name]
    vv = np.asarray(v, np.float32)
    rng = np.random.default_rng(K)
    B = rng.uniform(-1, 1, (n, K)).astype(np.float32)
    C0 = rng.uniform(-1, 1, (m, K)).astype(np.float32)
    C64, absC = oracle.spmm_f64(m, n, rp, ci, vv.astype(np.float64), B.astype(np.float64))
    M = swift.SwiftMatrix(m, n, rp, ci, vv, dtype=torch.float32)
    M.set_segment_items(8)
    C = _product(swift, M, torch.tensor(B, device="cuda"), K, swift.KERNEL_DEFAULT,
                 flags=swift.ACCUMULATE, C0=C0)
    assert gpu_common.rel_err(C - C0, C64, absC + np.abs(C0)) <= 1e-5
    M.close()


@pytest.mark.parametrize("name", ["powerlaw_3000x2500", "densecol_400x100", "random_300x200",
                                  "longrows_600x1000"])
def test_cut_rows_f64_bf16(swift, oracle, cuda, name):
    import torch
    m, n, rp, ci, v = CASES[name]
    rng = np.random.default_rng(3)
    # FP64, K = 64
    v64 = np.asarray(v, np.float64)
    B = rng.uniform(-1, 1, (n, 64))
    C64, absC = oracle.spmm_f64(m, n, rp, ci, v64, B)
    M = swift.SwiftMatrix(m, n, rp, ci, v64, dtype=torch.float64)
    M.set_segment_items(8)
    C = _product(swift, M, torch.tensor(B, device="cuda"), 64, swift.KERNEL_DEFAULT)
    assert not np.isnan(C).any()
    assert gpu_common.rel_err(C, C64, absC) <= 1e-12
    M.close()
    # BF16 storage, K = 128
    v32 = np.asarray(v, np.float32)
    B = rng.uniform(-1, 1, (n, 128)).astype(np.float32)
    C64, absC = oracle.spmm_f64(m, n, rp, ci, v32.astype(np.float64), B.astype(np.float64))
    M = swift.SwiftMatrix(m, n, rp, ci, v32, dtype=torch.float32)
    M.set_segment_items(8)
    M.prepare_values(swift.BF16)
    Bh = swift.convert_dense(torch.tensor(B, device="cuda"), swift.BF16)
    C = _product(swift, M, Bh, 128, swift.KERNEL_DEFAULT, dtype=swift.BF16)
    assert not np.isnan(C).any()
    assert gpu_common.rel_err(C, C64, absC) <= 1e-2
    M.close()


def merge_path_splits(rowptr, W):
    """Boundaries that cut a row, for segments of W items over (rows +
    entries): segment s starts at diagonal s*W, moved back to the start of
    the row it falls in when that row began less than W items earlier."""
    ptr = np.asarray(rowptr, np.int64)
    m, R = len(ptr) - 1, int(ptr[-1])
    nseg = -(-(m + R) // W)
    s = np.arange(nseg + 1, dtype=np.int64)
    d = np.minimum(s * W, m + R)
    lo = np.searchsorted(ptr[1:] + np.arange(m), d, side="left")
    e = d - lo
    lo_c = np.minimum(lo, max(m - 1, 0))
    move = (s > 0) & (s < nseg) & (lo < m) & (e > ptr[lo_c]) & (ptr[lo_c] + lo > d - W)
    e = np.where(move, ptr[lo_c], e)
    end_x, end_y = lo[1:], e[1:]
    cut = (end_x < m) & (end_y > ptr[np.minimum(end_x, max(m - 1, 0))])
    return int(cut.sum())


def test_merge_path_split_count(swift, oracle, cuda):
    """The carry slots the default table needs (nsplit * K values) follow
    from the merge-path rule: swift_spmm_noalloc's limit check sees the same
    count as a NumPy restatement of the rule."""
    import ctypes as C
    import torch
    d = swift.generate(2, scale=0.2, seed=1)
    M = swift.SwiftMatrix(d["m"], d["n"], d["rowptr"], d["colidx"], d["val"].astype(np.float32))
    inf = M.info()
    assert inf.rowview_nnz == inf.nnz  # the reside part is walked in the row view
    K = 256
    for W, fits in ((8, False), (512, True)):
        M.set_segment_items(W)
        nsplit = merge_path_splits(d["rowptr"], W)
        need = nsplit * K * 4
        print("W", W, "split boundaries", nsplit, "carry bytes", need, flush=True)
        assert (need <= 256 << 20) == fits
        B = torch.zeros((d["n"], K), device="cuda")
        out = torch.zeros((d["m"], K), device="cuda")
        rc = swift.lib().swift_spmm_noalloc(M.h, K, swift.ROW_MAJOR, swift.F32,
                                            swift.KERNEL_DEFAULT, 0, C.c_void_p(B.data_ptr()),
                                            C.c_void_p(out.data_ptr()), swift._stream())
        torch.cuda.synchronize()
        if fits:
            assert rc == 0
        else:
            assert rc == -10
            msg = swift.lib().swift_last_error().decode()
            assert "carry" in msg and ("(%d of them" % need) in msg, msg
    M.close()


@pytest.mark.parametrize("name", ["powerlaw_3000x2500", "uniform_64x64", "banded_mixed_4000"])
def test_cut_rows_repeatable(swift, oracle, cuda, name):
    """Rows whose row view holds every entry take no atomics: the fix-up sums
    a row's pieces in slot order, so two products agree bit for bit."""
    import torch
    m, n, rp, ci, v = CASES[name]
    M = swift.SwiftMatrix(m, n, rp, ci, np.asarray(v, np.float32))
    inf = M.info()
    assert inf.rowview_nnz == inf.nnz
    M.set_segment_items(8)
    Bt = torch.rand((n, 128), device="cuda") * 2 - 1
    a = _product(swift, M, Bt, 128, swift.KERNEL_DEFAULT)
    b = _product(swift, M, Bt, 128, swift.KERNEL_DEFAULT, fill=0.0)
    assert np.array_equal(a, b)
    M.close()


def test_cut_rows_two_streams(swift, oracle, cuda):
    """Products of one handle on two streams at once: each stream has carry
    slots of its own, so neither overwrites the other's pieces."""
    import torch
    m, n, rp, ci, v = CASES["powerlaw_3000x2500"]
    vv = np.asarray(v, np.float32)
    rng = np.random.default_rng(9)
    Bs = [rng.uniform(-1, 1, (n, 128)).astype(np.float32) for _ in range(2)]
    refs = [oracle.spmm_f64(m, n, rp, ci, vv.astype(np.float64), b.astype(np.float64)) for b in Bs]
    M = swift.SwiftMatrix(m, n, rp, ci, vv, dtype=torch.float32)
    M.set_segment_items(8)
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    Bt = [torch.tensor(b, device="cuda") for b in Bs]
    outs = [torch.full((m, 128), float("nan"), device="cuda") for _ in range(2)]
    torch.cuda.synchronize()
    for rep in range(3):
        for s, b, o in zip(streams, Bt, outs):
            with torch.cuda.stream(s):
                M.spmm(b, out=o, stream=s)
        torch.cuda.synchronize()
        for o, (C64, absC) in zip(outs, refs):
            got = o.cpu().numpy()
            assert not np.isnan(got).any()
            assert gpu_common.rel_err(got, C64, absC) <= 1e-5, rep
    M.close()


@pytest.mark.parametrize("K", [33, 1100])
@pytest.mark.parametrize("name", ["longrows_600x1000", "densecol_400x100"])
def test_cut_rows_fixup_shapes(swift, oracle, cuda, name, K):
    """The fix-up's other shapes: K = 33 (no 16-byte vectors: one element
    per thread, a masked K slice) and K = 1100 (more than 256 vectors per
    row: the workgroup loops over the row with one group of threads), on rows
    of more than 32 slots (workgroup path) and fewer (wave path)."""
    import torch
    m, n, rp, ci, v = CASES[name]
    vv = np.asarray(v, np.float32)
    B = np.random.default_rng(K).uniform(-1, 1, (n, K)).astype(np.float32)
    C64, absC = oracle.spmm_f64(m, n, rp, ci, vv.astype(np.float64), B.astype(np.float64))
    M = swift.SwiftMatrix(m, n, rp, ci, vv, dtype=torch.float32)
    M.set_segment_items(8)
    C = _product(swift, M, torch.tensor(B, device="cuda"), K, swift.KERNEL_DEFAULT)
    assert not np.isnan(C).any()
    assert gpu_common.rel_err(C, C64, absC) <= 1e-5
    M.close()
