"""Batched persistent launch (swift_batch_*) and the autotuner: every product
of the batch against the FP64 oracle (FP32 bound 1e-5 normalised by |A||B|)
and against separate swift_spmm launches; graph replays checked one by one
with eager work and new operands between them."""
import numpy as np
import pytest

import gpu_common
import matgen

pytestmark = pytest.mark.gpu

TOL = 1e-5


def make_batch(swift, count, seed=0, Ks=(32, 64, 128, 256)):
    """Uneven FP32 matrices: generator items (square, power-law rows) at small
    scale plus the edge-case matrices of the single-product tests."""
    import torch
    out = []
    edge = gpu_common.cases()
    names = ["empty_10x40", "uniform_64x64", "narrow_50x20", "wide_40x3000",
             "densecol_400x100", "unsorted_rows_64x96", "tall_2000x64", "powerlaw_3000x2500"]
    for i in range(count):
        if i % 3 == 2:
            m, n, rp, ci, v = edge[names[(i // 3) % len(names)]]
        else:
            d = swift.generate(5, scale=0.05, seed=seed, index=i)
            m, n, rp, ci, v = d["m"], d["n"], d["rowptr"], d["colidx"], d["val"]
        K = Ks[i % len(Ks)]
        v32 = np.asarray(v, np.float32)
        M = swift.SwiftMatrix(m, n, rp, ci, v32, dtype=torch.float32)
        out.append((M, K, (m, n, rp, ci, v32)))
    return out


def operands(items, seed):
    import torch
    rng = np.random.default_rng(seed)
    Bs, Bn, Cs = [], [], []
    for M, K, (m, n, *_rest) in items:
        b = rng.uniform(-1, 1, (n, K)).astype(np.float32)
        Bn.append(b)
        Bs.append(torch.tensor(b, device="cuda"))
        Cs.append(torch.full((m, K), 7.0, device="cuda"))  # stale values must go
    return Bs, Bn, Cs


def check_all(oracle, items, Bn, Cs):
    import torch
    torch.cuda.synchronize()
    for i, ((M, K, (m, n, rp, ci, v)), b, c) in enumerate(zip(items, Bn, Cs)):
        C64, absC = oracle.spmm_f64(m, n, rp, ci, v.astype(np.float64), b.astype(np.float64))
        err = gpu_common.rel_err(c.cpu().numpy(), C64, absC)
        assert err <= TOL, "matrix %d (%dx%d, K=%d): %.3g" % (i, m, n, K, err)


def test_batch_matches_oracle_and_separate(swift, oracle, cuda):
    import torch
    items = make_batch(swift, 24)
    print("batch of", len(items), flush=True)
    batch = swift.Batch([it[0] for it in items], [it[1] for it in items])
    p = batch.params()
    # segments of 64..512 items; a matrix walked in two halves of K has each
    # segment twice
    def units(it, W):
        inf = it[0].info()
        return -(-(inf.m + inf.rowview_nnz) // W) * (2 if inf.tuned_cols else 1)

    lo = sum(units(it, 512) for it in items)
    hi = sum(units(it, 64) for it in items)
    assert lo <= p["work_units"] <= hi
    assert p["depth"] in (64, 128, 256, 512) and p["block_threads"] == 256
    Bs, Bn, Cs = operands(items, 1)
    batch.spmm(Bs, Cs)
    check_all(oracle, items, Bn, Cs)
    # Batched and separate launches sum a row in different FP32 orders (split rows
    # end in atomics, so the order also varies between runs): on rows of ~2,600
    # entries two orders differ by ~1.6e-4, beyond an elementwise 1e-5.  Each is
    # within TOL of FP64 normalised by |A||B|, so they agree within 2*TOL.
    for (M, K, (m, n, rp, ci, v)), b, bn, c in zip(items, Bs, Bn, Cs):
        ref = M.spmm(b)
        torch.cuda.synchronize()
        _, absC = oracle.spmm_f64(m, n, rp, ci, v.astype(np.float64), bn.astype(np.float64))
        err = gpu_common.rel_err(c.cpu().numpy(), ref.cpu().numpy().astype(np.float64), absC)
        assert err <= 2 * TOL, "batched vs separate (%dx%d, K=%d): %.3g" % (m, n, K, err)
    batch.close()
    for it in items:
        it[0].close()


def test_graph_replays(swift, oracle, cuda):
    import torch
    items = make_batch(swift, 12, seed=3)
    batch = swift.Batch([it[0] for it in items], [it[1] for it in items])
    Bs, Bn, Cs = operands(items, 2)
    for rep in range(4):
        print("replay", rep, flush=True)
        for c in Cs:
            c.fill_(-3.0)
        batch.spmm(Bs, Cs, graph=True)
        check_all(oracle, items, Bn, Cs)
        # eager work between replays
        x = torch.randn(1 << 20, device="cuda")
        float(x.sum())
    # new operand tensors: the descriptor table is rewritten, the graph kept
    Bs2, Bn2, Cs2 = operands(items, 5)
    batch.spmm(Bs2, Cs2, graph=True)
    check_all(oracle, items, Bn2, Cs2)
    batch.spmm(Bs, Cs, graph=False)
    check_all(oracle, items, Bn, Cs)
    batch.close()
    for it in items:
        it[0].close()


def test_duplicate_matrix_in_batch(swift, oracle, cuda):
    import torch
    m, n, rp, ci, v = matgen.powerlaw_csr(700, 600, 21, max_len=300)
    v32 = np.asarray(v, np.float32)
    M = swift.SwiftMatrix(m, n, rp, ci, v32, dtype=torch.float32)
    items = [(M, 32, (m, n, rp, ci, v32)), (M, 128, (m, n, rp, ci, v32))]
    batch = swift.Batch([M, M], [32, 128])
    Bs, Bn, Cs = operands(items, 9)
    batch.spmm(Bs, Cs)
    check_all(oracle, items, Bn, Cs)
    batch.close()
    M.close()


def test_batch_rejects(swift, cuda):
    import torch
    m, n, rp, ci, v = matgen.random_csr(100, 80, 0.1, 4)
    M = swift.SwiftMatrix(m, n, rp, ci, np.asarray(v, np.float32), dtype=torch.float32)
    M64 = swift.SwiftMatrix(m, n, rp, ci, np.asarray(v, np.float64), dtype=torch.float64)
    with pytest.raises(swift.SwiftError):
        swift.Batch([M], [48])
    with pytest.raises(swift.SwiftError):
        swift.Batch([M64], [32])
    batch = swift.Batch([M], [32])
    B = torch.zeros((n * 32 + 1,), device="cuda")
    C = torch.zeros((m * 32 + 1,), device="cuda")
    with pytest.raises(swift.SwiftError):  # misaligned operands
        batch.spmm([B[1:]], [C[1:]])
    # re-segmenting a batched matrix invalidates the batch, loudly
    M.set_segment_items(64 if batch.params()["depth"] != 64 else 128)
    with pytest.raises(swift.SwiftError):
        batch.spmm([B[:n * 32]], [C[:m * 32]])
    batch.close()
    M.close()
    M64.close()


def test_autotune_single(swift, oracle, cuda):
    import torch
    m, n, rp, ci, v = matgen.powerlaw_csr(3000, 2500, 6, max_len=400)
    v32 = np.asarray(v, np.float32)
    M = swift.SwiftMatrix(m, n, rp, ci, v32, dtype=torch.float32)
    rng = np.random.default_rng(0)
    b = rng.uniform(-1, 1, (n, 64)).astype(np.float32)
    B = torch.tensor(b, device="cuda")
    C = torch.empty((m, 64), device="cuda")
    W = M.autotune(B, C)
    assert W in (0, 64, 128, 256, 512)  # 0: the default table of the B-row bytes
    C = M.spmm(B)
    torch.cuda.synchronize()
    C64, absC = oracle.spmm_f64(m, n, rp, ci, v32.astype(np.float64), b.astype(np.float64))
    assert gpu_common.rel_err(C.cpu().numpy(), C64, absC) <= TOL
    M.close()


# lane shape of the first K slice -> slot-row counts the autotuner may pick
# (0 = the rule: 16 for one-chunk 64-lane groups of 4/8-byte elements, 8 for
# other 32/64-lane groups and 16-lane groups of 2-byte elements, 4 below; 16
# only where a step still fits one staged batch of 64 entries)
SLOT_CHOICES = {8: {0, 8}, 16: {0, 4, 8, 16}, 32: {0, 4, 16}, 64: {0, 4, 8, 16}}


# (dtype, K, lane group of the first uncapped K slice, slice caps the autotuner
# may pick: 0 = the default, the first uncapped slice's full width and half)
@pytest.mark.parametrize("dt,K,lpe,widths", [("f32", 256, 64, (0, 256, 128)),
                                             ("bf16", 64, 16, (0, 64, 32)),
                                             ("f32", 100, 16, (0, 64, 32)),
                                             ("f64", 32, 16, (0, 32, 16)),
                                             ("f32", 128, 32, (0, 128, 64))])
def test_autotune_schedule_applied(swift, oracle, cuda, dt, K, lpe, widths):
    """swift_autotune over segment size x slot rows x lane width x walk on the benchmark
    structure at 1/10 scale: the choice is reported by swift_info, taken by
    the DEFAULT product of that K (before and after another K's product) and
    the product matches the oracle."""
    import torch
    print("case autotune", dt, K, flush=True)
    d = swift.generate(2, scale=0.1, seed=0)
    m, n, rp, ci = d["m"], d["n"], d["rowptr"], d["colidx"]
    f64 = dt == "f64"
    v = d["val"].astype(np.float64 if f64 else np.float32)
    M = swift.SwiftMatrix(m, n, rp, ci, v, dtype=torch.float64 if f64 else torch.float32)
    b = np.random.default_rng(K).uniform(-1, 1, (n, K)).astype(v.dtype)
    B = torch.tensor(b, device="cuda")
    code = swift.F64 if f64 else swift.F32
    if dt == "bf16":
        M.prepare_values(swift.BF16)
        B = swift.convert_dense(B, swift.BF16)
        code = swift.BF16
    C = torch.empty((m, K), device="cuda", dtype=torch.float64 if f64 else torch.float32)
    W = M.autotune(B, C)
    inf = M.info()
    print("tuned", W, inf.tuned_K, inf.tuned_slots, inf.tuned_kernel, inf.tuned_cols, flush=True)
    # W = 0: the default table of the B-row bytes
    assert inf.segment_items == W and W in (0, 64, 128, 256, 512) and inf.tuned_K == K
    assert inf.tuned_kernel in (swift.KERNEL_DEFAULT, swift.KERNEL_WINDOW)
    assert inf.tuned_slots in SLOT_CHOICES[lpe]
    # a width cap narrows the first slice's lane group (K in more launches)
    assert inf.tuned_cols in widths
    if inf.tuned_kernel == swift.KERNEL_WINDOW or inf.tuned_cols:
        assert inf.tuned_slots == 0
    C64, absC = oracle.spmm_f64(m, n, rp, ci, v.astype(np.float64), b.astype(np.float64))
    tol = 1e-12 if f64 else 1e-2 if dt == "bf16" else TOL
    out = torch.full_like(C, float("nan"))
    M.spmm(B, out=out, dtype=code)
    assert gpu_common.rel_err(out.cpu().numpy(), C64, absC) <= tol
    # another K: the rule applies there, and the tuned K keeps its choice
    M.spmm(B[:, :K // 2].contiguous(), dtype=code)
    out.fill_(float("nan"))
    M.spmm(B, out=out, dtype=code)
    assert gpu_common.rel_err(out.cpu().numpy(), C64, absC) <= tol
    M.close()


@pytest.mark.parametrize("mode", ["HALF", "FULL"])
def test_batch_lane_width(swift, oracle, cuda, mode):
    """Every matrix of K >= 64 walked in two halves of K/2 columns (two
    descriptors, each half's segments in turn), or none: launch and graph
    replays against the FP64 oracle."""
    code = getattr(swift, "BATCH_WIDTH_" + mode)
    swift.batch_set_lane_width(code)
    try:
        items = make_batch(swift, 16, seed=9)
        batch = swift.Batch([it[0] for it in items], [it[1] for it in items])
    finally:
        swift.batch_set_lane_width(swift.BATCH_WIDTH_TUNED)
    for M, K, _ in items:
        inf = M.info()
        want = K // 2 if mode == "HALF" and K >= 64 else 0
        assert inf.tuned_cols == want, (K, inf.tuned_cols)
    Bs, Bn, Cs = operands(items, 6)
    batch.spmm(Bs, Cs)
    check_all(oracle, items, Bn, Cs)
    for rep in range(2):
        for c in Cs:
            c.fill_(float("nan"))
        batch.spmm(Bs, Cs, graph=True)
        check_all(oracle, items, Bn, Cs)
    batch.close()
    for it in items:
        it[0].close()


def test_batch_lane_width_rejects(swift, cuda):
    with pytest.raises(swift.SwiftError):
        swift.batch_set_lane_width(7)


def test_batch_schedules(swift, oracle, cuda):
    """Every batched matrix runs the schedule its key was tuned to: a segment
    size and, per K, the rule or the batched walk's other slot-row count."""
    items = make_batch(swift, 16, seed=5)
    batch = swift.Batch([it[0] for it in items], [it[1] for it in items])
    for M, K, _ in items:
        inf = M.info()
        assert inf.segment_items in (64, 128, 256, 512)
        assert inf.tuned_K == K and inf.tuned_kernel == swift.KERNEL_DEFAULT
        assert inf.tuned_slots in ({0, 8} if K <= 64 else {0, 4})
        assert inf.tuned_cols in ({0, K // 2} if K >= 64 else {0})
    Bs, Bn, Cs = operands(items, 4)
    batch.spmm(Bs, Cs, graph=True)
    check_all(oracle, items, Bn, Cs)
    batch.close()
    for it in items:
        it[0].close()
