"""The CPU oracle, pinned against the reference's own CPU code.

Golden fixtures (tests/golden/*.npz) hold the outputs of the reference's CPU
headers compiled in place (oracle/_ref, written by tests/golden/make_golden.py).
Where those leave something open the checks use what the mathematics fixes:
the FP64 product against a dense matmul, the tile/reside decomposition
reproducing the CSR pattern entry for entry, and brute-force sorts.
"""
import glob
import os

import numpy as np
import pytest

import matgen

GOLD = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))
FORMAT_KEYS = ("tilenum", "regular_nnz", "reside_col", "reside_nnz", "segsum", "numblk",
               "sortrowidx", "tile_ptr", "tile_len", "tile_colidx", "tile_tall",
               "tile_format", "segmentoffset", "sortrowindex", "reside_cscptr",
               "reside_cscrowidx", "colorblk", "blkstart", "blkstop", "blkptr")


def load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(params=GOLD, ids=[os.path.basename(p)[:-4] for p in GOLD])
def gold(request):
    return load(request.param)


def test_golden_present():
    assert len(GOLD) >= 6


def test_csr2csc_matches_reference(oracle, gold):
    ptr, row, val, npc = oracle.csr2csc(int(gold["m"]), int(gold["n"]), gold["rowptr"],
                                        gold["colidx"], gold["val"])
    assert np.array_equal(ptr, gold["csc_ptr"])
    assert np.array_equal(row, gold["csc_row"])
    assert np.array_equal(val, gold["csc_val"])
    assert np.array_equal(npc, gold["nnzpercol"])


def test_col_sort_matches_reference(oracle, gold):
    n = int(gold["n"])
    perm, sn, sr, sv = oracle.col_sort(n, gold["nnzpercol"], gold["csc_ptr"], gold["csc_row"],
                                       gold["csc_val"])
    # the sorted counts do not depend on the tie rule
    assert np.array_equal(sn, gold["ref_sortnnz"])
    # stable tie rule: brute force
    want = np.array(sorted(range(n), key=lambda j: (gold["nnzpercol"][j], j)), np.int32)
    assert np.array_equal(perm, want)
    # the reference's own order differs only inside groups of equal counts
    ref_perm = gold["ref_colsort_perm"]
    assert sorted(ref_perm.tolist()) == list(range(n))
    assert np.array_equal(gold["nnzpercol"][ref_perm], sn)
    assert np.array_equal(sr, gold["sortrowidx_tmp"])


def test_format_matches_reference(oracle, gold):
    f = oracle.build(int(gold["m"]), int(gold["n"]), gold["rowptr"], gold["colidx"], gold["val"])
    for k in FORMAT_KEYS:
        assert np.array_equal(np.asarray(f[k]), np.asarray(gold["fmt_" + k])), k
    assert np.array_equal(f["sortval"], gold["fmt_sortval"])
    assert np.array_equal(f["reside_val"], gold["fmt_reside_val"])


def test_row_sort_matches_reference(oracle, gold):
    f = oracle.build(int(gold["m"]), int(gold["n"]), gold["rowptr"], gold["colidx"], gold["val"])
    f.row_idx_sort()
    assert np.array_equal(f["rowsorted"], gold["rowsorted"])
    assert np.array_equal(f["val_offset"], gold["val_offset"])
    # val_position: a permutation per group that sorts the keys, ties by lane
    R = f["regular_nnz"]
    vp = f["val_position"]
    for g0 in range(0, R, 256):
        g1 = min(R, g0 + 256)
        pos = vp[g0:g1]
        assert sorted(pos.tolist()) == list(range(g1 - g0))
        keys = f["sortrowidx"][g0:g1]
        assert np.array_equal(f["rowsorted"][g0:g1][pos], keys)
        order = np.argsort(pos)
        assert np.array_equal(order, np.lexsort((np.arange(g1 - g0), keys)))


def test_swift_cpu_matches_reference_bitwise(oracle, gold):
    f = oracle.build(int(gold["m"]), int(gold["n"]), gold["rowptr"], gold["colidx"], gold["val"])
    K = gold["B"].shape[1]
    sB = gold["B"][f["perm"]]
    c64 = f.swift_cpu(K, sB, np.float64)
    assert np.array_equal(c64, gold["C64"])
    c32 = f.swift_cpu(K, sB.astype(np.float32), np.float32)
    assert np.array_equal(c32, gold["C32"])


def test_swift_cpu_equals_dense_product(oracle, gold):
    m, n = int(gold["m"]), int(gold["n"])
    A = matgen.csr_to_dense(m, n, gold["rowptr"], gold["colidx"], gold["val"])
    want = A @ gold["B"]
    np.testing.assert_allclose(gold["C64"], want, rtol=1e-12, atol=1e-12)
    C, absC = oracle.spmm_f64(m, n, gold["rowptr"], gold["colidx"], gold["val"], gold["B"])
    np.testing.assert_allclose(C, want, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(absC, np.abs(A) @ np.abs(gold["B"]), rtol=1e-12, atol=1e-12)


def decompose(f):
    """(row, sorted col, value) of every entry the format stores."""
    n = f["n"]
    ents = []
    for t in range(f["tilenum"]):
        b = f["tile_ptr"][t]
        for l in range(32):
            ents.append((f["sortrowidx"][b + l], f["tile_colidx"][t] + l, f["sortval"][b + l]))
    for r in range(f["reside_col"]):
        for p in range(f["reside_cscptr"][r], f["reside_cscptr"][r + 1]):
            ents.append((f["reside_cscrowidx"][p], n - 1 - r, f["reside_val"][p]))
    return ents


@pytest.mark.parametrize("seed", range(4))
def test_format_invariants(oracle, seed):
    m, n, rp, ci, v = matgen.powerlaw_csr(257, 199, seed)
    f = oracle.build(m, n, rp, ci, v)
    nnz = len(ci)
    assert f["regular_nnz"] + f["reside_nnz"] == nnz
    assert f["regular_nnz"] == 32 * f["tilenum"]
    assert 0 <= f["reside_col"] <= 31
    # the decomposition reproduces the CSR pattern and values exactly
    got = sorted((int(r), int(f["perm"][c]), float(x)) for r, c, x in decompose(f))
    want = sorted((i, int(ci[e]), float(v[e])) for i in range(m) for e in range(rp[i], rp[i + 1]))
    assert got == want
    # a tile = 32 consecutive sorted columns at one depth
    sn, cptr = f["sortnnz"], np.concatenate([[0], np.cumsum(f["sortnnz"])])
    for t in range(f["tilenum"]):
        c0 = f["tile_colidx"][t]
        b = f["tile_ptr"][t]
        depths = set()
        for l in range(32):
            pos = np.where((f["sortrowidx_tmp"][cptr[c0 + l]:cptr[c0 + l + 1]] ==
                            f["sortrowidx"][b + l]))[0]
            assert len(pos) >= 1
            depths.add(tuple(pos.tolist()))
        assert len(depths) >= 1
    # formats and runs
    for t in range(f["tilenum"]):
        rows = f["sortrowidx"][32 * t:32 * t + 32]
        ch = int(np.sum(rows[1:] != rows[:-1]))
        assert f["tile_format"][t] == (0 if ch == 0 else (2 if ch < 16 else 1))
    assert f["segsum"] == f["countrow"][-1]
    assert int(np.sum(f["segmentoffset"])) == f["regular_nnz"] - f["segsum"]
    # balance blocks cover each reside column, none longer than 32
    for b in range(f["numblk"]):
        u = int(f["colorblk"][b])
        r = u & 0x7FFFFFFF
        if u >> 31:
            assert 0 < f["blkstop"][b] - f["blkstart"][b] <= 32
        else:
            assert f["reside_cscptr"][r + 1] - f["reside_cscptr"][r] <= 32


def test_no_reside_when_uniform_columns(oracle):
    # every column has the same count and N % 32 == 0: no irregular part
    m, n = 64, 64
    rows = [np.array([(i + d) % n for d in range(3)]) for i in range(m)]
    rp = np.concatenate([[0], np.cumsum([3] * m)]).astype(np.int32)
    ci = np.concatenate([np.sort(r) for r in rows]).astype(np.int32)
    f = oracle.build(m, n, rp, ci, np.ones(len(ci)))
    assert f["reside_nnz"] == 0 and f["reside_col"] == 0
    assert f["tilenum"] == 3 * 64 // 32


def test_narrow_matrix_all_reside(oracle):
    # N < 32: no tile fits, everything is irregular
    m, n, rp, ci, v = matgen.random_csr(50, 20, 0.3, 3)
    f = oracle.build(m, n, rp, ci, v)
    assert f["tilenum"] == 0
    assert f["reside_nnz"] == len(ci)
    assert f["reside_col"] == 20 or f["reside_col"] == int(np.count_nonzero(np.bincount(ci, minlength=n) > 0))


def test_empty_matrix(oracle):
    f = oracle.build(10, 40, np.zeros(11, np.int32), np.zeros(0, np.int32), np.zeros(0))
    assert f["tilenum"] == 0 and f["reside_nnz"] == 0 and f["numblk"] == 0


@pytest.mark.parametrize("seed", range(3))
def test_oracle_against_live_reference(oracle, seed):
    if not oracle.Ref.available(np.float64):
        pytest.skip("oracle/_ref not built (reference sources absent)")
    ref = oracle.Ref(np.float64)
    m, n, rp, ci, v = matgen.dense_column_csr(300 + seed, 90 + 7 * seed, 100 + seed)
    f = oracle.build(m, n, rp, ci, v)
    h, r = ref.format(m, n, f["sortrowidx_tmp"], f["sortval_tmp"], f["sortnnz"])
    for k in FORMAT_KEYS:
        assert np.array_equal(np.asarray(r[k]), np.asarray(f[k])), k
    ref.free(h)


def test_planted_mistake_is_caught(oracle):
    """Flip one tile entry of the oracle's output: the decomposition check and
    the Swift_cpu comparison both see it."""
    g = load(GOLD[0])
    f = oracle.build(int(g["m"]), int(g["n"]), g["rowptr"], g["colidx"], g["val"])
    bad = dict(f.a)
    bad["sortrowidx"] = bad["sortrowidx"].copy()
    bad["sortrowidx"][5] = (bad["sortrowidx"][5] + 1) % int(g["m"])
    assert not np.array_equal(bad["sortrowidx"], g["fmt_sortrowidx"])
    got = sorted((int(r), int(f["perm"][c])) for r, c, _ in decompose(bad))
    want = sorted((i, int(g["colidx"][e])) for i in range(int(g["m"]))
                  for e in range(g["rowptr"][i], g["rowptr"][i + 1]))
    assert got != want


@pytest.mark.parametrize("shape", [(1, 1), (31, 33), (64, 64), (100, 257), (300, 200), (10, 40)])
def test_block_stats_brute_force(oracle, shape):
    """Empty 32x32 blocks (test/MTX_info/main.cu:143-176) against a dense
    block reduction of the pattern."""
    m, n = shape
    mm, nn, rp, ci, _ = matgen.random_csr(m, n, 0.01, 7 + m)
    blocks, empty, sparsity = oracle.block_stats(m, n, rp, ci)
    A = matgen.csr_to_dense(m, n, rp, ci, np.ones(len(ci))) != 0
    rb, cb = -(-m // 32), -(-n // 32)
    P = np.zeros((rb * 32, cb * 32), bool)
    P[:m, :n] = A
    occupied = P.reshape(rb, 32, cb, 32).any(axis=(1, 3))
    assert blocks == rb * cb
    assert empty == int((~occupied).sum())
    assert sparsity == pytest.approx(len(ci) / (m * n))


def test_block_stats_known_blocks(oracle):
    """Entries placed in exactly three blocks of a 64 x 96 matrix (2 x 3 blocks)."""
    rp = np.cumsum(np.r_[0, [2 if i == 5 else 1 if i == 39 else 0 for i in range(64)]]).astype(np.int32)
    ci = np.array([0, 95, 40], np.int32)   # row 5: blocks (0,0), (0,2); row 39: block (1,1)
    blocks, empty, _ = oracle.block_stats(64, 96, rp, ci)
    assert (blocks, empty) == (6, 3)


@pytest.mark.parametrize("shape", [(1, 1), (64, 64), (100, 257), (97, 130)])
def test_block_density_brute_force(oracle, shape):
    """Dense 32x32 blocks against a dense block reduction: a random pattern plus
    one block filled to 60% and one to exactly the threshold."""
    m, n = shape
    rng = np.random.default_rng(m * 1000 + n)
    A = rng.random((m, n)) < 0.02
    if m >= 64 and n >= 64:
        A[0:32, 32:64] = rng.random((32, 32)) < 0.6
        blk = np.zeros(1024, bool)
        blk[:512] = True
        A[32:64, 0:32] = rng.permutation(blk).reshape(32, 32)
    rp = np.r_[0, np.cumsum(A.sum(1))].astype(np.int32)
    ci = np.concatenate([np.nonzero(r)[0] for r in A]).astype(np.int32) if A.any() else \
        np.zeros(0, np.int32)
    rb, cb = -(-m // 32), -(-n // 32)
    P = np.zeros((rb * 32, cb * 32), int)
    P[:m, :n] = A
    cnt = P.reshape(rb, 32, cb, 32).sum(axis=(1, 3))
    for thr in (1, 300, 512):
        d, dz, mx = oracle.block_density(m, n, rp, ci, thr)
        assert d == int((cnt >= thr).sum())
        assert dz == int(cnt[cnt >= thr].sum())
        assert mx == int(cnt.max())
    if m >= 64 and n >= 64:
        d, dz, _ = oracle.block_density(m, n, rp, ci, 512)
        assert d >= 2 and dz >= 2 * 512


def _grad_inputs(gold, K=24, seed=9):
    m, n = int(gold["m"]), int(gold["n"])
    rng = np.random.default_rng(seed)
    dC = rng.standard_normal((m, K))
    B = rng.standard_normal((n, K))
    return m, n, dC, B


def test_spmm_t_equals_dense_transpose_product(oracle, gold):
    """dB = A^T dC against a dense numpy A.T @ dC on the golden matrices."""
    m, n, dC, _ = _grad_inputs(gold)
    A = matgen.csr_to_dense(m, n, gold["rowptr"], gold["colidx"], gold["val"])
    got = oracle.spmm_t_f64(m, n, gold["rowptr"], gold["colidx"], gold["val"], dC)
    np.testing.assert_allclose(got, A.T @ dC, rtol=1e-12, atol=1e-12)


def test_sddmm_equals_rowwise_dots(oracle, gold):
    """dvals[e] = <dC[row_e], B[col_e]> in CSR order, against (dC @ B.T) read
    at the pattern."""
    m, n, dC, B = _grad_inputs(gold)
    rp, ci = gold["rowptr"], gold["colidx"]
    got = oracle.sddmm_f64(m, n, rp, ci, dC, B)
    rows = np.repeat(np.arange(m), np.diff(rp))
    want = (dC @ B.T)[rows, ci]
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


def test_backward_adjoint_identity(oracle):
    """<dC, A B> = <A^T dC, B> = sum_e v_e dvals_e, the identity the backward
    rests on, on a rectangular matrix with unsorted rows and duplicate-free
    random columns."""
    m, n, rp, ci, v = matgen.random_csr(70, 45, 0.12, 31, shuffle_rows=True)
    rng = np.random.default_rng(5)
    K = 7
    B, dC = rng.standard_normal((n, K)), rng.standard_normal((m, K))
    C, _ = oracle.spmm_f64(m, n, rp, ci, v, B)
    dB = oracle.spmm_t_f64(m, n, rp, ci, v, dC)
    dv = oracle.sddmm_f64(m, n, rp, ci, dC, B)
    a, b, c = float((dC * C).sum()), float((dB * B).sum()), float(v @ dv)
    assert abs(a - b) <= 1e-12 * max(1.0, abs(a)) and abs(a - c) <= 1e-12 * max(1.0, abs(a))


def test_planted_backward_mistakes_are_caught():
    """The dense checks above fail for the plausible mistakes: a product with
    A instead of A^T (on a square matrix, where the shapes still agree), row
    and column swapped in the SDDMM gather, and one dropped entry."""
    m, n, rp, ci, v = matgen.random_csr(40, 40, 0.15, 17)
    rng = np.random.default_rng(3)
    K = 5
    dC, B = rng.standard_normal((m, K)), rng.standard_normal((n, K))
    A = matgen.csr_to_dense(m, n, rp, ci, v)
    rows = np.repeat(np.arange(m), np.diff(rp))
    want_dB, want_dv = A.T @ dC, (dC @ B.T)[rows, ci]
    not_transposed = A @ dC
    assert not np.allclose(not_transposed, want_dB, rtol=1e-12, atol=1e-12)
    swapped = np.einsum("ek,ek->e", dC[ci], B[rows])
    assert not np.allclose(swapped, want_dv, rtol=1e-12, atol=1e-12)
    dropped = A.copy()
    dropped[rows[7], ci[7]] -= v[7]
    assert not np.allclose(dropped.T @ dC, want_dB, rtol=1e-12, atol=1e-12)
