"""swift_matrix_stats on the GPU against the oracle: block emptiness and
sparsity (test/MTX_info/main.cu:143-176, :484), the regular / reside part
sizes and tile formats of the oracle's Swift build, row/column extremes."""
import numpy as np
import pytest

import gpu_common

pytestmark = pytest.mark.gpu

CASES = gpu_common.cases()


def check(swift, oracle, m, n, rp, ci, v):
    import torch
    M = swift.SwiftMatrix(m, n, rp, ci, np.asarray(v, np.float32), dtype=torch.float32)
    st = M.stats()
    M.close()
    blocks, empty, sparsity = oracle.block_stats(m, n, rp, ci)
    assert st.blocks == blocks
    assert st.nonempty_blocks == blocks - empty
    assert st.empty_block_ratio == pytest.approx((empty / blocks) if blocks else 0.0)
    assert st.sparsity == pytest.approx(sparsity)
    ref = oracle.build(m, n, rp, ci, np.asarray(v, np.float64))
    assert (st.regular_nnz, st.reside_nnz, st.tilenum) == (ref["regular_nnz"], ref["reside_nnz"],
                                                            ref["tilenum"])
    fmt = np.bincount(ref["tile_format"], minlength=3) if ref["tilenum"] else np.zeros(3, int)
    assert list(st.tile_format_count) == [int(x) for x in fmt[:3]]
    lens = np.diff(np.asarray(rp, np.int64))
    assert st.max_row_nnz == (int(lens.max()) if m else 0)
    colcnt = np.bincount(np.asarray(ci, np.int64), minlength=n) if len(ci) else np.zeros(1, int)
    assert st.max_col_nnz == int(colcnt.max())
    dense, dense_nnz, max_block = oracle.block_density(m, n, rp, ci, 512)
    assert (st.dense_blocks, st.dense_block_nnz, st.max_block_nnz) == (dense, dense_nnz,
                                                                        max_block)
    return st


@pytest.mark.parametrize("name", list(CASES))
def test_stats_cases(swift, oracle, cuda, name):
    st = check(swift, oracle, *CASES[name])
    print(name, "empty_block_ratio", st.empty_block_ratio, "regular_ratio", st.regular_ratio,
          flush=True)


def test_stats_large(swift, oracle, cuda):
    """1/10 of the benchmark matrix: 100k x 100k, ~1.6M nonzeros."""
    d = swift.generate(2, scale=0.1, seed=0)
    st = check(swift, oracle, d["m"], d["n"], d["rowptr"], d["colidx"], d["val"])
    assert 0.0 < st.empty_block_ratio < 1.0


def test_stats_dense_blocks(swift, oracle, cuda):
    """Blocks at least half full: a banded matrix (bandwidth 12, every band
    block dense) and a random one around a 60%-filled 64 x 64 corner."""
    import matgen
    st = check(swift, oracle, *matgen.banded_mixed_csr(2000, 2000, 5, bw=12, lo=25, hi=25,
                                                      mixed=False))
    assert st.dense_blocks > 0 and st.max_block_nnz >= 512
    rng = np.random.default_rng(9)
    A = rng.random((300, 400)) < 0.01
    A[:64, :64] |= rng.random((64, 64)) < 0.6
    rp = np.r_[0, np.cumsum(A.sum(1))].astype(np.int32)
    ci = np.concatenate([np.nonzero(r)[0] for r in A]).astype(np.int32)
    st = check(swift, oracle, 300, 400, rp, ci, np.ones(len(ci)))
    assert st.dense_blocks >= 4


def test_stats_wide_bitmap(swift, oracle, cuda):
    """More column blocks than one bitmap pass holds (262144): 64 x 9,000,000."""
    import torch
    m, n = 64, 9_000_000
    rng = np.random.default_rng(3)
    rows = [np.sort(rng.choice(n, size=20, replace=False)) for _ in range(m)]
    rp = np.r_[0, np.cumsum([len(r) for r in rows])].astype(np.int32)
    ci = np.concatenate(rows).astype(np.int32)
    M = swift.SwiftMatrix(m, n, rp, ci, np.ones(len(ci), np.float32), dtype=torch.float32)
    st = M.stats()
    M.close()
    blocks = 2 * (n // 32 + (n % 32 > 0))
    occupied = len(set((i // 32, int(c) // 32) for i in range(m) for c in rows[i]))
    assert st.blocks == blocks and st.nonempty_blocks == occupied
    assert (st.dense_blocks, st.dense_block_nnz) == (0, 0) and st.max_block_nnz >= 1
