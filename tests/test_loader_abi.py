"""Host-side product code on the CPU: the C-ABI library loads and exports
every declared symbol, the .mtx loader matches the reference loader's outputs
(golden fixtures from mmio_allinone, src/auxiliary.h:167), and the seeded
generators are deterministic and have the documented shape."""
import glob
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MTX = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "mtx", "*.mtx")))


def header_functions():
    src = open(os.path.join(ROOT, "include", "swift_spmm.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"\b(swift_[a-z0-9_]+)\s*\(", src)
    return sorted(set(names))


def test_header_symbols_exported(swift):
    names = header_functions()
    assert len(names) >= 25
    L = swift.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(swift.EXPORTS) == names


def test_version(swift):
    assert "gfx950" in swift.version()


@pytest.mark.parametrize("path", MTX, ids=[os.path.basename(p) for p in MTX])
def test_loader_matches_reference(swift, oracle, path):
    with np.load(path[:-4] + ".npz", allow_pickle=False) as z:
        gold = {k: z[k] for k in z.files}
    rc, d = swift.mmio_load(path)
    assert rc == int(gold["rc"])
    rc2, d2 = oracle.mmio_load(path)
    assert rc2 == int(gold["rc"])
    if rc != 0:
        return
    for dd in (d, d2):
        assert dd["m"] == int(gold["m"]) and dd["n"] == int(gold["n"])
        assert dd["nnz"] == int(gold["nnz"]) and dd["sym"] == int(gold["sym"])
        for k in ("rowptr", "colidx", "val"):
            assert np.array_equal(dd[k], gold[k]), k


def test_loader_errors(swift, tmp_path):
    rc, _ = swift.mmio_load(str(tmp_path / "missing.mtx"))
    assert rc == -1
    p = tmp_path / "trunc.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real general\n3 3 4\n1 1 1.0\n2 2 2.0\n")
    rc, _ = swift.mmio_load(str(p))
    assert rc == -3
    p = tmp_path / "range.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real general\n3 3 1\n4 1 1.0\n")
    rc, _ = swift.mmio_load(str(p))
    assert rc == -3
    p = tmp_path / "array.mtx"
    p.write_text("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n")
    rc, _ = swift.mmio_load(str(p))
    assert rc == -2


def test_generator_deterministic(swift):
    a = swift.generate(1, seed=0)
    b = swift.generate(1, seed=0)
    c = swift.generate(1, seed=1)
    for k in ("rowptr", "colidx", "val"):
        assert np.array_equal(a[k], b[k])
    assert not np.array_equal(a["colidx"], c["colidx"])


def test_config1_shape(swift):
    d = swift.generate(1, seed=0)
    assert d["m"] == d["n"] == 4096 and d["k"] == 32
    assert abs(d["nnz"] - 80000) < 800
    lens = np.diff(d["rowptr"])
    assert lens.min() >= 1 and lens.max() <= 512
    assert np.all(np.abs(d["val"]) <= 1.0)
    for i in range(0, 4096, 97):  # columns ascending and distinct inside rows
        c = d["colidx"][d["rowptr"][i]:d["rowptr"][i + 1]]
        assert np.all(np.diff(c) > 0)


def test_config2_shape_small(swift):
    d = swift.generate(2, scale=0.01, seed=0)
    m = d["m"]
    assert m == 10000
    assert abs(d["nnz"] - 160000) < 0.02 * 160000
    rp, ci = d["rowptr"], d["colidx"]
    for i in range(0, m, 2):
        c = ci[rp[i]:rp[i + 1]]
        assert 8 <= len(c) <= 32 and np.all(np.abs(c - i) <= 32)


def test_config4_symmetric(swift):
    d = swift.generate(4, scale=0.01, seed=0)
    import scipy.sparse as sp
    A = sp.csr_matrix((d["val"], d["colidx"], d["rowptr"]), shape=(d["m"], d["n"]))
    assert abs(A - A.T).max() < 1e-15
    assert np.all(A.diagonal() > 0)
    deg = np.diff(d["rowptr"])
    np.testing.assert_allclose(A.diagonal(), 1.0 / deg, rtol=1e-12)


def test_dense_generator(swift):
    x = swift.generate_dense(10000, 0, -1.0, 1.0, seed=3)
    assert x.min() >= -1 and x.max() < 1 and abs(x.mean()) < 0.05
    y = swift.generate_dense(10000, 1, seed=3)
    assert abs(y.std() - 1) < 0.05
