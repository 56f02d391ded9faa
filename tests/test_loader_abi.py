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


def _same(a, b):
    """Loader results equal bit for bit (values compared as bytes: NaN too)."""
    (ra, da), (rb, db) = a, b
    if ra != rb:
        return False
    if ra != 0:
        return True
    return all(da[k] == db[k] for k in ("m", "n", "nnz", "sym")) and \
        all(da[k].tobytes() == db[k].tobytes() for k in ("rowptr", "colidx", "val"))


@pytest.mark.parametrize("threads", [2, 3, 8])
@pytest.mark.parametrize("path", MTX, ids=[os.path.basename(p) for p in MTX])
def test_parallel_loader_fixtures(swift, path, threads):
    """The parallel parser against the sequential (reference) reading on every
    fixture, error cases included, in both value types."""
    for dt in (swift.F64, swift.F32):
        assert _same(swift.mmio_load(path, threads=1, dtype=dt),
                     swift.mmio_load(path, threads=threads, dtype=dt))
    assert _same(swift.mmio_load(path), swift.mmio_load(path, threads=0))


def _write_mtx(path, field, sym, m, n, entries, rng, split=False, tail=""):
    """Writes entries [(i, j, value token(s))] with uneven whitespace: tabs,
    CRLF, blank lines, and (split) entries broken over lines."""
    out = ["%%%%MatrixMarket matrix coordinate %s %s\n" % (field, sym), "% a comment\n",
           "%d %d %d\n" % (m, n, len(entries))]
    for k, (i, j, v) in enumerate(entries):
        toks = [str(i), str(j)] + v
        seps = [rng.choice([" ", "  ", "\t"]) for _ in toks[1:]]
        if split and k % 7 == 3:
            seps[-1] = "\n"
        line = toks[0] + "".join(s + t for s, t in zip(seps, toks[1:]))
        out.append(line + rng.choice(["\n", "\r\n", " \n", "\n\n"]))
    out.append(tail)
    with open(path, "w", newline="") as f:
        f.write("".join(out))


@pytest.mark.parametrize("kind", ["real", "real_split", "pattern_sym", "integer", "complex_herm",
                                  "special_values", "trailing", "bad_token", "truncated",
                                  "out_of_range"])
def test_parallel_loader_generated(swift, tmp_path, kind):
    rng = np.random.default_rng(hash(kind) % 2**32)
    m, n, E = 3000, 2500, 60000
    I = rng.integers(1, m + 1, E)
    J = rng.integers(1, n + 1, E)
    field, sym, split, tail = "real", "general", False, ""
    vals = [["%.17g" % x] for x in rng.standard_normal(E)]
    if kind == "real_split":
        split = True
    elif kind == "pattern_sym":
        field, sym, n = "pattern", "symmetric", m
        J = rng.integers(1, m + 1, E)
        vals = [[] for _ in range(E)]
    elif kind == "integer":
        field = "integer"
        vals = [[str(int(x))] for x in rng.integers(-1000, 1000, E)]
    elif kind == "complex_herm":
        field, sym, n = "complex", "hermitian", m
        J = rng.integers(1, m + 1, E)
        vals = [["%.6e" % a, "%.6e" % b] for a, b in rng.standard_normal((E, 2))]
    elif kind == "special_values":
        pool = ["1e-310", "-0.0", "inf", "-inf", "nan", "0x1.8p+1", "+3.25", "1E5", "7"]
        vals = [[pool[k % len(pool)]] for k in range(E)]
    elif kind == "trailing":
        tail = "1 1 5.0\n2 2 zzz\n"
    elif kind == "bad_token":
        vals[E // 2] = ["1.5abc"]
    elif kind == "out_of_range":
        I[E - 5] = m + 1
    entries = list(zip(I.tolist(), J.tolist(), vals))
    p = str(tmp_path / (kind + ".mtx"))
    _write_mtx(p, field, sym, m, n, entries, rng, split=split, tail=tail)
    if kind == "truncated":
        txt = open(p).read()
        open(p, "w").write(txt[: int(len(txt) * 0.8)])
    seq = swift.mmio_load(p, threads=1)
    if kind in ("bad_token", "truncated", "out_of_range"):
        assert seq[0] == -3
    else:
        assert seq[0] == 0 and seq[1]["nnz"] >= E
    for t in (2, 5, 16):
        assert _same(seq, swift.mmio_load(p, threads=t)), t
    assert _same(swift.mmio_load(p, threads=1, dtype=swift.F32),
                 swift.mmio_load(p, threads=16, dtype=swift.F32))
