"""The C driver (host/swift_driver.cpp) in each study mode of the reference:
exit status 0 (every product passed ResultVerify) and result lines whose
tokens sit where the reference's parsers read them (e.g. vendor time at
token 10 and Swift time at token 12, DataProcess/CMA-compare/cma_compare.py)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "swift-high-performance-sparse-dense-matrix-multiplication-on-gpus_amd",
                      "swift_driver")
MTX = os.path.join(ROOT, "tests", "golden", "mtx", "general_real.mtx")


def run(args, tmp_path):
    if not os.path.exists(DRIVER):
        pytest.fail("swift_driver is not built")
    print("driver", " ".join(args), flush=True)
    p = subprocess.run([DRIVER, "-d", "0", "--seed", "1", "--reps", "2"] + args, cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    print(p.stdout[-2000:], p.stderr[-2000:], flush=True)
    assert p.returncode == 0, p.stderr
    return p.stdout


def lines(path):
    with open(path) as f:
        return [l.split() for l in f if l.strip()]


@pytest.mark.parametrize("prec", ["--fp32", "--fp64"])
@pytest.mark.parametrize("src", ["gen", "mtx"])
def test_main_mode(tmp_path, prec, src):
    args = ["--gen", "1", "--scale", "0.25", "32"] if src == "gen" else [MTX, "32"]
    run(args + [prec, "--out", "r.txt"], tmp_path)
    (t,) = lines(tmp_path / "r.txt")
    assert t[1] == "m" and t[9] == "rocSPARSE:" and t[11] == "Swift"
    assert float(t[10]) > 0 and float(t[12]) > 0


@pytest.mark.parametrize("prec", ["--fp32", "--fp64"])
def test_ksweep_mode(tmp_path, prec):
    run(["--gen", "1", "--scale", "0.25", prec, "--mode", "ksweep", "--ks", "24,48,96",
         "--out", "sweep"], tmp_path)
    for K in (24, 48, 96):
        (t,) = lines(tmp_path / ("sweep_K%d.txt" % K))
        assert int(t[6]) == K and t[11] == "final"
        assert float(t[10]) > 0 and float(t[12]) > 0


def test_cma_mode(tmp_path):
    run(["--gen", "1", "--scale", "0.25", "--fp32", "--mode", "cma", "--out", "c.txt", "32"],
        tmp_path)
    (t,) = lines(tmp_path / "c.txt")
    assert t[9] == "withoutCMA" and t[11] == "withCMA"
    assert float(t[10]) > 0 and float(t[12]) > 0


def test_irropt_mode(tmp_path):
    run(["--gen", "1", "--scale", "0.25", "--fp32", "--mode", "irropt", "--out", "i.txt", "32"],
        tmp_path)
    (t,) = lines(tmp_path / "i.txt")
    assert t[9] == "non-opt" and t[11] == "opt"
    assert float(t[10]) >= 0 and float(t[12]) >= 0


def test_preprocess_mode(tmp_path):
    run(["--gen", "1", "--scale", "0.25", "--fp32", "--mode", "preprocess", "--out", "p.txt",
         "32"], tmp_path)
    (t,) = lines(tmp_path / "p.txt")
    v = dict(zip(t[7::2], map(float, t[8::2])))
    assert set(v) == {"sort", "ratio", "block", "preprocess", "final", "overall_time"}
    assert v["preprocess"] == pytest.approx(float(t[8]) + float(t[12]), rel=1e-4)
    assert v["overall_time"] == pytest.approx(v["preprocess"] + v["final"], rel=1e-4)


def test_mtxinfo_mode(tmp_path):
    run(["--gen", "1", "--scale", "0.25", "--fp32", "--mode", "mtxinfo", "--out", "s.txt", "32"],
        tmp_path)
    (t,) = lines(tmp_path / "s.txt")
    assert t[11] == "sparsity" and t[13] == "emptyblkratio" and t[15] == "final"
    assert 0 < float(t[12]) < 1 and 0 <= float(t[14]) <= 1 and float(t[16]) > 0
    (r,) = lines(tmp_path / "s_irr_regular_ratio.txt")
    assert r[3] == "regulatpart" and r[6] == "irregularpart"
    assert int(r[4]) + int(r[7]) == int(r[2])
    assert float(r[5]) + float(r[8]) == pytest.approx(1.0, abs=1e-5)


@pytest.mark.parametrize("prec", ["--fp32", "--fp64"])
def test_vendor_mode(tmp_path, prec):
    """rocSPARSE alone: every CSR algorithm in both dense orders, each verified
    against the host golden, and the best per order (the vendor leg of
    bench.py)."""
    out = run(["--gen", "1", "--scale", "0.25", prec, "--mode", "vendor", "32"], tmp_path)
    (v,) = [l.split() for l in out.splitlines() if l.startswith("vendor rocSPARSE")]
    assert v[10] == "row-major" and v[12] == "col-major"
    assert v[14] == "row-alg" and v[16] == "col-alg"
    assert float(v[11]) > 0 and float(v[13]) > 0
    algs = [l.split() for l in out.splitlines() if l.startswith("vendor-alg")]
    assert [a[1] for a in algs] == ["default", "csr", "csr_row_split", "csr_nnz_split",
                                    "csr_merge_path"]
    assert float(v[11]) == min(float(a[3]) for a in algs if float(a[3]) > 0)
