"""The library's host code (Matrix Market loader with its threaded parser,
seeded generators) and the CPU oracle built with g++ under AddressSanitizer +
UndefinedBehaviorSanitizer, and the loader under ThreadSanitizer, then run over
the .mtx fixtures and generated matrices (tests/sanitize/host_check.cpp): the
threaded parser must equal the sequential one, generated matrices must round
trip through .mtx text, and the oracle's Swift_cpu must match its FP64 CSR
product -- with no sanitizer report."""
import glob
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "swift-high-performance-sparse-dense-matrix-multiplication-on-gpus_amd",
                    "csrc")
MTX = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "mtx", "*.mtx")))


def build_and_run(tmp_path, flags, oracle):
    if shutil.which("g++") is None:
        pytest.skip("g++ is not available")
    exe = str(tmp_path / "host_check")
    srcs = [os.path.join(ROOT, "tests", "sanitize", "host_check.cpp"),
            os.path.join(CSRC, "mmio.cpp"), os.path.join(CSRC, "gen.cpp")]
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-pthread",
           "-I", os.path.join(ROOT, "include")] + flags
    objs = []
    if oracle:
        obj = str(tmp_path / "swift_oracle.o")
        subprocess.run(["gcc", "-std=c11", "-O1", "-g", "-fno-omit-frame-pointer", "-c",
                        os.path.join(ROOT, "oracle", "swift_oracle.c"), "-o", obj] + flags,
                       check=True)
        objs = [obj]
        cmd += ["-DWITH_ORACLE", "-I", os.path.join(ROOT, "oracle")]
    subprocess.run(cmd + srcs + objs + ["-o", exe, "-lm"], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1",
               TSAN_OPTIONS="halt_on_error=1")
    p = subprocess.run([exe, str(tmp_path)] + MTX, capture_output=True, text=True, env=env,
                       timeout=600)
    print(p.stdout[-2000:], p.stderr[-4000:])
    assert p.returncode == 0 and p.stdout.strip().endswith("ok")


def test_address_and_ub_sanitizers(tmp_path):
    build_and_run(tmp_path, ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"],
                  oracle=True)


def test_thread_sanitizer_loader(tmp_path):
    build_and_run(tmp_path, ["-fsanitize=thread"], oracle=False)
