"""Host-side predicates on operand shapes (csrc/bounds.h), compiled with g++.

buffer_ok decides whether the walks read B through a buffer resource with
24-bit row offsets (__umul24(row, stride)).  Each case below sits on one side
of one of its three limits; the stride case is the one a rows-only check
misses: 100 rows of FP32 K = 5M are 2 GB short of the byte limit, yet one
row is 20 MB, past the 24 bits the multiply keeps."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "swift-high-performance-sparse-dense-matrix-multiplication-on-gpus_amd",
                    "csrc")

CASES = [  # (rows, ldb, element bytes, expected)
    (1000000, 128, 4, True),          # the 10^6-row benchmark operand, 512 MB
    (4400000, 128, 4, False),         # 2.25 GB: past 2^31 bytes
    ((1 << 24) - 1, 8, 2, True),      # last row count below 2^24
    (1 << 24, 8, 2, False),           # 2^24 rows
    (100, 5000000, 4, False),         # 20 MB rows: the stride leaves 24 bits
    (100, (1 << 22) - 1, 4, True),    # stride just below 2^24 bytes
    (100, 1 << 22, 4, False),         # stride of exactly 2^24 bytes
    (0, 128, 8, True),
]

PROGRAM = r"""
#include <cstdio>
#include "bounds.h"
int main() {
    struct { long long r, l; unsigned e; int want; } c[] = {%s};
    int bad = 0;
    for (auto &x : c) {
        const int got = swift::buffer_ok(x.r, x.l, x.e) ? 1 : 0;
        if (got != x.want) { std::printf("rows %%lld ldb %%lld elem %%u: %%d\n", x.r, x.l, x.e, got); bad++; }
    }
    std::printf(bad ? "fail\n" : "ok\n");
    return bad != 0;
}
"""


def test_buffer_ok_limits(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("g++ is not available")
    rows = ",".join("{%dLL,%dLL,%du,%d}" % (r, l, e, int(w)) for r, l, e, w in CASES)
    src = tmp_path / "bounds_check.cpp"
    src.write_text(PROGRAM % rows)
    exe = str(tmp_path / "bounds_check")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", CSRC, str(src), "-o", exe], check=True)
    p = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    print(p.stdout)
    assert p.returncode == 0 and p.stdout.strip() == "ok"
