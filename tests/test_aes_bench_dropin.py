"""The drop-in boundary: the reference's own bench/aes_bench.cpp, unchanged,
compiles against this repository's host API (headers named like the
reference's, aes-walsh-ckks_amd/compat, forwarding to host/ckks.hpp) and
links to its libraries.  The CPU test compiles it afresh where the
reference sources exist; the GPU test runs the binary oracle/Makefile
built (oracle/_ref/aes_bench) on the walsh15 and xboot14-sparse-1024 rows
and checks their RESULT lines (0 mismatches, as aes_bench.cpp:293-308
requires)."""
import os
import re
import subprocess

import pytest

from helpers import ROOT

REF = "/root/reference/bench/aes_bench.cpp"
PKG = os.path.join(ROOT, "aes-walsh-ckks_amd")
BIN = os.path.join(ROOT, "oracle", "_ref", "aes_bench")


@pytest.mark.skipif(not os.path.exists(REF), reason="reference sources absent")
def test_reference_aes_bench_compiles_unchanged(tmp_path):
    out = tmp_path / "aes_bench"
    r = subprocess.run(["g++", "-O0", "-std=c++20", "-fopenmp", "-w", "-I" + os.path.join(PKG, "compat"),
                        "-o", str(out), REF, "-L" + PKG, "-lckkshost", "-lmickks", "-Wl,-rpath," + PKG],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    assert out.exists()


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(BIN), reason="oracle/_ref/aes_bench not built (make -C oracle ref)")
@pytest.mark.parametrize("row,pattern", [
    ("walsh15", r"RESULT method=Walsh parameter=AESWalsh15 blocks=1024 "),
    ("xboot14-sparse-1024", r"RESULT method=XBOOT\* parameter=AESXBoot14 blocks=1024 "),
])
def test_reference_aes_bench_row(row, pattern):
    print("aes_bench", row, flush=True)
    r = subprocess.run([BIN, row], capture_output=True, text=True, timeout=600)
    print(r.stdout[-2000:], flush=True)
    assert r.returncode == 0, r.stderr[-3000:]
    m = re.search(pattern + r".* mismatches=(\d+)", r.stdout)
    assert m and int(m.group(1)) == 0
