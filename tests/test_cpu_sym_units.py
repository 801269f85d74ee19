"""Host-side check of the pair-symmetric update's work decomposition.

The update pass on the Hermitian half layout (k_sym_update) takes its units
from the table su_table builds (g12kba_amd/csrc/sym_units.cpp) and its
per-thread element arithmetic from sym_units.h.  tests/sym_units_check.cpp
compiles both with g++ and runs them as the kernel does, slot by slot through
its LDS regions, for N = 1..64: exact coverage of the stored elements and the
partial traces, bounds, block limits, and (N <= 12) C = X + P(X) exactly for a
random anti-Hermitian X (that the G2 right-hand side is X + P(X) is pinned
against the oracle by test_cpu_oracle.test_pair_symmetric_decomposition).  A
wrong sign of the mirrored family, a misplaced Hermitian image or an
off-by-one band fails it.
"""
import pathlib
import shutil
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
CSRC = ROOT / "g12kba_amd" / "csrc"


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_sym_update_units_cover_the_half_layout(tmp_path):
    exe = tmp_path / "sym_units_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-I", str(CSRC), str(ROOT / "tests" / "sym_units_check.cpp"),
                    str(CSRC / "sym_units.cpp"), "-o", str(exe)], check=True, capture_output=True, text=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stdout + out.stderr
