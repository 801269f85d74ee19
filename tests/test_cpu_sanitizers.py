"""Host-side C++ under AddressSanitizer + UndefinedBehaviorSanitizer.

oracle/asan_driver.cpp drives the boundary's host code (host_io.cpp: the
reference-format input parser on every parser fixture, the text writers, G0,
the domain-wall state, dipole, checkpoints, the Jacobi eigensolver) and the
oracle's drivers (element-wise switching + RK4 against the factorised run)
in a build with -fsanitize=address,undefined and no recovery: any invalid
access, leak or undefined operation fails the run.  The parsed values must
equal the fixtures (which come from the reference's own parser).
"""
import json
import os
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "oracle", "_build", "asan_driver")
FIELDS = ["rc", "HF", "EHM", "t1", "t2", "t3", "U", "time_steps", "dt_fixed", "quench_strength", "Ns", "Nq", "Nb",
          "q_type", "AS_rate", "AS_midpoint", "quench_rate", "quench_on_midpoint", "quench_off_midpoint"]


@pytest.fixture(scope="module")
def driver():
    if shutil.which("g++") is None:
        pytest.skip("no host compiler")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True, capture_output=True,
                   timeout=600)
    return DRIVER


def test_host_code_is_clean_under_sanitizers(driver):
    cases = json.load(open(os.path.join(ROOT, "tests", "golden", "parser_cases.json")))
    d = tempfile.mkdtemp()
    try:
        for name, c in cases.items():
            with open(os.path.join(d, "input_" + name), "w") as f:
                f.write(c["text"])
        env = dict(os.environ, OMP_NUM_THREADS="2", ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:exitcode=23",
                   UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
        r = subprocess.run([driver, d], capture_output=True, text=True, timeout=600, env=env)
        assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
        parsed = {}
        for line in r.stdout.splitlines():
            if line.startswith("parsed "):
                tok = line.split()
                parsed[tok[1]] = dict(zip(FIELDS, [float(x) for x in tok[2:]]))
        assert set(parsed) == set(cases)
        for name, c in cases.items():
            assert parsed[name]["rc"] == c["rc"]
            for k, v in c["values"].items():
                assert parsed[name][k] == pytest.approx(float(v), rel=0, abs=0), (name, k)
        # the writers and the checkpoint left their files
        for f in ("real_G1.txt", "imag_G1.txt", "dipole.txt", "progress.txt", "state.bin"):
            assert os.path.exists(os.path.join(d, f))
    finally:
        shutil.rmtree(d, ignore_errors=True)
