"""Resuming runs and the density guard's bookkeeping (GPU).

* The command-line drop-in split in two (``--steps k --save-state``, then
  ``--load-state``) equals the uninterrupted run: the final binary state bit
  for bit, and the text trajectory column for column.
* A run whose density guard trips mid-run (RK_solver.h:100-110): the
  trajectory stops at the failing record, the context's clock stops at the
  state it holds, later calls do nothing until a new state is set, and the
  CLI prints the reference's message and writes progress.txt after the last
  completed step (RK_solver.h:216-220).
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from tests.test_gpu_runs import CONFIG1_INPUT

pytestmark = pytest.mark.gpu

g = pytest.importorskip("g12kba_amd")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "g12kba_amd", "lib", "g12kba")


def run_cli(d, *args):
    print("cli", d, args, flush=True)
    r = subprocess.run([EXE, *args], cwd=d, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def short_config1(d, steps):
    with open(os.path.join(d, "input"), "w") as f:
        f.write(CONFIG1_INPUT.replace("time_steps\n3000\n", f"time_steps\n{steps}\n"))


def test_cli_split_run_resumes_bit_exact():
    a, b = tempfile.mkdtemp(), tempfile.mkdtemp()
    for d in (a, b):
        short_config1(d, 200)
    run_cli(a, "--save-state", "end.bin")
    run_cli(b, "--steps", "80", "--save-state", "mid.bin")
    first = {k: open(os.path.join(b, k)).read() for k in ("real_G1.txt", "imag_G1.txt", "dipole.txt", "progress.txt")}
    run_cli(b, "--load-state", "mid.bin", "--save-state", "end.bin")
    assert open(os.path.join(a, "end.bin"), "rb").read() == open(os.path.join(b, "end.bin"), "rb").read()
    for k in ("real_G1.txt", "imag_G1.txt"):
        whole = [ln.split() for ln in open(os.path.join(a, k)).read().splitlines()]
        head = [ln.split() for ln in first[k].splitlines()]
        tail = [ln.split() for ln in open(os.path.join(b, k)).read().splitlines()]
        assert len(whole) == len(head) == len(tail) == 64
        assert all(len(h) == 80 and len(t) == 120 and w == h + t for w, h, t in zip(whole, head, tail))
    assert first["dipole.txt"].split() + open(os.path.join(b, "dipole.txt")).read().split() == \
        open(os.path.join(a, "dipole.txt")).read().split()
    assert abs(float(first["progress.txt"]) - 40.0) < 1e-9
    assert abs(float(open(os.path.join(b, "progress.txt")).read()) - 100.0) < 1e-9


def rising_density_state():
    """Two sites, U = 0: G1 = 1.8 i v v^dagger with v = (sin 0.3, cos 0.3).
    Hopping moves the weight to site 0, so 2 Im G1[0,0] rises from 0.31 past
    3 (its maximum is 3.29) after about 40 steps of dt = 0.02."""
    v = np.array([np.sin(0.3), np.cos(0.3)])
    return 1.8j * np.outer(v, v)


def test_guard_trip_mid_run_stops_the_clock():
    p = g.make_params(Ns=2, Nb=1, U=0.0, dt=0.02, time_steps=200)
    G0 = rising_density_state()
    with g.Propagator(p) as prop:
        prop.set_state(G0)
        rc_ref, traj_ref = prop.rk4()  # the reference driver from t = 0
        prop.set_state(G0)
        rc, traj = prop.record(200)
        t, n = prop.time()
        rc2, more = prop.record(5)
        rc3, done = prop.advance(5, done=True)
        assert (t, n) == prop.time()
        d = tempfile.mkdtemp()
        prop.save(os.path.join(d, "trip.bin"))
    dens = 2 * traj[:, 0, 0].imag
    k = len(traj)
    assert rc == rc_ref == 1 and 10 < k < 200 and len(traj_ref) == k
    assert np.array_equal(traj, traj_ref)
    assert dens[-1] > 3 and (dens[:-1] <= 3).all()  # stops at the first failing record
    want_t = 0.0
    for _ in range(k - 1):  # accumulated as RK_solver.h:216
        want_t += p.dt
    assert n == k - 1 and t == want_t
    assert rc2 == rc3 == 1 and len(more) == 0 and done == 0
    with g.Propagator(p) as prop:
        prop.load(os.path.join(d, "trip.bin"))
        assert prop.time() == (t, n)
        assert np.array_equal(prop.get_G1(), traj[-1])


def test_cli_guard_message_and_progress():
    d = tempfile.mkdtemp()
    with open(os.path.join(d, "input"), "w") as f:
        # switching (21 steps of 0.1 at U = 0) moves the density past its first
        # maximum; it crosses 3 again about 120 production steps later
        f.write("HF\ntrue\nU\n0\ntime_steps\n200\ndt\n0.02\nNs\n2\nNb\n1\nt1\n1\nAS_rate\n0.5\nAS_midpoint\n10\n")
    G0 = rising_density_state()
    G0.astype(np.complex128).tofile(os.path.join(d, "g0.bin"))
    out = run_cli(d, "--g0", "g0.bin")
    assert "ERROR: Unphysical Number Density on site 1:" in out and "elapsed time:" in out
    dens = float(out.split("ERROR: Unphysical Number Density on site 1:")[1].split()[0])
    assert dens > 3
    cols = len(open(os.path.join(d, "imag_G1.txt")).readline().split())
    assert 10 < cols < 200
    t = 0.0
    for _ in range(cols - 1):
        t += 0.02
    assert abs(float(open(os.path.join(d, "progress.txt")).read()) - t / (200 * 0.02) * 100) < 1e-4
