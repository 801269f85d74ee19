"""Runs of many trajectories (g12_batch_*): every trajectory of a batch must
match the same trajectory run alone and the CPU oracle; hipGraph replay and
the tile autotuner must not change a single bit."""
import numpy as np
import pytest

from tests.helpers import oracle_cfg, physical_g2

pytestmark = pytest.mark.gpu

g = pytest.importorskip("g12kba_amd")
from oracle.oracle import Oracle  # noqa: E402

O = Oracle()


def ensemble(seed, n, time_steps, sizes=None, **over):
    """Uneven sizes and physics with one shared protocol (short switching)."""
    rng = np.random.default_rng(seed)
    ps = []
    for i in range(n):
        Ns, Nb = sizes[i] if sizes else (int(rng.integers(2, 7)), int(rng.integers(1, 4)))
        kw = dict(Ns=Ns, Nb=Nb, t1=1.0, t2=float(rng.uniform(0.1, 0.8)), U=float(rng.uniform(0.5, 3.0)),
                  dt=0.02, time_steps=time_steps, quench_strength=float(rng.uniform(0.1, 2.0)),
                  Nq=max(1, Ns // 4), quench_rate=2.0, quench_on_midpoint=0.1, quench_gate_time=0.04,
                  EHM=bool(rng.integers(0, 2)), V_law="coulomb", AS_rate=0.05, AS_midpoint=0.3, AS_tail=3)
        kw.update(over)
        ps.append(g.make_params(**kw))
    return ps


def iterations(p):
    """RK4 iterations of a production run: accumulated time, RK_solver.h:96,216."""
    n, t = 0, 0.0
    while t < p.time_steps * p.dt:
        t += p.dt
        n += 1
    return n


def start(p):
    N = p.Ns * p.Nb
    z = np.zeros((N * N, N * N), np.complex128)
    return g.make_G0(p), z, z


def run_single(p, G1, Gxx, Gxy, graphs=True, autotune=True):
    with g.Propagator(p) as prop:
        prop.configure(graphs=graphs, autotune=autotune)
        prop.set_state(G1, Gxx, Gxy)
        a = prop.adiabatic_switching()
        rc, traj = prop.rk4()
        return a, rc, traj, prop.get_state()


def run_batch(ps, states, graphs=True, autotune=True):
    with g.Batch(ps) as b:
        b.configure(graphs=graphs, autotune=autotune)
        for i, s in enumerate(states):
            b.set_state(i, *s)
        a = b.adiabatic_switching()
        rc, trajs = b.rk4()
        finals = [b.get_state(i) for i in range(len(ps))]
        return a, rc, trajs, finals, b.buckets()


def test_batch_matches_single_runs_and_oracle():
    # NP = 16 and NP = 32 buckets, an HF-only and a U = 0 trajectory (no
    # correlations, shorter switching), one pulse quench
    sizes = [(3, 2), (6, 1), (3, 3), (10, 2), (4, 2), (5, 1), (2, 2)]
    ps = ensemble(1, len(sizes), 16, sizes=sizes)
    ps[4].HF = 1
    ps[5].U = 0.0
    states = [start(p) for p in ps]
    a, rc, trajs, finals, buckets = run_batch(ps, states)
    assert rc == 0
    assert sorted(k[0] for k in buckets) == [16, 32]
    assert a[5] == 4 and a[0] > a[5]  # U = 0: the exact test holds from the second iteration
    for i, p in enumerate(ps):
        a1, rc1, traj1, fin1 = run_single(p, *states[i])
        assert a1 == a[i] and rc1 == 0
        assert np.array_equal(traj1, trajs[i]), f"trajectory {i}"
        for x, y in zip(fin1, finals[i]):
            assert np.array_equal(x, y), f"trajectory {i}"
        ref = O.fast_run(oracle_cfg(p), *states[i], V_profile=list(g.v_profile(p)), traj_every=1,
                         traj_cap=p.time_steps + 2)
        assert ref["rc"] == 0 and ref["as_steps"] == a[i]
        assert ref["traj"].shape == trajs[i].shape
        assert np.abs(ref["traj"] - trajs[i]).max() < 1e-12, f"trajectory {i}"


def test_graph_replay_is_bit_exact_across_windows():
    # 600 steps: three table windows (256 + 256 + 88 steps)
    p = g.make_params(Ns=4, Nb=2, t1=1.0, t2=0.5, U=1.0, dt=0.02, time_steps=600, quench_strength=1.0, Nq=2,
                      quench_rate=2.0, quench_on_midpoint=2.0, EHM=True, V_law="coulomb")
    rng = np.random.default_rng(2)
    N = 8
    A = (rng.standard_normal((N * N, N * N)) + 1j * rng.standard_normal((N * N, N * N))) * 1e-3
    G1, Gxx, Gxy = g.make_G0(p), A - A.conj().T, 0.5 * (A + A.T.conj()) * 1j
    out = []
    for graphs in (True, False):
        with g.Propagator(p) as prop:
            prop.configure(graphs=graphs)
            prop.set_state(G1, Gxx, Gxy)
            rc, traj = prop.rk4()
            out.append((rc, traj, prop.get_state(), prop.time()))
    (rc1, t1, s1, tm1), (rc2, t2, s2, tm2) = out
    assert rc1 == rc2 == 0 and len(t1) == iterations(p) == 601
    assert np.array_equal(t1, t2)
    for x, y in zip(s1, s2):
        assert np.array_equal(x, y)
    assert tm1 == tm2
    ref = O.rk4(oracle_cfg(p), G1, Gxx, Gxy, V_profile=g.v_profile(p))
    assert np.abs(ref[4] - t1).max() < 1e-12


def test_guard_stops_only_its_trajectory():
    sizes = [(3, 2), (4, 1), (2, 3)]
    ps = ensemble(3, 3, 12, sizes=sizes)
    states = [start(p) for p in ps]
    G1 = states[1][0].copy()
    G1[0, 0] = 2j  # 2 Im G1[0,0] = 4 > 3: recorded once, then stopped (RK_solver.h:100-110)
    states[1] = (G1, states[1][1], states[1][2])
    with g.Batch(ps) as b:
        for i, s in enumerate(states):
            b.set_state(i, *s)
        rc, trajs = b.rk4()
        assert b.guard() == [0, 1, 0]
        assert np.array_equal(b.get_G1(1), G1)
    assert rc == 1
    n = iterations(ps[0])
    assert [len(t) for t in trajs] == [n, 1, n]
    for i in (0, 2):
        with g.Propagator(ps[i]) as prop:
            prop.set_state(*states[i])
            rc1, traj1 = prop.rk4()
        assert rc1 == 0 and np.array_equal(traj1, trajs[i])


def test_autotune_changes_no_bits():
    sizes = [(8, 2), (5, 3), (12, 2), (4, 4)]
    ps = ensemble(4, 4, 6, sizes=sizes)
    states = [start(p) for p in ps]
    a1, rc1, tr1, f1, bk1 = run_batch(ps, states, autotune=True)
    a2, rc2, tr2, f2, bk2 = run_batch(ps, states, autotune=False)
    assert a1 == a2 and rc1 == rc2 == 0
    for x, y in zip(tr1, tr2):
        assert np.array_equal(x, y)
    for NP, n, ws, wf in bk1:
        w0 = {16: 1, 32: 2, 48: 4, 64: 4}[NP]
        assert ws in (w0, 2 * w0, 4 * w0) and wf in (w0, 2 * w0, 4 * w0)
    assert sum(k[1] for k in bk1) == 4


@pytest.mark.parametrize("prec", ["fp32_3m", "bf16x3"])
def test_mixed_precision_batch_error(prec):
    """Per-trajectory G1 error of a mixed-precision batch against the FP64 batch."""
    sizes = [(6, 2), (9, 1), (4, 3), (10, 2)]
    ps64 = ensemble(5, 4, 50, sizes=sizes)
    psm = ensemble(5, 4, 50, sizes=sizes, precision=prec)
    states = [start(p) for p in ps64]
    _, rc1, tr1, _, _ = run_batch(ps64, states)
    _, rc2, tr2, _, _ = run_batch(psm, states)
    assert rc1 == rc2 == 0
    for x, y in zip(tr1, tr2):
        dens = np.abs(np.diagonal(x - y, axis1=1, axis2=2).imag).max()
        assert 0 < dens < 1e-6


@pytest.mark.parametrize("prec", ["fp64", "fp32_3m"])
def test_batch_checkpoint_resume_is_bit_exact(tmp_path, prec):
    """Every trajectory of a batch saved mid-run and loaded into a new batch:
    continuing gives the uninterrupted batch's trajectories bit for bit."""
    sizes = [(6, 2), (9, 1), (4, 3)]
    ps = ensemble(6, 3, 40, sizes=sizes, precision=prec)
    states = [start(p) for p in ps]

    def new_batch():
        b = g.Batch(ps)
        for i, s in enumerate(states):
            b.set_state(i, *s)
        return b

    with new_batch() as b:
        b.adiabatic_switching()
        rc, full = b.record(25)
        end = [b.get_state(i) for i in range(len(ps))]
    with new_batch() as b:
        b.adiabatic_switching()
        rc1, first = b.record(10)
        for i in range(len(ps)):
            b.save(i, str(tmp_path / f"t{i}.bin"))
    with g.Batch(ps) as b:
        for i in range(len(ps)):
            b.load(i, str(tmp_path / f"t{i}.bin"))
        rc2, rest = b.record(15)
        end2 = [b.get_state(i) for i in range(len(ps))]
    assert rc == rc1 == rc2 == 0
    for i in range(len(ps)):
        assert np.array_equal(np.concatenate([first[i], rest[i]]), full[i]), f"trajectory {i}"
        for x, y in zip(end[i], end2[i]):
            assert np.array_equal(x, y), f"trajectory {i}"


@pytest.mark.parametrize("prec", ["fp64", "fp32_3m", "bf16x3"])
def test_full_tile_kernels_match_general_kernels(prec):
    """A bucket whose trajectories all have N = NP runs the full-tile kernels
    (no element bounds checks); sharing the bucket with a smaller N makes the
    same trajectory run the general ones.  The N = 48 trajectory must come out
    the same bit for bit (16 waves: the default at NP = 48, autotuner off)."""
    sizes = [(24, 2), (47, 1)]
    ps = ensemble(11, 2, 6, sizes=sizes, precision=prec)
    rng = np.random.default_rng(12)
    G1 = g.make_G0(ps[0])
    s48 = (G1, physical_g2(rng, 24, 2), physical_g2(rng, 24, 2))
    s47 = start(ps[1])
    a1, rc1, traj1, fin1, b1 = run_batch([ps[0]], [s48], autotune=False)
    a2, rc2, traj2, fin2, b2 = run_batch(ps, [s48, s47], autotune=False)
    assert rc1 == 0 and rc2 == 0
    assert [k[0] for k in b1] == [48] and [k[0] for k in b2] == [48]
    assert a1[0] == a2[0] and len(traj1[0]) == len(traj2[0]) > 0
    assert np.array_equal(traj1[0], traj2[0])
    for x, y in zip(fin1[0], fin2[0]):
        assert np.array_equal(x, y)
