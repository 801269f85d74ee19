"""Device-resident AS / RK4 drivers against the CPU oracle (FP64)."""
import numpy as np
import pytest

from tests.helpers import oracle_cfg, rand_c

pytestmark = pytest.mark.gpu

g = pytest.importorskip("g12kba_amd")
from oracle.oracle import Oracle  # noqa: E402
from oracle import fast  # noqa: E402

O = Oracle()


def small_state(rng, p, scale=1e-2):
    N = p.Ns * p.Nb
    G1 = g.make_G0(p)
    A = rand_c(rng, N * N, N * N, scale=scale)
    B = rand_c(rng, N * N, N * N, scale=scale)
    return G1, A - A.conj().T, B - B.conj().T


@pytest.mark.parametrize("ehm,law,qt", [(False, "exp", "full"), (True, "coulomb", "pulse"), (True, "exp", "full")])
@pytest.mark.parametrize("Ns,Nb", [(3, 2), (4, 2), (2, 3)])
def test_rk4_matches_oracle(Ns, Nb, ehm, law, qt):
    p = g.make_params(Ns=Ns, Nb=Nb, t1=1.0, t2=0.4, U=1.4, dt=0.02, time_steps=12, quench_strength=1.3,
                      Nq=max(1, Ns // 2), q_type=qt, quench_rate=2.0, quench_on_midpoint=0.05,
                      quench_off_midpoint=0.15, quench_gate_time=0.06, EHM=ehm, V_law=law)
    rng = np.random.default_rng(Ns * 100 + Nb)
    G1, Gxx, Gxy = small_state(rng, p)
    with g.Propagator(p) as prop:
        prop.set_state(G1, Gxx, Gxy)
        rc, traj = prop.rk4()
        G1g, Gxxg, Gxyg = prop.get_state()
    rc2, G1o, Gxxo, Gxyo, trajo = O.rk4(oracle_cfg(p), G1, Gxx, Gxy, V_profile=g.v_profile(p))
    assert rc == rc2 == 0
    assert traj.shape == trajo.shape
    assert np.abs(traj - trajo).max() < 1e-12
    for a, b in ((G1g, G1o), (Gxxg, Gxxo), (Gxyg, Gxyo)):
        assert np.abs(a - b).max() <= 1e-12 * max(1.0, np.abs(b).max())


@pytest.mark.parametrize("Ns,Nb", [(1, 1), (1, 2), (1, 3), (2, 1), (1, 8), (2, 8)])
def test_tiny_geometries_match_oracle(Ns, Nb):
    """Edge sizes N = 1..3 (one site: no hopping, no long-range shells; one
    orbital: no t2), padded to 16-wide tiles on the device, and the largest
    orbital count (Nb = 8)."""
    p = g.make_params(Ns=Ns, Nb=Nb, t1=1.0, t2=0.4, U=1.4, dt=0.02, time_steps=12, quench_strength=1.3,
                      Nq=max(1, Ns // 2), q_type="pulse", quench_rate=2.0, quench_on_midpoint=0.05,
                      quench_off_midpoint=0.15, quench_gate_time=0.06, EHM=True, V_law="coulomb")
    N = Ns * Nb
    rng = np.random.default_rng(N)
    _, Gxx, Gxy = small_state(rng, p)
    G1 = g.make_G0(p) + 0.1j * np.eye(N)
    with g.Propagator(p) as prop:
        prop.set_state(G1, Gxx, Gxy)
        rc, traj = prop.rk4()
        G1g, Gxxg, Gxyg = prop.get_state()
    rc2, G1o, Gxxo, Gxyo, trajo = O.rk4(oracle_cfg(p), G1, Gxx, Gxy, V_profile=g.v_profile(p))
    assert rc == rc2 == 0 and traj.shape == trajo.shape
    assert np.abs(traj - trajo).max() < 1e-12
    for a, b in ((G1g, G1o), (Gxxg, Gxxo), (Gxyg, Gxyo)):
        assert np.abs(a - b).max() <= 1e-12 * max(1.0, np.abs(b).max())


def test_adiabatic_switching_matches_oracle():
    p = g.make_params(Ns=3, Nb=2, t1=1.0, t2=0.5, U=1.0, AS_rate=0.05, AS_midpoint=0.3, AS_tail=4,
                      EHM=True, V_law="coulomb")
    rng = np.random.default_rng(11)
    G1, Gxx, Gxy = small_state(rng, p, scale=0.0)
    with g.Propagator(p) as prop:
        prop.set_state(G1, Gxx, Gxy)
        n = prop.adiabatic_switching()
        G1g, Gxxg, Gxyg = prop.get_state()
    G1o, Gxxo, Gxyo, n2 = O.adiabatic_switching(oracle_cfg(p), G1, Gxx, Gxy, V_profile=g.v_profile(p))
    assert n == n2
    assert np.abs(G1g - G1o).max() < 1e-12
    assert np.abs(Gxxg - Gxxo).max() < 1e-12 and np.abs(Gxyg - Gxyo).max() < 1e-12


def test_density_guard():
    """A state with 2 Im G1[0,0] > 3 is recorded once and RK4 returns 1 (RK_solver.h:100-110)."""
    p = g.make_params(Ns=3, Nb=1, U=1.0, dt=0.02, time_steps=10)
    G1 = g.make_G0(p)
    G1[0, 0] = 2j
    with g.Propagator(p) as prop:
        prop.set_state(G1)
        rc, traj = prop.rk4()
        after = prop.get_G1()
    assert rc == 1 and len(traj) == 1
    assert np.array_equal(traj[0], G1) and np.array_equal(after, G1)


def test_hf_only_freezes_g2():
    p = g.make_params(Ns=4, Nb=1, U=1.0, HF=True, dt=0.05, time_steps=6)
    rng = np.random.default_rng(3)
    G1, Gxx, Gxy = small_state(rng, p)
    with g.Propagator(p) as prop:
        prop.set_state(G1, Gxx, Gxy)
        rc, traj = prop.rk4()
        G1g, Gxxg, Gxyg = prop.get_state()
    rc2, G1o, _, _, _ = O.rk4(oracle_cfg(p), G1, Gxx, Gxy)
    assert rc == 0 and np.array_equal(Gxxg, Gxx) and np.array_equal(Gxyg, Gxy)
    assert np.abs(G1g - G1o).max() < 1e-13


def test_n32_stage_matches_factorised_oracle():
    """One RK4 step at the N = 32 production size against the numpy oracle."""
    p = g.make_params(Ns=16, Nb=2, t1=1.0, t2=0.3, U=2.0, dt=0.01, time_steps=1, EHM=True, V_law="coulomb")
    rng = np.random.default_rng(5)
    G1, Gxx, Gxy = small_state(rng, p, scale=1e-3)
    with g.Propagator(p) as prop:
        prop.set_state(G1, Gxx, Gxy)
        rc, _ = prop.rk4()
        G1g, Gxxg, Gxyg = prop.get_state()
    cfg = oracle_cfg(p)
    Ns, Nb = 16, 2
    rc2, G1o, Txx, Txy, _ = fast.rk4(cfg, G1, fast.ref_to_nat(Gxx, Ns, Nb), fast.ref_to_nat(Gxy, Ns, Nb),
                                     V_profile=list(g.v_profile(p)))
    assert rc == rc2 == 0
    assert np.abs(G1g - G1o).max() < 1e-13
    assert np.abs(Gxxg - fast.nat_to_ref(Txx, Ns, Nb)).max() < 1e-13
    assert np.abs(Gxyg - fast.nat_to_ref(Txy, Ns, Nb)).max() < 1e-13


def test_checkpoint_resume_is_bit_exact(tmp_path):
    """Binary checkpoint of (t, step, G1, G2) mid-run, loaded into a new
    context and continued: the trajectory equals the uninterrupted run."""
    p = g.make_params(Ns=4, Nb=2, t1=1.0, t2=0.4, U=1.2, dt=0.02, time_steps=40, quench_strength=1.0, Nq=2,
                      quench_rate=2.0, quench_on_midpoint=0.1, quench_gate_time=0.2, EHM=True, V_law="coulomb")
    rng = np.random.default_rng(21)
    G1, Gxx, Gxy = small_state(rng, p)
    with g.Propagator(p) as prop:
        prop.set_state(G1, Gxx, Gxy)
        rc, full = prop.record(30)
        end = prop.get_state()
    path = str(tmp_path / "state.bin")
    with g.Propagator(p) as prop:
        prop.set_state(G1, Gxx, Gxy)
        rc1, first = prop.record(12)
        prop.save(path)
    with g.Propagator(p) as prop:
        prop.load(path)
        assert prop.time()[1] == 12
        rc2, rest = prop.record(18)
        end2 = prop.get_state()
    assert rc == rc1 == rc2 == 0
    assert np.array_equal(np.concatenate([first, rest]), full)
    for a, b in zip(end, end2):
        assert np.array_equal(a, b)


def test_domain_wall_run_matches_oracle():
    """From init_G_test's domain wall (init_G.h:78-86): switching and 60 RK4
    steps against the factorised oracle; the wall melts (the dipole moves)."""
    p = g.make_params(Ns=6, Nb=2, t1=1.0, t2=0.4, U=1.2, dt=0.02, time_steps=60, quench_strength=0.0, Nq=0,
                      EHM=True, V_law="coulomb")
    N = p.Ns * p.Nb
    G1 = g.init_G_test(p)
    Z = np.zeros((N * N, N * N), np.complex128)
    with g.Propagator(p) as prop:
        prop.set_state(G1)
        nas = prop.adiabatic_switching()
        rc, traj = prop.rk4()
        G1g = prop.get_state()[0]
    ref = O.fast_run(oracle_cfg(p), G1, Z, Z, V_profile=list(g.v_profile(p)), traj_every=1, traj_cap=80)
    assert rc == ref["rc"] == 0 and nas == ref["as_steps"] and len(traj) == ref["rk_steps"]
    assert np.abs(traj - ref["traj"]).max() < 1e-11
    assert np.abs(G1g - ref["G1"]).max() < 1e-11
    d = g.dipole(traj)
    assert np.abs(d - d[0]).max() > 1e-3
