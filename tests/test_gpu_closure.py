"""Mechanics of the learned collision closure (GPU), without a trained network.

* The recorded collision term of a step is J = I + I^dagger of its first stage
  (G1_G2_t_step.h:41, collision_integral.h:27-69): against the oracle's
  collision integral of the state.
* An HF context given the exact J of a full run reproduces that run's G1: the
  full RK4 evaluates J at every stage input, the closure takes J at the step's
  two ends (their mean at the midpoint stages), so the two differ by O(dt^2);
  the difference falls about fourfold when dt halves, and is far below the
  effect of J itself (the same run without it).  Trace and anti-hermiticity of
  G1 hold to rounding.
* A given term covers exactly the steps it was set for.
"""
import numpy as np
import pytest

from tests.helpers import physical_g2

pytestmark = pytest.mark.gpu

g = pytest.importorskip("g12kba_amd")
from g12kba_amd import extrapolator as xp  # noqa: E402


def params(dt=0.02, steps=40):
    # quench gate open: the ramp acts within the run
    return g.make_params(Ns=5, Nb=2, t1=1.0, t2=0.4, U=1.6, dt=dt, time_steps=steps, quench_strength=0.8, Nq=2,
                         q_type="full", quench_rate=0.2, quench_on_midpoint=0.1, quench_gate_time=0.0, EHM=True,
                         V_law="coulomb")


def start(p, seed=3):
    rng = np.random.default_rng(seed)
    return g.make_G0(p), physical_g2(rng, p.Ns, p.Nb, 2e-2), physical_g2(rng, p.Ns, p.Nb, 2e-2)


def test_recorded_collision_term_matches_oracle():
    from oracle.oracle import Oracle
    p = params()
    G1, Gxx, Gxy = start(p)
    with g.Propagator(p) as prop:
        prop.set_state(G1, Gxx, Gxy)
        rc, traj, J = prop.record_collision(1)
    assert rc == 0 and len(J) == 1
    I = Oracle().collision(Gxy, Gxx, p.Ns, p.Nb, p.U, p.U * g.v_profile(p))
    want = I + I.conj().T
    assert np.abs(J[0] - want).max() <= 1e-12 * np.abs(want).max()
    assert np.abs(J[0] - J[0].conj().T).max() <= 1e-13 * np.abs(want).max()


def closure_vs_full(dt, steps):
    p = params(dt, steps)
    G1, Gxx, Gxy = start(p)
    with g.Propagator(p) as full:
        full.set_state(G1, Gxx, Gxy)
        rc, traj, J = full.record_collision(steps + 1)
        assert rc == 0
    with g.Propagator(xp.hf_params(p)) as hf:
        hf.set_state(G1)
        hf.set_collision(J[:steps + 1])
        rc, tc = hf.record(steps)
        assert rc == 0
        with pytest.raises(g.G12Error):  # the term is used up
            hf.record(1)
        hf.set_state(G1)
        hf.set_collision(None)
        _, t0 = hf.record(steps)  # the same run without the collision term
    return traj[:steps], tc, t0


def test_closure_with_exact_term_reproduces_full_run():
    full, clo, bare = closure_vs_full(0.02, 40)
    err = np.abs(clo - full).max()
    effect = np.abs(bare - full).max()
    assert err < 1e-3 * effect, (err, effect)
    tr = np.trace(clo, axis1=1, axis2=2)
    assert np.abs(tr - tr[0]).max() < 1e-13
    assert np.abs(clo + np.conj(np.transpose(clo, (0, 2, 1)))).max() < 1e-13
    full2, clo2, _ = closure_vs_full(0.01, 80)
    err2 = np.abs(clo2[::2] - full2[::2]).max()
    assert 2.5 < err / err2 < 6.5, (err, err2)  # second order in dt


def test_set_collision_only_on_uncorrelated_contexts():
    p = params()
    with g.Propagator(p) as full:
        with pytest.raises(g.G12Error):
            full.set_collision(np.zeros((3, 10, 10)))
    with g.Propagator(xp.hf_params(p)) as hf:
        with pytest.raises(g.G12Error):
            hf.record_collision(2)  # nothing to record
        with pytest.raises(g.G12Error):
            hf.set_collision(np.zeros((300, 10, 10)))  # more than a window
        hf.set_collision(np.zeros((3, 10, 10)))
        with pytest.raises(g.G12Error):
            hf.record(3)  # covers 2 steps
