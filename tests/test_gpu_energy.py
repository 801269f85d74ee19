"""Energy of GPU states: the device entry (G1 and the G2 pair diagonal read on
the device) against the host formula on the downloaded state, conservation of
the total along GPU runs (Nb = 1, constant h0), and batch = solo."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

g = pytest.importorskip("g12kba_amd")


def quench_shift(p, t):
    """Quench shift on sites < Nq at time t (the production protocol)."""
    if t < p.quench_gate_time:
        return 0.0
    on = g.fermi(t, p.quench_on_midpoint, p.quench_rate)
    if p.q_type == g.QUENCH_PULSE:
        on *= 1 - g.fermi(t, p.quench_off_midpoint, p.quench_rate)
    return on * p.quench_strength


def test_state_energy_matches_host_formula():
    p = g.make_params(Ns=8, Nb=2, t1=1.0, t2=0.5, U=1.2, dt=0.02, time_steps=400, quench_strength=0.7, Nq=3,
                      quench_rate=0.5, quench_on_midpoint=2.0, quench_gate_time=1.0, EHM=True, V_law="coulomb")
    with g.Propagator(p) as prop:
        prop.set_state(g.make_G0(p))
        prop.adiabatic_switching()
        prop.advance(120)
        t, n = prop.time()
        assert n == 120 and t > p.quench_gate_time
        E = prop.energy()
        G1, Gxx, Gxy = prop.get_state()
    w = quench_shift(p, t)
    assert w > 0.1
    ref = g.energy(p, G1, Gxx, Gxy, lam=1.0, w=w)
    assert abs(ref[3]) > 1e-3  # correlated state
    assert np.allclose(E, ref, rtol=1e-12, atol=1e-13 * np.abs(ref).max())


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-9), ("fp32_3m", 1e-6), ("bf16x3", 1e-5)])
def test_total_energy_conserved_on_gpu(prec, tol):
    """Nb = 1, EHM, no quench: after switching (which changes U and with it the
    energy) the total stays fixed over 600 production steps while the
    correlation energy moves; `tol` is relative to the total."""
    p = g.make_params(Ns=16, Nb=1, t1=1.0, U=1.5, dt=0.01, time_steps=600, q_type="none", EHM=True,
                      V_law="coulomb", precision=prec)
    with g.Propagator(p) as prop:
        G0 = g.make_G0(p)
        # a kick off the ground state: the hopping phase of a weak field
        ph = np.exp(0.3j * np.arange(p.Ns))
        prop.set_state(ph[:, None] * G0 * ph.conj()[None, :])
        prop.adiabatic_switching()
        E = [prop.energy()]
        for _ in range(12):
            assert prop.advance(50) == 0
            E.append(prop.energy())
    E = np.array(E)
    tot = E.sum(1)
    assert np.abs(E[:, 3] - E[0, 3]).max() > 1e-3
    assert np.abs(tot - tot[0]).max() < tol * abs(tot[0])


def test_batch_energy_matches_solo():
    ps = [g.make_params(Ns=6, Nb=2, U=1.0, dt=0.02, time_steps=50, EHM=True, V_law="coulomb"),
          g.make_params(Ns=10, Nb=1, U=2.0, dt=0.02, time_steps=50)]
    solo = []
    for p in ps:
        with g.Propagator(p) as prop:
            prop.set_state(g.make_G0(p))
            prop.adiabatic_switching()
            prop.advance(20)
            solo.append(prop.energy())
    b = g.Batch(ps)
    try:
        for i, p in enumerate(ps):
            b.set_state(i, g.make_G0(p))
        b.adiabatic_switching()
        b.advance(20)
        for i in range(len(ps)):
            assert np.array_equal(b.energy(i), solo[i])
    finally:
        b.close()
