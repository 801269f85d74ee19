"""Mixed-precision [H2,G2] contractions (FP32-3M, BF16 hi/lo split) against FP64.

The tile products change precision; G1, the collision integral, Psi/Pi and the
RK state stay FP64.  Bounds: routine-level relative error of the commutator
(1e-5 for FP32-3M, 1e-4 for BF16x3, relative to its largest entry), and the
G1 density error of a correlated N = 32 run against the FP64 run.
"""
import numpy as np
import pytest

from tests.helpers import rand_c

pytestmark = pytest.mark.gpu

g = pytest.importorskip("g12kba_amd")
from oracle import fast  # noqa: E402

BOUND = {"fp32_3m": 1e-5, "bf16x3": 1e-4}


@pytest.mark.parametrize("prec", ["fp32_3m", "bf16x3"])
@pytest.mark.parametrize("Ns,Nb", [(3, 2), (8, 2), (16, 2), (20, 3), (32, 2)])
def test_h2_commutator_mixed(Ns, Nb, prec):
    rng = np.random.default_rng(Ns + 100 * Nb)
    N = Ns * Nb
    G2 = rand_c(rng, N * N, N * N)
    h = rand_c(rng, N, N)
    want = fast.nat_to_ref(fast.comm2(fast.ref_to_nat(G2, Ns, Nb), h, h), Ns, Nb)
    got = g.h2xyG2xyxy_commutator(G2, h, h, Ns, Nb, precision=prec)
    err = np.abs(got - want).max() / np.abs(want).max()
    assert err < BOUND[prec], err
    assert err > 1e-12  # the reduced-precision path really ran


def _run(prec, steps):
    p = g.make_params(Ns=16, Nb=2, t1=1.0, t2=0.5, U=1.5, dt=0.01, time_steps=steps, quench_strength=1.0, Nq=4,
                      q_type="full", quench_rate=2.0, quench_on_midpoint=0.5, quench_gate_time=0.0,
                      AS_rate=0.5, AS_midpoint=3.0, EHM=True, V_law="coulomb", precision=prec)
    with g.Propagator(p) as prop:
        prop.set_state(g.make_G0(p))
        prop.adiabatic_switching()
        rc, traj = prop.rk4()
    assert rc == 0
    return traj


def test_mixed_run_density_error():
    steps = 400
    ref = _run("fp64", steps)
    dens_ref = np.diagonal(ref, axis1=1, axis2=2).imag
    for prec, bound in (("fp32_3m", 1e-6), ("bf16x3", 1e-5)):
        t = _run(prec, steps)
        dens = np.diagonal(t, axis1=1, axis2=2).imag
        err = np.abs(dens - dens_ref).max()
        derr = np.abs(g.dipole(t) - g.dipole(ref)).max()
        print(f"{prec}: max density error {err:.3e}, max dipole error {derr:.3e} over {steps} steps")
        assert err < bound
