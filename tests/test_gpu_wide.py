"""Sizes past the N <= 64 tiles (GPU): the wide-tile kernels (64 < N <= 128,
FP64) and more than eight orbitals per site (read_input.h:85-146 and
Main.cpp:48-50 take any Ns, Nb).

* N = 72 (Ns = 36, Nb = 2): two RK4 iterations of the EHM workload with the
  quench ramp acting, against the factorised oracle at 1e-10 relative (G1
  trajectory, both G2 tensors, every G2 element).
* N = 84 (Ns = 7, Nb = 12, orbital energies on all twelve orbitals): one
  iteration, the same bar; its tiles take two passes over their output blocks.
* N = 128 (Ns = 128, Nb = 1): the routine-level [H2, G2] (commutators.h:53-99)
  of a sum of product states T = sum_r a_r (x) b_r, whose commutator is
  sum_r (hx a_r - a_r hy) (x) b_r + a_r (x) (hx b_r - b_r hy) (the closed form
  is pinned against the oracle on the CPU), to 1e-12 relative.
"""
import numpy as np
import pytest

from tests.helpers import oracle_cfg, physical_g2, rand_c

pytestmark = pytest.mark.gpu

g = pytest.importorskip("g12kba_amd")


def relerr(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def wide_params(Ns, Nb, steps, **kw):
    # the quench ramp rises from t = 0 (gate open), so it acts at every stage
    return g.make_params(Ns=Ns, Nb=Nb, t1=1.0, t2=0.5, U=1.5, dt=0.01, time_steps=steps, quench_strength=1.0,
                         Nq=4, q_type="full", EHM=True, V_law="coulomb", quench_rate=2.0, quench_on_midpoint=0.01,
                         quench_gate_time=0.0, AS_rate=0.5, AS_midpoint=10.0, **kw)


@pytest.mark.parametrize("Ns,Nb,steps,eps", [(36, 2, 2, None), (7, 12, 1, [0.05 * k for k in range(12)])])
def test_wide_run_matches_oracle(Ns, Nb, steps, eps):
    from oracle.oracle import Oracle
    print("wide run", Ns, Nb, steps, flush=True)
    kw = {} if eps is None else dict(eps=eps)
    p = wide_params(Ns, Nb, steps, **kw)
    rng = np.random.default_rng(Ns * 100 + Nb)
    state = (g.make_G0(p), physical_g2(rng, Ns, Nb), physical_g2(rng, Ns, Nb))
    ref = Oracle().fast_run(oracle_cfg(p), *state, V_profile=list(g.v_profile(p)), eps=eps, run_as=False,
                            traj_every=1, traj_cap=steps + 2)
    assert ref["rc"] == 0
    with g.Propagator(p) as prop:
        prop.set_state(*state)
        rc, traj = prop.rk4()
        G1, Gxx, Gxy = prop.get_state()
    assert rc == 0 and len(traj) == ref["rk_steps"] == len(ref["traj"])
    assert relerr(traj, ref["traj"]) < 1e-10 and relerr(G1, ref["G1"]) < 1e-10
    assert relerr(Gxx, ref["Gxx"]) < 1e-10 and relerr(Gxy, ref["Gxy"]) < 1e-10
    # the run moved G2 well past rounding: the check is not of an unchanged state
    assert relerr(Gxy, state[2]) > 1e-4


def test_h2_commutator_n128_product_state():
    Ns, Nb, N = 128, 1, 128
    print("wide commutator", N, flush=True)
    rng = np.random.default_rng(128)
    a = [rand_c(rng, N, N) for _ in range(2)]
    b = [rand_c(rng, N, N) for _ in range(2)]
    hx, hy = rand_c(rng, N, N), rand_c(rng, N, N)
    # reference layout at Nb = 1: pair (p1, p2) -> p1 + p2 N, so the reference
    # array is T[p2][p1][q2][q1] of the natural T[p1][p2][q1][q2]
    outer = lambda x, y: np.einsum("ac,bd->badc", x, y)  # (x (x) y) in the reference layout
    T = outer(a[0], b[0])
    T += outer(a[1], b[1])
    got = g.h2xyG2xyxy_commutator(T.reshape(N * N, N * N), hx, hy, Ns, Nb).reshape(N, N, N, N)
    del T
    want = np.zeros_like(got)
    for ar, br in zip(a, b):
        want += outer(hx @ ar - ar @ hy, br)
        want += outer(ar, hx @ br - br @ hy)
    assert relerr(got, want) < 1e-12


def test_routines_n72_match_factorised_oracle():
    """Psi (both spins), Pi (both call orders) and the collision integral at
    N = 72 through the routine entries, which run the production table and
    epilogue code, against the factorised numpy oracle (pinned against the
    element-wise oracle on the CPU), to 1e-12 relative."""
    from oracle import fast
    Ns, Nb, N = 36, 2, 72
    print("wide routines", N, flush=True)
    rng = np.random.default_rng(72)
    G1 = rand_c(rng, N, N)
    V = rng.uniform(0.1, 1.0, 3)
    U = 1.1
    Wxy, Wxx = fast.w_tables(Ns, Nb, U, V)
    for W, fn in ((Wxx, g.Psi_xxxx), (Wxy, g.Psi_xyxy)):
        want = fast.psi_term(G1, W, fast.q_table(G1, W))
        assert relerr(fast.ref_to_nat(fn(G1, Ns, Nb, U, V), Ns, Nb), want) < 1e-12
        del want
    P, Q = rand_c(rng, N * N, N * N, scale=1e-2), rand_c(rng, N * N, N * N, scale=1e-2)
    DP, DQ = fast.d_table(fast.ref_to_nat(P, Ns, Nb)), fast.d_table(fast.ref_to_nat(Q, Ns, Nb))
    for A, B, DA, DB in ((P, Q, DP, DQ), (Q, P, DQ, DP)):  # both call orders (RK_solver.h:163-164)
        want = fast.pi_term(G1, fast.f_table(DA, DB, Wxy, Wxx))
        assert relerr(fast.ref_to_nat(g.Pi_xyxy(G1, A, B, U, Ns, Nb, V), Ns, Nb), want) < 1e-12
        del want
    F = fast.f_table(DP, DQ, Wxy, Wxx)  # collision_int(xyxy = P, xxxx = Q)
    assert relerr(g.collision_int(P, Q, Ns, Nb, U, V), F[np.arange(N), :, np.arange(N)]) < 1e-12
