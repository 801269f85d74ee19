"""CPU suite: pins the oracle and the host-side boundary.

* oracle vs golden vectors written from the reference's own routine headers
  (tests/golden/make_golden.py), and vs the live harness when it is built;
* textbook / brute-force pins: [h,G] as a matrix commutator, [H2,G2] against a
  dense kron(h,1)+kron(1,h) (pins the pair-index permutation), the factorised
  numpy oracle against the element-wise one;
* the closed-form RK4 iterate at U = 0 (stage coefficients, weights, AS/RK
  chaining), loop-count known answers, invariants;
* plants a mistake in the oracle and checks that the pins catch it.
"""
import glob
import os

import numpy as np
import pytest

from oracle import fast
from oracle.oracle import REF_SO, Oracle, Ref
from tests.helpers import GEOMS, product_state, product_state_commutator, rand_c

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
O = Oracle()
TOL = 1e-12


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def golden_files():
    return sorted(glob.glob(os.path.join(GOLD, "routines_*.npz")))


@pytest.mark.parametrize("path", golden_files(), ids=os.path.basename)
def test_oracle_matches_reference_golden(path):
    z = np.load(path)
    Ns, Nb = int(z["Ns"]), int(z["Nb"])
    V, eps, U = z["V"], z["eps"], float(z["U"])
    G1, P, Q = z["G1"], z["P"], z["Q"]
    assert rel(O.hf(z["h0"], G1, Ns, Nb, U, float(z["t1"]), float(z["t2"]), float(z["w"]), eps, V,
                    int(z["nq"])), z["hf"]) < TOL
    assert rel(O.collision(P, Q, Ns, Nb, U, V), z["collision"]) < TOL
    for ss, key in ((1, "psi_xxxx"), (0, "psi_xyxy")):
        ref = z[key]
        got = O.psi(G1, Ns, Nb, U, V, ss)
        assert (np.abs(got).max() == 0 == np.abs(ref).max()) or rel(got, ref) < TOL
    assert rel(O.pi(G1, P, Q, U, Ns, Nb, V), z["pi_PQ"]) < TOL
    assert rel(O.pi(G1, Q, P, U, Ns, Nb, V), z["pi_QP"]) < TOL
    K, R = O.g1_step(z["c1"], z["I1"], z["K1"], z["R1"], 2.0, Ns, Nb)
    assert rel(K, z["g1_K"]) < TOL and rel(R, z["g1_R"]) < TOL
    K, R = O.g2_step(z["c2"], z["s2"], z["p2"], z["K2"], z["R2"], 2.0, Ns, Nb)
    assert rel(K, z["g2_K"]) < TOL and rel(R, z["g2_R"]) < TOL
    assert np.array_equal(O.dipole(z["traj"]), z["dipole"])


def test_fermi_golden():
    z = np.load(os.path.join(GOLD, "fermi.npz"))
    for row, (t0, tau) in zip(z["f"], ((10, 0.5), (25, 3), (30, 2))):
        assert np.array_equal(row, [O.fermi(t, t0, tau) for t in z["t"]])


@pytest.mark.skipif(not os.path.exists(REF_SO), reason="reference harness not built here")
@pytest.mark.parametrize("Ns,Nb,nV", GEOMS)
def test_oracle_matches_live_reference(Ns, Nb, nV):
    r = Ref()
    rng = np.random.default_rng(77 + Ns * 10 + Nb)
    N = Ns * Nb
    G1, P, Q = rand_c(rng, N, N), rand_c(rng, N * N, N * N), rand_c(rng, N * N, N * N)
    V = rng.uniform(0.1, 1.0, nV)
    assert rel(O.pi(G1, P, Q, 0.7, Ns, Nb, V), r.pi(G1, P, Q, 0.7, Ns, Nb, V)) < TOL
    assert rel(O.collision(P, Q, Ns, Nb, 0.7, V), r.collision(P, Q, Ns, Nb, 0.7, V)) < TOL
    b = r.psi(G1, Ns, Nb, 0.7, V, 0)
    assert rel(O.psi(G1, Ns, Nb, 0.7, V, 0), b) < TOL


@pytest.mark.parametrize("N", [1, 4, 7])
def test_comm1_textbook(N):
    rng = np.random.default_rng(N)
    G, h = rand_c(rng, N, N), rand_c(rng, N, N)
    assert rel(O.comm1(G, h), h @ G - G @ h) < TOL


@pytest.mark.parametrize("Ns,Nb", [(2, 1), (3, 2), (2, 3), (4, 2)])
def test_comm2_kron_bruteforce(Ns, Nb):
    """[H2, M] = (h(x)1 + 1(x)h) M - M (h'(x)1 + 1(x)h') with the natural pair
    index p1*N + p2; the reference layout is related by fast.pair_perm."""
    rng = np.random.default_rng(Ns * 3 + Nb)
    N = Ns * Nb
    M = rand_c(rng, N * N, N * N)
    hx, hy = rand_c(rng, N, N), rand_c(rng, N, N)
    E = np.eye(N)
    Hx, Hy = np.kron(hx, E) + np.kron(E, hx), np.kron(hy, E) + np.kron(E, hy)
    Mn = fast.ref_to_nat(M, Ns, Nb).reshape(N * N, N * N)
    want = fast.nat_to_ref((Hx @ Mn - Mn @ Hy).reshape(N, N, N, N), Ns, Nb)
    assert rel(O.comm2(M, hx, hy, Ns, Nb), want) < TOL


@pytest.mark.parametrize("Ns,Nb", [(3, 2), (2, 5)])
def test_comm2_product_state_closed_form(Ns, Nb):
    """The closed form the N = 128 GPU check uses: for T = sum_r a_r (x) b_r,
    [H2, T] = sum_r (hx a_r - a_r hy) (x) b_r + a_r (x) (hx b_r - b_r hy);
    against the oracle's commutator, and failing with hy in place of hx on the
    left (a transposed-operand mistake)."""
    rng = np.random.default_rng(Ns * 7 + Nb)
    N = Ns * Nb
    T, a, b = product_state(rng, N)
    hx, hy = rand_c(rng, N, N), rand_c(rng, N, N)
    got = fast.ref_to_nat(O.comm2(fast.nat_to_ref(T, Ns, Nb), hx, hy, Ns, Nb), Ns, Nb)
    assert rel(got, product_state_commutator(a, b, hx, hy)) < TOL
    assert rel(got, product_state_commutator(a, b, hy, hy)) > 0.1


@pytest.mark.parametrize("Ns,Nb,nV", GEOMS[:4])
def test_factorised_oracle_matches_elementwise(Ns, Nb, nV):
    rng = np.random.default_rng(9 + Ns + Nb)
    N = Ns * Nb
    G1 = rand_c(rng, N, N)
    Mxx, Mxy = rand_c(rng, N * N, N * N, scale=0.1), rand_c(rng, N * N, N * N, scale=0.1)
    V, eps = rng.uniform(0.1, 1.0, nV), np.zeros(Nb)
    U, t1, t2, w, nq = 1.3, 1.0, 0.4, 0.3, 1
    h = O.hf(np.zeros((N, N)), G1, Ns, Nb, U, t1, t2, w, eps, V, nq)
    assert rel(fast.hf(G1, Ns, Nb, U, t1, t2, w, eps, V, nq), h) < TOL
    I = O.collision(Mxy, Mxx, Ns, Nb, U, V)
    K1, _ = O.g1_step(O.comm1(G1, h), I, np.zeros((N, N)), np.zeros((N, N)), 1.0, Ns, Nb)
    Ks = []
    for M, P, Q, ss in ((Mxx, Mxy, Mxx, 1), (Mxy, Mxx, Mxy, 0)):
        K, _ = O.g2_step(O.comm2(M, h, h, Ns, Nb), O.psi(G1, Ns, Nb, U, V, ss), O.pi(G1, P, Q, U, Ns, Nb, V),
                         np.zeros_like(M), np.zeros_like(M), 1.0, Ns, Nb)
        Ks.append(K)
    fK1, fKxx, fKxy, _ = fast.derivative(G1, fast.ref_to_nat(Mxx, Ns, Nb), fast.ref_to_nat(Mxy, Ns, Nb), Ns, Nb,
                                         U, t1, t2, w, eps, V, nq)
    assert rel(fK1, K1) < TOL
    assert rel(fast.nat_to_ref(fKxx, Ns, Nb), Ks[0]) < TOL
    assert rel(fast.nat_to_ref(fKxy, Ns, Nb), Ks[1]) < TOL


def _cfg(**kw):
    base = dict(Ns=3, Nb=2, t1=1.0, t2=0.5, U=0.0, dt=0.02, time_steps=40, quench_strength=0.0, Nq=0,
                q_type="full", HF=False, AS_rate=0.5, AS_midpoint=10.0, quench_rate=2.0, quench_on_midpoint=25.0,
                quench_off_midpoint=60.0)
    base.update(kw)
    return base


def _h0(Ns, Nb, t1, t2):
    return fast.hf(np.zeros((Ns * Nb, Ns * Nb)), Ns, Nb, 0.0, t1, t2, 0.0, None, [], 0)


def test_rk4_closed_form_at_U0():
    """At U = 0 and no quench, G1_pq(n) = R(-i dt (e_p - e_q))^n G1_pq(0) in the
    eigenbasis of h, R(z) = 1 + z + z^2/2 + z^3/6 + z^4/24; AS (dt = 0.1, 21
    steps for U = 0) then RK4 (dt_fixed)."""
    cfg = _cfg()
    N = 6
    rng = np.random.default_rng(1)
    G0 = 1j * np.diag(rng.uniform(0, 1, N)) + 0.2 * (lambda a: a - a.conj().T)(rand_c(rng, N, N))
    Z = np.zeros((N * N, N * N), complex)
    G1, Gxx, Gxy, nas = O.adiabatic_switching(cfg, G0, Z, Z)
    assert nas == 21
    rc, G1, _, _, traj = O.rk4(cfg, G1, Gxx, Gxy)
    assert rc == 0 and len(traj) == 40
    e, U = np.linalg.eigh(_h0(3, 2, 1.0, 0.5))

    def R(z):
        return 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24

    Gt = U.conj().T @ G0 @ U
    de = e[:, None] - e[None, :]
    want = Gt * R(-1j * 0.1 * de) ** 21 * R(-1j * 0.02 * de) ** 40
    assert np.abs(U @ want @ U.conj().T - G1).max() < 1e-12
    # every recorded step too
    for n in (0, 17, 39):
        w = Gt * R(-1j * 0.1 * de) ** 21 * R(-1j * 0.02 * de) ** n
        assert np.abs(U @ w @ U.conj().T - traj[n]).max() < 1e-12


def test_loop_counts():
    """Iteration counts of the AS and RK loops (AS.h:96-100, RK_solver.h:96,216)."""
    N = 1
    Z = np.zeros((1, 1), complex)
    G = np.array([[0.5j]])
    for (rate, mid, U), want in (((0.5, 10.0, 1.0), 304), ((3.0, 25.0, 0.5), 1373), ((0.5, 10.0, 0.0), 21)):
        cfg = _cfg(Ns=1, Nb=1, U=U, AS_rate=rate, AS_midpoint=mid)
        assert O.adiabatic_switching(cfg, G, Z, Z)[3] == want
    for (ts, dt), want in (((3000, 0.02), 3000), ((10000, 0.01), 10000), ((5000, 0.01), 5001),
                           ((4000, 0.02), 4000), ((20000, 0.01), 20001)):
        cfg = _cfg(Ns=1, Nb=1, U=0.0, time_steps=ts, dt=dt)
        rc, _, _, _, traj = O.rk4(cfg, G, Z, Z, traj_cap=1)
        n = O.L  # noqa: F841
        import ctypes
        from oracle.oracle import OrcParams, _ptr
        p, keep = O.params(cfg)
        g1 = G.copy()
        cnt = ctypes.c_int64(0)
        O.L.orc_rk4(ctypes.byref(p), _ptr(g1), _ptr(Z.copy()), _ptr(Z.copy()), None, 0, ctypes.byref(cnt))
        assert cnt.value == want, (ts, dt, cnt.value)


def test_invariants_correlated_run():
    """Trace of G1 is conserved to rounding and G1 stays anti-Hermitian."""
    cfg = _cfg(U=1.2, time_steps=30, quench_strength=1.0, Nq=1, quench_on_midpoint=0.2, AS_tail=2,
               AS_midpoint=0.2, gate_time=0.1)
    N = 6
    rng = np.random.default_rng(4)
    X = rand_c(rng, N, N)
    _, v = np.linalg.eigh(X + X.conj().T)
    G0 = 1j * v[:, :3] @ v[:, :3].conj().T
    Z = np.zeros((N * N, N * N), complex)
    G1, Gxx, Gxy, _ = O.adiabatic_switching(cfg, G0, Z, Z, V_profile=[0.5, 0.3])
    rc, G1, Gxx, Gxy, traj = O.rk4(cfg, G1, Gxx, Gxy, V_profile=[0.5, 0.3])
    assert rc == 0 and np.abs(Gxx).max() > 1e-6
    tr = np.array([np.trace(t) for t in traj])
    assert np.abs(tr - tr[0]).max() < 1e-13
    assert max(np.abs(t + t.conj().T).max() for t in traj) < 1e-13


def test_planted_mistakes_are_caught():
    """A dropped term, a sign flip or a transposed operand in the oracle's
    formulas fails the golden pins (mistakes are planted on CPU copies)."""
    z = np.load(golden_files()[-1])
    Ns, Nb = int(z["Ns"]), int(z["Nb"])
    good = O.pi(z["G1"], z["P"], z["Q"], float(z["U"]), Ns, Nb, z["V"])
    assert rel(good, z["pi_PQ"]) < TOL
    # transposed G2 pair (the argument swap of RK_solver.h:163-164) is detected
    assert rel(O.pi(z["G1"], z["Q"], z["P"], float(z["U"]), Ns, Nb, z["V"]), z["pi_PQ"]) > 1e-3
    # transposed G1 operand in Psi is detected
    assert rel(O.psi(z["G1"].T, Ns, Nb, float(z["U"]), z["V"], 0), z["psi_xyxy"]) > 1e-3
    # a dropped long-range shell is detected
    assert rel(O.collision(z["P"], z["Q"], Ns, Nb, float(z["U"]), z["V"][:-1]), z["collision"]) > 1e-3
    # a sign flip of the onsite Fock term is detected
    assert rel(O.hf(z["h0"], z["G1"], Ns, Nb, -float(z["U"]), float(z["t1"]), float(z["t2"]), float(z["w"]),
                    z["eps"], z["V"], int(z["nq"])), z["hf"]) > 1e-3


def _physical_run_state(Ns, Nb, U, V, steps=20, dt=0.05, seed=1):
    """G1 (anti-Hermitian, a filled random band) and both G2 tensors after
    `steps` Euler steps of the oracle's derivative from G2 = 0: a state with
    the symmetries every run keeps (G2 Hermitian, pair-symmetric)."""
    N = Ns * Nb
    rng = np.random.default_rng(seed)
    H = rand_c(rng, N, N)
    H = H + H.conj().T
    _, v = np.linalg.eigh(H)
    G1 = 1j * (v[:, : N // 2] @ v[:, : N // 2].conj().T)
    Txx = np.zeros((N,) * 4, complex)
    Txy = np.zeros_like(Txx)
    for _ in range(steps):
        K1, Kxx, Kxy, _h = fast.derivative(G1, Txx, Txy, Ns, Nb, U, 1.0, 0.4, 0.3, None, V, 1)
        G1, Txx, Txy = G1 + dt * K1, Txx + dt * Kxx, Txy + dt * Kxy
    return G1, Txx, Txy


def _pair_symmetric_rhs(G1, T, F, W, h, half=0.5):
    """The pair-symmetric path's right-hand side of one G2 tensor: X + P(X)
    with X the contiguous-tile (p2, q2) commutator terms plus
    g22 (ea[p2] - eb[q2]), e = i Qp[p1][q1][.] + F[p1][q1][.],
    ea = e + half g11 W[p1][.], eb = e + half g11 W[q1][.]."""
    Qp = fast.q_table(G1, W).transpose(1, 2, 0)
    X = np.einsum("bk,akcd->abcd", h, T) - np.einsum("abck,kd->abcd", T, h)
    e = 1j * Qp + F
    ea = e + half * G1[:, :, None] * W[:, None, :]
    eb = e + half * G1[:, :, None] * W.T[None, :, :]
    # X[p1,p2,q1,q2] += g[p2,q2] (ea[p1,q1,p2] - eb[p1,q1,q2])
    X = X + G1[None, :, None, :] * (ea.transpose(0, 2, 1)[:, :, :, None] - eb[:, None, :, :])
    return X + X.transpose(1, 0, 3, 2)


@pytest.mark.parametrize("Ns,Nb,U,V", [(3, 2, 1.5, [0.75, 0.5]), (4, 1, 2.0, []), (2, 3, 1.0, [0.5])])
def test_pair_symmetric_decomposition(Ns, Nb, U, V):
    """The mathematics of the GPU's pair-symmetric path: for a state with the
    run symmetries, the full G2 derivative (commutators.h:53-99, psi.h,
    pi.h, G1_G2_t_step.h:55-86 as the oracle states them) equals
    -i (X + P(X)) built from the contiguous-tile half alone.  Dropping the
    half weight of the W term, or a state without the symmetry, fails it."""
    G1, Txx, Txy = _physical_run_state(Ns, Nb, U, V)
    for T in (Txx, Txy):
        s = np.abs(T).max()
        assert s > 1e-4
        assert np.abs(T - T.transpose(1, 0, 3, 2)).max() < 1e-14 * s * 100
    K1, Kxx, Kxy, h = fast.derivative(G1, Txx, Txy, Ns, Nb, U, 1.0, 0.4, 0.3, None, V, 1)
    Wxy, Wxx = fast.w_tables(Ns, Nb, U, V)
    Dxx, Dxy = fast.d_table(Txx), fast.d_table(Txy)
    Fxx = fast.f_table(Dxy, Dxx, Wxy, Wxx)
    Fxy = fast.f_table(Dxx, Dxy, Wxy, Wxx)
    for T, F, W, K in ((Txx, Fxx, Wxx, Kxx), (Txy, Fxy, Wxy, Kxy)):
        assert np.abs(-1j * _pair_symmetric_rhs(G1, T, F, W, h) - K).max() < 1e-13 * np.abs(K).max()
        # planted mistake: the whole W term in X counts it twice (W_xx = 0 on
        # one site per orbital without long-range shells)
        if np.abs(W).max() > 0:
            assert np.abs(-1j * _pair_symmetric_rhs(G1, T, F, W, h, half=1.0) - K).max() > 1e-3 * np.abs(K).max()
    # a state without the exchange symmetry: the decomposition does not hold
    rng = np.random.default_rng(9)
    N = Ns * Nb
    A = rand_c(rng, N * N, N * N, scale=1e-2)
    Tn = (A + A.conj().T).reshape(N, N, N, N)
    _, Kn, _, _ = fast.derivative(G1, Tn, Txy, Ns, Nb, U, 1.0, 0.4, 0.3, None, V, 1)
    Fn = fast.f_table(fast.d_table(Txy), fast.d_table(Tn), Wxy, Wxx)
    assert np.abs(-1j * _pair_symmetric_rhs(G1, Tn, Fn, Wxx, h) - Kn).max() > 1e-3 * np.abs(Kn).max()


def test_accumulator_free_rk4_protocol():
    """The FP64 pair-symmetric update's RK4 bookkeeping: no accumulator; stage 0
    stores Y1 = S + (dt/2) K1, stage 1 stores Y2 = S + (dt/2) K2 apart, stage 2
    recovers K1 = (Y1 - S)/(dt/2) and K2 = (Y2 - S)/(dt/2), folds
    S' = S + dt (K1 + 2 K2 + 2 K3)/6 and stores Y3 = S + dt K3, and stage 3
    finishes S' + dt K4 / 6.  Over 200 steps of a stiff-ish linear system it
    equals textbook RK4 (RK_solver.h:120-214: c = 0, 1/2, 1/2, 1; weights 1, 2,
    2, 1) to a few ulp of the state; recovering K2 with the stage-0 coefficient
    dt instead of dt/2 (a plausible slip) does not."""
    rng = np.random.default_rng(3)
    n = 12
    H = rand_c(rng, n, n)
    H = H + H.conj().T
    f = lambda y: -1j * (H @ y)  # noqa: E731
    dt = 0.02

    def textbook(y):
        k1 = f(y)
        k2 = f(y + dt / 2 * k1)
        k3 = f(y + dt / 2 * k2)
        k4 = f(y + dt * k3)
        return y + dt / 6 * (k1 + 2 * k2 + 2 * k3 + k4)

    def protocol(y, c2=dt / 2):
        y1 = y + dt / 2 * f(y)        # stage 0: Y1 (no accumulator)
        y2 = y + dt / 2 * f(y1)       # stage 1: Y2 kept apart
        k1 = (y1 - y) / (dt / 2)      # stage 2: recovered terms
        k2 = (y2 - y) / c2
        k3 = f(y2)
        y3 = y + dt * k3
        s = y + dt * (k1 + 2 * k2 + 2 * k3) / 6
        return s + dt * f(y3) / 6     # stage 3

    y0 = rand_c(rng, n)
    a = b = c = y0
    for _ in range(200):
        a, b, c = textbook(a), protocol(b), protocol(c, c2=dt)
    scale = np.abs(a).max()
    assert np.abs(b - a).max() < 1e-12 * scale
    assert np.abs(c - a).max() > 1e-3 * scale
