"""Energy functional (no reference counterpart) on the CPU: the oracle's
formula pinned by a closed form, by Fock-space expectation values of Slater
determinants and by conservation along the oracle's own dynamics; the C-ABI
host entry against the oracle."""
import itertools

import numpy as np
import pytest

from oracle import fast

g = pytest.importorskip("g12kba_amd")


def chain_h0(Ns, Nb, t1, t2):
    return fast.hf(np.zeros((Ns * Nb,) * 2, complex), Ns, Nb, 0.0, t1, t2, 0.0, None, [], 0)


def test_one_body_energy_closed_form():
    """U = 0: a filled band of h0 has E = 2 x (sum of the occupied eigenvalues)."""
    Ns, Nb, t1, t2 = 7, 2, 1.0, 0.4
    ev, W = np.linalg.eigh(chain_h0(Ns, Nb, t1, t2))
    for nocc in (1, 5, 9):
        rho = W[:, :nocc] @ W[:, :nocc].conj().T
        E = fast.energy(1j * rho, None, None, Ns, Nb, 0.0, t1, t2, 0.0, None, [], 0)
        assert abs(E[0] - 2 * ev[:nocc].sum()) < 1e-12 and np.all(E[1:] == 0)


def fermion_ops(M):
    """Jordan-Wigner annihilators c_k on 2^M states (mode k = bit k)."""
    dim = 1 << M
    ops = []
    for k in range(M):
        c = np.zeros((dim, dim))
        for s in range(dim):
            if s >> k & 1:
                sign = (-1) ** bin(s & ((1 << k) - 1)).count("1")
                c[s ^ (1 << k), s] = sign
        ops.append(c)
    return ops


@pytest.mark.parametrize("Ns,Nb,nV,nocc", [(2, 2, 0, 2), (3, 1, 2, 1), (4, 1, 3, 2), (2, 2, 1, 1)])
def test_hartree_exchange_match_fock_space(Ns, Nb, nV, nocc):
    """<Phi|H|Phi> of a spin-symmetric Slater determinant with complex orbitals,
    H = sum_s h0 c+c + 1/2 sum_{s s'} W^{ss'}(X,Y) c+_Xs c+_Ys' c_Ys' c_Xs with
    W^{ss} = W_xx and W^{s,-s} = W_xy, equals one-body + Hartree + exchange."""
    rng = np.random.default_rng(Ns * 10 + Nb + nV)
    N, U, t1, t2 = Ns * Nb, 1.3, 1.0, 0.6
    V = list(U * rng.uniform(0.2, 0.8, nV))
    A = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    phi = np.linalg.qr(A)[0][:, :nocc]  # orthonormal complex orbitals
    rho = phi @ phi.conj().T  # rho[P][Q] = <c+_Q c_P>
    E = fast.energy(1j * rho, None, None, Ns, Nb, U, t1, t2, 0.0, None, V, 0)
    h0 = chain_h0(Ns, Nb, t1, t2)
    Wxy, Wxx = fast.w_tables(Ns, Nb, U, V)
    c = fermion_ops(2 * N)  # mode X + s*N
    cd = [m.T for m in c]
    H = np.zeros_like(c[0], dtype=complex)
    for s in range(2):
        for X, Y in itertools.product(range(N), repeat=2):
            if h0[X, Y] != 0:
                H += h0[X, Y] * cd[X + s * N] @ c[Y + s * N]
    for s, s2 in itertools.product(range(2), repeat=2):
        Wm = Wxx if s == s2 else Wxy
        for X, Y in itertools.product(range(N), repeat=2):
            if Wm[X, Y] != 0:
                a, b = X + s * N, Y + s2 * N
                H += 0.5 * Wm[X, Y] * cd[a] @ cd[b] @ c[b] @ c[a]
    vac = np.zeros(c[0].shape[0], complex)
    vac[0] = 1.0
    state = vac
    for s in range(2):
        for k in range(nocc):
            state = sum(phi[X, k] * (cd[X + s * N] @ state) for X in range(N))
    assert abs(np.vdot(state, state) - 1) < 1e-12
    assert abs(np.vdot(state, H @ state).real - E[:3].sum()) < 1e-12
    assert E[3] == 0


def rk4_run(G1, Txx, Txy, args, dt, n):
    out = []
    for _ in range(n):
        out.append(fast.energy(G1, Txx, Txy, *args))
        k1 = fast.derivative(G1, Txx, Txy, *args)[:3]
        s2 = [x + dt / 2 * k for x, k in zip((G1, Txx, Txy), k1)]
        k2 = fast.derivative(*s2, *args)[:3]
        s3 = [x + dt / 2 * k for x, k in zip((G1, Txx, Txy), k2)]
        k3 = fast.derivative(*s3, *args)[:3]
        s4 = [x + dt * k for x, k in zip((G1, Txx, Txy), k3)]
        k4 = fast.derivative(*s4, *args)[:3]
        G1, Txx, Txy = [x + dt / 6 * (a + 2 * b + 2 * cc + d) for x, a, b, cc, d in zip((G1, Txx, Txy), k1, k2, k3, k4)]
    return np.array(out)


def test_total_energy_conserved_by_the_dynamics():
    """Nb = 1, constant h0 (no quench), EHM: from a non-stationary Slater
    determinant with G2 = 0 the correlation energy builds up by O(0.1) while
    the total stays fixed to the integrator's error, which vanishes with dt
    (halving dt divides the drift by about 10 here, between third and fourth
    order over this window).  A correlation weight other than -1, or a
    dropped source term, leaves a drift that does not shrink with dt."""
    Ns, Nb, U = 5, 1, 1.5
    N = Ns * Nb
    V = [U / (1 + n) for n in range(1, Ns)]
    rng = np.random.default_rng(7)
    A = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    phi = np.linalg.qr(A)[0][:, :2]
    G1 = 1j * phi @ phi.conj().T
    Z = np.zeros((N,) * 4, complex)
    args = (Ns, Nb, U, 1.0, 0.5, 0.0, None, V, 0)
    drift = []
    for dt, n in ((0.04, 50), (0.02, 100)):
        E = rk4_run(G1, Z, Z, args, dt, n)
        tot = E.sum(1)
        assert np.abs(E[:, 3] - E[0, 3]).max() > 0.05
        drift.append(np.abs(tot - tot[0]).max())
        # the total is conserved, not the parts: with weight -1/2 on the
        # correlation term the drift would be of the order of its change
        half = E[:, :3].sum(1) + 0.5 * E[:, 3]
        assert np.abs(half - half[0]).max() > 1e3 * drift[-1]
    assert drift[1] < 1e-6
    assert drift[0] / drift[1] > 7


def test_c_abi_energy_matches_oracle():
    """g12_energy (host code, reference layouts) against the oracle on random
    states, with switching factor lam and a quench shift."""
    rng = np.random.default_rng(3)
    for Ns, Nb, ehm in ((4, 2, True), (5, 1, False), (3, 3, True)):
        N = Ns * Nb
        p = g.make_params(Ns=Ns, Nb=Nb, U=1.7, t1=0.9, t2=0.35, Nq=2, quench_strength=0.8, EHM=ehm,
                          V_law="coulomb", eps=[0.1 * k for k in range(Nb)])
        G1 = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
        Gxx = rng.standard_normal((N * N, N * N)) + 1j * rng.standard_normal((N * N, N * N))
        Gxy = rng.standard_normal((N * N, N * N)) + 1j * rng.standard_normal((N * N, N * N))
        lam, w = 0.6, 0.45
        V = list(lam * p.U * g.v_profile(p))
        ref = fast.energy(G1, fast.ref_to_nat(Gxx, Ns, Nb), fast.ref_to_nat(Gxy, Ns, Nb), Ns, Nb, lam * p.U, p.t1,
                          p.t2, w, [p.eps[k] for k in range(Nb)], V, p.Nq)
        E = g.energy(p, G1, Gxx, Gxy, lam=lam, w=w)
        assert np.allclose(E, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
        E0 = g.energy(p, G1, lam=lam, w=w)
        assert np.allclose(E0[:3], ref[:3], rtol=1e-12, atol=1e-12 * np.abs(ref).max()) and E0[3] == 0


def test_transposed_onsite_fock_breaks_conservation_for_two_orbitals(monkeypatch):
    """Nb = 2: the reference's onsite inter-orbital Fock term takes g[Z][X]
    instead of g[X][Z] (HF.h:83, reproduced by the drop-in).  Along its
    dynamics the total energy drifts at O(0.01); with the true exchange
    term it is conserved like the Nb = 1 case."""
    Ns, Nb, U = 3, 2, 1.5
    N = Ns * Nb
    rng = np.random.default_rng(11)
    A = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    phi = np.linalg.qr(A)[0][:, :3]
    G1 = 1j * phi @ phi.conj().T
    Z = np.zeros((N,) * 4, complex)
    args = (Ns, Nb, U, 1.0, 0.5, 0.0, None, [], 0)
    E = rk4_run(G1, Z, Z, args, 0.02, 150).sum(1)
    drift_ref = np.abs(E - E[0]).max()
    hf_ref = fast.hf

    def hf_exchange(G, Ns_, Nb_, U_, *rest):
        h = hf_ref(G, Ns_, Nb_, U_, *rest)
        site = np.arange(Ns_ * Nb_) % Ns_
        m = (site[:, None] == site[None, :]) & ~np.eye(Ns_ * Nb_, dtype=bool)
        h[m] += 1j * U_ * (G[m] - G.T[m])
        return h

    monkeypatch.setattr(fast, "hf", hf_exchange)
    E = rk4_run(G1, Z, Z, args, 0.02, 150).sum(1)
    drift_x = np.abs(E - E[0]).max()
    assert drift_ref > 1e-3 and drift_x < 1e-6
