"""Factorised FP64 CPU oracle in numpy (test infrastructure only).

Same mathematics as oracle/g12_oracle.cpp, evaluated through the small
partial-trace tables so that long runs at N = 32 finish on a CPU:

  D^s[X,Y,Z]  = M^s[(X,Z);(Y,Z)]                                   (N^3)
  F^s[X,Y,Z]  = -i sum_Z' ( W_xy(Z,Z') D^P[X,Y,Z'] + W_xx(Z,Z') D^Q[X,Y,Z'] )
                with (P,Q) = (xy,xx) for s = xx and (xx,xy) for s = xy
                (Pi_xyxy argument swap, G12KBA-GW/src/headers/RK_solver.h:163-164)
  I[X,Y]      = F^xx[X,Y,X]           (collision_integral.h:27-69)
  Q^s[Z,X,Y]  = sum_Z' g(X,Z') W^s(Z,Z') g(Z',Y)
  Pi^s[A;B]   = g(p2,q2) (F^s[p1,q1,p2] - F^s[p1,q1,q2])            (pi.h:25-95)
  Psi^s[A;B]  = g(p1,q1) g(p2,q2) (W^s(p1,p2) - W^s(q1,q2))
              + i g(p1,q1) (Q^s[p1,p2,q2] - Q^s[q1,p2,q2])
              + i g(p2,q2) (Q^s[p2,p1,q1] - Q^s[q2,p1,q1])           (psi.h:26-277)
  K^s         = -i ([H2, M^s] + Psi^s + Pi^s - conj(Pi^s[B^x;A^x]))  (G1_G2_t_step.h:55-86)
  K1          = -i ([h, G1] + I + I^dagger)                           (G1_G2_t_step.h:24-47)

G2 is held in the natural layout T[p1,p2,q1,q2] = M[(p1,p2);(q1,q2)] with the
one-particle index p = site + orb*Ns; ``ref_to_nat``/``nat_to_ref`` convert
from/to the reference layout (commutators.h:59-60).
"""
import math

import numpy as np


def pair_perm(Ns, Nb):
    """perm[p1*N+p2] = reference pair index of (p1, p2)."""
    N = Ns * Nb
    p = np.arange(N)
    site, orb = p % Ns, p // Ns
    i, j = site[:, None], site[None, :]
    a, b = orb[:, None], orb[None, :]
    return (i + j * Ns + (a + b * Nb) * Ns * Ns).reshape(-1)


def ref_to_nat(M, Ns, Nb):
    N = Ns * Nb
    pp = pair_perm(Ns, Nb)
    M = np.asarray(M).reshape(N * N, N * N)
    return M[np.ix_(pp, pp)].reshape(N, N, N, N)


def nat_to_ref(T, Ns, Nb):
    N = Ns * Nb
    pp = pair_perm(Ns, Nb)
    out = np.empty((N * N, N * N), dtype=T.dtype)
    out[np.ix_(pp, pp)] = T.reshape(N * N, N * N)
    return out


def w_tables(Ns, Nb, U, V):
    """(W_xy, W_xx) as N x N real matrices."""
    N = Ns * Nb
    s = np.arange(N) % Ns
    d = np.abs(s[:, None] - s[None, :])
    Wxy = np.zeros((N, N))
    Wxy[d == 0] = U
    for n in range(1, len(V) + 1):
        Wxy[d == n] = V[n - 1]
    Wxx = Wxy.copy()
    np.fill_diagonal(Wxx, 0.0)
    return Wxy, Wxx


def hf(G1, Ns, Nb, U, t1, t2, w, eps, V, quench_extent):
    """h_HFx_quench on a zero h (HF.h:57-104)."""
    N = Ns * Nb
    h = np.zeros((N, N), np.complex128)
    site = np.arange(N) % Ns
    orb = np.arange(N) // Ns
    diag = np.array([eps[o] if eps is not None else 0.0 for o in orb], dtype=np.complex128)
    diag = diag + np.where(site < quench_extent, w, 0.0)
    h[np.arange(N), np.arange(N)] += diag
    for X in range(N):
        i, a = site[X], orb[X]
        if i + 1 < Ns:
            Y = (i + 1) + a * Ns
            h[Y, X] -= t1
            h[X, Y] -= t1
        if a + 1 < Nb:
            Y = i + (a + 1) * Ns
            h[X, Y] -= t2
            h[Y, X] -= t2
    g = G1
    dg = np.diag(g)
    Wxy, _ = w_tables(Ns, Nb, U, V)
    same = site[:, None] == site[None, :]
    # onsite same-orbital Hartree
    h[np.arange(N), np.arange(N)] += -1j * U * dg
    # onsite other-orbital Hartree (x2) and long-range Hartree (x2)
    Wh = Wxy.copy()
    np.fill_diagonal(Wh, 0.0)
    h[np.arange(N), np.arange(N)] += -2j * (Wh @ dg)
    # onsite inter-orbital Fock: h[X,Z] += iU g[Z,X]  (transposed)
    m = same & ~np.eye(N, dtype=bool)
    h[m] += 1j * U * g.T[m]
    # long-range Fock: h[P,Q] += iV g[P,Q]
    lr = ~same
    h[lr] += 1j * Wxy[lr] * g[lr]
    return h


def comm2(T, hx, hy):
    """[H2, M] in the natural layout with row Hamiltonian hx, column hy.

    Each of the four terms is one (batched) matrix product on a reshaped view:
    hx on p1, hx on p2, hy on q1 (from the right, transposed), hy on q2."""
    N = T.shape[0]
    out = (hx @ T.reshape(N, N ** 3)).reshape(T.shape)
    out += np.matmul(hx, T.reshape(N, N, N * N)).reshape(T.shape)
    out -= (T.reshape(N ** 3, N) @ hy).reshape(T.shape)
    out -= np.matmul(hy.T, T.reshape(N * N, N, N)).reshape(T.shape)
    return out


def d_table(T):
    return np.einsum("xzyz->xyz", T)


def f_table(DP, DQ, Wxy, Wxx):
    return -1j * (DP @ Wxy.T + DQ @ Wxx.T)


def q_table(g, W):
    # Q[Z,X,Y] = sum_w g[X,w] W[Z,w] g[w,Y]
    return np.einsum("xw,zw,wy->zxy", g, W, g, optimize=True)


def pi_term(g, F):
    N = g.shape[0]
    # Pi[p1,p2,q1,q2] = g[p2,q2] * (F[p1,q1,p2] - F[p1,q1,q2])
    Fp2 = F.transpose(0, 2, 1)[:, :, :, None]  # [p1, p2, q1, 1] = F[p1,q1,p2]
    Fq2 = F[:, None, :, :]  # [p1, 1, q1, q2]
    return g[None, :, None, :] * (Fp2 - Fq2)


def psi_term(g, W, Q):
    quad = (g[:, None, :, None] * g[None, :, None, :]) * (W[:, :, None, None] - W[None, None, :, :])
    # i g[p1,q1] (Q[p1,p2,q2] - Q[q1,p2,q2])
    c1 = Q[:, :, None, :] - Q.transpose(1, 0, 2)[None, :, :, :]
    # i g[p2,q2] (Q[p2,p1,q1] - Q[q2,p1,q1])
    c2 = Q.transpose(1, 0, 2)[:, :, :, None] - Q.transpose(1, 2, 0)[:, None, :, :]
    return quad + 1j * g[:, None, :, None] * c1 + 1j * g[None, :, None, :] * c2


def derivative(G1, Txx, Txy, Ns, Nb, U, t1, t2, w, eps, V, quench_extent, corr=True):
    """Stage derivative (K1, Kxx, Kxy) for the stage input (G1, Txx, Txy)."""
    h = hf(G1, Ns, Nb, U, t1, t2, w, eps, V, quench_extent)
    K1 = h @ G1 - G1 @ h
    if not corr:
        return -1j * K1, np.zeros_like(Txx), np.zeros_like(Txy), h
    Wxy, Wxx = w_tables(Ns, Nb, U, V)
    Dxx, Dxy = d_table(Txx), d_table(Txy)
    Fxx = f_table(Dxy, Dxx, Wxy, Wxx)
    Fxy = f_table(Dxx, Dxy, Wxy, Wxx)
    N = Ns * Nb
    I = Fxx[np.arange(N), :, np.arange(N)]  # I[X,Y] = Fxx[X,Y,X]
    K1 = -1j * (K1 + I + I.conj().T)
    out = []
    for T, F, W in ((Txx, Fxx, Wxx), (Txy, Fxy, Wxy)):
        Q = q_table(G1, W)
        P = pi_term(G1, F)
        K = comm2(T, h, h) + psi_term(G1, W, Q) + P - P.transpose(3, 2, 1, 0).conj()
        out.append(-1j * K)
    return K1, out[0], out[1], h


def energy(G1, Txx, Txy, Ns, Nb, U, t1, t2, w, eps, V, quench_extent):
    """(one-body, Hartree, exchange, correlation) of a state, both spins, with
    rho = -i G1 and V the absolute V_n.  No reference counterpart (README.md:61-63
    lists energies as future work).  Pinned in tests/test_cpu_energy.py: the
    one-body part by the filled-band closed form, Hartree + exchange by
    <Phi|H|Phi> of Slater determinants in Fock space, and the correlation
    weight -1 by conservation of the total along the derivative above."""
    N = Ns * Nb
    h0 = hf(np.zeros((N, N), np.complex128), Ns, Nb, 0.0, t1, t2, w, eps, [0.0] * len(V), quench_extent)
    rho = -1j * np.asarray(G1)
    Wxy, Wxx = w_tables(Ns, Nb, U, V)
    d = np.diag(rho)
    e1 = 2 * np.trace(h0 @ rho).real
    eh = (d @ (Wxx + Wxy) @ d).real
    ef = -np.sum(Wxx * np.abs(rho) ** 2)
    ec = 0.0
    if Txx is not None:
        pd = lambda T: np.einsum("xyxy->xy", np.asarray(T).reshape(N, N, N, N))
        ec = -np.sum(Wxx * pd(Txx) + Wxy * pd(Txy)).real
    return np.array([e1, eh, ef, ec])


def fermi(t, t0, tau):
    return 1.0 / (1.0 + math.exp((t0 - t) / tau))


def rk4_step(cfg, G1, Txx, Txy, dt_full, stage_U, stage_q, V_profile, eps):
    Ns, Nb = cfg["Ns"], cfg["Nb"]
    corr = cfg["U"] != 0 and not cfg["HF"]
    A1 = np.zeros_like(G1)
    Axx = np.zeros_like(Txx)
    Axy = np.zeros_like(Txy)
    K1 = np.zeros_like(G1)
    Kxx = np.zeros_like(Txx)
    Kxy = np.zeros_like(Txy)
    dtk = dt_full
    for k in range(4):
        dtk = dt_full / 2 if k in (1, 2) else dt_full
        wk = 2.0 if k in (1, 2) else 1.0
        Y1 = G1 + K1 * dtk
        Yxx = Txx + Kxx * dtk
        Yxy = Txy + Kxy * dtk
        Ue = stage_U(k, dtk)
        V = [Ue * v for v in V_profile]
        w, nq = stage_q(k, dtk)
        K1, Kxx2, Kxy2, _ = derivative(Y1, Yxx, Yxy, Ns, Nb, Ue, cfg["t1"], cfg["t2"], w, eps,
                                       V, nq, corr)
        if corr:
            Kxx, Kxy = Kxx2, Kxy2
        A1 += wk * K1
        if corr:
            Axx += wk * Kxx
            Axy += wk * Kxy
    return G1 + A1 * dtk / 6.0, Txx + Axx * dtk / 6.0, Txy + Axy * dtk / 6.0


def adiabatic_switching(cfg, G1, Txx, Txy, V_profile=(), eps=None):
    count, steps, t = 0, 0, 0.0
    last = [None]
    tail = cfg.get("AS_tail", 20)
    adt = cfg.get("AS_dt", 0.1)
    while count < tail:
        if last[0] is not None and abs(last[0] - cfg["U"]) <= 0:
            count += 1

        def sU(k, dtk, t=t):
            ts = t if k == 0 else t + dtk
            last[0] = fermi(ts, cfg["AS_midpoint"], cfg["AS_rate"]) * cfg["U"]
            return last[0]

        G1, Txx, Txy = rk4_step(cfg, G1, Txx, Txy, adt, sU, lambda k, d: (0.0, 0), V_profile, eps)
        t += adt
        steps += 1
    return G1, Txx, Txy, steps


def rk4(cfg, G1, Txx, Txy, V_profile=(), eps=None, max_steps=None, record=True):
    """Production RK4 loop; returns (rc, G1, Txx, Txy, traj list)."""
    tmax = float(cfg["time_steps"]) * cfg["dt"]
    t, n = 0.0, 0
    traj = []
    gate = cfg.get("gate_time", 50.0)
    while t < tmax:
        if max_steps is not None and n >= max_steps:
            break
        if record:
            traj.append(G1.copy())
        n += 1
        dens = 2 * G1[0, 0].imag
        if abs(dens) > 3 or math.isnan(dens):
            return 1, G1, Txx, Txy, traj

        def sq(k, dtk, t=t):
            ts = t if k == 0 else t + dtk
            if ts < gate:
                return 0.0, 0
            on = fermi(ts, cfg["quench_on_midpoint"], cfg["quench_rate"])
            if cfg["q_type"] == "pulse":
                w = (on * (1 - fermi(ts, cfg["quench_off_midpoint"], cfg["quench_rate"]))) * cfg["quench_strength"]
            else:
                w = on * cfg["quench_strength"]
            return w, cfg["Nq"]

        G1, Txx, Txy = rk4_step(cfg, G1, Txx, Txy, cfg["dt"], lambda k, d: cfg["U"], sq,
                                V_profile, eps)
        t += cfg["dt"]
    return 0, G1, Txx, Txy, traj
