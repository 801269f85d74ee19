"""Shared inputs for the parity tests."""
import numpy as np


def rand_c(rng, *shape, scale=1.0):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) * scale


def anti_hermitian(rng, n, scale=1.0):
    a = rand_c(rng, n, n, scale=scale)
    return 0.5 * (a - a.conj().T)


def oracle_cfg(p):
    """Oracle run configuration from a g12kba_amd.Params."""
    return dict(Ns=p.Ns, Nb=p.Nb, t1=p.t1, t2=p.t2, U=p.U, dt=p.dt, time_steps=p.time_steps,
                quench_strength=p.quench_strength, Nq=p.Nq,
                q_type="pulse" if p.q_type == 1 else "full", HF=bool(p.HF),
                AS_rate=p.AS_rate, AS_midpoint=p.AS_midpoint, quench_rate=p.quench_rate,
                quench_on_midpoint=p.quench_on_midpoint, quench_off_midpoint=p.quench_off_midpoint,
                AS_dt=p.AS_dt, AS_tail=p.AS_tail, gate_time=p.quench_gate_time)


# (Ns, Nb, nV) geometries for routine-level parity; nV long-range shells
GEOMS = [(2, 1, 0), (3, 2, 2), (2, 3, 1), (4, 2, 3), (5, 1, 4), (3, 3, 2)]


def physical_g2(rng, Ns, Nb, scale=1e-3):
    """Random G2 in the reference layout with the symmetries of a G2 evolved
    from zero: Hermitian as an N^2 x N^2 matrix and symmetric under exchanging
    both particle pairs, T[p1][p2][q1][q2] = T[p2][p1][q2][q1] (reference pair
    index of (p1, p2) with p = site + orb*Ns: i + j*Ns + (a + b*Nb)*Ns^2)."""
    N = Ns * Nb
    site, orb = np.arange(N) % Ns, np.arange(N) // Ns
    pair = (site[:, None] + site[None, :] * Ns + (orb[:, None] + orb[None, :] * Nb) * Ns * Ns).ravel()
    A = rand_c(rng, N * N, N * N, scale=scale)
    T = (A + A.conj().T).reshape(N, N, N, N)
    T = 0.5 * (T + T.transpose(1, 0, 3, 2))
    M = np.empty((N * N, N * N), np.complex128)
    M[np.ix_(pair, pair)] = T.reshape(N * N, N * N)
    return M


def product_state(rng, N, rank=3):
    """T = sum_r a_r (x) b_r in the natural layout, T[p1][p2][q1][q2] =
    sum_r a_r[p1][q1] b_r[p2][q2] (complex128, N^4 elements)."""
    a = [rand_c(rng, N, N) for _ in range(rank)]
    b = [rand_c(rng, N, N) for _ in range(rank)]
    T = np.zeros((N, N, N, N), np.complex128)
    for ar, br in zip(a, b):
        T += np.einsum("ac,bd->abcd", ar, br)
    return T, a, b


def product_state_commutator(a, b, hx, hy):
    """[H2, T] = (hx(x)1 + 1(x)hx) T - T (hy(x)1 + 1(x)hy) of T = sum_r a_r (x) b_r:
    sum_r (hx a_r - a_r hy) (x) b_r + a_r (x) (hx b_r - b_r hy), natural layout."""
    N = hx.shape[0]
    out = np.zeros((N, N, N, N), np.complex128)
    for ar, br in zip(a, b):
        out += np.einsum("ac,bd->abcd", hx @ ar - ar @ hy, br)
        out += np.einsum("ac,bd->abcd", ar, hx @ br - br @ hy)
    return out
