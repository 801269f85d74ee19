"""Routine-level parity: each C-ABI entry point (GPU) against the CPU oracle.

Inputs are seeded random matrices at small N with non-empty long-range shells,
so every branch of the reference routines is exercised.  FP64 everywhere:
agreement to rounding (1e-12 relative) is required.  h_HF, the collision
integral, Psi, Pi and G2_step run the production kernels (the stage set-up's
G1 part and tables, the tile passes' source and update epilogues), so these
tests check the hot path's code routine by routine.
"""
import numpy as np
import pytest

from tests.helpers import GEOMS, rand_c

pytestmark = pytest.mark.gpu

g = pytest.importorskip("g12kba_amd")
from oracle.oracle import Oracle  # noqa: E402
from oracle import fast  # noqa: E402

O = Oracle()
TOL = 1e-12


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("Ns,Nb,nV", GEOMS)
def test_hf_hamiltonian(Ns, Nb, nV):
    rng = np.random.default_rng(100 + Ns * 10 + Nb)
    N = Ns * Nb
    G1, h0 = rand_c(rng, N, N), rand_c(rng, N, N)
    V, eps = rng.uniform(0.1, 1.0, nV), rng.standard_normal(Nb)
    a = g.h_HFx_quench(h0, G1, G1, Ns, Nb, 1.3, 1.0, 0.4, 0.7, eps, V, 1)
    b = O.hf(h0, G1, Ns, Nb, 1.3, 1.0, 0.4, 0.7, eps, V, 1)
    assert rel(a, b) < TOL


@pytest.mark.parametrize("N", [1, 5, 8, 33, 64])
def test_h1_commutator(N):
    rng = np.random.default_rng(N)
    G, h = rand_c(rng, N, N), rand_c(rng, N, N)
    assert rel(g.hG_commutator(G, h), h @ G - G @ h) < TOL


@pytest.mark.parametrize("Ns,Nb", [(2, 1), (3, 2), (2, 3), (5, 1), (4, 4), (7, 2), (8, 2), (16, 2), (24, 2)])
def test_h2_commutator(Ns, Nb):
    rng = np.random.default_rng(Ns * 7 + Nb)
    N = Ns * Nb
    G2 = rand_c(rng, N * N, N * N)
    hx, hy = rand_c(rng, N, N), rand_c(rng, N, N)
    a = g.h2xyG2xyxy_commutator(G2, hx, hy, Ns, Nb)
    if N <= 16:
        b = O.comm2(G2, hx, hy, Ns, Nb)
    else:
        b = fast.nat_to_ref(fast.comm2(fast.ref_to_nat(G2, Ns, Nb), hx, hy), Ns, Nb)
    assert rel(a, b) < TOL


@pytest.mark.parametrize("Ns,Nb,nV", GEOMS)
def test_collision(Ns, Nb, nV):
    rng = np.random.default_rng(200 + Ns * 10 + Nb)
    N = Ns * Nb
    P, Q = rand_c(rng, N * N, N * N), rand_c(rng, N * N, N * N)
    V = rng.uniform(0.1, 1.0, nV)
    assert rel(g.collision_int(P, Q, Ns, Nb, 0.9, V), O.collision(P, Q, Ns, Nb, 0.9, V)) < TOL


@pytest.mark.parametrize("Ns,Nb,nV", GEOMS)
@pytest.mark.parametrize("same_spin", [0, 1])
def test_psi(Ns, Nb, nV, same_spin):
    rng = np.random.default_rng(300 + Ns * 10 + Nb)
    N = Ns * Nb
    G1 = rand_c(rng, N, N)
    V = rng.uniform(0.1, 1.0, nV)
    fn = g.Psi_xxxx if same_spin else g.Psi_xyxy
    b = O.psi(G1, Ns, Nb, 1.1, V, same_spin)
    if np.abs(b).max() == 0:
        # W_xx = 0 (one site per orbital pair, no shells): the production
        # tables form Qp^xx as Qp^xy less its w = Z term, two roundings of one
        # product, so the result is zero to rounding of |G1|^2 U
        assert np.abs(fn(G1, Ns, Nb, 1.1, V)).max() <= 1e-14 * 1.1 * np.abs(G1).max() ** 2
    else:
        assert rel(fn(G1, Ns, Nb, 1.1, V), b) < TOL


@pytest.mark.parametrize("Ns,Nb,nV", GEOMS)
def test_pi(Ns, Nb, nV):
    rng = np.random.default_rng(400 + Ns * 10 + Nb)
    N = Ns * Nb
    G1, P, Q = rand_c(rng, N, N), rand_c(rng, N * N, N * N), rand_c(rng, N * N, N * N)
    V = rng.uniform(0.1, 1.0, nV)
    # both call orders of RK_solver.h:163-164
    assert rel(g.Pi_xyxy(G1, P, Q, 0.8, Ns, Nb, V), O.pi(G1, P, Q, 0.8, Ns, Nb, V)) < TOL
    assert rel(g.Pi_xyxy(G1, Q, P, 0.8, Ns, Nb, V), O.pi(G1, Q, P, 0.8, Ns, Nb, V)) < TOL


@pytest.mark.parametrize("Ns,Nb", [(2, 1), (3, 2), (4, 3)])
def test_steps(Ns, Nb):
    rng = np.random.default_rng(500 + Ns)
    N = Ns * Nb
    c, I, K, R = (rand_c(rng, N, N) for _ in range(4))
    a, b = g.G1_step(c, I, K, R, 2.0, Ns, Nb), O.g1_step(c, I, K, R, 2.0, Ns, Nb)
    assert rel(a[0], b[0]) < TOL and rel(a[1], b[1]) < TOL
    c, s, pi, K, R = (rand_c(rng, N * N, N * N) for _ in range(5))
    a, b = g.G2_step(c, s, pi, K, R, 2.0, Ns, Nb), O.g2_step(c, s, pi, K, R, 2.0, Ns, Nb)
    assert rel(a[0], b[0]) < TOL and rel(a[1], b[1]) < TOL
