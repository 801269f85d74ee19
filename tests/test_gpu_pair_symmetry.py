"""Pair-symmetric path (one contiguous tile-product pass and the update pass
that adds the pair-exchanged half) against the general two-pass path and the
oracle.

A state symmetric under exchanging both pairs and Hermitian as an N^2 x N^2
matrix, with G1 anti-Hermitian, has a pair-symmetric G2 right-hand side; the
library then runs the pair-symmetric path (g12_pair_symmetric reports it).
The environment variable G12KBA_PAIR_SYMMETRY=0, read when a context is
created, keeps the general path, so both run on the same input here:
* the N = 64 workload, two RK4 steps through the quench ramp, in every
  precision (FP64 within 1e-12 relative; the mixed modes within their
  own error bounds of each other);
* an uneven batch (N = 6 ... 60, every tile bucket, N not a multiple of 4 or
  16), adiabatic switching from G2 = 0 plus production steps, in every
  precision, and a batch whose N are all multiples of 16 (N = 16, 32, 48: the
  mixed runs' 8-wave update build with non-temporal stores in every block);
* path selection: non-symmetric, non-Hermitian or G1-not-anti-Hermitian states
  take the general path;
* layout conversion: a change of path converts every trajectory's state
  between the Hermitian half layout and the natural one; states read back
  exactly, and the steps after each change match a run that stayed on the
  path.
The oracle comparisons of the same path are test_gpu_runs.py (config 1 and 2
full runs, the N = 64 workload) and test_gpu_ensemble.py.
"""
import contextlib
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

g = pytest.importorskip("g12kba_amd")

from tests.helpers import anti_hermitian, physical_g2, rand_c  # noqa: E402


@contextlib.contextmanager
def general_path(on):
    """G12KBA_PAIR_SYMMETRY=0 while contexts are created (when on)."""
    old = os.environ.get("G12KBA_PAIR_SYMMETRY")
    if on:
        os.environ["G12KBA_PAIR_SYMMETRY"] = "0"
    try:
        yield
    finally:
        if old is None:
            os.environ.pop("G12KBA_PAIR_SYMMETRY", None)
        else:
            os.environ["G12KBA_PAIR_SYMMETRY"] = old


def relerr(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def n64_params(prec):
    return g.make_params(Ns=32, Nb=2, t1=1.0, t2=0.5, U=1.5, dt=0.01, time_steps=2, quench_strength=1.0, Nq=8,
                         q_type="full", EHM=True, V_law="coulomb", quench_rate=2.0, quench_on_midpoint=0.01,
                         quench_gate_time=0.0, AS_rate=0.5, AS_midpoint=10.0, precision=prec)


@pytest.fixture(scope="module")
def n64_state():
    rng = np.random.default_rng(6401)
    p = n64_params("fp64")
    return g.make_G0(p), physical_g2(rng, 32, 2), physical_g2(rng, 32, 2)


def run_solo(p, state, general):
    with general_path(general):
        with g.Propagator(p) as prop:
            prop.set_state(*state)
            sym = prop.pair_symmetric()
            rc, traj = prop.rk4()
            out = prop.get_state()
    return sym, rc, traj, out


@pytest.mark.parametrize("prec", ["fp64", "fp32_3m", "bf16x3"])
def test_paths_agree_n64(n64_state, prec):
    p = n64_params(prec)
    sym_a, rc_a, traj_a, (G1a, Gxxa, Gxya) = run_solo(p, n64_state, general=False)
    sym_b, rc_b, traj_b, (G1b, Gxxb, Gxyb) = run_solo(p, n64_state, general=True)
    print(f"n64 {prec}: paths {sym_a}/{sym_b}", flush=True)
    assert sym_a and not sym_b
    assert rc_a == rc_b == 0 and len(traj_a) == len(traj_b) == 2
    if prec == "fp64":
        assert relerr(G1a, G1b) < 1e-12 and relerr(traj_a, traj_b) < 1e-12
        assert relerr(Gxxa, Gxxb) < 1e-12 and relerr(Gxya, Gxyb) < 1e-12
    else:
        assert np.abs(np.diagonal(G1a - G1b).imag).max() < 1e-6
        assert relerr(Gxxa, Gxxb) < 1e-4 and relerr(Gxya, Gxyb) < 1e-4
    # the pair-symmetric path keeps the exchange symmetry exactly
    from oracle import fast
    T = fast.ref_to_nat(Gxxa, 32, 2).reshape(64, 64, 64, 64)
    assert np.array_equal(T, T.transpose(1, 0, 3, 2))


def uneven_params(prec):
    geoms = [(6, 1), (10, 2), (7, 3), (37, 1), (12, 4), (20, 3)]  # N = 6, 20, 21, 37, 48, 60
    out = []
    for k, (Ns, Nb) in enumerate(geoms):
        out.append(g.make_params(Ns=Ns, Nb=Nb, t1=1.0, t2=0.3 + 0.05 * k, U=1.0 + 0.3 * k, dt=0.02, time_steps=6,
                                 quench_strength=1.0, Nq=max(1, Ns // 4), q_type="full", EHM=(k % 2 == 0),
                                 V_law="coulomb", quench_rate=2.0, quench_on_midpoint=0.05, quench_gate_time=0.0,
                                 AS_rate=0.5, AS_midpoint=10.0, AS_tail=3, precision=prec))
    return out


def run_batch(params, general):
    with general_path(general):
        b = g.Batch(params)
        try:
            for i, q in enumerate(params):
                b.set_state(i, g.make_G0(q))
            sym = b.pair_symmetric()
            b.adiabatic_switching()
            rc, trajs = b.rk4()
            states = [b.get_state(i) for i in range(len(params))]
        finally:
            b.close()
    return sym, rc, trajs, states


@pytest.mark.parametrize("prec", ["fp64", "fp32_3m", "bf16x3"])
def test_paths_agree_uneven_batch(prec):
    params = uneven_params(prec)
    sym_a, rc_a, tr_a, st_a = run_batch(params, general=False)
    sym_b, rc_b, tr_b, st_b = run_batch(params, general=True)
    print(f"batch {prec}: paths {sym_a}/{sym_b}", flush=True)
    assert sym_a and not sym_b and rc_a == rc_b
    for i in range(len(params)):
        G1a, Gxxa, Gxya = st_a[i]
        G1b, Gxxb, Gxyb = st_b[i]
        assert np.abs(Gxxa).max() > 1e-8  # correlations have built up
        if prec == "fp64":
            assert relerr(G1a, G1b) < 1e-12 and relerr(tr_a[i], tr_b[i]) < 1e-12
            assert relerr(Gxxa, Gxxb) < 1e-10 and relerr(Gxya, Gxyb) < 1e-10
        else:
            assert np.abs(np.diagonal(G1a - G1b).imag).max() < 1e-6


def aligned_params(prec):
    geoms = [(8, 2), (16, 2), (16, 3)]  # N = 16, 32, 48: every N a multiple of 16
    return [g.make_params(Ns=Ns, Nb=Nb, t1=1.0, t2=0.35 + 0.05 * k, U=1.1 + 0.2 * k, dt=0.02, time_steps=5,
                          quench_strength=1.0, Nq=max(1, Ns // 4), q_type="full", EHM=True, V_law="coulomb",
                          quench_rate=2.0, quench_on_midpoint=0.05, quench_gate_time=0.0, AS_rate=0.5,
                          AS_midpoint=10.0, AS_tail=3, precision=prec) for k, (Ns, Nb) in enumerate(geoms)]


@pytest.mark.parametrize("prec", ["fp64", "fp32_3m", "bf16x3"])
def test_paths_agree_aligned_batch(prec):
    """A batch whose N are all multiples of 16 (the mixed runs take the
    update pass's 8-wave build and its non-temporal form for every block)
    against the general path."""
    params = aligned_params(prec)
    sym_a, rc_a, tr_a, st_a = run_batch(params, general=False)
    sym_b, rc_b, tr_b, st_b = run_batch(params, general=True)
    print(f"aligned batch {prec}: paths {sym_a}/{sym_b}", flush=True)
    assert sym_a and not sym_b and rc_a == rc_b
    for i in range(len(params)):
        G1a, Gxxa, Gxya = st_a[i]
        G1b, Gxxb, Gxyb = st_b[i]
        assert np.abs(Gxxa).max() > 1e-8
        if prec == "fp64":
            assert relerr(G1a, G1b) < 1e-12 and relerr(tr_a[i], tr_b[i]) < 1e-12
            assert relerr(Gxxa, Gxxb) < 1e-10 and relerr(Gxya, Gxyb) < 1e-10
        else:
            assert np.abs(np.diagonal(G1a - G1b).imag).max() < 1e-6


def test_path_selection():
    Ns, Nb = 3, 2
    N = Ns * Nb
    p = g.make_params(Ns=Ns, Nb=Nb, t1=1.0, t2=0.4, U=1.2, dt=0.02, time_steps=2, EHM=True, V_law="coulomb")
    rng = np.random.default_rng(5)
    G1 = g.make_G0(p)
    phys = physical_g2(rng, Ns, Nb)
    A = rand_c(rng, N * N, N * N, scale=1e-3)
    herm_only = A + A.conj().T            # Hermitian, not pair-symmetric
    with g.Propagator(p) as prop:
        assert prop.pair_symmetric()      # G2 = 0 at creation
        prop.set_state(G1, phys, phys)
        assert prop.pair_symmetric()
        prop.set_state(G1, herm_only, phys)
        assert not prop.pair_symmetric()
        prop.set_state(G1, 1j * phys, phys)   # pair-symmetric, anti-Hermitian
        assert not prop.pair_symmetric()
        prop.set_state(G1 + rand_c(rng, N, N, scale=1e-3), phys, phys)  # G1 not anti-Hermitian
        assert not prop.pair_symmetric()
        prop.set_state(G1 + anti_hermitian(rng, N, 1e-3), phys, None)
        assert prop.pair_symmetric()
        rc, _ = prop.rk4()
        assert rc == 0
    with general_path(True):
        with g.Propagator(p) as prop:
            assert not prop.pair_symmetric()
    hf = g.make_params(Ns=Ns, Nb=Nb, t1=1.0, t2=0.4, U=1.2, dt=0.02, time_steps=2, HF=True)
    with g.Propagator(hf) as prop:
        assert not prop.pair_symmetric()  # no G2 to propagate


def test_layout_follows_path_changes():
    """A run keeps its G2 states in the Hermitian half layout on the
    pair-symmetric path and in the natural layout on the general one; a
    g12_set_state that changes the path converts every other trajectory's
    state.  Every state reads back exactly through both changes, and the steps
    after each change agree with a run that stayed on that path."""
    geoms = [(6, 1), (10, 2), (7, 3)]  # N = 6, 20, 21
    params = [g.make_params(Ns=Ns, Nb=Nb, t1=1.0, t2=0.4, U=1.1 + 0.2 * k, dt=0.02, time_steps=3, EHM=True,
                            V_law="coulomb") for k, (Ns, Nb) in enumerate(geoms)]
    rng = np.random.default_rng(77)
    phys = [(g.make_G0(q), physical_g2(rng, Ns, Nb), physical_g2(rng, Ns, Nb)) for q, (Ns, Nb) in zip(params, geoms)]
    N1 = geoms[1][0] * geoms[1][1]
    A = rand_c(rng, N1 * N1, N1 * N1, scale=1e-3)
    herm_only = A + A.conj().T  # Hermitian, not pair-symmetric: the general path

    def same(b, i, st):
        G1, Gxx, Gxy = b.get_state(i)
        return np.array_equal(G1, st[0]) and np.array_equal(Gxx, st[1]) and np.array_equal(Gxy, st[2])

    def run(states, general):
        with general_path(general):
            b = g.Batch(params)
            try:
                for i, st in enumerate(states):
                    b.set_state(i, *st)
                sym = b.pair_symmetric()
                rc, _ = b.rk4()
                return sym, rc, [b.get_state(i) for i in range(len(params))]
            finally:
                b.close()

    mixed = [phys[0], (phys[1][0], herm_only, phys[1][2]), phys[2]]
    b = g.Batch(params)
    try:
        for i, st in enumerate(phys):
            b.set_state(i, *st)
        assert b.pair_symmetric()
        assert all(same(b, i, phys[i]) for i in range(3))  # half layout, read back exactly
        b.set_state(1, *mixed[1])  # the run turns general: trajectories 0 and 2 convert to natural
        assert not b.pair_symmetric()
        assert all(same(b, i, mixed[i]) for i in range(3))
        b.set_state(1, *phys[1])   # and back: every state converts to the half layout
        assert b.pair_symmetric()
        assert all(same(b, i, phys[i]) for i in range(3))
        b.set_state(1, *mixed[1])
        rc, _ = b.rk4()
        after_general = [b.get_state(i) for i in range(3)]
    finally:
        b.close()
    sym_ref, rc_ref, ref = run(mixed, general=True)
    assert rc == rc_ref == 0 and not sym_ref
    for a, r in zip(after_general, ref):
        for x, y in zip(a, r):
            assert relerr(x, y) < 1e-13
    sym_a, rc_a, st_a = run(phys, general=False)
    sym_b, rc_b, st_b = run(phys, general=True)
    assert sym_a and not sym_b and rc_a == rc_b == 0
    for a, r in zip(st_a, st_b):
        for x, y in zip(a, r):
            assert relerr(x, y) < 1e-12
