"""CPU suite for the library boundary: symbols, parser, writers, dipole, G0.

No GPU compute is called here (host-only entry points of libg12kba.so)."""
import json
import os
import re
import shutil
import tempfile

import numpy as np
import pytest

import g12kba_amd as g
from oracle.oracle import Oracle

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLD = os.path.join(HERE, "golden")


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "g12kba.h")).read()
    declared = set(re.findall(r"\b(g12_[a-z0-9_]+)\s*\(", hdr))
    assert declared, "no declarations found"
    for name in sorted(declared):
        assert hasattr(g.lib, name), f"{name} not exported"
    assert declared == set(g.EXPORTED)


@pytest.mark.parametrize("case", sorted(json.load(open(os.path.join(GOLD, "parser_cases.json")))))
def test_parser_matches_reference(case):
    c = json.load(open(os.path.join(GOLD, "parser_cases.json")))[case]
    d = tempfile.mkdtemp()
    try:
        path = os.path.join(d, "input")
        with open(path, "w") as f:
            f.write(c["text"])
        p = g.parse_input(path)
    finally:
        shutil.rmtree(d)
    v = c["values"]
    for k in ("t1", "t2", "t3", "U", "time_steps", "quench_strength", "Ns", "Nq", "Nb", "AS_rate",
              "AS_midpoint", "quench_rate", "quench_on_midpoint", "quench_off_midpoint", "HF", "EHM"):
        assert getattr(p, k) == v[k], k
    assert p.dt == v["dt_fixed"]
    # the reference keeps q_type "full" for none/invalid values
    assert p.q_type == (1 if v["q_type"] == 1 else 0)


def test_parser_missing_file_gives_zeros():
    p = g.parse_input("/nonexistent/input")
    assert p.Ns == 0 and p.U == 0 and p.dt == 0


def test_parser_extension_keys():
    d = tempfile.mkdtemp()
    try:
        path = os.path.join(d, "input")
        text = open(os.path.join(GOLD, "input_sample")).read() + "long_range_law\ncoulomb\nfilling\n3\nprecision\nbf16x3\n"
        open(path, "w").write(text)
        p = g.parse_input(path)
        assert p.V_law == g.VLAW_COULOMB and p.p_number == 3 and p.precision == g.PREC_BF16X3
        assert p.U == 0.5 and p.Ns == 4
    finally:
        shutil.rmtree(d)


def test_writers_byte_identical_to_reference():
    z = np.load(os.path.join(GOLD, "writers.npz"))
    want = json.load(open(os.path.join(GOLD, "writers.json")))
    d = tempfile.mkdtemp()
    try:
        g.write_outputs(d, z["traj"], z["p"], progress=-1)
        for fn, text in want.items():
            assert open(os.path.join(d, fn)).read() == text, fn
    finally:
        shutil.rmtree(d)


def test_dipole_matches_reference_golden():
    for path in sorted(os.listdir(GOLD)):
        if path.startswith("routines_"):
            z = np.load(os.path.join(GOLD, path))
            assert np.array_equal(g.dipole(z["traj"]), z["dipole"])


def test_dipole_closed_form():
    rng = np.random.default_rng(2)
    N = 6
    traj = rng.standard_normal((4, N, N)) + 1j * rng.standard_normal((4, N, N))
    want = np.zeros(4)
    for i in range(N // 2):
        want += 2 * (N / 2 - i - 0.5) * (traj[:, i, i].imag - traj[:, N - 1 - i, N - 1 - i].imag) / N
    assert np.abs(g.dipole(traj) - want).max() < 1e-14


def test_fermi_matches_oracle():
    o = Oracle()
    for t in (-5.0, 0.0, 9.9, 10.0, 50.0, 1e4):
        assert g.fermi(t, 10.0, 0.5) == o.fermi(t, 10.0, 0.5)


@pytest.mark.parametrize("Ns,Nb,t2", [(4, 2, 0.5), (16, 2, 0.3), (32, 2, 0.5), (24, 2, 0.4), (7, 3, 0.8)])
def test_g0_projector(Ns, Nb, t2):
    """G1(0)/i is the projector on the p lowest states of h(U=0): P^2 = P,
    Tr P = p, [h0, P] = 0, and h0 P = P h0 P has the p lowest eigenvalues."""
    p = g.make_params(Ns=Ns, Nb=Nb, t1=1.0, t2=t2)
    G = g.make_G0(p)
    P = -1j * G
    N = Ns * Nb
    h0 = np.zeros((N, N))
    for X in range(N):
        i, a = X % Ns, X // Ns
        if i + 1 < Ns:
            h0[X, X + 1] = h0[X + 1, X] = -1.0
        if a + 1 < Nb:
            h0[X, X + Ns] = h0[X + Ns, X] = -t2
    assert np.abs(P - P.conj().T).max() < 1e-13
    assert np.abs(P @ P - P).max() < 1e-12
    assert abs(np.trace(P) - Ns // 2) < 1e-12
    assert np.abs(h0 @ P - P @ h0).max() < 1e-11
    e = np.linalg.eigvalsh(h0)
    assert abs(np.trace(h0 @ P).real - e[: Ns // 2].sum()) < 1e-11


def test_g0_filling_override():
    p = g.make_params(Ns=6, Nb=2, t1=1.0, t2=0.5, p_number=6)
    assert abs(np.trace(-1j * g.make_G0(p)) - 6) < 1e-12


@pytest.mark.parametrize("Ns,Nb,t2", [(16, 2, 0.3), (7, 3, 0.8)])
def test_g0_fp32_emulation(Ns, Nb, t2):
    """The FP32-eigensolver emulation (the reference's LAPACKE_cheev precision,
    init_G.h:31-38) gives the same projector to FP32 accuracy on a
    non-degenerate Fermi level, and still a Hermitian one with trace p."""
    p = g.make_params(Ns=Ns, Nb=Nb, t1=1.0, t2=t2)
    G64 = g.make_G0(p)
    G32 = g.make_G0(p, fp32_emulation=True)
    P = -1j * G32
    assert np.abs(P - P.conj().T).max() < 1e-12
    assert abs(np.trace(P) - Ns // 2) < 1e-5
    assert np.abs(P @ P - P).max() < 1e-5
    d = np.abs(G32 - G64).max()
    assert 0 < d < 1e-5  # rounded through float, and only that far


@pytest.mark.parametrize("Ns,Nb", [(7, 1), (8, 2), (5, 3)])
def test_init_g_test_domain_wall(Ns, Nb):
    """init_G_test (init_G.h:78-86): G1[X][X] = i for every orbital of the
    first int(Ns/2) sites, zero elsewhere; a projector with trace Nb*int(Ns/2)."""
    p = g.make_params(Ns=Ns, Nb=Nb)
    G = g.init_G_test(p)
    N = Ns * Nb
    site = np.arange(N) % Ns
    expect = np.diag(np.where(site < Ns // 2, 1j, 0.0))
    assert np.array_equal(G, expect)
    P = -1j * G
    assert np.allclose(P @ P, P) and abs(np.trace(P) - Nb * (Ns // 2)) == 0


@pytest.mark.parametrize("bad", [dict(AS_rate=-0.5), dict(U=float("nan")), dict(t2=float("inf")), dict(AS_dt=0.0),
                                 dict(dt=-0.01, time_steps=-100), dict(AS_rate=1e6)])
def test_create_rejects_protocols_that_never_end(bad):
    """A falling or stalled switching ramp, a clock that never reaches its end
    and non-finite parameters are G12_E_ARG at create time, before any device
    work (AS_rate = 1e6: fermi(t) U would reach U exactly only after ~4e8
    switching steps)."""
    import ctypes
    import time
    p = g.make_params(Ns=3, Nb=1, U=1.0, dt=0.02, time_steps=10)
    for k, v in bad.items():
        setattr(p, k, v)
    h = ctypes.c_void_p()
    t0 = time.perf_counter()
    assert g.lib.g12_create(ctypes.byref(p), ctypes.byref(h)) == -1
    assert time.perf_counter() - t0 < 5
    assert g.lib.g12_batch_create(ctypes.byref(p), 1, ctypes.byref(h)) == -1


@pytest.mark.parametrize("ok", [dict(), dict(AS_rate=0.0, AS_midpoint=0.0), dict(dt=0.0)])
def test_create_accepts_what_the_reference_runs(ok):
    """Missing keys parse as 0 in the reference (read_input.h:53-83): AS_rate = 0
    (a step ramp) and dt = 0 (no production step) still run there.  Without a
    GPU creation stops at device selection (G12_E_HIP), after validation."""
    import ctypes
    p = g.make_params(Ns=3, Nb=1, U=1.0, dt=0.02, time_steps=10)
    for k, v in ok.items():
        setattr(p, k, v)
    h = ctypes.c_void_p()
    rc = g.lib.g12_create(ctypes.byref(p), ctypes.byref(h))
    if rc == 0:
        g.lib.g12_destroy(h)
    assert rc in (0, -2)


@pytest.mark.parametrize("Ns,Nb,prec,ok", [(65, 1, "fp32_3m", False), (33, 2, "bf16x3", False), (129, 1, "fp64", False),
                                           (1, 129, "fp64", False), (64, 2, "fp64", True), (36, 2, "fp64", True),
                                           (2, 12, "fp32_3m", True), (32, 2, "bf16x3", True)])
def test_create_size_limits(Ns, Nb, prec, ok):
    """N = Ns*Nb up to 128 in FP64 with any split into sites and orbitals
    (read_input.h:85-146 takes any Ns, Nb); the mixed-precision modes stop at
    N = 64.  Beyond: G12_E_SIZE before any device work."""
    import ctypes
    p = g.make_params(Ns=Ns, Nb=Nb, U=1.0, dt=0.02, time_steps=2, precision=prec)
    h = ctypes.c_void_p()
    rc = g.lib.g12_create(ctypes.byref(p), ctypes.byref(h))
    if rc == 0:
        g.lib.g12_destroy(h)
    if ok:
        assert rc in (0, -2)
    else:
        assert rc == -4
        assert g.lib.g12_batch_create(ctypes.byref(p), 1, ctypes.byref(h)) == -4
