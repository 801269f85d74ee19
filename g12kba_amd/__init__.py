"""MI355X-native HF-GKBA G1-G2 (GW) propagator — Python host mirror.

Thin ctypes layer over ``libg12kba.so`` (the C ABI declared in
``include/g12kba.h``).  The functions keep the reference routine names and
argument roles (G12KBA-GW/src/headers/*.h) so that tests read like the
reference's own call sites; arrays are numpy complex128 in the reference
layouts (G1: N x N; G2: N^2 x N^2 with the pair index
A = i + j*Ns + (a + b*Nb)*Ns^2).

The library is required: importing this package without a built
``g12kba_amd/lib/libg12kba.so`` raises, and there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("G12KBA_LIBRARY") or os.path.join(HERE, "lib", "libg12kba.so")

G12_OK, G12_GUARD = 0, 1
QUENCH_FULL, QUENCH_PULSE = 0, 1
VLAW_EXP, VLAW_COULOMB, VLAW_EXPLICIT = 0, 1, 2
PREC_FP64, PREC_FP32_3M, PREC_BF16X3 = 0, 1, 2
PRECISIONS = {"fp64": PREC_FP64, "fp32_3m": PREC_FP32_3M, "bf16x3": PREC_BF16X3}
MAX_N, MAX_N_MIXED, MAX_SHELLS = 128, 64, 128
MAX_NB = MAX_N

_dp = ctypes.POINTER(ctypes.c_double)


class G12Error(RuntimeError):
    pass


class Params(ctypes.Structure):
    """Mirror of ``g12_params`` (include/g12kba.h)."""
    _fields_ = [
        ("Ns", ctypes.c_int), ("Nb", ctypes.c_int),
        ("t1", ctypes.c_double), ("t2", ctypes.c_double), ("t3", ctypes.c_double),
        ("U", ctypes.c_double), ("time_steps", ctypes.c_int), ("dt", ctypes.c_double),
        ("quench_strength", ctypes.c_double), ("Nq", ctypes.c_int), ("q_type", ctypes.c_int),
        ("HF", ctypes.c_int), ("EHM", ctypes.c_int),
        ("AS_rate", ctypes.c_double), ("AS_midpoint", ctypes.c_double),
        ("quench_rate", ctypes.c_double), ("quench_on_midpoint", ctypes.c_double),
        ("quench_off_midpoint", ctypes.c_double), ("decay_rate", ctypes.c_double),
        ("V_law", ctypes.c_int), ("nV_profile", ctypes.c_int),
        ("V_profile", ctypes.c_double * MAX_SHELLS), ("eps", ctypes.c_double * MAX_NB),
        ("AS_dt", ctypes.c_double), ("AS_tail", ctypes.c_int),
        ("quench_gate_time", ctypes.c_double), ("p_number", ctypes.c_int),
        ("precision", ctypes.c_int),
    ]


class XPConfig(ctypes.Structure):
    _fields_ = [("hist", ctypes.c_int), ("horizon", ctypes.c_int), ("hidden", ctypes.c_int),
                ("dt", ctypes.c_double), ("seed", ctypes.c_uint64), ("noise", ctypes.c_double)]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
            "(make -C g12kba_amd/csrc); there is no CPU fallback")
    L = ctypes.CDLL(LIB_PATH)
    i, d, p = ctypes.c_int, ctypes.c_double, _dp
    i64, pi64 = ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)
    vp = ctypes.c_void_p
    sig = {
        "g12_params_default": [ctypes.POINTER(Params)],
        "g12_parse_input": [ctypes.c_char_p, ctypes.POINTER(Params)],
        "g12_make_g0": [ctypes.POINTER(Params), p, i],
        "g12_init_g_test": [ctypes.POINTER(Params), p],
        "g12_create": [ctypes.POINTER(Params), ctypes.POINTER(vp)],
        "g12_destroy": [vp],
        "g12_set_state": [vp, p, p, p],
        "g12_get_state": [vp, p, p, p],
        "g12_adiabatic_switching": [vp, ctypes.POINTER(i)],
        "g12_rk4": [vp, p, i64, pi64],
        "g12_rk4_advance": [vp, i64, pi64],
        "g12_rk4_record": [vp, i64, p, i64, pi64],
        "g12_rk4_record_collision": [vp, i64, p, p, i64, pi64],
        "g12_set_collision": [vp, p, i64],
        "g12_set_time": [vp, d, i64],
        "g12_get_time": [vp, p, pi64],
        "g12_timing_enable": [vp, i],
        "g12_timing_get": [vp, i, pi64, p],
        "g12_memory_bytes": [vp, pi64],
        "g12_synchronize": [],
        "g12_hf_hamiltonian": [p, p, p, i, i, d, d, d, d, p, p, i, i],
        "g12_h1_commutator": [p, p, p, i],
        "g12_h2_commutator": [p, p, p, p, i, i, i],
        "g12_h2_commutator_noalloc": [vp, vp, vp, vp, i, i, i, vp],
        "g12_collision": [p, p, p, i, i, d, p, i],
        "g12_psi_xxxx": [p, p, i, i, d, p, i],
        "g12_psi_xyxy": [p, p, i, i, d, p, i],
        "g12_pi": [p, p, p, p, d, i, i, p, i],
        "g12_g1_step": [p, p, p, p, d, i, i],
        "g12_g2_step": [p, p, p, p, p, d, i, i],
        "g12_dipole": [p, p, i64, i],
        "g12_energy": [ctypes.POINTER(Params), d, d, p, p, p, p],
        "g12_state_energy": [vp, p],
        "g12_batch_state_energy": [vp, i, p],
        "g12_write_outputs": [ctypes.c_char_p, p, i64, i, p, d],
        "g12_save_state": [vp, ctypes.c_char_p],
        "g12_load_state": [vp, ctypes.c_char_p],
        "g12_batch_save_state": [vp, i, ctypes.c_char_p],
        "g12_batch_load_state": [vp, i, ctypes.c_char_p],
        "g12_configure": [vp, i, i],
        "g12_pair_symmetric": [vp, ctypes.POINTER(i)],
        "g12_batch_pair_symmetric": [vp, ctypes.POINTER(i)],
        "g12_batch_create": [ctypes.POINTER(Params), i, ctypes.POINTER(vp)],
        "g12_batch_destroy": [vp],
        "g12_batch_size": [vp, ctypes.POINTER(i)],
        "g12_batch_memory_bytes": [vp, pi64],
        "g12_batch_set_state": [vp, i, p, p, p],
        "g12_batch_get_state": [vp, i, p, p, p],
        "g12_batch_adiabatic_switching": [vp, ctypes.POINTER(i)],
        "g12_batch_rk4": [vp, ctypes.POINTER(p), i64, pi64],
        "g12_batch_rk4_advance": [vp, i64],
        "g12_batch_rk4_record": [vp, i64, ctypes.POINTER(p), i64],
        "g12_batch_rk4_record_collision": [vp, i64, ctypes.POINTER(p), ctypes.POINTER(p), i64],
        "g12_batch_guard": [vp, ctypes.POINTER(i)],
        "g12_batch_timing_enable": [vp, i],
        "g12_batch_timing_get": [vp, i, pi64, p],
        "g12_batch_configure": [vp, i, i],
        "g12_batch_tune": [vp],
        "g12_batch_buckets": [vp, i, ctypes.POINTER(i), ctypes.POINTER(i), ctypes.POINTER(i),
                              ctypes.POINTER(i)],
    }
    fp = ctypes.POINTER(ctypes.c_float)
    xcfg = ctypes.POINTER(XPConfig)
    sig.update({
        "g12_xp_default_config": [xcfg],
        "g12_xp_create": [xcfg, ctypes.POINTER(vp)],
        "g12_xp_destroy": [vp],
        "g12_xp_param_count": [vp, pi64],
        "g12_xp_get_params": [vp, fp],
        "g12_xp_set_params": [vp, fp],
        "g12_xp_save": [vp, ctypes.c_char_p],
        "g12_xp_load": [vp, ctypes.c_char_p],
        "g12_xp_set_data": [vp, fp, fp, fp, i64, i],
        "g12_xp_train": [vp, i, i, d, fp],
        "g12_xp_loss_grad": [vp, fp, fp, i, p, fp],
        "g12_xp_forward": [vp, fp, i, fp],
        "g12_xp_features": [vp, p, i, i, i, p, fp, p, p],
        "g12_xp_rollout": [vp, p, i, i, i, p, p, i, i, p, p],
        "g12_xp_rollout_hermitian": [vp, p, i, i, i, p, p, i, i, p, p],
    })
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    L.g12_fermi.argtypes = [d, d, d]
    L.g12_fermi.restype = d
    L.g12_version.restype = ctypes.c_char_p
    return L


lib = _load()

# symbols declared in include/g12kba.h (checked by the CPU test-suite)
EXPORTED = [
    "g12_params_default", "g12_parse_input", "g12_make_g0", "g12_init_g_test", "g12_create", "g12_destroy",
    "g12_set_state", "g12_get_state", "g12_adiabatic_switching", "g12_rk4", "g12_rk4_advance",
    "g12_get_time", "g12_rk4_record", "g12_set_time", "g12_rk4_record_collision", "g12_set_collision", "g12_timing_enable", "g12_timing_get", "g12_memory_bytes", "g12_synchronize",
    "g12_hf_hamiltonian", "g12_h1_commutator", "g12_h2_commutator", "g12_h2_commutator_noalloc", "g12_collision",
    "g12_psi_xxxx", "g12_psi_xyxy", "g12_pi", "g12_g1_step", "g12_g2_step", "g12_fermi",
    "g12_dipole", "g12_energy", "g12_state_energy", "g12_batch_state_energy", "g12_write_outputs", "g12_save_state", "g12_load_state", "g12_batch_save_state", "g12_batch_load_state", "g12_version",
    "g12_configure", "g12_pair_symmetric", "g12_batch_pair_symmetric", "g12_batch_create", "g12_batch_destroy", "g12_batch_size",
    "g12_batch_memory_bytes", "g12_batch_set_state", "g12_batch_get_state",
    "g12_batch_adiabatic_switching", "g12_batch_rk4", "g12_batch_rk4_advance", "g12_batch_rk4_record", "g12_batch_rk4_record_collision", "g12_batch_guard",
    "g12_batch_timing_enable", "g12_batch_timing_get", "g12_batch_configure", "g12_batch_tune",
    "g12_batch_buckets", "g12_xp_default_config", "g12_xp_create", "g12_xp_destroy",
    "g12_xp_param_count", "g12_xp_get_params", "g12_xp_set_params", "g12_xp_save", "g12_xp_load",
    "g12_xp_set_data", "g12_xp_train", "g12_xp_loss_grad", "g12_xp_forward", "g12_xp_features",
    "g12_xp_rollout", "g12_xp_rollout_hermitian",
]


def _check(rc, what):
    if rc < 0:
        raise G12Error(f"{what} failed with code {rc}")
    return rc


def _c(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    if shape is not None and a.size != int(np.prod(shape)):
        raise ValueError(f"expected {shape} elements, got {a.shape}")
    return a


def _p(a):
    return None if a is None else a.ctypes.data_as(_dp)


def _vec(v):
    v = np.ascontiguousarray([] if v is None else v, dtype=np.float64)
    return v, (_p(v) if v.size else None)


# ---------------------------------------------------------------------------
# parameters


def default_params():
    p = Params()
    lib.g12_params_default(ctypes.byref(p))
    return p


def make_params(**kw):
    """Params from keyword overrides of the compiled defaults.

    ``q_type`` accepts "full"/"pulse"/"none"; ``V_law`` accepts
    "exp"/"coulomb"/"explicit"; ``precision`` accepts "fp64"/"fp32_3m"/"bf16x3".
    """
    p = default_params()
    q = kw.pop("q_type", None)
    if q is not None:
        if q == "none":
            p.quench_strength, p.Nq = 0.0, 0
            kw.pop("quench_strength", None)
            kw.pop("Nq", None)
        p.q_type = QUENCH_PULSE if q == "pulse" else QUENCH_FULL
    law = kw.pop("V_law", None)
    if law is not None:
        p.V_law = {"exp": VLAW_EXP, "coulomb": VLAW_COULOMB, "explicit": VLAW_EXPLICIT}.get(law, law)
    prof = kw.pop("V_profile", None)
    if prof is not None:
        prof = list(prof)
        p.nV_profile = len(prof)
        for k, v in enumerate(prof):
            p.V_profile[k] = v
    eps = kw.pop("eps", None)
    if eps is not None:
        for k, v in enumerate(eps):
            p.eps[k] = v
    prec = kw.pop("precision", None)
    if prec is not None:
        p.precision = PRECISIONS.get(prec, prec)
    for k, v in kw.items():
        if k in ("HF", "EHM"):
            v = int(bool(v))
        setattr(p, k, v)
    return p


def draw_ensemble(seed=20261203, count=256, max_n=None):
    """Seeded parameter draws of an ensemble: Ns uniform in {6..20}, Nb uniform
    in {1,2,3}, U uniform in [0.5, 3], t2 uniform in [0.1, 0.8], quench strength
    uniform in [0.1, 2] on Nq = Ns // 4 sites, long-range interactions with
    probability 0.5, t1 = 1 (numpy default_rng, drawn in that order; draws with
    Ns * Nb > max_n are skipped)."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < count:
        Ns = int(rng.integers(6, 21))
        Nb = int(rng.integers(1, 4))
        U = float(rng.uniform(0.5, 3.0))
        t2 = float(rng.uniform(0.1, 0.8))
        s = float(rng.uniform(0.1, 2.0))
        ehm = bool(rng.random() < 0.5)
        if max_n is not None and Ns * Nb > max_n:
            continue
        out.append(dict(Ns=Ns, Nb=Nb, U=U, t1=1.0, t2=t2, quench_strength=s, Nq=Ns // 4, EHM=ehm))
    return out


def ensemble_params(draws, **protocol):
    """Params of each trajectory of an ensemble: physics from each draw (dicts
    of Params fields, e.g. Ns, Nb, U, t1, t2, quench_strength, Nq, EHM), the
    time protocol and precision shared (keyword arguments)."""
    out = []
    for d in draws:
        kw = dict(protocol)
        kw.update(d)
        out.append(make_params(**kw))
    return out


def parse_input(path):
    p = default_params()
    _check(lib.g12_parse_input(path.encode(), ctypes.byref(p)), "g12_parse_input")
    return p


def v_profile(p):
    """The V_n / U profile the library uses for parameters p (n = 1..)."""
    if not p.EHM:
        return np.zeros(0)
    n = min(p.Ns - 1, MAX_SHELLS)
    if p.V_law == VLAW_EXPLICIT:
        return np.array([p.V_profile[k] for k in range(min(n, p.nV_profile))])
    d = np.arange(1, n + 1, dtype=np.float64)
    if p.V_law == VLAW_COULOMB:
        return 1.0 / (1.0 + d)
    return np.exp(-p.decay_rate * d)


def make_G0(p, fp32_emulation=False):
    """Non-interacting initial G1 (make_G0, init_G.h:57-75)."""
    N = p.Ns * p.Nb
    G1 = np.zeros((N, N), np.complex128)
    _check(lib.g12_make_g0(ctypes.byref(p), _p(G1), int(fp32_emulation)), "g12_make_g0")
    return G1


def init_G_test(p):
    """Domain-wall initial G1 (init_G_test, init_G.h:78-86): every orbital of
    the first int(Ns/2) sites occupied."""
    N = p.Ns * p.Nb
    G1 = np.zeros((N, N), np.complex128)
    _check(lib.g12_init_g_test(ctypes.byref(p), _p(G1)), "g12_init_g_test")
    return G1


# ---------------------------------------------------------------------------
# routine-level entry points (reference names, arguments and roles)


def h_HFx_quench(h, G1x, G1y, Ns, Nb, U, t1, t2, w, epsilon, V, quench_extent):
    """h += h_HF[G1x]  (HF.h:57-104); returns the updated h."""
    N = Ns * Nb
    h = _c(h, (N, N)).copy()
    eps, eps_p = _vec(epsilon)
    Vv, Vp = _vec(V)
    _check(lib.g12_hf_hamiltonian(_p(h), _p(_c(G1x, (N, N))), None, Ns, Nb, U, t1, t2, w, eps_p, Vp,
                                  Vv.size, quench_extent), "g12_hf_hamiltonian")
    return h


def hG_commutator(G, h):
    """[h, G] = hG - Gh  (commutators.h:28-51)."""
    G = _c(G)
    N = G.shape[0]
    out = np.zeros((N, N), np.complex128)
    _check(lib.g12_h1_commutator(_p(G), _p(_c(h, (N, N))), _p(out), N), "g12_h1_commutator")
    return out


def h2xyG2xyxy_commutator(G2, h1x, h1y, Ns, Nb, precision="fp64"):
    """[H2, G2] with H2 = h(x)1 + 1(x)h  (commutators.h:53-99)."""
    N = Ns * Nb
    G2 = _c(G2, (N * N, N * N))
    out = np.zeros((N * N, N * N), np.complex128)
    _check(lib.g12_h2_commutator(_p(G2), _p(_c(h1x, (N, N))), _p(_c(h1y, (N, N))), _p(out), Ns, Nb,
                                 PRECISIONS.get(precision, precision)), "g12_h2_commutator")
    return out


def collision_int(G2xyxy, G2xxxx, Ns, Nb, U, V):
    """Collision integral I  (collision_integral.h:27-69)."""
    N = Ns * Nb
    out = np.zeros((N, N), np.complex128)
    Vv, Vp = _vec(V)
    _check(lib.g12_collision(_p(out), _p(_c(G2xyxy, (N ** 2, N ** 2))), _p(_c(G2xxxx, (N ** 2, N ** 2))),
                             Ns, Nb, U, Vp, Vv.size), "g12_collision")
    return out


def Psi_xxxx(G1, Ns, Nb, U, V):
    N = Ns * Nb
    out = np.zeros((N * N, N * N), np.complex128)
    Vv, Vp = _vec(V)
    _check(lib.g12_psi_xxxx(_p(out), _p(_c(G1, (N, N))), Ns, Nb, U, Vp, Vv.size), "g12_psi_xxxx")
    return out


def Psi_xyxy(G1, Ns, Nb, U, V):
    N = Ns * Nb
    out = np.zeros((N * N, N * N), np.complex128)
    Vv, Vp = _vec(V)
    _check(lib.g12_psi_xyxy(_p(out), _p(_c(G1, (N, N))), Ns, Nb, U, Vp, Vv.size), "g12_psi_xyxy")
    return out


def Pi_xyxy(G1y, G2xyxy, G2xxxx, U, Ns, Nb, V):
    """GW polarisation term  (pi.h:25-95)."""
    N = Ns * Nb
    out = np.zeros((N * N, N * N), np.complex128)
    Vv, Vp = _vec(V)
    _check(lib.g12_pi(_p(out), _p(_c(G1y, (N, N))), _p(_c(G2xyxy, (N ** 2, N ** 2))),
                      _p(_c(G2xxxx, (N ** 2, N ** 2))), U, Ns, Nb, Vp, Vv.size), "g12_pi")
    return out


def G1_step(comm, I, K, RK_result, w, Ns, Nb):
    """Returns (K, RK_result + w K)  (G1_G2_t_step.h:24-47)."""
    N = Ns * Nb
    K = _c(K, (N, N)).copy()
    RK = _c(RK_result, (N, N)).copy()
    _check(lib.g12_g1_step(_p(_c(comm, (N, N))), _p(_c(I, (N, N))), _p(K), _p(RK), w, Ns, Nb), "g12_g1_step")
    return K, RK


def G2_step(comm, psi, pi, K, RK_result, w, Ns, Nb):
    """Returns (K, RK_result + w K)  (G1_G2_t_step.h:55-86)."""
    n = (Ns * Nb) ** 2
    K = _c(K, (n, n)).copy()
    RK = _c(RK_result, (n, n)).copy()
    _check(lib.g12_g2_step(_p(_c(comm, (n, n))), _p(_c(psi, (n, n))), _p(_c(pi, (n, n))), _p(K), _p(RK), w,
                           Ns, Nb), "g12_g2_step")
    return K, RK


def fermi(t, t0, tau):
    return lib.g12_fermi(t, t0, tau)


def dipole(traj):
    traj = _c(traj)
    T, N = traj.shape[0], traj.shape[1]
    p = np.zeros(T)
    _check(lib.g12_dipole(_p(p), _p(traj), T, N), "g12_dipole")
    return p


ENERGY_PARTS = ("one_body", "hartree", "exchange", "correlation")


def energy(p, G1, G2xx=None, G2xy=None, lam=1.0, w=0.0):
    """Energy of a state in the reference layouts, both spins: array of
    (one-body, Hartree, exchange, correlation); the total is their sum.  lam
    scales U and every V_n; w is the quench shift on sites < Nq.  Without G2
    the correlation part is 0."""
    N = p.Ns * p.Nb
    G1 = _c(G1, (N, N))
    Gxx = None if G2xx is None else _c(G2xx, (N ** 2, N ** 2))
    Gxy = None if G2xy is None else _c(G2xy, (N ** 2, N ** 2))
    E = np.zeros(4)
    _check(lib.g12_energy(ctypes.byref(p), lam, w, _p(G1), _p(Gxx), _p(Gxy), _p(E)), "g12_energy")
    return E


def write_outputs(directory, traj, dip, progress=-1.0):
    traj = _c(traj)
    dip = np.ascontiguousarray(dip, dtype=np.float64)
    _check(lib.g12_write_outputs(directory.encode(), _p(traj), traj.shape[0], traj.shape[1] if traj.ndim == 3 else 0,
                                 _p(dip), progress), "g12_write_outputs")


# ---------------------------------------------------------------------------
# device-resident propagation


class Propagator:
    """One trajectory on the GPU: Adiabatic_switching / RK4 (AS.h, RK_solver.h)."""

    def __init__(self, params):
        self.p = params
        self.N = params.Ns * params.Nb
        h = ctypes.c_void_p()
        _check(lib.g12_create(ctypes.byref(params), ctypes.byref(h)), "g12_create")
        self.h = h

    def close(self):
        if self.h:
            lib.g12_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_state(self, G1, G2xx=None, G2xy=None):
        N = self.N
        G1 = _c(G1, (N, N))
        Gxx = None if G2xx is None else _c(G2xx, (N ** 2, N ** 2))
        Gxy = None if G2xy is None else _c(G2xy, (N ** 2, N ** 2))
        _check(lib.g12_set_state(self.h, _p(G1), _p(Gxx), _p(Gxy)), "g12_set_state")

    def get_state(self):
        N = self.N
        G1 = np.zeros((N, N), np.complex128)
        Gxx = np.zeros((N * N, N * N), np.complex128)
        Gxy = np.zeros((N * N, N * N), np.complex128)
        _check(lib.g12_get_state(self.h, _p(G1), _p(Gxx), _p(Gxy)), "g12_get_state")
        return G1, Gxx, Gxy

    def get_G1(self):
        N = self.N
        G1 = np.zeros((N, N), np.complex128)
        _check(lib.g12_get_state(self.h, _p(G1), None, None), "g12_get_state")
        return G1

    def adiabatic_switching(self):
        n = ctypes.c_int(0)
        _check(lib.g12_adiabatic_switching(self.h, ctypes.byref(n)), "g12_adiabatic_switching")
        return n.value

    def rk4(self, traj_cap=None):
        """Returns (rc, trajectory[steps, N, N])."""
        if traj_cap is None:
            traj_cap = self.p.time_steps + 2
        traj = np.zeros((traj_cap, self.N, self.N), np.complex128)
        n = ctypes.c_int64(0)
        rc = _check(lib.g12_rk4(self.h, _p(traj), traj_cap, ctypes.byref(n)), "g12_rk4")
        return rc, traj[: min(n.value, traj_cap)]

    def advance(self, nsteps, done=False):
        """Advance nsteps from the current time; returns rc, or (rc, iterations
        begun) with done=True (fewer when the density guard trips)."""
        n = ctypes.c_int64(0)
        rc = _check(lib.g12_rk4_advance(self.h, nsteps, ctypes.byref(n)), "g12_rk4_advance")
        return (rc, n.value) if done else rc

    def record(self, nsteps):
        """Advance nsteps from the current time; returns (rc, G1 before each
        step begun: up to and including the failing one if the guard trips)."""
        traj = np.zeros((max(nsteps, 1), self.N, self.N), np.complex128)
        n = ctypes.c_int64(0)
        rc = _check(lib.g12_rk4_record(self.h, nsteps, _p(traj), nsteps, ctypes.byref(n)), "g12_rk4_record")
        return rc, traj[: n.value]

    def record_collision(self, nsteps):
        """record() that also returns the collision term J = I + I^dagger of each
        step run: (rc, G1 before each step, J of each step)."""
        traj = np.zeros((max(nsteps, 1), self.N, self.N), np.complex128)
        J = np.zeros((max(nsteps, 1), self.N, self.N), np.complex128)
        n = ctypes.c_int64(0)
        rc = _check(lib.g12_rk4_record_collision(self.h, nsteps, _p(traj), _p(J), nsteps, ctypes.byref(n)),
                    "g12_rk4_record_collision")
        ran = n.value - (1 if rc == G12_GUARD else 0)
        return rc, traj[: n.value], J[:ran]

    def set_collision(self, J):
        """Uncorrelated (HF) contexts: the collision term at the step times of
        the next len(J) - 1 production steps (None: off)."""
        if J is None:
            _check(lib.g12_set_collision(self.h, None, 0), "g12_set_collision")
            return
        J = _c(J)
        _check(lib.g12_set_collision(self.h, _p(J), J.shape[0]), "g12_set_collision")

    def set_time(self, t, steps):
        _check(lib.g12_set_time(self.h, float(t), int(steps)), "g12_set_time")

    def time(self):
        t = ctypes.c_double(0)
        n = ctypes.c_int64(0)
        lib.g12_get_time(self.h, ctypes.byref(t), ctypes.byref(n))
        return t.value, n.value

    def energy(self):
        """(one-body, Hartree, exchange, correlation) of the current state at full
        U and the quench shift of the current time."""
        E = np.zeros(4)
        _check(lib.g12_state_energy(self.h, _p(E)), "g12_state_energy")
        return E

    def timing(self, on=True):
        lib.g12_timing_enable(self.h, int(on))

    def timing_get(self, which):
        n = ctypes.c_int64(0)
        ms = ctypes.c_double(0)
        _check(lib.g12_timing_get(self.h, which, ctypes.byref(n), ctypes.byref(ms)), "g12_timing_get")
        return n.value, ms.value

    def memory_bytes(self):
        b = ctypes.c_int64(0)
        lib.g12_memory_bytes(self.h, ctypes.byref(b))
        return b.value

    def save(self, path):
        _check(lib.g12_save_state(self.h, path.encode()), "g12_save_state")

    def load(self, path):
        _check(lib.g12_load_state(self.h, path.encode()), "g12_load_state")

    def configure(self, graphs=None, autotune=None):
        """graphs: replay RK4 steps as hipGraphs; autotune: time tile wave counts."""
        g = -1 if graphs is None else int(bool(graphs))
        a = -1 if autotune is None else int(bool(autotune))
        _check(lib.g12_configure(self.h, g, a), "g12_configure")

    def pair_symmetric(self):
        """True when the context runs the pair-symmetric path (one tile-product
        pass per stage; the state is pair-symmetric and Hermitian)."""
        on = ctypes.c_int(0)
        _check(lib.g12_pair_symmetric(self.h, ctypes.byref(on)), "g12_pair_symmetric")
        return bool(on.value)


class Batch:
    """Many independent trajectories advanced together (one set of launches per
    RK4 stage for all of them).  `params` is a list of Params that share the
    time protocol and precision; sizes and physics may differ."""

    def __init__(self, params):
        self.params = list(params)
        n = len(self.params)
        arr = (Params * n)(*self.params)
        h = ctypes.c_void_p()
        _check(lib.g12_batch_create(arr, n, ctypes.byref(h)), "g12_batch_create")
        self.h = h
        self.n = n
        self.N = [q.Ns * q.Nb for q in self.params]

    def close(self):
        if self.h:
            lib.g12_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __len__(self):
        return self.n

    def set_state(self, i, G1, G2xx=None, G2xy=None):
        N = self.N[i]
        G1 = _c(G1, (N, N))
        Gxx = None if G2xx is None else _c(G2xx, (N ** 2, N ** 2))
        Gxy = None if G2xy is None else _c(G2xy, (N ** 2, N ** 2))
        _check(lib.g12_batch_set_state(self.h, i, _p(G1), _p(Gxx), _p(Gxy)), "g12_batch_set_state")

    def get_G1(self, i):
        N = self.N[i]
        G1 = np.zeros((N, N), np.complex128)
        _check(lib.g12_batch_get_state(self.h, i, _p(G1), None, None), "g12_batch_get_state")
        return G1

    def get_state(self, i):
        N = self.N[i]
        G1 = np.zeros((N, N), np.complex128)
        Gxx = np.zeros((N * N, N * N), np.complex128)
        Gxy = np.zeros((N * N, N * N), np.complex128)
        _check(lib.g12_batch_get_state(self.h, i, _p(G1), _p(Gxx), _p(Gxy)), "g12_batch_get_state")
        return G1, Gxx, Gxy

    def adiabatic_switching(self):
        steps = (ctypes.c_int * self.n)()
        _check(lib.g12_batch_adiabatic_switching(self.h, steps), "g12_batch_adiabatic_switching")
        return list(steps)

    def rk4(self, traj_cap=None, record=True):
        """Returns (rc, [trajectory_i[steps_i, N_i, N_i]])."""
        if traj_cap is None:
            traj_cap = self.params[0].time_steps + 2
        trajs = [np.zeros((traj_cap, N, N), np.complex128) if record else None for N in self.N]
        ptrs = (_dp * self.n)(*[_p(t) if t is not None else None for t in trajs])
        steps = (ctypes.c_int64 * self.n)()
        rc = _check(lib.g12_batch_rk4(self.h, ptrs if record else None, traj_cap, steps), "g12_batch_rk4")
        if not record:
            return rc, list(steps)
        return rc, [t[: min(s, traj_cap)] for t, s in zip(trajs, steps)]

    def advance(self, nsteps):
        return _check(lib.g12_batch_rk4_advance(self.h, nsteps), "g12_batch_rk4_advance")

    def record(self, nsteps, out=None, offset=0):
        """nsteps more production steps; G1 before each goes to out[i][offset:]
        (new arrays when out is None).  Returns (rc, out)."""
        if out is None:
            out = [np.zeros((nsteps, N, N), np.complex128) for N in self.N]
            offset = 0
        ptrs = (_dp * self.n)(*[_p(o[offset:]) for o in out])
        rc = _check(lib.g12_batch_rk4_record(self.h, nsteps, ptrs, nsteps), "g12_batch_rk4_record")
        return rc, out

    def record_collision(self, nsteps, out, jout, offset=0):
        """record() that also writes each correlated trajectory's collision
        terms J = I + I^dagger to jout[i][offset:] (jout[i] None: skipped)."""
        ptrs = (_dp * self.n)(*[None if o is None else _p(o[offset:]) for o in out])
        jptrs = (_dp * self.n)(*[None if j is None else _p(j[offset:]) for j in jout])
        rc = _check(lib.g12_batch_rk4_record_collision(self.h, nsteps, ptrs, jptrs, nsteps),
                    "g12_batch_rk4_record_collision")
        return rc

    def guard(self):
        f = (ctypes.c_int * self.n)()
        _check(lib.g12_batch_guard(self.h, f), "g12_batch_guard")
        return list(f)

    def timing(self, on=True):
        lib.g12_batch_timing_enable(self.h, int(on))

    def timing_get(self, which):
        n = ctypes.c_int64(0)
        ms = ctypes.c_double(0)
        _check(lib.g12_batch_timing_get(self.h, which, ctypes.byref(n), ctypes.byref(ms)),
               "g12_batch_timing_get")
        return n.value, ms.value

    def memory_bytes(self):
        b = ctypes.c_int64(0)
        lib.g12_batch_memory_bytes(self.h, ctypes.byref(b))
        return b.value

    def configure(self, graphs=None, autotune=None):
        g = -1 if graphs is None else int(bool(graphs))
        a = -1 if autotune is None else int(bool(autotune))
        _check(lib.g12_batch_configure(self.h, g, a), "g12_batch_configure")

    def tune(self):
        _check(lib.g12_batch_tune(self.h), "g12_batch_tune")

    def pair_symmetric(self):
        """Propagator.pair_symmetric of the batch (every correlated trajectory)."""
        on = ctypes.c_int(0)
        _check(lib.g12_batch_pair_symmetric(self.h, ctypes.byref(on)), "g12_batch_pair_symmetric")
        return bool(on.value)

    def save(self, i, path):
        """Binary checkpoint of trajectory i (with the batch's time and step)."""
        _check(lib.g12_batch_save_state(self.h, i, path.encode()), "g12_batch_save_state")

    def load(self, i, path):
        """Loads trajectory i; the batch's time and step become the file's."""
        _check(lib.g12_batch_load_state(self.h, i, path.encode()), "g12_batch_load_state")

    def energy(self, i):
        """Propagator.energy of trajectory i."""
        E = np.zeros(4)
        _check(lib.g12_batch_state_energy(self.h, i, _p(E)), "g12_batch_state_energy")
        return E

    def buckets(self):
        """[(NP, trajectories, waves_strided, waves_fused)] of the tile passes."""
        cap = 8
        a, b, c, d = [(ctypes.c_int * cap)() for _ in range(4)]
        n = _check(lib.g12_batch_buckets(self.h, cap, a, b, c, d), "g12_batch_buckets")
        return [(a[k], b[k], c[k], d[k]) for k in range(n)]


def synchronize():
    """Device-wide synchronisation of the library's GPU."""
    _check(lib.g12_synchronize(), "g12_synchronize")


def gpu_available():
    try:
        import torch  # noqa: F401  (device enumeration only)
        return torch.cuda.device_count() > 0
    except Exception:
        return False
