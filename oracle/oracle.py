"""ctypes loaders for the CPU oracle (test infrastructure only).

`liboracle.so` is the project's C++ restatement of the reference routines
(oracle/g12_oracle.cpp).  `libg12ref.so` (optional) wraps the reference's own
routine headers, compiled unmodified from /root/reference by oracle/Makefile;
it exists only where the reference sources were present at build time.

Only tests, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg may use
this module, and only as the checker.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "_build", "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libg12ref.so")

_dp = ctypes.POINTER(ctypes.c_double)


def _ptr(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("Ns", ctypes.c_int), ("Nb", ctypes.c_int),
        ("t1", ctypes.c_double), ("t2", ctypes.c_double), ("U", ctypes.c_double),
        ("dt", ctypes.c_double), ("time_steps", ctypes.c_int),
        ("quench_strength", ctypes.c_double), ("Nq", ctypes.c_int),
        ("q_pulse", ctypes.c_int), ("hf_only", ctypes.c_int),
        ("AS_rate", ctypes.c_double), ("AS_midpoint", ctypes.c_double),
        ("quench_rate", ctypes.c_double), ("quench_on", ctypes.c_double),
        ("quench_off", ctypes.c_double),
        ("AS_dt", ctypes.c_double), ("AS_tail", ctypes.c_int),
        ("gate_time", ctypes.c_double),
        ("nV", ctypes.c_int), ("V_profile", _dp), ("eps", _dp),
    ]


def _c(a):
    """complex128 C-contiguous view as the interleaved double buffer."""
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a


class Oracle:
    """The project's CPU restatement (reference layout, complex128)."""

    def __init__(self, path=ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        L = ctypes.CDLL(path)
        self.L = L
        d, i, p = ctypes.c_double, ctypes.c_int, _dp
        L.orc_hf.argtypes = [p, p, i, i, d, d, d, d, p, p, i, i]
        L.orc_comm1.argtypes = [p, p, p, i]
        L.orc_comm2.argtypes = [p, p, p, p, i, i]
        L.orc_collision.argtypes = [p, p, p, i, i, d, p, i]
        L.orc_psi.argtypes = [p, p, i, i, d, p, i, i]
        L.orc_pi.argtypes = [p, p, p, p, d, i, i, p, i]
        L.orc_g1_step.argtypes = [p, p, p, p, d, i, i]
        L.orc_g2_step.argtypes = [p, p, p, p, p, d, i, i]
        L.orc_fermi.argtypes = [d, d, d]
        L.orc_fermi.restype = d
        L.orc_dipole.argtypes = [p, p, ctypes.c_int64, i]
        L.orc_adiabatic_switching.argtypes = [ctypes.POINTER(OrcParams), p, p, p,
                                              ctypes.POINTER(ctypes.c_int)]
        L.orc_rk4.argtypes = [ctypes.POINTER(OrcParams), p, p, p, p, ctypes.c_int64,
                              ctypes.POINTER(ctypes.c_int64)]
        L.orc_rk4_fixed_step.argtypes = [ctypes.POINTER(OrcParams), p, p, p, d]
        L.orc_rk4_partial_step.argtypes = [ctypes.POINTER(OrcParams), p, p, p, d, i]
        L.orcf_run.argtypes = [ctypes.POINTER(OrcParams), p, p, p, i, ctypes.POINTER(i),
                               ctypes.POINTER(ctypes.c_int64), p, ctypes.c_int64, i, p,
                               ctypes.c_int64]

    # --- routines (same argument roles and order as the reference) ---
    def hf(self, h, G1, Ns, Nb, U, t1, t2, w, eps, V, quench_extent):
        h = _c(h).copy()
        eps = np.ascontiguousarray(eps, dtype=np.float64)
        V = np.ascontiguousarray(V, dtype=np.float64)
        self.L.orc_hf(_ptr(h), _ptr(_c(G1)), Ns, Nb, U, t1, t2, w, _ptr(eps),
                      _ptr(V) if V.size else None, V.size, quench_extent)
        return h

    def comm1(self, G, h):
        N = G.shape[0]
        out = np.zeros((N, N), np.complex128)
        self.L.orc_comm1(_ptr(_c(G)), _ptr(_c(h)), _ptr(out), N)
        return out

    def comm2(self, G2, hx, hy, Ns, Nb):
        out = np.zeros_like(_c(G2))
        self.L.orc_comm2(_ptr(_c(G2)), _ptr(_c(hx)), _ptr(_c(hy)), _ptr(out), Ns, Nb)
        return out

    def collision(self, P, Q, Ns, Nb, U, V):
        N = Ns * Nb
        out = np.zeros((N, N), np.complex128)
        V = np.ascontiguousarray(V, dtype=np.float64)
        self.L.orc_collision(_ptr(out), _ptr(_c(P)), _ptr(_c(Q)), Ns, Nb, U,
                             _ptr(V) if V.size else None, V.size)
        return out

    def psi(self, G1, Ns, Nb, U, V, same_spin):
        N = Ns * Nb
        out = np.zeros((N * N, N * N), np.complex128)
        V = np.ascontiguousarray(V, dtype=np.float64)
        self.L.orc_psi(_ptr(out), _ptr(_c(G1)), Ns, Nb, U, _ptr(V) if V.size else None,
                       V.size, int(same_spin))
        return out

    def pi(self, G1, P, Q, U, Ns, Nb, V):
        N = Ns * Nb
        out = np.zeros((N * N, N * N), np.complex128)
        V = np.ascontiguousarray(V, dtype=np.float64)
        self.L.orc_pi(_ptr(out), _ptr(_c(G1)), _ptr(_c(P)), _ptr(_c(Q)), U, Ns, Nb,
                      _ptr(V) if V.size else None, V.size)
        return out

    def g1_step(self, comm, I, K, RK, w, Ns, Nb):
        K = _c(K).copy()
        RK = _c(RK).copy()
        self.L.orc_g1_step(_ptr(_c(comm)), _ptr(_c(I)), _ptr(K), _ptr(RK), w, Ns, Nb)
        return K, RK

    def g2_step(self, comm, psi, pi, K, RK, w, Ns, Nb):
        K = _c(K).copy()
        RK = _c(RK).copy()
        self.L.orc_g2_step(_ptr(_c(comm)), _ptr(_c(psi)), _ptr(_c(pi)), _ptr(K), _ptr(RK),
                           w, Ns, Nb)
        return K, RK

    def fermi(self, t, t0, tau):
        return self.L.orc_fermi(t, t0, tau)

    def dipole(self, traj):
        traj = _c(traj)
        T, N = traj.shape[0], traj.shape[1]
        p = np.zeros(T)
        self.L.orc_dipole(_ptr(p), _ptr(traj), T, N)
        return p

    # --- drivers ---
    @staticmethod
    def params(cfg, V_profile=None, eps=None):
        """OrcParams from a run-configuration dict (g12kba_amd.config keys)."""
        keep = []
        p = OrcParams()
        p.Ns, p.Nb = cfg["Ns"], cfg["Nb"]
        p.t1, p.t2, p.U = cfg["t1"], cfg["t2"], cfg["U"]
        p.dt, p.time_steps = cfg["dt"], cfg["time_steps"]
        p.quench_strength, p.Nq = cfg["quench_strength"], cfg["Nq"]
        p.q_pulse = 1 if cfg["q_type"] == "pulse" else 0
        p.hf_only = int(bool(cfg["HF"]))
        p.AS_rate, p.AS_midpoint = cfg["AS_rate"], cfg["AS_midpoint"]
        p.quench_rate = cfg["quench_rate"]
        p.quench_on, p.quench_off = cfg["quench_on_midpoint"], cfg["quench_off_midpoint"]
        p.AS_dt, p.AS_tail = cfg.get("AS_dt", 0.1), cfg.get("AS_tail", 20)
        p.gate_time = cfg.get("gate_time", 50.0)
        if V_profile is not None and len(V_profile):
            v = np.ascontiguousarray(V_profile, dtype=np.float64)
            keep.append(v)
            p.nV, p.V_profile = v.size, _ptr(v)
        else:
            p.nV, p.V_profile = 0, None
        if eps is not None:
            e = np.ascontiguousarray(eps, dtype=np.float64)
            keep.append(e)
            p.eps = _ptr(e)
        else:
            p.eps = None
        return p, keep

    def adiabatic_switching(self, cfg, G1, Gxx, Gxy, V_profile=None, eps=None):
        p, keep = self.params(cfg, V_profile, eps)
        G1, Gxx, Gxy = _c(G1).copy(), _c(Gxx).copy(), _c(Gxy).copy()
        n = ctypes.c_int(0)
        self.L.orc_adiabatic_switching(ctypes.byref(p), _ptr(G1), _ptr(Gxx), _ptr(Gxy),
                                       ctypes.byref(n))
        return G1, Gxx, Gxy, n.value

    def rk4(self, cfg, G1, Gxx, Gxy, V_profile=None, eps=None, traj_cap=None):
        p, keep = self.params(cfg, V_profile, eps)
        N = cfg["Ns"] * cfg["Nb"]
        G1, Gxx, Gxy = _c(G1).copy(), _c(Gxx).copy(), _c(Gxy).copy()
        if traj_cap is None:
            traj_cap = cfg["time_steps"] + 2
        traj = np.zeros((traj_cap, N, N), np.complex128)
        n = ctypes.c_int64(0)
        rc = self.L.orc_rk4(ctypes.byref(p), _ptr(G1), _ptr(Gxx), _ptr(Gxy), _ptr(traj),
                            traj_cap, ctypes.byref(n))
        return rc, G1, Gxx, Gxy, traj[: min(n.value, traj_cap)]

    def partial_step(self, cfg, G1, Gxx, Gxy, dt, nstages, V_profile=None):
        p, keep = self.params(cfg, V_profile)
        G1, Gxx, Gxy = _c(G1).copy(), _c(Gxx).copy(), _c(Gxy).copy()
        self.L.orc_rk4_partial_step(ctypes.byref(p), _ptr(G1), _ptr(Gxx), _ptr(Gxy), dt, nstages)
        return G1, Gxx, Gxy

    def fast_run(self, cfg, G1, Gxx, Gxy, V_profile=None, eps=None, run_as=True, traj_every=0,
                 traj_cap=0, record_diag=False):
        """Factorised C++ oracle (oracle/g12_fast.cpp): AS (optional) + RK4.

        Returns dict(rc, G1, Gxx, Gxy, as_steps, rk_steps, traj, diag) with diag
        the Im G1[i,i] before every RK4 iteration (if record_diag)."""
        p, keep = self.params(cfg, V_profile, eps)
        N = cfg["Ns"] * cfg["Nb"]
        G1, Gxx, Gxy = _c(G1).copy(), _c(Gxx).copy(), _c(Gxy).copy()
        traj = np.zeros((max(traj_cap, 1), N, N), np.complex128)
        cap = cfg["time_steps"] + 2
        diag = np.zeros((cap, N)) if record_diag else None
        a = ctypes.c_int(0)
        n = ctypes.c_int64(0)
        rc = self.L.orcf_run(ctypes.byref(p), _ptr(G1), _ptr(Gxx), _ptr(Gxy), int(run_as), ctypes.byref(a),
                             ctypes.byref(n), _ptr(traj) if traj_cap else None, traj_cap, traj_every,
                             _ptr(diag) if record_diag else None, cap if record_diag else 0)
        nrec = 0 if not traj_every else min(traj_cap, (n.value + traj_every - 1) // traj_every)
        return dict(rc=rc, G1=G1, Gxx=Gxx, Gxy=Gxy, as_steps=a.value, rk_steps=n.value, traj=traj[:nrec],
                    diag=None if diag is None else diag[: n.value])

    def fixed_step(self, cfg, G1, Gxx, Gxy, dt):
        p, keep = self.params(cfg)
        G1, Gxx, Gxy = _c(G1).copy(), _c(Gxx).copy(), _c(Gxy).copy()
        self.L.orc_rk4_fixed_step(ctypes.byref(p), _ptr(G1), _ptr(Gxx), _ptr(Gxy), dt)
        return G1, Gxx, Gxy


class RefInputs(ctypes.Structure):
    _fields_ = [
        ("HF", ctypes.c_int), ("EHM", ctypes.c_int),
        ("t1", ctypes.c_double), ("t2", ctypes.c_double), ("t3", ctypes.c_double),
        ("U", ctypes.c_double), ("time_steps", ctypes.c_int),
        ("dt_fixed", ctypes.c_double), ("quench_strength", ctypes.c_double),
        ("Ns", ctypes.c_int), ("Nq", ctypes.c_int), ("Nb", ctypes.c_int),
        ("q_type", ctypes.c_int),
        ("AS_rate", ctypes.c_double), ("AS_midpoint", ctypes.c_double),
        ("quench_rate", ctypes.c_double), ("quench_on_midpoint", ctypes.c_double),
        ("quench_off_midpoint", ctypes.c_double),
    ]


class Ref:
    """The reference's own routine headers behind a C harness (optional)."""

    def __init__(self, path=REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = ctypes.CDLL(path)
        self.L = L
        d, i, p = ctypes.c_double, ctypes.c_int, _dp
        L.ref_hf.argtypes = [p, p, i, i, d, d, d, d, p, p, i, i]
        L.ref_collision.argtypes = [p, p, p, i, i, d, p, i]
        L.ref_psi.argtypes = [p, p, i, i, d, p, i, i]
        L.ref_pi.argtypes = [p, p, p, p, d, i, i, p, i]
        L.ref_g1_step.argtypes = [p, p, p, p, d, i, i]
        L.ref_g2_step.argtypes = [p, p, p, p, p, d, i, i]
        L.ref_fermi.argtypes = [d, d, d]
        L.ref_fermi.restype = d
        L.ref_dipole.argtypes = [p, p, ctypes.c_long, i]
        L.ref_parse.argtypes = [ctypes.c_char_p, ctypes.POINTER(RefInputs)]
        L.ref_write.argtypes = [ctypes.c_char_p, p, ctypes.c_long, i, p]

    def hf(self, h, G1, Ns, Nb, U, t1, t2, w, eps, V, quench_extent):
        h = _c(h).copy()
        eps = np.ascontiguousarray(eps, dtype=np.float64)
        V = np.ascontiguousarray(V, dtype=np.float64)
        self.L.ref_hf(_ptr(h), _ptr(_c(G1)), Ns, Nb, U, t1, t2, w, _ptr(eps), _ptr(V),
                      V.size, quench_extent)
        return h

    def collision(self, P, Q, Ns, Nb, U, V):
        N = Ns * Nb
        out = np.zeros((N, N), np.complex128)
        V = np.ascontiguousarray(V, dtype=np.float64)
        self.L.ref_collision(_ptr(out), _ptr(_c(P)), _ptr(_c(Q)), Ns, Nb, U, _ptr(V), V.size)
        return out

    def psi(self, G1, Ns, Nb, U, V, same_spin):
        N = Ns * Nb
        out = np.zeros((N * N, N * N), np.complex128)
        V = np.ascontiguousarray(V, dtype=np.float64)
        self.L.ref_psi(_ptr(out), _ptr(_c(G1)), Ns, Nb, U, _ptr(V), V.size, int(same_spin))
        return out

    def pi(self, G1, P, Q, U, Ns, Nb, V):
        N = Ns * Nb
        out = np.zeros((N * N, N * N), np.complex128)
        V = np.ascontiguousarray(V, dtype=np.float64)
        self.L.ref_pi(_ptr(out), _ptr(_c(G1)), _ptr(_c(P)), _ptr(_c(Q)), U, Ns, Nb, _ptr(V),
                      V.size)
        return out

    def g1_step(self, comm, I, K, RK, w, Ns, Nb):
        K = _c(K).copy()
        RK = _c(RK).copy()
        self.L.ref_g1_step(_ptr(_c(comm)), _ptr(_c(I)), _ptr(K), _ptr(RK), w, Ns, Nb)
        return K, RK

    def g2_step(self, comm, psi, pi, K, RK, w, Ns, Nb):
        K = _c(K).copy()
        RK = _c(RK).copy()
        self.L.ref_g2_step(_ptr(_c(comm)), _ptr(_c(psi)), _ptr(_c(pi)), _ptr(K), _ptr(RK),
                           w, Ns, Nb)
        return K, RK

    def fermi(self, t, t0, tau):
        return self.L.ref_fermi(t, t0, tau)

    def dipole(self, traj):
        traj = _c(traj)
        T, N = traj.shape[0], traj.shape[1]
        p = np.zeros(T)
        self.L.ref_dipole(_ptr(p), _ptr(traj), T, N)
        return p

    def parse(self, directory):
        out = RefInputs()
        rc = self.L.ref_parse(directory.encode(), ctypes.byref(out))
        return rc, {f: getattr(out, f) for f, _ in RefInputs._fields_}

    def write(self, directory, traj, p):
        traj = _c(traj)
        p = np.ascontiguousarray(p, dtype=np.float64)
        return self.L.ref_write(directory.encode(), _ptr(traj), traj.shape[0], traj.shape[1],
                                _ptr(p))
