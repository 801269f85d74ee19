"""CPU suite: pins of the oracle's drivers (adiabatic switching and production
RK4) that the routine-level golden vectors leave open.

* the factorised C++ oracle (oracle/g12_fast.cpp, ``Oracle.fast_run``: the
  reference for the N = 32 golden run and the N = 64 GPU parity tests) against
  the element-wise oracle's drivers (``orc_adiabatic_switching`` +
  ``orc_rk4``), with long-range shells on and off, full and pulse quenches
  crossing the gate, Nb = 1..3;
* the element-wise RK4 driver against a textbook classical RK4 written here
  from its definition (stage times t, t + dt/2, t + dt/2, t + dt; weights
  1, 2, 2, 1): at U = 0 with a tight-binding h(t) built here from the hopping
  rules (no oracle routine involved) and a gated pulse quench; and in
  Hartree-Fock mode, with the stage Hamiltonian from the golden-pinned
  ``Oracle.hf`` and the switching strength U_AS(t) = fermi(t) U evaluated per
  stage.  These pin the stage times of RK_solver.h:120-151 and AS.h:111-141.
* planted mistakes (a stage time moved to the step start, the quench gate
  dropped) make the textbook driver disagree with the oracle.
"""
import numpy as np
import pytest

from oracle.oracle import Oracle
from tests.helpers import physical_g2

O = Oracle()


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def cfg_of(**kw):
    base = dict(Ns=3, Nb=2, t1=1.0, t2=0.5, U=1.1, dt=0.02, time_steps=30, quench_strength=0.8, Nq=1,
                q_type="full", HF=False, AS_rate=0.05, AS_midpoint=0.1, quench_rate=0.05,
                quench_on_midpoint=0.2, quench_off_midpoint=0.45, AS_dt=0.1, AS_tail=3, gate_time=0.15)
    base.update(kw)
    return base


def start_state(rng, N, filled):
    X = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    _, v = np.linalg.eigh(X + X.conj().T)
    return 1j * v[:, :filled] @ v[:, :filled].conj().T


# (Ns, Nb, long-range profile, quench type)
FAST_CASES = [(4, 1, [], "full"), (4, 1, [0.5, 0.25], "pulse"), (3, 2, [0.5], "full"),
              (3, 2, [0.6, 0.2], "pulse"), (2, 3, [0.4], "pulse"), (2, 3, [], "full")]


@pytest.mark.parametrize("Ns,Nb,prof,qt", FAST_CASES)
def test_fast_run_matches_elementwise_drivers(Ns, Nb, prof, qt):
    """Oracle.fast_run (switching + production through the partial-trace
    tables) equals the element-wise drivers over the whole protocol, every
    recorded G1 and the final G2 included."""
    N = Ns * Nb
    rng = np.random.default_rng(100 + 10 * Ns + Nb + len(prof))
    cfg = cfg_of(Ns=Ns, Nb=Nb, q_type=qt, Nq=max(1, Ns // 2))
    G0 = start_state(rng, N, N // 2)
    Gxx, Gxy = physical_g2(rng, Ns, Nb, 1e-2), physical_g2(rng, Ns, Nb, 1e-2)
    G1a, Gxxa, Gxya, nas = O.adiabatic_switching(cfg, G0, Gxx, Gxy, V_profile=prof)
    rc, G1e, Gxxe, Gxye, traj = O.rk4(cfg, G1a, Gxxa, Gxya, V_profile=prof)
    f = O.fast_run(cfg, G0, Gxx, Gxy, V_profile=prof, run_as=True, traj_every=1, traj_cap=cfg["time_steps"] + 2)
    assert rc == f["rc"] == 0
    assert nas == f["as_steps"] and len(traj) == f["rk_steps"] == cfg["time_steps"]
    assert np.abs(Gxxe).max() > 1e-3  # the correlated part moved
    assert rel(f["traj"], traj) < 1e-13
    assert rel(f["G1"], G1e) < 1e-13
    assert rel(f["Gxx"], Gxxe) < 1e-13 and rel(f["Gxy"], Gxye) < 1e-13


# ---- textbook classical RK4 ---------------------------------------------------

def fermi(t, t0, tau):
    return 1.0 / (1.0 + np.exp((t0 - t) / tau))


def tight_binding(Ns, Nb, t1, t2, eps=None):
    """h0 of the chain: -t1 between neighbouring sites in one orbital, -t2
    between adjacent orbitals on one site, orbital energies on the diagonal;
    one-particle index X = site + orb * Ns."""
    N = Ns * Nb
    h = np.zeros((N, N), complex)
    for o in range(Nb):
        for s in range(Ns):
            X = s + o * Ns
            h[X, X] = 0.0 if eps is None else eps[o]
            if s + 1 < Ns:
                h[X, X + 1] = h[X + 1, X] = -t1
            if o + 1 < Nb:
                h[X, X + Ns] = h[X + Ns, X] = -t2
    return h


def quench_amp(cfg, ts):
    """Shift on the quenched sites at stage time ts: zero before the gate."""
    if ts < cfg["gate_time"]:
        return 0.0, 0
    on = fermi(ts, cfg["quench_on_midpoint"], cfg["quench_rate"])
    if cfg["q_type"] == "pulse":
        on = on * (1 - fermi(ts, cfg["quench_off_midpoint"], cfg["quench_rate"]))
    return on * cfg["quench_strength"], cfg["Nq"]


def rk4_textbook(f, y, t, dt):
    k1 = f(t, y)
    k2 = f(t + dt / 2, y + dt / 2 * k1)
    k3 = f(t + dt / 2, y + dt / 2 * k2)
    k4 = f(t + dt, y + dt * k3)
    return y + dt / 6 * (k1 + 2 * k2 + 2 * k3 + k4)


def production_textbook(cfg, f, y):
    """while (t < time_steps * dt): record, step, t += dt."""
    traj, t, tmax = [], 0.0, cfg["time_steps"] * cfg["dt"]
    while t < tmax:
        traj.append(y.copy())
        y = rk4_textbook(f, y, t, cfg["dt"])
        t += cfg["dt"]
    return y, np.array(traj)


def test_rk4_driver_matches_textbook_at_u0_with_gated_pulse():
    """U = 0: dG1/dt = -i [h0 + w(t) P_q, G1] with the pulse ramp switched on
    at the gate, which the run crosses."""
    Ns, Nb = 4, 2
    cfg = cfg_of(Ns=Ns, Nb=Nb, U=0.0, q_type="pulse", Nq=2, time_steps=40, dt=0.02, gate_time=0.31,
                 quench_on_midpoint=0.3, quench_off_midpoint=0.6, quench_rate=0.08, quench_strength=1.3)
    N = Ns * Nb
    h0 = tight_binding(Ns, Nb, cfg["t1"], cfg["t2"])
    site = np.arange(N) % Ns

    def f(t, G):
        w, nq = quench_amp(cfg, t)
        h = h0 + np.diag(np.where(site < nq, w, 0.0))
        return -1j * (h @ G - G @ h)

    G0 = start_state(np.random.default_rng(5), N, N // 2)
    Z = np.zeros((N * N, N * N), complex)
    rc, G1, _, _, traj = O.rk4(cfg, G0, Z, Z)
    want_G1, want_traj = production_textbook(cfg, f, G0)
    assert rc == 0 and len(traj) == len(want_traj) == 40
    assert rel(traj, want_traj) < 1e-13 and rel(G1, want_G1) < 1e-13
    # the gate is crossed inside the run and the ramp is not negligible there
    assert quench_amp(cfg, 0.3)[0] == 0.0 and quench_amp(cfg, 0.4)[0] > 0.5
    # planted mistakes: every stage at the step start, or no gate, disagree
    def f_start(t_stage_fn):
        def step(y, t, dt):
            k1 = f(t, y)
            k2 = f(t_stage_fn(t, dt), y + dt / 2 * k1)
            k3 = f(t_stage_fn(t, dt), y + dt / 2 * k2)
            k4 = f(t + dt, y + dt * k3)
            return y + dt / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
        y, t = G0.copy(), 0.0
        for _ in range(40):
            y = step(y, t, cfg["dt"])
            t += cfg["dt"]
        return y
    assert rel(f_start(lambda t, dt: t), G1) > 1e-9
    cfg_nogate = dict(cfg, gate_time=-1.0)
    _, G1_nogate, _, _, _ = O.rk4(cfg_nogate, G0, Z, Z)
    assert rel(G1_nogate, G1) > 1e-6


def test_drivers_match_textbook_in_hartree_fock_mode():
    """HF = true (G2 frozen, no collision term): dG1/dt = -i [h_HF[G1](t), G1]
    with h_HF from the golden-pinned Oracle.hf.  Switching: AS_dt = 0.1,
    U_AS(t) = fermi(t; mid, rate) U at each stage time (long-range shells
    scaled alike), no quench, 304 steps at (rate 0.5, midpoint 10).
    Production: full U, the quench gated at the stage time."""
    Ns, Nb = 3, 2
    N = Ns * Nb
    prof = np.array([0.5, 0.3])
    eps = np.zeros(Nb)
    cfg = cfg_of(Ns=Ns, Nb=Nb, U=1.0, HF=True, AS_rate=0.5, AS_midpoint=10.0, AS_tail=20, q_type="pulse",
                 Nq=2, time_steps=60, dt=0.02, gate_time=0.5, quench_on_midpoint=0.6, quench_off_midpoint=0.9,
                 quench_rate=0.1, quench_strength=1.1)

    def f_of(U_at, quench_at):
        def f(t, G):
            U = U_at(t)
            w, nq = quench_at(t)
            h = O.hf(np.zeros((N, N)), G, Ns, Nb, U, cfg["t1"], cfg["t2"], w, eps, U * prof, nq)
            return -1j * (h @ G - G @ h)
        return f

    f_as = f_of(lambda t: fermi(t, cfg["AS_midpoint"], cfg["AS_rate"]) * cfg["U"], lambda t: (0.0, 0))
    f_rk = f_of(lambda t: cfg["U"], lambda t: quench_amp(cfg, t))
    G0 = start_state(np.random.default_rng(11), N, 3)
    y, t = G0.copy(), 0.0
    for _ in range(304):  # switching length at (0.5, 10), pinned by the loop-count known answers
        y = rk4_textbook(f_as, y, t, cfg["AS_dt"])
        t += cfg["AS_dt"]
    Z = np.zeros((N * N, N * N), complex)
    G1a, _, _, nas = O.adiabatic_switching(cfg, G0, Z, Z, V_profile=prof)
    assert nas == 304
    assert rel(G1a, y) < 1e-12
    y, traj_want = production_textbook(cfg, f_rk, G1a)
    rc, G1e, _, _, traj = O.rk4(cfg, G1a, Z, Z, V_profile=prof)
    assert rc == 0 and len(traj) == len(traj_want)
    assert rel(traj, traj_want) < 1e-12 and rel(G1e, y) < 1e-12
    # the switching strength is evaluated per stage: a step-constant U_AS is caught
    y2, t = G0.copy(), 0.0
    for _ in range(304):
        u = fermi(t, cfg["AS_midpoint"], cfg["AS_rate"]) * cfg["U"]
        y2 = rk4_textbook(f_of(lambda _t, u=u: u, lambda _t: (0.0, 0)), y2, t, cfg["AS_dt"])
        t += cfg["AS_dt"]
    assert rel(y2, G1a) > 1e-8
