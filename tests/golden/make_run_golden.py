"""Full-run golden trajectories from the CPU oracle (build container only).

config1: Ns=4, Nb=2 (N=8), onsite U=1.0, t1=1, t2=0.5, AS (rate 0.5,
         midpoint 10), full quench 1.0 on 2 sites (rate 2, midpoint 25),
         dt=0.02, 3000 steps; element-wise oracle (oracle/g12_oracle.cpp).
config2: Ns=16, Nb=2 (N=32), EHM with V_n=U/(1+n), U=2.0, t1=1, t2=0.3, AS
         (0.5, 10), pulse quench 1.5 on 4 sites (rate 2, midpoints 30/60),
         dt=0.01, 10000 steps; factorised oracle (oracle/g12_fast.cpp), which
         the CPU suite checks against the element-wise one.
G1(0) is the numpy eigh projector on the Ns/2 lowest states of h(U=0) (the
config-1 Fermi level is degenerate, so G1(0) is part of the fixture).
Usage:  python tests/golden/make_run_golden.py config1|config2
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle.oracle import Oracle  # noqa: E402

CONFIGS = {
    "config1": dict(Ns=4, Nb=2, t1=1.0, t2=0.5, U=1.0, dt=0.02, time_steps=3000, quench_strength=1.0, Nq=2,
                    q_type="full", HF=False, AS_rate=0.5, AS_midpoint=10.0, quench_rate=2.0,
                    quench_on_midpoint=25.0, quench_off_midpoint=55.0, EHM=False),
    "config2": dict(Ns=16, Nb=2, t1=1.0, t2=0.3, U=2.0, dt=0.01, time_steps=10000, quench_strength=1.5, Nq=4,
                    q_type="pulse", HF=False, AS_rate=0.5, AS_midpoint=10.0, quench_rate=2.0,
                    quench_on_midpoint=30.0, quench_off_midpoint=60.0, EHM=True),
}


def g0(cfg):
    Ns, Nb = cfg["Ns"], cfg["Nb"]
    N = Ns * Nb
    h = np.zeros((N, N))
    for X in range(N):
        i, a = X % Ns, X // Ns
        if i + 1 < Ns:
            h[X, X + 1] = h[X + 1, X] = -cfg["t1"]
        if a + 1 < Nb:
            h[X, X + Ns] = h[X + Ns, X] = -cfg["t2"]
    _, v = np.linalg.eigh(h)
    p = Ns // 2
    return 1j * v[:, :p] @ v[:, :p].conj().T


def v_profile(cfg):
    if not cfg["EHM"]:
        return np.zeros(0)
    d = np.arange(1, cfg["Ns"], dtype=float)
    return 1.0 / (1.0 + d)


def dipole_from_diag(diag):
    N = diag.shape[1]
    p = np.zeros(diag.shape[0])
    for i in range(N // 2):
        p += 2 * (N / 2 - i - 0.5) * (diag[:, i] - diag[:, N - 1 - i]) / N
    return p


def main(name):
    o = Oracle()
    cfg = CONFIGS[name]
    N = cfg["Ns"] * cfg["Nb"]
    G1 = g0(cfg)
    Z = np.zeros((N * N, N * N), np.complex128)
    V = v_profile(cfg)
    t0 = time.time()
    if name == "config1":
        G1a, Gxx, Gxy, nas = o.adiabatic_switching(cfg, G1, Z, Z, V_profile=V)
        rc, G1e, Gxxe, Gxye, traj = o.rk4(cfg, G1a, Gxx, Gxy, V_profile=V)
        diag = np.array([np.diag(t).imag for t in traj])
        out = dict(G1_0=G1, G1_as=G1a, G2xx_as=Gxx, G2xy_as=Gxy, as_steps=nas, rk_steps=len(traj), rc=rc,
                   G1_end=G1e, G2xx_end=Gxxe, G2xy_end=Gxye, traj_every10=traj[::10], dipole=o.dipole(traj),
                   diag=diag)
    else:
        f = o.fast_run(cfg, G1, Z, Z, V_profile=V, traj_every=500, traj_cap=64, record_diag=True)
        rng = np.random.default_rng(0)
        idx = rng.choice(N ** 4, 2000, replace=False)
        out = dict(G1_0=G1, as_steps=f["as_steps"], rk_steps=f["rk_steps"], rc=f["rc"], G1_end=f["G1"],
                   traj_every500=f["traj"], dipole=dipole_from_diag(f["diag"]), diag_every10=f["diag"][::10],
                   G2_idx=idx, G2xx_sample=f["Gxx"].reshape(-1)[idx], G2xy_sample=f["Gxy"].reshape(-1)[idx],
                   G2xx_norm=np.linalg.norm(f["Gxx"]), G2xy_norm=np.linalg.norm(f["Gxy"]))
    out["seconds"] = time.time() - t0
    np.savez_compressed(os.path.join(HERE, f"run_{name}.npz"), **out)
    print(name, "done in", out["seconds"], "s; AS", out["as_steps"], "RK", out["rk_steps"], "rc", out["rc"])


if __name__ == "__main__":
    main(sys.argv[1])
