"""Writes the routine-level golden vectors from the reference's own headers.

Runs in the build container only (needs oracle/_ref/libg12ref.so, compiled
from /root/reference by oracle/Makefile).  Inputs are seeded random arrays;
outputs come from the reference routines psi.h, pi.h, collision_integral.h,
HF.h, G1_G2_t_step.h, fermi_func.h, dipole.h and the reference parser and
writers.  Usage:  python tests/golden/make_golden.py
"""
import os
import shutil
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle.oracle import Ref  # noqa: E402

GEOMS = [(2, 1, 0), (3, 2, 2), (2, 3, 1), (4, 2, 3)]


def rc(rng, *s):
    return rng.standard_normal(s) + 1j * rng.standard_normal(s)


def main():
    r = Ref()
    for Ns, Nb, nV in GEOMS:
        rng = np.random.default_rng(1000 + 100 * Ns + 10 * Nb + nV)
        N = Ns * Nb
        G1, h0 = rc(rng, N, N), rc(rng, N, N)
        P, Q = rc(rng, N * N, N * N), rc(rng, N * N, N * N)
        V, eps = rng.uniform(0.1, 1.0, nV), rng.standard_normal(Nb)
        U, t1, t2, w, nq = 1.3, 1.0, 0.4, 0.7, 1
        c1, I1, K1, R1 = (rc(rng, N, N) for _ in range(4))
        c2, s2, p2, K2, R2 = (rc(rng, N * N, N * N) for _ in range(5))
        out = dict(Ns=Ns, Nb=Nb, V=V, eps=eps, U=U, t1=t1, t2=t2, w=w, nq=nq, G1=G1, h0=h0, P=P, Q=Q,
                   c1=c1, I1=I1, K1=K1, R1=R1, c2=c2, s2=s2, p2=p2, K2=K2, R2=R2)
        out["hf"] = r.hf(h0, G1, Ns, Nb, U, t1, t2, w, eps, V, nq)
        out["collision"] = r.collision(P, Q, Ns, Nb, U, V)
        out["psi_xxxx"] = r.psi(G1, Ns, Nb, U, V, 1)
        out["psi_xyxy"] = r.psi(G1, Ns, Nb, U, V, 0)
        out["pi_PQ"] = r.pi(G1, P, Q, U, Ns, Nb, V)
        out["pi_QP"] = r.pi(G1, Q, P, U, Ns, Nb, V)
        out["g1_K"], out["g1_R"] = r.g1_step(c1, I1, K1, R1, 2.0, Ns, Nb)
        out["g2_K"], out["g2_R"] = r.g2_step(c2, s2, p2, K2, R2, 2.0, Ns, Nb)
        traj = rc(rng, 7, N, N)
        out["traj"] = traj
        out["dipole"] = r.dipole(traj)
        np.savez_compressed(os.path.join(HERE, f"routines_{Ns}x{Nb}_v{nV}.npz"), **out)
    ts = np.array([-3.0, 0.0, 9.7, 10.0, 25.3, 60.0, 1e3])
    np.savez_compressed(os.path.join(HERE, "fermi.npz"), t=ts,
                        f=np.array([[r.fermi(t, t0, tau) for t in ts] for t0, tau in ((10, 0.5), (25, 3), (30, 2))]))
    # parser: the reference's sample input plus edge-case variants
    cases = {}
    sample = os.path.join(HERE, "input_sample")
    variants = {
        "sample": open(sample).read(),
        "pulse_ehm": open(sample).read().replace("full", "pulse").replace("EHM\nfalse", "EHM\ntrue"),
        "none": open(sample).read().replace("full", "none"),
        "missing_U": open(sample).read().replace("U\n.5\n", ""),
        "no_trailing_newline": open(sample).read().rstrip("\n").replace("U\n.5\n", ""),
        "bad_flags": open(sample).read().replace("HF\nfalse", "HF\nyes").replace("full", "half"),
    }
    d = tempfile.mkdtemp()
    try:
        for name, text in variants.items():
            with open(os.path.join(d, "input"), "w") as f:
                f.write(text)
            rcode, vals = r.parse(d)
            cases[name] = dict(text=text, rc=rcode, values=vals)
        # writers: byte-exact text of write_to_cfile2D / write_to_rfile
        rng = np.random.default_rng(5)
        traj = rc(rng, 5, 3, 3) * np.array([1, 1e-7, 123456.789, 1e12, 0])[:, None, None]
        p = rng.standard_normal(5)
        r.write(d, traj, p)
        files = {fn: open(os.path.join(d, fn)).read() for fn in ("real_G1.txt", "imag_G1.txt", "dipole.txt")}
        np.savez_compressed(os.path.join(HERE, "writers.npz"), traj=traj, p=p)
        with open(os.path.join(HERE, "writers.json"), "w") as f:
            import json
            json.dump(files, f, indent=0)
    finally:
        shutil.rmtree(d)
    import json
    with open(os.path.join(HERE, "parser_cases.json"), "w") as f:
        json.dump(cases, f, indent=1)
    print("golden vectors written to", HERE)


if __name__ == "__main__":
    main()
