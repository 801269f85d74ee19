"""Parameter draws of the 256-trajectory ensemble workload.

Each trajectory: Ns uniform in {6..20}, Nb uniform in {1,2,3}, U uniform in
[0.5, 3.0], t2 uniform in [0.1, 0.8], quench strength uniform in [0.1, 2.0]
on Nq = floor(Ns/4) sites, long-range interactions with probability 0.5
(t1 = 1).  Drawn once from numpy's default_rng(20261203) in that order per
trajectory and committed as JSON, so every run and test uses identical draws.
Usage:  python tests/golden/make_ensemble.py
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SEED = 20261203
COUNT = 256


def draws(seed=SEED, count=COUNT, max_n=None):
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


def main():
    d = draws()
    with open(os.path.join(HERE, "ensemble_config4.json"), "w") as f:
        json.dump({"seed": SEED, "count": COUNT, "draws": d}, f, indent=0)
    n = np.array([x["Ns"] * x["Nb"] for x in d])
    print(f"{len(d)} draws, N in [{n.min()}, {n.max()}], E[N^4] = {np.mean(n ** 4.0):.3e}, "
          f"E[N^5] = {np.mean(n ** 5.0):.3e}")


if __name__ == "__main__":
    main()
