"""Parameter draws of the 256-trajectory ensemble workload.

Each trajectory: Ns uniform in {6..20}, Nb uniform in {1,2,3}, U uniform in
[0.5, 3.0], t2 uniform in [0.1, 0.8], quench strength uniform in [0.1, 2.0]
on Nq = floor(Ns/4) sites, long-range interactions with probability 0.5
(t1 = 1).  Drawn once with g12kba_amd.draw_ensemble (numpy default_rng,
seed 20261203) and committed as JSON, so every run and test uses identical
draws.  Usage:  python tests/golden/make_ensemble.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import g12kba_amd as g  # noqa: E402

SEED = 20261203
COUNT = 256


def main():
    d = g.draw_ensemble(SEED, COUNT)
    with open(os.path.join(HERE, "ensemble_config4.json"), "w") as f:
        json.dump({"seed": SEED, "count": COUNT, "draws": d}, f, indent=0)
    n = np.array([x["Ns"] * x["Nb"] for x in d])
    print(f"{len(d)} draws, N in [{n.min()}, {n.max()}], E[N^4] = {np.mean(n ** 4.0):.3e}, "
          f"E[N^5] = {np.mean(n ** 5.0):.3e}")


if __name__ == "__main__":
    main()
