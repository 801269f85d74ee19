"""The 256-trajectory ensemble workload (committed draws): batches covering
every padded tile size (16, 32, 48, 64) advance each trajectory bit-for-bit
as it advances alone, in FP64 and in the FP32-3M mode the ensemble runs in."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

g = pytest.importorskip("g12kba_amd")
HERE = os.path.dirname(os.path.abspath(__file__))
DRAWS = json.load(open(os.path.join(HERE, "golden", "ensemble_config4.json")))["draws"]


def pick(per_bucket=2):
    """The first draws of each padded tile size, onsite and long-range mixed."""
    out, seen = [], {}
    for d in DRAWS:
        NP = (d["Ns"] * d["Nb"] + 15) // 16 * 16
        if seen.get(NP, 0) < per_bucket:
            seen[NP] = seen.get(NP, 0) + 1
            out.append(d)
    return out


def protocol(precision, steps=3):
    # quench gate at 0 so the quench acts within the few steps run here
    return dict(dt=0.02, time_steps=steps, V_law="coulomb", q_type="full", quench_rate=2.0,
                quench_on_midpoint=0.02, quench_gate_time=0.0, precision=precision)


def test_draw_table():
    assert len(DRAWS) == 256
    n = np.array([d["Ns"] * d["Nb"] for d in DRAWS])
    assert n.min() >= 6 and n.max() <= 60
    assert all(d["Nq"] == d["Ns"] // 4 for d in DRAWS)
    assert {(n_ + 15) // 16 * 16 for n_ in n} == {16, 32, 48, 64}


@pytest.mark.parametrize("precision", ["fp64", "fp32_3m"])
def test_ensemble_batch_equals_single_runs(precision):
    sel = pick()
    ps = g.ensemble_params(sel, **protocol(precision))
    G0 = [g.make_G0(p) for p in ps]
    with g.Batch(ps) as b:
        assert sorted(k[0] for k in b.buckets()) == [16, 32, 48, 64]
        for i, G in enumerate(G0):
            b.set_state(i, G)
        rc, trajs = b.rk4()
        finals = [b.get_state(i) for i in range(len(ps))]
    assert rc == 0
    for i, p in enumerate(ps):
        with g.Propagator(p) as prop:
            prop.set_state(G0[i])
            rc1, traj1 = prop.rk4()
            fin1 = prop.get_state()
        assert rc1 == 0 and np.array_equal(traj1, trajs[i]), f"draw {i} N={p.Ns * p.Nb}"
        for x, y in zip(fin1, finals[i]):
            assert np.array_equal(x, y), f"draw {i} N={p.Ns * p.Nb}"
        tr = np.trace(trajs[i], axis1=1, axis2=2)
        assert np.abs(tr - tr[0]).max() < 1e-12
        assert np.abs(finals[i][1]).max() > 0  # correlations built up from G2 = 0
