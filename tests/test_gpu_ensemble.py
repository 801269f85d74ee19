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


def test_full_ensemble_fp32_3m_against_fp64_and_oracle():
    """All 256 committed draws in one batch (N = 6..60, every bucket), three
    production steps with the quench acting.  FP32-3M: every trajectory's
    densities within 1e-6 of the FP64 batch and its trace fixed to 1e-12.  FP64:
    one draw per padded size, two of them with N in 49..60 (general NP = 64
    tiles), against the factorised oracle at 1e-12 relative (G1 and both G2)."""
    from oracle.oracle import Oracle
    from tests.helpers import oracle_cfg
    ps32 = g.ensemble_params(DRAWS, **protocol("fp32_3m"))
    ps64 = g.ensemble_params(DRAWS, **protocol("fp64"))
    G0 = [g.make_G0(p) for p in ps64]
    Ns_all = np.array([p.Ns * p.Nb for p in ps64])
    checked = [int(np.flatnonzero(Ns_all <= 16)[0]), int(np.flatnonzero((Ns_all > 16) & (Ns_all <= 32))[0]),
               int(np.flatnonzero((Ns_all > 32) & (Ns_all <= 48))[0])]
    checked += [int(i) for i in np.flatnonzero((Ns_all >= 49) & (Ns_all <= 60))[:2]]
    assert len(checked) == 5 and Ns_all[checked[3]] < 64 and Ns_all[checked[4]] < 64
    out = {}
    for prec, ps in (("fp32_3m", ps32), ("fp64", ps64)):
        print(f"full ensemble {prec}: 256 trajectories", flush=True)
        with g.Batch(ps) as b:
            assert sorted(k[0] for k in b.buckets()) == [16, 32, 48, 64]
            for i, G in enumerate(G0):
                b.set_state(i, G)
            rc, trajs = b.rk4()
            assert rc == 0
            finals = {i: b.get_state(i) for i in checked} if prec == "fp64" else {}
            out[prec] = ([t.copy() for t in trajs], [b.get_G1(i) for i in range(len(ps))], finals)
    for i in range(len(ps64)):
        t32, t64 = out["fp32_3m"][0][i], out["fp64"][0][i]
        assert len(t32) == len(t64) == 3
        d32 = np.concatenate([np.diagonal(t32, axis1=1, axis2=2).imag, np.diagonal(out["fp32_3m"][1][i]).imag[None]])
        d64 = np.concatenate([np.diagonal(t64, axis1=1, axis2=2).imag, np.diagonal(out["fp64"][1][i]).imag[None]])
        assert np.abs(d32 - d64).max() <= 1e-6, f"draw {i} N={Ns_all[i]}"
        tr = np.trace(t32, axis1=1, axis2=2)
        assert np.abs(tr - tr[0]).max() < 1e-12 and abs(np.trace(out["fp32_3m"][1][i]) - tr[0]) < 1e-12
    O = Oracle()
    for i in checked:
        p = ps64[i]
        print(f"oracle check draw {i} N={Ns_all[i]}", flush=True)
        N = p.Ns * p.Nb
        Z = np.zeros((N * N, N * N), np.complex128)
        ref = O.fast_run(oracle_cfg(p), G0[i], Z, Z, V_profile=list(g.v_profile(p)), run_as=False)
        G1, Gxx, Gxy = out["fp64"][2][i]
        rel = lambda a, b: np.abs(a - b).max() / np.abs(b).max()  # noqa: E731
        assert ref["rc"] == 0 and ref["rk_steps"] == 3
        assert rel(G1, ref["G1"]) < 1e-12, f"draw {i} N={N}"
        assert rel(Gxx, ref["Gxx"]) < 1e-12 and rel(Gxy, ref["Gxy"]) < 1e-12, f"draw {i} N={N}"
