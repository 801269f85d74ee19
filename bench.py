#!/usr/bin/env python3
"""Benchmark: G2 RK4 steps/s of the HF-GKBA G1-G2 (GW) propagation on one MI355X.

Workload (N = 64): 1-D chain with Ns = 32 sites and Nb = 2 orbitals, long-range
interactions V_n = U/(1+n), U = 1.5, t1 = 1, t2 = 0.5, full quench of strength
1.0 on the first 8 sites, dt = 0.01, quench ramp midpoint 30 and rate 2.0.  G1 starts from the non-interacting state; the two G2 tensors
(2^24 complex entries each) hold a small seeded Hermitian, pair-exchange
symmetric perturbation.
A "step" is one RK4 step (4 stages) of the production protocol; inputs are
resident in HBM before the timed region.  The state is pair-symmetric and
Hermitian, so the library runs its pair-symmetric path (one tile-product pass
and one pair-exchange update pass per stage); `config.path` records the path.

Prints one JSON line (bench contract).  `value` times the steps as a run
executes them (one captured hipGraph per step); `roofline` covers the dominant
kernel, timed with HIP events on the library's stream over a second run of as
many steps, and `passes` both G2 passes; `cpu_baseline` times the project's
faithful-complexity C++ restatement (OpenMP) on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "G2 RK4 steps/s at N=64 (FP64 vs mixed @≤1e-6 G1 err); batched traj/s"
# HBM bytes per launch of each G2 pass (read + written, averaged over the four
# stages), measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE on this workload
# and calibrated with a plain copy kernel; TRAFFIC_SOURCE names the summary
# each comes from.  Passes without a measurement report null.
TRAFFIC_BYTES = {("fp64", "tile_fused"): 2.995e9, ("fp64", "tile_strided"): 1.099e9,
                 ("fp32_3m", "tile_fused"): 2.455e9, ("fp32_3m", "tile_strided"): 0.556e9,
                 ("bf16x3", "tile_fused"): 2.328e9, ("bf16x3", "tile_strided"): 0.574e9,
                 ("fp64", "tile_products"): 0.566e9, ("fp64", "pair_update"): 1.121e9,
                 ("fp32_3m", "tile_products"): 0.283e9, ("fp32_3m", "pair_update"): 0.869e9,
                 ("bf16x3", "tile_products"): 0.283e9, ("bf16x3", "pair_update"): 0.869e9}

TRAFFIC_SOURCE = {("fp64", False): "profiles/tile_kernels_n64_fp64_hbm_traffic.txt",
                  ("fp32_3m", False): "profiles/tile_kernels_n64_mixed_hbm_traffic.txt",
                  ("bf16x3", False): "profiles/tile_kernels_n64_mixed_hbm_traffic.txt",
                  ("fp64", True): "profiles/pair_symmetric_n64_hbm_traffic.txt",
                  ("fp32_3m", True): "profiles/pair_symmetric_n64_hbm_traffic.txt",
                  ("bf16x3", True): "profiles/pair_symmetric_n64_hbm_traffic.txt"}
HBM_PEAK_TBS = 8.0


def workload_params(g, precision="fp64", time_steps=5000):
    return g.make_params(Ns=32, Nb=2, t1=1.0, t2=0.5, U=1.5, dt=0.01, time_steps=time_steps,
                         quench_strength=1.0, Nq=8, q_type="full", EHM=True, V_law="coulomb",
                         quench_rate=2.0, quench_on_midpoint=30.0,
                         AS_rate=0.5, AS_midpoint=10.0, precision=precision)


def synthetic_state(g, p, seed=20261203):
    import numpy as np
    N = p.Ns * p.Nb
    rng = np.random.default_rng(seed)
    G1 = g.make_G0(p)
    # a physical perturbation: Hermitian as an N^2 x N^2 matrix and symmetric
    # under exchanging both particle pairs, as G2 evolved from zero is
    # reference pair index of (p1, p2), p = site + orb * Ns: i + j Ns + (a + b Nb) Ns^2
    site, orb = np.arange(N) % p.Ns, np.arange(N) // p.Ns
    pair = (site[:, None] + site[None, :] * p.Ns + (orb[:, None] + orb[None, :] * p.Nb) * p.Ns ** 2).ravel()

    def g2():
        A = (rng.standard_normal((N * N, N * N)) + 1j * rng.standard_normal((N * N, N * N))) * 1e-3
        T = (A + A.conj().T).reshape(N, N, N, N)  # [p1][p2][q1][q2]
        T = 0.5 * (T + T.transpose(1, 0, 3, 2))
        M = np.empty((N * N, N * N), np.complex128)
        M[np.ix_(pair, pair)] = T.reshape(N * N, N * N)
        return M
    return G1, g2(), g2()


def host_cpu():
    """CPU model, logical CPUs and physical cores of the host (Linux)."""
    model, cores = "unknown", set()
    try:
        phys = core = None
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "model name" and model == "unknown":
                model = v
            elif k == "physical id":
                phys = v
            elif k == "core id":
                core = v
                cores.add((phys, core))
    except OSError:
        pass
    return {"model": model, "nproc": os.cpu_count(), "physical_cores": len(cores) or None}


def cpu_baseline(p, G1, Gxx, Gxy, threads):
    """One RK4 stage of the same workload through the oracle (C++, OpenMP)."""
    os.environ["OMP_NUM_THREADS"] = str(threads)
    from oracle.oracle import Oracle  # checker / baseline only
    import g12kba_amd as g
    o = Oracle()
    cfg = dict(Ns=p.Ns, Nb=p.Nb, t1=p.t1, t2=p.t2, U=p.U, dt=p.dt, time_steps=p.time_steps,
               quench_strength=p.quench_strength, Nq=p.Nq, q_type="full", HF=False, AS_rate=p.AS_rate,
               AS_midpoint=p.AS_midpoint, quench_rate=p.quench_rate, quench_on_midpoint=p.quench_on_midpoint,
               quench_off_midpoint=p.quench_off_midpoint)
    t0 = time.perf_counter()
    o.partial_step(cfg, G1, Gxx, Gxy, p.dt, 1, V_profile=g.v_profile(p))
    t1 = time.perf_counter()
    return {"value": 1.0 / (4 * (t1 - t0)), "unit": "steps/s", "cores": threads, "kind": "port",
            "sample": f"1 of the 4 RK4 stages at N=64 (stage input, full derivative, RK accumulation), "
                      f"{t1 - t0:.1f} s, x4 per step",
            "host": host_cpu(),
            "threads_note": "OpenMP threads capped at 16, the host-CPU share of one GPU on the measuring box"}


PEAK = {"fp64": (78.6, "FP64 dense"), "fp32_3m": (157.3, "FP32 dense"), "bf16x3": (2500.0, "BF16 dense")}


def run_leg(g, precision, G1, Gxx, Gxy, steps, warmup):
    """Times `steps` RK4 steps of one precision mode as a run executes them by
    default (each step one captured hipGraph), then runs as many steps again
    with HIP events around every launch for the kernel breakdown; returns
    (ms/step of the timed steps, G1 after them, kernel timings)."""
    p = workload_params(g, precision)
    prop = g.Propagator(p)
    prop.set_state(G1, Gxx, Gxy)
    if warmup:
        prop.advance(warmup)
    g.synchronize()
    t0 = time.perf_counter()
    rc = prop.advance(steps)
    g.synchronize()
    t1 = time.perf_counter()
    if rc != 0:
        raise SystemExit(f"density guard tripped (rc={rc}, {precision})")
    G1_end = prop.get_G1()
    sym = prop.pair_symmetric()
    prop.timing(True)
    rc = prop.advance(steps)
    k = [prop.timing_get(i) for i in range(3)]
    prop.timing(False)
    prop.close()
    if rc != 0:
        raise SystemExit(f"density guard tripped (rc={rc}, {precision})")
    return 1e3 * (t1 - t0) / steps, G1_end, (k, sym)


ENSEMBLE_DRAWS = os.path.join(ROOT, "tests", "golden", "ensemble_config4.json")


def ensemble_protocol(precision):
    """Shared protocol of the 256-trajectory ensemble: dt = 0.02, 4000 steps,
    default switching (rate 0.5, midpoint 10), full quench ramp (rate 2,
    midpoint 30), V_n = U/(1+n) where long-range interactions are on."""
    return dict(dt=0.02, time_steps=4000, V_law="coulomb", q_type="full", quench_rate=2.0,
                quench_on_midpoint=30.0, AS_rate=0.5, AS_midpoint=10.0, precision=precision)


def switching_steps(g, p):
    """Length of the adiabatic switching (AS.h:87-214), as the library counts it."""
    count, n, have, last, t = 0, 0, False, 0.0, 0.0
    while count < p.AS_tail:
        if have and abs(last - p.U) <= 0:
            count += 1
        for k in range(4):
            dtk = p.AS_dt / 2 if k in (1, 2) else p.AS_dt
            last = g.fermi(t if k == 0 else t + dtk, p.AS_midpoint, p.AS_rate) * p.U
        have = True
        t += p.AS_dt
        n += 1
    return n


def production_steps(p):
    n, t = 0, 0.0
    while t < p.time_steps * p.dt:  # accumulated time (RK_solver.h:96,216)
        t += p.dt
        n += 1
    return n


def ensemble_run(g, precision, steps, warmup, extra=0):
    """Times `steps` RK4 steps of the whole ensemble (one batch), then runs
    `extra` more untimed; returns (ms/step, G1 of every trajectory after
    warmup + steps + extra, batch facts)."""
    draws = json.load(open(ENSEMBLE_DRAWS))["draws"]
    ps = g.ensemble_params(draws, **ensemble_protocol(precision))
    b = g.Batch(ps)
    for i, p in enumerate(ps):
        b.set_state(i, g.make_G0(p))
    b.tune()
    if warmup:
        b.advance(warmup)
    g.synchronize()
    t0 = time.perf_counter()
    rc = b.advance(steps)
    g.synchronize()
    t1 = time.perf_counter()
    if rc != 0:
        raise SystemExit(f"ensemble: density guard tripped ({precision})")
    if extra and b.advance(extra) != 0:
        raise SystemExit(f"ensemble: density guard tripped ({precision})")
    G1 = [b.get_G1(i) for i in range(len(ps))]
    facts = {"buckets": b.buckets(), "memory_GB": round(b.memory_bytes() / 1e9, 2),
             "as_steps": max(switching_steps(g, p) for p in ps), "rk_steps": production_steps(ps[0]),
             "N": [p.Ns * p.Nb for p in ps]}
    b.close()
    return 1e3 * (t1 - t0) / steps, G1, facts


def ensemble_leg(g, steps, warmup, check_fp64=True, err_steps=100):
    import numpy as np
    extra = max(0, err_steps - warmup - steps)
    ms, G1m, f = ensemble_run(g, "fp32_3m", steps, warmup, extra)
    progress("  ensemble FP32-3M done; FP64 reference")
    per_run = f["as_steps"] + f["rk_steps"]
    out = {"metric": "trajectories/s", "value": round(256.0 / (per_run * ms * 1e-3), 4),
           "unit": "trajectories/s", "trajectories": 256, "precision": "fp32_3m",
           "ms_per_step": round(ms, 3), "steps": steps, "warmup": warmup,
           "steps_per_trajectory": per_run, "as_steps": f["as_steps"], "rk_steps": f["rk_steps"],
           "derivation": "value = trajectories / (steps per trajectory x measured ms per ensemble step); "
                         "every step of a run (switching or production) does the same work",
           "algorithmic_flop_per_step": float(sum(256.0 * n ** 5 for n in f["N"])),
           "buckets": [{"NP": k[0], "trajectories": k[1], "waves_strided": k[2], "waves_fused": k[3]}
                       for k in f["buckets"]],
           "memory_GB": f["memory_GB"], "config": "256 seeded draws (tests/golden/ensemble_config4.json), "
                                                  "dt=0.02, 4000 steps, FP32-3M tile products"}
    out["achieved_tflops"] = round(out["algorithmic_flop_per_step"] / (ms * 1e-3) / 1e12, 3)
    if check_fp64:
        ms64, G1d, _ = ensemble_run(g, "fp64", steps, warmup, extra)
        errs = [float(np.abs(np.diagonal(a).imag - np.diagonal(b).imag).max()) for a, b in zip(G1m, G1d)]
        out["fp64_ms_per_step"] = round(ms64, 3)
        out["max_density_err_vs_fp64"] = max(errs)
        out["median_density_err_vs_fp64"] = float(np.median(errs))
        out["err_horizon_steps"] = warmup + steps + extra
        out["err_note"] = ("per trajectory: largest |Im G1_ii| difference after the horizon (production steps "
                           "from G1(0), G2 = 0), FP32-3M batch against the FP64 batch")
    return out


def extrapolator_leg(g, prefix=4000):
    """Learned collision closure on the N = 48 workload over its full horizon
    (switching, 4000 RK4 steps, then 16001 closure steps with RK4 fallback),
    judged by plain RK4 over the same horizon."""
    import numpy as np
    from g12kba_amd import extrapolator as xp
    if not os.path.exists(xp.DEFAULT_MODEL):
        return {"error": "trained network missing"}
    p = xp.config5_params()
    judge, t_judge = xp.rk4_run(p)
    progress(f"  judge done ({t_judge:.1f} s); closure run")
    out = {"workload": "Ns=24 Nb=2 (N=48) EHM V=U/(1+n) U=2 t1=1 t2=0.4 full quench 1.2 on 6 sites dt=0.01, "
                       f"{len(judge)} production steps (RK4 prefix {prefix}), quench gate and midpoint t=50",
           "judge_rk4_seconds": round(t_judge, 3), "rk4_steps_per_s": round((len(judge) + 304) / t_judge, 1)}
    with xp.Extrapolator() as m:
        m.load(xp.DEFAULT_MODEL)
        r = xp.closure_run(p, m, prefix=prefix)
        tr = r["traj"]
        post = r["by_net"][prefix:]
        dens = np.abs(np.diagonal(tr - judge, axis1=1, axis2=2).imag).max(axis=1)
        out["closure"] = {
            "thresholds": {"trace": 1e-6, "hermiticity": 1e-5,
                           "energy_drift": "0.25 |E_corr| of the state each closure segment starts from"},
            "largest_checks_accepted": r["checks"], "rk4_steps": r["rk_steps"],
            "network_steps": int(post.sum()), "post_prefix_coverage": round(float(post.mean()), 4),
            "rejected_chunks": r["rejected_chunks"], "seconds": round(r["seconds"], 3),
            "speedup_vs_rk4": round(t_judge / r["seconds"], 3),
            "max_density_err_vs_rk4": float(dens.max()),
            "max_dipole_err_vs_rk4": float(np.abs(g.dipole(tr) - g.dipole(judge)).max())}
        out["closure_steps_per_s"] = (round(r["closure_steps_per_s"], 1) if r["closure_steps_per_s"] else None)
        out["closure_phase_seconds"] = {k: round(v, 3) for k, v in r["closure_phase_seconds"].items()}
    return out


def mixed_error_run(g, rk_steps):
    """Mixed-precision error along a physical run of the N = 64 workload: from
    G1(0) and G2 = 0, adiabatic switching then `rk_steps` production steps in
    every precision; the G1 trajectories are compared with FP64's."""
    import numpy as np
    out, trajs = {}, {}
    for prec in ("fp64", "fp32_3m", "bf16x3"):
        progress(f"  error run {prec}")
        p = workload_params(g, prec, time_steps=rk_steps)
        with g.Propagator(p) as prop:
            prop.set_state(g.make_G0(p))
            t0 = time.perf_counter()
            nas = prop.adiabatic_switching()
            t1 = time.perf_counter()
            rc, traj = prop.rk4()
            t2 = time.perf_counter()
        if rc != 0:
            raise SystemExit(f"density guard tripped in the error run ({prec})")
        trajs[prec] = traj
        # switching and production timed apart; the switching time includes the
        # autotuner's one-time timing launches before the first step
        out[prec] = {"as_steps": nas, "rk_steps": len(traj), "seconds": round(t2 - t0, 2),
                     "as_seconds": round(t1 - t0, 3), "rk_seconds": round(t2 - t1, 3),
                     "rk_steps_per_s": round(len(traj) / (t2 - t1), 1)}
    d64 = np.diagonal(trajs["fp64"], axis1=1, axis2=2).imag
    p64 = g.dipole(trajs["fp64"])
    for prec in ("fp32_3m", "bf16x3"):
        d = np.diagonal(trajs[prec], axis1=1, axis2=2).imag
        out[prec]["max_density_err_vs_fp64"] = float(np.abs(d - d64).max())
        out[prec]["max_dipole_err_vs_fp64"] = float(np.abs(g.dipole(trajs[prec]) - p64).max())
    return out


def observables_leg(g, G1, Gxx, Gxy, reps=5):
    """Cost of the observables and full-precision I/O at the N = 64 workload:
    the device energy entry (G1 and the N^2 pair diagonals leave the device)
    and a binary checkpoint save and load of (t, G1, G2xx, G2xy)."""
    import shutil
    import tempfile
    p = workload_params(g, "fp64")
    out = {}
    with g.Propagator(p) as prop:
        prop.set_state(G1, Gxx, Gxy)
        prop.advance(2)
        prop.energy()  # first call: host buffers
        t0 = time.perf_counter()
        for _ in range(reps):
            prop.energy()
        out["energy_ms"] = round((time.perf_counter() - t0) * 1e3 / reps, 3)
        d = tempfile.mkdtemp()
        try:
            path = os.path.join(d, "state.bin")
            t0 = time.perf_counter()
            prop.save(path)
            t1 = time.perf_counter()
            prop.load(path)
            t2 = time.perf_counter()
            size = os.path.getsize(path)
        finally:
            shutil.rmtree(d, ignore_errors=True)
    out.update({"checkpoint_bytes": size, "save_seconds": round(t1 - t0, 3), "load_seconds": round(t2 - t1, 3),
                "save_GB_per_s": round(size / (t1 - t0) / 1e9, 2), "load_GB_per_s": round(size / (t2 - t1) / 1e9, 2),
                "note": "host file system of the run; includes the device layout permutation and PCIe copies"})
    return out


def update_bytes_per_element(precision):
    """Algorithmic HBM bytes (read, written) per stored G2 element of one
    pair-exchange update launch, averaged over the four stages.  Every stage
    reads the product pass's result X (16 B in FP64, 8 B in the mixed modes)
    and the state S (16 B; the mixed modes read its complex64 copy, 8 B, at
    stages 0 and 1).  Stages 0-2 write the next stage input Y (16 B / 8 B).
    Stage 2 folds the RK terms into S (S written).  The mixed modes store K1
    and K2 as complex64 (8 B each, written at stages 0 and 1, read at stage 2);
    FP64 keeps the stage-2 input apart instead (stage 1 writes it to the A
    buffer) and stage 2 recovers K1 and K2 from the two stored inputs (16 B
    each).  Stage 3 writes the new state S (and in the mixed modes its
    complex64 copy):
      FP64:  48, 48, 96, 48 B at stages 0-3 -> 40 B read + 20 B written;
      mixed: 32, 32, 64, 48 B               -> 24 B read + 20 B written."""
    return (40.0, 20.0) if precision == "fp64" else (24.0, 20.0)


def update_bytes_per_launch(N, precision):
    """Algorithmic bytes of one update launch over both tensors in the
    Hermitian half layout: N (N + 1) / 2 stored tiles of N^2 elements per
    tensor, of which the N (N - 1) / 2 lower elements of each of the N diagonal
    tiles are written as the Hermitian images of the upper ones (written
    only)."""
    rd, wr = update_bytes_per_element(precision)
    stored = N * (N + 1) // 2 * N * N
    images = N * (N * (N - 1) // 2)
    return 2.0 * ((stored - images) * (rd + wr) + images * wr)


def roofline(N, precision, kk, steps):
    """Rooflines of the two G2 passes from their evented launch times, and the
    dominant one (larger summed time).

    Pair-symmetric path: the product pass applies the p2 and q2 terms of
    [H2, G2] to the tiles t0 <= t1 of both tensors (the others are their
    negated conjugate transposes), 2 tensors x N (N + 1) / 2 tiles x 2 N^3
    complex MACs x 8 flop = 16 N^4 (N + 1) per launch (8 flop per complex MAC
    whatever the product scheme: 3M or BF16 split); the p1 and q1 terms are its
    pair exchange, which the update pass adds while it streams the RK update
    (HBM-bound: update_bytes_per_launch bytes per launch).  General
    path: the strided and the fused pass each apply two of the four terms to
    every tile, 32 N^5 flop.
    """
    k, sym = kk
    flops = 16.0 * N ** 4 * (N + 1) if sym else 32.0 * N ** 5
    peak, peak_name = PEAK[precision]

    def mfma(name, n_launch, ms_total):
        avg_ms = ms_total / max(n_launch, 1)
        achieved = flops / (avg_ms * 1e-3) / 1e12
        return {"bound": "mfma", "kernel": name, "achieved": round(achieved, 3), "peak": peak,
                "peak_type": peak_name, "unit": "TFLOP/s", "frac": round(achieved / peak, 4),
                "traffic": TRAFFIC_BYTES.get((precision, name)),
                "traffic_source": TRAFFIC_SOURCE.get((precision, sym)),
                "algorithmic_flop_per_launch": flops, "avg_launch_ms": round(avg_ms, 4)}

    def hbm(name, n_launch, ms_total):
        avg_ms = ms_total / max(n_launch, 1)
        nbytes = update_bytes_per_launch(N, precision)
        achieved = nbytes / (avg_ms * 1e-3) / 1e12
        return {"bound": "hbm", "kernel": name, "achieved": round(achieved, 3), "peak": HBM_PEAK_TBS,
                "unit": "TB/s", "frac": round(achieved / HBM_PEAK_TBS, 4),
                "traffic": TRAFFIC_BYTES.get((precision, name)),
                "traffic_source": TRAFFIC_SOURCE.get((precision, sym)),
                "algorithmic_bytes_per_launch": nbytes, "avg_launch_ms": round(avg_ms, 4)}

    if sym:
        passes = {"tile_products": mfma("tile_products", *k[0]), "pair_update": hbm("pair_update", *k[1])}
        names = ("tile_products", "pair_update")
    else:
        passes = {"tile_strided": mfma("tile_strided", *k[0]), "tile_fused": mfma("tile_fused", *k[1])}
        names = ("tile_strided", "tile_fused")
    dom = names[0] if k[0][1] >= k[1][1] else names[1]
    out = dict(passes[dom])
    out["path"] = "pair-symmetric" if sym else "general"
    out["passes"] = passes
    out["breakdown_ms_per_step"] = {names[0]: round(k[0][1] / steps, 3), names[1]: round(k[1][1] / steps, 3),
                                    "other": round(k[2][1] / steps, 3)}
    return out


T_START = time.time()


def progress(msg):
    """A progress line on stderr (stdout carries only the result line)."""
    print(f"[bench {time.time() - T_START:7.1f} s] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32_3m", "bf16x3"],
                    help="precision of the headline leg")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision legs")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-ensemble", action="store_true", help="skip the 256-trajectory ensemble leg")
    ap.add_argument("--ensemble-steps", type=int, default=20)
    ap.add_argument("--ensemble-err-steps", type=int, default=100,
                    help="production steps over which the ensemble's FP32-3M error against FP64 is taken")
    ap.add_argument("--no-extrapolator", action="store_true", help="skip the learned-extrapolator leg")
    ap.add_argument("--no-observables", action="store_true", help="skip the energy / checkpoint timing leg")
    ap.add_argument("--mixed-steps", type=int, default=5000,
                    help="production steps of the mixed-precision error run (5000: the full workload, whose "
                         "quench acts from t = 50)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build runs on exactly one GPU; --gpus {a.gpus} is not supported "
              "(trajectories do not shard across GPUs here)", file=sys.stderr)
        sys.exit(2)

    import g12kba_amd as g
    import numpy as np

    p = workload_params(g, "fp64")
    N = p.Ns * p.Nb
    G1, Gxx, Gxy = synthetic_state(g, p)
    legs = [a.precision] + ([] if a.no_mixed else [x for x in ("fp64", "fp32_3m", "bf16x3") if x != a.precision])
    res = {}
    for prec in legs:
        progress(f"N=64 leg {prec}: {a.warmup} warmup + {a.steps} timed steps")
        res[prec] = run_leg(g, prec, G1, Gxx, Gxy, a.steps, a.warmup)
    ms_step, G1_head, k_head = res[a.precision]
    if not np.isfinite(G1_head).all():
        raise SystemExit("non-finite state")
    mixed = {}
    if "fp64" in res:
        d64 = np.diagonal(res["fp64"][1]).imag
        for prec in ("fp32_3m", "bf16x3"):
            if prec in res:
                ms, G1m, km = res[prec]
                mixed[prec] = {"steps_per_s": round(1e3 / ms, 3), "ms_per_step": round(ms, 4),
                               "speedup_vs_fp64": round(res["fp64"][0] / ms, 3),
                               "max_density_err_vs_fp64": float(np.abs(np.diagonal(G1m).imag - d64).max()),
                               "err_horizon_steps": a.warmup + a.steps,
                               "roofline": roofline(N, prec, km, a.steps)}
    if mixed and a.mixed_steps > 0:
        progress(f"mixed-precision error run: switching + {a.mixed_steps} production steps in three precisions")
        mixed["error_run"] = mixed_error_run(g, a.mixed_steps)
    if not a.no_ensemble:
        progress("ensemble leg: 256 trajectories, FP32-3M and FP64")
    ensemble = None if a.no_ensemble else ensemble_leg(g, a.ensemble_steps, 1, err_steps=a.ensemble_err_steps)
    if not a.no_extrapolator:
        progress("extrapolator leg: RK4 judge and closure over the N=48 horizon")
    extrap = None if a.no_extrapolator else extrapolator_leg(g)
    if not a.no_observables:
        progress("observables leg")
    observables = None if a.no_observables else observables_leg(g, G1, Gxx, Gxy)
    cpu = None
    if not a.no_cpu:
        progress("CPU baseline")
        threads = a.cpu_threads or min(16, os.cpu_count() or 1)
        try:
            cpu = cpu_baseline(p, G1, Gxx, Gxy, threads)
        except Exception as e:  # oracle not built: report why
            cpu = {"value": None, "error": str(e)}
    out = {"metric": METRIC, "value": round(1e3 / ms_step, 3), "unit": "steps/s", "n_gpus": 1,
           "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(ms_step, 4), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64" if a.precision == "fp64" else a.precision,
           "data": "synthetic (non-interacting G1(0), seeded Hermitian pair-symmetric G2 perturbation)",
           "config": {"workload": "Ns=32 Nb=2 (N=64) EHM V=U/(1+n) U=1.5 t1=1 t2=0.5 full quench 1.0 on 8 sites "
                                  "dt=0.01", "N": N, "precision": a.precision,
                      "algorithmic_flop_per_step": 256.0 * N ** 5,
                      "path": "pair-symmetric" if k_head[1] else "general"},
           "roofline": roofline(N, a.precision, k_head, a.steps), "mixed": mixed, "ensemble": ensemble, "extrapolator": extrap,
           "observables": observables, "cpu_baseline": cpu}
    progress("done")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
