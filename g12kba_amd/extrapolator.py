"""Learned G1 extrapolator: network binding, training on trajectories the method
generates, and extrapolated runs with plain RK4 as fallback and judge.

The network (g12kba_amd/csrc/extrap.hip) sees one upper-triangle element of G1
at a time: `hist` past steps and the quench drive give the next `horizon`
steps.  An extrapolated run propagates a prefix with RK4, then rolls the
network out block by block; a block whose trace drift or anti-hermiticity
error exceeds its threshold is rejected and RK4 resumes from its last full
(G1, G2) state, which then also refills the history.  A plain RK4 run of the
whole horizon judges the result.  The reference has no counterpart (its README
names extrapolation as future work).

Command line:  python -m g12kba_amd.extrapolator train --out FILE [options]
"""
import argparse
import ctypes
import json
import os
import time

import numpy as np

from . import (Batch, G12Error, Propagator, XPConfig, _check, _p, draw_ensemble, ensemble_params, fermi, lib,
               make_G0, make_params)

_fp = ctypes.POINTER(ctypes.c_float)
HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_MODEL = os.path.join(HERE, "models", "xp_n48.bin")


def _f(a):
    return a.ctypes.data_as(_fp)


class Extrapolator:
    def __init__(self, hist=64, horizon=16, hidden=320, dt=0.01, seed=20261203):
        c = XPConfig()
        lib.g12_xp_default_config(ctypes.byref(c))
        c.hist, c.horizon, c.hidden, c.dt, c.seed = hist, horizon, hidden, dt, seed
        h = ctypes.c_void_p()
        _check(lib.g12_xp_create(ctypes.byref(c), ctypes.byref(h)), "g12_xp_create")
        self.h = h
        self.hist, self.horizon, self.hidden, self.dt = hist, horizon, hidden, dt
        self.din, self.dout = 2 * hist + 2 * (hist + horizon), 2 * horizon
        n = ctypes.c_int64(0)
        lib.g12_xp_param_count(self.h, ctypes.byref(n))
        self.n_params = n.value

    def dims(self):
        return dict(hist=self.hist, horizon=self.horizon, din=self.din, dout=self.dout, H=self.hidden)

    def close(self):
        if self.h:
            lib.g12_xp_destroy(self.h)
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

    def params(self):
        p = np.zeros(self.n_params, np.float32)
        _check(lib.g12_xp_get_params(self.h, _f(p)), "g12_xp_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.n_params:
            raise ValueError("parameter count")
        _check(lib.g12_xp_set_params(self.h, _f(p)), "g12_xp_set_params")

    def save(self, path):
        _check(lib.g12_xp_save(self.h, path.encode()), "g12_xp_save")

    def load(self, path):
        _check(lib.g12_xp_load(self.h, path.encode()), "g12_xp_load")

    def set_data(self, series, coef, drive):
        """series: complex (n, T); coef: (n,) drive coefficients; drive: (T,) profile."""
        s = np.ascontiguousarray(series, dtype=np.complex64)
        c = np.ascontiguousarray(coef, dtype=np.float32)
        d = np.ascontiguousarray(drive, dtype=np.float32)
        n, T = s.shape
        self._data = (s, c, d)
        _check(lib.g12_xp_set_data(self.h, s.view(np.float32).ctypes.data_as(_fp), _f(c), _f(d), n, T),
               "g12_xp_set_data")

    def train(self, steps, batch=4096, lr=1e-3, record_loss=True):
        loss = np.zeros(max(steps, 1), np.float32)
        _check(lib.g12_xp_train(self.h, steps, batch, lr, _f(loss) if record_loss else None), "g12_xp_train")
        return loss[:steps]

    def loss_grad(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        loss = ctypes.c_double(0)
        grad = np.zeros(self.n_params, np.float32)
        _check(lib.g12_xp_loss_grad(self.h, _f(X), _f(T), X.shape[0], ctypes.byref(loss), _f(grad)),
               "g12_xp_loss_grad")
        return loss.value, grad

    def forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.zeros((X.shape[0], self.dout), np.float32)
        _check(lib.g12_xp_forward(self.h, _f(X), X.shape[0], _f(Y)), "g12_xp_forward")
        return Y

    def features(self, hist, Ns, Nq, w):
        hist = np.ascontiguousarray(np.asarray(hist)[-self.hist:], dtype=np.complex128)
        N = hist.shape[1]
        n_el = N * (N + 1) // 2
        w = np.ascontiguousarray(w, dtype=np.float64)
        X = np.zeros((n_el, self.din), np.float32)
        ref = np.zeros(n_el, np.complex128)
        sc = np.zeros(n_el)
        _check(lib.g12_xp_features(self.h, _p(hist), N, Ns, Nq, _p(w), _f(X), _p(ref), _p(sc)), "g12_xp_features")
        return X, ref, sc

    def rollout(self, hist, Ns, Nq, w, nblocks, tr_ref=None, project=False):
        """Predicts nblocks * horizon matrices after `hist` (oldest first).
        w: quench shift for hist + nblocks * horizon steps; tr_ref: conserved
        trace (default: that of the last history matrix); project: restore
        Re G_ii = 0 and Im tr G = Im tr_ref in each block before feeding it
        back.  Returns (matrices, checks) with checks[b] = (max trace drift,
        max |G + G^dagger|) of the network's own output for block b."""
        hist = np.ascontiguousarray(np.asarray(hist)[-self.hist:], dtype=np.complex128)
        N = hist.shape[1]
        w = np.ascontiguousarray(w, dtype=np.float64)
        if w.size < self.hist + nblocks * self.horizon:
            raise ValueError("drive too short")
        out = np.zeros((nblocks * self.horizon, N, N), np.complex128)
        chk = np.zeros((nblocks, 2))
        tr = None if tr_ref is None else np.array([complex(tr_ref).real, complex(tr_ref).imag])
        _check(lib.g12_xp_rollout(self.h, _p(hist), N, Ns, Nq, _p(w), _p(tr), int(bool(project)), nblocks, _p(out),
                                  _p(chk)), "g12_xp_rollout")
        return out, chk


# ---------------------------------------------------------------------------
# protocol helpers


def quench_shift(p, steps):
    """Quench shift w at the start of each production step index in `steps`
    (time accumulated as the RK4 loop does; gate and ramps of RK_solver.h:141-151)."""
    steps = np.asarray(steps)
    hi = int(steps.max()) + 1 if steps.size else 0
    t, times = 0.0, np.zeros(max(hi, 1))
    for k in range(hi):
        times[k] = t
        t += p.dt
    out = np.zeros(steps.shape)
    for i, k in np.ndenumerate(steps):
        ts = times[k] if k >= 0 else k * p.dt
        if ts < p.quench_gate_time:
            continue
        on = fermi(ts, p.quench_on_midpoint, p.quench_rate)
        phi = on * (1 - fermi(ts, p.quench_off_midpoint, p.quench_rate)) if p.q_type == 1 else on
        out[i] = phi * p.quench_strength
    return out


def site_mask(p):
    N = p.Ns * p.Nb
    return np.array([1.0 if (x % p.Ns) < p.Nq else 0.0 for x in range(N)])


# ---------------------------------------------------------------------------
# training data from the method itself


def training_protocol(dt=0.01, steps=1600, gate=6.0, precision="fp64"):
    """Production protocol of the training trajectories: the quench (full,
    rate 2) switches on at t = gate so that its onset lies inside the window."""
    return dict(dt=dt, time_steps=steps, V_law="coulomb", q_type="full", quench_rate=2.0,
                quench_on_midpoint=gate, quench_gate_time=gate, AS_rate=0.5, AS_midpoint=10.0,
                precision=precision)


def series_from_trajectory(traj, p, rng, per_traj):
    """Upper-triangle element series of one trajectory (all diagonal elements
    and a random subset of the others, per_traj in total) and their drive
    coefficients strength * (m_p - m_q)."""
    N = traj.shape[1]
    pairs = [(a, b) for a in range(N) for b in range(a + 1, N)]
    k = max(0, min(len(pairs), per_traj - N))
    sel = [(a, a) for a in range(N)] + [pairs[i] for i in rng.choice(len(pairs), size=k, replace=False)]
    m = site_mask(p)
    series = np.stack([traj[:, a, b] for a, b in sel]).astype(np.complex64)
    coef = np.array([p.quench_strength * (m[a] - m[b]) for a, b in sel], np.float32)
    return series, coef


def build_training_set(n_traj=256, seed=20261204, max_n=48, steps=1600, per_traj=96, holdout=0, log=print):
    """Runs n_traj trajectories of the ensemble distribution (N <= max_n) as one
    batch (FP64: switching, then `steps` production steps) and returns
    (series, coef, drive, held_out) for Extrapolator.set_data, the last
    `holdout` trajectories being left out of the series and returned as
    [(params, trajectory)] for evaluation."""
    draws = draw_ensemble(seed, n_traj, max_n=max_n)
    proto = training_protocol(steps=steps)
    ps = ensemble_params(draws, **proto)
    t0 = time.time()
    T = steps
    trajs = [np.zeros((T, p.Ns * p.Nb, p.Ns * p.Nb), np.complex128) for p in ps]
    rc = 0
    with Batch(ps) as b:
        for i, p in enumerate(ps):
            b.set_state(i, make_G0(p))
        b.adiabatic_switching()
        log(f"switching done ({time.time() - t0:.1f} s)")
        chunk = 200
        for k in range(0, T, chunk):  # in chunks, reporting progress
            n = min(chunk, T - k)
            rc = max(rc, b.record(n, trajs, k)[0])
            log(f"recorded {k + n}/{T} steps ({time.time() - t0:.1f} s)")
        fail = b.guard()
    trajs = [tr[:f] if f else tr for tr, f in zip(trajs, fail)]
    log(f"training trajectories: {n_traj} x {T} steps in {time.time() - t0:.1f} s (guard trips: {sum(map(bool, fail))})")
    rng = np.random.default_rng(seed)
    S, C, held = [], [], []
    for i, (p, tr) in enumerate(zip(ps, trajs)):
        if len(tr) < T or not np.isfinite(tr).all():  # guard tripped: left out
            continue
        if i >= len(ps) - holdout:
            held.append((p, tr))
            continue
        s, c = series_from_trajectory(tr[:T], p, rng, per_traj)
        S.append(s)
        C.append(c)
    drive = quench_shift(ps[0], np.arange(T)) / ps[0].quench_strength  # shared profile
    return np.concatenate(S), np.concatenate(C), drive.astype(np.float32), held


def evaluate_blocks(model, held, windows=12, seed=7):
    """One-block prediction error from true histories on held-out trajectories,
    against holding the last value and linear extrapolation.  Returns per
    window the largest absolute element error of each predictor."""
    rng = np.random.default_rng(seed)
    H, F = model.hist, model.horizon
    rows = []
    for p, tr in held:
        w_all = quench_shift(p, np.arange(len(tr)))
        for k in rng.integers(H, len(tr) - F, size=windows):
            truth = tr[k:k + F]
            pred, _ = model.rollout(tr[k - H:k], p.Ns, p.Nq, w_all[k - H:k + F], 1)
            hold = np.repeat(tr[k - 1][None], F, 0)
            slope = tr[k - 1] - tr[k - 2]
            lin = tr[k - 1][None] + slope[None] * np.arange(1, F + 1)[:, None, None]
            rows.append((p.Ns * p.Nb, int(k), float(np.abs(pred - truth).max()), float(np.abs(hold - truth).max()),
                         float(np.abs(lin - truth).max())))
    return rows


# ---------------------------------------------------------------------------
# extrapolated runs


def extrapolated_run(p, model, prefix=4000, total=None, G1_0=None, trace_tol=1e-6, herm_tol=1e-5, resume=64,
                     blocks_per_call=8, project=False, log=None):
    """RK4 for `prefix` production steps, then network blocks with RK4 fallback
    up to `total` steps (default: the production length of p).  A block is
    rejected when the network's trace drift exceeds trace_tol or its
    anti-hermiticity error herm_tol; with `project` accepted blocks are
    projected onto the conserved trace and G1^dagger = -G1 (the checks then
    see each block's own drift).  Returns a dict with the G1 trajectory (total
    matrices: G1 before each step), which steps the network produced, the RK4
    steps spent and the wall time."""
    if total is None:
        total, t = 0, 0.0
        while t < p.time_steps * p.dt:
            t += p.dt
            total += 1
    N = p.Ns * p.Nb
    H, F = model.hist, model.horizon
    w_all = quench_shift(p, np.arange(total + H + F * blocks_per_call + 1))
    traj = np.zeros((total, N, N), np.complex128)
    by_net = np.zeros(total, bool)
    t_start = time.time()
    with Propagator(p) as prop:
        prop.set_state(make_G0(p) if G1_0 is None else G1_0)
        prop.adiabatic_switching()
        rc, rec = prop.record(prefix)
        if rc:
            raise G12Error("density guard tripped in the RK4 prefix")
        traj[:prefix] = rec
        t_r = prefix      # steps the RK4 state has completed (its G1 is G1(t_{t_r}))
        k = prefix        # next step to fill; the history is traj[k - H:k]
        rk_steps = prefix
        tr_true = np.trace(traj[prefix - 1])  # conserved by RK4 to round-off
        rejected = accepted = 0
        while k < total:
            nb = min(blocks_per_call, (total - k + F - 1) // F)
            pred, chk = model.rollout(traj[k - H:k], p.Ns, p.Nq, w_all[k - H:k + nb * F], nb, tr_ref=tr_true,
                                      project=project)
            ok = 0
            while ok < nb and chk[ok, 0] <= trace_tol and chk[ok, 1] <= herm_tol:
                ok += 1
            n_ok = min(ok * F, total - k)
            traj[k:k + n_ok] = pred[:n_ok]
            by_net[k:k + n_ok] = True
            accepted += ok
            k += n_ok
            if ok == nb:
                continue
            rejected += 1
            if log:
                log(f"  block at step {k} rejected: trace drift {chk[ok, 0]:.2e}, herm {chk[ok, 1]:.2e}")
            # RK4 resumes from its own state and refills `resume` steps past the
            # rejected block's start (the next history is then all RK4)
            stop = min(total, k + resume)
            n = stop - t_r
            rc, rec = prop.record(n)
            if rc:
                raise G12Error("density guard tripped during RK4 fallback")
            traj[t_r:stop] = rec
            by_net[t_r:stop] = False
            rk_steps += n
            t_r = k = stop
    return dict(traj=traj, by_net=by_net, rk_steps=rk_steps, accepted_blocks=accepted, rejected_blocks=rejected,
                seconds=time.time() - t_start, prefix=prefix, total=total)


def rk4_run(p, total=None, G1_0=None):
    """The judge: plain RK4 over the whole horizon (G1 before each step)."""
    t0 = time.time()
    with Propagator(p) as prop:
        prop.set_state(make_G0(p) if G1_0 is None else G1_0)
        prop.adiabatic_switching()
        if total is None:
            rc, traj = prop.rk4()
        else:
            rc, traj = prop.record(total)
    return traj, time.time() - t0


def config5_params(time_steps=20000, precision="fp64"):
    """Ns = 24, Nb = 2 (N = 48), V_n = U/(1+n), U = 2, t1 = 1, t2 = 0.4, full
    quench 1.2 on 6 sites (rate 2, midpoint 50, gate 50), dt = 0.01."""
    return make_params(Ns=24, Nb=2, t1=1.0, t2=0.4, U=2.0, dt=0.01, time_steps=time_steps, quench_strength=1.2,
                       Nq=6, q_type="full", quench_rate=2.0, quench_on_midpoint=50.0, EHM=True, V_law="coulomb",
                       AS_rate=0.5, AS_midpoint=10.0, precision=precision)


def _main():
    ap = argparse.ArgumentParser(prog="python -m g12kba_amd.extrapolator")
    sub = ap.add_subparsers(dest="cmd", required=True)
    tr = sub.add_parser("train", help="generate trajectories with the method and train the network")
    tr.add_argument("--out", default=DEFAULT_MODEL)
    tr.add_argument("--trajectories", type=int, default=256)
    tr.add_argument("--steps", type=int, default=1600)
    tr.add_argument("--per-traj", type=int, default=96)
    tr.add_argument("--train-steps", type=int, default=4000)
    tr.add_argument("--batch", type=int, default=4096)
    tr.add_argument("--lr", type=float, default=1e-3)
    tr.add_argument("--holdout", type=int, default=16, help="trajectories kept out of training for evaluation")
    tr.add_argument("--log", default=None, help="JSON summary file")
    a = ap.parse_args()
    if a.cmd == "train":
        S, C, D, held = build_training_set(a.trajectories, steps=a.steps, per_traj=a.per_traj, holdout=a.holdout,
                                           log=lambda s: print(s, flush=True))
        print(f"series: {S.shape}, drive max {D.max():.3f}", flush=True)
        with Extrapolator() as m:
            m.set_data(S, C, D)
            losses = []
            t0 = time.time()
            chunk = 250
            for i in range(0, a.train_steps, chunk):
                n = min(chunk, a.train_steps - i)
                lr = a.lr * (0.1 if i >= 0.75 * a.train_steps else 1.0)
                ls = m.train(n, a.batch, lr)
                losses.append(float(np.mean(ls)))
                print(f"step {i + n}: loss {losses[-1]:.3e} ({time.time() - t0:.1f} s)", flush=True)
            os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
            m.save(a.out)
            rows = evaluate_blocks(m, held) if held else []
        ev = {}
        if rows:
            r = np.array([x[2:] for x in rows])
            ev = {"windows": len(rows), "median_err": {"network": float(np.median(r[:, 0])),
                                                       "hold": float(np.median(r[:, 1])),
                                                       "linear": float(np.median(r[:, 2]))},
                  "network_better_than_linear": float(np.mean(r[:, 0] < r[:, 2])),
                  "network_better_than_hold": float(np.mean(r[:, 0] < r[:, 1]))}
            print(f"held-out blocks: {json.dumps(ev)}", flush=True)
        if a.log:
            with open(a.log, "w") as f:
                json.dump({"series": list(S.shape), "losses": losses, "train_steps": a.train_steps,
                           "batch": a.batch, "lr": a.lr, "trajectories": a.trajectories, "steps": a.steps,
                           "holdout": a.holdout, "evaluation": ev, "windows": rows}, f)


if __name__ == "__main__":
    _main()
