"""Launch-shape autotuner for the simulation kernels.

The library's defaults (lanes per simulation by a size rule, regrouping
interval, SSA chunk, workgroup size, iterations per launch) are rules of
thumb.  tune() instead times short probe runs of the actual model on the
actual device and keeps the fastest configuration:

1. The whole horizon, cut into windows and mixed in one launch.  One run of
   the library's default over `window_sims` simulations records every
   species at w * T / windows (w = 1 .. windows-1).  A probe of W windows
   (W divides `windows`) restarts the model from the states at w * T / W
   (and from the initial state) and runs each for T / W: a mass-action
   network is time-homogeneous, so a restarted simulation is the model from
   that point on, and the W windows together cover every regime of the run
   (start-up, transitions, the quasi-stationary phase) in proportion.  The
   windows are interleaved simulation by simulation in ONE launch, because a
   whole run's workgroups hold simulations at different points of their
   trajectories (a group that finishes a simulation starts its next one from
   the initial state while its neighbours are deep in theirs): the mix of
   leaping and SSA simulations inside a workgroup is what regrouping by mode
   separates.  W starts at `windows` and is cut until the library default's
   probe launch takes `min_probe_ms` (a short launch times mostly its start
   and its end).  The windows' costs differ (an SSA-heavy window can cost ten
   times a leaping one), and a launch whose last simulations are the
   costliest times mostly idle lanes: the calibration run's per-simulation
   counters rank the windows, and the probe hands out the costliest quarter
   first and the cheapest last.
2. Stage 1 times every lanes-per-simulation candidate (1, and G = 2..64 where
   the model fits the group kernel's LDS) with the library's defaults.
3. Stage 2 refines the winner: for G = 2..32 the regrouping interval
   (compact_interval, including off) and the SSA chunk, then a workgroup of
   two wavefronts instead of four; for G = 1 the workgroup size with and
   without regrouping, then the SSA chunk; for every G the iteration budget
   per launch.
4. Stage 4 times the rivals on fewer, longer windows -- the whole horizon
   when the estimate fits `confirm_ms` -- and keeps the fastest: the best
   configuration, the library's defaults at its G, the best one of its G on
   the other side of regrouping, and the defaults at every other G that came
   within 25 % in stage 1.  Short windows cut simulations off their
   neighbours' regimes and time many short simulations instead of a few long
   ones, which narrows the gaps between G (on the 100x200 network the probe
   put G = 16 and 32 within 1 % and G = 64 within 6 %; whole runs give 7.4
   and 5.5e9 firings/s at G = 16 and 32), and misjudges regrouping, which
   depends on that mixing.

A configuration's score is its firings over its kernel time (firings per
second of the probe launch).  Results are
memoised per (model signature, device, probe settings) in-process and
optionally in a JSON file.  Lanes per simulation selects one of several
statistically equivalent algorithms, so the trajectories of a seed depend on
it (reported with every result, ctl_stats.lanes_per_sim); the other options
change only scheduling, never a trajectory.

tune_batch() applies the same idea to a batch of uneven models: per kernel
class of the library's size rule, the class's largest model is probed as a
one-model batch (the batch kernels themselves) with the class's G and its
neighbours, and the fastest G is applied to the class
(ctl_batch_create_lanes).
"""
from __future__ import annotations

import hashlib
import json
import os
import time

import numpy as np

from . import CTL_E_LIMIT, Batch, CtlError, Model, Sim

CANDIDATES = (1, 2, 4, 8, 16, 32, 64)
_memo: dict = {}


def signature(left, right, x0, c, ts, feed=None) -> str:
    h = hashlib.sha1()
    for a in (left, right, x0, c, ts, feed if feed is not None else np.zeros(0)):
        a = np.ascontiguousarray(a)
        h.update(str(a.shape).encode())
        h.update(a.tobytes())
    return h.hexdigest()[:16]


def _fits_group(S: int, R: int) -> bool:
    Rp = (R + 64) // 64 * 64          # the padded reaction arrays at the widest group
    return (2 * S + 2 * Rp + max(0, (R + 1) // 2 - S)) * 4 <= 65536


def _key(opts: dict) -> str:
    return ",".join(f"{k}={v}" for k, v in sorted(opts.items()))


def _grid_states(left, right, x0, c, ts, feed, windows, window_sims, seed, gpu, force_ssa=False):
    """States at the start of each of `windows` equal windows of the horizon:
    [x0 (1 row)] + [window_sims rows at w*T/windows, w = 1 .. windows-1], from
    one run of the library's default configuration."""
    starts = [np.atleast_2d(np.asarray(x0))[:1].astype(np.uint32)]
    if windows <= 1:
        return starts
    T = float(ts[-1])
    t_w = np.array([T * w / windows for w in range(1, windows)])
    m = Model.from_arrays(left, right, np.atleast_2d(x0)[:1], np.atleast_2d(c)[:1], t_w, feed=feed)
    try:
        with Sim(m, 64, max(1, window_sims // 64), gpu=gpu, seed=seed ^ 0x5A5A, force_ssa=force_ssa) as s:
            s.run()
            X = s.samples()                          # [sims][windows-1][S], every species
    finally:
        m.close()
    starts += [np.ascontiguousarray(X[:, w, :]) for w in range(windows - 1)]
    return starts


def _mixed(starts, n, order=None):
    """n initial-state rows from the window starts.  Without `order`, row i
    takes start i % W (within a start, its rows in turn).  With `order` (the
    windows by decreasing cost), the rows come in four consecutive quarters,
    each interleaving a quarter of the windows, the costliest first: the
    persistent kernels fetch simulations in index order, so the launch ends
    on the cheapest windows instead of waiting on a few of the costliest."""
    W = len(starts)
    i = np.arange(n)
    w_of, k_of = i % W, i // W
    if order is not None:
        rank = np.empty(W, np.int64)
        rank[np.asarray(order)] = np.arange(W)
        r = rank[w_of]
        idx = np.lexsort((r, k_of, r // max(1, W // 4)))   # quarter, then round-robin inside it
        w_of, k_of = w_of[idx], k_of[idx]
    X = np.empty((n, starts[0].shape[1]), np.uint32)
    for w, st in enumerate(starts):
        sel = w_of == w
        X[sel] = st[k_of[sel] % st.shape[0]]
    return X


def _divisor_at_most(n: int, x: float) -> int:
    """Largest divisor of n that is <= x (at least 1)."""
    return max([d for d in range(1, n + 1) if n % d == 0 and d <= x] or [1])


def tune(left, right, x0, c, ts, *, feed=None, cs=None, gpu=0, probe_sims=65536, windows=64, min_probe_ms=250.0,
         window_sims=1024, candidates=CANDIDATES, refine=True, confirm_ms=15000.0, force_ssa=False, seed=0x7E57,
         cache_file=None, verbose=False) -> dict:
    """Returns {"lanes_per_sim": G, "options": Sim keyword arguments of the
    fastest configuration, "scores": {configuration: firings/s on the
    probe}, "confirm": {configuration: firings/s on the confirmation run}
    (empty when stage 4 did not run), "probe": {...}}.  force_ssa tunes the
    exact-SSA mode (every probe, the window states included, runs in it)."""
    left = np.asarray(left)
    R, S = left.shape
    ts = np.asarray(ts, np.float64)
    T = float(ts[-1])
    sig = signature(left, right, x0, c, ts, feed)
    force_ssa = bool(force_ssa)
    key = (sig, gpu, probe_sims, windows, min_probe_ms, tuple(candidates), refine, confirm_ms, force_ssa)
    if key in _memo:
        return _memo[key]
    disk_key = (f"{sig}@{gpu}:{probe_sims}:{windows}:{min_probe_ms:g}:{','.join(map(str, candidates))}:"
                f"{int(refine)}:{confirm_ms:g}:{int(force_ssa)}")
    if cache_file and os.path.exists(cache_file):
        with open(cache_file) as f:
            hit = json.load(f).get(disk_key)
        if hit:
            _memo[key] = hit
            return hit
    say = (lambda m: print(m, flush=True)) if verbose else (lambda m: None)
    windows = max(1, int(windows))
    window_sims = max(64, window_sims // 64 * 64)
    probe_sims = max(64, probe_sims // 64 * 64)
    t0 = time.perf_counter()
    states = _grid_states(left, right, x0, c, ts, feed, windows, window_sims, seed, gpu, force_ssa)

    def probe_model(W, order=None):
        """W windows of T/W each, starting at w*T/W, mixed in one model."""
        dt = T / W
        t_p = ts[ts <= dt * (1 + 1e-9)]
        if len(t_p) == 0:
            t_p = np.array([dt])
        X = _mixed(states[::windows // W], probe_sims, order)
        return Model.from_arrays(left, right, X, np.atleast_2d(c)[:1], t_p, feed=feed, cs=cs)

    def timed(m, opts, per_sim=False):
        with Sim(m, 64, probe_sims // 64, gpu=gpu, seed=seed, force_ssa=force_ssa, **opts) as s:
            s.run()
            st = s.stats()
            ss = s.sim_stats() if per_sim else None
        return st["firings"], st["kernel_ms"], ss

    # Calibration: fewer, longer windows until the library default's probe
    # launch takes min_probe_ms (short launches time mostly their start and end).
    W, calib = windows, []
    while True:
        m = probe_model(W)
        try:
            fir, ms, ss = timed(m, {}, per_sim=True)
        finally:
            m.close()
        calib.append({"windows": W, "kernel_ms": round(ms, 2)})
        if ms >= min_probe_ms or W == 1:
            break
        W = _divisor_at_most(windows, min(W - 1, W * ms / min_probe_ms))
    # Cost of a window's simulations in the calibration run (an SSA step
    # costs about a sixth of a leap), for the costliest-first order.
    cost = (ss["leaps"].astype(np.float64) + ss["ssa_steps"] / 6.0)
    w_cost = np.bincount(np.arange(probe_sims) % W, weights=cost, minlength=W)
    order = np.argsort(-w_cost, kind="stable")
    models = [probe_model(W, order)]
    say(f"# autotune: {W} windows of {T / W:g} time units over {probe_sims} simulations, costliest first "
        f"(calibration {calib}; {time.perf_counter() - t0:.1f} s to set up)")
    scores = {}
    confirmed = {}

    def score(opts: dict) -> float:
        k = _key(opts)
        if k in scores:
            return scores[k]
        t1 = time.perf_counter()
        try:
            fir, ms, _ = timed(models[0], opts)
        except CtlError as e:            # a shape the model does not allow (LDS)
            say(f"# autotune {k}: not possible ({e})")
            scores[k] = 0.0
            return 0.0
        rate = fir / max(ms, 1e-3) * 1e3
        scores[k] = rate
        say(f"# autotune {k}: {rate:.3e} firings/s ({ms:.1f} ms kernel, {time.perf_counter() - t1:.1f} s wall)")
        return rate

    W_conf = None
    try:
        best, best_rate, stage1 = None, -1.0, {}
        for G in candidates:
            if G > 1 and not _fits_group(S, R):
                continue
            o = {"lanes_per_sim": G}
            r = score(o)
            stage1[G] = r
            if r > best_rate:
                best, best_rate = o, r
        if refine and best is not None:
            G = best["lanes_per_sim"]
            if 1 < G < 64:
                grid = [{"compact_interval": -1}] + [{"compact_interval": ci, "ssa_chunk": ch}
                                                     for ci in (4, 8, 16) for ch in (4, 8, 12)]
            elif G == 1:
                grid = [{"block_size": 64, "compact_interval": -1}, {"block_size": 128, "compact_interval": -1},
                        {"block_size": 256, "compact_interval": -1}, {"block_size": 128, "compact_interval": 8},
                        {"block_size": 256, "compact_interval": 8}, {"block_size": 256, "compact_interval": 4}]
            else:
                grid = []
            for extra in grid:
                o = dict(best, **extra)
                r = score(o)
                if r > best_rate:
                    best, best_rate = o, r
            if 1 < G < 64 and best.get("compact_interval", 0) >= 0:
                o = dict(best, block_size=128)      # regrouping over two wavefronts instead of four
                r = score(o)
                if r > best_rate:
                    best, best_rate = o, r
            if G == 1:
                for ch in (5, 20):
                    o = dict(best, ssa_chunk=ch)
                    r = score(o)
                    if r > best_rate:
                        best, best_rate = o, r
            for it in ((1024, 8192) if G == 1 else (4096, 16384)):
                o = dict(best, iters_per_launch=it)
                r = score(o)
                if r > best_rate:
                    best, best_rate = o, r
            # Stage 4: the rivals run longer windows (the whole horizon when
            # that fits confirm_ms): the best configuration, the library's
            # defaults at its G, the best one of its G on the other side of
            # regrouping, and the defaults at every other G that came within
            # 25 % in stage 1.
            def regroups(o):
                ci = o.get("compact_interval", 0)
                return ci > 0 or (ci == 0 and 16 <= o["lanes_per_sim"] < 64 and not force_ssa)
            rivals = [best] + ([{"lanes_per_sim": G}] if best != {"lanes_per_sim": G} else [])
            if 1 < G < 64:
                other = None
                for k, r in scores.items():
                    o = {kv.split("=")[0]: int(kv.split("=")[1]) for kv in k.split(",")}
                    if (o.get("lanes_per_sim") == G and r > 0 and regroups(o) != regroups(best)
                            and o not in rivals):
                        if other is None or r > scores[_key(other)]:
                            other = o
                if other is not None:
                    rivals.append(other)
            for g, r in sorted(stage1.items()):
                if g != G and r >= 0.75 * stage1[G]:
                    rivals.append({"lanes_per_sim": g})
            probe_ms = fir / max(best_rate, 1e-9) * 1e3      # the default's firings at the best rate
            W_conf = next((d for d in range(1, W) if W % d == 0 and probe_ms * W / d <= confirm_ms), None)
            if confirm_ms > 0 and len(rivals) > 1 and W_conf is not None:
                # window v of W_conf spans windows v*r .. v*r+r-1 of the probe
                c_conf = w_cost.reshape(W_conf, W // W_conf).sum(axis=1)
                m = probe_model(W_conf, np.argsort(-c_conf, kind="stable"))
                try:
                    long_rate = {}
                    for o in rivals:
                        fir_c, ms_c, _ = timed(m, o)
                        long_rate[_key(o)] = confirmed[_key(o)] = fir_c / max(ms_c, 1e-3) * 1e3
                        say(f"# autotune confirm on {W_conf} window(s) {_key(o)}: {confirmed[_key(o)]:.3e} firings/s")
                    best = max(rivals, key=lambda o: long_rate[_key(o)])
                finally:
                    m.close()
    finally:
        for m in models:
            m.close()
    if best is None:
        raise CtlError(CTL_E_LIMIT, f"autotune: none of the candidates {tuple(candidates)} runs this model")
    res = {"lanes_per_sim": int(best["lanes_per_sim"]), "options": best, "scores": scores, "confirm": confirmed,
           "probe": {"sims": probe_sims, "windows": W, "t_end": T / W, "window_starts": [T * w / W for w in range(W)],
                     "calibration": calib, "confirm_windows": W_conf if confirmed else None,
                     "seconds": time.perf_counter() - t0}}
    _memo[key] = res
    if cache_file:
        disk = {}
        if os.path.exists(cache_file):
            with open(cache_file) as f:
                disk = json.load(f)
        disk[disk_key] = res
        with open(cache_file, "w") as f:
            json.dump(disk, f, indent=1)
    return res


def tune_batch(nets, replicas, *, gpu=0, probe_fraction=None, min_probe_ms=300.0, max_probe_replicas=8192,
               seed=0x7E58, verbose=False, log=None) -> dict:
    """Lanes per simulation for every model of a batch (for Batch(...,
    lanes=...)).  nets: dicts with left, right, x0, c, ts (and optionally
    feed, cs), replicas: per model.  The models are grouped by the library's
    size rule (the kernel class each would run in); for each class the model
    with the most estimated work (replicas x reactions) is probed as a
    one-model batch with at most max_probe_replicas replicas, at the class's
    G and at G/2 and 2G (where the model fits), and the fastest G is applied
    to every model of the class that fits it.  The probe runs the first
    probe_fraction of the model's horizon; by default (None) the fraction
    starts at 1/16 and grows (x4 or more) until the probe at the class's G
    takes min_probe_ms or covers the whole horizon, so that short runs are
    probed whole.  Returns {"lanes": int32 array, "classes": {rule G:
    {"model", "scores", "chosen", "fraction"}}}."""
    say = log or ((lambda m: print(m, flush=True)) if verbose else (lambda m: None))
    reps = np.asarray(replicas, np.int64)

    def model_of(n, ts=None):
        return Model.from_arrays(n["left"], n["right"], n["x0"], n["c"], n["ts"] if ts is None else ts,
                                 feed=n.get("feed"), cs=n.get("cs"))

    models = [model_of(n) for n in nets]
    try:
        with Batch(models, reps, gpu=gpu) as b:
            rule = np.array([b.model_lanes(i) for i in range(len(models))], np.int32)
    finally:
        for m in models:
            m.close()
    lanes = rule.copy()
    out = {}
    for G in sorted(set(int(g) for g in rule)):
        idx = np.nonzero(rule == G)[0]
        work = reps[idx] * np.array([np.asarray(nets[i]["left"]).shape[0] for i in idx])
        rep = int(idx[np.argmax(work)])
        n = nets[rep]
        S, R = np.asarray(n["left"]).shape[1], np.asarray(n["left"]).shape[0]
        ts = np.asarray(n["ts"], np.float64)
        nrep = int(min(reps[rep], max_probe_replicas))
        cand = [g for g in (G // 2, G, 2 * G) if 1 <= g <= 64]
        scores = {}

        def probe(g, frac):
            t_probe = ts[ts <= ts[-1] * frac * (1 + 1e-9)]
            m = model_of(n, ts=t_probe if len(t_probe) else ts[:1])
            try:
                with Batch([m], [nrep], gpu=gpu, seed=seed, lanes=[g]) as b:
                    b.run()
                    st = b.stats()
            finally:
                m.close()
            return st["firings"] / max(st["kernel_ms"], 1e-3) * 1e3, st["kernel_ms"]

        frac = probe_fraction if probe_fraction is not None else 1.0 / 16
        rate, ms = probe(G, frac)
        while probe_fraction is None and ms < min_probe_ms and frac < 1.0:
            frac = min(1.0, frac * max(4.0, 1.2 * min_probe_ms / max(ms, 1e-3)))
            rate, ms = probe(G, frac)
        scores[G] = rate
        say(f"# tune_batch class {G} (model {rep}, N={S}, M={R}, {frac:g} of the horizon): G={G} {rate:.3e} firings/s")
        for g in cand:
            if g == G:
                continue
            try:
                scores[g], _ = probe(g, frac)
            except CtlError as e:             # a G the model's LDS does not allow
                say(f"# tune_batch class {G}: G={g} not possible ({e})")
                continue
            say(f"# tune_batch class {G} (model {rep}, N={S}, M={R}, {frac:g} of the horizon): G={g} "
                f"{scores[g]:.3e} firings/s")
        chosen = max(scores, key=scores.get) if scores else G
        if chosen != G:
            # a group's LDS grows with species and reactions, so the chosen G
            # fits every model of the class no larger than the probed one
            for i in idx:
                Ri, Si = np.asarray(nets[i]["left"]).shape
                if chosen == 1 or (Si <= S and Ri <= R):
                    lanes[i] = chosen
        out[G] = {"model": rep, "scores": scores, "chosen": int(chosen), "fraction": frac}
    return {"lanes": lanes, "classes": out}
