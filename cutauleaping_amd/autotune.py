"""Launch-shape autotuner for the simulation kernels.

The library's defaults (lanes per simulation by a size rule, regrouping
interval, SSA chunk, workgroup size, iterations per launch) are rules of
thumb.  tune() instead times short probe runs of the actual model on the
actual device and keeps the fastest configuration:

1. Windows across the horizon.  One run of the library's default over
   `window_sims` simulations records every species at the window start
   times w * T / windows (w = 1 .. windows-1).  Probe window w restarts the
   model from those states (window 0 from the initial state) and runs
   `probe_fraction` of the horizon: a mass-action network is
   time-homogeneous, so a window is the model from that point on.  Early
   transients and late quasi-stationary phases (often SSA-heavy) both count.
2. Stage 1 times every lanes-per-simulation candidate (1, and G = 2..64 where
   the model fits the group kernel's LDS) with the library's defaults.
3. Stage 2 refines the winner: for G = 2..32 the regrouping interval
   (compact_interval, including off) and the SSA chunk, then a workgroup of
   two wavefronts instead of four; for G = 1 the workgroup size with and
   without regrouping, then the SSA chunk; for every G the iteration budget
   per launch.
4. For G = 2..32, the best configuration with regrouping and the best
   without are timed again on the first two windows made `confirm` times
   longer: regrouping pays once simulations have drifted apart in mode,
   which short windows under-represent (on the 100x200 network the short
   windows favour no regrouping, whole runs favour it by 10 %).

A configuration's score is its total firings over its total kernel time
across the windows (firings per second of the whole probe).  Results are
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

from . import Batch, CtlError, Model, Sim

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


def _window_starts(left, right, x0, c, ts, feed, windows, window_sims, seed, gpu):
    """Initial-state rows of each probe window: x0 for window 0, then the
    states of `window_sims` default-configuration simulations at w*T/windows."""
    starts = [np.atleast_2d(np.asarray(x0))[:1].astype(np.uint32)]
    if windows <= 1:
        return starts
    T = float(ts[-1])
    t_w = np.array([T * w / windows for w in range(1, windows)])
    m = Model.from_arrays(left, right, np.atleast_2d(x0)[:1], np.atleast_2d(c)[:1], t_w, feed=feed)
    try:
        with Sim(m, 64, max(1, window_sims // 64), gpu=gpu, seed=seed ^ 0x5A5A) as s:
            s.run()
            X = s.samples()                          # [sims][windows-1][S], every species
    finally:
        m.close()
    starts += [np.ascontiguousarray(X[:, w, :]) for w in range(windows - 1)]
    return starts


def tune(left, right, x0, c, ts, *, feed=None, cs=None, gpu=0, probe_sims=65536, probe_fraction=0.02,
         windows=4, window_sims=1024, candidates=CANDIDATES, refine=True, confirm=5, seed=0x7E57,
         cache_file=None, verbose=False) -> dict:
    """Returns {"lanes_per_sim": G, "options": Sim keyword arguments of the
    fastest configuration, "scores": {configuration: firings/s on the probe
    windows}, "confirm": {configuration: firings/s on the longer windows}
    (empty when stage 4 did not run), "probe": {...}}."""
    left = np.asarray(left)
    R, S = left.shape
    ts = np.asarray(ts, np.float64)
    sig = signature(left, right, x0, c, ts, feed)
    key = (sig, gpu, probe_sims, probe_fraction, windows, tuple(candidates), refine, confirm)
    if key in _memo:
        return _memo[key]
    disk_key = (f"{sig}@{gpu}:{probe_sims}:{probe_fraction}:{windows}:{','.join(map(str, candidates))}:"
                f"{int(refine)}:{confirm}")
    if cache_file and os.path.exists(cache_file):
        with open(cache_file) as f:
            hit = json.load(f).get(disk_key)
        if hit:
            _memo[key] = hit
            return hit
    say = (lambda m: print(m, flush=True)) if verbose else (lambda m: None)
    window_sims = max(64, window_sims // 64 * 64)
    probe_sims = max(window_sims, probe_sims // window_sims * window_sims)    # whole replica groups per row
    t0 = time.perf_counter()
    starts = _window_starts(left, right, x0, c, ts, feed, windows, window_sims, seed, gpu)
    dt = float(ts[-1]) * probe_fraction
    t_probe = ts[ts <= dt + 1e-12]
    if len(t_probe) == 0:
        t_probe = np.array([dt])
    models = [Model.from_arrays(left, right, st, np.atleast_2d(c)[:1], t_probe, feed=feed, cs=cs) for st in starts]
    say(f"# autotune: {len(models)} windows of {dt:g} time units, {probe_sims} simulations each "
        f"({time.perf_counter() - t0:.1f} s to set up)")
    scores = {}
    confirmed = {}

    def score(opts: dict) -> float:
        k = _key(opts)
        if k in scores:
            return scores[k]
        fir, ms = 0, 0.0
        t1 = time.perf_counter()
        for i, m in enumerate(models):
            try:
                with Sim(m, 64, probe_sims // 64, gpu=gpu, seed=seed + i, **opts) as s:
                    s.run()
                    st = s.stats()
            except CtlError as e:        # a shape the model does not allow (LDS)
                say(f"# autotune {k}: not possible ({e})")
                scores[k] = 0.0
                return 0.0
            fir += st["firings"]
            ms += st["kernel_ms"]
        rate = fir / max(ms, 1e-3) * 1e3
        scores[k] = rate
        say(f"# autotune {k}: {rate:.3e} firings/s ({ms:.1f} ms kernel, {time.perf_counter() - t1:.1f} s wall)")
        return rate

    try:
        best, best_rate = None, -1.0
        for G in candidates:
            if G > 1 and not _fits_group(S, R):
                continue
            o = {"lanes_per_sim": G}
            r = score(o)
            if r > best_rate:
                best, best_rate = o, r
        if refine and best is not None:
            G = best["lanes_per_sim"]
            if 1 < G < 64:
                grid = [{"compact_interval": -1}] + [{"compact_interval": ci, "ssa_chunk": ch}
                                                     for ci in (4, 8, 16) for ch in (3, 6, 12)]
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
            if confirm > 1 and 1 < G < 64:
                # Regrouping pays once simulations drift apart in mode, which short
                # windows under-represent: the best configuration with and the best
                # without it are timed again on windows `confirm` times longer.
                def regroups(o):
                    ci = o.get("compact_interval", 0)
                    return ci > 0 or (ci == 0 and G >= 16 and not o.get("force_ssa"))
                rivals = {}
                for k, r in scores.items():
                    o = {kv.split("=")[0]: int(kv.split("=")[1]) for kv in k.split(",")}
                    if o.get("lanes_per_sim") != G or r <= 0:
                        continue
                    side = regroups(o)
                    if side not in rivals or r > scores[_key(rivals[side])]:
                        rivals[side] = o
                if len(rivals) == 2:
                    t_long = ts[ts <= float(ts[-1]) * probe_fraction * confirm + 1e-12]
                    if len(t_long) == 0:
                        t_long = np.array([float(ts[-1]) * probe_fraction * confirm])
                    long_models = [Model.from_arrays(left, right, st, np.atleast_2d(c)[:1], t_long, feed=feed, cs=cs)
                                   for st in starts[:2]]
                    try:
                        long_rate = {}
                        for side, o in rivals.items():
                            fir, ms = 0, 0.0
                            for i, m in enumerate(long_models):
                                with Sim(m, 64, probe_sims // 64, gpu=gpu, seed=seed + i, **o) as s:
                                    s.run()
                                    st = s.stats()
                                fir += st["firings"]
                                ms += st["kernel_ms"]
                            long_rate[side] = fir / max(ms, 1e-3) * 1e3
                            confirmed[_key(o)] = long_rate[side]
                            say(f"# autotune confirm {_key(o)}: {long_rate[side]:.3e} firings/s")
                        best = rivals[max(long_rate, key=long_rate.get)]
                    finally:
                        for m in long_models:
                            m.close()
    finally:
        for m in models:
            m.close()
    res = {"lanes_per_sim": int(best["lanes_per_sim"]), "options": best, "scores": scores, "confirm": confirmed,
           "probe": {"sims": probe_sims, "fraction": probe_fraction, "t_end": float(t_probe[-1]),
                     "windows": len(starts), "window_starts": [float(ts[-1]) * w / windows for w in range(len(starts))],
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


def tune_batch(nets, replicas, *, gpu=0, probe_fraction=0.1, max_probe_replicas=8192, seed=0x7E58,
               verbose=False, log=None) -> dict:
    """Lanes per simulation for every model of a batch (for Batch(...,
    lanes=...)).  nets: dicts with left, right, x0, c, ts (and optionally
    feed, cs), replicas: per model.  The models are grouped by the library's
    size rule (the kernel class each would run in); for each class the model
    with the most estimated work (replicas x reactions) is probed as a
    one-model batch over the first probe_fraction of its horizon, with at
    most max_probe_replicas replicas, at the class's G and at G/2 and 2G
    (where the model fits), and the fastest G is applied to every model of
    the class that fits it.  Returns {"lanes": int32 array, "classes":
    {rule G: {"model", "scores", "chosen"}}}."""
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
        t_probe = ts[ts <= ts[-1] * probe_fraction + 1e-12]
        if len(t_probe) == 0:
            t_probe = ts[:1]
        nrep = int(min(reps[rep], max_probe_replicas))
        cand = [g for g in (G // 2, G, 2 * G) if 1 <= g <= 64]
        scores = {}
        m = model_of(n, ts=t_probe)
        try:
            for g in cand:
                try:
                    with Batch([m], [nrep], gpu=gpu, seed=seed, lanes=[g]) as b:
                        b.run()
                        st = b.stats()
                except CtlError as e:             # a G the model's LDS does not allow
                    say(f"# tune_batch class {G}: G={g} not possible ({e})")
                    continue
                scores[g] = st["firings"] / max(st["kernel_ms"], 1e-3) * 1e3
                say(f"# tune_batch class {G} (model {rep}, N={S}, M={R}): G={g} {scores[g]:.3e} firings/s")
        finally:
            m.close()
        chosen = max(scores, key=scores.get) if scores else G
        if chosen != G:
            # a group's LDS grows with species and reactions, so the chosen G
            # fits every model of the class no larger than the probed one
            for i in idx:
                Ri, Si = np.asarray(nets[i]["left"]).shape
                if chosen == 1 or (Si <= S and Ri <= R):
                    lanes[i] = chosen
        out[G] = {"model": rep, "scores": scores, "chosen": int(chosen)}
    return {"lanes": lanes, "classes": out}
