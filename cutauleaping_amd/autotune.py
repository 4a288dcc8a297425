"""Launch-shape autotuner for the simulation kernels.

The library's own choice of lanes per simulation (ctl_opts.lanes_per_sim =
0) is a size rule.  tune() instead times short probe runs of the actual
model on the actual device for every candidate and keeps the fastest:

* lanes per simulation: 1 (one lane per simulation, tau_leap_kernel) and
  G = 2..64 (G lanes per simulation, group_kernel<G>);
* for each, the library's default resident-slot count and iteration budget.

A probe runs `probe_sims` simulations (default: enough to fill the device
for the widest candidate) over the first `probe_fraction` of the sample
times and scores reaction firings per second of kernel time.  Results are
memoised per (model signature, device) in-process and optionally in a JSON
file.  Different lanes-per-simulation choices are different (statistically
equivalent) algorithms: the trajectories of a seed depend on the choice,
which is therefore reported with every result (ctl_stats.lanes_per_sim).
"""
from __future__ import annotations

import hashlib
import json
import os
import time

import numpy as np

from . import Model, Sim

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


def tune(left, right, x0, c, ts, *, feed=None, cs=None, gpu=0, probe_sims=65536, probe_fraction=0.05,
         candidates=CANDIDATES, seed=0x7E57, cache_file=None, verbose=False) -> dict:
    """Returns {"lanes_per_sim": G, "scores": {G: firings/s}, "probe": {...}}."""
    left = np.asarray(left)
    R, S = left.shape
    ts = np.asarray(ts, np.float64)
    key = (signature(left, right, x0, c, ts, feed), gpu)
    if key in _memo:
        return _memo[key]
    if cache_file and os.path.exists(cache_file):
        with open(cache_file) as f:
            disk = json.load(f)
        hit = disk.get(f"{key[0]}@{gpu}")
        if hit:
            hit["scores"] = {int(k): v for k, v in hit["scores"].items()}
            _memo[key] = hit
            return hit
    t_probe = ts * probe_fraction
    probe_model = Model.from_arrays(left, right, np.atleast_2d(x0)[:1], np.atleast_2d(c)[:1], t_probe,
                                    feed=feed, cs=cs)
    scores = {}
    blocks = max(1, probe_sims // 256)
    for G in candidates:
        if G > 1 and not _fits_group(S, R):
            continue
        t0 = time.perf_counter()
        with Sim(probe_model, 256, blocks, gpu=gpu, seed=seed, lanes_per_sim=G) as s:
            s.run()
            st = s.stats()
        rate = st["firings"] / max(st["kernel_ms"], 1e-3) * 1e3
        scores[G] = rate
        if verbose:
            print(f"# autotune G={G}: {rate:.3e} firings/s ({st['kernel_ms']:.1f} ms kernel, "
                  f"{time.perf_counter() - t0:.1f} s wall)", flush=True)
    probe_model.close()
    best = max(scores, key=scores.get)
    res = {"lanes_per_sim": int(best), "scores": scores,
           "probe": {"sims": blocks * 256, "fraction": probe_fraction, "t_end": float(t_probe[-1])}}
    _memo[key] = res
    if cache_file:
        disk = {}
        if os.path.exists(cache_file):
            with open(cache_file) as f:
                disk = json.load(f)
        disk[f"{key[0]}@{gpu}"] = res
        with open(cache_file, "w") as f:
            json.dump(disk, f, indent=1)
    return res
