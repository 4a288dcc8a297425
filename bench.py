"""Benchmark: reaction firings/sec of adaptive tau-leaping on one MI355X.

Workload (the 1-GPU configuration of the headline metric): a seeded synthetic
mass-action network with N=100 species and M=200 reactions (order 0-2,
1-2 products, rates log-uniform in [1e-6, 1], initial counts log-uniform in
[10, 1e5], mean-field int32-safety rejection), 2^20 simulations of identical
initial state and rates with distinct Philox streams, 200 samples to t_end=50.

The 2^20 simulations are processed in batches of 2^18 (--batch): a "step" is
one complete ctl_sim_run over one batch -- every simulation of the batch from
t=0 to t_end, inputs resident in HBM -- and step k takes the simulations
[k*batch, (k+1)*batch) of the workload (their global indices select their
Philox streams).  --steps 4 covers the whole workload; the default (one
step) keeps the GPU busy for about a minute.  A batch of 2^18 already holds
the device full (every resident group slot runs about 20 simulations in
turn), so the rate is that of the whole workload.
value = total firings / timed seconds (firings = accepted-leap reaction
counts + SSA events with time <= t_end).

Printed JSON also carries:
  roofline     VALU lane-op roofline of the simulation kernel: algorithmic
               lane-ops per launch (per-event charges below, from the kernel's
               own counters) / average launch time (HIP events on the library's
               stream), against 256 CU x 64 lanes x 2.4 GHz.
  cpu_baseline the oracle's reference-type restatement (OpenMP) on a bounded
               sample of the same workload, timed on this host.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--sims N] [--batch B] [--t-end T] [--lanes G]
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import numpy as np  # noqa: E402

VALU_PEAK = 256 * 64 * 2.4e9          # lane-ops/s (MI355X: 256 CUs, 64 lanes/clk, 2.4 GHz)
HBM_PEAK = 8.0e12                     # bytes/s


def lane_ops(stats, S, R, nL, nV, nF, SV):
    """Algorithmic VALU lane-ops of a run, from the per-event counters.
    leap attempt (accepted or switched): propensities + criticality (2R + 5 nL),
    mu/sigma (5 nV), tau candidates (8 S_V), a0c (R); every Poisson draw and
    every exponential gap of the small-rate Poisson process 40; every update
    pass 2 nF; SSA step: propensities (2R + 3 nL), search (R), uniforms + log
    (40), dense update (S)."""
    attempts = stats["leaps"] + stats["switches"]
    w = attempts * (2 * R + 5 * nL + 5 * nV + 8 * SV)
    w += stats["leaps"] * R
    w += (stats["poisson"] + stats["gaps"]) * 40
    w += (stats["leaps"] + stats["rejections"]) * 2 * nF
    w += stats["ssa_steps"] * (2 * R + 3 * nL + R + 40 + S)
    return float(w)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--sims", type=int, default=1 << 20, help="simulations in the workload")
    ap.add_argument("--batch", type=int, default=1 << 18, help="simulations per step")
    ap.add_argument("--t-end", type=float, default=50.0)
    ap.add_argument("--samples", type=int, default=200)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-sims", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--block-size", type=int, default=0)
    ap.add_argument("--slots", type=int, default=0)
    ap.add_argument("--iters", type=int, default=0)
    ap.add_argument("--ssa-chunk", type=int, default=0)
    ap.add_argument("--lanes", type=int, default=0, help="lanes per simulation (0: library's choice)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only copy BYTES device-to-device once (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (simulations are independent; "
              "multi-GPU sharding by input offset is documented, not launched).", file=sys.stderr)
        sys.exit(2)

    import cutauleaping_amd as ctl
    from cutauleaping_amd import networks

    if not os.path.exists(ctl.LIB_PATH):
        ctl.build()
    t0 = time.time()
    net = networks.mass_action(100, 200, seed=args.seed, t_end=args.t_end)
    ts = networks.sample_times(args.samples, args.t_end)
    print(f"# network seed={args.seed} attempt={net['attempt']} generated in {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)
    model = ctl.Model.from_arrays(net["left"], net["right"], net["x0"], net["c"], ts)
    info = model.info()
    st = model.structure()
    SV = int(len(np.unique(st["rav_flat"][:, 1]))) if info.nV else 0
    threads = 256
    batch = min(args.batch, args.sims)
    blocks = (batch + threads - 1) // threads
    n = threads * blocks
    nbatches = max(1, args.sims // n)
    sims = [ctl.Sim(model, threads, blocks, seed=0xC0FFEE, offset=(k % nbatches) * n, block_size=args.block_size,
                    slots=args.slots, iters_per_launch=args.iters, ssa_chunk=args.ssa_chunk,
                    lanes_per_sim=args.lanes, verbose=True)
            for k in range(min(nbatches, max(args.steps, args.warmup, 1)))]
    for w in range(args.warmup):
        sims[w % len(sims)].run()
        print(f"# warmup {w} done", file=sys.stderr, flush=True)
    fir = 0
    ker_ms = 0.0
    launches = 0
    work = 0.0
    t_start = time.perf_counter()
    for k in range(args.steps):
        sim = sims[k % len(sims)]
        sim.run()
        s = sim.stats()
        fir += s["firings"]
        ker_ms += s["kernel_ms"]
        launches += s["launches"]
        work += lane_ops(s, info.S, info.R, info.nL, info.nV, info.nF, SV)
        print(f"# step {k}: firings={s['firings']} leaps={s['leaps']} ssa_steps={s['ssa_steps']} "
              f"poisson={s['poisson']} gaps={s['gaps']} kernel_ms={s['kernel_ms']:.1f} launches={s['launches']}",
              file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t_start
    last = s
    sim_s = n * args.steps / elapsed
    kname = "tau_leap_kernel" if last["lanes_per_sim"] == 1 else f"group_kernel<{last['lanes_per_sim']}>"
    avg_launch_s = ker_ms / 1000.0 / max(launches, 1)
    achieved = work / args.steps / max(last["launches"], 1) / avg_launch_s
    hbm_bytes = n * (4 * info.S + 4 * info.SS * info.K + 64)   # inputs + samples + per-sim record
    out = {
        "metric": "reaction firings/sec on 1 GPU",
        "value": fir / elapsed,
        "unit": "firings/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1000.0 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"mass-action N={info.S} M={info.R} seed={args.seed}, {nbatches * n} simulations "
                               f"in batches of {n} (one batch per step), {info.K} samples to "
                               f"t_end={args.t_end:g}, Philox per simulation",
                   "simulations": nbatches * n, "batch": n, "species": info.S, "reactions": info.R,
                   "samples": info.K, "t_end": args.t_end, "seed": args.seed},
        "simulations_per_s": sim_s,
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": VALU_PEAK / 1e12,
                     "unit": "Tlane-op/s", "frac": achieved / VALU_PEAK, "traffic": None,
                     "kernel": kname, "avg_launch_ms": avg_launch_s * 1e3,
                     "launches_per_step": last["launches"],
                     "hbm": {"achieved_GBps": hbm_bytes / (ker_ms / 1000.0 / args.steps) / 1e9,
                             "peak_GBps": HBM_PEAK / 1e9,
                             "frac": hbm_bytes / (ker_ms / 1000.0 / args.steps) / HBM_PEAK}},
        "counters": {k: last[k] for k in ("firings", "leaps", "rejections", "switches", "ssa_steps",
                                          "poisson", "gaps", "absorbed", "lds_state", "slots", "block_size",
                                          "lanes_per_sim")},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(net, ts, args.cpu_sims)
    for x in sims:
        x.close()
    print(json.dumps(out), flush=True)


def copy_probe(nbytes):
    """One device-to-device copy of nbytes through the HIP runtime libctl uses
    (its blit kernel reads and writes exactly nbytes): the reference point
    for scaling FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import ctypes as C
    import cutauleaping_amd as ctl
    ctl.lib()
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    a, b = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), nbytes) or hip.hipMalloc(C.byref(b), nbytes):
        raise RuntimeError("hipMalloc failed")
    try:
        if hip.hipMemset(a, 1, nbytes) or hip.hipMemcpy(b, a, nbytes, 3) or hip.hipDeviceSynchronize():
            raise RuntimeError("copy failed")
    finally:
        hip.hipFree(a)
        hip.hipFree(b)
    print(json.dumps({"copy_probe_bytes": nbytes}), flush=True)


def cpu_baseline(net, ts, nsims):
    """The oracle's reference-type (double) restatement, OpenMP over simulations."""
    import oracle
    threads = min(os.cpu_count() or 1, 16)
    t0 = time.perf_counter()
    _, st = oracle.run(net["left"], net["right"], net["x0"], net["c"], ts, n=nsims, seed=0xC0FFEE,
                       precision=oracle.REFSEM, threads=threads)
    dt = time.perf_counter() - t0
    return {"value": float(st["firings"].sum()) / dt, "unit": "firings/s", "cores": threads,
            "kind": "port", "sample": f"{nsims} simulations of the same workload (full t_end), "
                                      f"{dt:.1f}s, {st['firings'].sum()} firings",
            "simulations_per_s": nsims / dt}


if __name__ == "__main__":
    main()
