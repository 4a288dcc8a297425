"""Benchmark: reaction firings/sec of adaptive tau-leaping on one MI355X.

Workload (the 1-GPU configuration of the headline metric): a seeded synthetic
mass-action network with N=100 species and M=200 reactions (order 0-2,
1-2 products, rates log-uniform in [1e-6, 1], initial counts log-uniform in
[10, 1e5], mean-field int32-safety rejection), 2^20 simulations of identical
initial state and rates with distinct Philox streams, 200 samples to t_end=50.

The 2^20 simulations are processed in slices of 2^16 (--slice): a "step" is
one complete ctl_sim_run over one slice -- every simulation of the slice from
t=0 to t_end, inputs resident in HBM -- and step k takes the simulations
[k*slice, (k+1)*slice) of the workload, k taken modulo the 16 slices (the
global indices select the Philox streams).  Sixteen steps cover the whole
workload, which is the default; warm-up steps run slices too.  A slice of
2^16 already holds the device full (16384 resident group slots of 16 lanes
run four simulations each in turn), so the rate is that of the workload.
value = total firings / timed seconds (firings = accepted-leap reaction
counts + SSA events with time <= t_end).

Printed JSON also carries:
  roofline     VALU lane-op roofline of the simulation kernel: algorithmic
               lane-ops per launch (per-event charges below, from the kernel's
               own counters) / average launch time (HIP events on the library's
               stream), against 256 CU x 64 lanes x 2.4 GHz.
  cpu_baseline the oracle's reference-type restatement (OpenMP) on a bounded
               sample of the same workload, timed on this host.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--sims N] [--slice B] [--t-end T] [--lanes G]
                [--workload c0|c1|c2|c3|c4]
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
# HBM bytes per simulation of this workload measured with rocprofv3 FETCH_SIZE / WRITE_SIZE,
# scaled by a 1 GiB copy (profiles/group_kernel_c1_hbm_traffic.md)
MEASURED_BYTES_PER_SIM = 84023


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


def batch_rooflines(models, batch, reps, secs):
    """VALU and HBM fractions of a batch run over `secs` seconds: lane_ops of
    every model from its own structure and per-simulation counters, plus its
    inputs, samples and per-simulation records, summed over the batch."""
    work, hbm = 0.0, 0.0
    for i, m in enumerate(models):
        info = m.info()
        sd = m.structure()
        SV = int(len(np.unique(sd["rav_flat"][:, 1]))) if info.nV else 0
        ss = batch.sim_stats(i)
        st = {k: int(ss[k].sum()) for k in ("leaps", "switches", "poisson", "gaps", "rejections", "ssa_steps")}
        work += lane_ops(st, info.S, info.R, info.nL, info.nV, info.nF, SV)
        hbm += int(reps[i]) * (4 * info.S + 4 * info.SS * info.K + 64)
    return {"valu": {"achieved": work / secs / 1e12, "peak": VALU_PEAK / 1e12, "unit": "Tlane-op/s",
                     "frac": work / secs / VALU_PEAK},
            "hbm": {"achieved_GBps": hbm / secs / 1e9, "peak_GBps": HBM_PEAK / 1e9, "frac": hbm / secs / HBM_PEAK}}


def model_rooflines(model, stats, kernel_ms, nsims):
    """VALU and HBM fractions of a single-model run: algorithmic lane-ops
    (lane_ops) and bytes (inputs, samples and per-simulation record) over the
    kernel time."""
    info = model.info()
    sd = model.structure()
    SV = int(len(np.unique(sd["rav_flat"][:, 1]))) if info.nV else 0
    work = lane_ops(stats, info.S, info.R, info.nL, info.nV, info.nF, SV)
    secs = kernel_ms / 1e3
    hbm = nsims * (4 * info.S + 4 * info.SS * info.K + 64)
    return {"valu": {"achieved": work / secs / 1e12, "peak": VALU_PEAK / 1e12, "unit": "Tlane-op/s",
                     "frac": work / secs / VALU_PEAK},
            "hbm": {"achieved_GBps": hbm / secs / 1e9, "peak_GBps": HBM_PEAK / 1e9, "frac": hbm / secs / HBM_PEAK}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--sims", type=int, default=1 << 20, help="simulations in the workload")
    ap.add_argument("--slice", type=int, default=1 << 16, help="simulations per step")
    ap.add_argument("--t-end", type=float, default=50.0)
    ap.add_argument("--samples", type=int, default=200)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-sims", type=int, default=4096, help="simulations of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--block-size", type=int, default=0)
    ap.add_argument("--slots", type=int, default=0)
    ap.add_argument("--iters", type=int, default=0)
    ap.add_argument("--ssa-chunk", type=int, default=0)
    ap.add_argument("--lanes", type=int, default=0, help="lanes per simulation (0: library's choice)")
    ap.add_argument("--compact", type=int, default=0,
                    help="iterations between regroupings of simulations by mode (0: library default, <0: off)")
    ap.add_argument("--c0-sims", type=int, default=1024, help="c0: simulations (the configuration runs 1024)")
    ap.add_argument("--workload", default="c1", choices=("c0", "c1", "c1ref", "c2", "c3", "c4", "mlp"),
                    help="c1: the headline (default); c1ref: C1's first slice through the reference's execution "
                         "design; c0: Schlogl; c2: batched uneven models; c3: surrogate-screened sweep; "
                         "c4: enzyme cascade (adaptive and force-SSA); mlp: the surrogate network's kernels alone")
    ap.add_argument("--mlp-rows", type=int, default=65536, help="mlp: rows (parameter sets) per pass")
    ap.add_argument("--models", type=int, default=8192, help="c2: models in the batch")
    ap.add_argument("--no-tune", action="store_true", help="c0, c2, c4: the library's defaults instead of the autotuner")
    ap.add_argument("--c3-train", type=int, default=65536)
    ap.add_argument("--c3-test", type=int, default=4096)
    ap.add_argument("--c3-replicas", type=int, default=256)
    ap.add_argument("--c3-judge", type=int, default=512)
    ap.add_argument("--c3-epochs", type=int, default=60)
    ap.add_argument("--phase-clock", action="store_true",
                    help="c1: run the profiling build (libctl_phases.so) and report the group kernel's time per phase")
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
    if args.workload != "c1":
        if "--steps" not in sys.argv:       # one complete run of the other workloads unless asked
            args.steps = 1
        if "--warmup" not in sys.argv:
            args.warmup = 0
        {"c0": bench_c0, "c1ref": bench_c1ref, "c2": bench_c2, "c3": bench_c3, "c4": bench_c4,
         "mlp": bench_mlp}[args.workload](args)
        return

    import cutauleaping_amd as ctl
    from cutauleaping_amd import networks

    if args.phase_clock:                 # the profiling build: same kernels plus the phase clock
        ctl.LIB_PATH = os.path.join(os.path.dirname(ctl.LIB_PATH), "libctl_phases.so")
    if not os.path.exists(ctl.LIB_PATH):
        raise SystemExit(f"bench.py: {ctl.LIB_PATH} is not built (python __graft_entry__.py build)")
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
    n = max(threads, min(args.slice, args.sims) // threads * threads)   # simulations per step
    nslices = max(1, args.sims // n)
    sims = {}

    def slice_sim(k):
        """The Sim of slice k: simulations [k*n, (k+1)*n) of the workload,
        their global indices (offset) selecting their Philox streams."""
        k %= nslices
        if k not in sims:
            sims[k] = ctl.Sim(model, threads, n // threads, seed=0xC0FFEE, offset=k * n,
                              block_size=args.block_size, slots=args.slots, iters_per_launch=args.iters,
                              ssa_chunk=args.ssa_chunk, lanes_per_sim=args.lanes,
                              compact_interval=args.compact)
        return sims[k]

    for k in range(min(nslices, args.steps + args.warmup)):
        slice_sim(k)                     # device allocation and upload, before the timed region
    for w in range(args.warmup):
        slice_sim(w).run()
        print(f"# warmup {w} done (slice {w % nslices})", file=sys.stderr, flush=True)
    fir = 0
    ker_ms = 0.0
    launches = 0
    work = 0.0
    timed = 0
    clock = np.zeros(16, np.uint64)
    if args.phase_clock:
        ctl._check(ctl.lib().ctl_phase_clock(0, clock.ctypes.data, 16, 1))     # zero the totals
    t_start = time.perf_counter()
    for k in range(args.steps):
        sim = slice_sim(k)
        sim.run()
        s = sim.stats()
        fir += s["firings"]
        ker_ms += s["kernel_ms"]
        launches += s["launches"]
        timed += n
        work += lane_ops(s, info.S, info.R, info.nL, info.nV, info.nF, SV)
        print(f"# step {k} (slice {k % nslices}): firings={s['firings']} leaps={s['leaps']} "
              f"ssa_steps={s['ssa_steps']} poisson={s['poisson']} gaps={s['gaps']} "
              f"kernel_ms={s['kernel_ms']:.1f} launches={s['launches']}", file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t_start
    last = s
    kname = "tau_leap_kernel" if last["lanes_per_sim"] == 1 else f"group_kernel<{last['lanes_per_sim']}>"
    avg_launch_s = ker_ms / 1000.0 / max(launches, 1)
    achieved = work / max(launches, 1) / avg_launch_s
    hbm_bytes = timed * (4 * info.S + 4 * info.SS * info.K + 64)   # inputs + samples + per-sim record
    covered = min(args.steps, nslices) * n
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
        "config": {"workload": f"mass-action N={info.S} M={info.R} seed={args.seed}, {nslices * n} simulations "
                               f"in {nslices} slices of {n}; step k runs slice k mod {nslices} from t=0 to "
                               f"t_end={args.t_end:g} ({info.K} samples), Philox per simulation",
                   "simulations": nslices * n, "simulations_per_step": n,
                   "simulations_timed": timed, "distinct_simulations_timed": covered,
                   "species": info.S, "reactions": info.R,
                   "samples": info.K, "t_end": args.t_end, "seed": args.seed},
        "simulations_per_s": timed / elapsed,
        "firings_timed": fir,
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": VALU_PEAK / 1e12,
                     "unit": "Tlane-op/s", "frac": achieved / VALU_PEAK,
                     "traffic": (MEASURED_BYTES_PER_SIM * timed / max(launches, 1)
                                 if last["lanes_per_sim"] > 1 and args.t_end == 50.0 and args.samples == 200 else None),
                     "traffic_unit": "bytes per launch",
                     "kernel": kname, "avg_launch_ms": avg_launch_s * 1e3,
                     "launches": launches,
                     "hbm": {"achieved_GBps": hbm_bytes / (ker_ms / 1000.0) / 1e9,
                             "peak_GBps": HBM_PEAK / 1e9,
                             "frac": hbm_bytes / (ker_ms / 1000.0) / HBM_PEAK}},
        "counters": {k: last[k] for k in ("firings", "leaps", "rejections", "switches", "ssa_steps",
                                          "poisson", "gaps", "absorbed", "lds_state", "slots", "block_size",
                                          "lanes_per_sim")},
    }
    if args.phase_clock:
        ctl._check(ctl.lib().ctl_phase_clock(0, clock.ctypes.data, 16, 0))
        tot = float(clock.sum())
        out["phases"] = {"unit": "share of the wavefronts' shader-clock ticks", "total_ticks": int(tot),
                         "ticks_per_leap": tot / max(last["leaps"] * args.steps, 1),
                         "share": {name: float(clock[i]) / max(tot, 1.0) for i, name in enumerate(PHASES)}}
    for x in sims.values():
        x.close()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(net, ts, args.cpu_sims)
    print(json.dumps(out), flush=True)


# phases of the group kernel's phase clock (ctl_phase_clock), in its order
PHASES = ("propensities", "tau1", "switch_tau2_pick", "counts_gaps", "poisson_queue", "update", "sampling",
          "ssa_open", "ssa_select", "ssa_update", "ssa_dependents", "before_leap", "before_ssa", "loop_fetch_finish",
          "regroup", "park")


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


_AFFINITY = os.sched_getaffinity(0) if hasattr(os, "sched_getaffinity") else None


def _restore_affinity():
    """OpenMP with OMP_PROC_BIND binds the calling thread to its first place
    (one core) for the rest of the process: give it back the CPUs it started
    with, so that later GPU runs and baselines see the whole share."""
    if _AFFINITY is not None:
        os.sched_setaffinity(0, _AFFINITY)


def host_cores():
    """Threads for the CPU baseline: the CPUs this process may run on (as
    at start-up), at most 16 (a one-GPU share of the host), with nproc and
    the lscpu model."""
    avail = len(_AFFINITY) if _AFFINITY is not None else (os.cpu_count() or 1)
    model, sockets, cps, tpc = "unknown", None, None, None
    try:
        import subprocess
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in txt.splitlines():
            k, _, v = line.partition(":")
            v = v.strip()
            if k.strip() == "Model name":
                model = v
            elif k.strip() == "Socket(s)":
                sockets = int(v) if v.isdigit() else None
            elif k.strip() == "Core(s) per socket":
                cps = int(v) if v.isdigit() else None
            elif k.strip() == "Thread(s) per core":
                tpc = int(v) if v.isdigit() else None
    except (OSError, ValueError, Exception):
        pass
    return min(avail, 16), {"nproc": avail, "cpu_count": os.cpu_count(), "lscpu_model": model,
                            "sockets": sockets, "cores_per_socket": cps, "threads_per_core": tpc}


def cpu_baseline(net, ts, nsims, locked=None, cs=None, force_ssa=False):
    """The oracle's reference-type (double) restatement of the per-simulation
    algorithm, OpenMP over simulations (dynamic schedule, one simulation per
    chunk), threads bound to cores (OMP_PROC_BIND=close, OMP_PLACES=cores)."""
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    import oracle
    threads, host = host_cores()
    t0 = time.perf_counter()
    try:
        _, st = oracle.run(net["left"], net["right"], net["x0"], net["c"], ts, n=nsims, seed=0xC0FFEE,
                           locked=locked, cs=cs, force_ssa=force_ssa, precision=oracle.REFSEM, threads=threads)
    finally:
        _restore_affinity()
    dt = time.perf_counter() - t0
    return {"value": float(st["firings"].sum()) / dt, "unit": "firings/s", "cores": threads,
            "kind": "port", "sample": f"{nsims} simulations of the same workload (full t_end), "
                                      f"{dt:.3f}s, {int(st['firings'].sum())} firings",
            "simulations_per_s": nsims / dt, "host": host}


def _line(metric, value, unit, steps, ms, dtype, config, **extra):
    out = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps, "warmup": 0,
           "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
           "data": "synthetic", "config": config}
    out.update(extra)
    print(json.dumps(out), flush=True)


def bench_c0(args):
    """Schlogl (the reference repository's CPU-runnable configuration): the
    four-reaction model of tests/golden/schlogl (A and B buffered), 1024
    simulations (--c0-sims) to t_end=10 with 100 samples, with the
    autotuner's launch shape (untimed; --no-tune, --lanes or --compact for
    the library's: one lane per simulation with its state in LDS); a step is
    one complete run of the simulations."""
    import cutauleaping_amd as ctl
    left = [[1, 2, 0], [0, 3, 0], [0, 0, 1], [0, 1, 0]]
    right = [[0, 3, 0], [1, 2, 0], [0, 1, 0], [0, 0, 1]]
    net = {"left": left, "right": right, "x0": [100000, 250, 200000], "c": [3e-7, 1e-4, 1e-3, 3.5]}
    ts = [round(0.1 * i, 10) for i in range(1, 101)]
    model = ctl.Model.from_arrays(left, right, net["x0"], net["c"], ts, feed=[1, 0, 1])
    n = max(64, args.c0_sims // 64 * 64)
    steps = max(args.steps, 1)
    fir, kms, launches = 0, 0.0, 0
    opts, tuned = {"lanes_per_sim": args.lanes, "compact_interval": args.compact}, None
    if args.lanes == 0 and args.compact == 0 and not args.no_tune:    # untimed: the autotuner's launch shape
        from cutauleaping_amd import autotune
        t0 = time.time()
        r = autotune.tune(left, right, net["x0"], net["c"], ts, feed=[1, 0, 1], probe_sims=n, window_sims=n)
        opts = dict(r["options"])
        tuned = {"options": opts, "seconds": round(time.time() - t0, 2), "confirm": r["confirm"]}
        print(f"# autotuned in {tuned['seconds']} s: {opts}", file=sys.stderr, flush=True)
    with ctl.Sim(model, 64, n // 64, seed=0xC0FFEE, **opts) as s:
        for _ in range(args.warmup):
            s.run()
        t1 = time.perf_counter()
        for _ in range(steps):
            s.run()
            st = s.stats()
            fir += st["firings"]
            kms += st["kernel_ms"]
            launches += st["launches"]
        dt = time.perf_counter() - t1
    extra = {"simulations_per_s": n * steps / dt, "kernel_ms_per_step": kms / steps,
             "counters": {k: st[k] for k in ("leaps", "ssa_steps", "switches", "poisson", "gaps", "launches",
                                             "slots", "block_size", "lds_state", "lanes_per_sim")},
             "roofline": model_rooflines(model, st, kms / steps, n)}
    if tuned:
        extra["autotune"] = tuned
    if not args.no_cpu:
        # about 10 s of host work: the rate per firing does not depend on how
        # many simulations run, so the sample is 256 times the GPU's 1024
        extra["cpu_baseline"] = cpu_baseline(net, ts, 256 * n, locked=[1, 0, 1])
    _line("reaction firings/sec on 1 GPU (Schlogl)", fir / dt, "firings/s", steps, dt * 1e3 / steps, "f32",
          {"workload": f"Schlogl (X with buffered A=1e5, B=2e5), {n} simulations, 100 samples to t_end=10",
           "simulations": n, "t_end": 10.0, "options": opts}, **extra)
    model.close()


def bench_c2(args):
    """Batched uneven models: networks.batched_models(models, seed=2) -- N
    log-uniform in 10..500, M in 10..1000, 64..4096 replicas each, rates and
    counts as the headline network, 10 sampled species, 100 samples to
    t_end -- in one batch: each model runs with the library's choice of lanes
    per simulation, the models of one choice form a kernel class, and the
    classes run concurrently.  A step is one complete run of the batch."""
    import cutauleaping_amd as ctl
    from cutauleaping_amd import networks
    t_end = args.t_end if args.t_end != 50.0 else 0.1
    t0 = time.time()
    procs = max(1, min(16, host_cores()[0]))
    nets = networks.batched_models(args.models, seed=2, t_end=t_end, processes=procs,
                                   log=lambda m: print(m, file=sys.stderr, flush=True))
    print(f"# {len(nets)} models generated in {time.time() - t0:.1f}s on {procs} processes", file=sys.stderr,
          flush=True)
    ts = networks.sample_times(100, t_end)
    reps = [n["replicas"] for n in nets]
    tuned = None
    if args.lanes == 0 and not args.no_tune:
        from cutauleaping_amd import autotune
        t0 = time.time()
        for n in nets:
            n["ts"] = ts
        tuned = autotune.tune_batch(nets, reps, log=lambda m: print(m, file=sys.stderr, flush=True))
        tuned["seconds"] = time.time() - t0
        print(f"# autotuned lanes per class in {tuned['seconds']:.1f}s: "
              f"{ {g: c['chosen'] for g, c in tuned['classes'].items()} }", file=sys.stderr, flush=True)
    models = [ctl.Model.from_arrays(n["left"], n["right"], n["x0"], n["c"], ts, cs=n["cs"]) for n in nets]
    fir = 0
    ms = 0.0
    with ctl.Batch(models, reps, seed=0xC0FFEE, lanes_per_sim=args.lanes, verbose=True,
                   lanes=None if tuned is None else tuned["lanes"]) as b:
        lanes = np.array([b.model_lanes(i) for i in range(len(models))])
        classes = {int(g): {"models": int((lanes == g).sum()), "simulations": int(np.asarray(reps)[lanes == g].sum())}
                   for g in np.unique(lanes)}
        print(f"# kernel classes (lanes per simulation: models, simulations): {classes}", file=sys.stderr, flush=True)
        for k in range(args.warmup):
            b.run()
        for k in range(args.steps):
            t1 = time.perf_counter()
            b.run()
            ms += (time.perf_counter() - t1) * 1e3
            st = b.stats()
            fir += st["firings"]
            print(f"# step {k}: firings={st['firings']} kernel_ms={st['kernel_ms']:.1f} launches={st['launches']}",
                  file=sys.stderr, flush=True)
        # the last step's per-model counters over its wall time (the classes run side by side)
        roof = batch_rooflines(models, b, reps, (ms / max(args.steps, 1)) / 1e3)
    sims = sum(reps)
    extra = {"roofline": roof}
    if not args.no_cpu:
        extra["cpu_baseline"] = c2_cpu_baseline(nets, ts)
    _line("reaction firings/sec on 1 GPU (batched uneven models)", fir / (ms / 1e3), "firings/s", args.steps,
          ms / args.steps, "f32",
          {"workload": f"{len(nets)} mass-action models, N 10..500, M 10..1000, {sims} simulations "
                       f"(64..4096 replicas each), 100 samples to t_end={t_end:g}, one batch of concurrent "
                       f"per-class persistent kernels", "models": len(nets), "simulations": sims, "t_end": t_end,
           "classes": classes,
           "autotune": None if tuned is None else
           {"seconds": tuned["seconds"], "classes": {str(g): {"probed_model": c["model"], "chosen": c["chosen"],
                                                              "scores": {str(k): v for k, v in c["scores"].items()}}
                                                     for g, c in tuned["classes"].items()}}},
          simulations_per_s=sims * args.steps / (ms / 1e3),
          counters={k: st[k] for k in ("leaps", "ssa_steps", "switches", "poisson", "gaps", "launches",
                                       "slots", "lanes_per_sim")},
          kernel_ms=st["kernel_ms"], **extra)
    for m in models:
        m.close()


def c2_cpu_baseline(nets, ts, every=64, per_model=16):
    """The oracle's REFSEM restatement over a sample of the batch: every
    64th model (128 models), its first min(replicas, 16) simulations,
    OpenMP over the simulations of a model on the host cores (about 7 s;
    every 128th model took 3.4 s)."""
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    import oracle
    threads, host = host_cores()
    fir, nsim, secs = 0, 0, 0.0
    for n in nets[::every]:
        k = min(int(n["replicas"]), per_model)
        t0 = time.perf_counter()
        try:
            _, st = oracle.run(n["left"], n["right"], n["x0"], n["c"], ts, cs=n["cs"], n=k, seed=0xC0FFEE,
                               precision=oracle.REFSEM, threads=threads)
        finally:
            _restore_affinity()
        secs += time.perf_counter() - t0
        fir += int(st["firings"].sum())
        nsim += k
    return {"value": fir / secs, "unit": "firings/s", "cores": threads, "kind": "port",
            "sample": f"every {every}th model of the batch ({len(nets[::every])} models), the first "
                      f"min(replicas, {per_model}) simulations of each ({nsim} simulations, {fir} firings, "
                      f"{secs:.1f}s)", "simulations_per_s": nsim / secs, "host": host}


def bench_c1ref(args):
    """A denominator shaped like the reference kernel: the first 65536
    simulations of the C1 workload through the reference's execution design,
    restated in this build -- one lane per simulation with its state in HBM,
    and one launch per step (a leap attempt, or a 100-step SSA burst for the
    lanes that switched) followed by a blocking 4-byte copy of the done
    counter (the reference runs three launches and that copy per step,
    kernel.cu:242-345).  It computes in FP32 where the reference computes
    propensities and tau in FP64, so it is faster than the reference kernel
    would be on this GPU.  A step is one complete run of the 65536."""
    import cutauleaping_amd as ctl
    from cutauleaping_amd import networks
    net = networks.mass_action(100, 200, seed=args.seed, t_end=50.0)
    ts = networks.sample_times(200, 50.0)
    model = ctl.Model.from_arrays(net["left"], net["right"], net["x0"], net["c"], ts)
    n = 65536
    fir, ms = 0, 0.0
    steps = max(args.steps, 1)
    with ctl.Sim(model, 256, n // 256, seed=0xC0FFEE, lanes_per_sim=1, iters_per_launch=1, ssa_chunk=100,
                 compact_interval=-1, block_size=-1, verbose=True) as s:      # a line per launch on stderr
        for k in range(steps):
            t1 = time.perf_counter()
            s.run()
            ms += (time.perf_counter() - t1) * 1e3
            st = s.stats()
            fir += st["firings"]
            print(f"# step {k}: firings={st['firings']} launches={st['launches']} kernel_ms={st['kernel_ms']:.1f}",
                  file=sys.stderr, flush=True)
    _line("reaction firings/sec on 1 GPU (reference-shaped execution)", fir / (ms / 1e3), "firings/s", steps,
          ms / steps, "f32",
          {"workload": "mass-action N=100 M=200 seed=1, the first 65536 simulations of the 2^20 workload, t_end=50; "
                       "one lane per simulation, state in HBM, one launch per leap attempt / 100-step SSA burst "
                       "with a blocking copy of the done counter after each", "simulations": n},
          simulations_per_s=n * steps / (ms / 1e3),
          counters={k: st[k] for k in ("leaps", "ssa_steps", "switches", "launches", "slots", "block_size",
                                       "lds_state", "lanes_per_sim")},
          kernel_ms=st["kernel_ms"], roofline=model_rooflines(model, st, st["kernel_ms"], n))
    model.close()


def bench_c4(args):
    """Enzyme cascade: networks.enzyme_cascade(20, seed=4) (61 species, 60
    reactions), 2^18 simulations, 100 samples to t_end=100, adaptive and
    force-SSA (two lines), each with the autotuner's launch shape (untimed;
    --no-tune or --lanes for the library's size rule); at least one untimed
    run before the timed ones."""
    import cutauleaping_amd as ctl
    from cutauleaping_amd import networks
    net = networks.enzyme_cascade(20, seed=4)
    t_end = args.t_end if args.t_end != 50.0 else 100.0
    ts = networks.sample_times(100, t_end)
    model = ctl.Model.from_arrays(net["left"], net["right"], net["x0"], net["c"], ts)
    n = min(args.sims, 1 << 18)
    for force in (False, True):
        fir, ms = 0, 0.0
        opts, tuned = {"lanes_per_sim": args.lanes}, None
        if args.lanes == 0 and not args.no_tune:     # untimed: the autotuner's launch shape for this mode
            from cutauleaping_amd import autotune
            t0 = time.time()
            r = autotune.tune(net["left"], net["right"], net["x0"], net["c"], ts, probe_sims=min(n, 65536),
                              force_ssa=force)
            opts = dict(r["options"])
            tuned = {"options": opts, "seconds": round(time.time() - t0, 2), "confirm": r["confirm"]}
            print(f"# autotuned force_ssa={int(force)} in {tuned['seconds']} s: {opts}", file=sys.stderr, flush=True)
        with ctl.Sim(model, 256, n // 256, seed=0xC0FFEE, force_ssa=force, verbose=True, **opts) as s:
            for k in range(max(args.warmup, 1)):      # untimed: the first run also loads the code object
                s.run()
            for k in range(args.steps):
                t1 = time.perf_counter()
                s.run()
                ms += (time.perf_counter() - t1) * 1e3
                st = s.stats()
                fir += st["firings"]
        extra = {"autotune": tuned} if tuned else {}
        if not args.no_cpu:     # about 10 s of host work per mode (both cost the host about the same per firing)
            extra["cpu_baseline"] = cpu_baseline(net, ts, args.cpu_sims * 64, force_ssa=force)
        _line("reaction firings/sec on 1 GPU (enzyme cascade)", fir / (ms / 1e3), "firings/s", args.steps,
              ms / args.steps, "f32",
              {"workload": f"enzyme cascade 20 stages (61 species, 60 reactions), {n} simulations, 100 samples "
                           f"to t_end={t_end:g}, force_ssa={int(force)}", "simulations": n, "force_ssa": force},
              simulations_per_s=n * args.steps / (ms / 1e3),
              counters={k: st[k] for k in ("leaps", "ssa_steps", "switches", "poisson", "gaps", "launches",
                                           "slots", "lanes_per_sim")},
              kernel_ms=st["kernel_ms"], roofline=model_rooflines(model, st, st["kernel_ms"], n), **extra)
    model.close()


def bench_c3(args):
    """Surrogate-screened parameter sweep (surrogate.sweep_study): training
    data from the simulator, a 5-member BF16-MFMA ensemble, then the held-out
    sweep timed plainly and screened, both judged by the simulator."""
    from cutauleaping_amd import surrogate as sg
    prob = sg.SweepProblem.synthetic(N=50, M=100, seed=5, t_end=5.0)
    say = lambda m: print(m, file=sys.stderr, flush=True)
    r = sg.sweep_study(prob, n_train=args.c3_train, n_test=args.c3_test, replicas=args.c3_replicas,
                       judge_replicas=args.c3_judge, members=5, epochs=args.c3_epochs, batch=256, lr=1e-3,
                       seed=1, log=say)
    _line("sweep time speed-up, surrogate-screened vs plain simulator", r["speedup"], "x", 1,
          r["screened_sweep_s"] * 1e3, "bf16",
          {"workload": f"mass-action N=50 M=100 seed=5, 32 uncertain rates (log-uniform, 2 decades), 10 observed "
                       f"species x 50 times to t_end=5; {args.c3_train} training sets x {args.c3_replicas} "
                       f"replicas; threshold chosen on {r['n_val']} validation sets; {args.c3_test} held-out test sets "
                       f"judged with {r['judge_replicas']} replicas; "
                       f"MLP 32-256-256-256-1000, 5 members", "n_train": args.c3_train, "n_test": args.c3_test},
          study=r)


def bench_mlp(args):
    """The surrogate's network alone (32-256-256-256-1000, BF16 operands, FP32
    accumulation on MFMA): one member's forward pass over --mlp-rows parameter
    sets on device buffers, timed with HIP events on the library's stream;
    then training steps of one batch of --mlp-rows rows (gather, forward,
    loss, backward, Adam), timed on the host around whole epochs.
    Algorithmic FLOPs: 2 * rows * sum(w_l * w_(l+1)) forward, twice that for
    the backward GEMMs (input errors and weight gradients; the first layer's
    input error is not computed, so slightly less)."""
    import ctypes as C
    import cutauleaping_amd as ctl
    widths = [32, 256, 256, 256, 1000]
    rows = args.mlp_rows
    macs = sum(a * b for a, b in zip(widths[:-1], widths[1:]))
    f_fwd = 2.0 * rows * macs
    f_bwd = 2.0 * rows * (2 * macs - widths[0] * widths[1])
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    rng = np.random.default_rng(1)
    x = rng.uniform(-1, 1, (rows, widths[0])).astype(np.float32)
    t = rng.standard_normal((rows, widths[-1])).astype(np.float32)
    net = ctl.MLP(widths, seed=1)
    xd, yd = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(xd), x.nbytes) or hip.hipMalloc(C.byref(yd), rows * widths[-1] * 4):
        raise SystemExit("bench.py: hipMalloc failed")
    try:
        if hip.hipMemcpy(xd, x.ctypes.data, x.nbytes, 1):
            raise SystemExit("bench.py: hipMemcpy failed")
        steps = max(args.steps, 5)
        for _ in range(max(args.warmup, 2)):
            net.forward_device(xd.value, rows, yd.value)
        ms = [net.forward_device(xd.value, rows, yd.value) for _ in range(steps)]
    finally:
        hip.hipFree(xd)
        hip.hipFree(yd)
    fwd_ms = float(np.mean(ms))
    fwd_tf = f_fwd / (fwd_ms * 1e-3) / 1e12
    net.set_data(x, t)
    net.train(2, rows, 1e-4, seed=1)                     # warm-up epochs (one step each)
    t0 = time.perf_counter()
    net.train(steps, rows, 1e-4, seed=2)
    step_ms = (time.perf_counter() - t0) * 1e3 / steps
    net.close()
    train_tf = (f_fwd + f_bwd) / (step_ms * 1e-3) / 1e12
    peak = 2500.0
    _line("surrogate MLP forward throughput (BF16 MFMA, FP32 accumulate)", fwd_tf, "TFLOP/s", steps, fwd_ms,
          "bf16", {"workload": f"MLP 32-256-256-256-1000, one member, {rows} rows per pass", "rows": rows},
          roofline={"bound": "mfma", "achieved": fwd_tf, "peak": peak, "unit": "TFLOP/s", "frac": fwd_tf / peak,
                    "traffic": None},
          flops_per_pass=f_fwd,
          train={"step_ms": step_ms, "tflops": train_tf, "frac": train_tf / peak,
                 "flops_per_step": f_fwd + f_bwd,
                 "note": "host-timed whole epochs of one step: includes the epoch's index upload and sync"})


if __name__ == "__main__":
    main()
