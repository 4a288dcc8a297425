"""GPU parity of the group kernel (lanes_per_sim = G > 1: G lanes cooperate
on one simulation, its state in LDS) through the C ABI against the oracle's
group algorithm with the same G: integer samples, final states and every
per-simulation counter bit-exact, plus statistics and properties at size."""
import os

import numpy as np
import pytest

import stats_util as su
from conftest import GOLDEN
from test_gpu_parity import SCH, SCH_TS, _compare, _random_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("G,force", [(2, False), (4, False), (4, True)])
def test_group_schlogl_bit_exact(gpu, oracle_lib, G, force):
    print(f"case: group schlogl G={G} force_ssa={force}", flush=True)
    m = gpu.Model.from_folder(os.path.join(GOLDEN, "schlogl"), 0, 1024)
    with gpu.Sim(m, 64, 16, seed=77, force_ssa=force, lanes_per_sim=G) as s:
        s.run()
        assert s.stats()["lanes_per_sim"] == G
        out, st, xf = oracle_lib.run(SCH["left"], SCH["right"], SCH["x0"], SCH["c"], SCH_TS, locked=SCH["feed"],
                                     n=1024, seed=77, force_ssa=force, final=True, group=G)
        _compare(s, out, st, xf)


@pytest.mark.parametrize("seed,N,M,n,per_sim,offset,G,geom", [
    (11, 6, 12, 300, True, 0, 4, {}),
    (12, 12, 30, 65, True, 17, 8, {"iters_per_launch": 5}),
    (13, 20, 40, 1, False, 1000, 16, {}),
    (14, 9, 9, 257, False, 3, 2, {"slots": 24, "iters_per_launch": 9}),
    (15, 30, 70, 129, True, 5, 32, {"iters_per_launch": 17}),
    (16, 40, 90, 77, False, 0, 64, {}),
    (17, 25, 60, 200, True, 11, 16, {"iters_per_launch": 7}),
])
def test_group_random_networks_bit_exact(gpu, oracle_lib, seed, N, M, n, per_sim, offset, G, geom):
    """Per-simulation rates and initial states, feed locks, sampled-species
    subsets, ragged sizes, every group width, input offsets, and small
    iteration budgets / slot counts that park simulations across launches
    (with SSA bursts held for a partner group flushed before parking)."""
    print(f"case: group random seed={seed} N={N} M={M} n={n} G={G} {geom}", flush=True)
    net, feed, x0, c, cs, ts = _random_case(seed, N, M, n, per_sim)
    m = gpu.Model.from_arrays(net["left"], net["right"], x0, c, ts, feed=feed, cs=cs)
    with gpu.Sim(m, n, 1, seed=seed * 1000 + 1, offset=offset, lanes_per_sim=G, **geom) as s:
        s.run()
        out, st, xf = oracle_lib.run(net["left"], net["right"], x0, c, ts, locked=feed, cs=cs, n=n,
                                     seed=seed * 1000 + 1, sim_base=offset, final=True, group=G)
        _compare(s, out, st, xf)


def test_group_high_order_and_absorbing_bit_exact(gpu, oracle_lib):
    print("case: group high order + absorbing", flush=True)
    left = [[3, 0, 0], [0, 2, 0], [1, 0, 0]]
    right = [[0, 0, 0], [0, 0, 1], [0, 0, 0]]
    x0 = np.array([[60, 40, 0], [5, 3, 0], [500, 200, 0], [2, 1, 7]], np.uint32).repeat(64, axis=0)
    c = [[0.01, 0.05, 0.2]]
    ts = np.linspace(0.05, 12.0, 60)
    for force in (False, True):
        m = gpu.Model.from_arrays(left, right, x0, c, ts)
        with gpu.Sim(m, 256, 1, seed=5, force_ssa=force, lanes_per_sim=2) as s:
            s.run()
            out, st, xf = oracle_lib.run(left, right, x0, c, ts, n=256, seed=5, force_ssa=force, final=True,
                                         group=2)
            assert st["absorbed"].sum() > 0
            _compare(s, out, st, xf)


@pytest.mark.parametrize("G", [16, 64])
def test_group_mass_action_100x200_subset_bit_exact(gpu, oracle_lib, G):
    """The benchmark network on 24 simulations at both ends of 2^20."""
    from cutauleaping_amd import networks
    print(f"case: group 100x200 network G={G}", flush=True)
    net = networks.mass_action(100, 200, seed=1)
    ts = networks.sample_times(200, 50.0)
    m = gpu.Model.from_arrays(net["left"], net["right"], net["x0"], net["c"], ts)
    for offset in (0, (1 << 20) - 24):
        with gpu.Sim(m, 24, 1, seed=0xC0FFEE, offset=offset, lanes_per_sim=G) as s:
            s.run()
            out, st, xf = oracle_lib.run(net["left"], net["right"], net["x0"], net["c"], ts, n=24, seed=0xC0FFEE,
                                         sim_base=offset, final=True, group=G)
            _compare(s, out, st, xf)


@pytest.mark.parametrize("force", [False, True])
def test_group_enzyme_cascade_bit_exact(gpu, oracle_lib, force):
    from cutauleaping_amd import networks
    print(f"case: group enzyme cascade force_ssa={force}", flush=True)
    net = networks.enzyme_cascade(20, seed=4)
    ts = networks.sample_times(100, 100.0)
    m = gpu.Model.from_arrays(net["left"], net["right"], net["x0"], net["c"], ts)
    with gpu.Sim(m, 128, 1, seed=31, force_ssa=force, lanes_per_sim=4) as s:
        s.run()
        out, st, xf = oracle_lib.run(net["left"], net["right"], net["x0"], net["c"], ts, n=128, seed=31,
                                     force_ssa=force, final=True, group=4)
        _compare(s, out, st, xf)
    n = 20
    assert np.all(xf[:, :n] + xf[:, 2 * n + 1:] == net["x0"][:n] + net["x0"][2 * n + 1:])


def test_group_replica_group_rows_bit_exact(gpu, oracle_lib):
    from cutauleaping_amd import networks
    print("case: group replica-group rows", flush=True)
    net = networks.mass_action(10, 20, seed=8, t_end=2.0)
    rng = np.random.default_rng(8)
    c = net["c"][None, :] * 10 ** rng.uniform(-1, 1, (16, 20))
    x0 = (net["x0"][None, :] * rng.uniform(0.5, 1.5, (16, 10))).astype(np.uint32)
    ts = np.linspace(0.1, 2.0, 20)
    m = gpu.Model.from_arrays(net["left"], net["right"], x0, c, ts)
    with gpu.Sim(m, 32, 16, seed=12, lanes_per_sim=8) as s:
        s.run()
        out, st = oracle_lib.run(net["left"], net["right"], x0, c, ts, n=512, seed=12, group=8)
        assert np.array_equal(s.samples(), out)
        mean, sd = s.moments(32)
    ref = out.reshape(16, 32, 20, 10).astype(float)
    assert np.allclose(mean, ref.mean(axis=1)) and np.allclose(sd, ref.std(axis=1), atol=1e-6)


def test_group_schlogl_statistics(gpu, oracle_lib):
    """16384 group-kernel simulations vs the exact master equation and vs the
    reference-type restatement."""
    print("case: group schlogl statistics 16384", flush=True)
    cme = np.load(os.path.join(GOLDEN, "schlogl_cme.npz"))
    m = gpu.Model.from_arrays(SCH["left"], SCH["right"], SCH["x0"], SCH["c"], SCH_TS, feed=SCH["feed"])
    with gpu.Sim(m, 256, 64, seed=2025, lanes_per_sim=4) as s:
        s.run()
        X = s.samples()[:, :, 1]
    n = X.shape[0]
    for k, tc in enumerate(cme["t_cdf"]):
        i = int(round(tc / 0.1)) - 1
        assert su.ks_cdf_D(X[:, i], cme["cdf"][k]) < 1.63 / np.sqrt(n) * 1.2, tc
    for i in (9, 49, 99):
        assert su.mean_ci_ok(X[:, i], cme["mean"][i]) and su.var_ci_ok(X[:, i], cme["sd"][i] ** 2)
    ref, _ = oracle_lib.run(SCH["left"], SCH["right"], SCH["x0"], SCH["c"], SCH_TS, locked=SCH["feed"],
                            n=4096, seed=99, precision=oracle_lib.REFSEM)
    ps = [su.ks_2samp_p(X[:, i], ref[:, i, 1]) for i in range(100)]
    assert su.holm_ok(ps), min(ps)


def test_group_mass_action_properties_65536(gpu, oracle_lib):
    """2^16 simulations of the benchmark network with the library's own
    choice of lanes per simulation: every simulation reaches t_end, counts
    stay in int32, and 8 simulations across the batch equal the oracle."""
    from cutauleaping_amd import networks
    print("case: group 100x200 network, 65536 sims", flush=True)
    net = networks.mass_action(100, 200, seed=1)
    ts = networks.sample_times(200, 50.0)
    m = gpu.Model.from_arrays(net["left"], net["right"], net["x0"], net["c"], ts)
    with gpu.Sim(m, 256, 256, seed=0xC0FFEE) as s:
        s.run()
        G = s.stats()["lanes_per_sim"]
        g = s.samples()
        gs = s.sim_stats()
        x, t = s.final_state()
    assert G > 1
    assert np.all(t >= 50.0) and np.all(x >= 0) and g.max() < 2 ** 31
    assert np.all(gs["firings"] > 0)
    for idx in np.linspace(0, 65535, 8).astype(int):
        out, st = oracle_lib.run(net["left"], net["right"], net["x0"], net["c"], ts, n=1, seed=0xC0FFEE,
                                 sim_base=int(idx), group=G)
        assert np.array_equal(g[idx], out[0]), idx
        assert gs["firings"][idx] == st["firings"][0]


@pytest.mark.parametrize("force", [False, True])
def test_group_batch_uneven_models_bit_exact(gpu, oracle_lib, force):
    """Uneven models (S 3..60, R 4..150) and replica counts in one launch of
    the group batch kernel (one 64-lane wavefront per simulation); each
    model's samples and counters equal the oracle's group algorithm (G = 64)
    with that model's stream base; a small iteration budget and few group
    slots park simulations across launches."""
    from cutauleaping_amd import networks
    print(f"case: group batch force_ssa={force}", flush=True)
    rng = np.random.default_rng(78)
    models, reps, nets = [], [], []
    for i in range(10):
        N = int(rng.integers(3, 61)); M = int(rng.integers(4, 151))
        net = networks.mass_action(N, M, seed=200 + i, t_end=0.5, check=False)
        net["x0"] = np.minimum(net["x0"], 3000).astype(np.uint32)
        ts = np.linspace(0.01, 0.5, int(rng.integers(5, 30)))
        cs = np.arange(min(N, 6), dtype=np.int32)
        models.append(gpu.Model.from_arrays(net["left"], net["right"], net["x0"], net["c"], ts, cs=cs))
        nets.append((net, ts, cs))
        reps.append(int(rng.integers(1, 40)))
    with gpu.Batch(models, reps, seed=4343, offset=7, iters_per_launch=40, slots=48, lanes_per_sim=64,
                   force_ssa=force) as b:
        b.run()
        assert b.stats()["lanes_per_sim"] == 64
        for i, (net, ts, cs) in enumerate(nets):
            base = b.stream_base(i)
            out, st = oracle_lib.run(net["left"], net["right"], net["x0"], net["c"], ts, cs=cs, n=reps[i],
                                     seed=4343, sim_base=7 + base, group=64, force_ssa=force)
            assert np.array_equal(b.samples(i), out), i
            gs = b.sim_stats(i)
            for f in ("firings", "leaps", "rejections", "switches", "ssa_steps", "poisson", "gaps", "events"):
                assert np.array_equal(gs[f].astype(np.int64), st[f].astype(np.int64)), (i, f)
        assert b.stats()["simulations"] == sum(reps)


def test_autotune_picks_a_measured_candidate(gpu, tmp_path):
    """The autotuner times every lanes-per-simulation candidate on a probe
    of windows tiling the horizon (interleaved in one launch, fewer and
    longer windows until the launch is long enough), refines the winner's
    regrouping interval, SSA chunk and iteration budget, confirms the
    best configuration against its rivals (the library's defaults, the other
    side of regrouping, other G within 25 %) on longer windows, and returns the fastest
    configuration measured; the result is memoised on disk."""
    from cutauleaping_amd import autotune, networks
    print("case: autotune", flush=True)
    net = networks.mass_action(100, 200, seed=1)
    ts = networks.sample_times(200, 50.0)
    cache = str(tmp_path / "tune.json")
    kw = dict(probe_sims=8192, windows=16, min_probe_ms=30.0, confirm_ms=3000.0, candidates=(1, 8, 16, 64),
              cache_file=cache)
    res = autotune.tune(net["left"], net["right"], net["x0"], net["c"], ts, verbose=True, **kw)
    assert res["lanes_per_sim"] in (1, 8, 16, 64) and res["options"]["lanes_per_sim"] == res["lanes_per_sim"]
    stage1 = [res["scores"][f"lanes_per_sim={g}"] for g in (1, 8, 16, 64)]
    assert min(stage1) > 0
    assert len(res["scores"]) > 4                          # stage 2 ran
    if res["confirm"]:                                     # stage 4: the longer windows decide
        assert len(res["confirm"]) >= 2
        assert res["confirm"][autotune._key(res["options"])] == max(res["confirm"].values())
        assert res["probe"]["windows"] % res["probe"]["confirm_windows"] == 0
    else:
        assert res["scores"][autotune._key(res["options"])] == max(res["scores"].values())
    W = res["probe"]["windows"]
    assert 16 % W == 0 and res["probe"]["calibration"][-1]["windows"] == W
    assert res["probe"]["window_starts"] == [50.0 * w / W for w in range(W)]
    cal = res["probe"]["calibration"]
    assert all(a["kernel_ms"] < 30.0 for a in cal[:-1])   # cut only while too short
    assert cal[-1]["kernel_ms"] >= 30.0 or W == 1
    autotune._memo.clear()
    again = autotune.tune(net["left"], net["right"], net["x0"], net["c"], ts, **kw)
    assert again["options"] == res["options"]


def test_autotune_batch_classes(gpu, oracle_lib):
    """tune_batch probes each kernel class's largest model at its G and the
    neighbouring G (over a fixed fraction of the horizon, or one grown until
    the probe is long enough), and returns lanes per model that the batch
    accepts; the batch then runs each model with those lanes, bit-exact
    against the oracle's algorithm for that G."""
    from cutauleaping_amd import autotune, networks
    print("case: autotune batch classes", flush=True)
    ts = networks.sample_times(20, 0.05)
    nets = []
    for i in (3, 17, 40, 51, 77, 90):
        n = networks.batched_model(i, seed=2, t_end=0.05)
        n["ts"] = ts
        n["replicas"] = min(n["replicas"], 256)
        nets.append(n)
    reps = [n["replicas"] for n in nets]
    res = autotune.tune_batch(nets, reps, probe_fraction=0.5, verbose=True)
    lanes = res["lanes"]
    assert lanes.shape == (len(nets),) and set(int(g) for g in lanes) <= {1, 2, 4, 8, 16, 32, 64}
    for G, cl in res["classes"].items():
        assert cl["chosen"] in cl["scores"] and cl["scores"][cl["chosen"]] == max(cl["scores"].values())
        assert cl["fraction"] == 0.5
    auto = autotune.tune_batch(nets, reps, min_probe_ms=5.0, verbose=True)    # horizon grown from 1/16
    for G, cl in auto["classes"].items():
        assert 1.0 / 16 <= cl["fraction"] <= 1.0 and cl["chosen"] in cl["scores"]
    models = [gpu.Model.from_arrays(n["left"], n["right"], n["x0"], n["c"], ts, cs=n["cs"]) for n in nets]
    with gpu.Batch(models, reps, seed=0xB00, lanes=lanes) as b:
        b.run()
        for i, n in enumerate(nets):
            assert b.model_lanes(i) == lanes[i]
            G = int(lanes[i])
            k = min(reps[i], 16)          # the first k replicas (each has its own stream)
            out, st = oracle_lib.run(n["left"], n["right"], n["x0"], n["c"], ts, cs=n["cs"], n=k, seed=0xB00,
                                     sim_base=b.stream_base(i), group=G if G > 1 else 0, threads=8)
            assert np.array_equal(b.samples(i)[:k], out), (i, G)
    for m in models:
        m.close()


@pytest.mark.parametrize("seed,G,compact,chunk,iters,slots,force,block", [
    (40, 2, 1, 1, 5, 0, False, 0),
    (42, 4, 3, 7, 11, 0, True, 0),
    (43, 8, 8, 6, 0, 0, False, 0),
    (49, 16, 2, 3, 9, 24, False, 0),
    (47, 32, 5, 100, 13, 0, False, 0),
    (40, 16, -1, 0, 0, 0, False, 0),
    (43, 16, 8, 6, 5, 0, False, 128),
    (49, 8, 4, 5, 7, 0, False, 64),
])
def test_group_regrouping_bit_exact(gpu, oracle_lib, seed, G, compact, chunk, iters, slots, force, block):
    """A workgroup's simulations are reassigned to its wavefronts by mode
    (leap / SSA) every `compact` iterations, SSA bursts running `chunk` steps
    per iteration; with small iteration budgets and slot counts that park
    simulations (and half-run bursts) across launches, and workgroups of
    four, two or one wavefront (one: burst pairing instead), every sample
    and counter equals the oracle's, which knows no scheduling."""
    print(f"case: group regrouping seed={seed} G={G} compact={compact} chunk={chunk} iters={iters} "
          f"slots={slots} force={force} block={block}", flush=True)
    net, feed, x0, c, cs, ts = _random_case(seed, 14, 30, 200, True)
    m = gpu.Model.from_arrays(net["left"], net["right"], x0, c, ts, feed=feed, cs=cs)
    with gpu.Sim(m, 200, 1, seed=seed, lanes_per_sim=G, force_ssa=force, compact_interval=compact,
                 ssa_chunk=chunk, iters_per_launch=iters, slots=slots, block_size=block) as s:
        s.run()
        if block:
            assert s.stats()["block_size"] == block
        out, st, xf = oracle_lib.run(net["left"], net["right"], x0, c, ts, locked=feed, cs=cs, n=200,
                                     seed=seed, force_ssa=force, final=True, group=G)
        if not force:       # both modes present, so regrouping has something to separate
            assert st["leaps"].sum() > 0 and st["ssa_steps"].sum() > 0
        _compare(s, out, st, xf)
    m.close()
