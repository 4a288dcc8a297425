"""The batched uneven workload at its configuration's sizes: models drawn
from the same generator as the benchmark (networks.batched_models, seed 2:
N log-uniform in 10..500 species, M in 10..1000 reactions) spanning the
range, the largest with N >= 400 and M >= 800, each model run with the
library's own lanes per simulation (so every kernel class of the batch runs
at once), a few replicas each.  Every model's samples and per-simulation
counters must equal the oracle's algorithm for that model's lanes per
simulation with that model's stream base."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pick(count=8192, seed=2):
    from cutauleaping_amd import networks
    N, M, _ = networks.batched_sizes(count, seed).T
    order = np.argsort(M, kind="stable")
    idx = [int(i) for i in order[np.linspace(0, len(order) - 1, 14).astype(int)]]
    big = [int(i) for i in np.nonzero((N >= 400) & (M >= 800))[0][:2]]
    return idx + big


@pytest.mark.parametrize("force", [False, True])
def test_c2_models_across_the_size_range_bit_exact(gpu, oracle_lib, force):
    from cutauleaping_amd import networks
    picks = _pick()
    print(f"case: C2 models {picks} force_ssa={force}", flush=True)
    t_end = 0.5
    ts = networks.sample_times(100, t_end)
    nets = [networks.batched_model(i, seed=2, t_end=t_end) for i in picks]
    assert max(n["left"].shape[1] for n in nets) >= 400 and max(n["left"].shape[0] for n in nets) >= 800
    models = [gpu.Model.from_arrays(n["left"], n["right"], n["x0"], n["c"], ts, cs=n["cs"]) for n in nets]
    reps = [3 + (k % 4) for k in range(len(nets))]
    with gpu.Batch(models, reps, seed=2468, offset=11, force_ssa=force, iters_per_launch=2000) as b:
        b.run()
        lanes = [b.model_lanes(i) for i in range(len(nets))]
        print(f"lanes per model: {lanes}", flush=True)
        assert len(set(lanes)) >= 3              # several kernel classes ran side by side
        for i, net in enumerate(nets):
            G = lanes[i]
            out, st = oracle_lib.run(net["left"], net["right"], net["x0"], net["c"], ts, cs=net["cs"], n=reps[i],
                                     seed=2468, sim_base=11 + b.stream_base(i), group=0 if G == 1 else G,
                                     force_ssa=force, threads=8)
            assert np.array_equal(b.samples(i), out), (i, picks[i], G)
            gs = b.sim_stats(i)
            for f in ("firings", "leaps", "rejections", "switches", "ssa_steps", "poisson", "gaps", "events"):
                assert np.array_equal(gs[f].astype(np.int64), st[f].astype(np.int64)), (i, picks[i], G, f)
        assert b.stats()["simulations"] == sum(reps)
    for m in models:
        m.close()


@pytest.mark.parametrize("G,iters", [(4, 37), (16, 23), (32, 5000)])
def test_class_kernel_parks_and_resumes_bit_exact(gpu, oracle_lib, G, iters):
    """One kernel class of G lanes per simulation (64/G groups of a wavefront
    on one model at a time) over uneven models and replica counts, with a
    small iteration budget that parks replicas (and held SSA bursts) between
    launches: samples and counters equal the oracle's group algorithm."""
    from cutauleaping_amd import networks
    print(f"case: class kernel G={G} iters={iters}", flush=True)
    rng = np.random.default_rng(91 + G)
    models, reps, nets = [], [], []
    for i in range(9):
        N = int(rng.integers(3, 61)); M = int(rng.integers(4, 151))
        net = networks.mass_action(N, M, seed=300 + i, t_end=0.5, check=False)
        net["x0"] = np.minimum(net["x0"], 3000).astype(np.uint32)
        ts = np.linspace(0.01, 0.5, int(rng.integers(5, 30)))
        cs = np.arange(min(N, 6), dtype=np.int32)
        models.append(gpu.Model.from_arrays(net["left"], net["right"], net["x0"], net["c"], ts, cs=cs))
        nets.append((net, ts, cs))
        reps.append(int(rng.integers(1, 70)))
    with gpu.Batch(models, reps, seed=777, offset=2, iters_per_launch=iters, lanes_per_sim=G) as b:
        b.run()
        assert b.stats()["lanes_per_sim"] == G and b.stats()["launches"] >= (2 if iters < 100 else 1)
        for i, (net, ts, cs) in enumerate(nets):
            assert b.model_lanes(i) == G
            out, st = oracle_lib.run(net["left"], net["right"], net["x0"], net["c"], ts, cs=cs, n=reps[i],
                                     seed=777, sim_base=2 + b.stream_base(i), group=G)
            assert np.array_equal(b.samples(i), out), i
            gs = b.sim_stats(i)
            for f in ("firings", "leaps", "rejections", "switches", "ssa_steps", "poisson", "gaps", "events"):
                assert np.array_equal(gs[f].astype(np.int64), st[f].astype(np.int64)), (i, f)
    for m in models:
        m.close()
