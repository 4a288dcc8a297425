"""Per-simulation feeds on the GPU (one-lane, group and batch kernels)
against the oracle with the same feeds: bit-exact samples, counters and
final states.  Feeds reset the fed species after every accepted leap and
not inside SSA bursts (2phase_n-tau-leaping.cu:2072-2079, 2240-2244)."""
import numpy as np
import pytest

from test_gpu_parity import _compare

pytestmark = pytest.mark.gpu


def _fed_case(seed, N, M, n):
    from cutauleaping_amd import networks
    net = networks.mass_action(N, M, seed=seed, t_end=2.0)
    rng = np.random.default_rng(seed)
    fs = np.sort(rng.choice(N, size=3, replace=False)).astype(np.int32)
    fv = rng.integers(5, 400, (n, 3)).astype(np.uint32)
    ts = np.sort(rng.uniform(0.02, 2.0, 25))
    return net, fs, fv, ts


@pytest.mark.parametrize("lanes,force", [(1, False), (1, True), (8, False), (8, True)])
def test_feeds_bit_exact(gpu, oracle_lib, lanes, force):
    print(f"case: feeds lanes={lanes} force_ssa={force}", flush=True)
    n = 130
    net, fs, fv, ts = _fed_case(51, 12, 30, n)
    m = gpu.Model.from_arrays(net["left"], net["right"], net["x0"], net["c"], ts)
    m.set_feeds(fs, fv)
    with gpu.Sim(m, n, 1, seed=77, force_ssa=force, lanes_per_sim=lanes, iters_per_launch=23) as s:
        s.run()
        out, st, xf = oracle_lib.run(net["left"], net["right"], net["x0"], net["c"], ts, n=n, seed=77,
                                     force_ssa=force, final=True, group=0 if lanes == 1 else lanes,
                                     feed_species=fs, feed_values=fv)
        _compare(s, out, st, xf)
    if not force:
        # a simulation that never switched to SSA ends on a leap, so at its feed values
        # (SSA steps apply no feed, 2phase_n-tau-leaping.cu:2240-2244)
        leaps_only = (st["absorbed"] == 0) & (st["switches"] == 0) & (st["leaps"] > 0)
        assert np.array_equal(xf[leaps_only][:, fs], fv[leaps_only].astype(np.int32))


@pytest.mark.parametrize("lanes", [1, 64])
def test_feeds_in_a_batch_bit_exact(gpu, oracle_lib, lanes):
    print(f"case: feeds in a batch lanes={lanes}", flush=True)
    models, nets, reps = [], [], [37, 64, 5]
    for k, (N, M) in enumerate(((6, 10), (15, 40), (9, 25))):
        net, fs, fv, ts = _fed_case(60 + k, N, M, reps[k])
        m = gpu.Model.from_arrays(net["left"], net["right"], net["x0"], net["c"], ts)
        m.set_feeds(fs, fv)
        models.append(m)
        nets.append((net, fs, fv, ts))
    with gpu.Batch(models, reps, seed=99, offset=3, lanes_per_sim=lanes) as b:
        b.run()
        for i, (net, fs, fv, ts) in enumerate(nets):
            out, st = oracle_lib.run(net["left"], net["right"], net["x0"], net["c"], ts, n=reps[i], seed=99,
                                     sim_base=3 + b.stream_base(i), group=0 if lanes == 1 else lanes,
                                     feed_species=fs, feed_values=fv)
            assert np.array_equal(b.samples(i), out), i
