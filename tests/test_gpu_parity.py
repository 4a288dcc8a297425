"""GPU parity of the one-lane-per-simulation kernel (lanes_per_sim=1)
through the C ABI against the CPU oracle.

Bit-exact (integer samples, final states and per-simulation counters) against
the oracle's FP32 replay of the same Philox streams; statistical agreement
with the exact Schlogl master equation, the reference-type restatement and
the exact SSA; properties at larger sizes.
"""
import os
import subprocess

import numpy as np
import pytest

import stats_util as su
from conftest import GOLDEN

pytestmark = pytest.mark.gpu

SCH = dict(left=[[1, 2, 0], [0, 3, 0], [0, 0, 1], [0, 1, 0]], right=[[0, 3, 0], [1, 2, 0], [0, 1, 0], [0, 0, 1]],
           x0=[100000, 250, 200000], c=[3e-7, 1e-4, 1e-3, 3.5], feed=[1, 0, 1])
SCH_TS = [round(0.1 * i, 10) for i in range(1, 101)]
STAT_FIELDS = ("firings", "leaps", "rejections", "switches", "ssa_steps", "poisson", "gaps", "events", "absorbed")


def _compare(gpu_sim, ref_out, ref_st, ref_xf):
    g = gpu_sim.samples()
    assert g.shape == ref_out.shape
    bad = np.argwhere(g != ref_out)
    assert bad.size == 0, f"{len(bad)} samples differ, first at {bad[:3].tolist()}"
    gs = gpu_sim.sim_stats()
    for f in STAT_FIELDS:
        assert np.array_equal(gs[f].astype(np.int64), ref_st[f].astype(np.int64)), f
    assert np.array_equal(gs["t_end"], ref_st["t_end"])
    x, t = gpu_sim.final_state()
    assert np.array_equal(x, ref_xf)


@pytest.mark.parametrize("force", [False, True])
def test_schlogl_reference_config_bit_exact(gpu, oracle_lib, force):
    """Config 0: Schlogl, 1024 simulations (64 threads x 16 blocks) from the input folder."""
    print(f"case: schlogl 1024 force_ssa={force}", flush=True)
    m = gpu.Model.from_folder(os.path.join(GOLDEN, "schlogl"), 0, 1024)
    with gpu.Sim(m, 64, 16, seed=77, force_ssa=force, lanes_per_sim=1) as s:
        s.run()
        out, st, xf = oracle_lib.run(SCH["left"], SCH["right"], SCH["x0"], SCH["c"], SCH_TS, locked=SCH["feed"],
                                     n=1024, seed=77, force_ssa=force, final=True)
        _compare(s, out, st, xf)


def _random_case(seed, N, M, n, per_sim):
    from cutauleaping_amd import networks
    net = networks.mass_action(N, M, seed=seed, t_end=2.0)
    rng = np.random.default_rng(seed)
    feed = (rng.random(N) < 0.2).astype(np.int32)
    x0 = net["x0"][None, :].repeat(n if per_sim else 1, axis=0)
    c = net["c"][None, :].repeat(n if per_sim else 1, axis=0)
    if per_sim:
        x0 = (x0 * rng.uniform(0.5, 1.5, x0.shape)).astype(np.uint32)
        c = c * rng.uniform(0.5, 2.0, c.shape)
    cs = np.sort(rng.choice(N, size=max(1, N // 2), replace=False)).astype(np.int32)
    ts = np.sort(rng.uniform(0.01, 2.0, 37))
    return net, feed, x0, c, cs, ts


@pytest.mark.parametrize("seed,N,M,n,per_sim,offset,geom", [
    (11, 6, 12, 300, True, 0, {}),
    (12, 12, 30, 65, True, 17, {"iters_per_launch": 5, "ssa_chunk": 3}),
    (13, 20, 40, 1, False, 1000, {}),
    (14, 9, 9, 257, False, 3, {"block_size": 64, "slots": 128, "iters_per_launch": 9, "ssa_chunk": 7}),
])
def test_random_networks_bit_exact(gpu, oracle_lib, seed, N, M, n, per_sim, offset, geom):
    """Per-simulation rates and initial states, feed locks, sampled-species
    subsets, ragged sizes, input offsets (Philox streams follow the global
    index) and launch geometries that split SSA bursts across relaunches."""
    print(f"case: random seed={seed} N={N} M={M} n={n} per_sim={per_sim} offset={offset} {geom}", flush=True)
    net, feed, x0, c, cs, ts = _random_case(seed, N, M, n, per_sim)
    m = gpu.Model.from_arrays(net["left"], net["right"], x0, c, ts, feed=feed, cs=cs)
    with gpu.Sim(m, n, 1, seed=seed * 1000 + 1, offset=offset, **geom, lanes_per_sim=1) as s:
        s.run()
        out, st, xf = oracle_lib.run(net["left"], net["right"], x0, c, ts, locked=feed, cs=cs, n=n,
                                     seed=seed * 1000 + 1, sim_base=offset, final=True)
        _compare(s, out, st, xf)


def test_high_order_and_absorbing_bit_exact(gpu, oracle_lib):
    """3X -> 0, 2Y -> Z, X -> 0: order-3 propensities, criticality with v > 1
    and absorption (a0 = 0) in both the leap and the SSA path."""
    print("case: high order + absorbing", flush=True)
    left = [[3, 0, 0], [0, 2, 0], [1, 0, 0]]
    right = [[0, 0, 0], [0, 0, 1], [0, 0, 0]]
    x0 = np.array([[60, 40, 0], [5, 3, 0], [500, 200, 0], [2, 1, 7]], np.uint32).repeat(64, axis=0)
    c = [[0.01, 0.05, 0.2]]
    ts = np.linspace(0.05, 12.0, 60)
    for force in (False, True):
        m = gpu.Model.from_arrays(left, right, x0, c, ts)
        with gpu.Sim(m, 256, 1, seed=5, force_ssa=force, lanes_per_sim=1) as s:
            s.run()
            out, st, xf = oracle_lib.run(left, right, x0, c, ts, n=256, seed=5, force_ssa=force, final=True)
            assert st["absorbed"].sum() > 0
            _compare(s, out, st, xf)


def test_mass_action_100x200_subset_bit_exact(gpu, oracle_lib):
    """The benchmark network (N=100, M=200, 200 samples to t_end=50) on 24
    simulations with their global indices spread over 2^20."""
    from cutauleaping_amd import networks
    print("case: 100x200 network, 24 sims", flush=True)
    net = networks.mass_action(100, 200, seed=1)
    ts = networks.sample_times(200, 50.0)
    m = gpu.Model.from_arrays(net["left"], net["right"], net["x0"], net["c"], ts)
    for offset in (0, (1 << 20) - 24):
        with gpu.Sim(m, 24, 1, seed=0xC0FFEE, offset=offset, lanes_per_sim=1) as s:
            s.run()
            out, st, xf = oracle_lib.run(net["left"], net["right"], net["x0"], net["c"], ts, n=24, seed=0xC0FFEE,
                                         sim_base=offset, final=True)
            _compare(s, out, st, xf)


def test_schlogl_statistics(gpu, oracle_lib):
    """16384 GPU simulations vs the exact master equation (KS at 10 times,
    mean / variance CIs), vs the reference-type restatement (two-sample KS at
    every sample time, Holm) and GPU exact SSA vs the master equation."""
    print("case: schlogl statistics 16384", flush=True)
    cme = np.load(os.path.join(GOLDEN, "schlogl_cme.npz"))
    m = gpu.Model.from_arrays(SCH["left"], SCH["right"], SCH["x0"], SCH["c"], SCH_TS, feed=SCH["feed"])
    for force in (False, True):
        with gpu.Sim(m, 256, 64, seed=2024 + force, force_ssa=force, lanes_per_sim=1) as s:
            s.run()
            X = s.samples()[:, :, 1]
        n = X.shape[0]
        for k, tc in enumerate(cme["t_cdf"]):
            i = int(round(tc / 0.1)) - 1
            assert su.ks_cdf_D(X[:, i], cme["cdf"][k]) < 1.63 / np.sqrt(n) * 1.2, (force, tc)
        for i in (9, 49, 99):
            assert su.mean_ci_ok(X[:, i], cme["mean"][i]) and su.var_ci_ok(X[:, i], cme["sd"][i] ** 2)
        if not force:
            ref, _ = oracle_lib.run(SCH["left"], SCH["right"], SCH["x0"], SCH["c"], SCH_TS, locked=SCH["feed"],
                                    n=4096, seed=99, precision=oracle_lib.REFSEM)
            ps = [su.ks_2samp_p(X[:, i], ref[:, i, 1]) for i in range(100)]
            assert su.holm_ok(ps), min(ps)


def test_mass_action_properties_65536(gpu, oracle_lib):
    """2^16 simulations of the benchmark network: every simulation reaches
    t_end, counts stay in int32, and 8 simulations picked across the batch
    equal the oracle replay of their own streams."""
    from cutauleaping_amd import networks
    print("case: 100x200 network, 65536 sims", flush=True)
    net = networks.mass_action(100, 200, seed=1)
    ts = networks.sample_times(200, 50.0)
    m = gpu.Model.from_arrays(net["left"], net["right"], net["x0"], net["c"], ts)
    with gpu.Sim(m, 256, 256, seed=0xC0FFEE, lanes_per_sim=1) as s:
        s.run()
        g = s.samples()
        gs = s.sim_stats()
        x, t = s.final_state()
    assert np.all(t >= 50.0) and np.all(x >= 0) and g.max() < 2 ** 31
    assert np.all(gs["firings"] > 0)
    for idx in np.linspace(0, 65535, 8).astype(int):
        out, st = oracle_lib.run(net["left"], net["right"], net["x0"], net["c"], ts, n=1, seed=0xC0FFEE,
                                 sim_base=int(idx))
        assert np.array_equal(g[idx], out[0]), idx
        assert gs["firings"][idx] == st["firings"][0]


def test_cli_end_to_end(gpu, oracle_lib, tmp_path):
    """The cuTauLeaping binary: dynamics files <prefix>_<gid> with the
    reference's row format, bit-exact with the oracle under CTL_SEED; the
    stdout variant (verbose=2); fitness on stdout."""
    print("case: CLI", flush=True)
    env = dict(os.environ, CTL_SEED="4242")
    out_dir = tmp_path / "out"
    cmd = [gpu.CLI_PATH, os.path.join(GOLDEN, "schlogl"), "32", "2", "0", "0", str(out_dir), "run", "0", "0"]
    p = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "Writing dynamics to file" in p.stdout
    files = sorted(os.listdir(out_dir))
    assert len(files) == 64 and "run_0" in files and "run_63" in files
    ref, _ = oracle_lib.run(SCH["left"], SCH["right"], SCH["x0"], SCH["c"], SCH_TS, locked=SCH["feed"], n=64, seed=4242)
    for g in (0, 17, 63):
        lines = (out_dir / f"run_{g}").read_text().splitlines()
        assert len(lines) == 100
        first = lines[0].split("\t")
        assert first[0] == "0.1" and len(first) == 4
        vals = np.array([[int(v) for v in l.split("\t")[1:]] for l in lines], np.uint32)
        assert np.array_equal(vals, ref[g])
    assert lines[-1].split("\t")[0] == "10"
    p2 = subprocess.run(cmd[:-1] + ["0", "2"], env=env, capture_output=True, text=True, timeout=300)
    assert p2.returncode == 0
    blocks = [b for b in p2.stdout.split("\n\n") if b.strip() and not b.startswith(" *") and "\t" in b]
    assert len(blocks) == 64
    # fitness on stdout: targets equal to the X mean of a 2-repetition group
    fold = tmp_path / "fit"
    subprocess.run(["cp", "-r", os.path.join(GOLDEN, "schlogl"), str(fold)], check=True)
    (fold / "p_rep").write_text("2\n")
    (fold / "ts_rep").write_text("1\n")
    (fold / "ts_numtgt").write_text("3\n")
    (fold / "tts_vector").write_text("64\n")
    (fold / "ts_matrix").write_text("".join(f"{t}\t100000\t300\t200000\n" for t in SCH_TS))
    p3 = subprocess.run([gpu.CLI_PATH, str(fold), "32", "2", "0", "0", str(tmp_path / "o2"), "fit", "2", "0"],
                        env=env, capture_output=True, text=True, timeout=300)
    assert p3.returncode == 0, p3.stdout + p3.stderr
    body = "".join(l for l in p3.stdout.splitlines(True) if not l.startswith("WARNING"))
    vals = [float(v) for v in body.split("\t") if v.strip()]
    assert len(vals) == 32
    # expected: mean over samples / species of |tgt - mean of the 2 repetitions| / tgt
    exp = []
    for th in range(32):
        m2 = (ref[th].astype(float) + ref[32 + th].astype(float)) / 2
        tgt = np.array([100000, 300, 200000], float)
        exp.append(np.sum(np.abs(tgt - m2) / tgt) / (100 * 3))
    assert np.allclose(vals, exp, rtol=1e-5, atol=1e-7)


@pytest.mark.parametrize("force", [False, True])
def test_enzyme_cascade_bit_exact(gpu, oracle_lib, force):
    """Config 4's network: a 20-stage enzyme cascade (61 species, 60
    reactions, rates over six decades), adaptive and force-SSA, 128
    simulations to t_end=100 with 100 samples."""
    from cutauleaping_amd import networks
    print(f"case: enzyme cascade force_ssa={force}", flush=True)
    net = networks.enzyme_cascade(20, seed=4)
    ts = networks.sample_times(100, 100.0)
    m = gpu.Model.from_arrays(net["left"], net["right"], net["x0"], net["c"], ts)
    with gpu.Sim(m, 128, 1, seed=31, force_ssa=force, lanes_per_sim=1) as s:
        s.run()
        out, st, xf = oracle_lib.run(net["left"], net["right"], net["x0"], net["c"], ts, n=128, seed=31,
                                     force_ssa=force, final=True)
        _compare(s, out, st, xf)
    # conservation per stage: E_k + C_k is constant
    n = 20
    assert np.all(xf[:, :n] + xf[:, 2 * n + 1:] == net["x0"][:n] + net["x0"][2 * n + 1:])


def test_moments_and_binary_output(gpu, tmp_path):
    """On-device per-group mean / sd equal numpy's over the same samples; the
    binary dump round-trips the [sim][sample][species] storage."""
    print("case: moments + binary output", flush=True)
    m = gpu.Model.from_arrays(SCH["left"], SCH["right"], SCH["x0"], SCH["c"], SCH_TS, feed=SCH["feed"])
    with gpu.Sim(m, 64, 8, seed=9, lanes_per_sim=1) as s:
        s.run()
        g = s.samples().astype(np.float64)
        mean, sd = s.moments(32)
        path = str(tmp_path / "samples.bin")
        s.write_samples_binary(path)
    ref = g.reshape(16, 32, 100, 3)
    assert np.allclose(mean, ref.mean(axis=1), rtol=1e-12, atol=1e-9)
    assert np.allclose(sd, ref.std(axis=1), rtol=1e-9, atol=1e-6)
    raw = np.fromfile(path, dtype=np.uint32)
    assert raw[0] == 0x534C5443 and raw[1] == 1 and tuple(raw[2:5]) == (512, 100, 3)
    assert np.array_equal(raw[8:].reshape(512, 100, 3), g.astype(np.uint32))


def test_batched_uneven_models_bit_exact(gpu, oracle_lib):
    """Twelve models of uneven size (S 3..40, R 4..80) and replica count
    (1..300) in one persistent launch; every model's samples and counters
    equal the oracle run with that model's stream base.  A small iteration
    budget forces relaunches mid-simulation."""
    from cutauleaping_amd import networks
    print("case: batched uneven models", flush=True)
    rng = np.random.default_rng(77)
    models, reps, nets = [], [], []
    for i in range(12):
        N = int(rng.integers(3, 41)); M = int(rng.integers(4, 81))
        net = networks.mass_action(N, M, seed=100 + i, t_end=1.0, check=False)
        net["x0"] = np.minimum(net["x0"], 3000).astype(np.uint32)
        ts = np.linspace(0.02, 1.0, int(rng.integers(5, 40)))
        cs = np.arange(min(N, 7), dtype=np.int32)
        models.append(gpu.Model.from_arrays(net["left"], net["right"], net["x0"], net["c"], ts, cs=cs))
        nets.append((net, ts, cs))
        reps.append(int(rng.integers(1, 301)))
    with gpu.Batch(models, reps, seed=4242, offset=5, iters_per_launch=50, ssa_chunk=4, lanes_per_sim=1) as b:
        b.run()
        for i, (net, ts, cs) in enumerate(nets):
            base = b.stream_base(i)
            out, st = oracle_lib.run(net["left"], net["right"], net["x0"], net["c"], ts, cs=cs, n=reps[i],
                                     seed=4242, sim_base=5 + base)
            g = b.samples(i)
            assert np.array_equal(g, out), i
            gs = b.sim_stats(i)
            for f in STAT_FIELDS:
                assert np.array_equal(gs[f].astype(np.int64), st[f].astype(np.int64)), (i, f)
        assert b.stats()["simulations"] == sum(reps)


def test_replica_group_rows_bit_exact(gpu, oracle_lib):
    """Rate and initial-state rows shared by groups of consecutive
    simulations (16 parameter sets x 32 replicas), as a parameter sweep uses
    them; per-group moments come from the device."""
    from cutauleaping_amd import networks
    print("case: replica-group rows", flush=True)
    net = networks.mass_action(10, 20, seed=8, t_end=2.0)
    rng = np.random.default_rng(8)
    c = net["c"][None, :] * 10 ** rng.uniform(-1, 1, (16, 20))
    x0 = (net["x0"][None, :] * rng.uniform(0.5, 1.5, (16, 10))).astype(np.uint32)
    ts = np.linspace(0.1, 2.0, 20)
    m = gpu.Model.from_arrays(net["left"], net["right"], x0, c, ts)
    with gpu.Sim(m, 32, 16, seed=12, lanes_per_sim=1) as s:
        s.run()
        out, st = oracle_lib.run(net["left"], net["right"], x0, c, ts, n=512, seed=12)
        assert np.array_equal(s.samples(), out)
        mean, sd = s.moments(32)
    ref = out.reshape(16, 32, 20, 10).astype(float)
    assert np.allclose(mean, ref.mean(axis=1)) and np.allclose(sd, ref.std(axis=1), atol=1e-6)


def test_cli_fitness_to_pref_allfit(gpu, oracle_lib, tmp_path):
    """fitness = 1 (kernel.cu:406-420): the fitness of each (group, thread)
    goes to ./pref_allfit in the working directory, one line per group of
    p_rep blocks, fixed notation with 20 decimals, tab-terminated values;
    the values equal the closed form over the oracle's samples."""
    print("case: CLI fitness=1 (pref_allfit)", flush=True)
    env = dict(os.environ, CTL_SEED="4242")
    fold = tmp_path / "fit"
    subprocess.run(["cp", "-r", os.path.join(GOLDEN, "schlogl"), str(fold)], check=True)
    (fold / "p_rep").write_text("2\n")
    (fold / "ts_rep").write_text("1\n")
    (fold / "ts_numtgt").write_text("3\n")
    (fold / "tts_vector").write_text("128\n")
    (fold / "ts_matrix").write_text("".join(f"{t}\t100000\t300\t200000\n" for t in SCH_TS))
    work = tmp_path / "cwd"
    work.mkdir()
    p = subprocess.run([gpu.CLI_PATH, str(fold), "32", "4", "0", "0", str(tmp_path / "o"), "fit", "1", "0"],
                       env=env, capture_output=True, text=True, timeout=300, cwd=str(work))
    assert p.returncode == 0, p.stdout + p.stderr
    assert " * Dumping fitness to file." in p.stdout
    lines = (work / "pref_allfit").read_text().split("\n")
    assert lines[-1] == "" and len(lines) == 3            # 4 blocks / p_rep 2 = 2 groups
    ref, _ = oracle_lib.run(SCH["left"], SCH["right"], SCH["x0"], SCH["c"], SCH_TS, locked=SCH["feed"], n=128,
                            seed=4242)
    tgt = np.array([100000, 300, 200000], float)
    for g, line in enumerate(lines[:2]):
        assert line.endswith("\t")
        fields = line.split("\t")[:-1]
        assert len(fields) == 32
        assert all("." in f and len(f.split(".")[1]) == 20 for f in fields)
        for th in range(32):
            m2 = (ref[(2 * g) * 32 + th].astype(float) + ref[(2 * g + 1) * 32 + th].astype(float)) / 2
            exp = np.sum(np.abs(tgt - m2) / tgt) / (100 * 3)
            assert abs(float(fields[th]) - exp) <= 1e-5 * exp + 1e-7, (g, th)


@pytest.mark.parametrize("compact,block_size,force,iters,slots", [
    (1, 0, False, 7, 300),
    (3, 0, True, 11, 0),
    (2, -1, False, 9, 0),
    (8, 0, False, 0, 0),
    (-1, 0, False, 0, 0),
])
def test_one_lane_regrouping_bit_exact(gpu, oracle_lib, compact, block_size, force, iters, slots):
    """The one-lane kernel regroups a workgroup's simulations by mode (leap /
    SSA) every `compact` iterations, a lane taking over another slot's state;
    with state in LDS (Schlogl) and in HBM (block_size -1), small iteration
    budgets and slot counts that park simulations across launches, every
    sample and counter equals the oracle's (which knows no scheduling)."""
    print(f"case: one-lane regrouping compact={compact} block_size={block_size} force={force} iters={iters}",
          flush=True)
    m = gpu.Model.from_arrays(SCH["left"], SCH["right"], SCH["x0"], SCH["c"], SCH_TS, feed=SCH["feed"])
    with gpu.Sim(m, 64, 16, seed=606, force_ssa=force, lanes_per_sim=1, compact_interval=compact,
                 block_size=block_size, iters_per_launch=iters, slots=slots) as s:
        s.run()
        out, st, xf = oracle_lib.run(SCH["left"], SCH["right"], SCH["x0"], SCH["c"], SCH_TS, locked=SCH["feed"],
                                     n=1024, seed=606, force_ssa=force, final=True)
        _compare(s, out, st, xf)
    m.close()
