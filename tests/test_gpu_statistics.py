"""Statistical agreement of the GPU kernels, at the benchmark configurations,
with the reference algorithm and with an exact SSA.

The reference kernel is Cao's adaptive tau-leaping with SSA fallback
(2phase_n-tau-leaping.cu:1568-2454).  Its statistical stand-in is the
oracle's REFSEM mode: the reference's own types (double propensities and
tau, float mu / sigma / g) and its per-reaction Poisson draws.  The exact SSA
is the direct method (force_ssa: every event drawn), run by the GPU kernel
itself (its SSA path is pinned bit for bit against the oracle's SSA in
test_gpu_parity.py / test_gpu_group.py) or by the oracle on the host.

Per (sample time, sampled species) column: two-sample Kolmogorov-Smirnov,
difference of means and difference of variances, each family of p-values
Holm-corrected at alpha = 0.01 over every non-constant column.  Counts must
never be negative (a negative int32 count would appear as >= 2^31 in the
uint32 samples).
"""
import os

import numpy as np
import pytest

import stats_util as su

pytestmark = pytest.mark.gpu

ALPHA = 0.01


def _threads():
    """Host threads for the oracle: the CPUs this process may use, at most 16."""
    try:
        return max(1, min(16, len(os.sched_getaffinity(0))))
    except AttributeError:
        return max(1, min(16, os.cpu_count() or 1))


def _holm_report(name, res):
    out = {}
    for fam in ("ks", "mean", "var"):
        p = np.asarray(res[fam])
        k = int(np.argmin(p))
        out[fam] = (float(p.min()), res["cols"][k], su.holm_ok(p, ALPHA))
    print(f"{name}: {len(res['ks'])} columns; " +
          "; ".join(f"{f} min p {v[0]:.2e} at (sample, species) {v[1]}" for f, v in out.items()), flush=True)
    return out


def _standardized_bias(A, B, cols):
    """|mean(A) - mean(B)| / pooled sd per column."""
    d = []
    for i, s in cols:
        a = A[:, i, s].astype(float)
        b = B[:, i, s].astype(float)
        sd = np.sqrt(0.5 * (a.var() + b.var()))
        d.append(abs(a.mean() - b.mean()) / sd if sd > 0 else 0.0)
    return np.asarray(d)


@pytest.fixture(scope="module")
def c1(gpu, oracle_lib):
    """The benchmark workload's network (100 species, 200 reactions, 200
    samples to t_end = 50): 4096 simulations of the production kernel (the
    library's own choice of lanes per simulation), 4096 exact-SSA simulations
    of the GPU's force-SSA path, and 2048 reference-algorithm simulations on
    the host, each from its own seed."""
    from cutauleaping_amd import networks
    net = networks.mass_action(100, 200, seed=1)
    ts = networks.sample_times(200, 50.0)
    m = gpu.Model.from_arrays(net["left"], net["right"], net["x0"], net["c"], ts)
    print("case: C1 statistics, 4096 GPU adaptive + 4096 GPU exact SSA", flush=True)
    with gpu.Sim(m, 256, 16, seed=0x5EED01) as s:
        s.run()
        G = s.stats()["lanes_per_sim"]
        kern = s.samples(layout=1)
        xk, _ = s.final_state()
    with gpu.Sim(m, 256, 16, seed=0x5EED02, force_ssa=True) as s:
        s.run()
        ssa = s.samples(layout=1)
        xs, _ = s.final_state()
    m.close()
    print(f"case: C1 statistics, 2048 reference-algorithm simulations on {_threads()} host threads", flush=True)
    ref, _, xr = oracle_lib.run(net["left"], net["right"], net["x0"], net["c"], ts, n=2048, seed=0x5EED03,
                                precision=oracle_lib.REFSEM, threads=_threads(), final=True)
    return dict(G=G, kern=kern, ssa=ssa, ref=ref, xk=xk, xs=xs, xr=xr)


def test_c1_kernel_matches_reference_algorithm(c1):
    """Production kernel against the reference algorithm: every species at
    every one of the 200 sample times, KS / means / variances Holm-clean."""
    assert c1["G"] > 1
    for a in (c1["kern"], c1["ssa"]):
        assert a.max() < 2 ** 31
    assert c1["xk"].min() >= 0 and c1["xs"].min() >= 0 and c1["xr"].min() >= 0
    res = su.compare_samples(c1["kern"], c1["ref"])
    rep = _holm_report(f"C1 kernel (G={c1['G']}) vs reference algorithm", res)
    assert len(res["ks"]) > 15000
    for fam, (pmin, col, ok) in rep.items():
        assert ok, (fam, pmin, col)


def test_c1_deviation_from_exact_ssa_is_the_reference_algorithms(c1):
    """Tau-leaping with eps = 0.03 is not exact: on this network both the
    kernel and the reference algorithm differ detectably from the exact SSA
    (species driven to a few copies, where the reference's sample
    interpolation records a falling species' leap-start value, and
    fast-relaxing species early on).  What the kernel must not add is a
    deviation of its own: column by column, its distance to the exact SSA
    (KS statistic, standardized mean difference) equals the reference
    algorithm's within the sampling noise of the three samples."""
    from scipy import stats
    kern, ref, ssa = c1["kern"], c1["ref"], c1["ssa"]
    res_k = su.compare_samples(kern, ssa)
    res_r = su.compare_samples(ref, ssa)
    cols = [c for c in res_k["cols"] if c in set(res_r["cols"])]
    nk, nr, ns = len(kern), len(ref), len(ssa)
    dk, dr = [], []
    for i, s in cols:
        dk.append(stats.ks_2samp(kern[:, i, s].astype(float), ssa[:, i, s].astype(float)).statistic)
        dr.append(stats.ks_2samp(ref[:, i, s].astype(float), ssa[:, i, s].astype(float)).statistic)
    dk, dr = np.asarray(dk), np.asarray(dr)
    # |D(k, S) - D(r, S)| <= D(k, r) (triangle inequality of the sup norm);
    # D(k, r) under equal laws stays below the Holm-level KS critical value
    m = len(cols)
    c_alpha = np.sqrt(-0.5 * np.log((ALPHA / m) / 2.0))
    bound = c_alpha * np.sqrt(1.0 / nk + 1.0 / nr)
    gap = np.abs(dk - dr)
    bk = _standardized_bias(kern, ssa, cols)
    br = _standardized_bias(ref, ssa, cols)
    nfail_k = int(np.sum(np.asarray(res_k["ks"]) < ALPHA / len(res_k["ks"])))
    nfail_r = int(np.sum(np.asarray(res_r["ks"]) < ALPHA / len(res_r["ks"])))
    print(f"C1 vs exact SSA over {m} columns: KS rejections (Bonferroni) kernel {nfail_k}, reference algorithm "
          f"{nfail_r}; max KS distance kernel {dk.max():.4f}, reference {dr.max():.4f}; max standardized mean "
          f"difference kernel {bk.max():.3f}, reference {br.max():.3f}; max |D_kernel - D_reference| "
          f"{gap.max():.4f} (bound {bound:.4f})", flush=True)
    assert gap.max() <= bound
    assert bk.max() <= br.max() + 4.0 * np.sqrt(1.0 / nk + 1.0 / nr + 2.0 / ns)


@pytest.fixture(scope="module")
def c4(gpu, oracle_lib):
    """The stiff enzyme cascade (20 stages, 61 species, 60 reactions, 100
    samples to t_end = 100): 16384 simulations each of the production kernel
    in adaptive mode and in force-SSA mode; 8192 simulations each of the
    reference algorithm and of the oracle's exact SSA on the host."""
    from cutauleaping_amd import networks
    net = networks.enzyme_cascade(20, seed=4)
    ts = networks.sample_times(100, 100.0)
    m = gpu.Model.from_arrays(net["left"], net["right"], net["x0"], net["c"], ts)
    out = {}
    for name, force, seed in (("adaptive", False, 0x5EED11), ("force", True, 0x5EED12)):
        print(f"case: C4 statistics, 16384 GPU simulations force_ssa={force}", flush=True)
        with gpu.Sim(m, 256, 64, seed=seed, force_ssa=force) as s:
            s.run()
            out[name] = s.samples(layout=1)
            out["G"] = s.stats()["lanes_per_sim"]
            out["x_" + name], _ = s.final_state()
    m.close()
    print("case: C4 statistics, 8192 reference-algorithm and 8192 exact-SSA simulations on the host", flush=True)
    out["ref"], _ = oracle_lib.run(net["left"], net["right"], net["x0"], net["c"], ts, n=8192, seed=0x5EED13,
                                   precision=oracle_lib.REFSEM, threads=_threads())
    out["ssa"], _ = oracle_lib.run(net["left"], net["right"], net["x0"], net["c"], ts, n=8192, seed=0x5EED14,
                                   force_ssa=True, threads=_threads())
    out["net"] = net
    return out


@pytest.mark.parametrize("a,b", [("adaptive", "ref"), ("adaptive", "ssa"), ("force", "ssa"), ("adaptive", "force")])
def test_c4_statistics(c4, a, b):
    """Adaptive kernel against the reference algorithm and against the exact
    SSA, the kernel's force-SSA mode against the exact SSA, and the kernel's
    two modes against each other: KS / means / variances Holm-clean at every
    species and sample time; conservation of every enzyme (E_k + C_k) and of
    the substrate mass in every final state."""
    A, B = c4[a], c4[b]
    assert A.max() < 2 ** 31 and B.max() < 2 ** 31
    res = su.compare_samples(A, B)
    rep = _holm_report(f"C4 {a} (G={c4['G']}) vs {b}", res)
    assert len(res["ks"]) > 500
    for fam, (pmin, col, ok) in rep.items():
        assert ok, (fam, pmin, col)
    net = c4["net"]
    n = 20
    for key in ("x_adaptive", "x_force"):
        x = c4[key].astype(np.int64)
        assert x.min() >= 0
        assert np.all(x[:, :n] + x[:, 2 * n + 1:] == net["x0"][:n].astype(np.int64) + net["x0"][2 * n + 1:])
        total = x[:, n:2 * n + 1].sum(axis=1) + x[:, 2 * n + 1:].sum(axis=1)
        assert np.all(total == int(net["x0"][n:2 * n + 1].sum()))
