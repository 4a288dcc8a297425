"""Pins the oracle's group algorithm (G lanes per simulation) with exact
laws and invariants that do not restate it: the Schlogl master equation,
multinomial channel choice (which fixes the lane partial sums, their scan
and the walk inside the chosen lane, for both the SSA selection and the
critical-reaction pick of a leap), binomial isomerisation with
conservation, feed-locked species, and agreement in distribution with the
reference-type restatement on random networks."""
import os

import numpy as np
import pytest
from scipy import stats

import stats_util as su
from conftest import GOLDEN

SCHLOGL_L = [[1, 2, 0], [0, 3, 0], [0, 0, 1], [0, 1, 0]]
SCHLOGL_R = [[0, 3, 0], [1, 2, 0], [0, 1, 0], [0, 0, 1]]
SCHLOGL_X0 = [100000, 250, 200000]
SCHLOGL_C = [3e-7, 1e-4, 1e-3, 3.5]


@pytest.mark.parametrize("G,force", [(2, False), (4, False), (8, True)])
def test_group_schlogl_against_master_equation(oracle_lib, G, force):
    cme = np.load(os.path.join(GOLDEN, "schlogl_cme.npz"))
    n = 4096
    out, st = oracle_lib.run(SCHLOGL_L, SCHLOGL_R, SCHLOGL_X0, SCHLOGL_C, cme["t"], locked=[1, 0, 1],
                             n=n, seed=300 + G, force_ssa=force, group=G)
    assert np.all(out[:, :, 0] == 100000) and np.all(out[:, :, 2] == 200000)
    X = out[:, :, 1]
    crit = 1.63 / np.sqrt(n)
    for k, tc in enumerate(cme["t_cdf"]):
        i = int(round(tc / 0.1)) - 1
        assert su.ks_cdf_D(X[:, i], cme["cdf"][k]) < crit * 1.2, tc
    for i in (9, 49, 99):
        assert su.mean_ci_ok(X[:, i], cme["mean"][i])
        assert su.var_ci_ok(X[:, i], cme["sd"][i] ** 2)
    if not force:
        assert st["leaps"].sum() > 0 and st["switches"].sum() > 0


def _channels(M, seed):
    """X -> Y_j for j = 0..M-1 (species 0 is X, species 1 + j is Y_j)."""
    rng = np.random.default_rng(seed)
    left = np.zeros((M, M + 1), np.int32)
    right = np.zeros((M, M + 1), np.int32)
    left[:, 0] = 1
    right[np.arange(M), 1 + np.arange(M)] = 1
    c = 10 ** rng.uniform(-1, 1, M)
    return left, right, c


def _chi2_p(counts, p):
    tot = counts.sum()
    exp = tot * p
    return stats.chi2.sf(((counts - exp) ** 2 / exp).sum(), len(p) - 1)


@pytest.mark.parametrize("G", [4, 8, 16, 64])
def test_group_ssa_selection_is_multinomial(oracle_lib, G):
    """Exact SSA through the group selection: every X molecule ends in
    channel j with probability c_j / sum c.  37 channels over G lanes put
    several reactions in every lane (G <= 16), so a wrong lane prefix, scan
    order or in-lane walk biases the channel counts; G = 64 leaves 27 lanes
    without reactions, whose zero partials the selection must skip."""
    M = 37
    left, right, c = _channels(M, 11)
    x0 = np.zeros(M + 1, np.uint32)
    x0[0] = 3000
    out, st, xf = oracle_lib.run(left, right, x0, c, [30.0], n=400, seed=17, force_ssa=True, group=G,
                                 final=True)
    assert np.all(xf[:, 0] == 0)
    counts = xf[:, 1:].sum(axis=0).astype(float)
    assert counts.sum() == 400 * 3000
    assert _chi2_p(counts, c / c.sum()) > 1e-3


@pytest.mark.parametrize("G", [4, 8, 64])
def test_group_critical_pick_is_multinomial(oracle_lib, G):
    """X0 = 9 < nc: every channel is critical, tau' is infinite and each leap
    fires exactly one reaction chosen by the critical pick, an exact SSA.
    Channel 0's weight enters the pick walk first (the reference's walk
    starts from a_0, T.cu:1993-2006); here it is critical as well."""
    M = 29
    left, right, c = _channels(M, 12)
    x0 = np.zeros(M + 1, np.uint32)
    x0[0] = 9
    out, st, xf = oracle_lib.run(left, right, x0, c, [40.0], n=6000, seed=18, group=G, final=True)
    assert np.all(xf[:, 0] == 0)
    assert st["ssa_steps"].sum() == 0 and st["leaps"].sum() == 6000 * 9
    counts = xf[:, 1:].sum(axis=0).astype(float)
    assert _chi2_p(counts, c / c.sum()) > 1e-3


def test_group_isomerisation_binomial_and_conservation(oracle_lib):
    N = 3000
    for force, G in ((True, 2), (False, 4)):
        out, st, xf = oracle_lib.run([[1, 0], [0, 1]], [[0, 1], [1, 0]], [N, 0], [1.0, 2.0], [5.0, 6.0],
                                     n=6000, seed=19, final=True, force_ssa=force, group=G)
        assert np.all(xf.sum(axis=1) == N)
        a = out[:, -1, 0].astype(float)
        if force:
            cdf = stats.binom.cdf(np.arange(N + 1), N, 2 / 3)
            assert su.ks_cdf_D(a, cdf) < 1.63 / np.sqrt(6000) * 1.3
        else:
            assert abs(a.mean() / (N * 2 / 3) - 1) < 0.01
            assert abs(a.var() / (N * 2 / 9) - 1) < 0.3


@pytest.mark.parametrize("G", [8, 16])
def test_group_random_network_matches_reference_types(oracle_lib, G):
    """A random 12-species, 40-reaction network: the group algorithm and the
    reference-type restatement agree in distribution (two-sample KS on every
    species and sample time, Holm-corrected); feed-locked species stay put."""
    from cutauleaping_amd import networks
    net = networks.mass_action(12, 40, seed=23, t_end=4.0)
    feed = np.zeros(12, np.int32)
    feed[[2, 7]] = 1
    ts = np.linspace(0.5, 4.0, 8)
    a, sa = oracle_lib.run(net["left"], net["right"], net["x0"], net["c"], ts, locked=feed, n=3000, seed=1,
                           group=G)
    b, sb = oracle_lib.run(net["left"], net["right"], net["x0"], net["c"], ts, locked=feed, n=3000, seed=2,
                           precision=1)
    assert sa["leaps"].sum() > 0 and sa["poisson"].sum() > 0 and sa["gaps"].sum() > 0
    for s in (2, 7):
        assert np.all(a[:, :, s] == net["x0"][s])
    ps = [su.ks_2samp_p(a[:, i, s], b[:, i, s]) for i in range(len(ts)) for s in range(12)
          if np.ptp(np.concatenate([a[:, i, s], b[:, i, s]])) > 0]
    assert su.holm_ok(ps), min(ps)


def test_group_rejects_bad_arguments(oracle_lib):
    with pytest.raises(ValueError):
        oracle_lib.run([[1]], [[0]], [10], [1.0], [1.0], n=4, group=3)
    with pytest.raises(ValueError):
        oracle_lib.run([[1]], [[0]], [10], [1.0], [1.0], n=4, group=4, precision=1)
