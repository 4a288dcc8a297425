"""Pins the oracle's model compile and simulator (CPU).

Compile: the Schlogl flats and HOR / HORtype worked by hand from
2phase_n-tau-leaping.cu:219-335, 402-432, plus the sequential HORtype rule.
Simulator: exact answers that do not restate the algorithm -- the Schlogl
chemical master equation (tests/golden/schlogl_cme.npz), the binomial law of
pure decay and isomerisation, the Poisson law of birth-death, the CME of
dimerisation / trimerisation decay (orders 2 and 3) -- plus invariants
(feed-locked species constant, conservation laws, absorption semantics).
"""
import os

import numpy as np
import pytest
from scipy import stats
from scipy.sparse import csr_matrix
from scipy.sparse.linalg import expm_multiply

import stats_util as su
from conftest import GOLDEN

SCHLOGL_L = [[1, 2, 0], [0, 3, 0], [0, 0, 1], [0, 1, 0]]
SCHLOGL_R = [[0, 3, 0], [1, 2, 0], [0, 1, 0], [0, 0, 1]]
SCHLOGL_X0 = [100000, 250, 200000]
SCHLOGL_C = [3e-7, 1e-4, 1e-3, 3.5]


def test_schlogl_compile(oracle_lib):
    m = oracle_lib.compile_model(SCHLOGL_L, SCHLOGL_R, [1, 0, 1])
    assert (m["nL"], m["nV"], m["nF"]) == (5, 8, 4)
    assert m["hor"].tolist() == [3, 3, 1] and m["hort"].tolist() == [1, 3, 1]
    # left flat, reaction-major (r, s, v)
    got = [(r, s, v) for r in range(4) for s, v in zip(m["Ls"][m["Lptr"][r]:m["Lptr"][r + 1]],
                                                         m["Lv"][m["Lptr"][r]:m["Lptr"][r + 1]])]
    assert got == [(0, 0, 1), (0, 1, 2), (1, 1, 3), (2, 2, 1), (3, 1, 1)]
    # rav: species-major net change, feed-locked species included (as the reference builds it)
    rav = [(r, s, v) for s in range(3) for r, v in zip(m["Vr"][m["Vptr"][s]:m["Vptr"][s + 1]],
                                                         m["Vv"][m["Vptr"][s]:m["Vptr"][s + 1]])]
    assert rav == [(0, 0, -1), (1, 0, 1), (0, 1, 1), (1, 1, -1), (2, 1, 1), (3, 1, -1), (2, 2, -1), (3, 2, 1)]
    # fear: locked columns zeroed
    assert m["Fs"].tolist() == [1, 1, 1, 1] and m["Fv"].tolist() == [1, -1, 1, -1]


def test_hortype_sequential_rule(oracle_lib):
    # species 0: r0 = 2X (order 2, coef 2) then r1 = X+Y+Z (order 3, coef 1):
    # the higher order overwrites HORtype with 1 (the reference's sequential rule).
    left = [[2, 0, 0], [1, 1, 1]]
    right = [[0, 0, 0], [0, 0, 0]]
    m = oracle_lib.compile_model(left, right)
    assert m["hor"].tolist() == [3, 3, 3] and m["hort"].tolist() == [1, 1, 1]
    # reversed order: order-3 first, then 2X raises the type to 2 (max rule)
    m = oracle_lib.compile_model(left[::-1], right)
    assert m["hor"].tolist() == [3, 3, 3] and m["hort"].tolist() == [2, 1, 1]


@pytest.fixture(scope="module")
def cme():
    return np.load(os.path.join(GOLDEN, "schlogl_cme.npz"))


@pytest.mark.parametrize("precision,force", [(0, False), (1, False), (0, True)])
def test_schlogl_against_master_equation(oracle_lib, cme, precision, force):
    n = 4096
    out, st = oracle_lib.run(SCHLOGL_L, SCHLOGL_R, SCHLOGL_X0, SCHLOGL_C, cme["t"], locked=[1, 0, 1],
                             n=n, seed=123 + precision + 10 * force, precision=precision, force_ssa=force)
    assert np.all(out[:, :, 0] == 100000) and np.all(out[:, :, 2] == 200000)
    X = out[:, :, 1]
    crit = 1.63 / np.sqrt(n)              # 99% KS critical value
    for k, tc in enumerate(cme["t_cdf"]):
        i = int(round(tc / 0.1)) - 1
        assert su.ks_cdf_D(X[:, i], cme["cdf"][k]) < crit * 1.2, tc
    for i in (9, 49, 99):
        assert su.mean_ci_ok(X[:, i], cme["mean"][i])
        assert su.var_ci_ok(X[:, i], cme["sd"][i] ** 2)
    if force:
        assert st["leaps"].sum() == 0
    else:
        # both paths are exercised (low-mode trajectories may never leap)
        assert st["leaps"].sum() > 0 and st["switches"].sum() > 0


def _cme_decay(order, x0, c, ts):
    """Exact distribution of v X -> 0 (pure decay of order v) at times ts."""
    x = np.arange(x0 + 1, dtype=float)
    from scipy.special import comb
    a = c * comb(x, order)
    rows, cols, vals = [], [], []
    for i in range(x0 + 1):
        if i >= order and a[i] > 0:
            rows.append(i - order); cols.append(i); vals.append(a[i])
        rows.append(i); cols.append(i); vals.append(-a[i])
    Q = csr_matrix((vals, (rows, cols)), shape=(x0 + 1, x0 + 1))
    p0 = np.zeros(x0 + 1); p0[x0] = 1
    return np.array([expm_multiply(Q * t, p0) for t in ts])


@pytest.mark.parametrize("order,x0,c", [(1, 400, 0.3), (2, 300, 0.002), (3, 120, 2e-4)])
def test_decay_orders_against_exact(oracle_lib, order, x0, c):
    """The exact-SSA mode (force_ssa) reproduces the CME of v X -> 0 for
    v = 1, 2, 3, which pins h(x, v) = C(x, v) of the propensity."""
    ts = np.array([0.5, 1.0, 2.0, 4.0])
    n = 4000
    left, right = [[order]], [[0]]
    P = _cme_decay(order, x0, c, ts)
    for precision in (0, 1):
        out, st = oracle_lib.run(left, right, [x0], [c], ts, n=n, seed=77 + order, force_ssa=True,
                                 precision=precision)
        for i in range(len(ts)):
            D = su.ks_cdf_D(out[:, i, 0], np.cumsum(P[i]))
            assert D < 1.63 / np.sqrt(n) * 1.3, (order, precision, ts[i], D)


@pytest.mark.parametrize("order,x0,c", [(1, 400, 0.37), (2, 300, 0.0023)])
def test_decay_adaptive_replay_matches_reference_types(oracle_lib, order, x0, c):
    """Adaptive tau-leaping: the FP32 replay and the reference-type (double)
    restatement agree in distribution at every sample time.  (Neither matches
    the exact law closely at coarse leaps: a decreasing species records its
    leap-start value, the reference's saturating conversion, T.cu:1538-1541.)
    The rate is chosen so that the deterministic first-order leaps
    (tau = eps / c) do not end on a sample time, where FP32 and FP64 rounding
    would decide which leap records the sample."""
    ts = np.linspace(0.3, 4.1, 16)
    a, _ = oracle_lib.run([[order]], [[0]], [x0], [c], ts, n=4000, seed=5, precision=0)
    b, _ = oracle_lib.run([[order]], [[0]], [x0], [c], ts, n=4000, seed=6, precision=1)
    ps = [su.ks_2samp_p(a[:, i, 0], b[:, i, 0]) for i in range(len(ts))]
    assert su.holm_ok(ps), min(ps)


def test_birth_death_stationary_poisson(oracle_lib):
    """0 -> X (k), X -> 0 (g): stationary Poisson(k/g).  Exact SSA matches the
    law; adaptive tau-leaping matches the mean and the variance to within the
    known explicit-leap inflation 1/(1 - g tau/2) (a few percent at eps=0.03).
    The rates are non-integer: at an exact balance k = g X (integer rates)
    the reference drops the species from tau' (it needs mu != 0 AND sigma != 0,
    T.cu:1903-1909), tau' = FLT_MAX and the trajectory freezes."""
    k, g = 203.7, 1.13
    ts = np.array([20.0, 25.0])
    pois = stats.poisson.cdf(np.arange(700), k / g)
    out, st = oracle_lib.run([[0], [1]], [[1], [0]], [5], [k, g], ts, n=8000, seed=3, force_ssa=True)
    x = out[:, -1, 0].astype(float)
    assert su.ks_cdf_D(x, pois) < 1.63 / np.sqrt(8000) * 1.3
    out, st = oracle_lib.run([[0], [1]], [[1], [0]], [5], [k, g], ts, n=8000, seed=3)
    x = out[:, -1, 0].astype(float)
    # tolerance: 1% on the mean (the leap-start value recorded for decreasing
    # leaps, T.cu:1538-1541, shifts it up by ~0.6% here), 15% on the variance
    assert abs(x.mean() / (k / g) - 1) < 0.01
    assert 0.95 < x.var() / (k / g) < 1.15
    assert st["leaps"].mean() > 10       # the leap path, not only SSA


def test_exact_balance_freezes_like_the_reference(oracle_lib):
    """Integer rates with k = g X: once X hits the balance point tau' is
    FLT_MAX (reference quirk, see above), so the late samples collapse."""
    out, _ = oracle_lib.run([[0], [1]], [[1], [0]], [5], [200.0, 1.0], [20.0, 25.0], n=2000, seed=3)
    assert out[:, -1, 0].astype(float).var() < 20.0


def test_isomerisation_binomial_and_conservation(oracle_lib):
    """A <-> B, c1 = 1, c2 = 2: stationary A ~ Binomial(N, 2/3).  Exact SSA
    matches the law; both modes conserve A + B in the final state (the
    interpolated samples of a leap need not be conserved); adaptive
    tau-leaping holds the mean to 1% and the variance to 30% (leap-start
    recording of decreasing species, T.cu:1538-1541, inflates it)."""
    N = 3000
    for force in (True, False):
        out, st, xf = oracle_lib.run([[1, 0], [0, 1]], [[0, 1], [1, 0]], [N, 0], [1.0, 2.0], [5.0, 6.0],
                                     n=6000, seed=9, final=True, force_ssa=force)
        assert np.all(xf.sum(axis=1) == N)
        a = out[:, -1, 0].astype(float)
        if force:
            cdf = stats.binom.cdf(np.arange(N + 1), N, 2 / 3)
            assert su.ks_cdf_D(a, cdf) < 1.63 / np.sqrt(6000) * 1.3
        else:
            assert abs(a.mean() / (N * 2 / 3) - 1) < 0.01
            assert abs(a.var() / (N * 2 / 9) - 1) < 0.3


def test_absorption_repeats_last_recorded_sample(oracle_lib):
    # X -> 0 at a high rate: absorbs long before t_max; every later sample
    # repeats the last recorded one (T.cu:1665-1697)
    ts = np.linspace(0.1, 10, 100)
    out, st = oracle_lib.run([[1]], [[0]], [30], [5.0], ts, n=200, seed=4)
    assert np.all(st["absorbed"] == 1)
    assert np.all(st["t_end"] == 10.0 + float(np.float32(0.1)))
    assert np.all(out[:, -1, 0] == 0) or np.all(out[:, -1, 0] <= out[:, 0, 0])
    for row in out[:, :, 0]:
        assert np.all(np.diff(row.astype(int)) <= 0)


def test_feed_locked_and_nonnegative_random_network(oracle_lib):
    from cutauleaping_amd import networks
    rng = np.random.default_rng(5)
    net = networks.mass_action(8, 16, seed=21, t_end=5.0)
    feed = (rng.random(8) < 0.25).astype(np.int32)
    out, st = oracle_lib.run(net["left"], net["right"], net["x0"], net["c"], np.linspace(0.05, 5, 100),
                             locked=feed, n=256, seed=8)
    for s in np.nonzero(feed)[0]:
        assert np.all(out[:, :, s] == net["x0"][s])
    assert out.max() < 2 ** 31


def test_small_lambda_process_matches_independent_draws(oracle_lib):
    """The FP32 replay samples reactions with tau*a_j < 1 through one shared
    Poisson process; the reference-type restatement draws every reaction
    independently.  Many slow reactions, two-sample KS on every species and
    sample time (Holm-corrected)."""
    rng = np.random.default_rng(17)
    S, M = 6, 40
    left = np.zeros((M, S), np.int32)
    right = np.zeros((M, S), np.int32)
    for j in range(M):
        if j % 4:
            left[j, rng.integers(S)] = 1
        right[j, rng.integers(S)] = 1
    c = 10 ** rng.uniform(-3, -0.5, M)
    x0 = rng.integers(50, 400, S)
    ts = np.linspace(0.5, 20, 8)
    a, sa = oracle_lib.run(left, right, x0, c, ts, n=3000, seed=1, precision=0)
    b, sb = oracle_lib.run(left, right, x0, c, ts, n=3000, seed=2, precision=1)
    assert sa["gaps"].sum() > 0 and sa["leaps"].sum() > 0
    ps = [su.ks_2samp_p(a[:, i, s], b[:, i, s]) for i in range(len(ts)) for s in range(S)]
    assert su.holm_ok(ps), min(ps)
