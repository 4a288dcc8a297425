"""Pins the oracle's per-simulation feeds (read_feeds,
2phase_n-tau-leaping.cu:459-567).  After every accepted leap each fed species
is reset to its simulation's value (T.cu:2072-2079).  SSA steps apply no
feed: with USE_FEAR_SSA defined (T.cu:38) the SSA kernels update through the
fear matrix (T.cu:2240-2244, 2420-2424) and their feed blocks sit in the
unused #else branch (T.cu:2253-2258, 2434-2440).  The feed row is the
simulation's own row (the loader's per-simulation layout, T.cu:529; the
reference kernel reads row tid, T.cu:2077, which repeats block 0's rows in
every block -- a deliberate deviation)."""
import numpy as np
import pytest
from scipy import stats

import stats_util as su


@pytest.mark.parametrize("precision,group", [(0, 0), (1, 0), (0, 4)])
def test_feeds_are_not_applied_inside_ssa(oracle_lib, precision, group):
    """Force-SSA: A -> B with A fed at a is a pure death of A, since no SSA
    step resets A: B(t) ~ Binomial(a, 1 - exp(-c t)) exactly, A + B = a."""
    c, a = 0.05, 100
    ts = np.array([1.0, 2.0, 4.0])
    n = 4000
    out, st, xf = oracle_lib.run([[1, 0]], [[0, 1]], [a, 0], [c], ts, n=n, seed=31 + group, force_ssa=True,
                                 precision=precision, group=group, final=True, feed_species=[0], feed_values=[[a]])
    assert np.all(xf[:, 0] + xf[:, 1] == a)
    B = out[:, :, 1].astype(float)
    for i, t in enumerate(ts):
        # SSA samples hold the state just before the sample time
        D = su.ks_cdf_D(B[:, i], stats.binom.cdf(np.arange(a + 1), a, 1.0 - np.exp(-c * t)))
        assert D < 1.63 / np.sqrt(n) * 1.3, (t, D)


@pytest.mark.parametrize("precision,group", [(0, 0), (1, 0), (0, 4)])
def test_fed_reactant_makes_a_poisson_birth_process_under_leaps(oracle_lib, precision, group):
    """Adaptive mode where every step is a leap (large counts): the feed resets
    A to a after each leap, so B gains Poisson(c a tau) per leap and
    E[B(t)] = B0 + c a t; the final A is the feed value."""
    c, a, b0 = 1e-3, 100000, 100000
    ts = np.array([1.0, 2.0, 4.0])
    n = 4000
    out, st, xf = oracle_lib.run([[1, 0]], [[0, 1]], [a, b0], [c], ts, n=n, seed=41 + group, precision=precision,
                                 group=group, final=True, feed_species=[0], feed_values=[[a]])
    assert st["switches"].sum() == 0 and st["leaps"].min() > 0
    assert np.all(xf[:, 0] == a)
    B = out[:, :, 1].astype(float)
    for i, t in enumerate(ts):
        assert su.mean_ci_ok(B[:, i], b0 + c * a * t), (t, B[:, i].mean(), b0 + c * a * t)


def test_feed_rows_per_simulation(oracle_lib):
    """Each simulation uses its own feed row: with leaps only (large counts),
    A -> B ends with A at the simulation's own feed value."""
    n = 64
    rng = np.random.default_rng(4)
    fv = rng.integers(50000, 100000, (n, 1)).astype(np.uint32)
    ts = np.array([1.0, 2.0])
    for group in (0, 8):
        out, st, xf = oracle_lib.run([[1, 0]], [[0, 1]], [100000, 100000], [1e-3], ts, n=n, seed=9, final=True,
                                     group=group, feed_species=[0], feed_values=fv)
        assert st["switches"].sum() == 0
        assert np.array_equal(xf[:, 0], fv[:, 0].astype(np.int32))


def test_feed_arguments_are_checked(oracle_lib):
    with pytest.raises(ValueError):
        oracle_lib.run([[1, 0]], [[0, 1]], [10, 0], [1.0], [1.0], n=4, feed_species=[2], feed_values=[[5]])
    with pytest.raises(ValueError):
        oracle_lib.run([[1, 0]], [[0, 1]], [10, 0], [1.0], [1.0], n=4, feed_species=[0],
                       feed_values=[[5], [6], [7]])
