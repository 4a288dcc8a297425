"""Pins the oracle's per-simulation feeds (read_feeds, 2phase_n-tau-leaping.cu:
459-567; applied after every leap, T.cu:2072-2079, and every SSA step,
T.cu:2253-2258): a fed species is reset to its simulation's value after each
event, so A -> B with A fed at a is a pure birth process of B at rate c*a,
whose count at time t is Poisson(c*a*t) exactly."""
import numpy as np
import pytest
from scipy import stats

import stats_util as su


@pytest.mark.parametrize("precision,group", [(0, 0), (1, 0), (0, 4)])
def test_fed_reactant_makes_a_poisson_birth_process(oracle_lib, precision, group):
    c, a = 0.05, 100
    ts = np.array([1.0, 2.0, 4.0])
    n = 4000
    for force in (True, False):
        out, st, xf = oracle_lib.run([[1, 0]], [[0, 1]], [a, 0], [c], ts, n=n, seed=31 + group, force_ssa=force,
                                     precision=precision, group=group, final=True,
                                     feed_species=[0], feed_values=[[a]])
        assert np.all(xf[:, 0] == a)
        B = out[:, :, 1].astype(float)
        for i, t in enumerate(ts):
            lam = c * a * t
            if force:
                # SSA samples hold the state just before the sample time: exact Poisson
                D = su.ks_cdf_D(B[:, i], stats.poisson.cdf(np.arange(200), lam))
                assert D < 1.63 / np.sqrt(n) * 1.3, (t, D)
            else:
                assert su.mean_ci_ok(B[:, i], lam), (t, B[:, i].mean(), lam)


def test_feed_rows_per_simulation(oracle_lib):
    """Each simulation uses its own feed row (the reference's per-thread rows)."""
    from cutauleaping_amd import networks
    net = networks.mass_action(8, 16, seed=41, t_end=2.0)
    n = 64
    rng = np.random.default_rng(4)
    fs = np.array([1, 5], np.int32)
    fv = rng.integers(10, 500, (n, 2)).astype(np.uint32)
    ts = np.linspace(0.1, 2.0, 20)
    for group in (0, 8):
        out, st, xf = oracle_lib.run(net["left"], net["right"], net["x0"], net["c"], ts, n=n, seed=9, final=True,
                                     group=group, feed_species=fs, feed_values=fv)
        absorbed = st["absorbed"] == 1
        assert np.array_equal(xf[~absorbed][:, fs], fv[~absorbed].astype(np.int32))


def test_feed_arguments_are_checked(oracle_lib):
    with pytest.raises(ValueError):
        oracle_lib.run([[1, 0]], [[0, 1]], [10, 0], [1.0], [1.0], n=4, feed_species=[2], feed_values=[[5]])
    with pytest.raises(ValueError):
        oracle_lib.run([[1, 0]], [[0, 1]], [10, 0], [1.0], [1.0], n=4, feed_species=[0],
                       feed_values=[[5], [6], [7]])
