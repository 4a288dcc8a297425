"""Statistical helpers of the tests (two-sample KS, CIs, Holm correction)."""
import numpy as np
from scipy import stats


def ks_2samp_p(a, b):
    return stats.ks_2samp(np.asarray(a, float), np.asarray(b, float)).pvalue


def ks_cdf_D(x, cdf):
    """Max |empirical CDF - exact CDF| over the support 0..len(cdf)-1 (discrete)."""
    x = np.sort(np.asarray(x))
    emp = np.searchsorted(x, np.arange(len(cdf)), side="right") / len(x)
    return float(np.max(np.abs(emp - cdf)))


def holm_ok(pvalues, alpha=0.01):
    """True when Holm's step-down procedure rejects no hypothesis, i.e. the
    smallest p-value is at least alpha / m."""
    p = np.asarray(pvalues, float)
    return bool(p.min() >= alpha / len(p))


def mean_ci_ok(x, mu, z=3.29):
    """|mean(x) - mu| within z standard errors (z = 3.29: two-sided 99.9%)."""
    x = np.asarray(x, float)
    se = x.std(ddof=1) / np.sqrt(len(x))
    return abs(x.mean() - mu) <= z * max(se, 1e-12)


def var_ci_ok(x, var, z=3.29):
    """Sample variance within z standard errors of var (uses the 4th moment)."""
    x = np.asarray(x, float)
    n = len(x)
    m4 = np.mean((x - x.mean()) ** 4)
    se = np.sqrt(max(m4 - var ** 2, 1e-12) / n)
    return abs(x.var(ddof=1) - var) <= z * se
