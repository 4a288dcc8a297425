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


def _norm_p(z):
    """Two-sided normal p-value of z."""
    return float(2.0 * stats.norm.sf(abs(z)))


def mean_diff_p(a, b):
    """Two-sample (Welch) z-test of equal means: the difference of the sample
    means over its standard error (large samples)."""
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    se = np.sqrt(a.var(ddof=1) / len(a) + b.var(ddof=1) / len(b))
    d = a.mean() - b.mean()
    if se == 0.0:
        return 1.0 if d == 0.0 else 0.0
    return _norm_p(d / se)


def var_diff_p(a, b):
    """Two-sample z-test of equal variances: the difference of the sample
    variances over its standard error, each from the sample's fourth central
    moment, se^2 = (m4 - s^4) / n (large samples, any distribution)."""
    def mv(x):
        x = np.asarray(x, float)
        c = x - x.mean()
        s2 = float(np.mean(c * c)) * len(x) / (len(x) - 1)
        m4 = float(np.mean(c ** 4))
        return s2, max(m4 - s2 * s2, 0.0) / len(x)
    va, sa = mv(a)
    vb, sb = mv(b)
    se = np.sqrt(sa + sb)
    if se == 0.0:
        return 1.0 if va == vb else 0.0
    return _norm_p((va - vb) / se)


def compare_samples(A, B, skip_constant=True):
    """Per (sample time, species) column of two sample arrays [sims, K, SS]:
    two-sample KS, mean and variance p-values.  Columns where both samples
    hold one and the same value carry no information and are skipped.
    Returns dict(ks=[...], mean=[...], var=[...], cols=[(i, s), ...])."""
    A = np.asarray(A)
    B = np.asarray(B)
    out = {"ks": [], "mean": [], "var": [], "cols": []}
    for i in range(A.shape[1]):
        for s in range(A.shape[2]):
            a = A[:, i, s].astype(float)
            b = B[:, i, s].astype(float)
            if skip_constant and a.min() == a.max() == b.min() == b.max():
                continue
            out["ks"].append(ks_2samp_p(a, b))
            out["mean"].append(mean_diff_p(a, b))
            out["var"].append(var_diff_p(a, b))
            out["cols"].append((i, s))
    return out
