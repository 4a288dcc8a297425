"""The learned surrogate end to end on the GPU at test scale: data from the
simulator (replica-group rows, on-device moments), ensemble training with
the BF16-MFMA kernels, the screened sweep and its fallback to the
simulator, judged against the simulator."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def study(gpu):
    from cutauleaping_amd import surrogate as sg
    print("case: surrogate study (test scale)", flush=True)
    prob = sg.SweepProblem.synthetic(N=50, M=100, seed=5, t_end=1.0)
    res = sg.sweep_study(prob, n_train=1024, n_test=64, replicas=64, members=3, epochs=60, batch=128,
                         lr=2e-3, quantiles=(0.9, 1.0), seed=3, n_val=128)
    return prob, res


def test_training_learns_the_sweep(study):
    prob, r = study
    assert np.isfinite(r["final_loss"]) and r["final_loss"] < 0.5      # standardised targets start near 1
    # the network alone beats the trivial predictor (the training mean) on
    # held-out sets (the validation sets for the network, the test sets for
    # the trivial predictor: both are fresh draws from the same range)
    assert r["surrogate_only_val_err_mean"] < 0.7 * r["trivial_err_mean"]
    assert r["surrogate_only_val_err_sd"] < r["trivial_err_sd"]


def test_screened_sweep_threshold_from_validation(study):
    """The threshold is a quantile of the validation sets' spread, chosen by
    the validation errors; the test sets are answered at that threshold only."""
    prob, r = study
    p90, p100 = r["validation"]
    assert 0.0 < p90["fallback_fraction"] <= 0.5 and p100["fallback_fraction"] == 0.0
    assert r["chosen"]["quantile"] in (0.9, 1.0)
    assert r["test"]["threshold"] == r["chosen"]["threshold"]
    assert 0.0 <= r["test"]["fallback_fraction"] <= 1.0
    assert r["test"]["seconds"] > 0 and r["plain_sweep_s"] > 0
    # the surrogate's own answer (the network, without fallback) is far faster than simulating
    assert r["plain_sweep_s"] > 5 * r["test"]["surrogate_s"]


def test_fallback_thresholds(gpu):
    """threshold 0: every set is simulated, so the screened answer equals the
    plain simulator run with the same streams; threshold inf: none is."""
    from cutauleaping_amd import surrogate as sg
    prob = sg.SweepProblem.synthetic(N=50, M=100, seed=5, t_end=1.0)
    rng = np.random.default_rng(0)
    u = prob.sample_u(256, rng)
    m, s, _, _ = prob.simulate(u, 64, seed=11)
    y = prob.features(m, s)
    sur = sg.Surrogate((prob.n_in, 64, 64, prob.n_out), members=2, seed=2)
    sur.fit(u, y, epochs=5, batch=64)
    ut = prob.sample_u(16, rng)
    ref_m, ref_s, _, _ = prob.simulate(ut, 64, offset=999, seed=21)
    all_sim = sg.screened_sweep(prob, sur, ut, replicas=64, threshold=0.0, offset=999, seed=21)
    assert all_sim["fallback"].all()
    assert np.array_equal(all_sim["mean"], ref_m) and np.array_equal(all_sim["sd"], ref_s)
    none = sg.screened_sweep(prob, sur, ut, replicas=64, threshold=np.inf)
    assert not none["fallback"].any() and none["simulator_seconds"] == 0.0
    sur.close()
