"""GPU: the allocation-free moment entry point (ctl_sim_moments_noalloc)
against ctl_sim_moments through the C ABI.  Both run the same two reduction
passes, so their results must agree bit for bit: on two calls in a row (the
second on a smaller input, so the shared device scratch still holds the
first call's partial sums) and after the allocating entry point has run.
A request larger than the scratch is rejected with CTL_E_LIMIT."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _model(gpu, K, seed=8):
    from cutauleaping_amd import networks
    net = networks.mass_action(10, 20, seed=seed, t_end=2.0)
    ts = np.linspace(2.0 / K, 2.0, K)
    return gpu.Model.from_arrays(net["left"], net["right"], net["x0"], net["c"], ts)


def test_noalloc_moments_equal_the_allocating_entry_point(gpu):
    print("case: noalloc moments, 4096 sims group 1024, then 300 sims group 100", flush=True)
    m = _model(gpu, 20)
    # first call: 4 groups of 1024 replicas (4 partial chunks each)
    with gpu.Sim(m, 256, 16, seed=3) as s:
        s.run()
        mean_n, sd_n = s.moments(1024, noalloc=True)
        mean_a, sd_a = s.moments(1024)
        X = s.samples().astype(np.float64).reshape(4, 1024, 20, 10)
    assert np.array_equal(mean_n, mean_a) and np.array_equal(sd_n, sd_a)
    assert np.allclose(mean_n, X.mean(axis=1), rtol=1e-12, atol=1e-9)
    assert np.allclose(sd_n, X.std(axis=1), rtol=1e-9, atol=1e-6)
    # second call, smaller input: 3 groups of 100 replicas (one chunk each)
    with gpu.Sim(m, 100, 3, seed=4) as s:
        s.run()
        mean_n2, sd_n2 = s.moments(100, noalloc=True)
        mean_a2, sd_a2 = s.moments(100)
        X2 = s.samples().astype(np.float64).reshape(3, 100, 20, 10)
    assert np.array_equal(mean_n2, mean_a2) and np.array_equal(sd_n2, sd_a2)
    assert np.allclose(mean_n2, X2.mean(axis=1), rtol=1e-12, atol=1e-9)
    assert np.allclose(sd_n2, X2.std(axis=1), rtol=1e-9, atol=1e-6)
    m.close()


def test_noalloc_moments_reject_sizes_beyond_the_scratch(gpu):
    print("case: noalloc moments, 2048 sims x 1000 samples, group 1 (beyond the scratch)", flush=True)
    m = _model(gpu, 1000, seed=9)
    with gpu.Sim(m, 256, 8, seed=5) as s:
        s.run()
        with pytest.raises(gpu.CtlError) as e:
            s.moments(1, noalloc=True)
        assert e.value.code == gpu.CTL_E_LIMIT
        mean, sd = s.moments(1)                  # the allocating entry point still serves it
        assert np.array_equal(mean[:, :, :], s.samples().astype(np.float64)) and not sd.any()
        mean_n, sd_n = s.moments(2048, noalloc=True)   # one group fits
        mean_a, sd_a = s.moments(2048)
        assert np.array_equal(mean_n, mean_a) and np.array_equal(sd_n, sd_a)
    m.close()
