"""The surrogate MLP's BF16-MFMA kernels against the numpy oracle
(oracle/mlp.py): forward outputs, loss and gradients of every parameter, on
ragged shapes that exercise every bounds check of the GEMM tiles, plus
training and the device-buffer entry point."""
import numpy as np
import pytest

from oracle import mlp as om

pytestmark = pytest.mark.gpu

# Tolerances: operands are identical BF16 values on both sides; the GPU
# accumulates in FP32 in MFMA order, the oracle in float64, and a hidden
# activation that lands on a BF16 rounding boundary may round the other way.
FWD_REL = 5e-3      # relative L2 of outputs
GRAD_REL = 2e-2     # relative L2 of each parameter block's gradient
GRAD_COS = 0.999    # cosine similarity of each block


def _blocks(widths):
    offs, n = om.layout(widths)
    out = []
    for (wo, bo), a, b in zip(offs, widths[:-1], widths[1:]):
        out.append((wo, wo + a * b))
        out.append((bo, bo + b))
    return out


def _init(ctl, widths, seed):
    net = ctl.MLP(widths, seed=seed)
    p = net.get_params()
    # biases start at zero on the device: give them values so the bias paths are tested
    rng = np.random.default_rng(seed)
    offs, n = om.layout(widths)
    for (wo, bo), b in zip(offs, widths[1:]):
        p[bo:bo + b] = rng.uniform(-0.3, 0.3, b).astype(np.float32)
    net.set_params(p)
    return net, p


@pytest.mark.parametrize("widths,rows", [
    ([32, 256, 256, 256, 1000], 300),
    ([5, 37, 3], 1),
    ([7, 65, 129, 33], 97),
    ([32, 64, 10], 4096),
])
def test_forward_matches_oracle(ctl, gpu, widths, rows):
    net, p = _init(ctl, widths, seed=rows)
    assert np.array_equal(net.get_params(), p)
    x = np.random.default_rng(1).uniform(-1, 1, (rows, widths[0])).astype(np.float32)
    y = net.forward(x)
    want = om.forward(p, widths, x, bf16=True)
    assert y.shape == want.shape
    assert np.all(np.isfinite(y))
    rel = np.linalg.norm(y - want) / np.linalg.norm(want)
    assert rel < FWD_REL, rel
    net.close()


@pytest.mark.parametrize("widths,rows", [
    ([32, 256, 256, 256, 1000], 256),
    ([6, 40, 33, 9], 71),
])
def test_gradient_matches_oracle(ctl, gpu, widths, rows):
    net, p = _init(ctl, widths, seed=11 + rows)
    rng = np.random.default_rng(rows)
    x = rng.uniform(-1, 1, (rows, widths[0])).astype(np.float32)
    t = rng.standard_normal((rows, widths[-1])).astype(np.float32)
    loss, g = net.gradient(x, t)
    lw, gw = om.gradient(p, widths, x, t, bf16=True)
    assert loss == pytest.approx(lw, rel=1e-2)
    for lo, hi in _blocks(widths):
        a, b = g[lo:hi].astype(np.float64), gw[lo:hi]
        nb = np.linalg.norm(b)
        assert nb > 0
        rel = np.linalg.norm(a - b) / nb
        cos = float(a @ b / (np.linalg.norm(a) * nb))
        assert rel < GRAD_REL and cos > GRAD_COS, (lo, hi, rel, cos)
    net.close()


def test_training_fits_smooth_map(ctl, gpu):
    """Adam on the device data drives the loss of a smooth 8 -> 16 map down
    by more than 20x, and the trained net generalises to new inputs."""
    widths = [8, 128, 128, 16]
    rng = np.random.default_rng(5)
    A = rng.standard_normal((8, 16)) / 3
    f = lambda x: np.tanh(x @ A) + 0.5 * np.sin(2 * x[:, :1])
    x = rng.uniform(-1, 1, (8192, 8)).astype(np.float32)
    t = f(x).astype(np.float32)
    net = ctl.MLP(widths, seed=3)
    net.set_data(x, t)
    losses = net.train(40, 128, 2e-3, seed=1)
    assert np.all(np.isfinite(losses))
    assert losses[-1] < losses[0] / 20, losses
    xv = rng.uniform(-1, 1, (1024, 8)).astype(np.float32)
    err = np.mean((net.forward(xv) - f(xv)) ** 2)
    assert err < 3 * losses[-1] + 1e-3
    # the oracle agrees with the trained device parameters
    want = om.forward(net.get_params(), widths, xv, bf16=True)
    assert np.linalg.norm(net.forward(xv) - want) / np.linalg.norm(want) < FWD_REL
    net.close()


def test_forward_device_matches_host(ctl, gpu):
    """Device buffers come from the HIP runtime libctl itself links (a second
    runtime in the process, such as the one bundled with a torch wheel,
    cannot share pointers with it)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    H2D, D2H = 1, 2
    widths = [32, 256, 256, 256, 1000]
    net = ctl.MLP(widths, seed=9)
    x = np.random.default_rng(2).uniform(-1, 1, (70000, 32)).astype(np.float32)   # > one 65536 chunk
    y = np.zeros((x.shape[0], widths[-1]), np.float32)
    xd, yd = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(xd), x.nbytes) == 0
    assert hip.hipMalloc(C.byref(yd), y.nbytes) == 0
    try:
        assert hip.hipMemcpy(xd, x.ctypes.data, x.nbytes, H2D) == 0
        ms = net.forward_device(xd.value, x.shape[0], yd.value)
        assert ms > 0
        assert hip.hipMemcpy(y.ctypes.data, yd, y.nbytes, D2H) == 0
    finally:
        hip.hipFree(xd)
        hip.hipFree(yd)
    assert np.array_equal(y, net.forward(x))
    net.close()


def test_bad_arguments(ctl, gpu):
    net = ctl.MLP([4, 8, 2], seed=1)
    with pytest.raises(ValueError):
        net.set_params(np.zeros(3, np.float32))
    assert ctl.lib().ctl_mlp_params(net._h, None, 0, 1) == ctl.CTL_E_ARGS
    with pytest.raises(ctl.CtlError):
        net.train(1, 4, 1e-3)      # no training data
    net.close()
