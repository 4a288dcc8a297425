"""Pins the surrogate MLP oracle (oracle/mlp.py) with checks that do not
restate it: central finite differences fix the backward pass, torch's
bfloat16 conversion fixes the rounding."""
import numpy as np
import pytest
import torch

from oracle import mlp as om


def test_bf16_round_matches_torch():
    rng = np.random.default_rng(3)
    x = np.concatenate([
        rng.standard_normal(100000).astype(np.float32) * np.float32(10.0) ** rng.integers(-30, 30, 100000),
        np.array([0.0, -0.0, 1.0, 1.00390625, 1.01171875, 65504.0, 3.4e38, -3.4e38, np.inf, -np.inf,
                  1e-40, -1e-42], np.float32),
    ]).astype(np.float32)
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    got = om.bf16_round(x)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    # ties go to even: 1 + 2^-8 lies halfway between 1 and 1 + 2^-7
    assert om.bf16_round(np.float32(1.0 + 2.0 ** -8)) == np.float32(1.0)
    assert om.bf16_round(np.float32(1.0 + 3 * 2.0 ** -8)) == np.float32(1.0 + 2.0 ** -6)
    nan = om.bf16_round(np.array([np.nan], np.float32))
    assert np.isnan(nan[0])


def test_layout_counts():
    offs, n = om.layout([32, 256, 256, 256, 1000])
    assert n == 32 * 256 + 256 + 2 * (256 * 256 + 256) + 256 * 1000 + 1000
    assert offs[0] == (0, 32 * 256)
    assert offs[1][0] == 32 * 256 + 256


@pytest.mark.parametrize("widths", [[3, 5, 2], [4, 7, 6, 3], [6, 8, 8, 8, 5]])
def test_gradient_finite_differences(widths):
    rng = np.random.default_rng(sum(widths))
    _, n = om.layout(widths)
    p = rng.standard_normal(n) * 0.7
    x = rng.standard_normal((9, widths[0]))
    t = rng.standard_normal((9, widths[-1]))
    lval, g = om.gradient(p, widths, x, t, bf16=False)
    assert lval == pytest.approx(om.loss(p, widths, x, t, bf16=False), rel=1e-12)
    h = 1e-6
    fd = np.empty(n)
    for i in range(n):
        q = p.copy()
        q[i] += h
        lp = om.loss(q, widths, x, t, bf16=False)
        q[i] -= 2 * h
        lm = om.loss(q, widths, x, t, bf16=False)
        fd[i] = (lp - lm) / (2 * h)
    assert np.max(np.abs(fd - g)) < 1e-6 * max(1.0, np.max(np.abs(g)))


def test_bf16_forward_close_to_exact():
    """BF16 operands change the output by about the BF16 unit (2^-8), not more."""
    widths = [32, 64, 64, 10]
    rng = np.random.default_rng(7)
    _, n = om.layout(widths)
    p = (rng.standard_normal(n) * 0.2).astype(np.float32)
    x = rng.uniform(-1, 1, (64, 32)).astype(np.float32)
    y16 = om.forward(p, widths, x, bf16=True)
    y64 = om.forward(p, widths, x, bf16=False)
    rel = np.linalg.norm(y16 - y64) / np.linalg.norm(y64)
    assert 1e-5 < rel < 2e-2
