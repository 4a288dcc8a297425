"""Host-side logic that needs no GPU: the autotuner's probe construction and
the benchmark's CPU-baseline bookkeeping."""
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)

from cutauleaping_amd import autotune  # noqa: E402


def _starts(W, rows=3, S=2):
    """Window w's rows are 100*w + k (k = row within the window); window 0 is one row."""
    out = [np.zeros((1, S), np.uint32)]
    for w in range(1, W):
        out.append((100 * w + np.arange(rows, dtype=np.uint32))[:, None].repeat(S, axis=1))
    return out


def test_mixed_round_robin():
    X = autotune._mixed(_starts(4), 16)[:, 0]
    assert list(X[:8]) == [0, 100, 200, 300, 0, 101, 201, 301]
    assert list(X[8:]) == [0, 102, 202, 302, 0, 100, 200, 300]      # rows of a window wrap around


def test_mixed_costliest_quarter_first():
    """With an order, every window keeps its n / W rows, the windows of the
    first quarter of the order fill the first quarter of the rows, and inside
    a quarter the windows alternate row by row."""
    W, n = 8, 64
    order = [5, 2, 7, 1, 0, 3, 6, 4]
    X = autotune._mixed(_starts(W), n, order)[:, 0]
    win = X // 100
    assert all((win == w).sum() == n // W for w in range(W))
    q = n // 4
    for k in range(4):
        assert set(win[k * q:(k + 1) * q]) == set(order[2 * k:2 * k + 2])
        assert list(win[k * q:k * q + 4]) == [order[2 * k], order[2 * k + 1]] * 2
    # the same rows as the round-robin layout, only reordered
    plain = autotune._mixed(_starts(W), n)[:, 0]
    assert sorted(X) == sorted(plain)


def test_divisor_at_most():
    assert [autotune._divisor_at_most(64, x) for x in (64, 63, 40, 16, 3.2, 1, 0.3)] == [64, 32, 32, 16, 2, 1, 1]
    assert autotune._divisor_at_most(48, 20) == 16


def test_cpu_baseline_keeps_the_process_affinity():
    """The OpenMP runtime binds the calling thread to one core when
    OMP_PROC_BIND is set; a second baseline in the same process must still
    see (and use) every core the process started with."""
    code = r"""
import os, sys
sys.path.insert(0, %r)
os.environ["OMP_PROC_BIND"] = "close"; os.environ["OMP_PLACES"] = "cores"
import bench
from cutauleaping_amd import networks
net = networks.enzyme_cascade(4, seed=4)
ts = networks.sample_times(10, 5.0)
before = len(os.sched_getaffinity(0))
a = bench.cpu_baseline(net, ts, 64)
b = bench.cpu_baseline(net, ts, 64, force_ssa=True)
print(before, len(os.sched_getaffinity(0)), a["cores"], b["cores"])
""" % ROOT
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    before, after, ca, cb = map(int, out.stdout.split()[-4:])
    assert after == before and ca == cb == min(before, 16)
