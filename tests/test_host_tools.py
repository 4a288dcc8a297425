"""Host-side logic that needs no GPU: the autotuner's probe construction and
the benchmark's CPU-baseline bookkeeping."""
import os
import subprocess
import sys

import numpy as np
import pytest

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


class _FakeModel:
    """Model.from_arrays stand-in: keeps the initial-state rows and sample times."""
    made = []

    def __init__(self, x0, t):
        self.x0, self.t = np.atleast_2d(x0), np.asarray(t, np.float64)

    @classmethod
    def from_arrays(cls, left, right, x0, c, t, feed=None, cs=None):
        m = cls(x0, t)
        cls.made.append(m)
        return m

    def close(self):
        pass


def _fake_sim(rate_of, ms_per_unit):
    """Sim stand-in: firings proportional to the horizon, kernel time from
    rate_of(options) (firings/s), per-simulation counters by row."""

    class FakeSim:
        calls = []

        def __init__(self, m, threads, blocks, gpu=0, seed=0, force_ssa=False, **o):
            self.m, self.n, self.o = m, threads * blocks, o
            FakeSim.calls.append((float(m.t[-1]), dict(o)))

        def __enter__(self):
            return self

        def __exit__(self, *a):
            return False

        def run(self):
            return self

        def samples(self):
            K, S = len(self.m.t), self.m.x0.shape[1]
            return np.ones((self.n, K, S), np.uint32) * np.arange(1, K + 1, dtype=np.uint32)[None, :, None]

        def stats(self):
            fir = 1e9 * float(self.m.t[-1]) * ms_per_unit / 100.0
            return {"firings": fir, "kernel_ms": fir / rate_of(self.o) * 1e3,
                    "lanes_per_sim": self.o.get("lanes_per_sim", 16)}

        def sim_stats(self):
            ss = np.zeros(self.n, dtype=[("leaps", "u4"), ("ssa_steps", "u8")])
            ss["leaps"] = np.arange(self.n) % 5
            ss["ssa_steps"] = 6 * (np.arange(self.n) % 3)
            return ss

    return FakeSim


def test_tune_control_flow(monkeypatch, tmp_path):
    """Calibration cuts the windows until the probe is long enough; stage 4
    re-times the best configuration, the library's defaults at its G, the
    other side of regrouping and every G within 25 % over the whole horizon,
    and its fastest wins even against the probe's ranking; the result is
    memoised on disk."""
    G_RATE = {1: 1e8, 2: 2e8, 4: 4e8, 8: 8e8, 16: 1.2e9, 32: 1.15e9, 64: 5e8}

    def probe_rate(o):                     # the probe's view: no regrouping looks best
        r = G_RATE[o.get("lanes_per_sim", 16)]
        return r * (1.05 if o.get("compact_interval", 0) < 0 else 1.0)

    FakeSim = _fake_sim(probe_rate, ms_per_unit=10.0)
    whole = {}

    def whole_rate(o):                     # the whole horizon: the defaults (regrouping) win
        r = G_RATE[o.get("lanes_per_sim", 16)]
        if o.get("compact_interval", 0) < 0:
            return r * 0.9
        return r * (1.0 if "ssa_chunk" not in o else 0.95)

    class Sim(FakeSim):
        def stats(self):
            st = FakeSim.stats(self)
            if float(self.m.t[-1]) >= 50.0 - 1e-9:       # a run over the whole horizon
                st["kernel_ms"] = st["firings"] / whole_rate(self.o) * 1e3
                whole[autotune._key(self.o)] = whole_rate(self.o)
            return st

    monkeypatch.setattr(autotune, "Sim", Sim)
    monkeypatch.setattr(autotune, "Model", _FakeModel)
    autotune._memo.clear()
    S, R = 100, 200
    left = np.zeros((R, S), np.int32)
    ts = np.linspace(0.25, 50.0, 200)
    cache = str(tmp_path / "tune.json")
    res = autotune.tune(left, left, np.ones(S), np.ones(R), ts, probe_sims=4096, windows=64,
                        min_probe_ms=250.0, cache_file=cache)
    cal = res["probe"]["calibration"]
    assert cal[0]["windows"] == 64 and len(cal) > 1 and all(a["kernel_ms"] < 250.0 for a in cal[:-1])
    assert cal[-1]["kernel_ms"] >= 250.0 and 64 % res["probe"]["windows"] == 0
    assert res["probe"]["confirm_windows"] == 1
    conf = res["confirm"]
    assert "lanes_per_sim=16" in conf                           # the defaults at the best G
    assert "compact_interval=-1,lanes_per_sim=16" in conf       # the probe's favourite
    assert "lanes_per_sim=32" in conf                           # another G within 25 %
    assert res["options"] == {"lanes_per_sim": 16}              # the whole horizon decides
    assert all(conf[k] == pytest.approx(whole[k], rel=1e-9) for k in conf) and len(conf) == 4
    autotune._memo.clear()
    monkeypatch.setattr(autotune, "Sim", None)                  # a disk hit runs nothing
    again = autotune.tune(left, left, np.ones(S), np.ones(R), ts, probe_sims=4096, windows=64,
                          min_probe_ms=250.0, cache_file=cache)
    assert again["options"] == res["options"]
    autotune._memo.clear()


def test_tune_batch_grows_the_probe_horizon(monkeypatch):
    """Per kernel class the probe horizon starts at 1/16 and grows until the
    probe at the class's G takes min_probe_ms (or covers the horizon); the
    neighbouring G are timed over that horizon and the fastest is applied to
    the class's models no larger than the probed one."""
    ms_per_unit = {16: 400.0, 4: 20.0}             # kernel ms per time unit of the probed model, by class

    class FakeBatch:
        def __init__(self, models, reps, gpu=0, seed=0, lanes=None):
            self.models, self.lanes = models, lanes

        def __enter__(self):
            return self

        def __exit__(self, *a):
            return False

        def model_lanes(self, i):
            return 16 if self.models[i].x0.shape[1] > 50 else 4

        def run(self):
            return self

        def stats(self):
            m, g = self.models[0], self.lanes[0]
            rule = 16 if m.x0.shape[1] > 50 else 4
            T = float(m.t[-1])
            speed = {rule: 1.0, rule // 2: 0.6, rule * 2: 1.3}[g]
            return {"firings": 1e9 * T, "kernel_ms": ms_per_unit[rule] * T / speed}

    monkeypatch.setattr(autotune, "Batch", FakeBatch)
    monkeypatch.setattr(autotune, "Model", _FakeModel)
    ts = np.linspace(0.01, 1.0, 100)
    nets = []
    for S, R in ((100, 200), (90, 150), (20, 30)):
        nets.append({"left": np.zeros((R, S), np.int32), "right": np.zeros((R, S), np.int32),
                     "x0": np.ones(S), "c": np.ones(R), "ts": ts})
    res = autotune.tune_batch(nets, [1024, 1024, 1024], min_probe_ms=300.0)
    big, small = res["classes"][16], res["classes"][4]
    # class 16: 400 ms per unit; 1/16 of the horizon ends on the sample at 0.06 (24 ms), so the
    # horizon grows by 1.2 * 300 / 24 = 15 at once, to 0.9375 (the sample at 0.93: 372 ms)
    assert big["fraction"] == pytest.approx(1.0 / 16 * 1.2 * 300 / 24.0)
    assert small["fraction"] == 1.0                 # 20 ms per unit: grown to the whole horizon
    assert big["chosen"] == 32 and small["chosen"] == 8
    assert list(res["lanes"]) == [32, 32, 8]        # model 1 is no larger than the probed model 0
