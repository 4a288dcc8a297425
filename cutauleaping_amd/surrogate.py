"""Learned surrogate for parameter sweeps of a reaction network.

A sweep asks, for many settings of uncertain rate constants, for the mean
and standard deviation trajectories of observed species.  The plain answer
runs `replicas` stochastic simulations per setting on the GPU and reduces
them on the device (Sim.moments).  The surrogate answer is an ensemble of
MLPs (cutauleaping_amd.MLP, BF16-MFMA kernels) mapping the log10 rate
multipliers to those trajectories; settings on which the ensemble members
disagree by more than a threshold fall back to the simulator, which is also
the judge of the surrogate's error on held-out settings.

Problem (defaults): a seeded mass-action network with N=50 species and
M=100 reactions (networks.mass_action), 32 uncertain constants
c_j = c_j^nom * 10^u_j with u_j ~ U(-1, 1), 10 observed species at 50
equally spaced times.  Network features: u (32 values).  Targets: log1p of
the per-time, per-species mean and (population) standard deviation over
the replicas, 2 * 50 * 10 = 1000 values, standardised per output.

The reference has no counterpart: it simulates every setting (its
fitness/parameter-estimation drivers call the simulator once per
candidate, kernel.cu:242-345).
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import MLP, Model, Sim
from . import networks

WIDTHS = (32, 256, 256, 256, 1000)


@dataclass
class SweepProblem:
    left: np.ndarray
    right: np.ndarray
    x0: np.ndarray
    c_nom: np.ndarray
    uncertain: np.ndarray      # reaction indices whose constants vary
    observed: np.ndarray       # sampled species indices
    ts: np.ndarray             # sample times (last = t_end)
    seed: int = 0
    info: dict = field(default_factory=dict)

    @classmethod
    def synthetic(cls, N=50, M=100, seed=5, n_uncertain=32, n_observed=10, K=50, t_end=5.0,
                  decades=1.0) -> "SweepProblem":
        """Draws the network (int32-safe at the nominal constants and with
        every uncertain constant raised by 10^decades), the uncertain
        reactions and the observed species from `seed`."""
        for sub in range(100):
            net = networks.mass_action(N, M, seed=seed * 1000 + sub, t_end=t_end)
            rng = np.random.default_rng([seed, sub, 7])
            unc = np.sort(rng.choice(M, size=n_uncertain, replace=False))
            obs = np.sort(rng.choice(N, size=n_observed, replace=False))
            hi = net["c"].copy()
            hi[unc] *= 10.0 ** decades
            if networks._mean_field_ok(net["left"], net["right"], net["x0"], hi, t_end):
                break
        else:
            raise RuntimeError("no int32-safe sweep network found")
        ts = np.arange(1, K + 1, dtype=np.float64) * (t_end / K)
        return cls(net["left"], net["right"], net["x0"], net["c"], unc, obs, ts, seed,
                   dict(N=N, M=M, sub_seed=seed * 1000 + sub, attempt=net["attempt"], decades=decades))

    @property
    def n_in(self) -> int:
        return len(self.uncertain)

    @property
    def n_out(self) -> int:
        return 2 * len(self.ts) * len(self.observed)

    def sample_u(self, n: int, rng) -> np.ndarray:
        d = self.info.get("decades", 1.0)
        return rng.uniform(-d, d, size=(n, self.n_in)).astype(np.float32)

    def rates(self, u) -> np.ndarray:
        u = np.atleast_2d(np.asarray(u, np.float64))
        c = np.tile(self.c_nom, (u.shape[0], 1))
        c[:, self.uncertain] *= 10.0 ** u
        return c

    def simulate(self, u, replicas: int, *, offset: int = 0, seed: int = 0x5EEDC0DE, gpu: int = 0, **opts):
        """Runs `replicas` simulations per row of u in one launch; returns
        (mean, sd, stats, seconds) with mean/sd shaped [rows, K, observed]."""
        u = np.atleast_2d(u)
        rows = u.shape[0]
        if replicas % 64:
            raise ValueError("replicas must be a multiple of 64")
        model = Model.from_arrays(self.left, self.right, self.x0, self.rates(u), self.ts, cs=self.observed)
        t0 = time.perf_counter()
        with Sim(model, 64, rows * replicas // 64, gpu=gpu, offset=offset, seed=seed, **opts) as s:
            s.run()
            mean, sd = s.moments(replicas)
            st = s.stats()
        dt = time.perf_counter() - t0
        model.close()
        return mean, sd, st, dt

    def features(self, mean, sd) -> np.ndarray:
        """[rows, n_out] float32 log1p targets: all means, then all sds."""
        rows = mean.shape[0]
        return np.concatenate([np.log1p(mean).reshape(rows, -1), np.log1p(sd).reshape(rows, -1)],
                              axis=1).astype(np.float32)

    def unfeatures(self, y):
        rows = y.shape[0]
        h = self.n_out // 2
        K, SS = len(self.ts), len(self.observed)
        y = np.asarray(y, np.float64)
        m = np.expm1(np.maximum(y[:, :h], 0.0)).reshape(rows, K, SS)
        s = np.expm1(np.maximum(y[:, h:], 0.0)).reshape(rows, K, SS)
        return m, s


def relative_error(hat, ref) -> np.ndarray:
    """Per-row relative L2 error ||hat - ref|| / ||ref|| over all times and species."""
    hat = hat.reshape(hat.shape[0], -1)
    ref = ref.reshape(ref.shape[0], -1)
    return np.linalg.norm(hat - ref, axis=1) / np.maximum(np.linalg.norm(ref, axis=1), 1e-12)


class Surrogate:
    """An ensemble of BF16-MFMA MLPs over standardised targets."""

    def __init__(self, widths=WIDTHS, members: int = 5, gpu: int = 0, seed: int = 1):
        self.widths = tuple(int(w) for w in widths)
        self.nets = [MLP(self.widths, gpu=gpu, seed=seed * 101 + k) for k in range(members)]
        self.mu = None
        self.sigma = None

    def fit(self, u, y, *, epochs=200, batch=256, lr=1e-3, seed=0):
        """Trains every member on (u, y) (y from SweepProblem.features);
        returns the per-epoch losses, [members, epochs]."""
        y = np.asarray(y, np.float32)
        self.mu = y.mean(axis=0)
        self.sigma = np.maximum(y.std(axis=0), 1e-3).astype(np.float32)
        z = ((y - self.mu) / self.sigma).astype(np.float32)
        losses = []
        for k, net in enumerate(self.nets):
            net.set_data(u, z)
            losses.append(net.train(epochs, batch, lr, seed=seed * 31 + k))
        return np.stack(losses)

    def predict(self, u):
        """(y_mean [rows, n_out], spread [rows]): the ensemble mean in target
        units and the RMS over outputs of the members' standard deviation,
        also in target (log1p) units, i.e. about a relative spread."""
        zs = np.stack([net.forward(u) for net in self.nets])
        ys = zs * self.sigma + self.mu
        spread = np.sqrt(np.mean(ys.var(axis=0), axis=1))
        return ys.mean(axis=0), spread

    def close(self):
        for n in self.nets:
            n.close()


def screened_sweep(problem: SweepProblem, surrogate: Surrogate, u, *, replicas: int, threshold: float,
                   offset: int = 0, seed: int = 0x5EEDC0DE, **opts):
    """Answers a sweep with the surrogate, simulating the rows whose ensemble
    spread exceeds `threshold`.  Returns dict(mean, sd, fallback (bool rows),
    seconds, surrogate_seconds, simulator_seconds)."""
    t0 = time.perf_counter()
    y, spread = surrogate.predict(u)
    mean, sd = problem.unfeatures(y)
    t1 = time.perf_counter()
    fb = spread > threshold
    sim_s = 0.0
    if fb.any():
        m2, s2, _, sim_s = problem.simulate(u[fb], replicas, offset=offset, seed=seed, **opts)
        mean[fb] = m2
        sd[fb] = s2
    t2 = time.perf_counter()
    return dict(mean=mean, sd=sd, fallback=fb, spread=spread, seconds=t2 - t0, surrogate_seconds=t1 - t0,
                simulator_seconds=sim_s)


def sweep_study(problem: SweepProblem, *, n_train=8192, n_test=512, replicas=256, judge_replicas=None,
                members=5, epochs=150, batch=256, lr=1e-3, quantiles=(0.9, 0.95, 0.98, 0.99, 1.0), seed=1,
                gpu=0, log=None, sim_opts=None, n_val=1024, target=0.05, policy=0.95) -> dict:
    """Data generation, ensemble training, threshold selection on validation
    sets and a judged, timed sweep of held-out test sets.

    Training: n_train parameter sets x `replicas` simulations, reduced to
    mean/sd on the device.  Validation: n_val further sets, judged with
    `judge_replicas` simulations each and answered plainly with `replicas`
    (the fallback's answer).  For each candidate threshold (a quantile of the
    ensemble spread over the validation sets; quantile 1.0 = no fallback) the
    screened answer's errors are measured on the validation sets; the chosen
    threshold is the `policy` quantile when both its mean errors are <=
    target, else the one with the fewest fallbacks that meets the target,
    else the most accurate.  Test: n_test new sets, judged likewise, answered
    by the plain sweep (timed) and by the screened sweep at the chosen
    threshold only (timed); the reported speed-up and errors are those of the
    test sets.  Errors are relative L2 per set over all times and species, of
    the means and of the standard deviations, averaged over the sets."""
    sim_opts = dict(sim_opts or {})
    judge_replicas = judge_replicas or replicas
    say = log or (lambda *a: None)
    rng = np.random.default_rng(seed)
    u_tr = problem.sample_u(n_train, rng)
    u_te = problem.sample_u(n_test, rng)
    u_val = problem.sample_u(n_val, rng)

    def simulate_all(u, reps, sd_seed, what):
        chunk = max(1, (1 << 19) // reps)           # sets per launch: about 2^19 simulations
        means, sds, fir = [], [], 0
        for i in range(0, len(u), chunk):
            m, s, st, _ = problem.simulate(u[i:i + chunk], reps, offset=i * reps, seed=sd_seed, gpu=gpu,
                                           **sim_opts)
            means.append(m)
            sds.append(s)
            fir += st["firings"]
            say(f"# {what} {min(i + chunk, len(u))}/{len(u)} sets")
        return np.concatenate(means), np.concatenate(sds), fir

    t0 = time.perf_counter()
    m_tr, s_tr, gen_fir = simulate_all(u_tr, replicas, seed, "training data")
    gen_s = time.perf_counter() - t0
    y_tr = problem.features(m_tr, s_tr)
    sur = Surrogate(WIDTHS if problem.n_in == 32 and problem.n_out == 1000 else
                    (problem.n_in, 256, 256, 256, problem.n_out), members=members, gpu=gpu, seed=seed)
    t0 = time.perf_counter()
    losses = sur.fit(u_tr, y_tr, epochs=epochs, batch=batch, lr=lr, seed=seed)
    train_s = time.perf_counter() - t0
    say(f"# trained {members} members in {train_s:.1f}s, final losses {losses[:, -1].round(4).tolist()}")

    # threshold selection on the validation sets
    vm, vs, _ = simulate_all(u_val, judge_replicas, seed + 4, "validation judge")
    pvm, pvs, _ = simulate_all(u_val, replicas, seed + 5, "validation plain")
    y_val, spread_val = sur.predict(u_val)
    svm, svs = problem.unfeatures(y_val)
    points = []
    for q in quantiles:
        thr = float(np.quantile(spread_val, q)) if q < 1.0 else float("inf")
        fb = spread_val > thr
        e_m = relative_error(np.where(fb[:, None, None], pvm, svm), vm)
        e_s = relative_error(np.where(fb[:, None, None], pvs, svs), vs)
        points.append(dict(quantile=q, threshold=thr, fallback_fraction=float(fb.mean()),
                           err_mean=float(e_m.mean()), err_sd=float(e_s.mean()),
                           err_mean_p95=float(np.percentile(e_m, 95)), err_sd_p95=float(np.percentile(e_s, 95))))
        say(f"# validation quantile {q}: fallback {points[-1]['fallback_fraction']:.3f}, "
            f"errors {points[-1]['err_mean']:.4f} / {points[-1]['err_sd']:.4f}")
    ok = [p for p in points if p["err_mean"] <= target and p["err_sd"] <= target]
    pol = [p for p in ok if p["quantile"] == policy]
    chosen = pol[0] if pol else (min(ok, key=lambda p: p["fallback_fraction"]) if ok else
                                 min(points, key=lambda p: p["err_sd"]))

    # the test sets: judge, plain sweep (timed), screened sweep at the chosen threshold (timed)
    jm, js, _ = simulate_all(u_te, judge_replicas, seed + 1, "test judge")
    pm, ps, _, plain_s = problem.simulate(u_te, replicas, offset=0, seed=seed + 2, gpu=gpu, **sim_opts)
    say(f"# plain sweep of {n_test} sets x {replicas} replicas: {plain_s:.2f}s")
    res = screened_sweep(problem, sur, u_te, replicas=replicas, threshold=chosen["threshold"], offset=0,
                         seed=seed + 3, gpu=gpu, **sim_opts)
    e_m = relative_error(res["mean"], jm)
    e_s = relative_error(res["sd"], js)
    test = dict(quantile=chosen["quantile"], threshold=chosen["threshold"], seconds=res["seconds"],
                surrogate_s=res["surrogate_seconds"], fallback_s=res["simulator_seconds"],
                fallback_fraction=float(res["fallback"].mean()), speedup=plain_s / res["seconds"],
                err_mean=float(e_m.mean()), err_sd=float(e_s.mean()),
                err_mean_p95=float(np.percentile(e_m, 95)), err_sd_p95=float(np.percentile(e_s, 95)))
    say(f"# test at validation quantile {chosen['quantile']}: speed-up {test['speedup']:.1f}x, fallback "
        f"{test['fallback_fraction']:.3f}, errors {test['err_mean']:.4f} / {test['err_sd']:.4f}")
    # the trivial predictor (the training mean) for scale, and the plain sweep's own error
    mm, ms = problem.unfeatures(np.repeat(y_tr.mean(axis=0, keepdims=True), n_test, axis=0))
    sur.close()
    return dict(
        n_train=n_train, n_test=n_test, n_val=n_val, replicas=replicas, judge_replicas=judge_replicas,
        members=members, generation_s=gen_s, generation_firings=int(gen_fir), train_s=train_s,
        final_loss=float(losses[:, -1].mean()), plain_sweep_s=plain_s, validation=points, chosen=chosen,
        test=test, speedup=test["speedup"], screened_sweep_s=test["seconds"],
        fallback_fraction=test["fallback_fraction"], err_mean=test["err_mean"], err_sd=test["err_sd"],
        meets_target=bool(test["err_mean"] <= target and test["err_sd"] <= target),
        surrogate_only_val_err_mean=points[-1]["err_mean"] if quantiles[-1] == 1.0 else None,
        surrogate_only_val_err_sd=points[-1]["err_sd"] if quantiles[-1] == 1.0 else None,
        trivial_err_mean=float(relative_error(mm, jm).mean()),
        trivial_err_sd=float(relative_error(ms, js).mean()),
        plain_err_mean=float(relative_error(pm, jm).mean()),
        plain_err_sd=float(relative_error(ps, js).mean()),
    )
