"""Seeded synthetic reaction networks for parity tests and benchmarks.

mass_action(N, M, seed)
    Random mass-action network: every reaction has order k ~ U{0,1,2} with k
    reactant species drawn uniformly with replacement (a repeated species is a
    coefficient of 2), m ~ U{1,2} product species drawn uniformly with
    replacement, rate constant c = 10^U(-6, 0) and initial counts
    floor(10^U(1, 5)).  A drawn network is rejected (and the next attempt of
    the same seed is drawn) when its mean-field ODE leaves the int32-safe
    range: any species above 1e8 before t_end.  The accepted attempt number is
    returned with the network so the draw can be reproduced.

enzyme_cascade(stages, seed)
    E_k + S_k <-> C_k -> E_k + S_{k+1}, k = 1..stages, with the last product
    S_{stages+1}; binding / unbinding log-uniform up to 1e3, catalysis 1e-3,
    enzymes 5-20, S_1 in [1e5, 1e6].

schlogl()
    The Schlogl model of the reference repository's schlogl.xml.
"""
from __future__ import annotations

import numpy as np


class _OutOfBudget(Exception):
    pass


def _mean_field_ok(left, right, x0, c, t_end, limit=1e8, max_evals=20000):
    """Integrates dx/dt = (right-left)^T a(x) with a_j = c_j prod C(x, v)
    (orders 0-2: one reactant, a pair of distinct reactants, or a dimer).
    A system the integrator cannot cross in max_evals evaluations (too stiff
    to finish cheaply) counts as safe when every count it reached stayed
    below limit / 100, and as unsafe otherwise."""
    from scipy.integrate import solve_ivp
    from scipy.sparse import csr_matrix

    V = (right - left).astype(np.float64)
    R, S = left.shape
    if left.sum(axis=1).max(initial=0) > 2:
        raise ValueError("mean-field check supports reaction orders 0-2")
    # reactant slots: p[j], q[j] species (or S for "none", whose count is 1);
    # dimers (coefficient 2) use x (x - 1) / 2
    p = np.full(R, S)
    q = np.full(R, S)
    dimer = np.zeros(R, bool)
    for j in range(R):
        nz = np.nonzero(left[j])[0]
        if len(nz) == 1 and left[j, nz[0]] == 2:
            p[j] = q[j] = nz[0]
            dimer[j] = True
        elif len(nz) >= 1:
            p[j] = nz[0]
            if len(nz) == 2:
                q[j] = nz[1]
    VT = csr_matrix(V.T)
    evals = [0]
    peak = [float(np.max(x0))]

    def f(t, x):
        evals[0] += 1
        peak[0] = max(peak[0], float(np.max(x)))
        if evals[0] > max_evals:
            raise _OutOfBudget
        xe = np.append(np.maximum(x, 0.0), 1.0)
        xp, xq = xe[p], xe[q]
        xq = np.where(dimer, np.maximum(xq - 1.0, 0.0) / 2.0, xq)
        return VT @ (c * xp * xq)

    def blowup(t, x):
        return limit - np.max(x)
    blowup.terminal = True

    try:
        sol = solve_ivp(f, (0.0, t_end), x0.astype(np.float64), method="LSODA", events=blowup,
                        rtol=1e-4, atol=1e-2)
    except _OutOfBudget:
        return peak[0] < limit / 100
    if sol.status == -1:
        return False
    return bool(np.all(np.isfinite(sol.y)) and np.max(sol.y) < limit)


def mass_action(N=100, M=200, seed=1, t_end=50.0, max_attempts=1000, check=True):
    """Returns dict(left, right, x0, c, attempt)."""
    for attempt in range(max_attempts):
        rng = np.random.default_rng([seed, attempt])
        left = np.zeros((M, N), np.int32)
        right = np.zeros((M, N), np.int32)
        for j in range(M):
            k = rng.integers(0, 3)
            for s in rng.integers(0, N, size=k):
                left[j, s] += 1
            m = rng.integers(1, 3)
            for s in rng.integers(0, N, size=m):
                right[j, s] += 1
        c = 10.0 ** rng.uniform(-6.0, 0.0, size=M)
        x0 = np.floor(10.0 ** rng.uniform(1.0, 5.0, size=N)).astype(np.uint32)
        if not check or _mean_field_ok(left, right, x0, c, t_end):
            return dict(left=left, right=right, x0=x0, c=c, attempt=attempt)
    raise RuntimeError("no int32-safe network found")


def enzyme_cascade(stages=20, seed=4):
    """Species order: E_1..E_n, S_1..S_{n+1}, C_1..C_n (3n+1 species, 3n reactions)."""
    rng = np.random.default_rng(seed)
    n = stages
    S = 3 * n + 1
    E = lambda k: k
    Sub = lambda k: n + k
    Cx = lambda k: 2 * n + 1 + k
    left = np.zeros((3 * n, S), np.int32)
    right = np.zeros((3 * n, S), np.int32)
    c = np.zeros(3 * n)
    for k in range(n):
        b, u, cat = 3 * k, 3 * k + 1, 3 * k + 2
        left[b, E(k)] = 1; left[b, Sub(k)] = 1; right[b, Cx(k)] = 1          # E + S -> C
        left[u, Cx(k)] = 1; right[u, E(k)] = 1; right[u, Sub(k)] = 1         # C -> E + S
        left[cat, Cx(k)] = 1; right[cat, E(k)] = 1; right[cat, Sub(k + 1)] = 1  # C -> E + S'
        c[b] = 10.0 ** rng.uniform(-3.0, 3.0)
        c[u] = 10.0 ** rng.uniform(-3.0, 3.0)
        c[cat] = 1e-3
    x0 = np.zeros(S, np.uint32)
    for k in range(n):
        x0[E(k)] = rng.integers(5, 21)
    x0[Sub(0)] = int(10.0 ** rng.uniform(5.0, 6.0))
    return dict(left=left, right=right, x0=x0, c=c)


def schlogl():
    """Schlogl (A, X, B with A, B buffered) as in the reference's schlogl.xml."""
    left = np.array([[1, 2, 0], [0, 3, 0], [0, 0, 1], [0, 1, 0]], np.int32)
    right = np.array([[0, 3, 0], [1, 2, 0], [0, 1, 0], [0, 0, 1]], np.int32)
    return dict(left=left, right=right, x0=np.array([100000, 250, 200000], np.uint32),
                c=np.array([3e-7, 1e-4, 1e-3, 3.5]), feed=np.array([1, 0, 1], np.int32))


def sample_times(K, t_end):
    return np.array([t_end * (i + 1) / K for i in range(K)], dtype=np.float64)


def _batched_size(rng):
    N = int(np.floor(10.0 ** rng.uniform(1.0, np.log10(500.0))))
    M = int(np.floor(10.0 ** rng.uniform(1.0, 3.0)))
    replicas = int(np.floor(10.0 ** rng.uniform(np.log10(64.0), np.log10(4096.0))))
    return N, M, replicas


def batched_sizes(count=8192, seed=2):
    """(N, M, replicas) of models 0..count-1 of batched_models(count, seed),
    without drawing their networks: an int array [count, 3]."""
    return np.array([_batched_size(np.random.default_rng([seed, i, 0xB7])) for i in range(count)], np.int64)


def batched_model(i, seed=2, t_end=1.0, check=True, observed=10):
    """Model i of batched_models(..., seed): the same dict as its list entry."""
    return _batched_one((i, seed, t_end, check, observed))


def _batched_one(args):
    i, seed, t_end, check, observed = args
    rng = np.random.default_rng([seed, i, 0xB7])
    N, M, replicas = _batched_size(rng)
    net = mass_action(N, M, seed=seed * 100003 + i, t_end=t_end, check=check)
    net["replicas"] = replicas
    net["cs"] = np.sort(rng.choice(N, size=min(observed, N), replace=False)).astype(np.int32)
    return net


def batched_models(count=8192, seed=2, t_end=1.0, check=True, observed=10, processes=1, log=None):
    """Uneven batch of independent mass-action models: model i draws
    N = floor(10^U(1, log10 500)) species, M = floor(10^U(1, 3)) reactions and
    floor(10^U(log10 64, log10 4096)) replicas, then its network as
    mass_action(N, M, seed * 100003 + i); `observed` sampled species per model.
    Returns a list of dicts (left, right, x0, c, attempt, replicas, cs).
    processes > 1 draws the models in spawned workers whose BLAS / OpenMP
    pools are one thread each (forked workers inherit the parent's thread
    pools and stall); log(msg) reports progress every 128 models."""
    jobs = [(i, seed, t_end, check, observed) for i in range(count)]
    out = []
    if processes > 1:
        import multiprocessing as mp
        import os
        keys = ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS")
        saved = {k: os.environ.get(k) for k in keys}
        os.environ.update({k: "1" for k in keys})
        try:
            with mp.get_context("spawn").Pool(processes) as pool:
                for net in pool.imap(_batched_one, jobs, chunksize=8):
                    out.append(net)
                    if log and len(out) % 128 == 0:
                        log(f"# {len(out)}/{count} models drawn")
        finally:
            for k, v in saved.items():
                if v is None:
                    os.environ.pop(k, None)
                else:
                    os.environ[k] = v
        return out
    for j in jobs:
        out.append(_batched_one(j))
        if log and len(out) % 128 == 0:
            log(f"# {len(out)}/{count} models drawn")
    return out
