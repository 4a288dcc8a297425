"""CPU oracle for the tau-leaping hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
may import this package; the product never does.  It wraps
``oracle/libctl_oracle.so`` (built by ``oracle/Makefile``), a plain-C
restatement of cuTauLeaping's per-simulation algorithm
(2phase_n-tau-leaping.cu:1269-2454, kernel.cu:242-345).

Every function of the C library is pinned by a check in tests/ that does not
restate it: Philox by rocRAND-generated known answers, log/exp by libm,
the Poisson sampler by scipy goodness of fit, the model compile by
hand-worked Schlogl values, the simulator by the exact Schlogl master
equation and the exact SSA.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libctl_oracle.so")

REPLAY = 0
REFSEM = 1


class orc_net(C.Structure):
    _fields_ = [("S", C.c_int), ("R", C.c_int),
                ("left", C.POINTER(C.c_int32)), ("right", C.POINTER(C.c_int32)),
                ("locked", C.POINTER(C.c_int32))]


class orc_job(C.Structure):
    _fields_ = [("n", C.c_int),
                ("x0", C.POINTER(C.c_uint32)), ("x0_rows", C.c_int),
                ("c", C.POINTER(C.c_double)), ("c_rows", C.c_int),
                ("ts", C.POINTER(C.c_double)), ("K", C.c_int),
                ("cs", C.POINTER(C.c_int32)), ("SS", C.c_int),
                ("seed_lo", C.c_uint32), ("seed_hi", C.c_uint32),
                ("sim_base", C.c_uint32),
                ("nc", C.c_int), ("eps", C.c_float), ("ssa_steps", C.c_int),
                ("ssa_factor", C.c_float), ("force_ssa", C.c_int),
                ("precision", C.c_int), ("threads", C.c_int),
                ("nfs", C.c_int), ("fs", C.POINTER(C.c_int32)), ("fv", C.POINTER(C.c_uint32)), ("fv_rows", C.c_int),
                ("group", C.c_int)]


class orc_stat(C.Structure):
    _fields_ = [("firings", C.c_uint64), ("leaps", C.c_uint64), ("rejections", C.c_uint64),
                ("switches", C.c_uint64), ("ssa_steps", C.c_uint64), ("poisson", C.c_uint64),
                ("gaps", C.c_uint64), ("t_end", C.c_double), ("events", C.c_uint32), ("absorbed", C.c_uint32)]


class orc_model(C.Structure):
    _fields_ = [("S", C.c_int), ("R", C.c_int), ("nL", C.c_int), ("nV", C.c_int), ("nF", C.c_int)] + [
        (n, C.POINTER(C.c_int)) for n in ("Lptr", "Ls", "Lv", "Vptr", "Vr", "Vv", "Fptr", "Fs", "Fv",
                                          "FSptr", "FSr", "FSv")
    ] + [("hor", C.POINTER(C.c_ubyte)), ("hort", C.POINTER(C.c_ubyte))]


STAT_DTYPE = np.dtype([("firings", "u8"), ("leaps", "u8"), ("rejections", "u8"), ("switches", "u8"),
                       ("ssa_steps", "u8"), ("poisson", "u8"), ("gaps", "u8"), ("t_end", "f8"), ("events", "u4"),
                       ("absorbed", "u4")], align=True)

_lib = None


def build() -> None:
    """Compile libctl_oracle.so with oracle/Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        L.orc_run.argtypes = [C.POINTER(orc_net), C.POINTER(orc_job), C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_run.restype = C.c_int
        L.orc_compile.argtypes = [C.POINTER(orc_net), C.POINTER(orc_model)]
        L.orc_compile.restype = C.c_int
        L.orc_model_free.argtypes = [C.POINTER(orc_model)]
        L.orc_philox.argtypes = [C.POINTER(C.c_uint32)] * 3
        L.orc_logf.argtypes = [C.c_float]
        L.orc_logf.restype = C.c_float
        L.orc_expf.argtypes = [C.c_float]
        L.orc_expf.restype = C.c_float
        L.orc_log1m.argtypes = [C.c_float]
        L.orc_log1m.restype = C.c_float
        L.orc_poisson_stream.argtypes = [C.c_uint32] * 4 + [C.c_float, C.c_int, C.c_void_p]
        L.orc_uniform_stream.argtypes = [C.c_uint32] * 4 + [C.c_int, C.c_int, C.c_void_p]
        _lib = L
    return _lib


def _ptr(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def _net(left, right, locked):
    left = np.ascontiguousarray(left, dtype=np.int32)
    right = np.ascontiguousarray(right, dtype=np.int32)
    R, S = left.shape
    keep = [left, right]
    lk = None
    if locked is not None:
        lk = np.ascontiguousarray(locked, dtype=np.int32)
        keep.append(lk)
    net = orc_net(S, R, _ptr(left, C.c_int32), _ptr(right, C.c_int32),
                  _ptr(lk, C.c_int32) if lk is not None else C.POINTER(C.c_int32)())
    return net, keep


def compile_model(left, right, locked=None) -> dict:
    """Model compile of the oracle (flats, HOR, HORtype) as numpy arrays."""
    net, keep = _net(left, right, locked)
    m = orc_model()
    if lib().orc_compile(C.byref(net), C.byref(m)) != 0:
        raise ValueError("oracle: model compile failed")
    S, R = m.S, m.R
    arr = lambda p, n: np.ctypeslib.as_array(p, shape=(n,)).copy() if n > 0 else np.zeros(0, np.int32)
    out = dict(S=S, R=R, nL=m.nL, nV=m.nV, nF=m.nF,
               Lptr=arr(m.Lptr, R + 1), Ls=arr(m.Ls, m.nL), Lv=arr(m.Lv, m.nL),
               Vptr=arr(m.Vptr, S + 1), Vr=arr(m.Vr, m.nV), Vv=arr(m.Vv, m.nV),
               Fptr=arr(m.Fptr, R + 1), Fs=arr(m.Fs, m.nF), Fv=arr(m.Fv, m.nF),
               hor=np.ctypeslib.as_array(m.hor, shape=(S,)).copy(),
               hort=np.ctypeslib.as_array(m.hort, shape=(S,)).copy())
    lib().orc_model_free(C.byref(m))
    return out


def run(left, right, x0, c, ts, *, locked=None, cs=None, n=None, seed=1, sim_base=0,
        nc=10, eps=0.03, ssa_steps=100, ssa_factor=10.0, force_ssa=False,
        precision=REPLAY, threads=0, final=False, group=0, feed_species=None, feed_values=None):
    """Runs n simulations; returns (samples[n, K, SS] uint32, stats structured
    array) and, with final=True, also the final states [n, S] int32.
    group=G > 0 runs the group algorithm with G lanes per simulation;
    feed_species / feed_values ([rows, nfs]) enable per-simulation feeds."""
    x0 = np.ascontiguousarray(np.atleast_2d(x0), dtype=np.uint32)
    c = np.ascontiguousarray(np.atleast_2d(c), dtype=np.float64)
    ts = np.ascontiguousarray(ts, dtype=np.float64)
    R, S = np.asarray(left).shape
    if n is None:
        n = max(x0.shape[0], c.shape[0])
    if cs is None:
        cs = np.arange(S)
    cs = np.ascontiguousarray(cs, dtype=np.int32)
    net, keep = _net(left, right, locked)
    job = orc_job(n, _ptr(x0, C.c_uint32), x0.shape[0], _ptr(c, C.c_double), c.shape[0],
                  _ptr(ts, C.c_double), len(ts), _ptr(cs, C.c_int32), len(cs),
                  seed & 0xffffffff, (seed >> 32) & 0xffffffff, sim_base,
                  nc, eps, ssa_steps, ssa_factor, int(bool(force_ssa)), precision, threads,
                  0, C.POINTER(C.c_int32)(), C.POINTER(C.c_uint32)(), 0, group)
    if feed_species is not None and len(feed_species):
        fsa = np.ascontiguousarray(feed_species, dtype=np.int32)
        fva = np.ascontiguousarray(np.atleast_2d(feed_values), dtype=np.uint32)
        assert fva.shape[1] == len(fsa)
        job.nfs, job.fs, job.fv, job.fv_rows = len(fsa), _ptr(fsa, C.c_int32), _ptr(fva, C.c_uint32), fva.shape[0]
    out = np.zeros((n, len(ts), len(cs)), dtype=np.uint32)
    st = np.zeros(n, dtype=STAT_DTYPE)
    assert st.dtype.itemsize == C.sizeof(orc_stat)
    xf = np.zeros((n, S), dtype=np.int32)
    rc = lib().orc_run(C.byref(net), C.byref(job), out.ctypes.data, st.ctypes.data, xf.ctypes.data)
    if rc != 0:
        raise ValueError(f"oracle: orc_run failed ({rc})")
    return (out, st, xf) if final else (out, st)


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox(c, k, o)
    return list(o)


def logf(x: float) -> float:
    return lib().orc_logf(x)


def expf(x: float) -> float:
    return lib().orc_expf(x)


def log1m(v: float) -> float:
    return lib().orc_log1m(v)


def poisson_stream(lam: float, count: int, seed=7, sim=0, ev=0) -> np.ndarray:
    out = np.zeros(count, dtype=np.int32)
    lib().orc_poisson_stream(seed & 0xffffffff, seed >> 32, sim, ev, lam, count, out.ctypes.data)
    return out


def uniform_stream(count: int, open_closed=True, seed=7, sim=0, ev=0) -> np.ndarray:
    out = np.zeros(count, dtype=np.float32)
    lib().orc_uniform_stream(seed & 0xffffffff, seed >> 32, sim, ev, int(open_closed), count,
                             out.ctypes.data)
    return out
