"""MI355X-native tau-leaping (cuTauLeaping's hot path) -- Python host mirror.

The product is ``libctl.so`` (hand-written HIP for gfx950 behind the C ABI of
``include/ctl.h``) and the ``cuTauLeaping`` CLI next to it.  This module binds
that ABI with ctypes; it never falls back to a CPU path: if the library is
missing, importing a handle raises, and every simulation call needs a GPU.

    m = Model.from_folder("tests/golden/schlogl", offset=0, n=1024)
    with Sim(m, threads=64, blocks=16, seed=1) as s:
        s.run()
        samples = s.samples()        # [sim, sample, sampled species] uint32

Reference interface mirrored (2phase_n-tau-leaping.cu / kernel.cu):
  read_stoichiometry + read_quantities_offset + read_parameters_offset +
  read_instants  -> Model.from_folder
  allocate_space + setup_random + PrepareVariables -> Sim(...)
  the host do-while of kernel.cu:242-345           -> Sim.run
  WriteDynamics2 / PrintDynamics2                  -> Sim.write_dynamics
  calculateFitness                                 -> Sim.fitness
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libctl.so")
CLI_PATH = os.path.join(HERE, "cuTauLeaping")
REPO = os.path.dirname(HERE)

CTL_OK = 0
CTL_E_ARGS = -1
CTL_E_DEVICE = -2
CTL_E_NOMEM = -3
CTL_E_ROWS = -4
CTL_E_WIDTH = -5
CTL_E_NOFILE = -6
CTL_E_TTS = -10
CTL_E_STATE = -11
CTL_E_LIMIT = -12

# Every entry point declared in include/ctl.h.
EXPORTS = (
    "ctl_opts_default", "ctl_model_load_folder", "ctl_model_from_arrays", "ctl_model_info_get",
    "ctl_model_structure", "ctl_model_load_fitness", "ctl_model_destroy", "ctl_sim_create",
    "ctl_sim_run", "ctl_sim_samples", "ctl_sim_samples_device", "ctl_sim_stats",
    "ctl_sim_sim_stats", "ctl_sim_final_state", "ctl_write_dynamics", "ctl_fitness",
    "ctl_sim_destroy", "ctl_last_error", "ctl_version", "ctl_device_count", "ctl_sim_moments",
    "ctl_write_samples_binary", "ctl_batch_create", "ctl_batch_run", "ctl_batch_samples",
    "ctl_batch_sim_stats", "ctl_batch_stats", "ctl_batch_stream_base", "ctl_batch_destroy",
    "ctl_mlp_create", "ctl_mlp_num_params", "ctl_mlp_params", "ctl_mlp_forward", "ctl_mlp_forward_device",
    "ctl_mlp_gradient", "ctl_mlp_set_data", "ctl_mlp_train", "ctl_mlp_destroy",
    "ctl_model_load_feeds", "ctl_model_set_feeds", "ctl_sim_moments_noalloc", "ctl_batch_model_lanes",
    "ctl_batch_create_lanes", "ctl_phase_clock",
)


class CtlError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"ctl error {code}: {msg}")
        self.code = code


class ctl_model_desc(C.Structure):
    _fields_ = [("S", C.c_int32), ("R", C.c_int32),
                ("left", C.POINTER(C.c_int32)), ("right", C.POINTER(C.c_int32)),
                ("feed", C.POINTER(C.c_int32)),
                ("x0", C.POINTER(C.c_uint32)), ("x0_rows", C.c_int64),
                ("c", C.POINTER(C.c_double)), ("c_rows", C.c_int64),
                ("t_vector", C.POINTER(C.c_double)), ("K", C.c_int32),
                ("cs_vector", C.POINTER(C.c_int32)), ("SS", C.c_int32)]


class ctl_opts(C.Structure):
    _fields_ = [("threads", C.c_int32), ("blocks", C.c_int32), ("gpu", C.c_int32),
                ("offset", C.c_int64), ("force_ssa", C.c_int32), ("nc", C.c_int32),
                ("eps", C.c_float), ("ssa_steps", C.c_int32), ("ssa_factor", C.c_float),
                ("seed", C.c_uint64), ("verbose", C.c_int32), ("block_size", C.c_int32),
                ("slots", C.c_int32), ("iters_per_launch", C.c_int32), ("ssa_chunk", C.c_int32),
                ("lanes_per_sim", C.c_int32), ("compact_interval", C.c_int32)]


class ctl_stats(C.Structure):
    _fields_ = [("firings", C.c_uint64), ("leaps", C.c_uint64), ("rejections", C.c_uint64),
                ("switches", C.c_uint64), ("ssa_steps", C.c_uint64), ("poisson", C.c_uint64),
                ("gaps", C.c_uint64), ("simulations", C.c_uint64), ("absorbed", C.c_uint64),
                ("kernel_ms", C.c_double), ("wall_ms", C.c_double), ("launches", C.c_int32),
                ("slots", C.c_int32), ("block_size", C.c_int32), ("lds_state", C.c_int32),
                ("lanes_per_sim", C.c_int32)]


class ctl_model_info(C.Structure):
    _fields_ = [("S", C.c_int32), ("R", C.c_int32), ("K", C.c_int32), ("SS", C.c_int32),
                ("nL", C.c_int32), ("nV", C.c_int32), ("nF", C.c_int32),
                ("x0_rows", C.c_int64), ("c_rows", C.c_int64), ("row_offset", C.c_int64),
                ("t_max", C.c_double)]


SIM_STAT_DTYPE = np.dtype([("firings", "u8"), ("ssa_steps", "u8"), ("poisson", "u8"),
                           ("gaps", "u8"), ("t_end", "f8"), ("leaps", "u4"), ("rejections", "u4"),
                           ("switches", "u4"), ("events", "u4"), ("absorbed", "u4"),
                           ("reserved", "u4")])

_lib = None


def build(verbose: bool = False) -> None:
    """Compiles libctl.so and the CLI for gfx950 (cutauleaping_amd/csrc/Makefile)."""
    jobs = str(min(8, os.cpu_count() or 1))
    out = None if verbose else subprocess.DEVNULL
    subprocess.run(["make", "-s", "-j", jobs, "-C", os.path.join(HERE, "csrc")], check=True,
                   stdout=out)


def lib():
    """The loaded libctl.so; raises when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with cutauleaping_amd.build() "
                              "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        vp, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
        L.ctl_opts_default.argtypes = [C.POINTER(ctl_opts)]
        L.ctl_opts_default.restype = None
        L.ctl_model_load_folder.argtypes = [C.c_char_p, i64, i64, C.c_int, C.POINTER(vp)]
        L.ctl_model_from_arrays.argtypes = [C.POINTER(ctl_model_desc), C.POINTER(vp)]
        L.ctl_model_info_get.argtypes = [vp, C.POINTER(ctl_model_info)]
        L.ctl_model_structure.argtypes = [vp, vp, vp, vp, vp, vp]
        L.ctl_model_load_fitness.argtypes = [vp, C.c_char_p, i32, i32, C.c_int]
        L.ctl_model_load_feeds.argtypes = [vp, C.c_char_p, i64, i64, C.c_int]
        L.ctl_model_set_feeds.argtypes = [vp, vp, i32, vp, i64]
        L.ctl_model_destroy.argtypes = [vp]
        L.ctl_model_destroy.restype = None
        L.ctl_sim_create.argtypes = [vp, C.POINTER(ctl_opts), C.POINTER(vp)]
        L.ctl_sim_run.argtypes = [vp]
        L.ctl_sim_samples.argtypes = [vp, vp, C.c_size_t, C.c_int]
        L.ctl_sim_samples_device.argtypes = [vp, C.POINTER(vp), C.POINTER(C.c_size_t)]
        L.ctl_sim_stats.argtypes = [vp, C.POINTER(ctl_stats)]
        L.ctl_sim_sim_stats.argtypes = [vp, vp, C.c_size_t]
        L.ctl_sim_final_state.argtypes = [vp, vp, vp]
        L.ctl_write_dynamics.argtypes = [vp, C.c_char_p, C.c_int]
        L.ctl_fitness.argtypes = [vp, vp, C.c_size_t]
        L.ctl_sim_moments.argtypes = [vp, i32, vp, vp, C.c_size_t]
        L.ctl_sim_moments_noalloc.argtypes = [vp, i32, vp, vp, C.c_size_t]
        L.ctl_write_samples_binary.argtypes = [vp, C.c_char_p]
        L.ctl_batch_create.argtypes = [vp, vp, i32, C.POINTER(ctl_opts), C.POINTER(vp)]
        L.ctl_batch_create_lanes.argtypes = [vp, vp, vp, i32, C.POINTER(ctl_opts), C.POINTER(vp)]
        L.ctl_phase_clock.argtypes = [i32, vp, i32, i32]
        L.ctl_batch_run.argtypes = [vp]
        L.ctl_batch_samples.argtypes = [vp, i32, vp, C.c_size_t]
        L.ctl_batch_sim_stats.argtypes = [vp, i32, vp, C.c_size_t]
        L.ctl_batch_stats.argtypes = [vp, C.POINTER(ctl_stats)]
        L.ctl_batch_stream_base.argtypes = [vp, i32, C.POINTER(i64)]
        L.ctl_batch_model_lanes.argtypes = [vp, i32, C.POINTER(i32)]
        L.ctl_batch_destroy.argtypes = [vp]
        L.ctl_batch_destroy.restype = None
        L.ctl_sim_destroy.argtypes = [vp]
        L.ctl_sim_destroy.restype = None
        L.ctl_mlp_create.argtypes = [vp, i32, i32, C.c_uint64, C.POINTER(vp)]
        L.ctl_mlp_num_params.argtypes = [vp, C.POINTER(i64)]
        L.ctl_mlp_params.argtypes = [vp, vp, C.c_size_t, i32]
        L.ctl_mlp_forward.argtypes = [vp, vp, i64, vp]
        L.ctl_mlp_forward_device.argtypes = [vp, vp, i64, vp, C.POINTER(C.c_float)]
        L.ctl_mlp_gradient.argtypes = [vp, vp, vp, i32, vp, C.c_size_t, C.POINTER(C.c_float)]
        L.ctl_mlp_set_data.argtypes = [vp, vp, vp, i64]
        L.ctl_mlp_train.argtypes = [vp, i32, i32, C.c_float, C.c_uint64, vp]
        L.ctl_mlp_destroy.argtypes = [vp]
        L.ctl_mlp_destroy.restype = None
        L.ctl_last_error.restype = C.c_char_p
        L.ctl_version.restype = C.c_char_p
        L.ctl_device_count.restype = C.c_int
        _lib = L
    return _lib


def _check(rc: int) -> None:
    if rc != 0:
        raise CtlError(rc, lib().ctl_last_error().decode(errors="replace"))


def device_count() -> int:
    return lib().ctl_device_count()


def default_opts() -> ctl_opts:
    o = ctl_opts()
    lib().ctl_opts_default(C.byref(o))
    return o


class Model:
    """A compiled reaction network plus its per-simulation inputs (C handle)."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle) if not isinstance(handle, C.c_void_p) else handle

    @classmethod
    def from_folder(cls, folder: str, offset: int = 0, n: int = 0, verbose: bool = False) -> "Model":
        h = C.c_void_p()
        _check(lib().ctl_model_load_folder(folder.encode(), offset, n, int(verbose), C.byref(h)))
        return cls(h)

    @classmethod
    def from_arrays(cls, left, right, x0, c, t_vector, feed=None, cs=None) -> "Model":
        left = np.ascontiguousarray(left, dtype=np.int32)
        right = np.ascontiguousarray(right, dtype=np.int32)
        R, S = left.shape
        x0 = np.ascontiguousarray(np.atleast_2d(x0), dtype=np.uint32)
        c = np.ascontiguousarray(np.atleast_2d(c), dtype=np.float64)
        ts = np.ascontiguousarray(t_vector, dtype=np.float64)
        if x0.shape[1] != S or c.shape[1] != R:
            raise ValueError("x0 / c widths must match species / reactions")
        fd = None if feed is None else np.ascontiguousarray(feed, dtype=np.int32)
        csv = None if cs is None else np.ascontiguousarray(cs, dtype=np.int32)
        P = lambda a, t: a.ctypes.data_as(C.POINTER(t)) if a is not None else C.POINTER(t)()
        d = ctl_model_desc(S, R, P(left, C.c_int32), P(right, C.c_int32), P(fd, C.c_int32),
                           P(x0, C.c_uint32), x0.shape[0], P(c, C.c_double), c.shape[0],
                           P(ts, C.c_double), len(ts), P(csv, C.c_int32),
                           0 if csv is None else len(csv))
        h = C.c_void_p()
        _check(lib().ctl_model_from_arrays(C.byref(d), C.byref(h)))
        return cls(h)

    def info(self) -> ctl_model_info:
        i = ctl_model_info()
        _check(lib().ctl_model_info_get(self._h, C.byref(i)))
        return i

    def structure(self) -> dict:
        i = self.info()
        hor = np.zeros(i.S, np.uint8)
        hort = np.zeros(i.S, np.uint8)
        lf = np.zeros((i.nL, 3), np.int32)
        rf = np.zeros((i.nV, 3), np.int32)
        ff = np.zeros((i.nF, 3), np.int32)
        _check(lib().ctl_model_structure(self._h, hor.ctypes.data, hort.ctypes.data, lf.ctypes.data,
                                         rf.ctypes.data, ff.ctypes.data))
        return dict(hor=hor, hort=hort, left_flat=lf, rav_flat=rf, fear_flat=ff)

    def load_feeds(self, folder: str, offset: int = 0, n: int = 0) -> None:
        """Per-simulation feeds from folder/MX_feed (off unless loaded)."""
        _check(lib().ctl_model_load_feeds(self._h, folder.encode(), offset, n, 0))

    def set_feeds(self, species, values) -> None:
        """Species species[u] is reset to values[row][u] after every leap and
        SSA step (rows follow the x0 rule: 1, one per simulation, or groups)."""
        sp = np.ascontiguousarray(species, dtype=np.int32)
        v = np.ascontiguousarray(np.atleast_2d(values), dtype=np.uint32)
        if len(sp) and v.shape[1] != len(sp):
            raise ValueError("values must have one column per feed species")
        _check(lib().ctl_model_set_feeds(self._h, sp.ctypes.data, len(sp), v.ctypes.data, v.shape[0] if len(sp) else 0))

    def load_fitness(self, folder: str, threads: int, blocks: int) -> None:
        _check(lib().ctl_model_load_fitness(self._h, folder.encode(), threads, blocks, 0))

    def close(self) -> None:
        if self._h:
            lib().ctl_model_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Sim:
    """Device state for threads*blocks simulations of a Model (C handle)."""

    def __init__(self, model: Model, threads: int, blocks: int, *, gpu=0, offset=0, force_ssa=False,
                 seed=0x5EEDC0DE, nc=10, eps=0.03, ssa_steps=100, ssa_factor=10.0, block_size=0,
                 slots=0, iters_per_launch=0, ssa_chunk=0, lanes_per_sim=0, compact_interval=0, verbose=False):
        o = default_opts()
        o.threads, o.blocks, o.gpu, o.offset = threads, blocks, gpu, offset
        o.force_ssa = int(bool(force_ssa))
        o.seed = seed
        o.nc, o.eps, o.ssa_steps, o.ssa_factor = nc, eps, ssa_steps, ssa_factor
        o.block_size, o.slots, o.iters_per_launch, o.ssa_chunk = block_size, slots, iters_per_launch, ssa_chunk
        o.lanes_per_sim = lanes_per_sim
        o.compact_interval = compact_interval
        o.verbose = int(bool(verbose))
        self.model = model
        self.n = threads * blocks
        self.opts = o
        info = model.info()
        self.S, self.R, self.K, self.SS = info.S, info.R, info.K, info.SS
        self._h = C.c_void_p()
        _check(lib().ctl_sim_create(model._h, C.byref(o), C.byref(self._h)))

    def run(self) -> "Sim":
        _check(lib().ctl_sim_run(self._h))
        return self

    def samples(self, layout: int = 1) -> np.ndarray:
        """layout 1: [sim, K, SS]; layout 0: the reference's [K, SS, sim]."""
        out = np.empty(self.n * self.K * self.SS, dtype=np.uint32)
        _check(lib().ctl_sim_samples(self._h, out.ctypes.data, out.size, layout))
        return out.reshape((self.n, self.K, self.SS) if layout == 1 else (self.K, self.SS, self.n))

    def stats(self) -> dict:
        s = ctl_stats()
        _check(lib().ctl_sim_stats(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in ctl_stats._fields_}

    def sim_stats(self) -> np.ndarray:
        out = np.zeros(self.n, dtype=SIM_STAT_DTYPE)
        assert out.dtype.itemsize == 64
        _check(lib().ctl_sim_sim_stats(self._h, out.ctypes.data, self.n))
        return out

    def final_state(self):
        x = np.zeros((self.n, self.S), np.int32)
        t = np.zeros(self.n, np.float64)
        _check(lib().ctl_sim_final_state(self._h, x.ctypes.data, t.ctypes.data))
        return x, t

    def write_dynamics(self, path_prefix: str) -> None:
        _check(lib().ctl_write_dynamics(self._h, path_prefix.encode(), 0))

    def fitness(self, groups: int) -> np.ndarray:
        out = np.zeros(self.opts.threads * groups, np.float64)
        _check(lib().ctl_fitness(self._h, out.ctypes.data, out.size))
        return out

    def moments(self, group: int, noalloc: bool = False):
        """Per-group mean and standard deviation, [n/group, K, SS] each (on the
        device); noalloc=True uses the allocation-free entry point."""
        g = self.n // group
        mean = np.zeros((g, self.K, self.SS), np.float64)
        sd = np.zeros_like(mean)
        fn = lib().ctl_sim_moments_noalloc if noalloc else lib().ctl_sim_moments
        _check(fn(self._h, group, mean.ctypes.data, sd.ctypes.data, mean.size))
        return mean, sd

    def write_samples_binary(self, path: str) -> None:
        _check(lib().ctl_write_samples_binary(self._h, path.encode()))

    def close(self) -> None:
        if self._h:
            lib().ctl_sim_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many models of uneven size in one persistent launch (C handle)."""

    def __init__(self, models, replicas, *, gpu=0, offset=0, force_ssa=False, seed=0x5EEDC0DE, nc=10,
                 eps=0.03, ssa_steps=100, ssa_factor=10.0, block_size=0, slots=0, iters_per_launch=0,
                 ssa_chunk=0, lanes_per_sim=0, lanes=None, verbose=False):
        """lanes: optional lanes per simulation for each model (0 = the
        library's choice), e.g. from autotune.tune_batch."""
        o = default_opts()
        o.verbose = int(bool(verbose))
        o.gpu, o.offset, o.force_ssa, o.seed = gpu, offset, int(bool(force_ssa)), seed
        o.nc, o.eps, o.ssa_steps, o.ssa_factor = nc, eps, ssa_steps, ssa_factor
        o.block_size, o.slots, o.iters_per_launch, o.ssa_chunk = block_size, slots, iters_per_launch, ssa_chunk
        o.lanes_per_sim = lanes_per_sim
        self.models = list(models)
        self.replicas = np.ascontiguousarray(replicas, dtype=np.int64)
        self.info = [m.info() for m in self.models]
        handles = (C.c_void_p * len(self.models))(*[m._h.value for m in self.models])
        self._h = C.c_void_p()
        if lanes is None:
            _check(lib().ctl_batch_create(handles, self.replicas.ctypes.data, len(self.models), C.byref(o),
                                          C.byref(self._h)))
        else:
            lv = np.ascontiguousarray(lanes, dtype=np.int32)
            if lv.shape != (len(self.models),):
                raise ValueError("lanes must give one value per model")
            _check(lib().ctl_batch_create_lanes(handles, self.replicas.ctypes.data, lv.ctypes.data,
                                                len(self.models), C.byref(o), C.byref(self._h)))

    def run(self) -> "Batch":
        _check(lib().ctl_batch_run(self._h))
        return self

    def samples(self, model: int) -> np.ndarray:
        i = self.info[model]
        out = np.empty(int(self.replicas[model]) * i.K * i.SS, np.uint32)
        _check(lib().ctl_batch_samples(self._h, model, out.ctypes.data, out.size))
        return out.reshape(int(self.replicas[model]), i.K, i.SS)

    def sim_stats(self, model: int) -> np.ndarray:
        out = np.zeros(int(self.replicas[model]), dtype=SIM_STAT_DTYPE)
        _check(lib().ctl_batch_sim_stats(self._h, model, out.ctypes.data, out.size))
        return out

    def stream_base(self, model: int) -> int:
        b = C.c_int64()
        _check(lib().ctl_batch_stream_base(self._h, model, C.byref(b)))
        return b.value

    def model_lanes(self, model: int) -> int:
        """Lanes per simulation the model runs with (its kernel class)."""
        g = C.c_int32()
        _check(lib().ctl_batch_model_lanes(self._h, model, C.byref(g)))
        return g.value

    def stats(self) -> dict:
        s = ctl_stats()
        _check(lib().ctl_batch_stats(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in ctl_stats._fields_}

    def close(self) -> None:
        if self._h:
            lib().ctl_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MLP:
    """BF16-MFMA multilayer perceptron on one GPU (C handle): ReLU hidden
    layers, linear output, FP32 master weights trained with Adam."""

    def __init__(self, widths, *, gpu=0, seed=1):
        self.widths = [int(w) for w in widths]
        w = np.ascontiguousarray(self.widths, dtype=np.int32)
        self._h = C.c_void_p()
        _check(lib().ctl_mlp_create(w.ctypes.data, len(w) - 1, gpu, seed, C.byref(self._h)))
        n = C.c_int64()
        _check(lib().ctl_mlp_num_params(self._h, C.byref(n)))
        self.nparam = n.value

    def get_params(self) -> np.ndarray:
        out = np.empty(self.nparam, np.float32)
        _check(lib().ctl_mlp_params(self._h, out.ctypes.data, out.size, 0))
        return out

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.nparam:
            raise ValueError(f"expected {self.nparam} parameters, got {p.size}")
        _check(lib().ctl_mlp_params(self._h, p.ctypes.data, p.size, 1))

    def forward(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, self.widths[0])
        y = np.empty((x.shape[0], self.widths[-1]), np.float32)
        _check(lib().ctl_mlp_forward(self._h, x.ctypes.data, x.shape[0], y.ctypes.data))
        return y

    def forward_device(self, x_ptr: int, rows: int, y_ptr: int) -> float:
        """Forward pass on device buffers (raw pointers); returns device ms."""
        ms = C.c_float()
        _check(lib().ctl_mlp_forward_device(self._h, C.c_void_p(x_ptr), rows, C.c_void_p(y_ptr), C.byref(ms)))
        return ms.value

    def gradient(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, self.widths[0])
        t = np.ascontiguousarray(t, dtype=np.float32).reshape(x.shape[0], self.widths[-1])
        g = np.empty(self.nparam, np.float32)
        loss = C.c_float()
        _check(lib().ctl_mlp_gradient(self._h, x.ctypes.data, t.ctypes.data, x.shape[0], g.ctypes.data, g.size,
                                      C.byref(loss)))
        return loss.value, g

    def set_data(self, x, t) -> None:
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, self.widths[0])
        t = np.ascontiguousarray(t, dtype=np.float32).reshape(x.shape[0], self.widths[-1])
        _check(lib().ctl_mlp_set_data(self._h, x.ctypes.data, t.ctypes.data, x.shape[0]))

    def train(self, epochs: int, batch: int, lr: float, seed: int = 0) -> np.ndarray:
        losses = np.empty(epochs, np.float32)
        _check(lib().ctl_mlp_train(self._h, epochs, batch, lr, seed, losses.ctypes.data))
        return losses

    def close(self) -> None:
        if self._h:
            lib().ctl_mlp_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
