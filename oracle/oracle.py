"""ctypes access to the CPU checkers.  TEST INFRASTRUCTURE ONLY.

* OracleGrid — the restatement in oracle/gpk_oracle.cpp (liboracle.so).
* RefGrid    — the reference's own GPKGrid compiled from its sources
               (oracle/_ref/libgpkref.so, built by oracle/Makefile).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import
this module; the product package never does.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from gpk_amd import _abi

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libgpkref.so")
REF_CLI = os.path.join(HERE, "_ref", "gpk")

MODES = {"strict32": 0, "fp32": 1, "fp64": 2, "strict64": 3}


def build(ref_root="/root/reference"):
    """Compile the restatement and, when the reference tree is present, the
    reference libraries (make -C oracle)."""
    subprocess.check_call(["make", "-s", "-C", HERE, "all", "REF=" + ref_root])


class OgStats(C.Structure):
    _fields_ = [(n, C.c_int64) for n in (
        "band", "candidates", "accepted", "deferred", "exchanges", "eroded",
        "flip_global", "flip_depth", "flip_ready0", "bfs_visits",
        "bgr_proposals", "bgr_pairs", "bgr_depth", "guarded")]


_olib = None
_rlib = None


def oracle_lib():
    global _olib
    if _olib is None:
        if not os.path.exists(ORACLE_SO):
            build()
        L = C.CDLL(ORACLE_SO)
        vp = C.c_void_p
        L.og_create.restype = vp
        L.og_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float,
                                C.c_float, C.POINTER(_abi.ModelDesc)]
        L.og_destroy.argtypes = [vp]
        L.og_set.argtypes = [vp, C.c_int, vp, C.c_size_t]
        L.og_get.argtypes = [vp, C.c_int, vp, C.c_size_t]
        L.og_step.restype = C.c_double
        L.og_step.argtypes = [vp, C.c_double, C.c_double]
        L.og_stage.argtypes = [vp, C.c_int]
        L.og_max_dt.restype = C.c_double
        L.og_max_dt.argtypes = [vp]
        L.og_time.restype = C.c_double
        L.og_time.argtypes = [vp]
        L.og_set_time.argtypes = [vp, C.c_double]
        L.og_step_id.restype = C.c_uint64
        L.og_step_id.argtypes = [vp]
        L.og_real_size.argtypes = [vp]
        L.og_simple_move.argtypes = [vp, C.c_int, C.c_int32]
        L.og_phase_volume.restype = C.c_double
        L.og_phase_volume.argtypes = [vp, C.c_int32]
        L.og_n_grains.argtypes = [vp]
        L.og_solute_sum.restype = C.c_double
        L.og_solute_sum.argtypes = [vp]
        L.og_voronoi_detached.restype = C.c_int64
        L.og_voronoi_detached.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint32]
        L.og_stats_get.argtypes = [vp, C.POINTER(OgStats)]
        L.og_voronoi.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint32, vp, vp]
        L.og_gauss.argtypes = [C.c_float, vp, C.c_int, C.POINTER(C.c_int), vp, vp, vp,
                               C.POINTER(C.c_int)]
        _olib = L
    return _olib


def ref_available():
    return os.path.exists(REF_SO)


def ref_lib():
    global _rlib
    if _rlib is None:
        L = C.CDLL(REF_SO)
        vp = C.c_void_p
        L.rf_create.restype = vp
        L.rf_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float,
                                C.POINTER(_abi.ModelDesc)]
        L.rf_destroy.argtypes = [vp]
        L.rf_set.argtypes = [vp, C.c_int, vp, C.c_size_t]
        L.rf_get.argtypes = [vp, C.c_int, vp, C.c_size_t]
        L.rf_step.restype = C.c_double
        L.rf_step.argtypes = [vp, C.c_float, C.c_float]
        L.rf_run.argtypes = [vp, C.c_double, C.c_float]
        L.rf_time.restype = C.c_double
        L.rf_time.argtypes = [vp]
        L.rf_set_time.argtypes = [vp, C.c_double]
        L.rf_counters.argtypes = [vp, C.POINTER(C.c_int64)]
        L.rf_max_dt.restype = C.c_double
        L.rf_max_dt.argtypes = [vp]
        L.rf_phase_volume.restype = C.c_double
        L.rf_phase_volume.argtypes = [vp, C.c_int32]
        L.rf_n_grains.argtypes = [vp]
        L.rf_simple_move.argtypes = [vp, C.c_int, C.c_int32]
        _rlib = L
    return _rlib


def _dtype(kind, real):
    return {"i32": np.int32, "u8": np.uint8, "real": real}[kind]


class _GridBase:
    real = np.float32

    def __init__(self, Nz, Ny, Nx):
        self.Nz, self.Ny, self.Nx = Nz, Ny, Nx
        self.N = Nz * Ny * Nx

    def shape(self):
        return (self.Nz, self.Ny, self.Nx) if self.Nz > 1 else (self.Ny, self.Nx)

    def set(self, name, arr):
        fid, kind = _abi.FIELDS[name]
        a = np.ascontiguousarray(np.asarray(arr).reshape(-1), dtype=_dtype(kind, self.real))
        assert a.size == self.N
        rc = self._set(fid, a.ctypes.data, a.nbytes)
        if rc != 0:
            raise ValueError("set %s failed (%d)" % (name, rc))

    def get(self, name):
        fid, kind = _abi.FIELDS[name]
        a = np.empty(self.N, dtype=_dtype(kind, self.real))
        rc = self._get(fid, a.ctypes.data, a.nbytes)
        if rc != 0:
            raise ValueError("get %s failed (%d)" % (name, rc))
        return a.reshape(self.shape())


class OracleGrid(_GridBase):
    """The restatement; mode in {"strict32", "fp32", "fp64", "strict64"}."""

    def __init__(self, Nz, Ny, Nx, dx, model, D=0.0, sigma=2.0, mode="strict32"):
        super().__init__(Nz, Ny, Nx)
        self.L = oracle_lib()
        self.mode = mode
        self.model = model
        self.h = self.L.og_create(MODES[mode], Nz, Ny, Nx, dx, D, sigma, C.byref(model))
        if not self.h:
            raise ValueError("og_create failed")
        self.real = np.float32 if self.L.og_real_size(self.h) == 4 else np.float64

    def __del__(self):
        if getattr(self, "h", None):
            self.L.og_destroy(self.h)
            self.h = None

    def _set(self, f, p, b):
        return self.L.og_set(self.h, f, p, b)

    def _get(self, f, p, b):
        return self.L.og_get(self.h, f, p, b)

    def step(self, dt=-1.0, cap=1e30):
        return self.L.og_step(self.h, dt, cap)

    def stage(self, which):
        self.L.og_stage(self.h, which)

    def max_dt(self):
        return self.L.og_max_dt(self.h)

    @property
    def time(self):
        return self.L.og_time(self.h)

    @time.setter
    def time(self, t):
        self.L.og_set_time(self.h, t)

    @property
    def step_id(self):
        return self.L.og_step_id(self.h)

    def stats(self):
        s = OgStats()
        self.L.og_stats_get(self.h, C.byref(s))
        return {n: getattr(s, n) for n, _ in OgStats._fields_}

    def simple_move(self, cell, new_label):
        return bool(self.L.og_simple_move(self.h, cell, new_label))

    def phase_volume(self, label):
        return self.L.og_phase_volume(self.h, label)

    def n_grains(self):
        return self.L.og_n_grains(self.h)

    def solute_sum(self):
        return self.L.og_solute_sum(self.h)


class RefGrid(_GridBase):
    """The compiled reference GPKGrid (float)."""

    def __init__(self, Nz, Ny, Nx, dx, model, D=0.0, sigma=2.0):
        super().__init__(Nz, Ny, Nx)
        self.L = ref_lib()
        self.h = self.L.rf_create(Nz, Ny, Nx, dx, D, sigma, C.byref(model))
        if not self.h:
            raise ValueError("rf_create failed")

    def __del__(self):
        if getattr(self, "h", None):
            self.L.rf_destroy(self.h)
            self.h = None

    def _set(self, f, p, b):
        return self.L.rf_set(self.h, f, p, b)

    def _get(self, f, p, b):
        return self.L.rf_get(self.h, f, p, b)

    def step(self, dt=-1.0, cap=1e30):
        return self.L.rf_step(self.h, dt, cap)

    def run(self, t_end, dt=-1.0):
        return self.L.rf_run(self.h, t_end, dt)

    @property
    def time(self):
        return self.L.rf_time(self.h)

    @time.setter
    def time(self, t):
        self.L.rf_set_time(self.h, t)

    def counters(self):
        a = (C.c_int64 * 3)()
        self.L.rf_counters(self.h, a)
        return {"accepted": a[0], "deferred": a[1], "exchanges": a[2]}

    def max_dt(self):
        return self.L.rf_max_dt(self.h)

    def phase_volume(self, label):
        return self.L.rf_phase_volume(self.h, label)

    def n_grains(self):
        return self.L.rf_n_grains(self.h)

    def simple_move(self, cell, new_label):
        return bool(self.L.rf_simple_move(self.h, cell, new_label))


def voronoi(Nz, Ny, Nx, n_seeds, seed):
    """Brute-force restatement of voronoi_init (demos/gpk.cpp:51-189)."""
    L = oracle_lib()
    N = Nz * Ny * Nx
    phi = np.empty(N, np.int32)
    g = np.empty(N, np.float32)
    rc = L.og_voronoi(Nz, Ny, Nx, n_seeds, seed, phi.ctypes.data, g.ctypes.data)
    if rc != 0:
        raise ValueError("voronoi failed (%d)" % rc)
    return phi, g


def gauss(sigma):
    """GaussKernel weights and the fused geometry kernels (float)."""
    L = oracle_lib()
    cap = 256
    w = np.zeros(cap, np.float32)
    w0 = np.zeros(cap, np.float32)
    w1 = np.zeros(cap, np.float32)
    w2 = np.zeros(cap, np.float32)
    r = C.c_int()
    rg = C.c_int()
    rc = L.og_gauss(sigma, w.ctypes.data, cap, C.byref(r), w0.ctypes.data, w1.ctypes.data,
                    w2.ctypes.data, C.byref(rg))
    if rc != 0:
        raise ValueError("gauss failed")
    n = 2 * r.value + 1
    m = 2 * rg.value + 1
    return w[:n].copy(), r.value, w0[:m].copy(), w1[:m].copy(), w2[:m].copy(), rg.value


def voronoi_detached(Nz, Ny, Nx, n_seeds, seed):
    """Cells of the nearest-site map outside their label's largest face
    component (the cells voronoi_init re-attaches)."""
    return int(oracle_lib().og_voronoi_detached(Nz, Ny, Nx, n_seeds, seed))
