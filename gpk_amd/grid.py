"""Python mirror of the reference ``gpk::GPKGrid`` API over the C ABI.

``GPKGrid(Nz, Ny, Nx, dx, thermo, D=0, gaussian_sigma=2)`` follows
include/gpk_core.hpp:237-264: the same constructor arguments, ``step``,
``run`` and diagnostics.  State lives on the GPU; the public fields
(``phi``, ``alpha``, ``c``, ``T``, ``g``, ``kappa_f`` ...) are numpy views
downloaded on access and uploaded on assignment.
"""
import ctypes as C

import numpy as np

from . import _abi
from ._lib import GPKError, check, lib

_FIELD_ALIASES = {
    "phi": "phi", "alpha": "alpha", "c": "c", "T": "T", "g": "g", "rho": "rho",
    "nx_f": "nx", "ny_f": "ny", "nz_f": "nz", "kappa_f": "kappa", "vn_f": "vn",
    "phi_comp": "phi_comp", "g_comp": "g_comp", "mask": "mask",
}


class GPKGrid:
    """One simulation domain on one GPU (Nz == 1 selects 2-D)."""

    def __init__(self, Nz, Ny, Nx, dx, thermo, D=0.0, gaussian_sigma=2.0,
                 precision="strict32", device=0):
        L = lib()
        self.Nz, self.Ny, self.Nx = int(Nz), int(Ny), int(Nx)
        self.dx = float(dx)
        self.N = self.Nz * self.Ny * self.Nx
        self.precision = precision
        self.device = int(device)
        self.real = np.float64 if precision == "fp64" else np.float32
        self._model = thermo
        gd = _abi.GridDesc()
        gd.Nz, gd.Ny, gd.Nx = self.Nz, self.Ny, self.Nx
        gd.dx, gd.D, gd.gaussian_sigma = dx, D, gaussian_sigma
        gd.precision = _abi.PRECISIONS[precision]
        gd.device = device
        h = C.c_void_p()
        check(L.gpk_create(C.byref(gd), C.byref(thermo), C.byref(h)))
        self._h = h
        self.topology_accepted_last = 0
        self.topology_deferred_last = 0
        self.gaussian_remap_exchanges_last = 0
        self.last_stats = None

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().gpk_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- fields ----
    def _shape(self):
        return (self.Nz, self.Ny, self.Nx) if self.Nz > 1 else (self.Ny, self.Nx)

    def _dtype(self, kind):
        return {"i32": np.int32, "u8": np.uint8, "real": self.real}[kind]

    @property
    def shape(self):
        return self._shape()

    def get(self, name):
        fid, kind = _abi.FIELDS[_FIELD_ALIASES.get(name, name)]
        out = np.empty(self.N, dtype=self._dtype(kind))
        check(lib().gpk_download(self._h, fid, out.ctypes.data, out.nbytes))
        return out.reshape(self._shape())

    def set(self, name, value):
        fid, kind = _abi.FIELDS[_FIELD_ALIASES.get(name, name)]
        arr = np.broadcast_to(np.asarray(value, dtype=self._dtype(kind)), (self.N,)
                              if np.ndim(value) <= 1 else np.shape(value))
        arr = np.ascontiguousarray(arr.reshape(-1), dtype=self._dtype(kind))
        if arr.size != self.N:
            raise GPKError(_abi.GPK_E_INVALID, "field %s needs %d values" % (name, self.N))
        check(lib().gpk_upload(self._h, fid, arr.ctypes.data, arr.nbytes))

    def __getattr__(self, name):
        if name in _FIELD_ALIASES:
            return self.get(name)
        raise AttributeError(name)

    def __setattr__(self, name, value):
        if name in _FIELD_ALIASES and "_h" in self.__dict__:
            self.set(name, value)
        else:
            object.__setattr__(self, name, value)

    def upload_device(self, name, ptr, nbytes):
        fid, _ = _abi.FIELDS[_FIELD_ALIASES.get(name, name)]
        check(lib().gpk_upload_device(self._h, fid, C.c_void_p(ptr), nbytes))

    def download_device(self, name, ptr, nbytes):
        fid, _ = _abi.FIELDS[_FIELD_ALIASES.get(name, name)]
        check(lib().gpk_download_device(self._h, fid, C.c_void_p(ptr), nbytes))

    # ---- stepping (src/gpk_simulation.cpp:31-72) ----
    def step(self, dt=-1.0, dt_cap=1.0e30):
        used = C.c_double()
        st = _abi.StepStats()
        check(lib().gpk_step(self._h, dt, dt_cap, C.byref(used), C.byref(st)))
        self.topology_accepted_last = st.accepted
        self.topology_deferred_last = st.deferred
        self.gaussian_remap_exchanges_last = st.exchanges
        self.last_stats = st
        return used.value

    def step_many(self, nsteps, dt=-1.0, dt_cap=1.0e30):
        st = _abi.StepStats()
        check(lib().gpk_step_many(self._h, int(nsteps), dt, dt_cap, C.byref(st)))
        return st

    def run(self, t_end, dt=-1.0, cb=None, cb_interval=100):
        steps = 0
        while self.time < t_end:
            remaining = float(np.float32(t_end - self.time))
            self.step(dt, remaining)
            steps += 1
            if cb is not None and steps % cb_interval == 0:
                cb(self, steps)
        return steps

    def set_model(self, thermo):
        check(lib().gpk_set_model(self._h, C.byref(thermo)))
        self._model = thermo

    @property
    def time(self):
        t = C.c_double()
        check(lib().gpk_get_time(self._h, C.byref(t)))
        return t.value

    @time.setter
    def time(self, t):
        check(lib().gpk_set_time(self._h, float(t)))

    # ---- diagnostics ----
    def max_dt(self):
        v = C.c_double()
        check(lib().gpk_max_dt(self._h, C.byref(v)))
        return v.value

    def phase_volume(self, label):
        v = C.c_double()
        check(lib().gpk_phase_volume(self._h, int(label), C.byref(v)))
        return v.value

    def n_grains(self):
        v = C.c_int()
        check(lib().gpk_n_grains(self._h, C.byref(v)))
        return v.value

    def mean_grain_radius(self):
        v = C.c_double()
        check(lib().gpk_mean_grain_radius(self._h, C.byref(v)))
        return v.value

    def solid_fraction(self):
        v = C.c_float()
        check(lib().gpk_solid_fraction(self._h, C.byref(v)))
        return v.value

    def interface_fraction(self):
        v = C.c_float()
        check(lib().gpk_interface_fraction(self._h, C.byref(v)))
        return v.value

    def solute_sum(self):
        v = C.c_double()
        check(lib().gpk_solute_sum(self._h, C.byref(v)))
        return v.value

    def velocity_field_ptr(self, theta, vout_ptr):
        """Projected velocity for theta = (eps4, eps6, M0) into a device buffer
        of N reals (gpk_velocity_field)."""
        th = (C.c_double * 3)(*[float(t) for t in theta])
        check(lib().gpk_velocity_field(self._h, th, C.c_void_p(vout_ptr)))

    def velocity_adjoint_ptr(self, theta, gvel_ptr):
        """dL/dtheta from dL/dv in a device buffer of N reals (gpk_velocity_adjoint)."""
        th = (C.c_double * 3)(*[float(t) for t in theta])
        out = (C.c_double * 3)()
        check(lib().gpk_velocity_adjoint(self._h, th, C.c_void_p(gvel_ptr), out))
        return (out[0], out[1], out[2])

    def device_bytes(self):
        v = C.c_size_t()
        check(lib().gpk_device_bytes(self._h, C.byref(v)))
        return v.value

    def topology_simple_move(self, cell, new_label):
        v = C.c_int()
        check(lib().gpk_topology_simple_move(self._h, int(cell), int(new_label), C.byref(v)))
        return bool(v.value)

    def voronoi_init(self, n_seeds, seed=42):
        check(lib().gpk_voronoi_init(self._h, int(n_seeds), int(seed)))

    def sync(self):
        check(lib().gpk_sync(self._h))

    def set_profiling(self, on=True):
        check(lib().gpk_set_profiling(self._h, 1 if on else 0))

    def kernel_timing(self, which=0):
        ms, n, b = C.c_double(), C.c_int64(), C.c_double()
        check(lib().gpk_kernel_timing(self._h, which, C.byref(ms), C.byref(n), C.byref(b)))
        return ms.value, n.value, b.value

    def idx(self, z, y, x):
        return (z * self.Ny + y) * self.Nx + x


def voronoi_host(Nz, Ny, Nx, n_seeds, seed=42):
    """Host Voronoi generator (same labels as voronoi_init, demos/gpk.cpp:51-189)."""
    N = Nz * Ny * Nx
    phi = np.empty(N, np.int32)
    g = np.empty(N, np.float32)
    check(lib().gpk_voronoi_host(Nz, Ny, Nx, int(n_seeds), int(seed), phi.ctypes.data,
                                 g.ctypes.data))
    return phi, g
