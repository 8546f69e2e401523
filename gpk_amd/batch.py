"""Batched 2-D instances (gpk_batch_* in include/gpk_c.h).

A parameter sweep of small 2-D grids -- the reference runs such sweeps one
GPKGrid after another (demos/gpk.cpp:1121-1259) -- held packed on one GPU and
advanced by one persistent kernel per step.  Instance b is an Ny[b] x Nx[b]
grid; all share dx, sigma, precision and the model.  Fields are exchanged as
one list of per-instance (Ny, Nx) arrays.
"""
import ctypes as C

import numpy as np

from . import _abi
from ._lib import GPKError, check, lib

_KIND = {"i32": np.int32, "u8": np.uint8}


class GPKBatch:
    def __init__(self, shapes, dx, thermo, D=0.0, gaussian_sigma=2.0, precision="fp32", device=0):
        self.shapes = [(int(ny), int(nx)) for ny, nx in shapes]
        self.n = len(self.shapes)
        self.sizes = np.array([ny * nx for ny, nx in self.shapes], np.int64)
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)])
        self.Ntot = int(self.offsets[-1])
        self.precision = precision
        self.real = np.float64 if precision == "fp64" else np.float32
        Ny = (C.c_int * self.n)(*[s[0] for s in self.shapes])
        Nx = (C.c_int * self.n)(*[s[1] for s in self.shapes])
        h = C.c_void_p()
        check(lib().gpk_batch_create(self.n, Ny, Nx, dx, D, gaussian_sigma, _abi.PRECISIONS[precision],
                                     device, C.byref(thermo), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().gpk_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _dtype(self, kind):
        return _KIND.get(kind, self.real)

    def set(self, name, arrays):
        fid, kind = _abi.FIELDS[name]
        dt = self._dtype(kind)
        if len(arrays) != self.n:
            raise GPKError(_abi.GPK_E_INVALID, "need one array per instance")
        flat = np.concatenate([np.ascontiguousarray(np.broadcast_to(np.asarray(a, dt), s)).ravel()
                               for a, s in zip(arrays, self.shapes)])
        check(lib().gpk_batch_upload(self._h, fid, flat.ctypes.data, flat.nbytes))

    def get(self, name):
        fid, kind = _abi.FIELDS[name]
        out = np.empty(self.Ntot, self._dtype(kind))
        check(lib().gpk_batch_download(self._h, fid, out.ctypes.data, out.nbytes))
        return [out[self.offsets[b]:self.offsets[b + 1]].reshape(self.shapes[b]) for b in range(self.n)]

    def step(self, nsteps=1, dt=-1.0, dt_cap=1.0e30):
        check(lib().gpk_batch_step(self._h, int(nsteps), dt, dt_cap))

    def stats(self):
        s = _abi.BatchStats()
        check(lib().gpk_batch_stats_get(self._h, C.byref(s)))
        return {n: getattr(s, n) for n, _ in _abi.BatchStats._fields_}

    def instances(self):
        """(time, last step ms, band cells) per instance."""
        t = np.empty(self.n, np.float64)
        ms = np.empty(self.n, np.float64)
        nb = np.empty(self.n, np.int32)
        check(lib().gpk_batch_instances(self._h, t.ctypes.data, ms.ctypes.data, nb.ctypes.data))
        return t, ms, nb

    def autotune(self):
        w, n, ms = C.c_int(), C.c_int(), C.c_double()
        check(lib().gpk_batch_autotune(self._h, C.byref(w), C.byref(n), C.byref(ms)))
        return w.value, n.value, ms.value

    def set_launch(self, workgroup_size, workgroups):
        check(lib().gpk_batch_set_launch(self._h, int(workgroup_size), int(workgroups)))


def lamellar_instance(side, width, period, rng, p_perturb=0.3):
    """side x side labels: stripes of label 1, `width` cells wide every `period`
    cells along x in label 0; each stripe edge of each row moves out by one
    cell with probability p_perturb (seeded Bernoulli)."""
    x = np.arange(side)
    row = ((x % period) < width).astype(np.int32)
    phi = np.tile(row, (side, 1))
    left = np.flatnonzero((row == 1) & (np.roll(row, 1) == 0))
    right = np.flatnonzero((row == 1) & (np.roll(row, -1) == 0))
    for y in range(side):
        for e in left:
            if rng.random() < p_perturb:
                phi[y, (e - 1) % side] = 1
        for e in right:
            if rng.random() < p_perturb:
                phi[y, (e + 1) % side] = 1
    return phi


def lamellar_sweep(n, seed=0, side_range=(64, 384), width_range=(6, 20), ratio_range=(2.0, 12.0)):
    """Instance shapes and labels of the lamellar sweep: side ~ U_int[64, 384],
    width t ~ U_int[6, 20], period ratio r ~ U[2, 12], period round(r t)."""
    rng = np.random.default_rng(seed)
    shapes, phis, params = [], [], []
    for _ in range(n):
        s = int(rng.integers(side_range[0], side_range[1] + 1))
        t = int(rng.integers(width_range[0], width_range[1] + 1))
        r = float(rng.uniform(*ratio_range))
        lam = max(t + 2, int(round(r * t)))
        shapes.append((s, s))
        phis.append(lamellar_instance(s, t, lam, rng))
        params.append((s, t, lam))
    return shapes, phis, params
