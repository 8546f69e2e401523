"""gpk_velocity_adjoint_noalloc against gpk_velocity_adjoint, through the C ABI.

The allocation-free entry point runs the same kernels with its scratch in a
static device array, so its gradient must equal the allocating entry
point's bit for bit.  Two calls in a row on one GPU, the second on a smaller
grid (a smaller band), then a band beyond the scratch's capacity, which is
refused with GPK_E_INVALID before any kernel reads the scratch.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import gpk_amd
from gpk_amd import _abi
from gpk_amd._lib import lib

pytestmark = pytest.mark.gpu

THETA = (0.03, 0.01, 1.0)


def grid(phi, prec):
    nz, ny, nx = phi.shape
    G = gpk_amd.GPKGrid(nz, ny, nx, 1.0, _abi.anisotropic_growth(1.0, 1.0, THETA[0], THETA[1]),
                        0.0, 1.5, precision=prec)
    G.set("phi", phi)
    G.set("g", np.zeros(phi.shape, G.real))
    return G


def sphere(n, r):
    z, y, x = np.mgrid[0:n, 0:n, 0:n]
    c = n // 2
    return (((z - c) ** 2 + (y - c) ** 2 + (x - c) ** 2) <= r * r).astype(np.int32)


def disc(n, r):
    y, x = np.mgrid[0:n, 0:n]
    return (((y - n // 2) ** 2 + (x - n // 2) ** 2) <= r * r).astype(np.int32)[None]


def adjoint(G, gvel, noalloc):
    th = (C.c_double * 3)(*THETA)
    out = (C.c_double * 3)()
    fn = lib().gpk_velocity_adjoint_noalloc if noalloc else lib().gpk_velocity_adjoint
    rc = fn(G._h, th, C.c_void_p(gvel.data_ptr()), out)
    return rc, (out[0], out[1], out[2])


def gvel_for(G, seed):
    dt = torch.float64 if G.real == np.float64 else torch.float32
    g = torch.Generator(device="cpu").manual_seed(seed)
    return torch.randn(G.N, generator=g, dtype=dt).to("cuda:0")


def test_noalloc_matches_allocating_entry_point_twice():
    cases = [(sphere(96, 30), "fp64", 1), (disc(256, 50), "fp32", 2)]   # second: smaller band
    bands = []
    for phi, prec, seed in cases:
        print("noalloc adjoint", phi.shape, prec, flush=True)
        G = grid(phi, prec)
        gv = gvel_for(G, seed)
        rc_n, g_n = adjoint(G, gv, True)
        assert rc_n == 0, lib().gpk_last_error()
        bands.append(int(G.get("mask").astype(bool).sum()))   # band of the velocity chain
        rc_a, g_a = adjoint(G, gv, False)
        assert rc_a == 0, lib().gpk_last_error()
        assert g_n == g_a, (g_n, g_a)
        assert all(np.isfinite(g_n)) and any(v != 0.0 for v in g_n)
        # and once more from the same handle: the static scratch keeps the
        # last call's contents, which the kernels overwrite before reading
        rc_n2, g_n2 = adjoint(G, gv, True)
        assert rc_n2 == 0 and g_n2 == g_a
        G.close()
    assert bands[1] < bands[0]


def test_noalloc_refuses_a_band_beyond_its_scratch():
    # 176^3 Voronoi map at the 256^3 / 20,000-grain density: about 2.3 million
    # band cells, more than the 16 MiB scratch holds (about 2.09 million)
    n = 176
    print("noalloc adjoint oversize", n, flush=True)
    G = gpk_amd.GPKGrid(n, n, n, 1.0, _abi.anisotropic_growth(1.0, 1.0, THETA[0], THETA[1]), 0.0, 1.5,
                        precision="fp32")
    G.voronoi_init(6500, 42)
    gv = gvel_for(G, 3)
    rc, _ = adjoint(G, gv, True)
    assert rc == _abi.GPK_E_INVALID
    assert b"scratch" in lib().gpk_last_error()
    G.close()
    # the scratch still serves a grid that fits
    G = grid(disc(128, 30), "fp64")
    gv = gvel_for(G, 4)
    rc_n, g_n = adjoint(G, gv, True)
    rc_a, g_a = adjoint(G, gv, False)
    assert rc_n == 0 and rc_a == 0 and g_n == g_a
    G.close()
