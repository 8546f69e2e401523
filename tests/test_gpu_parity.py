"""GPU parity: every precision mode of the HIP path against the same mode of
the oracle restatement, bit for bit, step by step (labels, alpha, g, c, mask,
phi_comp, normals, curvature, velocity, dt and the transaction counters).

strict32 is the reference's arithmetic (the oracle's strict32 mode is pinned
to the compiled reference and its fixtures by the CPU tests); fp32 and fp64
use the separable geometry order defined by the oracle's fp32/fp64 modes.
"""
import numpy as np
import pytest

import gpk_amd
from gpk_amd import _abi
from oracle import oracle as O
from tests.util import (FIELDS_ALL, dislocation_rho, elastic_solidification_setup, precipitation_setup,
                        rex_setup, solidification_setup)

pytestmark = pytest.mark.gpu


def lockstep(o, G, steps, fields=FIELDS_ALL, check_every=1):
    for s in range(steps):
        d1 = o.step()
        d2 = G.step()
        assert d1 == d2, ("dt", s, d1, d2)
        st = o.stats()
        gs = G.last_stats
        assert st["band"] == gs.band_cells, ("band", s)
        assert (st["candidates"], st["accepted"], st["deferred"], st["exchanges"], st["eroded"]) == \
            (gs.candidates, gs.accepted, gs.deferred, gs.exchanges, gs.eroded), ("counters", s)
        if s % check_every == 0 or s == steps - 1:
            for f in fields:
                a, b = o.get(f), G.get(f)
                assert np.array_equal(a, b), (f, s, int(np.count_nonzero(a != b)))
    assert o.time == G.time


def pair(dims, model, D, sigma, mode, init, dx=1.0):
    o = O.OracleGrid(*dims, dx, model, D, sigma, mode)
    G = gpk_amd.GPKGrid(*dims, dx, model, D, sigma, precision=mode)
    for name, arr in init.items():
        o.set(name, np.asarray(arr).astype(o.real) if name not in ("phi",) else arr)
        G.set(name, np.asarray(arr).astype(G.real) if name not in ("phi",) else arr)
    return o, G


def test_grain_growth_2d_256_strict32_100_steps():
    # 256^2, 200 Voronoi grains (seed 42), sigma 1.5, dt at the stability limit
    phi, g = O.voronoi(1, 256, 256, 200, 42)
    o, G = pair((1, 256, 256), _abi.isotropic(1.0, 1.0), 0.0, 1.5, "strict32", {"phi": phi, "g": g})
    lockstep(o, G, 100)


@pytest.mark.parametrize("mode", ["fp32", "fp64"])
def test_grain_growth_2d_256_fast_modes(mode):
    phi, g = O.voronoi(1, 256, 256, 200, 42)
    o, G = pair((1, 256, 256), _abi.isotropic(1.0, 1.0), 0.0, 1.5, mode, {"phi": phi, "g": g})
    lockstep(o, G, 60, check_every=5)


@pytest.mark.parametrize("mode", ["strict32", "fp32", "fp64"])
def test_grain_growth_3d(mode):
    # 48^3 at the 256^3 / 20000-grain density (~839 cells per grain)
    phi, g = O.voronoi(48, 48, 48, 132, 42)
    o, G = pair((48, 48, 48), _abi.isotropic(1.0, 1.0), 0.0, 1.5, mode, {"phi": phi, "g": g})
    lockstep(o, G, 40 if mode == "strict32" else 25, check_every=4)


@pytest.mark.parametrize("mode", ["strict32", "fp32", "fp64"])
def test_solidification_partitioned_3d(mode):
    model, phi, c, T = solidification_setup(32)
    o, G = pair((32, 32, 32), model, 1e-9, 1.5, mode, {"phi": phi, "c": c, "T": T}, dx=1e-6)
    lockstep(o, G, 40, check_every=4)
    assert o.solute_sum() == G.solute_sum()


def test_grain_growth_sigma2_3d_strict32():
    phi, g = O.voronoi(32, 32, 32, 40, 9)
    o, G = pair((32, 32, 32), _abi.isotropic(1.0, 1.0), 0.0, 2.0, "strict32", {"phi": phi, "g": g})
    lockstep(o, G, 24, check_every=3)


@pytest.mark.parametrize("mode", ["fp32", "fp64"])
@pytest.mark.parametrize("sigma", [1.0, 2.0, 2.6667])
def test_fast_modes_other_radii_3d(mode, sigma):
    # 3-D row-table geometry at window radii 6, 9 and 12 (sigma 1.5 is radius 8)
    phi, g = O.voronoi(32, 32, 32, 40, 11)
    o, G = pair((32, 32, 32), _abi.isotropic(1.0, 1.0), 0.0, sigma, mode, {"phi": phi, "g": g})
    lockstep(o, G, 24, check_every=4)


@pytest.mark.parametrize("mode", ["strict32", "fp32", "fp64"])
def test_broad_kernel_geometry_3d(mode):
    # sigma 5 cells: 3-D geometry radius 20 > 18, per-label separable
    # smoothing + fourth-order differences (src/gpk_geometry.cpp:185-390)
    phi, g = O.voronoi(32, 36, 40, 40, 3)
    o, G = pair((32, 36, 40), _abi.isotropic(1.0, 1.0), 0.0, 5.0, mode, {"phi": phi, "g": g})
    lockstep(o, G, 20, check_every=4)


@pytest.mark.parametrize("mode", ["strict32", "fp64"])
def test_broad_kernel_partial_boxes_3d(mode):
    # sigma 4.6 cells (radius 17, geometry radius 19 > 18) on 64 x 80 x 96 with
    # 30 grains: most labels' smoothing boxes are strictly smaller than the
    # grid along y and x and some wrap across the periodic faces, so the
    # box-restricted passes are exercised beyond the whole-grid case
    phi, g = O.voronoi(64, 80, 96, 30, 11)
    o, G = pair((64, 80, 96), _abi.isotropic(1.0, 1.0), 0.0, 4.6, mode, {"phi": phi, "g": g})
    lockstep(o, G, 6, check_every=2)


@pytest.mark.parametrize("mode", ["strict32", "fp64"])
def test_recrystallisation_2d(mode):
    o, G = pair((1, 80, 80), _abi.recrystallisation(1e-9, 0.5, 2e6), 0.0, 1.5, mode, {"phi": rex_setup()},
                dx=1e-6)
    lockstep(o, G, 60, check_every=5)


@pytest.mark.parametrize("mode", ["strict32", "fp32", "fp64"])
def test_dislocation_sweep_and_recovery_2d(mode):
    model = _abi.dislocation(1e-9, 0.5, 4e10, 2.5e-10, 0.5, recovery_rate=50.0, rho_rex=1e12)
    o, G = pair((1, 80, 80), model, 0.0, 1.5, mode, {"phi": rex_setup(), "rho": dislocation_rho()}, dx=1e-6)
    lockstep(o, G, 60, fields=FIELDS_ALL + ("rho",), check_every=5)


@pytest.mark.parametrize("mode", ["strict32", "fp64"])
def test_precipitation_phase_selection_2d(mode):
    model, init = precipitation_setup()
    o, G = pair((1, 80, 80), model, 2e-9, 1.5, mode, init, dx=0.5e-6)
    lockstep(o, G, 50, check_every=5)


@pytest.mark.parametrize("mode", ["strict32", "fp32", "fp64"])
def test_elastic_coupling_2d(mode):
    model, init = elastic_solidification_setup()
    o, G = pair((1, 80, 80), model, 1e-9, 1.5, mode, init, dx=1e-6)
    lockstep(o, G, 50, fields=FIELDS_ALL + ("E_el",), check_every=5)


def test_aniso_ostwald_lamella_2d_strict32():
    model = _abi.aniso_ostwald(1e-8, 0.5, 1.0, 4.0, 1.0, 0.1, 1e8)
    phi = np.zeros((96, 128), np.int32)
    phi[40:52, :] = 1
    phi[39, ::9] = 1
    phi[52, 4::11] = 1
    c = np.full(phi.shape, 0.1, np.float32)
    o, G = pair((1, 96, 128), model, 0.0, 1.5, "strict32", {"phi": phi, "c": c}, dx=1e-6)
    lockstep(o, G, 60, check_every=5)


def test_tensor_aniso_with_bias_2d_strict32():
    model = _abi.tensor_aniso(1e-8, 0.5, 1.0, 2.0, 1.0, dG_bias=2e4)
    y, x = np.mgrid[0:80, 0:80]
    phi = ((y - 40) ** 2 / 225.0 + (x - 40) ** 2 / 100.0 <= 1).astype(np.int32)
    o, G = pair((1, 80, 80), model, 0.0, 1.5, "strict32", {"phi": phi}, dx=1e-6)
    lockstep(o, G, 40, check_every=4)


def test_plain_diffusion_2d_strict32():
    model = _abi.solidification(M0=5e-8, sigma0=0.2, eps4=0.0)
    y, x = np.mgrid[0:64, 0:64]
    phi = ((y - 32) ** 2 + (x - 32) ** 2 <= 36).astype(np.int32)
    c = np.full((64, 64), 0.02, np.float32)
    T = np.full((64, 64), 903.0, np.float32)
    o, G = pair((1, 64, 64), model, 1e-9, 1.5, "strict32", {"phi": phi, "c": c, "T": T}, dx=1e-6)
    lockstep(o, G, 40, check_every=4)


def test_anisotropic_growth_2d_labels_and_tolerance():
    # the 2-D cos(4 theta) law calls cosf/atan2f: device and host libm may
    # differ in the last place, so velocities are compared to 1e-5 relative
    # and the labels, band and counters must still agree exactly
    phi, g = O.voronoi(1, 128, 128, 60, 42)
    o, G = pair((1, 128, 128), _abi.anisotropic_growth(1.0, 1.0, 0.05), 0.0, 1.5, "strict32",
                {"phi": phi, "g": g})
    for s in range(30):
        o.step()
        G.step()
        assert np.array_equal(o.get("phi"), G.get("phi")), s
        vo, vg = o.get("vn").astype(np.float64), G.get("vn").astype(np.float64)
        assert np.all(np.abs(vo - vg) <= 1e-5 * np.maximum(np.abs(vo).max(), 1e-30)), s


@pytest.mark.parametrize("mode", ["strict32", "fp64"])
def test_flip_searches_through_workgroup_and_device_memory(mode, monkeypatch):
    # wavefront searches hand over after 10 visited cells and workgroup
    # searches after 30, so on this 48^3 map (25 steps) about 550 searches go
    # to the workgroup search and about 50 on to the device-memory search
    # (and bounds searches that outgrow the workgroup wait for their exact
    # turn); labels and counters must still equal the oracle's
    monkeypatch.setenv("GPK_FLIP_SEARCH_CAPS", "10,30")
    phi, g = O.voronoi(48, 48, 48, 132, 42)
    o, G = pair((48, 48, 48), _abi.isotropic(1.0, 1.0), 0.0, 1.5, mode, {"phi": phi, "g": g})
    lockstep(o, G, 25, check_every=5)


@pytest.mark.parametrize("dims,mode", [((18, 22, 30), "strict32"), ((18, 22, 30), "fp64"), ((1, 54, 70), "fp32")])
def test_grain_growth_x_not_multiple_of_4(dims, mode):
    # Nx = 30 / 70: the interface mask takes its one-cell-per-thread form (the
    # four-cells-per-thread form needs Nx a multiple of 4)
    phi, g = O.voronoi(*dims, 24, 9)
    o, G = pair(dims, _abi.isotropic(1.0, 1.0), 0.0, 1.5, mode, {"phi": phi, "g": g})
    lockstep(o, G, 15, check_every=3)


@pytest.mark.parametrize("mode", ["fp32", "fp64"])
def test_fast_modes_generic_kernel_large_labels(mode):
    # labels above 2^16 take the generic (non-tiled) separable kernel; the
    # operation order is the same, so parity with the oracle is bit-exact
    phi, g = O.voronoi(20, 24, 28, 30, 4)
    phi = phi + 70000
    o, G = pair((20, 24, 28), _abi.isotropic(1.0, 1.0), 0.0, 1.5, mode, {"phi": phi, "g": g})
    lockstep(o, G, 20, check_every=4)


@pytest.mark.parametrize("sigma", [1.0, 2.6667])
def test_fast_mode_other_radii_2d(sigma):
    phi, g = O.voronoi(1, 128, 96, 40, 6)
    o, G = pair((1, 128, 96), _abi.isotropic(1.0, 1.0), 0.0, sigma, "fp64", {"phi": phi, "g": g})
    lockstep(o, G, 30, check_every=5)


def test_projection_generic_path_labels_above_2_15():
    # labels >= 2^15 take the untiled projection; same order, same bits
    phi, g = O.voronoi(1, 96, 96, 30, 8)
    phi = phi + 40000
    o, G = pair((1, 96, 96), _abi.isotropic(1.0, 1.0), 0.0, 1.5, "strict32", {"phi": phi, "g": g})
    lockstep(o, G, 30, check_every=5)


@pytest.mark.parametrize("mode", ["fp32", "fp64"])
def test_geometry_tile_hash_overflow_path(mode):
    # random labels: every cell is an interface cell and a geometry tile holds
    # far more distinct residents than its hash; those cells take the in-kernel
    # path that reads window rows from the grid (same arithmetic)
    rng = np.random.default_rng(3)
    phi = rng.integers(0, 3000, size=(1, 96, 96)).astype(np.int32).reshape(96, 96)
    o, G = pair((1, 96, 96), _abi.isotropic(1.0, 1.0), 0.0, 1.5, mode, {"phi": phi})
    lockstep(o, G, 2)


@pytest.mark.parametrize("mode", ["fp32", "fp64"])
def test_geometry_tile_hash_overflow_path_3d(mode):
    # the same in 3-D: a 16x8x16 tile of random labels holds about 1,700
    # residents, so most of its cells take the row-from-the-grid path, next to
    # residents that fit the hash and partial tiles at the 24-cell faces
    rng = np.random.default_rng(5)
    phi = rng.integers(0, 5000, size=(24, 24, 24)).astype(np.int32)
    o, G = pair((24, 24, 24), _abi.isotropic(1.0, 1.0), 0.0, 1.5, mode, {"phi": phi})
    lockstep(o, G, 2)
