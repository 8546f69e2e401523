"""The oracle restatement (strict32) against the compiled reference sources
(oracle/_ref/libgpkref.so), bit for bit, over several models (CPU).

Skipped only where oracle/_ref could not be built (no reference tree)."""
import numpy as np
import pytest

from gpk_amd import _abi
from oracle import oracle as O
from tests.util import (dislocation_rho, elastic_solidification_setup, precipitation_setup, rex_setup,
                        solidification_setup)

pytestmark = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")

FIELDS = ("phi", "alpha", "g", "c", "mask", "phi_comp", "kappa", "nx", "ny", "nz", "vn", "g_comp")


def run_pair(dims, model, D, sigma, steps, init, dx=1.0):
    nz, ny, nx = dims
    o = O.OracleGrid(nz, ny, nx, dx, model, D, sigma, "strict32")
    r = O.RefGrid(nz, ny, nx, dx, model, D, sigma)
    for name, arr in init.items():
        o.set(name, arr)
        r.set(name, arr)
    for s in range(steps):
        d1 = o.step()
        d2 = r.step()
        assert d1 == d2, ("dt", s)
        for f in FIELDS:
            assert np.array_equal(o.get(f), r.get(f)), (f, s)
        st = o.stats()
        rc = r.counters()
        assert (st["accepted"], st["deferred"], st["exchanges"]) == \
            (rc["accepted"], rc["deferred"], rc["exchanges"]), s
    assert o.time == r.time


def test_grain_growth_2d():
    phi, g = O.voronoi(1, 96, 96, 40, 42)
    run_pair((1, 96, 96), _abi.isotropic(1.0, 1.0), 0.0, 1.5, 60, {"phi": phi, "g": g})


def test_grain_growth_3d():
    phi, g = O.voronoi(24, 24, 24, 16, 5)
    run_pair((24, 24, 24), _abi.isotropic(1.0, 1.0), 0.0, 1.5, 25, {"phi": phi, "g": g})


def test_anisotropic_growth_2d_eps4():
    phi, g = O.voronoi(1, 64, 64, 20, 11)
    run_pair((1, 64, 64), _abi.anisotropic_growth(1.0, 1.0, 0.05), 0.0, 1.5, 40, {"phi": phi, "g": g})


def test_solidification_partitioned_3d():
    model, phi, c, T = solidification_setup(24, r=4)
    run_pair((24, 24, 24), model, 1e-9, 1.5, 30, {"phi": phi, "c": c, "T": T}, dx=1e-6)


def test_solidification_plain_diffusion_2d():
    model = _abi.solidification(M0=5e-8, sigma0=0.2, eps4=0.04)
    y, x = np.mgrid[0:64, 0:64]
    phi = ((y - 32) ** 2 + (x - 32) ** 2 <= 36).astype(np.int32)
    c = np.full((64, 64), 0.02, np.float32)
    T = np.full((64, 64), 933.0 - 10.0 - 20.0, np.float32)
    run_pair((1, 64, 64), model, 1e-9, 1.5, 40, {"phi": phi, "c": c, "T": T}, dx=1e-6)


def test_aniso_ostwald_lamella_2d():
    model = _abi.aniso_ostwald(1e-8, 0.5, 1.0, 4.0, 1.0, 0.1, 1e8)
    phi = np.zeros((64, 80), np.int32)
    phi[28:38, :] = 1
    phi[27, ::7] = 1
    c = np.full(phi.shape, 0.1, np.float32)
    run_pair((1, 64, 80), model, 0.0, 1.5, 40, {"phi": phi, "c": c}, dx=1e-6)


def test_tensor_aniso_with_bias_2d():
    model = _abi.tensor_aniso(1e-8, 0.5, 1.0, 2.0, 1.0, dG_bias=2e4)
    y, x = np.mgrid[0:64, 0:64]
    phi = ((y - 32) ** 2 / 144.0 + (x - 32) ** 2 / 64.0 <= 1).astype(np.int32)
    run_pair((1, 64, 64), model, 0.0, 1.5, 40, {"phi": phi}, dx=1e-6)


def test_topology_simple_move_matches_reference():
    phi, g = O.voronoi(1, 48, 48, 20, 3)
    m = _abi.isotropic(1.0, 1.0)
    o = O.OracleGrid(1, 48, 48, 1.0, m, 0.0, 1.5, "strict32")
    r = O.RefGrid(1, 48, 48, 1.0, m, 0.0, 1.5)
    o.set("phi", phi)
    r.set("phi", phi)
    rng = np.random.default_rng(0)
    p = phi.ravel()
    for cell in rng.integers(0, p.size, 400):
        for nl in {int(p[(cell + 1) % p.size]), int(p[(cell - 48) % p.size]), int(p[cell]), 999}:
            assert o.simple_move(int(cell), nl) == r.simple_move(int(cell), nl)


def test_broad_kernel_geometry_3d():
    # sigma 5 cells: geometry radius 20 > 18 selects the per-label separable
    # smoothing and fourth-order differences (src/gpk_geometry.cpp:185-390)
    phi, g = O.voronoi(32, 36, 40, 40, 3)
    run_pair((32, 36, 40), _abi.isotropic(1.0, 1.0), 0.0, 5.0, 20, {"phi": phi, "g": g})


def test_recrystallisation_2d():
    # RecrystallisationThermo: E_stored drives the deformed matrix (label 0)
    # into the recrystallised discs
    phi = rex_setup()
    run_pair((1, 80, 80), _abi.recrystallisation(1e-9, 0.5, 2e6), 0.0, 1.5, 60, {"phi": phi}, dx=1e-6)


def test_dislocation_sweep_and_recovery_2d():
    # DislocationThermo: rho coupling in the driving force, swept cells take
    # rho_rex, static recovery rho *= exp(-r dt) every step
    model = _abi.dislocation(1e-9, 0.5, 4e10, 2.5e-10, 0.5, recovery_rate=50.0, rho_rex=1e12)
    nz, ny, nx = 1, 80, 80
    init = {"phi": rex_setup(), "rho": dislocation_rho()}
    o = O.OracleGrid(nz, ny, nx, 1e-6, model, 0.0, 1.5, "strict32")
    r = O.RefGrid(nz, ny, nx, 1e-6, model, 0.0, 1.5)
    for name, arr in init.items():
        o.set(name, arr)
        r.set(name, arr)
    accepted = 0
    for s in range(60):
        assert o.step() == r.step(), s
        accepted += o.stats()["accepted"]
        for f in FIELDS + ("rho",):
            assert np.array_equal(o.get(f), r.get(f)), (f, s)
    assert accepted > 0
    rho = o.get("rho")
    assert rho.min() < 1e12 and rho.max() < 1.5e14   # swept and recovered


def test_precipitation_phase_selection_2d():
    # PrecipitationThermo: per-cell lowest-G product phase replaces the
    # geometric competitor (src/gpk_physics.cpp:79-95)
    model, init = precipitation_setup()
    run_pair((1, 80, 80), model, 2e-9, 1.5, 50, init, dx=0.5e-6)


def test_elastic_coupling_2d():
    # eigenstrain smoothing + energy coupling (src/gpk_physics.cpp:10-63, 122-127)
    model, init = elastic_solidification_setup()
    o = O.OracleGrid(1, 80, 80, 1e-6, model, 1e-9, 1.5, "strict32")
    r = O.RefGrid(1, 80, 80, 1e-6, model, 1e-9, 1.5)
    for name, arr in init.items():
        o.set(name, arr)
        r.set(name, arr)
    for s in range(50):
        assert o.step() == r.step(), s
        for f in FIELDS + ("E_el",):
            assert np.array_equal(o.get(f), r.get(f)), (f, s)
    assert o.get("E_el").max() > 0
