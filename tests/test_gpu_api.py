"""C ABI behaviour on the device: round trips, error codes, diagnostics,
the topology certificate and the device Voronoi generator."""
import ctypes as C

import numpy as np
import pytest

import gpk_amd
from gpk_amd import _abi, _lib
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_upload_download_round_trip_all_writable_fields():
    for prec, real in (("strict32", np.float32), ("fp64", np.float64)):
        G = gpk_amd.GPKGrid(4, 6, 5, 1.0, _abi.isotropic(1.0, 1.0), precision=prec)
        rng = np.random.default_rng(1)
        phi = rng.integers(0, 7, size=(4, 6, 5)).astype(np.int32)
        G.set("phi", phi)
        assert np.array_equal(G.get("phi"), phi)
        for f in ("alpha", "c", "T", "g", "rho"):
            v = rng.random((4, 6, 5)).astype(real)
            G.set(f, v)
            assert np.array_equal(G.get(f), v)
        # initial work fields are zero, as the reference constructs them
        for f in ("kappa", "nx", "vn"):
            assert not G.get(f).any()
        G.close()


def test_initial_values_match_reference_constructor():
    G = gpk_amd.GPKGrid(1, 8, 8, 1.0, _abi.isotropic(1.0, 1.0))
    assert not G.get("phi").any()
    assert np.all(G.get("alpha") == 1)
    assert np.all(G.get("T") == 800)
    assert G.time == 0.0


def test_label_range_is_checked():
    G = gpk_amd.GPKGrid(1, 8, 8, 1.0, _abi.isotropic(1.0, 1.0))
    bad = np.zeros((8, 8), np.int32)
    bad[0, 0] = -3
    with pytest.raises(gpk_amd.GPKError) as e:
        G.set("phi", bad)
    assert e.value.code == _abi.GPK_E_LABEL_RANGE


def test_rejected_labels_leave_the_grid_unchanged():
    # a rejected upload (negative or >= 2^26 labels) keeps the previous labels,
    # and the grid still steps exactly as the oracle from those labels
    phi, g = O.voronoi(1, 48, 48, 20, 3)
    m = _abi.isotropic(1.0, 1.0)
    G = gpk_amd.GPKGrid(1, 48, 48, 1.0, m, 0.0, 1.5, precision="fp64")
    o = O.OracleGrid(1, 48, 48, 1.0, m, 0.0, 1.5, "fp64")
    for X in (G, o):
        X.set("phi", phi)
        X.set("g", g)
    for badval in (-5, 1 << 26, (1 << 31) - 1):
        bad = phi.reshape(48, 48).copy()
        bad[7, 9] = badval
        with pytest.raises(gpk_amd.GPKError) as e:
            G.set("phi", bad)
        assert e.value.code == _abi.GPK_E_LABEL_RANGE
        assert np.array_equal(G.get("phi").ravel(), phi)
    for _ in range(5):
        assert G.step() == o.step()
    assert np.array_equal(G.get("phi"), o.get("phi"))
    assert G.n_grains() == o.n_grains()
    G.close()


def test_size_mismatch_is_an_error():
    L = _lib.lib()
    G = gpk_amd.GPKGrid(1, 8, 8, 1.0, _abi.isotropic(1.0, 1.0))
    buf = np.zeros(10, np.int32)
    assert L.gpk_upload(G._h, 0, buf.ctypes.data, buf.nbytes) == _abi.GPK_E_INVALID
    assert L.gpk_download(G._h, 99, buf.ctypes.data, buf.nbytes) == _abi.GPK_E_INVALID


def test_topology_simple_move_matches_oracle():
    phi, g = O.voronoi(1, 48, 48, 20, 3)
    m = _abi.isotropic(1.0, 1.0)
    o = O.OracleGrid(1, 48, 48, 1.0, m, 0.0, 1.5, "strict32")
    G = gpk_amd.GPKGrid(1, 48, 48, 1.0, m, 0.0, 1.5)
    o.set("phi", phi)
    G.set("phi", phi)
    rng = np.random.default_rng(0)
    p = phi.ravel()
    for cell in rng.integers(0, p.size, 150):
        for nl in {int(p[(cell + 1) % p.size]), int(p[(cell - 48) % p.size]), int(p[cell]), 999}:
            assert o.simple_move(int(cell), nl) == G.topology_simple_move(int(cell), nl), (cell, nl)


def test_topology_simple_move_large_components():
    """Connectivity searches that outgrow the LDS visited set (more than
    6144 cells) and finish in the device-memory search.  Label 0 fills full
    periodic rows; the cell c = (40, 50) of the first row is label 1.
    One row: label 1 minus c is one cylinder (the top and bottom branches of
    c meet around the torus), so c may take label 0.  Two rows: removing c
    splits label 1 into two bands, so it may not."""
    m = _abi.isotropic(1.0, 1.0)
    for rows, expect in (((40,), True), ((40, 80), False)):
        phi = np.ones((128, 128), np.int32)
        for r in rows:
            phi[r, :] = 0
        phi[40, 50] = 1
        o = O.OracleGrid(1, 128, 128, 1.0, m, 0.0, 1.5, "strict32")
        G = gpk_amd.GPKGrid(1, 128, 128, 1.0, m, 0.0, 1.5)
        o.set("phi", phi)
        G.set("phi", phi)
        cell = 40 * 128 + 50
        assert o.simple_move(cell, 0) == expect
        assert G.topology_simple_move(cell, 0) == expect, rows
        G.close()


# every case has cells outside their label's largest face component, so the
# device connectivity normalisation re-attaches cells in each of them
# (test_capi.test_voronoi_cases_have_detached_cells)
VORONOI_CASES = [((1, 96, 96), 200, 42), ((24, 24, 24), 60, 42), ((40, 36, 32), 300, 5),
                 ((1, 64, 64), 1500, 3), ((20, 20, 20), 900, 7), ((16, 24, 20), 700, 1)]


@pytest.mark.parametrize("dims,ng,seed", VORONOI_CASES)
def test_device_voronoi_matches_oracle(dims, ng, seed):
    G = gpk_amd.GPKGrid(*dims, 1.0, _abi.isotropic(1.0, 1.0))
    G.voronoi_init(ng, seed)
    phi, g = O.voronoi(*dims, ng, seed)
    assert np.array_equal(G.get("phi").ravel(), phi)
    assert np.array_equal(G.get("g").ravel(), g)
    assert np.all(G.get("alpha") == 1)


def test_diagnostics_match_oracle():
    phi, g = O.voronoi(1, 96, 96, 40, 42)
    m = _abi.isotropic(1.0, 1.0)
    o = O.OracleGrid(1, 96, 96, 1.0, m, 0.0, 1.5, "strict32")
    G = gpk_amd.GPKGrid(1, 96, 96, 1.0, m, 0.0, 1.5)
    for X in (o, G):
        X.set("phi", phi)
        X.set("g", g)
        for _ in range(30):
            X.step()
    assert G.n_grains() == o.n_grains()
    for lab in (1, 5, 17):
        assert G.phase_volume(lab) == pytest.approx(o.phase_volume(lab), rel=1e-12)
    mask = o.get("mask")
    assert G.interface_fraction() == np.float32(np.count_nonzero(mask)) / np.float32(mask.size)
    assert G.max_dt() == pytest.approx(o.max_dt(), rel=0)
