"""Full-size checks at the benchmark shapes, through properties that do not
need the oracle (which cannot step 256^3 in test time):

* 256^3, 20,000 Voronoi grains, sigma 1.5, FP64 and FP32:
  - the interface mask equals the periodic face-neighbour criterion
    (src/gpk_geometry.cpp:42-82) recomputed from the labels;
  - dt equals 0.125 dx / max|v| over the band (src/gpk_simulation.cpp:11-25);
  - alpha stays in [1e-6, 1]; time is the sum of the dt used;
  - the grain count equals the number of distinct labels;
  - two runs from the same state give identical labels, alpha and time
    (no order dependence in the transaction kernels).
* 128^3 partitioned solidification (r = 6 seed): the FP64 solute sum is
  conserved to 1e-13 relative over 40 steps (src/gpk_transport.cpp:48-135).
"""
import numpy as np
import pytest

import gpk_amd
from gpk_amd import _abi
from tests.util import solidification_setup

pytestmark = pytest.mark.gpu

N = 256


def face_mask(phi):
    m = np.zeros(phi.shape, bool)
    for ax in range(3):
        for sh in (1, -1):
            m |= phi != np.roll(phi, sh, axis=ax)
    return m


def run(prec, steps):
    G = gpk_amd.GPKGrid(N, N, N, 1.0, _abi.isotropic(1.0, 1.0), 0.0, 1.5, precision=prec)
    G.voronoi_init(20000, 42)
    dts = []
    for _ in range(steps):
        dts.append(G.step())
    return G, dts


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_grain_growth_256_invariants(prec):
    print("256^3", prec, flush=True)
    G, dts = run(prec, 4)
    phi = G.get("phi")
    mask = G.get("mask").astype(bool)
    assert np.array_equal(mask, face_mask(phi))
    alpha = G.get("alpha")
    tol = 1e-6 if prec == "fp64" else np.float32(1e-6)
    assert alpha.min() >= tol and alpha.max() <= 1.0
    assert G.time == sum(dts)
    assert G.n_grains() == len(np.unique(phi))
    # dt of one more step from the velocities that step computes
    dt = G.step()
    vn = G.get("vn")
    band = G.get("mask").astype(bool)
    vmax = np.abs(vn).max()
    if prec == "fp64":
        assert dt == 0.125 * 1.0 / vmax
    else:
        assert dt == float(np.float32(np.float32(0.125) * np.float32(1.0) / np.float32(vmax)))
    assert G.last_stats.band_cells == int(face_mask(phi).sum())
    assert band.sum() > 0.3 * N ** 3
    G.close()


def test_grain_growth_256_repeatable():
    print("256^3 repeatability", flush=True)
    G1, _ = run("fp64", 3)
    G2, _ = run("fp64", 3)
    assert np.array_equal(G1.get("phi"), G2.get("phi"))
    assert np.array_equal(G1.get("alpha"), G2.get("alpha"))
    assert G1.time == G2.time
    G1.close()
    G2.close()


def test_solidification_128_mass_conservation():
    print("128^3 solidification", flush=True)
    model, phi, c, T = solidification_setup(128)
    G = gpk_amd.GPKGrid(128, 128, 128, 1e-6, model, 1e-9, 1.5, precision="fp64")
    G.set("phi", phi)
    G.set("c", c)
    G.set("T", T)
    M0 = G.solute_sum()
    for _ in range(40):
        G.step()
    assert abs(G.solute_sum() - M0) <= 1e-13 * M0
    assert G.solid_fraction() > (phi.sum() / phi.size) * 0.5
    G.close()


def test_fp64_labels_vs_strict32_128():
    # P2: plain FP64 against the reference's float labels.  GPU strict32 is
    # bit-exact with the compiled reference (oracle strict32 == reference,
    # GPU == oracle), so GPU fp64 against GPU strict32 counts how far FP64
    # labels move from the reference's at the 256^3 grain density
    # (128^3, 2,500 grains, sigma 1.5, 50 steps)
    n, grains = 128, 2500
    grids = {}
    for prec in ("strict32", "fp64"):
        print("128^3 P2", prec, flush=True)
        G = gpk_amd.GPKGrid(n, n, n, 1.0, _abi.isotropic(1.0, 1.0), 0.0, 1.5, precision=prec)
        G.voronoi_init(grains, 42)
        grids[prec] = G
    assert np.array_equal(grids["strict32"].get("phi"), grids["fp64"].get("phi"))
    diffs = {}
    for s in range(1, 51):
        for G in grids.values():
            G.step()
        if s in (1, 10, 25, 50):
            diffs[s] = int(np.count_nonzero(grids["strict32"].get("phi") != grids["fp64"].get("phi")))
    print("fp64 vs strict32 label differences by step:", diffs, flush=True)
    for G in grids.values():
        G.close()
    assert diffs[1] == 0 and diffs[10] == 0
    assert diffs[50] <= 1e-4 * n ** 3
