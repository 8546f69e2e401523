"""The oracle against the reference's own golden data (CPU).

* T1: weighted-mean curvature of digitised circles (9 significant digits),
  paper/data/validate_theory_foundations/t1_curvature_error.csv.
* G13-G15: 360^2 curved bicrystals after 800 steps (flips, balanced remap and
  extinction all active) and their (time, area) trajectories,
  paper/data/validate_extended_base_models/g1{3,4,5}_*.csv.
* Voronoi initial labels written by the reference CLI (demos/gpk.cpp:51-189).
"""
import glob
import os

import numpy as np
import pytest

from gpk_amd import _abi
from oracle import oracle as O
from tests.util import GOLDEN, bicrystal_shape, circle, load_bicrystal, weighted_kappa


def test_t1_curvature_matches_reference_csv():
    z = np.load(os.path.join(GOLDEN, "t1_curvature.npz"))
    m = _abi.isotropic(1.0, 1.0)
    for R, text in zip(z["R"], z["kappa_text"]):
        N = 4 * int(R) + 24
        sigma = float(np.float32(2.0) * np.float32(R) / np.float32(15.0))
        o = O.OracleGrid(1, N, N, 1.0, m, 0.0, sigma, "strict32")
        o.set("phi", circle(N, int(R)))
        o.step(0.0)
        k = weighted_kappa(o.get("mask"), o.get("phi"), o.get("nx"), o.get("ny"), o.get("kappa"))
        assert "%.9g" % k == str(text), (R, k, text)


@pytest.mark.parametrize("key", ["g13", "g14", "g15"])
def test_curved_bicrystal_matches_reference_csv(key):
    before, after, t_text, a_text = load_bicrystal(key)
    phi0 = bicrystal_shape(key)
    assert np.array_equal(phi0, before)
    dx = float(np.float32(0.1e-6))
    o = O.OracleGrid(1, 360, 360, dx, _abi.isotropic(1e-8, 0.324), 0.0, 2.0, "strict32")
    o.set("phi", phi0)
    ts, areas = [], []
    for snap in range(11):
        if snap > 0:
            for _ in range(80):
                o.step()
        v = o.phase_volume(1)
        if v <= 1e-12:
            break
        ts.append("%.8e" % o.time)
        areas.append("%.8e" % (v * dx * dx))
    assert np.array_equal(o.get("phi"), after)
    assert ts == [str(t) for t in t_text]
    assert areas == [str(a) for a in a_text]


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "voronoi_*.npz"))))
def test_voronoi_matches_reference_cli(path):
    z = np.load(path)
    nz, ny, nx, ng, seed = (int(v) for v in z["dims"])
    phi, g = O.voronoi(nz, ny, nx, ng, seed)
    assert np.array_equal(phi.reshape(nz, ny, nx), z["phi"])
    assert [("%.6f" % v) for v in g] == [str(t) for t in z["g_text"]]
