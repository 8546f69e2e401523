"""The HIP path (strict32) against the reference's own golden data: T1
curvature of digitised circles and the G13-G15 curved bicrystals after 800
steps, with their (time, area) trajectories."""
import os

import numpy as np
import pytest

import gpk_amd
from gpk_amd import _abi
from tests.util import GOLDEN, bicrystal_shape, circle, load_bicrystal, weighted_kappa

pytestmark = pytest.mark.gpu


def test_t1_curvature_gpu():
    z = np.load(os.path.join(GOLDEN, "t1_curvature.npz"))
    m = _abi.isotropic(1.0, 1.0)
    for R, text in zip(z["R"], z["kappa_text"]):
        N = 4 * int(R) + 24
        sigma = float(np.float32(2.0) * np.float32(R) / np.float32(15.0))
        G = gpk_amd.GPKGrid(1, N, N, 1.0, m, 0.0, sigma, precision="strict32")
        G.set("phi", circle(N, int(R)))
        G.step(0.0)
        k = weighted_kappa(G.get("mask"), G.get("phi"), G.get("nx"), G.get("ny"), G.get("kappa"))
        assert "%.9g" % k == str(text), (R, k, text)
        G.close()


@pytest.mark.parametrize("key", ["g13", "g14", "g15"])
def test_curved_bicrystal_gpu(key):
    before, after, t_text, a_text = load_bicrystal(key)
    dx = float(np.float32(0.1e-6))
    G = gpk_amd.GPKGrid(1, 360, 360, dx, _abi.isotropic(1e-8, 0.324), 0.0, 2.0, precision="strict32")
    G.set("phi", bicrystal_shape(key))
    ts, areas = [], []
    for snap in range(11):
        if snap > 0:
            for _ in range(80):
                G.step()
        v = G.phase_volume(1)
        if v <= 1e-12:
            break
        ts.append("%.8e" % G.time)
        areas.append("%.8e" % (v * dx * dx))
    assert np.array_equal(G.get("phi"), after)
    assert ts == [str(t) for t in t_text]
    assert areas == [str(a) for a in a_text]
