"""FP64 guard band of the FP32 mode's flip decision.

In the FP32 mode alpha_new = alpha + v * (dt / dx) is evaluated in float; a
value within 2.5e-7 of the threshold 1e-6 is evaluated again in double and
decided there (strict32 keeps the reference's float decision,
src/gpk_transitions.cpp:262-291).  The case below is built so that the float
evaluation rounds onto the threshold (a pure-FP32 flip) while the exact value
of the same float inputs lies above it: one band cell of a 2-D disc gets
alpha = nextafter(1e-6f, 1) and a step with a hand-picked tiny dt that
brings v * dt to about -0.6 ulp.  The expected outcome is checked with numpy
float32/float64 arithmetic, independently of both implementations.
"""
import numpy as np
import pytest

from gpk_amd import _abi
from oracle.oracle import OracleGrid

N = 64
EPS = np.float32(1e-6)


def disc():
    y, x = np.mgrid[0:N, 0:N]
    return (((y - 31.5) ** 2 + (x - 29.5) ** 2) < 18.0 ** 2).astype(np.int32)


def pick_case(vn, mask, phi, phi_comp):
    """A band cell with v < 0 and a float32 dt for which float32 arithmetic
    puts alpha_new on the threshold while the exact value lies above it."""
    a0 = np.nextafter(EPS, np.float32(1))
    ulp = np.float64(a0) - np.float64(EPS)
    cells = np.flatnonzero(mask.ravel() & (phi.ravel() != phi_comp.ravel()) & (vn.ravel() < -1e-3))
    for i in cells:
        v = np.float32(vn.ravel()[i])
        dt0 = -0.6 * ulp / float(v)
        for k in range(-40, 41):
            dt = np.float32(dt0 * (1 + k * 0.01))
            idt = np.float32(dt / np.float32(1.0))
            an32 = np.float32(a0 + np.float32(v * idt))
            exact = np.float64(a0) + np.float64(v) * (np.float64(dt) / 1.0)
            if an32 <= EPS and exact > np.float64(EPS) and abs(float(an32) - float(EPS)) < 2.5e-7:
                return int(i), a0, dt, exact
    raise AssertionError("no guard-band case found")


def oracle_setup():
    O = OracleGrid(1, N, N, 1.0, _abi.isotropic(1.0, 1.0), 0.0, 1.5, mode="fp32")
    O.set("phi", disc())
    O.set("alpha", np.ones(N * N, np.float32))
    O.step(0.0)
    return O


def test_oracle_guard_band_decides_in_double():
    O = oracle_setup()
    vn, mask = O.get("vn"), O.get("mask").astype(bool)
    phi, phi_comp = O.get("phi"), O.get("phi_comp")
    i, a0, dt, exact = pick_case(vn, mask, phi, phi_comp)
    alpha = O.get("alpha").ravel().copy()
    alpha[i] = a0
    O.set("alpha", alpha)
    O.step(float(dt))
    st = O.stats()
    assert st["guarded"] == 1
    # pure float32 would have flipped the cell; the double decision keeps it
    assert O.get("phi").ravel()[i] == phi.ravel()[i]
    assert st["candidates"] == 0 and st["accepted"] == 0
    assert O.get("alpha").ravel()[i] == np.float32(exact)


def test_strict32_keeps_the_float_decision():
    # strict32 (the reference's arithmetic) has no guard band
    O = OracleGrid(1, N, N, 1.0, _abi.isotropic(1.0, 1.0), 0.0, 1.5, mode="strict32")
    O.set("phi", disc())
    O.set("alpha", np.ones(N * N, np.float32))
    O.step(0.0)
    O.step(0.0)
    assert O.stats()["guarded"] == 0


@pytest.mark.gpu
def test_gpu_guard_band_matches_oracle():
    import gpk_amd
    O = oracle_setup()
    G = gpk_amd.GPKGrid(1, N, N, 1.0, _abi.isotropic(1.0, 1.0), 0.0, 1.5, precision="fp32")
    G.set("phi", disc())
    G.set("alpha", np.ones(N * N, np.float32))
    G.step(0.0)
    vn = G.get("vn")
    assert np.array_equal(vn, O.get("vn"))
    i, a0, dt, exact = pick_case(vn, G.get("mask").astype(bool), G.get("phi"), G.get("phi_comp"))
    alpha = G.get("alpha").ravel().copy()
    alpha[i] = a0
    G.set("alpha", alpha)
    O.set("alpha", alpha)
    G.step(float(dt))
    O.step(float(dt))
    assert G.last_stats.guarded == 1 == O.stats()["guarded"]
    assert np.array_equal(G.get("phi"), O.get("phi"))
    assert np.array_equal(G.get("alpha").view(np.uint32), O.get("alpha").view(np.uint32))
    assert G.get("alpha").ravel()[i] == np.float32(exact)
    G.close()
