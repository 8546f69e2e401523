"""Batched 2-D instances (one persistent kernel per step over a packed
instance table) against the oracle stepping each instance on its own, bit for
bit in every precision mode: labels, alpha, orientation, mask, competitor and
each instance's time.  Workloads: lamellar stripes with seeded one-cell edge
perturbations under AnisoOstwaldThermo with c = c_sat (pure capillary flow,
demos/gpk.cpp:1105-1143), D = 0; Voronoi grains under isotropic growth with
the grain-network topology test."""
import numpy as np
import pytest

from gpk_amd import _abi
from gpk_amd.batch import GPKBatch, lamellar_sweep
from oracle import oracle as O

pytestmark = pytest.mark.gpu

C_SAT = np.float32(0.1)


def model():
    return _abi.aniso_ostwald(1.0, 1.0, 1.0, 0.5, 1.0, float(C_SAT), 1e8)


def make_batch(shapes, phis, mode, sigma=1.5):
    B = GPKBatch(shapes, 1.0, model(), 0.0, sigma, precision=mode)
    B.set("phi", phis)
    B.set("c", [np.full(s, C_SAT, np.float32) for s in shapes])
    B.set("g", [np.zeros(s, np.float32) for s in shapes])
    B.set("alpha", [np.ones(s, np.float32) for s in shapes])
    return B


def oracle_run(shape, phi, mode, steps, sigma=1.5):
    o = O.OracleGrid(1, shape[0], shape[1], 1.0, model(), 0.0, sigma, mode)
    o.set("phi", phi)
    o.set("c", np.full(shape, C_SAT, np.float32).astype(o.real))
    for _ in range(steps):
        o.step()
    return o


@pytest.mark.parametrize("mode", ["strict32", "fp32", "fp64"])
def test_batch_matches_oracle_per_instance(mode):
    shapes, phis, params = lamellar_sweep(10, seed=3, side_range=(40, 96), width_range=(4, 10),
                                          ratio_range=(2.0, 5.0))
    steps = 40
    B = make_batch(shapes, phis, mode)
    print("batch", mode, params, flush=True)
    B.step(steps)
    t, ms, nb = B.instances()
    got = {f: B.get(f) for f in ("phi", "alpha", "g", "mask", "phi_comp")}
    st = B.stats()
    assert st["accepted"] > 0 and st["exchanges"] >= 0
    for b, (shape, phi) in enumerate(zip(shapes, phis)):
        o = oracle_run(shape, phi, mode, steps)
        for f in ("phi", "alpha", "g", "mask", "phi_comp"):
            ref = o.get(f)
            assert np.array_equal(got[f][b], ref), (b, f, int(np.count_nonzero(got[f][b] != ref)))
        assert t[b] == o.time, (b, t[b], o.time)
    B.close()


def test_batch_launch_shapes_agree_and_autotune_restores_state():
    shapes, phis, _ = lamellar_sweep(24, seed=5, side_range=(32, 80), width_range=(4, 9),
                                     ratio_range=(2.0, 4.0))
    ref = None
    for wg, nwg in ((256, 3), (512, 24), (1024, 7)):
        B = make_batch(shapes, phis, "fp32")
        B.set_launch(wg, nwg)
        B.step(25)
        cur = (B.get("phi"), B.get("alpha"), B.instances()[0])
        if ref is None:
            ref = cur
        else:
            for b in range(len(shapes)):
                assert np.array_equal(cur[0][b], ref[0][b]) and np.array_equal(cur[1][b], ref[1][b])
            assert np.array_equal(cur[2], ref[2])
        B.close()
    B = make_batch(shapes, phis, "fp32")
    before = (B.get("phi"), B.get("alpha"), B.instances()[0])
    wg, nwg, ms = B.autotune()
    print("autotune", wg, nwg, ms, flush=True)
    assert wg in (256, 512, 1024) and nwg >= 1 and ms > 0
    after = (B.get("phi"), B.get("alpha"), B.instances()[0])
    for b in range(len(shapes)):
        assert np.array_equal(before[0][b], after[0][b]) and np.array_equal(before[1][b], after[1][b])
    assert np.array_equal(before[2], after[2])
    B.step(25)
    for b in range(len(shapes)):
        assert np.array_equal(B.get("phi")[b], ref[0][b])
    B.close()


GRAIN_CASES = [(40, 48, 40), (56, 56, 70), (64, 40, 50), (33, 71, 45), (48, 48, 90)]


@pytest.mark.parametrize("mode", ["strict32", "fp32", "fp64"])
def test_batch_grain_network_matches_oracle(mode):
    """Isotropic grain growth (grain-network topology test) from Voronoi
    grains: flips refused by the 3x3 count or a periodic connectivity search,
    and unresolved grains eroded every 8th step, all as in the oracle."""
    shapes = [(ny, nx) for ny, nx, _ in GRAIN_CASES]
    phis, gs = [], []
    for k, (ny, nx, ns) in enumerate(GRAIN_CASES):
        p, g = O.voronoi(1, ny, nx, ns, 11 + k)
        phis.append(p.reshape(ny, nx))
        gs.append(g.reshape(ny, nx))
    steps = 40
    thermo = _abi.isotropic(1.0, 1.0)
    B = GPKBatch(shapes, 1.0, thermo, 0.0, 1.5, precision=mode)
    B.set("phi", phis)
    B.set("g", gs)
    B.set("alpha", [np.ones(s, np.float32) for s in shapes])
    print("batch grain network", mode, GRAIN_CASES, flush=True)
    B.step(steps)
    t = B.instances()[0]
    got = {f: B.get(f) for f in ("phi", "alpha", "g", "mask", "phi_comp")}
    st = B.stats()
    tot = {"accepted": 0, "deferred": 0, "eroded": 0, "exchanges": 0}
    for b, (shape, phi) in enumerate(zip(shapes, phis)):
        o = O.OracleGrid(1, shape[0], shape[1], 1.0, thermo, 0.0, 1.5, mode)
        o.set("phi", phi)
        o.set("g", gs[b].astype(o.real))
        for _ in range(steps):
            o.step()
            ost = o.stats()
            for key in tot:
                tot[key] += ost[key]
        for f in ("phi", "alpha", "g", "mask", "phi_comp"):
            ref = o.get(f)
            assert np.array_equal(got[f][b], ref), (b, f, int(np.count_nonzero(got[f][b] != ref)))
        assert t[b] == o.time, (b, t[b], o.time)
    # every path of the decision ran: refusals, erosion and remap exchanges
    assert tot["deferred"] > 0 and tot["eroded"] > 0 and tot["exchanges"] > 0
    for key in tot:
        assert st[key] == tot[key], (key, st[key], tot[key])
    B.close()


def test_batch_rejects_unsupported_models():
    with pytest.raises(Exception):
        GPKBatch([(32, 32)], 1.0, _abi.dislocation(1.0, 1.0, 1.0, 1.0), 0.0, 1.5, precision="fp32")
    with pytest.raises(Exception):
        GPKBatch([(32, 32)], 1.0, model(), 1e-9, 1.5, precision="fp32")


def test_batch_grain_network_larger_instances():
    """Grain growth on larger instances (up to 128^2, longer connectivity
    searches through bigger grains), FP64, against the oracle bit for bit."""
    cases = [(128, 128, 220), (96, 160, 200), (120, 72, 150)]
    shapes = [(ny, nx) for ny, nx, _ in cases]
    phis, gs = [], []
    for k, (ny, nx, ns) in enumerate(cases):
        p, g = O.voronoi(1, ny, nx, ns, 77 + k)
        phis.append(p.reshape(ny, nx))
        gs.append(g.reshape(ny, nx))
    steps = 24
    thermo = _abi.isotropic(1.0, 1.0)
    B = GPKBatch(shapes, 1.0, thermo, 0.0, 1.5, precision="fp64")
    B.set("phi", phis)
    B.set("g", gs)
    B.set("alpha", [np.ones(s, np.float32) for s in shapes])
    print("batch grain network, larger instances", cases, flush=True)
    B.step(steps)
    got = {f: B.get(f) for f in ("phi", "alpha", "g")}
    st = B.stats()
    deferred = 0
    for b, (shape, phi) in enumerate(zip(shapes, phis)):
        o = O.OracleGrid(1, shape[0], shape[1], 1.0, thermo, 0.0, 1.5, "fp64")
        o.set("phi", phi)
        o.set("g", gs[b].astype(o.real))
        for _ in range(steps):
            o.step()
            deferred += o.stats()["deferred"]
        for f in ("phi", "alpha", "g"):
            assert np.array_equal(got[f][b], o.get(f)), (b, f)
    assert deferred > 0 and st["deferred"] == deferred
    B.close()


def test_batch_full_scale_lamellar_fp32_guard_band():
    # config-4 scale: instances of side 300-384 with the sweep's stripe
    # parameters (t in [6, 20], period ratio in [2, 12]), 150 steps in FP32
    # with the FP64 guard band, each instance against the oracle's fp32 mode
    shapes, phis, params = lamellar_sweep(5, seed=11, side_range=(300, 384))
    steps = 150
    print("full-scale batch", params, flush=True)
    B = make_batch(shapes, phis, "fp32")
    B.step(steps)
    t, _, _ = B.instances()
    got = {f: B.get(f) for f in ("phi", "alpha", "g")}
    st = B.stats()
    assert st["accepted"] > 0
    for b, (shape, phi) in enumerate(zip(shapes, phis)):
        o = oracle_run(shape, phi, "fp32", steps)
        for f in ("phi", "alpha", "g"):
            ref = o.get(f)
            assert np.array_equal(got[f][b], ref), (b, f, int(np.count_nonzero(got[f][b] != ref)))
        assert t[b] == o.time, (b, t[b], o.time)
    print("accepted", st["accepted"], "exchanges", st["exchanges"], "guarded", st["guarded"], flush=True)
    B.close()


@pytest.mark.parametrize("mode", ["fp32", "fp64"])
@pytest.mark.parametrize("sigma", [1.2, 2.0])
def test_batch_other_geometry_radii(mode, sigma):
    # sigma 1.2 (geometry radius 7) and 2.0 (radius 9) take the cell-by-cell
    # geometry next to the tiled projection and remap scoring (radius 3)
    shapes, phis, params = lamellar_sweep(6, seed=11, side_range=(40, 80), width_range=(4, 10),
                                          ratio_range=(2.0, 5.0))
    steps = 25
    B = make_batch(shapes, phis, mode, sigma)
    print("batch", mode, sigma, params, flush=True)
    B.step(steps)
    t, ms, nb = B.instances()
    got = {f: B.get(f) for f in ("phi", "alpha", "g", "mask", "phi_comp")}
    assert B.stats()["accepted"] > 0
    for b, (shape, phi) in enumerate(zip(shapes, phis)):
        o = oracle_run(shape, phi, mode, steps, sigma)
        for f in ("phi", "alpha", "g", "mask", "phi_comp"):
            ref = o.get(f)
            assert np.array_equal(got[f][b], ref), (b, f, int(np.count_nonzero(got[f][b] != ref)))
        assert t[b] == o.time, (b, t[b], o.time)
    B.close()
