"""Ordered flip transactions in the regime where most candidates are refused.

Late in a 3-D grain-growth run, cells whose flip would disconnect a grain
stay candidates step after step (a refused candidate keeps alpha = 1e-6,
src/gpk_transitions.cpp:305-308), and most of them need a global
connectivity search.  The GPU decides such candidates before their turn
in the serial order, from searches on views that bound the undecided
earlier candidates.  These tests start the oracle from a late GPU state
and compare every following step bit for bit: labels, alpha, the accepted
and refused counts and the remap exchanges.
"""
import numpy as np
import pytest

import gpk_amd
from gpk_amd import _abi
from oracle import oracle as O
from tests.test_gpu_parity import lockstep

pytestmark = pytest.mark.gpu


def late_state(n, grains, prec, steps):
    G = gpk_amd.GPKGrid(n, n, n, 1.0, _abi.isotropic(1.0, 1.0), 0.0, 1.5, precision=prec)
    phi, g = O.voronoi(n, n, n, grains, 42)
    G.set("phi", phi)
    G.set("g", g.astype(G.real))
    refused = 0
    for _ in range(steps):
        G.step()
        refused = G.last_stats.deferred
    return G, refused


@pytest.mark.parametrize("prec", ["fp64", "strict32"])
def test_late_stage_lockstep_64(prec):
    # 64^3 at the 256^3 / 20,000-grain density (312 grains); 152 steps
    # (a multiple of 8, so the extinction cadence of both grids agrees)
    n, grains, warm = 64, 312, 152
    print("late-stage lockstep", prec, n, warm, flush=True)
    G, refused = late_state(n, grains, prec, warm)
    print("refused candidates at step %d: %d" % (warm, refused), flush=True)
    assert refused > 500
    o = O.OracleGrid(n, n, n, 1.0, _abi.isotropic(1.0, 1.0), 0.0, 1.5, prec)
    for f in ("phi", "alpha", "g"):
        o.set(f, G.get(f))
    o.time = G.time
    # g_comp keeps stale values outside the band (a full-grid work field of
    # the reference), which a freshly made oracle grid does not share
    lockstep(o, G, 12, fields=("phi", "alpha", "g", "mask", "kappa", "nx", "ny", "nz", "vn"), check_every=3)
    G.close()


def test_refusal_heavy_2d_lockstep_strict32():
    # 2-D: 200 grains on 256^2 for 400 steps in the oracle's and the GPU's
    # own hands from the same Voronoi start (the refusal count grows as the
    # grains coarsen); labels, alpha and counters equal at every step
    phi, g = O.voronoi(1, 256, 256, 200, 42)
    m = _abi.isotropic(1.0, 1.0)
    o = O.OracleGrid(1, 256, 256, 1.0, m, 0.0, 1.5, "strict32")
    G = gpk_amd.GPKGrid(1, 256, 256, 1.0, m, 0.0, 1.5, precision="strict32")
    for X in (o, G):
        X.set("phi", phi)
        X.set("g", g)
    lockstep(o, G, 400, fields=("phi", "alpha"), check_every=50)
    assert o.stats()["deferred"] > 100   # refusals in the last step (oracle: about 250)
    G.close()
