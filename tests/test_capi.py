"""C ABI checks that need no GPU: the library loads, exports every symbol
include/gpk_c.h declares, reports errors through status codes, and its
host-side Voronoi generator matches the oracle."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import gpk_amd
from gpk_amd import _abi, _lib
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "gpk_c.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gpk_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared()
    assert len(names) >= 25
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_lib.EXPORTS)


def test_struct_layouts_match_header():
    # sizes of the POD structs crossing the ABI
    assert C.sizeof(_abi.ModelDesc) == 2 * 4 + 21 * 4 + 6 * 4 + 32 * 4
    assert C.sizeof(_abi.GridDesc) == 8 * 4
    assert C.sizeof(_abi.StepStats) == 6 * 8 + 2 * 8 + 8 + 8
    # four counters, three int32 (padded to 8), the guard-band counter
    assert C.sizeof(_abi.BatchStats) == 4 * 8 + 3 * 4 + 4 + 3 * 8


def test_model_defaults_agree_with_python_mirror():
    L = _lib.lib()
    for kind in range(8):
        a = _abi.ModelDesc()
        L.gpk_model_defaults(kind, C.byref(a))
        b = _abi.model_defaults(kind)
        for name, _ in _abi.ModelDesc._fields_:
            va, vb = getattr(a, name), getattr(b, name)
            if isinstance(va, C.Array):
                va, vb = list(va), list(vb)
            assert va == pytest.approx(vb), (kind, name)


def test_errors_are_status_codes():
    L = _lib.lib()
    h = C.c_void_p()
    gd = _abi.GridDesc(1, 8, 8, 1.0, 0.0, 1.5, 0, 0)
    m = _abi.isotropic(1.0, 1.0)
    rc = L.gpk_create(None, C.byref(m), C.byref(h))
    assert rc == _abi.GPK_E_INVALID
    assert b"null" in L.gpk_last_error()
    gd.Nx = 0
    assert L.gpk_create(C.byref(gd), C.byref(m), C.byref(h)) == _abi.GPK_E_INVALID
    gd.Nx = 8
    bad = _abi.isotropic(1.0, 1.0)
    bad.kind = 17
    assert L.gpk_create(C.byref(gd), C.byref(bad), C.byref(h)) == _abi.GPK_E_UNSUPPORTED_MODEL
    assert L.gpk_step(None, -1.0, 1e30, None, None) == _abi.GPK_E_INVALID


def test_voronoi_host_matches_oracle():
    for dims, ng, seed in (((1, 64, 48), 30, 7), ((16, 20, 24), 40, 3), ((24, 24, 24), 60, 42),
                           ((1, 96, 96), 200, 42), ((20, 30, 25), 333, 9)):
        phi, g = gpk_amd.voronoi_host(*dims, ng, seed)
        po, go = O.voronoi(*dims, ng, seed)
        assert np.array_equal(phi, po), dims
        assert np.array_equal(g, go), dims


@pytest.mark.parametrize("dims,ng,seed", [((1, 96, 96), 200, 42), ((24, 24, 24), 60, 42),
                                          ((40, 36, 32), 300, 5), ((1, 64, 64), 1500, 3),
                                          ((20, 20, 20), 900, 7), ((16, 24, 20), 700, 1)])
def test_voronoi_cases_have_detached_cells(dims, ng, seed):
    # the device Voronoi tests use these cases: each needs re-attachment
    assert O.voronoi_detached(*dims, ng, seed) > 0
    phi, g = gpk_amd.voronoi_host(*dims, ng, seed)
    po, go = O.voronoi(*dims, ng, seed)
    assert np.array_equal(phi.ravel(), po) and np.array_equal(g.ravel(), go)


def test_voronoi_rejects_bad_seed_count():
    with pytest.raises(gpk_amd.GPKError):
        gpk_amd.voronoi_host(1, 8, 8, 0, 1)
    with pytest.raises(gpk_amd.GPKError):
        gpk_amd.voronoi_host(1, 8, 8, 65, 1)


def test_bench_refuses_multi_gpu():
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
