"""Digital-topology certificate (src/gpk_connectivity.cpp:10-195) pins (CPU).

* the six canonical transactions of demos/validate_topology_transactions.cpp:29-80
  (2-D smooth advance, split, merger, extinction; 3-D split, extinction);
* exhaustive enumeration of every binary 3x3 neighbourhood of a 2-D cell
  (2^8 patterns, two labels), oracle against the compiled reference.
"""
import itertools

import numpy as np
import pytest

from gpk_amd import _abi
from oracle import oracle as O


def grid(dims, phi):
    m = _abi.isotropic(1.0, 1.0)
    o = O.OracleGrid(*dims, 1.0, m, 0.0, 1.5, "strict32")
    o.set("phi", phi)
    return o


CASES = []


def case(name, dims, expect):
    def deco(fn):
        CASES.append((name, dims, fn, expect))
        return fn
    return deco


# the six canonical transactions (demos/validate_topology_transactions.cpp:29-80)
@case("2D smooth boundary", (1, 9, 9), True)
def _smooth():
    phi = np.zeros((9, 9), np.int32)
    phi[:, :5] = 1
    return phi, 4 * 9 + 4, 0


@case("2D component split", (1, 11, 11), False)
def _split2():
    phi = np.zeros((11, 11), np.int32)
    phi[3:8, 2:5] = 1
    phi[3:8, 6:9] = 1
    phi[5, 5] = 1
    return phi, 5 * 11 + 5, 0


@case("2D premature merger", (1, 9, 9), False)
def _merge2():
    phi = np.zeros((9, 9), np.int32)
    phi[4, 3] = 1
    phi[4, 5] = 1
    return phi, 4 * 9 + 4, 1


@case("2D terminal extinction", (1, 9, 9), True)
def _ext2():
    phi = np.zeros((9, 9), np.int32)
    phi[4, 4] = 1
    return phi, 4 * 9 + 4, 0


@case("3D component split", (9, 9, 9), False)
def _split3():
    phi = np.zeros((9, 9, 9), np.int32)
    phi[3:6, 3:6, 1:4] = 2
    phi[3:6, 3:6, 5:8] = 2
    phi[4, 4, 4] = 2
    return phi, (4 * 9 + 4) * 9 + 4, 0


@case("3D terminal extinction", (7, 7, 7), True)
def _ext3():
    phi = np.zeros((7, 7, 7), np.int32)
    phi[3, 3, 3] = 3
    return phi, (3 * 7 + 3) * 7 + 3, 0


@pytest.mark.parametrize("name,dims,fn,expect", CASES, ids=[c[0] for c in CASES])
def test_canonical_transactions(name, dims, fn, expect):
    phi, cell, nl = fn()
    o = grid(dims, phi)
    assert o.simple_move(cell, nl) == expect


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")
def test_all_binary_3x3_patterns_match_reference():
    m = _abi.isotropic(1.0, 1.0)
    ring = [(-1, -1), (-1, 0), (-1, 1), (0, 1), (1, 1), (1, 0), (1, -1), (0, -1)]
    for bits in itertools.product((0, 1), repeat=8):
        for centre in (1, 2):
            phi = np.full((7, 7), 3, np.int32)
            phi[2:5, 2:5] = 1
            for (dy, dx), b in zip(ring, bits):
                phi[3 + dy, 3 + dx] = 1 if b else 2
            phi[3, 3] = centre
            o = O.OracleGrid(1, 7, 7, 1.0, m, 0.0, 1.5, "strict32")
            r = O.RefGrid(1, 7, 7, 1.0, m, 0.0, 1.5)
            o.set("phi", phi)
            r.set("phi", phi)
            for nl in (1, 2, 3):
                assert o.simple_move(24, nl) == r.simple_move(24, nl), (bits, centre, nl)
