"""Shared set-ups for the parity tests (initial conditions of the benchmark and deck workloads)."""
import os

import numpy as np

from gpk_amd import _abi

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

FIELDS_ALL = ("phi", "alpha", "g", "mask", "phi_comp", "kappa", "nx", "ny", "nz", "vn", "g_comp", "c")


def load_bicrystal(key):
    z = np.load(os.path.join(GOLDEN, key + ".npz"))
    ny, nx = (int(v) for v in z["shape"])
    before = np.unpackbits(z["before"])[: ny * nx].reshape(ny, nx).astype(np.int32)
    after = np.unpackbits(z["after"])[: ny * nx].reshape(ny, nx).astype(np.int32)
    return before, after, list(z["t_text"]), list(z["area_text"])


def bicrystal_shape(key, n=360):
    y, x = np.mgrid[0:n, 0:n]
    dy, dx = y - n // 2, x - n // 2
    if key == "g13":
        return (dy * dy + dx * dx <= 72 * 72).astype(np.int32)
    if key == "g14":
        return ((dy / 88.0) ** 2 + (dx / 48.0) ** 2 <= 1.0).astype(np.int32)
    return ((np.abs(dy) <= 64) & (np.abs(dx) <= 64)).astype(np.int32)


def circle(N, R):
    y, x = np.mgrid[0:N, 0:N]
    return ((y - N // 2) ** 2 + (x - N // 2) ** 2 <= R * R).astype(np.int32)


def weighted_kappa(mask, phi, nx, ny, kappa):
    """Weighted mean curvature of the T1 check (demos/validate_theory_foundations.cpp:126-135)."""
    sel = np.flatnonzero((mask.astype(bool) & (phi == 1)).ravel())
    nxv = nx.ravel()[sel].astype(np.float64)
    nyv = ny.ravel()[sel].astype(np.float64)
    dom = np.maximum(np.abs(nxv), np.abs(nyv))
    cut = np.where(dom > 1e-8, 1.0 / np.where(dom > 1e-8, dom, 1.0), 1.0)
    kv = kappa.ravel()[sel].astype(np.float64)
    # sequential double sums in index order, as the reference loop
    wk = 0.0
    w = 0.0
    for a, b in zip(kv, cut):
        wk += a * b
        w += b
    return wk / w


def solidification_setup(N, rng_seed=7, c0=0.02, k=0.14, undercooling=0.5, r=6):
    """128^3-style solute-coupled solidification start: a solid sphere of radius r
    at the centre with c = c0*k inside, c0 + U(-1e-3, 1e-3) outside; the
    dimensionless undercooling Delta maps to dT = Delta*|m|c0(1-k)/k."""
    m_liq, T_melt = -500.0, 933.0
    model = _abi.solidification(M0=5e-8, sigma0=0.2, eps4=0.02, m_liq=m_liq, k_part=k,
                                T_melt=T_melt, L=1e8, solute_rejection=True)
    zz, yy, xx = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    c = N // 2
    inside = (zz - c) ** 2 + (yy - c) ** 2 + (xx - c) ** 2 <= r * r
    phi = np.zeros((N, N, N), np.int32)
    phi[inside] = 1
    rng = np.random.default_rng(rng_seed)
    conc = (c0 + rng.uniform(-1e-3, 1e-3, size=phi.shape)).astype(np.float32)
    conc[inside] = np.float32(c0 * k)
    dT = undercooling * abs(m_liq) * c0 * (1 - k) / k
    T = np.full(phi.shape, T_melt + m_liq * c0 - dT, np.float32)
    return model, phi, conc, T


def rex_setup(N=80):
    """Deformed matrix (label 0) with three recrystallised discs (labels 1-3)."""
    y, x = np.mgrid[0:N, 0:N]
    phi = np.zeros((N, N), np.int32)
    for k, (cy, cx, r) in enumerate([(20, 20, 7), (55, 30, 9), (40, 62, 6)]):
        phi[(y - cy) ** 2 + (x - cx) ** 2 <= r * r] = k + 1
    return phi


def dislocation_rho(N=80):
    """Dislocation density varying along x, 1e14 m^-2 +- 50 %."""
    x = np.arange(N)[None, :].repeat(N, 0)
    return (1e14 * (1 + 0.5 * np.sin(x / 7.0))).astype(np.float32)


def precipitation_setup(N=80):
    """Matrix phase 0 with discs of phases 1 and 2; c varies along x, T = 880 K
    (inputs/precipitation.ini coefficients)."""
    y, x = np.mgrid[0:N, 0:N]
    phi = rex_setup(N)
    phi[phi == 3] = 1
    c = (0.06 + 0.01 * np.sin(x / 9.0)).astype(np.float32)
    T = np.full((N, N), 880.0, np.float32)
    model = _abi.precipitation(2e-8, 0.18, [0.0, -8.0e6, -7.5e6], [0.0, 2.0e7, -1.0e7],
                               [0.0, 2.5e8, 1.8e8], [0.0, 1.0e4, 5.0e3], 900.0)
    return model, {"phi": phi, "c": c, "T": T}


def elastic_solidification_setup(N=80):
    """Solid disc r = 8 with c = 0.006 in liquid c = 0.02, 20 K undercooled;
    elastic coupling E = 5e9, eigenstrain 0.05 (c - 0.02), 12 smoothing sweeps."""
    y, x = np.mgrid[0:N, 0:N]
    phi = (((y - 40) ** 2 + (x - 40) ** 2) <= 64).astype(np.int32)
    c = np.full((N, N), 0.02, np.float32)
    c[phi == 1] = 0.006
    T = np.full((N, N), 903.0, np.float32)
    model = _abi.with_elastic(_abi.solidification(M0=5e-8, sigma0=0.2, eps4=0.0), 5e9, 0.05, 0.02, 12)
    return model, {"phi": phi, "c": c, "T": T}
