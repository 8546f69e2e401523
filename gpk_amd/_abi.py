"""ctypes mirror of include/gpk_c.h (structs, enums, flags).

Kept in one place so the product bindings and the test-side checkers agree
on the POD layouts that cross the C ABI.
"""
import ctypes as C
import math

# status codes
GPK_OK = 0
GPK_E_INVALID = -1
GPK_E_HIP = -2
GPK_E_UNSUPPORTED_MODEL = -3
GPK_E_LABEL_RANGE = -4
GPK_E_OOM = -5
GPK_E_NO_DEVICE = -6
GPK_E_STALLED = -7

# precision modes
STRICT32 = 0
FP32 = 1
FP64 = 2
PRECISIONS = {"strict32": STRICT32, "fp32": FP32, "fp64": FP64}

# fields: (id, kind) with kind "i32", "u8" or "real"
FIELDS = {
    "phi": (0, "i32"), "alpha": (1, "real"), "c": (2, "real"), "T": (3, "real"),
    "g": (4, "real"), "rho": (5, "real"), "nx": (6, "real"), "ny": (7, "real"),
    "nz": (8, "real"), "kappa": (9, "real"), "vn": (10, "real"),
    "phi_comp": (11, "i32"), "g_comp": (12, "real"), "mask": (13, "u8"), "E_el": (14, "real"),
}

# model kinds
ISOTROPIC = 0
ANISO_GROWTH = 1
SOLIDIFICATION = 2
TENSOR_ANISO = 3
ANISO_OSTWALD = 4
RECRYSTALLISATION = 5
DISLOCATION = 6
PRECIPITATION = 7
MAX_PHASES = 8

FLAG_GRAIN_NETWORK = 0x1
FLAG_SOLUTE_REJECTION = 0x2
FLAG_STOCHASTIC = 0x4
FLAG_DISLOCATION = 0x8
FLAG_ELASTIC = 0x10


class ModelDesc(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("flags", C.c_uint32),
        ("sigma0", C.c_float), ("eps4", C.c_float), ("eps6", C.c_float),
        ("M0", C.c_float), ("k_partition", C.c_float), ("noise_scale", C.c_float),
        ("kB", C.c_float),
        ("m_liq", C.c_float), ("k_part", C.c_float), ("T_melt", C.c_float), ("L", C.c_float),
        ("bxx", C.c_float), ("byy", C.c_float), ("bzz", C.c_float),
        ("c_sat", C.c_float), ("Omega", C.c_float), ("dG_bias", C.c_float),
        ("E_stored", C.c_float), ("disloc_coeff", C.c_float), ("rho_rex", C.c_float),
        ("recovery_rate", C.c_float),
        ("elastic_modulus", C.c_float), ("eigenstrain_coeff", C.c_float), ("c_ref", C.c_float),
        ("elastic_smooth_iters", C.c_int32),
        ("n_phases", C.c_int32), ("T_ref", C.c_float),
        ("pa", C.c_float * 8), ("pb", C.c_float * 8), ("pc2", C.c_float * 8), ("pdT", C.c_float * 8),
    ]


class GridDesc(C.Structure):
    _fields_ = [
        ("Nz", C.c_int32), ("Ny", C.c_int32), ("Nx", C.c_int32),
        ("dx", C.c_float), ("D", C.c_float), ("gaussian_sigma", C.c_float),
        ("precision", C.c_int32), ("device", C.c_int32),
    ]


class BatchStats(C.Structure):
    """gpk_batch_stats (include/gpk_c.h)."""
    _fields_ = [("band_cells", C.c_int64), ("accepted", C.c_int64), ("exchanges", C.c_int64),
                ("proposals", C.c_int64), ("instances", C.c_int32), ("workgroup_size", C.c_int32),
                ("workgroups", C.c_int32), ("guarded", C.c_int64), ("deferred", C.c_int64),
                ("eroded", C.c_int64)]


class StepStats(C.Structure):
    _fields_ = [
        ("band_cells", C.c_int64), ("candidates", C.c_int64), ("accepted", C.c_int64),
        ("deferred", C.c_int64), ("exchanges", C.c_int64), ("eroded", C.c_int64),
        ("dt", C.c_double), ("time", C.c_double), ("step_id", C.c_int64), ("guarded", C.c_int64),
    ]


def model_defaults(kind):
    """Defaults of the reference ThermoModel base (include/gpk_core.hpp:73-123)
    and of each model's constructor (include/gpk_thermo.hpp)."""
    m = ModelDesc()
    m.kind = kind
    m.flags = 0
    m.sigma0 = 0.5
    m.eps4 = 0.0
    m.eps6 = 0.0
    m.M0 = 1e-8
    m.k_partition = 1.0
    m.noise_scale = 1.0
    m.kB = 1.380649e-23
    m.m_liq, m.k_part, m.T_melt, m.L = -500.0, 0.3, 933.0, 1e8
    m.bxx = m.byy = m.bzz = 1.0
    m.c_sat, m.Omega, m.dG_bias = 0.0, 0.0, 0.0
    m.elastic_smooth_iters = 8
    m.T_ref = 900.0
    if kind in (ISOTROPIC, ANISO_GROWTH):
        m.flags = FLAG_GRAIN_NETWORK
    if kind == SOLIDIFICATION:
        m.M0, m.sigma0, m.eps4 = 5e-8, 0.2, 0.04
    return m


def isotropic(M0, sigma0=0.5):
    """IsotropicGrowthThermo(M0, sigma0) (include/gpk_thermo.hpp:18-27)."""
    m = model_defaults(ISOTROPIC)
    m.M0, m.sigma0, m.eps4 = M0, sigma0, 0.0
    return m


def anisotropic_growth(M0, sigma0, eps4, eps6=0.0):
    """AnisotropicGrowthThermo(M0, sigma0, eps4) (include/gpk_thermo.hpp:33-45);
    eps6 is a 2-D six-fold extension (0 reproduces the reference law)."""
    m = model_defaults(ANISO_GROWTH)
    m.M0, m.sigma0, m.eps4, m.eps6 = M0, sigma0, eps4, eps6
    return m


def solidification(M0=5e-8, sigma0=0.2, eps4=0.04, m_liq=-500.0, k_part=0.3,
                   T_melt=933.0, L=1e8, solute_rejection=False):
    """SolidificationThermo (include/gpk_thermo.hpp:52-85); solute_rejection
    sets ThermoModel::solute_rejection with k_partition = k_part, as the
    solidification driver does (demos/gpk.cpp:329-330)."""
    m = model_defaults(SOLIDIFICATION)
    m.M0, m.sigma0, m.eps4 = M0, sigma0, eps4
    m.m_liq, m.k_part, m.T_melt, m.L = m_liq, k_part, T_melt, L
    if solute_rejection:
        m.flags |= FLAG_SOLUTE_REJECTION
        m.k_partition = k_part
    return m


def aniso_ostwald(M0, sigma0, bxx, byy, bzz, c_sat, Omega):
    """AnisoOstwaldThermo (include/gpk_thermo.hpp:284-325)."""
    m = model_defaults(ANISO_OSTWALD)
    m.M0, m.sigma0 = M0, sigma0
    m.bxx, m.byy, m.bzz, m.c_sat, m.Omega = bxx, byy, bzz, c_sat, Omega
    return m


def tensor_aniso(M0, sigma0, bxx=1.0, byy=1.0, bzz=1.0, dG_bias=0.0):
    """TensorAnisotropyThermo (include/gpk_thermo.hpp:229-263)."""
    m = model_defaults(TENSOR_ANISO)
    m.M0, m.sigma0 = M0, sigma0
    m.bxx, m.byy, m.bzz, m.dG_bias = bxx, byy, bzz, dG_bias
    return m


def recrystallisation(M0, sigma0, E_stored):
    """RecrystallisationThermo(M0, sigma0, E_stored) (include/gpk_thermo.hpp:90-111):
    label 0 is the deformed matrix, labels >= 1 recrystallised grains."""
    m = model_defaults(RECRYSTALLISATION)
    m.M0, m.sigma0, m.E_stored = M0, sigma0, E_stored
    return m


def dislocation(M0, sigma0, mu, b, alphaT=0.5, recovery_rate=0.0, rho_rex=0.0):
    """DislocationThermo (include/gpk_thermo.hpp:135-170): stored energy
    disloc_coeff * rho with disloc_coeff = 0.5 alphaT mu b^2 in float, a swept
    boundary leaves rho_rex behind, static recovery at recovery_rate."""
    f = lambda v: C.c_float(v).value  # noqa: E731
    coeff = f(f(f(f(0.5 * f(alphaT)) * f(mu)) * f(b)) * f(b))
    m = model_defaults(DISLOCATION)
    m.M0, m.sigma0 = M0, sigma0
    m.flags |= FLAG_DISLOCATION
    m.disloc_coeff, m.recovery_rate, m.rho_rex = coeff, recovery_rate, rho_rex
    return m


def precipitation(M0, sigma0, a, b, c2, dT, T_ref=900.0):
    """PrecipitationThermo (include/gpk_thermo.hpp:333-387): phases 0..n-1 with
    G_q = a_q + b_q c + c2_q c^2 + dT_q (T - T_ref); the product phase of a
    cell is its lowest-G unlike face label."""
    n = len(a)
    if not (len(b) == len(c2) == len(dT) == n) or n > MAX_PHASES:
        raise ValueError("precipitation: 1..%d phases with four coefficient lists" % MAX_PHASES)
    m = model_defaults(PRECIPITATION)
    m.M0, m.sigma0, m.n_phases, m.T_ref = M0, sigma0, n, T_ref
    for q in range(n):
        m.pa[q], m.pb[q], m.pc2[q], m.pdT[q] = a[q], b[q], c2[q], dT[q]
    return m


def with_elastic(m, modulus, eigenstrain_coeff, c_ref, smooth_iters=8):
    """Switch on the elastic coupling of any model (ThermoModel::elastic_enabled,
    src/gpk_physics.cpp:10-63)."""
    m.flags |= FLAG_ELASTIC
    m.elastic_modulus, m.eigenstrain_coeff, m.c_ref = modulus, eigenstrain_coeff, c_ref
    m.elastic_smooth_iters = smooth_iters
    return m


def half_pi_f32():
    return float(C.c_float(math.pi * 0.5).value)
