// gpk_device.hpp — device-side building blocks of the GPK per-step kinetics
// for CDNA4 (gfx950): grid indexing, digital-topology bit kernels, the
// constitutive laws as POD-parameterised device functions.
//
// Reference algorithm citations are to the GPK source tree.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/gpk_c.h"

namespace gpkd {

// ---------------------------------------------------------------------------
// Grid shape; flat index i = (z*Ny + y)*Nx + x (include/gpk_core.hpp:12-14).
// ---------------------------------------------------------------------------
struct Dims {
    int Nz, Ny, Nx, N;
};

__device__ __forceinline__ void decode(const Dims& d, int i, int& z, int& y, int& x) {
    x = i % d.Nx;
    const int t = i / d.Nx;
    y = t % d.Ny;
    z = t / d.Ny;
}
__device__ __forceinline__ int lin(const Dims& d, int z, int y, int x) {
    return (z * d.Ny + y) * d.Nx + x;
}
__device__ __forceinline__ int wrapi(int v, int n) {
    while (v < 0) v += n;
    while (v >= n) v -= n;
    return v;
}
// face neighbours in the reference order x-, x+, y-, y+, z-, z+
__device__ __forceinline__ void faces(const Dims& d, int i, int f[6]) {
    int z, y, x;
    decode(d, i, z, y, x);
    const int xm = x == 0 ? d.Nx - 1 : x - 1, xp = x == d.Nx - 1 ? 0 : x + 1;
    const int ym = y == 0 ? d.Ny - 1 : y - 1, yp = y == d.Ny - 1 ? 0 : y + 1;
    const int zm = z == 0 ? d.Nz - 1 : z - 1, zp = z == d.Nz - 1 ? 0 : z + 1;
    f[0] = lin(d, z, y, xm); f[1] = lin(d, z, y, xp);
    f[2] = lin(d, z, ym, x); f[3] = lin(d, z, yp, x);
    f[4] = lin(d, zm, y, x); f[5] = lin(d, zp, y, x);
}

// ---------------------------------------------------------------------------
// Face-connected component counting inside the 3^d cube (bit-parallel flood).
// Cube index = lz*9 + ly*3 + lx (2-D: ly*3 + lx).  Counts the components of
// `occ` (centre excluded by the caller) that contain a face neighbour of the
// centre (src/gpk_connectivity.cpp:23-77 with full_connectivity == false).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t grow3(uint32_t c, bool d3) {
    // masks of positions with lx==0 / lx==2 / ly==0 / ly==2 across all layers
    const uint32_t X0 = 0x1249249u;       // bits 0,3,6,...,24
    const uint32_t X2 = X0 << 2;
    const uint32_t Y0 = 0x01C0E07u;       // bits 0-2, 9-11, 18-20
    const uint32_t Y2 = Y0 << 6;
    uint32_t g = c | ((c << 1) & ~X0) | ((c >> 1) & ~X2) | ((c << 3) & ~Y0) | ((c >> 3) & ~Y2);
    if (d3) g |= (c << 9) | (c >> 9);
    return g & 0x7FFFFFFu;
}

__device__ __forceinline__ int topo_count(uint32_t occ, bool d3) {
    const uint32_t facemask = d3 ? ((1u << 4) | (1u << 10) | (1u << 12) | (1u << 14) |
                                    (1u << 16) | (1u << 22))
                                 : ((1u << 1) | (1u << 3) | (1u << 5) | (1u << 7));
    int n = 0;
    uint32_t rem = occ;
    while (rem) {
        uint32_t comp = rem & (~rem + 1u);
        while (true) {
            const uint32_t nc = grow3(comp, d3) & occ;
            if (nc == comp) break;
            comp = nc;
        }
        rem &= ~comp;
        if (comp & facemask) ++n;
    }
    return n;
}

// 5^d box (bounded certificate, src/gpk_connectivity.cpp:124-173): box index
// = lz*25 + ly*5 + lx, 125 bits held in two 64-bit words.
struct Bits128 {
    uint64_t lo, hi;
};
__device__ __forceinline__ Bits128 b_or(Bits128 a, Bits128 b) { return {a.lo | b.lo, a.hi | b.hi}; }
__device__ __forceinline__ Bits128 b_and(Bits128 a, Bits128 b) { return {a.lo & b.lo, a.hi & b.hi}; }
__device__ __forceinline__ Bits128 b_andn(Bits128 a, Bits128 b) { return {a.lo & ~b.lo, a.hi & ~b.hi}; }
__device__ __forceinline__ bool b_eq(Bits128 a, Bits128 b) { return a.lo == b.lo && a.hi == b.hi; }
__device__ __forceinline__ Bits128 b_shl(Bits128 a, int s) {
    if (s == 0) return a;
    if (s >= 64) return {0ull, a.lo << (s - 64)};
    return {a.lo << s, (a.hi << s) | (a.lo >> (64 - s))};
}
__device__ __forceinline__ Bits128 b_shr(Bits128 a, int s) {
    if (s == 0) return a;
    if (s >= 64) return {a.hi >> (s - 64), 0ull};
    return {(a.lo >> s) | (a.hi << (64 - s)), a.hi >> s};
}
__device__ __forceinline__ void b_set(Bits128& a, int bit) {
    if (bit < 64) a.lo |= 1ull << bit; else a.hi |= 1ull << (bit - 64);
}
__device__ __forceinline__ bool b_test(const Bits128& a, int bit) {
    return bit < 64 ? ((a.lo >> bit) & 1ull) : ((a.hi >> (bit - 64)) & 1ull);
}

struct BoxMasks {
    Bits128 X0, X4, Y0, Y4, ALL;
};
__device__ __forceinline__ BoxMasks box_masks(bool d3) {
    BoxMasks m{{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}};
    const int nz = d3 ? 5 : 1;
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < 5; ++y)
            for (int x = 0; x < 5; ++x) {
                const int b = z * 25 + y * 5 + x;
                b_set(m.ALL, b);
                if (x == 0) b_set(m.X0, b);
                if (x == 4) b_set(m.X4, b);
                if (y == 0) b_set(m.Y0, b);
                if (y == 4) b_set(m.Y4, b);
            }
    return m;
}
__device__ __forceinline__ Bits128 grow5(Bits128 c, const BoxMasks& m, bool d3) {
    Bits128 g = c;
    g = b_or(g, b_andn(b_shl(c, 1), m.X0));
    g = b_or(g, b_andn(b_shr(c, 1), m.X4));
    g = b_or(g, b_andn(b_shl(c, 5), m.Y0));
    g = b_or(g, b_andn(b_shr(c, 5), m.Y4));
    if (d3) {
        g = b_or(g, b_shl(c, 25));
        g = b_or(g, b_shr(c, 25));
    }
    return b_and(g, m.ALL);
}

// ---------------------------------------------------------------------------
// Constitutive laws (include/gpk_thermo.hpp:18-85, 229-325; src/gpk_thermo.cpp:9-30)
// ---------------------------------------------------------------------------
// PrecipitationThermo::G (include/gpk_thermo.hpp:358-360)
template <class R>
__device__ __forceinline__ R law_gibbs(const gpk_model_desc& d, int q, R cv, R Tv) {
    return R(d.pa[q]) + R(d.pb[q]) * cv + R(d.pc2[q]) * cv * cv + R(d.pdT[q]) * (Tv - R(d.T_ref));
}

// PrecipitationThermo::select_product_phase (include/gpk_thermo.hpp:371-383)
// over the six face labels in face order.  The reference's candidate list is
// the distinct unlike face labels in first-occurrence order; a repeated label
// has the same dG and never wins the strict "<" a second time, and the parent
// label is skipped, so scanning the faces directly gives the same answer.
template <class R>
__device__ __forceinline__ int law_select_phase(const gpk_model_desc& d, int32_t pp, R cv, R Tv,
                                                const int32_t (&fl)[6]) {
    R best = R(1e30f);
    int best_q = -1;
    if (pp < 0 || pp >= d.n_phases) return -1;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const int32_t q = fl[k];
        if (q < 0 || q >= d.n_phases || q == pp) continue;
        const R dG = law_gibbs<R>(d, q, cv, Tv) - law_gibbs<R>(d, pp, cv, Tv);
        if (dG < best) { best = dG; best_q = q; }
    }
    return best_q;
}

template <class R>
__device__ __forceinline__ R law_driving_force(const gpk_model_desc& d, int32_t pp, int32_t pq,
                                               R cv, R Tv) {
    switch (d.kind) {
    case GPK_MODEL_SOLIDIFICATION: {
        const R T_liq = R(d.T_melt) + R(d.m_liq) * cv;
        const R dT = T_liq - Tv;
        const R dG = (R(d.L) / R(d.T_melt)) * dT;
        if (pp == 0 && pq == 1) return -dG;
        if (pp == 1 && pq == 0) return dG;
        return R(0);
    }
    case GPK_MODEL_TENSOR_ANISO:
        if (d.dG_bias == 0.f) return R(0);
        return (pp != 0) ? R(d.dG_bias) : -R(d.dG_bias);
    case GPK_MODEL_ANISO_OSTWALD:
        if (pp == 0) return R(d.Omega) * (R(d.c_sat) - cv);
        return R(d.Omega) * (cv - R(d.c_sat));
    case GPK_MODEL_RECRYSTALLISATION:   // include/gpk_thermo.hpp:100-108
        if (pp == 0 && pq >= 1) return -R(d.E_stored);
        if (pp >= 1 && pq == 0) return R(d.E_stored);
        return R(0);
    case GPK_MODEL_PRECIPITATION:       // include/gpk_thermo.hpp:362-369
        if (pp < 0 || pp >= d.n_phases || pq < 0 || pq >= d.n_phases) return R(0);
        return law_gibbs<R>(d, pq, cv, Tv) - law_gibbs<R>(d, pp, cv, Tv);
    default:
        return R(0);
    }
}

template <class R>
__device__ __forceinline__ R law_interface_composition(const gpk_model_desc& d, int32_t pp,
                                                       int32_t pq, R cp, R cq) {
    if (d.kind == GPK_MODEL_ANISO_OSTWALD) {
        if (pp == 0) return cp;
        if (pq == 0) return cq;
    }
    return cp;
}

__device__ __forceinline__ float dcos(float v) { return cosf(v); }
__device__ __forceinline__ double dcos(double v) { return cos(v); }
__device__ __forceinline__ float dsin(float v) { return sinf(v); }
__device__ __forceinline__ double dsin(double v) { return sin(v); }
__device__ __forceinline__ float datan2(float a, float b) { return atan2f(a, b); }
__device__ __forceinline__ double datan2(double a, double b) { return atan2(a, b); }
__device__ __forceinline__ float dsqrt(float v) { return sqrtf(v); }
__device__ __forceinline__ double dsqrt(double v) { return sqrt(v); }
__device__ __forceinline__ float dfabs(float v) { return fabsf(v); }
__device__ __forceinline__ double dfabs(double v) { return fabs(v); }

template <class R> struct Lim;
template <> struct Lim<float> {
    static constexpr float aniso_k = 1e-12f;
};
template <> struct Lim<double> {
    static constexpr double aniso_k = 1e-12;
};

template <class R>
__device__ __forceinline__ R law_energy(const gpk_model_desc& d, R nx, R ny, R nz, R gv) {
    if (d.kind == GPK_MODEL_TENSOR_ANISO || d.kind == GPK_MODEL_ANISO_OSTWALD) {
        const R cg = dcos(gv), sg = dsin(gv);
        const R ncx = nx * cg + ny * sg;
        const R ncy = -nx * sg + ny * cg;
        const R ncz = nz;
        const R K = R(d.bxx) * ncx * ncx + R(d.byy) * ncy * ncy + R(d.bzz) * ncz * ncz;
        const R Kc = K < Lim<R>::aniso_k ? Lim<R>::aniso_k : K;
        return R(d.sigma0) * dsqrt(Kc);
    }
    if (d.eps4 == 0.f && d.eps6 == 0.f) return R(d.sigma0);
    const R cg = dcos(gv), sg = dsin(gv);
    const R ncx = nx * cg + ny * sg;
    const R ncy = -nx * sg + ny * cg;
    const R ncz = nz;
    if (ncz == R(0)) {
        const R th = datan2(ncy, ncx);
        R s = R(1) + R(d.eps4) * dcos(R(4) * th);
        if (d.eps6 != 0.f) s += R(d.eps6) * dcos(R(6) * th);
        return R(d.sigma0) * s;
    }
    const R aniso = ncx * ncx * ncx * ncx + ncy * ncy * ncy * ncy + ncz * ncz * ncz * ncz - R(3) / R(5);
    return R(d.sigma0) * (R(1) + R(d.eps4) * aniso);
}

// ---------------------------------------------------------------------------
// Counter-based normal deviate (src/gpk_transitions.cpp:17-35).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t mix64(uint64_t x) {
    x ^= x >> 30; x *= 0xbf58476d1ce4e5b9ULL;
    x ^= x >> 27; x *= 0x94d049bb133111ebULL;
    x ^= x >> 31;
    return x;
}
__device__ __forceinline__ float u01(uint64_t x) {
    const uint64_t m = (x >> 11) | 1ULL;
    return float(double(m & ((1ULL << 53) - 1)) * (1.0 / double(1ULL << 53)));
}
__device__ __forceinline__ float normal_dev(int i, uint64_t step, uint64_t salt) {
    const uint64_t a = mix64((uint64_t(i) + 0x9e3779b97f4a7c15ULL) ^ (step << 1) ^ salt);
    const uint64_t b = mix64((uint64_t(i) + 0xd2b74407b1ce6e93ULL) ^ (step << 3) ^ (salt + 17ULL));
    float u1 = u01(a);
    u1 = u1 < 1e-12f ? 1e-12f : u1;
    const float u2 = u01(b);
    return sqrtf(-2.f * logf(u1)) * cosf(2.f * 3.14159265358979323846f * u2);
}

} // namespace gpkd
