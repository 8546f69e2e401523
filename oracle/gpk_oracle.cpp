/*
 * gpk_oracle.cpp — CPU restatement of GPK's per-step kinetics.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP
 * library: only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg load it.  The product path (gpk_amd/, libgpk_hip.so) never links or
 * calls it.
 *
 * It restates the reference algorithm (paths relative to the GPK source tree)
 * in four arithmetic modes:
 *   mode 0  STRICT32  float fields, direct masked geometry stencil in the
 *                     reference's operation order: bit-exact with the
 *                     reference step (pinned against the compiled reference
 *                     and its CSV fixtures by tests/).
 *   mode 1  FAST32    float fields, separable geometry: window rows as
 *                     "same label" bit masks summed through chunk tables,
 *                     FMA over y and z; everything else as STRICT32.
 *   mode 2  FP64      double fields and weights, separable geometry.
 *   mode 3  STRICT64  double fields, direct geometry order.
 * The GPU library implements modes 0-2 with identical operation order, so
 * each GPU mode is compared bit for bit with the same oracle mode.
 *
 * Build: g++ -O2 -fopenmp -ffp-contract=off (no -march): the host compiler
 * must not contract a*b+c, matching the reference build
 * (CMakeLists.txt:7-28 has no -march/-ffast-math by default).
 */
#include "../include/gpk_c.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <queue>
#include <random>
#include <unordered_map>
#include <unordered_set>
#include <utility>
#include <vector>

#include <omp.h>

namespace {

// ---------------------------------------------------------------------------
// Mode constants.  Float modes use the reference's float literals; double
// modes use the same decimal values in double.
// ---------------------------------------------------------------------------
template <class R> struct K;
template <> struct K<float> {
    static constexpr float flip_eps = 1e-6f;       // src/gpk_transitions.cpp:292,310
    static constexpr float grad_eps = 1e-8f;       // src/gpk_geometry.cpp:140
    static constexpr float kappa_eps = 1e-6f;      // src/gpk_geometry.cpp:149
    static constexpr float gain_eps = 1e-7f;       // src/gpk_transitions.cpp:99
    static constexpr float vmax0 = 1e-30f;         // src/gpk_simulation.cpp:12
    static constexpr float big = 1e30f;            // src/gpk_simulation.cpp:23
    static constexpr float c_iface = 0.125f, c_diff = 0.4f;
    static constexpr float aniso_k = 1e-12f;       // include/gpk_thermo.hpp:261
    static constexpr float cl_floor = 1e-6f;       // src/gpk_transport.cpp:68,110
};
template <> struct K<double> {
    static constexpr double flip_eps = 1e-6;
    static constexpr double grad_eps = 1e-8;
    static constexpr double kappa_eps = 1e-6;
    static constexpr double gain_eps = 1e-7;
    static constexpr double vmax0 = 1e-30;
    static constexpr double big = 1e30;
    static constexpr double c_iface = 0.125, c_diff = 0.4;
    static constexpr double aniso_k = 1e-12;
    static constexpr double cl_floor = 1e-6;
};

inline float  fma_r(float a, float b, float c)   { return std::fmaf(a, b, c); }
inline double fma_r(double a, double b, double c) { return std::fma(a, b, c); }

// ---------------------------------------------------------------------------
// Normalised truncated Gaussian (include/gpk_core.hpp:37-55).  The radius is
// always computed in float so both precisions use the same windows.
// ---------------------------------------------------------------------------
template <class R> struct Gauss {
    std::vector<R> w;
    int radius = 0;
    float sigma_cells = 0.f;
};

Gauss<float> make_gauss_f(float sigma) {
    Gauss<float> k;
    k.sigma_cells = std::max(sigma, 0.5f);
    const float s = k.sigma_cells;
    k.radius = static_cast<int>(std::ceil(3.5f * s));
    const int n = 2 * k.radius + 1;
    k.w.resize(n);
    float sum = 0.f;
    for (int i = 0; i < n; ++i) {
        const float x = float(i - k.radius);
        k.w[i] = std::exp(-x * x / (2.f * s * s));
        sum += k.w[i];
    }
    for (auto& v : k.w) v /= sum;
    return k;
}

Gauss<double> make_gauss_d(float sigma) {
    Gauss<double> k;
    k.sigma_cells = std::max(sigma, 0.5f);
    const double s = k.sigma_cells;
    k.radius = static_cast<int>(std::ceil(3.5f * k.sigma_cells));
    const int n = 2 * k.radius + 1;
    k.w.resize(n);
    double sum = 0.0;
    for (int i = 0; i < n; ++i) {
        const double x = double(i - k.radius);
        k.w[i] = std::exp(-x * x / (2.0 * s * s));
        sum += k.w[i];
    }
    for (auto& v : k.w) v /= sum;
    return k;
}

template <class R> Gauss<R> make_gauss(float sigma);
template <> Gauss<float> make_gauss<float>(float s) { return make_gauss_f(s); }
template <> Gauss<double> make_gauss<double>(float s) { return make_gauss_d(s); }

// Gaussian convolved with 4th-order centred first/second difference stencils
// (src/gpk_geometry.cpp:12-36).
template <class R>
void geometry_kernels(const Gauss<R>& g, std::vector<R>& w0, std::vector<R>& w1,
                      std::vector<R>& w2, int& rg) {
    rg = g.radius + 2;
    const int width = 2 * rg + 1;
    w0.assign(width, R(0)); w1.assign(width, R(0)); w2.assign(width, R(0));
    const int o1[4] = {-2, -1, 1, 2};
    const int o2[5] = {-2, -1, 0, 1, 2};
    R c1[4], c2[5];
    if (sizeof(R) == sizeof(float)) {
        const float f1[4] = {1.f/12.f, -8.f/12.f, 8.f/12.f, -1.f/12.f};
        const float f2[5] = {-1.f/12.f, 16.f/12.f, -30.f/12.f, 16.f/12.f, -1.f/12.f};
        for (int k = 0; k < 4; ++k) c1[k] = R(f1[k]);
        for (int k = 0; k < 5; ++k) c2[k] = R(f2[k]);
    } else {
        const double d1[4] = {1.0/12.0, -8.0/12.0, 8.0/12.0, -1.0/12.0};
        const double d2[5] = {-1.0/12.0, 16.0/12.0, -30.0/12.0, 16.0/12.0, -1.0/12.0};
        for (int k = 0; k < 4; ++k) c1[k] = R(d1[k]);
        for (int k = 0; k < 5; ++k) c2[k] = R(d2[k]);
    }
    auto at = [&](int o) -> R {
        if (o < -g.radius || o > g.radius) return R(0);
        return g.w[o + g.radius];
    };
    for (int o = -rg; o <= rg; ++o) {
        const int j = o + rg;
        w0[j] = at(o);
        for (int k = 0; k < 4; ++k) w1[j] += c1[k] * at(o - o1[k]);
        for (int k = 0; k < 5; ++k) w2[j] += c2[k] * at(o - o2[k]);
    }
}

// ---------------------------------------------------------------------------
// Row-sum tables of the separable modes (FP32, FP64).  The "same label" bits
// of one window row are packed into m with bit dx set when window position dx
// (0 = leftmost) carries the resident label.  The row is cut into
// nch = ceil(W/9) chunks of C = ceil(W/nch) consecutive dx; chunk k's table
// holds, for each bit pattern, the sum of its three weight sets over the set
// bits in ascending dx order.  A row sum adds the chunk values in ascending
// chunk order.
// ---------------------------------------------------------------------------
template <class R>
struct RowLut {
    int W = 0, C = 0, nch = 0;
    int start[8], len[8], base[8];
    std::vector<R> t;   // base[k] + idx*3 + a
};

template <class R>
RowLut<R> make_row_lut(const std::vector<R>& w0, const std::vector<R>& w1, const std::vector<R>& w2) {
    RowLut<R> L;
    L.W = static_cast<int>(w0.size());
    L.nch = (L.W + 8) / 9;
    L.C = (L.W + L.nch - 1) / L.nch;
    int off = 0;
    for (int k = 0; k < L.nch; ++k) {
        L.start[k] = k * L.C;
        L.len[k] = std::min(L.C, L.W - L.start[k]);
        L.base[k] = off;
        off += (1 << L.len[k]) * 3;
    }
    L.t.assign(off, R(0));
    const std::vector<R>* w[3] = {&w0, &w1, &w2};
    for (int k = 0; k < L.nch; ++k)
        for (int idx = 0; idx < (1 << L.len[k]); ++idx)
            for (int a = 0; a < 3; ++a) {
                R sum = R(0);
                for (int j = 0; j < L.len[k]; ++j)
                    if ((idx >> j) & 1) sum += (*w[a])[L.start[k] + j];
                L.t[L.base[k] + idx * 3 + a] = sum;
            }
    return L;
}

// ---------------------------------------------------------------------------
// Constitutive laws (include/gpk_thermo.hpp:18-85, 229-325; src/gpk_thermo.cpp:9-30)
// ---------------------------------------------------------------------------
template <class R>
struct Law {
    gpk_model_desc d{};

    R driving_force(int32_t pp, int32_t pq, R cv, R Tv) const {
        switch (d.kind) {
        case GPK_MODEL_SOLIDIFICATION: {
            const R T_liq = R(d.T_melt) + R(d.m_liq) * cv;
            const R dT = T_liq - Tv;
            const R dG = (R(d.L) / R(d.T_melt)) * dT;
            if (pp == 0 && pq == 1) return -dG;
            if (pp == 1 && pq == 0) return +dG;
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
            return gibbs(pq, cv, Tv) - gibbs(pp, cv, Tv);
        default:
            return R(0);
        }
    }

    // include/gpk_thermo.hpp:358-360
    R gibbs(int q, R cv, R Tv) const {
        return R(d.pa[q]) + R(d.pb[q]) * cv + R(d.pc2[q]) * cv * cv + R(d.pdT[q]) * (Tv - R(d.T_ref));
    }
    // include/gpk_thermo.hpp:371-383
    int select_phase(int32_t pp, R cv, R Tv, const int32_t* cand, int nc) const {
        R best = R(1e30f);
        int best_q = -1;
        if (pp < 0 || pp >= d.n_phases) return -1;
        for (int k = 0; k < nc; ++k) {
            const int32_t q = cand[k];
            if (q < 0 || q >= d.n_phases || q == pp) continue;
            const R dG = gibbs(q, cv, Tv) - gibbs(pp, cv, Tv);
            if (dG < best) { best = dG; best_q = q; }
        }
        return best_q;
    }

    R interface_composition(int32_t pp, int32_t pq, R cp, R cq) const {
        if (d.kind == GPK_MODEL_ANISO_OSTWALD) {
            if (pp == 0) return cp;
            if (pq == 0) return cq;
            return cp;
        }
        return cp;
    }

    R energy(R nx, R ny, R nz, R gv) const {
        if (d.kind == GPK_MODEL_TENSOR_ANISO || d.kind == GPK_MODEL_ANISO_OSTWALD) {
            const R cg = std::cos(gv), sg = std::sin(gv);
            const R ncx = nx * cg + ny * sg;
            const R ncy = -nx * sg + ny * cg;
            const R ncz = nz;
            const R Kq = R(d.bxx) * ncx * ncx + R(d.byy) * ncy * ncy + R(d.bzz) * ncz * ncz;
            return R(d.sigma0) * std::sqrt(std::max(Kq, R(K<R>::aniso_k)));
        }
        if (d.eps4 == 0.f && d.eps6 == 0.f) return R(d.sigma0);
        const R cg = std::cos(gv), sg = std::sin(gv);
        const R ncx = nx * cg + ny * sg;
        const R ncy = -nx * sg + ny * cg;
        const R ncz = nz;
        if (ncz == R(0)) {
            const R th = std::atan2(ncy, ncx);
            R s = R(1) + R(d.eps4) * std::cos(R(4) * th);
            if (d.eps6 != 0.f) s += R(d.eps6) * std::cos(R(6) * th);
            return R(d.sigma0) * s;
        }
        const R aniso = ncx*ncx*ncx*ncx + ncy*ncy*ncy*ncy + ncz*ncz*ncz*ncz - R(3) / R(5);
        return R(d.sigma0) * (R(1) + R(d.eps4) * aniso);
    }
};

// Counter-based normal deviate (src/gpk_transitions.cpp:17-35).
inline uint64_t mix64(uint64_t x) {
    x ^= x >> 30; x *= 0xbf58476d1ce4e5b9ULL;
    x ^= x >> 27; x *= 0x94d049bb133111ebULL;
    x ^= x >> 31;
    return x;
}
inline float u01(uint64_t x) {
    const uint64_t m = (x >> 11) | 1ULL;
    return float((m & ((1ULL << 53) - 1)) * (1.0 / double(1ULL << 53)));
}
inline float normal_dev(int i, uint64_t step, uint64_t salt) {
    const uint64_t a = mix64((uint64_t(i) + 0x9e3779b97f4a7c15ULL) ^ (step << 1) ^ salt);
    const uint64_t b = mix64((uint64_t(i) + 0xd2b74407b1ce6e93ULL) ^ (step << 3) ^ (salt + 17ULL));
    const float u1 = std::max(1e-12f, u01(a));
    const float u2 = u01(b);
    return std::sqrt(-2.f * std::log(u1)) * std::cos(2.f * float(M_PI) * u2);
}

struct Stats {
    int64_t band = 0, candidates = 0, accepted = 0, deferred = 0, exchanges = 0, eroded = 0;
    // flip-transaction structure (GPU scheduling design data)
    int64_t flip_global = 0;        // candidates that needed a global BFS or count
    int64_t flip_depth = 0;         // longest chain of adjacent candidates in index order
    int64_t flip_ready0 = 0;        // candidates with no earlier candidate neighbour
    int64_t bfs_visits = 0;         // cells visited by global BFS
    int64_t bgr_proposals = 0, bgr_pairs = 0, bgr_depth = 0;
    int64_t guarded = 0;            // fp32 mode: decisions taken again in double
};

// ---------------------------------------------------------------------------
// The grid.
// ---------------------------------------------------------------------------
template <class R, bool kDirect>
struct Grid {
    int Nz, Ny, Nx, N;
    R dx, D;
    float dx_f, D_f, sigma_in;
    Law<R> law;
    Gauss<R> gauss;
    float vel_sigma, remap_sigma;
    std::vector<R> w0, w1, w2;
    int rg = 0;
    RowLut<R> lut;
    std::vector<double> proj_w, remap_w;   // weight by integer r^2
    int proj_r = 1, remap_r = 1;

    std::vector<int32_t> phi, phi_comp;
    std::vector<R> alpha, c, T, g, rho, nx, ny, nz, kappa, vn, g_comp, E_el;
    std::vector<uint8_t> mask;
    std::vector<int> active;
    std::vector<R> alpha_new, eps_star, eps_smooth, solid_old;
    std::vector<uint8_t> below;     // alpha_new at or below the flip threshold
    double time = 0.0;
    uint64_t step_id = 0;
    Stats st;

    Grid(int Nz_, int Ny_, int Nx_, float dx_, float D_, float sigma, const gpk_model_desc& m)
        : Nz(Nz_), Ny(Ny_), Nx(Nx_), N(Nz_ * Ny_ * Nx_), dx(R(dx_)), D(R(D_)),
          dx_f(dx_), D_f(D_), sigma_in(sigma) {
        law.d = m;
        gauss = make_gauss<R>(sigma);
        geometry_kernels(gauss, w0, w1, w2, rg);
        // row tables of the separable modes (64-bit window rows); the broad
        // 3-D path and the direct stencil do not use them
        if (2 * rg + 1 <= 64) lut = make_row_lut(w0, w1, w2);
        // src/gpk_grid.cpp:24-27 (GaussKernel clamps sigma to >= 0.5)
        vel_sigma = std::max(Nz > 1 ? std::clamp(0.375f * sigma, 0.85f, 3.00f) : 0.85f, 0.5f);
        remap_sigma = std::max(Nz > 1 ? std::clamp(0.125f * sigma, 0.70f, 1.00f) : 0.70f, 0.5f);
        proj_r = std::max(1, static_cast<int>(std::ceil(3.f * vel_sigma)));
        remap_r = std::max(1, static_cast<int>(std::ceil(3.f * remap_sigma)));
        proj_w = weight_table(vel_sigma, proj_r);
        remap_w = weight_table(remap_sigma, remap_r);
        phi.assign(N, 0); phi_comp.assign(N, 0);
        alpha.assign(N, R(1)); c.assign(N, R(0)); T.assign(N, R(800)); g.assign(N, R(0));
        rho.assign(N, R(0));
        nx.assign(N, R(0)); ny.assign(N, R(0)); nz.assign(N, R(0));
        kappa.assign(N, R(0)); vn.assign(N, R(0)); g_comp.assign(N, R(0)); E_el.assign(N, R(0));
        mask.assign(N, 0);
        alpha_new.assign(N, R(0));
        eps_star.assign(N, R(0)); eps_smooth.assign(N, R(0));
    }

    // exp(-r^2 / (2 s^2)) by integer r^2: float expf promoted to double in the
    // float modes (src/gpk_physics.cpp:184-185), exp in double otherwise.
    std::vector<double> weight_table(float s, int r) const {
        const int n = 3 * r * r + 1;
        std::vector<double> t(n);
        if (sizeof(R) == sizeof(float)) {
            const float inv = 0.5f / (s * s);
            for (int r2 = 0; r2 < n; ++r2) {
                const float r2f = static_cast<float>(r2);
                t[r2] = static_cast<double>(std::exp(-r2f * inv));
            }
        } else {
            const double sd = s;
            const double inv = 0.5 / (sd * sd);
            for (int r2 = 0; r2 < n; ++r2) t[r2] = std::exp(-double(r2) * inv);
        }
        return t;
    }

    inline int idx(int z, int y, int x) const { return (z * Ny + y) * Nx + x; }
    inline void decode(int i, int& z, int& y, int& x) const {
        x = i % Nx; y = (i / Nx) % Ny; z = i / (Ny * Nx);
    }
    // face neighbours in the reference order: x-, x+, y-, y+, z-, z+
    inline void faces(int i, int out[6]) const {
        int z, y, x; decode(i, z, y, x);
        const int xm = x == 0 ? Nx - 1 : x - 1, xp = x == Nx - 1 ? 0 : x + 1;
        const int ym = y == 0 ? Ny - 1 : y - 1, yp = y == Ny - 1 ? 0 : y + 1;
        const int zm = z == 0 ? Nz - 1 : z - 1, zp = z == Nz - 1 ? 0 : z + 1;
        out[0] = idx(z, y, xm); out[1] = idx(z, y, xp);
        out[2] = idx(z, ym, x); out[3] = idx(z, yp, x);
        out[4] = idx(zm, y, x); out[5] = idx(zp, y, x);
    }
    static inline int wrap(int v, int n) { while (v < 0) v += n; while (v >= n) v -= n; return v; }

    // -- src/gpk_geometry.cpp:42-82 ------------------------------------------
    void interface_mask() {
        #pragma omp parallel for schedule(static)
        for (int i = 0; i < N; ++i) {
            int f[6]; faces(i, f);
            const int32_t v = phi[i];
            mask[i] = uint8_t((v != phi[f[0]]) | (v != phi[f[1]]) | (v != phi[f[2]]) |
                              (v != phi[f[3]]) | (v != phi[f[4]]) | (v != phi[f[5]]));
        }
        active.clear();
        for (int i = 0; i < N; ++i) if (mask[i]) active.push_back(i);
    }

    // -- src/gpk_geometry.cpp:88-183 -------------------------------------------
    // -- src/gpk_geometry.cpp:185-333: 3-D grids whose geometry radius exceeds
    // 18 (src/gpk_grid.cpp:44).  Per interface label, in ascending order: the
    // 0/1 indicator smoothed by the plain Gaussian in x, then y, then z
    // (k-ordered sums, separate multiply and add), fourth-order centred
    // differences at the label's interface cells, normal from 1/|grad I|.
    bool broad() const { return Nz > 1 && rg > 18; }
    void broad_geometry() {
        std::vector<int32_t> labs;
        for (int i = 0; i < N; ++i) if (mask[i]) labs.push_back(phi[i]);
        std::sort(labs.begin(), labs.end());
        labs.erase(std::unique(labs.begin(), labs.end()), labs.end());
        std::fill(nx.begin(), nx.end(), R(0));
        std::fill(ny.begin(), ny.end(), R(0));
        std::fill(nz.begin(), nz.end(), R(0));
        std::fill(kappa.begin(), kappa.end(), R(0));
        const int kr = gauss.radius;
        const R idx12 = R(1) / (R(12) * dx);
        const R idx_sq = R(1) / (dx * dx);
        const R idx12_sq = idx_sq / R(12);
        const R idx144_sq = idx_sq / R(144);
        const R klim = R(1) / (R(gauss.sigma_cells) * dx);
        std::vector<R> t1(N), t2(N), sm(N);
        for (const int32_t p : labs) {
            #pragma omp parallel for schedule(static)
            for (int zy = 0; zy < Nz * Ny; ++zy) {
                const int base = zy * Nx;
                for (int x = 0; x < Nx; ++x) {
                    R acc = R(0);
                    for (int k = 0; k <= 2 * kr; ++k) {
                        const int xx = wrap(x + k - kr, Nx);
                        acc += gauss.w[k] * (phi[base + xx] == p ? R(1) : R(0));
                    }
                    t1[base + x] = acc;
                }
            }
            #pragma omp parallel for schedule(static)
            for (int zy = 0; zy < Nz * Ny; ++zy) {
                const int z = zy / Ny, y = zy % Ny;
                for (int x = 0; x < Nx; ++x) {
                    R acc = R(0);
                    for (int k = 0; k <= 2 * kr; ++k)
                        acc += gauss.w[k] * t1[idx(z, wrap(y + k - kr, Ny), x)];
                    t2[idx(z, y, x)] = acc;
                }
            }
            #pragma omp parallel for schedule(static)
            for (int zy = 0; zy < Nz * Ny; ++zy) {
                const int z = zy / Ny, y = zy % Ny;
                for (int x = 0; x < Nx; ++x) {
                    R acc = R(0);
                    for (int k = 0; k <= 2 * kr; ++k)
                        acc += gauss.w[k] * t2[idx(wrap(z + k - kr, Nz), y, x)];
                    sm[idx(z, y, x)] = acc;
                }
            }
            const int na = static_cast<int>(active.size());
            #pragma omp parallel for schedule(static)
            for (int a = 0; a < na; ++a) {
                const int i = active[a];
                if (phi[i] != p) continue;
                int z, y, x; decode(i, z, y, x);
                const int xo[4] = {wrap(x - 2, Nx), wrap(x - 1, Nx), wrap(x + 1, Nx), wrap(x + 2, Nx)};
                const int yo[4] = {wrap(y - 2, Ny), wrap(y - 1, Ny), wrap(y + 1, Ny), wrap(y + 2, Ny)};
                const int zo[4] = {wrap(z - 2, Nz), wrap(z - 1, Nz), wrap(z + 1, Nz), wrap(z + 2, Nz)};
                auto S = [&](int zz, int yy, int xx) { return sm[idx(zz, yy, xx)]; };
                const R I0 = sm[i];
                const R Ix = (S(z, y, xo[0]) - R(8) * S(z, y, xo[1]) + R(8) * S(z, y, xo[2]) - S(z, y, xo[3])) * idx12;
                const R Iy = (S(z, yo[0], x) - R(8) * S(z, yo[1], x) + R(8) * S(z, yo[2], x) - S(z, yo[3], x)) * idx12;
                const R Iz = (S(zo[0], y, x) - R(8) * S(zo[1], y, x) + R(8) * S(zo[2], y, x) - S(zo[3], y, x)) * idx12;
                const R Ixx = (-S(z, y, xo[3]) + R(16) * S(z, y, xo[2]) - R(30) * I0 + R(16) * S(z, y, xo[1])
                               - S(z, y, xo[0])) * idx12_sq;
                const R Iyy = (-S(z, yo[3], x) + R(16) * S(z, yo[2], x) - R(30) * I0 + R(16) * S(z, yo[1], x)
                               - S(z, yo[0], x)) * idx12_sq;
                const R Izz = (-S(zo[3], y, x) + R(16) * S(zo[2], y, x) - R(30) * I0 + R(16) * S(zo[1], y, x)
                               - S(zo[0], y, x)) * idx12_sq;
                const R dc[4] = {R(1), R(-8), R(8), R(-1)};
                R sxy = 0, sxz = 0, syz = 0;
                for (int aa = 0; aa < 4; ++aa)
                    for (int bb = 0; bb < 4; ++bb) {
                        const R w = dc[aa] * dc[bb];
                        sxy += w * S(z, yo[bb], xo[aa]);
                        sxz += w * S(zo[bb], y, xo[aa]);
                        syz += w * S(zo[bb], yo[aa], x);
                    }
                const R Ixy = sxy * idx144_sq, Ixz = sxz * idx144_sq, Iyz = syz * idx144_sq;
                const R gsq = Ix * Ix + Iy * Iy + Iz * Iz;
                if (gsq * dx * dx > R(K<R>::grad_eps)) {
                    const R inv = R(1) / std::sqrt(gsq);
                    nx[i] = -Ix * inv; ny[i] = -Iy * inv; nz[i] = -Iz * inv;
                    const R trH = Ixx + Iyy + Izz;
                    const R gHg = Ix * Ix * Ixx + Iy * Iy * Iyy + Iz * Iz * Izz
                                + R(2) * (Ix * Iy * Ixy + Ix * Iz * Ixz + Iy * Iz * Iyz);
                    R kap = (gHg - gsq * trH) / (gsq * std::sqrt(gsq));
                    if (std::fabs(kap * dx) < R(K<R>::kappa_eps)) kap = R(0);
                    kappa[i] = std::clamp(kap, -klim, klim);
                }
            }
        }
        const int na = static_cast<int>(active.size());
        #pragma omp parallel for schedule(static)
        for (int a = 0; a < na; ++a) {
            const int i = active[a];
            phi_comp[i] = competitor(i);
        }
    }

    void geometry() {
        if (broad()) {
            broad_geometry();
            return;
        }
        std::fill(nx.begin(), nx.end(), R(0));
        std::fill(ny.begin(), ny.end(), R(0));
        std::fill(nz.begin(), nz.end(), R(0));
        std::fill(kappa.begin(), kappa.end(), R(0));
        const int r = rg;
        const R inv_dx = R(1) / dx;
        const R inv_dx2 = inv_dx * inv_dx;
        const R klim = R(1) / (R(gauss.sigma_cells) * dx);
        const int na = static_cast<int>(active.size());
        #pragma omp parallel for schedule(static)
        for (int a = 0; a < na; ++a) {
            const int i = active[a];
            const int32_t res = phi[i];
            int z0, y0, x0; decode(i, z0, y0, x0);
            R Ix = 0, Iy = 0, Iz = 0, Ixx = 0, Iyy = 0, Izz = 0, Ixy = 0, Ixz = 0, Iyz = 0;
            if (kDirect) {
                const int dzb = Nz > 1 ? -r : 0, dze = Nz > 1 ? r : 0;
                for (int dz = dzb; dz <= dze; ++dz) {
                    const int z = wrap(z0 + dz, Nz);
                    const int jz = dz + r;
                    const R gz0 = Nz > 1 ? w0[jz] : R(1);
                    const R gz1 = Nz > 1 ? w1[jz] : R(0);
                    const R gz2 = Nz > 1 ? w2[jz] : R(0);
                    for (int dy = -r; dy <= r; ++dy) {
                        const int y = wrap(y0 + dy, Ny);
                        const int jy = dy + r;
                        const R gy0 = w0[jy], gy1 = w1[jy], gy2 = w2[jy];
                        for (int ddx = -r; ddx <= r; ++ddx) {
                            const int x = wrap(x0 + ddx, Nx);
                            if (phi[idx(z, y, x)] != res) continue;
                            const int jx = ddx + r;
                            const R gx0 = w0[jx], gx1 = w1[jx], gx2 = w2[jx];
                            Ix += gx1 * gy0 * gz0; Iy += gx0 * gy1 * gz0;
                            Ixx += gx2 * gy0 * gz0; Iyy += gx0 * gy2 * gz0; Ixy += gx1 * gy1 * gz0;
                            if (Nz > 1) {
                                Iz += gx0 * gy0 * gz1; Izz += gx0 * gy0 * gz2;
                                Ixz += gx1 * gy0 * gz1; Iyz += gx0 * gy1 * gz1;
                            }
                        }
                    }
                }
            } else {
                // separable: masked row sums over x, FMA over y, FMA over z
                if (lut.nch == 0) std::abort();   // window rows wider than 64 cells
                const int dzb = Nz > 1 ? -r : 0, dze = Nz > 1 ? r : 0;
                for (int dz = dzb; dz <= dze; ++dz) {
                    const int z = wrap(z0 + dz, Nz);
                    R S00 = 0, S01 = 0, S02 = 0, S10 = 0, S11 = 0, S20 = 0;
                    for (int dy = -r; dy <= r; ++dy) {
                        const int y = wrap(y0 + dy, Ny);
                        const int jy = dy + r;
                        uint64_t m = 0;
                        for (int ddx = -r; ddx <= r; ++ddx) {
                            const int x = wrap(x0 + ddx, Nx);
                            if (phi[idx(z, y, x)] == res) m |= 1ull << (ddx + r);
                        }
                        R R0 = 0, R1 = 0, R2 = 0;
                        for (int k = 0; k < lut.nch; ++k) {
                            const int sh = lut.start[k];
                            const int ix = static_cast<int>((m >> sh) & ((1ull << lut.len[k]) - 1));
                            const R* t = lut.t.data() + lut.base[k] + ix * 3;
                            if (k == 0) { R0 = t[0]; R1 = t[1]; R2 = t[2]; }
                            else { R0 += t[0]; R1 += t[1]; R2 += t[2]; }
                        }
                        S00 = fma_r(w0[jy], R0, S00); S01 = fma_r(w0[jy], R1, S01);
                        S02 = fma_r(w0[jy], R2, S02); S10 = fma_r(w1[jy], R0, S10);
                        S11 = fma_r(w1[jy], R1, S11); S20 = fma_r(w2[jy], R0, S20);
                    }
                    if (Nz > 1) {
                        const int jz = dz + r;
                        const R gz0 = w0[jz], gz1 = w1[jz], gz2 = w2[jz];
                        Ix = fma_r(gz0, S01, Ix); Iy = fma_r(gz0, S10, Iy);
                        Ixx = fma_r(gz0, S02, Ixx); Iyy = fma_r(gz0, S20, Iyy);
                        Ixy = fma_r(gz0, S11, Ixy);
                        Iz = fma_r(gz1, S00, Iz); Izz = fma_r(gz2, S00, Izz);
                        Ixz = fma_r(gz1, S01, Ixz); Iyz = fma_r(gz1, S10, Iyz);
                    } else {
                        Ix = S01; Iy = S10; Ixx = S02; Iyy = S20; Ixy = S11;
                    }
                }
            }
            Ix *= inv_dx; Iy *= inv_dx; Iz *= inv_dx;
            Ixx *= inv_dx2; Iyy *= inv_dx2; Izz *= inv_dx2;
            Ixy *= inv_dx2; Ixz *= inv_dx2; Iyz *= inv_dx2;
            const R gsq = Ix * Ix + Iy * Iy + Iz * Iz;
            if (gsq * dx * dx > R(K<R>::grad_eps)) {
                const R gr = std::sqrt(gsq);
                nx[i] = -Ix / gr; ny[i] = -Iy / gr; nz[i] = -Iz / gr;
                const R trH = Ixx + Iyy + Izz;
                const R gHg = Ix * Ix * Ixx + Iy * Iy * Iyy + Iz * Iz * Izz
                            + R(2) * (Ix * Iy * Ixy + Ix * Iz * Ixz + Iy * Iz * Iyz);
                R kap = (gHg - gsq * trH) / (gsq * gr);
                if (std::fabs(kap * dx) < R(K<R>::kappa_eps)) kap = R(0);
                kappa[i] = std::clamp(kap, -klim, klim);
            }
        }
        // competitor: most face-supported unlike label (src/gpk_geometry.cpp:154-182)
        #pragma omp parallel for schedule(static)
        for (int a = 0; a < na; ++a) {
            const int i = active[a];
            phi_comp[i] = competitor(i);
        }
    }

    int32_t competitor(int i) const {
        int z, y, x; decode(i, z, y, x);
        int f[6]; faces(i, f);
        const R score[6] = {-nx[i], nx[i], -ny[i], ny[i], -nz[i], nz[i]};
        const int nface = Nz > 1 ? 6 : 4;
        int best = -1, best_support = -1; R best_score = R(-2);
        for (int k = 0; k < nface; ++k) {
            const int32_t q = phi[f[k]];
            if (q == phi[i]) continue;
            int support = 0;
            for (int dz = (Nz > 1 ? -1 : 0); dz <= (Nz > 1 ? 1 : 0); ++dz)
            for (int dy = -1; dy <= 1; ++dy)
            for (int ddx = -1; ddx <= 1; ++ddx) {
                if (ddx == 0 && dy == 0 && dz == 0) continue;
                int xx = x + ddx; if (xx < 0) xx += Nx; else if (xx >= Nx) xx -= Nx;
                int yy = y + dy; if (yy < 0) yy += Ny; else if (yy >= Ny) yy -= Ny;
                int zz = z + dz; if (zz < 0) zz += Nz; else if (zz >= Nz) zz -= Nz;
                support += phi[idx(zz, yy, xx)] == q;
            }
            if (support > best_support || (support == best_support && score[k] > best_score)) {
                best = k; best_support = support; best_score = score[k];
            }
        }
        return best >= 0 ? phi[f[best]] : phi[i];
    }

    // -- src/gpk_physics.cpp:67-148 ---------------------------------------------
    void velocities() {
        std::fill(vn.begin(), vn.end(), R(0));
        const int na = static_cast<int>(active.size());
        #pragma omp parallel for schedule(static)
        for (int a = 0; a < na; ++a) {
            const int i = active[a];
            int f[6]; faces(i, f);
            // src/gpk_physics.cpp:79-95
            if (law.d.kind == GPK_MODEL_PRECIPITATION) {
                int32_t cand[6]; int nc = 0;
                for (int k = 0; k < 6; ++k) {
                    const int32_t q = phi[f[k]];
                    if (q == phi[i]) continue;
                    bool seen = false;
                    for (int t = 0; t < nc; ++t) if (cand[t] == q) { seen = true; break; }
                    if (!seen) cand[nc++] = q;
                }
                const int sel = law.select_phase(phi[i], c[i], T[i], cand, nc);
                if (sel >= 0) phi_comp[i] = sel;
            }
            const int32_t q = phi_comp[i];
            const R gp = g[i];
            R gq = gp, cq = c[i];
            for (int k = 0; k < 6; ++k)
                if (phi[f[k]] == q) { gq = g[f[k]]; cq = c[f[k]]; break; }
            g_comp[i] = gq;
            const R ci = law.interface_composition(phi[i], q, c[i], cq);
            R dG = law.driving_force(phi[i], q, ci, T[i]);
            const R sig = law.energy(nx[i], ny[i], nz[i], gp);
            const R mob = R(law.d.M0);
            // src/gpk_physics.cpp:122-127
            if (law.d.flags & GPK_FLAG_ELASTIC) {
                R ep = E_el[i];
                for (int k = 0; k < 6; ++k)
                    if (phi[f[k]] == q) { ep = E_el[f[k]]; break; }
                dG += (ep - E_el[i]);
            }
            // src/gpk_physics.cpp:131-140
            if (law.d.flags & GPK_FLAG_DISLOCATION) {
                R rp = rho[i];
                for (int k = 0; k < 6; ++k)
                    if (phi[f[k]] == q) { rp = rho[f[k]]; break; }
                dG += R(law.d.disloc_coeff) * (rp - rho[i]);
            }
            vn[i] = mob * (dG - sig * kappa[i]);
        }
    }

    // -- src/gpk_physics.cpp:152-193 --------------------------------------------
    void projection() {
        std::vector<R> out(vn);
        const int r = proj_r;
        const int na = static_cast<int>(active.size());
        #pragma omp parallel for schedule(static)
        for (int a = 0; a < na; ++a) {
            const int i = active[a];
            if (phi[i] == phi_comp[i]) continue;
            const int32_t lo = std::min(phi[i], phi_comp[i]);
            const int32_t hi = std::max(phi[i], phi_comp[i]);
            const R oi = phi[i] == lo ? R(1) : R(-1);
            int z0, y0, x0; decode(i, z0, y0, x0);
            double num = 0.0, den = 0.0;
            const int dzb = Nz > 1 ? -r : 0, dze = Nz > 1 ? r : 0;
            for (int dz = dzb; dz <= dze; ++dz)
            for (int dy = -r; dy <= r; ++dy)
            for (int ddx = -r; ddx <= r; ++ddx) {
                int z = (z0 + dz) % Nz; if (z < 0) z += Nz;
                int y = (y0 + dy) % Ny; if (y < 0) y += Ny;
                int x = (x0 + ddx) % Nx; if (x < 0) x += Nx;
                const int j = idx(z, y, x);
                if (!mask[j] || phi[j] == phi_comp[j] ||
                    std::min(phi[j], phi_comp[j]) != lo ||
                    std::max(phi[j], phi_comp[j]) != hi) continue;
                const double w = proj_w[ddx * ddx + dy * dy + dz * dz];
                const R oj = phi[j] == lo ? R(1) : R(-1);
                num += w * oj * vn[j];
                den += w;
            }
            if (den > 0.0) out[i] = oi * static_cast<R>(num / den);
        }
        vn.swap(out);
    }

    // -- src/gpk_simulation.cpp:11-25 ---------------------------------------------
    R max_dt() const {
        R vmax = R(K<R>::vmax0);
        for (int i : active) vmax = std::max(vmax, std::fabs(vn[i]));
        const R nb = Nz > 1 ? R(6) : R(4);
        const R dti = R(K<R>::c_iface) * dx / vmax;
        const R dtd = D > R(0) ? R(K<R>::c_diff) * dx * dx / (nb * D) : R(K<R>::big);
        return std::min(dti, dtd);
    }

    // -- src/gpk_connectivity.cpp:10-195 ------------------------------------------
    int topo_number(int cell, int32_t label) const {
        int z0, y0, x0; decode(cell, z0, y0, x0);
        const int lzc = Nz > 1 ? 3 : 1, lzm = Nz > 1 ? 1 : 0;
        const int centre = lzm * 9 + 4;
        bool occ[27] = {}, seen[27] = {};
        for (int lz = 0; lz < lzc; ++lz) {
            int z = z0 + (Nz > 1 ? lz - 1 : 0); if (z < 0) z += Nz; else if (z >= Nz) z -= Nz;
            for (int ly = 0; ly < 3; ++ly) {
                int y = y0 + ly - 1; if (y < 0) y += Ny; else if (y >= Ny) y -= Ny;
                for (int lx = 0; lx < 3; ++lx) {
                    int x = x0 + lx - 1; if (x < 0) x += Nx; else if (x >= Nx) x -= Nx;
                    occ[lz * 9 + ly * 3 + lx] = phi[idx(z, y, x)] == label;
                }
            }
        }
        occ[centre] = false;
        int ncomp = 0;
        for (int s = 0; s < 27; ++s) {
            if (s / 9 >= lzc || !occ[s] || seen[s]) continue;
            int q[27], h = 0, t = 0;
            q[t++] = s; seen[s] = true;
            bool touches = false;
            while (h < t) {
                const int cur = q[h++];
                const int cz = cur / 9, cy = (cur % 9) / 3, cx = cur % 3;
                if (std::abs(cz - lzm) + std::abs(cy - 1) + std::abs(cx - 1) == 1) touches = true;
                const int dz6[6] = {0, 0, 0, 0, -1, 1}, dy6[6] = {0, 0, -1, 1, 0, 0},
                          dx6[6] = {-1, 1, 0, 0, 0, 0};
                for (int k = 0; k < 6; ++k) {
                    const int zz = cz + dz6[k], yy = cy + dy6[k], xx = cx + dx6[k];
                    if (zz < 0 || zz >= lzc || yy < 0 || yy >= 3 || xx < 0 || xx >= 3) continue;
                    const int nxt = zz * 9 + yy * 3 + xx;
                    if (occ[nxt] && !seen[nxt]) { seen[nxt] = true; q[t++] = nxt; }
                }
            }
            if (touches) ++ncomp;
        }
        return ncomp;
    }

    bool global_connected(int cell, int32_t label, bool excl, int64_t* visits) const {
        int f[6]; faces(cell, f);
        const int nface = Nz > 1 ? 6 : 4;
        std::vector<int> br;
        for (int k = 0; k < nface; ++k) if (phi[f[k]] == label) br.push_back(f[k]);
        if (br.size() <= 1) return true;
        std::unordered_set<int> vis;
        std::unordered_set<int> targets(br.begin() + 1, br.end());
        std::vector<int> stack{br.front()};
        vis.insert(br.front());
        while (!stack.empty()) {
            const int cur = stack.back(); stack.pop_back();
            if (visits) ++*visits;
            int nf[6]; faces(cur, nf);
            for (int k = 0; k < nface; ++k) {
                const int nxt = nf[k];
                if (vis.count(nxt) || phi[nxt] != label || (excl && nxt == cell)) continue;
                vis.insert(nxt);
                if (targets.erase(nxt) > 0 && targets.empty()) return true;
                stack.push_back(nxt);
            }
        }
        return targets.empty();
    }

    bool bounded_connected(int cell, int32_t label, bool excl) const {
        int z0, y0, x0; decode(cell, z0, y0, x0);
        const int zc = Nz > 1 ? 5 : 1, zm = Nz > 1 ? 2 : 0;
        const int lc = zm * 25 + 2 * 5 + 2;
        bool occ[125] = {}, seen[125] = {};
        for (int lz = 0; lz < zc; ++lz) {
            const int z = wrap(z0 + (Nz > 1 ? lz - 2 : 0), Nz);
            for (int ly = 0; ly < 5; ++ly) {
                const int y = wrap(y0 + ly - 2, Ny);
                for (int lx = 0; lx < 5; ++lx) {
                    const int x = wrap(x0 + lx - 2, Nx);
                    occ[lz * 25 + ly * 5 + lx] = phi[idx(z, y, x)] == label;
                }
            }
        }
        if (excl) occ[lc] = false;
        const int bl[6] = {lc - 1, lc + 1, lc - 5, lc + 5, lc - 25, lc + 25};
        const int nface = Nz > 1 ? 6 : 4;
        int br[6], nb = 0;
        for (int k = 0; k < nface; ++k) if (occ[bl[k]]) br[nb++] = bl[k];
        if (nb <= 1) return true;
        int q[125], h = 0, t = 0;
        q[t++] = br[0]; seen[br[0]] = true;
        const int dz6[6] = {0, 0, 0, 0, -1, 1}, dy6[6] = {0, 0, -1, 1, 0, 0}, dx6[6] = {-1, 1, 0, 0, 0, 0};
        while (h < t) {
            const int cur = q[h++];
            const int cz = cur / 25, cy = (cur % 25) / 5, cx = cur % 5;
            for (int k = 0; k < nface; ++k) {
                const int zz = cz + dz6[k], yy = cy + dy6[k], xx = cx + dx6[k];
                if (zz < 0 || zz >= zc || yy < 0 || yy >= 5 || xx < 0 || xx >= 5) continue;
                const int nxt = zz * 25 + yy * 5 + xx;
                if (occ[nxt] && !seen[nxt]) { seen[nxt] = true; q[t++] = nxt; }
            }
        }
        for (int k = 1; k < nb; ++k) if (!seen[br[k]]) return false;
        return true;
    }

    // returns decision; *global set if the decision needed global information
    bool simple_move(int cell, int32_t nl, int oc, int ncnt, bool local_only,
                     bool* global = nullptr, int64_t* visits = nullptr) const {
        (void)ncnt;
        if (cell < 0 || cell >= N) return false;
        const int32_t ol = phi[cell];
        if (ol == nl) return false;
        const int oo = topo_number(cell, ol);
        const bool ext = oc <= 1;
        if (global && (oo == 0 || (oo > 1 && !local_only))) *global = true;
        if (oo == 0 && !ext) return false;
        if (oo > 1 && !(local_only ? bounded_connected(cell, ol, true)
                                   : global_connected(cell, ol, true, visits))) return false;
        const int no = topo_number(cell, nl);
        if (global && no > 1 && !local_only) *global = true;
        if (no == 0) return false;
        if (no > 1 && !(local_only ? bounded_connected(cell, nl, false)
                                   : global_connected(cell, nl, false, visits))) return false;
        return true;
    }

    std::unordered_map<int32_t, int> histogram() const {
        std::unordered_map<int32_t, int> h;
        for (int32_t v : phi) ++h[v];
        return h;
    }

    // longest chain bookkeeping: level of a cell = 1 + max level of earlier
    // marked cells within Chebyshev distance `reach`
    int chain_level(int cell, int reach, const std::unordered_map<int, int>& lev) const {
        int z0, y0, x0; decode(cell, z0, y0, x0);
        int best = 0;
        const int rz = Nz > 1 ? reach : 0;
        for (int dz = -rz; dz <= rz; ++dz)
        for (int dy = -reach; dy <= reach; ++dy)
        for (int ddx = -reach; ddx <= reach; ++ddx) {
            const int j = idx(wrap(z0 + dz, Nz), wrap(y0 + dy, Ny), wrap(x0 + ddx, Nx));
            auto it = lev.find(j);
            if (it != lev.end()) best = std::max(best, it->second);
        }
        return best + 1;
    }

    // -- src/gpk_transitions.cpp:38-178 -------------------------------------------
    int64_t balanced_remap(int64_t accepted_last) {
        int64_t exchanges = 0;
        if (accepted_last == 0) return 0;
        interface_mask();
        std::fill(eps_star.begin(), eps_star.end(), -std::numeric_limits<R>::infinity());
        std::fill(eps_smooth.begin(), eps_smooth.end(), -std::numeric_limits<R>::infinity());
        std::copy(phi.begin(), phi.end(), phi_comp.begin());
        const int r = remap_r;
        const int na = static_cast<int>(active.size());
        #pragma omp parallel for schedule(static)
        for (int a = 0; a < na; ++a) {
            const int i = active[a];
            int z0, y0, x0; decode(i, z0, y0, x0);
            int f[6]; faces(i, f);
            const int nface = Nz > 1 ? 6 : 4;
            int32_t cand[6]; int nc = 0;
            for (int k = 0; k < nface; ++k) {
                const int32_t v = phi[f[k]];
                if (v == phi[i]) continue;
                bool known = false; for (int q = 0; q < nc; ++q) known |= cand[q] == v;
                if (!known) cand[nc++] = v;
            }
            double resident = 0.0, best = -1.0; int32_t best_label = phi[i];
            double sup[6] = {0, 0, 0, 0, 0, 0};
            const int dzb = Nz > 1 ? -r : 0, dze = Nz > 1 ? r : 0;
            for (int dz = dzb; dz <= dze; ++dz)
            for (int dy = -r; dy <= r; ++dy)
            for (int ddx = -r; ddx <= r; ++ddx) {
                int z = (z0 + dz) % Nz; if (z < 0) z += Nz;
                int y = (y0 + dy) % Ny; if (y < 0) y += Ny;
                int x = (x0 + ddx) % Nx; if (x < 0) x += Nx;
                const int32_t v = phi[idx(z, y, x)];
                const double w = remap_w[ddx * ddx + dy * dy + dz * dz];
                if (v == phi[i]) resident += w;
                for (int q = 0; q < nc; ++q) if (v == cand[q]) sup[q] += w;
            }
            eps_star[i] = static_cast<R>(resident);
            for (int q = 0; q < nc; ++q) if (sup[q] > best) { best = sup[q]; best_label = cand[q]; }
            eps_smooth[i] = static_cast<R>(best);
            phi_comp[i] = best_label;
        }

        struct Prop { R gain; int cell; };
        struct Lists { std::vector<Prop> fwd, rev; };
        std::map<std::pair<int32_t, int32_t>, Lists> grouped;
        for (int i : active) {
            const int32_t p = phi[i], q = phi_comp[i];
            const R gain = eps_smooth[i] - eps_star[i];
            if (p == q || !(gain > R(K<R>::gain_eps))) continue;
            const auto key = std::make_pair(std::min(p, q), std::max(p, q));
            if (p == key.first) grouped[key].fwd.push_back({gain, i});
            else grouped[key].rev.push_back({gain, i});
            ++st.bgr_proposals;
        }
        auto rank = [](const Prop& a, const Prop& b) {
            if (a.gain != b.gain) return a.gain > b.gain;
            return a.cell < b.cell;
        };
        auto counts = histogram();
        const int nface = Nz > 1 ? 6 : 4;
        auto target_g = [&](int cell, int32_t tgt) -> R {
            int f[6]; faces(cell, f);
            for (int k = 0; k < nface; ++k) if (phi[f[k]] == tgt) return g[f[k]];
            return g[cell];
        };
        auto mismatch = [&](int cell, int32_t prop, int pc, int32_t pp) {
            int f[6]; faces(cell, f);
            int m = 0;
            for (int k = 0; k < nface; ++k) {
                const int32_t q = f[k] == pc ? pp : phi[f[k]];
                m += q != prop;
            }
            return m;
        };
        std::unordered_map<int, int> lev;
        for (auto& item : grouped) {
            auto& fw = item.second.fwd;
            auto& rv = item.second.rev;
            std::sort(fw.begin(), fw.end(), rank);
            std::sort(rv.begin(), rv.end(), rank);
            const size_t cnt = std::min(fw.size(), rv.size());
            for (size_t k = 0; k < cnt; ++k) {
                const int a = fw[k].cell, b = rv[k].cell;
                const int32_t lo = item.first.first, hi = item.first.second;
                ++st.bgr_pairs;
                {
                    const int la = chain_level(a, 2, lev), lb = chain_level(b, 2, lev);
                    const int l = std::max(la, lb);
                    lev[a] = l; lev[b] = l;
                    st.bgr_depth = std::max<int64_t>(st.bgr_depth, l);
                }
                if (phi[a] != lo || phi[b] != hi) continue;
                const R ga = target_g(a, hi);
                const R gb = target_g(b, lo);
                const int before = mismatch(a, lo, b, hi) + mismatch(b, hi, a, lo);
                const int after = mismatch(a, hi, b, lo) + mismatch(b, lo, a, hi);
                if (after >= before) continue;
                if (!simple_move(a, hi, counts[lo], counts[hi], true)) continue;
                phi[a] = hi;
                if (!simple_move(b, lo, counts[hi] + 1, counts[lo] - 1, true)) {
                    phi[a] = lo;
                    continue;
                }
                phi[b] = lo; g[a] = ga; g[b] = gb; alpha[a] = alpha[b] = R(1);
                if (law.d.flags & GPK_FLAG_DISLOCATION) rho[a] = rho[b] = R(law.d.rho_rex);
                exchanges += 2;
            }
        }
        interface_mask();
        return exchanges;
    }

    // -- src/gpk_transitions.cpp:180-249 --------------------------------------------
    int64_t complete_unresolved() {
        if (!(law.d.flags & GPK_FLAG_GRAIN_NETWORK)) return 0;
        if ((step_id & 7ULL) != 0ULL) return 0;
        auto counts = histogram();
        const double s = std::max(gauss.sigma_cells, 1.f);
        const double vol = Nz > 1 ? (4.0 * M_PI * s * s * s / 3.0) : (M_PI * s * s);
        const int min_cells = std::max(2, static_cast<int>(std::ceil(vol)));
        std::unordered_set<int32_t> unres;
        for (const auto& it : counts) if (it.second < min_cells) unres.insert(it.first);
        if (unres.empty() || unres.size() == counts.size()) return 0;
        const int nface = Nz > 1 ? 6 : 4;
        int64_t eroded = 0;
        while (true) {
            std::vector<int> cells; std::vector<int32_t> tg; std::vector<R> tgg;
            for (int i = 0; i < N; ++i) {
                if (!unres.count(phi[i])) continue;
                int z, y, x; decode(i, z, y, x);
                int f[6]; faces(i, f);
                int best = -1, bs = -1;
                for (int k = 0; k < nface; ++k) {
                    const int32_t q = phi[f[k]];
                    if (unres.count(q)) continue;
                    int sup = 0;
                    for (int dz = (Nz > 1 ? -1 : 0); dz <= (Nz > 1 ? 1 : 0); ++dz)
                    for (int dy = -1; dy <= 1; ++dy)
                    for (int ddx = -1; ddx <= 1; ++ddx) {
                        const int zz = (z + dz + Nz) % Nz, yy = (y + dy + Ny) % Ny, xx = (x + ddx + Nx) % Nx;
                        sup += phi[idx(zz, yy, xx)] == q;
                    }
                    if (sup > bs) { best = k; bs = sup; }
                }
                if (best >= 0) { cells.push_back(i); tg.push_back(phi[f[best]]); tgg.push_back(g[f[best]]); }
            }
            if (cells.empty()) break;
            for (size_t k = 0; k < cells.size(); ++k) {
                phi[cells[k]] = tg[k]; g[cells[k]] = tgg[k]; alpha[cells[k]] = R(1);
                if (law.d.flags & GPK_FLAG_DISLOCATION) rho[cells[k]] = R(law.d.rho_rex);
            }
            eroded += static_cast<int64_t>(cells.size());
        }
        interface_mask();
        return eroded;
    }

    // -- src/gpk_transitions.cpp:251-326 ---------------------------------------------
    void alpha_and_flips(R dt) {
        const R idt = dt / dx;
        const bool noise = (law.d.flags & GPK_FLAG_STOCHASTIC) && dt > R(0);
        R amp = R(0);
        if (noise) {
            const float cv = dx_f * dx_f * dx_f;
            const float inv_dt = float(dt) > 0.f ? 1.f / float(dt) : 0.f;
            amp = R(law.d.noise_scale * std::sqrt(2.f * law.d.M0 * law.d.kB * inv_dt
                                                   / std::max(cv, 1e-30f)));
        }
        std::copy(alpha.begin(), alpha.end(), alpha_new.begin());
        const int na = static_cast<int>(active.size());
        // fp32 mode: an FP32 alpha_new within flip_guard of the threshold is
        // recomputed and decided in double (the error of the FP32 value is at
        // most about 2^-23 (|alpha| + |v dt/dx|) <= 1.4e-7)
        const bool guard = sizeof(R) == 4 && !kDirect;
        constexpr float flip_guard = 2.5e-7f;
        below.assign(N, 0);
        int64_t nguard = 0;
        #pragma omp parallel for schedule(static) reduction(+ : nguard)
        for (int a = 0; a < na; ++a) {
            const int i = active[a];
            R v = vn[i];
            if (noise)
                v += amp * std::sqrt(std::max(T[i], R(1))) * R(normal_dev(i, step_id, 0xabcddcbaULL));
            alpha_new[i] = std::min(alpha[i] + v * idt, R(1));
            bool b = alpha_new[i] <= R(K<R>::flip_eps);
            if (guard && std::fabs(float(alpha_new[i]) - K<float>::flip_eps) < flip_guard) {
                const double a64 = std::min(double(alpha[i]) + double(v) * (double(dt) / double(dx)), 1.0);
                alpha_new[i] = R(a64);
                b = a64 <= double(K<float>::flip_eps);
                ++nguard;
            }
            below[i] = b ? 1 : 0;
        }
        st.guarded = nguard;
        auto counts = histogram();
        st.accepted = 0; st.deferred = 0; st.candidates = 0;
        st.flip_global = 0; st.flip_depth = 0; st.flip_ready0 = 0;
        const bool network = law.d.flags & GPK_FLAG_GRAIN_NETWORK;
        std::unordered_map<int, int> lev;
        for (int i : active) {
            if (!(below[i] && phi_comp[i] != phi[i])) {
                alpha[i] = alpha_new[i];
                continue;
            }
            ++st.candidates;
            {
                const int l = chain_level(i, 1, lev);
                lev[i] = l;
                if (l == 1) ++st.flip_ready0;
                st.flip_depth = std::max<int64_t>(st.flip_depth, l);
            }
            const int32_t ol = phi[i], nl = phi_comp[i];
            bool glob = false;
            if (!network || simple_move(i, nl, counts[ol], counts[nl], false, &glob, &st.bfs_visits)) {
                phi[i] = nl; g[i] = g_comp[i]; alpha[i] = R(1);
                if (law.d.flags & GPK_FLAG_DISLOCATION) rho[i] = R(law.d.rho_rex);
                --counts[ol]; ++counts[nl];
                ++st.accepted;
            } else {
                alpha[i] = std::max(alpha_new[i], R(K<R>::flip_eps));
                ++st.deferred;
            }
            if (glob) ++st.flip_global;
        }
        st.exchanges = balanced_remap(st.accepted);
        if (Nz > 1 && st.exchanges > 0) {
            const int64_t first = st.exchanges;
            st.exchanges = balanced_remap(st.accepted) + first;
        }
        st.eroded = complete_unresolved();
    }

    // -- src/gpk_transport.cpp:11-35 -----------------------------------------------
    void diffusion(R dt) {
        if (D <= R(0)) return;
        std::vector<R> out(N);
        const R dx2 = dx * dx;
        #pragma omp parallel for schedule(static)
        for (int i = 0; i < N; ++i) {
            int f[6]; faces(i, f);
            R lap = c[f[1]] + c[f[0]] + c[f[3]] + c[f[2]];
            if (Nz > 1) lap += c[f[5]] + c[f[4]] - R(6) * c[i];
            else lap -= R(4) * c[i];
            const R De = mask[i] ? D * alpha[i] : D;
            out[i] = c[i] + dt * De * lap / dx2;
        }
        c.swap(out);
    }

    // -- src/gpk_transport.cpp:48-135 ------------------------------------------------
    void solid_fraction(std::vector<R>& s) const {
        s.resize(N);
        for (int i = 0; i < N; ++i) s[i] = phi[i] != 0 ? alpha[i] : R(1) - alpha[i];
    }
    int part_faces(int i, int nb[6]) const {
        int z, y, x; decode(i, z, y, x);
        int n = 0;
        nb[n++] = idx(z, y, x == Nx - 1 ? 0 : x + 1);
        nb[n++] = idx(z, y, x == 0 ? Nx - 1 : x - 1);
        nb[n++] = idx(z, y == Ny - 1 ? 0 : y + 1, x);
        nb[n++] = idx(z, y == 0 ? Ny - 1 : y - 1, x);
        if (Nz > 1) {
            nb[n++] = idx(z == Nz - 1 ? 0 : z + 1, y, x);
            nb[n++] = idx(z == 0 ? Nz - 1 : z - 1, y, x);
        }
        return n;
    }
    void solute_rejection() {
        const R k = R(law.d.k_partition);
        if (k >= R(1) || solid_old.size() != static_cast<size_t>(N)) return;
        std::vector<R> sn; solid_fraction(sn);
        std::vector<R> delta(N, R(0));
        for (int i = 0; i < N; ++i) {
            const R ds = sn[i] - solid_old[i];
            if (ds <= R(0)) continue;
            const R s0 = std::clamp(solid_old[i], R(0), R(1));
            const R cL = c[i] / std::max(R(1) - s0 * (R(1) - k), R(K<R>::cl_floor));
            R q = (R(1) - k) * cL * std::min(ds, R(1));
            q = std::min(q, c[i]);
            if (q <= R(0)) continue;
            int nb[6]; const int nn = part_faces(i, nb);
            R wsum = R(0), w[6];
            for (int a = 0; a < nn; ++a) {
                const int j = nb[a];
                w[a] = phi[j] == 0 ? std::max(R(1) - sn[j], R(0)) : R(0);
                wsum += w[a];
            }
            if (wsum <= R(0)) continue;
            delta[i] -= q;
            for (int a = 0; a < nn; ++a) if (w[a] > R(0)) delta[nb[a]] += q * w[a] / wsum;
        }
        for (int i = 0; i < N; ++i) c[i] += delta[i];
    }
    void diffusion_partitioned(R dt) {
        if (D <= R(0)) return;
        const R k = R(law.d.k_partition);
        std::vector<R> s, cL(N), out(N);
        solid_fraction(s);
        for (int i = 0; i < N; ++i) cL[i] = c[i] / std::max(R(1) - s[i] * (R(1) - k), R(K<R>::cl_floor));
        const R coef = dt * D / (dx * dx);
        #pragma omp parallel for schedule(static)
        for (int i = 0; i < N; ++i) {
            int nb[6]; const int nn = part_faces(i, nb);
            const R li = R(1) - s[i];
            R flux = R(0);
            for (int a = 0; a < nn; ++a) {
                const int j = nb[a];
                const R lf = std::min(li, R(1) - s[j]);
                if (lf > R(0)) flux += lf * (cL[j] - cL[i]);
            }
            out[i] = c[i] + coef * flux;
        }
        c.swap(out);
    }

    // -- src/gpk_simulation.cpp:31-54 ---------------------------------------------------
    R step(R dt, R cap) {
        interface_mask();
        st.band = static_cast<int64_t>(active.size());
        geometry();
        elastic();
        velocities();
        projection();
        const R stable = max_dt();
        if (dt < R(0)) dt = stable; else dt = std::min(dt, stable);
        dt = std::min(dt, cap);
        const bool part = law.d.flags & GPK_FLAG_SOLUTE_REJECTION;
        if (part) solid_fraction(solid_old);
        alpha_and_flips(dt);
        if (part) { solute_rejection(); diffusion_partitioned(dt); }
        else diffusion(dt);
        recovery(dt);
        time += double(dt);
        ++step_id;
        return dt;
    }

    // -- src/gpk_physics.cpp:10-63 -------------------------------------------------
    void elastic() {
        const auto& d = law.d;
        if (!(d.flags & GPK_FLAG_ELASTIC) || d.elastic_modulus <= 0.f || d.eigenstrain_coeff == 0.f) {
            std::fill(E_el.begin(), E_el.end(), R(0));
            return;
        }
        std::vector<R> es(N), sm(N), tmp(N);
        for (int i = 0; i < N; ++i) { es[i] = R(d.eigenstrain_coeff) * (c[i] - R(d.c_ref)); sm[i] = es[i]; }
        const int iters = std::max(1, int(d.elastic_smooth_iters));
        for (int it = 0; it < iters; ++it) {
            for (int i = 0; i < N; ++i) {
                int f[6]; faces(i, f);
                R avg = sm[f[0]] + sm[f[1]] + sm[f[2]] + sm[f[3]];
                if (Nz > 1) { avg += sm[f[4]] + sm[f[5]]; avg /= R(6); }
                else avg /= R(4);
                tmp[i] = R(0.5) * sm[i] + R(0.5) * avg;
            }
            sm.swap(tmp);
        }
        for (int i = 0; i < N; ++i) {
            const R dd = es[i] - sm[i];
            E_el[i] = R(0.5) * R(d.elastic_modulus) * dd * dd;
        }
    }

    // -- src/gpk_transport.cpp:142-152 ---------------------------------------------
    void recovery(R dt) {
        if (!(law.d.flags & GPK_FLAG_DISLOCATION) || law.d.recovery_rate <= 0.f || dt <= R(0)) return;
        const R decay = sizeof(R) == 4 ? R(std::exp(-law.d.recovery_rate * float(dt)))
                                       : R(std::exp(-double(law.d.recovery_rate) * double(dt)));
        for (int i = 0; i < N; ++i) rho[i] *= decay;
    }

    // -- diagnostics (src/gpk_simulation.cpp:78-123) --------------------------------------
    double phase_volume(int32_t label) const {
        double v = 0.0;
        for (int i = 0; i < N; ++i) {
            if (phi[i] == label) v += alpha[i];
            if (mask[i] && phi_comp[i] == label && phi_comp[i] != phi[i]) v += 1.0 - alpha[i];
        }
        return v;
    }
    int n_grains() const {
        std::vector<int32_t> t(phi);
        std::sort(t.begin(), t.end());
        return static_cast<int>(std::unique(t.begin(), t.end()) - t.begin());
    }
    double solute_sum() const {
        // canonical FP64 order: sequential sums over 1024-cell chunks, then over
        // 1024-chunk groups, then over groups
        const int64_t C = 1024;
        const int64_t nch = (N + C - 1) / C;
        std::vector<double> p1(nch);
        for (int64_t b = 0; b < nch; ++b) {
            double s = 0.0;
            for (int64_t i = b * C; i < std::min<int64_t>(N, (b + 1) * C); ++i) s += double(c[i]);
            p1[b] = s;
        }
        const int64_t ng = (nch + C - 1) / C;
        double tot = 0.0;
        for (int64_t b = 0; b < ng; ++b) {
            double s = 0.0;
            for (int64_t i = b * C; i < std::min<int64_t>(nch, (b + 1) * C); ++i) s += p1[i];
            tot += s;
        }
        return tot;
    }
};

// ---------------------------------------------------------------------------
// Voronoi initial condition (demos/gpk.cpp:51-189), brute force.
// ---------------------------------------------------------------------------
int voronoi(int Nz, int Ny, int Nx, int n_seeds, uint32_t seed, int32_t* phi, float* g,
            int64_t* detached = nullptr) {
    const int N = Nz * Ny * Nx;
    if (n_seeds <= 0 || n_seeds > N) return GPK_E_INVALID;
    std::mt19937 rng(seed);
    std::uniform_int_distribution<int> rz(0, Nz - 1), ry(0, Ny - 1), rx(0, Nx - 1);
    std::uniform_real_distribution<float> rg(0.f, float(M_PI) * 0.5f);
    std::vector<int> sz(n_seeds), sy(n_seeds), sx(n_seeds);
    std::vector<float> sg(n_seeds);
    std::unordered_set<int> used;
    for (int k = 0; k < n_seeds; ++k) {
        int lin;
        do {
            sz[k] = rz(rng); sy[k] = ry(rng); sx[k] = rx(rng);
            lin = (sz[k] * Ny + sy[k]) * Nx + sx[k];
        } while (!used.insert(lin).second);
        sg[k] = rg(rng);
    }
    auto idx = [&](int z, int y, int x) { return (z * Ny + y) * Nx + x; };
    #pragma omp parallel for schedule(static)
    for (int z = 0; z < Nz; ++z)
    for (int y = 0; y < Ny; ++y)
    for (int x = 0; x < Nx; ++x) {
        long long bd = LLONG_MAX; int best = 0;
        for (int k = 0; k < n_seeds; ++k) {
            long long a = std::abs(z - sz[k]), b = std::abs(y - sy[k]), e = std::abs(x - sx[k]);
            a = std::min(a, (long long)Nz - a); b = std::min(b, (long long)Ny - b); e = std::min(e, (long long)Nx - e);
            const long long d = a * a + b * b + e * e;
            if (d < bd) { bd = d; best = k; }
        }
        const int i = idx(z, y, x);
        phi[i] = best + 1; g[i] = sg[best];
    }
    // keep the largest face component of each label, re-attach the rest layer by layer
    auto nbrs = [&](int c, int o[6]) {
        const int x = c % Nx, y = (c / Nx) % Ny, z = c / (Ny * Nx);
        const int xm = x == 0 ? Nx - 1 : x - 1, xp = x == Nx - 1 ? 0 : x + 1;
        const int ym = y == 0 ? Ny - 1 : y - 1, yp = y == Ny - 1 ? 0 : y + 1;
        const int zm = z == 0 ? Nz - 1 : z - 1, zp = z == Nz - 1 ? 0 : z + 1;
        o[0] = idx(z, y, xm); o[1] = idx(z, y, xp); o[2] = idx(z, ym, x);
        o[3] = idx(z, yp, x); o[4] = idx(zm, y, x); o[5] = idx(zp, y, x);
    };
    const int nface = Nz > 1 ? 6 : 4;
    std::vector<uint8_t> vis(N, 0);
    std::vector<int> comp_of(N, -1);
    std::vector<int64_t> comp_size;
    std::vector<int32_t> comp_label;
    std::vector<int> largest(n_seeds + 1, -1);
    for (int s = 0; s < N; ++s) {
        if (vis[s]) continue;
        const int32_t lab = phi[s];
        const int id = static_cast<int>(comp_size.size());
        int64_t size = 0;
        std::queue<int> q; q.push(s); vis[s] = 1;
        while (!q.empty()) {
            const int cc = q.front(); q.pop();
            comp_of[cc] = id; ++size;
            int f[6]; nbrs(cc, f);
            for (int k = 0; k < nface; ++k)
                if (!vis[f[k]] && phi[f[k]] == lab) { vis[f[k]] = 1; q.push(f[k]); }
        }
        comp_size.push_back(size); comp_label.push_back(lab);
        const int prev = largest[lab];
        if (prev < 0 || size > comp_size[prev]) largest[lab] = id;
    }
    std::vector<uint8_t> att(N, 0);
    size_t natt = 0;
    for (int i = 0; i < N; ++i)
        if (largest[comp_label[comp_of[i]]] == comp_of[i]) { att[i] = 1; ++natt; }
    if (detached) *detached = int64_t(N) - int64_t(natt);
    while (natt < static_cast<size_t>(N)) {
        std::vector<int> cells; std::vector<int32_t> tg; std::vector<float> og;
        for (int cc = 0; cc < N; ++cc) {
            if (att[cc]) continue;
            int f[6]; nbrs(cc, f);
            int32_t cand[6]; int sup[6] = {0, 0, 0, 0, 0, 0}, src[6]; int nc = 0;
            for (int k = 0; k < nface; ++k) {
                if (!att[f[k]]) continue;
                const int32_t v = phi[f[k]];
                int q = 0; while (q < nc && cand[q] != v) ++q;
                if (q == nc) { cand[q] = v; src[q] = f[k]; ++nc; }
                ++sup[q];
            }
            if (nc == 0) continue;
            int b = 0;
            for (int q = 1; q < nc; ++q)
                if (sup[q] > sup[b] || (sup[q] == sup[b] && cand[q] < cand[b])) b = q;
            cells.push_back(cc); tg.push_back(cand[b]); og.push_back(g[src[b]]);
        }
        if (cells.empty()) return GPK_E_STALLED;
        for (size_t k = 0; k < cells.size(); ++k) { phi[cells[k]] = tg[k]; g[cells[k]] = og[k]; att[cells[k]] = 1; }
        natt += cells.size();
    }
    return GPK_OK;
}

// ---------------------------------------------------------------------------
// Type-erased handle
// ---------------------------------------------------------------------------
struct Base {
    virtual ~Base() = default;
    virtual int set(int f, const void* p, size_t b) = 0;
    virtual int get(int f, void* p, size_t b) = 0;
    virtual double step(double dt, double cap) = 0;
    virtual double max_dt_now() = 0;
    virtual Stats stats() const = 0;
    virtual double& t() = 0;
    virtual uint64_t& sid() = 0;
    virtual int simple_move(int cell, int32_t nl) = 0;
    virtual double phase_volume(int32_t l) const = 0;
    virtual int n_grains() const = 0;
    virtual double solute_sum() const = 0;
    virtual int real_size() const = 0;
    virtual int n() const = 0;
    virtual void stage(int which) = 0;
};

template <class R, bool kD>
struct Impl final : Base {
    Grid<R, kD> G;
    Impl(int a, int b, int c, float dx, float D, float s, const gpk_model_desc& m) : G(a, b, c, dx, D, s, m) {}
    std::vector<R>* rfield(int f) {
        switch (f) {
        case GPK_FIELD_ALPHA: return &G.alpha; case GPK_FIELD_C: return &G.c;
        case GPK_FIELD_T: return &G.T; case GPK_FIELD_G: return &G.g;
        case GPK_FIELD_RHO: return &G.rho; case GPK_FIELD_NX: return &G.nx;
        case GPK_FIELD_NY: return &G.ny; case GPK_FIELD_NZ: return &G.nz;
        case GPK_FIELD_KAPPA: return &G.kappa; case GPK_FIELD_VN: return &G.vn;
        case GPK_FIELD_G_COMP: return &G.g_comp; case GPK_FIELD_E_EL: return &G.E_el;
        default: return nullptr;
        }
    }
    int set(int f, const void* p, size_t b) override {
        if (f == GPK_FIELD_PHI || f == GPK_FIELD_PHI_COMP) {
            auto& v = f == GPK_FIELD_PHI ? G.phi : G.phi_comp;
            if (b != v.size() * 4) return GPK_E_INVALID;
            std::memcpy(v.data(), p, b); return GPK_OK;
        }
        if (f == GPK_FIELD_MASK) {
            if (b != G.mask.size()) return GPK_E_INVALID;
            std::memcpy(G.mask.data(), p, b); return GPK_OK;
        }
        auto* v = rfield(f);
        if (!v || b != v->size() * sizeof(R)) return GPK_E_INVALID;
        std::memcpy(v->data(), p, b); return GPK_OK;
    }
    int get(int f, void* p, size_t b) override {
        if (f == GPK_FIELD_PHI || f == GPK_FIELD_PHI_COMP) {
            auto& v = f == GPK_FIELD_PHI ? G.phi : G.phi_comp;
            if (b != v.size() * 4) return GPK_E_INVALID;
            std::memcpy(p, v.data(), b); return GPK_OK;
        }
        if (f == GPK_FIELD_MASK) {
            if (b != G.mask.size()) return GPK_E_INVALID;
            std::memcpy(p, G.mask.data(), b); return GPK_OK;
        }
        auto* v = rfield(f);
        if (!v || b != v->size() * sizeof(R)) return GPK_E_INVALID;
        std::memcpy(p, v->data(), b); return GPK_OK;
    }
    double step(double dt, double cap) override { return double(G.step(R(dt), R(cap))); }
    double max_dt_now() override { return double(G.max_dt()); }
    Stats stats() const override { return G.st; }
    double& t() override { return G.time; }
    uint64_t& sid() override { return G.step_id; }
    int simple_move(int cell, int32_t nl) override {
        if (cell < 0 || cell >= G.N) return 0;
        const int32_t ol = G.phi[cell];
        int oc = 0, nc = 0;
        for (int32_t v : G.phi) { oc += v == ol; nc += v == nl; }
        return G.simple_move(cell, nl, oc, nc, false) ? 1 : 0;
    }
    double phase_volume(int32_t l) const override { return G.phase_volume(l); }
    int n_grains() const override { return G.n_grains(); }
    double solute_sum() const override { return G.solute_sum(); }
    int real_size() const override { return int(sizeof(R)); }
    int n() const override { return G.N; }
    void stage(int which) override {
        // 0: mask, 1: mask+geometry, 2: +velocities, 3: +projection
        G.interface_mask();
        G.st.band = static_cast<int64_t>(G.active.size());
        if (which >= 1) G.geometry();
        if (which >= 2) G.velocities();
        if (which >= 3) G.projection();
    }
};

} // namespace

// ---------------------------------------------------------------------------
// C API of the oracle (used from tests through ctypes)
// ---------------------------------------------------------------------------
extern "C" {

struct og_stats {
    int64_t band, candidates, accepted, deferred, exchanges, eroded;
    int64_t flip_global, flip_depth, flip_ready0, bfs_visits;
    int64_t bgr_proposals, bgr_pairs, bgr_depth;
    int64_t guarded;
};

void* og_create(int mode, int Nz, int Ny, int Nx, float dx, float D, float sigma,
                const gpk_model_desc* m) {
    if (!m || Nz <= 0 || Ny <= 0 || Nx <= 0) return nullptr;
    try {
        switch (mode) {
        case 0: return new Impl<float, true>(Nz, Ny, Nx, dx, D, sigma, *m);
        case 1: return new Impl<float, false>(Nz, Ny, Nx, dx, D, sigma, *m);
        case 2: return new Impl<double, false>(Nz, Ny, Nx, dx, D, sigma, *m);
        case 3: return new Impl<double, true>(Nz, Ny, Nx, dx, D, sigma, *m);
        default: return nullptr;
        }
    } catch (...) { return nullptr; }
}
void og_destroy(void* h) { delete static_cast<Base*>(h); }
int og_set(void* h, int f, const void* p, size_t b) { return static_cast<Base*>(h)->set(f, p, b); }
int og_get(void* h, int f, void* p, size_t b) { return static_cast<Base*>(h)->get(f, p, b); }
double og_step(void* h, double dt, double cap) { return static_cast<Base*>(h)->step(dt, cap); }
void og_stage(void* h, int which) { static_cast<Base*>(h)->stage(which); }
double og_max_dt(void* h) { return static_cast<Base*>(h)->max_dt_now(); }
double og_time(void* h) { return static_cast<Base*>(h)->t(); }
void og_set_time(void* h, double t) { static_cast<Base*>(h)->t() = t; }
uint64_t og_step_id(void* h) { return static_cast<Base*>(h)->sid(); }
int og_real_size(void* h) { return static_cast<Base*>(h)->real_size(); }
int og_simple_move(void* h, int cell, int32_t nl) { return static_cast<Base*>(h)->simple_move(cell, nl); }
double og_phase_volume(void* h, int32_t l) { return static_cast<Base*>(h)->phase_volume(l); }
int og_n_grains(void* h) { return static_cast<Base*>(h)->n_grains(); }
double og_solute_sum(void* h) { return static_cast<Base*>(h)->solute_sum(); }
void og_stats_get(void* h, og_stats* o) {
    const Stats s = static_cast<Base*>(h)->stats();
    o->band = s.band; o->candidates = s.candidates; o->accepted = s.accepted;
    o->deferred = s.deferred; o->exchanges = s.exchanges; o->eroded = s.eroded;
    o->flip_global = s.flip_global; o->flip_depth = s.flip_depth; o->flip_ready0 = s.flip_ready0;
    o->bfs_visits = s.bfs_visits; o->bgr_proposals = s.bgr_proposals; o->bgr_pairs = s.bgr_pairs;
    o->bgr_depth = s.bgr_depth;
    o->guarded = s.guarded;
}
int og_voronoi(int Nz, int Ny, int Nx, int n_seeds, uint32_t seed, int32_t* phi, float* g) {
    return voronoi(Nz, Ny, Nx, n_seeds, seed, phi, g);
}
// cells outside the largest face component of their nearest-site label,
// i.e. the cells the normalisation re-attaches
int64_t og_voronoi_detached(int Nz, int Ny, int Nx, int n_seeds, uint32_t seed) {
    std::vector<int32_t> phi(size_t(Nz) * Ny * Nx);
    std::vector<float> g(phi.size());
    int64_t det = -1;
    if (voronoi(Nz, Ny, Nx, n_seeds, seed, phi.data(), g.data(), &det) != 0) return -1;
    return det;
}
// float GaussKernel weights, for closed-form checks (sum = 1 etc.)
int og_gauss(float sigma, float* w, int cap, int* radius, float* w0, float* w1, float* w2, int* rg) {
    const Gauss<float> k = make_gauss_f(sigma);
    *radius = k.radius;
    if (2 * k.radius + 1 > cap) return GPK_E_INVALID;
    std::copy(k.w.begin(), k.w.end(), w);
    std::vector<float> a, b, c; int r = 0;
    geometry_kernels(k, a, b, c, r);
    *rg = r;
    if (2 * r + 1 > cap) return GPK_E_INVALID;
    std::copy(a.begin(), a.end(), w0); std::copy(b.begin(), b.end(), w1); std::copy(c.begin(), c.end(), w2);
    return GPK_OK;
}

} // extern "C"
