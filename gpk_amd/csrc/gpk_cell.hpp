// gpk_cell.hpp — per-cell device functions of the step shared by the
// grid-wide kernels and the batched-instance kernel: stencil sums, normals,
// curvature, competitor and constitutive velocity of one interface cell.
#pragma once
#include "gpk_device.hpp"

namespace gpkd {

// ---------------------------------------------------------------------------
// Geometry (src/gpk_geometry.cpp:88-183) fused with the competitor choice and
// the constitutive velocity (src/gpk_physics.cpp:67-148).
//   kDirect: the reference's masked tensor stencil, accumulated in dz,dy,dx
//            order from the host-built product table tab[t*9+c] = (gx*gy)*gz
//            (bit-identical to forming the products inline).
//   else   : separable masked row sums over x, FMA over y then z.
// ---------------------------------------------------------------------------
struct GeomParams {
    int rg;                 // stencil radius
    int W;                  // 2*rg+1
    float dx_f;
    float sigma_cells;
};

template <class R> struct Eps;
template <> struct Eps<float> {
    static constexpr float grad = 1e-8f, kappa = 1e-6f, flip = 1e-6f, gain = 1e-7f;
    static constexpr float vmax0 = 1e-30f, big = 1e30f, c_iface = 0.125f, c_diff = 0.4f;
    static constexpr float cl_floor = 1e-6f;
};
// FP64 guard band of the FP32 mode's flip decision.  The FP32 value
// alpha + v * (dt / dx) differs from the exact value of the same FP32 inputs
// by at most about 2^-23 (|alpha| + |v dt / dx|) <= 1.4e-7 (|alpha| <= 1 and
// |v dt / dx| <= 1/8 by the timestep rule), so every cell whose FP32 value
// lies within flip_guard = 2.5e-7 of the threshold is decided again in FP64;
// the decision then equals the exact one for the FP32 state.
constexpr float flip_guard = 2.5e-7f;

// alpha_new = min(alpha + v dt/dx, 1) and whether it is at or below the flip
// threshold (src/gpk_transitions.cpp:262-291); with `guard`, values inside the
// band are recomputed and decided in FP64 (FP32 mode only)
template <class R>
__device__ __forceinline__ bool alpha_step(R alpha, R v, R idt, double dt, double dx, bool guard,
                                           R& an, bool& guarded);

template <> struct Eps<double> {
    static constexpr double grad = 1e-8, kappa = 1e-6, flip = 1e-6, gain = 1e-7;
    static constexpr double vmax0 = 1e-30, big = 1e30, c_iface = 0.125, c_diff = 0.4;
    static constexpr double cl_floor = 1e-6;
};

template <class R>
__device__ __forceinline__ bool alpha_step(R alpha, R v, R idt, double dt, double dx, bool guard,
                                           R& an, bool& guarded) {
    an = alpha + v * idt;
    an = R(1) < an ? R(1) : an;
    guarded = false;
    if (sizeof(R) == 4 && guard) {
        const float d = float(an) - Eps<float>::flip;
        if ((d < 0.f ? -d : d) < flip_guard) {
            double a64 = double(alpha) + double(v) * (dt / dx);
            a64 = 1.0 < a64 ? 1.0 : a64;
            an = R(a64);
            guarded = true;
            return a64 <= double(Eps<float>::flip);
        }
    }
    return an <= Eps<R>::flip;
}

__device__ __forceinline__ float fmar(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmar(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Row-sum tables of the separable modes: window-row "same label" bits m (bit
// dx set when window position dx carries the resident label) cut into nch
// chunks of at most C consecutive dx; chunk k covers dx in
// [start_k, start_k + len_k) and its table entry idx*3 + a holds the sum of
// weight set a over the set bits in ascending dx.  The row sum adds the chunk
// values in ascending chunk order.
struct RowLutDesc {
    int W, nch;
    int start[8], len[8], base[8];
};

template <class R>
__device__ __forceinline__ void row_sum(const RowLutDesc& L, const R* __restrict__ t, uint64_t m,
                                        R& R0, R& R1, R& R2) {
    for (int k = 0; k < L.nch; ++k) {
        const int ix = int((m >> L.start[k]) & ((1ull << L.len[k]) - 1ull));
        const R* e = t + L.base[k] + ix * 3;
        if (k == 0) { R0 = e[0]; R1 = e[1]; R2 = e[2]; }
        else { R0 += e[0]; R1 += e[1]; R2 += e[2]; }
    }
}

template <class R>
struct LutRef {
    RowLutDesc L;
    const R* t;
};
template <class R>
__device__ __forceinline__ void row_sum(const LutRef<R>& lr, uint64_t m, R& R0, R& R1, R& R2) {
    row_sum<R>(lr.L, lr.t, m, R0, R1, R2);
}

// LawTag<1>: the model is known to be the plain isotropic law (kind
// ISOTROPIC, eps4 = eps6 = 0, no elastic or dislocation term), for which the
// general law gives dG = 0 and sigma = sigma0; the kernel is compiled without
// the other laws.  LawTag<0>: the law is chosen from the descriptor at run time.
template <int L> struct LawTag {};

// competitor (src/gpk_geometry.cpp:154-182) and constitutive velocity
// (src/gpk_physics.cpp:97-146) of band cell i (slot a) from its normal and
// curvature
template <class R, int DIM, class CubeAt, int LAW = 0>
__device__ __forceinline__ void comp_vel_cell(
    const Dims& d, const gpk_model_desc& model, int i, int a, int z0, int y0, int x0, int32_t res,
    R nxv, R nyv, R nzv, R kap, CubeAt cube_at, const int32_t* __restrict__ phi,
    const R* __restrict__ g, const R* __restrict__ c, const R* __restrict__ T,
    const R* __restrict__ rho, const R* __restrict__ E_el, R* __restrict__ vraw,
    int32_t* __restrict__ phi_comp, R* __restrict__ g_comp, LawTag<LAW> = LawTag<0>{}) {
    (void)cube_at;
    constexpr bool plain = LAW == 1;

    // competitor: most-supported unlike face label, ties by outward-normal
    // alignment, then first face (src/gpk_geometry.cpp:154-182)
    int f[6];
    faces(d, i, f);
    const R score[6] = {-nxv, nxv, -nyv, nyv, -nzv, nzv};
    const int nface = DIM == 3 ? 6 : 4;
    int32_t cube[27];
    {
        int t = 0;
        const int zb = DIM == 3 ? -1 : 0, ze = DIM == 3 ? 1 : 0;
        for (int dz = zb; dz <= ze; ++dz) {
            int zz = z0 + dz; if (zz < 0) zz += d.Nz; else if (zz >= d.Nz) zz -= d.Nz;
            for (int dy = -1; dy <= 1; ++dy) {
                int yy = y0 + dy; if (yy < 0) yy += d.Ny; else if (yy >= d.Ny) yy -= d.Ny;
                for (int ddx = -1; ddx <= 1; ++ddx, ++t) {
                    int xx = x0 + ddx; if (xx < 0) xx += d.Nx; else if (xx >= d.Nx) xx -= d.Nx;
                    cube[t] = cube_at(t, zz, yy, xx);
                }
            }
        }
    }
    const int ncube = DIM == 3 ? 27 : 9, ccentre = DIM == 3 ? 13 : 4;
    int best = -1, best_support = -1;
    R best_score = R(-2);
    for (int k = 0; k < nface; ++k) {
        const int32_t q = phi[f[k]];
        if (q == res) continue;
        int support = 0;
        for (int t = 0; t < ncube; ++t) support += (t != ccentre) && (cube[t] == q);
        if (support > best_support || (support == best_support && score[k] > best_score)) {
            best = k; best_support = support; best_score = score[k];
        }
    }
    int32_t qid = best >= 0 ? phi[f[best]] : res;
    // phase selection (src/gpk_physics.cpp:79-95): the lowest-G unlike face
    // label replaces the geometric competitor
    if (!plain && model.kind == GPK_MODEL_PRECIPITATION) {
        int32_t fl[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) fl[k] = phi[f[k]];
        const int sel = law_select_phase<R>(model, res, c[i], T[i], fl);
        if (sel >= 0) qid = sel;
    }
    phi_comp[i] = qid;

    // constitutive velocity v = M (dG - sigma(n) kappa) (src/gpk_physics.cpp:97-146)
    const R gpv = g[i];
    R gq = gpv;
    int fq = -1;
    for (int k = 0; k < 6; ++k)
        if (phi[f[k]] == qid) { fq = f[k]; break; }
    if (fq >= 0) gq = g[fq];
    g_comp[i] = gq;
    R dG, sig;
    if (plain) {
        dG = R(0);
        sig = R(model.sigma0);
    } else {
        const R cq = fq >= 0 ? c[fq] : c[i];
        const R ci = law_interface_composition<R>(model, res, qid, c[i], cq);
        dG = law_driving_force<R>(model, res, qid, ci, T[i]);
        sig = law_energy<R>(model, nxv, nyv, nzv, gpv);
    }
    const R mob = R(model.M0);
    // elastic energy of the product side minus the parent's
    // (src/gpk_physics.cpp:122-127)
    if (!plain && (model.flags & GPK_FLAG_ELASTIC)) {
        R ep = E_el[i];
        for (int k = 0; k < 6; ++k)
            if (phi[f[k]] == qid) { ep = E_el[f[k]]; break; }
        dG += (ep - E_el[i]);
    }
    // stored dislocation energy of the product side minus the parent's
    // (src/gpk_physics.cpp:131-140)
    if (!plain && (model.flags & GPK_FLAG_DISLOCATION)) {
        R rp = rho[i];
        for (int k = 0; k < 6; ++k)
            if (phi[f[k]] == qid) { rp = rho[f[k]]; break; }
        dG += R(model.disloc_coeff) * (rp - rho[i]);
    }
    vraw[a] = mob * (dG - sig * kap);
}

// normals, curvature (src/gpk_geometry.cpp:136-151), competitor (154-182) and
// the constitutive velocity of band cell i (slot a) from its stencil sums
template <class R, int DIM, class CubeAt, int LAW = 0>
__device__ __forceinline__ void finish_cell(
    const Dims& d, const GeomParams& gp, const gpk_model_desc& model, int i, int a, int z0, int y0,
    int x0, int32_t res, R Ix, R Iy, R Iz, R Ixx, R Iyy, R Izz, R Ixy, R Ixz, R Iyz, CubeAt cube_at,
    const int32_t* __restrict__ phi, const R* __restrict__ g, const R* __restrict__ c,
    const R* __restrict__ T, const R* __restrict__ rho, const R* __restrict__ E_el, R* __restrict__ onx,
    R* __restrict__ ony,
    R* __restrict__ onz,
    R* __restrict__ okappa, R* __restrict__ vraw, int32_t* __restrict__ phi_comp,
    R* __restrict__ g_comp, LawTag<LAW> law = LawTag<0>{}) {
    const R dx = R(gp.dx_f);
    const R inv_dx = R(1) / dx;
    const R inv_dx2 = inv_dx * inv_dx;
    const R klim = R(1) / (R(gp.sigma_cells) * dx);
    Ix *= inv_dx; Iy *= inv_dx; Iz *= inv_dx;
    Ixx *= inv_dx2; Iyy *= inv_dx2; Izz *= inv_dx2;
    Ixy *= inv_dx2; Ixz *= inv_dx2; Iyz *= inv_dx2;
    const R gsq = Ix * Ix + Iy * Iy + Iz * Iz;
    R nxv = 0, nyv = 0, nzv = 0, kap = 0;
    if (gsq * dx * dx > Eps<R>::grad) {
        const R gr = dsqrt(gsq);
        nxv = -Ix / gr; nyv = -Iy / gr; nzv = -Iz / gr;
        const R trH = Ixx + Iyy + Izz;
        const R gHg = Ix * Ix * Ixx + Iy * Iy * Iyy + Iz * Iz * Izz
                    + R(2) * (Ix * Iy * Ixy + Ix * Iz * Ixz + Iy * Iz * Iyz);
        R kv = (gHg - gsq * trH) / (gsq * gr);
        if (dfabs(kv * dx) < Eps<R>::kappa) kv = R(0);
        kap = kv < -klim ? -klim : (klim < kv ? klim : kv);
    }
    onx[a] = nxv; ony[a] = nyv; onz[a] = nzv; okappa[a] = kap;
    comp_vel_cell<R, DIM>(d, model, i, a, z0, y0, x0, res, nxv, nyv, nzv, kap, cube_at, phi, g, c, T,
                          rho, E_el, vraw, phi_comp, g_comp, law);
}


// stencil sums of band cell i (slot a) straight from the label field, then
// finish_cell; all pointers are relative to the grid the Dims describe
template <class R, int DIM, bool kDirect>
__device__ __forceinline__ void geometry_cell(
    const Dims& d, const GeomParams& gp, const gpk_model_desc& model, int a, int i,
    const int32_t* __restrict__ phi, const R* __restrict__ w0, const R* __restrict__ w1,
    const R* __restrict__ w2, const R* __restrict__ tab, const RowLutDesc& lutd,
    const R* __restrict__ g, const R* __restrict__ c, const R* __restrict__ T,
    const R* __restrict__ rho, const R* __restrict__ E_el,
    R* __restrict__ onx, R* __restrict__ ony, R* __restrict__ onz, R* __restrict__ okappa,
    R* __restrict__ vraw, int32_t* __restrict__ phi_comp, R* __restrict__ g_comp) {
    const int32_t res = phi[i];
    int z0, y0, x0;
    decode(d, i, z0, y0, x0);
    const int r = gp.rg, W = gp.W;
    const LutRef<R> lut{lutd, tab};
    R Ix = 0, Iy = 0, Iz = 0, Ixx = 0, Iyy = 0, Izz = 0, Ixy = 0, Ixz = 0, Iyz = 0;
    if (kDirect) {
        int t = 0;
        const int dzb = DIM == 3 ? -r : 0, dze = DIM == 3 ? r : 0;
        for (int dz = dzb; dz <= dze; ++dz) {
            const int z = wrapi(z0 + dz, d.Nz);
            for (int dy = -r; dy <= r; ++dy) {
                const int y = wrapi(y0 + dy, d.Ny);
                const int rowbase = (z * d.Ny + y) * d.Nx;
                for (int ddx = -r; ddx <= r; ++ddx, ++t) {
                    const int x = wrapi(x0 + ddx, d.Nx);
                    if (phi[rowbase + x] != res) continue;
                    if (DIM == 3) {
                        const R* p = tab + t * 9;
                        Ix += p[0]; Iy += p[1]; Ixx += p[2]; Iyy += p[3]; Ixy += p[4];
                        Iz += p[5]; Izz += p[6]; Ixz += p[7]; Iyz += p[8];
                    } else {
                        const R* p = tab + t * 5;
                        Ix += p[0]; Iy += p[1]; Ixx += p[2]; Iyy += p[3]; Ixy += p[4];
                    }
                }
            }
        }
    } else {
        const int dzb = DIM == 3 ? -r : 0, dze = DIM == 3 ? r : 0;
        for (int dz = dzb; dz <= dze; ++dz) {
            const int z = wrapi(z0 + dz, d.Nz);
            R S00 = 0, S01 = 0, S02 = 0, S10 = 0, S11 = 0, S20 = 0;
            for (int dy = -r; dy <= r; ++dy) {
                const int y = wrapi(y0 + dy, d.Ny);
                const int rowbase = (z * d.Ny + y) * d.Nx;
                const int jy = dy + r;
                uint64_t m = 0;
                for (int ddx = -r; ddx <= r; ++ddx) {
                    const int x = wrapi(x0 + ddx, d.Nx);
                    if (phi[rowbase + x] == res) m |= 1ull << (ddx + r);
                }
                R R0, R1, R2;
                row_sum(lut, m, R0, R1, R2);
                S00 = fmar(w0[jy], R0, S00); S01 = fmar(w0[jy], R1, S01);
                S02 = fmar(w0[jy], R2, S02); S10 = fmar(w1[jy], R0, S10);
                S11 = fmar(w1[jy], R1, S11); S20 = fmar(w2[jy], R0, S20);
            }
            if (DIM == 3) {
                const int jz = dz + r;
                const R gz0 = w0[jz], gz1 = w1[jz], gz2 = w2[jz];
                Ix = fmar(gz0, S01, Ix); Iy = fmar(gz0, S10, Iy);
                Ixx = fmar(gz0, S02, Ixx); Iyy = fmar(gz0, S20, Iyy);
                Ixy = fmar(gz0, S11, Ixy);
                Iz = fmar(gz1, S00, Iz); Izz = fmar(gz2, S00, Izz);
                Ixz = fmar(gz1, S01, Ixz); Iyz = fmar(gz1, S10, Iyz);
            } else {
                Ix = S01; Iy = S10; Ixx = S02; Iyy = S20; Ixy = S11;
            }
        }
    }
    (void)W;
    auto cube_at = [&](int t, int zz, int yy, int xx) {
        (void)t;
        return phi[lin(d, zz, yy, xx)];
    };
    finish_cell<R, DIM>(d, gp, model, i, a, z0, y0, x0, res, Ix, Iy, Iz, Ixx, Iyy, Izz, Ixy, Ixz, Iyz,
                        cube_at, phi, g, c, T, rho, E_el, onx, ony, onz, okappa, vraw, phi_comp, g_comp);
}



// pair-projected velocity of band cell i (slot a), src/gpk_physics.cpp:152-193
template <class R, int DIM>
__device__ __forceinline__ R project_value(const Dims& d, int r, int a, int i,
                                           const int* __restrict__ slot,
                                           const uint8_t* __restrict__ mask,
                                           const int32_t* __restrict__ phi,
                                           const int32_t* __restrict__ phi_comp,
                                           const double* __restrict__ wtab,
                                           const R* __restrict__ vraw) {
    const int32_t p = phi[i], q = phi_comp[i];
    R out = vraw[a];
    if (p != q) {
        const int32_t lo = p < q ? p : q, hi = p < q ? q : p;
        const R oi = p == lo ? R(1) : R(-1);
        int z0, y0, x0;
        decode(d, i, z0, y0, x0);
        double num = 0.0, den = 0.0;
        const int dzb = DIM == 3 ? -r : 0, dze = DIM == 3 ? r : 0;
        for (int dz = dzb; dz <= dze; ++dz) {
            const int z = wrapi(z0 + dz, d.Nz);
            for (int dy = -r; dy <= r; ++dy) {
                const int y = wrapi(y0 + dy, d.Ny);
                for (int ddx = -r; ddx <= r; ++ddx) {
                    const int x = wrapi(x0 + ddx, d.Nx);
                    const int j = lin(d, z, y, x);
                    if (!mask[j]) continue;
                    const int32_t pj = phi[j], qj = phi_comp[j];
                    if (pj == qj) continue;
                    const int32_t lj = pj < qj ? pj : qj, hj = pj < qj ? qj : pj;
                    if (lj != lo || hj != hi) continue;
                    const double w = wtab[ddx * ddx + dy * dy + dz * dz];
                    const R oj = pj == lo ? R(1) : R(-1);
                    num += w * double(oj) * double(vraw[slot[j]]);
                    den += w;
                }
            }
        }
        if (den > 0.0) out = oi * R(num / den);
    }
    return out;
}

// remap scores of band cell i (src/gpk_transitions.cpp:50-81): returns the
// gain eps_smooth - eps_star and the best candidate label in *best_label
template <class R, int DIM>
__device__ __forceinline__ R remap_score_cell(const Dims& d, int r, int i,
                                              const int32_t* __restrict__ phi,
                                              const double* __restrict__ wtab,
                                              int32_t& best_label) {
    const int32_t p = phi[i];
    int f[6];
    faces(d, i, f);
    const int nface = DIM == 3 ? 6 : 4;
    int32_t cand[6];
    int nc = 0;
    for (int k = 0; k < nface; ++k) {
        const int32_t v = phi[f[k]];
        if (v == p) continue;
        bool known = false;
        for (int q = 0; q < nc; ++q) known |= cand[q] == v;
        if (!known) cand[nc++] = v;
    }
    int z0, y0, x0;
    decode(d, i, z0, y0, x0);
    double resident = 0.0, sup[6] = {0, 0, 0, 0, 0, 0};
    const int dzb = DIM == 3 ? -r : 0, dze = DIM == 3 ? r : 0;
    for (int dz = dzb; dz <= dze; ++dz) {
        const int z = wrapi(z0 + dz, d.Nz);
        for (int dy = -r; dy <= r; ++dy) {
            const int y = wrapi(y0 + dy, d.Ny);
            for (int ddx = -r; ddx <= r; ++ddx) {
                const int x = wrapi(x0 + ddx, d.Nx);
                const int32_t v = phi[lin(d, z, y, x)];
                const double w = wtab[ddx * ddx + dy * dy + dz * dz];
                if (v == p) resident += w;
                for (int q = 0; q < nc; ++q) if (v == cand[q]) sup[q] += w;
            }
        }
    }
    double best = -1.0;
    int32_t bl = p;
    for (int q = 0; q < nc; ++q) if (sup[q] > best) { best = sup[q]; bl = cand[q]; }
    best_label = bl;
    const R es = R(resident), eb = R(best);
    return eb - es;
}

__device__ __forceinline__ bool bounded_cert_masks(bool d3, Bits128 bo, Bits128 bn, int old_count);

// bounded (5^d box) certificate for relabelling `cell` from ol to nl, labels
// read through label_at (src/gpk_connectivity.cpp:10-195 with local_only)
template <class LabelAt>
__device__ bool bounded_cert(const Dims& d, bool d3, int cell, int32_t ol, int32_t nl, int old_count,
                             LabelAt label_at) {
    int z0, y0, x0;
    decode(d, cell, z0, y0, x0);
    Bits128 bo{0, 0}, bn{0, 0};
    const int nz = d3 ? 5 : 1;
    for (int lz = 0; lz < nz; ++lz) {
        const int z = wrapi(z0 + (d3 ? lz - 2 : 0), d.Nz);
        for (int ly = 0; ly < 5; ++ly) {
            const int y = wrapi(y0 + ly - 2, d.Ny);
            for (int lx = 0; lx < 5; ++lx) {
                const int x = wrapi(x0 + lx - 2, d.Nx);
                const int cc = lin(d, z, y, x);
                const int32_t lab = label_at(cc);
                const int b = lz * 25 + ly * 5 + lx;
                if (lab == ol) b_set(bo, b);
                if (lab == nl) b_set(bn, b);
            }
        }
    }
    return bounded_cert_masks(d3, bo, bn, old_count);
}

// the certificate's decision from the box masks of the old (bo) and new (bn)
// label, bit lz * 25 + ly * 5 + lx
__device__ __forceinline__ bool bounded_cert_masks(bool d3, Bits128 bo, Bits128 bn, int old_count) {
    const int lc = (d3 ? 2 : 0) * 25 + 12;
    // 3^d cube occupancy from the box
    uint32_t co = 0, cn = 0;
    {
        int t = 0;
        const int zb = d3 ? -1 : 0, ze = d3 ? 1 : 0;
        for (int dz = zb; dz <= ze; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int ddx = -1; ddx <= 1; ++ddx, ++t) {
                    const int b = lc + dz * 25 + dy * 5 + ddx;
                    if (b_test(bo, b)) co |= 1u << t;
                    if (b_test(bn, b)) cn |= 1u << t;
                }
    }
    const int centre = d3 ? 13 : 4;
    co &= ~(1u << centre);
    cn &= ~(1u << centre);
    const BoxMasks M = box_masks(d3);
    const int bl[6] = {lc - 1, lc + 1, lc - 5, lc + 5, lc - 25, lc + 25};
    const int nface = d3 ? 6 : 4;
    auto connected = [&](Bits128 occ) -> bool {
        int br[6], nb = 0;
        for (int k = 0; k < nface; ++k) if (b_test(occ, bl[k])) br[nb++] = bl[k];
        if (nb <= 1) return true;
        Bits128 comp{0, 0};
        b_set(comp, br[0]);
        while (true) {
            const Bits128 g = b_and(grow5(comp, M, d3), occ);
            if (b_eq(g, comp)) break;
            comp = g;
        }
        for (int k = 1; k < nb; ++k) if (!b_test(comp, br[k])) return false;
        return true;
    };
    const int oo = topo_count(co, d3);
    if (oo == 0 && !(old_count <= 1)) return false;
    if (oo > 1) {
        Bits128 occ = bo;
        // centre excluded for the old label
        Bits128 cm{0, 0};
        b_set(cm, lc);
        occ = b_andn(occ, cm);
        if (!connected(occ)) return false;
    }
    const int no = topo_count(cn, d3);
    if (no == 0) return false;
    if (no > 1 && !connected(bn)) return false;
    return true;
}

} // namespace gpkd
