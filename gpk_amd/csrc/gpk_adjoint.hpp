// gpk_adjoint.hpp — hand-written adjoint of the velocity chain
//
//   v_raw_j  = M (dG_j - sigma(n_j; eps4, eps6) kappa_j)      (src/gpk_physics.cpp:97-146)
//   v_out    = P v_raw                                         (src/gpk_physics.cpp:152-193)
//
// with sigma = sigma0 (1 + eps4 a4(n) + eps6 a6(n)): in 2-D (crystal-frame
// n_z == 0) a4 = cos 4 theta, a6 = cos 6 theta; in 3-D a4 = sum n^4 - 3/5,
// a6 = 0 (src/gpk_thermo.cpp:9-30; the cos 6 theta term is an extension).
// P is the pair projection: for a band cell i with phi_i != phi_comp_i,
//   (P v)_i = o_i sum_j w_ij o_j v_j / den_i,  den_i = sum_j w_ij,
// over band cells j of the (2r+1)^d window with the same unordered label pair
// and phi_j != phi_comp_j, o = +1 on the pair's low label and -1 on its high
// label; other band cells pass v_raw through.  The window relation is
// symmetric, so
//   (P^T g)_j = g_j                                  if phi_j == phi_comp_j
//             = o_j sum_i w_ij h_i,  h_i = o_i g_i / den_i   otherwise,
// and dL/dtheta = sum_j (P^T g)_j dv_raw_j/dtheta.  Normals and curvature
// depend on the labels only; they and den are recomputed, not stored.
#pragma once

#include "gpk_kernels.hpp"

namespace gpkd {

// d sigma / d eps4 and d sigma / d eps6 divided by sigma0, on the same branch
// structure as law_energy.
template <class R>
__device__ __forceinline__ void law_aniso_basis(R nx, R ny, R nz, R gv, R& a4, R& a6) {
    const R cg = dcos(gv), sg = dsin(gv);
    const R ncx = nx * cg + ny * sg;
    const R ncy = -nx * sg + ny * cg;
    const R ncz = nz;
    if (ncz == R(0)) {
        const R th = datan2(ncy, ncx);
        a4 = dcos(R(4) * th);
        a6 = dcos(R(6) * th);
        return;
    }
    a4 = ncx * ncx * ncx * ncx + ncy * ncy * ncy * ncy + ncz * ncz * ncz * ncz - R(3) / R(5);
    a6 = R(0);
}

// window sum over same-pair projected cells; fn(j, w) per contributing cell
template <int DIM, class Fn>
__device__ __forceinline__ void pair_window(const Dims& d, int r, int i, int32_t lo, int32_t hi,
                                            const uint8_t* __restrict__ mask,
                                            const int32_t* __restrict__ phi,
                                            const int32_t* __restrict__ phi_comp,
                                            const double* __restrict__ wtab, Fn fn) {
    int z0, y0, x0;
    decode(d, i, z0, y0, x0);
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
                fn(j, pj == lo, wtab[ddx * ddx + dy * dy + dz * dz]);
            }
        }
    }
}

// h_a = o_i g_i / den_i for projected band cells, 0 otherwise
template <class R, int DIM>
__global__ void __launch_bounds__(256)
k_adj_h(Dims d, int r, int nb, const int* __restrict__ band, const uint8_t* __restrict__ mask,
        const int32_t* __restrict__ phi, const int32_t* __restrict__ phi_comp,
        const double* __restrict__ wtab, const R* __restrict__ gvel, double* __restrict__ h) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nb) return;
    const int i = band[a];
    const int32_t p = phi[i], q = phi_comp[i];
    if (p == q) {
        h[a] = 0.0;
        return;
    }
    const int32_t lo = p < q ? p : q, hi = p < q ? q : p;
    double den = 0.0;
    pair_window<DIM>(d, r, i, lo, hi, mask, phi, phi_comp, wtab,
                     [&](int, bool, double w) { den += w; });
    h[a] = (p == lo ? 1.0 : -1.0) * double(gvel[i]) / den;
}

// per band cell: (P^T g)_j times dv_raw_j/d(eps4, eps6, M0); one FP64 partial
// per workgroup and parameter, reduced in a fixed tree order
template <class R, int DIM>
__global__ void __launch_bounds__(256)
k_adj_terms(Dims d, int r, int nb, const int* __restrict__ band, const int* __restrict__ slot,
            const uint8_t* __restrict__ mask, const int32_t* __restrict__ phi,
            const int32_t* __restrict__ phi_comp, const double* __restrict__ wtab,
            const R* __restrict__ gvel, const double* __restrict__ h,
            const R* __restrict__ bnx, const R* __restrict__ bny, const R* __restrict__ bnz,
            const R* __restrict__ bkappa, const R* __restrict__ g, gpk_model_desc model,
            double* __restrict__ partial) {
    __shared__ double red[3][256];
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    double t4 = 0.0, t6 = 0.0, tm = 0.0;
    if (a < nb) {
        const int j = band[a];
        const int32_t p = phi[j], q = phi_comp[j];
        double pt;
        if (p == q) {
            pt = double(gvel[j]);
        } else {
            const int32_t lo = p < q ? p : q, hi = p < q ? q : p;
            double s = 0.0;
            pair_window<DIM>(d, r, j, lo, hi, mask, phi, phi_comp, wtab,
                             [&](int i, bool, double w) { s += w * h[slot[i]]; });
            pt = (p == lo ? 1.0 : -1.0) * s;
        }
        R a4, a6;
        law_aniso_basis<R>(bnx[a], bny[a], bnz[a], g[j], a4, a6);
        const R kap = bkappa[a];
        const R sig = law_energy<R>(model, bnx[a], bny[a], bnz[a], g[j]);
        const R ms0k = R(model.M0) * R(model.sigma0) * kap;
        t4 = pt * double(-ms0k * a4);
        t6 = pt * double(-ms0k * a6);
        tm = pt * double(-(sig * kap));
    }
    red[0][threadIdx.x] = t4;
    red[1][threadIdx.x] = t6;
    red[2][threadIdx.x] = tm;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int k = 0; k < 3; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < 3) partial[3 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// one workgroup: out[k] = sum over blocks of partial[3 b + k], fixed order
__global__ void __launch_bounds__(256)
k_adj_final(int nblocks, const double* __restrict__ partial, double* __restrict__ out) {
    __shared__ double red[3][256];
    double s[3] = {0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nblocks; b += 256)
        for (int k = 0; k < 3; ++k) s[k] += partial[3 * b + k];
    for (int k = 0; k < 3; ++k) red[k][threadIdx.x] = s[k];
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st)
            for (int k = 0; k < 3; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x < 3) out[threadIdx.x] = red[threadIdx.x][0];
}

} // namespace gpkd
