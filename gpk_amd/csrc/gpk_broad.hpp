// gpk_broad.hpp — geometry of 3-D grids whose stencil radius exceeds 18
// (src/gpk_grid.cpp:44, src/gpk_geometry.cpp:185-390).
//
// For each label present on the interface, in ascending label order: the 0/1
// indicator of the label smoothed by the plain Gaussian (radius kr) along x,
// then y, then z -- each output a k-ordered sum of separately rounded
// products -- and, at the label's interface cells, fourth-order centred
// differences of the smoothed field give the gradient and Hessian, hence the
// normal (times 1/|grad|) and the curvature.  Competitor and velocity follow
// once every label is done.
#pragma once

#include "gpk_cell.hpp"

namespace gpkd {

__global__ void k_broad_flag_labels(int nb, const int* __restrict__ band, const int32_t* __restrict__ phi,
                                    uint8_t* __restrict__ flags) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a < nb) flags[phi[band[a]]] = 1;
}

// x pass straight from the labels: acc += w[k] * [phi == p] in k order
// (adding +0 for an unlike cell leaves acc unchanged)
template <class R>
__global__ void __launch_bounds__(256)
k_broad_x(Dims d, int32_t p, int kr, const R* __restrict__ w, const int32_t* __restrict__ phi,
          R* __restrict__ out) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < d.N; i += gridDim.x * blockDim.x) {
        const int x = i % d.Nx, row = i - x;
        R acc = R(0);
        for (int k = 0; k <= 2 * kr; ++k) {
            const int xx = wrapi(x + k - kr, d.Nx);
            const R ind = phi[row + xx] == p ? R(1) : R(0);
            acc += w[k] * ind;
        }
        out[i] = acc;
    }
}

// y or z pass: acc += w[k] * in[shifted] in k order
template <class R, int AXIS>
__global__ void __launch_bounds__(256)
k_broad_yz(Dims d, int kr, const R* __restrict__ w, const R* __restrict__ in, R* __restrict__ out) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < d.N; i += gridDim.x * blockDim.x) {
        int z, y, x;
        decode(d, i, z, y, x);
        R acc = R(0);
        for (int k = 0; k <= 2 * kr; ++k) {
            const int j = AXIS == 1 ? lin(d, z, wrapi(y + k - kr, d.Ny), x)
                                    : lin(d, wrapi(z + k - kr, d.Nz), y, x);
            acc += w[k] * in[j];
        }
        out[i] = acc;
    }
}

// normals and curvature at the interface cells of label p
template <class R>
__global__ void __launch_bounds__(256)
k_broad_fd(Dims d, int32_t p, int nb, const int* __restrict__ band, const int32_t* __restrict__ phi,
           const R* __restrict__ sm, float dx_f, float sigma_cells, R* __restrict__ onx,
           R* __restrict__ ony, R* __restrict__ onz, R* __restrict__ okappa) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nb) return;
    const int i = band[a];
    if (phi[i] != p) return;
    int z, y, x;
    decode(d, i, z, y, x);
    const R dx = R(dx_f);
    const R idx12 = R(1) / (R(12) * dx);
    const R idx_sq = R(1) / (dx * dx);
    const R idx12_sq = idx_sq / R(12);
    const R idx144_sq = idx_sq / R(144);
    const int xo[4] = {wrapi(x - 2, d.Nx), wrapi(x - 1, d.Nx), wrapi(x + 1, d.Nx), wrapi(x + 2, d.Nx)};
    const int yo[4] = {wrapi(y - 2, d.Ny), wrapi(y - 1, d.Ny), wrapi(y + 1, d.Ny), wrapi(y + 2, d.Ny)};
    const int zo[4] = {wrapi(z - 2, d.Nz), wrapi(z - 1, d.Nz), wrapi(z + 1, d.Nz), wrapi(z + 2, d.Nz)};
    auto S = [&](int zz, int yy, int xx) { return sm[lin(d, zz, yy, xx)]; };
    const R I0 = sm[i];
    const R Ix = (S(z, y, xo[0]) - R(8) * S(z, y, xo[1]) + R(8) * S(z, y, xo[2]) - S(z, y, xo[3])) * idx12;
    const R Iy = (S(z, yo[0], x) - R(8) * S(z, yo[1], x) + R(8) * S(z, yo[2], x) - S(z, yo[3], x)) * idx12;
    const R Iz = (S(zo[0], y, x) - R(8) * S(zo[1], y, x) + R(8) * S(zo[2], y, x) - S(zo[3], y, x)) * idx12;
    const R Ixx = (-S(z, y, xo[3]) + R(16) * S(z, y, xo[2]) - R(30) * I0 + R(16) * S(z, y, xo[1]) -
                   S(z, y, xo[0])) * idx12_sq;
    const R Iyy = (-S(z, yo[3], x) + R(16) * S(z, yo[2], x) - R(30) * I0 + R(16) * S(z, yo[1], x) -
                   S(z, yo[0], x)) * idx12_sq;
    const R Izz = (-S(zo[3], y, x) + R(16) * S(zo[2], y, x) - R(30) * I0 + R(16) * S(zo[1], y, x) -
                   S(zo[0], y, x)) * idx12_sq;
    const R dc[4] = {R(1), R(-8), R(8), R(-1)};
    R sxy = 0, sxz = 0, syz = 0;
    for (int u = 0; u < 4; ++u)
        for (int v = 0; v < 4; ++v) {
            const R w = dc[u] * dc[v];
            sxy += w * S(z, yo[v], xo[u]);
            sxz += w * S(zo[v], y, xo[u]);
            syz += w * S(zo[v], yo[u], x);
        }
    const R Ixy = sxy * idx144_sq, Ixz = sxz * idx144_sq, Iyz = syz * idx144_sq;
    const R gsq = Ix * Ix + Iy * Iy + Iz * Iz;
    R nxv = 0, nyv = 0, nzv = 0, kap = 0;
    if (gsq * dx * dx > Eps<R>::grad) {
        const R inv = R(1) / dsqrt(gsq);
        nxv = -Ix * inv; nyv = -Iy * inv; nzv = -Iz * inv;
        const R trH = Ixx + Iyy + Izz;
        const R gHg = Ix * Ix * Ixx + Iy * Iy * Iyy + Iz * Iz * Izz +
                      R(2) * (Ix * Iy * Ixy + Ix * Iz * Ixz + Iy * Iz * Iyz);
        R kv = (gHg - gsq * trH) / (gsq * dsqrt(gsq));
        if (dfabs(kv * dx) < Eps<R>::kappa) kv = R(0);
        const R klim = R(1) / (R(sigma_cells) * dx);
        kap = kv < -klim ? -klim : (klim < kv ? klim : kv);
    }
    onx[a] = nxv; ony[a] = nyv; onz[a] = nzv; okappa[a] = kap;
}

// competitor and velocity of every band cell from the stored normals
template <class R>
__global__ void __launch_bounds__(256)
k_broad_finish(Dims d, gpk_model_desc model, int nb, const int* __restrict__ band,
               const int32_t* __restrict__ phi, const R* __restrict__ g, const R* __restrict__ c,
               const R* __restrict__ T, const R* __restrict__ rho, const R* __restrict__ E_el,
               const R* __restrict__ onx, const R* __restrict__ ony,
               const R* __restrict__ onz, const R* __restrict__ okappa, R* __restrict__ vraw,
               int32_t* __restrict__ phi_comp, R* __restrict__ g_comp) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nb) return;
    const int i = band[a];
    int z0, y0, x0;
    decode(d, i, z0, y0, x0);
    auto cube_at = [&](int t, int zz, int yy, int xx) {
        (void)t;
        return phi[lin(d, zz, yy, xx)];
    };
    comp_vel_cell<R, 3>(d, model, i, a, z0, y0, x0, phi[i], onx[a], ony[a], onz[a], okappa[a], cube_at,
                        phi, g, c, T, rho, E_el, vraw, phi_comp, g_comp);
}

} // namespace gpkd
