// gpk_broad.hpp — geometry of 3-D grids whose stencil radius exceeds 18
// (src/gpk_grid.cpp:44, src/gpk_geometry.cpp:185-390).
//
// For each label present on the interface: the 0/1 indicator of the label
// smoothed by the plain Gaussian (radius kr) along x, then y, then z -- each
// output a k-ordered sum of separately rounded products -- and, at the
// label's interface cells, fourth-order centred differences of the smoothed
// field give the gradient and Hessian, hence the normal (times 1/|grad|) and
// the curvature.  The reference smooths the whole grid once per label; here
// each label's passes cover only the box its interface cells and the
// difference stencil need, and all labels of a chunk run in one launch per
// pass (every band cell belongs to exactly one label, so the label order of
// the reference does not change any value).  Competitor and velocity follow
// once every label is done.
#pragma once

#include "gpk_cell.hpp"

namespace gpkd {

__global__ void k_broad_flag_labels(int nb, const int* __restrict__ band, const int32_t* __restrict__ phi,
                                    uint8_t* __restrict__ flags) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a < nb) flags[phi[band[a]]] = 1;
}

// fourth-order centred differences of the smoothed indicator S at (z, y, x):
// gradient and Hessian, hence normal (-grad / |grad|) and curvature
// (src/gpk_geometry.cpp:292-389); S(zz, yy, xx) takes wrapped coordinates
template <class R, class SF>
__device__ __forceinline__ void fd_cell(const Dims& d, int z, int y, int x, SF S, float dx_f,
                                        float sigma_cells, R& onx, R& ony, R& onz, R& okappa) {
    const R dx = R(dx_f);
    const R idx12 = R(1) / (R(12) * dx);
    const R idx_sq = R(1) / (dx * dx);
    const R idx12_sq = idx_sq / R(12);
    const R idx144_sq = idx_sq / R(144);
    const int xo[4] = {wrapi(x - 2, d.Nx), wrapi(x - 1, d.Nx), wrapi(x + 1, d.Nx), wrapi(x + 2, d.Nx)};
    const int yo[4] = {wrapi(y - 2, d.Ny), wrapi(y - 1, d.Ny), wrapi(y + 1, d.Ny), wrapi(y + 2, d.Ny)};
    const int zo[4] = {wrapi(z - 2, d.Nz), wrapi(z - 1, d.Nz), wrapi(z + 1, d.Nz), wrapi(z + 2, d.Nz)};
    const R I0 = S(z, y, x);
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
    onx = nxv; ony = nyv; onz = nzv; okappa = kap;
}

// ---------------------------------------------------------------------------
// Box-restricted passes.  Label p's smoothed field is needed only at its
// interface cells and their +-2 difference stencil: the box B (band-cell
// bounding box + 2).  The z pass therefore writes B, the y pass B widened by
// kr in z, the x pass B widened by kr in z and y; each output value is the
// same k-ordered sum as over the whole grid.  An interval that would reach
// its full periodic side becomes the whole side.  Each label's three boxes
// live contiguously in three scratch arrays at the offsets of its BroadBox.
// ---------------------------------------------------------------------------
struct BroadBox {
    int32_t label;
    int zs[2], zn[2];      // z interval: [0] the box B, [1] B widened by kr
    int ys[2], yn[2];      // y interval: [0] B, [1] B widened by kr
    int xs, xn;            // x interval of B
    long long off[3];      // start of the label's z-, y- and x-pass outputs in its chunk
};

// position of grid coordinate v inside an interval starting at s, along a
// periodic side of length n
__device__ __forceinline__ int wrap_into(int v, int s, int n) { return wrapi(v - s, n); }

// label index of flattened position g in pass `which` (boxes sorted by offset)
__device__ __forceinline__ int box_of(const BroadBox* __restrict__ bx, int nbox, int which, long long g) {
    int lo = 0, hi = nbox - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (bx[mid].off[which] <= g) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__global__ void k_broad_label_index(int ng, const int* __restrict__ llist, int* __restrict__ lidx) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < ng) lidx[llist[t]] = t;
}

__global__ void k_broad_bbox_init(int ng, int* __restrict__ bb) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < 6 * ng) bb[e] = (e & 1) ? -1 : 0x7fffffff;
}

// bounding box of each listed label's interface cells: bb[6 t .. 6 t + 5] =
// zmin, zmax, ymin, ymax, xmin, xmax.  Consecutive band cells mostly share a
// label, so each wavefront reduces per label before its atomics.
__device__ __forceinline__ int wave_min_i32(int v) {
    for (int o = 32; o > 0; o >>= 1) { const int u = __shfl_xor(v, o); v = u < v ? u : v; }
    return v;
}
__device__ __forceinline__ int wave_max_i32(int v) {
    for (int o = 32; o > 0; o >>= 1) { const int u = __shfl_xor(v, o); v = u > v ? u : v; }
    return v;
}

__global__ void k_broad_bbox(Dims d, int nb, const int* __restrict__ band, const int32_t* __restrict__ phi,
                             const int* __restrict__ lidx, int* __restrict__ bb) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    int t = -1, z = 0, y = 0, x = 0;
    if (a < nb) {
        const int i = band[a];
        decode(d, i, z, y, x);
        t = lidx[phi[i]];
    }
    unsigned long long todo = __ballot(t >= 0);
    while (todo) {
        const int leader = __ffsll((long long)todo) - 1;
        const int tt = __shfl(t, leader);
        const bool mine = t == tt;
        const int zmin = wave_min_i32(mine ? z : 0x7fffffff), zmax = wave_max_i32(mine ? z : -1);
        const int ymin = wave_min_i32(mine ? y : 0x7fffffff), ymax = wave_max_i32(mine ? y : -1);
        const int xmin = wave_min_i32(mine ? x : 0x7fffffff), xmax = wave_max_i32(mine ? x : -1);
        if (lane == leader) {
            int* b = bb + 6 * tt;
            atomicMin(b + 0, zmin); atomicMax(b + 1, zmax);
            atomicMin(b + 2, ymin); atomicMax(b + 3, ymax);
            atomicMin(b + 4, xmin); atomicMax(b + 5, xmax);
        }
        todo &= ~__ballot(mine);
    }
}

template <class R>
__global__ void __launch_bounds__(256)
k_broad_box_x(Dims d, int kr, int nbox, const BroadBox* __restrict__ bx, long long total,
              const R* __restrict__ w, const int32_t* __restrict__ phi, R* __restrict__ outx) {
    for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < total;
         g += (long long)gridDim.x * blockDim.x) {
        const BroadBox& B = bx[box_of(bx, nbox, 2, g)];
        const long long l = g - B.off[2];
        const int lx = int(l % B.xn), ly = int((l / B.xn) % B.yn[1]), lz = int(l / ((long long)B.xn * B.yn[1]));
        const int z = wrapi(B.zs[1] + lz, d.Nz), y = wrapi(B.ys[1] + ly, d.Ny), x = wrapi(B.xs + lx, d.Nx);
        const int32_t* row = phi + lin(d, z, y, 0);
        int xx = wrapi(x - kr, d.Nx);
        R acc = R(0);
        for (int k = 0; k <= 2 * kr; ++k) {
            const R ind = row[xx] == B.label ? R(1) : R(0);
            acc += w[k] * ind;
            if (++xx == d.Nx) xx = 0;
        }
        outx[g] = acc;
    }
}

template <class R>
__global__ void __launch_bounds__(256)
k_broad_box_y(Dims d, int kr, int nbox, const BroadBox* __restrict__ bx, long long total,
              const R* __restrict__ w, const R* __restrict__ inx, R* __restrict__ outy) {
    for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < total;
         g += (long long)gridDim.x * blockDim.x) {
        const BroadBox& B = bx[box_of(bx, nbox, 1, g)];
        const long long l = g - B.off[1];
        const int lx = int(l % B.xn), ly = int((l / B.xn) % B.yn[0]), lz = int(l / ((long long)B.xn * B.yn[0]));
        const int y = wrapi(B.ys[0] + ly, d.Ny);
        const R* src = inx + B.off[2] + (long long)lz * B.yn[1] * B.xn + lx;
        // the widened interval is whole (index wraps) or starts kr before B
        const bool whole = B.yn[1] == d.Ny;
        int yy = whole ? wrapi(y - kr, d.Ny) : ly;
        R acc = R(0);
        for (int k = 0; k <= 2 * kr; ++k) {
            acc += w[k] * src[(long long)yy * B.xn];
            if (++yy == d.Ny && whole) yy = 0;
        }
        outy[g] = acc;
    }
}

template <class R>
__global__ void __launch_bounds__(256)
k_broad_box_z(Dims d, int kr, int nbox, const BroadBox* __restrict__ bx, long long total,
              const R* __restrict__ w, const R* __restrict__ iny, R* __restrict__ outz) {
    for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < total;
         g += (long long)gridDim.x * blockDim.x) {
        const BroadBox& B = bx[box_of(bx, nbox, 0, g)];
        const long long l = g - B.off[0];
        const int lx = int(l % B.xn), ly = int((l / B.xn) % B.yn[0]), lz = int(l / ((long long)B.xn * B.yn[0]));
        const int z = wrapi(B.zs[0] + lz, d.Nz);
        const long long plane = (long long)B.yn[0] * B.xn;
        const R* src = iny + B.off[1] + (long long)ly * B.xn + lx;
        const bool whole = B.zn[1] == d.Nz;
        int zz = whole ? wrapi(z - kr, d.Nz) : lz;
        R acc = R(0);
        for (int k = 0; k <= 2 * kr; ++k) {
            acc += w[k] * src[zz * plane];
            if (++zz == d.Nz && whole) zz = 0;
        }
        outz[g] = acc;
    }
}

// normals and curvature of the band cells whose label lies in the chunk
// [c0, c1) of the label list, from their label's smoothed box
template <class R>
__global__ void __launch_bounds__(256)
k_broad_box_fd(Dims d, int nb, const int* __restrict__ band, const int32_t* __restrict__ phi,
               const int* __restrict__ lidx, int c0, int c1, const BroadBox* __restrict__ bx,
               const R* __restrict__ sz, float dx_f, float sigma_cells, R* __restrict__ onx,
               R* __restrict__ ony, R* __restrict__ onz, R* __restrict__ okappa) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nb) return;
    const int i = band[a];
    const int t = lidx[phi[i]];
    if (t < c0 || t >= c1) return;
    const BroadBox& B = bx[t];
    int z, y, x;
    decode(d, i, z, y, x);
    const R* base = sz + B.off[0];
    auto S = [&](int zz, int yy, int xx) {
        const int lz = wrap_into(zz, B.zs[0], d.Nz), ly = wrap_into(yy, B.ys[0], d.Ny),
                  lx = wrap_into(xx, B.xs, d.Nx);
        return base[((long long)lz * B.yn[0] + ly) * B.xn + lx];
    };
    fd_cell<R>(d, z, y, x, S, dx_f, sigma_cells, onx[a], ony[a], onz[a], okappa[a]);
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
