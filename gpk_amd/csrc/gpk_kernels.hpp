// gpk_kernels.hpp — HIP kernels of the GPK per-step kinetics (gfx950).
//
// Data layout in HBM: full-grid SoA fields (phi int32, alpha/c/T/g/g_comp
// real, mask uint8, phi_comp int32) plus band-compact arrays indexed by the
// band slot a (sorted interface list `band[a]`, inverse map `slot[cell]`):
// normals, curvature, raw and projected velocity.  All reductions that feed
// a decision are exact (integer, max) or sequential per output element, so
// every kernel is deterministic and reproduces the CPU restatement's
// operation order.
#pragma once
#include <type_traits>

#include "gpk_device.hpp"
#include "gpk_cell.hpp"
#include "gpk_tiles.hpp"

namespace gpkd {

// RETRY: undecided after an inconclusive bounds search; it is queued again
// only when every earlier transaction on its labels is decided
enum : int { UNDEC = 0, ACCEPT = 1, REJECT = 2, QUEUED = 3, RETRY = 4 };
__device__ __forceinline__ bool undecided(int s) { return s != ACCEPT && s != REJECT; }

// LDS written by some lanes of a wavefront and then read by others
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int ld_acq(const int* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_rel(int* p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------
// Interface mask (src/gpk_geometry.cpp:42-82): mask[i] = any of the six
// periodic face neighbours carries a different label.
// ---------------------------------------------------------------------------
__global__ void k_mask(Dims d, const int32_t* __restrict__ phi, uint8_t* __restrict__ mask) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < d.N; i += gridDim.x * blockDim.x) {
        int f[6];
        faces(d, i, f);
        const int32_t v = phi[i];
        mask[i] = uint8_t((v != phi[f[0]]) | (v != phi[f[1]]) | (v != phi[f[2]]) |
                          (v != phi[f[3]]) | (v != phi[f[4]]) | (v != phi[f[5]]));
    }
}

// The same mask four cells at a time along x (Nx a multiple of 4): the four
// labels of a group and of its four face-neighbour groups in y and z are
// 16-byte loads, the two x neighbours outside the group two 4-byte loads, and
// the four mask bytes one 4-byte store.
__global__ void k_mask4(Dims d, const int32_t* __restrict__ phi, uint8_t* __restrict__ mask) {
    const int qx = d.Nx >> 2;
    const int nq = d.N >> 2;
    for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < nq; g += gridDim.x * blockDim.x) {
        const int row = g / qx;                   // z * Ny + y
        const int x0 = (g - row * qx) << 2;
        const int z = row / d.Ny, y = row - z * d.Ny;
        const int ym = y == 0 ? d.Ny - 1 : y - 1, yp = y == d.Ny - 1 ? 0 : y + 1;
        const int zm = z == 0 ? d.Nz - 1 : z - 1, zp = z == d.Nz - 1 ? 0 : z + 1;
        const int base = row * d.Nx;
        const int4 c = *reinterpret_cast<const int4*>(phi + base + x0);
        const int4 a = *reinterpret_cast<const int4*>(phi + (z * d.Ny + ym) * d.Nx + x0);
        const int4 b = *reinterpret_cast<const int4*>(phi + (z * d.Ny + yp) * d.Nx + x0);
        const int4 e = *reinterpret_cast<const int4*>(phi + (zm * d.Ny + y) * d.Nx + x0);
        const int4 f = *reinterpret_cast<const int4*>(phi + (zp * d.Ny + y) * d.Nx + x0);
        const int32_t l = phi[base + (x0 == 0 ? d.Nx - 1 : x0 - 1)];
        const int32_t r = phi[base + (x0 + 4 == d.Nx ? 0 : x0 + 4)];
        const uint32_t m0 = (c.x != l) | (c.x != c.y) | (c.x != a.x) | (c.x != b.x) | (c.x != e.x) | (c.x != f.x);
        const uint32_t m1 = (c.y != c.x) | (c.y != c.z) | (c.y != a.y) | (c.y != b.y) | (c.y != e.y) | (c.y != f.y);
        const uint32_t m2 = (c.z != c.y) | (c.z != c.w) | (c.z != a.z) | (c.z != b.z) | (c.z != e.z) | (c.z != f.z);
        const uint32_t m3 = (c.w != c.z) | (c.w != r) | (c.w != a.w) | (c.w != b.w) | (c.w != e.w) | (c.w != f.w);
        *reinterpret_cast<uint32_t*>(mask + base + x0) = m0 | (m1 << 8) | (m2 << 16) | (m3 << 24);
    }
}

// ---------------------------------------------------------------------------
// Ordered stream compaction: out = indices i (ascending) with flags[i] != 0.
// Block b owns items [4096 b, 4096 b + 4096), thread t the 16 consecutive
// items at 16 t; count per block, one-block exclusive scan of the block
// counts, then each thread writes its set items at its block offset plus the
// prefix of its block's thread counts.
// ---------------------------------------------------------------------------
constexpr int CMP_T = 256, CMP_V = 16, CMP_CHUNK = CMP_T * CMP_V;

__device__ __forceinline__ uint32_t flag_bits16(const uint8_t* __restrict__ f, long long base, long long n) {
    uint32_t bits = 0;
    if (base + CMP_V <= n) {
        const uint4 v = *reinterpret_cast<const uint4*>(f + base);
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int j = 0; j < CMP_V; ++j) bits |= ((w[j >> 2] >> (8 * (j & 3))) & 0xFFu) ? (1u << j) : 0u;
    } else {
        for (int j = 0; j < CMP_V; ++j)
            if (base + j < n && f[base + j]) bits |= 1u << j;
    }
    return bits;
}

// inclusive prefix sum over the block's threads in thread order; *total gets
// the block sum (LDS: one int per wavefront + the total)
template <int NT>
__device__ __forceinline__ int block_incl_scan(int v, int* wsum, int* total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int u = __shfl_up(v, off);
        if (lane >= off) v += u;
    }
    if (lane == 63) wsum[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int k = 0; k < NT / 64; ++k) { const int t = wsum[k]; wsum[k] = acc; acc += t; }
        *total = acc;
    }
    __syncthreads();
    return v + wsum[w];
}

__global__ void __launch_bounds__(CMP_T) k_cmp_count(long long n, const uint8_t* __restrict__ f,
                                                     int* __restrict__ bcnt) {
    __shared__ int wsum[CMP_T / 64], tot;
    const long long base = (long long)blockIdx.x * CMP_CHUNK + (long long)threadIdx.x * CMP_V;
    const int c = __popc(flag_bits16(f, base, n));
    (void)block_incl_scan<CMP_T>(c, wsum, &tot);
    if (threadIdx.x == 0) bcnt[blockIdx.x] = tot;
}

// one block: bcnt (counts) -> exclusive offsets, *total = sum
__global__ void __launch_bounds__(1024) k_cmp_scan(int nb, int* __restrict__ bcnt, int* __restrict__ total) {
    __shared__ int wsum[16], tot;
    constexpr int PER = 32;
    if (nb <= 1024 * PER) {
        // each thread a contiguous run of at most PER counts, loaded together:
        // one block scan of the run sums, then the run's prefixes written
        const int per = (nb + 1023) / 1024;
        const int b0 = int(threadIdx.x) * per;
        int v[PER];
        int sum = 0;
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            v[u] = (u < per && b0 + u < nb) ? bcnt[b0 + u] : 0;
            sum += v[u];
        }
        const int incl = block_incl_scan<1024>(sum, wsum, &tot);
        int run = incl - sum;
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            if (u < per && b0 + u < nb) bcnt[b0 + u] = run;
            run += v[u];
        }
        if (threadIdx.x == 0) *total = tot;
        return;
    }
    int running = 0;
    for (int b0 = 0; b0 < nb; b0 += 1024) {
        const int b = b0 + threadIdx.x;
        const int v = b < nb ? bcnt[b] : 0;
        const int incl = block_incl_scan<1024>(v, wsum, &tot);
        if (b < nb) bcnt[b] = running + incl - v;
        running += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = running;
}

// (slot, when given: slot[out[o]] = o, the inverse map of the list)
__global__ void __launch_bounds__(CMP_T) k_cmp_scatter(long long n, const uint8_t* __restrict__ f,
                                                       const int* __restrict__ boff, int* __restrict__ out,
                                                       int* __restrict__ slot) {
    __shared__ int wsum[CMP_T / 64], tot;
    const long long base = (long long)blockIdx.x * CMP_CHUNK + (long long)threadIdx.x * CMP_V;
    uint32_t bits = flag_bits16(f, base, n);
    const int c = __popc(bits);
    int o = boff[blockIdx.x] + block_incl_scan<CMP_T>(c, wsum, &tot) - c;
    while (bits) {
        const int j = __ffs(bits) - 1;
        out[o] = int(base + j);
        if (slot) slot[base + j] = o;
        ++o;
        bits &= bits - 1u;
    }
}

// several device ints set by one launch
struct IntSets {
    static constexpr int CAP = 8;
    int* p[CAP];
    int v[CAP];
    int n;
};
__global__ void k_set_ints(IntSets a) {
    if (int(threadIdx.x) < a.n) a.p[threadIdx.x][0] = a.v[threadIdx.x];
}

template <class R, int DIM, bool kDirect>
__global__ void __launch_bounds__(256)
k_geometry(Dims d, GeomParams gp, gpk_model_desc model, int nband,
           const int* __restrict__ band, const int32_t* __restrict__ phi,
           const R* __restrict__ w0, const R* __restrict__ w1, const R* __restrict__ w2,
           const R* __restrict__ tab, RowLutDesc lutd,
           const R* __restrict__ g, const R* __restrict__ c, const R* __restrict__ T,
           const R* __restrict__ rho, const R* __restrict__ E_el,
           R* __restrict__ onx, R* __restrict__ ony, R* __restrict__ onz,
           R* __restrict__ okappa, R* __restrict__ vraw,
           int32_t* __restrict__ phi_comp, R* __restrict__ g_comp) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nband) return;
    geometry_cell<R, DIM, kDirect>(d, gp, model, a, band[a], phi, w0, w1, w2, tab, lutd, g, c, T, rho, E_el, onx,
                                   ony, onz, okappa, vraw, phi_comp, g_comp);
}

// ---------------------------------------------------------------------------
// Pair projection (src/gpk_physics.cpp:152-193) + |v| max for the timestep
// (src/gpk_simulation.cpp:11-25).  Same-pair, orientation-signed Gaussian
// average over the (2r+1)^d window, FP64 accumulation in window order.
// ---------------------------------------------------------------------------
template <class R, int DIM>
__global__ void __launch_bounds__(256)
k_project(Dims d, int r, int nband, const int* __restrict__ band, const int* __restrict__ slot,
          const uint8_t* __restrict__ mask, const int32_t* __restrict__ phi,
          const int32_t* __restrict__ phi_comp, const double* __restrict__ wtab,
          const R* __restrict__ vraw, R* __restrict__ vout,
          typename UBits<R>::T* __restrict__ vmax_bits) {
    using U = typename UBits<R>::T;
    __shared__ U bmax;
    if (threadIdx.x == 0) bmax = 0;
    __syncthreads();
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    U mb = 0;
    if (a < nband) {
        const R out = project_value<R, DIM>(d, r, a, band[a], slot, mask, phi, phi_comp, wtab, vraw);
        vout[a] = out;
        mb = abs_bits(out);
    }
    block_max_to<U>(mb, bmax, vmax_bits);
}

template <class R>
__global__ void k_vmax_masked(int nband, const int* __restrict__ band, const uint8_t* __restrict__ mask,
                              const R* __restrict__ vn, typename UBits<R>::T* __restrict__ vmax_bits) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a < nband && mask[band[a]]) atomic_max_abs(vmax_bits, vn[a]);
}

// ---------------------------------------------------------------------------
// Projection for radii other than the tiled one (3-D sigma above 2.27 gives
// r = 4..9): full-grid pair keys and velocities are written once per step,
// then each band cell walks its window rows branch-free in the reference
// order.  Same sums as project_value (labels < 2^15: the key packs the pair).
// ---------------------------------------------------------------------------
template <class R>
__global__ void k_proj_keys(int n, const uint8_t* __restrict__ mask, const int32_t* __restrict__ phi,
                            const int32_t* __restrict__ phi_comp, const int* __restrict__ slot,
                            const R* __restrict__ vraw, uint32_t* __restrict__ key, R* __restrict__ vfull) {
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        uint32_t k = 0xFFFFFFFFu;
        R v = R(0);
        if (mask[j]) {
            const uint32_t p = uint32_t(phi[j]), q = uint32_t(phi_comp[j]);
            if (p != q) {
                const uint32_t lo = p < q ? p : q, hi = p < q ? q : p;
                k = (((lo << 16) | hi) << 1) | (p == lo ? 0u : 1u);
                v = vraw[slot[j]];
            }
        }
        key[j] = k;
        vfull[j] = v;
    }
}

template <class R, int DIM>
__global__ void __launch_bounds__(256)
k_project_keyed(Dims d, int r, int nband, const int* __restrict__ band, const uint32_t* __restrict__ key,
                const R* __restrict__ vfull, const R* __restrict__ vraw, const double* __restrict__ wtab,
                R* __restrict__ vout, typename UBits<R>::T* __restrict__ vmax_bits) {
    using U = typename UBits<R>::T;
    __shared__ U bmax;
    if (threadIdx.x == 0) bmax = 0;
    __syncthreads();
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    U mb = 0;
    if (a < nband) {
        const int i = band[a];
        const uint32_t ki = key[i];
        R out = vraw[a];
        if (ki != 0xFFFFFFFFu) {
            const uint32_t pair = ki >> 1;
            const R oi = (ki & 1u) ? R(-1) : R(1);
            int z0, y0, x0;
            decode(d, i, z0, y0, x0);
            double num = 0.0, den = 0.0;
            const int dzb = DIM == 3 ? -r : 0, dze = DIM == 3 ? r : 0;
            const int xstart = wrapi(x0 - r, d.Nx);
            for (int dz = dzb; dz <= dze; ++dz) {
                const int z = wrapi(z0 + dz, d.Nz);
                for (int dy = -r; dy <= r; ++dy) {
                    const int y = wrapi(y0 + dy, d.Ny);
                    const uint32_t* krow = key + lin(d, z, y, 0);
                    const R* vrow = vfull + lin(d, z, y, 0);
                    const double* wr = wtab + dy * dy + dz * dz;
                    int x = xstart;
                    for (int ddx = -r; ddx <= r; ++ddx) {
                        const uint32_t kj = krow[x];
                        const bool hit = (kj >> 1) == pair;
                        const double w = wr[ddx * ddx];
                        // (-w) v == -(w v) exactly; a miss adds +-0
                        const double t = __longlong_as_double(__double_as_longlong(w * double(vrow[x])) ^
                                                              (static_cast<long long>(kj & 1u) << 63));
                        num = add_if(hit, t, num);
                        den = add_if(hit, w, den);
                        if (++x == d.Nx) x = 0;
                    }
                }
            }
            if (den > 0.0) out = oi * R(num / den);
        }
        vout[a] = out;
        mb = abs_bits(out);
    }
    block_max_to<U>(mb, bmax, vmax_bits);
}

// resident / best-candidate Gaussian scores of the balanced remap
// (src/gpk_transitions.cpp:53-91) from LDS
// Remap-score cache: a band cell's scores (gain, best label) depend only on
// the labels of its (2r+1)^d window, so they are kept per cell and reused
// until a label inside that window changes.  Every label write marks the
// window-sized box around the written cell dirty; a scoring pass recomputes
// dirty or unscored cells and a cleanup pass drops the cache of cells that
// became dirty while off the band.
__global__ void k_score_cache_cleanup(int n, uint8_t* __restrict__ valid, uint8_t* __restrict__ dirty) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        if (dirty[i]) { dirty[i] = 0; valid[i] = 0; }
}

// scalar block shared by the step kernels (device memory)
struct Scalars {
    unsigned long long vmax_bits;   // |v| max as bits (float bits in low word for FP32)
    double dt;                      // timestep of the current step
    double time;                    // simulation time
    int band;                       // |A| of the geometry band
    int band2;                      // size of the latest recomputed band
    int ncand;                      // flip candidates
    int undecided;                  // undecided transactions
    int qcount[2][8];               // search queues per band slab (double-buffered per round)
    int ovcount[2];                 // wavefront searches handed to the workgroup search, per round
    int fbcount;                    // BFS fallback queue
    int accepted, deferred;
    int nprop;                      // remap proposals
    int npairs;
    int exchanges;
    int eroded;
    int layer_changed;
    int nsel;                       // generic selection count
    int present, unresolved;        // label census
    int changed;                    // a component-labelling round lowered some id
    int guarded;                    // flip decisions taken again in FP64 (FP32 mode)
    int bfs_done;                   // workgroups of k_flip_bfs finished (last one runs the fallback)
};

template <class R>
__global__ void k_vmax_init(Scalars* s) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        if (sizeof(R) == 4) s->vmax_bits = __float_as_uint(Eps<float>::vmax0);
        else s->vmax_bits = (unsigned long long)__double_as_longlong(Eps<double>::vmax0);
    }
}

// dt = min(request or stable, cap) (src/gpk_simulation.cpp:31-40)
template <class R>
__global__ void k_dt(Scalars* s, R dt_req, R cap, float dx_f, float D_f, int dim3) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    R vmax;
    if (sizeof(R) == 4) vmax = R(__uint_as_float((unsigned int)s->vmax_bits));
    else vmax = R(__longlong_as_double((long long)s->vmax_bits));
    const R dx = R(dx_f), D = R(D_f);
    const R nb = dim3 ? R(6) : R(4);
    const R dti = R(Eps<R>::c_iface) * dx / vmax;
    const R dtd = D > R(0) ? R(Eps<R>::c_diff) * dx * dx / (nb * D) : R(Eps<R>::big);
    const R stable = dtd < dti ? dtd : dti;
    R dt = dt_req < R(0) ? stable : (stable < dt_req ? stable : dt_req);
    dt = cap < dt ? cap : dt;
    s->dt = double(dt);
}

__global__ void k_time_advance(Scalars* s) {
    if (threadIdx.x == 0 && blockIdx.x == 0) s->time += s->dt;
}

// ---------------------------------------------------------------------------
// alpha integration and flip candidates (src/gpk_transitions.cpp:251-292)
// ---------------------------------------------------------------------------
template <class R>
__global__ void k_alpha(int nband, const int* __restrict__ band, const Scalars* __restrict__ s,
                        float dx_f, const R* __restrict__ vn, R* __restrict__ alpha,
                        const R* __restrict__ T, const int32_t* __restrict__ phi,
                        const int32_t* __restrict__ phi_comp, uint8_t* __restrict__ cflag,
                        R* __restrict__ anew_s, int noise, float noise_amp, uint64_t step_id,
                        int guard, Scalars* __restrict__ sw) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nband) return;
    const int i = band[a];
    const R dt = R(s->dt);
    const R idt = dt / R(dx_f);
    R v = vn[a];
    if (noise) {
        const R Tm = T[i] < R(1) ? R(1) : T[i];
        v += R(noise_amp) * dsqrt(Tm) * R(normal_dev(i, step_id, 0xabcddcbaULL));
    }
    R an;
    bool guarded;
    const bool below = alpha_step<R>(alpha[i], v, idt, double(dt), double(R(dx_f)), guard != 0, an, guarded);
    if (guarded) atomicAdd(&sw->guarded, 1);
    const bool cand = below && phi_comp[i] != phi[i];
    if (!cand) alpha[i] = an;
    cflag[a] = cand ? 1 : 0;
    anew_s[a] = an;
}

// noise amplitude needs dt on the host side of the formula; computed on device
template <class R>
__global__ void k_noise_amp(const Scalars* s, float dx_f, float M0, float kB, float scale,
                            float* out) {
    if (threadIdx.x || blockIdx.x) return;
    const float dt = float(R(s->dt));
    const float cv = dx_f * dx_f * dx_f;
    const float inv_dt = dt > 0.f ? 1.f / dt : 0.f;
    const float m = cv < 1e-30f ? 1e-30f : cv;
    *out = scale * sqrtf(2.f * M0 * kB * inv_dt / m);
}

// candidate table: cell, old/new label, new orientation, alpha_new.  With
// ordered transactions (Pk given) each candidate's cell holds ~k in phi until
// k_flip_apply writes its label back: labels lie in [0, 2^26), so a negative
// label word names the candidate, and a neighbour's view costs one load of
// phi plus, for candidate cells only, loads from the candidate table.
template <class R>
__global__ void k_cand_init(int m, const int* __restrict__ sel, const int* __restrict__ band,
                            int32_t* __restrict__ phi, const int32_t* __restrict__ phi_comp,
                            const R* __restrict__ g_comp, const R* __restrict__ anew_s,
                            int* __restrict__ Ccell, int* __restrict__ Cold, int* __restrict__ Cnew,
                            R* __restrict__ Cg, R* __restrict__ Canew, int* __restrict__ state,
                            unsigned long long* __restrict__ Pk) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    const int sa = sel[k];
    const int cell = band[sa];
    const int32_t ol = phi[cell], nl = phi_comp[cell];
    Ccell[k] = cell; Cold[k] = ol; Cnew[k] = nl;
    Cg[k] = g_comp[cell]; Canew[k] = anew_s[sa];
    state[k] = UNDEC;
    if (Pk) {
        phi[cell] = ~k;
        Pk[2 * k] = ((unsigned long long)(uint32_t)ol << 32) | (uint32_t)k;
        Pk[2 * k + 1] = ((unsigned long long)(uint32_t)nl << 32) | (uint32_t)k;
    }
}

__global__ void k_cand_pos(int n2, const unsigned long long* __restrict__ Ps,
                           const int* __restrict__ Cold, int* __restrict__ pos_old,
                           int* __restrict__ pos_new) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n2) return;
    const unsigned long long key = Ps[p];
    const int k = int(uint32_t(key));
    const int32_t lab = int32_t(uint32_t(key >> 32));
    if (lab == Cold[k]) pos_old[k] = p; else pos_new[k] = p;
}

// label of `cell` as seen by transaction k (earlier accepted flips applied);
// phi holds ~j at the cell of candidate j while the transactions run
__device__ __forceinline__ int32_t view_label(int cell, int k, const int32_t* phi, const int* state,
                                              const int* Cold, const int* Cnew) {
    const int32_t v = phi[cell];
    if (v >= 0) return v;
    const int j = ~v;
    if (j < k && ld_acq(state + j) == ACCEPT) return Cnew[j];
    return Cold[j];
}

// Views of label L bounding every outcome of the earlier transactions not
// yet decided (each can only add its cell to L or remove it from L):
//   VIEW_UPPER: an undecided earlier transaction into L counts as accepted,
//               so the cells read as L include every cell of L at time k;
//   VIEW_LOWER: an undecided earlier transaction out of L counts as
//               accepted, so every cell read as L is a cell of L at time k.
// Face connectivity grows with the cell set: branches disconnected in the
// upper view are disconnected at time k, connected in the lower view are
// connected at time k.  A transaction decided while a search runs moves a
// cell from its bound value to its exact one, which keeps both properties.
enum : int { VIEW_EXACT = 0, VIEW_UPPER = 1, VIEW_LOWER = 2 };
__device__ __forceinline__ int32_t view_label_b(int mode, int cell, int k, int32_t L, const int32_t* phi,
                                                const int* state, const int* Cold, const int* Cnew) {
    const int32_t v = phi[cell];
    if (v >= 0) return v;
    const int j = ~v;
    if (j < k) {
        const int sj = ld_acq(state + j);
        if (sj == ACCEPT) return Cnew[j];
        if (mode == VIEW_UPPER && sj != REJECT && Cnew[j] == L) return L;
        if (mode == VIEW_LOWER && sj != REJECT && Cold[j] == L) return -1;
    }
    return Cold[j];
}

// The tests of one queued transaction, written as one loop around a single
// search call so that the search is inlined once (mode is a run-time view).
// Exact turns test the old label (centre excluded), then the new label,
// stopping at the first not connected.  Bounds turns test each label in the
// upper view and, when that connects, in the lower view: one test
// disconnected in the upper view refuses, all connected in the lower view
// accept, anything else waits (RETRY).  A search that outgrew its visited set
// (-1) makes the verdict -1 (searched again by a larger search), or, with
// overflow_waits, leaves a bounds turn waiting.  search(mode, t) tests the
// old label (t == 0) or the new label (t == 1).
template <class F>
__device__ __forceinline__ int flip_verdict(int flags, bool overflow_waits, F&& search) {
    const bool bounds = (flags & 4) != 0;
    bool fail = false, open = false, big = false;
    int ok = 1, t = 0, phase = 0;
    while (t < 2) {
        if (!(flags & (1 << t))) { ++t; continue; }
        const int mode = !bounds ? VIEW_EXACT : (phase == 0 ? VIEW_UPPER : VIEW_LOWER);
        const int r = search(mode, t);
        if (!bounds) {
            ok = r;
            if (ok != 1) break;
            ++t;
        } else if (phase == 0) {
            if (r == 0) { fail = true; break; }
            if (r < 0) {
                if (overflow_waits) open = true;
                else big = true;
                ++t;
            } else {
                phase = 1;
            }
        } else {
            if (r < 0 && !overflow_waits) big = true;
            else if (r != 1) open = true;
            phase = 0;
            ++t;
        }
    }
    if (!bounds) return ok < 0 ? -1 : (ok ? ACCEPT : REJECT);
    return fail ? REJECT : (big ? -1 : (open ? RETRY : ACCEPT));
}

struct FlipCtx {
    Dims d;
    int dim3;
    int m;
    const int* Ccell; const int* Cold; const int* Cnew;
    int* state;
    const int32_t* phi;
    const unsigned long long* Ps;
    const int* pos_old; const int* pos_new;
    const int* count;          // label counts at the start of the flip pass
    int* queue[2]; int* qflags;
    int slab;                  // candidates per band slab: queue segment x holds slab x's entries
    int wave_cap, wg_cap;      // visited cells before a wavefront / workgroup search hands over
    Scalars* s;
};

// a transaction's verdict, counted when it is taken so that the host's
// last look at the undecided count also reads the accepted / refused totals
__device__ __forceinline__ void decide(FlipCtx& c, int k, int v) {
    st_rel(c.state + k, v);
    atomicAdd(v == ACCEPT ? &c.s->accepted : &c.s->deferred, 1);
    atomicSub(&c.s->undecided, 1);
}

// earlier transactions touching label L (the label-ordered prefix of Ps up
// to pos): count change of the accepted ones, and the undecided ones split
// into those that would add a cell to L and those that would remove one
// The earlier candidates on label L are the run of Ps (sorted by label,
// then index) that ends at pos - 1.  Both walks below read it backwards in
// steps of 8 entries whose keys, then states and new labels, are loaded
// together, and handle them in order; an entry before the run (or before
// index 0) reads as another label, so nothing past the run is used.
constexpr int PREFIX_STEP = 8;
__device__ __forceinline__ void prefix_keys(const FlipCtx& c, int q0, unsigned long long key[PREFIX_STEP]) {
#pragma unroll
    for (int u = 0; u < PREFIX_STEP; ++u) key[u] = q0 - u >= 0 ? c.Ps[q0 - u] : ~0ull;
}
__device__ __forceinline__ bool key_on(unsigned long long key, int32_t L) {
    return int32_t(uint32_t(key >> 32)) == L;   // ~0 reads as label -1, never a label
}

struct LabelPrefix { int delta, add, rem; };
__device__ __forceinline__ LabelPrefix label_prefix(const FlipCtx& c, int pos, int32_t L) {
    LabelPrefix r{0, 0, 0};
    for (int q0 = pos - 1; q0 >= 0; q0 -= PREFIX_STEP) {
        unsigned long long key[PREFIX_STEP];
        prefix_keys(c, q0, key);
        int sj[PREFIX_STEP], cn[PREFIX_STEP];
#pragma unroll
        for (int u = 0; u < PREFIX_STEP; ++u) {
            sj[u] = REJECT; cn[u] = -1;
            if (key_on(key[u], L)) {
                const int j = int(uint32_t(key[u]));
                sj[u] = ld_acq(c.state + j);
                cn[u] = c.Cnew[j];
            }
        }
#pragma unroll
        for (int u = 0; u < PREFIX_STEP; ++u) {
            if (!key_on(key[u], L)) return r;
            const bool adds = cn[u] == L;
            if (sj[u] == ACCEPT) r.delta += adds ? 1 : -1;
            else if (sj[u] != REJECT) { if (adds) ++r.add; else ++r.rem; }
        }
    }
    return r;
}

// all earlier transactions touching label L decided?
__device__ __forceinline__ bool label_prefix_ready(const FlipCtx& c, int pos, int32_t L) {
    for (int q0 = pos - 1; q0 >= 0; q0 -= PREFIX_STEP) {
        unsigned long long key[PREFIX_STEP];
        prefix_keys(c, q0, key);
        int sj[PREFIX_STEP];
#pragma unroll
        for (int u = 0; u < PREFIX_STEP; ++u)
            sj[u] = key_on(key[u], L) ? ld_acq(c.state + int(uint32_t(key[u]))) : REJECT;
#pragma unroll
        for (int u = 0; u < PREFIX_STEP; ++u) {
            if (!key_on(key[u], L)) return true;
            if (undecided(sj[u])) return false;
        }
    }
    return true;
}

// One round of the ordered flip transaction (src/gpk_transitions.cpp:288-313,
// certificate src/gpk_connectivity.cpp:175-194).  A candidate's topological
// numbers are taken once every earlier candidate in its 3^d cube sharing
// one of its labels is decided.  Its count rule and connectivity searches
// read the labels at its turn in the serial order; while earlier candidates
// on the tested label are undecided, they are bounded instead:
//   - the count of the old label lies in [decided - removals, decided +
//     additions]: a bound above 1 refuses, a bound at most 1 passes;
//   - a search is queued for k_flip_bfs on the bounding views (flag 4).
// A bounds search that settles nothing leaves the candidate RETRY, queued
// again (exactly) once every earlier candidate on its labels is decided.
constexpr int FLIP_RESPEC = 4;

// a verdict taken where another thread may look at the same candidate: only
// the thread whose claim succeeds counts it
__device__ __forceinline__ int decide_claim(FlipCtx& c, int k, int seen, int v) {
    if (atomicCAS(c.state + k, seen, v) != seen) return 0;
    atomicAdd(v == ACCEPT ? &c.s->accepted : &c.s->deferred, 1);
    atomicSub(&c.s->undecided, 1);
    return 1;
}

// One look at candidate k in round `round`: 0 nothing (decided already,
// waiting, or claimed elsewhere), 1 decided here, 2 claimed for a search
// (state QUEUED, qflags written, *prev = the state it had); the caller runs
// the search or queues it.
__device__ int flip_try(FlipCtx& c, int k, int round, int* prev) {
    const int sk = ld_acq(c.state + k);
    *prev = sk;
    if (sk != UNDEC && sk != RETRY) return 0;
    const Dims& d = c.d;
    const int cell = c.Ccell[k];
    const int32_t ol = c.Cold[k], nl = c.Cnew[k];
    int z0, y0, x0;
    decode(d, cell, z0, y0, x0);
    // the label words of the whole 3^d cube (position t = (dz+1)*9 +
    // (dy+1)*3 + (dx+1); 2-D: (dy+1)*3 + (dx+1)) are loaded together first
    const int centre = c.dim3 ? 13 : 4;
    const int npos = c.dim3 ? 27 : 9;
    int32_t lw[27];
#pragma unroll
    for (int t = 0; t < 27; ++t) {
        lw[t] = 0;
        if (t < npos && t != centre) {
            const int dz = c.dim3 ? t / 9 - 1 : 0, dy = (t / 3) % 3 - 1, ddx = t % 3 - 1;
            int z = z0 + dz; if (z < 0) z += d.Nz; else if (z >= d.Nz) z -= d.Nz;
            int y = y0 + dy; if (y < 0) y += d.Ny; else if (y >= d.Ny) y -= d.Ny;
            int x = x0 + ddx; if (x < 0) x += d.Nx; else if (x >= d.Nx) x -= d.Nx;
            lw[t] = c.phi[lin(d, z, y, x)];
        }
    }
    uint32_t occ_o = 0, occ_n = 0;
#pragma unroll
    for (int t = 0; t < 27; ++t) {
        if (t >= npos || t == centre) continue;
        int32_t lab = lw[t];
        if (lab < 0) {
            const int j = ~lab;
            const int jo = c.Cold[j];
            lab = jo;
            if (j < k) {
                const int jn = c.Cnew[j];
                const int sj = ld_acq(c.state + j);
                if (jo == ol || jo == nl || jn == ol || jn == nl) {
                    if (undecided(sj)) return 0;   // not ready
                }
                if (sj == ACCEPT) lab = jn;
            }
        }
        if (lab == ol) occ_o |= 1u << t;
        if (lab == nl) occ_n |= 1u << t;
    }
    const int oo = topo_count(occ_o, c.dim3 != 0);
    const int no = topo_count(occ_n, c.dim3 != 0);
    if (no == 0) return decide_claim(c, k, sk, REJECT);
    const bool need_old = oo != 1;      // count (oo == 0) or BFS (oo > 1)
    const bool need_new = no > 1;
    if (!need_old && !need_new) return decide_claim(c, k, sk, ACCEPT);
    // a retried candidate tries its bounds again every FLIP_RESPEC rounds
    // (they narrow as earlier candidates are decided), else waits for its
    // exact turn; qflags bits 8 and up hold the round of its last search
    const bool respec = sk == RETRY && round - (c.qflags[k] >> 8) >= FLIP_RESPEC;
    if (sk == RETRY && !respec) {
        if (need_old && !label_prefix_ready(c, c.pos_old[k], ol)) return 0;
        if (need_new && !label_prefix_ready(c, c.pos_new[k], nl)) return 0;
    }
    bool ready = true;
    LabelPrefix po{0, 0, 0};
    if (need_old) {
        po = label_prefix(c, c.pos_old[k], ol);
        ready = po.add + po.rem == 0;
    }
    if (need_new && ready) ready = label_prefix_ready(c, c.pos_new[k], nl);
    if (oo == 0) {
        // src/gpk_connectivity.cpp:179-185: a vanishing old label passes only
        // for the label's last cell
        const int cnt = c.count[ol] + po.delta;
        if (cnt - po.rem > 1) return decide_claim(c, k, sk, REJECT);
        if (cnt + po.add > 1) return 0;          // depends on undecided candidates
        if (!need_new) return decide_claim(c, k, sk, ACCEPT);
    }
    int flags = (oo > 1 ? 1 : 0) | (need_new ? 2 : 0);
    if (!ready) {
        if (sk == RETRY && !respec) return 0;
        flags |= 4;
    }
    if (atomicCAS(c.state + k, sk, QUEUED) != sk) return 0;
    c.qflags[k] = flags | ((round < (1 << 22) ? round : (1 << 22)) << 8);
    return 2;
}

__global__ void __launch_bounds__(256) k_flip_round(FlipCtx c, int round) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= c.m) return;
    int prev;
    const bool queued = flip_try(c, k, round, &prev) == 2;
    // queue entries reserved per slab with one atomic per slab met by the
    // wavefront (its candidates are consecutive: one slab, or two)
    const int x = k / c.slab;
    const unsigned long long lt = (1ull << (threadIdx.x & 63)) - 1ull;
    unsigned long long todo = __ballot(queued);
    while (todo) {
        const int leader = __builtin_ffsll((long long)todo) - 1;
        const int xl = __shfl(x, leader);
        const unsigned long long grp = __ballot(queued && x == xl);
        int base = 0;
        if ((threadIdx.x & 63) == leader) base = atomicAdd(&c.s->qcount[round & 1][xl], __popcll(grp));
        base = __shfl(base, leader);
        if (queued && x == xl) c.queue[round & 1][xl * c.slab + base + __popcll(grp & lt)] = k;
        todo &= ~grp;
    }
}

// View labels (view_label_b) of the six face neighbours g6 of a frontier
// cell for transaction k: the label words of all faces are loaded before any
// is used, then the table entries of the candidates among them (label word
// ~j; Cold and Cnew do not change while the rounds run) are loaded together
// and resolved.  Faces past nface and inactive lanes read 0x7fffffff.
__device__ __forceinline__ void face_view_labels(const FlipCtx& c, int mode, int k, int32_t L, bool act,
                                                 int nface, const int g6[6], int32_t lab[6]) {
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        lab[q] = 0x7fffffff;
        if (act && q < nface) lab[q] = c.phi[g6[q]];
    }
    int sj[6], cn[6], co[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        const int j = ~lab[q];
        sj[q] = REJECT; cn[q] = -1; co[q] = -1;
        if (lab[q] < 0) {
            co[q] = c.Cold[j];
            if (j < k) {
                sj[q] = ld_acq(c.state + j);
                cn[q] = c.Cnew[j];
            }
        }
    }
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        if (lab[q] < 0) {
            int32_t r = co[q];
            if (sj[q] == ACCEPT) r = cn[q];
            else if (mode == VIEW_UPPER && sj[q] != REJECT && cn[q] == L) r = L;
            else if (mode == VIEW_LOWER && sj[q] != REJECT && co[q] == L) r = -1;
            lab[q] = r;
        }
    }
}

// Colour verdict of a multi-source search: 1 when the recorded contacts join
// all nb colours into one group, 0 when some group has no frontier cell left
// (it is closed: disconnected), -1 undecided.  Contact bit a*8+b joins
// colours a and b; the groups are the transitive closure of the contacts,
// found by three squarings of the 6x6 reachability rows (chains of up to
// eight contacts; six colours need five).  Uniform values, no indexing.
__device__ __forceinline__ int colour_verdict(int nb, unsigned long long adj, uint32_t present) {
    const uint32_t all = (1u << nb) - 1u;
    uint32_t row[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) row[q] = (uint32_t(adj >> (q * 8)) & 0x3fu) | (1u << q);
#pragma unroll
    for (int it = 0; it < 3; ++it) {
        uint32_t nr[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            uint32_t a = row[q];
#pragma unroll
            for (int p = 0; p < 6; ++p) a |= ((row[q] >> p) & 1u) ? row[p] : 0u;
            nr[q] = a;
        }
#pragma unroll
        for (int q = 0; q < 6; ++q) row[q] = nr[q];
    }
    if ((row[0] & all) == all) return 1;
    bool closed = false;
#pragma unroll
    for (int q = 0; q < 6; ++q) closed |= q < nb && (row[q] & present & all) == 0u;
    return closed ? 0 : -1;
}

// ---------------------------------------------------------------------------
// Multi-source face-connectivity test on label L around `cell` as seen by
// transaction k: are all face branches of L (centre excluded when `excl`)
// in one connected component?  Same answer as the reference's early-exit
// periodic search (src/gpk_connectivity.cpp:79-115), found by growing every
// branch at once and stopping when all colours have met or one merged colour
// group can grow no further.  One workgroup; LDS hash of visited cells.
// ---------------------------------------------------------------------------
constexpr int BFS_HCAP = 4096;      // hash slots (power of two; 44 KB of LDS in all: three workgroups per CU)
constexpr int BFS_HMAX = 3072;      // max insertions before overflow
constexpr int BFS_FCAP = 1024;      // frontier capacity

struct BfsShared {
    unsigned long long hash[BFS_HCAP];
    int fr[2][BFS_FCAP];
    uint8_t frc[2][BFS_FCAP];
    int fcount[2];
    int inserted;
    unsigned long long adj;
    uint32_t pres[2]; // colours present in each frontier
    int result;       // -1 running, 0 disconnected, 1 connected
    int overflow;
    int nbranch;
};

__device__ __forceinline__ uint32_t hslot(int cell) {
    return (uint32_t(cell) * 2654435761u) >> (32 - 12);
}

// returns 1 connected, 0 disconnected, -1 overflow; mode picks the view of
// the undecided earlier transactions (view_label_b)
__device__ __forceinline__ int bfs_connected_lds(BfsShared& S, const FlipCtx& c, int cell, int32_t L,
                                                 bool excl, int k, int mode) {
    const Dims& d = c.d;
    const int nface = c.dim3 ? 6 : 4;
    for (int h = threadIdx.x; h < BFS_HCAP; h += blockDim.x) S.hash[h] = ~0ull;
    if (threadIdx.x == 0) {
        S.fcount[0] = 0; S.fcount[1] = 0; S.inserted = 0; S.adj = 0ull; S.pres[0] = 0u; S.pres[1] = 0u;
        S.result = -1; S.overflow = 0; S.nbranch = 0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int f[6];
        faces(d, cell, f);
        int32_t fl[6];
        face_view_labels(c, mode, k, L, true, nface, f, fl);
        int nb = 0;
        for (int q = 0; q < nface; ++q) {
            if (fl[q] != L) continue;
            const int col = nb++;
            const unsigned long long val = ((unsigned long long)(uint32_t)f[q] << 8) | col;
            uint32_t h = hslot(f[q]);
            while (true) {
                const unsigned long long cur = S.hash[h];
                if (cur == ~0ull) {
                    S.hash[h] = val; ++S.inserted;
                    S.fr[0][S.fcount[0]] = f[q]; S.frc[0][S.fcount[0]] = uint8_t(col);
                    ++S.fcount[0];
                    S.pres[0] |= 1u << col;
                    break;
                }
                if (int(cur >> 8) == f[q]) {
                    const int oc = int(cur & 0xff);
                    S.adj |= (1ull << (oc * 8 + col)) | (1ull << (col * 8 + oc));
                    break;
                }
                h = (h + 1) & (BFS_HCAP - 1);
            }
        }
        S.nbranch = nb;
        if (nb <= 1) S.result = 1;
    }
    __syncthreads();
    int cur = 0;
    while (true) {
        if (threadIdx.x == 0 && S.result < 0) {
            if (S.overflow || S.fcount[cur] > BFS_FCAP) S.result = -2;
            else S.result = colour_verdict(S.nbranch, S.adj, S.pres[cur]);
        }
        __syncthreads();
        if (S.result != -1) break;
        const int nxt = cur ^ 1;
        if (threadIdx.x == 0) { S.fcount[nxt] = 0; S.pres[nxt] = 0u; }
        __syncthreads();
        const int nf = S.fcount[cur] < BFS_FCAP ? S.fcount[cur] : BFS_FCAP;
        for (int e = threadIdx.x; e < nf; e += blockDim.x) {
            const int cc = S.fr[cur][e];
            const int col = S.frc[cur][e];
            int f[6];
            faces(d, cc, f);
            int32_t lab[6];
            face_view_labels(c, mode, k, L, true, nface, f, lab);
            // first hash probes of all faces issued together
            bool want[6];
            unsigned long long prv[6];
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                want[q] = q < nface && !(excl && f[q] == cell) && lab[q] == L;
                prv[q] = 0ull;
                if (want[q])
                    prv[q] = atomicCAS(&S.hash[hslot(f[q])], ~0ull, ((unsigned long long)(uint32_t)f[q] << 8) | col);
            }
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                if (!want[q]) continue;
                const int nbc = f[q];
                const unsigned long long val = ((unsigned long long)(uint32_t)nbc << 8) | col;
                uint32_t h = hslot(nbc);
                unsigned long long prev = prv[q];
                for (int probe = 0; probe < BFS_HCAP; ++probe) {
                    if (probe > 0) prev = atomicCAS(&S.hash[h], ~0ull, val);
                    if (prev == ~0ull) {
                        const int ins = atomicAdd(&S.inserted, 1);
                        const int pos = atomicAdd(&S.fcount[nxt], 1);
                        if (ins >= c.wg_cap || pos >= BFS_FCAP) {
                            S.overflow = 1;
                        } else {
                            S.fr[nxt][pos] = nbc; S.frc[nxt][pos] = uint8_t(col);
                            atomicOr(&S.pres[nxt], 1u << col);
                        }
                        break;
                    }
                    if (int(prev >> 8) == nbc) {
                        const int oc = int(prev & 0xff);
                        if (oc != col)
                            atomicOr(&S.adj, (1ull << (oc * 8 + col)) | (1ull << (col * 8 + oc)));
                        break;
                    }
                    h = (h + 1) & (BFS_HCAP - 1);
                }
            }
        }
        __syncthreads();
        if (threadIdx.x == 0 && S.fcount[nxt] > BFS_FCAP) S.overflow = 1;
        cur = nxt;
        __syncthreads();
    }
    const int r = S.result;
    __syncthreads();
    return r == -2 ? -1 : r;
}


// The same test run by one wavefront, for the small components most
// searches meet: a 2048-slot LDS hash of visited cells (32-bit entries: cell
// << 3 | colour, so grids below 2^29 cells), frontiers of up to 256 cells,
// the frontier appends ordered by ballot prefix within the wavefront, colour
// contacts and presence held in registers (uniform over the wavefront), so
// no workgroup barrier is needed.  Returns 1 connected, 0 disconnected, -1
// when the search outgrows the hash or a frontier.  10 KB per wavefront:
// four 40 KB workgroups per CU.
constexpr int WB_HCAP = 2048;       // hash slots (power of two)
constexpr int WB_HBITS = 11;        // log2(WB_HCAP)
constexpr int WB_HMAX = 1536;       // insertions before overflow
constexpr int WB_FCAP = 256;        // frontier capacity
constexpr int WB_WAVES = 4;         // wavefronts (searches) per 256-thread workgroup
constexpr int WB_MAXN = 1 << 29;    // grids this hash can key
constexpr uint32_t WB_EMPTY = 0xffffffffu;

struct WaveBfs {
    alignas(16) uint32_t hash[WB_HCAP];   // (cell << 3) | colour, WB_EMPTY empty
    uint32_t fr[2][WB_FCAP];   // (cell << 3) | colour
};

// OR over the wavefront of a value that few lanes hold (colour contacts):
// the lanes holding one are visited through the ballot, each read with one
// lane read per word
__device__ __forceinline__ unsigned long long wave_or_sparse(unsigned long long v) {
    unsigned long long any = __ballot(v != 0ull), acc = 0ull;
    while (any) {
        const int l = __builtin_ffsll((long long)any) - 1;
        const uint32_t lo = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(v)), l));
        const uint32_t hi = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(v >> 32)), l));
        acc |= ((unsigned long long)hi << 32) | lo;
        any &= any - 1ull;
    }
    return acc;
}

// OR over the wavefront of a 6-bit colour set, one ballot per colour
__device__ __forceinline__ uint32_t wave_or6(uint32_t v) {
    uint32_t acc = 0u;
#pragma unroll
    for (int q = 0; q < 6; ++q) acc |= __ballot((v >> q) & 1u) != 0ull ? (1u << q) : 0u;
    return acc;
}

__device__ __forceinline__ int bfs_connected_wave(WaveBfs& S, const FlipCtx& c, int cell, int32_t L, bool excl,
                                                  int k, int mode) {
    const Dims& d = c.d;
    const int nface = c.dim3 ? 6 : 4;
    const int lane = threadIdx.x & 63;
    const unsigned long long lt = (1ull << lane) - 1ull;
    for (int h = lane; h < WB_HCAP / 4; h += 64)
        reinterpret_cast<uint4*>(S.hash)[h] = make_uint4(WB_EMPTY, WB_EMPTY, WB_EMPTY, WB_EMPTY);
    // seeds: the face neighbours of label L, one colour each in face order (a
    // neighbour met twice, on grids of side 2, is one cell of two colours)
    int f[6];
    faces(d, cell, f);
    int32_t fl[6];
    face_view_labels(c, mode, k, L, true, nface, f, fl);
    int seed[6], nb = 0;
    unsigned long long adj = 0ull;
    for (int q = 0; q < nface; ++q) {
        if (fl[q] != L) continue;
        seed[nb++] = f[q];
    }
    if (nb <= 1) return 1;
    int nfront = 0;
    uint32_t present = 0u;
    wave_sync();
    for (int q = 0; q < nb; ++q) {
        int dup = -1;
        for (int p = 0; p < q; ++p) if (seed[p] == seed[q] && dup < 0) dup = p;
        if (dup >= 0) {
            adj |= (1ull << (dup * 8 + q)) | (1ull << (q * 8 + dup));
            continue;
        }
        if (lane == 0) {
            uint32_t h = (uint32_t(seed[q]) * 2654435761u) >> (32 - WB_HBITS);
            while (S.hash[h] != WB_EMPTY) h = (h + 1) & (WB_HCAP - 1);
            S.hash[h] = (uint32_t(seed[q]) << 3) | uint32_t(q);
            S.fr[0][nfront] = (uint32_t(seed[q]) << 3) | uint32_t(q);
        }
        ++nfront;
        present |= 1u << q;
    }
    wave_sync();
    int cur = 0, inserted = nfront;
    while (true) {
        const int v = colour_verdict(nb, adj, present);
        if (v >= 0) return v;
        const int nxt = cur ^ 1;
        int ncount = 0;
        bool ovf = false;
        unsigned long long myadj = 0ull;
        uint32_t mypres = 0u;
        for (int e0 = 0; e0 < nfront; e0 += 64) {
            const int e = e0 + lane;
            const bool act = e < nfront;
            const uint32_t fe = act ? S.fr[cur][e] : 0u;
            const int cc = int(fe >> 3);
            const int col = int(fe & 7u);
            int g6[6];
            faces(d, cc, g6);
            int32_t lab[6];
            face_view_labels(c, mode, k, L, act, nface, g6, lab);
            // first hash probes of all faces issued together; collisions then
            // probe on, face by face
            bool want[6], ins[6];
            uint32_t hq[6], prv[6];
            const uint32_t ctag = uint32_t(col);
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                want[q] = act && q < nface && !(excl && g6[q] == cell) && lab[q] == L;
                ins[q] = false;
                hq[q] = (uint32_t(g6[q]) * 2654435761u) >> (32 - WB_HBITS);
                prv[q] = 0u;
                if (want[q]) prv[q] = atomicCAS(&S.hash[hq[q]], WB_EMPTY, (uint32_t(g6[q]) << 3) | ctag);
            }
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                if (!want[q]) continue;
                const int nbc = g6[q];
                const uint32_t val = (uint32_t(nbc) << 3) | ctag;
                uint32_t prev = prv[q];
                uint32_t h = hq[q];
                int probe = 1;
                while (true) {
                    if (prev == WB_EMPTY) { ins[q] = true; break; }
                    if (int(prev >> 3) == nbc) {
                        const int oc = int(prev & 7u);
                        if (oc != col) myadj |= (1ull << (oc * 8 + col)) | (1ull << (col * 8 + oc));
                        break;
                    }
                    if (probe++ == WB_HCAP) { ovf = true; break; }
                    h = (h + 1) & (WB_HCAP - 1);
                    prev = atomicCAS(&S.hash[h], WB_EMPTY, val);
                }
            }
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                const unsigned long long bal = __ballot(ins[q]);
                const int pos = ncount + __popcll(bal & lt);
                if (ins[q]) {
                    if (pos < WB_FCAP) {
                        S.fr[nxt][pos] = (uint32_t(g6[q]) << 3) | ctag;
                    }
                    mypres |= 1u << col;
                }
                ncount += __popcll(bal);
            }
        }
        inserted += ncount;
        adj |= wave_or_sparse(myadj);
        present = wave_or6(mypres);
        if (__ballot(ovf) != 0ull || ncount > WB_FCAP || inserted > c.wave_cap) return -1;
        nfront = ncount;
        cur = nxt;
        wave_sync();
    }
}

// One wavefront per queued candidate (WB_WAVES per workgroup).  Candidates
// are numbered in band (grid) order and queued per slab of the band, one
// slab for each of the eight XCDs: workgroups are dispatched to the XCDs in
// turn, so workgroup b searches slab b % 8 and each XCD's L2 holds the labels
// of one region of the grid (the launch has a multiple of 8 workgroups).  Exact turns
// and bounds searches as k_flip_bfs; a search that outgrows the wavefront's
// hash queues the candidate (still QUEUED) for k_flip_bfs's workgroup
// search.  Block 0 clears the counts that the next round fills.  (Entries are
// assigned statically: a wavefront taking the next entry through a global
// counter measured slower, the counter's atomics costing more than the
// balance gained.)
__global__ void __launch_bounds__(256) k_flip_bfs_wave(FlipCtx c, int round, int* __restrict__ ovq,
                                                        int* __restrict__ ovcount) {
    __shared__ WaveBfs S[WB_WAVES];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (blockIdx.x == 0 && threadIdx.x < 8) c.s->qcount[(round + 1) & 1][threadIdx.x] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) ovcount[(round + 1) & 1] = 0;
    const int x = blockIdx.x & 7, nbx = gridDim.x >> 3;
    const int n = c.s->qcount[round & 1][x];
    const int* q = c.queue[round & 1] + x * c.slab;
    for (int qi = (blockIdx.x >> 3) * WB_WAVES + w; qi < n; qi += nbx * WB_WAVES) {
        const int k = q[qi];
        const int flags = c.qflags[k];
        const int cell = c.Ccell[k];
        // ACCEPT, REJECT, RETRY, or -1: to the workgroup search (every search
        // of a grid too large for the wavefront hash's keys goes there)
        const int verdict = flip_verdict(flags, false, [&](int mode, int t) {
            if (c.d.N >= WB_MAXN) return -1;
            return bfs_connected_wave(S[w], c, cell, t == 0 ? c.Cold[k] : c.Cnew[k], t == 0, k, mode);
        });
        if (lane == 0) {
            if (verdict == ACCEPT || verdict == REJECT) decide(c, k, verdict);
            else if (verdict == RETRY) st_rel(c.state + k, RETRY);
            else ovq[atomicAdd(&ovcount[round & 1], 1)] = k;
        }
    }
}

// Large-component fallback: the same multi-source test with the visited set
// in device memory and frontiers of grid capacity.  A cell's mark packs the
// transaction tag and the colour ((tag << 3) | colour), so no clearing is
// needed between tests.  Runs as a single workgroup over the fallback queue.
__device__ __forceinline__ int claim_mark(int* mark, int cell, int tag, int col, int* other) {
    int old = __hip_atomic_load(mark + cell, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    while (true) {
        if ((old >> 3) == tag) { *other = old & 7; return 0; }
        const int prev = atomicCAS(mark + cell, old, (tag << 3) | col);
        if (prev == old) return 1;
        old = prev;
    }
}

__device__ void flip_bfs_fallback(FlipCtx& c, const int* __restrict__ fbq, int* __restrict__ mark,
                                  int* __restrict__ fr0, int* __restrict__ fr1, int* __restrict__ tag_ctr) {
    __shared__ int fcount[2];
    __shared__ uint32_t pres[2];   // colours present in each frontier
    __shared__ unsigned long long adj;
    __shared__ int result, nbranch, tag;
    const Dims& d = c.d;
    const int nface = c.dim3 ? 6 : 4;
    const int n = ld_acq(&c.s->fbcount);
    for (int qi = 0; qi < n; ++qi) {
        const int k = ld_acq(fbq + qi);
        const int flags = c.qflags[k];
        const int cell = c.Ccell[k];
        int ok = 1;
        for (int pass = 0; pass < 2 && ok == 1; ++pass) {
            if (pass == 0 && !(flags & 1)) continue;
            if (pass == 1 && !(flags & 2)) continue;
            const int32_t L = pass == 0 ? c.Cold[k] : c.Cnew[k];
            const bool excl = pass == 0;
            __syncthreads();
            if (threadIdx.x == 0) {
                tag = ++(*tag_ctr);
                fcount[0] = 0; fcount[1] = 0; pres[0] = 0u; pres[1] = 0u; adj = 0ull; result = -1;
                int f[6];
                faces(d, cell, f);
                int nb = 0;
                for (int q = 0; q < nface; ++q) {
                    if (view_label(f[q], k, c.phi, c.state, c.Cold, c.Cnew) != L) continue;
                    const int col = nb++;
                    int oc = 0;
                    if (claim_mark(mark, f[q], tag, col, &oc)) {
                        fr0[fcount[0]++] = f[q];
                        pres[0] |= 1u << col;
                    } else adj |= (1ull << (oc * 8 + col)) | (1ull << (col * 8 + oc));
                }
                nbranch = nb;
                if (nb <= 1) result = 1;
            }
            __syncthreads();
            int cur = 0;
            while (true) {
                if (threadIdx.x == 0 && result < 0) {
                    result = colour_verdict(nbranch, adj, pres[cur]);
                    fcount[cur ^ 1] = 0;
                    pres[cur ^ 1] = 0u;
                }
                __syncthreads();
                if (result != -1) break;
                const int nxt = cur ^ 1;
                const int* frc = cur == 0 ? fr0 : fr1;
                int* frn = nxt == 0 ? fr0 : fr1;
                const int nf = fcount[cur];
                const int tg = tag;
                for (int e = threadIdx.x; e < nf; e += blockDim.x) {
                    const int cc = frc[e];
                    const int col = __hip_atomic_load(mark + cc, __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_AGENT) & 7;
                    int f[6];
                    faces(d, cc, f);
                    for (int q = 0; q < nface; ++q) {
                        const int nbc = f[q];
                        if (excl && nbc == cell) continue;
                        if (view_label(nbc, k, c.phi, c.state, c.Cold, c.Cnew) != L) continue;
                        int oc = 0;
                        if (claim_mark(mark, nbc, tg, col, &oc)) {
                            frn[atomicAdd(&fcount[nxt], 1)] = nbc;
                            atomicOr(&pres[nxt], 1u << col);
                        } else if (oc != col) {
                            atomicOr(&adj, (1ull << (oc * 8 + col)) | (1ull << (col * 8 + oc)));
                        }
                    }
                }
                __threadfence();
                __syncthreads();
                cur = nxt;
            }
            ok = result;
            __syncthreads();
        }
        if (threadIdx.x == 0) decide(c, k, ok ? ACCEPT : REJECT);
        __syncthreads();
    }
    if (threadIdx.x == 0) c.s->fbcount = 0;
}

// One workgroup per queued candidate (the candidates whose wavefront search
// outgrew its hash); candidates whose search outgrows this LDS visited set
// go to the fallback queue, which the last workgroup to finish runs with the
// device-memory search before the kernel ends.
__global__ void __launch_bounds__(256) k_flip_bfs(FlipCtx c, const int* __restrict__ queue,
                                                   const int* __restrict__ qcount, int* __restrict__ fbq,
                                                   int* __restrict__ mark, int* __restrict__ fr0,
                                                   int* __restrict__ fr1, int* __restrict__ tag_ctr) {
    __shared__ BfsShared S;
    __shared__ int last;
    const int n = *qcount;
    // nothing queued: no search can add to the fallback queue, which the last
    // launch that searched left empty, so every workgroup leaves at once
    if (n == 0) return;
    for (int qi = blockIdx.x; qi < n; qi += gridDim.x) {
        const int k = queue[qi];
        const int flags = c.qflags[k];
        const int cell = c.Ccell[k];
        // an exact turn whose search outgrows this visited set goes to the
        // fallback queue; a bounds turn whose search does waits (RETRY)
        const int verdict = flip_verdict(flags, true, [&](int mode, int t) {
            return bfs_connected_lds(S, c, cell, t == 0 ? c.Cold[k] : c.Cnew[k], t == 0, k, mode);
        });
        if (threadIdx.x == 0) {
            if (verdict < 0) {
                const int p = atomicAdd(&c.s->fbcount, 1);
                st_rel(fbq + p, k);
            } else if (verdict == RETRY) {
                st_rel(c.state + k, RETRY);
            } else {
                decide(c, k, verdict);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        __threadfence();
        last = atomicAdd(&c.s->bfs_done, 1) == int(gridDim.x) - 1;
    }
    __syncthreads();
    if (last) {
        __threadfence();
        flip_bfs_fallback(c, fbq, mark, fr0, fr1, tag_ctr);
        if (threadIdx.x == 0) st_rel(&c.s->bfs_done, 0);
    }
}

// every candidate's old label back in phi (a transaction abandoned on error)
__global__ void k_flip_restore(int m, const int* __restrict__ Ccell, const int* __restrict__ Cold,
                               int32_t* __restrict__ phi) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < m) phi[Ccell[k]] = Cold[k];
}

// apply the decided transactions (src/gpk_transitions.cpp:300-312)
template <class R>
__global__ void k_flip_apply(int m, const int* __restrict__ Ccell, const int* __restrict__ Cold,
                             const int* __restrict__ Cnew, const R* __restrict__ Cg,
                             const R* __restrict__ Canew, const int* __restrict__ state,
                             int32_t* __restrict__ phi, R* __restrict__ g, R* __restrict__ alpha,
                             int* __restrict__ count, Scalars* s,
                             int all_accept, R* __restrict__ rho, R rho_rex, int sweep) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    bool acc = false, def = false;
    if (k < m) {
        const int cell = Ccell[k];
        if (all_accept || state[k] == ACCEPT) {
            phi[cell] = Cnew[k];
            g[cell] = Cg[k];
            alpha[cell] = R(1);
            if (sweep) rho[cell] = rho_rex;   // src/gpk_transitions.cpp:308
            atomicSub(count + Cold[k], 1);
            atomicAdd(count + Cnew[k], 1);
            acc = true;
        } else {
            phi[cell] = Cold[k];   // the cell's label word back from ~k
            const R an = Canew[k];
            alpha[cell] = an < Eps<R>::flip ? Eps<R>::flip : an;
            def = true;
        }
    }
    // one counter update per wavefront (transactions count their verdicts
    // when they take them)
    const unsigned long long ba = __ballot(acc), bd = __ballot(def);
    if (all_accept && (threadIdx.x & 63) == 0) {
        if (ba) atomicAdd(&s->accepted, __popcll(ba));
        if (bd) atomicAdd(&s->deferred, __popcll(bd));
    }
}

// ---------------------------------------------------------------------------
// Balanced Gaussian remap (src/gpk_transitions.cpp:38-178)
// ---------------------------------------------------------------------------
template <class R, int DIM>
__global__ void __launch_bounds__(256)
k_remap_score(Dims d, int r, int nband, const int* __restrict__ band,
              const int32_t* __restrict__ phi, const double* __restrict__ wtab,
              int32_t* __restrict__ phi_comp, uint8_t* __restrict__ pflag,
              R* __restrict__ gain_s) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nband) return;
    const int i = band[a];
    int32_t bl;
    const R gain = remap_score_cell<R, DIM>(d, r, i, phi, wtab, bl);
    phi_comp[i] = bl;
    pflag[a] = (phi[i] != bl && gain > Eps<R>::gain) ? 1 : 0;
    gain_s[a] = gain;
}

__global__ void k_copy_i32(int n, const int32_t* __restrict__ src, int32_t* __restrict__ dst) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// the same copy 16 bytes per thread (n a multiple of 4, 16-byte aligned buffers)
__global__ void k_copy_i32x4(int n4, const int4* __restrict__ src, int4* __restrict__ dst) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// proposal keys: first pass by gain (descending) then cell, second by
// (low, high, direction)
template <class R>
__global__ void k_prop_keys(int np, const int* __restrict__ sel, const int* __restrict__ band,
                            const R* __restrict__ gain_s, const int32_t* __restrict__ phi,
                            const int32_t* __restrict__ phi_comp, int* __restrict__ pcell,
                            unsigned long long* __restrict__ kgain, int* __restrict__ idx) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= np) return;
    const int sa = sel[k];
    const int cell = band[sa];
    pcell[k] = cell;
    idx[k] = k;
    const R gn = gain_s[sa];
    unsigned long long gb;
    if (sizeof(R) == 4) gb = (unsigned long long)(~__float_as_uint(float(gn)));
    else gb = ~(unsigned long long)__double_as_longlong(double(gn));
    kgain[k] = gb;
}

// group key ((low << hb) | high) << 1 | dir, labels below 2^hb
__global__ void k_prop_group_keys(int np, int hb, const int* __restrict__ order,
                                  const int* __restrict__ pcell, const int32_t* __restrict__ phi,
                                  const int32_t* __restrict__ phi_comp,
                                  unsigned long long* __restrict__ kgrp) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= np) return;
    const int cell = pcell[order[s]];
    const int32_t p = phi[cell], q = phi_comp[cell];
    const uint32_t lo = uint32_t(p < q ? p : q), hi = uint32_t(p < q ? q : p);
    const uint32_t dir = p == int32_t(lo) ? 0u : 1u;
    kgrp[s] = ((unsigned long long)lo << hb | hi) << 1 | dir;
}

// segments of equal (low, high, dir): start index per position
__global__ void k_seg_bounds(int np, const unsigned long long* __restrict__ kg,
                             int* __restrict__ seg_start, int* __restrict__ seg_end) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= np) return;
    const bool first = s == 0 || kg[s] != kg[s - 1];
    const bool last = s == np - 1 || kg[s] != kg[s + 1];
    if (first) {
        int e = s;
        while (e + 1 < np && kg[e + 1] == kg[s]) ++e;
        for (int t = s; t <= e; ++t) { seg_start[t] = s; seg_end[t] = e; }
    }
    (void)last;
}

// forward entry at rank r pairs with the reverse entry of the same rank
__global__ void k_pair_flags(int np, const unsigned long long* __restrict__ kg,
                             const int* __restrict__ seg_start, const int* __restrict__ seg_end,
                             uint8_t* __restrict__ pflag) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= np) return;
    int f = 0;
    if ((kg[s] & 1ull) == 0ull) {
        const int e = seg_end[s];
        const int rnk = s - seg_start[s];
        if (e + 1 < np && (kg[e + 1] >> 1) == (kg[s] >> 1)) {
            const int rs = e + 1;
            const int re = seg_end[rs];
            if (rnk <= re - rs) f = 1;
        }
    }
    pflag[s] = uint8_t(f);
}

// pair q: the q-th flagged forward entry (plist, in sorted order)
template <class R>
__global__ void k_pair_emit(int npairs, int hb, const unsigned long long* __restrict__ kg,
                            const int* __restrict__ seg_start, const int* __restrict__ seg_end,
                            const int* __restrict__ plist,
                            const int* __restrict__ order, const int* __restrict__ pcell,
                            int* __restrict__ Pa, int* __restrict__ Pb, int* __restrict__ Plo,
                            int* __restrict__ Phi, int* __restrict__ pstate,
                            int* __restrict__ pid) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= npairs) return;
    const int s = plist[q];
    const int rnk = s - seg_start[s];
    const int rs = seg_end[s] + 1;
    const int a = pcell[order[s]], b = pcell[order[rs + rnk]];
    const unsigned long long key = kg[s] >> 1;
    Pa[q] = a; Pb[q] = b;
    Plo[q] = int(uint32_t(key >> hb));
    Phi[q] = int(uint32_t(key & ((1ull << hb) - 1)));
    pstate[q] = UNDEC;
    pid[a] = q; pid[b] = q;
}

struct PairCtx {
    Dims d;
    int dim3;
    int np;
    const int* Pa; const int* Pb; const int* Plo; const int* Phi;
    int* pstate;
    const int* pid;
    const int32_t* phi;
    const void* g;          // R*
    void* Pg;               // R*, 2 per pair (ga, gb)
    const int* count;
    Scalars* s;
};

__device__ __forceinline__ float ld_agent(const float* p) {
    return __uint_as_float(__hip_atomic_load(reinterpret_cast<const unsigned int*>(p),
                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}
__device__ __forceinline__ double ld_agent(const double* p) {
    return __longlong_as_double((long long)__hip_atomic_load(
        reinterpret_cast<const unsigned long long*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}
__device__ __forceinline__ void st_agent(float* p, float v) {
    __hip_atomic_store(reinterpret_cast<unsigned int*>(p), __float_as_uint(v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(double* p, double v) {
    __hip_atomic_store(reinterpret_cast<unsigned long long*>(p),
                       (unsigned long long)__double_as_longlong(v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

// A pair's state word is UNDEC, or its verdict | (round tag << 4), the tag
// being the round (from 1) that took it.  A verdict taken in the current
// round counts as undecided: a pair reads only verdicts of earlier rounds,
// whose orientations (Pg) earlier launches wrote, so no release / acquire
// between the waves of one launch is needed (the launch boundary orders them).
__device__ __forceinline__ bool pair_settled(int st, int tag) { return st != UNDEC && (st >> 4) < tag; }
__device__ __forceinline__ int pair_verdict(int st) { return st & 15; }

template <class R>
__device__ __forceinline__ void pair_view(const PairCtx& c, int cell, int q, int tag, int32_t& lab, R& gv) {
    const int p = c.pid[cell];
    const R* g = static_cast<const R*>(c.g);
    int sp = UNDEC;
    if (p >= 0 && p < q) sp = ld_acq(c.pstate + p);
    if (pair_settled(sp, tag) && pair_verdict(sp) == ACCEPT) {
        const R* Pg = static_cast<const R*>(c.Pg);
        if (c.Pa[p] == cell) { lab = c.Phi[p]; gv = ld_agent(Pg + 2 * p); }
        else { lab = c.Plo[p]; gv = ld_agent(Pg + 2 * p + 1); }
        return;
    }
    lab = c.phi[cell];
    gv = g[cell];
}

// Readiness set of each pair: the earlier pairs that own a cell inside the
// 5^d boxes around a and b and share one of its two labels.  The set depends
// only on the pair table, so it is listed once per sweep, one wavefront per
// pair (lane l looks at box cells l, l + 64, ...), in box order.  Entry
// q * PAIR_DEPS holds the count, or -1 when the list does not fit (then the
// round kernel scans the boxes itself).
constexpr int PAIR_DEPS = 16;

__global__ void __launch_bounds__(256) k_pair_deps(PairCtx c, int* __restrict__ deps) {
    const int q = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (q >= c.np) return;
    const Dims& d = c.d;
    const bool d3 = c.dim3 != 0;
    const int a = c.Pa[q], b = c.Pb[q];
    const int32_t lo = c.Plo[q], hi = c.Phi[q];
    const int nbox = d3 ? 125 : 25;
    int za, ya, xa, zb, yb, xb;
    decode(d, a, za, ya, xa);
    decode(d, b, zb, yb, xb);
    int n = 0;
    for (int k0 = 0; k0 < 2 * nbox; k0 += 64) {
        const int k = k0 + lane;
        bool hit = false;
        int p = -1;
        if (k < 2 * nbox) {
            const int side = k >= nbox ? 1 : 0;
            const int r = k - side * nbox;
            const int lz = d3 ? r / 25 : 2, ly = (r / 5) % 5, lx = r % 5;
            const int z0 = side ? zb : za, y0 = side ? yb : ya, x0 = side ? xb : xa;
            const int cell = lin(d, wrapi(z0 + lz - 2, d.Nz), wrapi(y0 + ly - 2, d.Ny), wrapi(x0 + lx - 2, d.Nx));
            p = c.pid[cell];
            if (p >= 0 && p < q) {
                const int32_t pl = c.Plo[p], ph = c.Phi[p];
                hit = pl == lo || pl == hi || ph == lo || ph == hi;
            }
        }
        const unsigned long long m = __ballot(hit);
        const int before = __popcll(m & ((1ull << lane) - 1ull));
        if (hit && n + before < PAIR_DEPS - 1) deps[q * PAIR_DEPS + 1 + n + before] = p;
        n += __popcll(m);
    }
    if (lane == 0) deps[q * PAIR_DEPS] = n < PAIR_DEPS ? n : -1;
}

// One wavefront per pair q (src/gpk_transitions.cpp:129-170 on the versioned
// view).  Lanes check the readiness list, then gather the versioned label and
// orientation of every cell of the 5^d boxes around a and b in parallel; the
// certificates' box masks come from ballots, and lane 0 takes the decision.
template <class R>
__global__ void __launch_bounds__(256) k_pair_round(PairCtx c, const int* __restrict__ deps, int tag) {
    constexpr int WPB = 4;                       // wavefronts (pairs) per block
    __shared__ int32_t blab[WPB][2][128];        // versioned labels, box A / box B
    __shared__ R bg[WPB][2][128];                // versioned orientations
    __shared__ int bcell[WPB][2][128];           // grid index of each box position
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int q = blockIdx.x * WPB + w;
    if (q >= c.np) return;
    if (ld_acq(c.pstate + q) != UNDEC) return;
    const Dims& d = c.d;
    const bool d3 = c.dim3 != 0;
    const int a = c.Pa[q], b = c.Pb[q];
    const int32_t lo = c.Plo[q], hi = c.Phi[q];
    const int nbox = d3 ? 125 : 25;
    int za, ya, xa, zb, yb, xb;
    decode(d, a, za, ya, xa);
    decode(d, b, zb, yb, xb);
    // readiness: earlier pairs sharing a label inside the 5^d boxes of a and b
    const int nd = deps[q * PAIR_DEPS];
    bool wait = false;
    if (nd >= 0) {
        if (lane < nd) wait = !pair_settled(ld_acq(c.pstate + deps[q * PAIR_DEPS + 1 + lane]), tag);
    } else {
        for (int k = lane; k < 2 * nbox; k += 64) {
            const int side = k >= nbox ? 1 : 0;
            const int r = k - side * nbox;
            const int lz = d3 ? r / 25 : 2, ly = (r / 5) % 5, lx = r % 5;
            const int z0 = side ? zb : za, y0 = side ? yb : ya, x0 = side ? xb : xa;
            const int cell = lin(d, wrapi(z0 + lz - 2, d.Nz), wrapi(y0 + ly - 2, d.Ny), wrapi(x0 + lx - 2, d.Nx));
            const int p = c.pid[cell];
            if (p >= 0 && p < q) {
                const int32_t pl = c.Plo[p], ph = c.Phi[p];
                if ((pl == lo || pl == hi || ph == lo || ph == hi) && !pair_settled(ld_acq(c.pstate + p), tag))
                    wait = true;
            }
        }
    }
    if (__ballot(wait) != 0ull) return;
    // versioned view of both boxes
    for (int k = lane; k < 2 * 128; k += 64) {
        const int side = k >> 7, r = k & 127;
        if (r < nbox) {
            const int lz = d3 ? r / 25 : 2, ly = (r / 5) % 5, lx = r % 5;
            const int z0 = side ? zb : za, y0 = side ? yb : ya, x0 = side ? xb : xa;
            const int cell = lin(d, wrapi(z0 + lz - 2, d.Nz), wrapi(y0 + ly - 2, d.Ny), wrapi(x0 + lx - 2, d.Nx));
            int32_t l;
            R gv;
            pair_view<R>(c, cell, q, tag, l, gv);
            blab[w][side][r] = l;
            bg[w][side][r] = gv;
            bcell[w][side][r] = cell;
        }
    }
    wave_sync();
    // box masks of the two certificates: relabel a lo -> hi with a's own
    // label read as lo; relabel b hi -> lo with a (already moved) and b read
    // as hi (src/gpk_transitions.cpp:151-158)
    Bits128 boA{0, 0}, bnA{0, 0}, boB{0, 0}, bnB{0, 0};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int r = lane + 64 * h;
        bool oA = false, nA = false, oB = false, nB = false;
        if (r < nbox) {
            const int ca = bcell[w][0][r], cb = bcell[w][1][r];
            const int32_t la = ca == a ? lo : blab[w][0][r];
            const int32_t lb = cb == a ? hi : (cb == b ? hi : blab[w][1][r]);
            oA = la == lo; nA = la == hi;
            oB = lb == hi; nB = lb == lo;
        }
        const unsigned long long mo = __ballot(oA), mn = __ballot(nA);
        const unsigned long long po = __ballot(oB), pn = __ballot(nB);
        if (h == 0) { boA.lo = mo; bnA.lo = mn; boB.lo = po; bnB.lo = pn; }
        else { boA.hi = mo; bnA.hi = mn; boB.hi = po; bnB.hi = pn; }
    }
    if (lane != 0) return;
    R* Pg = static_cast<R*>(c.Pg);
    const int nface = d3 ? 6 : 4;
    const int lc = (d3 ? 2 : 0) * 25 + 12;
    const int fo[6] = {lc - 1, lc + 1, lc - 5, lc + 5, lc - 25, lc + 25};
    const int32_t la = blab[w][0][lc], lb = blab[w][1][lc];
    int verdict = REJECT;
    R ga = bg[w][0][lc], gb = bg[w][1][lc];
    if (la == lo && lb == hi) {
        // target orientations (first face neighbour carrying the target label)
        for (int k = 0; k < nface; ++k)
            if (blab[w][0][fo[k]] == hi) { ga = bg[w][0][fo[k]]; break; }
        for (int k = 0; k < nface; ++k)
            if (blab[w][1][fo[k]] == lo) { gb = bg[w][1][fo[k]]; break; }
        auto mism = [&](int side, int32_t prop, int pc, int32_t pp) {
            int m = 0;
            for (int k = 0; k < nface; ++k) {
                const int32_t l = bcell[w][side][fo[k]] == pc ? pp : blab[w][side][fo[k]];
                m += l != prop;
            }
            return m;
        };
        const int before = mism(0, lo, b, hi) + mism(1, hi, a, lo);
        const int after = mism(0, hi, b, lo) + mism(1, lo, a, hi);
        if (after < before && bounded_cert_masks(d3, boA, bnA, c.count[lo]) &&
            bounded_cert_masks(d3, boB, bnB, c.count[hi] + 1)) {
            verdict = ACCEPT;
            st_agent(Pg + 2 * q, ga);
            st_agent(Pg + 2 * q + 1, gb);
        }
    }
    st_rel(c.pstate + q, verdict | (tag << 4));
    if (verdict == ACCEPT) atomicAdd(&c.s->exchanges, 2);
    atomicSub(&c.s->undecided, 1);
}

template <class R>
__global__ void k_pair_apply(int np, const int* __restrict__ Pa, const int* __restrict__ Pb,
                             const int* __restrict__ Plo, const int* __restrict__ Phi,
                             const int* __restrict__ pstate, const R* __restrict__ Pg,
                             int32_t* __restrict__ phi, R* __restrict__ g, R* __restrict__ alpha,
                             int* __restrict__ pid, Scalars* s, R* __restrict__ rho, R rho_rex,
                             int sweep) {
    (void)s;   // exchanges are counted by the rounds that accept them
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= np) return;
    const int a = Pa[q], b = Pb[q];
    pid[a] = -1; pid[b] = -1;
    if (pair_verdict(pstate[q]) != ACCEPT) return;
    phi[a] = Phi[q]; phi[b] = Plo[q];
    g[a] = Pg[2 * q]; g[b] = Pg[2 * q + 1];
    alpha[a] = R(1); alpha[b] = R(1);
    if (sweep) { rho[a] = rho_rex; rho[b] = rho_rex; }   // src/gpk_transitions.cpp:173
}

// remap-score cache: the (2r+1)^d box around every relabelled cell is marked
// dirty, one thread per (entry, cell of the entry, box position); entry k
// relabelled its cells when st[k] == ACCEPT (st null: every entry did).
// Flips have one cell per entry (c0), pairs two (c0, c1).
__global__ void k_mark_dirty(Dims d, int dim3, int r, int n, int ncell, const int* __restrict__ c0,
                             const int* __restrict__ c1, const int* __restrict__ st,
                             uint8_t* __restrict__ dirty) {
    const int W = 2 * r + 1;
    const long long box = dim3 ? (long long)W * W * W : (long long)W * W;
    const long long total = (long long)n * ncell * box;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        const long long e = t / box;
        const int o = int(t - e * box);
        const int k = int(e / ncell);
        if (st && (st[k] & 15) != ACCEPT) continue;   // pair states carry a round tag
        const int cell = (e - (long long)k * ncell) == 0 ? c0[k] : c1[k];
        int z0, y0, x0;
        decode(d, cell, z0, y0, x0);
        const int dz = dim3 ? o / (W * W) - r : 0, dy = (o / W) % W - r, dx = o % W - r;
        dirty[lin(d, wrapi(z0 + dz, d.Nz), wrapi(y0 + dy, d.Ny), wrapi(x0 + dx, d.Nx))] = 1;
    }
}

// The same marks one box row (W cells along x) per thread, with 32-bit
// index arithmetic (the host launches it when n * ncell * rows < 2^31).
__global__ void k_mark_dirty_rows(Dims d, int dim3, int r, int n, int ncell, const int* __restrict__ c0,
                                  const int* __restrict__ c1, const int* __restrict__ st,
                                  uint8_t* __restrict__ dirty) {
    const int W = 2 * r + 1;
    const int rows = dim3 ? W * W : W;
    const int total = n * ncell * rows;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
        const int e = t / rows;
        const int o = t - e * rows;
        const int k = e / ncell;
        if (st && (st[k] & 15) != ACCEPT) continue;   // pair states carry a round tag
        const int cell = e - k * ncell == 0 ? c0[k] : c1[k];
        int z0, y0, x0;
        decode(d, cell, z0, y0, x0);
        const int dz = dim3 ? o / W - r : 0, dy = (dim3 ? o % W : o) - r;
        const int base = lin(d, wrapi(z0 + dz, d.Nz), wrapi(y0 + dy, d.Ny), 0);
        for (int dx = -r; dx <= r; ++dx) dirty[base + wrapi(x0 + dx, d.Nx)] = 1;
    }
}

// ---------------------------------------------------------------------------
// Unresolved-grain extinction (src/gpk_transitions.cpp:180-249)
// ---------------------------------------------------------------------------
__global__ void k_label_census(int nlab, const int* __restrict__ count, int min_cells,
                               uint8_t* __restrict__ unres, Scalars* s) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nlab) return;
    const int n = count[l];
    const bool u = n > 0 && n < min_cells;
    unres[l] = u ? 1 : 0;
    if (n > 0) atomicAdd(&s->present, 1);
    if (u) atomicAdd(&s->unresolved, 1);
}

__global__ void k_unres_flags(int N, const int32_t* __restrict__ phi,
                              const uint8_t* __restrict__ unres, uint8_t* __restrict__ flag) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x)
        flag[i] = unres[phi[i]];
}

// layer-by-layer erosion from an immutable snapshot per layer (single workgroup)
template <class R>
__global__ void __launch_bounds__(1024)
k_erode(Dims d, int dim3, int n, const int* __restrict__ cells, const uint8_t* __restrict__ unres,
        int32_t* __restrict__ phi, R* __restrict__ g, R* __restrict__ alpha,
        int* __restrict__ count, int* __restrict__ tgt, R* __restrict__ tgg, Scalars* s,
        R* __restrict__ rho, R rho_rex, int sweep, int rr, uint8_t* __restrict__ dirty) {
    __shared__ int changed;
    const int nface = dim3 ? 6 : 4;
    while (true) {
        if (threadIdx.x == 0) changed = 0;
        __syncthreads();
        for (int e = threadIdx.x; e < n; e += blockDim.x) {
            const int i = cells[e];
            tgt[e] = -1;
            if (!unres[phi[i]]) continue;
            int z, y, x;
            decode(d, i, z, y, x);
            int f[6];
            faces(d, i, f);
            int best = -1, bs = -1;
            for (int k = 0; k < nface; ++k) {
                const int32_t q = phi[f[k]];
                if (unres[q]) continue;
                int sup = 0;
                for (int dz = (dim3 ? -1 : 0); dz <= (dim3 ? 1 : 0); ++dz)
                    for (int dy = -1; dy <= 1; ++dy)
                        for (int ddx = -1; ddx <= 1; ++ddx) {
                            const int zz = (z + dz + d.Nz) % d.Nz, yy = (y + dy + d.Ny) % d.Ny,
                                      xx = (x + ddx + d.Nx) % d.Nx;
                            sup += phi[lin(d, zz, yy, xx)] == q;
                        }
                if (sup > bs) { best = k; bs = sup; }
            }
            if (best >= 0) {
                tgt[e] = phi[f[best]];
                tgg[e] = g[f[best]];
            }
        }
        __syncthreads();
        for (int e = threadIdx.x; e < n; e += blockDim.x) {
            if (tgt[e] < 0) continue;
            const int i = cells[e];
            atomicSub(count + phi[i], 1);
            atomicAdd(count + tgt[e], 1);
            phi[i] = tgt[e];
            g[i] = tgg[e];
            alpha[i] = R(1);
            if (sweep) rho[i] = rho_rex;   // src/gpk_transitions.cpp:245
            mark_dirty_box(d, dim3, i, rr, dirty);
            atomicAdd(&changed, 1);
        }
        __syncthreads();
        const int ch = changed;
        if (threadIdx.x == 0) s->eroded += ch;
        __syncthreads();
        if (ch == 0) break;
    }
}

// ---------------------------------------------------------------------------
// Solute transport (src/gpk_transport.cpp:11-135)
// ---------------------------------------------------------------------------
template <class R>
__global__ void k_diffusion(Dims d, int dim3, const R* __restrict__ c, const uint8_t* __restrict__ mask,
                            const R* __restrict__ alpha, const Scalars* __restrict__ s, float dx_f,
                            float D_f, R* __restrict__ cout) {
    const R dt = R(s->dt), dx = R(dx_f), D = R(D_f);
    const R dx2 = dx * dx;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < d.N; i += gridDim.x * blockDim.x) {
        int f[6];
        faces(d, i, f);
        R lap = c[f[1]] + c[f[0]] + c[f[3]] + c[f[2]];
        if (dim3) lap += c[f[5]] + c[f[4]] - R(6) * c[i];
        else lap -= R(4) * c[i];
        const R De = mask[i] ? D * alpha[i] : D;
        cout[i] = c[i] + dt * De * lap / dx2;
    }
}

template <class R>
__device__ __forceinline__ R solid_frac(int32_t p, R a) { return p != 0 ? a : R(1) - a; }

template <class R>
__global__ void k_solid_fraction(int N, const int32_t* __restrict__ phi, const R* __restrict__ alpha,
                                 R* __restrict__ s) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x)
        s[i] = solid_frac<R>(phi[i], alpha[i]);
}

// faces in the transport order x+, x-, y+, y-, z+, z-
__device__ __forceinline__ int part_faces(const Dims& d, int dim3, int i, int nb[6]) {
    int z, y, x;
    decode(d, i, z, y, x);
    int n = 0;
    nb[n++] = lin(d, z, y, x == d.Nx - 1 ? 0 : x + 1);
    nb[n++] = lin(d, z, y, x == 0 ? d.Nx - 1 : x - 1);
    nb[n++] = lin(d, z, y == d.Ny - 1 ? 0 : y + 1, x);
    nb[n++] = lin(d, z, y == 0 ? d.Ny - 1 : y - 1, x);
    if (dim3) {
        nb[n++] = lin(d, z == d.Nz - 1 ? 0 : z + 1, y, x);
        nb[n++] = lin(d, z == 0 ? d.Nz - 1 : z - 1, y, x);
    }
    return n;
}

// per-source rejected solute q and liquid weight sum (src/gpk_transport.cpp:64-89)
template <class R>
__global__ void k_reject_source(Dims d, int dim3, float k_f, const int32_t* __restrict__ phi,
                                const R* __restrict__ alpha, const R* __restrict__ s_old,
                                const R* __restrict__ c, R* __restrict__ q_out,
                                R* __restrict__ wsum_out) {
    const R k = R(k_f);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < d.N; i += gridDim.x * blockDim.x) {
        R qv = R(0), ws = R(0);
        const R sn = solid_frac<R>(phi[i], alpha[i]);
        const R ds = sn - s_old[i];
        if (ds > R(0)) {
            R s0 = s_old[i];
            s0 = s0 < R(0) ? R(0) : (R(1) < s0 ? R(1) : s0);
            R den = R(1) - s0 * (R(1) - k);
            den = den < Eps<R>::cl_floor ? Eps<R>::cl_floor : den;
            const R cL = c[i] / den;
            const R dsc = R(1) < ds ? R(1) : ds;
            R q = (R(1) - k) * cL * dsc;
            q = c[i] < q ? c[i] : q;
            if (q > R(0)) {
                int nb[6];
                const int nn = part_faces(d, dim3, i, nb);
                for (int a = 0; a < nn; ++a) {
                    const int j = nb[a];
                    R w = R(0);
                    if (phi[j] == 0) {
                        const R l = R(1) - solid_frac<R>(phi[j], alpha[j]);
                        w = l < R(0) ? R(0) : l;
                    }
                    ws += w;
                }
                if (ws > R(0)) qv = q;
            }
        }
        q_out[i] = qv;
        wsum_out[i] = ws;
    }
}

// ordered gather: delta[j] accumulated over sources in ascending source index,
// each source in its own neighbour order — the serial scatter's sequence
template <class R>
__global__ void k_reject_gather(Dims d, int dim3, const int32_t* __restrict__ phi,
                                const R* __restrict__ alpha, const R* __restrict__ q,
                                const R* __restrict__ wsum, R* __restrict__ c) {
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < d.N; j += gridDim.x * blockDim.x) {
        int nbj[6];
        const int nn = part_faces(d, dim3, j, nbj);
        // no rejecting source around j (the common case): the delta is +0
        bool any = q[j] > R(0);
        for (int a = 0; a < nn; ++a) any |= q[nbj[a]] > R(0);
        if (!any) {
            c[j] += R(0);
            continue;
        }
        int src[7];
        int ns = 0;
        src[ns++] = j;
        for (int a = 0; a < nn; ++a) {
            bool dup = false;
            for (int b = 0; b < ns; ++b) dup |= src[b] == nbj[a];
            if (!dup) src[ns++] = nbj[a];
        }
        for (int x = 1; x < ns; ++x)
            for (int y = x; y > 0 && src[y - 1] > src[y]; --y) {
                const int t = src[y]; src[y] = src[y - 1]; src[y - 1] = t;
            }
        R wj = R(0);
        if (phi[j] == 0) {
            const R l = R(1) - solid_frac<R>(phi[j], alpha[j]);
            wj = l < R(0) ? R(0) : l;
        }
        R delta = R(0);
        for (int e = 0; e < ns; ++e) {
            const int i = src[e];
            const R qi = q[i];
            if (qi <= R(0)) continue;
            if (i == j) delta -= qi;
            if (wj > R(0)) {
                int nbi[6];
                const int ni = part_faces(d, dim3, i, nbi);
                for (int a = 0; a < ni; ++a)
                    if (nbi[a] == j) delta += qi * wj / wsum[i];
            }
        }
        c[j] += delta;
    }
}

template <class R>
__global__ void k_liquid_comp(int N, float k_f, const int32_t* __restrict__ phi,
                              const R* __restrict__ alpha, const R* __restrict__ c,
                              R* __restrict__ s, R* __restrict__ cL) {
    const R k = R(k_f);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        const R sv = solid_frac<R>(phi[i], alpha[i]);
        s[i] = sv;
        R den = R(1) - sv * (R(1) - k);
        den = den < Eps<R>::cl_floor ? Eps<R>::cl_floor : den;
        cL[i] = c[i] / den;
    }
}

template <class R>
__global__ void k_diffusion_part(Dims d, int dim3, const R* __restrict__ s, const R* __restrict__ cL,
                                 const R* __restrict__ c, const Scalars* __restrict__ sc,
                                 float dx_f, float D_f, R* __restrict__ cout) {
    const R dt = R(sc->dt), dx = R(dx_f), D = R(D_f);
    const R coef = dt * D / (dx * dx);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < d.N; i += gridDim.x * blockDim.x) {
        int nb[6];
        const int nn = part_faces(d, dim3, i, nb);
        const R li = R(1) - s[i];
        R flux = R(0);
        for (int a = 0; a < nn; ++a) {
            const int j = nb[a];
            const R lj = R(1) - s[j];
            const R lf = lj < li ? lj : li;
            if (lf > R(0)) flux += lf * (cL[j] - cL[i]);
        }
        cout[i] = c[i] + coef * flux;
    }
}

// ---------------------------------------------------------------------------
// Label table maintenance and diagnostics
// ---------------------------------------------------------------------------
// per-label cell counts: each thread walks 64 consecutive cells and issues one
// atomic per run of equal labels (neighbouring cells mostly share a label)
__global__ void k_label_hist(int N, const int32_t* __restrict__ phi, int* __restrict__ count) {
    const long long b = (long long)(blockIdx.x * blockDim.x + threadIdx.x) * 64;
    if (b >= N) return;
    const int e = b + 64 < N ? int(b + 64) : N;
    int cur = phi[b], run = 0;
    for (int i = int(b); i < e; ++i) {
        const int v = phi[i];
        if (v != cur) { atomicAdd(count + cur, run); cur = v; run = 0; }
        ++run;
    }
    atomicAdd(count + cur, run);
}

__global__ void k_fill_i32(int n, int32_t* p, int32_t v) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_u8(int n, uint8_t* p, uint8_t v) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}
template <class R>
__global__ void k_fill_r(int n, R* p, R v) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// elastic coupling (src/gpk_physics.cpp:10-63): eigenstrain, Jacobi-type
// smoothing sweeps, energy density 0.5 E (eps* - smooth(eps*))^2
template <class R>
__global__ void k_elastic_eps(int n, const R* __restrict__ c, float coeff, float c_ref,
                              R* __restrict__ eps_star, R* __restrict__ eps_smooth) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const R e = R(coeff) * (c[i] - R(c_ref));
        eps_star[i] = e;
        eps_smooth[i] = e;
    }
}
template <class R>
__global__ void k_elastic_smooth(Dims d, int dim3, const R* __restrict__ src, R* __restrict__ dst) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < d.N; i += gridDim.x * blockDim.x) {
        int f[6];
        faces(d, i, f);
        R avg = src[f[0]] + src[f[1]] + src[f[2]] + src[f[3]];
        if (dim3) {
            avg += src[f[4]] + src[f[5]];
            avg /= R(6);
        } else {
            avg /= R(4);
        }
        dst[i] = R(0.5) * src[i] + R(0.5) * avg;
    }
}
template <class R>
__global__ void k_elastic_energy(int n, const R* __restrict__ eps_star, const R* __restrict__ eps_smooth,
                                 float modulus, R* __restrict__ E_el) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const R dd = eps_star[i] - eps_smooth[i];
        E_el[i] = R(0.5) * R(modulus) * dd * dd;
    }
}

template <class R>
__global__ void k_scale_r(int n, R* __restrict__ p, R f) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] *= f;
}

__global__ void k_minmax_i32(int N, const int32_t* __restrict__ p, int* __restrict__ out) {
    int mn = 0x7fffffff, mx = -0x7fffffff - 1;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        const int v = p[i];
        mn = v < mn ? v : mn;
        mx = v > mx ? v : mx;
    }
    atomicMin(out, mn);
    atomicMax(out + 1, mx);
}

// band-compact field -> full grid (zeros outside the band)
template <class R>
__global__ void k_scatter_band(int nband, const int* __restrict__ band, const R* __restrict__ src,
                               R* __restrict__ dst) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a < nband) dst[band[a]] = src[a];
}

// FP64 sum in the canonical chunk order: sequential over each 1024-cell
// chunk, then over each group of 1024 chunk sums, then over the groups.  Each
// level stages its operands in LDS with coalesced loads and one thread per
// chunk (group) adds them in order.
template <class R>
__global__ void __launch_bounds__(256) k_sum_chunks(int n, const R* __restrict__ v, double* __restrict__ out) {
    __shared__ double st[4][1025];          // 4 chunks per block; the pad spreads the rows over banks
    const int nch = (n + 1023) / 1024;
    const int ch0 = blockIdx.x * 4;
    for (int t = threadIdx.x; t < 4096; t += blockDim.x) {
        const int k = t >> 10;
        const long long i = (long long)(ch0 + k) * 1024 + (t & 1023);
        st[k][t & 1023] = i < n ? double(v[i]) : 0.0;
    }
    __syncthreads();
    const int k = threadIdx.x;
    if (k < 4 && ch0 + k < nch) {
        const int cnt = n - (ch0 + k) * 1024 < 1024 ? n - (ch0 + k) * 1024 : 1024;
        double s = 0.0;
        for (int i = 0; i < cnt; ++i) s += st[k][i];
        out[ch0 + k] = s;
    }
}

// group sums: block g adds chunk sums [1024 g, 1024 g + 1024) in order
__global__ void __launch_bounds__(256) k_sum_groups(int nch, const double* __restrict__ p1,
                                                    double* __restrict__ gsum) {
    __shared__ double st[1024];
    const int b = blockIdx.x * 1024;
    const int cnt = nch - b < 1024 ? nch - b : 1024;
    for (int t = threadIdx.x; t < cnt; t += blockDim.x) st[t] = p1[b + t];
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < cnt; ++i) s += st[i];
        gsum[blockIdx.x] = s;
    }
}

__global__ void k_sum_tail(int ng, const double* __restrict__ gsum, double* __restrict__ out) {
    if (threadIdx.x || blockIdx.x) return;
    double tot = 0.0;
    for (int g = 0; g < ng; ++g) tot += gsum[g];
    *out = tot;
}

// phase volume terms per cell (src/gpk_simulation.cpp:78-86), chunk-summed
template <class R>
__global__ void __launch_bounds__(256)
k_phase_terms(int N, int32_t label, const int32_t* __restrict__ phi,
              const int32_t* __restrict__ phi_comp, const uint8_t* __restrict__ mask,
              const R* __restrict__ alpha, double* __restrict__ out) {
    // two chunks per block; per cell the two terms in order, a term that does
    // not apply staged as +0.0 (the partial sums are never -0, so adding it
    // changes nothing)
    __shared__ double st[2][2049];
    const int nch = (N + 1023) / 1024;
    const int ch0 = blockIdx.x * 2;
    for (int t = threadIdx.x; t < 2048; t += blockDim.x) {
        const int k = t >> 10;
        const long long i = (long long)(ch0 + k) * 1024 + (t & 1023);
        double t1 = 0.0, t2 = 0.0;
        if (i < N) {
            const double a = double(alpha[i]);
            if (phi[i] == label) t1 = a;
            if (mask[i] && phi_comp[i] == label && phi_comp[i] != phi[i]) t2 = 1.0 - a;
        }
        st[k][2 * (t & 1023)] = t1;
        st[k][2 * (t & 1023) + 1] = t2;
    }
    __syncthreads();
    const int k = threadIdx.x;
    if (k < 2 && ch0 + k < nch) {
        const int cnt = N - (ch0 + k) * 1024 < 1024 ? N - (ch0 + k) * 1024 : 1024;
        double s = 0.0;
        for (int i = 0; i < 2 * cnt; ++i) s += st[k][i];
        out[ch0 + k] = s;
    }
}

__global__ void k_count_u8(int N, const uint8_t* __restrict__ m, int* __restrict__ out) {
    int c = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) c += m[i] != 0;
    atomicAdd(out, c);
}
__global__ void k_count_pos(int n, const int* __restrict__ cnt, int* __restrict__ out) {
    int c = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) c += cnt[i] > 0;
    atomicAdd(out, c);
}

// ---------------------------------------------------------------------------
// Voronoi nearest site (demos/gpk.cpp:72-89): exact integer periodic squared
// distance, lowest site index on ties (the reference's strict < scan in site
// order, demos/gpk.cpp:51-120).
// ---------------------------------------------------------------------------

// Nearest site from a bucket grid of the sites: a site in a bucket
// at periodic bucket distance k >= ring + 1 along some axis lies at least
// ring * bmin + 1 cells away along that axis, so the ring search stops once
// the best squared distance is below that bound (a site at the bound is
// still visited and the lowest index wins ties, as in the plain scan).
// bstart / bsites: the sites of each bucket in ascending index order.
struct SiteBuckets {
    int Bz, By, Bx;       // bucket extent per axis
    int bz, by, bx;       // buckets per axis
    int bmin;             // smallest bucket extent over the searched axes
};
__global__ void k_voronoi_buckets(Dims d, SiteBuckets sb, const int* __restrict__ sz,
                                  const int* __restrict__ sy, const int* __restrict__ sx,
                                  const int* __restrict__ bstart, const int* __restrict__ bsites,
                                  int32_t* __restrict__ phi) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.N) return;
    int z, y, x;
    decode(d, i, z, y, x);
    const bool d3 = d.Nz > 1;
    const int cz = z / sb.Bz, cy = y / sb.By, cx = x / sb.Bx;
    long long bd = 0x7FFFFFFFFFFFFFFFLL;
    int best = 0;
    for (int ring = 0;; ++ring) {
        const int rz = d3 ? ring : 0;
        for (int dz = -rz; dz <= rz; ++dz)
            for (int dy = -ring; dy <= ring; ++dy)
                for (int dxb = -ring; dxb <= ring; ++dxb) {
                    const int az = dz < 0 ? -dz : dz, ay = dy < 0 ? -dy : dy, ax = dxb < 0 ? -dxb : dxb;
                    const int m = az > ay ? (az > ax ? az : ax) : (ay > ax ? ay : ax);
                    if (m != ring) continue;
                    const int qz = ((cz + dz) % sb.bz + sb.bz) % sb.bz;
                    const int qy = ((cy + dy) % sb.by + sb.by) % sb.by;
                    const int qx = ((cx + dxb) % sb.bx + sb.bx) % sb.bx;
                    const int q = (qz * sb.by + qy) * sb.bx + qx;
                    for (int t = bstart[q]; t < bstart[q + 1]; ++t) {
                        const int k = bsites[t];
                        long long a = z - sz[k], b = y - sy[k], e = x - sx[k];
                        a = a < 0 ? -a : a;
                        b = b < 0 ? -b : b;
                        e = e < 0 ? -e : e;
                        a = a < d.Nz - a ? a : d.Nz - a;
                        b = b < d.Ny - b ? b : d.Ny - b;
                        e = e < d.Nx - e ? e : d.Nx - e;
                        const long long dd = a * a + b * b + e * e;
                        if (dd < bd || (dd == bd && k < best)) {
                            bd = dd;
                            best = k;
                        }
                    }
                }
        const bool all = 2 * ring + 1 >= sb.bx && 2 * ring + 1 >= sb.by && (!d3 || 2 * ring + 1 >= sb.bz);
        if (all) break;
        const long long lb = (long long)ring * sb.bmin + 1;
        if (bd < lb * lb) break;
    }
    phi[i] = best + 1;
}

// ---------------------------------------------------------------------------
// Voronoi connectivity normalisation (demos/gpk.cpp:122-189): keep the
// largest face-connected component of every label (ties: the component whose
// first cell in flat order comes first), then attach the other cells layer by
// layer to the most frequent attached face-neighbour label (ties: the smaller
// label), taking the orientation of the first face neighbour with that label.
// ---------------------------------------------------------------------------
__global__ void k_site_g(int n, const int32_t* __restrict__ phi, const float* __restrict__ sg,
                         float* __restrict__ g32) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        g32[i] = sg[phi[i] - 1];
}

__global__ void k_cc_init(int n, int* __restrict__ comp) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) comp[i] = i;
}

// one propagation round: comp[i] = min over i and its same-label face
// neighbours, then one pointer jump.  Values only decrease and always name a
// cell of i's component, so the fixed point is the component's first cell.
__global__ void k_cc_step(Dims d, int dim3, const int32_t* __restrict__ phi, int* __restrict__ comp,
                          int* __restrict__ changed) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < d.N; i += gridDim.x * blockDim.x) {
        int f[6];
        faces(d, i, f);
        const int32_t p = phi[i];
        const int c0 = __hip_atomic_load(comp + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int m = c0;
        const int nf = dim3 ? 6 : 4;
        for (int k = 0; k < nf; ++k) {
            if (phi[f[k]] != p) continue;
            const int cj = __hip_atomic_load(comp + f[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            m = cj < m ? cj : m;
        }
        const int cm = __hip_atomic_load(comp + m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        m = cm < m ? cm : m;
        if (m < c0 && atomicMin(comp + i, m) > m) atomicOr(changed, 1);
    }
}

__global__ void k_cc_size(int n, const int* __restrict__ comp, int* __restrict__ csz) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        atomicAdd(csz + comp[i], 1);
}

// best[label] = max over the label's components of (size, -first cell)
__global__ void k_cc_best(int n, const int32_t* __restrict__ phi, const int* __restrict__ comp,
                          const int* __restrict__ csz, unsigned long long* __restrict__ best) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        if (comp[i] == i)
            atomicMax(best + phi[i], ((unsigned long long)(uint32_t)csz[i] << 32) | (0xFFFFFFFFull - uint32_t(i)));
}

__global__ void k_cc_attached(int n, const int32_t* __restrict__ phi, const int* __restrict__ comp,
                              const unsigned long long* __restrict__ best, uint8_t* __restrict__ att,
                              uint8_t* __restrict__ detached) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t root = 0xFFFFFFFFu - uint32_t(best[phi[i]] & 0xFFFFFFFFull);
        const bool a = uint32_t(comp[i]) == root;
        att[i] = a ? 1 : 0;
        detached[i] = a ? 0 : 1;
    }
}

// one attachment layer over the detached list, from the state before the layer
__global__ void k_attach_scan(Dims d, int dim3, int nu, const int* __restrict__ ul,
                              const uint8_t* __restrict__ att, const int32_t* __restrict__ phi,
                              const float* __restrict__ g32, int32_t* __restrict__ tg,
                              float* __restrict__ og, uint8_t* __restrict__ ok) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nu) return;
    const int cc = ul[u];
    ok[u] = 0;
    if (att[cc]) return;
    int f[6];
    faces(d, cc, f);
    const int nf = dim3 ? 6 : 4;
    int32_t cand[6];
    int sup[6], src[6], nc = 0;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        if (k >= nf || !att[f[k]]) continue;
        const int32_t v = phi[f[k]];
        int q = nc;
#pragma unroll
        for (int t = 0; t < 6; ++t)
            if (t < nc && cand[t] == v && q == nc) q = t;
        if (q == nc) { cand[nc] = v; src[nc] = f[k]; sup[nc] = 0; ++nc; }
#pragma unroll
        for (int t = 0; t < 6; ++t)
            if (t == q) ++sup[t];
    }
    if (nc == 0) return;
    int b = 0;
#pragma unroll
    for (int q = 1; q < 6; ++q)
        if (q < nc && (sup[q] > sup[b] || (sup[q] == sup[b] && cand[q] < cand[b]))) b = q;
    tg[u] = cand[b];
    og[u] = g32[src[b]];
    ok[u] = 1;
}

__global__ void k_attach_apply(int nu, const int* __restrict__ ul, const int32_t* __restrict__ tg,
                               const float* __restrict__ og, const uint8_t* __restrict__ ok,
                               int32_t* __restrict__ phi, float* __restrict__ g32,
                               uint8_t* __restrict__ att, int* __restrict__ applied) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nu || !ok[u]) return;
    const int cc = ul[u];
    phi[cc] = tg[u];
    g32[cc] = og[u];
    att[cc] = 1;
    atomicAdd(applied, 1);
}

template <class R>
__global__ void k_widen_g(int n, const float* __restrict__ g32, R* __restrict__ g) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) g[i] = R(g32[i]);
}

} // namespace gpkd
