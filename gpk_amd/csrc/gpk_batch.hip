// gpk_batch.hip — many small 2-D grids stepped together.
//
// The reference runs a parameter sweep one GPKGrid after another
// (demos/gpk.cpp:1121-1259, the breakup driver's d-loop).  Here every
// instance of a sweep lives in one packed set of arrays, described by an
// instance table, and ONE persistent kernel advances all of them by a step:
// each workgroup takes the next instance from a queue (largest first) and
// runs the whole GPKGrid::step (src/gpk_simulation.cpp:31-54) for it with
// workgroup barriers between the stages.  Per-instance dt and time stay on
// the device, so a step is two launches (queue reset + step) and is captured
// once in a hipGraph and replayed.
//
// Stages per instance, each in the reference's order and arithmetic (the
// per-cell functions are the ones the grid-wide kernels use):
//   interface band (src/gpk_geometry.cpp:42-82), geometry + competitor +
//   velocity (88-183, src/gpk_physics.cpp:67-148), pair projection
//   (src/gpk_physics.cpp:152-193), dt (src/gpk_simulation.cpp:11-40),
//   alpha and flips (src/gpk_transitions.cpp:251-313), balanced Gaussian
//   remap (38-178), time advance.
// Models with the grain-network topology test decide their flip candidates
// one after another in index order (src/gpk_transitions.cpp:288-313): each
// candidate's 3x3 topological numbers, label count and, for a label with
// several branches, a periodic face-connectivity search run by the whole
// workgroup; every 8th step erodes unresolved grains (180-249).
// Scope: 2-D instances, D = 0, no noise, no dislocation or elastic terms;
// labels below 1024 per instance.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <numeric>
#include <vector>

#include "gpk_cell.hpp"
#include "gpk_host.hpp"
#include "gpk_tiles.hpp"

using namespace gpkd;
using namespace gpkh;

namespace {

constexpr int BATCH_LMAX = 1024;   // labels per instance (LDS histogram)

struct Inst {
    int Ny, Nx, N, off;   // off: first cell in the packed arrays
    int soff;             // first entry of the instance's sort scratch
    int spow;             // sort scratch length (power of two >= N)
};

struct InstStat {
    double time, dt;
    long long band_total, accepted_total, exchanges_total, proposals_total, guarded_total;
    long long deferred_total, eroded_total;
    long long step_id;    // steps taken by the instance
    int band, accepted, exchanges, proposals;
    long long t0, t1;     // wall clock at the step's start and end
};

struct Prop {
    uint32_t hdr;     // (lo << 11) | (hi << 1) | dir, dir 0 = resident is lo
    int32_t cell;
    uint64_t gk;      // ~orderable(gain): ascending = gain descending
};

template <class R>
struct BatchArgs {
    int n_inst;
    const Inst* inst;
    const int* order;
    int* next;
    InstStat* st;
    GeomParams gp;
    gpk_model_desc model;
    float dx_f;
    R dt_req, cap;
    int proj_r, remap_r;
    int network;          // grain-network topology test on flips
    int min_cells;        // grains below this size are unresolved
    const R *w0, *w1, *w2, *tab;
    RowLutDesc lutd;
    const double *projw, *remapw;
    int32_t* phi;
    R* alpha;
    R* g;
    const R* c;
    const R* T;
    uint8_t* mask;
    int32_t* phi_comp;
    R* g_comp;
    int* band;
    int* slot;
    R *nx, *ny, *nz, *kappa, *vraw, *vn, *gain;
    uint8_t* pflag;       // remap proposal flag per band slot (tiled scoring)
    Prop* props;
};

// LDS of the tiled stages (FP32 / FP64 modes): geometry at radius 8
// (sigma 1.5-2.0 cells), projection and remap scoring at radius 3; the
// stages run one after another, so they share one block
template <class R>
union BatchTileSmem {
    GeoTileSmem<R, 2, 8> geo;
    ProjTileSmem<R, 2, 3> proj;
    RemapTileSmem<2, 3> remap;
};
struct NoTileSmem {};

__device__ __forceinline__ uint64_t orderable(float v) {
    const uint32_t b = __float_as_uint(v);
    return (b & 0x80000000u) ? uint64_t(~b) : uint64_t(b | 0x80000000u);
}
__device__ __forceinline__ uint64_t orderable(double v) {
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    return (b >> 63) ? ~b : (b | (1ull << 63));
}
__device__ __forceinline__ bool prop_less(const Prop& a, const Prop& b) {
    if (a.hdr != b.hdr) return a.hdr < b.hdr;
    if (a.gk != b.gk) return a.gk < b.gk;
    return a.cell < b.cell;
}

// workgroup helpers ---------------------------------------------------------
template <int WG>
struct Shared {
    int wave_cnt[WG / 64];
    int ibuf[4];
    double rbuf[WG / 64];
    int hist[BATCH_LMAX];
    int inst;
    // connectivity search state
    int bfs_cnt[2];
    uint32_t bfs_present[2];
    uint32_t bfs_adj;
    int bfs_res, bfs_nb;
};

// exclusive prefix of a 0/1 flag over the workgroup, in thread order
template <int WG>
__device__ __forceinline__ int block_scan_flag(Shared<WG>& sh, bool f, int& total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t bal = __ballot(f);
    const int before = __popcll(bal & ((1ull << lane) - 1ull));
    if (lane == 0) sh.wave_cnt[wave] = __popcll(bal);
    __syncthreads();
    int off = 0, tot = 0;
    for (int w = 0; w < WG / 64; ++w) {
        const int c = sh.wave_cnt[w];
        off += w < wave ? c : 0;
        tot += c;
    }
    __syncthreads();
    total = tot;
    return off + before;
}

template <int WG>
__device__ __forceinline__ int block_sum_int(Shared<WG>& sh, int v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0) sh.wave_cnt[threadIdx.x >> 6] = v;
    __syncthreads();
    int t = 0;
    for (int w = 0; w < WG / 64; ++w) t += sh.wave_cnt[w];
    __syncthreads();
    return t;
}

template <int WG, class R>
__device__ __forceinline__ R block_max(Shared<WG>& sh, R v) {
    for (int o = 32; o > 0; o >>= 1) {
        const R u = __shfl_xor(v, o);
        v = v < u ? u : v;
    }
    if ((threadIdx.x & 63) == 0) sh.rbuf[threadIdx.x >> 6] = double(v);
    __syncthreads();
    R m = R(sh.rbuf[0]);
    for (int w = 1; w < WG / 64; ++w) {
        const R u = R(sh.rbuf[w]);
        m = m < u ? u : m;
    }
    __syncthreads();
    return m;
}

// interface mask of every cell and the ordered band list (+ slot map)
template <int WG>
__device__ int band_compact(Shared<WG>& sh, const Dims& d, const int32_t* __restrict__ phi,
                            uint8_t* __restrict__ mask, int* __restrict__ band, int* __restrict__ slot,
                            bool with_list) {
    int nb = 0;
    for (int base = 0; base < d.N; base += WG) {
        const int i = base + int(threadIdx.x);
        bool m = false;
        if (i < d.N) {
            int f[6];
            faces(d, i, f);
            const int32_t v = phi[i];
            m = (v != phi[f[0]]) | (v != phi[f[1]]) | (v != phi[f[2]]) | (v != phi[f[3]]);
            mask[i] = uint8_t(m);
        }
        if (with_list) {
            int tot;
            const int pos = block_scan_flag<WG>(sh, m, tot);
            if (m) {
                band[nb + pos] = i;
                slot[i] = nb + pos;
            }
            nb += tot;
        }
    }
    __syncthreads();
    return nb;
}

// Are the face branches of label L around `cell` (the cell itself left out
// when excl) one connected component of L on the periodic instance?  The
// reference answers with a depth-first search from the first branch towards
// the others (src/gpk_connectivity.cpp:79-115).  Here every branch grows at
// once, one breadth-first level per pass of the workgroup; colours that meet
// are joined, and the search ends when all colours are joined (connected) or
// a joined group has no frontier left (its component is closed without the
// others: disconnected) -- the same answer.  mark[c] = (tag << 3) | colour
// for cells visited by search `tag`; the frontiers hold (cell << 3) | colour
// and never exceed the instance size, as every cell is claimed once.
template <int WG>
__device__ bool batch_connected(Shared<WG>& sh, const Dims& d, const int32_t* __restrict__ phi, int cell,
                                int32_t L, bool excl, int* __restrict__ mark, int* __restrict__ fr0,
                                int* __restrict__ fr1, int tag) {
    const int tid = threadIdx.x;
    if (tid == 0) {
        int f[6];
        faces(d, cell, f);
        int nb = 0;
        for (int q = 0; q < 4; ++q) {
            if (phi[f[q]] != L) continue;
            __hip_atomic_store(mark + f[q], (tag << 3) | nb, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            fr0[nb] = (f[q] << 3) | nb;
            ++nb;
        }
        sh.bfs_nb = nb;
        sh.bfs_cnt[0] = nb;
        sh.bfs_present[0] = (1u << nb) - 1u;
        sh.bfs_adj = 0u;
        sh.bfs_res = nb <= 1 ? 1 : -1;
    }
    __syncthreads();
    int cur = 0;
    while (true) {
        // thread 0 checks the level just finished and clears the counters of
        // the next frontier (last read two levels ago) in the same turn
        if (tid == 0) {
            sh.bfs_cnt[cur ^ 1] = 0;
            sh.bfs_present[cur ^ 1] = 0u;
        }
        if (tid == 0 && sh.bfs_res < 0) {
            const int nb = sh.bfs_nb;
            const uint32_t adj = sh.bfs_adj;
            int comp[4] = {0, 1, 2, 3};
            bool changed = true;
            while (changed) {
                changed = false;
                for (int a = 0; a < nb; ++a)
                    for (int b = 0; b < nb; ++b)
                        if (((adj >> (a * 4 + b)) & 1u) && comp[a] != comp[b]) {
                            const int m = comp[a] < comp[b] ? comp[a] : comp[b];
                            comp[a] = m;
                            comp[b] = m;
                            changed = true;
                        }
            }
            bool all = true;
            for (int q = 1; q < nb; ++q) all &= comp[q] == comp[0];
            if (all) {
                sh.bfs_res = 1;
            } else {
                const uint32_t present = sh.bfs_present[cur];
                for (int q = 0; q < nb; ++q) {
                    bool open = false;
                    for (int p = 0; p < nb; ++p) open |= comp[p] == comp[q] && ((present >> p) & 1u);
                    if (!open) sh.bfs_res = 0;
                }
            }
        }
        __syncthreads();
        if (sh.bfs_res >= 0) break;
        const int nxt = cur ^ 1;
        const int* fin = cur ? fr1 : fr0;
        int* fout = cur ? fr0 : fr1;
        const int n = sh.bfs_cnt[cur];
        for (int e = tid; e < n; e += WG) {
            const int v = fin[e];
            const int col = v & 7;
            const int want = (tag << 3) | col;
            int f[6];
            faces(d, v >> 3, f);
            for (int q = 0; q < 4; ++q) {
                const int c2 = f[q];
                if ((excl && c2 == cell) || phi[c2] != L) continue;
                int old = __hip_atomic_load(mark + c2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                while (true) {
                    if ((old >> 3) == tag) {
                        const int oc = old & 7;
                        if (oc != col) atomicOr(&sh.bfs_adj, (1u << (oc * 4 + col)) | (1u << (col * 4 + oc)));
                        break;
                    }
                    const int prev = atomicCAS(mark + c2, old, want);
                    if (prev == old) {
                        fout[atomicAdd(&sh.bfs_cnt[nxt], 1)] = (c2 << 3) | col;
                        atomicOr(&sh.bfs_present[nxt], 1u << col);
                        break;
                    }
                    old = prev;
                }
            }
        }
        __syncthreads();
        cur = nxt;
    }
    const bool ok = sh.bfs_res == 1;
    __syncthreads();
    return ok;
}

// 3x3 occupancy of labels ol and nl around cell (centre left out), bit
// (dy + 1) * 3 + (dx + 1)
__device__ __forceinline__ void cube_masks(const Dims& d, const int32_t* __restrict__ phi, int cell, int32_t ol,
                                           int32_t nl, uint32_t& co, uint32_t& cn) {
    const int y0 = cell / d.Nx, x0 = cell - y0 * d.Nx;
    co = 0u;
    cn = 0u;
    for (int dy = -1; dy <= 1; ++dy) {
        const int y = wrapi(y0 + dy, d.Ny);
        for (int dx = -1; dx <= 1; ++dx) {
            if (dy == 0 && dx == 0) continue;
            const int32_t l = phi[y * d.Nx + wrapi(x0 + dx, d.Nx)];
            const int t = (dy + 1) * 3 + dx + 1;
            co |= l == ol ? 1u << t : 0u;
            cn |= l == nl ? 1u << t : 0u;
        }
    }
}

template <class R, int WG, bool kDirect, class TS>
__device__ void step_instance(const BatchArgs<R>& A, Shared<WG>& sh, TS& tl, int b) {
    const Inst I = A.inst[b];
    const Dims d{1, I.Ny, I.Nx, I.N};
    const int tid = threadIdx.x;
    int32_t* phi = A.phi + I.off;
    R* alpha = A.alpha + I.off;
    R* g = A.g + I.off;
    const R* c = A.c + I.off;
    const R* T = A.T + I.off;
    uint8_t* mask = A.mask + I.off;
    int32_t* phi_comp = A.phi_comp + I.off;
    R* g_comp = A.g_comp + I.off;
    int* band = A.band + I.off;
    int* slot = A.slot + I.off;
    R* bnx = A.nx + I.off;
    R* bny = A.ny + I.off;
    R* bnz = A.nz + I.off;
    R* bk = A.kappa + I.off;
    R* vraw = A.vraw + I.off;
    R* vn = A.vn + I.off;
    R* gain = A.gain + I.off;
    Prop* props = A.props + I.soff;
    if (tid == 0) A.st[b].t0 = (long long)wall_clock64();

    // band (src/gpk_geometry.cpp:42-82)
    const int nb = band_compact<WG>(sh, d, phi, mask, band, slot, true);
    // geometry, competitor, constitutive velocity: in FP32 / FP64 at radius 8
    // the instance's tiles one after another out of LDS (the grid kernels'
    // tile code, same arithmetic as the per-cell path), else cell by cell
    bool tiled_geo = false;
    if constexpr (!kDirect) {
        if (A.gp.rg == 8) {
            using S = TiledLds<R, 2, 8>;
            const int ntx = (d.Nx + S::TX - 1) / S::TX, nty = (d.Ny + S::TY - 1) / S::TY;
            geometry_tile_lut(tl.geo, A.tab);
            for (int t = 0; t < ntx * nty; ++t)
                geometry_tile<R, 2, 8, 0>(tl.geo, t, d, A.gp, A.model, ntx, nty, phi, mask, slot, A.w0, A.w1, A.w2,
                                          g, c, T, static_cast<const R*>(nullptr), static_cast<const R*>(nullptr),
                                          bnx, bny, bnz, bk, vraw, phi_comp, g_comp);
            tiled_geo = true;
        }
    }
    if (!tiled_geo)
        for (int a = tid; a < nb; a += WG)
            geometry_cell<R, 2, kDirect>(d, A.gp, A.model, a, band[a], phi, A.w0, A.w1, A.w2, A.tab, A.lutd, g,
                                         c, T, static_cast<const R*>(nullptr), static_cast<const R*>(nullptr), bnx, bny, bnz, bk, vraw, phi_comp, g_comp);
    __syncthreads();
    // pair projection and the interface speed bound
    R vm = R(Eps<R>::vmax0);
    bool tiled_proj = false;
    if constexpr (!kDirect) {
        if (A.proj_r == 3) {
            using H = HaloGeom<2, 3>;
            const int ntx = (d.Nx + H::TX - 1) / H::TX, nty = (d.Ny + H::TY - 1) / H::TY;
            typename UBits<R>::T mb = 0;
            for (int t = 0; t < ntx * nty; ++t) {
                const auto m = project_tile<R, 2, 3>(tl.proj, t, d, ntx, nty, mask, phi, phi_comp, slot, A.projw,
                                                     vraw, vn);
                mb = mb < m ? m : mb;
            }
            const R av = bits_abs<R>(mb);
            vm = vm < av ? av : vm;
            tiled_proj = true;
        }
    }
    if (!tiled_proj)
        for (int a = tid; a < nb; a += WG) {
            const R v = project_value<R, 2>(d, A.proj_r, a, band[a], slot, mask, phi, phi_comp, A.projw, vraw);
            vn[a] = v;
            const R av = dfabs(v);
            vm = vm < av ? av : vm;
        }
    const R vmax = block_max<WG, R>(sh, vm);
    // dt (src/gpk_simulation.cpp:11-40), D = 0
    const R dx = R(A.dx_f);
    const R dti = R(Eps<R>::c_iface) * dx / vmax;
    const R stable = R(Eps<R>::big) < dti ? R(Eps<R>::big) : dti;
    R dt = A.dt_req < R(0) ? stable : (stable < A.dt_req ? stable : A.dt_req);
    dt = A.cap < dt ? A.cap : dt;
    // alpha and flips (src/gpk_transitions.cpp:251-313)
    // (FP32 mode: decisions within the guard band taken again in FP64)
    const R idt = R(double(dt)) / dx;
    int acc = 0, gd = 0, deferred = 0;
    if (!A.network) {
        // without the topology test every candidate is accepted
        for (int a = tid; a < nb; a += WG) {
            const int i = band[a];
            R an;
            bool guarded;
            const bool below = alpha_step<R>(alpha[i], vn[a], idt, double(dt), double(dx), !kDirect, an, guarded);
            gd += guarded ? 1 : 0;
            if (below && phi_comp[i] != phi[i]) {
                phi[i] = phi_comp[i];
                g[i] = g_comp[i];
                alpha[i] = R(1);
                ++acc;
            } else {
                alpha[i] = an;
            }
        }
    } else {
        // candidates in band (= index) order; alpha_new stored for all.  The
        // band list is not read again in this step before the remap rebuilds
        // it, so its front takes each candidate's old label (read here, before
        // any flip, since a candidate is the only writer of its own label)
        int* cand = reinterpret_cast<int*>(props);   // sort scratch: 4 ints per cell
        int32_t* cand_old = band;
        int* mark = cand + I.spow;
        int* fr0 = cand + 2 * I.spow;
        int* fr1 = cand + 3 * I.spow;
        int nc = 0;
        for (int base = 0; base < nb; base += WG) {
            const int a = base + tid;
            bool f = false;
            int i = 0;
            if (a < nb) {
                i = band[a];
                R an;
                bool guarded;
                const bool below =
                    alpha_step<R>(alpha[i], vn[a], idt, double(dt), double(dx), !kDirect, an, guarded);
                gd += guarded ? 1 : 0;
                alpha[i] = an;
                f = below && phi_comp[i] != phi[i];
            }
            const int32_t pi = f ? phi[i] : 0;
            int tot;
            const int pos = block_scan_flag<WG>(sh, f, tot);   // every band[a] of the pass read before
            if (f) {
                cand[nc + pos] = i;
                cand_old[nc + pos] = pi;
            }
            nc += tot;
        }
        // label counts, kept current as flips are accepted
        for (int l = tid; l < BATCH_LMAX; l += WG) sh.hist[l] = 0;
        __syncthreads();
        for (int i = tid; i < d.N; i += WG) atomicAdd(&sh.hist[phi[i]], 1);
        __syncthreads();
        // one candidate after another, each on the labels the earlier ones
        // left.  Every thread evaluates the 3x3 numbers (the same values) so
        // that the connectivity searches are entered by the whole workgroup;
        // only thread 0 reads the label counts, which it alone updates, so
        // one barrier per candidate orders the updates.  (With oo == 0 the
        // count rule may refuse a candidate whose new-label search then ran
        // for nothing; the verdict is the same.)
        int tag = 0;
        for (int e = 0; e < nc; ++e) {
            // ol from the list: thread 0 may already have relabelled cell i
            // while a slower wavefront starts this iteration
            const int i = cand[e];
            const int32_t ol = cand_old[e], nl = phi_comp[i];
            uint32_t co, cn;
            cube_masks(d, phi, i, ol, nl, co, cn);
            const int oo = topo_count(co, false), no = topo_count(cn, false);
            // src/gpk_connectivity.cpp:179-206: a new label that touches the
            // cell, connected branches of both labels, and (below) a
            // vanishing old label only when it is the label's last cell
            bool ok = no > 0;
            if (ok && (oo > 1 || no > 1) && tag == 0) {
                for (int c2 = tid; c2 < d.N; c2 += WG) mark[c2] = 0;
                __syncthreads();
            }
            if (ok && oo > 1) ok = batch_connected<WG>(sh, d, phi, i, ol, true, mark, fr0, fr1, ++tag);
            if (ok && no > 1) ok = batch_connected<WG>(sh, d, phi, i, nl, false, mark, fr0, fr1, ++tag);
            if (tid == 0) {
                ok = ok && !(oo == 0 && sh.hist[ol] > 1);
                if (ok) {
                    phi[i] = nl;
                    g[i] = g_comp[i];
                    alpha[i] = R(1);
                    --sh.hist[ol];
                    ++sh.hist[nl];
                } else {
                    const R an = alpha[i];
                    alpha[i] = an < Eps<R>::flip ? Eps<R>::flip : an;
                }
            }
            if (tid == 0) {
                acc += ok ? 1 : 0;
                deferred += ok ? 0 : 1;
            }
            __syncthreads();
        }
    }
    const int accepted = block_sum_int<WG>(sh, acc);
    const int nguard = block_sum_int<WG>(sh, gd);
    const int ndeferred = block_sum_int<WG>(sh, deferred);

    // balanced Gaussian remap (src/gpk_transitions.cpp:38-178), one sweep in 2-D
    int exchanges = 0, np = 0;
    if (accepted > 0) {
        const int nb2 = band_compact<WG>(sh, d, phi, mask, band, slot, true);
        for (int i = tid; i < d.N; i += WG) phi_comp[i] = phi[i];
        for (int l = tid; l < BATCH_LMAX; l += WG) sh.hist[l] = 0;
        __syncthreads();
        for (int i = tid; i < d.N; i += WG) atomicAdd(&sh.hist[phi[i]], 1);
        // scores (tiled in FP32 / FP64 at radius 3: gain and proposal flag per
        // band slot, best label in phi_comp) and the proposal list in band order
        bool tiled_remap = false;
        if constexpr (!kDirect) {
            if (A.remap_r == 3) {
                using H = HaloGeom<2, 3>;
                const int ntx = (d.Nx + H::TX - 1) / H::TX, nty = (d.Ny + H::TY - 1) / H::TY;
                uint8_t* pflag = A.pflag + I.off;
                const ScoreCache<R> nocache{nullptr, nullptr, nullptr, nullptr};
                for (int t = 0; t < ntx * nty; ++t)
                    remap_score_tile<R, 2, 3, false>(tl.remap, t, d, ntx, nty, mask, phi, slot, A.remapw, phi_comp,
                                                     pflag, gain, nocache);
                __syncthreads();
                for (int base = 0; base < nb2; base += WG) {
                    const int a = base + tid;
                    bool f = false;
                    Prop pr{};
                    if (a < nb2 && pflag[a]) {
                        const int i = band[a];
                        const int32_t p = phi[i], bl = phi_comp[i];
                        const int32_t lo = p < bl ? p : bl, hi = p < bl ? bl : p;
                        pr.hdr = (uint32_t(lo) << 11) | (uint32_t(hi) << 1) | (p == lo ? 0u : 1u);
                        pr.cell = i;
                        pr.gk = ~orderable(gain[a]);
                        f = true;
                    }
                    int tot;
                    const int pos = block_scan_flag<WG>(sh, f, tot);
                    if (f) props[np + pos] = pr;
                    np += tot;
                }
                tiled_remap = true;
            }
        }
        for (int base = 0; base < nb2 && !tiled_remap; base += WG) {
            const int a = base + tid;
            bool f = false;
            Prop pr{};
            if (a < nb2) {
                const int i = band[a];
                int32_t bl;
                const R gn = remap_score_cell<R, 2>(d, A.remap_r, i, A.phi + I.off, A.remapw, bl);
                phi_comp[i] = bl;
                gain[a] = gn;
                const int32_t p = phi[i];
                f = p != bl && gn > Eps<R>::gain;
                if (f) {
                    const int32_t lo = p < bl ? p : bl, hi = p < bl ? bl : p;
                    pr.hdr = (uint32_t(lo) << 11) | (uint32_t(hi) << 1) | (p == lo ? 0u : 1u);
                    pr.cell = i;
                    pr.gk = ~orderable(gn);
                }
            }
            int tot;
            const int pos = block_scan_flag<WG>(sh, f, tot);
            if (f) props[np + pos] = pr;
            np += tot;
        }
        __syncthreads();
        if (np > 1) {
            int P = 1;
            while (P < np) P <<= 1;
            for (int k = np + tid; k < P; k += WG) props[k] = Prop{0xFFFFFFFFu, 0x7FFFFFFF, ~0ull};
            __syncthreads();
            for (int k = 2; k <= P; k <<= 1)
                for (int j = k >> 1; j > 0; j >>= 1) {
                    for (int t = tid; t < P; t += WG) {
                        const int u = t ^ j;
                        if (u > t) {
                            const Prop x = props[t], y = props[u];
                            const bool up = (t & k) == 0;
                            if (up ? prop_less(y, x) : prop_less(x, y)) {
                                props[t] = y;
                                props[u] = x;
                            }
                        }
                    }
                    __syncthreads();
                }
        }
        // pairs in (lo, hi) order, k-th forward with k-th reverse, each
        // decided on the labels left by the pairs before it
        if (tid == 0 && np > 1) {
            auto lab = [&](int cc) -> int32_t { return phi[cc]; };
            int s = 0;
            while (s < np) {
                const uint32_t key = props[s].hdr >> 1;
                int e0 = s;
                while (e0 < np && (props[e0].hdr >> 1) == key && (props[e0].hdr & 1u) == 0u) ++e0;
                int e1 = e0;
                while (e1 < np && (props[e1].hdr >> 1) == key) ++e1;
                const int32_t lo = int32_t(key >> 10), hi = int32_t(key & 1023u);
                const int cnt = min(e0 - s, e1 - e0);
                for (int k = 0; k < cnt; ++k) {
                    const int ca = props[s + k].cell, cb = props[e0 + k].cell;
                    if (phi[ca] != lo || phi[cb] != hi) continue;
                    int fa[6], fb[6];
                    faces(d, ca, fa);
                    faces(d, cb, fb);
                    R ga = g[ca], gb = g[cb];
                    for (int q = 0; q < 4; ++q) if (phi[fa[q]] == hi) { ga = g[fa[q]]; break; }
                    for (int q = 0; q < 4; ++q) if (phi[fb[q]] == lo) { gb = g[fb[q]]; break; }
                    auto mism = [&](const int* f, int32_t prop, int pc, int32_t pp) {
                        int m = 0;
                        for (int q = 0; q < 4; ++q) m += ((f[q] == pc ? pp : phi[f[q]]) != prop);
                        return m;
                    };
                    const int before = mism(fa, lo, cb, hi) + mism(fb, hi, ca, lo);
                    const int after = mism(fa, hi, cb, lo) + mism(fb, lo, ca, hi);
                    if (after >= before) continue;
                    if (!bounded_cert(d, false, ca, lo, hi, sh.hist[lo], lab)) continue;
                    phi[ca] = hi;
                    if (!bounded_cert(d, false, cb, hi, lo, sh.hist[hi] + 1, lab)) {
                        phi[ca] = lo;
                        continue;
                    }
                    phi[cb] = lo;
                    g[ca] = ga;
                    g[cb] = gb;
                    alpha[ca] = R(1);
                    alpha[cb] = R(1);
                    exchanges += 2;
                }
                s = e1;
            }
        }
        __syncthreads();
    }
    // unresolved-grain extinction, every 8th step (src/gpk_transitions.cpp:180-249)
    int eroded = 0;
    if (A.network && (A.st[b].step_id & 7LL) == 0) {
        for (int l = tid; l < BATCH_LMAX; l += WG) sh.hist[l] = 0;
        __syncthreads();
        for (int i = tid; i < d.N; i += WG) atomicAdd(&sh.hist[phi[i]], 1);
        __syncthreads();
        int pres = 0, unr = 0;
        for (int l = tid; l < BATCH_LMAX; l += WG) {
            const int h = sh.hist[l];
            pres += h > 0 ? 1 : 0;
            unr += (h > 0 && h < A.min_cells) ? 1 : 0;
        }
        pres = block_sum_int<WG>(sh, pres);
        unr = block_sum_int<WG>(sh, unr);
        if (unr > 0 && unr != pres) {
            // layer by layer: targets from the labels before the layer
            int* tgt = reinterpret_cast<int*>(props);
            R* tgg = reinterpret_cast<R*>(tgt + 2 * I.spow);
            auto unres = [&](int32_t l) { const int h = sh.hist[l]; return h > 0 && h < A.min_cells; };
            while (true) {
                for (int i = tid; i < d.N; i += WG) {
                    int t = -1;
                    if (unres(phi[i])) {
                        const int y0 = i / d.Nx, x0 = i - y0 * d.Nx;
                        int f[6];
                        faces(d, i, f);
                        int bs = -1;
                        for (int k = 0; k < 4; ++k) {
                            const int32_t q = phi[f[k]];
                            if (unres(q)) continue;
                            int sup = 0;
                            for (int dy = -1; dy <= 1; ++dy) {
                                const int y = wrapi(y0 + dy, d.Ny);
                                for (int dx = -1; dx <= 1; ++dx)
                                    sup += phi[y * d.Nx + wrapi(x0 + dx, d.Nx)] == q ? 1 : 0;
                            }
                            if (sup > bs) {
                                bs = sup;
                                t = q;
                                tgg[i] = g[f[k]];
                            }
                        }
                    }
                    tgt[i] = t;
                }
                __syncthreads();
                int ch = 0;
                for (int i = tid; i < d.N; i += WG) {
                    const int t = tgt[i];
                    if (t < 0) continue;
                    phi[i] = t;
                    g[i] = tgg[i];
                    alpha[i] = R(1);
                    ++ch;
                }
                ch = block_sum_int<WG>(sh, ch);
                eroded += ch;
                if (ch == 0) break;
            }
        }
    }
    if (accepted > 0 || eroded > 0) band_compact<WG>(sh, d, phi, mask, band, slot, false);
    if (tid == 0) {
        InstStat& S = A.st[b];
        S.dt = double(dt);
        S.time += double(dt);
        S.band = nb;
        S.accepted = accepted;
        S.exchanges = exchanges;
        S.proposals = np;
        S.band_total += nb;
        S.accepted_total += accepted;
        S.exchanges_total += exchanges;
        S.proposals_total += np;
        S.guarded_total += nguard;
        S.deferred_total += ndeferred;
        S.eroded_total += eroded;
        S.step_id += 1;
        S.t1 = (long long)wall_clock64();
    }
    __syncthreads();
}

__global__ void k_batch_reset(int* next) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *next = 0;
}

// persistent: workgroups take instances from the queue until it is empty
template <class R, int WG, bool kDirect>
__global__ void __launch_bounds__(WG) k_batch_step(BatchArgs<R> A) {
    __shared__ Shared<WG> sh;
    __shared__ typename std::conditional<kDirect, NoTileSmem, BatchTileSmem<R>>::type tl;
    while (true) {
        if (threadIdx.x == 0) sh.inst = atomicAdd(A.next, 1);
        __syncthreads();
        const int q = sh.inst;
        __syncthreads();
        if (q >= A.n_inst) break;
        step_instance<R, WG, kDirect>(A, sh, tl, A.order[q]);
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct BatchBase {
    virtual ~BatchBase() = default;
    virtual void upload(int field, const void* src, size_t bytes) = 0;
    virtual void download(int field, void* dst, size_t bytes) = 0;
    virtual void step(int nsteps, float dt, float cap) = 0;
    virtual void stats(gpk_batch_stats* s) = 0;
    virtual void inst_info(double* time, double* last_ms, int* band) = 0;
    virtual void autotune(int* wg, int* nwg, double* best_ms) = 0;
    virtual void set_launch(int wg, int nwg) = 0;
};

template <class R, bool kDirect>
struct Batch final : BatchBase {
    int n_inst = 0;
    long long Ntot = 0, Stot = 0;
    std::vector<Inst> hinst;
    float dx_f, sigma_f;
    gpk_model_desc model;
    int device;
    hipStream_t stream = nullptr;
    Weights W;
    int proj_r = 3, remap_r = 3;
    RowLutDesc lutd{};
    DevBuf inst, order, next, st;
    DevBuf phi, alpha, g, c, T, mask, phi_comp, g_comp, band, slot;
    DevBuf nx, ny, nz, kappa, vraw, vn, gain, pflag, props;
    DevBuf wt0, wt1, wt2, tab, projw, remapw;
    int wg = 256, nwg = 0;
    double clock_hz = 1e8;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    float g_dt = 0.f, g_cap = 0.f;
    int g_wg = 0, g_nwg = 0;

    Batch(int n, const int* Ny, const int* Nx, float dx, float sigma, const gpk_model_desc& m, int dev)
        : n_inst(n), dx_f(dx), sigma_f(sigma), model(m), device(dev) {
        HIPCHK(hipSetDevice(device));
        HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        hinst.resize(n);
        for (int b = 0; b < n; ++b) {
            Inst& I = hinst[b];
            I.Ny = Ny[b];
            I.Nx = Nx[b];
            I.N = Ny[b] * Nx[b];
            I.off = int(Ntot);
            int P = 1;
            while (P < I.N) P <<= 1;
            I.soff = int(Stot);
            I.spow = P;
            Ntot += I.N;
            Stot += P;
        }
        if (Ntot > INT32_MAX / 2 || Stot > INT32_MAX / 2)
            throw GpkError(GPK_E_INVALID, "batch too large for 32-bit cell offsets");
        W = sizeof(R) == 4 ? gauss_float(sigma_f) : gauss_double(sigma_f);
        const float vel_sigma = 0.85f, remap_sigma = 0.70f;   // 2-D values
        proj_r = std::max(1, static_cast<int>(std::ceil(3.f * vel_sigma)));
        remap_r = std::max(1, static_cast<int>(std::ceil(3.f * remap_sigma)));
        allocate(vel_sigma, remap_sigma);
        int rate_khz = 0;
        if (hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, device) == hipSuccess &&
            rate_khz > 0)
            clock_hz = double(rate_khz) * 1e3;
        int cus = 0;
        HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
        wg = 256;
        nwg = std::min(n_inst, 2 * cus);
    }

    ~Batch() override {
        if (stream) (void)hipStreamSynchronize(stream);
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        if (stream) (void)hipStreamDestroy(stream);
    }

    void allocate(float vel_sigma, float remap_sigma) {
        const size_t n = size_t(Ntot), rs = sizeof(R);
        inst.alloc(sizeof(Inst) * n_inst);
        HIPCHK(hipMemcpy(inst.p, hinst.data(), sizeof(Inst) * n_inst, hipMemcpyHostToDevice));
        // largest instances first
        std::vector<int> ord(n_inst);
        std::iota(ord.begin(), ord.end(), 0);
        std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return hinst[a].N > hinst[b].N; });
        order.alloc(sizeof(int) * n_inst);
        HIPCHK(hipMemcpy(order.p, ord.data(), sizeof(int) * n_inst, hipMemcpyHostToDevice));
        next.alloc(16);
        st.alloc(sizeof(InstStat) * n_inst);
        HIPCHK(hipMemset(st.p, 0, sizeof(InstStat) * n_inst));
        phi.alloc(n * 4); alpha.alloc(n * rs); g.alloc(n * rs); c.alloc(n * rs); T.alloc(n * rs);
        mask.alloc(n); phi_comp.alloc(n * 4); g_comp.alloc(n * rs); band.alloc(n * 4); slot.alloc(n * 4);
        nx.alloc(n * rs); ny.alloc(n * rs); nz.alloc(n * rs); kappa.alloc(n * rs); vraw.alloc(n * rs);
        vn.alloc(n * rs); gain.alloc(n * rs); pflag.alloc(n);
        props.alloc(size_t(Stot) * sizeof(Prop));
        HIPCHK(hipMemset(phi.p, 0, n * 4));
        HIPCHK(hipMemset(phi_comp.p, 0, n * 4));
        HIPCHK(hipMemset(mask.p, 0, n));
        HIPCHK(hipMemset(g.p, 0, n * rs));
        HIPCHK(hipMemset(c.p, 0, n * rs));
        HIPCHK(hipMemset(T.p, 0, n * rs));
        HIPCHK(hipMemset(g_comp.p, 0, n * rs));
        {
            std::vector<R> one(n, R(1));
            HIPCHK(hipMemcpy(alpha.p, one.data(), n * rs, hipMemcpyHostToDevice));
        }
        std::vector<R> w0(W.w0.begin(), W.w0.end()), w1(W.w1.begin(), W.w1.end()),
            w2(W.w2.begin(), W.w2.end());
        const size_t Wd = w0.size();
        wt0.alloc(Wd * rs); wt1.alloc(Wd * rs); wt2.alloc(Wd * rs);
        HIPCHK(hipMemcpy(wt0.p, w0.data(), Wd * rs, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(wt1.p, w1.data(), Wd * rs, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(wt2.p, w2.data(), Wd * rs, hipMemcpyHostToDevice));
        std::vector<R> t;
        if (kDirect) t = direct_table<R>(w0, w1, w2, false);
        else lutd = make_row_lut<R>(w0, w1, w2, t);
        tab.alloc(t.size() * rs);
        HIPCHK(hipMemcpy(tab.p, t.data(), t.size() * rs, hipMemcpyHostToDevice));
        const bool single = sizeof(R) == 4;
        const auto pw = radial_table(vel_sigma, proj_r, single);
        const auto rw = radial_table(remap_sigma, remap_r, single);
        projw.alloc(pw.size() * 8); remapw.alloc(rw.size() * 8);
        HIPCHK(hipMemcpy(projw.p, pw.data(), pw.size() * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(remapw.p, rw.data(), rw.size() * 8, hipMemcpyHostToDevice));
    }

    DevBuf* field_buf(int f, size_t& esize) {
        esize = sizeof(R);
        switch (f) {
        case GPK_FIELD_PHI: esize = 4; return &phi;
        case GPK_FIELD_PHI_COMP: esize = 4; return &phi_comp;
        case GPK_FIELD_MASK: esize = 1; return &mask;
        case GPK_FIELD_ALPHA: return &alpha;
        case GPK_FIELD_G: return &g;
        case GPK_FIELD_C: return &c;
        case GPK_FIELD_T: return &T;
        case GPK_FIELD_G_COMP: return &g_comp;
        default: throw GpkError(GPK_E_INVALID, "field not held by a batch");
        }
    }

    void upload(int f, const void* src, size_t bytes) override {
        size_t es;
        DevBuf* b = field_buf(f, es);
        if (bytes != size_t(Ntot) * es) throw GpkError(GPK_E_INVALID, "batch field size mismatch");
        if (f == GPK_FIELD_PHI) {
            const int32_t* p = static_cast<const int32_t*>(src);
            for (long long i = 0; i < Ntot; ++i)
                if (p[i] < 0 || p[i] >= BATCH_LMAX)
                    throw GpkError(GPK_E_LABEL_RANGE, "batch labels must lie in [0, 1024)");
        }
        HIPCHK(hipMemcpyAsync(b->p, src, bytes, hipMemcpyHostToDevice, stream));
        HIPCHK(hipStreamSynchronize(stream));
    }

    void download(int f, void* dst, size_t bytes) override {
        size_t es;
        DevBuf* b = field_buf(f, es);
        if (bytes != size_t(Ntot) * es) throw GpkError(GPK_E_INVALID, "batch field size mismatch");
        HIPCHK(hipMemcpyAsync(dst, b->p, bytes, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
    }

    BatchArgs<R> args(float dt, float cap) {
        BatchArgs<R> A{};
        A.n_inst = n_inst;
        A.inst = inst.as<Inst>();
        A.order = order.as<int>();
        A.next = next.as<int>();
        A.st = st.as<InstStat>();
        A.gp = GeomParams{W.rg, 2 * W.rg + 1, dx_f, W.sigma_cells};
        A.model = model;
        A.dx_f = dx_f;
        A.dt_req = R(dt);
        A.cap = R(cap);
        A.proj_r = proj_r;
        A.remap_r = remap_r;
        A.network = (model.flags & GPK_FLAG_GRAIN_NETWORK) ? 1 : 0;
        {
            const double sc = std::max(W.sigma_cells, 1.f);
            A.min_cells = std::max(2, static_cast<int>(std::ceil(M_PI * sc * sc)));
        }
        A.w0 = wt0.as<R>(); A.w1 = wt1.as<R>(); A.w2 = wt2.as<R>(); A.tab = tab.as<R>();
        A.lutd = lutd;
        A.projw = projw.as<double>(); A.remapw = remapw.as<double>();
        A.phi = phi.as<int32_t>(); A.alpha = alpha.as<R>(); A.g = g.as<R>(); A.c = c.as<R>();
        A.T = T.as<R>(); A.mask = mask.as<uint8_t>(); A.phi_comp = phi_comp.as<int32_t>();
        A.g_comp = g_comp.as<R>(); A.band = band.as<int>(); A.slot = slot.as<int>();
        A.nx = nx.as<R>(); A.ny = ny.as<R>(); A.nz = nz.as<R>(); A.kappa = kappa.as<R>();
        A.vraw = vraw.as<R>(); A.vn = vn.as<R>(); A.gain = gain.as<R>(); A.pflag = pflag.as<uint8_t>();
        A.props = props.as<Prop>();
        return A;
    }

    void launch(const BatchArgs<R>& A, int wgs, int nw) {
        k_batch_reset<<<1, 64, 0, stream>>>(next.as<int>());
        switch (wgs) {
        case 256: k_batch_step<R, 256, kDirect><<<nw, 256, 0, stream>>>(A); break;
        case 512: k_batch_step<R, 512, kDirect><<<nw, 512, 0, stream>>>(A); break;
        case 1024: k_batch_step<R, 1024, kDirect><<<nw, 1024, 0, stream>>>(A); break;
        default: throw GpkError(GPK_E_INVALID, "workgroup size must be 256, 512 or 1024");
        }
        HIPCHK(hipGetLastError());
    }

    // one step = queue reset + persistent step kernel, captured once
    void ensure_graph(float dt, float cap) {
        if (gexec && g_dt == dt && g_cap == cap && g_wg == wg && g_nwg == nwg) return;
        if (gexec) { HIPCHK(hipGraphExecDestroy(gexec)); gexec = nullptr; }
        if (graph) { HIPCHK(hipGraphDestroy(graph)); graph = nullptr; }
        const BatchArgs<R> A = args(dt, cap);
        HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        try {
            launch(A, wg, nwg);
        } catch (...) {
            hipGraph_t tmp = nullptr;
            (void)hipStreamEndCapture(stream, &tmp);
            if (tmp) (void)hipGraphDestroy(tmp);
            throw;
        }
        HIPCHK(hipStreamEndCapture(stream, &graph));
        HIPCHK(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
        g_dt = dt; g_cap = cap; g_wg = wg; g_nwg = nwg;
    }

    void step(int nsteps, float dt, float cap) override {
        ensure_graph(dt, cap);
        for (int s = 0; s < nsteps; ++s) HIPCHK(hipGraphLaunch(gexec, stream));
        HIPCHK(hipStreamSynchronize(stream));
    }

    std::vector<InstStat> pull_stats() {
        std::vector<InstStat> h(n_inst);
        HIPCHK(hipMemcpyAsync(h.data(), st.p, sizeof(InstStat) * n_inst, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        return h;
    }

    void stats(gpk_batch_stats* s) override {
        const auto h = pull_stats();
        std::memset(s, 0, sizeof(*s));
        for (const auto& x : h) {
            s->band_cells += x.band_total;
            s->accepted += x.accepted_total;
            s->exchanges += x.exchanges_total;
            s->proposals += x.proposals_total;
            s->guarded += x.guarded_total;
            s->deferred += x.deferred_total;
            s->eroded += x.eroded_total;
        }
        s->instances = n_inst;
        s->workgroup_size = wg;
        s->workgroups = nwg;
    }

    void inst_info(double* time, double* last_ms, int* bnd) override {
        const auto h = pull_stats();
        for (int b = 0; b < n_inst; ++b) {
            if (time) time[b] = h[b].time;
            if (last_ms) last_ms[b] = 1e3 * double(h[b].t1 - h[b].t0) / clock_hz;
            if (bnd) bnd[b] = h[b].band;
        }
    }

    void set_launch(int w, int nw) override {
        if (w != 256 && w != 512 && w != 1024) throw GpkError(GPK_E_INVALID, "workgroup size 256/512/1024");
        if (nw < 1) throw GpkError(GPK_E_INVALID, "workgroup count must be positive");
        wg = w;
        nwg = std::min(nw, n_inst);
    }

    // times two steps per launch shape on a copy of the state, keeps the
    // fastest, then restores the state (labels, alpha, g, stats)
    void autotune(int* wg_out, int* nwg_out, double* best_ms) override {
        const size_t n = size_t(Ntot), rs = sizeof(R);
        DevBuf s_phi, s_alpha, s_g, s_st;
        s_phi.alloc(n * 4); s_alpha.alloc(n * rs); s_g.alloc(n * rs); s_st.alloc(sizeof(InstStat) * n_inst);
        HIPCHK(hipMemcpyAsync(s_phi.p, phi.p, n * 4, hipMemcpyDeviceToDevice, stream));
        HIPCHK(hipMemcpyAsync(s_alpha.p, alpha.p, n * rs, hipMemcpyDeviceToDevice, stream));
        HIPCHK(hipMemcpyAsync(s_g.p, g.p, n * rs, hipMemcpyDeviceToDevice, stream));
        HIPCHK(hipMemcpyAsync(s_st.p, st.p, sizeof(InstStat) * n_inst, hipMemcpyDeviceToDevice, stream));
        int cus = 0;
        HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        double best = 1e300;
        int bw = wg, bn = nwg;
        const BatchArgs<R> A = args(-1.f, 1e30f);
        const int sizes[3] = {256, 512, 1024};
        const int mult[4] = {1, 2, 4, 8};
        for (int wi = 0; wi < 3; ++wi)
            for (int mi = 0; mi < 4; ++mi) {
                const int w = sizes[wi];
                const int nw = std::min(n_inst, mult[mi] * cus * 256 / w);
                if (nw < 1) continue;
                launch(A, w, nw);   // warm-up (state restored below)
                HIPCHK(hipEventRecord(e0, stream));
                launch(A, w, nw);
                launch(A, w, nw);
                HIPCHK(hipEventRecord(e1, stream));
                HIPCHK(hipEventSynchronize(e1));
                float ms = 0.f;
                HIPCHK(hipEventElapsedTime(&ms, e0, e1));
                if (0.5 * ms < best) { best = 0.5 * ms; bw = w; bn = nw; }
                HIPCHK(hipMemcpyAsync(phi.p, s_phi.p, n * 4, hipMemcpyDeviceToDevice, stream));
                HIPCHK(hipMemcpyAsync(alpha.p, s_alpha.p, n * rs, hipMemcpyDeviceToDevice, stream));
                HIPCHK(hipMemcpyAsync(g.p, s_g.p, n * rs, hipMemcpyDeviceToDevice, stream));
                HIPCHK(hipMemcpyAsync(st.p, s_st.p, sizeof(InstStat) * n_inst, hipMemcpyDeviceToDevice, stream));
            }
        HIPCHK(hipStreamSynchronize(stream));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        wg = bw;
        nwg = bn;
        if (wg_out) *wg_out = bw;
        if (nwg_out) *nwg_out = bn;
        if (best_ms) *best_ms = best;
    }
};

} // namespace

struct gpk_batch {
    std::unique_ptr<BatchBase> b;
};

extern "C" {

int gpk_batch_create(int n_inst, const int* Ny, const int* Nx, float dx, float D, float gaussian_sigma,
                     int precision, int device, const gpk_model_desc* model, gpk_batch** out) {
    return guarded([&] {
        if (!out || !Ny || !Nx || !model || n_inst <= 0) throw GpkError(GPK_E_INVALID, "bad batch arguments");
        *out = nullptr;
        for (int b = 0; b < n_inst; ++b)
            if (Ny[b] < 3 || Nx[b] < 3) throw GpkError(GPK_E_INVALID, "instances need Ny, Nx >= 3");
        for (int b = 0; b < n_inst; ++b)
            if ((long long)Ny[b] * Nx[b] >= (1LL << 28))
                throw GpkError(GPK_E_INVALID, "batched instances hold fewer than 2^28 cells each");
        if (D != 0.f) throw GpkError(GPK_E_UNSUPPORTED_MODEL, "batched instances run with D = 0");
        if (model->flags & (GPK_FLAG_SOLUTE_REJECTION | GPK_FLAG_STOCHASTIC | GPK_FLAG_DISLOCATION |
                            GPK_FLAG_ELASTIC) ||
            model->kind == GPK_MODEL_PRECIPITATION || model->kind == GPK_MODEL_DISLOCATION)
            throw GpkError(GPK_E_UNSUPPORTED_MODEL,
                           "batched instances: models without solute rejection, noise, stored energy "
                           "or elastic terms");
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
            throw GpkError(GPK_E_NO_DEVICE, "gpk: no HIP device visible");
        auto h = std::make_unique<gpk_batch>();
        switch (precision) {
        case GPK_PREC_STRICT32:
            h->b.reset(new Batch<float, true>(n_inst, Ny, Nx, dx, gaussian_sigma, *model, device));
            break;
        case GPK_PREC_FP32:
            h->b.reset(new Batch<float, false>(n_inst, Ny, Nx, dx, gaussian_sigma, *model, device));
            break;
        case GPK_PREC_FP64:
            h->b.reset(new Batch<double, false>(n_inst, Ny, Nx, dx, gaussian_sigma, *model, device));
            break;
        default: throw GpkError(GPK_E_INVALID, "unknown precision");
        }
        *out = h.release();
    });
}

int gpk_batch_destroy(gpk_batch* h) {
    return guarded([&] { delete h; });
}

#define NEED_B \
    if (!h) throw GpkError(GPK_E_INVALID, "null batch handle")

int gpk_batch_upload(gpk_batch* h, int field, const void* host, size_t bytes) {
    return guarded([&] { NEED_B; h->b->upload(field, host, bytes); });
}
int gpk_batch_download(gpk_batch* h, int field, void* host, size_t bytes) {
    return guarded([&] { NEED_B; h->b->download(field, host, bytes); });
}
int gpk_batch_step(gpk_batch* h, int nsteps, float dt, float dt_cap) {
    return guarded([&] {
        NEED_B;
        if (nsteps < 0) throw GpkError(GPK_E_INVALID, "negative step count");
        h->b->step(nsteps, dt, dt_cap);
    });
}
int gpk_batch_stats_get(gpk_batch* h, gpk_batch_stats* s) {
    return guarded([&] { NEED_B; if (!s) throw GpkError(GPK_E_INVALID, "null"); h->b->stats(s); });
}
int gpk_batch_instances(gpk_batch* h, double* time, double* last_step_ms, int* band_cells) {
    return guarded([&] { NEED_B; h->b->inst_info(time, last_step_ms, band_cells); });
}
int gpk_batch_autotune(gpk_batch* h, int* workgroup_size, int* workgroups, double* step_ms) {
    return guarded([&] { NEED_B; h->b->autotune(workgroup_size, workgroups, step_ms); });
}
int gpk_batch_set_launch(gpk_batch* h, int workgroup_size, int workgroups) {
    return guarded([&] { NEED_B; h->b->set_launch(workgroup_size, workgroups); });
}

} // extern "C"
