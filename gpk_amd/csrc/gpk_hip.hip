// gpk_hip.hip — host orchestration and C ABI of libgpk_hip.so.
//
// One handle = one grid on one device with one HIP stream.  The per-step
// pipeline (GPKGrid::step, src/gpk_simulation.cpp:31-54) runs as a sequence
// of hand-written kernels on that stream; the host reads back only the few
// counts that size the next launch (band size, candidates, proposals) and
// the undecided count of the ordered transactions.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <queue>
#include <random>
#include <stdexcept>
#include <string>
#include <thread>
#include <type_traits>
#include <unordered_set>
#include <vector>

#include "gpk_kernels.hpp"
#include "gpk_adjoint.hpp"
#include "gpk_broad.hpp"
#include "gpk_host.hpp"
#include "gpk_sort.hpp"

using namespace gpkd;
using namespace gpkh;

namespace gpkh {
std::string& last_error() {
    thread_local std::string msg;
    return msg;
}
} // namespace gpkh

namespace {


// ---------------------------------------------------------------------------
// Voronoi generator (demos/gpk.cpp:51-189).  Site draws use std::mt19937 and
// the libstdc++ distributions.  On a grid the nearest-site labels come from
// k_voronoi_buckets and the connectivity normalisation runs on the device;
// gpk_voronoi_host does both on the host (the functions below).
// ---------------------------------------------------------------------------
struct Sites {
    std::vector<int> z, y, x;
    std::vector<float> g;
};

Sites draw_sites(int Nz, int Ny, int Nx, int n, uint32_t seed) {
    Sites S;
    std::mt19937 rng(seed);
    std::uniform_int_distribution<int> rz(0, Nz - 1), ry(0, Ny - 1), rx(0, Nx - 1);
    std::uniform_real_distribution<float> rg(0.f, float(M_PI) * 0.5f);
    S.z.resize(n); S.y.resize(n); S.x.resize(n); S.g.resize(n);
    std::unordered_set<long long> used;
    used.reserve(size_t(2) * n);
    for (int k = 0; k < n; ++k) {
        long long lin;
        do {
            S.z[k] = rz(rng); S.y[k] = ry(rng); S.x[k] = rx(rng);
            lin = (static_cast<long long>(S.z[k]) * Ny + S.y[k]) * Nx + S.x[k];
        } while (!used.insert(lin).second);
        S.g[k] = rg(rng);
    }
    return S;
}

// host nearest-site search over a bucket grid; exact: a site in a bucket at
// periodic bucket distance k >= ring+1 along some axis lies at least
// ring*Bmin + 1 cells away along that axis, so the search stops once the best
// squared distance is below that bound (ties at the bound keep searching and
// the lowest site index wins, as in the reference's strict < scan)
void nearest_host(int Nz, int Ny, int Nx, const Sites& S, int32_t* phi) {
    const int n = static_cast<int>(S.z.size());
    const long long N = 1LL * Nz * Ny * Nx;
    const double per = std::max(1.0, double(N) / n);
    const int B = std::max(1, Nz == 1 ? static_cast<int>(std::sqrt(per)) : static_cast<int>(std::cbrt(per)));
    const int Bz = Nz == 1 ? Nz : std::min(B, Nz), By = std::min(B, Ny), Bx = std::min(B, Nx);
    const int bz = (Nz + Bz - 1) / Bz, by = (Ny + By - 1) / By, bx = (Nx + Bx - 1) / Bx;
    // smallest bucket extent over the axes that are searched
    int Bmin = std::min(Ny - (by - 1) * By, By);
    Bmin = std::min(Bmin, std::min(Nx - (bx - 1) * Bx, Bx));
    if (Nz > 1) Bmin = std::min(Bmin, std::min(Nz - (bz - 1) * Bz, Bz));
    Bmin = std::max(Bmin, 1);
    std::vector<std::vector<int>> bucket(size_t(bz) * by * bx);
    for (int k = 0; k < n; ++k)
        bucket[(size_t(S.z[k] / Bz) * by + S.y[k] / By) * bx + S.x[k] / Bx].push_back(k);
    auto worker = [&](int z0, int z1) {
        for (int z = z0; z < z1; ++z)
            for (int y = 0; y < Ny; ++y)
                for (int x = 0; x < Nx; ++x) {
                    long long bd = LLONG_MAX;
                    int best = 0;
                    const int cz = z / Bz, cy = y / By, cx = x / Bx;
                    for (int ring = 0;; ++ring) {
                        const int rz = Nz == 1 ? 0 : ring;
                        for (int dz = -rz; dz <= rz; ++dz)
                            for (int dy = -ring; dy <= ring; ++dy)
                                for (int dxb = -ring; dxb <= ring; ++dxb) {
                                    if (std::max({std::abs(dz), std::abs(dy), std::abs(dxb)}) != ring) continue;
                                    const int qz = ((cz + dz) % bz + bz) % bz;
                                    const int qy = ((cy + dy) % by + by) % by;
                                    const int qx = ((cx + dxb) % bx + bx) % bx;
                                    for (int k : bucket[(size_t(qz) * by + qy) * bx + qx]) {
                                        long long a = std::abs(z - S.z[k]), b = std::abs(y - S.y[k]),
                                                  e = std::abs(x - S.x[k]);
                                        a = std::min(a, (long long)Nz - a);
                                        b = std::min(b, (long long)Ny - b);
                                        e = std::min(e, (long long)Nx - e);
                                        const long long d = a * a + b * b + e * e;
                                        if (d < bd || (d == bd && k < best)) { bd = d; best = k; }
                                    }
                                }
                        const bool covered_all = (2 * ring + 1 >= bx) && (2 * ring + 1 >= by) &&
                                                 (Nz == 1 || 2 * ring + 1 >= bz);
                        if (covered_all) break;
                        const long long lb = 1LL * ring * Bmin + 1;
                        if (bd < lb * lb) break;
                    }
                    phi[(size_t(z) * Ny + y) * Nx + x] = best + 1;
                }
    };
    const int nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (Nz >= nt) {
        std::vector<std::thread> th;
        const int step = (Nz + nt - 1) / nt;
        for (int t = 0; t < nt; ++t) th.emplace_back(worker, t * step, std::min(Nz, (t + 1) * step));
        for (auto& t : th) t.join();
    } else {
        worker(0, Nz);
    }
}

// keep the largest face component of each label; attach the rest layer by layer
int normalise_components(int Nz, int Ny, int Nx, int n_seeds, const Sites& S, int32_t* phi,
                         float* g) {
    const int N = Nz * Ny * Nx;
    for (int i = 0; i < N; ++i) g[i] = S.g[phi[i] - 1];
    auto idx = [&](int z, int y, int x) { return (z * Ny + y) * Nx + x; };
    auto nbrs = [&](int c, int o[6]) {
        const int x = c % Nx, y = (c / Nx) % Ny, z = c / (Ny * Nx);
        const int xm = x == 0 ? Nx - 1 : x - 1, xp = x == Nx - 1 ? 0 : x + 1;
        const int ym = y == 0 ? Ny - 1 : y - 1, yp = y == Ny - 1 ? 0 : y + 1;
        const int zm = z == 0 ? Nz - 1 : z - 1, zp = z == Nz - 1 ? 0 : z + 1;
        o[0] = idx(z, y, xm); o[1] = idx(z, y, xp); o[2] = idx(z, ym, x);
        o[3] = idx(z, yp, x); o[4] = idx(zm, y, x); o[5] = idx(zp, y, x);
    };
    const int nface = Nz > 1 ? 6 : 4;
    std::vector<int> comp(N, -1);
    std::vector<long long> csize;
    std::vector<int32_t> clabel;
    std::vector<int> largest(size_t(n_seeds) + 1, -1);
    std::vector<int> queue;
    queue.reserve(1024);
    for (int s = 0; s < N; ++s) {
        if (comp[s] >= 0) continue;
        const int id = static_cast<int>(csize.size());
        const int32_t lab = phi[s];
        queue.clear();
        queue.push_back(s);
        comp[s] = id;
        for (size_t h = 0; h < queue.size(); ++h) {
            int f[6];
            nbrs(queue[h], f);
            for (int k = 0; k < nface; ++k)
                if (comp[f[k]] < 0 && phi[f[k]] == lab) { comp[f[k]] = id; queue.push_back(f[k]); }
        }
        csize.push_back(static_cast<long long>(queue.size()));
        clabel.push_back(lab);
        const int prev = largest[lab];
        if (prev < 0 || csize[id] > csize[prev]) largest[lab] = id;
    }
    std::vector<uint8_t> att(N, 0);
    long long natt = 0;
    for (int i = 0; i < N; ++i)
        if (largest[clabel[comp[i]]] == comp[i]) { att[i] = 1; ++natt; }
    std::vector<int> cells;
    std::vector<int32_t> tg;
    std::vector<float> og;
    while (natt < N) {
        cells.clear(); tg.clear(); og.clear();
        for (int cc = 0; cc < N; ++cc) {
            if (att[cc]) continue;
            int f[6];
            nbrs(cc, f);
            int32_t cand[6];
            int sup[6] = {0, 0, 0, 0, 0, 0}, src[6], nc = 0;
            for (int k = 0; k < nface; ++k) {
                if (!att[f[k]]) continue;
                const int32_t v = phi[f[k]];
                int q = 0;
                while (q < nc && cand[q] != v) ++q;
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
        for (size_t k = 0; k < cells.size(); ++k) {
            phi[cells[k]] = tg[k]; g[cells[k]] = og[k]; att[cells[k]] = 1;
        }
        natt += static_cast<long long>(cells.size());
    }
    return GPK_OK;
}

// ---------------------------------------------------------------------------
// Engine interface
// ---------------------------------------------------------------------------
struct EngineBase {
    virtual ~EngineBase() = default;
    virtual void upload(int field, const void* src, size_t bytes, bool device) = 0;
    virtual void download(int field, void* dst, size_t bytes, bool device) = 0;
    virtual void step(float dt, float cap, gpk_step_stats* st) = 0;
    virtual void set_model(const gpk_model_desc& m) = 0;
    virtual double time() = 0;
    virtual void set_time(double t) = 0;
    virtual double max_dt() = 0;
    virtual double phase_volume(int32_t l) = 0;
    virtual int n_grains() = 0;
    virtual int interface_cells() = 0;
    virtual double mean_grain_radius() = 0;
    virtual float solid_fraction() = 0;
    virtual float interface_fraction() = 0;
    virtual double solute_sum() = 0;
    virtual size_t device_bytes() const = 0;
    virtual int simple_move(int cell, int32_t nl) = 0;
    virtual void voronoi(int n_seeds, uint32_t seed) = 0;
    virtual void sync() = 0;
    virtual void set_profiling(int on) = 0;
    virtual void timing(int which, double* ms, int64_t* n, double* bytes) = 0;
    virtual int64_t steps() const = 0;
    virtual void velocity_field(const double theta[3], void* vout) = 0;
    virtual void velocity_adjoint(const double theta[3], const void* gvel, double grad[3]) = 0;
    virtual void velocity_adjoint_noalloc(const double theta[3], const void* gvel, double grad[3]) = 0;
};

// size of the allocation-free adjoint's static device scratch (below)
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(16) << 20;
void scratch_init(int device);
void* scratch_of(int device);

template <class R, bool kDirect>
struct Engine final : EngineBase {
    // ---- configuration ----
    int Nz, Ny, Nx, N, dim3;
    float dx_f, D_f, sigma_f;
    gpk_model_desc model;
    int device;
    hipStream_t stream = nullptr;
    Dims dims;
    Weights W;
    int proj_r = 1, remap_r = 1;
    float vel_sigma = 0.85f, remap_sigma = 0.7f;
    int64_t step_id = 0;
    int nlab = 0;           // label table size (max label + 1)
    RowLutDesc lutd{};      // row tables of the separable modes

    // ---- device state ----
    DevBuf phi, mask, phi_comp, slot, pid, mark;
    DevBuf alpha, c, T, g, rho, g_comp, work1, work2, work3, work4;
    DevBuf bandA, bandB, nx, ny, nz, kappa, vraw, vn;
    DevBuf cflag, anew_s, sel;
    DevBuf Ccell, Cold, Cnew, Cg, Canew, state, pos_old, pos_new, qflags, queue0, queue1, fbq, ovq;
    DevBuf Pk, Ps;
    DevBuf gain_s, pcell, kgain, idx, order, kgrp, kgrp2, order2, seg_start, seg_end,
        pflag2, pidx;
    // radix sort: ping-pong keys and payloads, (digit, block) counts, total
    DevBuf rs_keys, rs_vals, rs_cnt, rs_total;
    DevBuf Pa, Pb, Plo, Phi, pstate, Pg, pdeps, cmp_cnt, pkey, pvfull;
    bool mask_fresh = false;    // mask matches the current labels
    int flip_batch = 4, pair_batch = 4;   // rounds launched before the first host check
    // visited cells after which a wavefront / workgroup connectivity search
    // hands its candidate to the next larger search: the capacities, unless
    // GPK_FLIP_SEARCH_CAPS="wave,workgroup" lowers them (tests use it to run
    // the larger searches on small grids)
    int wave_cap = WB_HMAX, wg_cap = BFS_HMAX;
    void read_search_caps() {
        const char* e = std::getenv("GPK_FLIP_SEARCH_CAPS");
        int a = 0, b = 0;
        if (e && std::sscanf(e, "%d,%d", &a, &b) == 2) {
            wave_cap = std::max(1, std::min(a, WB_HMAX));
            wg_cap = std::max(1, std::min(b, BFS_HMAX));
        }
    }
    // fallback connectivity searches take tag = ++tag_ctr (at most two per
    // candidate) and mark cells with (tag << 3) | colour; tag_bound bounds the
    // counter from the host so the marks are cleared before it reaches 2^27
    long long tag_bound = 0;
    void reserve_tags(long long n) {
        if (tag_bound + n >= (1LL << 27)) {
            fill<int32_t>(mark, N, 0);
            set_int(tagctr.as<int>(), 0);
            tag_bound = 0;
        }
        tag_bound += n;
    }
    DevBuf count, unres, ulist, utgt, utgg, uflag;
    DevBuf wt0, wt1, wt2, tab, projw, remapw;
    DevBuf E_el, eps_a, eps_b, eps_c;   // elastic energy density and eigenstrain work arrays
    // remap-score cache (tiled scoring path): per-cell gain / best label,
    // validity and the dirty marks left by label writes
    DevBuf rs_gain, rs_best, rs_valid, rs_dirty;
    bool rs_reset = true;
    bool el_nonzero = false;
    // broad-kernel 3-D geometry (stencil radius > 18): Gaussian taps, label
    // flags / list of the interface labels, per-label boxes and pass buffers
    bool broad = false;
    DevBuf kw, lflag, llist, blidx, bbb, bboxes, bsx, bsy, bsz;
    std::vector<int32_t> hlabels;
    std::vector<BroadBox> hboxes;
    DevBuf scal, dsum, tagctr, minmax, noise_amp;
    Scalars* hs = nullptr;      // pinned host mirror of the scalar block
    int geo_band = 0;           // size of the geometry band (bandA)
    bool fields_valid = false;

    // profiling: one event pair per geometry launch, read after the run so
    // the timed steps are not synchronised by the measurement
    bool profiling = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> evpool;
    size_t evused = 0;
    double geo_ms = 0.0;
    int64_t geo_launches = 0;
    double geo_flops = 0.0;     // algorithmic flops of the measured launches
    double geo_band_sum = 0.0;

    Engine(const gpk_grid_desc& gd, const gpk_model_desc& m)
        : Nz(gd.Nz), Ny(gd.Ny), Nx(gd.Nx), N(gd.Nz * gd.Ny * gd.Nx), dim3(gd.Nz > 1 ? 1 : 0),
          dx_f(gd.dx), D_f(gd.D), sigma_f(gd.gaussian_sigma), model(m), device(gd.device) {
        HIPCHK(hipSetDevice(device));
        HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        dims = Dims{Nz, Ny, Nx, N};
        W = (sizeof(R) == 4) ? gauss_float(sigma_f) : gauss_double(sigma_f);
        vel_sigma = std::max(Nz > 1 ? std::clamp(0.375f * sigma_f, 0.85f, 3.00f) : 0.85f, 0.5f);
        remap_sigma = std::max(Nz > 1 ? std::clamp(0.125f * sigma_f, 0.70f, 1.00f) : 0.70f, 0.5f);
        proj_r = std::max(1, static_cast<int>(std::ceil(3.f * vel_sigma)));
        remap_r = std::max(1, static_cast<int>(std::ceil(3.f * remap_sigma)));
        broad = dim3 && W.rg > 18;   // src/gpk_grid.cpp:44
        scratch_init(device);
        allocate();
        HIPCHK(hipEventCreate(&ev0));
        HIPCHK(hipEventCreate(&ev1));
    }

    ~Engine() override {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto& e : evpool) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
        if (hs) (void)hipHostFree(hs);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }

    int64_t steps() const override { return step_id; }

    size_t device_bytes() const override {
        size_t t = 0;
        const DevBuf* all[] = {&phi, &mask, &phi_comp, &slot, &pid, &mark, &alpha, &c, &T,
                               &g, &rho, &g_comp, &work1, &work2, &work3, &work4, &bandA, &bandB,
                               &nx, &ny, &nz, &kappa, &vraw, &vn, &cflag, &anew_s, &sel, &Ccell,
                               &Cold, &Cnew, &Cg, &Canew, &state, &pos_old, &pos_new, &qflags,
                               &queue0, &queue1, &fbq, &ovq, &Pk, &Ps, &gain_s, &pcell, &kgain,
                               &idx, &order, &kgrp, &kgrp2, &order2, &seg_start, &seg_end, &pflag2,
                               &pidx, &Pa, &Pb, &Plo, &Phi, &pstate, &Pg, &pdeps, &cmp_cnt, &pkey, &pvfull, &count, &unres, &ulist,
                               &utgt, &utgg, &uflag, &wt0, &wt1, &wt2, &tab, &projw, &remapw,
                               &scal, &dsum, &tagctr, &minmax, &noise_amp, &rs_keys, &rs_vals, &rs_cnt, &rs_total, &kw, &lflag, &E_el, &eps_a,
                               &rs_gain, &rs_best, &rs_valid, &rs_dirty,
                               &eps_b, &eps_c,
                               &llist, &blidx, &bbb, &bboxes, &bsx, &bsy, &bsz};
        for (auto* b : all) t += b->bytes;
        return t;
    }

    template <class T>
    void fill(DevBuf& b, long long n, T v) {
        if (sizeof(T) == 4 && std::is_integral<T>::value)
            k_fill_i32<<<gsblk(n, 256), 256, 0, stream>>>(int(n), b.as<int32_t>(), int32_t(v));
        else if (sizeof(T) == 1)
            k_fill_u8<<<gsblk(n, 256), 256, 0, stream>>>(int(n), b.as<uint8_t>(), uint8_t(v));
        else
            k_fill_r<T><<<gsblk(n, 256), 256, 0, stream>>>(int(n), b.as<T>(), v);
    }

    void allocate() {
        const size_t n = size_t(N), rs = sizeof(R);
        phi.alloc(n * 4); mask.alloc(n); phi_comp.alloc(n * 4); slot.alloc(n * 4);
        pid.alloc(n * 4); mark.alloc(n * 4);
        alpha.alloc(n * rs); c.alloc(n * rs); T.alloc(n * rs); g.alloc(n * rs); rho.alloc(n * rs);
        E_el.alloc(n * rs);
        rs_gain.alloc(n * rs); rs_best.alloc(n * 4); rs_valid.alloc(n); rs_dirty.alloc(n);
        g_comp.alloc(n * rs); work1.alloc(n * rs); work2.alloc(n * rs); work3.alloc(n * rs);
        work4.alloc(n * rs);
        bandA.alloc(n * 4); bandB.alloc(n * 4);
        nx.alloc(n * rs); ny.alloc(n * rs); nz.alloc(n * rs); kappa.alloc(n * rs);
        vraw.alloc(n * rs); vn.alloc(n * rs);
        cflag.alloc(n); anew_s.alloc(n * rs); sel.alloc(n * 4);
        Ccell.alloc(n * 4); Cold.alloc(n * 4); Cnew.alloc(n * 4); Cg.alloc(n * rs);
        Canew.alloc(n * rs); state.alloc(n * 4); pos_old.alloc(n * 4); pos_new.alloc(n * 4);
        qflags.alloc(n * 4); queue0.alloc(n * 4); queue1.alloc(n * 4); fbq.alloc(n * 4); ovq.alloc(n * 4);
        Pk.alloc(2 * n * 8); Ps.alloc(2 * n * 8);
        gain_s.alloc(n * rs); pcell.alloc(n * 4); kgain.alloc(n * 8);
        idx.alloc(n * 4); order.alloc(n * 4); kgrp.alloc(n * 8); kgrp2.alloc(n * 8);
        order2.alloc(n * 4); seg_start.alloc(n * 4); seg_end.alloc(n * 4); pflag2.alloc(n * 4);
        pidx.alloc(n * 4);
        Pa.alloc(n * 4); Pb.alloc(n * 4); Plo.alloc(n * 4); Phi.alloc(n * 4); pstate.alloc(n * 4);
        Pg.alloc(2 * n * rs);
        ulist.alloc(n * 4); utgt.alloc(n * 4); utgg.alloc(n * rs); uflag.alloc(n);
        scal.alloc(sizeof(Scalars)); dsum.alloc(((n + 1023) / 1024 + (n + 1048575) / 1048576 + 16) * 8);
        tagctr.alloc(16); minmax.alloc(16); noise_amp.alloc(16);
        HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&hs), sizeof(Scalars), hipHostMallocDefault));
        read_search_caps();
        std::memset(hs, 0, sizeof(Scalars));
        // weights
        std::vector<R> w0(W.w0.begin(), W.w0.end()), w1(W.w1.begin(), W.w1.end()),
            w2(W.w2.begin(), W.w2.end());
        const int Wd = 2 * W.rg + 1;
        wt0.alloc(Wd * rs); wt1.alloc(Wd * rs); wt2.alloc(Wd * rs);
        HIPCHK(hipMemcpy(wt0.p, w0.data(), Wd * rs, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(wt1.p, w1.data(), Wd * rs, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(wt2.p, w2.data(), Wd * rs, hipMemcpyHostToDevice));
        if (broad) {
            // plain Gaussian taps: w0 of the geometry stencil is the Gaussian
            // padded by two zeros on each side
            const int nk = 2 * W.radius + 1;
            kw.alloc(nk * rs);
            HIPCHK(hipMemcpy(kw.p, w0.data() + (W.rg - W.radius), nk * rs, hipMemcpyHostToDevice));
        }
        // direct-stencil product table, (gx*gy)*gz in the reference's order
        if (broad) {
            tab.alloc(16);
        } else if (kDirect) {
            const std::vector<R> t = direct_table<R>(w0, w1, w2, dim3 != 0);
            tab.alloc(t.size() * rs);
            HIPCHK(hipMemcpy(tab.p, t.data(), t.size() * rs, hipMemcpyHostToDevice));
        } else {
            if (Wd > 64)
                throw GpkError(GPK_E_INVALID, "fp32/fp64 geometry needs a stencil of at most 64 cells "
                                              "(radius <= 31); use strict32 for wider 2-D stencils");
            std::vector<R> t;
            lutd = make_row_lut<R>(w0, w1, w2, t);
            tab.alloc(t.size() * rs);
            HIPCHK(hipMemcpy(tab.p, t.data(), t.size() * rs, hipMemcpyHostToDevice));
        }
        const bool single = sizeof(R) == 4;
        const auto pw = radial_table(vel_sigma, proj_r, single);
        const auto rw = radial_table(remap_sigma, remap_r, single);
        projw.alloc(pw.size() * 8); remapw.alloc(rw.size() * 8);
        HIPCHK(hipMemcpy(projw.p, pw.data(), pw.size() * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(remapw.p, rw.data(), rw.size() * 8, hipMemcpyHostToDevice));
        // radix sort ping-pong buffers and per-(digit, block) counts, sized for
        // the largest sort (the 2m flip keys, m <= N)
        rs_keys.alloc(2 * n * 8);
        rs_vals.alloc(n * 4);
        rs_cnt.alloc(size_t(RS_BINS) * size_t((2 * n + RS_CHUNK - 1) / RS_CHUNK) * 4);
        rs_total.alloc(16);
        // block counts of the ordered compaction, sized for the whole grid
        cmp_cnt.alloc(size_t((n + CMP_CHUNK - 1) / CMP_CHUNK) * 4);
        // full-grid pair keys of the projection when it cannot use the tile
        if (proj_r != 3) { pkey.alloc(n * 4); pvfull.alloc(n * rs); }
        // initial values (src/gpk_grid.cpp:14-21)
        fill<int32_t>(phi, N, 0);
        fill<uint8_t>(mask, N, 0);
        fill<int32_t>(phi_comp, N, 0);
        fill<int32_t>(pid, N, -1);
        fill<int32_t>(mark, N, 0);
        fill<R>(alpha, N, R(1));
        fill<R>(c, N, R(0));
        fill<R>(T, N, R(800));
        fill<R>(g, N, R(0));
        fill<R>(rho, N, R(0));
        fill<R>(E_el, N, R(0));
        fill<R>(g_comp, N, R(0));
        HIPCHK(hipMemsetAsync(scal.p, 0, sizeof(Scalars), stream));
        HIPCHK(hipMemsetAsync(tagctr.p, 0, 16, stream));
        tag_bound = 0;
        nlab = 1;
        count.alloc(size_t(nlab + 1) * 4);
        unres.alloc(size_t(nlab + 1));
        rebuild_counts();
        HIPCHK(hipStreamSynchronize(stream));
    }

    // ---- scalar helpers ----
    Scalars* ds() { return scal.as<Scalars>(); }
    void pull_scalars() {
        HIPCHK(hipMemcpyAsync(hs, scal.p, sizeof(Scalars), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
    }
    void set_int(int* dev, int v) {
        k_fill_i32<<<1, 1, 0, stream>>>(1, dev, v);
    }
    // several device ints set by one launch
    void set_ints(std::initializer_list<std::pair<int*, int>> l) {
        IntSets a{};
        for (const auto& e : l) {
            if (a.n == IntSets::CAP) throw GpkError(GPK_E_INVALID, "set_ints: too many values");
            a.p[a.n] = e.first;
            a.v[a.n] = e.second;
            ++a.n;
        }
        k_set_ints<<<1, 64, 0, stream>>>(a);
    }

    void rebuild_counts() {
        fill<int32_t>(count, nlab, 0);
        k_label_hist<<<nblk((N + 63) / 64, 256), 256, 0, stream>>>(N, phi.as<int32_t>(), count.as<int>());
    }

    // range of the labels at src (N int32 on the device); a label outside
    // [0, 2^26) is rejected before anything sized by the label range is touched
    void check_labels(const int32_t* src) {
        int init[2] = {INT_MAX, INT_MIN};
        HIPCHK(hipMemcpyAsync(minmax.p, init, 8, hipMemcpyHostToDevice, stream));
        k_minmax_i32<<<gsblk(N, 256), 256, 0, stream>>>(N, src, minmax.as<int>());
        int mm[2];
        HIPCHK(hipMemcpyAsync(mm, minmax.p, 8, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        if (mm[0] < 0 || mm[1] >= GPK_MAX_LABEL)
            throw GpkError(GPK_E_LABEL_RANGE, "labels must lie in [0, 2^26); found [" +
                                                  std::to_string(mm[0]) + ", " + std::to_string(mm[1]) + "]");
        const int need = mm[1] + 1;
        if (need > nlab || count.bytes < size_t(need) * 4) {
            nlab = need;
            count.alloc(size_t(nlab) * 4);
            unres.alloc(size_t(nlab));
        }
        nlab = std::max(nlab, need);
    }

    // ---- field access ----
    DevBuf* field_buf(int f, int& kind) {
        kind = 2;   // real
        switch (f) {
        case GPK_FIELD_PHI: kind = 0; return &phi;
        case GPK_FIELD_PHI_COMP: kind = 0; return &phi_comp;
        case GPK_FIELD_MASK: kind = 1; return &mask;
        case GPK_FIELD_ALPHA: return &alpha;
        case GPK_FIELD_C: return &c;
        case GPK_FIELD_T: return &T;
        case GPK_FIELD_G: return &g;
        case GPK_FIELD_RHO: return &rho;
        case GPK_FIELD_E_EL: return &E_el;
        case GPK_FIELD_G_COMP: return &g_comp;
        case GPK_FIELD_NX: kind = 3; return &nx;
        case GPK_FIELD_NY: kind = 3; return &ny;
        case GPK_FIELD_NZ: kind = 3; return &nz;
        case GPK_FIELD_KAPPA: kind = 3; return &kappa;
        case GPK_FIELD_VN: kind = 3; return &vn;
        default: return nullptr;
        }
    }
    size_t elem(int kind) const { return kind == 0 ? 4 : kind == 1 ? 1 : sizeof(R); }

    void upload(int f, const void* src, size_t bytes, bool device) override {
        int kind;
        DevBuf* b = field_buf(f, kind);
        if (!b || kind == 3) throw GpkError(GPK_E_INVALID, "field is not writable");
        if (!src || bytes != size_t(N) * elem(kind)) throw GpkError(GPK_E_INVALID, "upload size mismatch");
        const hipMemcpyKind mk = device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
        if (f == GPK_FIELD_PHI) {
            // labels are staged and range-checked first: on GPK_E_LABEL_RANGE
            // the grid keeps its previous labels, counts and label tables
            HIPCHK(hipMemcpyAsync(work4.p, src, bytes, mk, stream));
            check_labels(work4.as<int32_t>());
            HIPCHK(hipMemcpyAsync(phi.p, work4.p, bytes, hipMemcpyDeviceToDevice, stream));
            rebuild_counts();
            rs_reset = true;   // labels written from outside: no cached remap score holds
        } else {
            HIPCHK(hipMemcpyAsync(b->p, src, bytes, mk, stream));
        }
        mask_fresh = false;
        HIPCHK(hipStreamSynchronize(stream));
    }

    void download(int f, void* dst, size_t bytes, bool device) override {
        int kind;
        DevBuf* b = field_buf(f, kind);
        if (!b) throw GpkError(GPK_E_INVALID, "unknown field");
        if (!dst || bytes != size_t(N) * elem(kind)) throw GpkError(GPK_E_INVALID, "download size mismatch");
        const hipMemcpyKind mk = device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
        if (kind == 3) {
            // band-compact field: scatter into zeros over the full grid
            fill<R>(work4, N, R(0));
            if (geo_band > 0)
                k_scatter_band<R><<<nblk(geo_band, 256), 256, 0, stream>>>(geo_band, bandA.as<int>(),
                                                                           b->as<R>(), work4.as<R>());
            HIPCHK(hipMemcpyAsync(dst, work4.p, bytes, mk, stream));
        } else {
            HIPCHK(hipMemcpyAsync(dst, b->p, bytes, mk, stream));
        }
        HIPCHK(hipStreamSynchronize(stream));
    }

    void set_model(const gpk_model_desc& m) override {
        if (m.kind != model.kind) throw GpkError(GPK_E_INVALID, "model kind cannot change");
        model = m;
    }
    double time() override {
        pull_scalars();
        return hs->time;
    }
    void set_time(double t) override {
        HIPCHK(hipMemcpyAsync(&ds()->time, &t, 8, hipMemcpyHostToDevice, stream));
        HIPCHK(hipStreamSynchronize(stream));
    }
    void sync() override { HIPCHK(hipStreamSynchronize(stream)); }
    void set_profiling(int on) override {
        profiling = on != 0;
        evused = 0;
        geo_ms = 0.0;
        geo_launches = 0;
        geo_flops = 0.0;
        geo_band_sum = 0.0;
    }
    void record_geo_begin() {
        if (evused == evpool.size()) {
            hipEvent_t a, b;
            HIPCHK(hipEventCreate(&a));
            HIPCHK(hipEventCreate(&b));
            evpool.emplace_back(a, b);
        }
        HIPCHK(hipEventRecord(evpool[evused].first, stream));
    }
    void record_geo_end(int nb) {
        HIPCHK(hipEventRecord(evpool[evused].second, stream));
        ++evused;
        // reference-formulation flops of the geometry stencil per band cell:
        // 2 flops (multiply-add) per window cell and accumulated component
        const double W = 2.0 * W_rg() + 1.0;
        const double per = dim3 ? 2.0 * W * W * W * 9.0 : 2.0 * W * W * 5.0;
        geo_flops += per * nb;
        geo_band_sum += nb;
    }
    double W_rg() const { return double(W.rg); }
    // which = 0: geometry kernel.  avg_ms, launches, and (through bytes) the
    // algorithmic flops per launch of the geometry stencil
    void timing(int which, double* ms, int64_t* n, double* bytes) override {
        (void)which;
        HIPCHK(hipStreamSynchronize(stream));
        double tot = 0.0;
        for (size_t e = 0; e < evused; ++e) {
            float v = 0.f;
            HIPCHK(hipEventElapsedTime(&v, evpool[e].first, evpool[e].second));
            tot += v;
        }
        geo_ms = tot;
        geo_launches = int64_t(evused);
        if (ms) *ms = geo_launches ? geo_ms / double(geo_launches) : 0.0;
        if (n) *n = geo_launches;
        if (bytes) *bytes = geo_launches ? geo_flops / double(geo_launches) : 0.0;
    }

    // ---- geometry launch: LDS-tiled separable kernel when the labels fit
    // 16 bits and the radius has an instantiation, else the generic kernel
    template <int DIM, int RG>
    void launch_tiled(const GeomParams& gp) {
        using S = TiledLds<R, DIM, RG>;
        const int ntx = (Nx + S::TX - 1) / S::TX, nty = (Ny + S::TY - 1) / S::TY,
                  ntz = (Nz + S::TZ - 1) / S::TZ;
        const int ntiles = ntx * nty * ntz;
        const int grid = 8 * ((ntiles + 7) / 8);
        const bool plain_iso = model.kind == GPK_MODEL_ISOTROPIC && model.eps4 == 0.f &&
                               model.eps6 == 0.f &&
                               !(model.flags & (GPK_FLAG_ELASTIC | GPK_FLAG_DISLOCATION));
        auto kern = plain_iso ? k_geometry_tiled<R, DIM, RG, 1> : k_geometry_tiled<R, DIM, RG, 0>;
        kern<<<grid, 256, 0, stream>>>(
            dims, gp, model, ntx, nty, ntz, phi.as<int32_t>(), mask.as<uint8_t>(), slot.as<int>(),
            wt0.as<R>(), wt1.as<R>(), wt2.as<R>(), tab.as<R>(), g.as<R>(), c.as<R>(), T.as<R>(), rho.as<R>(), E_el.as<R>(),
            nx.as<R>(), ny.as<R>(), nz.as<R>(), kappa.as<R>(), vraw.as<R>(), phi_comp.as<int32_t>(),
            g_comp.as<R>());
    }
    template <int DIM>
    bool try_tiled(const GeomParams& gp) {
        switch (W.rg) {
        case 6: launch_tiled<DIM, 6>(gp); return true;
        case 8: launch_tiled<DIM, 8>(gp); return true;
        case 9: launch_tiled<DIM, 9>(gp); return true;
        case 12: launch_tiled<DIM, 12>(gp); return true;
        default: return false;
        }
    }
    // src/gpk_geometry.cpp:185-390: per interface label, ascending
    void broad_geometry(int nb) {
        const int L = std::max(nlab, 1);
        if (lflag.bytes < size_t(L)) { lflag.alloc(size_t(L)); llist.alloc(size_t(L) * 4); }
        k_fill_u8<<<gsblk(L, 256), 256, 0, stream>>>(L, lflag.as<uint8_t>(), 0);
        k_broad_flag_labels<<<nblk(nb, 256), 256, 0, stream>>>(nb, bandA.as<int>(), phi.as<int32_t>(),
                                                               lflag.as<uint8_t>());
        compact_async(lflag.as<uint8_t>(), L, llist.as<int>());
        pull_scalars();
        const int ng = hs->nsel;
        hlabels.resize(ng);
        if (ng > 0) {
            HIPCHK(hipMemcpyAsync(hlabels.data(), llist.p, size_t(ng) * 4, hipMemcpyDeviceToHost, stream));
            HIPCHK(hipStreamSynchronize(stream));
        }
        const int kr = W.radius;
        if (ng > 0) {
            if (blidx.bytes < size_t(L) * 4) blidx.alloc(size_t(L) * 4);
            if (bbb.bytes < size_t(ng) * 24) bbb.alloc(size_t(ng) * 24);
            k_broad_label_index<<<nblk(ng, 256), 256, 0, stream>>>(ng, llist.as<int>(), blidx.as<int>());
            k_broad_bbox_init<<<nblk(6 * ng, 256), 256, 0, stream>>>(ng, bbb.as<int>());
            k_broad_bbox<<<nblk(nb, 256), 256, 0, stream>>>(dims, nb, bandA.as<int>(), phi.as<int32_t>(),
                                                            blidx.as<int>(), bbb.as<int>());
            std::vector<int> hb(size_t(ng) * 6);
            HIPCHK(hipMemcpyAsync(hb.data(), bbb.p, hb.size() * 4, hipMemcpyDeviceToHost, stream));
            HIPCHK(hipStreamSynchronize(stream));
            // boxes: interface bounding box + 2 (difference stencil), widened by
            // kr for the y and x passes; a side reaching its period is whole
            struct Iv { int s, n; };
            auto box = [](int lo, int hi, int n) { return hi - lo + 1 >= n ? Iv{0, n} : Iv{lo, hi - lo + 1}; };
            auto widen = [](Iv v, int k, int n) { return v.n + 2 * k >= n ? Iv{0, n} : Iv{v.s - k, v.n + 2 * k}; };
            hboxes.resize(ng);
            std::vector<long long> sz(size_t(ng) * 3);
            for (int t = 0; t < ng; ++t) {
                const int* b = &hb[size_t(t) * 6];
                const Iv Bz = box(b[0] - 2, b[1] + 2, Nz), By = box(b[2] - 2, b[3] + 2, Ny),
                         Bx = box(b[4] - 2, b[5] + 2, Nx);
                const Iv Wz = widen(Bz, kr, Nz), Wy = widen(By, kr, Ny);
                BroadBox& X = hboxes[t];
                X.label = hlabels[t];
                X.zs[0] = Bz.s; X.zn[0] = Bz.n; X.zs[1] = Wz.s; X.zn[1] = Wz.n;
                X.ys[0] = By.s; X.yn[0] = By.n; X.ys[1] = Wy.s; X.yn[1] = Wy.n;
                X.xs = Bx.s; X.xn = Bx.n;
                sz[size_t(t) * 3 + 0] = (long long)Bz.n * By.n * Bx.n;
                sz[size_t(t) * 3 + 1] = (long long)Wz.n * By.n * Bx.n;
                sz[size_t(t) * 3 + 2] = (long long)Wz.n * Wy.n * Bx.n;
            }
            // chunks of consecutive labels whose x-pass boxes fit the scratch cap;
            // offsets are relative to the chunk
            const long long cap = std::max<long long>(N, 1LL << 24);
            std::vector<int> cut{0};
            std::vector<long long> tot;   // per chunk: z, y, x totals
            long long o0 = 0, o1 = 0, o2 = 0;
            for (int t = 0; t < ng; ++t) {
                if (t > cut.back() && o2 + sz[size_t(t) * 3 + 2] > cap) {
                    tot.insert(tot.end(), {o0, o1, o2});
                    cut.push_back(t);
                    o0 = o1 = o2 = 0;
                }
                BroadBox& X = hboxes[t];
                X.off[0] = o0; X.off[1] = o1; X.off[2] = o2;
                o0 += sz[size_t(t) * 3 + 0]; o1 += sz[size_t(t) * 3 + 1]; o2 += sz[size_t(t) * 3 + 2];
            }
            tot.insert(tot.end(), {o0, o1, o2});
            cut.push_back(ng);
            long long m0 = 0, m1 = 0, m2 = 0;
            for (size_t c = 0; c + 1 < cut.size(); ++c) {
                m0 = std::max(m0, tot[3 * c]); m1 = std::max(m1, tot[3 * c + 1]); m2 = std::max(m2, tot[3 * c + 2]);
            }
            if (bsz.bytes < size_t(m0) * sizeof(R)) bsz.alloc(size_t(m0) * sizeof(R));
            if (bsy.bytes < size_t(m1) * sizeof(R)) bsy.alloc(size_t(m1) * sizeof(R));
            if (bsx.bytes < size_t(m2) * sizeof(R)) bsx.alloc(size_t(m2) * sizeof(R));
            if (bboxes.bytes < size_t(ng) * sizeof(BroadBox)) bboxes.alloc(size_t(ng) * sizeof(BroadBox));
            HIPCHK(hipMemcpyAsync(bboxes.p, hboxes.data(), size_t(ng) * sizeof(BroadBox), hipMemcpyHostToDevice,
                                  stream));
            const BroadBox* dbx = bboxes.as<BroadBox>();
            auto grid_of = [](long long n) { return int(std::max(1LL, std::min<long long>((n + 255) / 256, 65536))); };
            for (size_t c = 0; c + 1 < cut.size(); ++c) {
                const int c0 = cut[c], nbox = cut[c + 1] - cut[c];
                k_broad_box_x<R><<<grid_of(tot[3 * c + 2]), 256, 0, stream>>>(
                    dims, kr, nbox, dbx + c0, tot[3 * c + 2], kw.as<R>(), phi.as<int32_t>(), bsx.as<R>());
                k_broad_box_y<R><<<grid_of(tot[3 * c + 1]), 256, 0, stream>>>(
                    dims, kr, nbox, dbx + c0, tot[3 * c + 1], kw.as<R>(), bsx.as<R>(), bsy.as<R>());
                k_broad_box_z<R><<<grid_of(tot[3 * c]), 256, 0, stream>>>(
                    dims, kr, nbox, dbx + c0, tot[3 * c], kw.as<R>(), bsy.as<R>(), bsz.as<R>());
                k_broad_box_fd<R><<<nblk(nb, 256), 256, 0, stream>>>(
                    dims, nb, bandA.as<int>(), phi.as<int32_t>(), blidx.as<int>(), c0, cut[c + 1], dbx,
                    bsz.as<R>(), dx_f, W.sigma_cells, nx.as<R>(), ny.as<R>(), nz.as<R>(), kappa.as<R>());
            }
        }
        k_broad_finish<R><<<nblk(nb, 256), 256, 0, stream>>>(
            dims, model, nb, bandA.as<int>(), phi.as<int32_t>(), g.as<R>(), c.as<R>(), T.as<R>(), rho.as<R>(), E_el.as<R>(), nx.as<R>(),
            ny.as<R>(), nz.as<R>(), kappa.as<R>(), vraw.as<R>(), phi_comp.as<int32_t>(), g_comp.as<R>());
        HIPCHK(hipGetLastError());
    }

    void launch_geometry(int nb, const GeomParams& gp) {
        if (broad) {
            broad_geometry(nb);
            return;
        }
        if (!kDirect && (dim3 ? try_tiled<3>(gp) : try_tiled<2>(gp))) return;
        if (dim3)
            k_geometry<R, 3, kDirect><<<nblk(nb, 256), 256, 0, stream>>>(
                dims, gp, model, nb, bandA.as<int>(), phi.as<int32_t>(), wt0.as<R>(), wt1.as<R>(),
                wt2.as<R>(), tab.as<R>(), lutd, g.as<R>(), c.as<R>(), T.as<R>(), rho.as<R>(), E_el.as<R>(), nx.as<R>(), ny.as<R>(),
                nz.as<R>(), kappa.as<R>(), vraw.as<R>(), phi_comp.as<int32_t>(), g_comp.as<R>());
        else
            k_geometry<R, 2, kDirect><<<nblk(nb, 256), 256, 0, stream>>>(
                dims, gp, model, nb, bandA.as<int>(), phi.as<int32_t>(), wt0.as<R>(), wt1.as<R>(),
                wt2.as<R>(), tab.as<R>(), lutd, g.as<R>(), c.as<R>(), T.as<R>(), rho.as<R>(), E_el.as<R>(), nx.as<R>(), ny.as<R>(),
                nz.as<R>(), kappa.as<R>(), vraw.as<R>(), phi_comp.as<int32_t>(), g_comp.as<R>());
    }

    template <int DIM, class F>
    void launch_small_tiled(F&& f) {
        const int TZ = SmallTile<DIM>::TZ, TY = SmallTile<DIM>::TY, TX = SmallTile<DIM>::TX;
        const int ntx = (Nx + TX - 1) / TX, nty = (Ny + TY - 1) / TY, ntz = (Nz + TZ - 1) / TZ;
        const int ntiles = ntx * nty * ntz;
        f(ntx, nty, ntz, 8 * ((ntiles + 7) / 8));
    }

    // ---- ordered compaction of a byte-flag array (count, scan, scatter);
    // leaves the count in the device scalars (nsel) without reading it back
    // (and, given slot_map, slot_map[out[a]] = a)
    void compact_async(const uint8_t* flags, long long n, int* out, int* slot_map = nullptr) {
        const int nb = int((n + CMP_CHUNK - 1) / CMP_CHUNK);
        if (cmp_cnt.bytes < size_t(nb) * 4) cmp_cnt.alloc(size_t(nb) * 4);
        k_cmp_count<<<nb, CMP_T, 0, stream>>>(n, flags, cmp_cnt.as<int>());
        k_cmp_scan<<<1, 1024, 0, stream>>>(nb, cmp_cnt.as<int>(), &ds()->nsel);
        k_cmp_scatter<<<nb, CMP_T, 0, stream>>>(n, flags, cmp_cnt.as<int>(), out, slot_map);
    }

    // ---- stable LSD radix sort of n 64-bit keys by bits [begin, end), kin
    // -> kout, int payloads vin -> vout when given (gpk_sort.hpp); kin and
    // vin are read only, the passes alternate between kout and rs_keys so
    // that the last one lands in kout
    void radix_sort(const unsigned long long* kin, unsigned long long* kout, const int* vin, int* vout, int n,
                    int begin, int end) {
        if (n <= 0) return;
        const int passes = (end - begin + 7) / 8;
        if (passes <= 0) {
            HIPCHK(hipMemcpyAsync(kout, kin, size_t(n) * 8, hipMemcpyDeviceToDevice, stream));
            if (vin) HIPCHK(hipMemcpyAsync(vout, vin, size_t(n) * 4, hipMemcpyDeviceToDevice, stream));
            return;
        }
        const int nb = (n + RS_CHUNK - 1) / RS_CHUNK;
        if (rs_cnt.bytes < size_t(RS_BINS) * nb * 4 || rs_keys.bytes < size_t(n) * 8 ||
            (vin && rs_vals.bytes < size_t(n) * 4))
            throw GpkError(GPK_E_INVALID, "radix sort larger than its buffers");
        const unsigned long long* ks = kin;
        const int* vs = vin;
        for (int p = 0; p < passes; ++p) {
            const int shift = begin + 8 * p;
            const int bits = std::min(8, end - shift);
            const bool to_out = (passes - 1 - p) % 2 == 0;
            unsigned long long* kd = to_out ? kout : rs_keys.as<unsigned long long>();
            // a keys-only sort (vin null) keeps both payload pointers null
            int* vd = vin ? (to_out ? vout : rs_vals.as<int>()) : nullptr;
            k_rs_count<<<nb, RS_T, 0, stream>>>(n, ks, shift, bits, rs_cnt.as<int>());
            k_cmp_scan<<<1, 1024, 0, stream>>>(RS_BINS * nb, rs_cnt.as<int>(), rs_total.as<int>());
            k_rs_scatter<<<nb, RS_T, 0, stream>>>(n, ks, kd, vs, vd, shift, bits, rs_cnt.as<int>());
            ks = kd;
            vs = vd;
        }
    }

    // ---- interface mask (src/gpk_geometry.cpp:42-82); mask_fresh records
    // that mask matches the current labels, so a compaction can skip it
    void update_mask() {
        if (dims.Nx % 4 == 0)
            k_mask4<<<gsblk(N / 4, 256), 256, 0, stream>>>(dims, phi.as<int32_t>(), mask.as<uint8_t>());
        else
            k_mask<<<gsblk(N, 256), 256, 0, stream>>>(dims, phi.as<int32_t>(), mask.as<uint8_t>());
        mask_fresh = true;
    }

    // ---- band compaction: mask + sorted interface list (+ slot map) ----
    int band_compact(DevBuf& list, bool with_slot) {
        if (!mask_fresh) update_mask();
        compact_async(mask.as<uint8_t>(), N, list.as<int>(), with_slot ? slot.as<int>() : nullptr);
        pull_scalars();
        return hs->nsel;
    }

    int select_flags(const uint8_t* flags, int n, int* out) {
        if (n <= 0) return 0;
        compact_async(flags, n, out);
        pull_scalars();
        return hs->nsel;
    }

    int sweep_rho() const { return (model.flags & GPK_FLAG_DISLOCATION) ? 1 : 0; }
    // remap-score cache invalidation around relabelled cells
    void mark_dirty(int n, int ncell, const int* c0, const int* c1, const int* st) {
        if (n <= 0) return;
        const long long W = 2 * remap_r + 1;
        const long long rows = (long long)n * ncell * (dim3 ? W * W : W);
        if (rows < (1LL << 31)) {
            const int grid = int(std::max(1LL, std::min<long long>((rows + 255) / 256, 65536)));
            k_mark_dirty_rows<<<grid, 256, 0, stream>>>(dims, dim3, remap_r, n, ncell, c0, c1, st,
                                                          rs_dirty.as<uint8_t>());
            return;
        }
        const long long total = rows * W;
        const int grid = int(std::max(1LL, std::min<long long>((total + 255) / 256, 65536)));
        k_mark_dirty<<<grid, 256, 0, stream>>>(dims, dim3, remap_r, n, ncell, c0, c1, st,
                                                 rs_dirty.as<uint8_t>());
    }
    // the FP32 mode decides flips near the threshold again in FP64; strict32
    // keeps the reference's float decision
    static constexpr bool guard_band() { return sizeof(R) == 4 && !kDirect; }

    // src/gpk_physics.cpp:10-63; depends on c only, so it runs before the
    // fused geometry + velocity kernels that read E_el
    void elastic() {
        const bool on = (model.flags & GPK_FLAG_ELASTIC) && model.elastic_modulus > 0.f &&
                        model.eigenstrain_coeff != 0.f;
        if (!on) {
            if (el_nonzero) {
                k_fill_r<R><<<gsblk(N, 256), 256, 0, stream>>>(N, E_el.as<R>(), R(0));
                el_nonzero = false;
            }
            return;
        }
        if (!eps_a.p) { eps_a.alloc(size_t(N) * sizeof(R)); eps_b.alloc(size_t(N) * sizeof(R)); eps_c.alloc(size_t(N) * sizeof(R)); }
        const int gb = gsblk(N, 256);
        k_elastic_eps<R><<<gb, 256, 0, stream>>>(N, c.as<R>(), model.eigenstrain_coeff, model.c_ref,
                                                 eps_a.as<R>(), eps_b.as<R>());
        const int iters = std::max(1, int(model.elastic_smooth_iters));
        for (int it = 0; it < iters; ++it) {
            k_elastic_smooth<R><<<gb, 256, 0, stream>>>(dims, dim3, eps_b.as<R>(), eps_c.as<R>());
            eps_b.swap(eps_c);
        }
        k_elastic_energy<R><<<gb, 256, 0, stream>>>(N, eps_a.as<R>(), eps_b.as<R>(), model.elastic_modulus,
                                                    E_el.as<R>());
        el_nonzero = true;
    }

    // pair projection of vraw into vn (src/gpk_physics.cpp:152-193), vmax on the side
    void project(int nb) {
        k_vmax_init<R><<<1, 1, 0, stream>>>(ds());
        // the tiled projection packs (low, high, orientation) in 32 bits: labels < 2^15
        if (nb > 0 && nlab <= 32768 && proj_r == 3) {
            if (dim3) launch_small_tiled<3>([&](int ntx, int nty, int ntz, int grid) {
                k_project_tiled<R, 3, 3><<<grid, PROJ_THREADS, 0, stream>>>(
                    dims, ntx, nty, ntz, mask.as<uint8_t>(), phi.as<int32_t>(), phi_comp.as<int32_t>(),
                    slot.as<int>(), projw.as<double>(), vraw.as<R>(), vn.as<R>(),
                    reinterpret_cast<typename UBits<R>::T*>(&ds()->vmax_bits));
            });
            else launch_small_tiled<2>([&](int ntx, int nty, int ntz, int grid) {
                k_project_tiled<R, 2, 3><<<grid, PROJ_THREADS, 0, stream>>>(
                    dims, ntx, nty, ntz, mask.as<uint8_t>(), phi.as<int32_t>(), phi_comp.as<int32_t>(),
                    slot.as<int>(), projw.as<double>(), vraw.as<R>(), vn.as<R>(),
                    reinterpret_cast<typename UBits<R>::T*>(&ds()->vmax_bits));
            });
        } else if (nb > 0 && nlab <= 32768) {
            if (pkey.bytes < size_t(N) * 4) { pkey.alloc(size_t(N) * 4); pvfull.alloc(size_t(N) * sizeof(R)); }
            k_proj_keys<R><<<gsblk(N, 256), 256, 0, stream>>>(N, mask.as<uint8_t>(), phi.as<int32_t>(),
                                                              phi_comp.as<int32_t>(), slot.as<int>(), vraw.as<R>(),
                                                              pkey.as<uint32_t>(), pvfull.as<R>());
            if (dim3)
                k_project_keyed<R, 3><<<nblk(nb, 256), 256, 0, stream>>>(
                    dims, proj_r, nb, bandA.as<int>(), pkey.as<uint32_t>(), pvfull.as<R>(), vraw.as<R>(),
                    projw.as<double>(), vn.as<R>(), reinterpret_cast<typename UBits<R>::T*>(&ds()->vmax_bits));
            else
                k_project_keyed<R, 2><<<nblk(nb, 256), 256, 0, stream>>>(
                    dims, proj_r, nb, bandA.as<int>(), pkey.as<uint32_t>(), pvfull.as<R>(), vraw.as<R>(),
                    projw.as<double>(), vn.as<R>(), reinterpret_cast<typename UBits<R>::T*>(&ds()->vmax_bits));
        } else if (nb > 0) {
            if (dim3)
                k_project<R, 3><<<nblk(nb, 256), 256, 0, stream>>>(
                    dims, proj_r, nb, bandA.as<int>(), slot.as<int>(), mask.as<uint8_t>(),
                    phi.as<int32_t>(), phi_comp.as<int32_t>(), projw.as<double>(), vraw.as<R>(),
                    vn.as<R>(), reinterpret_cast<typename UBits<R>::T*>(&ds()->vmax_bits));
            else
                k_project<R, 2><<<nblk(nb, 256), 256, 0, stream>>>(
                    dims, proj_r, nb, bandA.as<int>(), slot.as<int>(), mask.as<uint8_t>(),
                    phi.as<int32_t>(), phi_comp.as<int32_t>(), projw.as<double>(), vraw.as<R>(),
                    vn.as<R>(), reinterpret_cast<typename UBits<R>::T*>(&ds()->vmax_bits));
        }
    }

    // ---- velocity chain for the adjoint: geometry + law + projection on the
    // current labels with (eps4, eps6, M0) = theta; leaves the derived band
    // fields (mask, normals, curvature, phi_comp, vraw, vn) for these labels ----
    int velocity_chain(const double theta[3]) {
        if ((model.kind != GPK_MODEL_ISOTROPIC && model.kind != GPK_MODEL_ANISO_GROWTH) ||
            (model.flags & (GPK_FLAG_ELASTIC | GPK_FLAG_DISLOCATION)))
            throw GpkError(GPK_E_UNSUPPORTED_MODEL,
                           "velocity adjoint: isotropic or anisotropic-growth models without "
                           "elastic or stored-energy coupling");
        const gpk_model_desc saved = model;
        model.eps4 = float(theta[0]);
        model.eps6 = float(theta[1]);
        model.M0 = float(theta[2]);
        int nb = 0;
        try {
            nb = band_compact(bandA, true);
            geo_band = nb;
            if (nb > 0) launch_geometry(nb, GeomParams{W.rg, 2 * W.rg + 1, dx_f, W.sigma_cells});
            project(nb);
        } catch (...) {
            model = saved;
            throw;
        }
        model = saved;
        return nb;
    }

    void velocity_field(const double theta[3], void* vout) override {
        const int nb = velocity_chain(theta);
        k_fill_r<R><<<gsblk(N, 256), 256, 0, stream>>>(N, static_cast<R*>(vout), R(0));
        if (nb > 0)
            k_scatter_band<R><<<nblk(nb, 256), 256, 0, stream>>>(nb, bandA.as<int>(), vn.as<R>(),
                                                                 static_cast<R*>(vout));
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(stream));
    }

    DevBuf adj_h, adj_part, adj_out;
    // transposed projection and parameter terms of the adjoint over the nb
    // band cells velocity_chain left: hbuf (nb doubles) and part (3 doubles
    // per 256 band cells) are written before they are read, outp (3 doubles)
    // receives the fixed-order sum
    void adjoint_terms(const double theta[3], const void* gvel, int nb, double* hbuf, double* part,
                       double* outp, double grad[3]) {
        const int nblocks = nblk(nb, 256);
        gpk_model_desc m = model;
        m.eps4 = float(theta[0]);
        m.eps6 = float(theta[1]);
        m.M0 = float(theta[2]);
        const R* gv = static_cast<const R*>(gvel);
        if (dim3) {
            k_adj_h<R, 3><<<nblocks, 256, 0, stream>>>(dims, proj_r, nb, bandA.as<int>(), mask.as<uint8_t>(),
                                                        phi.as<int32_t>(), phi_comp.as<int32_t>(),
                                                        projw.as<double>(), gv, hbuf);
            k_adj_terms<R, 3><<<nblocks, 256, 0, stream>>>(
                dims, proj_r, nb, bandA.as<int>(), slot.as<int>(), mask.as<uint8_t>(), phi.as<int32_t>(),
                phi_comp.as<int32_t>(), projw.as<double>(), gv, hbuf, nx.as<R>(), ny.as<R>(),
                nz.as<R>(), kappa.as<R>(), g.as<R>(), m, part);
        } else {
            k_adj_h<R, 2><<<nblocks, 256, 0, stream>>>(dims, proj_r, nb, bandA.as<int>(), mask.as<uint8_t>(),
                                                        phi.as<int32_t>(), phi_comp.as<int32_t>(),
                                                        projw.as<double>(), gv, hbuf);
            k_adj_terms<R, 2><<<nblocks, 256, 0, stream>>>(
                dims, proj_r, nb, bandA.as<int>(), slot.as<int>(), mask.as<uint8_t>(), phi.as<int32_t>(),
                phi_comp.as<int32_t>(), projw.as<double>(), gv, hbuf, nx.as<R>(), ny.as<R>(),
                nz.as<R>(), kappa.as<R>(), g.as<R>(), m, part);
        }
        k_adj_final<<<1, 256, 0, stream>>>(nblocks, part, outp);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(grad, outp, 24, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
    }

    void velocity_adjoint(const double theta[3], const void* gvel, double grad[3]) override {
        const double th[3] = {theta[0], theta[1], theta[2]};
        const int nb = velocity_chain(th);
        grad[0] = grad[1] = grad[2] = 0.0;
        if (nb == 0) {
            HIPCHK(hipStreamSynchronize(stream));
            return;
        }
        const int nblocks = nblk(nb, 256);
        if (adj_h.bytes < size_t(nb) * 8) adj_h.alloc(size_t(nb) * 8);
        if (adj_part.bytes < size_t(nblocks) * 24) adj_part.alloc(size_t(nblocks) * 24);
        if (!adj_out.p) adj_out.alloc(24);
        adjoint_terms(th, gvel, nb, adj_h.as<double>(), adj_part.as<double>(), adj_out.as<double>(), grad);
    }

    // The same adjoint with its scratch in the library's static device array
    // (layout: result at 0, partials from 256, terms after the partials,
    // each 256-byte aligned).  Nothing is allocated: the handle's buffers are
    // sized at creation, and the per-label separable geometry of wide 3-D
    // stencils (which sizes its buffers per call) is refused.
    void velocity_adjoint_noalloc(const double theta[3], const void* gvel, double grad[3]) override {
        if (broad)
            throw GpkError(GPK_E_UNSUPPORTED_MODEL,
                           "allocation-free adjoint: 3-D geometry radius above 18 is not supported");
        unsigned char* base = static_cast<unsigned char*>(scratch_of(device));
        if (!base) throw GpkError(GPK_E_HIP, "allocation-free adjoint: scratch not set up on this device");
        const double th[3] = {theta[0], theta[1], theta[2]};
        const int nb = velocity_chain(th);
        grad[0] = grad[1] = grad[2] = 0.0;
        if (nb == 0) {
            HIPCHK(hipStreamSynchronize(stream));
            return;
        }
        const size_t nblocks = size_t(nblk(nb, 256));
        const size_t off_part = 256;
        const size_t off_h = off_part + (nblocks * 24 + 255) / 256 * 256;
        const size_t need = off_h + size_t(nb) * 8;
        if (need > NOALLOC_SCRATCH_BYTES) {
            HIPCHK(hipStreamSynchronize(stream));
            throw GpkError(GPK_E_INVALID, "allocation-free adjoint: " + std::to_string(nb) +
                                              " band cells need " + std::to_string(need) +
                                              " bytes of scratch, more than its " +
                                              std::to_string(NOALLOC_SCRATCH_BYTES));
        }
        adjoint_terms(th, gvel, nb, reinterpret_cast<double*>(base + off_h),
                      reinterpret_cast<double*>(base + off_part), reinterpret_cast<double*>(base), grad);
    }

    // ---- the step (src/gpk_simulation.cpp:31-54) ----
    void step(float dt_req, float cap, gpk_step_stats* st) override {
        const int nb = band_compact(bandA, true);
        geo_band = nb;
        elastic();
        // geometry + competitor + constitutive velocity
        GeomParams gp{W.rg, 2 * W.rg + 1, dx_f, W.sigma_cells};
        if (nb > 0) {
            if (profiling) record_geo_begin();
            launch_geometry(nb, gp);
            if (profiling) record_geo_end(nb);
        }
        project(nb);
        k_dt<R><<<1, 1, 0, stream>>>(ds(), R(dt_req), R(cap), dx_f, D_f, dim3);
        const bool partition = (model.flags & GPK_FLAG_SOLUTE_REJECTION) != 0;
        if (partition)
            k_solid_fraction<R><<<gsblk(N, 256), 256, 0, stream>>>(N, phi.as<int32_t>(), alpha.as<R>(),
                                                                   work1.as<R>());
        // counters
        set_ints({{&ds()->accepted, 0}, {&ds()->deferred, 0}, {&ds()->exchanges, 0}, {&ds()->eroded, 0},
                  {&ds()->guarded, 0}});
        int m = 0;
        if (nb > 0) m = flips(nb);
        // without the topology test every candidate is accepted
        // (src/gpk_transitions.cpp:299-303): the count is known here
        // (with it, the transaction's last read of the scalars holds the count)
        int accepted = m;
        if ((model.flags & GPK_FLAG_GRAIN_NETWORK) && m > 0) accepted = hs->accepted;
        int exchanges = 0;
        if (accepted > 0) {
            exchanges = remap_sweep();
            if (dim3 && exchanges > 0) exchanges += remap_sweep();
        }
        if ((model.flags & GPK_FLAG_GRAIN_NETWORK) && (uint64_t(step_id) & 7ULL) == 0ULL) extinction();
        if (partition) {
            transport_partitioned();
        } else if (D_f > 0.f) {
            k_diffusion<R><<<gsblk(N, 256), 256, 0, stream>>>(dims, dim3, c.as<R>(), mask.as<uint8_t>(),
                                                              alpha.as<R>(), ds(), dx_f, D_f, work2.as<R>());
            c.swap(work2);
        }
        // static recovery of rho (src/gpk_transport.cpp:142-152); the decay
        // factor uses the host libm, as the reference does
        if ((model.flags & GPK_FLAG_DISLOCATION) && model.recovery_rate > 0.f) {
            pull_scalars();
            const R dtr = R(hs->dt);
            if (dtr > R(0)) {
                const R decay = sizeof(R) == 4 ? R(std::exp(-model.recovery_rate * float(dtr)))
                                               : R(std::exp(-double(model.recovery_rate) * double(dtr)));
                k_scale_r<R><<<gsblk(N, 256), 256, 0, stream>>>(N, rho.as<R>(), decay);
            }
        }
        k_time_advance<<<1, 1, 0, stream>>>(ds());
        ++step_id;
        if (st) {
            pull_scalars();
            st->band_cells = nb;
            st->candidates = m;
            st->accepted = hs->accepted;
            st->deferred = hs->deferred;
            st->exchanges = exchanges;
            st->eroded = hs->eroded;
            st->guarded = hs->guarded;
            st->dt = hs->dt;
            st->time = hs->time;
            st->step_id = step_id;
        }
    }

    // ---- alpha integration + ordered flip transaction ----
    int flips(int nb) {
        const bool noise = (model.flags & GPK_FLAG_STOCHASTIC) != 0;
        float amp = 0.f;
        if (noise) {
            k_noise_amp<R><<<1, 1, 0, stream>>>(ds(), dx_f, model.M0, model.kB, model.noise_scale,
                                                noise_amp.as<float>());
            HIPCHK(hipMemcpyAsync(&amp, noise_amp.p, 4, hipMemcpyDeviceToHost, stream));
            HIPCHK(hipStreamSynchronize(stream));
        }
        k_alpha<R><<<nblk(nb, 256), 256, 0, stream>>>(nb, bandA.as<int>(), ds(), dx_f, vn.as<R>(),
                                                      alpha.as<R>(), T.as<R>(), phi.as<int32_t>(),
                                                      phi_comp.as<int32_t>(), cflag.as<uint8_t>(),
                                                      anew_s.as<R>(), noise ? 1 : 0, amp,
                                                      uint64_t(step_id), guard_band() ? 1 : 0, ds());
        const int m = select_flags(cflag.as<uint8_t>(), nb, sel.as<int>());
        if (m == 0) return 0;
        const bool network = (model.flags & GPK_FLAG_GRAIN_NETWORK) != 0;
        k_cand_init<R><<<nblk(m, 256), 256, 0, stream>>>(
            m, sel.as<int>(), bandA.as<int>(), phi.as<int32_t>(), phi_comp.as<int32_t>(), g_comp.as<R>(),
            anew_s.as<R>(), Ccell.as<int>(), Cold.as<int>(), Cnew.as<int>(), Cg.as<R>(), Canew.as<R>(),
            state.as<int>(), network ? Pk.as<unsigned long long>() : nullptr);
        if (network) {
            // (label << 32 | k) keys are generated in k order, so a stable sort
            // on the label bits alone orders them by (label, k)
            int hibit = 32;
            while ((1LL << (hibit - 32)) < nlab) ++hibit;
            radix_sort(Pk.as<unsigned long long>(), Ps.as<unsigned long long>(), nullptr, nullptr, 2 * m, 32,
                       std::min(64, hibit + 1));
            k_cand_pos<<<nblk(2 * m, 256), 256, 0, stream>>>(2 * m, Ps.as<unsigned long long>(),
                                                              Cold.as<int>(), pos_old.as<int>(),
                                                              pos_new.as<int>());
            k_fill_i32<<<1, 64, 0, stream>>>(16, &ds()->qcount[0][0], 0);
            set_ints({{&ds()->undecided, m}, {&ds()->fbcount, 0},
                      {&ds()->ovcount[0], 0}, {&ds()->ovcount[1], 0}});
            FlipCtx fc;
            fc.d = dims; fc.dim3 = dim3; fc.m = m;
            fc.Ccell = Ccell.as<int>(); fc.Cold = Cold.as<int>(); fc.Cnew = Cnew.as<int>();
            fc.state = state.as<int>(); fc.phi = phi.as<int32_t>();
            fc.Ps = Ps.as<unsigned long long>(); fc.pos_old = pos_old.as<int>();
            fc.pos_new = pos_new.as<int>(); fc.count = count.as<int>();
            fc.queue[0] = queue0.as<int>(); fc.queue[1] = queue1.as<int>();
            fc.qflags = qflags.as<int>(); fc.s = ds(); fc.slab = (m + 7) / 8;
            fc.wave_cap = wave_cap; fc.wg_cap = wg_cap;
            reserve_tags(2LL * m);
            int round = 0;
            // one wavefront per queued search (four per workgroup, four 40 KB
            // workgroups per CU), then the workgroup search for the few that
            // outgrew a wavefront's hash; a multiple of 8 workgroups, so that
            // every band slab has its own
            const int wave_grid = 8 * std::max(1, std::min((m + 8 * WB_WAVES - 1) / (8 * WB_WAVES), 128));
            const int bfs_grid = std::max(1, std::min(m, 128));
            // the first batch is as long as the last step's transaction: rounds
            // past the end are cheap no-ops, host checks are not
            int batch = flip_batch;
            while (true) {
                for (int r = 0; r < batch; ++r, ++round) {
                    k_flip_round<<<nblk(m, 256), 256, 0, stream>>>(fc, round);
                    k_flip_bfs_wave<<<wave_grid, 256, 0, stream>>>(fc, round, ovq.as<int>(), ds()->ovcount);
                    k_flip_bfs<<<bfs_grid, 256, 0, stream>>>(fc, ovq.as<int>(), &ds()->ovcount[round & 1],
                                                             fbq.as<int>(), mark.as<int>(),
                                                             reinterpret_cast<int*>(work3.p),
                                                             reinterpret_cast<int*>(work4.p), tagctr.as<int>());
                }
                pull_scalars();
                if (hs->undecided <= 0) break;
                if (round > 4 * (m + 8)) {
                    // labels back from the candidate words before the error
                    k_flip_restore<<<nblk(m, 256), 256, 0, stream>>>(m, Ccell.as<int>(), Cold.as<int>(),
                                                                      phi.as<int32_t>());
                    HIPCHK(hipStreamSynchronize(stream));
                    throw GpkError(GPK_E_HIP, "flip transaction did not converge");
                }
                batch = 4;
            }
            flip_batch = round > flip_batch ? round : std::max(4, flip_batch - 2);
        }
        mask_fresh = false;
        k_flip_apply<R><<<nblk(m, 256), 256, 0, stream>>>(
            m, Ccell.as<int>(), Cold.as<int>(), Cnew.as<int>(), Cg.as<R>(), Canew.as<R>(), state.as<int>(),
            phi.as<int32_t>(), g.as<R>(), alpha.as<R>(), count.as<int>(), ds(),
            network ? 0 : 1, rho.as<R>(), R(model.rho_rex), sweep_rho());
        mark_dirty(m, 1, Ccell.as<int>(), nullptr, network ? state.as<int>() : nullptr);
        return m;
    }

    // ---- balanced Gaussian remap sweep (src/gpk_transitions.cpp:38-178) ----
    int remap_sweep() {
        set_int(&ds()->exchanges, 0);
        const bool tiled = nlab <= 65536 && remap_r == 3;
        const int nb = band_compact(bandB, tiled);
        if (N % 4 == 0)
            k_copy_i32x4<<<gsblk(N / 4, 256), 256, 0, stream>>>(N / 4, reinterpret_cast<const int4*>(phi.p),
                                                                 reinterpret_cast<int4*>(phi_comp.p));
        else
            k_copy_i32<<<gsblk(N, 256), 256, 0, stream>>>(N, phi.as<int32_t>(), phi_comp.as<int32_t>());
        if (rs_reset || !tiled) {
            fill<uint8_t>(rs_valid, N, 0);
            fill<uint8_t>(rs_dirty, N, 0);
            rs_reset = !tiled;
        }
        const ScoreCache<R> sc{rs_gain.as<R>(), rs_best.as<int32_t>(), rs_valid.as<uint8_t>(),
                               rs_dirty.as<uint8_t>()};
        if (nb > 0 && tiled) {
            if (dim3) launch_small_tiled<3>([&](int ntx, int nty, int ntz, int grid) {
                k_remap_score_tiled<R, 3, 3><<<grid, 256, 0, stream>>>(
                    dims, ntx, nty, ntz, mask.as<uint8_t>(), phi.as<int32_t>(), slot.as<int>(),
                    remapw.as<double>(), phi_comp.as<int32_t>(), cflag.as<uint8_t>(), gain_s.as<R>(), sc);
            });
            else launch_small_tiled<2>([&](int ntx, int nty, int ntz, int grid) {
                k_remap_score_tiled<R, 2, 3><<<grid, 256, 0, stream>>>(
                    dims, ntx, nty, ntz, mask.as<uint8_t>(), phi.as<int32_t>(), slot.as<int>(),
                    remapw.as<double>(), phi_comp.as<int32_t>(), cflag.as<uint8_t>(), gain_s.as<R>(), sc);
            });
            k_score_cache_cleanup<<<gsblk(N, 256), 256, 0, stream>>>(N, rs_valid.as<uint8_t>(),
                                                                     rs_dirty.as<uint8_t>());
        } else if (nb > 0) {
            if (dim3)
                k_remap_score<R, 3><<<nblk(nb, 256), 256, 0, stream>>>(
                    dims, remap_r, nb, bandB.as<int>(), phi.as<int32_t>(), remapw.as<double>(),
                    phi_comp.as<int32_t>(), cflag.as<uint8_t>(), gain_s.as<R>());
            else
                k_remap_score<R, 2><<<nblk(nb, 256), 256, 0, stream>>>(
                    dims, remap_r, nb, bandB.as<int>(), phi.as<int32_t>(), remapw.as<double>(),
                    phi_comp.as<int32_t>(), cflag.as<uint8_t>(), gain_s.as<R>());
        }
        const int np = select_flags(cflag.as<uint8_t>(), nb, sel.as<int>());
        int npairs = 0;
        if (np > 0) {
            k_prop_keys<R><<<nblk(np, 256), 256, 0, stream>>>(np, sel.as<int>(), bandB.as<int>(),
                                                              gain_s.as<R>(), phi.as<int32_t>(),
                                                              phi_comp.as<int32_t>(), pcell.as<int>(),
                                                              kgain.as<unsigned long long>(), idx.as<int>());
            // order (low, high, dir) ascending, gain descending, cell ascending
            // (the reference's std::map of per-pair lists sorted by gain, then
            // cell): a stable radix sort of the proposals (cell order) by the
            // group key, then a rank sort by gain within each group
            int hb = 1;
            while ((1LL << hb) < nlab) ++hb;
            k_prop_group_keys<<<nblk(np, 256), 256, 0, stream>>>(np, hb, idx.as<int>(), pcell.as<int>(),
                                                                  phi.as<int32_t>(), phi_comp.as<int32_t>(),
                                                                  kgrp.as<unsigned long long>());
            radix_sort(kgrp.as<unsigned long long>(), kgrp2.as<unsigned long long>(), idx.as<int>(),
                       order.as<int>(), np, 0, 2 * hb + 1);
            k_seg_bounds<<<nblk(np, 256), 256, 0, stream>>>(np, kgrp2.as<unsigned long long>(),
                                                             seg_start.as<int>(), seg_end.as<int>());
            k_seg_rank<<<nblk(np, 256), 256, 0, stream>>>(np, order.as<int>(), kgain.as<unsigned long long>(),
                                                           seg_start.as<int>(), seg_end.as<int>(),
                                                           order2.as<int>());
            k_pair_flags<<<nblk(np, 256), 256, 0, stream>>>(np, kgrp2.as<unsigned long long>(),
                                                             seg_start.as<int>(), seg_end.as<int>(),
                                                             pflag2.as<uint8_t>());
            npairs = select_flags(pflag2.as<uint8_t>(), np, pidx.as<int>());
            if (npairs > 0) {
                k_pair_emit<R><<<nblk(npairs, 256), 256, 0, stream>>>(
                    npairs, hb, kgrp2.as<unsigned long long>(), seg_start.as<int>(), seg_end.as<int>(),
                    pidx.as<int>(), order2.as<int>(), pcell.as<int>(), Pa.as<int>(),
                    Pb.as<int>(), Plo.as<int>(), Phi.as<int>(), pstate.as<int>(), pid.as<int>());
                set_int(&ds()->undecided, npairs);
                PairCtx pc;
                pc.d = dims; pc.dim3 = dim3; pc.np = npairs;
                pc.Pa = Pa.as<int>(); pc.Pb = Pb.as<int>(); pc.Plo = Plo.as<int>(); pc.Phi = Phi.as<int>();
                pc.pstate = pstate.as<int>(); pc.pid = pid.as<int>(); pc.phi = phi.as<int32_t>();
                pc.g = g.p; pc.Pg = Pg.p; pc.count = count.as<int>(); pc.s = ds();
                if (pdeps.bytes < size_t(npairs) * PAIR_DEPS * 4) pdeps.alloc(size_t(npairs) * PAIR_DEPS * 4);
                k_pair_deps<<<nblk(npairs, 4), 256, 0, stream>>>(pc, pdeps.as<int>());
                int rounds = 0, batch = pair_batch;
                while (true) {
                    for (int r = 0; r < batch; ++r, ++rounds)
                        k_pair_round<R><<<nblk(npairs, 4), 256, 0, stream>>>(pc, pdeps.as<int>(), rounds + 1);
                    pull_scalars();
                    if (hs->undecided <= 0) break;
                    if (rounds > 4 * (npairs + 8))
                        throw GpkError(GPK_E_HIP, "remap transaction did not converge");
                    batch = 4;
                }
                pair_batch = rounds > pair_batch ? rounds : std::max(4, pair_batch - 1);
                mask_fresh = false;
                k_pair_apply<R><<<nblk(npairs, 256), 256, 0, stream>>>(
                    npairs, Pa.as<int>(), Pb.as<int>(), Plo.as<int>(), Phi.as<int>(), pstate.as<int>(),
                    Pg.as<R>(), phi.as<int32_t>(), g.as<R>(), alpha.as<R>(), pid.as<int>(), ds(),
                    rho.as<R>(), R(model.rho_rex), sweep_rho());
                mark_dirty(npairs, 2, Pa.as<int>(), Pb.as<int>(), pstate.as<int>());
            }
        }
        if (!mask_fresh) update_mask();
        // the pair rounds' last read of the scalars holds the exchange count
        return npairs == 0 ? 0 : hs->exchanges;
    }

    // ---- unresolved-grain extinction (src/gpk_transitions.cpp:180-249) ----
    void extinction() {
        const double s = std::max(W.sigma_cells, 1.f);
        const double vol = dim3 ? (4.0 * M_PI * s * s * s / 3.0) : (M_PI * s * s);
        const int min_cells = std::max(2, static_cast<int>(std::ceil(vol)));
        set_ints({{&ds()->present, 0}, {&ds()->unresolved, 0}});
        k_label_census<<<nblk(nlab, 256), 256, 0, stream>>>(nlab, count.as<int>(), min_cells,
                                                            unres.as<uint8_t>(), ds());
        pull_scalars();
        if (hs->unresolved == 0 || hs->unresolved == hs->present) return;
        k_unres_flags<<<gsblk(N, 256), 256, 0, stream>>>(N, phi.as<int32_t>(), unres.as<uint8_t>(),
                                                         uflag.as<uint8_t>());
        const int n = select_flags(uflag.as<uint8_t>(), N, ulist.as<int>());
        mask_fresh = false;
        if (n > 0)
            k_erode<R><<<1, 1024, 0, stream>>>(dims, dim3, n, ulist.as<int>(), unres.as<uint8_t>(),
                                               phi.as<int32_t>(), g.as<R>(), alpha.as<R>(), count.as<int>(),
                                               utgt.as<int>(), utgg.as<R>(), ds(), rho.as<R>(),
                                               R(model.rho_rex), sweep_rho(), remap_r,
                                               rs_dirty.as<uint8_t>());
        update_mask();
    }

    // ---- partitioned solute transport (src/gpk_transport.cpp:48-135) ----
    void transport_partitioned() {
        const float k = model.k_partition;
        if (k < 1.f) {
            // work1 holds the solid fraction before the flip pass
            k_reject_source<R><<<gsblk(N, 256), 256, 0, stream>>>(dims, dim3, k, phi.as<int32_t>(),
                                                                  alpha.as<R>(), work1.as<R>(), c.as<R>(),
                                                                  work2.as<R>(), work3.as<R>());
            k_reject_gather<R><<<gsblk(N, 256), 256, 0, stream>>>(dims, dim3, phi.as<int32_t>(),
                                                                  alpha.as<R>(), work2.as<R>(),
                                                                  work3.as<R>(), c.as<R>());
        }
        if (D_f > 0.f) {
            k_liquid_comp<R><<<gsblk(N, 256), 256, 0, stream>>>(N, k, phi.as<int32_t>(), alpha.as<R>(),
                                                                c.as<R>(), work1.as<R>(), work2.as<R>());
            k_diffusion_part<R><<<gsblk(N, 256), 256, 0, stream>>>(dims, dim3, work1.as<R>(), work2.as<R>(),
                                                                   c.as<R>(), ds(), dx_f, D_f, work3.as<R>());
            c.swap(work3);
        }
    }

    // ---- diagnostics ----
    // GPKGrid::max_dt() (src/gpk_simulation.cpp:11-25) reads vn over the
    // latest compacted band; vn is zero outside the geometry band
    double max_dt() override {
        k_vmax_init<R><<<1, 1, 0, stream>>>(ds());
        if (geo_band > 0)
            k_vmax_masked<R><<<nblk(geo_band, 256), 256, 0, stream>>>(
                geo_band, bandA.as<int>(), mask.as<uint8_t>(), vn.as<R>(),
                reinterpret_cast<typename UBits<R>::T*>(&ds()->vmax_bits));
        const double saved_dt = 0.0;
        (void)saved_dt;
        DevBuf tmp;
        tmp.alloc(sizeof(Scalars));
        HIPCHK(hipMemcpyAsync(tmp.p, scal.p, sizeof(Scalars), hipMemcpyDeviceToDevice, stream));
        k_dt<R><<<1, 1, 0, stream>>>(tmp.as<Scalars>(), R(-1), R(Eps<R>::big), dx_f, D_f, dim3);
        Scalars out;
        HIPCHK(hipMemcpyAsync(&out, tmp.p, sizeof(Scalars), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        return out.dt;
    }

    double mean_grain_radius() override {
        // src/gpk_simulation.cpp:88-106
        const int ng = n_grains();
        if (ng == 0) return 0.0;
        const double mv = double(N) / ng;
        const double Rr = dim3 ? std::cbrt(3.0 * mv / (4.0 * M_PI)) : std::sqrt(mv / M_PI);
        return Rr * double(dx_f);
    }
    float solid_fraction() override {
        return static_cast<float>(phase_volume(1) / static_cast<double>(N));
    }
    float interface_fraction() override {
        return static_cast<float>(interface_cells()) / static_cast<float>(N);
    }

    // second and third levels of the canonical FP64 sum over the chunk sums
    // at dsum[0, nch); blocks until the value is on the host
    double sum_groups(int nch) {
        const int ng = (nch + 1023) / 1024;
        double* p1 = dsum.as<double>();
        k_sum_groups<<<ng, 256, 0, stream>>>(nch, p1, p1 + nch);
        k_sum_tail<<<1, 1, 0, stream>>>(ng, p1 + nch, p1 + nch + ng);
        double v = 0.0;
        HIPCHK(hipMemcpyAsync(&v, p1 + nch + ng, 8, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        return v;
    }

    double phase_volume(int32_t label) override {
        const int nch = (N + 1023) / 1024;
        k_phase_terms<R><<<nblk(nch, 2), 256, 0, stream>>>(N, label, phi.as<int32_t>(),
                                                             phi_comp.as<int32_t>(), mask.as<uint8_t>(),
                                                             alpha.as<R>(), dsum.as<double>());
        const double v = sum_groups(nch);
        HIPCHK(hipStreamSynchronize(stream));
        return v;
    }

    int n_grains() override {
        set_int(&ds()->present, 0);
        k_count_pos<<<gsblk(nlab, 256), 256, 0, stream>>>(nlab, count.as<int>(), &ds()->present);
        pull_scalars();
        return hs->present;
    }

    int interface_cells() override {
        set_int(&ds()->nsel, 0);
        k_count_u8<<<gsblk(N, 256), 256, 0, stream>>>(N, mask.as<uint8_t>(), &ds()->nsel);
        pull_scalars();
        return hs->nsel;
    }

    double solute_sum() override {
        const int nch = (N + 1023) / 1024;
        k_sum_chunks<R><<<nblk(nch, 4), 256, 0, stream>>>(N, c.as<R>(), dsum.as<double>());
        const double v = sum_groups(nch);
        HIPCHK(hipStreamSynchronize(stream));
        return v;
    }

    int simple_move(int cell, int32_t nl) override {
        if (cell < 0 || cell >= N) return 0;
        // a one-candidate transaction evaluated with the flip machinery
        int32_t ol = 0;
        HIPCHK(hipMemcpyAsync(&ol, phi.as<int32_t>() + cell, 4, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        if (ol == nl) return 0;
        if (nl < 0 || nl >= nlab) {
            // a label absent from the grid has no face branch: new_object == 0
            return 0;
        }
        int one[3] = {cell, ol, nl};
        HIPCHK(hipMemcpyAsync(Ccell.p, &one[0], 4, hipMemcpyHostToDevice, stream));
        HIPCHK(hipMemcpyAsync(Cold.p, &one[1], 4, hipMemcpyHostToDevice, stream));
        HIPCHK(hipMemcpyAsync(Cnew.p, &one[2], 4, hipMemcpyHostToDevice, stream));
        set_int(state.as<int>(), UNDEC);
        unsigned long long keys[2] = {((unsigned long long)(uint32_t)ol << 32),
                                      ((unsigned long long)(uint32_t)nl << 32)};
        if (ol > nl) std::swap(keys[0], keys[1]);
        HIPCHK(hipMemcpyAsync(Ps.p, keys, 16, hipMemcpyHostToDevice, stream));
        int po = ol < nl ? 0 : 1, pn = 1 - po;
        HIPCHK(hipMemcpyAsync(pos_old.p, &po, 4, hipMemcpyHostToDevice, stream));
        HIPCHK(hipMemcpyAsync(pos_new.p, &pn, 4, hipMemcpyHostToDevice, stream));
        set_ints({{&ds()->undecided, 1}, {&ds()->qcount[0][0], 0}, {&ds()->qcount[1][0], 0}, {&ds()->fbcount, 0}});
        FlipCtx fc;
        fc.d = dims; fc.dim3 = dim3; fc.m = 1;
        fc.Ccell = Ccell.as<int>(); fc.Cold = Cold.as<int>(); fc.Cnew = Cnew.as<int>();
        fc.state = state.as<int>(); fc.phi = phi.as<int32_t>();
        fc.Ps = Ps.as<unsigned long long>(); fc.pos_old = pos_old.as<int>(); fc.pos_new = pos_new.as<int>();
        fc.count = count.as<int>(); fc.queue[0] = queue0.as<int>(); fc.queue[1] = queue1.as<int>();
        fc.qflags = qflags.as<int>(); fc.s = ds(); fc.slab = 1;
        fc.wave_cap = wave_cap; fc.wg_cap = wg_cap;
        reserve_tags(2);
        k_flip_round<<<1, 64, 0, stream>>>(fc, 0);
        k_flip_bfs<<<1, 256, 0, stream>>>(fc, queue0.as<int>(), &ds()->qcount[0][0], fbq.as<int>(), mark.as<int>(),
                                          reinterpret_cast<int*>(work3.p),
                                          reinterpret_cast<int*>(work4.p), tagctr.as<int>());
        int st = 0;
        HIPCHK(hipMemcpyAsync(&st, state.p, 4, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        return st == ACCEPT ? 1 : 0;
    }

    // connectivity normalisation on the device (demos/gpk.cpp:122-189); the
    // work buffers are free outside a step
    void normalise_device(int n_seeds, const Sites& S) {
        DevBuf sg, g32;
        sg.alloc(size_t(n_seeds) * 4);
        g32.alloc(size_t(N) * 4);
        HIPCHK(hipMemcpyAsync(sg.p, S.g.data(), size_t(n_seeds) * 4, hipMemcpyHostToDevice, stream));
        k_site_g<<<gsblk(N, 256), 256, 0, stream>>>(N, phi.as<int32_t>(), sg.as<float>(), g32.as<float>());
        int* comp = reinterpret_cast<int*>(work3.p);
        int* csz = reinterpret_cast<int*>(work4.p);
        k_cc_init<<<gsblk(N, 256), 256, 0, stream>>>(N, comp);
        int it = 0;
        while (true) {
            set_int(&ds()->changed, 0);
            for (int r = 0; r < 4; ++r, ++it)
                k_cc_step<<<gsblk(N, 256), 256, 0, stream>>>(dims, dim3, phi.as<int32_t>(), comp, &ds()->changed);
            pull_scalars();
            if (hs->changed == 0) break;
            if (it > N + 8) throw GpkError(GPK_E_STALLED, "Voronoi component labelling did not converge");
        }
        fill<int32_t>(work4, N, 0);
        k_cc_size<<<gsblk(N, 256), 256, 0, stream>>>(N, comp, csz);
        unsigned long long* best = Pk.as<unsigned long long>();
        k_fill_r<unsigned long long><<<gsblk(n_seeds + 1, 256), 256, 0, stream>>>(n_seeds + 1, best, 0ull);
        k_cc_best<<<gsblk(N, 256), 256, 0, stream>>>(N, phi.as<int32_t>(), comp, csz, best);
        uint8_t* att = cflag.as<uint8_t>();
        uint8_t* det = uflag.as<uint8_t>();
        k_cc_attached<<<gsblk(N, 256), 256, 0, stream>>>(N, phi.as<int32_t>(), comp, best, att, det);
        const int nu = select_flags(det, N, sel.as<int>());
        int remaining = nu;
        while (remaining > 0) {
            set_int(&ds()->nsel, 0);
            k_attach_scan<<<nblk(nu, 256), 256, 0, stream>>>(dims, dim3, nu, sel.as<int>(), att,
                                                             phi.as<int32_t>(), g32.as<float>(),
                                                             queue0.as<int32_t>(), reinterpret_cast<float*>(work1.p),
                                                             reinterpret_cast<uint8_t*>(qflags.p));
            k_attach_apply<<<nblk(nu, 256), 256, 0, stream>>>(nu, sel.as<int>(), queue0.as<int32_t>(),
                                                              reinterpret_cast<float*>(work1.p),
                                                              reinterpret_cast<uint8_t*>(qflags.p),
                                                              phi.as<int32_t>(), g32.as<float>(), att,
                                                              &ds()->nsel);
            pull_scalars();
            if (hs->nsel == 0) throw GpkError(GPK_E_STALLED, "Voronoi connectivity normalization stalled");
            remaining -= hs->nsel;
        }
        k_widen_g<R><<<gsblk(N, 256), 256, 0, stream>>>(N, g32.as<float>(), g.as<R>());
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(stream));
    }

    void voronoi(int n_seeds, uint32_t seed) override {
        if (n_seeds <= 0 || n_seeds > N)
            throw GpkError(GPK_E_INVALID, "Voronoi seed count must lie between one and the cell count");
        rs_reset = true;
        mask_fresh = false;
        Sites S = draw_sites(Nz, Ny, Nx, n_seeds, seed);
        // nearest site on the device
        DevBuf sz, sy, sx;
        sz.alloc(size_t(n_seeds) * 4); sy.alloc(size_t(n_seeds) * 4); sx.alloc(size_t(n_seeds) * 4);
        HIPCHK(hipMemcpyAsync(sz.p, S.z.data(), size_t(n_seeds) * 4, hipMemcpyHostToDevice, stream));
        HIPCHK(hipMemcpyAsync(sy.p, S.y.data(), size_t(n_seeds) * 4, hipMemcpyHostToDevice, stream));
        HIPCHK(hipMemcpyAsync(sx.p, S.x.data(), size_t(n_seeds) * 4, hipMemcpyHostToDevice, stream));
        // bucket grid of the sites (the same grid as the host search), sites of
        // each bucket in ascending index order
        SiteBuckets sb{};
        {
            const double per = std::max(1.0, double(N) / n_seeds);
            const int B = std::max(1, Nz == 1 ? static_cast<int>(std::sqrt(per)) : static_cast<int>(std::cbrt(per)));
            sb.Bz = Nz == 1 ? Nz : std::min(B, Nz);
            sb.By = std::min(B, Ny);
            sb.Bx = std::min(B, Nx);
            sb.bz = (Nz + sb.Bz - 1) / sb.Bz;
            sb.by = (Ny + sb.By - 1) / sb.By;
            sb.bx = (Nx + sb.Bx - 1) / sb.Bx;
            int bmin = std::min(Ny - (sb.by - 1) * sb.By, sb.By);
            bmin = std::min(bmin, std::min(Nx - (sb.bx - 1) * sb.Bx, sb.Bx));
            if (Nz > 1) bmin = std::min(bmin, std::min(Nz - (sb.bz - 1) * sb.Bz, sb.Bz));
            sb.bmin = std::max(bmin, 1);
        }
        const size_t nbk = size_t(sb.bz) * sb.by * sb.bx;
        std::vector<int> bstart(nbk + 1, 0);
        std::vector<int> bsites(static_cast<size_t>(n_seeds));
        auto bucket_of = [&](int k) {
            return (size_t(S.z[k] / sb.Bz) * sb.by + S.y[k] / sb.By) * sb.bx + S.x[k] / sb.Bx;
        };
        for (int k = 0; k < n_seeds; ++k) ++bstart[bucket_of(k) + 1];
        for (size_t q = 0; q < nbk; ++q) bstart[q + 1] += bstart[q];
        {
            std::vector<int> fillp(bstart.begin(), bstart.end() - 1);
            for (int k = 0; k < n_seeds; ++k) bsites[size_t(fillp[bucket_of(k)]++)] = k;
        }
        DevBuf dbs, dbi;
        dbs.alloc((nbk + 1) * 4);
        dbi.alloc(size_t(n_seeds) * 4);
        HIPCHK(hipMemcpyAsync(dbs.p, bstart.data(), (nbk + 1) * 4, hipMemcpyHostToDevice, stream));
        HIPCHK(hipMemcpyAsync(dbi.p, bsites.data(), size_t(n_seeds) * 4, hipMemcpyHostToDevice, stream));
        k_voronoi_buckets<<<nblk(N, 256), 256, 0, stream>>>(dims, sb, sz.as<int>(), sy.as<int>(), sx.as<int>(),
                                                           dbs.as<int>(), dbi.as<int>(), phi.as<int32_t>());
        HIPCHK(hipGetLastError());
        normalise_device(n_seeds, S);
        fill<R>(alpha, N, R(1));
        check_labels(phi.as<int32_t>());
        rebuild_counts();
        HIPCHK(hipStreamSynchronize(stream));
    }
};

// ---- scratch of the allocation-free adjoint (gpk_velocity_adjoint_noalloc) ----
// The adjoint's per-band-cell terms (nb doubles), per-workgroup partial sums
// (3 doubles per 256 band cells) and its 3-value result.  Every GPU the
// library runs on holds one copy until the process ends.  Host code uses
// the per-device address from hipGetSymbolAddress, never the array's name.
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex scratch_mu;
void* scratch_dev[SCRATCH_MAX_DEVICES] = {};

// once per device, with that device current (called from the engine's set-up)
void scratch_init(int device) {
    if (device < 0 || device >= SCRATCH_MAX_DEVICES)
        throw GpkError(GPK_E_INVALID, "device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> lk(scratch_mu);
    if (scratch_dev[device]) return;
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw GpkError(GPK_E_HIP, "no device address for the adjoint scratch");
    scratch_dev[device] = p;
}

void* scratch_of(int device) {
    std::lock_guard<std::mutex> lk(scratch_mu);
    return scratch_dev[device];
}

EngineBase* make_engine(const gpk_grid_desc& gd, const gpk_model_desc& m) {
    switch (gd.precision) {
    case GPK_PREC_STRICT32: return new Engine<float, true>(gd, m);
    case GPK_PREC_FP32: return new Engine<float, false>(gd, m);
    case GPK_PREC_FP64: return new Engine<double, false>(gd, m);
    default: throw GpkError(GPK_E_INVALID, "unknown precision mode");
    }
}


} // namespace

struct gpk_handle {
    std::unique_ptr<EngineBase> eng;
};

extern "C" {

const char* gpk_last_error(void) { return last_error().c_str(); }
const char* gpk_version(void) { return "gpk_hip 0.1 (gfx950)"; }

void gpk_model_defaults(int32_t kind, gpk_model_desc* m) {
    if (!m) return;
    std::memset(m, 0, sizeof(*m));
    m->kind = kind;
    m->sigma0 = 0.5f; m->eps4 = 0.f; m->eps6 = 0.f; m->M0 = 1e-8f;
    m->k_partition = 1.f; m->noise_scale = 1.f; m->kB = 1.380649e-23f;
    m->m_liq = -500.f; m->k_part = 0.3f; m->T_melt = 933.f; m->L = 1e8f;
    m->bxx = m->byy = m->bzz = 1.f;
    if (kind == GPK_MODEL_ISOTROPIC || kind == GPK_MODEL_ANISO_GROWTH) m->flags = GPK_FLAG_GRAIN_NETWORK;
    if (kind == GPK_MODEL_SOLIDIFICATION) { m->M0 = 5e-8f; m->sigma0 = 0.2f; m->eps4 = 0.04f; }
    m->elastic_smooth_iters = 8;   // include/gpk_core.hpp:99
    m->T_ref = 900.f;
}

int gpk_create(const gpk_grid_desc* grid, const gpk_model_desc* model, gpk_handle** out) {
    return guarded([&] {
        if (!grid || !model || !out) throw GpkError(GPK_E_INVALID, "null argument");
        *out = nullptr;
        if (grid->Nz <= 0 || grid->Ny <= 0 || grid->Nx <= 0)
            throw GpkError(GPK_E_INVALID, "grid dimensions must be positive");
        if (1LL * grid->Nz * grid->Ny * grid->Nx >= (1LL << 31))
            throw GpkError(GPK_E_INVALID, "grid exceeds 2^31 cells");
        if (!(grid->dx > 0.f)) throw GpkError(GPK_E_INVALID, "dx must be positive");
        if (model->kind < GPK_MODEL_ISOTROPIC || model->kind > GPK_MODEL_PRECIPITATION)
            throw GpkError(GPK_E_UNSUPPORTED_MODEL, "unknown model kind");
        if (model->kind == GPK_MODEL_PRECIPITATION && (model->n_phases < 1 || model->n_phases > GPK_MAX_PHASES))
            throw GpkError(GPK_E_INVALID, "precipitation needs 1..GPK_MAX_PHASES phases");
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
            throw GpkError(GPK_E_NO_DEVICE, "no HIP device visible");
        if (grid->device < 0 || grid->device >= ndev) throw GpkError(GPK_E_INVALID, "device ordinal out of range");
        auto h = std::make_unique<gpk_handle>();
        h->eng.reset(make_engine(*grid, *model));
        *out = h.release();
    });
}

int gpk_destroy(gpk_handle* h) {
    return guarded([&] { delete h; });
}

#define NEED_H \
    if (!h) throw GpkError(GPK_E_INVALID, "null handle")

int gpk_upload(gpk_handle* h, int field, const void* host, size_t bytes) {
    return guarded([&] { NEED_H; h->eng->upload(field, host, bytes, false); });
}
int gpk_download(gpk_handle* h, int field, void* host, size_t bytes) {
    return guarded([&] { NEED_H; h->eng->download(field, host, bytes, false); });
}
int gpk_upload_device(gpk_handle* h, int field, const void* p, size_t bytes) {
    return guarded([&] { NEED_H; h->eng->upload(field, p, bytes, true); });
}
int gpk_download_device(gpk_handle* h, int field, void* p, size_t bytes) {
    return guarded([&] { NEED_H; h->eng->download(field, p, bytes, true); });
}

int gpk_step(gpk_handle* h, float dt, float dt_cap, double* dt_used, gpk_step_stats* stats) {
    return guarded([&] {
        NEED_H;
        gpk_step_stats st{};
        h->eng->step(dt, dt_cap, &st);
        if (dt_used) *dt_used = st.dt;
        if (stats) *stats = st;
    });
}

int gpk_step_many(gpk_handle* h, int nsteps, float dt, float dt_cap, gpk_step_stats* stats) {
    return guarded([&] {
        NEED_H;
        if (nsteps < 0) throw GpkError(GPK_E_INVALID, "nsteps must be non-negative");
        gpk_step_stats acc{}, st{};
        for (int s = 0; s < nsteps; ++s) {
            h->eng->step(dt, dt_cap, &st);
            acc.band_cells += st.band_cells; acc.candidates += st.candidates;
            acc.accepted += st.accepted; acc.deferred += st.deferred;
            acc.exchanges += st.exchanges; acc.eroded += st.eroded; acc.guarded += st.guarded;
            acc.dt = st.dt; acc.time = st.time; acc.step_id = st.step_id;
        }
        if (stats) *stats = acc;
    });
}

int gpk_run(gpk_handle* h, double t_end, float dt, int* steps_taken) {
    return guarded([&] {
        NEED_H;
        int n = 0;
        double t = h->eng->time();
        while (t < t_end) {
            const float remaining = static_cast<float>(t_end - t);
            gpk_step_stats st{};
            h->eng->step(dt, remaining, &st);
            t = st.time;
            ++n;
        }
        if (steps_taken) *steps_taken = n;
    });
}

int gpk_velocity_field(gpk_handle* h, const double theta[3], void* vout_dev) {
    return guarded([&] {
        NEED_H;
        if (!theta || !vout_dev) throw GpkError(GPK_E_INVALID, "null theta or output");
        h->eng->velocity_field(theta, vout_dev);
    });
}

int gpk_velocity_adjoint(gpk_handle* h, const double theta[3], const void* gvel_dev, double grad[3]) {
    return guarded([&] {
        NEED_H;
        if (!theta || !gvel_dev || !grad) throw GpkError(GPK_E_INVALID, "null theta, gradient or output");
        h->eng->velocity_adjoint(theta, gvel_dev, grad);
    });
}

int gpk_velocity_adjoint_noalloc(gpk_handle* h, const double theta[3], const void* gvel_dev, double grad[3]) {
    return guarded([&] {
        NEED_H;
        if (!theta || !gvel_dev || !grad) throw GpkError(GPK_E_INVALID, "null theta, gradient or output");
        h->eng->velocity_adjoint_noalloc(theta, gvel_dev, grad);
    });
}

int gpk_set_model(gpk_handle* h, const gpk_model_desc* m) {
    return guarded([&] {
        NEED_H;
        if (!m) throw GpkError(GPK_E_INVALID, "null model");
        h->eng->set_model(*m);
    });
}

int gpk_get_time(gpk_handle* h, double* t) {
    return guarded([&] { NEED_H; if (!t) throw GpkError(GPK_E_INVALID, "null"); *t = h->eng->time(); });
}
int gpk_set_time(gpk_handle* h, double t) {
    return guarded([&] { NEED_H; h->eng->set_time(t); });
}
int gpk_max_dt(gpk_handle* h, double* dt) {
    return guarded([&] { NEED_H; if (!dt) throw GpkError(GPK_E_INVALID, "null"); *dt = h->eng->max_dt(); });
}
int gpk_phase_volume(gpk_handle* h, int32_t label, double* v) {
    return guarded([&] { NEED_H; if (!v) throw GpkError(GPK_E_INVALID, "null"); *v = h->eng->phase_volume(label); });
}
int gpk_n_grains(gpk_handle* h, int* n) {
    return guarded([&] { NEED_H; if (!n) throw GpkError(GPK_E_INVALID, "null"); *n = h->eng->n_grains(); });
}
int gpk_mean_grain_radius(gpk_handle* h, double* r) {
    return guarded([&] { NEED_H; if (!r) throw GpkError(GPK_E_INVALID, "null"); *r = h->eng->mean_grain_radius(); });
}
int gpk_solid_fraction(gpk_handle* h, float* f) {
    return guarded([&] { NEED_H; if (!f) throw GpkError(GPK_E_INVALID, "null"); *f = h->eng->solid_fraction(); });
}
int gpk_interface_fraction(gpk_handle* h, float* f) {
    return guarded([&] { NEED_H; if (!f) throw GpkError(GPK_E_INVALID, "null"); *f = h->eng->interface_fraction(); });
}
int gpk_solute_sum(gpk_handle* h, double* s) {
    return guarded([&] { NEED_H; if (!s) throw GpkError(GPK_E_INVALID, "null"); *s = h->eng->solute_sum(); });
}
int gpk_device_bytes(gpk_handle* h, size_t* b) {
    return guarded([&] { NEED_H; if (!b) throw GpkError(GPK_E_INVALID, "null"); *b = h->eng->device_bytes(); });
}
int gpk_topology_simple_move(gpk_handle* h, int cell, int32_t new_label, int* result) {
    return guarded([&] {
        NEED_H;
        if (!result) throw GpkError(GPK_E_INVALID, "null");
        *result = h->eng->simple_move(cell, new_label);
    });
}
int gpk_voronoi_init(gpk_handle* h, int n_seeds, uint32_t seed) {
    return guarded([&] { NEED_H; h->eng->voronoi(n_seeds, seed); });
}
int gpk_voronoi_host(int Nz, int Ny, int Nx, int n_seeds, uint32_t seed, int32_t* phi, float* g) {
    return guarded([&] {
        if (!phi || !g || Nz <= 0 || Ny <= 0 || Nx <= 0) throw GpkError(GPK_E_INVALID, "bad argument");
        const long long N = 1LL * Nz * Ny * Nx;
        if (n_seeds <= 0 || n_seeds > N)
            throw GpkError(GPK_E_INVALID, "Voronoi seed count must lie between one and the cell count");
        Sites S = draw_sites(Nz, Ny, Nx, n_seeds, seed);
        nearest_host(Nz, Ny, Nx, S, phi);
        const int rc = normalise_components(Nz, Ny, Nx, n_seeds, S, phi, g);
        if (rc != GPK_OK) throw GpkError(rc, "Voronoi connectivity normalization stalled");
    });
}
int gpk_sync(gpk_handle* h) {
    return guarded([&] { NEED_H; h->eng->sync(); });
}
int gpk_set_profiling(gpk_handle* h, int enable) {
    return guarded([&] { NEED_H; h->eng->set_profiling(enable); });
}
int gpk_kernel_timing(gpk_handle* h, int which, double* avg_ms, int64_t* launches, double* bytes) {
    return guarded([&] { NEED_H; h->eng->timing(which, avg_ms, launches, bytes); });
}

} // extern "C"
