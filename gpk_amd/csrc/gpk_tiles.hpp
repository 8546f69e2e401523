// gpk_tiles.hpp — LDS-tiled geometry, pair projection and remap scoring of
// one tile, as device functions run by a whole workgroup.  The grid-wide
// kernels (gpk_kernels.hpp) run one tile per workgroup; the batched 2-D
// kernel (gpk_batch.hip) walks an instance's tiles with one workgroup.
#pragma once
#include <type_traits>

#include "gpk_device.hpp"
#include "gpk_cell.hpp"

namespace gpkd {

// ---------------------------------------------------------------------------
// Separable geometry from per-tile row masks (FP32 / FP64 modes).
// One workgroup per TZ x TY x TX tile.  The tile's interface cells are listed
// and their resident labels deduplicated in an LDS hash.  For a batch of up to
// MAXRES residents, every staged row (tile + RG halo, SX <= 64 cells) is read
// by one wavefront (or half of one when SX <= 32) and turned into one "same
// label" mask per resident with a single compare + ballot (bit = staged
// column; 32-bit words when SX <= 32, else 64-bit).  Each interface cell
// then takes its window row mask as (row mask >> lx) and reads the row sums
// from the chunk tables; y and z follow with FMA.  Same operation order as
// k_geometry<R,DIM,false>.  Tiles are assigned XCD-major (tile t =
// (id % 8) * per_xcd + id / 8) so neighbouring tiles share one XCD's L2.
// ---------------------------------------------------------------------------
// 3-D tiles up to radius 8 are 16 cells wide in x, so that a staged row
// (16 + 2 * 8 cells) is one 32-bit mask and one wavefront load covers two
// staged rows; wider radii keep 32-cell-wide tiles with 64-bit masks.
template <int DIM, int RG> struct TileShape;
template <int RG> struct TileShape<3, RG> {
    static constexpr int TZ = RG <= 8 ? 16 : 8, TY = 8, TX = RG <= 8 ? 16 : 32;
};
template <int RG> struct TileShape<2, RG> { static constexpr int TZ = 1, TY = 64, TX = 32; };

constexpr int GEO_HCAP = 256;    // resident hash slots per tile (resident indices fit 8 bits)

template <class R, int DIM, int RG>
struct TiledLds {
    static constexpr int TZ = TileShape<DIM, RG>::TZ, TY = TileShape<DIM, RG>::TY, TX = TileShape<DIM, RG>::TX;
    static constexpr int HZ = DIM == 3 ? RG : 0;
    static constexpr int SZ = TZ + 2 * HZ, SY = TY + 2 * RG, SX = TX + 2 * RG;
    // row mask word: 32 bits when a staged row fits them (then one wavefront
    // load stages RPL = 2 rows, one per half-wavefront)
    static constexpr bool NARROW = SX <= 32;
    using M = typename std::conditional<NARROW, uint32_t, unsigned long long>::type;
    static constexpr int RPL = NARROW ? 2 : 1;
    // residents per mask batch: as many as keep a 256-thread workgroup's LDS
    // within a quarter of the CU's 160 KB at radius 8 (four workgroups per CU)
    static constexpr int MAXRES = NARROW ? (sizeof(R) == 8 ? 5 : 7) : (sizeof(R) == 8 ? 3 : 4);
    static constexpr int MAXB = (GEO_HCAP + MAXRES - 1) / MAXRES;   // batches per tile
    static constexpr int NROWS = SZ * SY;
    static constexpr int W = 2 * RG + 1;
    static constexpr int NCH = (W + 8) / 9;
    static constexpr int C = (W + NCH - 1) / NCH;
    static constexpr int lut_entries() {
        int n = 0;
        for (int k = 0; k < NCH; ++k) {
            const int len = (W - k * C) < C ? (W - k * C) : C;
            n += (1 << len) * 3;
        }
        return n;
    }
    static constexpr int NLUT = lut_entries();
    // LDS stride of one table entry (3 sums): FP32 entries are padded to 16
    // bytes so that one 128-bit LDS read returns all three
    static constexpr int LS = sizeof(R) == 4 ? 4 : 3;
    static constexpr int NLUT_LDS = NLUT / 3 * LS;
    static_assert(SX <= 64, "a staged row must fit one wavefront");
};

// stencil sums of one interface cell from its window row masks
// (rowmask(jz, jy): bit jx set when window position jx carries the resident)
// stencil sums of one interface cell from its window row masks
// (rowmask(jz, jy): bit jx set when window position jx carries the resident),
// or of two cells at z0 and z0 + 1 with the same resident (3-D, `pair`): the
// y sums of a window plane are then formed once for both cells -- plane jz of
// the upper cell is plane jz + 1 of the lower -- and each cell adds its own
// planes in its own window order, so every value is the one the cell alone
// would get.
template <class R, int DIM, int RG, int LAW, class RowMask>
__device__ __forceinline__ void tiled_cell(
    const Dims& d, const GeomParams& gp, const gpk_model_desc& model, int x0, int y0, int z0, bool pair,
    int32_t p, const R* lut, const R* __restrict__ w0, const R* __restrict__ w1,
    const R* __restrict__ w2, RowMask rowmask,
    const int32_t* __restrict__ phi, const int* __restrict__ slot, const R* __restrict__ g,
    const R* __restrict__ c, const R* __restrict__ T, const R* __restrict__ rho,
    const R* __restrict__ E_el, R* __restrict__ onx, R* __restrict__ ony,
    R* __restrict__ onz, R* __restrict__ okappa, R* __restrict__ vraw,
    int32_t* __restrict__ phi_comp, R* __restrict__ g_comp) {
    using S = TiledLds<R, DIM, RG>;
    R Ix = 0, Iy = 0, Iz = 0, Ixx = 0, Iyy = 0, Izz = 0, Ixy = 0, Ixz = 0, Iyz = 0;
    R Jx = 0, Jy = 0, Jz = 0, Jxx = 0, Jyy = 0, Jzz = 0, Jxy = 0, Jxz = 0, Jyz = 0;   // upper cell
    // one extra plane when any lane of the wavefront carries a pair
    const int nzw = DIM == 3 ? S::W + (__ballot(pair) != 0ull ? 1 : 0) : 1;
    for (int jz = 0; jz < nzw; ++jz) {
        R S00 = 0, S01 = 0, S02 = 0, S10 = 0, S11 = 0, S20 = 0;
        // row loop unrolled fully in FP32 and by 8 in FP64 (measured: the
        // FP64 build runs out of registers with the full unroll)
        constexpr int UNR = sizeof(R) == 4 ? S::W : 8;
#pragma unroll UNR
        for (int jy = 0; jy < S::W; ++jy) {
            const auto m = rowmask(jz, jy);
            using MW = decltype(m + 0u);
            R R0, R1, R2;
            int base = 0;
#pragma unroll
            for (int k = 0; k < S::NCH; ++k) {
                const int start = k * S::C;
                const int len = (S::W - start) < S::C ? (S::W - start) : S::C;
                const int ix = int((MW(m) >> start) & ((MW(1) << len) - MW(1)));
                R t0, t1, t2;
                if constexpr (S::LS == 4) {
                    const float4 v = *reinterpret_cast<const float4*>(lut + base + ix * 4);
                    t0 = v.x; t1 = v.y; t2 = v.z;
                } else {
                    const R* t = lut + base + ix * S::LS;
                    t0 = t[0]; t1 = t[1]; t2 = t[2];
                }
                if (k == 0) { R0 = t0; R1 = t1; R2 = t2; }
                else { R0 += t0; R1 += t1; R2 += t2; }
                base += (1 << len) * S::LS;
            }
            const R gy0 = w0[jy], gy1 = w1[jy], gy2 = w2[jy];
            S00 = fmar(gy0, R0, S00); S01 = fmar(gy0, R1, S01);
            S02 = fmar(gy0, R2, S02); S10 = fmar(gy1, R0, S10);
            S11 = fmar(gy1, R1, S11); S20 = fmar(gy2, R0, S20);
        }
        if (DIM == 3) {
            if (jz < S::W) {
                const R gz0 = w0[jz], gz1 = w1[jz], gz2 = w2[jz];
                Ix = fmar(gz0, S01, Ix); Iy = fmar(gz0, S10, Iy);
                Ixx = fmar(gz0, S02, Ixx); Iyy = fmar(gz0, S20, Iyy);
                Ixy = fmar(gz0, S11, Ixy);
                Iz = fmar(gz1, S00, Iz); Izz = fmar(gz2, S00, Izz);
                Ixz = fmar(gz1, S01, Ixz); Iyz = fmar(gz1, S10, Iyz);
            }
            if (jz >= 1) {
                const R gz0 = w0[jz - 1], gz1 = w1[jz - 1], gz2 = w2[jz - 1];
                Jx = fmar(gz0, S01, Jx); Jy = fmar(gz0, S10, Jy);
                Jxx = fmar(gz0, S02, Jxx); Jyy = fmar(gz0, S20, Jyy);
                Jxy = fmar(gz0, S11, Jxy);
                Jz = fmar(gz1, S00, Jz); Jzz = fmar(gz2, S00, Jzz);
                Jxz = fmar(gz1, S01, Jxz); Jyz = fmar(gz1, S10, Jyz);
            }
        } else {
            Ix = S01; Iy = S10; Ixx = S02; Iyy = S20; Ixy = S11;
        }
    }
    auto cube_at = [&](int t, int zz, int yy, int xx) -> int32_t {
        (void)t;
        return phi[lin(d, zz, yy, xx)];
    };
    for (int h = 0; h < (pair ? 2 : 1); ++h) {
        const int zc = z0 + h;
        const int i = lin(d, zc, y0, x0);
        const bool u = h != 0;
        finish_cell<R, DIM>(d, gp, model, i, slot[i], zc, y0, x0, p, u ? Jx : Ix, u ? Jy : Iy, u ? Jz : Iz,
                            u ? Jxx : Ixx, u ? Jyy : Iyy, u ? Jzz : Izz, u ? Jxy : Ixy, u ? Jxz : Ixz,
                            u ? Jyz : Iyz, cube_at, phi, g, c, T, rho, E_el, onx, ony, onz, okappa, vraw,
                            phi_comp, g_comp, LawTag<LAW>{});
    }
}

// LDS of one geometry tile (rm aliases the tile's interface list and each
// listed cell's resident index before the first batch: 2 + 2 bytes per cell)
template <class R, int DIM, int RG>
struct GeoTileSmem {
    using S = TiledLds<R, DIM, RG>;
    static constexpr int NT = S::TZ * S::TY * S::TX;
    using M = typename S::M;
    static constexpr int NLC = NT * 4 / int(sizeof(M));
    static constexpr int NRM = S::NROWS * S::MAXRES > NLC ? S::NROWS * S::MAXRES : NLC;
    M rm[NRM];                      // row masks of one resident batch
    alignas(16) R lut[S::NLUT_LDS];
    uint16_t blist[NT];             // (pair << 15) | (slot in the batch << 12) | tile cell, grouped by batch
    int hkey[GEO_HCAP];
    uint8_t hidx[GEO_HCAP];         // resident index of a hash slot (< GEO_HCAP)
    int resl[GEO_HCAP];
    int bcount[S::MAXB + 1], boff[S::MAXB + 1];
    int zoff[S::SZ], yoff[S::SY];   // wrapped plane / row offsets of the staged rows
    int nlist, nres;
};

// row-sum tables into the tile LDS (they stay valid across tiles)
template <class R, int DIM, int RG>
__device__ __forceinline__ void geometry_tile_lut(GeoTileSmem<R, DIM, RG>& sm, const R* __restrict__ lut_g) {
    using S = TiledLds<R, DIM, RG>;
    for (int e = threadIdx.x; e < S::NLUT; e += blockDim.x) sm.lut[e / 3 * S::LS + e % 3] = lut_g[e];
}

// the geometry of every interface cell of tile `tile` (tiles numbered x
// fastest), run by the whole workgroup; the tables must have been loaded
template <class R, int DIM, int RG, int LAW>
__device__ __forceinline__ void geometry_tile(
    GeoTileSmem<R, DIM, RG>& sm, int tile, const Dims& d, const GeomParams& gp, const gpk_model_desc& model,
    int ntx, int nty, const int32_t* __restrict__ phi, const uint8_t* __restrict__ mask,
    const int* __restrict__ slot, const R* __restrict__ w0, const R* __restrict__ w1, const R* __restrict__ w2,
    const R* __restrict__ g, const R* __restrict__ c, const R* __restrict__ T, const R* __restrict__ rho,
    const R* __restrict__ E_el, R* __restrict__ onx, R* __restrict__ ony, R* __restrict__ onz,
    R* __restrict__ okappa, R* __restrict__ vraw, int32_t* __restrict__ phi_comp, R* __restrict__ g_comp) {
    using S = TiledLds<R, DIM, RG>;
    using SM = GeoTileSmem<R, DIM, RG>;
    constexpr int NT = SM::NT;
    using M = typename S::M;
    static_assert(NT <= 4096 && S::MAXRES <= 8, "tile index, batch slot and pair flag must fit 16 bits");
    M* const rm = sm.rm;
    short* const list = reinterpret_cast<short*>(rm);
    unsigned short* const cres = reinterpret_cast<unsigned short*>(rm) + NT;
    const R* const lut = sm.lut;
    uint16_t* const blist = sm.blist;
    int* const hkey = sm.hkey;
    uint8_t* const hidx = sm.hidx;
    int* const resl = sm.resl;
    int* const bcount = sm.bcount;
    int* const boff = sm.boff;
    int* const zoff = sm.zoff;
    int* const yoff = sm.yoff;
    int& nlist = sm.nlist;
    int& nres = sm.nres;
    __syncthreads();   // the workgroup's previous tile may still read the LDS
    const int tx = tile % ntx, ty = (tile / ntx) % nty, tz = tile / (ntx * nty);
    const int x0t = tx * S::TX, y0t = ty * S::TY, z0t = tz * S::TZ;

    if (threadIdx.x == 0) { nlist = 0; nres = 0; }
    for (int e = threadIdx.x; e < GEO_HCAP; e += blockDim.x) hkey[e] = -1;
    for (int e = threadIdx.x; e <= S::MAXB; e += blockDim.x) bcount[e] = 0;
    for (int e = threadIdx.x; e < S::SZ; e += blockDim.x) zoff[e] = wrapi(z0t + e - S::HZ, d.Nz) * d.Ny * d.Nx;
    for (int e = threadIdx.x; e < S::SY; e += blockDim.x) yoff[e] = wrapi(y0t + e - RG, d.Ny) * d.Nx;
    __syncthreads();
    // interface cells of the tile and their distinct resident labels, one
    // thread per z column of the tile; in 3-D the column's band cells are
    // paired greedily from the bottom: a band cell heads a pair with the band
    // cell above it when both carry the same label and it is not itself the
    // upper cell of a pair (zpair: 1 head, 2 taken; the batch list's memory,
    // not used before the list is built)
    uint8_t* const zpair = reinterpret_cast<uint8_t*>(blist);
    static_assert(sizeof(sm.blist) >= NT, "pair flags alias the batch list");
    for (int col = threadIdx.x; col < S::TY * S::TX; col += blockDim.x) {
        const int lx = col % S::TX, ly = col / S::TX;
        const int x = x0t + lx, y = y0t + ly;
        int plab = -1;          // label of the previous band cell if it is free to head a pair
        for (int lz = 0; lz < S::TZ; ++lz) {
            const int e = lz * (S::TY * S::TX) + col;
            const int z = z0t + lz;
            int lab = -1;
            if (x < d.Nx && y < d.Ny && z < d.Nz) {
                const int i = lin(d, z, y, x);
                if (mask[i]) {
                    list[atomicAdd(&nlist, 1)] = short(e);
                    const int p = phi[i];
                    lab = p;
                    int h = int((uint32_t(p) * 2654435761u) >> 23) & (GEO_HCAP - 1);
                    for (int probe = 0; probe < GEO_HCAP; ++probe) {
                        const int prev = atomicCAS(&hkey[h], -1, p);
                        if (prev == -1 || prev == p) break;
                        h = (h + 1) & (GEO_HCAP - 1);
                    }
                }
            }
            if (DIM == 3) {
                uint8_t f = 0;
                if (lab >= 0 && plab == lab) {
                    f = 2;                       // upper cell of the pair headed at lz - 1
                    zpair[e - S::TY * S::TX] = 1;
                    plab = -1;
                } else {
                    plab = lab;
                }
                zpair[e] = f;
            }
        }
    }
    __syncthreads();
    for (int h = threadIdx.x; h < GEO_HCAP; h += blockDim.x) {
        if (hkey[h] != -1) {
            const int k = atomicAdd(&nres, 1);
            hidx[h] = uint8_t(k);
            resl[k] = hkey[h];
        }
    }
    __syncthreads();
    const int n = nlist;
    for (int q = threadIdx.x; q < n; q += blockDim.x) {
        const int e = list[q];
        const int lx = e % S::TX, ly = (e / S::TX) % S::TY, lz = e / (S::TX * S::TY);
        const int p = phi[lin(d, z0t + lz, y0t + ly, x0t + lx)];
        int h = int((uint32_t(p) * 2654435761u) >> 23) & (GEO_HCAP - 1);
        unsigned short r = 0xFFFFu;     // 0xFFFF: the hash was full, no batch
        for (int probe = 0; probe < GEO_HCAP; ++probe) {
            if (hkey[h] == p) { r = (unsigned short)hidx[h]; break; }
            if (hkey[h] == -1) break;
            h = (h + 1) & (GEO_HCAP - 1);
        }
        // 3-D pairs from the column pass (bit 8: head of a pair, bit 9:
        // taken by one)
        unsigned short fl = 0;
        if (DIM == 3 && r != 0xFFFFu) {
            const uint8_t zf = zpair[e];
            fl = zf == 1 ? 0x100 : (zf == 2 ? 0x200 : 0);
        }
        cres[q] = r == 0xFFFFu ? r : (unsigned short)(r | fl);
        if (!(fl & 0x200)) atomicAdd(&bcount[r == 0xFFFFu ? S::MAXB : r / S::MAXRES], 1);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int o = 0;
        for (int b = 0; b <= S::MAXB; ++b) { boff[b] = o; o += bcount[b]; }
    }
    // the hash is no longer read: its slots count each batch's pairs (from
    // the front of the batch) and single cells (from the back), so that
    // wavefronts run pairs or singles but rarely both
    static_assert(128 + S::MAXB < GEO_HCAP, "batch counters must fit the hash");
    int* const fcnt = hkey;
    int* const bcnt = hkey + 128;
    for (int e = threadIdx.x; e <= S::MAXB; e += blockDim.x) { fcnt[e] = 0; bcnt[e] = 0; }
    __syncthreads();
    for (int q = threadIdx.x; q < n; q += blockDim.x) {
        const uint32_t cr = cres[q];
        if (cr != 0xFFFFu && (cr & 0x200u)) continue;     // handled with its lz - 1 neighbour
        const uint32_t r = cr == 0xFFFFu ? cr : (cr & 0xFFu);
        const uint32_t pr = cr == 0xFFFFu ? 0u : ((cr >> 8) & 1u);
        const int b = r == 0xFFFFu ? S::MAXB : int(r) / S::MAXRES;
        const uint32_t t = r == 0xFFFFu ? 0u : r % S::MAXRES;
        const int pos = pr ? boff[b] + atomicAdd(&fcnt[b], 1) : boff[b] + bcount[b] - 1 - atomicAdd(&bcnt[b], 1);
        blist[pos] = uint16_t((pr << 15) | (t << 12) | uint32_t(list[q]));
    }
    const int nr = nres;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
    // staged column of this lane: with 32-bit masks each half-wavefront
    // stages its own row (h = which one)
    const int sl = S::NARROW ? (lane & 31) : lane;
    const int h = S::NARROW ? (lane >> 5) : 0;
    const int xw = sl < S::SX ? wrapi(x0t + sl - RG, d.Nx) : 0;
    constexpr int INFL = 4;                  // wavefront loads in flight
    constexpr int RPP = INFL * S::RPL;       // staged rows per wavefront pass
    for (int b0 = 0; b0 < nr; b0 += S::MAXRES) {
        const int nbat = nr - b0 < S::MAXRES ? nr - b0 : S::MAXRES;
        const int bb = b0 / S::MAXRES;
        __syncthreads();
        // row masks: each wavefront pass stages RPP consecutive rows (RPL per
        // load), one ballot per load and resident of the batch; the mask
        // words are gathered into lanes (k * MAXRES + t) * RPL + row-of-load
        // by branch-free selects and stored by one LDS write per lane
        static_assert(INFL * S::MAXRES * S::RPL <= 64, "one lane per mask word of a pass");
        int rv[S::MAXRES];
#pragma unroll
        for (int t = 0; t < S::MAXRES; ++t) rv[t] = t < nbat ? resl[b0 + t] : -2;   // -2: no label
        for (int r0 = wave * RPP; r0 < S::NROWS; r0 += RPP * nwave) {
            int v[INFL];
#pragma unroll
            for (int k = 0; k < INFL; ++k) {
                const int r = r0 + k * S::RPL + h;
                v[k] = -1;
                if (r < S::NROWS && sl < S::SX) v[k] = phi[zoff[r / S::SY] + yoff[r % S::SY] + xw];
            }
            int mw0 = 0, mw1 = 0;   // mask words (64-bit masks: low, high halves)
            const int qk = lane / S::RPL;
            const bool odd = (lane & 1) != 0;
#pragma unroll
            for (int k = 0; k < INFL; ++k)
#pragma unroll
                for (int t = 0; t < S::MAXRES; ++t) {
                    const unsigned long long msk = __ballot(v[k] == rv[t]);
                    const bool mine = qk == k * S::MAXRES + t;
                    if constexpr (S::NARROW) {
                        const uint32_t w = odd ? uint32_t(msk >> 32) : uint32_t(msk);
                        mw0 = mine ? int(w) : mw0;
                    } else {
                        mw0 = mine ? int(uint32_t(msk)) : mw0;
                        mw1 = mine ? int(uint32_t(msk >> 32)) : mw1;
                    }
                }
            const int kr = qk / S::MAXRES, tr = qk - kr * S::MAXRES;
            const int r = r0 + kr * S::RPL + (odd && S::NARROW ? 1 : 0);
            if (lane < INFL * S::MAXRES * S::RPL && r < S::NROWS) {
                if constexpr (S::NARROW) rm[r * S::MAXRES + tr] = M(uint32_t(mw0));
                else rm[r * S::MAXRES + tr] = (M(uint32_t(mw1)) << 32) | M(uint32_t(mw0));
            }
        }
        __syncthreads();
        const int bend = boff[bb] + bcount[bb];
        for (int u = boff[bb] + threadIdx.x; u < bend; u += blockDim.x) {
            const uint32_t pk = blist[u];
            const int e = int(pk & 0xFFFu);
            const int t = int((pk >> 12) & 7u);
            const bool pr = (pk >> 15) & 1u;
            const int lx = e % S::TX, ly = (e / S::TX) % S::TY, lz = e / (S::TX * S::TY);
            const M* rows = rm + (lz * S::SY + ly) * S::MAXRES + t;
            tiled_cell<R, DIM, RG, LAW>(d, gp, model, x0t + lx, y0t + ly, z0t + lz, pr, resl[b0 + t], lut, w0, w1, w2,
                                   [&](int jz, int jy) {
                                       // the extra plane of a pair; clamped for the
                                       // single cells of the same wavefront (unused)
                                       const int jc = jz < S::SZ - lz ? jz : S::SZ - 1 - lz;
                                       return rows[(jc * S::SY + jy) * S::MAXRES] >> lx;
                                   },
                                   phi, slot, g, c, T, rho, E_el, onx, ony, onz, okappa, vraw, phi_comp, g_comp);
        }
    }
    // residents that did not fit the hash: window rows read from the grid
    __syncthreads();
    for (int u = boff[S::MAXB] + threadIdx.x; u < boff[S::MAXB] + bcount[S::MAXB]; u += blockDim.x) {
        const int e = int(blist[u] & 0xFFFu);
        const int lx = e % S::TX, ly = (e / S::TX) % S::TY, lz = e / (S::TX * S::TY);
        const int x0 = x0t + lx, y0 = y0t + ly, z0 = z0t + lz;
        const int32_t p = phi[lin(d, z0, y0, x0)];
        tiled_cell<R, DIM, RG, LAW>(d, gp, model, x0, y0, z0, false, p, lut, w0, w1, w2,
                               [&](int jz, int jy) {
                                   const int z = DIM == 3 ? wrapi(z0 + jz - RG, d.Nz) : z0;
                                   const int y = wrapi(y0 + jy - RG, d.Ny);
                                   unsigned long long mm = 0;
                                   for (int jx = 0; jx < S::W; ++jx)
                                       if (phi[lin(d, z, y, wrapi(x0 + jx - RG, d.Nx))] == p)
                                           mm |= 1ull << jx;
                                   return mm;
                               },
                               phi, slot, g, c, T, rho, E_el, onx, ony, onz, okappa, vraw, phi_comp, g_comp);
    }
}

// LAW = 1: compiled for the plain isotropic law only (LawTag<1>).  At radius 8
// (sigma 1.5) that instance fits 128 VGPRs and 40 KB of LDS, i.e. four
// workgroups per CU; the register bound is asked of the compiler only there,
// since other radii would spill.
template <class R, int DIM, int RG, int LAW>
__global__ void __launch_bounds__(256, LAW == 1 && RG == 8 ? 4 : 1)
k_geometry_tiled(Dims d, GeomParams gp, gpk_model_desc model, int ntx, int nty, int ntz,
                 const int32_t* __restrict__ phi, const uint8_t* __restrict__ mask,
                 const int* __restrict__ slot, const R* __restrict__ w0,
                 const R* __restrict__ w1, const R* __restrict__ w2,
                 const R* __restrict__ lut_g,
                 const R* __restrict__ g, const R* __restrict__ c, const R* __restrict__ T,
                 const R* __restrict__ rho, const R* __restrict__ E_el,
                 R* __restrict__ onx, R* __restrict__ ony, R* __restrict__ onz,
                 R* __restrict__ okappa, R* __restrict__ vraw,
                 int32_t* __restrict__ phi_comp, R* __restrict__ g_comp) {
    __shared__ GeoTileSmem<R, DIM, RG> sm;
    const int ntiles = ntx * nty * ntz;
    const int per_xcd = (ntiles + 7) / 8;
    const int tile = (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
    if (tile >= ntiles) return;
    geometry_tile_lut(sm, lut_g);
    geometry_tile<R, DIM, RG, LAW>(sm, tile, d, gp, model, ntx, nty, phi, mask, slot, w0, w1, w2, g, c, T, rho,
                                   E_el, onx, ony, onz, okappa, vraw, phi_comp, g_comp);
}

__device__ __forceinline__ void atomic_max_abs(unsigned int* p, float v) {
    atomicMax(p, __float_as_uint(fabsf(v)));
}
__device__ __forceinline__ void atomic_max_abs(unsigned long long* p, double v) {
    atomicMax(p, (unsigned long long)__double_as_longlong(fabs(v)));
}
template <class R> struct UBits;
template <> struct UBits<float> { using T = unsigned int; };
template <> struct UBits<double> { using T = unsigned long long; };
__device__ __forceinline__ unsigned int abs_bits(float v) { return __float_as_uint(fabsf(v)); }
__device__ __forceinline__ unsigned long long abs_bits(double v) {
    return (unsigned long long)__double_as_longlong(fabs(v));
}
// the value back from abs_bits
template <class R> __device__ __forceinline__ R bits_abs(typename UBits<R>::T b);
template <> __device__ __forceinline__ float bits_abs<float>(unsigned int b) { return __uint_as_float(b); }
template <> __device__ __forceinline__ double bits_abs<double>(unsigned long long b) {
    return __longlong_as_double((long long)b);
}

// Workgroup maximum of |v| bits folded into one atomicMax on dst (the bits
// of a non-negative float or double order as the values do): wavefront
// shuffles, one LDS atomic per wavefront, one device atomic per workgroup,
// instead of one device atomic per wavefront on a single address.  Called by
// every thread of the workgroup; lds must hold 0 and have been published by
// a barrier before the call.
template <class U>
__device__ __forceinline__ void block_max_to(U v, U& lds, U* __restrict__ dst) {
    for (int o = 32; o > 0; o >>= 1) {
        const U u = __shfl_xor(v, o);
        v = v < u ? u : v;
    }
    if ((threadIdx.x & 63) == 0) atomicMax(&lds, v);
    __syncthreads();
    if (threadIdx.x == 0) atomicMax(dst, lds);
}


// ---------------------------------------------------------------------------
// LDS-tiled projection and remap scoring (labels < 2^16).  A workgroup stages
// a TZ x TY x TX tile plus an R-cell halo; each interface cell of the tile then
// runs the reference's window sum in the reference order out of LDS.
// ---------------------------------------------------------------------------
template <int DIM> struct SmallTile;
template <> struct SmallTile<3> { static constexpr int TZ = 4, TY = 8, TX = 32; };
template <> struct SmallTile<2> { static constexpr int TZ = 1, TY = 32, TX = 32; };

template <int DIM, int RH>
struct HaloGeom {
    static constexpr int TZ = SmallTile<DIM>::TZ, TY = SmallTile<DIM>::TY, TX = SmallTile<DIM>::TX;
    static constexpr int HZ = DIM == 3 ? RH : 0;
    static constexpr int SZ = TZ + 2 * HZ, SY = TY + 2 * RH, SX = TX + 2 * RH;
    static constexpr int NS = SZ * SY * SX;
    static constexpr int NT = TZ * TY * TX;
};

__device__ __forceinline__ int xcd_tile(int ntiles) {
    const int per_xcd = (ntiles + 7) / 8;
    return (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
}

// acc + v when hit, acc otherwise, as one select and one FMA: fma(v, 1, acc)
// rounds acc + v exactly as the add does, and fma(v, 0, acc) adds +-0, which
// leaves acc unchanged (the window sums start at +0 and never hold -0)
__device__ __forceinline__ double add_if(bool hit, double v, double acc) {
    return __builtin_fma(v, hit ? 1.0 : 0.0, acc);
}

// same-pair projection of v (src/gpk_physics.cpp:152-193) from LDS
constexpr int PROJ_THREADS = 512;   // 16 wavefronts per CU at two tiles' LDS

template <class R, int DIM, int RP>
struct ProjTileSmem {
    using H = HaloGeom<DIM, RP>;
    uint32_t key[H::NS];
    R vv[H::NS];
    short list[H::NT];
    int zoff[H::SZ], yoff[H::SY], xoff[H::SX];   // wrapped offsets of the staged box
    int nlist;
};

// projected velocity of every interface cell of tile `tile`, run by the
// whole workgroup; returns the thread's largest |v| as bits
template <class R, int DIM, int RP>
__device__ __forceinline__ typename UBits<R>::T project_tile(
    ProjTileSmem<R, DIM, RP>& sm, int tile, const Dims& d, int ntx, int nty, const uint8_t* __restrict__ mask,
    const int32_t* __restrict__ phi, const int32_t* __restrict__ phi_comp, const int* __restrict__ slot,
    const double* __restrict__ wtab_g, const R* __restrict__ vraw, R* __restrict__ vout) {
    using H = HaloGeom<DIM, RP>;
    constexpr int W = 2 * RP + 1;
    uint32_t* const key = sm.key;
    R* const vv = sm.vv;
    short* const list = sm.list;
    int* const zoff = sm.zoff;
    int* const yoff = sm.yoff;
    int* const xoff = sm.xoff;
    int& nlist = sm.nlist;
    __syncthreads();   // the workgroup's previous tile may still read the LDS
    const int tx = tile % ntx, ty = (tile / ntx) % nty, tz = tile / (ntx * nty);
    const int x0t = tx * H::TX, y0t = ty * H::TY, z0t = tz * H::TZ;
    if (threadIdx.x == 0) nlist = 0;
    for (int e = threadIdx.x; e < H::SZ; e += blockDim.x) zoff[e] = wrapi(z0t + e - H::HZ, d.Nz) * d.Ny * d.Nx;
    for (int e = threadIdx.x; e < H::SY; e += blockDim.x) yoff[e] = wrapi(y0t + e - RP, d.Ny) * d.Nx;
    for (int e = threadIdx.x; e < H::SX; e += blockDim.x) xoff[e] = wrapi(x0t + e - RP, d.Nx);
    __syncthreads();
    // key = (low << 16 | high) for interface cells whose two labels differ,
    // all ones otherwise (labels below 2^15, so the packed pair fits 31 bits
    // and never equals all ones); the velocity is staged with the cell's
    // orientation in the pair, so a window cell is one key compare.
    // Four staged cells per thread are loaded at once: mask, labels and band
    // slot together, then the velocity of the interface ones.
    constexpr int INFL = 4;
    for (int e0 = threadIdx.x; e0 < H::NS; e0 += INFL * blockDim.x) {
        int j[INFL], sl[INFL];
        uint32_t p[INFL], q[INFL];
        uint8_t mk[INFL];
#pragma unroll
        for (int k = 0; k < INFL; ++k) {
            const int e = e0 + k * blockDim.x;
            const int ee = e < H::NS ? e : 0;
            const int lx = ee % H::SX, ly = (ee / H::SX) % H::SY, lz = ee / (H::SX * H::SY);
            j[k] = zoff[lz] + yoff[ly] + xoff[lx];
        }
#pragma unroll
        for (int k = 0; k < INFL; ++k) {
            mk[k] = mask[j[k]];
            p[k] = uint32_t(phi[j[k]]);
            q[k] = uint32_t(phi_comp[j[k]]);
            sl[k] = slot[j[k]];
        }
        R v[INFL];
#pragma unroll
        for (int k = 0; k < INFL; ++k) v[k] = mk[k] ? vraw[sl[k]] : R(0);
#pragma unroll
        for (int k = 0; k < INFL; ++k) {
            const int e = e0 + k * blockDim.x;
            if (e < H::NS) {
                uint32_t kk = 0xFFFFFFFFu;
                R vs = v[k];
                if (mk[k] && p[k] != q[k]) {
                    const uint32_t lo = p[k] < q[k] ? p[k] : q[k], hi = p[k] < q[k] ? q[k] : p[k];
                    kk = (lo << 16) | hi;
                    // staged with its orientation: w (-v) == -(w v) exactly
                    vs = p[k] == lo ? vs : -vs;
                }
                key[e] = kk;
                vv[e] = vs;
            }
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < H::NT; e += blockDim.x) {
        const int lx = e % H::TX, ly = (e / H::TX) % H::TY, lz = e / (H::TX * H::TY);
        const int x = x0t + lx, y = y0t + ly, z = z0t + lz;
        if (x < d.Nx && y < d.Ny && z < d.Nz && mask[lin(d, z, y, x)]) list[atomicAdd(&nlist, 1)] = short(e);
    }
    __syncthreads();
    const int n = nlist;
    typename UBits<R>::T mb = 0;
    for (int qq = threadIdx.x; qq < n; qq += blockDim.x) {
        const int e = list[qq];
        const int lx = e % H::TX, ly = (e / H::TX) % H::TY, lz = e / (H::TX * H::TY);
        const int c = ((lz + H::HZ) * H::SY + (ly + RP)) * H::SX + (lx + RP);
        const uint32_t ki = key[c];
        R out = vv[c];
        if (ki != 0xFFFFFFFFu) {
            const uint32_t pair = ki;
            // orientation of the centre: + when its label is the pair's low one
            const int ic = lin(d, z0t + lz, y0t + ly, x0t + lx);
            const R oi = phi[ic] < phi_comp[ic] ? R(1) : R(-1);
            double num = 0.0, den = 0.0;
            const int nzw = DIM == 3 ? W : 1;
            for (int jz = 0; jz < nzw; ++jz) {
                const int dz = DIM == 3 ? jz - RP : 0;
                for (int jy = 0; jy < W; ++jy) {
                    const int dy = jy - RP;
                    const int rowb = ((lz + H::HZ + dz) * H::SY + (ly + RP + dy)) * H::SX + lx;
                    // the row's labels and velocities first, then the sums in
                    // window order; a miss adds +-0, which leaves the sums
                    // unchanged (they never hold -0)
                    uint32_t kj[W];
                    R vj[W];
#pragma unroll
                    for (int jx = 0; jx < W; ++jx) { kj[jx] = key[rowb + jx]; vj[jx] = vv[rowb + jx]; }
                    const double* wr = wtab_g + dy * dy + dz * dz;
#pragma unroll
                    for (int jx = 0; jx < W; ++jx) {
                        const int dxx = jx - RP;
                        const bool hit = kj[jx] == pair;
                        const double w = wr[dxx * dxx];
                        const double t = w * double(vj[jx]);
                        num = add_if(hit, t, num);
                        den = add_if(hit, w, den);
                    }
                }
            }
            if (den > 0.0) out = oi * R(num / den);
        }
        const int i = lin(d, z0t + lz, y0t + ly, x0t + lx);
        vout[slot[i]] = out;
        const auto b = abs_bits(out);
        mb = mb < b ? b : mb;
    }
    return mb;
}

template <class R, int DIM, int RP>
__global__ void __launch_bounds__(PROJ_THREADS)
k_project_tiled(Dims d, int ntx, int nty, int ntz, const uint8_t* __restrict__ mask,
                const int32_t* __restrict__ phi, const int32_t* __restrict__ phi_comp,
                const int* __restrict__ slot, const double* __restrict__ wtab_g,
                const R* __restrict__ vraw, R* __restrict__ vout,
                typename UBits<R>::T* __restrict__ vmax_bits) {
    __shared__ ProjTileSmem<R, DIM, RP> sm;
    __shared__ typename UBits<R>::T bmax;
    const int ntiles = ntx * nty * ntz;
    const int tile = xcd_tile(ntiles);
    if (tile >= ntiles) return;
    if (threadIdx.x == 0) bmax = 0;
    const auto mb = project_tile<R, DIM, RP>(sm, tile, d, ntx, nty, mask, phi, phi_comp, slot, wtab_g, vraw, vout);
    block_max_to(mb, bmax, vmax_bits);
}

template <class R>
struct ScoreCache {
    R* gain;
    int32_t* best;
    uint8_t* valid;     // null: no cache
    uint8_t* dirty;
};

__device__ __forceinline__ void mark_dirty_box(const Dims& d, int dim3, int cell, int r, uint8_t* dirty) {
    if (!dirty) return;
    int z0, y0, x0;
    decode(d, cell, z0, y0, x0);
    const int rz = dim3 ? r : 0;
    for (int dz = -rz; dz <= rz; ++dz) {
        const int z = wrapi(z0 + dz, d.Nz);
        for (int dy = -r; dy <= r; ++dy) {
            const int y = wrapi(y0 + dy, d.Ny);
            for (int ddx = -r; ddx <= r; ++ddx) dirty[lin(d, z, y, wrapi(x0 + ddx, d.Nx))] = 1;
        }
    }
}

// Gaussian label scores over one cell's (2RS+1)^d window, read from the
// staged labels: acc[0] sums the weights of the resident label p, acc[1 + t]
// those of candidate label L[1 + t] for t < NC, each in window order.  NC is
// the wavefront's largest candidate count, so every lane walks the same
// accumulators; a lane with fewer candidates holds -1 there, which matches no
// label.
template <int NC, int DIM, int RS, class H>
__device__ __forceinline__ void remap_window(const uint16_t* lab, const double* __restrict__ wtab_g,
                                             int lz, int ly, int lx, int32_t p, const int32_t (&L)[7],
                                             double (&acc)[7]) {
    constexpr int W = 2 * RS + 1;
    const int nzw = DIM == 3 ? W : 1;
    for (int jz = 0; jz < nzw; ++jz) {
        const int dz = DIM == 3 ? jz - RS : 0;
        for (int jy = 0; jy < W; ++jy) {
            const int dy = jy - RS;
            const int rowb = ((lz + H::HZ + dz) * H::SY + (ly + RS + dy)) * H::SX + lx;
            const double* wr = wtab_g + dy * dy + dz * dz;
            int32_t v[W];
#pragma unroll
            for (int jx = 0; jx < W; ++jx) v[jx] = lab[rowb + jx];
#pragma unroll
            for (int jx = 0; jx < W; ++jx) {
                const int dxx = jx - RS;
                const double w = wr[dxx * dxx];
                acc[0] = add_if(v[jx] == p, w, acc[0]);
#pragma unroll
                for (int t = 0; t < NC; ++t) acc[1 + t] = add_if(v[jx] == L[1 + t], w, acc[1 + t]);
            }
        }
    }
}

template <int DIM, int RS>
struct RemapTileSmem {
    using H = HaloGeom<DIM, RS>;
    uint16_t lab[H::NS];
    short list[H::NT];
    short order[H::NT];      // the listed cells by candidate count
    uint8_t ncl[H::NT];      // candidate count of each listed cell
    int zoff[H::SZ], yoff[H::SY], xoff[H::SX];   // wrapped offsets of the staged box
    int nlist;
    int ncnt[8];             // listed cells per candidate count, then their next slot
};

// remap scores of every interface cell of tile `tile` (labels < 2^16), run by
// the whole workgroup; BYCOUNT: the cells in order of their candidate count
// (pays where many workgroups share a CU; a workgroup that walks an
// instance's tiles alone waits on the extra barriers instead)
template <class R, int DIM, int RS, bool BYCOUNT = true>
__device__ __forceinline__ void remap_score_tile(
    RemapTileSmem<DIM, RS>& sm, int tile, const Dims& d, int ntx, int nty, const uint8_t* __restrict__ mask,
    const int32_t* __restrict__ phi, const int* __restrict__ slot, const double* __restrict__ wtab_g,
    int32_t* __restrict__ phi_comp, uint8_t* __restrict__ pflag, R* __restrict__ gain_s, const ScoreCache<R>& sc) {
    using H = HaloGeom<DIM, RS>;
    uint16_t* const lab = sm.lab;
    short* const list = sm.list;
    int* const zoff = sm.zoff;
    int* const yoff = sm.yoff;
    int* const xoff = sm.xoff;
    int& nlist = sm.nlist;
    __syncthreads();   // the workgroup's previous tile may still read the LDS
    const int tx = tile % ntx, ty = (tile / ntx) % nty, tz = tile / (ntx * nty);
    const int x0t = tx * H::TX, y0t = ty * H::TY, z0t = tz * H::TZ;
    if (threadIdx.x == 0) nlist = 0;
    __syncthreads();
    // interface cells whose window is unchanged since their scores were
    // computed take the cached scores; the others are listed for evaluation
    for (int e = threadIdx.x; e < H::NT; e += blockDim.x) {
        const int lx = e % H::TX, ly = (e / H::TX) % H::TY, lz = e / (H::TX * H::TY);
        const int x = x0t + lx, y = y0t + ly, z = z0t + lz;
        if (!(x < d.Nx && y < d.Ny && z < d.Nz)) continue;
        const int i = lin(d, z, y, x);
        if (!mask[i]) continue;
        if (sc.valid && sc.valid[i] && !sc.dirty[i]) {
            const int32_t bl = sc.best[i];
            const R gain = sc.gain[i];
            phi_comp[i] = bl;
            const int a = slot[i];
            pflag[a] = (phi[i] != bl && gain > Eps<R>::gain) ? 1 : 0;
            gain_s[a] = gain;
        } else {
            list[atomicAdd(&nlist, 1)] = short(e);
        }
    }
    __syncthreads();
    const int n = nlist;
    if (n == 0) return;
    for (int e = threadIdx.x; e < H::SZ; e += blockDim.x) zoff[e] = wrapi(z0t + e - H::HZ, d.Nz) * d.Ny * d.Nx;
    for (int e = threadIdx.x; e < H::SY; e += blockDim.x) yoff[e] = wrapi(y0t + e - RS, d.Ny) * d.Nx;
    for (int e = threadIdx.x; e < H::SX; e += blockDim.x) xoff[e] = wrapi(x0t + e - RS, d.Nx);
    __syncthreads();
    for (int e = threadIdx.x; e < H::NS; e += blockDim.x) {
        const int lx = e % H::SX, ly = (e / H::SX) % H::SY, lz = e / (H::SX * H::SY);
        lab[e] = uint16_t(phi[zoff[lz] + yoff[ly] + xoff[lx]]);
    }
    __syncthreads();
    const int nface = DIM == 3 ? 6 : 4;
    // candidates of a listed cell: its distinct unlike face labels, in face order
    auto candidates = [&](int e, int32_t& p, int32_t (&cand)[6]) {
        const int lx = e % H::TX, ly = (e / H::TX) % H::TY, lz = e / (H::TX * H::TY);
        const int c = ((lz + H::HZ) * H::SY + (ly + RS)) * H::SX + (lx + RS);
        p = lab[c];
        const int fo[6] = {-1, 1, -H::SX, H::SX, -H::SX * H::SY, H::SX * H::SY};
        int nc = 0;
        for (int k = 0; k < nface; ++k) {
            const int32_t v = lab[c + fo[k]];
            if (v == p) continue;
            bool known = false;
            for (int t = 0; t < nc; ++t) known |= cand[t] == v;
            if (!known) cand[nc++] = v;
        }
        return nc;
    };
    // the listed cells in order of their candidate count (counting sort), so
    // that the lanes of a wavefront walk the same number of accumulators
    // (a wavefront walks its largest count; most cells have one candidate)
    if constexpr (BYCOUNT) {
        if (threadIdx.x < 8) sm.ncnt[threadIdx.x] = 0;
        __syncthreads();
        for (int q = threadIdx.x; q < n; q += blockDim.x) {
            int32_t p, cand[6];
            const int nc = candidates(list[q], p, cand);
            sm.ncl[q] = uint8_t(nc);
            atomicAdd(&sm.ncnt[nc], 1);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int o = 0;
            for (int k = 0; k < 8; ++k) { const int c = sm.ncnt[k]; sm.ncnt[k] = o; o += c; }
        }
        __syncthreads();
        for (int q = threadIdx.x; q < n; q += blockDim.x) sm.order[atomicAdd(&sm.ncnt[sm.ncl[q]], 1)] = list[q];
        __syncthreads();
    }
    for (int q = threadIdx.x; q < n; q += blockDim.x) {
        const int e = BYCOUNT ? sm.order[q] : list[q];
        const int lx = e % H::TX, ly = (e / H::TX) % H::TY, lz = e / (H::TX * H::TY);
        int32_t p, cand[6];
        const int nc = candidates(e, p, cand);
        // acc[0] = resident score, acc[1 + t] = support of candidate t; the
        // number of accumulators walked is the wavefront's largest nc + 1
        int32_t L[7];
        L[0] = p;
#pragma unroll
        for (int t = 0; t < 6; ++t) L[1 + t] = t < nc ? cand[t] : -1;
        int ncw = nc;
        for (int off = 32; off > 0; off >>= 1) {
            const int o = __shfl_xor(ncw, off);
            ncw = o > ncw ? o : ncw;
        }
        double acc[7] = {0, 0, 0, 0, 0, 0, 0};
        switch (ncw) {     // wavefront-uniform
        case 0: remap_window<0, DIM, RS, H>(lab, wtab_g, lz, ly, lx, p, L, acc); break;
        case 1: remap_window<1, DIM, RS, H>(lab, wtab_g, lz, ly, lx, p, L, acc); break;
        case 2: remap_window<2, DIM, RS, H>(lab, wtab_g, lz, ly, lx, p, L, acc); break;
        case 3: remap_window<3, DIM, RS, H>(lab, wtab_g, lz, ly, lx, p, L, acc); break;
        case 4: remap_window<4, DIM, RS, H>(lab, wtab_g, lz, ly, lx, p, L, acc); break;
        case 5: remap_window<5, DIM, RS, H>(lab, wtab_g, lz, ly, lx, p, L, acc); break;
        default: remap_window<6, DIM, RS, H>(lab, wtab_g, lz, ly, lx, p, L, acc); break;
        }
        const double resident = acc[0];
        double sup[6];
#pragma unroll
        for (int t = 0; t < 6; ++t) sup[t] = acc[1 + t];
        double best = -1.0;
        int32_t bl = p;
        for (int t = 0; t < nc; ++t) if (sup[t] > best) { best = sup[t]; bl = cand[t]; }
        const int i = lin(d, z0t + lz, y0t + ly, x0t + lx);
        phi_comp[i] = bl;
        const R gain = R(best) - R(resident);
        const int a = slot[i];
        pflag[a] = (p != bl && gain > Eps<R>::gain) ? 1 : 0;
        gain_s[a] = gain;
        if (sc.valid) {
            sc.best[i] = bl;
            sc.gain[i] = gain;
            sc.valid[i] = 1;
            sc.dirty[i] = 0;
        }
    }
}

template <class R, int DIM, int RS>
__global__ void __launch_bounds__(256, 8)
k_remap_score_tiled(Dims d, int ntx, int nty, int ntz, const uint8_t* __restrict__ mask,
                    const int32_t* __restrict__ phi, const int* __restrict__ slot,
                    const double* __restrict__ wtab_g, int32_t* __restrict__ phi_comp,
                    uint8_t* __restrict__ pflag, R* __restrict__ gain_s, ScoreCache<R> sc) {
    __shared__ RemapTileSmem<DIM, RS> sm;
    const int ntiles = ntx * nty * ntz;
    const int tile = xcd_tile(ntiles);
    if (tile >= ntiles) return;
    remap_score_tile<R, DIM, RS>(sm, tile, d, ntx, nty, mask, phi, slot, wtab_g, phi_comp, pflag, gain_s, sc);
}

} // namespace gpkd
