// gpk_sort.hpp — stable LSD radix sort of 64-bit keys (optionally with int
// payloads) and a rank sort within segments, for the flip and remap
// transactions.
//
// One pass sorts by an 8-bit digit (or narrower on the last pass):
//   k_rs_count   per 4096-key block, a 256-bin histogram of the digit, stored
//                digit-major (count[d * nblocks + b]);
//   k_cmp_scan   exclusive prefix of those counts (blocks in order within a
//                digit, digits in order): the first output slot of every
//                (digit, block);
//   k_rs_scatter each block walks its keys in order, 256 at a time; a key's
//                slot is its (digit, block) base + the earlier keys of its
//                digit in this block: same-digit lanes of a wavefront found
//                by one ballot per digit bit, wavefronts in order through
//                per-digit LDS counters.
// Every pass keeps the input order among equal digits, so the passes from
// the lowest digit up sort stably by the whole bit range.
#pragma once

#include "gpk_device.hpp"

namespace gpkd {

constexpr int RS_T = 256, RS_ITEMS = 16, RS_CHUNK = RS_T * RS_ITEMS, RS_BINS = 256;
static_assert(RS_T == RS_BINS, "one thread per digit bin");

__global__ void __launch_bounds__(RS_T) k_rs_count(int n, const unsigned long long* __restrict__ keys, int shift,
                                                   int bits, int* __restrict__ cnt) {
    __shared__ int h[RS_BINS];
    h[threadIdx.x] = 0;
    __syncthreads();
    const unsigned long long mask = (1ull << bits) - 1ull;
    const int base = blockIdx.x * RS_CHUNK;
    // the thread's keys loaded together, then counted
    unsigned long long kr[RS_ITEMS];
#pragma unroll
    for (int it = 0; it < RS_ITEMS; ++it) {
        const int i = base + it * RS_T + int(threadIdx.x);
        kr[it] = i < n ? keys[i] : 0ull;
    }
#pragma unroll
    for (int it = 0; it < RS_ITEMS; ++it) {
        const int i = base + it * RS_T + int(threadIdx.x);
        if (i < n) atomicAdd(&h[int((kr[it] >> shift) & mask)], 1);
    }
    __syncthreads();
    cnt[threadIdx.x * gridDim.x + blockIdx.x] = h[threadIdx.x];
}

__global__ void __launch_bounds__(RS_T) k_rs_scatter(int n, const unsigned long long* __restrict__ kin,
                                                     unsigned long long* __restrict__ kout,
                                                     const int* __restrict__ vin, int* __restrict__ vout,
                                                     int shift, int bits, const int* __restrict__ off) {
    constexpr int NW = RS_T / 64;
    __shared__ int run[RS_BINS];          // next slot of each digit for this block
    __shared__ int wbase[NW][RS_BINS];    // per wavefront: its keys of a digit, then their first slot
    run[threadIdx.x] = off[threadIdx.x * gridDim.x + blockIdx.x];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const unsigned long long lt = (1ull << lane) - 1ull;
    const unsigned long long mask = (1ull << bits) - 1ull;
    const int base = blockIdx.x * RS_CHUNK;
    // the thread's keys and payloads loaded together before the ranked passes
    unsigned long long kr[RS_ITEMS];
    int vr[RS_ITEMS];
#pragma unroll
    for (int it = 0; it < RS_ITEMS; ++it) {
        const int i = base + it * RS_T + int(threadIdx.x);
        kr[it] = i < n ? kin[i] : 0ull;
        vr[it] = (vin && i < n) ? vin[i] : 0;
    }
#pragma unroll
    for (int it = 0; it < RS_ITEMS; ++it) {
        if (base + it * RS_T >= n) break;          // uniform over the block
        for (int q = threadIdx.x; q < NW * RS_BINS; q += RS_T) (&wbase[0][0])[q] = 0;
        __syncthreads();
        const int i = base + it * RS_T + int(threadIdx.x);
        const bool valid = i < n;
        unsigned long long key = 0ull;
        int d = 0;
        if (valid) {
            key = kr[it];
            d = int((key >> shift) & mask);
        }
        unsigned long long peers = __ballot(valid);
        for (int b = 0; b < bits; ++b) {
            const bool set = (d >> b) & 1;
            const unsigned long long bb = __ballot(valid && set);
            peers &= set ? bb : ~bb;
        }
        const int rank = __popcll(peers & lt);
        if (valid && rank == 0) wbase[w][d] = __popcll(peers);
        __syncthreads();
        {
            const int dg = threadIdx.x;
            int acc = run[dg];
            for (int ww = 0; ww < NW; ++ww) {
                const int c = wbase[ww][dg];
                wbase[ww][dg] = acc;
                acc += c;
            }
            run[dg] = acc;
        }
        __syncthreads();
        if (valid) {
            const int pos = wbase[w][d] + rank;
            kout[pos] = key;
            if (vin) vout[pos] = vr[it];
        }
        __syncthreads();
    }
}

// Within each segment [seg_start, seg_end] of a list already sorted by its
// segment key: order the entries by sec[item] ascending, ties by their
// position (so a stable sort by the segment key then sec, input order last).
// out[seg_start + rank] = item; one thread per entry, O(segment length).
__global__ void __launch_bounds__(256) k_seg_rank(int n, const int* __restrict__ item,
                                                  const unsigned long long* __restrict__ sec,
                                                  const int* __restrict__ seg_start,
                                                  const int* __restrict__ seg_end, int* __restrict__ out) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int e = item[s];
    const unsigned long long k = sec[e];
    const int a = seg_start[s], b = seg_end[s];
    int rank = 0;
    for (int t = a; t <= b; ++t) {
        const unsigned long long kt = sec[item[t]];
        rank += (kt < k || (kt == k && t < s)) ? 1 : 0;
    }
    out[a + rank] = e;
}

} // namespace gpkd
