// LSD radix sort of (u64 key, u32 value) pairs and an exclusive scan, hand-written for
// gfx950.  Replaces the host-side thrust::sort_by_key of LBVHIndex::build
// (src/liblvr2/reconstruction/cuda/LBVHIndex.cu:138, 532).
//
// One pass = 8 key bits: (1) per-tile digit histograms, (2) one exclusive scan over the
// digit-major histogram table, (3) a stable scatter in which each 256-thread block ranks
// its 4096 keys with wave-level match masks (8 ballots per key) and per-wave digit counts.
#pragma once

#include <algorithm>

#include "common.hpp"

namespace lvrhip {

constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;
constexpr int RS_RADIX = 256;

__global__ __launch_bounds__(RS_THREADS) void rs_hist_kernel(const uint64_t* __restrict__ keys, int64_t n, int shift,
                                                              uint32_t* __restrict__ ghist, int ntiles)
{
    __shared__ uint32_t h[RS_RADIX];
    const int tid = threadIdx.x;
    h[tid] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
#pragma unroll 4
    for (int r = 0; r < RS_ITEMS; r++) {
        int64_t i = base + (int64_t)r * RS_THREADS + tid;
        if (i < n) atomicAdd(&h[(uint32_t)(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    ghist[(int64_t)tid * ntiles + blockIdx.x] = h[tid];
}

// One 4096-entry segment of an in-place exclusive scan by a 1024-thread block: entries
// [base, base + 4096) get carry + their exclusive prefix within the segment; returns the
// segment's total (on every thread).
constexpr int SCAN_SEG = 4096;
__device__ __forceinline__ uint32_t scan_segment(uint32_t* __restrict__ a, int64_t len, int64_t base, uint32_t carry,
                                                 uint32_t* wsum)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    uint32_t v[4];
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        int64_t i = base + (int64_t)tid * 4 + j;
        v[j] = i < len ? a[i] : 0u;
        s += v[j];
    }
    // inclusive wave scan of per-thread sums
    uint32_t incl = s;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(incl, o, 64);
        if (lane >= o) incl += y;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    uint32_t wpre = 0, tot = 0;
    for (int w = 0; w < 16; w++) {
        uint32_t x = wsum[w];
        if (w < wave) wpre += x;
        tot += x;
    }
    uint32_t run = carry + wpre + incl - s;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        int64_t i = base + (int64_t)tid * 4 + j;
        if (i < len) a[i] = run;
        run += v[j];
    }
    __syncthreads();  // wsum is reused by the next segment
    return tot;
}

// In-place exclusive scan of a[0..len) with one 1024-thread block, one segment per step.
__global__ __launch_bounds__(1024) void scan_exclusive_kernel(uint32_t* __restrict__ a, int64_t len,
                                                              uint32_t* __restrict__ total)
{
    __shared__ uint32_t wsum[16];
    uint32_t carry = 0;
    for (int64_t base = 0; base < len; base += SCAN_SEG) carry += scan_segment(a, len, base, carry, wsum);
    if (threadIdx.x == 0 && total) *total = carry;
}

// Multi-block form for long tables (the radix sort's digit table: 256 entries per 4096 keys):
// (1) each block sums its segment, (2) one block scans the segment sums, (3) each block scans
// its segment from its carry.
__global__ __launch_bounds__(1024) void scan_segsum_kernel(const uint32_t* __restrict__ a, int64_t len,
                                                           uint32_t* __restrict__ sums)
{
    __shared__ uint32_t wsum[16];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t base = (int64_t)blockIdx.x * SCAN_SEG;
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        int64_t i = base + (int64_t)tid * 4 + j;
        s += i < len ? a[i] : 0u;
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) wsum[wave] = s;
    __syncthreads();
    if (tid == 0) {
        uint32_t t = 0;
        for (int w = 0; w < 16; w++) t += wsum[w];
        sums[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(1024) void scan_segapply_kernel(uint32_t* __restrict__ a, int64_t len,
                                                             const uint32_t* __restrict__ carry)
{
    __shared__ uint32_t wsum[16];
    scan_segment(a, len, (int64_t)blockIdx.x * SCAN_SEG, carry[blockIdx.x], wsum);
}

// Entries the radix sort's table must hold for n keys: the digit table, then the segment sums.
inline size_t rs_hist_size(int64_t n)
{
    const int64_t len = (int64_t)RS_RADIX * ceil_div(std::max<int64_t>(n, 1), RS_TILE);
    return (size_t)(len + ceil_div(len, SCAN_SEG));
}

// Exclusive scan of a[0..len) in place; sums holds ceil(len / 4096) entries.
inline void scan_exclusive(uint32_t* a, int64_t len, uint32_t* sums, hipStream_t st)
{
    const int64_t nseg = ceil_div(len, SCAN_SEG);
    if (nseg <= 4) {
        scan_exclusive_kernel<<<1, 1024, 0, st>>>(a, len, nullptr);
        LVRHIP_LAUNCH_CHECK();
        return;
    }
    scan_segsum_kernel<<<(unsigned)nseg, 1024, 0, st>>>(a, len, sums);
    LVRHIP_LAUNCH_CHECK();
    scan_exclusive_kernel<<<1, 1024, 0, st>>>(sums, nseg, nullptr);
    LVRHIP_LAUNCH_CHECK();
    scan_segapply_kernel<<<(unsigned)nseg, 1024, 0, st>>>(a, len, sums);
    LVRHIP_LAUNCH_CHECK();
}

__global__ __launch_bounds__(RS_THREADS) void rs_scatter_kernel(const uint64_t* __restrict__ kin,
                                                                 const uint32_t* __restrict__ vin,
                                                                 uint64_t* __restrict__ kout,
                                                                 uint32_t* __restrict__ vout, int64_t n, int shift,
                                                                 const uint32_t* __restrict__ goff, int ntiles)
{
    __shared__ uint32_t off[RS_RADIX];
    __shared__ uint32_t wcnt[RS_THREADS / 64][RS_RADIX];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    off[tid] = goff[(int64_t)tid * ntiles + blockIdx.x];
#pragma unroll
    for (int w = 0; w < RS_THREADS / 64; w++) wcnt[w][tid] = 0;
    __syncthreads();
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
    for (int r = 0; r < RS_ITEMS; r++) {
        const int64_t i = base + (int64_t)r * RS_THREADS + tid;
        const bool valid = i < n;
        uint64_t key = valid ? kin[i] : 0ull;
        uint32_t val = valid ? vin[i] : 0u;
        uint32_t d = (uint32_t)(key >> shift) & 255u;
        uint64_t peers = ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            bool bit = (d >> b) & 1u;
            uint64_t bal = ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        const int rank = __popcll(peers & lt_mask);
        if (valid && (peers & lt_mask) == 0) wcnt[wave][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pre = off[d] + (uint32_t)rank;
            for (int w = 0; w < wave; w++) pre += wcnt[w][d];
            kout[pre] = key;
            vout[pre] = val;
        }
        __syncthreads();
        uint32_t add = 0;
#pragma unroll
        for (int w = 0; w < RS_THREADS / 64; w++) {
            add += wcnt[w][tid];
            wcnt[w][tid] = 0;
        }
        off[tid] += add;
        __syncthreads();
    }
}

// Sort n pairs by key bits [0, 64).  tmp buffers must hold n entries; hist must hold
// rs_hist_size(n) entries.  The result ends in (keys, vals).
// Stable sort of (key, value) pairs by key bits [begin_bit, end_bit), 8 bits per pass.
inline void radix_sort_pairs(uint64_t* keys, uint32_t* vals, uint64_t* ktmp, uint32_t* vtmp, int64_t n,
                             uint32_t* hist, hipStream_t st, int end_bit = 64, int begin_bit = 0)
{
    if (n <= 1) return;
    const int ntiles = (int)ceil_div(n, RS_TILE);
    uint64_t* kin = keys;
    uint32_t* vin = vals;
    uint64_t* kout = ktmp;
    uint32_t* vout = vtmp;
    int passes = 0;
    for (int shift = begin_bit; shift < end_bit; shift += 8, passes++) {
        rs_hist_kernel<<<ntiles, RS_THREADS, 0, st>>>(kin, n, shift, hist, ntiles);
        LVRHIP_LAUNCH_CHECK();
        scan_exclusive(hist, (int64_t)RS_RADIX * ntiles, hist + (size_t)RS_RADIX * ntiles, st);
        rs_scatter_kernel<<<ntiles, RS_THREADS, 0, st>>>(kin, vin, kout, vout, n, shift, hist, ntiles);
        LVRHIP_LAUNCH_CHECK();
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    if (passes & 1) {
        LVRHIP_CHECK(hipMemcpyAsync(keys, kin, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, st));
        LVRHIP_CHECK(hipMemcpyAsync(vals, vin, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
    }
}

}  // namespace lvrhip
