// lvrhip: C ABI over the gfx950 kernels (declarations and reference citations in
// include/lvrhip.h).  One handle = one point cloud resident on one device:
// Morton-sorted SoA points, the LBVH, and the normals once computed.
#include "common.hpp"
#include "radix_sort.hpp"
#include "lbvh.hpp"
#include "knn_tile.hpp"
#include "sinks.hpp"
#include "knn_bestfirst.hpp"
#include "adjoint.hpp"
#include "lattice.hpp"
#include "mesh.hpp"

#include "lvrhip.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

using namespace lvrhip;

namespace {

thread_local std::string g_last_error;

// Every device allocation of the library.  With LVRHIP_FAIL_ALLOC_AFTER=n in the environment
// it refuses, on the host, every allocation after the first n made since the variable took its
// current value; with LVRHIP_FAIL_ALLOC_ABOVE=b every single allocation of more than b bytes (a
// device with less memory).  A refusal returns hipErrorOutOfMemory as a full device does: callers
// can exercise the out-of-memory contract (the error text LargeScaleReconstruction.tcc:656-662
// matches) without filling a device.
hipError_t dev_malloc(void** p, size_t bytes)
{
    if (const char* v = std::getenv("LVRHIP_FAIL_ALLOC_ABOVE")) {
        if ((unsigned long long)bytes > std::strtoull(v, nullptr, 10)) return hipErrorOutOfMemory;
    }
    if (const char* v = std::getenv("LVRHIP_FAIL_ALLOC_AFTER")) {
        static std::mutex mu;
        static std::string seen;
        static long long made = 0;
        std::lock_guard<std::mutex> lk(mu);
        if (seen != v) {
            seen = v;
            made = 0;
        }
        if (made >= std::atoll(v)) return hipErrorOutOfMemory;
        made++;
    }
    return hipMalloc(p, bytes);
}

template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t cap = 0;
    void reserve(size_t n)
    {
        if (n <= cap) return;
        release();
        if (n) LVRHIP_CHECK(dev_malloc(reinterpret_cast<void**>(&p), n * sizeof(T)));
        cap = n;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    ~DBuf() { release(); }
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
};

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev)
    {
        LVRHIP_CHECK(hipGetDevice(&prev));
        if (prev != dev) LVRHIP_CHECK(hipSetDevice(dev));
    }
    ~DeviceGuard()
    {
        int cur = -1;
        if (hipGetDevice(&cur) == hipSuccess && cur != prev && prev >= 0) (void)hipSetDevice(prev);
    }
};

inline uint32_t host_fbits(float f)
{
    uint32_t b;
    std::memcpy(&b, &f, 4);
    return b;
}

// ---- small kernels of the host pipeline ---------------------------------------------------------

__global__ void leaf_area_kernel(const float4* __restrict__ lo, const float4* __restrict__ hi, int nleaf,
                                 int stride, int nsample, float* __restrict__ out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nsample) return;
    int64_t l = (int64_t)i * stride;
    if (l >= nleaf) l = nleaf - 1;
    float4 a = lo[l], b = hi[l];
    float e[3] = {b.x - a.x, b.y - a.y, b.z - a.z};
    // two largest extents
    float m0 = fmaxf(e[0], fmaxf(e[1], e[2]));
    float m2 = fminf(e[0], fminf(e[1], e[2]));
    float m1 = e[0] + e[1] + e[2] - m0 - m2;
    int cnt = __float_as_int(a.w);
    out[i] = cnt > 0 ? (m0 * m1) / (float)cnt : 0.0f;
}

__global__ void morton_queries_kernel(const float* __restrict__ q, int64_t m, const MortonFrame* __restrict__ f,
                                      uint64_t* __restrict__ keys, uint32_t* __restrict__ vals)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    keys[i] = morton63(*f, q[3 * i], q[3 * i + 1], q[3 * i + 2]);
    vals[i] = (uint32_t)i;
}

__global__ void gather_queries_kernel(const float* __restrict__ q, int64_t m, const uint32_t* __restrict__ order,
                                      float* __restrict__ qx, float* __restrict__ qy, float* __restrict__ qz)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    uint32_t j = order[i];
    qx[i] = q[3 * (int64_t)j];
    qy[i] = q[3 * (int64_t)j + 1];
    qz[i] = q[3 * (int64_t)j + 2];
}

// Ordered compaction of 0/1 flags (the next adaptive round's query list):
// per-4096 block counts, one scan of the block counts, then a block-local stable scatter.
constexpr int CF_THREADS = 256;
constexpr int CF_ITEMS = 16;
constexpr int CF_TILE = CF_THREADS * CF_ITEMS;

__global__ __launch_bounds__(CF_THREADS) void flag_block_sums_kernel(const uint32_t* __restrict__ flags, int64_t n,
                                                                     const uint32_t* __restrict__ n_dev,
                                                                     uint32_t* __restrict__ bsum)
{
    __shared__ uint32_t ws[CF_THREADS / kWave];
    if (n_dev) n = (int64_t)*n_dev < n ? (int64_t)*n_dev : n;
    const int64_t base = (int64_t)blockIdx.x * CF_TILE;
    uint32_t c = 0;
    for (int r = 0; r < CF_ITEMS; r++) {
        const int64_t i = base + (int64_t)r * CF_THREADS + threadIdx.x;
        if (i < n && flags[i]) c++;
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) c += shfl_xor(c, m);
    if (lane_id() == 0) ws[threadIdx.x / kWave] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < CF_THREADS / kWave; w++) t += ws[w];
        bsum[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(CF_THREADS) void compact_scatter_kernel(const uint32_t* __restrict__ flags, int64_t n,
                                                                     const uint32_t* __restrict__ n_dev,
                                                                     const uint32_t* __restrict__ boff,
                                                                     const uint32_t* __restrict__ qsel,
                                                                     uint32_t* __restrict__ out)
{
    __shared__ uint32_t wc[CF_THREADS / kWave];
    if (n_dev) n = (int64_t)*n_dev < n ? (int64_t)*n_dev : n;
    const int64_t base = (int64_t)blockIdx.x * CF_TILE;
    const int wave = threadIdx.x / kWave;
    uint32_t run = boff[blockIdx.x];
    for (int r = 0; r < CF_ITEMS; r++) {
        const int64_t i = base + (int64_t)r * CF_THREADS + threadIdx.x;
        const bool f = i < n && flags[i] != 0;
        const uint64_t m = ballot(f);
        if (lane_id() == 0) wc[wave] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t pre = 0, tot = 0;
        for (int w = 0; w < CF_THREADS / kWave; w++) {
            if (w < wave) pre += wc[w];
            tot += wc[w];
        }
        if (f) out[run + pre + (uint32_t)mbcnt(m)] = qsel ? qsel[i] : (uint32_t)i;
        run += tot;
        __syncthreads();
    }
}

// interpolateSurfaceNormals (AdaptiveKSearchSurface.tcc:312-327): the point's own normal
// plus its ki nearest (self included), summed in ascending (d^2, index) order like the
// reference's kSearch result, normalised, into a fresh array.  The kNN pass stored the ki list
// unordered, so G lanes per point recompute one neighbour's key each and rank the list in LDS,
// INTERP_POINTS points per block in turns of INTERP_THREADS / G; then one lane per point adds the
// normals in rank order (the ordered sum is sequential: a lane per point keeps the block's lanes
// busy where one lane of each group would leave the others idle).
constexpr int INTERP_THREADS = 256;
constexpr int INTERP_POINTS = 64;

template <int G>
__global__ __launch_bounds__(INTERP_THREADS) void interpolate_kernel(
    const float* __restrict__ nx, const float* __restrict__ ny, const float* __restrict__ nz,
    const float4* __restrict__ pts4, const uint32_t* __restrict__ ki_list, int kiw, int64_t n, float* __restrict__ ox,
    float* __restrict__ oy, float* __restrict__ oz)
{
    constexpr int PTS = INTERP_THREADS / G;  // points ranked per turn
    static_assert(INTERP_POINTS % PTS == 0 && INTERP_POINTS <= INTERP_THREADS, "whole turns");
    __shared__ uint64_t skey[INTERP_THREADS];
    __shared__ uint32_t sord[INTERP_POINTS][G];
    __shared__ int snv[INTERP_POINTS];
    const int grp = threadIdx.x / G, j = threadIdx.x % G, base = grp * G;
    const int64_t first = (int64_t)blockIdx.x * INTERP_POINTS;
    for (int turn = 0; turn < INTERP_POINTS / PTS; turn++) {
        const int lp = turn * PTS + grp;  // the point's slot in the block
        const int64_t i = first + lp;
        // positions with their point index in w (one 16-byte load per neighbour)
        const float4 q = i < n ? pts4[i] : make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(~0u));
        const bool valid = i < n && __float_as_uint(q.w) != ~0u;  // padding slots have no normal
        uint32_t sidx = 0;
        uint64_t key = ~0ull;
        bool real = false;
        if (valid && j < kiw) {
            sidx = ki_list[(size_t)i * kiw + j];
            if (sidx < n) {
                const float4 p = pts4[sidx];
                key = make_key(d2_contract(q.x, q.y, q.z, p.x, p.y, p.z), __float_as_uint(p.w));
                real = true;
            }
        }
        skey[threadIdx.x] = key;
        // entries that name a point rank first (their keys are distinct and below ~0)
        const int lane0 = (int)(threadIdx.x % kWave) - j;  // the group's first lane in the wave
        const uint64_t gmask = G == 64 ? ~0ull : ((1ull << G) - 1ull);
        const int nv = __popcll((ballot(real) >> lane0) & gmask);
        __syncthreads();
        if (real) {
            int rank = 0;
            for (int t = 0; t < kiw; t++) rank += skey[base + t] < key ? 1 : 0;
            sord[lp][rank] = sidx;
        }
        if (j == 0) snv[lp] = valid ? nv : -1;
        __syncthreads();
    }
    // one lane per point adds the normals in rank order, as the reference's loop does
    if (threadIdx.x < INTERP_POINTS) {
        const int lp = threadIdx.x;
        const int64_t i = first + lp;
        const int nv = snv[lp];
        if (nv >= 0) {
            float x = nx[i], y = ny[i], z = nz[i];
            for (int r = 0; r < nv; r++) {
                const uint32_t t = sord[lp][r];
                x += nx[t];
                y += ny[t];
                z += nz[t];
            }
            normalize3(x, y, z);
            ox[i] = x;
            oy[i] = y;
            oz[i] = z;
        }
    }
}

// interpolateSurfaceNormals for ki above 64: row i (input order) holds the ki nearest of point i
// in ascending key order (self included), so the sum runs in that order, after the point's own
// normal, exactly as interpolate_kernel's ranked sum does.  Output at the point's sorted slot.
__global__ void interpolate_rows_kernel(int64_t n, int k, const int32_t* __restrict__ idx,
                                        const uint32_t* __restrict__ rank, const float* __restrict__ nx,
                                        const float* __restrict__ ny, const float* __restrict__ nz,
                                        float* __restrict__ ox, float* __restrict__ oy, float* __restrict__ oz)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t si = rank[i];
    float x = nx[si], y = ny[si], z = nz[si];
    for (int j = 0; j < k; j++) {
        const int32_t id = idx[(size_t)i * k + j];
        if (id < 0) continue;
        const uint32_t s = rank[id];
        x += nx[s];
        y += ny[s];
        z += nz[s];
    }
    normalize3(x, y, z);
    ox[si] = x;
    oy[si] = y;
    oz[si] = z;
}

// the points in input order (xyz), from the sorted SoA
__global__ void input_order_points_kernel(const float4* __restrict__ pts4, const uint32_t* __restrict__ rank,
                                          int64_t n, float* __restrict__ q)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 p = pts4[rank[i]];
    q[3 * i] = p.x;
    q[3 * i + 1] = p.y;
    q[3 * i + 2] = p.z;
}

__global__ void gather_normals_kernel(const float* __restrict__ in, const uint32_t* __restrict__ pidx, int64_t n,
                                      float* __restrict__ nx, float* __restrict__ ny, float* __restrict__ nz)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = pidx[i];
    if (j == ~0u) {
        nx[i] = ny[i] = nz[i] = 0.0f;
        return;
    }
    nx[i] = in[3 * (int64_t)j];
    ny[i] = in[3 * (int64_t)j + 1];
    nz[i] = in[3 * (int64_t)j + 2];
}

__global__ void scatter_normals_kernel(const float* __restrict__ nx, const float* __restrict__ ny,
                                       const float* __restrict__ nz, const uint32_t* __restrict__ pidx, int64_t n,
                                       float* __restrict__ out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = pidx[i];
    if (j == ~0u) return;
    out[3 * (int64_t)j] = nx[i];
    out[3 * (int64_t)j + 1] = ny[i];
    out[3 * (int64_t)j + 2] = nz[i];
}

// Adaptive neighbourhood size per point, input order: min(kn << round, kmax).
__global__ void adaptive_k_kernel(const uint8_t* __restrict__ kround, const uint32_t* __restrict__ pidx, int64_t n,
                                  int kn, int64_t kmax, int32_t* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = pidx[i];
    if (j == ~0u) return;
    const int64_t k = (int64_t)kn << kround[i];
    out[j] = (int32_t)(k < kmax ? k : kmax);
}

// Autotuning sample: every `every`-th run of `run` consecutive real positions (sort order).
__global__ void sample_runs_kernel(const uint32_t* __restrict__ qreal, int64_t m, int run, int every,
                                   uint32_t* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int64_t pos = (i / run) * run * every + i % run;
    out[i] = qreal ? qreal[pos] : (uint32_t)pos;
}

// Sort each row's t keys (one wave per row) into (index, d^2); pad to k.
__global__ __launch_bounds__(kWave) void sort_rows_kernel(const uint64_t* __restrict__ keys, int64_t m, int k, int t,
                                                          int32_t* __restrict__ idx, float* __restrict__ d2)
{
    const int64_t r = blockIdx.x;
    if (r >= m) return;
    const uint64_t* row = keys + (size_t)r * k;
    for (int j = lane_id(); j < k; j += kWave) {
        if (j < t) {
            const uint64_t kj = row[j];
            int rank = 0;
            for (int i = 0; i < t; i++) rank += row[i] < kj ? 1 : 0;
            idx[(size_t)r * k + rank] = (int32_t)(uint32_t)kj;
            if (d2) d2[(size_t)r * k + rank] = bitsf((uint32_t)(kj >> 32));
        } else {
            idx[(size_t)r * k + j] = -1;
            if (d2) d2[(size_t)r * k + j] = INFINITY;
        }
    }
}

// radiusSearch from sorted kSearch rows: the neighbours with d^2 < r^2 are a prefix of each row;
// the rest become padding (-1 / +inf) and count[r] = the prefix length.
__global__ void radius_filter_kernel(int64_t m, int k, float r2, int32_t* __restrict__ idx, float* __restrict__ d2,
                                     int32_t* __restrict__ count)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= m) return;
    int found = 0;
    for (int j = 0; j < k; j++) {
        const size_t o = (size_t)r * k + j;
        const bool in = idx[o] >= 0 && d2[o] < r2;
        if (in) {
            found++;
        } else {
            idx[o] = -1;
            d2[o] = INFINITY;
        }
    }
    if (count) count[r] = found;
}

// getInterpolatedNormal (PointsetSurface.tcc:61-74) from sorted kSearch rows of t <= ki input
// indices; normals in sorted order, found through the inverse permutation.
__global__ void mean_normal_rows_kernel(int64_t m, int k, int t, const int32_t* __restrict__ idx,
                                        const uint32_t* __restrict__ rank, const float* __restrict__ nx,
                                        const float* __restrict__ ny, const float* __restrict__ nz, int ki,
                                        float* __restrict__ out)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= m) return;
    float rx = 0.0f, ry = 1.0f, rz = 0.0f;
    for (int j = 0; j < t; j++) {
        const int32_t id = idx[(size_t)r * k + j];
        if (id < 0) continue;  // padding (a cloud with NaN points)
        const uint32_t s = rank[id];
        float x = nx[s], y = ny[s], z = nz[s];
        if (!(x == 0.0f && y == 0.0f && z == 0.0f)) {
            const float l = sqrtf(x * x + y * y + z * z);
            x /= l;
            y /= l;
            z /= l;
        }
        rx += x;
        ry += y;
        rz += z;
    }
    const float c = (float)ki;
    out[3 * r] = rx / c;
    out[3 * r + 1] = ry / c;
    out[3 * r + 2] = rz / c;
}

// the centroid of each row's neighbours, for the tile kernel's rows (the input position stays when
// the row is empty)
__global__ void centroid_rows_kernel(int64_t m, int k, const int32_t* __restrict__ idx,
                                     const uint32_t* __restrict__ rank, const float4* __restrict__ pts,
                                     const float* __restrict__ q, float* __restrict__ out)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= m) return;
    float cx = 0.0f, cy = 0.0f, cz = 0.0f;
    int nk = 0;
    for (int j = 0; j < k; j++) {
        const int32_t id = idx[(size_t)r * k + j];
        if (id < 0) continue;
        const float4 p = pts[rank[id]];
        cx += p.x;
        cy += p.y;
        cz += p.z;
        nk++;
    }
    if (nk == 0) {
        out[3 * r] = q[3 * r];
        out[3 * r + 1] = q[3 * r + 1];
        out[3 * r + 2] = q[3 * r + 2];
        return;
    }
    const float c = (float)nk;
    out[3 * r] = cx / c;
    out[3 * r + 1] = cy / c;
    out[3 * r + 2] = cz / c;
}

// Best-first kNN launch for t <= BF_KMAX: the list length is the smallest of 8, 10, 16, 32 that
// holds t (10: the signed distance's usual kd).
// Query coordinates: SoA (qs = 1) or AoS (qs = 3, qy = qx + 1, qz = qx + 2).  With a queue the
// grid is min(chunks, resident) waves; spill holds spill_cap entries per lane for each of them.
template <class Out>
void launch_bf(const BvhView& bv, const float* qx, const float* qy, const float* qz, int qs, int64_t m, int t,
               uint64_t empty, const Out& out, hipStream_t st, unsigned long long* stats, uint32_t* queue,
               int64_t resident, uint2* spill, int spill_cap)
{
    if (m == 0) return;
    int64_t grid = ceil_div(m, BF_THREADS);
    if (queue) {
        grid = std::min(grid, resident);
        fill_u32_kernel<<<1, 256, 0, st>>>(queue, QUEUE_SEGS * QUEUE_STRIDE, 0u);
        LVRHIP_LAUNCH_CHECK();
    }
    const unsigned g = (unsigned)grid;
    if (t <= 8)
        bf_knn_kernel<8, Out><<<g, BF_THREADS, 0, st>>>(bv, qx, qy, qz, qs, m, t, empty, out, stats, queue, spill,
                                                        spill_cap);
    else if (t <= 10)
        bf_knn_kernel<10, Out><<<g, BF_THREADS, 0, st>>>(bv, qx, qy, qz, qs, m, t, empty, out, stats, queue, spill,
                                                         spill_cap);
    else if (t <= 16)
        bf_knn_kernel<16, Out><<<g, BF_THREADS, 0, st>>>(bv, qx, qy, qz, qs, m, t, empty, out, stats, queue, spill,
                                                         spill_cap);
    else
        bf_knn_kernel<32, Out><<<g, BF_THREADS, 0, st>>>(bv, qx, qy, qz, qs, m, t, empty, out, stats, queue, spill,
                                                         spill_cap);
    LVRHIP_LAUNCH_CHECK();
}

// The wave-per-query pass over list[0, *count) (the queries the lane pass left).
template <class Out>
void launch_bf_wave(const BvhView& bv, const float* qx, const float* qy, const float* qz, int qs,
                    const uint32_t* list, const uint32_t* count, int t, uint64_t empty, const Out& out,
                    hipStream_t st, unsigned long long* stats)
{
    const unsigned grid = BF_WAVE_BLOCKS;
    if (t <= 8)
        bf_knn_wave_kernel<8, Out><<<grid, BF_THREADS, 0, st>>>(bv, qx, qy, qz, qs, list, count, t, empty, out, stats);
    else if (t <= 10)
        bf_knn_wave_kernel<10, Out><<<grid, BF_THREADS, 0, st>>>(bv, qx, qy, qz, qs, list, count, t, empty, out,
                                                                 stats);
    else if (t <= 16)
        bf_knn_wave_kernel<16, Out><<<grid, BF_THREADS, 0, st>>>(bv, qx, qy, qz, qs, list, count, t, empty, out,
                                                                 stats);
    else
        bf_knn_wave_kernel<32, Out><<<grid, BF_THREADS, 0, st>>>(bv, qx, qy, qz, qs, list, count, t, empty, out,
                                                                 stats);
    LVRHIP_LAUNCH_CHECK();
}

// ki list without normals (used when ki > 2*kn, so the first normal round cannot supply it);
// unordered, like the one written by SinkNormals.
template <int P>
struct SinkKiOnly {
    static constexpr int kTargets = 2;
    uint32_t* out;
    int kw;
    uint32_t qpos_;
    __device__ __forceinline__ void begin(int, int, int64_t, uint32_t qpos, float, float, float) { qpos_ = qpos; }
    __device__ __forceinline__ void accept(int tau, int g, int, float, float, float, uint64_t, uint32_t sp, uint32_t* cnt)
    {
        if (tau != 0) return;
        const uint32_t slot = atomicAdd(&cnt[g], 1u);
        if (slot < (uint32_t)kw) out[(size_t)qpos_ * kw + slot] = sp;
    }
    __device__ __forceinline__ void finish(int, int, bool, uint32_t, float, float, float, const uint32_t*, const uint32_t*) {}
};

// hipEvent pairs around kernels on the handle's stream, accumulated per timer.
struct Profiler {
    bool on = false;
    double ms[LVRHIP_T_COUNT] = {};
    int64_t launches[LVRHIP_T_COUNT] = {};
    struct Open {
        int timer;
        hipEvent_t a, b;
    };
    std::vector<Open> open;
    std::vector<hipEvent_t> pool;

    hipEvent_t ev()
    {
        hipEvent_t e;
        if (!pool.empty()) {
            e = pool.back();
            pool.pop_back();
        } else {
            LVRHIP_CHECK(hipEventCreate(&e));
        }
        return e;
    }
    int start(int timer, hipStream_t st)
    {
        if (!on) return -1;
        Open o{timer, ev(), ev()};
        LVRHIP_CHECK(hipEventRecord(o.a, st));
        open.push_back(o);
        return (int)open.size() - 1;
    }
    void stop(int tok, hipStream_t st)
    {
        if (tok < 0) return;
        LVRHIP_CHECK(hipEventRecord(open[tok].b, st));
    }
    void collect()
    {
        for (auto& o : open) {
            LVRHIP_CHECK(hipEventSynchronize(o.b));
            float t = 0.0f;
            LVRHIP_CHECK(hipEventElapsedTime(&t, o.a, o.b));
            ms[o.timer] += t;
            launches[o.timer] += 1;
            pool.push_back(o.a);
            pool.push_back(o.b);
        }
        open.clear();
    }
    ~Profiler()
    {
        for (auto& o : open) {
            (void)hipEventDestroy(o.a);
            (void)hipEventDestroy(o.b);
        }
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

// With a work queue the grid is capped at `resident` waves (the ones the device holds at once):
// they drain the queue; without, one wave per run of tiles.
template <int P, template <int> class SinkT>
void launch_tile_p(const BvhView& bv, const QueryView& qv, const TileParams& prm, const SinkT<P>& sink,
                   hipStream_t st, int64_t resident)
{
    constexpr int Q = kWave / P;
    const int64_t ntiles = ceil_div(qv.nq, Q);
    int64_t grid = ceil_div(ntiles, prm.tiles_per_wave);
    if (grid == 0) return;
    if (prm.queue) {
        grid = std::min(grid, resident);
        fill_u32_kernel<<<1, 256, 0, st>>>(prm.queue, QUEUE_SEGS * QUEUE_STRIDE, 0u);
        LVRHIP_LAUNCH_CHECK();
    }
    knn_tile_kernel<P, SinkT<P>><<<(unsigned)grid, kWave, 0, st>>>(bv, qv, prm, sink);
    LVRHIP_LAUNCH_CHECK();
}

// ---- the no-allocation signed distances (lvrhip_distances_device_noalloc) -----------------------
// Its scratch lives in one file-scope device array, `scratch`, that every device holds from the
// moment the library's code object is loaded until the process ends.  Layout (byte offsets):
// redo flags and redo list (NA_MAX_QUERIES entries each), compaction block sums, two counters,
// the work queue, and the lane search's stack spill (NA_SPILL entries per lane for NA_WAVES
// waves).  Host code never uses the array's name: it asks hipGetSymbolAddress for the address
// on each device once (na_scratch) and passes that pointer to the kernels.
constexpr int64_t NA_MAX_QUERIES = 1 << 20;
constexpr int NA_SPILL = 16;
constexpr int64_t NA_WAVES = 2048;
constexpr size_t na_align(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t NA_FLAGS = 0;
constexpr size_t NA_LIST = na_align(NA_FLAGS + 4 * (size_t)NA_MAX_QUERIES);
constexpr size_t NA_BSUM = na_align(NA_LIST + 4 * (size_t)NA_MAX_QUERIES);
constexpr size_t NA_COUNT = na_align(NA_BSUM + 4 * (size_t)(NA_MAX_QUERIES / CF_TILE + 2));
constexpr size_t NA_QUEUE = na_align(NA_COUNT + 4 * 4);
constexpr size_t NA_SPILL_OFF = na_align(NA_QUEUE + 4 * (size_t)(QUEUE_SEGS * QUEUE_STRIDE));
constexpr size_t NA_BYTES = na_align(NA_SPILL_OFF + sizeof(uint2) * (size_t)NA_SPILL * BF_THREADS * NA_WAVES);
static_assert(NA_BYTES == LVRHIP_NOALLOC_SCRATCH_BYTES, "the header states the scratch size");

static __device__ __align__(256) uint4 scratch[NA_BYTES / sizeof(uint4)];

constexpr int NA_MAX_DEVICES = 64;
std::mutex g_na_mu;
char* g_na_ptr[NA_MAX_DEVICES] = {};

// The device address of `scratch` on the current device (looked up once per device).
char* na_scratch(int device)
{
    if (device < 0 || device >= NA_MAX_DEVICES) throw Error(LVRHIP_EINVAL, "device index out of range");
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (!g_na_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw Error(LVRHIP_EHIP, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
        g_na_ptr[device] = static_cast<char*>(p);
    }
    return g_na_ptr[device];
}

}  // namespace

struct lvrhip_surface {
    int device = 0;
    hipStream_t st = nullptr;
    int64_t n = 0;     // points
    int64_t npad = 0;  // sorted positions (each scan padded to a multiple of LEAF)
    int nscans = 1;
    std::vector<int64_t> h_off, h_pstart;  // per scan: input offset, sorted start (nscans + 1 entries)
    int64_t min_scan = 0;                  // smallest non-empty scan
    const float* src = nullptr;            // device AoS points the build reads
    DBuf<float> own_src;                   // the copy of host-given points
    int kn = 10, ki = 10, kd = 5, method = LVRHIP_METHOD_PCA;
    float flip[3] = {1e6f, 1e6f, 1e6f};
    bool scan_flips = false;
    int lanes_per_query = 8;
    int tiles_per_wave = 12;
    float r2_margin = 1.15f;
    bool merge_retry = false;  // rounds 2..5 in one four-target search
    // launch shape of the signed-distance searches: lattice vertices are far fewer than points
    // and vary more in cost per tile, so fewer consecutive tiles per wave balance better
    int sdf_tiles_per_wave = 2;
    float sdf_r2_margin = 1.3f;
    float sdf_grow_min = 2.0f;
    // kNN tile launches: waves take runs of tiles from a work queue (dynamic balance), the grid
    // capped at what the device holds at once
    bool use_queue = true;
    int64_t resident_waves = 4096;
    DBuf<uint32_t> d_queue;
    // searches with at most BF_KMAX neighbours (distances, kSearch, radius search, interpolated
    // normals at positions) run best-first unless the tile kernel is forced
    int search_mode = LVRHIP_SEARCH_AUTO;
    int64_t bf_resident = 4096;  // best-first waves the device holds at once
    DBuf<uint32_t> s_redo, s_redo_list;
    DBuf<uint2> s_spill;  // the lane search's stack entries beyond its LDS part

    DBuf<float> px, py, pz;
    DBuf<float4> pts4;  // the same points as (x, y, z, original index bits): one load per point
    DBuf<uint32_t> pidx, rank, qreal;
    DBuf<int64_t> d_off, d_pstart;
    DBuf<uint16_t> scan_of;
    DBuf<int> roots;
    DBuf<float> d_flips;
    DBuf<uint32_t> sbox;
    DBuf<MortonFrame> frames;
    DBuf<float> area;  // median leaf area per point: the first radius guess
    DBuf<float4> leaf_lo, leaf_hi, nlo, nhi;
    DBuf<float4> leaf_agg;  // per leaf: its point sums about the box centre (3 float4)
    DBuf<int> fr_ref;       // per scan: the FRONTIER top-level subtrees (frontier_kernel)
    DBuf<float4> fr_lo, fr_hi;
    DBuf<BvhNode> nodes;
    DBuf<BvhNode4> nodes4;
    DBuf<int> leaf_parent;
    DBuf<float4> ru_lo[2], ru_hi[2];  // the refit's range-union levels, in turn
    int nleaf = 0;

    DBuf<float> nx, ny, nz, tx, ty, tz;
    bool has_normals = false;
    bool interpolated = false;  // nx/ny/nz hold interpolated normals
    // reverse pass
    DBuf<float> g_p, g_n, adj;
    DBuf<uint32_t> a_flags, a_sel, a_list;
    DBuf<uint32_t> ki_list;
    int kiw = 0;
    bool ki_valid = false;
    DBuf<uint8_t> kround;
    DBuf<float> mom;  // per position: moments of its k-set about the point (9 floats)
    DBuf<uint32_t> s_bsum;
    // device counters, allocated and zeroed with the handle: [r] = queries of adaptive round r
    // (r = 2..5), [6] temporary list lengths of the merged rounds and the adjoint, [7] the
    // lattice's, [8] queries a best-first search left to the tile kernel, [9] / [10] the mesh's
    // face and vertex counts, [11] the most moves of one mesh vertex, [12] a move round's vertices
    static constexpr int CNT_SLOTS = 16;
    DBuf<uint32_t> d_cnt;
    uint32_t* counter(int i) const
    {
        if (!d_cnt.p || i < 0 || i >= CNT_SLOTS) throw Error(LVRHIP_ESTATE, "internal: device counter unavailable");
        return d_cnt.p + i;
    }
    int64_t nq_first = 0;  // queries of the first round of the last normals pass
    int64_t n_round[5] = {0, 0, 0, 0, 0};
    DBuf<int> err;
    Profiler prof;
    // work counters, KS_GROUPS groups x KS_COUNT: normals rounds 1..5, merged later rounds,
    // SDF, other searches (kSearch, the ki list, the adjoint)
    static constexpr int KS_GROUPS = 8;
    DBuf<unsigned long long> kstats;

    hipStream_t gst = nullptr;  // private stream for graphs when the handle's stream is the null stream
    bool own_gst = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    hipStream_t graph_stream = nullptr;

    uint32_t* h_pin = nullptr;  // pinned host words for readbacks that must not stage (hipHostMalloc)

    ~lvrhip_surface()
    {
        if (h_pin) (void)hipHostFree(h_pin);
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        if (own_gst && gst) (void)hipStreamDestroy(gst);
    }

    unsigned long long* kstats_ptr(int group)
    {
        if (!prof.on) return nullptr;
        if (!kstats.p) {
            kstats.reserve(KS_GROUPS * KS_COUNT);
            LVRHIP_CHECK(hipMemsetAsync(kstats.p, 0, KS_GROUPS * KS_COUNT * sizeof(unsigned long long), st));
        }
        return kstats.p + group * KS_COUNT;
    }

    // scratch
    DBuf<uint64_t> s_k0, s_k1;
    DBuf<uint32_t> s_v0, s_v1, s_hist, s_flags, s_list0, s_list1, s_spos, s_ord, s_sample;
    DBuf<float> s_qx, s_qy, s_qz, s_f;
    DBuf<float> s_hint;  // per sorted point: the k-th d^2 of its last adaptive round
    DBuf<int32_t> s_idx;

    BvhView view() const
    {
        BvhView v;
        v.px = px.p;
        v.py = py.p;
        v.pz = pz.p;
        v.pidx = pidx.p;
        v.rank = rank.p;
        v.pts4 = pts4.p;
        v.leaf_lo = leaf_lo.p;
        v.leaf_hi = leaf_hi.p;
        v.leaf_agg = leaf_agg.p;
        v.nodes = nodes.p;
        v.nodes4 = nodes4.p;
        v.nleaf = nleaf;
        v.n = (uint32_t)npad;
        v.roots = nscans > 1 ? roots.p : nullptr;
        v.fr_ref = nleaf > 0 ? fr_ref.p : nullptr;
        v.fr_lo = fr_lo.p;
        v.fr_hi = fr_hi.p;
        v.scan_of = nscans > 1 ? scan_of.p : nullptr;
        return v;
    }

    TileParams params(int t0, int t1, int stats_group) const
    {
        TileParams prm{};
        prm.t0 = t0;
        prm.t1 = t1;
        prm.tiles_per_wave = tiles_per_wave;
        prm.r2_init = 0.0f;
        prm.area_pp = area.p;
        prm.r2_margin = r2_margin;
        prm.grow_min = 1.01f;
        prm.err = err.p;
        prm.stats = nullptr;
        (void)stats_group;
        return prm;
    }

    // The round-1 query list: every real sorted position (NULL = identity, one scan).
    const uint32_t* all_queries() const { return nscans > 1 ? qreal.p : nullptr; }

    // A neighbourhood size as the reference's kSearch returns it: at most the points of the
    // scan.  Scans of a multi-scan cloud must each hold at least k points.
    int clamp_k(int64_t k) const
    {
        if (nscans == 1) return (int)std::min<int64_t>(k, n);
        if (k > min_scan)
            throw Error(LVRHIP_EINVAL, "every scan needs at least " + std::to_string(k) +
                                           " points for this k (smallest scan: " + std::to_string(min_scan) + ")");
        return (int)k;
    }

    void read_err_and_counts()
    {
        int e = 0;
        uint32_t c[8] = {};
        LVRHIP_CHECK(hipMemcpyAsync(&e, err.p, sizeof(int), hipMemcpyDeviceToHost, st));
        if (d_cnt.p) LVRHIP_CHECK(hipMemcpyAsync(c, d_cnt.p, sizeof(c), hipMemcpyDeviceToHost, st));
        LVRHIP_CHECK(hipStreamSynchronize(st));
        if (d_cnt.p) {
            n_round[0] = nq_first;
            for (int r = 2; r <= 5; r++) n_round[r - 1] = c[r];
        }
        if (e != 0) {
            fill_u32_kernel<<<1, 64, 0, st>>>((uint32_t*)err.p, 1, 0u);
            LVRHIP_LAUNCH_CHECK();
            LVRHIP_CHECK(hipStreamSynchronize(st));
            throw Error(LVRHIP_ECAPACITY, e == 5 ? "kNN traversal stack overflow" : "kNN radius search did not converge");
        }
    }
    void check_err() { read_err_and_counts(); }

    // max_waves (with the work queue): the grid's cap, below the waves the device holds for a
    // query list that is usually near empty (the launch then costs less than a full grid).
    template <template <int> class SinkT>
    void tile(const QueryView& qv, const TileParams& prm_in, int P, SinkT<2> s2, SinkT<4> s4, SinkT<8> s8,
              int64_t max_waves = 0)
    {
        TileParams prm = prm_in;
        if (use_queue) {
            d_queue.reserve(QUEUE_SEGS * QUEUE_STRIDE);
            prm.queue = d_queue.p;
        }
        const int64_t cap = max_waves > 0 ? std::min(max_waves, resident_waves) : resident_waves;
        switch (P) {
            case 2: launch_tile_p<2, SinkT>(view(), qv, prm, s2, st, cap); break;
            case 4: launch_tile_p<4, SinkT>(view(), qv, prm, s4, st, cap); break;
            case 8: launch_tile_p<8, SinkT>(view(), qv, prm, s8, st, cap); break;
            default: throw Error(LVRHIP_EINVAL, "lanes per query must be 2, 4 or 8");
        }
    }

    // Scan table and every buffer the build writes (sized once, so that the build itself
    // allocates nothing and can be captured in a graph).
    void init(const float* d_xyz, const int64_t* offsets, int ns)
    {
        nscans = ns;
        h_off.assign(offsets, offsets + ns + 1);
        h_pstart.assign(ns + 1, 0);
        min_scan = 0;
        for (int s = 0; s < ns; s++) {
            const int64_t c = h_off[s + 1] - h_off[s];
            h_pstart[s + 1] = h_pstart[s] + ceil_div(c, LEAF) * LEAF;
            if (c > 0 && (min_scan == 0 || c < min_scan)) min_scan = c;
        }
        n = h_off[ns];
        npad = h_pstart[ns];
        src = d_xyz;
        err.reserve(1);
        fill_u32_kernel<<<1, 64, 0, st>>>((uint32_t*)err.p, 1, 0u);
        LVRHIP_LAUNCH_CHECK();
        d_cnt.reserve(CNT_SLOTS);
        fill_u32_kernel<<<1, 64, 0, st>>>(d_cnt.p, CNT_SLOTS, 0u);
        LVRHIP_LAUNCH_CHECK();
        if (!h_pin) LVRHIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&h_pin), 64, hipHostMallocDefault));
        if (n == 0) return;
        d_off.reserve(ns + 1);
        d_pstart.reserve(ns + 1);
        LVRHIP_CHECK(hipMemcpyAsync(d_off.p, h_off.data(), (ns + 1) * sizeof(int64_t), hipMemcpyHostToDevice, st));
        LVRHIP_CHECK(
            hipMemcpyAsync(d_pstart.p, h_pstart.data(), (ns + 1) * sizeof(int64_t), hipMemcpyHostToDevice, st));
        nleaf = (int)(npad / LEAF);
        const int nint = std::max(nleaf - 1, 1);
        sbox.reserve(6 * (size_t)ns);
        frames.reserve(ns);
        area.reserve(1);
        s_k0.reserve(std::max(n, (int64_t)nleaf));
        s_k1.reserve(std::max(n, npad));
        s_v0.reserve(n);
        s_v1.reserve(n);
        s_hist.reserve(rs_hist_size(n));
        px.reserve(npad);
        pts4.reserve(npad);
        py.reserve(npad);
        pz.reserve(npad);
        pidx.reserve(npad);
        rank.reserve(n);
        qreal.reserve(n);
        if (ns > 1) {
            scan_of.reserve(npad);
            roots.reserve(ns);
        }
        leaf_lo.reserve(nleaf);
        leaf_hi.reserve(nleaf);
        leaf_agg.reserve((size_t)3 * nleaf);
        fr_ref.reserve((size_t)FRONTIER * nscans);
        fr_lo.reserve((size_t)FRONTIER * nscans);
        fr_hi.reserve((size_t)FRONTIER * nscans);
        leaf_parent.reserve(nleaf);
        nodes.reserve(nint);
        nodes4.reserve(nint);
        nlo.reserve(nint);
        nhi.reserve(nint);
        s_flags.reserve(std::max<int64_t>(nint, n));
        s_f.reserve(MEDIAN_MAX);
        for (int b = 0; b < 2; b++) {
            ru_lo[b].reserve(nleaf);
            ru_hi[b].reserve(nleaf);
        }
    }

    // LBVH build from `src`, entirely on the device (no host synchronisation).
    void build()
    {
        if (n <= 0) return;
        const int64_t N = n;
        const int ns = nscans;
        const int tok = prof.start(LVRHIP_T_BUILD, st);
        box_init_kernel<<<(unsigned)ceil_div(6 * ns, 256), 256, 0, st>>>(sbox.p, ns);
        LVRHIP_LAUNCH_CHECK();
        const int64_t waves = ceil_div(N, (int64_t)kWave * SB_ITEMS);
        scan_bbox_kernel<<<(unsigned)ceil_div(waves, 4), 256, 0, st>>>(src, N, d_off.p, ns, sbox.p);
        LVRHIP_LAUNCH_CHECK();
        scan_frames_kernel<<<(unsigned)ceil_div(ns, 256), 256, 0, st>>>(sbox.p, ns, frames.p);
        LVRHIP_LAUNCH_CHECK();
        // sort by (scan, Morton code)
        const unsigned gb = (unsigned)ceil_div(N, 256);
        morton_scan_kernel<<<gb, 256, 0, st>>>(src, N, d_off.p, ns, frames.p, s_k0.p, s_v0.p);
        LVRHIP_LAUNCH_CHECK();
        radix_sort_pairs(s_k0.p, s_v0.p, s_k1.p, s_v1.p, N, s_hist.p, st);
        const unsigned gp = (unsigned)ceil_div(npad, 256);
        gather_padded_kernel<<<gp, 256, 0, st>>>(src, d_off.p, d_pstart.p, ns, npad, s_v0.p, s_k0.p, px.p, py.p, pz.p,
                                                 pidx.p, rank.p, s_k1.p, qreal.p, ns > 1 ? scan_of.p : nullptr,
                                                 pts4.p);
        LVRHIP_LAUNCH_CHECK();
        // leaves and hierarchy
        leaf_boxes_kernel<<<gp, 256, 0, st>>>(px.p, py.p, pz.p, npad, s_k1.p, nleaf, leaf_lo.p, leaf_hi.p, s_k0.p);
        LVRHIP_LAUNCH_CHECK();
        leaf_sums_kernel<<<(unsigned)ceil_div(nleaf, 256), 256, 0, st>>>(px.p, py.p, pz.p, npad, leaf_lo.p, leaf_hi.p,
                                                                         nleaf, leaf_agg.p);
        LVRHIP_LAUNCH_CHECK();
        const int nint = std::max(nleaf - 1, 1);
        fill_u32_kernel<<<(unsigned)ceil_div(nleaf, 256), 256, 0, st>>>((uint32_t*)leaf_parent.p, nleaf, ~0u);
        LVRHIP_LAUNCH_CHECK();
        if (nleaf > 1) {
            karras_kernel<<<(unsigned)ceil_div(nleaf - 1, 256), 256, 0, st>>>(s_k0.p, nleaf, nodes.p, leaf_parent.p);
            LVRHIP_LAUNCH_CHECK();
            // the refit as range unions, level by level (lbvh.hpp)
            refit_leaf_slots_kernel<<<(unsigned)ceil_div(nleaf, 256), 256, 0, st>>>(nleaf, leaf_parent.p, leaf_lo.p,
                                                                                   leaf_hi.p, nodes.p);
            LVRHIP_LAUNCH_CHECK();
            const float4* tlo = leaf_lo.p;
            const float4* thi = leaf_hi.p;
            for (int j = 1, cur = 0; (1 << j) <= nleaf; j++, cur ^= 1) {
                const int half = 1 << (j - 1);
                range_union_level_kernel<<<(unsigned)ceil_div(nleaf - 2 * half + 1, 256), 256, 0, st>>>(
                    nleaf, half, tlo, thi, ru_lo[cur].p, ru_hi[cur].p);
                LVRHIP_LAUNCH_CHECK();
                tlo = ru_lo[cur].p;
                thi = ru_hi[cur].p;
                refit_level_kernel<<<(unsigned)ceil_div(nint, 256), 256, 0, st>>>(nint, j, tlo, thi, nodes.p, nlo.p,
                                                                                 nhi.p);
                LVRHIP_LAUNCH_CHECK();
            }
        }
        if (ns > 1) {
            scan_roots_kernel<<<(unsigned)ceil_div(ns, 256), 256, 0, st>>>(d_pstart.p, ns, leaf_parent.p, nodes.p,
                                                                          roots.p);
            LVRHIP_LAUNCH_CHECK();
        }
        if (nleaf > 1) {
            node4_kernel<<<(unsigned)ceil_div(nleaf - 1, 256), 256, 0, st>>>(nodes.p, nleaf - 1, nodes4.p);
            LVRHIP_LAUNCH_CHECK();
        }
        frontier_kernel<<<(unsigned)ns, FRONTIER, 0, st>>>(nodes.p, leaf_lo.p, leaf_hi.p, nleaf,
                                                          ns > 1 ? roots.p : nullptr, fr_ref.p, fr_lo.p, fr_hi.p);
        LVRHIP_LAUNCH_CHECK();
        // density scale for the first radius guess: median leaf area per point
        const int nsample = std::min(nleaf, MEDIAN_MAX);
        const int stride = std::max(1, nleaf / nsample);
        leaf_area_kernel<<<(unsigned)ceil_div(nsample, 256), 256, 0, st>>>(leaf_lo.p, leaf_hi.p, nleaf, stride,
                                                                          nsample, s_f.p);
        LVRHIP_LAUNCH_CHECK();
        median_kernel<<<1, 1024, 0, st>>>(s_f.p, nsample, area.p);
        LVRHIP_LAUNCH_CHECK();
        fill_u32_kernel<<<1, 64, 0, st>>>((uint32_t*)err.p, 1, 0u);
        LVRHIP_LAUNCH_CHECK();
        prof.stop(tok, st);
    }

    // out = the entries i < min(nq, *nq_dev) with flags[i] != 0 (qsel[i] or i), in order;
    // *count_dev = their count
    void compact_flags(const uint32_t* flags, int64_t nq, const uint32_t* nq_dev, const uint32_t* qsel, uint32_t* out,
                       uint32_t* count_dev)
    {
        s_bsum.reserve(std::max<int64_t>(ceil_div(nq, CF_TILE), 1) + 1);
        compact_flags_with(flags, nq, nq_dev, qsel, out, count_dev, s_bsum.p);
    }
    // the same with the caller's block-sum buffer (ceil(nq / CF_TILE) + 1 entries)
    void compact_flags_with(const uint32_t* flags, int64_t nq, const uint32_t* nq_dev, const uint32_t* qsel,
                            uint32_t* out, uint32_t* count_dev, uint32_t* bsum)
    {
        const int64_t nb = std::max<int64_t>(ceil_div(nq, CF_TILE), 1);
        flag_block_sums_kernel<<<(unsigned)nb, CF_THREADS, 0, st>>>(flags, nq, nq_dev, bsum);
        LVRHIP_LAUNCH_CHECK();
        scan_exclusive_kernel<<<1, 1024, 0, st>>>(bsum, nb, count_dev);
        LVRHIP_LAUNCH_CHECK();
        compact_scatter_kernel<<<(unsigned)nb, CF_THREADS, 0, st>>>(flags, nq, nq_dev, bsum, qsel, out);
        LVRHIP_LAUNCH_CHECK();
    }

    // Buffers of a normals pass (host side only: no GPU work).
    void prepare_normals()
    {
        kiw = (ki > 0 && ki <= 2 * kn) ? clamp_k(ki) : 0;
        for (int r = 1; r <= 5; r++) (void)clamp_k((int64_t)kn << r);
        nx.reserve(npad);
        ny.reserve(npad);
        nz.reserve(npad);
        kround.reserve(npad);
        mom.reserve((size_t)npad * 9);
        if (kiw > 0) ki_list.reserve((size_t)npad * kiw);
        s_flags.reserve(n);
        s_list0.reserve(n);
        s_list1.reserve(n);
        a_sel.reserve(n);
        a_list.reserve(n);
        s_hint.reserve(npad);
    }

    // calculateSurfaceNormals' adaptive rounds (AdaptiveKSearchSurface.tcc:187-284) for the
    // queries q0[0, nq0) (NULL: all points): round r searches k = kn * 2^r for the queries the
    // previous round's bounding-box test rejected; the counts stay on the device, so the
    // sequence enqueues without host synchronisation.
    void enqueue_normals(const uint32_t* q0, int64_t nq0)
    {
        uint32_t* lists[2] = {s_list0.p, s_list1.p};
        nq_first = nq0;
        if (nq0 == 0) return;
        const int last = merge_retry ? 1 : 5;  // rounds searched one at a time
        for (int r = 1; r <= last; r++) {
            const int t0 = clamp_k((int64_t)kn << r);
            TileParams prm = params(t0, r == 1 ? kiw : 0, 0);
            // the later rounds' query lists are sparse: a run's consecutive tiles lie far apart, so
            // its shared leaf list does not pay, and one tile per run balances their costlier tiles;
            // round 1 keeps at least two runs per resident wave (a small cloud has few tiles)
            if (r > 1)
                prm.tiles_per_wave = 1;
            else
                prm.tiles_per_wave = (int)std::clamp<int64_t>(
                    ceil_div(nq0, kWave / lanes_per_query) / (2 * resident_waves), 1, tiles_per_wave);
            // each round's k-th d^2 per point sizes the next round's first radius: k doubles, and
            // on a surface the count grows like d^2 (x 1.15 of margin; the tuned r2_margin only
            // sizes the radius carried from tile to tile)
            prm.hint_in = r > 1 ? s_hint.p : nullptr;
            prm.hint_scale = 2.0f * 1.15f;
            prm.hint_out = r < last ? s_hint.p : nullptr;
            prm.leaf_hint = 1;
            prm.stats = kstats_ptr(r - 1);
            const uint32_t* qsel = r == 1 ? q0 : lists[(r - 1) & 1];
            const uint32_t* nq_dev = r == 1 ? nullptr : counter(r);
            QueryView qv{px.p, py.p, pz.p, qsel, (uint32_t)nq0, nq_dev};
            auto fill = [&](auto& s) {
                s.mom = mom.p;
                s.retry = r < 5 ? s_flags.p : nullptr;
                s.ki_list = ki_list.p;
                s.kiw = prm.t1;
            };
            auto mk = [&](auto tag) {
                SinkNormals<decltype(tag)::value> s{};
                fill(s);
                return s;
            };
            auto mk1 = [&](auto tag) {  // the rounds without a ki list: one target
                SinkMoments<decltype(tag)::value> s{};
                fill(s);
                return s;
            };
            const int tok = prof.start(r == 1 ? LVRHIP_T_NORMALS_FIRST : LVRHIP_T_NORMALS_RETRY, st);
            if (prm.t1 > 0)
                tile<SinkNormals>(qv, prm, lanes_per_query, mk(std::integral_constant<int, 2>()),
                                  mk(std::integral_constant<int, 4>()), mk(std::integral_constant<int, 8>()));
            else
                tile<SinkMoments>(qv, prm, lanes_per_query, mk1(std::integral_constant<int, 2>()),
                                  mk1(std::integral_constant<int, 4>()), mk1(std::integral_constant<int, 8>()));
            prof.stop(tok, st);
            PlaneFitParams pf{};
            pf.qsel = qsel;
            pf.nq = nq0;
            pf.nq_dev = nq_dev;
            pf.retry = r < 5 ? s_flags.p : nullptr;
            pf.mom = mom.p;
            pf.px = px.p;
            pf.py = py.p;
            pf.pz = pz.p;
            pf.method = method;
            pf.kcount = t0;
            pf.fx = flip[0];
            pf.fy = flip[1];
            pf.fz = flip[2];
            pf.flips = (scan_flips && nscans > 1) ? d_flips.p : nullptr;
            pf.scan_of = nscans > 1 ? scan_of.p : nullptr;
            pf.nx = nx.p;
            pf.ny = ny.p;
            pf.nz = nz.p;
            pf.kround = kround.p;
            pf.round = r;
            plane_fit_kernel<<<(unsigned)ceil_div(nq0, 256), 256, 0, st>>>(pf);
            LVRHIP_LAUNCH_CHECK();
            // rejected queries -> the next round's list and count
            if (r < 5) compact_flags(s_flags.p, nq0, nq_dev, qsel, lists[r & 1], counter(r + 1));
        }
        if (merge_retry) enqueue_merged_rounds(lists[1], counter(2), nq0);
    }

    // Rounds 2..5 for the queries round 1 rejected, in one search with four targets
    // (SinkLevels); the plane fit of the round-5 points; the points that stop at round 2..4
    // searched again at their round's k (SinkNormals) and fitted.
    void enqueue_merged_rounds(const uint32_t* qsel, const uint32_t* nq_dev, int64_t nq0)
    {
        TileParams prm = params(clamp_k((int64_t)kn << 5), clamp_k((int64_t)kn << 4), 1);
        prm.t2 = clamp_k((int64_t)kn << 3);
        prm.t3 = clamp_k((int64_t)kn << 2);
        prm.tiles_per_wave = 1;  // sparse lists, as in the rounds searched one at a time
        prm.stats = kstats_ptr(5);
        uint32_t* redo = s_flags.p;
        {
            QueryView qv{px.p, py.p, pz.p, qsel, (uint32_t)nq0, nq_dev};
            auto mk = [&](auto tag) {
                constexpr int P = decltype(tag)::value;
                SinkLevels<P> s{};
                s.mom = mom.p;
                s.kround = kround.p;
                s.redo = redo;
                return s;
            };
            const int tok = prof.start(LVRHIP_T_NORMALS_RETRY, st);
            tile<SinkLevels>(qv, prm, lanes_per_query, mk(std::integral_constant<int, 2>()),
                             mk(std::integral_constant<int, 4>()), mk(std::integral_constant<int, 8>()));
            prof.stop(tok, st);
        }
        plane_fit(qsel, nq0, nq_dev, redo, prm.t0, 5);
        fill_u32_kernel<<<1, 64, 0, st>>>(counter(3), 3, 0u);
        LVRHIP_LAUNCH_CHECK();
        round_count_kernel<<<(unsigned)ceil_div(nq0, 256), 256, 0, st>>>(qsel, nq_dev, nq0, kround.p, counter(0));
        LVRHIP_LAUNCH_CHECK();
        // points whose search stops at round 2, 3 or 4
        for (int r = 2; r <= 4; r++) {
            select_round_kernel<<<(unsigned)ceil_div(nq0, 256), 256, 0, st>>>(redo, nq0, (uint32_t)r, a_sel.p);
            LVRHIP_LAUNCH_CHECK();
            compact_flags(a_sel.p, nq0, nq_dev, qsel, a_list.p, counter(6));
            const int t0 = clamp_k((int64_t)kn << r);
            TileParams pr = params(t0, 0, 1);
            pr.tiles_per_wave = 1;
            pr.stats = kstats_ptr(r - 1);
            QueryView qv{px.p, py.p, pz.p, a_list.p, (uint32_t)nq0, counter(6)};
            auto mk = [&](auto tag) {
                constexpr int P = decltype(tag)::value;
                SinkNormals<P> s{};
                s.mom = mom.p;
                s.retry = nullptr;
                s.ki_list = ki_list.p;
                s.kiw = 0;
                return s;
            };
            const int tok = prof.start(LVRHIP_T_NORMALS_RETRY, st);
            tile<SinkNormals>(qv, pr, lanes_per_query, mk(std::integral_constant<int, 2>()),
                              mk(std::integral_constant<int, 4>()), mk(std::integral_constant<int, 8>()));
            prof.stop(tok, st);
            plane_fit(a_list.p, nq0, counter(6), nullptr, t0, r);
        }
    }

    void plane_fit(const uint32_t* qsel, int64_t nq, const uint32_t* nq_dev, const uint32_t* skip, int kcount,
                   int round)
    {
        PlaneFitParams pf{};
        pf.qsel = qsel;
        pf.nq = nq;
        pf.nq_dev = nq_dev;
        pf.retry = skip;
        pf.mom = mom.p;
        pf.px = px.p;
        pf.py = py.p;
        pf.pz = pz.p;
        pf.method = method;
        pf.kcount = kcount;
        pf.kn = kn;
        pf.kmax = nscans == 1 ? n : ((int64_t)1 << 40);
        pf.fx = flip[0];
        pf.fy = flip[1];
        pf.fz = flip[2];
        pf.flips = (scan_flips && nscans > 1) ? d_flips.p : nullptr;
        pf.scan_of = nscans > 1 ? scan_of.p : nullptr;
        pf.nx = nx.p;
        pf.ny = ny.p;
        pf.nz = nz.p;
        pf.kround = kround.p;
        pf.round = round;
        plane_fit_kernel<<<(unsigned)ceil_div(nq, 256), 256, 0, st>>>(pf);
        LVRHIP_LAUNCH_CHECK();
    }

    void calculate_normals()
    {
        drop_graph();
        has_normals = false;
        interpolated = false;
        ki_valid = false;
        for (auto& c : n_round) c = 0;
        if (n == 0) {
            has_normals = true;
            return;
        }
        prepare_normals();
        enqueue_normals(all_queries(), n);
        read_err_and_counts();
        prof.collect();
        has_normals = true;
        ki_valid = kiw > 0;
    }

    // the ki list when the first normals round could not supply it (ki > 2 kn)
    void enqueue_ki_list()
    {
        const int kw = clamp_k(ki);
        kiw = kw;
        ki_list.reserve((size_t)npad * kw);
        TileParams prm = params(kw, 0, 3);
        prm.stats = kstats_ptr(7);
        QueryView qv{px.p, py.p, pz.p, all_queries(), (uint32_t)n, nullptr};
        auto mk = [&](auto tag) {
            constexpr int P = decltype(tag)::value;
            SinkKiOnly<P> s{};
            s.out = ki_list.p;
            s.kw = kw;
            return s;
        };
        tile<SinkKiOnly>(qv, prm, lanes_per_query, mk(std::integral_constant<int, 2>()),
                         mk(std::integral_constant<int, 4>()), mk(std::integral_constant<int, 8>()));
    }

    // ki above 64 runs through sorted kSearch rows (interpolate_rows): single-scan handles, and
    // not inside a captured graph (the rows are allocated per call)
    bool rows_interpolation() const { return std::min<int64_t>(ki, nscans > 1 ? min_scan : n) > 64; }
    void check_interpolate_graph() const
    {
        if (rows_interpolation()) throw Error(LVRHIP_EINVAL, "a captured graph takes ki <= 64");
    }

    void interpolate_rows()
    {
        require_one_scan("interpolation with ki above 64");
        const int k = clamp_k(ki);
        DBuf<float> q;
        q.reserve(3 * (size_t)n);
        input_order_points_kernel<<<(unsigned)ceil_div(n, 256), 256, 0, st>>>(pts4.p, rank.p, n, q.p);
        LVRHIP_LAUNCH_CHECK();
        DBuf<int32_t> idx;
        idx.reserve((size_t)n * k);
        knn_device(q.p, n, k, idx.p, nullptr);
        tx.reserve(npad);
        ty.reserve(npad);
        tz.reserve(npad);
        // padding slots keep what they hold
        LVRHIP_CHECK(hipMemcpyAsync(tx.p, nx.p, (size_t)npad * sizeof(float), hipMemcpyDeviceToDevice, st));
        LVRHIP_CHECK(hipMemcpyAsync(ty.p, ny.p, (size_t)npad * sizeof(float), hipMemcpyDeviceToDevice, st));
        LVRHIP_CHECK(hipMemcpyAsync(tz.p, nz.p, (size_t)npad * sizeof(float), hipMemcpyDeviceToDevice, st));
        const int tok = prof.start(LVRHIP_T_INTERPOLATE, st);
        interpolate_rows_kernel<<<(unsigned)ceil_div(n, 256), 256, 0, st>>>(n, k, idx.p, rank.p, nx.p, ny.p, nz.p,
                                                                           tx.p, ty.p, tz.p);
        LVRHIP_LAUNCH_CHECK();
        prof.stop(tok, st);
        LVRHIP_CHECK(hipStreamSynchronize(st));
    }

    // interpolateSurfaceNormals (tcc:293-335) from nx/ny/nz into ox/oy/oz
    void enqueue_interpolate(float* ox, float* oy, float* oz)
    {
        const int64_t R = npad;
        const int tok = prof.start(LVRHIP_T_INTERPOLATE, st);
        const unsigned blocks = (unsigned)ceil_div(R, INTERP_POINTS);
        if (kiw <= 16) {
            interpolate_kernel<16><<<blocks, INTERP_THREADS, 0, st>>>(nx.p, ny.p, nz.p, pts4.p, ki_list.p, kiw, R, ox,
                                                                    oy, oz);
        } else if (kiw <= 32) {
            interpolate_kernel<32><<<blocks, INTERP_THREADS, 0, st>>>(nx.p, ny.p, nz.p, pts4.p, ki_list.p, kiw, R, ox,
                                                                    oy, oz);
        } else {
            interpolate_kernel<64><<<blocks, INTERP_THREADS, 0, st>>>(nx.p, ny.p, nz.p, pts4.p, ki_list.p, kiw, R, ox,
                                                                    oy, oz);
        }
        LVRHIP_LAUNCH_CHECK();
        prof.stop(tok, st);
    }

    void interpolate_normals()
    {
        if (!has_normals) throw Error(LVRHIP_ESTATE, "interpolate_normals called before calculate_normals");
        drop_graph();
        if (ki <= 0 || n == 0) return;
        if (rows_interpolation()) {
            interpolate_rows();
        } else {
            if (!(ki_valid && kiw == clamp_k(ki))) {
                enqueue_ki_list();
                check_err();
                ki_valid = true;
            }
            tx.reserve(npad);
            ty.reserve(npad);
            tz.reserve(npad);
            enqueue_interpolate(tx.p, ty.p, tz.p);
        }
        if (prof.on) prof.collect();
        interpolated = true;
        std::swap(nx.p, tx.p);
        std::swap(nx.cap, tx.cap);
        std::swap(ny.p, ty.p);
        std::swap(ny.cap, ty.cap);
        std::swap(nz.p, tz.p);
        std::swap(nz.cap, tz.cap);
    }

    void scatter_normals(const float* ax, const float* ay, const float* az, float* d_out)
    {
        scatter_normals_kernel<<<(unsigned)ceil_div(npad, 256), 256, 0, st>>>(ax, ay, az, pidx.p, npad, d_out);
        LVRHIP_LAUNCH_CHECK();
    }

    // The whole per-cloud pipeline, enqueue only: [LBVH build,] adaptive normals,
    // interpolation, normals to d_out in input order.
    void enqueue_pipeline(float* d_out, bool rebuild)
    {
        prepare_normals();
        if (rebuild) build();
        enqueue_normals(all_queries(), n);
        if (ki > 0) {
            if (kiw != clamp_k(ki)) enqueue_ki_list();
            tx.reserve(npad);
            ty.reserve(npad);
            tz.reserve(npad);
            enqueue_interpolate(tx.p, ty.p, tz.p);
            scatter_normals(tx.p, ty.p, tz.p, d_out);
        } else {
            scatter_normals(nx.p, ny.p, nz.p, d_out);
        }
    }

    hipStream_t graph_st()
    {
        if (st) return st;
        if (!gst) {
            LVRHIP_CHECK(hipStreamCreateWithFlags(&gst, hipStreamNonBlocking));
            own_gst = true;
        }
        return gst;
    }

    void drop_graph()
    {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        gexec = nullptr;
        graph = nullptr;
    }

    void graph_capture(float* d_out, bool rebuild)
    {
        if (prof.on) throw Error(LVRHIP_ESTATE, "switch profiling off before capturing a graph");
        if (n == 0) throw Error(LVRHIP_EINVAL, "empty point cloud");
        if (ki > 0) check_interpolate_graph();
        prepare_normals();
        drop_graph();
        hipStream_t saved = st;
        st = graph_st();
        try {
            // one eager pass sizes every buffer, so the captured pass allocates nothing
            enqueue_pipeline(d_out, rebuild);
            read_err_and_counts();
            LVRHIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            try {
                enqueue_pipeline(d_out, rebuild);
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(st, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            LVRHIP_CHECK(hipStreamEndCapture(st, &graph));
            LVRHIP_CHECK(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
            graph_stream = st;
        } catch (...) {
            st = saved;
            throw;
        }
        st = saved;
        has_normals = false;  // the handle's eager normals are not the graph's output
        ki_valid = false;
    }

    void graph_launch()
    {
        if (!gexec) throw Error(LVRHIP_ESTATE, "no graph captured");
        LVRHIP_CHECK(hipGraphLaunch(gexec, graph_stream));
    }

    void synchronize()
    {
        hipStream_t saved = st;
        if (gexec) st = graph_stream;
        try {
            read_err_and_counts();
        } catch (...) {
            st = saved;
            throw;
        }
        st = saved;
        if (!gexec) LVRHIP_CHECK(hipStreamSynchronize(st));
    }

    // Project autotuner: times the adaptive normals pass on a sample (every 16th run of 4096
    // consecutive points in sort order) for each (lanes per query, tiles per wave, radius
    // margin) and keeps the fastest.
    void autotune(int* outP, int* outT, float* outM)
    {
        if (n == 0) return;
        if (prof.on) throw Error(LVRHIP_ESTATE, "switch profiling off before autotuning");
        drop_graph();
        prepare_normals();
        const int RUN = 4096, EVERY = 16;
        const int64_t runs = ceil_div(n, (int64_t)RUN);
        int64_t m = 0;
        for (int64_t r = 0; r < runs; r += EVERY) m += std::min<int64_t>(RUN, n - r * RUN);
        s_sample.reserve(m);
        sample_runs_kernel<<<(unsigned)ceil_div(m, 256), 256, 0, st>>>(all_queries(), m, RUN, EVERY, s_sample.p);
        LVRHIP_LAUNCH_CHECK();
        hipEvent_t a, b;
        LVRHIP_CHECK(hipEventCreate(&a));
        LVRHIP_CHECK(hipEventCreate(&b));
        const int Ps[2] = {4, 8};
        const int Ts[4] = {4, 8, 12, 16};
        const float Ms[3] = {1.15f, 1.3f, 1.6f};
        float best = INFINITY;
        int bp = lanes_per_query, bt = tiles_per_wave;
        float bm = r2_margin;
        try {
            enqueue_normals(s_sample.p, m);  // warm-up
            read_err_and_counts();
            for (int P : Ps)
                for (int T : Ts)
                    for (float M : Ms) {
                        lanes_per_query = P;
                        tiles_per_wave = T;
                        r2_margin = M;
                        float t = INFINITY;
                        for (int rep = 0; rep < 2; rep++) {
                            LVRHIP_CHECK(hipEventRecord(a, st));
                            enqueue_normals(s_sample.p, m);
                            LVRHIP_CHECK(hipEventRecord(b, st));
                            read_err_and_counts();
                            float ms = 0.0f;
                            LVRHIP_CHECK(hipEventElapsedTime(&ms, a, b));
                            t = std::min(t, ms);
                        }
                        if (t < best) {
                            best = t;
                            bp = P;
                            bt = T;
                            bm = M;
                        }
                    }
        } catch (...) {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
            throw;
        }
        lanes_per_query = bp;
        tiles_per_wave = bt;
        r2_margin = bm;
        // the later rounds merged into one search, or searched one at a time
        try {
            const bool cur = merge_retry;
            merge_retry = !cur;
            float t = INFINITY;
            for (int rep = 0; rep < 2; rep++) {
                LVRHIP_CHECK(hipEventRecord(a, st));
                enqueue_normals(s_sample.p, m);
                LVRHIP_CHECK(hipEventRecord(b, st));
                read_err_and_counts();
                float ms = 0.0f;
                LVRHIP_CHECK(hipEventElapsedTime(&ms, a, b));
                t = std::min(t, ms);
            }
            if (!(t < best)) merge_retry = cur;
        } catch (...) {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
            throw;
        }
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        has_normals = false;
        ki_valid = false;
        for (auto& c : n_round) c = 0;
        if (outP) *outP = bp;
        if (outT) *outT = bt;
        if (outM) *outM = bm;
    }

    // Morton-sort m device queries; fills s_qx/s_qy/s_qz and s_v0 (sorted -> caller index).
    void sort_queries(const float* d_q, int64_t m)
    {
        s_k0.reserve(std::max<int64_t>(m, 1));
        s_k1.reserve(std::max<int64_t>(m, 1));
        s_v0.reserve(std::max<int64_t>(m, 1));
        s_v1.reserve(std::max<int64_t>(m, 1));
        s_hist.reserve(rs_hist_size(m));
        const unsigned gb = (unsigned)ceil_div(m, 256);
        morton_queries_kernel<<<gb, 256, 0, st>>>(d_q, m, frames.p, s_k0.p, s_v0.p);
        LVRHIP_LAUNCH_CHECK();
        // the top 32 bits of the 63-bit code (cells of 1/1,600 of the box per axis) order the
        // queries well enough for the searches' locality in half the passes
        radix_sort_pairs(s_k0.p, s_v0.p, s_k1.p, s_v1.p, m, s_hist.p, st, 64, 32);
        s_qx.reserve(m);
        s_qy.reserve(m);
        s_qz.reserve(m);
        gather_queries_kernel<<<gb, 256, 0, st>>>(d_q, m, s_v0.p, s_qx.p, s_qy.p, s_qz.p);
        LVRHIP_LAUNCH_CHECK();
    }

    void require_one_scan(const char* what) const
    {
        if (nscans > 1) throw Error(LVRHIP_EINVAL, std::string(what) + " needs a single-scan handle");
    }

    void distances_device(const float* d_q, int64_t m, float voxel, float* d_proj, float* d_eucl, uint8_t* d_inv)
    {
        if (!has_normals) throw Error(LVRHIP_ESTATE, "distances called before calculate_normals");
        require_one_scan("distances");
        if (m == 0) return;
        if (n == 0) throw Error(LVRHIP_ESTATE, "distances on an empty point cloud");
        drop_graph();
        const int t0 = clamp_k(kd);
        s_k1.reserve(std::max<size_t>((size_t)m * t0, (size_t)m));
        s_spos.reserve((size_t)m * t0);
        s_ord.reserve((size_t)m * t0);
        const int tsort = prof.start(LVRHIP_T_SDF_SORT, st);
        sort_queries(d_q, m);
        prof.stop(tsort, st);
        const bool bf = best_first(t0);
        const uint32_t* qsel = nullptr;
        const uint32_t* nq_dev = nullptr;
        if (bf) {
            // best-first for every vertex; the tile kernel redoes the ones whose stacks overflowed
            s_redo.reserve(m);
            BfSdfOut o{px.p, py.p, pz.p, nx.p, ny.p, nz.p, s_v0.p, voxel, d_proj, d_eucl, d_inv, s_redo.p};
            const int tok = prof.start(LVRHIP_T_SDF_KERNEL, st);
            enqueue_bf(m, t0, BF_EMPTY, o, 6);
            prof.stop(tok, st);
            qsel = s_redo_list.p;
            nq_dev = counter(8);
        }
        TileParams prm = params(t0, 0, 2);
        prm.tiles_per_wave = sdf_tiles_per_wave;
        prm.r2_margin = sdf_r2_margin;
        prm.grow_min = sdf_grow_min;
        prm.stats = kstats_ptr(6);
        QueryView qv{s_qx.p, s_qy.p, s_qz.p, qsel, (uint32_t)m, nq_dev};
        auto mk = [&](auto tag) {
            constexpr int P = decltype(tag)::value;
            SinkSdf<P> s{};
            s.px = px.p;
            s.py = py.p;
            s.pz = pz.p;
            s.nx = nx.p;
            s.ny = ny.p;
            s.nz = nz.p;
            s.qorig = s_v0.p;
            s.keys = s_k1.p;
            s.spos = s_spos.p;
            s.ord = s_ord.p;
            s.kw = t0;
            s.voxel = voxel;
            s.out_proj = d_proj;
            s.out_eucl = d_eucl;
            s.out_invalid = d_inv;
            s.npts = (uint32_t)npad;
            return s;
        };
        const int tok = prof.start(bf ? LVRHIP_T_SDF_REDO : LVRHIP_T_SDF_KERNEL, st);
        // after the best-first passes the list is usually empty: a wave per CU drains it
        tile<SinkSdf>(qv, prm, lanes_per_query, mk(std::integral_constant<int, 2>()),
                      mk(std::integral_constant<int, 4>()), mk(std::integral_constant<int, 8>()),
                      bf ? resident_waves / 16 : 0);
        prof.stop(tok, st);
        check_err();
        prof.collect();
    }

    // distances_device without allocating: the queries in the caller's order (no Morton sort), the
    // best-first searches with their scratch in the static array at scr, and no tile kernel (a
    // query that outgrows the wave search's stack is reported as LVRHIP_ECAPACITY).
    void distances_noalloc(const float* d_q, int64_t m, float voxel, float* d_proj, float* d_eucl, uint8_t* d_inv,
                           char* scr)
    {
        if (!has_normals) throw Error(LVRHIP_ESTATE, "distances called before calculate_normals");
        require_one_scan("distances");
        if (m == 0) return;
        if (n == 0) throw Error(LVRHIP_ESTATE, "distances on an empty point cloud");
        if (m > NA_MAX_QUERIES)
            throw Error(LVRHIP_EINVAL, "the no-allocation distances take at most " + std::to_string(NA_MAX_QUERIES) +
                                           " queries per call (the static scratch size)");
        const int t0 = clamp_k(kd);
        if (t0 > BF_KMAX)
            throw Error(LVRHIP_EINVAL, "the no-allocation distances take kd <= " + std::to_string(BF_KMAX));
        uint32_t* flags = reinterpret_cast<uint32_t*>(scr + NA_FLAGS);
        uint32_t* list = reinterpret_cast<uint32_t*>(scr + NA_LIST);
        uint32_t* bsum = reinterpret_cast<uint32_t*>(scr + NA_BSUM);
        uint32_t* cnt = reinterpret_cast<uint32_t*>(scr + NA_COUNT);
        uint32_t* queue = reinterpret_cast<uint32_t*>(scr + NA_QUEUE);
        uint2* spill = reinterpret_cast<uint2*>(scr + NA_SPILL_OFF);
        BfSdfOut o{px.p, py.p, pz.p, nx.p, ny.p, nz.p, nullptr, voxel, d_proj, d_eucl, d_inv, flags};
        // every flag is written by the lane search; the queue is cleared by launch_bf
        launch_bf(view(), d_q, d_q + 1, d_q + 2, 3, m, t0, BF_EMPTY, o, st, nullptr, queue,
                  std::min(bf_resident, NA_WAVES), spill, NA_SPILL);
        compact_flags_with(flags, m, nullptr, nullptr, list, cnt, bsum);
        launch_bf_wave(view(), d_q, d_q + 1, d_q + 2, 3, list, cnt, t0, BF_EMPTY, o, st, nullptr);
        compact_flags_with(flags, m, nullptr, nullptr, list, cnt + 1, bsum);
        LVRHIP_CHECK(hipMemcpyAsync(h_pin, cnt + 1, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        LVRHIP_CHECK(hipStreamSynchronize(st));
        const uint32_t left = h_pin[0];
        if (left != 0)
            throw Error(LVRHIP_ECAPACITY, std::to_string(left) + " queries outgrew the search stacks of the "
                                              "no-allocation distances; use lvrhip_distances_device");
    }

    // ---- the PointsetGrid lattice (lattice.hpp) -------------------------------------------------
    DBuf<uint64_t> g_k0, g_k1, g_cells, g_cells2, g_first;
    DBuf<uint32_t> g_v0, g_v1, g_flags, g_idx, g_gid, g_vidx, g_vid, g_hist;
    DBuf<float> grid_vtx;
    DBuf<int32_t> grid_cells, grid_cidx;
    int64_t grid_nv = 0, grid_nc = 0;

    void lat_sort(uint64_t* k, uint32_t* v, int64_t m, int end_bit = 64)
    {
        g_hist.reserve(rs_hist_size(m));
        radix_sort_pairs(k, v, g_k1.p, g_v1.p, m, g_hist.p, st, end_bit);
    }

    // unique non-sentinel keys of the sorted keys[0, m) into out; returns their count
    int64_t lat_unique(const uint64_t* keys, int64_t m, uint64_t* out)
    {
        g_flags.reserve(m);
        g_idx.reserve(m);
        lat_first_flags_kernel<<<(unsigned)ceil_div(m, 256), 256, 0, st>>>(keys, m, g_flags.p);
        LVRHIP_LAUNCH_CHECK();
        compact_flags(g_flags.p, m, nullptr, nullptr, g_idx.p, counter(7));
        lat_gather_keys_kernel<<<(unsigned)ceil_div(m, 256), 256, 0, st>>>(keys, g_idx.p, counter(7), out);
        LVRHIP_LAUNCH_CHECK();
        uint32_t c = 0;
        LVRHIP_CHECK(hipMemcpyAsync(&c, counter(7), sizeof(c), hipMemcpyDeviceToHost, st));
        LVRHIP_CHECK(hipStreamSynchronize(st));
        return c;
    }

    // box: the grid's bounding box (min xyz, max xyz), or NULL for the cloud's own box
    void build_grid(float voxel, bool extrude, const float* box = nullptr)
    {
        require_one_scan("the lattice");
        if (!(voxel > 0.0f) || std::isinf(voxel)) throw Error(LVRHIP_EINVAL, "voxel size must be positive");
        drop_graph();
        grid_nv = grid_nc = 0;
        if (n == 0) return;
        LatBox b;
        if (box) {
            for (int j = 0; j < 3; j++) {
                if (!(box[j] <= box[3 + j]) || std::isinf(box[j]) || std::isinf(box[3 + j]))
                    throw Error(LVRHIP_EINVAL, "the grid box must be finite with min <= max");
                b.lo[j] = box[j];
                b.hi[j] = box[3 + j];
            }
        } else {
            // the cloud's box (the surface's bounding box, PointsetSurface.tcc:52-56)
            uint32_t enc[6];
            LVRHIP_CHECK(hipMemcpyAsync(enc, sbox.p, sizeof(enc), hipMemcpyDeviceToHost, st));
            LVRHIP_CHECK(hipStreamSynchronize(st));
            for (int j = 0; j < 6; j++) {
                const uint32_t e = enc[j];
                uint32_t bits = (e & 0x80000000u) ? (e & 0x7FFFFFFFu) : ~e;
                float f;
                std::memcpy(&f, &bits, 4);
                (j < 3 ? b.lo[j] : b.hi[j - 3]) = f;
            }
        }
        b.voxel = voxel;
        for (int j = 0; j < 3; j++) {
            const double a = std::floor((double)b.lo[j] / voxel), c = std::floor((double)b.hi[j] / voxel);
            if (!(a > -(double)(LAT_BIAS - 2) && c < (double)(LAT_BIAS - 2)))
                throw Error(LVRHIP_EINVAL, "lattice index out of range: voxel too small for the cloud's extent");
        }
        const int64_t R = npad;
        int64_t m = R;
        g_k0.reserve(m);
        g_k1.reserve(m);
        g_v0.reserve(m);
        g_v1.reserve(m);
        lat_point_cells_kernel<<<(unsigned)ceil_div(m, 256), 256, 0, st>>>(px.p, py.p, pz.p, pidx.p, R, b, g_k0.p,
                                                                          g_v0.p);
        LVRHIP_LAUNCH_CHECK();
        lat_sort(g_k0.p, g_v0.p, m);
        g_cells.reserve(m);
        int64_t nc = lat_unique(g_k0.p, m, g_cells.p);
        if (extrude && nc > 0) {
            m = 27 * nc;
            g_k0.reserve(m);
            g_k1.reserve(m);
            g_v0.reserve(m);
            g_v1.reserve(m);
            lat_extrude_kernel<<<(unsigned)ceil_div(m, 256), 256, 0, st>>>(g_cells.p, nc, b, g_k0.p, g_v0.p);
            LVRHIP_LAUNCH_CHECK();
            lat_sort(g_k0.p, g_v0.p, m);
            g_cells2.reserve(m);
            nc = lat_unique(g_k0.p, m, g_cells2.p);
            std::swap(g_cells.p, g_cells2.p);
            std::swap(g_cells.cap, g_cells2.cap);
        }
        if (nc == 0) return;
        if (8 * nc >= (int64_t)0xFFFFFFFFu) throw Error(LVRHIP_ECAPACITY, "lattice too large");
        // corners: sorted by corner key, stable in order = cell rank * 8 + slot
        m = 8 * nc;
        g_k0.reserve(m);
        g_k1.reserve(m);
        g_v0.reserve(m);
        g_v1.reserve(m);
        lat_corners_kernel<<<(unsigned)ceil_div(m, 256), 256, 0, st>>>(g_cells.p, nc, g_k0.p, g_v0.p);
        LVRHIP_LAUNCH_CHECK();
        lat_sort(g_k0.p, g_v0.p, m);
        g_flags.reserve(m);
        lat_first_flags_kernel<<<(unsigned)ceil_div(m, 256), 256, 0, st>>>(g_k0.p, m, g_flags.p);
        LVRHIP_LAUNCH_CHECK();
        const int64_t nb = ceil_div(m, (int64_t)LAT_TILE);
        s_bsum.reserve(nb + 1);
        flag_block_sums_kernel<<<(unsigned)nb, CF_THREADS, 0, st>>>(g_flags.p, m, nullptr, s_bsum.p);
        LVRHIP_LAUNCH_CHECK();
        scan_exclusive_kernel<<<1, 1024, 0, st>>>(s_bsum.p, nb, counter(7));
        LVRHIP_LAUNCH_CHECK();
        g_gid.reserve(m);
        lat_group_ids_kernel<<<(unsigned)nb, LAT_THREADS, 0, st>>>(g_flags.p, m, s_bsum.p, g_gid.p);
        LVRHIP_LAUNCH_CHECK();
        uint32_t nvu = 0;
        LVRHIP_CHECK(hipMemcpyAsync(&nvu, counter(7), sizeof(nvu), hipMemcpyDeviceToHost, st));
        LVRHIP_CHECK(hipStreamSynchronize(st));
        const int64_t nv = nvu;
        // vertex ids in first-use order
        g_first.reserve(nv);
        g_vidx.reserve(nv);
        lat_first_order_kernel<<<(unsigned)ceil_div(m, 256), 256, 0, st>>>(g_flags.p, g_gid.p, g_v0.p, m, g_first.p,
                                                                          g_vidx.p);
        LVRHIP_LAUNCH_CHECK();
        lat_sort(g_first.p, g_vidx.p, nv, 32);
        g_vid.reserve(nv);
        lat_vertex_ids_kernel<<<(unsigned)ceil_div(nv, 256), 256, 0, st>>>(g_vidx.p, nv, g_vid.p);
        LVRHIP_LAUNCH_CHECK();
        grid_cells.reserve(m);
        lat_cells_kernel<<<(unsigned)ceil_div(m, 256), 256, 0, st>>>(g_gid.p, g_v0.p, g_vid.p, m, grid_cells.p);
        LVRHIP_LAUNCH_CHECK();
        grid_vtx.reserve(3 * nv);
        lat_vertices_kernel<<<(unsigned)ceil_div(nv, 256), 256, 0, st>>>(g_first.p, g_cells.p, nv, voxel, grid_vtx.p);
        LVRHIP_LAUNCH_CHECK();
        grid_cidx.reserve(3 * nc);
        lat_cell_index_kernel<<<(unsigned)ceil_div(nc, 256), 256, 0, st>>>(g_cells.p, nc, grid_cidx.p);
        LVRHIP_LAUNCH_CHECK();
        LVRHIP_CHECK(hipStreamSynchronize(st));
        grid_nv = nv;
        grid_nc = nc;
    }

    // ---- marching cubes over the lattice (mesh.hpp) -------------------------------------------
    DBuf<uint8_t> mc_case, mc_keep, mc_border;
    DBuf<uint32_t> mc_ntri, mc_off, mc_sums, mc_fcell, mc_vals, mc_moves, mc_list;
    DBuf<uint32_t> mc_nbr;  // per mesh vertex: the neighbours of its last centroid search
    DBuf<uint64_t> mc_keys;
    DBuf<float> mesh_vtx, mc_q, mc_c;
    DBuf<int32_t> mesh_faces;
    int64_t mesh_nv = 0, mesh_nf = 0, mesh_moves = 0;

    uint32_t read_counter(int i)
    {
        LVRHIP_CHECK(hipMemcpyAsync(h_pin, counter(i), sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        LVRHIP_CHECK(hipStreamSynchronize(st));
        return h_pin[0];
    }

    // decomposition: LVRHIP_MESH_MC (FastBox cells) or LVRHIP_MESH_PMC (BilinearFastBox cells and
    // the border-edge moves with 5 neighbours, FastReconstruction.tcc:155-165)
    void build_mesh(const float* d_dist, const uint8_t* d_inv, int decomposition)
    {
        require_one_scan("the mesh");
        drop_graph();
        mesh_nv = mesh_nf = mesh_moves = 0;
        const int64_t nc = grid_nc;
        if (nc == 0) return;
        if (grid_nv * 3 >= (int64_t)0xFFFFFFFFu) throw Error(LVRHIP_ECAPACITY, "lattice too large for the mesh");
        const int pmc = decomposition == LVRHIP_MESH_PMC ? 1 : 0;
        mc_case.reserve(nc);
        mc_keep.reserve(nc);
        mc_ntri.reserve(nc);
        mc_off.reserve(nc);
        mc_sums.reserve(ceil_div(nc, SCAN_SEG) + 1);
        const unsigned gc = (unsigned)ceil_div(nc, 256);
        mc_cells_kernel<<<gc, 256, 0, st>>>(grid_cells.p, nc, grid_vtx.p, d_dist, d_inv, pmc, mc_case.p, mc_keep.p,
                                            mc_ntri.p);
        LVRHIP_LAUNCH_CHECK();
        LVRHIP_CHECK(hipMemcpyAsync(mc_off.p, mc_ntri.p, (size_t)nc * sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
        scan_exclusive(mc_off.p, nc, mc_sums.p, st);
        mc_total_kernel<<<1, 64, 0, st>>>(mc_off.p, mc_ntri.p, nc, counter(9));
        LVRHIP_LAUNCH_CHECK();
        const int64_t nf = read_counter(9);
        if (nf == 0) return;
        const int64_t m = 3 * nf;  // triangle corners
        if (m >= (int64_t)0xFFFFFFFFu) throw Error(LVRHIP_ECAPACITY, "mesh too large");
        mc_keys.reserve(m);
        mc_vals.reserve(m);
        mc_fcell.reserve(nf);
        g_k1.reserve(m);
        g_v1.reserve(m);
        mc_emit_kernel<<<gc, 256, 0, st>>>(grid_cells.p, nc, mc_case.p, mc_keep.p, mc_off.p, mc_keys.p, mc_vals.p,
                                           mc_fcell.p);
        LVRHIP_LAUNCH_CHECK();
        lat_sort(mc_keys.p, mc_vals.p, m, 32);  // by lattice edge, stable in corner order
        // one vertex per lattice edge: the group of its first use
        g_flags.reserve(m);
        lat_first_flags_kernel<<<(unsigned)ceil_div(m, 256), 256, 0, st>>>(mc_keys.p, m, g_flags.p);
        LVRHIP_LAUNCH_CHECK();
        const int64_t nb = ceil_div(m, (int64_t)LAT_TILE);
        s_bsum.reserve(nb + 1);
        flag_block_sums_kernel<<<(unsigned)nb, CF_THREADS, 0, st>>>(g_flags.p, m, nullptr, s_bsum.p);
        LVRHIP_LAUNCH_CHECK();
        scan_exclusive_kernel<<<1, 1024, 0, st>>>(s_bsum.p, nb, counter(10));
        LVRHIP_LAUNCH_CHECK();
        g_gid.reserve(m);
        lat_group_ids_kernel<<<(unsigned)nb, LAT_THREADS, 0, st>>>(g_flags.p, m, s_bsum.p, g_gid.p);
        LVRHIP_LAUNCH_CHECK();
        const int64_t nv = read_counter(10);
        // vertex ids in first-use order
        g_first.reserve(nv);
        g_vidx.reserve(nv);
        lat_first_order_kernel<<<(unsigned)ceil_div(m, 256), 256, 0, st>>>(g_flags.p, g_gid.p, mc_vals.p, m, g_first.p,
                                                                          g_vidx.p);
        LVRHIP_LAUNCH_CHECK();
        lat_sort(g_first.p, g_vidx.p, nv, 32);
        g_vid.reserve(nv);
        lat_vertex_ids_kernel<<<(unsigned)ceil_div(nv, 256), 256, 0, st>>>(g_vidx.p, nv, g_vid.p);
        LVRHIP_LAUNCH_CHECK();
        mesh_faces.reserve(m);
        lat_cells_kernel<<<(unsigned)ceil_div(m, 256), 256, 0, st>>>(g_gid.p, mc_vals.p, g_vid.p, m, mesh_faces.p);
        LVRHIP_LAUNCH_CHECK();
        mesh_vtx.reserve(3 * nv);
        mc_vertices_kernel<<<(unsigned)ceil_div(nv, 256), 256, 0, st>>>(g_first.p, nv, mc_fcell.p, mc_off.p, mc_case.p,
                                                                        mc_keep.p, grid_cells.p, grid_vtx.p, d_dist,
                                                                        mesh_vtx.p);
        LVRHIP_LAUNCH_CHECK();
        LVRHIP_CHECK(hipStreamSynchronize(st));
        mesh_nv = nv;
        mesh_nf = nf;
        if (pmc) optimize_planar_faces(nc, 5);
    }

    // BilinearFastBox::optimizePlanarFaces(mesh, kc) over every cell (BilinearFastBox.tcc:140-212):
    // border edges from the edge multiplicities, per-vertex move counts, then one round of centroid
    // moves per count level.
    void optimize_planar_faces(int64_t nc, int kc)
    {
        const int64_t nf = mesh_nf, nv = mesh_nv, m = 3 * nf;
        mc_keys.reserve(m);
        mc_vals.reserve(m);
        g_k1.reserve(m);
        g_v1.reserve(m);
        mc_face_edges_kernel<<<(unsigned)ceil_div(m, 256), 256, 0, st>>>(mesh_faces.p, nf, mc_keys.p, mc_vals.p);
        LVRHIP_LAUNCH_CHECK();
        lat_sort(mc_keys.p, mc_vals.p, m);
        mc_border.reserve(m);
        mc_border_kernel<<<(unsigned)ceil_div(m, 256), 256, 0, st>>>(mc_keys.p, mc_vals.p, m, mc_border.p);
        LVRHIP_LAUNCH_CHECK();
        mc_moves.reserve(nv);
        LVRHIP_CHECK(hipMemsetAsync(mc_moves.p, 0, (size_t)nv * sizeof(uint32_t), st));
        LVRHIP_CHECK(hipMemsetAsync(counter(11), 0, 2 * sizeof(uint32_t), st));
        mc_moves_kernel<<<(unsigned)ceil_div(nc, 256), 256, 0, st>>>(mc_off.p, mc_ntri.p, nc, mc_border.p,
                                                                    mesh_faces.p, mc_moves.p, counter(11));
        LVRHIP_LAUNCH_CHECK();
        const uint32_t rounds = read_counter(11);
        mc_nbr.reserve((size_t)nv * kc);
        for (uint32_t r = 1; r <= rounds; r++) {
            LVRHIP_CHECK(hipMemsetAsync(counter(12), 0, sizeof(uint32_t), st));
            mc_list.reserve(nv);
            mc_q.reserve(3 * nv);
            mc_round_list_kernel<<<(unsigned)ceil_div(nv, 256), 256, 0, st>>>(mc_moves.p, nv, r, mesh_vtx.p,
                                                                             mc_list.p, mc_q.p, counter(12));
            LVRHIP_LAUNCH_CHECK();
            const int64_t mr = read_counter(12);
            if (mr == 0) break;
            mc_c.reserve(3 * mr);
            // a vertex's later rounds start from its previous round's neighbours
            centroids_device(mc_q.p, mr, kc, mc_c.p, mc_list.p, mc_nbr.p, r > 1);
            mc_round_scatter_kernel<<<(unsigned)ceil_div(mr, 256), 256, 0, st>>>(mc_list.p, mr, mc_c.p, mesh_vtx.p);
            LVRHIP_LAUNCH_CHECK();
            mesh_moves += mr;
        }
        LVRHIP_CHECK(hipStreamSynchronize(st));
    }

    // Reverse pass of distances: d_grad (n x 3, input order) += d(sum_v w_v proj_v)/d(points)
    // with neighbour sets, adaptive k and flip signs held fixed.
    void distances_backward_device(const float* d_q, int64_t m, const float* d_w, float* d_grad)
    {
        if (!has_normals) throw Error(LVRHIP_ESTATE, "distances_backward called before calculate_normals");
        require_one_scan("distances_backward");
        if (method != LVRHIP_METHOD_PCA) throw Error(LVRHIP_EINVAL, "the adjoint is provided for the PCA method only");
        if (interpolated) throw Error(LVRHIP_ESTATE, "the adjoint needs uninterpolated normals (ki = 0)");
        if (m == 0 || n == 0) return;
        drop_graph();
        const int64_t R = npad;
        const int t0 = clamp_k(kd);
        s_k1.reserve(std::max<size_t>((size_t)m * t0, (size_t)m));
        s_spos.reserve((size_t)m * t0);
        s_ord.reserve((size_t)m * t0);
        g_p.reserve(3 * (size_t)R);
        g_n.reserve(3 * (size_t)R);
        adj.reserve((size_t)R * ADJ_STRIDE);
        a_flags.reserve(R);
        a_sel.reserve(R);
        a_list.reserve(R);
        zero_f32_kernel<<<(unsigned)ceil_div(3 * R, 256), 256, 0, st>>>(g_p.p, 3 * R);
        LVRHIP_LAUNCH_CHECK();
        zero_f32_kernel<<<(unsigned)ceil_div(3 * R, 256), 256, 0, st>>>(g_n.p, 3 * R);
        LVRHIP_LAUNCH_CHECK();
        sort_queries(d_q, m);
        // 1. SDF reverse: dL/dp (through the means) and dL/dn of every vertex's kd neighbours
        {
            TileParams prm = params(t0, 0, 2);
            prm.tiles_per_wave = sdf_tiles_per_wave;
            prm.r2_margin = sdf_r2_margin;
            prm.grow_min = sdf_grow_min;
            prm.stats = kstats_ptr(6);
            QueryView qv{s_qx.p, s_qy.p, s_qz.p, nullptr, (uint32_t)m, nullptr};
            auto mk = [&](auto tag) {
                constexpr int P = decltype(tag)::value;
                SinkSdfAdjoint<P> s{};
                s.px = px.p;
                s.py = py.p;
                s.pz = pz.p;
                s.nx = nx.p;
                s.ny = ny.p;
                s.nz = nz.p;
                s.qorig = s_v0.p;
                s.w = d_w;
                s.keys = s_k1.p;
                s.spos = s_spos.p;
                s.ord = s_ord.p;
                s.kw = t0;
                s.gpx = g_p.p;
                s.gpy = g_p.p + R;
                s.gpz = g_p.p + 2 * R;
                s.gnx = g_n.p;
                s.gny = g_n.p + R;
                s.gnz = g_n.p + 2 * R;
                s.npts = (uint32_t)R;
                return s;
            };
            tile<SinkSdfAdjoint>(qv, prm, lanes_per_query, mk(std::integral_constant<int, 2>()),
                                 mk(std::integral_constant<int, 4>()), mk(std::integral_constant<int, 8>()));
        }
        // 2. dL/dC per point with a normal gradient
        pca_adjoint_prep_kernel<<<(unsigned)ceil_div(R, 256), 256, 0, st>>>(
            R, px.p, py.p, pz.p, nx.p, ny.p, nz.p, mom.p, kround.p, kn, nscans == 1 ? n : (int64_t)1 << 40, g_n.p,
            g_n.p + R, g_n.p + 2 * R, adj.p, a_flags.p);
        LVRHIP_LAUNCH_CHECK();
        // 3. PCA reverse, one launch per adaptive round (the forward's k-sets)
        for (int r = 1; r <= 5; r++) {
            select_round_kernel<<<(unsigned)ceil_div(R, 256), 256, 0, st>>>(a_flags.p, R, (uint32_t)(1 + r), a_sel.p);
            LVRHIP_LAUNCH_CHECK();
            compact_flags(a_sel.p, R, nullptr, nullptr, a_list.p, counter(6));
            TileParams prm = params(clamp_k((int64_t)kn << r), 0, 3);
            prm.stats = kstats_ptr(7);
            QueryView qv{px.p, py.p, pz.p, a_list.p, (uint32_t)R, counter(6)};
            auto mk = [&](auto tag) {
                constexpr int P = decltype(tag)::value;
                SinkPcaAdjoint<P> s{};
                s.adj = adj.p;
                s.gpx = g_p.p;
                s.gpy = g_p.p + R;
                s.gpz = g_p.p + 2 * R;
                return s;
            };
            tile<SinkPcaAdjoint>(qv, prm, lanes_per_query, mk(std::integral_constant<int, 2>()),
                                 mk(std::integral_constant<int, 4>()), mk(std::integral_constant<int, 8>()));
        }
        accumulate_grad_kernel<<<(unsigned)ceil_div(R, 256), 256, 0, st>>>(g_p.p, g_p.p + R, g_p.p + 2 * R, pidx.p, R,
                                                                         d_grad);
        LVRHIP_LAUNCH_CHECK();
        check_err();
        prof.collect();
    }

    bool best_first(int t) const { return search_mode != LVRHIP_SEARCH_TILE && t >= 1 && t <= BF_KMAX; }

    // Best-first search of the m sorted queries s_q*: one lane per query for all of them, then one
    // wave per query for those whose lane stack overflowed; the queries both left (redo flags
    // set by Out) end up in s_redo_list[0, *counter(8)), for the tile kernel.  Enqueue only.
    template <class Out>
    void enqueue_bf(int64_t m, int t0, uint64_t empty, const Out& o, int stats_group)
    {
        s_redo.reserve(m);
        s_redo_list.reserve(m);
        // A small batch goes to the wave-per-query search at once: the lane search would run it on
        // a few waves, each as slow as its slowest lane, while every query here gets a wave.
        if (search_mode == LVRHIP_SEARCH_WAVE || (search_mode == LVRHIP_SEARCH_AUTO && m <= BF_WAVE_BLOCKS)) {
            fill_u32_kernel<<<(unsigned)ceil_div(m, 256), 256, 0, st>>>(s_redo.p, m, 1u);
            LVRHIP_LAUNCH_CHECK();
        } else {
            uint32_t* q = nullptr;
            if (use_queue) {
                d_queue.reserve(QUEUE_SEGS * QUEUE_STRIDE);
                q = d_queue.p;
            }
            // one spill area per wave of the grid
            const int64_t waves = q ? std::min(ceil_div(m, BF_THREADS), bf_resident) : ceil_div(m, BF_THREADS);
            s_spill.reserve((size_t)waves * BF_SPILL * BF_THREADS);
            launch_bf(view(), s_qx.p, s_qy.p, s_qz.p, 1, m, t0, empty, o, st, kstats_ptr(stats_group), q, bf_resident,
                      s_spill.p, BF_SPILL);
        }
        compact_flags(s_redo.p, m, nullptr, nullptr, s_redo_list.p, counter(8));
        launch_bf_wave(view(), s_qx.p, s_qy.p, s_qz.p, 1, s_redo_list.p, counter(8), t0, empty, o, st,
                       kstats_ptr(stats_group == 6 ? 5 : stats_group));
        compact_flags(s_redo.p, m, nullptr, nullptr, s_redo_list.p, counter(8));
    }

    // Number of queries the last enqueue_bf left to the tile kernel (synchronises).
    uint32_t bf_redo_count()
    {
        uint32_t c = 0;
        LVRHIP_CHECK(hipMemcpyAsync(&c, counter(8), sizeof(c), hipMemcpyDeviceToHost, st));
        LVRHIP_CHECK(hipStreamSynchronize(st));
        return c;
    }

    // Sorted kSearch rows (input index, d^2; padded with -1 / +inf) for m device queries.  With
    // r2 >= 0 (radiusSearch) a row keeps only the neighbours with d^2 < r2, and d_count[row] is
    // their number.
    void knn_device(const float* d_q, int64_t m, int k, int32_t* d_idx, float* d_d2, float r2 = -1.0f,
                    int32_t* d_count = nullptr)
    {
        require_one_scan("kSearch");
        if (m == 0) return;
        drop_graph();
        const int t0 = (int)std::min<int64_t>(k, n);
        if (t0 == 0) {
            sort_rows_kernel<<<(unsigned)m, kWave, 0, st>>>(nullptr, m, k, 0, d_idx, d_d2);
            LVRHIP_LAUNCH_CHECK();
            if (d_count) LVRHIP_CHECK(hipMemsetAsync(d_count, 0, (size_t)m * sizeof(int32_t), st));
            LVRHIP_CHECK(hipStreamSynchronize(st));
            return;
        }
        sort_queries(d_q, m);
        if (best_first(t0)) {
            s_redo.reserve(m);
            const uint64_t empty = r2 >= 0.0f ? ((uint64_t)host_fbits(r2) << 32) : BF_EMPTY;
            BfRowsOut o{s_v0.p, k, d_idx, d_d2, d_count, s_redo.p};
            enqueue_bf(m, t0, empty, o, 7);
            if (bf_redo_count() == 0) {
                check_err();
                return;
            }
        }
        // the tile kernel for every query (also when a best-first stack overflowed)
        DBuf<uint64_t> keys;
        keys.reserve((size_t)m * k);
        DBuf<float> d2tmp;
        if (r2 >= 0.0f && !d_d2) {
            d2tmp.reserve((size_t)m * k);
            d_d2 = d2tmp.p;
        }
        TileParams prm = params(t0, 0, 3);
        prm.stats = kstats_ptr(7);
        QueryView qv{s_qx.p, s_qy.p, s_qz.p, nullptr, (uint32_t)m, nullptr};
        auto mk = [&](auto tag) {
            constexpr int P = decltype(tag)::value;
            SinkList<P> s{};
            s.out = keys.p;
            s.qorig = s_v0.p;
            s.kw = k;
            return s;
        };
        tile<SinkList>(qv, prm, lanes_per_query, mk(std::integral_constant<int, 2>()),
                       mk(std::integral_constant<int, 4>()), mk(std::integral_constant<int, 8>()));
        sort_rows_kernel<<<(unsigned)m, kWave, 0, st>>>(keys.p, m, k, t0, d_idx, d_d2);
        LVRHIP_LAUNCH_CHECK();
        if (r2 >= 0.0f) {
            radius_filter_kernel<<<(unsigned)ceil_div(m, 256), 256, 0, st>>>(m, k, r2, d_idx, d_d2, d_count);
            LVRHIP_LAUNCH_CHECK();
        }
        check_err();
    }

    // The centroid of the kc cloud points nearest to each of m device positions (the vertex move
    // of BilinearFastBox::optimizePlanarFaces, BilinearFastBox.tcc:168-204); d_out in caller order.
    // vid / nbr (optional): each query's mesh vertex and, per vertex, the neighbours of its last
    // search, saved here and, with prefill, the start of this search.
    void centroids_device(const float* d_q, int64_t m, int kc, float* d_out, const uint32_t* vid = nullptr,
                          uint32_t* nbr = nullptr, bool prefill = false)
    {
        require_one_scan("the mesh");
        if (m == 0) return;
        if (n == 0) throw Error(LVRHIP_ESTATE, "the mesh needs a non-empty point cloud");
        const int t0 = clamp_k(kc);
        if (best_first(t0)) {
            sort_queries(d_q, m);
            s_redo.reserve(m);
            BfCentroidOut o{pts4.p, s_v0.p, d_out, s_redo.p, vid, nbr, prefill, (uint32_t)npad};
            enqueue_bf(m, t0, BF_EMPTY, o, 7);
            if (bf_redo_count() == 0) {
                check_err();
                return;
            }
        }
        DBuf<int32_t> idx;
        idx.reserve((size_t)m * t0);
        const int saved = search_mode;
        search_mode = LVRHIP_SEARCH_TILE;
        try {
            knn_device(d_q, m, t0, idx.p, nullptr);
        } catch (...) {
            search_mode = saved;
            throw;
        }
        search_mode = saved;
        centroid_rows_kernel<<<(unsigned)ceil_div(m, 256), 256, 0, st>>>(m, t0, idx.p, rank.p, pts4.p, d_q, d_out);
        LVRHIP_LAUNCH_CHECK();
        LVRHIP_CHECK(hipStreamSynchronize(st));
    }

    // PointsetSurface::getInterpolatedNormal (PointsetSurface.tcc:61-74) at m device positions.
    void interpolated_normals_device(const float* d_q, int64_t m, float* d_out)
    {
        if (!has_normals) throw Error(LVRHIP_ESTATE, "interpolated normals requested before calculate_normals");
        require_one_scan("getInterpolatedNormal");
        if (ki < 1) throw Error(LVRHIP_EINVAL, "getInterpolatedNormal needs ki >= 1");
        if (m == 0) return;
        if (n == 0) throw Error(LVRHIP_ESTATE, "getInterpolatedNormal on an empty point cloud");
        drop_graph();
        const int t0 = clamp_k(ki);
        if (best_first(t0)) {
            sort_queries(d_q, m);
            s_redo.reserve(m);
            BfMeanNormalOut o{nx.p, ny.p, nz.p, s_v0.p, ki, d_out, s_redo.p};
            enqueue_bf(m, t0, BF_EMPTY, o, 7);
            if (bf_redo_count() == 0) {
                check_err();
                return;
            }
        }
        DBuf<int32_t> idx;
        idx.reserve((size_t)m * t0);
        const int saved = search_mode;
        search_mode = LVRHIP_SEARCH_TILE;
        try {
            knn_device(d_q, m, t0, idx.p, nullptr);
        } catch (...) {
            search_mode = saved;
            throw;
        }
        search_mode = saved;
        mean_normal_rows_kernel<<<(unsigned)ceil_div(m, 256), 256, 0, st>>>(m, t0, t0, idx.p, rank.p, nx.p, ny.p,
                                                                          nz.p, ki, d_out);
        LVRHIP_LAUNCH_CHECK();
        LVRHIP_CHECK(hipStreamSynchronize(st));
    }
};

namespace {

template <typename F>
int guarded(F&& f)
{
    try {
        f();
        return LVRHIP_OK;
    } catch (const Error& e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_last_error = "host out of memory";
        return LVRHIP_ENOMEM;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return LVRHIP_EHIP;
    }
}

#define REQUIRE(cond, msg)                                           \
    do {                                                             \
        if (!(cond)) throw Error(LVRHIP_EINVAL, std::string(msg));   \
    } while (0)

// the same check in a function that returns the status itself
#define REQUIRE_C(cond, msg)              \
    do {                                  \
        if (!(cond)) {                    \
            g_last_error = (msg);         \
            return LVRHIP_EINVAL;         \
        }                                 \
    } while (0)

}  // namespace

extern "C" {

const char* lvrhip_last_error(void) { return g_last_error.c_str(); }

int lvrhip_create_scans_device(const float* d_xyz, const int64_t* scan_offsets, int nscans, int device, void* stream,
                               lvrhip_t* out)
{
    return guarded([&] {
        REQUIRE(out != nullptr, "out handle is NULL");
        REQUIRE(scan_offsets != nullptr, "scan offsets pointer is NULL");
        REQUIRE(nscans >= 1 && nscans <= 65535, "number of scans must be in [1, 65535]");
        REQUIRE(scan_offsets[0] == 0, "scan offsets must start at 0");
        int64_t padded = 0;
        for (int s = 0; s < nscans; s++) {
            REQUIRE(scan_offsets[s + 1] >= scan_offsets[s], "scan offsets must be non-decreasing");
            padded += (scan_offsets[s + 1] - scan_offsets[s] + LEAF - 1) / LEAF * LEAF;
        }
        const int64_t n = scan_offsets[nscans];
        REQUIRE(padded < (int64_t)0xFFFFFFFFu - 64, "more than 2^32-65 points (with padding)");
        REQUIRE(n == 0 || d_xyz != nullptr, "points pointer is NULL");
        DeviceGuard dg(device);
        auto* h = new lvrhip_surface;
        h->device = device;
        h->st = (hipStream_t)stream;
        try {
            int cus = 0;
            if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
            {
                h->resident_waves = (int64_t)cus * 16;
                // 12 best-first waves per CU (3 per SIMD: their registers; LDS would allow 20)
                h->bf_resident = (int64_t)cus * 12;
            }
            h->init(d_xyz, scan_offsets, nscans);
            h->build();
            LVRHIP_CHECK(hipStreamSynchronize(h->st));
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

int lvrhip_create_device(const float* d_xyz, size_t n, int device, void* stream, lvrhip_t* out)
{
    REQUIRE_C(n < (size_t)0xFFFFFFFFu - 64, "more than 2^32-65 points");
    const int64_t off[2] = {0, (int64_t)n};
    return lvrhip_create_scans_device(d_xyz, off, 1, device, stream, out);
}

int lvrhip_create_scans(const float* xyz, const int64_t* scan_offsets, int nscans, int device, lvrhip_t* out)
{
    return guarded([&] {
        REQUIRE(out != nullptr, "out handle is NULL");
        REQUIRE(scan_offsets != nullptr && nscans >= 1, "scan table missing");
        const int64_t n = scan_offsets[nscans];
        REQUIRE(n >= 0, "negative point count");
        REQUIRE(n == 0 || xyz != nullptr, "points pointer is NULL");
        DeviceGuard dg(device);
        auto own = new DBuf<float>;
        try {
            own->reserve(std::max<size_t>(3 * (size_t)n, 1));
            if (n) LVRHIP_CHECK(hipMemcpy(own->p, xyz, 3 * (size_t)n * sizeof(float), hipMemcpyHostToDevice));
            int rc = lvrhip_create_scans_device(own->p, scan_offsets, nscans, device, nullptr, out);
            if (rc != LVRHIP_OK) throw Error(rc, g_last_error);
            // the handle keeps the device copy as the source of later rebuilds
            std::swap((*out)->own_src.p, own->p);
            std::swap((*out)->own_src.cap, own->cap);
        } catch (...) {
            delete own;
            throw;
        }
        delete own;
    });
}

int lvrhip_create(const float* xyz, size_t n, int device, lvrhip_t* out)
{
    REQUIRE_C(n < (size_t)0xFFFFFFFFu - 64, "more than 2^32-65 points");
    const int64_t off[2] = {0, (int64_t)n};
    return lvrhip_create_scans(xyz, off, 1, device, out);
}

int lvrhip_destroy(lvrhip_t h)
{
    return guarded([&] {
        if (!h) return;
        DeviceGuard dg(h->device);
        (void)hipStreamSynchronize(h->st);
        delete h;
    });
}

int lvrhip_set_k(lvrhip_t h, int kn, int ki, int kd)
{
    return guarded([&] {
        REQUIRE(h, "NULL handle");
        REQUIRE(kn >= 1 && kn <= (1 << 20), "kn out of range");
        REQUIRE(ki >= 0 && ki <= (1 << 20), "ki out of range");
        REQUIRE(kd >= 1 && kd <= (1 << 20), "kd out of range");
        if (ki != h->ki || kn != h->kn) h->ki_valid = false;
        h->kn = kn;
        h->ki = ki;
        h->kd = kd;
    });
}

int lvrhip_set_flip_point(lvrhip_t h, float x, float y, float z)
{
    return guarded([&] {
        REQUIRE(h, "NULL handle");
        h->flip[0] = x;
        h->flip[1] = y;
        h->flip[2] = z;
        h->scan_flips = false;
    });
}

int lvrhip_set_method(lvrhip_t h, int method)
{
    return guarded([&] {
        REQUIRE(h, "NULL handle");
        REQUIRE(method == LVRHIP_METHOD_LS_PLANE || method == LVRHIP_METHOD_ITERATIVE || method == LVRHIP_METHOD_PCA,
                "method must be 0, 2 or 3");
        h->method = method;
    });
}

int lvrhip_set_stream(lvrhip_t h, void* stream)
{
    return guarded([&] {
        REQUIRE(h, "NULL handle");
        h->st = (hipStream_t)stream;
    });
}

int lvrhip_calculate_normals(lvrhip_t h)
{
    return guarded([&] {
        REQUIRE(h, "NULL handle");
        DeviceGuard dg(h->device);
        h->calculate_normals();
    });
}

int lvrhip_interpolate_normals(lvrhip_t h)
{
    return guarded([&] {
        REQUIRE(h, "NULL handle");
        DeviceGuard dg(h->device);
        h->interpolate_normals();
    });
}

int lvrhip_set_normals(lvrhip_t h, const float* normals)
{
    return guarded([&] {
        REQUIRE(h, "NULL handle");
        REQUIRE(h->n == 0 || normals, "normals pointer is NULL");
        DeviceGuard dg(h->device);
        const int64_t N = h->n;
        h->drop_graph();
        h->ki_valid = false;
        h->has_normals = false;
        h->interpolated = false;
        if (N > 0) {
            const int64_t R = h->npad;
            h->nx.reserve(R);
            h->ny.reserve(R);
            h->nz.reserve(R);
            DBuf<float> d;
            d.reserve(3 * (size_t)N);
            LVRHIP_CHECK(hipMemcpyAsync(d.p, normals, 3 * (size_t)N * sizeof(float), hipMemcpyHostToDevice, h->st));
            gather_normals_kernel<<<(unsigned)ceil_div(R, 256), 256, 0, h->st>>>(d.p, h->pidx.p, R, h->nx.p, h->ny.p,
                                                                               h->nz.p);
            LVRHIP_LAUNCH_CHECK();
            LVRHIP_CHECK(hipStreamSynchronize(h->st));
        }
        h->has_normals = true;
    });
}

int lvrhip_get_normals_device(lvrhip_t h, float* d_out)
{
    return guarded([&] {
        REQUIRE(h, "NULL handle");
        REQUIRE(h->has_normals, "no normals: call lvrhip_calculate_normals first");
        if (h->n == 0) return;
        REQUIRE(d_out, "output pointer is NULL");
        DeviceGuard dg(h->device);
        h->scatter_normals(h->nx.p, h->ny.p, h->nz.p, d_out);
    });
}

int lvrhip_get_normals(lvrhip_t h, float* out)
{
    return guarded([&] {
        REQUIRE(h, "NULL handle");
        REQUIRE(h->has_normals, "no normals: call lvrhip_calculate_normals first");
        if (h->n == 0) return;
        REQUIRE(out, "output pointer is NULL");
        DeviceGuard dg(h->device);
        DBuf<float> d;
        d.reserve(3 * (size_t)h->n);
        int rc = lvrhip_get_normals_device(h, d.p);
        if (rc != LVRHIP_OK) throw Error(rc, g_last_error);
        LVRHIP_CHECK(hipMemcpyAsync(out, d.p, 3 * (size_t)h->n * sizeof(float), hipMemcpyDeviceToHost, h->st));
        LVRHIP_CHECK(hipStreamSynchronize(h->st));
    });
}

int lvrhip_knn(lvrhip_t h, const float* q, size_t m, int k, int32_t* out_idx, float* out_d2)
{
    return guarded([&] {
        REQUIRE(h, "NULL handle");
        REQUIRE(k >= 1 && k <= (1 << 16), "k out of range");
        REQUIRE(m == 0 || (q && out_idx), "NULL query or output pointer");
        REQUIRE(m < (size_t)0x7FFFFFFF, "too many queries");
        if (m == 0) return;
        DeviceGuard dg(h->device);
        DBuf<float> dq;
        DBuf<int32_t> di;
        DBuf<float> dd;
        dq.reserve(3 * m);
        di.reserve(m * (size_t)k);
        if (out_d2) dd.reserve(m * (size_t)k);
        LVRHIP_CHECK(hipMemcpyAsync(dq.p, q, 3 * m * sizeof(float), hipMemcpyHostToDevice, h->st));
        h->knn_device(dq.p, (int64_t)m, k, di.p, out_d2 ? dd.p : nullptr);
        LVRHIP_CHECK(hipMemcpyAsync(out_idx, di.p, m * (size_t)k * sizeof(int32_t), hipMemcpyDeviceToHost, h->st));
        if (out_d2)
            LVRHIP_CHECK(hipMemcpyAsync(out_d2, dd.p, m * (size_t)k * sizeof(float), hipMemcpyDeviceToHost, h->st));
        LVRHIP_CHECK(hipStreamSynchronize(h->st));
    });
}

int lvrhip_radius_search(lvrhip_t h, const float* q, size_t m, int k, float r, int32_t* out_idx, float* out_d2,
                         int32_t* out_count)
{
    return guarded([&] {
        REQUIRE(h, "NULL handle");
        REQUIRE(k >= 1 && k <= (1 << 16), "k out of range");
        REQUIRE(r >= 0.0f, "radius must be a non-negative number");
        REQUIRE(m == 0 || (q && out_idx), "NULL query or output pointer");
        REQUIRE(m < (size_t)0x7FFFFFFF, "too many queries");
        if (m == 0) return;
        DeviceGuard dg(h->device);
        DBuf<float> dq, dd;
        DBuf<int32_t> di, dc;
        dq.reserve(3 * m);
        di.reserve(m * (size_t)k);
        dd.reserve(m * (size_t)k);
        dc.reserve(m);
        LVRHIP_CHECK(hipMemcpyAsync(dq.p, q, 3 * m * sizeof(float), hipMemcpyHostToDevice, h->st));
        h->knn_device(dq.p, (int64_t)m, k, di.p, dd.p, r * r, dc.p);
        LVRHIP_CHECK(hipMemcpyAsync(out_idx, di.p, m * (size_t)k * sizeof(int32_t), hipMemcpyDeviceToHost, h->st));
        if (out_d2)
            LVRHIP_CHECK(hipMemcpyAsync(out_d2, dd.p, m * (size_t)k * sizeof(float), hipMemcpyDeviceToHost, h->st));
        if (out_count)
            LVRHIP_CHECK(hipMemcpyAsync(out_count, dc.p, m * sizeof(int32_t), hipMemcpyDeviceToHost, h->st));
        LVRHIP_CHECK(hipStreamSynchronize(h->st));
    });
}

int lvrhip_interpolated_normals(lvrhip_t h, const float* q, size_t m, float* out_xyz)
{
    return guarded([&] {
        REQUIRE(h, "NULL handle");
        REQUIRE(m == 0 || (q && out_xyz), "NULL query or output pointer");
        REQUIRE(m < (size_t)0x7FFFFFFF, "too many queries");
        DeviceGuard dg(h->device);
        if (m == 0) {
            h->interpolated_normals_device(nullptr, 0, nullptr);
            return;
        }
        DBuf<float> dq, dn;
        dq.reserve(3 * m);
        dn.reserve(3 * m);
        LVRHIP_CHECK(hipMemcpyAsync(dq.p, q, 3 * m * sizeof(float), hipMemcpyHostToDevice, h->st));
        h->interpolated_normals_device(dq.p, (int64_t)m, dn.p);
        LVRHIP_CHECK(hipMemcpyAsync(out_xyz, dn.p, 3 * m * sizeof(float), hipMemcpyDeviceToHost, h->st));
        LVRHIP_CHECK(hipStreamSynchronize(h->st));
    });
}

int lvrhip_set_search_mode(lvrhip_t h, int mode)
{
    return guarded([&] {
        REQUIRE(h, "NULL handle");
        REQUIRE(mode == LVRHIP_SEARCH_AUTO || mode == LVRHIP_SEARCH_TILE || mode == LVRHIP_SEARCH_WAVE,
                "search mode must be 0, 1 or 2");
        h->search_mode = mode;
    });
}

int lvrhip_distances_device(lvrhip_t h, const float* d_q, size_t m, float voxel, float* d_proj, float* d_eucl,
                            uint8_t* d_invalid)
{
    return guarded([&] {
        REQUIRE(h, "NULL handle");
        REQUIRE(m == 0 || d_q, "NULL query pointer");
        REQUIRE(m < (size_t)0xFFFFFFFFu, "too many queries");
        DeviceGuard dg(h->device);
        h->distances_device(d_q, (int64_t)m, voxel, d_proj, d_eucl, d_invalid);
    });
}

int lvrhip_distances_device_noalloc(lvrhip_t h, const float* d_q, size_t m, float voxel, float* d_proj,
                                    float* d_eucl, uint8_t* d_invalid)
{
    return guarded([&] {
        REQUIRE(h, "NULL handle");
        REQUIRE(m == 0 || d_q, "NULL query pointer");
        REQUIRE(m <= (size_t)NA_MAX_QUERIES, "the no-allocation distances take at most " +
                                                 std::to_string(NA_MAX_QUERIES) + " queries per call");
        DeviceGuard dg(h->device);
        h->distances_noalloc(d_q, (int64_t)m, voxel, d_proj, d_eucl, d_invalid, na_scratch(h->device));
    });
}

int lvrhip_distances(lvrhip_t h, const float* q, size_t m, float voxel, float* out_proj, float* out_eucl,
                     uint8_t* out_invalid)
{
    return guarded([&] {
        REQUIRE(h, "NULL handle");
        REQUIRE(m == 0 || q, "NULL query pointer");
        if (m == 0) return;
        DeviceGuard dg(h->device);
        DBuf<float> dq, dp, de;
        DBuf<uint8_t> di;
        dq.reserve(3 * m);
        dp.reserve(m);
        de.reserve(m);
        di.reserve(m);
        LVRHIP_CHECK(hipMemcpyAsync(dq.p, q, 3 * m * sizeof(float), hipMemcpyHostToDevice, h->st));
        h->distances_device(dq.p, (int64_t)m, voxel, dp.p, de.p, di.p);
        if (out_proj) LVRHIP_CHECK(hipMemcpyAsync(out_proj, dp.p, m * sizeof(float), hipMemcpyDeviceToHost, h->st));
        if (out_eucl) LVRHIP_CHECK(hipMemcpyAsync(out_eucl, de.p, m * sizeof(float), hipMemcpyDeviceToHost, h->st));
        if (out_invalid)
            LVRHIP_CHECK(hipMemcpyAsync(out_invalid, di.p, m * sizeof(uint8_t), hipMemcpyDeviceToHost, h->st));
        LVRHIP_CHECK(hipStreamSynchronize(h->st));
    });
}

int lvrhip_distances_backward_device(lvrhip_t h, const float* d_q, size_t m, const float* d_grad_proj,
                                     float* d_grad_xyz)
{
    return guarded([&] {
        REQUIRE(h, "NULL handle");
        REQUIRE(m == 0 || (d_q && d_grad_proj), "NULL query or gradient pointer");
        REQUIRE(h->n == 0 || d_grad_xyz, "NULL output gradient pointer");
        REQUIRE(m < (size_t)0xFFFFFFFFu, "too many queries");
        DeviceGuard dg(h->device);
        h->distances_backward_device(d_q, (int64_t)m, d_grad_proj, d_grad_xyz);
    });
}

int lvrhip_normals_stats(lvrhip_t h, int64_t n_round[5])
{
    return guarded([&] {
        REQUIRE(h && n_round, "NULL argument");
        for (int i = 0; i < 5; i++) n_round[i] = h->n_round[i];
    });
}

int lvrhip_get_adaptive_k(lvrhip_t h, int32_t* out_k)
{
    return guarded([&] {
        REQUIRE(h, "NULL handle");
        REQUIRE(h->has_normals && h->kround.p, "no adaptive rounds: call lvrhip_calculate_normals first");
        if (h->n == 0) return;
        REQUIRE(out_k, "output pointer is NULL");
        DeviceGuard dg(h->device);
        DBuf<int32_t> d;
        d.reserve((size_t)h->n);
        adaptive_k_kernel<<<(unsigned)ceil_div(h->npad, 256), 256, 0, h->st>>>(
            h->kround.p, h->pidx.p, h->npad, h->kn, h->nscans == 1 ? h->n : ((int64_t)1 << 40), d.p);
        LVRHIP_LAUNCH_CHECK();
        LVRHIP_CHECK(hipMemcpyAsync(out_k, d.p, (size_t)h->n * sizeof(int32_t), hipMemcpyDeviceToHost, h->st));
        LVRHIP_CHECK(hipStreamSynchronize(h->st));
    });
}

int lvrhip_grid_build(lvrhip_t h, float voxel, int extrude, size_t* out_nv, size_t* out_nc)
{
    return guarded([&] {
        REQUIRE(h, "NULL handle");
        DeviceGuard dg(h->device);
        h->build_grid(voxel, extrude != 0);
        if (out_nv) *out_nv = (size_t)h->grid_nv;
        if (out_nc) *out_nc = (size_t)h->grid_nc;
    });
}

int lvrhip_grid_build_box(lvrhip_t h, float voxel, int extrude, const float* box, size_t* out_nv, size_t* out_nc)
{
    return guarded([&] {
        REQUIRE(h && box, "NULL argument");
        DeviceGuard dg(h->device);
        h->build_grid(voxel, extrude != 0, box);
        if (out_nv) *out_nv = (size_t)h->grid_nv;
        if (out_nc) *out_nc = (size_t)h->grid_nc;
    });
}

int lvrhip_grid_get_device(lvrhip_t h, float* d_vertices, int32_t* d_cells, int32_t* d_cell_index)
{
    return guarded([&] {
        REQUIRE(h, "NULL handle");
        DeviceGuard dg(h->device);
        if (d_vertices && h->grid_nv)
            LVRHIP_CHECK(hipMemcpyAsync(d_vertices, h->grid_vtx.p, 3 * (size_t)h->grid_nv * sizeof(float),
                                        hipMemcpyDeviceToDevice, h->st));
        if (d_cells && h->grid_nc)
            LVRHIP_CHECK(hipMemcpyAsync(d_cells, h->grid_cells.p, 8 * (size_t)h->grid_nc * sizeof(int32_t),
                                        hipMemcpyDeviceToDevice, h->st));
        if (d_cell_index && h->grid_nc)
            LVRHIP_CHECK(hipMemcpyAsync(d_cell_index, h->grid_cidx.p, 3 * (size_t)h->grid_nc * sizeof(int32_t),
                                        hipMemcpyDeviceToDevice, h->st));
    });
}

int lvrhip_grid_get(lvrhip_t h, float* vertices, int32_t* cells, int32_t* cell_index)
{
    return guarded([&] {
        REQUIRE(h, "NULL handle");
        DeviceGuard dg(h->device);
        if (vertices && h->grid_nv)
            LVRHIP_CHECK(hipMemcpyAsync(vertices, h->grid_vtx.p, 3 * (size_t)h->grid_nv * sizeof(float),
                                        hipMemcpyDeviceToHost, h->st));
        if (cells && h->grid_nc)
            LVRHIP_CHECK(hipMemcpyAsync(cells, h->grid_cells.p, 8 * (size_t)h->grid_nc * sizeof(int32_t),
                                        hipMemcpyDeviceToHost, h->st));
        if (cell_index && h->grid_nc)
            LVRHIP_CHECK(hipMemcpyAsync(cell_index, h->grid_cidx.p, 3 * (size_t)h->grid_nc * sizeof(int32_t),
                                        hipMemcpyDeviceToHost, h->st));
        LVRHIP_CHECK(hipStreamSynchronize(h->st));
    });
}

int lvrhip_mesh_build(lvrhip_t h, const float* d_dist, const uint8_t* d_invalid, size_t nv, int decomposition,
                      size_t* nverts, size_t* nfaces)
{
    return guarded([&] {
        REQUIRE(h, "NULL handle");
        REQUIRE(nv == (size_t)h->grid_nv, "the distances must hold one value per lattice vertex (" +
                                              std::to_string(h->grid_nv) + "), got " + std::to_string(nv));
        REQUIRE(h->grid_nc == 0 || d_dist, "NULL distance pointer");
        REQUIRE(decomposition == LVRHIP_MESH_MC || decomposition == LVRHIP_MESH_PMC,
                "decomposition must be LVRHIP_MESH_MC or LVRHIP_MESH_PMC");
        DeviceGuard dg(h->device);
        h->build_mesh(d_dist, d_invalid, decomposition);
        if (nverts) *nverts = (size_t)h->mesh_nv;
        if (nfaces) *nfaces = (size_t)h->mesh_nf;
    });
}

int lvrhip_mesh_build_host(lvrhip_t h, const float* dist, const uint8_t* invalid, size_t nv, int decomposition,
                           size_t* nverts, size_t* nfaces)
{
    return guarded([&] {
        REQUIRE(h, "NULL handle");
        REQUIRE(nv == (size_t)h->grid_nv, "the distances must hold one value per lattice vertex (" +
                                              std::to_string(h->grid_nv) + "), got " + std::to_string(nv));
        REQUIRE(h->grid_nc == 0 || dist, "NULL distance pointer");
        REQUIRE(decomposition == LVRHIP_MESH_MC || decomposition == LVRHIP_MESH_PMC,
                "decomposition must be LVRHIP_MESH_MC or LVRHIP_MESH_PMC");
        DeviceGuard dg(h->device);
        DBuf<float> dd;
        DBuf<uint8_t> di;
        if (nv) {
            dd.reserve(nv);
            LVRHIP_CHECK(hipMemcpyAsync(dd.p, dist, nv * sizeof(float), hipMemcpyHostToDevice, h->st));
            if (invalid) {
                di.reserve(nv);
                LVRHIP_CHECK(hipMemcpyAsync(di.p, invalid, nv, hipMemcpyHostToDevice, h->st));
            }
        }
        h->build_mesh(dd.p, invalid ? di.p : nullptr, decomposition);
        if (nverts) *nverts = (size_t)h->mesh_nv;
        if (nfaces) *nfaces = (size_t)h->mesh_nf;
    });
}

int lvrhip_mesh_get_device(lvrhip_t h, float* d_vertices, int32_t* d_faces)
{
    return guarded([&] {
        REQUIRE(h, "NULL handle");
        DeviceGuard dg(h->device);
        if (d_vertices && h->mesh_nv)
            LVRHIP_CHECK(hipMemcpyAsync(d_vertices, h->mesh_vtx.p, 3 * (size_t)h->mesh_nv * sizeof(float),
                                        hipMemcpyDeviceToDevice, h->st));
        if (d_faces && h->mesh_nf)
            LVRHIP_CHECK(hipMemcpyAsync(d_faces, h->mesh_faces.p, 3 * (size_t)h->mesh_nf * sizeof(int32_t),
                                        hipMemcpyDeviceToDevice, h->st));
    });
}

int lvrhip_mesh_get(lvrhip_t h, float* vertices, int32_t* faces)
{
    return guarded([&] {
        REQUIRE(h, "NULL handle");
        DeviceGuard dg(h->device);
        if (vertices && h->mesh_nv)
            LVRHIP_CHECK(hipMemcpyAsync(vertices, h->mesh_vtx.p, 3 * (size_t)h->mesh_nv * sizeof(float),
                                        hipMemcpyDeviceToHost, h->st));
        if (faces && h->mesh_nf)
            LVRHIP_CHECK(hipMemcpyAsync(faces, h->mesh_faces.p, 3 * (size_t)h->mesh_nf * sizeof(int32_t),
                                        hipMemcpyDeviceToHost, h->st));
        LVRHIP_CHECK(hipStreamSynchronize(h->st));
    });
}

int lvrhip_size(lvrhip_t h, size_t* n)
{
    return guarded([&] {
        REQUIRE(h && n, "NULL argument");
        *n = (size_t)h->n;
    });
}

int lvrhip_set_profiling(lvrhip_t h, int on)
{
    return guarded([&] {
        REQUIRE(h, "NULL handle");
        DeviceGuard dg(h->device);
        h->prof.collect();
        h->prof.on = on != 0;
    });
}

int lvrhip_get_kernel_stats(lvrhip_t h, int64_t out[128], int reset)
{
    return guarded([&] {
        REQUIRE(h && out, "NULL argument");
        DeviceGuard dg(h->device);
        unsigned long long v[lvrhip_surface::KS_GROUPS * KS_COUNT] = {};
        if (h->kstats.p) {
            LVRHIP_CHECK(hipMemcpyAsync(v, h->kstats.p, sizeof(v), hipMemcpyDeviceToHost, h->st));
            LVRHIP_CHECK(hipStreamSynchronize(h->st));
            if (reset) {
                LVRHIP_CHECK(hipMemsetAsync(h->kstats.p, 0, sizeof(v), h->st));
                LVRHIP_CHECK(hipStreamSynchronize(h->st));
            }
        }
        for (int i = 0; i < lvrhip_surface::KS_GROUPS * KS_COUNT; i++) out[i] = (int64_t)v[i];
    });
}

int lvrhip_set_tuning(lvrhip_t h, int lanes_per_query, int tiles_per_wave, float r2_margin)
{
    return guarded([&] {
        REQUIRE(h, "NULL handle");
        if (lanes_per_query > 0) {
            REQUIRE(lanes_per_query == 2 || lanes_per_query == 4 || lanes_per_query == 8,
                    "lanes per query must be 2, 4 or 8");
            h->lanes_per_query = lanes_per_query;
        }
        if (tiles_per_wave > 0) {
            REQUIRE(tiles_per_wave <= 4096, "tiles per wave out of range");
            h->tiles_per_wave = tiles_per_wave;
        }
        if (r2_margin > 0.0f) {
            REQUIRE(r2_margin >= 1.0f && r2_margin <= 1e6f, "radius margin out of range");
            h->r2_margin = r2_margin;
        }
    });
}

int lvrhip_get_timings(lvrhip_t h, double* ms, int64_t* launches, int reset)
{
    return guarded([&] {
        REQUIRE(h, "NULL handle");
        DeviceGuard dg(h->device);
        h->prof.collect();
        for (int i = 0; i < LVRHIP_T_COUNT; i++) {
            if (ms) ms[i] = h->prof.ms[i];
            if (launches) launches[i] = h->prof.launches[i];
            if (reset) {
                h->prof.ms[i] = 0.0;
                h->prof.launches[i] = 0;
            }
        }
    });
}

int lvrhip_num_scans(lvrhip_t h, int* nscans)
{
    return guarded([&] {
        REQUIRE(h && nscans, "NULL argument");
        *nscans = h->nscans;
    });
}

int lvrhip_set_scan_flip_points(lvrhip_t h, const float* xyz)
{
    return guarded([&] {
        REQUIRE(h && xyz, "NULL argument");
        DeviceGuard dg(h->device);
        h->drop_graph();
        h->d_flips.reserve(3 * (size_t)h->nscans);
        LVRHIP_CHECK(
            hipMemcpyAsync(h->d_flips.p, xyz, 3 * (size_t)h->nscans * sizeof(float), hipMemcpyHostToDevice, h->st));
        LVRHIP_CHECK(hipStreamSynchronize(h->st));
        h->scan_flips = h->nscans > 1;
        if (h->nscans == 1) {
            h->flip[0] = xyz[0];
            h->flip[1] = xyz[1];
            h->flip[2] = xyz[2];
        }
    });
}

int lvrhip_rebuild(lvrhip_t h)
{
    return guarded([&] {
        REQUIRE(h, "NULL handle");
        DeviceGuard dg(h->device);
        h->has_normals = false;
        h->ki_valid = false;
        h->build();
        LVRHIP_CHECK(hipStreamSynchronize(h->st));
        h->prof.collect();
    });
}

int lvrhip_graph_capture(lvrhip_t h, float* d_out_normals, int rebuild)
{
    return guarded([&] {
        REQUIRE(h, "NULL handle");
        REQUIRE(d_out_normals, "output pointer is NULL");
        DeviceGuard dg(h->device);
        h->graph_capture(d_out_normals, rebuild != 0);
    });
}

int lvrhip_graph_launch(lvrhip_t h)
{
    return guarded([&] {
        REQUIRE(h, "NULL handle");
        DeviceGuard dg(h->device);
        h->graph_launch();
    });
}

int lvrhip_synchronize(lvrhip_t h)
{
    return guarded([&] {
        REQUIRE(h, "NULL handle");
        DeviceGuard dg(h->device);
        h->synchronize();
    });
}

int lvrhip_autotune(lvrhip_t h, int* lanes_per_query, int* tiles_per_wave, float* r2_margin)
{
    return guarded([&] {
        REQUIRE(h, "NULL handle");
        DeviceGuard dg(h->device);
        h->autotune(lanes_per_query, tiles_per_wave, r2_margin);
    });
}

int lvrhip_set_sdf_tuning(lvrhip_t h, int tiles_per_wave, float r2_margin)
{
    return guarded([&] {
        REQUIRE(h, "NULL handle");
        if (tiles_per_wave > 0) {
            REQUIRE(tiles_per_wave <= 4096, "tiles per wave out of range");
            h->sdf_tiles_per_wave = tiles_per_wave;
        }
        if (r2_margin > 0.0f) {
            REQUIRE(r2_margin >= 1.0f && r2_margin <= 1e6f, "radius margin out of range");
            h->sdf_r2_margin = r2_margin;
        }
    });
}

int lvrhip_set_retry_mode(lvrhip_t h, int merged)
{
    return guarded([&] {
        REQUIRE(h, "NULL handle");
        h->drop_graph();
        h->merge_retry = merged != 0;
    });
}

int lvrhip_get_retry_mode(lvrhip_t h, int* merged)
{
    return guarded([&] {
        REQUIRE(h && merged, "NULL argument");
        *merged = h->merge_retry ? 1 : 0;
    });
}

int lvrhip_get_tuning(lvrhip_t h, int* lanes_per_query, int* tiles_per_wave, float* r2_margin)
{
    return guarded([&] {
        REQUIRE(h, "NULL handle");
        if (lanes_per_query) *lanes_per_query = h->lanes_per_query;
        if (tiles_per_wave) *tiles_per_wave = h->tiles_per_wave;
        if (r2_margin) *r2_margin = h->r2_margin;
    });
}

}  // extern "C"
