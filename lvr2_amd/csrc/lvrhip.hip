// lvrhip: C ABI over the gfx950 kernels (declarations and reference citations in
// include/lvrhip.h).  One handle = one point cloud resident on one device:
// Morton-sorted SoA points, the LBVH, and the normals once computed.
#include "common.hpp"
#include "radix_sort.hpp"
#include "lbvh.hpp"
#include "knn_tile.hpp"
#include "sinks.hpp"

#include "lvrhip.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

using namespace lvrhip;

namespace {

thread_local std::string g_last_error;

template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t cap = 0;
    void reserve(size_t n)
    {
        if (n <= cap) return;
        release();
        if (n) LVRHIP_CHECK(hipMalloc(&p, n * sizeof(T)));
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

__global__ void morton_queries_kernel(const float* __restrict__ q, int64_t m, MortonFrame f,
                                      uint64_t* __restrict__ keys, uint32_t* __restrict__ vals)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    keys[i] = morton63(f, q[3 * i], q[3 * i + 1], q[3 * i + 2]);
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
                                                                     uint32_t* __restrict__ bsum)
{
    __shared__ uint32_t ws[CF_THREADS / kWave];
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
                                                                     const uint32_t* __restrict__ boff,
                                                                     const uint32_t* __restrict__ qsel,
                                                                     uint32_t* __restrict__ out)
{
    __shared__ uint32_t wc[CF_THREADS / kWave];
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
// reference's kSearch result, normalised, into a fresh array.  G lanes per point: the
// kNN pass stored the ki list unordered, so each lane recomputes one neighbour's key, the
// list is ranked in LDS (and written back sorted), then one lane sums in order.
constexpr int INTERP_THREADS = 256;

template <int G>
__global__ __launch_bounds__(INTERP_THREADS) void interpolate_kernel(
    const float* __restrict__ nx, const float* __restrict__ ny, const float* __restrict__ nz,
    const float* __restrict__ px, const float* __restrict__ py, const float* __restrict__ pz,
    const uint32_t* __restrict__ pidx, uint32_t* __restrict__ ki_list, int kiw, int64_t n, float* __restrict__ ox,
    float* __restrict__ oy, float* __restrict__ oz)
{
    __shared__ uint64_t skey[INTERP_THREADS];
    __shared__ uint32_t sord[INTERP_THREADS];
    const int grp = threadIdx.x / G, j = threadIdx.x % G, base = grp * G;
    const int64_t i = (int64_t)blockIdx.x * (INTERP_THREADS / G) + grp;
    const bool valid = i < n;
    uint32_t s = 0;
    uint64_t key = ~0ull;
    if (valid && j < kiw) {
        s = ki_list[(size_t)i * kiw + j];
        if (s < n) key = make_key(d2_contract(px[i], py[i], pz[i], px[s], py[s], pz[s]), pidx[s]);
        skey[base + j] = key;
    }
    __syncthreads();
    if (valid && j < kiw) {
        int rank = 0;
        for (int t = 0; t < kiw; t++) rank += skey[base + t] < key ? 1 : 0;
        sord[base + rank] = s;
        ki_list[(size_t)i * kiw + rank] = s;
    }
    __syncthreads();
    if (valid && j == 0) {
        float x = nx[i], y = ny[i], z = nz[i];
        for (int r = 0; r < kiw; r++) {
            const uint32_t t = sord[base + r];
            if (t >= n) continue;
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

__global__ void gather_normals_kernel(const float* __restrict__ in, const uint32_t* __restrict__ pidx, int64_t n,
                                      float* __restrict__ nx, float* __restrict__ ny, float* __restrict__ nz)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t j = pidx[i];
    nx[i] = in[3 * j];
    ny[i] = in[3 * j + 1];
    nz[i] = in[3 * j + 2];
}

__global__ void scatter_normals_kernel(const float* __restrict__ nx, const float* __restrict__ ny,
                                       const float* __restrict__ nz, const uint32_t* __restrict__ pidx, int64_t n,
                                       float* __restrict__ out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t j = pidx[i];
    out[3 * j] = nx[i];
    out[3 * j + 1] = ny[i];
    out[3 * j + 2] = nz[i];
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

// ki list without normals (used when ki > 2*kn, so the first normal round cannot supply it);
// unordered, like the one written by SinkNormals.
template <int P>
struct SinkKiOnly {
    uint32_t* out;
    int kw;
    uint32_t qpos_;
    __device__ void begin(int, int, int64_t, uint32_t qpos, float, float, float) { qpos_ = qpos; }
    __device__ void accept(int tau, int g, int, float, float, float, uint64_t, uint32_t sp, uint32_t* cnt)
    {
        if (tau != 0) return;
        const uint32_t slot = atomicAdd(&cnt[g], 1u);
        if (slot < (uint32_t)kw) out[(size_t)qpos_ * kw + slot] = sp;
    }
    __device__ void finish(int, int, bool, uint32_t, float, float, float, const uint32_t*, const uint32_t*) {}
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

template <int P, template <int> class SinkT>
void launch_tile_p(const BvhView& bv, const QueryView& qv, const TileParams& prm, const SinkT<P>& sink,
                   hipStream_t st)
{
    constexpr int Q = kWave / P;
    const int64_t ntiles = ceil_div(qv.nq, Q);
    const int64_t grid = ceil_div(ntiles, prm.tiles_per_wave);
    if (grid == 0) return;
    knn_tile_kernel<P, SinkT<P>><<<(unsigned)grid, kWave, 0, st>>>(bv, qv, prm, sink);
    LVRHIP_LAUNCH_CHECK();
}

}  // namespace

struct lvrhip_surface {
    int device = 0;
    hipStream_t st = nullptr;
    int64_t n = 0;
    int kn = 10, ki = 10, kd = 5, method = LVRHIP_METHOD_PCA;
    float flip[3] = {1e6f, 1e6f, 1e6f};
    int lanes_per_query = 8;
    int tiles_per_wave = 8;
    float r2_margin = 1.3f;

    DBuf<float> px, py, pz;
    DBuf<uint32_t> pidx, rank;
    DBuf<float4> leaf_lo, leaf_hi;
    DBuf<BvhNode> nodes;
    int nleaf = 0;
    float bbox[6] = {0, 0, 0, 0, 0, 0};
    MortonFrame frame{};
    float area_pp = 0.0f;

    DBuf<float> nx, ny, nz, tx, ty, tz;
    bool has_normals = false;
    DBuf<uint32_t> ki_list;
    int kiw = 0;
    bool ki_valid = false;
    DBuf<uint8_t> kround;
    DBuf<float> mom;  // per point: moments of its k-set about the point (9 floats)
    DBuf<uint32_t> s_bsum;
    int64_t n_round[5] = {0, 0, 0, 0, 0};
    DBuf<int> err;
    Profiler prof;
    DBuf<unsigned long long> kstats;  // 4 groups x KS_COUNT

    unsigned long long* kstats_ptr(int group)
    {
        if (!prof.on) return nullptr;
        if (!kstats.p) {
            kstats.reserve(4 * KS_COUNT);
            LVRHIP_CHECK(hipMemsetAsync(kstats.p, 0, 4 * KS_COUNT * sizeof(unsigned long long), st));
        }
        return kstats.p + group * KS_COUNT;
    }

    // scratch
    DBuf<uint64_t> s_k0, s_k1;
    DBuf<uint32_t> s_v0, s_v1, s_hist, s_flags, s_scan, s_list0, s_list1, s_spos, s_ord;
    DBuf<float> s_qx, s_qy, s_qz, s_f;
    DBuf<int32_t> s_idx;

    BvhView view() const
    {
        BvhView v;
        v.px = px.p;
        v.py = py.p;
        v.pz = pz.p;
        v.pidx = pidx.p;
        v.leaf_lo = leaf_lo.p;
        v.leaf_hi = leaf_hi.p;
        v.nodes = nodes.p;
        v.nleaf = nleaf;
        v.n = (uint32_t)n;
        return v;
    }

    float r2_init(int k) const
    {
        float r2 = area_pp * (float)k / 3.14159265f * r2_margin;
        return r2 > 1e-30f ? r2 : 1e-30f;
    }

    void check_err()
    {
        int e = 0;
        LVRHIP_CHECK(hipMemcpyAsync(&e, err.p, sizeof(int), hipMemcpyDeviceToHost, st));
        LVRHIP_CHECK(hipStreamSynchronize(st));
        if (e != 0) {
            int z = 0;
            LVRHIP_CHECK(hipMemcpyAsync(err.p, &z, sizeof(int), hipMemcpyHostToDevice, st));
            LVRHIP_CHECK(hipStreamSynchronize(st));
            throw Error(LVRHIP_ECAPACITY, e == 5 ? "kNN traversal stack overflow" : "kNN radius search did not converge");
        }
    }

    template <template <int> class SinkT>
    void tile(const QueryView& qv, const TileParams& prm, int P, SinkT<2> s2, SinkT<4> s4, SinkT<8> s8)
    {
        switch (P) {
            case 2: launch_tile_p<2, SinkT>(view(), qv, prm, s2, st); break;
            case 4: launch_tile_p<4, SinkT>(view(), qv, prm, s4, st); break;
            case 8: launch_tile_p<8, SinkT>(view(), qv, prm, s8, st); break;
            default: throw Error(LVRHIP_EINVAL, "lanes per query must be 2, 4 or 8");
        }
    }

    void build(const float* d_xyz)
    {
        if (n <= 0) return;
        const int64_t N = n;
        // bounding box
        const int nparts = 256;
        s_f.reserve(6 * nparts + 6);
        bbox_partial_kernel<<<nparts, 256, 0, st>>>(d_xyz, N, s_f.p);
        LVRHIP_LAUNCH_CHECK();
        bbox_final_kernel<<<1, 64, 0, st>>>(s_f.p, nparts, s_f.p + 6 * nparts);
        LVRHIP_LAUNCH_CHECK();
        LVRHIP_CHECK(hipMemcpyAsync(bbox, s_f.p + 6 * nparts, 6 * sizeof(float), hipMemcpyDeviceToHost, st));
        LVRHIP_CHECK(hipStreamSynchronize(st));
        float ext = std::max(bbox[3] - bbox[0], std::max(bbox[4] - bbox[1], bbox[5] - bbox[2]));
        frame.lo[0] = bbox[0];
        frame.lo[1] = bbox[1];
        frame.lo[2] = bbox[2];
        frame.scale = ext > 0.0f ? 2097151.0f / ext : 1.0f;
        // Morton sort
        s_k0.reserve(N);
        s_k1.reserve(N);
        pidx.reserve(N);
        s_v1.reserve(N);
        s_hist.reserve((size_t)RS_RADIX * ceil_div(N, RS_TILE));
        const unsigned gb = (unsigned)ceil_div(N, 256);
        morton_aos_kernel<<<gb, 256, 0, st>>>(d_xyz, N, frame, s_k0.p, pidx.p);
        LVRHIP_LAUNCH_CHECK();
        radix_sort_pairs(s_k0.p, pidx.p, s_k1.p, s_v1.p, N, s_hist.p, st);
        px.reserve(N);
        py.reserve(N);
        pz.reserve(N);
        rank.reserve(N);
        gather_sorted_kernel<<<gb, 256, 0, st>>>(d_xyz, N, pidx.p, px.p, py.p, pz.p, rank.p);
        LVRHIP_LAUNCH_CHECK();
        // leaves and hierarchy
        nleaf = (int)ceil_div(N, LEAF);
        leaf_lo.reserve(nleaf);
        leaf_hi.reserve(nleaf);
        s_k1.reserve(nleaf);
        leaf_boxes_kernel<<<gb, 256, 0, st>>>(px.p, py.p, pz.p, N, s_k0.p, nleaf, leaf_lo.p, leaf_hi.p, s_k1.p);
        LVRHIP_LAUNCH_CHECK();
        const int nint = std::max(nleaf - 1, 1);
        nodes.reserve(nint);
        DBuf<int> leaf_parent;
        leaf_parent.reserve(nleaf);
        fill_u32_kernel<<<(unsigned)ceil_div(nleaf, 256), 256, 0, st>>>((uint32_t*)leaf_parent.p, nleaf, ~0u);
        LVRHIP_LAUNCH_CHECK();
        DBuf<float4> nlo, nhi;
        nlo.reserve(nint);
        nhi.reserve(nint);
        s_flags.reserve(nint);
        fill_u32_kernel<<<(unsigned)ceil_div(nint, 256), 256, 0, st>>>(s_flags.p, nint, 0u);
        LVRHIP_LAUNCH_CHECK();
        if (nleaf > 1) {
            karras_kernel<<<(unsigned)ceil_div(nleaf - 1, 256), 256, 0, st>>>(s_k1.p, nleaf, nodes.p, leaf_parent.p);
            LVRHIP_LAUNCH_CHECK();
            refit_kernel<<<(unsigned)ceil_div(nleaf, 256), 256, 0, st>>>(nleaf, leaf_parent.p, leaf_lo.p, leaf_hi.p,
                                                                        nodes.p, nlo.p, nhi.p, s_flags.p);
            LVRHIP_LAUNCH_CHECK();
        }
        // density scale for the first radius guess: median leaf area per point
        const int nsample = std::min(nleaf, 8192);
        const int stride = std::max(1, nleaf / nsample);
        s_f.reserve(std::max<size_t>(nsample, 6 * nparts + 6));
        leaf_area_kernel<<<(unsigned)ceil_div(nsample, 256), 256, 0, st>>>(leaf_lo.p, leaf_hi.p, nleaf, stride,
                                                                          nsample, s_f.p);
        LVRHIP_LAUNCH_CHECK();
        std::vector<float> a(nsample);
        LVRHIP_CHECK(hipMemcpyAsync(a.data(), s_f.p, nsample * sizeof(float), hipMemcpyDeviceToHost, st));
        LVRHIP_CHECK(hipStreamSynchronize(st));
        std::nth_element(a.begin(), a.begin() + nsample / 2, a.end());
        area_pp = a[nsample / 2];
        err.reserve(1);
        int z = 0;
        LVRHIP_CHECK(hipMemcpyAsync(err.p, &z, sizeof(int), hipMemcpyHostToDevice, st));
        LVRHIP_CHECK(hipStreamSynchronize(st));
    }

    // out = the entries i < nq with flags[i] != 0 (qsel[i] or i), in order; *count_dev = count
    void compact_flags(const uint32_t* flags, int64_t nq, const uint32_t* qsel, uint32_t* out, uint32_t* count_dev)
    {
        const int64_t nb = ceil_div(nq, CF_TILE);
        s_bsum.reserve(nb + 1);
        flag_block_sums_kernel<<<(unsigned)nb, CF_THREADS, 0, st>>>(flags, nq, s_bsum.p);
        LVRHIP_LAUNCH_CHECK();
        scan_exclusive_kernel<<<1, 1024, 0, st>>>(s_bsum.p, nb, count_dev);
        LVRHIP_LAUNCH_CHECK();
        compact_scatter_kernel<<<(unsigned)nb, CF_THREADS, 0, st>>>(flags, nq, s_bsum.p, qsel, out);
        LVRHIP_LAUNCH_CHECK();
    }

    void calculate_normals()
    {
        const int64_t N = n;
        has_normals = false;
        ki_valid = false;
        for (auto& c : n_round) c = 0;
        if (N == 0) {
            has_normals = true;
            return;
        }
        nx.reserve(N);
        ny.reserve(N);
        nz.reserve(N);
        kround.reserve(N);
        mom.reserve((size_t)N * 9);
        kiw = (ki > 0 && ki <= 2 * kn) ? (int)std::min<int64_t>(ki, N) : 0;
        if (kiw > 0) ki_list.reserve((size_t)N * kiw);
        s_flags.reserve(N);
        s_scan.reserve(N + 1);
        s_list0.reserve(N);
        s_list1.reserve(N);
        const uint32_t* qsel = nullptr;
        int64_t nq = N;
        uint32_t* lists[2] = {s_list0.p, s_list1.p};
        for (int r = 1; r <= 5; r++) {
            const int64_t k = (int64_t)kn << r;
            const int t0 = (int)std::min<int64_t>(k, N);
            TileParams prm{};
            prm.t0 = t0;
            prm.t1 = (r == 1) ? kiw : 0;
            prm.tiles_per_wave = tiles_per_wave;
            prm.r2_init = r2_init(t0);
            prm.r2_margin = r2_margin;
            prm.err = err.p;
            prm.stats = kstats_ptr(r == 1 ? 0 : 1);
            QueryView qv{px.p, py.p, pz.p, qsel, (uint32_t)nq};
            auto mk = [&](auto tag) {
                constexpr int P = decltype(tag)::value;
                SinkNormals<P> s{};
                s.mom = mom.p;
                s.retry = r < 5 ? s_flags.p : nullptr;
                s.ki_list = ki_list.p;
                s.kiw = prm.t1;
                return s;
            };
            const int tok = prof.start(r == 1 ? LVRHIP_T_NORMALS_FIRST : LVRHIP_T_NORMALS_RETRY, st);
            tile<SinkNormals>(qv, prm, lanes_per_query, mk(std::integral_constant<int, 2>()),
                              mk(std::integral_constant<int, 4>()), mk(std::integral_constant<int, 8>()));
            plane_fit_kernel<<<(unsigned)ceil_div(nq, 256), 256, 0, st>>>(
                qsel, nq, r < 5 ? s_flags.p : nullptr, mom.p, px.p, py.p, pz.p, method, t0, flip[0], flip[1], flip[2],
                nx.p, ny.p, nz.p, kround.p, r);
            LVRHIP_LAUNCH_CHECK();
            n_round[r - 1] = nq;
            if (r == 5) {
                prof.stop(tok, st);
                break;
            }
            // failing queries -> next list
            uint32_t* out = lists[r & 1];
            compact_flags(s_flags.p, nq, qsel, out, s_scan.p + N);
            prof.stop(tok, st);
            uint32_t cnt = 0;
            LVRHIP_CHECK(hipMemcpyAsync(&cnt, s_scan.p + N, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
            LVRHIP_CHECK(hipStreamSynchronize(st));
            if (cnt == 0) break;
            qsel = out;
            nq = cnt;
        }
        check_err();
        prof.collect();
        has_normals = true;
        ki_valid = kiw > 0;
    }

    void ensure_ki_list()
    {
        const int64_t N = n;
        const int kw = (int)std::min<int64_t>(ki, N);
        if (ki_valid && kiw == kw) return;
        kiw = kw;
        ki_list.reserve((size_t)N * kw);
        TileParams prm{};
        prm.t0 = kw;
        prm.t1 = 0;
        prm.tiles_per_wave = tiles_per_wave;
        prm.r2_init = r2_init(kw);
        prm.r2_margin = r2_margin;
        prm.err = err.p;
        prm.stats = kstats_ptr(3);
        QueryView qv{px.p, py.p, pz.p, nullptr, (uint32_t)N};
        auto mk = [&](auto tag) {
            constexpr int P = decltype(tag)::value;
            SinkKiOnly<P> s{};
            s.out = ki_list.p;
            s.kw = kw;
            return s;
        };
        tile<SinkKiOnly>(qv, prm, lanes_per_query, mk(std::integral_constant<int, 2>()),
                         mk(std::integral_constant<int, 4>()), mk(std::integral_constant<int, 8>()));
        check_err();
        ki_valid = true;
    }

    void interpolate_normals()
    {
        if (!has_normals) throw Error(LVRHIP_ESTATE, "interpolate_normals called before calculate_normals");
        if (ki <= 0 || n == 0) return;
        if (std::min<int64_t>(ki, n) > 64) throw Error(LVRHIP_EINVAL, "ki above 64 is not supported");
        ensure_ki_list();
        const int64_t N = n;
        tx.reserve(N);
        ty.reserve(N);
        tz.reserve(N);
        const int tok = prof.start(LVRHIP_T_INTERPOLATE, st);
        if (kiw <= 16) {
            interpolate_kernel<16><<<(unsigned)ceil_div(N, INTERP_THREADS / 16), INTERP_THREADS, 0, st>>>(
                nx.p, ny.p, nz.p, px.p, py.p, pz.p, pidx.p, ki_list.p, kiw, N, tx.p, ty.p, tz.p);
        } else if (kiw <= 32) {
            interpolate_kernel<32><<<(unsigned)ceil_div(N, INTERP_THREADS / 32), INTERP_THREADS, 0, st>>>(
                nx.p, ny.p, nz.p, px.p, py.p, pz.p, pidx.p, ki_list.p, kiw, N, tx.p, ty.p, tz.p);
        } else {
            interpolate_kernel<64><<<(unsigned)ceil_div(N, INTERP_THREADS / 64), INTERP_THREADS, 0, st>>>(
                nx.p, ny.p, nz.p, px.p, py.p, pz.p, pidx.p, ki_list.p, kiw, N, tx.p, ty.p, tz.p);
        }
        LVRHIP_LAUNCH_CHECK();
        prof.stop(tok, st);
        if (prof.on) prof.collect();
        std::swap(nx.p, tx.p);
        std::swap(nx.cap, tx.cap);
        std::swap(ny.p, ty.p);
        std::swap(ny.cap, ty.cap);
        std::swap(nz.p, tz.p);
        std::swap(nz.cap, tz.cap);
    }

    // Morton-sort m device queries; fills s_qx/s_qy/s_qz and s_v0 (sorted -> caller index).
    void sort_queries(const float* d_q, int64_t m)
    {
        s_k0.reserve(std::max<int64_t>(m, 1));
        s_k1.reserve(std::max<int64_t>(m, 1));
        s_v0.reserve(std::max<int64_t>(m, 1));
        s_v1.reserve(std::max<int64_t>(m, 1));
        s_hist.reserve((size_t)RS_RADIX * ceil_div(std::max<int64_t>(m, 1), RS_TILE));
        const unsigned gb = (unsigned)ceil_div(m, 256);
        morton_queries_kernel<<<gb, 256, 0, st>>>(d_q, m, frame, s_k0.p, s_v0.p);
        LVRHIP_LAUNCH_CHECK();
        radix_sort_pairs(s_k0.p, s_v0.p, s_k1.p, s_v1.p, m, s_hist.p, st);
        s_qx.reserve(m);
        s_qy.reserve(m);
        s_qz.reserve(m);
        gather_queries_kernel<<<gb, 256, 0, st>>>(d_q, m, s_v0.p, s_qx.p, s_qy.p, s_qz.p);
        LVRHIP_LAUNCH_CHECK();
    }

    void distances_device(const float* d_q, int64_t m, float voxel, float* d_proj, float* d_eucl, uint8_t* d_inv)
    {
        if (!has_normals) throw Error(LVRHIP_ESTATE, "distances called before calculate_normals");
        if (m == 0) return;
        if (n == 0) throw Error(LVRHIP_ESTATE, "distances on an empty point cloud");
        const int t0 = (int)std::min<int64_t>(kd, n);
        s_k1.reserve(std::max<size_t>((size_t)m * t0, (size_t)m));
        s_spos.reserve((size_t)m * t0);
        s_ord.reserve((size_t)m * t0);
        const int tsort = prof.start(LVRHIP_T_SDF_SORT, st);
        sort_queries(d_q, m);
        prof.stop(tsort, st);
        TileParams prm{};
        prm.t0 = t0;
        prm.t1 = 0;
        prm.tiles_per_wave = tiles_per_wave;
        prm.r2_init = r2_init(t0);
        prm.r2_margin = r2_margin;
        prm.err = err.p;
        prm.stats = kstats_ptr(2);
        QueryView qv{s_qx.p, s_qy.p, s_qz.p, nullptr, (uint32_t)m};
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
            s.npts = (uint32_t)n;
            return s;
        };
        const int tok = prof.start(LVRHIP_T_SDF_KERNEL, st);
        tile<SinkSdf>(qv, prm, lanes_per_query, mk(std::integral_constant<int, 2>()),
                      mk(std::integral_constant<int, 4>()), mk(std::integral_constant<int, 8>()));
        prof.stop(tok, st);
        check_err();
        prof.collect();
    }

    void knn_device(const float* d_q, int64_t m, int k, int32_t* d_idx, float* d_d2)
    {
        if (m == 0) return;
        const int t0 = (int)std::min<int64_t>(k, n);
        if (t0 > 0) {
            sort_queries(d_q, m);
            DBuf<uint64_t> keys;
            keys.reserve((size_t)m * k);
            TileParams prm{};
            prm.t0 = t0;
            prm.t1 = 0;
            prm.tiles_per_wave = tiles_per_wave;
            prm.r2_init = r2_init(t0);
            prm.r2_margin = r2_margin;
            prm.err = err.p;
            prm.stats = kstats_ptr(3);
            QueryView qv{s_qx.p, s_qy.p, s_qz.p, nullptr, (uint32_t)m};
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
            check_err();
        } else {
            sort_rows_kernel<<<(unsigned)m, kWave, 0, st>>>(nullptr, m, k, 0, d_idx, d_d2);
            LVRHIP_LAUNCH_CHECK();
            LVRHIP_CHECK(hipStreamSynchronize(st));
        }
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

}  // namespace

extern "C" {

const char* lvrhip_last_error(void) { return g_last_error.c_str(); }

int lvrhip_create_device(const float* d_xyz, size_t n, int device, void* stream, lvrhip_t* out)
{
    return guarded([&] {
        REQUIRE(out != nullptr, "out handle is NULL");
        REQUIRE(n < (size_t)0xFFFFFFFFu - 64, "more than 2^32-65 points");
        REQUIRE(n == 0 || d_xyz != nullptr, "points pointer is NULL");
        DeviceGuard dg(device);
        auto* h = new lvrhip_surface;
        h->device = device;
        h->st = (hipStream_t)stream;
        h->n = (int64_t)n;
        try {
            h->err.reserve(1);
            h->build(d_xyz);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

int lvrhip_create(const float* xyz, size_t n, int device, lvrhip_t* out)
{
    return guarded([&] {
        REQUIRE(out != nullptr, "out handle is NULL");
        REQUIRE(n == 0 || xyz != nullptr, "points pointer is NULL");
        DeviceGuard dg(device);
        DBuf<float> d;
        d.reserve(std::max<size_t>(3 * n, 1));
        if (n) LVRHIP_CHECK(hipMemcpy(d.p, xyz, 3 * n * sizeof(float), hipMemcpyHostToDevice));
        int rc = lvrhip_create_device(d.p, n, device, nullptr, out);
        if (rc != LVRHIP_OK) throw Error(rc, g_last_error);
        LVRHIP_CHECK(hipDeviceSynchronize());
    });
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
        h->ki_valid = false;
        h->has_normals = false;
        if (N > 0) {
            h->nx.reserve(N);
            h->ny.reserve(N);
            h->nz.reserve(N);
            DBuf<float> d;
            d.reserve(3 * (size_t)N);
            LVRHIP_CHECK(hipMemcpyAsync(d.p, normals, 3 * (size_t)N * sizeof(float), hipMemcpyHostToDevice, h->st));
            gather_normals_kernel<<<(unsigned)ceil_div(N, 256), 256, 0, h->st>>>(d.p, h->pidx.p, N, h->nx.p, h->ny.p,
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
        scatter_normals_kernel<<<(unsigned)ceil_div(h->n, 256), 256, 0, h->st>>>(h->nx.p, h->ny.p, h->nz.p, h->pidx.p,
                                                                                h->n, d_out);
        LVRHIP_LAUNCH_CHECK();
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

int lvrhip_normals_stats(lvrhip_t h, int64_t n_round[5])
{
    return guarded([&] {
        REQUIRE(h && n_round, "NULL argument");
        for (int i = 0; i < 5; i++) n_round[i] = h->n_round[i];
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

int lvrhip_get_kernel_stats(lvrhip_t h, int64_t out[64], int reset)
{
    return guarded([&] {
        REQUIRE(h && out, "NULL argument");
        DeviceGuard dg(h->device);
        unsigned long long v[4 * KS_COUNT] = {};
        if (h->kstats.p) {
            LVRHIP_CHECK(hipMemcpyAsync(v, h->kstats.p, sizeof(v), hipMemcpyDeviceToHost, h->st));
            LVRHIP_CHECK(hipStreamSynchronize(h->st));
            if (reset) {
                LVRHIP_CHECK(hipMemsetAsync(h->kstats.p, 0, sizeof(v), h->st));
                LVRHIP_CHECK(hipStreamSynchronize(h->st));
            }
        }
        for (int i = 0; i < 4 * KS_COUNT; i++) out[i] = (int64_t)v[i];
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

}  // extern "C"
