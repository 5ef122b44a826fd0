// LBVH over Morton-sorted points, built entirely on the device.
//
// Reference: lbvh::LBVHIndex::build (src/liblvr2/reconstruction/cuda/LBVHIndex.cu:75-324)
// builds a bottom-up (Apetrei-style) hierarchy over per-point AABBs after a host sort.
// Here: points are sorted by 63-bit Morton code (21 bits/axis, cubic normalisation to the
// cloud's bounding box), cut into fixed leaves of LEAF consecutive points, and a Karras
// (2012) radix tree is built top-down over the leaves' codes, with leaf indices
// breaking code ties; node boxes are then range unions of leaf boxes (each node covers a
// contiguous leaf range), level by level with no atomics.  Each internal node stores both
// children's boxes, so a traversal step reads one 64-byte node.
#pragma once

#include "common.hpp"

namespace lvrhip {

constexpr int LEAF = 16;

struct alignas(16) BvhNode {
    float4 a;  // lo0.x lo0.y lo0.z hi0.x
    float4 b;  // hi0.y hi0.z lo1.x lo1.y
    float4 c;  // lo1.z hi1.x hi1.y hi1.z
    int4 r;    // child0, child1 (>= 0 internal node, < 0 leaf = -1 - id), parent, far end of the leaf range
};

__device__ __forceinline__ int leaf_ref(int leaf) { return -1 - leaf; }

// The same hierarchy with every node's grandchildren as its children, for traversals that
// descend two levels per step: child k (0..3) has the box lo = (f[6k], f[6k+1], f[6k+2]),
// hi = (f[6k+3], f[6k+4], f[6k+5]) of f = the 24 floats of b.  A child of the binary node
// that is a leaf takes the first of its two slots; the second is empty (EMPTY4, an inverted
// box that no distance test passes).
struct alignas(16) BvhNode4 {
    float4 b[6];
    int4 r;
};
constexpr int EMPTY4 = (int)0x80000000;

// ---- Morton codes ---------------------------------------------------------------------

__device__ __forceinline__ uint64_t expand21(uint32_t v)
{
    uint64_t x = v & 0x1FFFFFu;
    x = (x | (x << 32)) & 0x1F00000000FFFFull;
    x = (x | (x << 16)) & 0x1F0000FF0000FFull;
    x = (x | (x << 8)) & 0x100F00F00F00F00Full;
    x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

struct MortonFrame {
    float lo[3];
    float scale;  // cells per unit, identical for the three axes
};

__device__ __forceinline__ uint64_t morton63(const MortonFrame& f, float x, float y, float z)
{
    const float top = 2097151.0f;  // 2^21 - 1
    float u = fminf(fmaxf((x - f.lo[0]) * f.scale, 0.0f), top);
    float v = fminf(fmaxf((y - f.lo[1]) * f.scale, 0.0f), top);
    float w = fminf(fmaxf((z - f.lo[2]) * f.scale, 0.0f), top);
    return (expand21((uint32_t)u) << 2) | (expand21((uint32_t)v) << 1) | expand21((uint32_t)w);
}

// ---- scans --------------------------------------------------------------------------------
// A cloud is one or more scans: input points [off[s], off[s+1]) belong to scan s.  In the
// sorted arrays each scan starts on a leaf boundary (pstart[s], a multiple of LEAF) and its
// last leaf is padded with NaN points, so no leaf holds points of two scans.

// largest s in [0, ns) with start[s] <= i (start ascending, empty scans repeat a value)
__device__ __forceinline__ int scan_of_pos(const int64_t* __restrict__ start, int ns, int64_t i)
{
    int lo = 0, hi = ns - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= i)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// float <-> order-preserving u32 (for atomicMin / atomicMax on boxes)
__device__ __forceinline__ uint32_t ord_enc(float f)
{
    const uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float ord_dec(uint32_t e)
{
    return __uint_as_float((e & 0x80000000u) ? (e & 0x7FFFFFFFu) : ~e);
}

// per scan box (encoded): box[6 s + 0..2] = min, [6 s + 3..5] = max; cleared by box_init_kernel
__global__ void box_init_kernel(uint32_t* __restrict__ box, int ns)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 6 * ns) box[i] = (i % 6) < 3 ? 0xFFFFFFFFu : 0u;
}

constexpr int SB_ITEMS = 16;  // points per lane

// One wave per 64 * SB_ITEMS consecutive input points: reduced in registers and merged with one
// atomic per coordinate when they belong to one scan, per point otherwise.
__global__ __launch_bounds__(256) void scan_bbox_kernel(const float* __restrict__ xyz, int64_t n,
                                                        const int64_t* __restrict__ off, int ns,
                                                        uint32_t* __restrict__ box)
{
    const int lane = lane_id();
    const int64_t w0 = (((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave) * (kWave * SB_ITEMS);
    if (w0 >= n) return;
    const int64_t w1 = w0 + kWave * SB_ITEMS < n ? w0 + kWave * SB_ITEMS : n;
    const int s0 = scan_of_pos(off, ns, w0), s1 = scan_of_pos(off, ns, w1 - 1);
    if (s0 == s1) {
        float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int it = 0; it < SB_ITEMS; it++) {
            const int64_t i = w0 + (int64_t)it * kWave + lane;
            if (i < w1) {
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    const float v = xyz[3 * i + j];
                    mn[j] = fminf(mn[j], v);
                    mx[j] = fmaxf(mx[j], v);
                }
            }
        }
#pragma unroll
        for (int j = 0; j < 3; j++) {
            mn[j] = wave_min(mn[j]);
            mx[j] = wave_max(mx[j]);
        }
        if (lane < 3) {
            atomicMin(&box[6 * s0 + lane], ord_enc(lane == 0 ? mn[0] : lane == 1 ? mn[1] : mn[2]));
            atomicMax(&box[6 * s0 + 3 + lane], ord_enc(lane == 0 ? mx[0] : lane == 1 ? mx[1] : mx[2]));
        }
    } else {
        for (int it = 0; it < SB_ITEMS; it++) {
            const int64_t i = w0 + (int64_t)it * kWave + lane;
            if (i < w1) {
                const int sc = scan_of_pos(off, ns, i);
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    const float v = xyz[3 * i + j];
                    if (v == v) {
                        atomicMin(&box[6 * sc + j], ord_enc(v));
                        atomicMax(&box[6 * sc + 3 + j], ord_enc(v));
                    }
                }
            }
        }
    }
}

// Morton frame of each scan: its box corner, a cubic cell size over the largest extent.
__global__ void scan_frames_kernel(const uint32_t* __restrict__ box, int ns, MortonFrame* __restrict__ frames)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    float lo[3], ext = 0.0f;
    for (int j = 0; j < 3; j++) {
        lo[j] = ord_dec(box[6 * s + j]);
        const float e = ord_dec(box[6 * s + 3 + j]) - lo[j];
        ext = e > ext ? e : ext;  // empty scans: NaN / -inf extents leave ext = 0
    }
    MortonFrame f;
    for (int j = 0; j < 3; j++) f.lo[j] = lo[j] == lo[j] && fabsf(lo[j]) < INFINITY ? lo[j] : 0.0f;
    f.scale = ext > 0.0f ? 2097151.0f / ext : 1.0f;
    frames[s] = f;
}

// Sort keys of the input points: the 63-bit Morton code in the scan's own frame; with more
// than one scan, the scan index in the top 16 bits above the code's leading 48 bits.
__global__ __launch_bounds__(256) void morton_scan_kernel(const float* __restrict__ xyz, int64_t n,
                                                          const int64_t* __restrict__ off, int ns,
                                                          const MortonFrame* __restrict__ frames,
                                                          uint64_t* __restrict__ keys, uint32_t* __restrict__ vals)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int sc = ns > 1 ? scan_of_pos(off, ns, i) : 0;
    const uint64_t m = morton63(frames[sc], xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
    keys[i] = ns > 1 ? (((uint64_t)sc << 48) | (m >> 15)) : m;
    vals[i] = (uint32_t)i;
}

// Padded sorted SoA layout, one thread per padded position p: the point (NaN for padding),
// its input index (~0 for padding), the inverse permutation, the sort key, the list of real
// positions in sort order, and the scan of each position.
__global__ __launch_bounds__(256) void gather_padded_kernel(
    const float* __restrict__ xyz, const int64_t* __restrict__ off, const int64_t* __restrict__ pstart, int ns,
    int64_t npad, const uint32_t* __restrict__ order, const uint64_t* __restrict__ skeys, float* __restrict__ px,
    float* __restrict__ py, float* __restrict__ pz, uint32_t* __restrict__ pidx, uint32_t* __restrict__ rank,
    uint64_t* __restrict__ pcode, uint32_t* __restrict__ qreal, uint16_t* __restrict__ scan_of,
    float4* __restrict__ pts4)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npad) return;
    const int sc = ns > 1 ? scan_of_pos(pstart, ns, p) : 0;
    const int64_t local = p - pstart[sc];
    const int64_t cnt = off[sc + 1] - off[sc];
    if (local < cnt) {
        const int64_t r = off[sc] + local;
        const uint32_t j = order[r];
        px[p] = xyz[3 * (int64_t)j];
        py[p] = xyz[3 * (int64_t)j + 1];
        pz[p] = xyz[3 * (int64_t)j + 2];
        pidx[p] = j;
        rank[j] = (uint32_t)p;
        pcode[p] = skeys[r];
        qreal[r] = (uint32_t)p;
        pts4[p] = make_float4(xyz[3 * (int64_t)j], xyz[3 * (int64_t)j + 1], xyz[3 * (int64_t)j + 2], __uint_as_float(j));
    } else {
        px[p] = py[p] = pz[p] = __builtin_nanf("");
        pidx[p] = ~0u;
        pts4[p] = make_float4(__builtin_nanf(""), __builtin_nanf(""), __builtin_nanf(""), __uint_as_float(~0u));
        pcode[p] = cnt > 0 ? skeys[off[sc] + cnt - 1] : 0ull;
    }
    if (scan_of) scan_of[p] = (uint16_t)sc;
}

// Leaf boxes: one lane per point, LEAF lanes per leaf; also the leaf codes.
__global__ __launch_bounds__(256) void leaf_boxes_kernel(const float* __restrict__ px, const float* __restrict__ py,
                                                         const float* __restrict__ pz, int64_t n,
                                                         const uint64_t* __restrict__ codes, int nleaf,
                                                         float4* __restrict__ leaf_lo, float4* __restrict__ leaf_hi,
                                                         uint64_t* __restrict__ leaf_code)
{
    static_assert(LEAF >= 2 && LEAF <= kWave && (LEAF & (LEAF - 1)) == 0, "leaf = power of two lanes");
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    // padding slots (NaN) are neither in the box nor in the count
    bool valid = i < n && px[i] == px[i];
    float x = valid ? px[i] : INFINITY, y = valid ? py[i] : INFINITY, z = valid ? pz[i] : INFINITY;
    float X = valid ? px[i] : -INFINITY, Y = valid ? py[i] : -INFINITY, Z = valid ? pz[i] : -INFINITY;
    const uint64_t vb = ballot(valid);
    const int cnt = __popcll((vb >> (lane_id() & ~(LEAF - 1))) & ((1ull << LEAF) - 1ull));
#pragma unroll
    for (int m = LEAF / 2; m >= 1; m >>= 1) {
        x = fminf(x, shfl_xor(x, m));
        y = fminf(y, shfl_xor(y, m));
        z = fminf(z, shfl_xor(z, m));
        X = fmaxf(X, shfl_xor(X, m));
        Y = fmaxf(Y, shfl_xor(Y, m));
        Z = fmaxf(Z, shfl_xor(Z, m));
    }
    int64_t leaf = i / LEAF;
    if ((i % LEAF) == 0 && leaf < nleaf) {
        leaf_lo[leaf] = make_float4(x, y, z, __int_as_float(cnt));
        leaf_hi[leaf] = make_float4(X, Y, Z, 0.0f);
        leaf_code[leaf] = codes[i];
    }
}

// Point sums of every leaf about its box centre c = (lo + hi) / 2, for consumers that take a
// whole leaf at once: agg[3 l] = (sum (p - c), count), agg[3 l + 1] = (xx, xy, xz, yy),
// agg[3 l + 2] = (yz, zz, 0, 0) of sum (p - c)(p - c)^T.  One thread per leaf.
__device__ __forceinline__ float leaf_centre(float lo, float hi) { return (lo + hi) * 0.5f; }

__global__ __launch_bounds__(256) void leaf_sums_kernel(const float* __restrict__ px, const float* __restrict__ py,
                                                        const float* __restrict__ pz, int64_t n,
                                                        const float4* __restrict__ leaf_lo,
                                                        const float4* __restrict__ leaf_hi, int nleaf,
                                                        float4* __restrict__ agg)
{
    const int leaf = blockIdx.x * blockDim.x + threadIdx.x;
    if (leaf >= nleaf) return;
    const float4 lo = leaf_lo[leaf], hi = leaf_hi[leaf];
    const float cx = leaf_centre(lo.x, hi.x), cy = leaf_centre(lo.y, hi.y), cz = leaf_centre(lo.z, hi.z);
    float sx = 0, sy = 0, sz = 0, xx = 0, xy = 0, xz = 0, yy = 0, yz = 0, zz = 0;
    int cnt = 0;
    for (int j = 0; j < LEAF; j++) {
        const int64_t i = (int64_t)leaf * LEAF + j;
        if (i >= n || !(px[i] == px[i])) continue;  // padding slots
        const float rx = px[i] - cx, ry = py[i] - cy, rz = pz[i] - cz;
        sx += rx;
        sy += ry;
        sz += rz;
        xx = fmaf(rx, rx, xx);
        xy = fmaf(rx, ry, xy);
        xz = fmaf(rx, rz, xz);
        yy = fmaf(ry, ry, yy);
        yz = fmaf(ry, rz, yz);
        zz = fmaf(rz, rz, zz);
        cnt++;
    }
    agg[3 * (int64_t)leaf] = make_float4(sx, sy, sz, (float)cnt);
    agg[3 * (int64_t)leaf + 1] = make_float4(xx, xy, xz, yy);
    agg[3 * (int64_t)leaf + 2] = make_float4(yz, zz, 0.0f, 0.0f);
}

// Karras 2012: internal node i of the radix tree over nl sorted leaf codes.
__device__ __forceinline__ int karras_delta(const uint64_t* codes, int nl, int i, int j)
{
    if (j < 0 || j >= nl) return -1;
    uint64_t a = codes[i], b = codes[j];
    if (a != b) return __clzll((long long)(a ^ b));
    return 64 + __clz((int)((uint32_t)i ^ (uint32_t)j));
}

__global__ __launch_bounds__(256) void karras_kernel(const uint64_t* __restrict__ codes, int nl,
                                                     BvhNode* __restrict__ nodes, int* __restrict__ leaf_parent)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nl - 1) return;
    const int dl = karras_delta(codes, nl, i, i + 1), dr = karras_delta(codes, nl, i, i - 1);
    const int d = dl > dr ? 1 : -1;
    const int dmin = karras_delta(codes, nl, i, i - d);
    int lmax = 2;
    while (karras_delta(codes, nl, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (karras_delta(codes, nl, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = karras_delta(codes, nl, i, j);
    int s = 0;
    int t = l;
    do {
        t = (t + 1) >> 1;
        if (karras_delta(codes, nl, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int gamma = i + s * d + (d < 0 ? -1 : 0);
    const int lo = min(i, j), hi = max(i, j);
    int c0, c1;
    if (lo == gamma) {
        c0 = leaf_ref(gamma);
        leaf_parent[gamma] = i;
    } else {
        c0 = gamma;
        nodes[gamma].r.z = i;
    }
    if (hi == gamma + 1) {
        c1 = leaf_ref(gamma + 1);
        leaf_parent[gamma + 1] = i;
    } else {
        c1 = gamma + 1;
        nodes[gamma + 1].r.z = i;
    }
    nodes[i].r.x = c0;
    nodes[i].r.y = c1;
    nodes[i].r.w = j;  // the node covers leaves [min(i, j), max(i, j)]
    if (i == 0) nodes[0].r.z = -1;
}

// Refit without atomics.  Every node of the Karras tree covers a contiguous leaf range
// [min(i, r.w), max(i, r.w)], so its box is the union of the leaf boxes over that range: a range
// union, answered from the doubling table T_j[i] = union of leaves [i, i + 2^j) as
// T_j[first] U T_j[last - 2^j + 1] with j = floor(log2(length)).  Level j is built from level j - 1
// (two buffers in turn), then the nodes whose length has that j take their box and write it into
// their parent's child slot; leaves write theirs first.  Every slot has one writer, and min / max
// give the same box in any order, so the result is the bottom-up refit's.

// one child's box in its parent's a / b / c fields (left: a.xyzw b.xy, right: b.zw c.xyzw)
__device__ __forceinline__ void put_child_box(BvhNode* nd, bool left, const float4& lo, const float4& hi)
{
    if (left) {
        nd->a = make_float4(lo.x, lo.y, lo.z, hi.x);
        nd->b.x = hi.y;
        nd->b.y = hi.z;
    } else {
        nd->b.z = lo.x;
        nd->b.w = lo.y;
        nd->c = make_float4(lo.z, hi.x, hi.y, hi.z);
    }
}

__global__ __launch_bounds__(256) void refit_leaf_slots_kernel(int nl, const int* __restrict__ leaf_parent,
                                                               const float4* __restrict__ leaf_lo,
                                                               const float4* __restrict__ leaf_hi, BvhNode* nodes)
{
    const int leaf = blockIdx.x * blockDim.x + threadIdx.x;
    if (leaf >= nl || nl < 2) return;
    const int p = leaf_parent[leaf];
    if (p < 0) return;
    put_child_box(&nodes[p], nodes[p].r.x == leaf_ref(leaf), leaf_lo[leaf], leaf_hi[leaf]);
}

// level j of the doubling table from level j - 1 (half = 2^(j-1)); entries i <= nl - 2 * half
__global__ __launch_bounds__(256) void range_union_level_kernel(int nl, int half, const float4* __restrict__ lo_in,
                                                                const float4* __restrict__ hi_in,
                                                                float4* __restrict__ lo_out, float4* __restrict__ hi_out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > nl - 2 * half) return;
    const float4 a = lo_in[i], b = lo_in[i + half], c = hi_in[i], d = hi_in[i + half];
    lo_out[i] = make_float4(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z), 0.0f);
    hi_out[i] = make_float4(fmaxf(c.x, d.x), fmaxf(c.y, d.y), fmaxf(c.z, d.z), 0.0f);
}

// the internal nodes whose leaf count has floor(log2) == j: own box, and the parent's child slot
__global__ __launch_bounds__(256) void refit_level_kernel(int nint, int j, const float4* __restrict__ tlo,
                                                          const float4* __restrict__ thi, BvhNode* nodes,
                                                          float4* __restrict__ node_lo, float4* __restrict__ node_hi)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nint) return;
    const int4 r = nodes[i].r;
    const int first = min(i, r.w), last = max(i, r.w);
    const int len = last - first + 1;
    if (31 - __clz(len) != j) return;
    const int second = last - (1 << j) + 1;
    const float4 a = tlo[first], b = tlo[second], c = thi[first], d = thi[second];
    const float4 lo = make_float4(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z), 0.0f);
    const float4 hi = make_float4(fmaxf(c.x, d.x), fmaxf(c.y, d.y), fmaxf(c.z, d.z), 0.0f);
    node_lo[i] = lo;
    node_hi[i] = hi;
    if (r.z >= 0) put_child_box(&nodes[r.z], nodes[r.z].r.x == i, lo, hi);
}

__global__ void fill_u32_kernel(uint32_t* p, int64_t n, uint32_t v)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// BvhNode4 of every internal node.  One thread per node.
__global__ __launch_bounds__(256) void node4_kernel(const BvhNode* __restrict__ nodes, int nint,
                                                    BvhNode4* __restrict__ node4)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nint) return;
    const BvhNode nd = nodes[i];
    float f[24];
    int r[4];
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const int c = k ? nd.r.y : nd.r.x;
        if (c >= 0) {
            const BvhNode cn = nodes[c];
            const float g[12] = {cn.a.x, cn.a.y, cn.a.z, cn.a.w, cn.b.x, cn.b.y,
                                 cn.b.z, cn.b.w, cn.c.x, cn.c.y, cn.c.z, cn.c.w};
#pragma unroll
            for (int j = 0; j < 12; j++) f[12 * k + j] = g[j];
            r[2 * k] = cn.r.x;
            r[2 * k + 1] = cn.r.y;
        } else {
            const float g[6] = {k ? nd.b.z : nd.a.x, k ? nd.b.w : nd.a.y, k ? nd.c.x : nd.a.z,
                                k ? nd.c.y : nd.a.w, k ? nd.c.z : nd.b.x, k ? nd.c.w : nd.b.y};
#pragma unroll
            for (int j = 0; j < 6; j++) f[12 * k + j] = g[j];
#pragma unroll
            for (int j = 0; j < 3; j++) {
                f[12 * k + 6 + j] = INFINITY;
                f[12 * k + 9 + j] = -INFINITY;
            }
            r[2 * k] = c;
            r[2 * k + 1] = EMPTY4;
        }
    }
    BvhNode4 o;
#pragma unroll
    for (int j = 0; j < 6; j++) o.b[j] = make_float4(f[4 * j], f[4 * j + 1], f[4 * j + 2], f[4 * j + 3]);
    o.r = make_int4(r[0], r[1], r[2], r[3]);
    node4[i] = o;
}

// The top of each scan's subtree cut FRONTIER_LEVELS levels below its root: slot j of level
// L + 1 holds child (j & 1) of slot j >> 1 of level L (a leaf stays in its even slot, the odd
// one stays empty), so the FRONTIER slots of the last level cover the subtree's leaves
// exactly once.  A traversal tests these boxes in one wave-wide step instead of descending
// the top levels one dependent load at a time.  One block of FRONTIER threads per scan.
constexpr int FRONTIER_LEVELS = 8;
constexpr int FRONTIER = 1 << FRONTIER_LEVELS;
constexpr int FR_NONE = (int)0x80000000;  // an empty slot

__global__ __launch_bounds__(FRONTIER) void frontier_kernel(const BvhNode* __restrict__ nodes,
                                                            const float4* __restrict__ leaf_lo,
                                                            const float4* __restrict__ leaf_hi, int nleaf,
                                                            const int* __restrict__ roots, int* __restrict__ fr_ref,
                                                            float4* __restrict__ fr_lo, float4* __restrict__ fr_hi)
{
    __shared__ int ref[FRONTIER];
    __shared__ float4 lo[FRONTIER], hi[FRONTIER];
    const int t = threadIdx.x;
    const int root = roots ? roots[blockIdx.x] : (nleaf > 1 ? 0 : leaf_ref(0));
    if (t == 0) {
        ref[0] = root;
        if (root < 0 && -1 - root >= nleaf) {
            ref[0] = FR_NONE;  // an empty scan (its nominal leaf lies past the end)
        } else if (root < 0) {
            lo[0] = leaf_lo[-1 - root];
            hi[0] = leaf_hi[-1 - root];
        } else {
            const BvhNode nd = nodes[root];
            lo[0] = make_float4(fminf(nd.a.x, nd.b.z), fminf(nd.a.y, nd.b.w), fminf(nd.a.z, nd.c.x), 0.0f);
            hi[0] = make_float4(fmaxf(nd.a.w, nd.c.y), fmaxf(nd.b.x, nd.c.z), fmaxf(nd.b.y, nd.c.w), 0.0f);
        }
    } else {
        ref[t] = FR_NONE;
    }
    __syncthreads();
    for (int level = 0; level < FRONTIER_LEVELS; level++) {
        const int width = 2 << level;  // slots of the next level
        int r = FR_NONE;
        float4 l = make_float4(0, 0, 0, 0), h = make_float4(0, 0, 0, 0);
        if (t < width) {
            const int src = ref[t >> 1];
            if (src >= 0) {
                const BvhNode nd = nodes[src];
                if (t & 1) {
                    r = nd.r.y;
                    l = make_float4(nd.b.z, nd.b.w, nd.c.x, 0.0f);
                    h = make_float4(nd.c.y, nd.c.z, nd.c.w, 0.0f);
                } else {
                    r = nd.r.x;
                    l = make_float4(nd.a.x, nd.a.y, nd.a.z, 0.0f);
                    h = make_float4(nd.a.w, nd.b.x, nd.b.y, 0.0f);
                }
            } else if (!(t & 1)) {  // a leaf (or an empty slot) moves down unchanged
                r = src;
                l = lo[t >> 1];
                h = hi[t >> 1];
            }
        }
        __syncthreads();
        if (t < width) {
            ref[t] = r;
            lo[t] = l;
            hi[t] = h;
        }
        __syncthreads();
    }
    const size_t o = (size_t)blockIdx.x * FRONTIER + t;
    fr_ref[o] = ref[t];
    fr_lo[o] = lo[t];
    fr_hi[o] = hi[t];
}

// Root of each scan's subtree: scan s covers leaves [pstart[s] / LEAF, pstart[s+1] / LEAF), a
// contiguous run sharing the scan's key prefix, which in a radix tree is exactly the leaf
// range of one node (the lowest ancestor of its first leaf that reaches its last leaf).
__global__ void scan_roots_kernel(const int64_t* __restrict__ pstart, int ns, const int* __restrict__ leaf_parent,
                                  const BvhNode* __restrict__ nodes, int* __restrict__ roots)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    const int ls = (int)(pstart[s] / LEAF), le = (int)(pstart[s + 1] / LEAF) - 1;
    if (le <= ls) {
        roots[s] = leaf_ref(ls);  // one leaf (or an empty scan, never queried)
        return;
    }
    int p = leaf_parent[ls];
    while (p >= 0) {
        const int4 r = nodes[p].r;
        const int lo = p < r.w ? p : r.w, hi = p < r.w ? r.w : p;
        if (lo <= ls && hi >= le) break;
        p = r.z;
    }
    roots[s] = p >= 0 ? p : 0;
}

// Median of n <= 8192 samples (bitonic sort in LDS, one block of 1024 threads): out = the
// element of rank n / 2.
constexpr int MEDIAN_MAX = 8192;
__global__ __launch_bounds__(1024) void median_kernel(const float* __restrict__ v, int n, float* __restrict__ out)
{
    __shared__ float a[MEDIAN_MAX];
    for (int i = threadIdx.x; i < MEDIAN_MAX; i += blockDim.x) a[i] = i < n ? v[i] : INFINITY;
    __syncthreads();
    for (int k = 2; k <= MEDIAN_MAX; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < MEDIAN_MAX; i += blockDim.x) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const bool up = (i & k) == 0;
                    const float x = a[i], y = a[ixj];
                    if ((x > y) == up) {
                        a[i] = y;
                        a[ixj] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
    if (threadIdx.x == 0) out[0] = a[n / 2];
}

}  // namespace lvrhip
