// LBVH over Morton-sorted points, built entirely on the device.
//
// Reference: lbvh::LBVHIndex::build (src/liblvr2/reconstruction/cuda/LBVHIndex.cu:75-324)
// builds a bottom-up (Apetrei-style) hierarchy over per-point AABBs after a host sort.
// Here: points are sorted by 63-bit Morton code (21 bits/axis, cubic normalisation to the
// cloud's bounding box), cut into fixed leaves of LEAF consecutive points, and a Karras
// (2012) radix tree is built top-down over the leaves' codes, with leaf indices
// breaking code ties; child boxes are then refitted bottom-up with one atomic flag per
// internal node.  Each internal node stores both children's boxes, so a traversal step
// reads one 64-byte node.
#pragma once

#include "common.hpp"

namespace lvrhip {

constexpr int LEAF = 16;

struct alignas(16) BvhNode {
    float4 a;  // lo0.x lo0.y lo0.z hi0.x
    float4 b;  // hi0.y hi0.z lo1.x lo1.y
    float4 c;  // lo1.z hi1.x hi1.y hi1.z
    int4 r;    // child0, child1 (>= 0 internal node, < 0 leaf = -1 - id), parent, unused
};

__device__ __forceinline__ int leaf_ref(int leaf) { return -1 - leaf; }

// ---- bounding box ----------------------------------------------------------------------

__global__ __launch_bounds__(256) void bbox_partial_kernel(const float* __restrict__ xyz, int64_t n,
                                                           float* __restrict__ part)
{
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
        for (int j = 0; j < 3; j++) {
            float v = xyz[3 * i + j];
            mn[j] = fminf(mn[j], v);
            mx[j] = fmaxf(mx[j], v);
        }
    }
    __shared__ float s[4][6];
#pragma unroll
    for (int j = 0; j < 3; j++) {
        mn[j] = wave_min(mn[j]);
        mx[j] = wave_max(mx[j]);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
        for (int j = 0; j < 3; j++) {
            s[wave][j] = mn[j];
            s[wave][3 + j] = mx[j];
        }
    __syncthreads();
    if (threadIdx.x < 6) {
        float v = s[0][threadIdx.x];
        for (int w = 1; w < 4; w++) v = threadIdx.x < 3 ? fminf(v, s[w][threadIdx.x]) : fmaxf(v, s[w][threadIdx.x]);
        part[6 * blockIdx.x + threadIdx.x] = v;
    }
}

__global__ void bbox_final_kernel(const float* __restrict__ part, int nparts, float* __restrict__ out)
{
    if (threadIdx.x < 6) {
        float v = threadIdx.x < 3 ? INFINITY : -INFINITY;
        for (int i = 0; i < nparts; i++) {
            float x = part[6 * i + threadIdx.x];
            v = threadIdx.x < 3 ? fminf(v, x) : fmaxf(v, x);
        }
        out[threadIdx.x] = v;
    }
}

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

// Codes of AoS points, values = their index.
__global__ __launch_bounds__(256) void morton_aos_kernel(const float* __restrict__ xyz, int64_t n, MortonFrame f,
                                                         uint64_t* __restrict__ keys, uint32_t* __restrict__ vals)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    keys[i] = morton63(f, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
    vals[i] = (uint32_t)i;
}

// Sorted SoA copy plus the inverse permutation.
__global__ __launch_bounds__(256) void gather_sorted_kernel(const float* __restrict__ xyz, int64_t n,
                                                            const uint32_t* __restrict__ order,
                                                            float* __restrict__ px, float* __restrict__ py,
                                                            float* __restrict__ pz, uint32_t* __restrict__ rank)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t j = order[i];
    px[i] = xyz[3 * (int64_t)j];
    py[i] = xyz[3 * (int64_t)j + 1];
    pz[i] = xyz[3 * (int64_t)j + 2];
    if (rank) rank[j] = (uint32_t)i;
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
    bool valid = i < n;
    float x = valid ? px[i] : INFINITY, y = valid ? py[i] : INFINITY, z = valid ? pz[i] : INFINITY;
    float X = valid ? px[i] : -INFINITY, Y = valid ? py[i] : -INFINITY, Z = valid ? pz[i] : -INFINITY;
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
        int64_t cnt = n - leaf * LEAF;
        if (cnt > LEAF) cnt = LEAF;
        leaf_lo[leaf] = make_float4(x, y, z, __int_as_float((int)cnt));
        leaf_hi[leaf] = make_float4(X, Y, Z, 0.0f);
        leaf_code[leaf] = codes[i];
    }
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
    if (i == 0) nodes[0].r.z = -1;
}

// Bottom-up refit: a thread per leaf climbs; the second arrival at a node writes both
// child boxes into it (acquire/release at agent scope order the two climbers).
__global__ __launch_bounds__(256) void refit_kernel(int nl, const int* __restrict__ leaf_parent,
                                                    const float4* __restrict__ leaf_lo,
                                                    const float4* __restrict__ leaf_hi, BvhNode* nodes,
                                                    float4* node_lo, float4* node_hi, uint32_t* flags)
{
    int leaf = blockIdx.x * blockDim.x + threadIdx.x;
    if (leaf >= nl || nl < 2) return;
    int p = leaf_parent[leaf];
    while (p >= 0) {
        uint32_t old = __hip_atomic_fetch_add(&flags[p], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (old == 0) return;
        int4 r = nodes[p].r;
        float4 l0, h0, l1, h1;
        if (r.x < 0) {
            l0 = leaf_lo[-1 - r.x];
            h0 = leaf_hi[-1 - r.x];
        } else {
            l0 = node_lo[r.x];
            h0 = node_hi[r.x];
        }
        if (r.y < 0) {
            l1 = leaf_lo[-1 - r.y];
            h1 = leaf_hi[-1 - r.y];
        } else {
            l1 = node_lo[r.y];
            h1 = node_hi[r.y];
        }
        nodes[p].a = make_float4(l0.x, l0.y, l0.z, h0.x);
        nodes[p].b = make_float4(h0.y, h0.z, l1.x, l1.y);
        nodes[p].c = make_float4(l1.z, h1.x, h1.y, h1.z);
        node_lo[p] = make_float4(fminf(l0.x, l1.x), fminf(l0.y, l1.y), fminf(l0.z, l1.z), 0.0f);
        node_hi[p] = make_float4(fmaxf(h0.x, h1.x), fmaxf(h0.y, h1.y), fmaxf(h0.z, h1.z), 0.0f);
        p = r.z;
    }
}

__global__ void fill_u32_kernel(uint32_t* p, int64_t n, uint32_t v)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace lvrhip
