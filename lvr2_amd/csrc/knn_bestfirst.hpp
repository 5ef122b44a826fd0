// Exact kNN for small k by best-first traversal, one lane per query.
//
// The tile kernel (knn_tile.hpp) serves large neighbourhoods (the adaptive normals' k = 2kn ..
// 32kn) well, but a signed-distance query needs only kd (5 .. 30) neighbours, and for lattice
// vertices just off a very dense surface its radius attempts must close in on the k-th distance
// to within micrometres, tile after tile.  A best-first search needs no radius: each lane keeps
// its t best keys (fp32 d^2, point index) sorted in registers and prunes the tree with the t-th,
// which tightens as points are found.  It replaces AdaptiveKSearchSurface::distance's kSearch
// (AdaptiveKSearchSurface.tcc:414-457, called from PointsetGrid.tcc:123-149) and
// SearchKDTree::kSearch / radiusSearch (KDTree.tcc:331-387) for k <= BF_KMAX.
//
//   1. A greedy descent from the root to the leaf whose box is nearest at every level; its
//      points seed the list, so the bound is finite before anything is pushed.
//   2. Best-first from the root over BvhNode4 (four grandchildren per node): children within
//      the bound are pushed farthest first, each with its box distance, onto a per-lane LDS
//      stack; a popped entry beyond the (meanwhile tightened) bound is dropped.
// The few queries whose lane stack overflows (hundreds of leaves within their bound) are searched
// again by a whole wave each (bf_knn_wave_kernel), and the tile kernel takes any that outgrow
// that stack too.
// Box distances are monotone lower bounds of the fp32 d^2 of every point in the box, so a box
// is skipped only when none of its points can enter the list: the result is the exact t
// smallest keys, identical to the tile kernel's.
#pragma once

#include "common.hpp"
#include "knn_tile.hpp"
#include "lbvh.hpp"
#include "sinks.hpp"

namespace lvrhip {

constexpr int BF_THREADS = 64;  // one wave per workgroup
constexpr int BF_STACK = 16;    // stack entries per lane (ref, box d^2) in LDS: 8 KB per wave
constexpr int BF_SPILL = 48;    // further entries per lane in global memory (deep traversals only)

// A lane's traversal stack: CAP entries of (node or leaf reference, box d^2 bits).
// The first BF_STACK entries in LDS (the pointer carries the LDS address space, so that the
// two parts are never addressed through one generic pointer), the rest in global memory.
typedef __attribute__((address_space(3))) uint2 lds_uint2;
struct BfLdsStack {
    lds_uint2* s;  // this lane's LDS entries, interleaved with the wave's other lanes
    uint2* spill;  // this lane's spill entries, interleaved likewise
    int cap;       // BF_STACK + the spill entries per lane
    __device__ __forceinline__ int capacity() const { return cap; }
    __device__ __forceinline__ uint2 get(int i) const
    {
        uint2 v;
        if (i < BF_STACK) {
            const lds_uint2& e = s[i * BF_THREADS];
            v.x = e.x;
            v.y = e.y;
        } else {
            v = spill[(size_t)(i - BF_STACK) * BF_THREADS];
        }
        return v;
    }
    __device__ __forceinline__ void put(int i, uint2 v) const
    {
        if (i < BF_STACK) {
            lds_uint2& e = s[i * BF_THREADS];
            e.x = v.x;
            e.y = v.y;
        } else {
            spill[(size_t)(i - BF_STACK) * BF_THREADS] = v;
        }
    }
};

constexpr int BF_KMAX = 32;
// key of an empty list slot: above every key of a finite d^2 (and of d^2 = inf with a real index)
constexpr uint64_t BF_EMPTY = 0x7F800000FFFFFFFFull;

// The t smallest keys seen so far, ascending in key[K - t, K).  The slots below K - t hold 0,
// which no candidate displaces (candidates enter by a strict comparison), so the t-th key is
// always key[K - 1] and every index into the arrays is a compile-time constant.
template <int K>
struct BfList {
    uint64_t key[K];
    uint32_t sp[K];  // sorted position of each key's point (~0: empty)

    __device__ __forceinline__ void init(int t, uint64_t empty)
    {
#pragma unroll
        for (int j = 0; j < K; j++) {
            key[j] = j < K - t ? 0ull : empty;
            sp[j] = ~0u;
        }
    }
    __device__ __forceinline__ uint64_t worst() const { return key[K - 1]; }
    // every one of the t slots holds a point (unfilled slots hold the largest keys)
    __device__ __forceinline__ bool full() const { return sp[K - 1] != ~0u; }
    // d^2 of the t-th key: the pruning bound
    __device__ __forceinline__ float bound() const { return bitsf((uint32_t)(key[K - 1] >> 32)); }
    __device__ __forceinline__ bool contains(uint64_t c) const
    {
        bool f = false;
#pragma unroll
        for (int j = 0; j < K; j++) f = f || key[j] == c;
        return f;
    }

    // insert c (c < worst()), dropping the largest
    __device__ __forceinline__ void insert(uint64_t c, uint32_t s)
    {
        bool ltp = c < key[0];
        uint64_t pk = key[0];
        uint32_t ps = sp[0];
        key[0] = ltp ? c : pk;
        sp[0] = ltp ? s : ps;
#pragma unroll
        for (int j = 1; j < K; j++) {
            const bool lt = c < key[j];
            const uint64_t kj = key[j];
            const uint32_t sj = sp[j];
            key[j] = lt ? (ltp ? pk : c) : kj;
            sp[j] = lt ? (ltp ? ps : s) : sj;
            pk = kj;
            ps = sj;
            ltp = lt;
        }
    }
};

// The leaf's points whose key beats the current t-th enter the list smallest first: each round
// finds the smallest remaining candidate (compile-time indices only) and inserts it, so a wave
// runs as many insertion rounds as its busiest lane has improving points, not one per point.
// kDedup: the list may already hold some of the leaf's points (a search started from a known
// neighbour set); those are not inserted twice.
template <int K, bool kDedup = false>
__device__ __forceinline__ void bf_scan_leaf(const BvhView& bv, int leaf, float qx, float qy, float qz, BfList<K>& L)
{
    const float4* p = bv.pts4 + (size_t)leaf * LEAF;
    float4 v[LEAF];
#pragma unroll
    for (int j = 0; j < LEAF; j++) v[j] = p[j];
    uint64_t c[LEAF];
    const uint64_t w = L.worst();
#pragma unroll
    for (int j = 0; j < LEAF; j++) {
        const float d = d2_contract(qx, qy, qz, v[j].x, v[j].y, v[j].z);
        const uint64_t key = make_key(d, fbits(v[j].w));
        // padding slots have NaN coordinates: never a neighbour
        c[j] = (d == d && key < w) ? key : ~0ull;
    }
    // Keys are unique (they end in the point's index), so each round takes the smallest key above
    // the previous round's instead of clearing the taken slot: the same sequence of insertions.
    uint64_t lo = 0;
    for (int round = 0; round < LEAF; round++) {
        uint64_t best = ~0ull;
        int bj = 0;
#pragma unroll
        for (int j = 0; j < LEAF; j++) {
            if (c[j] >= lo && c[j] < best) {
                best = c[j];
                bj = j;
            }
        }
        if (!(best < L.worst())) break;  // the rest cannot enter (~0 marks none)
        if (!kDedup || !L.contains(best)) L.insert(best, (uint32_t)(leaf * LEAF + bj));
        lo = best + 1;  // best < worst <= ~0: no wrap
    }
}

__device__ __forceinline__ float bf_child_d2(const float* f, int k, float qx, float qy, float qz)
{
    return box_d2(f[6 * k], f[6 * k + 1], f[6 * k + 2], f[6 * k + 3], f[6 * k + 4], f[6 * k + 5], qx, qy, qz, qx, qy,
                  qz);
}

// The search of one lane; returns false if its stack overflowed (the list is then incomplete).
template <int K, bool kDedup, class Stack>
__device__ __forceinline__ bool bf_search(const BvhView& bv, const Stack& stk, float qx, float qy, float qz,
                                          BfList<K>& L, int& leaves, int& nodes)
{
    const int root = bv.nleaf > 1 ? 0 : leaf_ref(0);
    // 1. greedy descent: the seed leaf
    int ref = root;
    while (ref >= 0) {
        const BvhNode4 nd = bv.nodes4[ref];
        const float* f = reinterpret_cast<const float*>(nd.b);
        const int rr[4] = {nd.r.x, nd.r.y, nd.r.z, nd.r.w};
        float best = INFINITY;
        int next = rr[0];  // slot 0 always holds a child
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const float d = bf_child_d2(f, k, qx, qy, qz);
            if (rr[k] != EMPTY4 && d < best) {
                best = d;
                next = rr[k];
            }
        }
        ref = next;
    }
    const int seed = -1 - ref;
    bf_scan_leaf<K, kDedup>(bv, seed, qx, qy, qz, L);
    leaves++;
    if (root < 0) return true;
    // 2. best-first from the root.  Each pass pops (expanding nodes) until the lane holds a leaf
    // to scan, then the lanes scan their leaves together: leaf scans, the costly part, run with
    // most lanes of the wave active instead of interleaving with other lanes' node expansions.
    int sp = 0;
    uint2 e = make_uint2((uint32_t)root, 0u);  // the entry to process next
    bool have = true;                           // e holds an entry (else pop one)
    for (;;) {
        int leaf = -1;
        for (;;) {
            if (!have) {
                if (sp == 0) break;
                e = stk.get(--sp);
            }
            have = false;
            const float bound = L.bound();
            if (bitsf(e.y) > bound) continue;
            const int r = (int)e.x;
            if (r < 0) {
                if (-1 - r != seed) {
                    leaf = -1 - r;
                    break;
                }
                continue;
            }
            nodes++;
            const BvhNode4 nd = bv.nodes4[r];
            const float* f = reinterpret_cast<const float*>(nd.b);
            int c[4] = {nd.r.x, nd.r.y, nd.r.z, nd.r.w};
            float d[4];
            int nv = 0;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                d[k] = bf_child_d2(f, k, qx, qy, qz);
                if (c[k] == EMPTY4 || !(d[k] <= bound)) {
                    c[k] = EMPTY4;
                    d[k] = -1.0f;  // sorts last
                } else {
                    nv++;
                }
            }
            // descending by distance: the valid children first, the nearest of them last
            auto cswap = [&](int x, int y) {
                if (d[x] < d[y]) {
                    const float td = d[x];
                    d[x] = d[y];
                    d[y] = td;
                    const int tc = c[x];
                    c[x] = c[y];
                    c[y] = tc;
                }
            };
            cswap(0, 1);
            cswap(2, 3);
            cswap(0, 2);
            cswap(1, 3);
            cswap(1, 2);
            // the nearest child is processed next without a trip through the stack; the others
            // are pushed farthest first
#pragma unroll
            for (int k = 0; k < 4; k++) {
                if (k + 1 < nv) {
                    if (sp == stk.capacity()) return false;
                    stk.put(sp, make_uint2((uint32_t)c[k], fbits(d[k])));
                    sp++;
                } else if (k + 1 == nv) {
                    e = make_uint2((uint32_t)c[k], fbits(d[k]));
                    have = true;
                }
            }
        }
        if (leaf < 0) break;  // stack empty: the list is final
        bf_scan_leaf<K, kDedup>(bv, leaf, qx, qy, qz, L);
        leaves++;
    }
    return true;
}

// One lane per query i < m (queries in Morton order for locality): Out(i, list, t, ok, q)
// consumes the result.  empty: the key an unfilled slot holds (BF_EMPTY, or (r^2, 0) to keep
// only points with d^2 < r^2).
// Outputs that start a query's search from a known neighbour set declare kPrefill = true and
// prefill_list(i, L, t, x, y, z).
template <class O, class = void>
struct bf_prefills : std::false_type {
};
template <class O>
struct bf_prefills<O, std::void_t<decltype(O::kPrefill)>> : std::integral_constant<bool, O::kPrefill> {
};

// stats (optional, KS_* counters of knn_tile.hpp): waves (KS_TILES), the largest leaf count of a
// lane summed over waves (KS_ATTEMPTS: a wave runs as long as its slowest lane), points scanned
// (KS_CANDIDATES), nodes expanded (KS_POPS), queries left to the tile kernel (KS_REFINE),
// queries resolved (KS_RESOLVED), shader-clock cycles per wave (KS_CY_TOTAL).
template <int K, class Out>
__device__ __forceinline__ void bf_stats(unsigned long long* stats, bool valid, bool ok, int leaves, int nodes,
                                         unsigned long long c0)
{
    int lmax = leaves, lsum = leaves, nsum = nodes, bad = (valid && !ok) ? 1 : 0, good = (valid && ok) ? 1 : 0;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        lmax = max(lmax, shfl_xor(lmax, o));
        lsum += shfl_xor(lsum, o);
        nsum += shfl_xor(nsum, o);
        bad += shfl_xor(bad, o);
        good += shfl_xor(good, o);
    }
    if (lane_id() == 0) {
        atomicAdd(&stats[KS_TILES], 1ull);
        atomicAdd(&stats[KS_ATTEMPTS], (unsigned long long)lmax);
        atomicAdd(&stats[KS_CANDIDATES], (unsigned long long)lsum * LEAF);
        atomicAdd(&stats[KS_POPS], (unsigned long long)nsum);
        atomicAdd(&stats[KS_REFINE], (unsigned long long)bad);
        atomicAdd(&stats[KS_RESOLVED], (unsigned long long)good);
        atomicAdd(&stats[KS_CY_TOTAL], cycles() - c0);
    }
}

// stats (optional, KS_* counters of knn_tile.hpp): waves (KS_TILES), the largest leaf count of a
// lane summed over waves (KS_ATTEMPTS: a wave runs as long as its slowest lane), points scanned
// (KS_CANDIDATES), nodes expanded (KS_POPS), queries left to the next pass (KS_REFINE),
// queries resolved (KS_RESOLVED), shader-clock cycles per wave (KS_CY_TOTAL).
//
// One lane per query (queries in Morton order), chunks of 64 consecutive queries per wave.  With
// a work queue (QUEUE_SEGS counters QUEUE_STRIDE apart, zeroed before the launch; the grid is
// then the waves the device holds at once) the chunks are dealt like the tile kernel's runs:
// blocks of QUEUE_CHUNK consecutive chunks to the XCDs in turn, each XCD taking from its own
// counter and then helping the others, so neighbouring chunks share an L2 and the costly regions
// of a cloud are spread over all XCDs.  Without a queue, wave b takes chunk b.
//
// Query i is (qx[i qs], qy[i qs], qz[i qs]): qs = 1 for SoA coordinates, 3 for an AoS buffer
// (qy = qx + 1, qz = qx + 2).  spill: spill_cap entries per lane for each wave of the grid.
template <int K, class Out>
__global__ __launch_bounds__(BF_THREADS) void bf_knn_kernel(const BvhView bv, const float* __restrict__ qx,
                                                            const float* __restrict__ qy,
                                                            const float* __restrict__ qz, int qs, int64_t m,
                                                            int t, uint64_t empty, Out out,
                                                            unsigned long long* stats, uint32_t* queue, uint2* spill,
                                                            int spill_cap)
{
    __shared__ uint2 stk[BF_STACK][BF_THREADS];
    uint2* const my_spill = spill + (size_t)blockIdx.x * spill_cap * BF_THREADS + threadIdx.x;
    const int lane = threadIdx.x;
    const int64_t nchunks = (m + BF_THREADS - 1) / BF_THREADS;
    const int xcd = (int)(blockIdx.x % QUEUE_SEGS);
    int seg_k = 0;
    bool once = false;
    for (;;) {
        int64_t chunk = -1;
        if (queue) {
            while (seg_k < QUEUE_SEGS) {
                const int seg = (xcd + seg_k) % QUEUE_SEGS;
                uint32_t r = 0;
                if (lane == 0) r = atomicAdd(queue + seg * QUEUE_STRIDE, 1u);
                r = (uint32_t)__builtin_amdgcn_readlane((int)r, 0);
                const int64_t cand = ((int64_t)(r / QUEUE_CHUNK) * QUEUE_SEGS + seg) * QUEUE_CHUNK + r % QUEUE_CHUNK;
                if (cand < nchunks) {
                    chunk = cand;
                    break;
                }
                seg_k++;  // this XCD's chunks only grow with r: all taken
            }
        } else if (!once) {
            chunk = blockIdx.x;
            once = true;
        }
        if (chunk < 0 || chunk >= nchunks) break;
        const unsigned long long c0 = stats ? cycles() : 0ull;
        const int64_t i = chunk * BF_THREADS + lane;
        int leaves = 0, nodes = 0;
        bool ok = true;
        if (i < m) {  // no barrier in the search: lanes are independent
            const float x = qx[i * qs], y = qy[i * qs], z = qz[i * qs];
            BfList<K> L;
            L.init(t, empty);
            if constexpr (bf_prefills<Out>::value) out.template prefill_list<K>(i, L, t, x, y, z);
            ok = bf_search<K, bf_prefills<Out>::value>(
                bv, BfLdsStack{(lds_uint2*)&stk[0][lane], my_spill, BF_STACK + spill_cap}, x, y, z, L, leaves,
                nodes);
            out.template emit<K>(i, L, t, ok, x, y, z);
            ok = ok && L.full();
        }
        if (stats) bf_stats<K, Out>(stats, i < m, ok, leaves, nodes, c0);
    }
}

// Wave-cooperative best-first search for the queries list[0, *count): the ones whose lane search
// overflowed its stack (a very dense or a far-off neighbourhood, hundreds of leaves deep).  One
// wave per query: a greedy descent to a seed leaf; then every step pops up to 64 stack entries
// at once, the lanes expand their nodes (children within the bound pushed) or scan their leaves,
// and the t smallest keys of the list and the scanned points are extracted by t rounds of a
// wave-wide minimum.  Lane j holds the j-th key of the list in a register.  The serial chain is
// the tree depth plus leaves / 64 steps, not one step per node or leaf.  A query whose stack
// outgrows BF_WAVE_STACK entries is flagged for the tile kernel.
constexpr int BF_WAVE_STACK = 4096;
constexpr int BF_WAVE_BLOCKS = 1024;  // waves striding over the list

__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const uint64_t w = shfl_xor(v, o);
        v = w < v ? w : v;
    }
    return v;
}

template <int K, class Out>
__global__ __launch_bounds__(BF_THREADS) void bf_knn_wave_kernel(const BvhView bv, const float* __restrict__ qx,
                                                                 const float* __restrict__ qy,
                                                                 const float* __restrict__ qz, int qs,
                                                                 const uint32_t* __restrict__ list,
                                                                 const uint32_t* __restrict__ count, int t,
                                                                 uint64_t empty, Out out, unsigned long long* stats)
{
    __shared__ uint2 stk[BF_WAVE_STACK];
    __shared__ uint64_t fkey[BF_KMAX];
    __shared__ uint32_t fsp[BF_KMAX];
    const int lane = threadIdx.x;
    const int64_t n = (int64_t)*count;
    const int root = bv.nleaf > 1 ? 0 : leaf_ref(0);
    for (int64_t j = blockIdx.x; j < n; j += gridDim.x) {
        const unsigned long long c0 = stats ? cycles() : 0ull;
        const int64_t i = list[j];
        const float x = qx[i * qs], y = qy[i * qs], z = qz[i * qs];
        // lane j < t holds the j-th smallest key so far (empty: none)
        uint64_t bk = empty;
        uint32_t bs = ~0u;
        uint64_t worst = empty;  // the t-th key (uniform)
        int leaves = 0, nodes = 0;
        // seed leaf: greedy descent (every lane computes the same path)
        int ref = root;
        while (ref >= 0) {
            const BvhNode4 nd = bv.nodes4[ref];
            const float* f = reinterpret_cast<const float*>(nd.b);
            const int rr[4] = {nd.r.x, nd.r.y, nd.r.z, nd.r.w};
            float bd = INFINITY;
            int next = rr[0];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const float d = bf_child_d2(f, k, x, y, z);
                if (rr[k] != EMPTY4 && d < bd) {
                    bd = d;
                    next = rr[k];
                }
            }
            ref = next;
        }
        const int seed = -1 - ref;
        if (root >= 0 && lane == 0) stk[0] = make_uint2((uint32_t)root, 0u);
        int sp = root >= 0 ? 1 : 0;
        bool ok = true;
        bool first = true;
        while (first || sp > 0) {
            // this step's entries: the seed leaf first, then up to 64 popped entries
            int leaf = -1, node = -1;
            if (first) {
                if (lane == 0) leaf = seed;
            } else {
                const int w = min(BF_THREADS, sp);
                if (lane < w) {
                    const uint2 e = stk[sp - 1 - lane];
                    if (!(bitsf(e.y) > bitsf((uint32_t)(worst >> 32)))) {
                        const int r = (int)e.x;
                        if (r < 0) {
                            if (-1 - r != seed) leaf = -1 - r;
                        } else {
                            node = r;
                        }
                    }
                }
                sp -= w;
            }
            first = false;
            const float bound = bitsf((uint32_t)(worst >> 32));
            // nodes: children within the bound onto the stack
            int c[4] = {EMPTY4, EMPTY4, EMPTY4, EMPTY4};
            float d[4] = {0, 0, 0, 0};
            if (node >= 0) {
                nodes++;
                const BvhNode4 nd = bv.nodes4[node];
                const float* f = reinterpret_cast<const float*>(nd.b);
                const int rr[4] = {nd.r.x, nd.r.y, nd.r.z, nd.r.w};
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    d[k] = bf_child_d2(f, k, x, y, z);
                    if (rr[k] != EMPTY4 && d[k] <= bound) c[k] = rr[k];
                }
            }
            int npush = 0;
#pragma unroll
            for (int k = 0; k < 4; k++) npush += __popcll(ballot(c[k] != EMPTY4));
            if (sp + npush > BF_WAVE_STACK) {
                ok = false;
                break;
            }
            __syncthreads();  // the pops above read the slots the pushes overwrite
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint64_t m4 = ballot(c[k] != EMPTY4);
                if (c[k] != EMPTY4) stk[sp + mbcnt(m4)] = make_uint2((uint32_t)c[k], fbits(d[k]));
                sp += __popcll(m4);
            }
            __syncthreads();
            // leaves: the scanned keys below the t-th are candidates (~0: none)
            uint64_t ck[LEAF];
            bool any = false;
            if (leaf >= 0) {
                leaves++;
                const float4* pp = bv.pts4 + (size_t)leaf * LEAF;
#pragma unroll
                for (int q = 0; q < LEAF; q++) {
                    const float4 v = pp[q];
                    const float dd = d2_contract(x, y, z, v.x, v.y, v.z);
                    const uint64_t key = make_key(dd, fbits(v.w));
                    ck[q] = (dd == dd && key < worst) ? key : ~0ull;
                    any = any || ck[q] != ~0ull;
                }
            } else {
#pragma unroll
                for (int q = 0; q < LEAF; q++) ck[q] = ~0ull;
            }
            if (ballot(any) == 0) continue;
            // the new list: t rounds, each taking the smallest remaining key of the old list
            // (lane j's bk) and the candidates, to lane r
            uint64_t nk = empty;
            uint32_t ns = ~0u;
            bool own = lane < t && bk != empty;  // this lane's list entry is still unplaced
            for (int r = 0; r < t; r++) {
                uint64_t lm = own ? bk : ~0ull;
                int li = LEAF;  // LEAF: the list entry
#pragma unroll
                for (int q = 0; q < LEAF; q++) {
                    if (ck[q] < lm) {
                        lm = ck[q];
                        li = q;
                    }
                }
                const uint64_t gm = wave_min_u64(lm);
                if (gm == ~0ull || !(gm < empty)) break;  // nothing left below the bound
                const bool mine = lm == gm;  // keys are unique: exactly one lane
                const int src = __ffsll((long long)ballot(mine)) - 1;
                uint32_t s_ = ~0u;
                if (mine) {
                    if (li == LEAF) {
                        s_ = bs;
                        own = false;
                    } else {
                        s_ = (uint32_t)(leaf * LEAF + li);
#pragma unroll
                        for (int q = 0; q < LEAF; q++)
                            if (q == li) ck[q] = ~0ull;
                    }
                }
                s_ = (uint32_t)shfl((int)s_, src);
                if (lane == r) {
                    nk = gm;
                    ns = s_;
                }
            }
            bk = nk;
            bs = ns;
            // the t-th key (empty while fewer than t are known)
            worst = (uint64_t)shfl((long long)bk, t - 1);
        }
        // the list to lane 0 through LDS
        if (lane < t) {
            fkey[lane] = bk;
            fsp[lane] = bs;
        }
        __syncthreads();
        if (lane == 0) {
            BfList<K> L;
#pragma unroll
            for (int q = 0; q < K; q++) {
                const int o = q - (K - t);
                L.key[q] = q < K - t ? 0ull : fkey[o];
                L.sp[q] = q < K - t ? ~0u : fsp[o];
            }
            out.template emit<K>(i, L, t, ok, x, y, z);
        }
        if (stats) {
            const int lsum = wave_sum_int(leaves), nsum = wave_sum_int(nodes);
            if (lane == 0) {
                atomicAdd(&stats[KS_TILES], 1ull);
                atomicAdd(&stats[KS_CANDIDATES], (unsigned long long)lsum * LEAF);
                atomicAdd(&stats[KS_POPS], (unsigned long long)nsum);
                atomicAdd(&stats[KS_RESOLVED], ok ? 1ull : 0ull);
                atomicAdd(&stats[KS_REFINE], ok ? 0ull : 1ull);
                atomicAdd(&stats[KS_CY_TOTAL], cycles() - c0);
            }
        }
        __syncthreads();
    }
}

// Signed distance from the sorted list: positions and normals summed in ascending key order
// (tcc:435-449).  Queries whose search overflowed are flagged in redo[i] for the tile kernel.
struct BfSdfOut {
    const float* px;
    const float* py;
    const float* pz;
    const float* nx;
    const float* ny;
    const float* nz;
    const uint32_t* qorig;  // query position -> caller's index (NULL: the same)
    float voxel;
    float* out_proj;
    float* out_eucl;
    uint8_t* out_invalid;
    uint32_t* redo;

    template <int K>
    __device__ __forceinline__ void emit(int64_t i, const BfList<K>& L, int t, bool ok, float qx, float qy,
                                         float qz) const
    {
        // an unfilled slot (NaN coordinates) sorts last: the tile kernel handles such queries
        ok = ok && L.full();
        redo[i] = ok ? 0u : 1u;
        if (!ok) return;
        float ax = 0, ay = 0, az = 0, bx = 0, by = 0, bz = 0;
        float v[K][6];
#pragma unroll
        for (int j = 0; j < K; j++) {
            if (j >= K - t) {
                const uint32_t s = L.sp[j];
                v[j][0] = px[s];
                v[j][1] = py[s];
                v[j][2] = pz[s];
                v[j][3] = nx[s];
                v[j][4] = ny[s];
                v[j][5] = nz[s];
            }
        }
#pragma unroll
        for (int j = 0; j < K; j++) {
            if (j >= K - t) {
                ax += v[j][0];
                ay += v[j][1];
                az += v[j][2];
                bx += v[j][3];
                by += v[j][4];
                bz += v[j][5];
            }
        }
        sdf_emit(ax, ay, az, bx, by, bz, t, qx, qy, qz, voxel, qorig ? qorig[i] : (uint32_t)i, out_proj, out_eucl,
                 out_invalid);
    }
};

// kSearch rows in the caller's order: row j < t the j-th key (index, d^2), the rest -1 / +inf.
// With a radius (empty = (r^2, 0)) the unfilled slots are padding too, and count[row] = the
// number of neighbours found.
struct BfRowsOut {
    const uint32_t* qorig;
    int k;
    int32_t* idx;
    float* d2;
    int32_t* count;  // optional
    uint32_t* redo;

    template <int K>
    __device__ __forceinline__ void emit(int64_t i, const BfList<K>& L, int t, bool ok, float, float, float) const
    {
        redo[i] = ok ? 0u : 1u;
        if (!ok) return;
        const size_t row = qorig[i];
        int found = 0;
#pragma unroll
        for (int j = 0; j < K; j++) {
            if (j >= K - t) {
                const int o = j - (K - t);
                const bool real = L.sp[j] != ~0u;
                idx[row * k + o] = real ? (int32_t)(uint32_t)L.key[j] : -1;
                if (d2) d2[row * k + o] = real ? bitsf((uint32_t)(L.key[j] >> 32)) : INFINITY;
                found += real ? 1 : 0;
            }
        }
        for (int o = t; o < k; o++) {
            idx[row * k + o] = -1;
            if (d2) d2[row * k + o] = INFINITY;
        }
        if (count) count[row] = found;
    }
};

// PointsetSurface::getInterpolatedNormal (PointsetSurface.tcc:61-74): a Normal starts as
// (0, 1, 0) (Normal.hpp:62); each of the ki nearest points' normals is normalised
// (Normal(x, y, z), Normal.tcc:49-54) and added; the sum is divided by ki and returned as is
// (the copy into the return value does not normalise again).
struct BfMeanNormalOut {
    const float* nx;
    const float* ny;
    const float* nz;
    const uint32_t* qorig;
    int ki;
    float* out;  // caller order, 3 floats per query
    uint32_t* redo;

    template <int K>
    __device__ __forceinline__ void emit(int64_t i, const BfList<K>& L, int t, bool ok, float, float, float) const
    {
        ok = ok && L.full();
        redo[i] = ok ? 0u : 1u;
        if (!ok) return;
        float rx = 0.0f, ry = 1.0f, rz = 0.0f;
#pragma unroll
        for (int j = 0; j < K; j++) {
            if (j >= K - t) {
                const uint32_t s = L.sp[j];
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
        }
        const float c = (float)ki;
        rx /= c;
        ry /= c;
        rz /= c;
        const size_t o = 3 * (size_t)qorig[i];
        out[o] = rx;
        out[o + 1] = ry;
        out[o + 2] = rz;
    }
};

// The centroid of the neighbours (BilinearFastBox::optimizePlanarFaces, BilinearFastBox.tcc:172-186):
// a zero vector plus each neighbour's position in ascending order, divided by their number.
struct BfCentroidOut {
    const float4* pts;  // sorted points
    const uint32_t* qorig;
    float* out;  // caller order, 3 floats per query
    uint32_t* redo;
    // optional: the query's vertex (caller index -> vertex id) and, per vertex, the t neighbours
    // (sorted positions) of its last search, which a later search of the vertex starts from
    // (prefill): their keys at the new position bound the search at once.  The result is the
    // same exact neighbour set.
    const uint32_t* vid;
    uint32_t* nbr;
    bool prefill;
    uint32_t npts;  // sorted positions
    static constexpr bool kPrefill = true;

    template <int K>
    __device__ __forceinline__ void prefill_list(int64_t i, BfList<K>& L, int t, float qx, float qy, float qz) const
    {
        if (!prefill || !vid || !nbr) return;
        const size_t v = vid[qorig[i]];
        for (int j = 0; j < t; j++) {
            const uint32_t s = nbr[v * t + j];
            if (s >= npts) continue;
            const float4 p = pts[s];
            const float d = d2_contract(qx, qy, qz, p.x, p.y, p.z);
            const uint64_t key = make_key(d, fbits(p.w));
            if (d == d && key < L.worst() && !L.contains(key)) L.insert(key, s);
        }
    }

    template <int K>
    __device__ __forceinline__ void emit(int64_t i, const BfList<K>& L, int t, bool ok, float qx, float qy,
                                         float qz) const
    {
        redo[i] = ok ? 0u : 1u;
        if (!ok) return;
        if (vid && nbr) {
            const size_t v = vid[qorig[i]];
#pragma unroll
            for (int j = 0; j < K; j++)  // compile-time slots: the list stays in registers
                if (j >= K - t) nbr[v * t + (j - (K - t))] = L.sp[j];
        }
        float cx = 0.0f, cy = 0.0f, cz = 0.0f;
        int nk = 0;
#pragma unroll
        for (int j = 0; j < K; j++) {
            if (j >= K - t && L.sp[j] != ~0u) {
                const float4 p = pts[L.sp[j]];
                cx += p.x;
                cy += p.y;
                cz += p.z;
                nk++;
            }
        }
        const size_t o = 3 * (size_t)qorig[i];
        if (nk == 0) {  // nothing found: the position stays
            out[o] = qx;
            out[o + 1] = qy;
            out[o + 2] = qz;
            return;
        }
        const float c = (float)nk;
        out[o] = cx / c;
        out[o + 1] = cy / c;
        out[o + 2] = cz / c;
    }
};

}  // namespace lvrhip
