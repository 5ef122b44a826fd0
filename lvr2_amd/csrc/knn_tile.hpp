// Exact kNN for tiles of Morton-ordered queries, one wavefront per tile.
//
// Replaces the per-thread register-queue kNN of the reference (lbvh/query.cuh:45-167,
// static_priorityqueue.cuh:18-93, query_knn_kernels.cu:9-48) and the CPU kd-tree search
// it must agree with (KDTree.tcc:99-126, 186-248).
//
// A wave owns Q = 64/P consecutive queries, P lanes per query.  For a radius R it
//   1. walks the LBVH once for the whole tile (wave-parallel, LDS stack), collecting in LDS
//      the leaves whose box is within R of the tile's box;
//   2. stages the points of those leaves in LDS (chunks of CAP points; when they fit in one
//      chunk they stay resident for all passes);
//   3. gives every query its own list of the chunk's leaves whose box is within that
//      query's current bound, and scans only those, the P lanes of the query splitting
//      each leaf's points;
//   4. per query, counts candidates with d^2 <= R^2 into bins linear in d^2.  If fewer
//      than t candidates are in range the query stays pending and the tile retries with
//      2R; otherwise the bin holding the t-th smallest key is found, refined by radix
//      passes over the 64-bit key (d^2 bits, point index) until it holds at most BUF keys,
//      and a collection pass (bounded by that bin's upper edge) hands every key below the
//      bin to the sink, buffers the bin, and ranks the buffer exactly.
// Every point within R of a query is a candidate (box distances are monotone lower bounds of
// the fp32 d^2), so a query with >= t in-range candidates gets the exact t smallest keys.
// Two or four targets are selected at once (e.g. k for the plane fit and ki for the
// interpolation list, or the four neighbourhood sizes of the later adaptive rounds).  If a tile's leaf list outgrows LDS, each pass re-walks the tree and streams the
// leaves chunk by chunk instead.
#pragma once

#include <type_traits>

#include "common.hpp"
#include "lbvh.hpp"

namespace lvrhip {

// LDS budget: a wave of 8 queries (P = 8) with one or two targets needs 12.4 / 13.0 KB, so that
// twelve waves (three per SIMD) fit a CU's 160 KB; with the register budget of three waves
// (168 VGPRs) that occupancy hides the LDS and L2 round trips of the passes better than bigger
// chunks and finer bins at two waves per SIMD.
// Histogram per query: HIST_WORDS 32-bit words.  A pass over a tile's leaf list scans at most
// TL_CAP leaves per query, fewer than 65,536 points, so its bins fit 16-bit counters, two to a
// word (2 HIST_WORDS bins, HIST_BITS + 1 key bits per refinement pass; used by one-target sinks);
// a streamed pass (no bound on its points) counts HIST_WORDS bins in whole words.
constexpr int HIST_WORDS = 64;
constexpr int HIST_BITS = 6;  // log2(HIST_WORDS)
constexpr int CAP = 256;       // candidate slots per chunk
constexpr int CHUNK_LEAVES = CAP / LEAF;
constexpr int TL_CAP = 448;    // leaves kept per tile (a run's leaf list lives on the stack: <= STACK_CAP)
constexpr int STACK_CAP = 448;
// A pop of w entries pushes at most 4 per entry (3 w more slots).  Wide pops keep STACK_HEADROOM
// slots free for the depth-first pops of one entry that follow when the stack is nearly full.
constexpr int STACK_HEADROOM = 64;
static_assert(TL_CAP <= STACK_CAP, "a run's leaf list is kept on the stack");
static_assert(TL_CAP * LEAF < 65536, "16-bit bins: a list pass scans fewer than 65,536 points per query");
static_assert(STACK_CAP >= FRONTIER, "a traversal starts with up to FRONTIER subtrees on the stack");
static_assert(CHUNK_LEAVES % (kWave / LEAF) == 0 && CHUNK_LEAVES <= kWave, "a chunk is staged in whole wave steps");
constexpr int MAX_ATTEMPTS = 1024;    // per tile (a safety net: each cluster of queries grows x4 per
                                      // attempt after CLUSTER_GUARD of its own)
constexpr int CLUSTER_SHRINK = 12;    // a cluster's first attempts that may shrink R or cap its growth
constexpr int CLUSTER_GUARD = 24;
constexpr int QUEUE_SEGS = 8;     // work-queue segments: one per XCD
constexpr int QUEUE_STRIDE = 16;  // their counters 64 bytes apart
constexpr int QUEUE_CHUNK = 32;   // consecutive runs dealt to one segment at a time
constexpr float RUN_R2_MARGIN = 2.0f;  // radius^2 of a run's leaf list over its second tile's first

struct BvhView {
    const float* px;
    const float* py;
    const float* pz;
    const uint32_t* pidx;  // original index of each sorted point
    const uint32_t* rank;  // sorted position of each original index
    const float4* pts4;    // x, y, z and original index (as bits) of each sorted point
    const float4* leaf_lo;
    const float4* leaf_hi;
    const float4* leaf_agg;  // per leaf: its point sums about the box centre (leaf_sums_kernel)
    const BvhNode* nodes;
    const BvhNode4* nodes4;  // the same nodes with their grandchildren (node4_kernel)
    int nleaf;
    uint32_t n;
    const int* roots;         // per scan: root node of its subtree (NULL: one scan, root 0)
    const int* fr_ref;        // per scan: FRONTIER top-level subtrees (NULL: start at the root)
    const float4* fr_lo;
    const float4* fr_hi;
    const uint16_t* scan_of;  // per query position: its scan (NULL: one scan)
};

struct QueryView {
    const float* qx;  // query coordinates indexed by query position
    const float* qy;
    const float* qz;
    const uint32_t* qsel;  // optional: tile slot -> query position
    uint32_t nq;           // number of tile slots
    const uint32_t* nq_dev;  // optional: the number of tile slots, read on the device
};

struct TileParams {
    int t0, t1, t2, t3;  // targets, already clamped to n: t0 the largest (it decides whether a
                         // radius suffices), the others <= t0 (0 for none); a sink uses
                         // Sink::kTargets of them
    int tiles_per_wave;
    float r2_init;
    const float* area_pp;       // optional device value, area per point: first radius^2 =
                                // area_pp * t0 / pi * r2_margin (else r2_init)
    float r2_margin;            // next tile's radius^2 = margin * largest k-th d^2 of this tile
    float grow_min;             // smallest growth of the disc term R^2 - dmin^2 after a failed attempt
    int* err;                   // set to a nonzero code on capacity overflow
    unsigned long long* stats;  // optional counters, see KS_* below
    uint32_t* queue;            // optional work queue (QUEUE_SEGS counters QUEUE_STRIDE apart,
                                // zeroed before the launch): waves take the next run of
                                // tiles_per_wave tiles from it; else runs are strided over the grid
    const float* hint_in;       // optional, per query position: the d^2 of its t-th key for an earlier
                                // target t; a tile's first radius^2 is hint_scale times the largest
                                // over its queries (else the radius carried from the previous tile)
    float hint_scale;
    float* hint_out;            // optional, per query position: the d^2 of the t0-th key of every
                                // query resolved (may be hint_in: a query reads it before)
    int leaf_hint;              // the queries are the cloud's own sorted points (query position =
                                // sorted position): a run's first tile sizes its radius from the
                                // boxes of its queries' leaves instead of the cloud's mean density
};

enum {
    KS_TILES = 0,   // tiles processed
    KS_ATTEMPTS,    // radius attempts (>= tiles)
    KS_RESIDENT,    // attempts whose candidates fit one LDS chunk
    KS_CANDIDATES,  // (query, leaf) pairs scanned by the histogram pass, in points
    KS_REFINE,      // refinement passes (wave-level)
    KS_CHUNKS,      // chunk loads over all passes
    KS_RESOLVED,    // queries resolved
    KS_POPS,        // stack entries expanded by the traversal
    // shader-clock cycles per phase, summed over waves
    KS_CY_TRAV,     // tile traversal
    KS_CY_LOAD,     // chunk staging
    KS_CY_GLIST,    // per-query leaf lists
    KS_CY_HIST,     // histogram pass (scan only)
    KS_CY_SELECT,   // bin search + refinement passes
    KS_CY_COLLECT,  // collection pass (scan only)
    KS_CY_FINISH,   // boundary ranking + sink finish
    KS_CY_TOTAL,    // whole wave
    KS_COUNT
};

__device__ __forceinline__ unsigned long long cycles() { return (unsigned long long)clock64(); }

// Wave-local counters, flushed with one atomic per counter at the end of the wave.
struct WaveStats {
    unsigned long long c[KS_COUNT];
    bool timing;
    __device__ void clear()
    {
#pragma unroll
        for (int i = 0; i < KS_COUNT; i++) c[i] = 0;
    }
    __device__ __forceinline__ unsigned long long tick() const { return timing ? cycles() : 0ull; }
    __device__ __forceinline__ void tock(int i, unsigned long long t0)
    {
        if (timing) c[i] += cycles() - t0;
    }
    __device__ __forceinline__ void flush(unsigned long long* dst)
    {
        if (!dst || lane_id() != 0) return;
#pragma unroll
        for (int i = 0; i < KS_COUNT; i++)
            if (c[i]) atomicAdd(&dst[i], c[i]);
    }
};

template <int Q, int BUF, int NT>
struct TileSmem {
    // x, y, z, original index (as bits); slots [CAP, CAP + LEAF) form an all-NaN leaf (chunk
    // slot CHUNK_LEAVES) that pads the per-query lists to whole steps of the scan
    float4 cand[CAP + LEAF];
    float4 lbox[CHUNK_LEAVES][2];  // boxes of the chunk's leaves (lo, hi)
    uint32_t hist[Q][HIST_WORDS];
    int stack[STACK_CAP];
    int tleaf[TL_CAP];                    // the tile's leaves (or the current chunk's, when streaming)
    uint8_t glist[Q][CHUNK_LEAVES + 4];   // per query: chunk leaf slots within its bound, then 4 pads
    uint8_t flist[Q][CHUNK_LEAVES];       // per query: chunk leaf slots taken whole (collection)
    uint64_t bkey[NT][Q][BUF];  // boundary keys (their low half, the original index, finds the point)
    uint32_t bcnt[NT][Q];
    uint32_t scnt[NT][Q];  // per-query sink slot counters
};

// ---- traversal ---------------------------------------------------------------------------

__device__ __forceinline__ float box_d2(float lx, float ly, float lz, float hx, float hy, float hz, float tlx,
                                        float tly, float tlz, float thx, float thy, float thz)
{
    // gap per axis, then (gx^2 + gy^2) + gz^2 -- the same association as d2_contract
    float gx = fmaxf(fmaxf(lx - thx, tlx - hx), 0.0f);
    float gy = fmaxf(fmaxf(ly - thy, tly - hy), 0.0f);
    float gz = fmaxf(fmaxf(lz - thz, tlz - hz), 0.0f);
    return (gx * gx + gy * gy) + gz * gz;
}

struct TileBox {
    float lx, ly, lz, hx, hy, hz;
};

// Appends to sm.tleaf[0..cap) the leaves whose box is within sqrt(r2) of the tile box,
// continuing the traversal held on the LDS stack (sp entries).  Stops when the stack is
// empty or the list is full; returns the list length (wave-uniform).  sp = -1 on overflow.
// With rej != NULL, *rej is lowered to the smallest box distance of a child the traversal
// rejected (a lower bound of the distance of every point outside the list).  With
// store = false the leaves are only counted (up to cap).
template <class Sm>
__device__ __forceinline__ int fill_leaves(Sm& sm, const BvhView& bv, const TileBox& tb, float r2, int& sp, int cap,
                                           WaveStats& ws, float* rej = nullptr, bool store = true)
{
    constexpr int NONE = 0x7fffffff;
    const int lane = lane_id();
    int nl = 0;
    float rmin = INFINITY;
    while (sp > 0) {
        const int room = cap - nl;
        if (room == 0) break;
        // Wide pops (up to 64 entries, as many as leave STACK_HEADROOM slots free after their
        // pushes); below that, depth-first pops of one entry (up to 4 pushes), whose growth is
        // bounded by the tree depth.  A popped node's four grandchildren are tested at once
        // (BvhNode4).
        const int free_slots = STACK_CAP - sp;
        const int wide = min(min(kWave, sp), (free_slots - STACK_HEADROOM) / 3);
        const int w = wide >= 1 ? wide : (free_slots >= 4 ? 1 : 0);
        if (w <= 0) {
            sp = -1;
            return nl;
        }
        ws.c[KS_POPS] += (unsigned)w;
        const int ref = lane < w ? sm.stack[sp - 1 - lane] : NONE;
        sp -= w;
        __syncthreads();
        int o[4] = {NONE, NONE, NONE, NONE};
        if (ref != NONE) {
            if (ref < 0) {
                if (-1 - ref < bv.nleaf) o[0] = ref;  // a leaf pushed back: already tested
            } else if (ref < bv.nleaf - 1) {
                const BvhNode4 nd = bv.nodes4[ref];
                const float* f = reinterpret_cast<const float*>(nd.b);
                const int rr[4] = {nd.r.x, nd.r.y, nd.r.z, nd.r.w};
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const float d = box_d2(f[6 * k], f[6 * k + 1], f[6 * k + 2], f[6 * k + 3], f[6 * k + 4],
                                           f[6 * k + 5], tb.lx, tb.ly, tb.lz, tb.hx, tb.hy, tb.hz);
                    if (d <= r2 && rr[k] != EMPTY4) o[k] = rr[k];
                    else rmin = fminf(rmin, d);
                }
            }
        }
        // leaf outputs in the order (output slot, lane); the ones past the list's room and the
        // internal nodes go on the stack
        uint64_t bl[4];
        int nleaves = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            bl[k] = ballot(o[k] != NONE && o[k] < 0);
            nleaves += __popcll(bl[k]);
        }
        const int take = min(nleaves, room);
        bool push[4];
        int before = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const bool lk = o[k] != NONE && o[k] < 0;
            const int rk = before + mbcnt(bl[k]);
            before += __popcll(bl[k]);
            if (store && lk && rk < take) sm.tleaf[nl + rk] = -1 - o[k];
            push[k] = o[k] != NONE && (!lk || rk >= take);
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint64_t bp = ballot(push[k]);
            if (push[k]) sm.stack[sp + mbcnt(bp)] = o[k];
            sp += __popcll(bp);
        }
        nl += take;
        __syncthreads();
    }
    if (rej) *rej = fminf(*rej, uniformf(wave_min(rmin)));
    return nl;
}
template <class Sm>
__device__ __forceinline__ int start_traversal(Sm& sm, const BvhView& bv, int cur, const TileBox& tb, float r2,
                                               float* rej = nullptr)
{
    const int lane = lane_id();
    if (!bv.fr_ref) {
        const int root = bv.roots ? uniform(bv.roots[cur]) : (bv.nleaf > 1 ? 0 : -1);
        if (lane == 0) sm.stack[0] = root;
        __syncthreads();
        return 1;
    }
    const size_t f0 = (size_t)cur * FRONTIER;
    int sp = 0;
    float rmin = INFINITY;
#pragma unroll
    for (int k = 0; k < FRONTIER / kWave; k++) {
        const size_t j = f0 + k * kWave + lane;
        const int ref = bv.fr_ref[j];
        bool hit = false;
        if (ref != FR_NONE) {
            const float4 lo = bv.fr_lo[j], hi = bv.fr_hi[j];
            const float d = box_d2(lo.x, lo.y, lo.z, hi.x, hi.y, hi.z, tb.lx, tb.ly, tb.lz, tb.hx, tb.hy, tb.hz);
            hit = d <= r2;
            if (!hit) rmin = fminf(rmin, d);
        }
        const uint64_t m = ballot(hit);
        if (hit) sm.stack[sp + mbcnt(m)] = ref;
        sp += __popcll(m);
    }
    if (rej) *rej = fminf(*rej, uniformf(wave_min(rmin)));
    __syncthreads();
    return sp;
}

// The leaves of a run's list (leaf indices in sm.stack[0, nrun), a superset of every leaf within
// sqrt(r2) of the tile box) whose box is within sqrt(r2) of the tile box, into sm.tleaf: exactly
// the leaves a traversal for (tb, r2) finds, without walking the tree.  Returns their count;
// *rej as in fill_leaves (over the list's rejected leaves).
template <class Sm>
__device__ __forceinline__ int filter_leaves(Sm& sm, const BvhView& bv, int nrun, const TileBox& tb, float r2,
                                             float* rej)
{
    const int lane = lane_id();
    int nl = 0;
    float rmin = INFINITY;
    for (int j0 = 0; j0 < nrun; j0 += kWave) {
        const int j = j0 + lane;
        bool hit = false;
        int leaf = 0;
        if (j < nrun) {
            leaf = sm.stack[j];
            const float4 lo = bv.leaf_lo[leaf], hi = bv.leaf_hi[leaf];
            const float d = box_d2(lo.x, lo.y, lo.z, hi.x, hi.y, hi.z, tb.lx, tb.ly, tb.lz, tb.hx, tb.hy, tb.hz);
            hit = d <= r2;
            if (!hit) rmin = fminf(rmin, d);
        }
        const uint64_t m = ballot(hit);
        if (hit) sm.tleaf[nl + mbcnt(m)] = leaf;
        nl += __popcll(m);
    }
    if (rej) *rej = fminf(*rej, uniformf(wave_min(rmin)));
    __syncthreads();
    return nl;
}

// Stage the points of leaves sm.tleaf[base, base + nl) in LDS; slot = j * LEAF + offset.
// Slots past the end of the cloud get NaN coordinates (never in range).
template <class Sm>
__device__ __forceinline__ void load_chunk(Sm& sm, const BvhView& bv, int base, int nl)
{
    constexpr int LPI = kWave / LEAF;             // leaves per iteration
    constexpr int IT = CHUNK_LEAVES / LPI;         // iterations for a full chunk
    const int lane = lane_id();
    float4 v[IT];
    // all loads first, then the LDS stores: IT global loads in flight per lane
#pragma unroll
    for (int it = 0; it < IT; it++) {
        const int jl = it * LPI + lane / LEAF;
        v[it] = make_float4(__builtin_nanf(""), __builtin_nanf(""), __builtin_nanf(""), 0.0f);
        if (jl < nl) {
            const uint32_t spos = (uint32_t)sm.tleaf[base + jl] * LEAF + (lane % LEAF);
            if (spos < bv.n) v[it] = bv.pts4[spos];
        }
    }
    float4 bl = make_float4(0, 0, 0, 0), bh = make_float4(0, 0, 0, 0);
    if (lane < nl) {
        const int leaf = sm.tleaf[base + lane];
        bl = bv.leaf_lo[leaf];
        bh = bv.leaf_hi[leaf];
    }
#pragma unroll
    for (int it = 0; it < IT; it++) {
        const int jl = it * LPI + lane / LEAF;
        if (jl < nl) sm.cand[jl * LEAF + (lane % LEAF)] = v[it];
    }
    if (lane < nl) {
        sm.lbox[lane][0] = bl;
        sm.lbox[lane][1] = bh;
    }
    __syncthreads();
}

// sorted position of candidate slot c of the chunk whose leaves start at sm.tleaf[base]
template <class Sm>
__device__ __forceinline__ uint32_t slot_spos(const Sm& sm, int base, int c)
{
    const uint32_t u = (uint32_t)c;  // slots are non-negative: unsigned division is a shift
    return (uint32_t)sm.tleaf[base + (int)(u / LEAF)] * LEAF + u % LEAF;
}

// Per query group g: the chunk's leaves (sm.tleaf[base, base+nl)) whose box lies within
// sqrt(rq) of the query point, as chunk slots in sm.glist[g]; returns their count
// (replicated over the group's lanes).  rq < 0 disables the query.
// With rin >= 0, leaves whose farthest corner is closer than sqrt(rin) are left out too:
// every point of such a leaf has d^2 < rin (the fp32 corner distance bounds the fp32 point
// distance from above, all roundings being monotone).
template <int P, class Sm>
__device__ __forceinline__ int build_glist(Sm& sm, const BvhView& bv, int base, int nl, int g, int s, float qx,
                                           float qy, float qz, float rq, float rin = -1.0f)
{
    constexpr uint64_t GM = (P == 64) ? ~0ull : ((1ull << P) - 1ull);
    int len = 0;
    for (int j0 = 0; j0 < nl; j0 += P) {
        const int j = j0 + s;
        bool hit = false;
        if (j < nl && rq >= 0.0f) {
            const int leaf = sm.tleaf[base + j];
            const float4 lo = sm.lbox[j][0], hi = sm.lbox[j][1];
            hit = box_d2(lo.x, lo.y, lo.z, hi.x, hi.y, hi.z, qx, qy, qz, qx, qy, qz) <= rq;
            if (hit && rin > 0.0f) {
                const float fx = fmaxf(fabsf(qx - lo.x), fabsf(qx - hi.x));
                const float fy = fmaxf(fabsf(qy - lo.y), fabsf(qy - hi.y));
                const float fz = fmaxf(fabsf(qz - lo.z), fabsf(qz - hi.z));
                hit = !((fx * fx + fy * fy) + fz * fz < rin);
            }
        }
        const uint32_t gm = (uint32_t)((ballot(hit) >> (g * P)) & GM);
        if (hit) sm.glist[g][len + __popc(gm & ((1u << s) - 1u))] = (uint8_t)j;
        len += __popc(gm);
    }
    // pads: the NaN leaf, so that scan_glist reads whole steps without bounds tests
    for (int e = s; e < 4; e += P) sm.glist[g][len + e] = (uint8_t)CHUNK_LEAVES;
    __syncthreads();
    return len;
}

// build_glist for a collection pass whose sink takes whole leaves: the leaves within sqrt(rq)
// whose farthest corner is closer than sqrt(rfull) (every point of them has d^2 < rfull) go
// to sm.flist[g] (nfull of them), the others to sm.glist[g] (the return value).
template <int P, class Sm>
__device__ __forceinline__ int build_glist_split(Sm& sm, int nl, int g, int s, float qx, float qy, float qz, float rq,
                                                 float rfull, int& nfull)
{
    constexpr uint64_t GM = (P == 64) ? ~0ull : ((1ull << P) - 1ull);
    int len = 0;
    nfull = 0;
    for (int j0 = 0; j0 < nl; j0 += P) {
        const int j = j0 + s;
        bool hit = false, full = false;
        if (j < nl && rq >= 0.0f) {
            const float4 lo = sm.lbox[j][0], hi = sm.lbox[j][1];
            hit = box_d2(lo.x, lo.y, lo.z, hi.x, hi.y, hi.z, qx, qy, qz, qx, qy, qz) <= rq;
            if (hit && rfull > 0.0f) {
                const float fx = fmaxf(fabsf(qx - lo.x), fabsf(qx - hi.x));
                const float fy = fmaxf(fabsf(qy - lo.y), fabsf(qy - hi.y));
                const float fz = fmaxf(fabsf(qz - lo.z), fabsf(qz - hi.z));
                full = (fx * fx + fy * fy) + fz * fz < rfull;
            }
        }
        const bool part = hit && !full;
        const uint32_t below = (1u << s) - 1u;
        const uint32_t pm = (uint32_t)((ballot(part) >> (g * P)) & GM);
        const uint32_t fm = (uint32_t)((ballot(full) >> (g * P)) & GM);
        if (part) sm.glist[g][len + __popc(pm & below)] = (uint8_t)j;
        if (full) sm.flist[g][nfull + __popc(fm & below)] = (uint8_t)j;
        len += __popc(pm);
        nfull += __popc(fm);
    }
    for (int e = s; e < 4; e += P) sm.glist[g][len + e] = (uint8_t)CHUNK_LEAVES;
    __syncthreads();
    return len;
}

// Sinks that take a whole leaf at once in the collection pass declare kLeafAccept = true and
// accept_leaf(g, s, count, centre, sums, box).
template <class S, class = void>
struct leaf_accepting : std::false_type {
};
template <class S>
struct leaf_accepting<S, std::void_t<decltype(S::kLeafAccept)>> : std::integral_constant<bool, S::kLeafAccept> {
};

// The candidates of one radius attempt: either the tile's leaf list (ntl >= 0; resident when
// it fits one chunk) or, after a list overflow, a fresh traversal per pass.
struct Cands {
    int ntl;        // leaves in sm.tleaf, or -1 when streaming
    bool resident;  // all candidates already staged in sm.cand
    int staged;     // first list entry of the chunk staged in LDS (-1: none)
    bool reverse;   // the next pass walks the chunks backwards (it starts with the staged one)
};

// Calls f(nl, base) once per staged chunk of nl leaves starting at sm.tleaf[base].
// The kernel's helpers are all inlined: a call would move the structs it passes by
// reference (views, counters) into scratch memory.
template <class Sm, class F>
__device__ __forceinline__ bool for_each_chunk(Sm& sm, const BvhView& bv, const TileBox& tb, float R2, int cur,
                                               Cands& cs, WaveStats& ws, F&& f)
{
    if (cs.ntl >= 0) {
        if (cs.resident) {
            f(cs.ntl, 0);
            __syncthreads();
            return true;
        }
        // passes alternate direction, so each starts with the chunk the last one left staged
        const int nch = (cs.ntl + CHUNK_LEAVES - 1) / CHUNK_LEAVES;
        for (int i = 0; i < nch; i++) {
            const int base = (cs.reverse ? nch - 1 - i : i) * CHUNK_LEAVES;
            const int nl = min(CHUNK_LEAVES, cs.ntl - base);
            if (base != cs.staged) {
                const auto t0 = ws.tick();
                load_chunk(sm, bv, base, nl);
                ws.tock(KS_CY_LOAD, t0);
                ws.c[KS_CHUNKS] += 1;
                cs.staged = base;
            }
            f(nl, base);
            __syncthreads();
        }
        cs.reverse = !cs.reverse;
        if (cs.ntl <= CHUNK_LEAVES) cs.resident = true;
        return true;
    }
    int sp = start_traversal(sm, bv, cur, tb, R2);
    for (;;) {
        auto t0 = ws.tick();
        const int nl = fill_leaves(sm, bv, tb, R2, sp, CHUNK_LEAVES, ws);
        ws.tock(KS_CY_TRAV, t0);
        if (sp < 0) return false;
        t0 = ws.tick();
        load_chunk(sm, bv, 0, nl);
        ws.tock(KS_CY_LOAD, t0);
        ws.c[KS_CHUNKS] += 1;
        f(nl, 0);
        __syncthreads();
        if (sp == 0) return true;
    }
}

// Scan the points of query group g's listed leaves: fn(p, slot) for each, the P lanes of
// the group taking every P-th point of a leaf (LEAF / P independent LDS reads per lane).
// The reads of a step are unconditional (a list's last step runs into its NaN pads), so they
// are all in flight before the first use; fn must ignore points with NaN coordinates.
template <int P, class Sm, class Fn>
__device__ __forceinline__ void scan_glist(const Sm& sm, int g, int s, int len, Fn&& fn)
{
    constexpr int U = LEAF / P;  // points per lane per leaf
    constexpr int L = U >= 4 ? 1 : 4 / U;  // leaves per step: at least 4 LDS reads in flight
    static_assert(LEAF % P == 0, "lanes per query must divide the leaf size");
    // the next step's list entries are read one step ahead: a step reads up to 2 L - 2 entries past
    // the row's last leaf, which the row's 4 pads (the NaN leaf) cover
    static_assert(2 * L - 2 <= 4, "the lists carry 4 pads");
    int nxt[L];
#pragma unroll
    for (int l = 0; l < L; l++) nxt[l] = sm.glist[g][l];
    for (int j = 0; j < len; j += L) {
        int slot0[L];
        float4 p[L][U];
#pragma unroll
        for (int l = 0; l < L; l++) slot0[l] = nxt[l] * LEAF + s;
#pragma unroll
        for (int l = 0; l < L; l++)
#pragma unroll
            for (int u = 0; u < U; u++) p[l][u] = sm.cand[slot0[l] + u * P];
#pragma unroll
        for (int l = 0; l < L; l++) nxt[l] = sm.glist[g][j + L + l];
#pragma unroll
        for (int l = 0; l < L; l++)
#pragma unroll
            for (int u = 0; u < U; u++) fn(p[l][u], slot0[l] + u * P);
    }
}

// ---- histogram helpers ----------------------------------------------------------------------

// Locate the bin holding the t-th smallest counted key of query group g (P lanes).
// Returns total count; if t <= total: bin, below (count in lower bins), cnt (in that bin).
// Bins b of a histogram in 16-bit halves (hsh = 1: word b / 2, half b % 2) or whole words (hsh = 0).
__device__ __forceinline__ void hist_add(uint32_t* h, int b, uint32_t hsh)
{
    atomicAdd(&h[b >> hsh], 1u << ((b & (int)hsh) << 4));
}
template <bool kHalf>
__device__ __forceinline__ int hist_bin(const uint32_t (*hist)[HIST_WORDS], int g, int b)
{
    if constexpr (kHalf) return (int)((hist[g][b >> 1] >> ((b & 1) << 4)) & 0xffffu);
    return (int)hist[g][b];
}

template <int P, int Q, int NBT>
__device__ __forceinline__ int find_bin(const uint32_t (*hist)[HIST_WORDS], int g, int s, int t, int& bin,
                                        int& below, int& cnt)
{
    constexpr bool kHalf = NBT == 2 * HIST_WORDS;
    static_assert(NBT == HIST_WORDS || kHalf, "whole words or 16-bit halves");
    constexpr int BPL = NBT / P;
    static_assert(NBT % P == 0, "bins per lane");
    int sum = 0;
#pragma unroll
    for (int j = 0; j < BPL; j++) sum += hist_bin<kHalf>(hist, g, s * BPL + j);
    const int incl = lane_group_scan_int<P>(sum, s);
    const int total = lane_group_sum_int<P>(sum);
    const int excl = incl - sum;
    bool mine = (t >= 1) && (excl < t) && (t <= incl);
    int b_ = 0, below_ = 0, cnt_ = 0;
    if (mine) {
        int run = excl;
#pragma unroll
        for (int j = 0; j < BPL; j++) {
            const int b = s * BPL + j;
            const int v = hist_bin<kHalf>(hist, g, b);
            if (run < t && t <= run + v) {
                b_ = b;
                below_ = run;
                cnt_ = v;
            }
            run += v;
        }
    }
    // one lane of the group at most owns the t-th key (the others hold zeros)
    bin = lane_group_sum_int<P>(b_);
    below = lane_group_sum_int<P>(below_);
    cnt = lane_group_sum_int<P>(cnt_);
    return total;
}

__device__ __forceinline__ int vbin(float d2, float inv_w, int nb)
{
    float f = d2 * inv_w;
    int b = (int)f;
    return b < nb - 1 ? b : nb - 1;
}

// Smallest non-negative float x with vbin(x, inv_w, nb) >= b (b >= 1).
__device__ __forceinline__ float vbin_floor(int b, float inv_w, int nb)
{
    float x = (float)b / inv_w;
    if (!(x >= 0.0f) || isinf(x)) x = 0.0f;
    for (int it = 0; it < 256 && x > 0.0f && vbin(bitsf(fbits(x) - 1u), inv_w, nb) >= b; it++)
        x = bitsf(fbits(x) - 1u);
    for (int it = 0; it < 256 && vbin(x, inv_w, nb) < b; it++) x = bitsf(fbits(x) + 1u);
    return x;
}

// ---- the kernel -----------------------------------------------------------------------------

// Per-target selection state (replicated over the P lanes of a query).
struct Target {
    int t;          // wanted count
    int below;      // keys below klo
    int cnt;        // keys in [klo, khi]
    uint64_t klo, khi;
    float dlo, dhi;  // float d2 of klo / khi (fast path of the collection test)
    bool refine;
};

// Sinks with an `lds` member get the candidate staging area, free once a query's collection
// and ranking are done (their finish may use it).
template <class S>
__device__ __forceinline__ auto give_lds(S& s, float4* p, int) -> decltype(s.lds = p, void())
{
    s.lds = p;
}
template <class S>
__device__ __forceinline__ void give_lds(S&, float4*, long)
{
}

// The register budget: three waves per SIMD where the LDS allows them (P = 8 with one or two
// targets), else up to two.
template <int P, class Sink>
constexpr bool tile_three_waves()
{
    return P == 8 && Sink::kTargets <= 2;
}

template <int P, class Sink>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(tile_three_waves<P, Sink>() ? 3 : 1,
                                                                       tile_three_waves<P, Sink>() ? 3 : 2))) void knn_tile_kernel(
    BvhView bv, QueryView qv, TileParams prm, Sink sink)
{
    constexpr int Q = kWave / P;
    constexpr int BUF = Sink::kTargets == 1 ? 24 : 16;  // boundary keys per query and target
    constexpr int NT = Sink::kTargets;
    static_assert(NT == 1 || NT == 2 || NT == 4, "one, two or four targets");
    __shared__ TileSmem<Q, BUF, NT> sm;
    const int lane = lane_id();
    const int g = lane / P, s = lane % P;
    float r2 = fmaxf(prm.r2_init, 1e-30f);
    WaveStats ws;
    ws.clear();
    ws.timing = prm.stats != nullptr;
    const auto t_wave = ws.tick();
    give_lds(sink, sm.cand, 0);
    if (lane < LEAF) sm.cand[CAP + lane] = make_float4(__builtin_nanf(""), __builtin_nanf(""), __builtin_nanf(""), 0.0f);
    const uint32_t nq = qv.nq_dev ? uniform(*qv.nq_dev) : qv.nq;
    if (prm.area_pp) r2 = fmaxf(uniformf(*prm.area_pp) * (float)prm.t0 / 3.14159265f * prm.r2_margin, 1e-30f);
    const float r2_first = r2;
    const int64_t ntiles = ((int64_t)nq + Q - 1) / Q;

    // Runs of tiles_per_wave consecutive tiles, taken from the work queue (every wave ends once
    // it is drained) or strided over the grid (a grid sized for an upper bound of the
    // device-side count leaves the surplus waves idle).  The queue deals blocks of QUEUE_CHUNK
    // consecutive runs to the QUEUE_SEGS XCDs in turn (workgroups are dealt to the XCDs in
    // turn too), each XCD with its own counter: neighbouring tiles, which read the same leaves,
    // share one L2, and the runs are still taken roughly in order (a costly region is not left
    // to the end).  A wave whose XCD's blocks are drained helps the others in turn.
    const int64_t nruns = (ntiles + prm.tiles_per_wave - 1) / prm.tiles_per_wave;
    const int xcd = (int)(blockIdx.x % QUEUE_SEGS);
    int seg_k = 0;
    int64_t run = -1;
    for (;;) {
        if (prm.queue) {
            run = -1;
            while (seg_k < QUEUE_SEGS) {
                const int seg = (xcd + seg_k) % QUEUE_SEGS;
                uint32_t r = 0;
                if (lane == 0) r = atomicAdd(prm.queue + seg * QUEUE_STRIDE, 1u);
                r = (uint32_t)__builtin_amdgcn_readlane((int)r, 0);
                const int64_t cand = ((int64_t)(r / QUEUE_CHUNK) * QUEUE_SEGS + seg) * QUEUE_CHUNK + r % QUEUE_CHUNK;
                if (cand < nruns) {
                    run = cand;
                    break;
                }
                seg_k++;  // this XCD's runs only grow with r: all taken
            }
            if (run < 0) break;
        } else {
            run = run < 0 ? (int64_t)blockIdx.x : run + gridDim.x;
        }
        if (run * prm.tiles_per_wave >= ntiles) break;
        // every run starts from the same radius guess: results do not depend on which wave
        // takes which run (the radius only decides the order in which candidates are seen)
        r2 = r2_first;
        float disc_hint = 0.0f;  // R^2 - dmin^2 that held t0 points for the last resolved queries
        // The run's leaf list (sm.stack[0, nrun)): built once, at its second tile, for the box of
        // its remaining queries (of one scan) and RUN_R2_MARGIN times that tile's first radius^2;
        // attempts within it filter the list instead of walking the tree.  A traversal reuses
        // the stack and ends it.
        bool run_ok = false;
        int nrun = 0, run_scan = 0;
        float run_r2 = 0.0f;
        TileBox rbox;
    for (int tt = 0; tt < prm.tiles_per_wave; tt++) {
        const int64_t tile = run * prm.tiles_per_wave + tt;
        const int64_t qbase = tile * Q;
        if (tile >= ntiles) break;
        const int64_t qi = qbase + g;
        const bool valid = qi < nq;
        const uint32_t qpos = valid ? (qv.qsel ? qv.qsel[qi] : (uint32_t)qi) : 0u;
        const float qx = valid ? qv.qx[qpos] : 0.0f;
        const float qy = valid ? qv.qy[qpos] : 0.0f;
        const float qz = valid ? qv.qz[qpos] : 0.0f;
        const int scan = (valid && bv.scan_of) ? (int)bv.scan_of[qpos] : 0;
        bool active = valid && prm.t0 > 0;
        ws.c[KS_TILES] += 1;
        if (prm.hint_in) {
            // the earlier target's neighbourhood of each query, grown for this one
            const float h = uniformf(wave_max(valid ? prm.hint_in[qpos] : 0.0f));
            if (h > 0.0f && h < 1e37f) r2 = fmaxf(h * prm.hint_scale, 1e-30f);
        } else if (prm.leaf_hint && tt == 0) {
            // a leaf's 16 points on a surface patch: its box's two largest extents a, b span
            // about (16/9) a b of surface (a 4 x 4 lattice spans 3 x 3 spacings), so t0 points
            // lie within pi R^2 = t0 (16/9) a b / 16
            float est = 0.0f;
            if (valid) {
                const uint32_t leaf = qpos / (uint32_t)LEAF;
                const float4 lo = bv.leaf_lo[leaf], hi = bv.leaf_hi[leaf];
                const float ex = hi.x - lo.x, ey = hi.y - lo.y, ez = hi.z - lo.z;
                const float a = fmaxf(fmaxf(ex, ey), ez), c = fminf(fminf(ex, ey), ez);
                const float b = (ex + ey + ez) - a - c;
                est = (float)prm.t0 * a * b * (1.0f / (9.0f * 3.14159265f)) * prm.r2_margin;
            }
            const float h = uniformf(wave_max(est));
            if (h > 0.0f && h < 1e37f) r2 = h;
        }
        if (tt == 1 && prm.tiles_per_wave > 2) {
            const auto t0 = ws.tick();
            const int64_t q0 = qbase, q1 = min((run + 1) * prm.tiles_per_wave * Q, (int64_t)nq);
            const uint32_t p0 = qv.qsel ? qv.qsel[q0] : (uint32_t)q0;
            run_scan = bv.scan_of ? uniform((int)bv.scan_of[p0]) : 0;
            float lx = INFINITY, ly = INFINITY, lz = INFINITY, hx = -INFINITY, hy = -INFINITY, hz = -INFINITY;
            for (int64_t i = q0 + lane; i < q1; i += kWave) {
                const uint32_t pi = qv.qsel ? qv.qsel[i] : (uint32_t)i;
                if (bv.scan_of && (int)bv.scan_of[pi] != run_scan) continue;
                const float x = qv.qx[pi], y = qv.qy[pi], z = qv.qz[pi];
                lx = fminf(lx, x);
                ly = fminf(ly, y);
                lz = fminf(lz, z);
                hx = fmaxf(hx, x);
                hy = fmaxf(hy, y);
                hz = fmaxf(hz, z);
            }
            rbox.lx = uniformf(wave_min(lx));
            rbox.ly = uniformf(wave_min(ly));
            rbox.lz = uniformf(wave_min(lz));
            rbox.hx = uniformf(wave_max(hx));
            rbox.hy = uniformf(wave_max(hy));
            rbox.hz = uniformf(wave_max(hz));
            run_r2 = fmaxf(r2, 1e-30f) * RUN_R2_MARGIN;
            int sp = start_traversal(sm, bv, run_scan, rbox, run_r2);
            const int nl = fill_leaves(sm, bv, rbox, run_r2, sp, TL_CAP, ws);
            run_ok = sp == 0;  // a list that fits
            if (run_ok) {
                for (int j = lane; j < nl; j += kWave) sm.stack[j] = sm.tleaf[j];
                nrun = nl;
            }
            __syncthreads();
            ws.tock(KS_CY_TRAV, t0);
        }
        sink.begin(g, s, qi, qpos, qx, qy, qz);
        if (s == 0) {
#pragma unroll
            for (int tau = 0; tau < NT; tau++) sm.scnt[tau][g] = 0u;
        }
        float kth_max = 0.0f;
        // radius^2 bracket of the tile: r2_floor left a query of the tile short, r2_ceil was
        // shrunk from (too many leaves); growth stays below r2_ceil, shrinking above r2_floor
        float r2_floor = 0.0f, r2_ceil = INFINITY;
        int shrinks = 0;
        int prev_lead = -1;
        int cl_att = 0;  // attempts of the current cluster

        for (int attempt = 0; attempt < MAX_ATTEMPTS; attempt++) {
            const uint64_t pending = ballot(active);
            if (pending == 0) break;
            // One scan per attempt (the scan of the lowest pending lane): neighbourhoods never
            // cross scans, each scan's points form one subtree of the hierarchy.  And only the
            // pending queries within R (per axis) of that lane's query: a tile of scattered
            // queries (a sparse query list) would otherwise span a box far larger than R.
            const int lead = __ffsll((long long)pending) - 1;
            if (lead != prev_lead) {  // a new cluster: its own bracket
                r2_floor = 0.0f;
                r2_ceil = INFINITY;
                shrinks = 0;
                cl_att = 0;
                prev_lead = lead;
            }
            cl_att++;
            const int cur = bv.scan_of ? __builtin_amdgcn_readlane(scan, lead) : 0;
            const float lx = readlane_f(qx, lead), ly = readlane_f(qy, lead), lz = readlane_f(qz, lead);
            const float reach = sqrtf(r2);
            const bool near = fabsf(qx - lx) <= reach && fabsf(qy - ly) <= reach && fabsf(qz - lz) <= reach;
            const bool act = active && scan == cur && near;
            // tile box of the pending queries of that scan
            TileBox tb;
            tb.lx = uniformf(wave_min(act ? qx : INFINITY));
            tb.ly = uniformf(wave_min(act ? qy : INFINITY));
            tb.lz = uniformf(wave_min(act ? qz : INFINITY));
            tb.hx = uniformf(wave_max(act ? qx : -INFINITY));
            tb.hy = uniformf(wave_max(act ? qy : -INFINITY));
            tb.hz = uniformf(wave_max(act ? qz : -INFINITY));
            const float R2 = r2;
            ws.c[KS_ATTEMPTS] += 1;

            // --- candidate leaves of the tile -------------------------------------------------
            Cands cs;
            float rej = INFINITY;  // nearest box the traversal left out
            {
                const auto t0 = ws.tick();
                int sp = 0, nl = 0;
                if (run_ok && cur == run_scan && R2 <= run_r2 && tb.lx >= rbox.lx && tb.ly >= rbox.ly &&
                    tb.lz >= rbox.lz && tb.hx <= rbox.hx && tb.hy <= rbox.hy && tb.hz <= rbox.hz) {
                    nl = filter_leaves(sm, bv, nrun, tb, R2, &rej);
                } else {
                    run_ok = false;  // the traversal takes the stack
                    sp = start_traversal(sm, bv, cur, tb, R2, &rej);
                    nl = fill_leaves(sm, bv, tb, R2, sp, TL_CAP, ws, &rej);
                }
                if (sp < 0) {
                    if (lane == 0) atomicMax(prm.err, 5);
                    return;
                }
                const bool overflow = sp > 0;  // the list is full and the traversal unfinished
                // A radius far larger than the targets need (carried from a sparser tile, the
                // first guess of a run in a dense region, or a growth step that overshot for a
                // query off a dense surface) would put tens of thousands of points per query
                // through the passes: with far more leaves in range than the tile's targets can
                // use (the queries' t0 points each, their spread over the tile box widening the
                // union), shrink it by the leaf count (counts grow like R^2 - r2_floor) before
                // any point is scanned.  At most four times per tile.
                if (shrinks < 4 && R2 > r2_floor && cl_att <= CLUSTER_SHRINK) {
                    const float ext = fmaxf(fmaxf(tb.hx - tb.lx, tb.hy - tb.ly), tb.hz - tb.lz);
                    const float spread = 1.0f + ext / sqrtf(R2);
                    const float want = ((float)prm.t0 / (float)LEAF + 1.0f) * spread * spread + (float)Q;
                    const int big = (int)fminf(8.0f * want, 65536.0f);
                    int total = nl;
                    if (overflow) total = fill_leaves(sm, bv, tb, R2, sp, 4 * big, ws, nullptr, false) + nl;
                    if (total > big) {
                        // geometric steps: for a query off a surface the count grows faster
                        r2 = r2_floor + (R2 - r2_floor) * fmaxf(sqrtf(2.0f * want / (float)total), 0.125f);
                        r2_ceil = R2;
                        shrinks++;
                        ws.tock(KS_CY_TRAV, t0);
                        __syncthreads();
                        continue;
                    }
                }
                ws.tock(KS_CY_TRAV, t0);
                cs.ntl = overflow ? -1 : nl;  // the passes then traverse again, chunk by chunk
                cs.resident = false;
                cs.staged = -1;
                cs.reverse = false;
            }

            // --- pass 1: value histogram over [0, R2] ---------------------------------------
            // 16-bit bins over a leaf list, whole words when streaming.  One-target sinks only: their
            // large k puts more keys in a bin (the finer bins spare refinement passes); with a ki
            // target the wider bin search cost more than it saved.
            const bool narrow = NT == 1 && cs.ntl >= 0;
            const uint32_t hsh = narrow ? 1u : 0u;
            const int nb = HIST_WORDS << hsh;
            const int nbb = HIST_BITS + (int)hsh;
            const float inv_w = (float)nb / R2;
            auto locate = [&](int t, int& b, int& below, int& cnt) {
                if constexpr (NT == 1) {
                    if (narrow) return find_bin<P, Q, 2 * HIST_WORDS>(sm.hist, g, s, t, b, below, cnt);
                }
                return find_bin<P, Q, HIST_WORDS>(sm.hist, g, s, t, b, below, cnt);
            };
            for (int b = s; b < HIST_WORDS; b += P) sm.hist[g][b] = 0u;
            __syncthreads();
            uint32_t* myhist = &sm.hist[g][0];
            float dmin = INFINITY;  // nearest in-range candidate of this lane
            bool fine = for_each_chunk(sm, bv, tb, R2, cur, cs, ws, [&](int nl, int base) {
                auto t0 = ws.tick();
                const int len = build_glist<P>(sm, bv, base, nl, g, s, qx, qy, qz, act ? R2 : -1.0f);
                ws.tock(KS_CY_GLIST, t0);
                t0 = ws.tick();
                scan_glist<P>(sm, g, s, len, [&](const float4& p, int) {
                    const float d2 = d2_contract(qx, qy, qz, p.x, p.y, p.z);
                    if (d2 <= R2) {
                        hist_add(myhist, vbin(d2, inv_w, nb), hsh);
                        dmin = fminf(dmin, d2);
                    }
                });
                ws.tock(KS_CY_HIST, t0);
                if (prm.stats) ws.c[KS_CANDIDATES] += (unsigned)(LEAF * wave_sum_int(s == 0 ? len : 0));
            });
            if (!fine) {
                if (lane == 0) atomicMax(prm.err, 5);
                return;
            }
            if (cs.resident) ws.c[KS_RESIDENT] += 1;

            // --- targets -----------------------------------------------------------------------
            const auto t_sel = ws.tick();
            Target tg[NT];
            bool ok = false;
            int in_range = 0;
            {
                int b, below, cnt;
                const int total = locate(prm.t0, b, below, cnt);
                in_range = total;
                ok = act && total >= prm.t0;
#pragma unroll
                for (int tau = 0; tau < NT; tau++) {
                    const int t = tau == 0 ? prm.t0 : tau == 1 ? prm.t1 : tau == 2 ? prm.t2 : prm.t3;
                    tg[tau].t = t;
                    tg[tau].refine = false;
                    tg[tau].below = 0;
                    tg[tau].cnt = 0;
                    tg[tau].klo = ~0ull;
                    tg[tau].khi = ~0ull;
                    if (tau >= 1) {
                        if (t > 0) locate(t, b, below, cnt);
                    }
                    if (ok && t > 0) {
                        const float lo = b == 0 ? 0.0f : vbin_floor(b, inv_w, nb);
                        tg[tau].klo = (uint64_t)fbits(lo) << 32;
                        if (b == nb - 1 || !(inv_w > 0.0f)) {
                            tg[tau].khi = ((uint64_t)fbits(R2) << 32) | 0xffffffffull;
                        } else {
                            const float hi = vbin_floor(b + 1, inv_w, nb);
                            tg[tau].khi = ((uint64_t)fbits(hi) << 32) - 1ull;
                        }
                        tg[tau].below = below;
                        tg[tau].cnt = cnt;
                        tg[tau].refine = cnt > BUF;
                    }
                }
            }

            // --- refinement passes (radix over the 64-bit key), one target at a time -----------
#pragma unroll
            for (int tau = 0; tau < NT; tau++) {
                for (int guard = 0; guard < 16; guard++) {
                    const bool need = ok && tg[tau].refine;
                    if (ballot(need) == 0) break;
                    ws.c[KS_REFINE] += 1;
                    const uint64_t span = tg[tau].khi - tg[tau].klo;
                    const int nbits = span ? 64 - __clzll((long long)span) : 0;
                    const int shift = nbits > nbb ? nbits - nbb : 0;
                    const float rq = need ? bitsf((uint32_t)(tg[tau].khi >> 32)) : -1.0f;
                    const float rin = need ? bitsf((uint32_t)(tg[tau].klo >> 32)) : -1.0f;
                    for (int b = s; b < HIST_WORDS; b += P) sm.hist[g][b] = 0u;
                    __syncthreads();
                    // a streamed pass walks the tree only as far as the widest refined bin reaches
                    const float rq_all = fminf(uniformf(wave_max(rq)), R2);
                    fine = for_each_chunk(sm, bv, tb, rq_all, cur, cs, ws, [&](int nl, int base) {
                        const int len = build_glist<P>(sm, bv, base, nl, g, s, qx, qy, qz, rq, rin);
                        scan_glist<P>(sm, g, s, len, [&](const float4& p, int) {
                            const float d2 = d2_contract(qx, qy, qz, p.x, p.y, p.z);
                            // outside the bin's d^2 range (rq <= R^2; -1 when not refining)
                            if (!(d2 <= rq && d2 >= rin)) return;
                            const uint64_t key = make_key(d2, __float_as_uint(p.w));
                            if (key >= tg[tau].klo && key <= tg[tau].khi)
                                hist_add(myhist, (int)((key - tg[tau].klo) >> shift), hsh);
                        });
                    });
                    if (!fine) {
                        if (lane == 0) atomicMax(prm.err, 5);
                        return;
                    }
                    int b, below, cnt;
                    locate(tg[tau].t - tg[tau].below, b, below, cnt);
                    if (need) {
                        const uint64_t nlo = tg[tau].klo + ((uint64_t)b << shift);
                        uint64_t nhi = nlo + ((shift >= 64) ? ~0ull : ((1ull << shift) - 1ull));
                        if (nhi > tg[tau].khi || nhi < nlo) nhi = tg[tau].khi;
                        tg[tau].klo = nlo;
                        tg[tau].khi = nhi;
                        tg[tau].below += below;
                        tg[tau].cnt = cnt;
                        tg[tau].refine = cnt > BUF;
                    }
                }
            }
            float rc = -1.0f;  // collection bound: the largest boundary bin edge
#pragma unroll
            for (int tau = 0; tau < NT; tau++) {
                tg[tau].dlo = bitsf((uint32_t)(tg[tau].klo >> 32));
                tg[tau].dhi = bitsf((uint32_t)(tg[tau].khi >> 32));
                if (ok && tg[tau].t > 0) rc = fmaxf(rc, tg[tau].dhi);
            }

            ws.tock(KS_CY_SELECT, t_sel);

            // --- collection pass, bounded by the last bin's upper edge ---------------------------------
            if (s == 0) {
#pragma unroll
                for (int tau = 0; tau < NT; tau++) sm.bcnt[tau][g] = 0;
            }
            __syncthreads();
            // tau 0 takes the leaves that lie wholly below its boundary bin whole, when no other
            // target needs their points one by one
            constexpr bool kLeaves = leaf_accepting<Sink>::value;
            const bool whole = kLeaves && (NT < 2 || prm.t1 <= 0) && (NT < 3 || prm.t2 <= 0) && (NT < 4 || prm.t3 <= 0);
            const float rfull = (whole && ok) ? tg[0].dlo : -1.0f;
            auto collect = [&](int nl, int base) {
                auto t0 = ws.tick();
                int nfull = 0;
                const int len = whole ? build_glist_split<P>(sm, nl, g, s, qx, qy, qz, rc, rfull, nfull)
                                      : build_glist<P>(sm, bv, base, nl, g, s, qx, qy, qz, rc);
                ws.tock(KS_CY_GLIST, t0);
                t0 = ws.tick();
                if constexpr (kLeaves) {
                    for (int j = s; j < nfull; j += P) {
                        const int js = sm.flist[g][j];
                        const int leaf = sm.tleaf[base + js];
                        const float4 lo = sm.lbox[js][0], hi = sm.lbox[js][1];
                        const float4* a = bv.leaf_agg + 3 * (size_t)leaf;
                        sink.accept_leaf(g, s, a[0], a[1], a[2], lo, hi);
                    }
                }
                scan_glist<P>(sm, g, s, len, [&](const float4& p, int c) {
                    const float d2 = d2_contract(qx, qy, qz, p.x, p.y, p.z);
                    if (!(d2 <= rc)) return;  // beyond every target's boundary bin (rc <= R2)
#pragma unroll
                    for (int tau = 0; tau < NT; tau++) {
                        if (tg[tau].t <= 0) continue;
                        if (d2 < tg[tau].dlo) {
                            sink.accept(tau, g, s, p.x, p.y, p.z, make_key(d2, __float_as_uint(p.w)),
                                        slot_spos(sm, base, c), sm.scnt[tau]);
                        } else if (d2 <= tg[tau].dhi) {
                            const uint64_t key = make_key(d2, __float_as_uint(p.w));
                            if (key < tg[tau].klo) {
                                sink.accept(tau, g, s, p.x, p.y, p.z, key, slot_spos(sm, base, c), sm.scnt[tau]);
                            } else if (key <= tg[tau].khi) {
                                const uint32_t slot = atomicAdd(&sm.bcnt[tau][g], 1u);
                                if (slot < (uint32_t)BUF) sm.bkey[tau][g][slot] = key;
                            }
                        }
                    }
                });
                ws.tock(KS_CY_COLLECT, t0);
            };
            fine = for_each_chunk(sm, bv, tb, fminf(uniformf(wave_max(rc)), R2), cur, cs, ws, collect);
            if (!fine) {
                if (lane == 0) atomicMax(prm.err, 5);
                return;
            }

            // --- exact ranking inside the boundary bins --------------------------------------------
            const auto t_fin = ws.tick();
            float kth = -1.0f;  // d^2 of the t0-th key, on the lane that ranked it
            if (ok) {
#pragma unroll
                for (int tau = 0; tau < NT; tau++) {
                    if (tg[tau].t <= 0) continue;
                    const int nb = min((int)sm.bcnt[tau][g], BUF);
                    const int need = tg[tau].t - tg[tau].below;
                    for (int j = s; j < nb; j += P) {
                        const uint64_t kj = sm.bkey[tau][g][j];
                        int rank = 0;
                        for (int i = 0; i < nb; i++) rank += sm.bkey[tau][g][i] < kj ? 1 : 0;
                        if (rank < need) {
                            const uint32_t sp_ = bv.rank[(uint32_t)kj];
                            sink.accept(tau, g, s, bv.px[sp_], bv.py[sp_], bv.pz[sp_], kj, sp_, sm.scnt[tau]);
                        }
                        if (tau == 0 && rank == need - 1) kth = bitsf((uint32_t)(kj >> 32));
                    }
                }
            }
            if (prm.hint_out && kth >= 0.0f) prm.hint_out[qpos] = kth;
            __syncthreads();
            sink.finish(g, s, ok, qpos, qx, qy, qz, sm.scnt[0], sm.scnt[NT > 1 ? 1 : 0]);
            ws.tock(KS_CY_FINISH, t_fin);
            // the query's nearest in-range d^2 (dmin over its lanes)
            float dm = dmin;
            dm = lane_group_reduce<P>(dm, [](float a, float b) { return fminf(a, b); });
            // Next tile's radius: per resolved query dmin^2 + (kth^2 - dmin^2) * margin, the
            // margin applied to the part of the radius that grows with the count (for a query
            // off a surface, R^2 - dmin^2 is what the covered disc grows with); the tile's largest.
            const float est = (ok && kth >= 0.0f) ? dm + (kth - dm) * prm.r2_margin : 0.0f;
            kth_max = fmaxf(kth_max, uniformf(wave_max(est)));
            const float disc_ok = uniformf(wave_max((ok && kth >= 0.0f) ? kth - dm : 0.0f));
            if (disc_ok > 0.0f) disc_hint = disc_ok;
            ws.c[KS_RESOLVED] += (unsigned)__popcll(ballot(ok && s == 0));
            active = active && !ok;
            if (ballot(act && !ok) != 0) {  // queries of this scan still short of t: grow R
                // Each short query estimates the radius^2 that holds t0 points from what it saw:
                // the count in range grows like R^2 - dmin^2 (a surface seen from a distance
                // dmin), and the smallest growth applies to that disc term; with nothing in
                // range, the nearest box the traversal left out plus the disc the last resolved
                // queries needed (else x1.1, else x4).  Growing R^2 itself by a fixed factor
                // overshoots badly for a query far from a dense surface (the disc it then
                // covers holds thousands of points).
                float need = 0.0f;
                if (act && !ok) {
                    if (in_range > 0) {
                        const float disc = R2 - dm;
                        need = dm + disc * fmaxf((float)prm.t0 / (float)in_range * 1.25f, fmaxf(prm.grow_min, 1.5f));
                    } else if (rej < INFINITY) {
                        need = disc_hint > 0.0f ? rej + 1.5f * disc_hint : rej * 1.1f;
                    } else {
                        need = R2 * 4.0f;
                    }
                }
                float nr2 = uniformf(wave_max(need));
                nr2 = fminf(fmaxf(nr2, R2 * 1.01f), R2 * 16.0f);
                if (cl_att > CLUSTER_GUARD) nr2 = R2 * 4.0f;  // a guard: plain doubling of R
                if (!(nr2 > R2) || nr2 < 1e-30f) nr2 = fmaxf(R2 * 4.0f, 1e-30f);
                if (nr2 > 1e37f) nr2 = INFINITY;
                // below a radius already found too wide, unless shrinking is over
                if (shrinks < 4 && nr2 > r2_ceil && cl_att <= CLUSTER_SHRINK)
                    nr2 = fmaxf(R2 + (r2_ceil - R2) * 0.5f, R2 * 1.01f);
                r2_floor = R2;
                r2 = nr2;
            }
            __syncthreads();
        }
        if (ballot(active) != 0) {
            if (lane == 0) atomicMax(prm.err, 6);
        }
        // radius guess for the next tile of this run
        if (kth_max > 0.0f) r2 = kth_max;
        r2 = fmaxf(r2, 1e-30f);
    }
    }
    ws.tock(KS_CY_TOTAL, t_wave);
    ws.flush(prm.stats);
}

}  // namespace lvrhip
