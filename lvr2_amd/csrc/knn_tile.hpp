// Exact kNN for tiles of Morton-ordered queries, one wavefront per tile.
//
// Replaces the per-thread register-queue kNN of the reference (lbvh/query.cuh:45-167,
// static_priorityqueue.cuh:18-93, query_knn_kernels.cu:9-48) and the CPU kd-tree search
// it must agree with (KDTree.tcc:99-126, 186-248).
//
// A wave owns Q = 64/P consecutive queries, P lanes per query.  For a radius R it
//   1. walks the LBVH once for the whole tile (wave-parallel, LDS stack), collecting the
//      leaves whose box is within R of the tile's box;
//   2. stages the candidate points in LDS (chunks of CAP points; when the candidates fit in
//      one chunk they stay resident for all passes);
//   3. per query, counts candidates with d^2 <= R^2 into 32 bins linear in d^2 (one LDS
//      column per lane).  If fewer than t candidates are in range the query keeps R
//      pending and the tile retries with 2R; otherwise the bin holding the t-th smallest
//      key is found, refined by radix passes over the 64-bit key (d^2 bits, point index)
//      until it holds at most BUF keys, and a collection pass hands every key below that
//      bin to the sink, buffers the bin, and ranks the buffer exactly.
// Every point within R of a query is a candidate (the box distance is a monotone lower
// bound of the fp32 d^2), so a query with >= t in-range candidates gets the exact t
// smallest keys.  Two targets are selected at once (e.g. k for the plane fit and ki for
// the interpolation list).
#pragma once

#include "common.hpp"
#include "lbvh.hpp"

namespace lvrhip {

constexpr int NB = 32;            // histogram bins
constexpr int CAP = 512;          // candidate slots per chunk
constexpr int CHUNK_LEAVES = CAP / LEAF;
constexpr int STACK_CAP = 512;
constexpr int MAX_ATTEMPTS = 80;

struct BvhView {
    const float* px;
    const float* py;
    const float* pz;
    const uint32_t* pidx;  // original index of each sorted point
    const float4* leaf_lo;
    const float4* leaf_hi;
    const BvhNode* nodes;
    int nleaf;
    uint32_t n;
};

struct QueryView {
    const float* qx;  // query coordinates indexed by query position
    const float* qy;
    const float* qz;
    const uint32_t* qsel;  // optional: tile slot -> query position
    uint32_t nq;           // number of tile slots
};

struct TileParams {
    int t0, t1;  // targets (t1 <= t0; t1 = 0 for none), already clamped to n
    int tiles_per_wave;
    float r2_init;
    float r2_margin;  // next tile's radius^2 = margin * largest k-th d^2 of this tile
    int* err;  // set to a nonzero code on capacity overflow
};

template <int Q, int BUF>
struct TileSmem {
    float4 cand[CAP];  // x, y, z, original index (as bits)
    uint32_t cspos[CAP];
    uint32_t hist[NB][kWave];
    int stack[STACK_CAP];
    int chunk_leaf[CHUNK_LEAVES];
    uint64_t bkey[2][Q][BUF];
    uint32_t bspos[2][Q][BUF];
    uint32_t bcnt[2][Q];
    uint32_t scnt[2][Q];  // per-query sink slot counters
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

// Fills sm.chunk_leaf with up to CHUNK_LEAVES leaves whose box is within sqrt(r2) of the
// tile box, continuing the traversal held on the LDS stack.  Returns the number of
// leaves (wave-uniform).  sp is the stack size (wave-uniform), -1 on overflow.
template <class Sm>
__device__ int fill_chunk(Sm& sm, const BvhView& bv, const TileBox& tb, float r2, int& sp)
{
    const int lane = lane_id();
    int nl = 0;
    while (sp > 0) {
        const int room = CHUNK_LEAVES - nl;
        if (room == 0) break;
        int w = min(min(kWave, sp), STACK_CAP - sp);
        if (w <= 0) {
            sp = -1;
            return nl;
        }
        int ref = lane < w ? sm.stack[sp - 1 - lane] : 0x7fffffff;
        sp -= w;
        __syncthreads();
        // outputs: up to two per lane
        int o0 = 0x7fffffff, o1 = 0x7fffffff;
        if (ref != 0x7fffffff) {
            if (ref < 0) {
                if (-1 - ref < bv.nleaf) o0 = ref;  // a leaf pushed back: already tested
            } else if (ref < bv.nleaf - 1) {
                const BvhNode nd = bv.nodes[ref];
                float d0 = box_d2(nd.a.x, nd.a.y, nd.a.z, nd.a.w, nd.b.x, nd.b.y, tb.lx, tb.ly, tb.lz, tb.hx, tb.hy,
                                  tb.hz);
                float d1 = box_d2(nd.b.z, nd.b.w, nd.c.x, nd.c.y, nd.c.z, nd.c.w, tb.lx, tb.ly, tb.lz, tb.hx, tb.hy,
                                  tb.hz);
                if (d0 <= r2) o0 = nd.r.x;
                if (d1 <= r2) o1 = nd.r.y;
            }
        }
        const bool l0 = o0 < 0, l1 = o1 < 0;  // leaf outputs
        const uint64_t bl0 = ballot(l0), bl1 = ballot(l1);
        const int nleaves = __popcll(bl0) + __popcll(bl1);
        const int take = min(nleaves, room);
        // leaf order: all first outputs, then all second outputs
        const int r0 = mbcnt(bl0), r1 = __popcll(bl0) + mbcnt(bl1);
        bool push0 = (o0 != 0x7fffffff) && (!l0 || r0 >= take);
        bool push1 = (o1 != 0x7fffffff) && (!l1 || r1 >= take);
        if (l0 && r0 < take) sm.chunk_leaf[nl + r0] = -1 - o0;
        if (l1 && r1 < take) sm.chunk_leaf[nl + r1] = -1 - o1;
        const uint64_t bp0 = ballot(push0), bp1 = ballot(push1);
        if (push0) sm.stack[sp + mbcnt(bp0)] = o0;
        if (push1) sm.stack[sp + __popcll(bp0) + mbcnt(bp1)] = o1;
        sp += __popcll(bp0) + __popcll(bp1);
        nl += take;
        __syncthreads();
    }
    return nl;
}

// Stage the points of the chunk's leaves in LDS; slot = leaf_slot * LEAF + offset.
// Slots past the end of the cloud get NaN coordinates (never in range).
template <class Sm>
__device__ void load_chunk(Sm& sm, const BvhView& bv, int nl)
{
    const int lane = lane_id();
    for (int j = 0; j < nl; j += 2) {
        const int slot_leaf = j + (lane >> 5);
        if (slot_leaf < nl) {
            const int leaf = sm.chunk_leaf[slot_leaf];
            const uint32_t spos = (uint32_t)leaf * LEAF + (lane & 31);
            const int slot = slot_leaf * LEAF + (lane & 31);
            if (spos < bv.n) {
                sm.cand[slot] = make_float4(bv.px[spos], bv.py[spos], bv.pz[spos], __uint_as_float(bv.pidx[spos]));
                sm.cspos[slot] = spos;
            } else {
                sm.cand[slot] = make_float4(__builtin_nanf(""), __builtin_nanf(""), __builtin_nanf(""), 0.0f);
                sm.cspos[slot] = 0;
            }
        }
    }
    __syncthreads();
}

// ---- histogram helpers ----------------------------------------------------------------------

// Locate the bin holding the t-th smallest counted key of query group g (P lanes).
// Returns total count; if t <= total: bin, below (count in lower bins), cnt (in that bin).
template <int P>
__device__ __forceinline__ int find_bin(const uint32_t (*hist)[kWave], int g, int s, int t, int& bin, int& below,
                                        int& cnt)
{
    constexpr int BPL = NB / P;
    static_assert(NB % P == 0, "bins per lane");
    int sum = 0;
#pragma unroll
    for (int j = 0; j < BPL; j++) {
        const int b = s * BPL + j;
#pragma unroll
        for (int c = 0; c < P; c++) sum += (int)hist[b][g * P + c];
    }
    int incl = sum;
#pragma unroll
    for (int o = 1; o < P; o <<= 1) {
        int y = shfl(incl, lane_id() - o);
        if (s >= o) incl += y;
    }
    const int total = shfl(incl, g * P + P - 1);
    const int excl = incl - sum;
    bool mine = (t >= 1) && (excl < t) && (t <= incl);
    int b_ = 0, below_ = 0, cnt_ = 0;
    if (mine) {
        int run = excl;
#pragma unroll
        for (int j = 0; j < BPL; j++) {
            const int b = s * BPL + j;
            int v = 0;
#pragma unroll
            for (int c = 0; c < P; c++) v += (int)hist[b][g * P + c];
            if (run < t && t <= run + v) {
                b_ = b;
                below_ = run;
                cnt_ = v;
            }
            run += v;
        }
    }
    const uint64_t owners = ballot(mine);
    const uint64_t gm = (owners >> (g * P)) & ((P == 64) ? ~0ull : ((1ull << P) - 1));
    const int owner = g * P + (gm ? __ffsll((long long)gm) - 1 : 0);
    bin = shfl(b_, owner);
    below = shfl(below_, owner);
    cnt = shfl(cnt_, owner);
    return total;
}

__device__ __forceinline__ int vbin(float d2, float inv_w)
{
    float f = d2 * inv_w;
    int b = (int)f;
    return b < NB - 1 ? b : NB - 1;
}

// Smallest non-negative float x with vbin(x, inv_w) >= b (b >= 1).
__device__ float vbin_floor(int b, float inv_w)
{
    float x = (float)b / inv_w;
    if (!(x >= 0.0f) || isinf(x)) x = 0.0f;
    for (int it = 0; it < 256 && x > 0.0f && vbin(bitsf(fbits(x) - 1u), inv_w) >= b; it++) x = bitsf(fbits(x) - 1u);
    for (int it = 0; it < 256 && vbin(x, inv_w) < b; it++) x = bitsf(fbits(x) + 1u);
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

template <int P, class Sink>
__global__ __launch_bounds__(kWave) void knn_tile_kernel(BvhView bv, QueryView qv, TileParams prm, Sink sink)
{
    constexpr int Q = kWave / P;
    constexpr int BUF = (Q >= 64) ? 8 : (Q >= 32 ? 12 : 16);
    __shared__ TileSmem<Q, BUF> sm;
    const int lane = lane_id();
    const int g = lane / P, s = lane % P;
    const int root = bv.nleaf > 1 ? 0 : -1;  // internal node 0, or the single leaf
    float r2 = fmaxf(prm.r2_init, 1e-30f);

    for (int tt = 0; tt < prm.tiles_per_wave; tt++) {
        const int64_t tile = (int64_t)blockIdx.x * prm.tiles_per_wave + tt;
        const int64_t qbase = tile * Q;
        if (qbase >= qv.nq) break;
        const int64_t qi = qbase + g;
        const bool valid = qi < qv.nq;
        const uint32_t qpos = valid ? (qv.qsel ? qv.qsel[qi] : (uint32_t)qi) : 0u;
        const float qx = valid ? qv.qx[qpos] : 0.0f;
        const float qy = valid ? qv.qy[qpos] : 0.0f;
        const float qz = valid ? qv.qz[qpos] : 0.0f;
        bool active = valid && prm.t0 > 0;
        sink.begin(g, s, qi, qpos, qx, qy, qz);
        if (s == 0) {
            sm.scnt[0][g] = 0u;
            sm.scnt[1][g] = 0u;
        }
        float kth_max = 0.0f;

        for (int attempt = 0; attempt < MAX_ATTEMPTS; attempt++) {
            if (ballot(active) == 0) break;
            // tile box of the pending queries
            TileBox tb;
            tb.lx = uniformf(wave_min(active ? qx : INFINITY));
            tb.ly = uniformf(wave_min(active ? qy : INFINITY));
            tb.lz = uniformf(wave_min(active ? qz : INFINITY));
            tb.hx = uniformf(wave_max(active ? qx : -INFINITY));
            tb.hy = uniformf(wave_max(active ? qy : -INFINITY));
            tb.hz = uniformf(wave_max(active ? qz : -INFINITY));
            const float R2 = r2;
            const float inv_w = (float)NB / R2;

            // --- pass 1: value histogram over [0, R2] ---------------------------------------
            int sp = 0, nl = 0;
            bool resident = false;
#pragma unroll
            for (int b = 0; b < NB; b++) sm.hist[b][lane] = 0u;
            if (lane == 0) sm.stack[0] = root >= 0 ? 0 : -1;  // leaf 0 -> ref -1
            sp = 1;
            __syncthreads();
            bool first_chunk = true;
            for (;;) {
                nl = fill_chunk(sm, bv, tb, R2, sp);
                if (sp < 0) break;
                if (first_chunk && sp == 0) resident = true;
                first_chunk = false;
                load_chunk(sm, bv, nl);
                if (active) {
                    const int C = nl * LEAF;
                    for (int c = s; c < C; c += P) {
                        const float4 p = sm.cand[c];
                        const float d2 = d2_contract(qx, qy, qz, p.x, p.y, p.z);
                        if (d2 <= R2) atomicAdd(&sm.hist[vbin(d2, inv_w)][lane], 1u);
                    }
                }
                __syncthreads();
                if (sp == 0) break;
            }
            if (sp < 0) {
                if (lane == 0) atomicMax(prm.err, 5);
                return;
            }

            // --- targets -----------------------------------------------------------------------
            Target tg[2];
            bool ok = false;
            {
                int b, below, cnt;
                const int total = find_bin<P>(sm.hist, g, s, prm.t0, b, below, cnt);
                ok = active && total >= prm.t0;
#pragma unroll
                for (int tau = 0; tau < 2; tau++) {
                    const int t = tau == 0 ? prm.t0 : prm.t1;
                    tg[tau].t = t;
                    tg[tau].refine = false;
                    tg[tau].below = 0;
                    tg[tau].cnt = 0;
                    tg[tau].klo = ~0ull;
                    tg[tau].khi = ~0ull;
                    if (tau == 1) {
                        if (t > 0) find_bin<P>(sm.hist, g, s, t, b, below, cnt);
                    }
                    if (ok && t > 0) {
                        const float lo = b == 0 ? 0.0f : vbin_floor(b, inv_w);
                        tg[tau].klo = (uint64_t)fbits(lo) << 32;
                        if (b == NB - 1 || !(inv_w > 0.0f)) {
                            tg[tau].khi = ((uint64_t)fbits(R2) << 32) | 0xffffffffull;
                        } else {
                            const float hi = vbin_floor(b + 1, inv_w);
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
            for (int tau = 0; tau < 2; tau++) {
                for (int guard = 0; guard < 16; guard++) {
                    const bool need = ok && tg[tau].refine;
                    if (ballot(need) == 0) break;
                    const uint64_t span = tg[tau].khi - tg[tau].klo;
                    const int nbits = span ? 64 - __clzll((long long)span) : 0;
                    const int shift = nbits > 5 ? nbits - 5 : 0;
#pragma unroll
                    for (int b = 0; b < NB; b++) sm.hist[b][lane] = 0u;
                    __syncthreads();
                    auto count = [&](int C) {
                        if (!need) return;
                        for (int c = s; c < C; c += P) {
                            const float4 p = sm.cand[c];
                            const float d2 = d2_contract(qx, qy, qz, p.x, p.y, p.z);
                            if (!(d2 <= R2)) continue;
                            const uint64_t key = make_key(d2, __float_as_uint(p.w));
                            if (key >= tg[tau].klo && key <= tg[tau].khi)
                                atomicAdd(&sm.hist[(int)((key - tg[tau].klo) >> shift)][lane], 1u);
                        }
                    };
                    if (resident) {
                        count(nl * LEAF);
                        __syncthreads();
                    } else {
                        if (lane == 0) sm.stack[0] = root >= 0 ? 0 : -1;
                        sp = 1;
                        __syncthreads();
                        for (;;) {
                            nl = fill_chunk(sm, bv, tb, R2, sp);
                            load_chunk(sm, bv, nl);
                            count(nl * LEAF);
                            __syncthreads();
                            if (sp <= 0) break;
                        }
                    }
                    int b, below, cnt;
                    find_bin<P>(sm.hist, g, s, tg[tau].t - tg[tau].below, b, below, cnt);
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
#pragma unroll
            for (int tau = 0; tau < 2; tau++) {
                tg[tau].dlo = bitsf((uint32_t)(tg[tau].klo >> 32));
                tg[tau].dhi = bitsf((uint32_t)(tg[tau].khi >> 32));
            }

            // --- collection pass -------------------------------------------------------------------
            if (s == 0) {
                sm.bcnt[0][g] = 0;
                sm.bcnt[1][g] = 0;
            }
            __syncthreads();
            auto collect = [&](int C) {
                if (!ok) return;
                for (int c = s; c < C; c += P) {
                    const float4 p = sm.cand[c];
                    const float d2 = d2_contract(qx, qy, qz, p.x, p.y, p.z);
                    if (!(d2 <= R2)) continue;
#pragma unroll
                    for (int tau = 0; tau < 2; tau++) {
                        if (tg[tau].t <= 0) continue;
                        if (d2 < tg[tau].dlo) {
                            sink.accept(tau, g, s, p.x, p.y, p.z, make_key(d2, __float_as_uint(p.w)), sm.cspos[c],
                                        sm.scnt[tau]);
                        } else if (d2 <= tg[tau].dhi) {
                            const uint64_t key = make_key(d2, __float_as_uint(p.w));
                            if (key < tg[tau].klo) {
                                sink.accept(tau, g, s, p.x, p.y, p.z, key, sm.cspos[c], sm.scnt[tau]);
                            } else if (key <= tg[tau].khi) {
                                const uint32_t slot = atomicAdd(&sm.bcnt[tau][g], 1u);
                                if (slot < (uint32_t)BUF) {
                                    sm.bkey[tau][g][slot] = key;
                                    sm.bspos[tau][g][slot] = sm.cspos[c];
                                }
                            }
                        }
                    }
                }
            };
            if (resident) {
                collect(nl * LEAF);
                __syncthreads();
            } else {
                if (lane == 0) sm.stack[0] = root >= 0 ? 0 : -1;
                sp = 1;
                __syncthreads();
                for (;;) {
                    nl = fill_chunk(sm, bv, tb, R2, sp);
                    load_chunk(sm, bv, nl);
                    collect(nl * LEAF);
                    __syncthreads();
                    if (sp <= 0) break;
                }
            }

            // --- exact ranking inside the boundary bins --------------------------------------------
            float kth = 0.0f;
            if (ok) {
#pragma unroll
                for (int tau = 0; tau < 2; tau++) {
                    if (tg[tau].t <= 0) continue;
                    const int nb = min((int)sm.bcnt[tau][g], BUF);
                    const int need = tg[tau].t - tg[tau].below;
                    for (int j = s; j < nb; j += P) {
                        const uint64_t kj = sm.bkey[tau][g][j];
                        int rank = 0;
                        for (int i = 0; i < nb; i++) rank += sm.bkey[tau][g][i] < kj ? 1 : 0;
                        if (rank < need) {
                            const uint32_t sp_ = sm.bspos[tau][g][j];
                            sink.accept(tau, g, s, bv.px[sp_], bv.py[sp_], bv.pz[sp_], kj, sp_, sm.scnt[tau]);
                        }
                        if (tau == 0 && rank == need - 1) kth = bitsf((uint32_t)(kj >> 32));
                    }
                }
            }
            __syncthreads();
            sink.finish(g, s, ok, qpos, qx, qy, qz, sm.scnt[0], sm.scnt[1]);
            // largest kth distance of the tile: next tile's radius estimate
            kth_max = fmaxf(kth_max, uniformf(wave_max(ok ? kth : 0.0f)));
            active = active && !ok;
            if (ballot(active) != 0) {
                float nr2 = R2 * 4.0f;
                if (!(nr2 > R2) || nr2 < 1e-30f) nr2 = fmaxf(R2 * 4.0f, 1e-30f);
                if (nr2 > 1e37f) nr2 = INFINITY;
                r2 = nr2;
            }
            __syncthreads();
        }
        if (ballot(active) != 0) {
            if (lane == 0) atomicMax(prm.err, 6);
        }
        // radius guess for the next tile of this wave
        if (kth_max > 0.0f) r2 = kth_max * prm.r2_margin;
        r2 = fmaxf(r2, 1e-30f);
    }
}

}  // namespace lvrhip
