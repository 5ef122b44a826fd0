// Consumers of the exact kNN sets produced by knn_tile_kernel.
//
//   SinkNormals: fused plane fit per query point (AdaptiveKSearchSurface.tcc:158-284):
//                boundingBoxOK on the k-set, the selected method's plane normal, the flip
//                towards the flip point, plus the sorted ki-prefix used by the
//                interpolation (tcc:293-335).
//   SinkSdf:     AdaptiveKSearchSurface::distance (tcc:414-457) + the invalid rule of
//                PointsetGrid::calcDistanceValues (PointsetGrid.tcc:138-143).
//   SinkList:    raw (d^2, index) keys for SearchTree::kSearch.
//
// Sinks see each selected neighbour once, on one of the P lanes of its query; finish()
// runs on all lanes (it may shuffle within the query's lane group).
#pragma once

#include "common.hpp"
#include "eig3.hpp"

namespace lvrhip {

// Sum / minimum / maximum over a query's P lanes (every lane of the wave active).
template <int P>
__device__ __forceinline__ float group_sum(float v)
{
    return lane_group_reduce<P>(v, [](float a, float b) { return a + b; });
}
template <int P>
__device__ __forceinline__ float group_min(float v)
{
    return lane_group_reduce<P>(v, [](float a, float b) { return fminf(a, b); });
}
template <int P>
__device__ __forceinline__ float group_max(float v)
{
    return lane_group_reduce<P>(v, [](float a, float b) { return fmaxf(a, b); });
}

// Sort the n (key, payload) pairs of one query by key (keys are unique) with the P lanes of
// the group: each lane ranks its elements against all others.
template <int P>
__device__ __forceinline__ void group_rank_sort(int s, int n, const uint64_t* keys, const uint32_t* payload,
                                                uint32_t* out)
{
    for (int j = s; j < n; j += P) {
        const uint64_t kj = keys[j];
        int rank = 0;
        for (int i = 0; i < n; i++) rank += keys[i] < kj ? 1 : 0;
        out[rank] = payload[j];
    }
}

// BaseVector::normalize (geometry/BaseVector.hpp:130-143) in fp32.
__device__ __forceinline__ void normalize3(float& x, float& y, float& z)
{
    if (!(x == 0.0f && y == 0.0f && z == 0.0f)) {
        float l = sqrtf(x * x + y * y + z * z);
        x /= l;
        y /= l;
        z /= l;
    }
}

// Plane normal from the moments of {q} + k-set about q:
// m = (sx sy sz sxx sxy sxz syy syz szz), kcount = |k-set|.
__device__ inline void plane_from_moments(int method, int kcount, float qx, float qy, float qz, const float* m,
                                          float& n0, float& n1, float& n2)
{
    const float sx = m[0], sy = m[1], sz = m[2], sxx = m[3], sxy = m[4], sxz = m[5], syy = m[6], syz = m[7],
                szz = m[8];
    if (method == 2) {
        // calcPlaneIterative (tcc:522-607): second moments about the query point
        const float xx = sxx, xy = sxy, xz = sxz, yy = syy, yz = syz, zz = szz;
        const float det_x = yy * zz - yz * yz;
        const float det_y = xx * zz - xz * xz;
        const float det_z = xx * yy - xy * xy;
        float dx, dy, dz;
        if (det_x >= det_y && det_x >= det_z) {
            dx = 1.0f;
            dy = (xz * yz - xy * zz) / det_x;
            dz = (xy * yz - xz * yy) / det_x;
        } else if (det_y >= det_x && det_y >= det_z) {
            dx = (yz * xz - xy * zz) / det_y;
            dy = 1.0f;
            dz = (xy * xz - yz * xx) / det_y;
        } else {
            dx = (yz * xy - xz * yy) / det_z;
            dy = (xz * xy - yz * xx) / det_z;
            dz = 1.0f;
        }
        const float inv = 1.0f / sqrtf(dx * dx + dy * dy + dz * dz);
        n0 = dx * inv;
        n1 = dy * inv;
        n2 = dz * inv;
    } else if (method == 0) {
        // calcPlane (tcc:467-520): least squares y = c0 + c1 x + c2 z on the k-set, solved in
        // fp64 from the query-centred moments; the normal is the cross product of two tangents
        // that start at the query point, 100 units along x and z (tcc:497-504)
        const double c = (double)kcount;
        double A[3][3] = {{c, sx, sz}, {sx, sxx, sxz}, {sz, sxz, szz}};
        double b[3] = {sy, sxy, syz};
        solve3(A, b);
        const double c1 = b[1], c2 = b[2];
        const double c0 = (double)qy + b[0] - c1 * qx - c2 * qz;
        const float eps = 100.0f;
        const float C0 = (float)c0, C1 = (float)c1, C2 = (float)c2;
        const float z1 = C0 + C1 * (qx + eps) + C2 * qz;
        const float z2 = C0 + C1 * qx + C2 * (qz + eps);
        const float d1x = (qx + eps) - qx, d1y = z1 - qy, d1z = qz - qz;
        const float d2x = qx - qx, d2y = z2 - qy, d2z = (qz + eps) - qz;
        n0 = d1y * d2z - d1z * d2y;
        n1 = d1z * d2x - d1x * d2z;
        n2 = d1x * d2y - d1y * d2x;
        normalize3(n0, n1, n2);
    } else {
        // calcPlaneIPCAExact (tcc:609-673): covariance of {q} + k-set, smallest eigenvector.
        // The covariance is formed in fp64 from the fp32 moments: its cancellation
        // (S2 / n - m m^T) would otherwise add eps32 * |m|^2 / l_max to the matrix.
        const double cnt = (double)(kcount + 1);
        const double mx = sx / cnt, my = sy / cnt, mz = sz / cnt;
        double C[6];
        C[0] = sxx / cnt - mx * mx;
        C[1] = sxy / cnt - mx * my;
        C[2] = sxz / cnt - mx * mz;
        C[3] = syy / cnt - my * my;
        C[4] = syz / cnt - my * mz;
        C[5] = szz / cnt - mz * mz;
        smallest_eigvec(C, n0, n1, n2);
    }
}

// Fused into the kNN pass: the k-set's bounding box test (boundingBoxOK,
// AdaptiveKSearchSurface.tcc:337-350) and its moments about the query point; the plane fit
// itself runs in plane_fit_kernel.  Also the sorted ki-prefix for the interpolation.
template <int P>
struct SinkNormals {
    static constexpr int kTargets = 2;
    // parameters
    float* mom;            // [qpos * 9] moments of the k-set about q
    uint32_t* retry;       // per query slot: 1 if the bbox test failed (NULL in the last round)
    uint32_t* ki_list;     // out [qpos * kiw + j]: the ki nearest, unordered
    int kiw;
    // per-lane state
    uint32_t qpos_, qi_;
    float qx_, qy_, qz_;
    float sx, sy, sz, sxx, sxy, sxz, syy, syz, szz;
    float mnx, mny, mnz, mxx, mxy, mxz;

    __device__ __forceinline__ void begin(int g, int s, int64_t qi, uint32_t qpos, float qx, float qy, float qz)
    {
        qpos_ = qpos;
        qi_ = (uint32_t)qi;
        qx_ = qx;
        qy_ = qy;
        qz_ = qz;
        sx = sy = sz = sxx = sxy = sxz = syy = syz = szz = 0.0f;
        mnx = mny = mnz = INFINITY;
        mxx = mxy = mxz = -INFINITY;
    }

    __device__ __forceinline__ void accept(int tau, int g, int s, float x, float y, float z, uint64_t key, uint32_t spos,
                           uint32_t* cnt)
    {
        if (tau == 0) {
            const float rx = x - qx_, ry = y - qy_, rz = z - qz_;
            sx += rx;
            sy += ry;
            sz += rz;
            sxx = fmaf(rx, rx, sxx);
            sxy = fmaf(rx, ry, sxy);
            sxz = fmaf(rx, rz, sxz);
            syy = fmaf(ry, ry, syy);
            syz = fmaf(ry, rz, syz);
            szz = fmaf(rz, rz, szz);
            mnx = min_num(mnx, x);
            mny = min_num(mny, y);
            mnz = min_num(mnz, z);
            mxx = max_num(mxx, x);
            mxy = max_num(mxy, y);
            mxz = max_num(mxz, z);
        } else {
            const uint32_t slot = atomicAdd(&cnt[g], 1u);
            if (slot < (uint32_t)kiw) ki_list[(size_t)qpos_ * kiw + slot] = spos;
        }
    }

    // A leaf all of whose points are in the k-set: its point sums about the box centre c
    // (leaf_sums_kernel) moved to q: sum (p - q) = S1 + n d and
    // sum (p - q)(p - q)^T = S2 + d S1^T + S1 d^T + n d d^T with d = c - q; its box is the
    // exact bounding box of its points.
    static constexpr bool kLeafAccept = true;
    __device__ __forceinline__ void accept_leaf(int, int, const float4& a, const float4& b, const float4& c,
                                                const float4& lo, const float4& hi)
    {
        const float dx = leaf_centre(lo.x, hi.x) - qx_, dy = leaf_centre(lo.y, hi.y) - qy_,
                    dz = leaf_centre(lo.z, hi.z) - qz_;
        const float n = a.w;
        const float ndx = n * dx, ndy = n * dy, ndz = n * dz;
        sx += a.x + ndx;
        sy += a.y + ndy;
        sz += a.z + ndz;
        sxx += b.x + fmaf(2.0f * dx, a.x, ndx * dx);
        sxy += b.y + fmaf(dx, a.y, fmaf(dy, a.x, ndx * dy));
        sxz += b.z + fmaf(dx, a.z, fmaf(dz, a.x, ndx * dz));
        syy += b.w + fmaf(2.0f * dy, a.y, ndy * dy);
        syz += c.x + fmaf(dy, a.z, fmaf(dz, a.y, ndy * dz));
        szz += c.y + fmaf(2.0f * dz, a.z, ndz * dz);
        mnx = min_num(mnx, lo.x);
        mny = min_num(mny, lo.y);
        mnz = min_num(mnz, lo.z);
        mxx = max_num(mxx, hi.x);
        mxy = max_num(mxy, hi.y);
        mxz = max_num(mxz, hi.z);
    }

    __device__ __forceinline__ void finish(int g, int s, bool ok, uint32_t qpos, float, float, float, const uint32_t*,
                           const uint32_t* cnt1)
    {
        sx = group_sum<P>(sx);
        sy = group_sum<P>(sy);
        sz = group_sum<P>(sz);
        sxx = group_sum<P>(sxx);
        sxy = group_sum<P>(sxy);
        sxz = group_sum<P>(sxz);
        syy = group_sum<P>(syy);
        syz = group_sum<P>(syz);
        szz = group_sum<P>(szz);
        mnx = group_min<P>(mnx);
        mny = group_min<P>(mny);
        mnz = group_min<P>(mnz);
        mxx = group_max<P>(mxx);
        mxy = group_max<P>(mxy);
        mxz = group_max<P>(mxz);
        if (ok && s == 0) {
            const float dx = mxx - mnx, dy = mxy - mny, dz = mxz - mnz;
            const float e = 0.05f;
            const bool bb_ok = dx >= e * dy && dx >= e * dz && dy >= e * dx && dy >= e * dz && dz >= e * dx &&
                               dz >= e * dy;
            if (retry) retry[qi_] = bb_ok ? 0u : 1u;
            float* m = mom + (size_t)qpos * 9;
            m[0] = sx;
            m[1] = sy;
            m[2] = sz;
            m[3] = sxx;
            m[4] = sxy;
            m[5] = sxz;
            m[6] = syy;
            m[7] = syz;
            m[8] = szz;
        }
    }
};

// SinkNormals for the adaptive rounds after the first, which carry no ki list: one target,
// so the tile kernel's per-target state (and its LDS) halves and the boundary buffer holds more
// keys (the selection of k = 32 kn then rarely needs a refinement pass).
template <int P>
struct SinkMoments : SinkNormals<P> {
    static constexpr int kTargets = 1;
};

// The later adaptive rounds in one search (AdaptiveKSearchSurface.tcc:187-216, rounds 2..5):
// targets t0 = 32 kn (round 5), t1 = 16 kn, t2 = 8 kn, t3 = 4 kn (round 2).  Each round's k-set
// is a prefix of the next one's, so one selection yields every round's bounding box; the
// point's adaptive round is the first whose box passes boundingBoxOK, else round 5, as in the
// round-by-round loop.  Only round 5's moments are accumulated (the registers of four moment
// sets would halve the occupancy); a point that stops earlier is flagged in `redo` with its
// round and searched again at that round's k.
template <int P>
struct SinkLevels {
    static constexpr int kTargets = 4;
    float* mom;       // out [qpos * 9]: moments of the 32 kn-set about q (round-5 points)
    uint8_t* kround;  // out: the adaptive round
    uint32_t* redo;   // out [query slot]: the round (2..4) to search again, 0 for round 5
    uint32_t qi_;
    float qx_, qy_, qz_;
    float m[9];
    float lo[4][3], hi[4][3];

    __device__ __forceinline__ void begin(int, int, int64_t qi, uint32_t, float qx, float qy, float qz)
    {
        qi_ = (uint32_t)qi;
        qx_ = qx;
        qy_ = qy;
        qz_ = qz;
#pragma unroll
        for (int i = 0; i < 9; i++) m[i] = 0.0f;
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
            for (int i = 0; i < 3; i++) {
                lo[t][i] = INFINITY;
                hi[t][i] = -INFINITY;
            }
    }

    __device__ __forceinline__ void accept(int tau, int, int, float x, float y, float z, uint64_t, uint32_t,
                                           uint32_t*)
    {
        lo[tau][0] = min_num(lo[tau][0], x);
        lo[tau][1] = min_num(lo[tau][1], y);
        lo[tau][2] = min_num(lo[tau][2], z);
        hi[tau][0] = max_num(hi[tau][0], x);
        hi[tau][1] = max_num(hi[tau][1], y);
        hi[tau][2] = max_num(hi[tau][2], z);
        if (tau == 0) {
            const float rx = x - qx_, ry = y - qy_, rz = z - qz_;
            m[0] += rx;
            m[1] += ry;
            m[2] += rz;
            m[3] = fmaf(rx, rx, m[3]);
            m[4] = fmaf(rx, ry, m[4]);
            m[5] = fmaf(rx, rz, m[5]);
            m[6] = fmaf(ry, ry, m[6]);
            m[7] = fmaf(ry, rz, m[7]);
            m[8] = fmaf(rz, rz, m[8]);
        }
    }

    __device__ __forceinline__ void finish(int, int s, bool ok, uint32_t qpos, float, float, float, const uint32_t*,
                                           const uint32_t*)
    {
        bool pass[4];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const float dx = group_max<P>(hi[t][0]) - group_min<P>(lo[t][0]);
            const float dy = group_max<P>(hi[t][1]) - group_min<P>(lo[t][1]);
            const float dz = group_max<P>(hi[t][2]) - group_min<P>(lo[t][2]);
            const float e = 0.05f;
            pass[t] = dx >= e * dy && dx >= e * dz && dy >= e * dx && dy >= e * dz && dz >= e * dx && dz >= e * dy;
        }
#pragma unroll
        for (int i = 0; i < 9; i++) m[i] = group_sum<P>(m[i]);
        if (!ok || s != 0) return;
        // rounds 2, 3, 4 are targets 3, 2, 1; round 5 (target 0) is taken whatever its box
        const int round = pass[3] ? 2 : pass[2] ? 3 : pass[1] ? 4 : 5;
        kround[qpos] = (uint8_t)round;
        redo[qi_] = round < 5 ? (uint32_t)round : 0u;
        if (round == 5) {
            float* o = mom + (size_t)qpos * 9;
#pragma unroll
            for (int i = 0; i < 9; i++) o[i] = m[i];
        }
    }
};

// Adaptive-round counts of the queries of the merged later rounds: cnt[r] += #(kround >= r),
// r = 3..5 (cnt[2], the queries themselves, comes from the compaction).
__global__ __launch_bounds__(256) void round_count_kernel(const uint32_t* __restrict__ qsel,
                                                          const uint32_t* __restrict__ nq_dev, int64_t nq,
                                                          const uint8_t* __restrict__ kround, uint32_t* __restrict__ cnt)
{
    __shared__ uint32_t c[3];
    if (threadIdx.x < 3) c[threadIdx.x] = 0u;
    __syncthreads();
    const int64_t n = nq_dev ? (int64_t)*nq_dev : nq;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const int r = kround[qsel ? qsel[i] : (uint32_t)i];
        if (r >= 3) atomicAdd(&c[0], 1u);
        if (r >= 4) atomicAdd(&c[1], 1u);
        if (r >= 5) atomicAdd(&c[2], 1u);
    }
    __syncthreads();
    if (threadIdx.x < 3 && c[threadIdx.x]) atomicAdd(&cnt[3 + threadIdx.x], c[threadIdx.x]);
}

// Plane fit + flip (AdaptiveKSearchSurface.tcc:225-274) for the queries of one adaptive round
// whose neighbourhood passed boundingBoxOK (or all of them in the last round).
struct PlaneFitParams {
    const uint32_t* qsel;     // query slot -> sorted position (NULL: identity)
    int64_t nq;               // query slots (upper bound when nq_dev is set)
    const uint32_t* nq_dev;   // optional device-side count of query slots
    const uint32_t* retry;    // per slot: 1 = bbox test failed, plane fit deferred (NULL: last round)
    const float* mom;         // [spos * 9] moments of the k-set about the point
    const float* px;
    const float* py;
    const float* pz;
    int method;
    int kcount;               // |k-set| of this round; < 0: min(kn << kround[q], kmax) per query
    int kn;
    int64_t kmax;
    float fx, fy, fz;         // flip point (one scan)
    const float* flips;       // per scan flip points (x, y, z), overrides fx/fy/fz
    const uint16_t* scan_of;  // per sorted position: scan
    float* nx;
    float* ny;
    float* nz;
    uint8_t* kround;
    int round;
};

__global__ __launch_bounds__(256) void plane_fit_kernel(PlaneFitParams a)
{
    const int64_t qi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nq = a.nq_dev ? (int64_t)*a.nq_dev : a.nq;
    if (qi >= nq) return;
    if (a.retry && a.retry[qi]) return;
    const uint32_t qpos = a.qsel ? a.qsel[qi] : (uint32_t)qi;
    const float qx = a.px[qpos], qy = a.py[qpos], qz = a.pz[qpos];
    float m[9];
#pragma unroll
    for (int i = 0; i < 9; i++) m[i] = a.mom[(size_t)qpos * 9 + i];
    float n0, n1, n2;
    int kc = a.kcount;
    if (kc < 0) {
        const int64_t k = (int64_t)a.kn << a.kround[qpos];
        kc = (int)(k < a.kmax ? k : a.kmax);
    }
    plane_from_moments(a.method, kc, qx, qy, qz, m, n0, n1, n2);
    float fx = a.fx, fy = a.fy, fz = a.fz;
    if (a.flips) {
        const int sc = a.scan_of ? (int)a.scan_of[qpos] : 0;
        fx = a.flips[3 * sc];
        fy = a.flips[3 * sc + 1];
        fz = a.flips[3 * sc + 2];
    }
    // flip towards the flip point (tcc:268-274); Normal::operator- renormalises
    const float ddx = fx - qx, ddy = fy - qy, ddz = fz - qz;
    if (n0 * ddx + n1 * ddy + n2 * ddz < 0.0f) {
        n0 = -n0;
        n1 = -n1;
        n2 = -n2;
        normalize3(n0, n1, n2);
    }
    a.nx[qpos] = n0;
    a.ny[qpos] = n1;
    a.nz[qpos] = n2;
    if (a.kround && a.kcount >= 0) a.kround[qpos] = (uint8_t)a.round;
}

// AdaptiveKSearchSurface::distance (tcc:435-457) from the neighbours' means, with the invalid
// rule of PointsetGrid::calcDistanceValues (PointsetGrid.tcc:138-143): shared by the tile
// kernel's SDF sink and the best-first search.
__device__ __forceinline__ void sdf_emit(float ax, float ay, float az, float bx, float by, float bz, int cnt, float qx,
                                         float qy, float qz, float voxel, uint32_t oi, float* out_proj,
                                         float* out_eucl, uint8_t* out_invalid)
{
    const float c = (float)cnt;
    bx /= c;
    by /= c;
    bz /= c;
    ax /= c;
    ay /= c;
    az /= c;
    if (!(bx == 0.0f && by == 0.0f && bz == 0.0f)) {  // BaseVector::normalize
        const float l = sqrtf(bx * bx + by * by + bz * bz);
        bx /= l;
        by /= l;
        bz /= l;
    }
    const float dx = qx - ax, dy = qy - ay, dz = qz - az;
    const float proj = dx * bx + dy * by + dz * bz;
    const float eucl = sqrtf(dx * dx + dy * dy + dz * dz);
    if (out_proj) out_proj[oi] = proj;
    if (out_eucl) out_eucl[oi] = eucl;
    // the comparison is made in double
    if (out_invalid) out_invalid[oi] = ((double)eucl > 1.75 * (double)voxel) ? 1 : 0;
}

template <int P>
struct SinkSdf {
    static constexpr int kTargets = 2;
    const float* px;
    const float* py;
    const float* pz;
    const float* nx;
    const float* ny;
    const float* nz;
    const uint32_t* qorig;  // query position -> caller's index
    uint64_t* keys;         // scratch [qpos * kw + j]
    uint32_t* spos;         // scratch
    uint32_t* ord;          // scratch, sorted spos
    int kw;
    float voxel;
    float* out_proj;
    float* out_eucl;
    uint8_t* out_invalid;
    uint32_t npts;
    uint32_t qpos_;
    float4* lds;  // the kernel's candidate staging area (CAP float4), free in finish

    __device__ __forceinline__ void begin(int g, int s, int64_t qi, uint32_t qpos, float, float, float) { qpos_ = qpos; }

    __device__ __forceinline__ void accept(int tau, int g, int s, float, float, float, uint64_t key, uint32_t sp, uint32_t* cnt)
    {
        if (tau != 0) return;
        const uint32_t slot = atomicAdd(&cnt[g], 1u);
        if (slot < (uint32_t)kw) {
            keys[(size_t)qpos_ * kw + slot] = key;
            spos[(size_t)qpos_ * kw + slot] = sp;
        }
    }

    // means of the ordered sums, signed distance, invalid flag (tcc:435-457, PointsetGrid.tcc:138)
    __device__ __forceinline__ void emit(float ax, float ay, float az, float bx, float by, float bz, int cnt,
                                         uint32_t qpos, float qx, float qy, float qz)
    {
        sdf_emit(ax, ay, az, bx, by, bz, cnt, qx, qy, qz, voxel, qorig ? qorig[qpos] : qpos, out_proj, out_eucl,
                 out_invalid);
    }

    __device__ __forceinline__ void finish(int g, int s, bool ok, uint32_t qpos, float qx, float qy, float qz,
                                           const uint32_t* cnt0, const uint32_t*)
    {
        constexpr int Q = kWave / P;
        constexpr int SLOT = CAP * 16 / Q;  // staging bytes per query
        constexpr int ENT = SLOT / 32;      // neighbours per query: an 8-byte key + 6 floats each
        const int cnt = min((int)cnt0[g], kw);
        const size_t row = (size_t)qpos * kw;
        // ordered fp32 sums (AdaptiveKSearchSurface.tcc:435-449) in ascending key order
        float ax = 0, ay = 0, az = 0, bx = 0, by = 0, bz = 0;
        if (kw <= ENT && lds) {
            // rank in LDS and gather the neighbours' points and normals with all P lanes
            char* base = reinterpret_cast<char*>(lds) + g * SLOT;
            uint64_t* kk = reinterpret_cast<uint64_t*>(base);
            float* d = reinterpret_cast<float*>(base + 8 * ENT);
            if (ok)
                for (int j = s; j < cnt; j += P) kk[j] = keys[row + j];
            __syncthreads();
            if (ok)
                for (int j = s; j < cnt; j += P) {
                    const uint64_t kj = kk[j];
                    int rank = 0;
                    for (int i = 0; i < cnt; i++) rank += kk[i] < kj ? 1 : 0;
                    const uint32_t sp = spos[row + j];
                    const bool in = sp < npts;  // (always, for keys of staged points)
                    float* o = d + 6 * rank;
                    o[0] = in ? px[sp] : 0.0f;
                    o[1] = in ? py[sp] : 0.0f;
                    o[2] = in ? pz[sp] : 0.0f;
                    o[3] = in ? nx[sp] : 0.0f;
                    o[4] = in ? ny[sp] : 0.0f;
                    o[5] = in ? nz[sp] : 0.0f;
                }
            __syncthreads();
            if (!ok || s != 0) return;
            for (int j = 0; j < cnt; j++) {
                ax += d[6 * j];
                ay += d[6 * j + 1];
                az += d[6 * j + 2];
                bx += d[6 * j + 3];
                by += d[6 * j + 4];
                bz += d[6 * j + 5];
            }
        } else {
            if (ok) group_rank_sort<P>(s, cnt, keys + row, spos + row, ord + row);
            __syncthreads();
            if (!ok || s != 0) return;
            const uint32_t* o = ord + row;
            for (int j = 0; j < cnt; j++) {
                const uint32_t sp = o[j];
                if (sp >= npts) continue;
                ax += px[sp];
                ay += py[sp];
                az += pz[sp];
                bx += nx[sp];
                by += ny[sp];
                bz += nz[sp];
            }
        }
        emit(ax, ay, az, bx, by, bz, cnt, qpos, qx, qy, qz);
    }
};

template <int P>
struct SinkList {
    static constexpr int kTargets = 2;
    uint64_t* out;          // [qorig * kw + slot], unsorted
    const uint32_t* qorig;
    int kw;
    uint32_t row_;

    __device__ __forceinline__ void begin(int g, int s, int64_t qi, uint32_t qpos, float, float, float)
    {
        row_ = qorig ? qorig[qpos] : qpos;
    }
    __device__ __forceinline__ void accept(int tau, int g, int s, float, float, float, uint64_t key, uint32_t, uint32_t* cnt)
    {
        if (tau != 0) return;
        const uint32_t slot = atomicAdd(&cnt[g], 1u);
        if (slot < (uint32_t)kw) out[(size_t)row_ * kw + slot] = key;
    }
    __device__ __forceinline__ void finish(int, int, bool, uint32_t, float, float, float, const uint32_t*, const uint32_t*) {}
};

}  // namespace lvrhip
