// Adjoint (reverse mode) of the Hoppe signed distance with respect to the point positions,
// through the PCA normals.
//
// Forward (AdaptiveKSearchSurface::distance, tcc:414-457, and calcPlaneIPCAExact,
// tcc:609-673): for a vertex v with kd nearest points N(v),
//     c = mean_{j in N(v)} p_j,  mt = mean_{j in N(v)} n_j,  m = mt / |mt|,  proj = (v - c) . m
// and every normal n_j = s_j u0(C_j), the eigenvector of the smallest eigenvalue of the
// covariance C_j of the n = k_j + 1 samples {p_j} + kNN_{k_j}(p_j) (p_j itself twice), with
// the sign s_j of the flip.  Neighbour sets, adaptive k and flip signs are piecewise constant
// in the positions: the adjoint treats them as fixed, and recomputes the neighbour lists with
// the forward's kNN kernel instead of storing them.
//
// Reverse, for upstream weights w_v = dL/dproj_v:
//     dL/dp_j  += -w_v m / kd                                    (through c)
//     dL/dn_j  +=  w_v (I - m m^T)(v - c) / (|mt| kd)            (through m)
//     dL/dC_j   =  s_j sum_{i=1,2} (g_j . u_i) / (l_0 - l_i) sym(u_i u_0^T)   (g_j = dL/dn_j)
//     dL/dx_s  +=  (2 / n) dL/dC_j (x_s - mu_j) for every sample x_s of C_j (mean mu_j)
// the last from dC = (1/n) sum_s (dx_s (x_s - mu)^T + (x_s - mu) dx_s^T).
#pragma once

#include "common.hpp"
#include "eig3.hpp"
#include "sinks.hpp"

namespace lvrhip {

// SDF reverse pass: the forward's kd-NN and ordered means, then dL/dp and dL/dn of the kd
// neighbours (atomics into sorted-position arrays).
template <int P>
struct SinkSdfAdjoint {
    static constexpr int kTargets = 2;
    const float* px;
    const float* py;
    const float* pz;
    const float* nx;
    const float* ny;
    const float* nz;
    const uint32_t* qorig;    // query position -> caller's index
    const float* w;           // dL/dproj per caller index
    uint64_t* keys;           // scratch [qpos * kw + j]
    uint32_t* spos;           // scratch
    uint32_t* ord;            // scratch, sorted spos
    int kw;
    float* gpx;               // dL/dp, sorted positions
    float* gpy;
    float* gpz;
    float* gnx;               // dL/dn, sorted positions
    float* gny;
    float* gnz;
    uint32_t npts;
    uint32_t qpos_;

    __device__ __forceinline__ void begin(int, int, int64_t, uint32_t qpos, float, float, float) { qpos_ = qpos; }

    __device__ __forceinline__ void accept(int tau, int g, int, float, float, float, uint64_t key, uint32_t sp,
                                           uint32_t* cnt)
    {
        if (tau != 0) return;
        const uint32_t slot = atomicAdd(&cnt[g], 1u);
        if (slot < (uint32_t)kw) {
            keys[(size_t)qpos_ * kw + slot] = key;
            spos[(size_t)qpos_ * kw + slot] = sp;
        }
    }

    __device__ __forceinline__ void finish(int g, int s, bool ok, uint32_t qpos, float qx, float qy, float qz,
                                           const uint32_t* cnt0, const uint32_t*)
    {
        const int cnt = min((int)cnt0[g], kw);
        if (ok)
            group_rank_sort<P>(s, cnt, keys + (size_t)qpos * kw, spos + (size_t)qpos * kw, ord + (size_t)qpos * kw);
        __syncthreads();
        if (!ok || s != 0 || cnt == 0) return;
        const uint32_t oi = qorig ? qorig[qpos] : qpos;
        const float wv = w[oi];
        if (wv == 0.0f) return;
        // the forward's ordered fp32 means (tcc:435-449)
        float ax = 0, ay = 0, az = 0, bx = 0, by = 0, bz = 0;
        const uint32_t* o = ord + (size_t)qpos * kw;
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
        const float c = (float)cnt;
        ax /= c;
        ay /= c;
        az /= c;
        bx /= c;
        by /= c;
        bz /= c;
        const float len = sqrtf(bx * bx + by * by + bz * bz);
        if (!(len > 0.0f)) return;  // zero mean normal: the forward leaves it unnormalised, no gradient
        const float mx = bx / len, my = by / len, mz = bz / len;
        const float dx = qx - ax, dy = qy - ay, dz = qz - az;
        const float dm = dx * mx + dy * my + dz * mz;
        // (I - m m^T)(v - c) / (|mt| kd)
        const float tn = wv / (len * c);
        const float hx = (dx - dm * mx) * tn, hy = (dy - dm * my) * tn, hz = (dz - dm * mz) * tn;
        const float tp = -wv / c;
        for (int j = 0; j < cnt; j++) {
            const uint32_t sp = o[j];
            if (sp >= npts) continue;
            atomicAdd(&gpx[sp], tp * mx);
            atomicAdd(&gpy[sp], tp * my);
            atomicAdd(&gpz[sp], tp * mz);
            atomicAdd(&gnx[sp], hx);
            atomicAdd(&gny[sp], hy);
            atomicAdd(&gnz[sp], hz);
        }
    }
};

// Per sorted position with dL/dn != 0: dL/dC (symmetric, 6 values), the sample mean mu and
// the factor 2/n, for SinkPcaAdjoint; flags[j] = 1 + adaptive round (0: no gradient).
constexpr int ADJ_STRIDE = 10;  // cxx cxy cxz cyy cyz czz, mux muy muz, 2/n

__global__ __launch_bounds__(256) void pca_adjoint_prep_kernel(
    int64_t n, const float* __restrict__ px, const float* __restrict__ py, const float* __restrict__ pz,
    const float* __restrict__ nx, const float* __restrict__ ny, const float* __restrict__ nz,
    const float* __restrict__ mom, const uint8_t* __restrict__ kround, int kn, int64_t kmax,
    const float* __restrict__ gnx,
    const float* __restrict__ gny, const float* __restrict__ gnz, float* __restrict__ adj, uint32_t* __restrict__ flags)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const double g0 = gnx[j], g1 = gny[j], g2 = gnz[j];
    const float qx = px[j];
    if (!(qx == qx) || (g0 == 0.0 && g1 == 0.0 && g2 == 0.0)) {  // padding, or no gradient
        flags[j] = 0u;
        return;
    }
    const int r = kround[j];
    const int64_t k = ((int64_t)kn << r) < kmax ? ((int64_t)kn << r) : kmax;  // the forward's clamped k
    const double cnt = (double)(k + 1);
    const float* m = mom + (size_t)j * 9;
    const double mx = m[0] / cnt, my = m[1] / cnt, mz = m[2] / cnt;
    double a[3][3], v[3][3];
    a[0][0] = m[3] / cnt - mx * mx;
    a[0][1] = a[1][0] = m[4] / cnt - mx * my;
    a[0][2] = a[2][0] = m[5] / cnt - mx * mz;
    a[1][1] = m[6] / cnt - my * my;
    a[1][2] = a[2][1] = m[7] / cnt - my * mz;
    a[2][2] = m[8] / cnt - mz * mz;
    jacobi3_dev<double>(a, v);
    const double l[3] = {a[0][0], a[1][1], a[2][2]};
    const int i0 = smallest_col_dev<double>(l[0], l[1], l[2]);
    const double u0[3] = {v[0][i0], v[1][i0], v[2][i0]};
    // the flip's sign: the forward normal is +-u0
    const double sgn = (u0[0] * nx[j] + u0[1] * ny[j] + u0[2] * nz[j]) < 0.0 ? -1.0 : 1.0;
    double cb[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int i = 0; i < 3; i++) {
        if (i == i0) continue;
        const double gap = l[i0] - l[i];
        if (!(fabs(gap) > 1e-300)) continue;  // exactly degenerate: no defined derivative
        const double ui[3] = {v[0][i], v[1][i], v[2][i]};
        const double f = sgn * (g0 * ui[0] + g1 * ui[1] + g2 * ui[2]) / gap;
        for (int a_ = 0; a_ < 3; a_++)
            for (int b_ = 0; b_ < 3; b_++) cb[a_][b_] += 0.5 * f * (ui[a_] * u0[b_] + u0[a_] * ui[b_]);
    }
    float* o = adj + (size_t)j * ADJ_STRIDE;
    o[0] = (float)cb[0][0];
    o[1] = (float)cb[0][1];
    o[2] = (float)cb[0][2];
    o[3] = (float)cb[1][1];
    o[4] = (float)cb[1][2];
    o[5] = (float)cb[2][2];
    // sample mean: mu = q + S1 / n (moments are about q)
    o[6] = (float)((double)qx + mx);
    o[7] = (float)((double)py[j] + my);
    o[8] = (float)((double)pz[j] + mz);
    o[9] = (float)(2.0 / cnt);
    flags[j] = 1u + (uint32_t)r;
}

// flags[i] == want for the query slots i < n (round selection of the PCA reverse pass)
__global__ void select_round_kernel(const uint32_t* __restrict__ flags, int64_t n, uint32_t want,
                                    uint32_t* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = flags[i] == want ? 1u : 0u;
}

// PCA reverse pass: the forward's k_j-NN of each selected point again; every sample x_s of
// C_j (the k-set, and the point itself once more as the seed) receives (2/n) dL/dC (x_s - mu).
template <int P>
struct SinkPcaAdjoint {
    static constexpr int kTargets = 2;
    const float* adj;
    float* gpx;
    float* gpy;
    float* gpz;
    float c0, c1, c2, c3, c4, c5, ux, uy, uz, f;

    __device__ __forceinline__ void begin(int, int, int64_t, uint32_t qpos, float, float, float)
    {
        const float* o = adj + (size_t)qpos * ADJ_STRIDE;
        c0 = o[0];
        c1 = o[1];
        c2 = o[2];
        c3 = o[3];
        c4 = o[4];
        c5 = o[5];
        ux = o[6];
        uy = o[7];
        uz = o[8];
        f = o[9];
    }

    __device__ __forceinline__ void push(uint32_t sp, float x, float y, float z) const
    {
        const float rx = x - ux, ry = y - uy, rz = z - uz;
        atomicAdd(&gpx[sp], f * (c0 * rx + c1 * ry + c2 * rz));
        atomicAdd(&gpy[sp], f * (c1 * rx + c3 * ry + c4 * rz));
        atomicAdd(&gpz[sp], f * (c2 * rx + c4 * ry + c5 * rz));
    }

    __device__ __forceinline__ void accept(int tau, int, int, float x, float y, float z, uint64_t, uint32_t sp,
                                           uint32_t*)
    {
        if (tau == 0) push(sp, x, y, z);
    }

    __device__ __forceinline__ void finish(int, int s, bool ok, uint32_t qpos, float qx, float qy, float qz,
                                           const uint32_t*, const uint32_t*)
    {
        if (ok && s == 0) push(qpos, qx, qy, qz);  // the seed sample
    }
};

// grad_xyz[pidx[j]] += (gpx, gpy, gpz)[j]
__global__ void accumulate_grad_kernel(const float* __restrict__ gpx, const float* __restrict__ gpy,
                                       const float* __restrict__ gpz, const uint32_t* __restrict__ pidx, int64_t n,
                                       float* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = pidx[i];
    if (j == ~0u) return;
    out[3 * (int64_t)j] += gpx[i];
    out[3 * (int64_t)j + 1] += gpy[i];
    out[3 * (int64_t)j + 2] += gpz[i];
}

__global__ void zero_f32_kernel(float* __restrict__ p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

}  // namespace lvrhip
