// Small dense solvers used per query point.
//
// smallest_eigvec: eigenvector of the smallest eigenvalue of a symmetric 3x3 matrix, the
// quantity calcPlaneIPCAExact takes from Eigen::EigenSolver<Matrix3f>
// (AdaptiveKSearchSurface.tcc:642-659).  Cyclic Jacobi rotations in fp32; when the two
// smallest eigenvalues are within EIG_FP64_GAP of the spread the decomposition is redone in
// fp64.  The fp32 pass is backward stable: its eigenvector error is about
// eps32 / ((l_mid - l_min) / l_max), 2e-6 rad at the threshold.
#pragma once

#include "common.hpp"

namespace lvrhip {

template <typename T>
__device__ __forceinline__ void jacobi3_dev(T a[3][3], T v[3][3])
{
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) v[i][j] = (i == j) ? T(1) : T(0);
    for (int sweep = 0; sweep < 24; sweep++) {
        if (a[0][1] == T(0) && a[0][2] == T(0) && a[1][2] == T(0)) break;
#pragma unroll
        for (int r = 0; r < 3; r++) {
            const int p = r == 2 ? 1 : 0;
            const int q = r == 0 ? 1 : 2;
            const T apq = a[p][q];
            if (apq == T(0)) continue;
            const T g = T(100) * fabs(apq);
            if (sweep > 3 && fabs(a[p][p]) + g == fabs(a[p][p]) && fabs(a[q][q]) + g == fabs(a[q][q])) {
                a[p][q] = a[q][p] = T(0);
                continue;
            }
            const T theta = (a[q][q] - a[p][p]) / (T(2) * apq);
            T t;
            if (fabs(theta) > T(1e30))
                t = T(1) / (T(2) * theta);
            else
                t = (theta >= T(0) ? T(1) : T(-1)) / (fabs(theta) + sqrt(theta * theta + T(1)));
            const T c = T(1) / sqrt(t * t + T(1));
            const T s = t * c;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const T akp = a[k][p], akq = a[k][q];
                a[k][p] = c * akp - s * akq;
                a[k][q] = s * akp + c * akq;
            }
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const T apk = a[p][k], aqk = a[q][k];
                a[p][k] = c * apk - s * aqk;
                a[q][k] = s * apk + c * aqk;
            }
            a[p][q] = a[q][p] = T(0);
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const T vkp = v[k][p], vkq = v[k][q];
                v[k][p] = c * vkp - s * vkq;
                v[k][q] = s * vkp + c * vkq;
            }
        }
    }
}

// Column rule of calcPlaneIPCAExact (tcc:647-659).
template <typename T>
__device__ __forceinline__ int smallest_col_dev(T w0, T w1, T w2)
{
    if (w0 < w1 && w0 < w2) return 0;
    if (w1 < w2) return 1;
    return 2;
}

constexpr float EIG_FP64_GAP = 3e-2f;

// C: xx xy xz yy yz zz (formed in fp64; the fp32 pass sees it rounded)
__device__ inline void smallest_eigvec(const double* C, float& n0, float& n1, float& n2)
{
    float a[3][3] = {{(float)C[0], (float)C[1], (float)C[2]},
                     {(float)C[1], (float)C[3], (float)C[4]},
                     {(float)C[2], (float)C[4], (float)C[5]}};
    float v[3][3];
    jacobi3_dev<float>(a, v);
    const float w0 = a[0][0], w1 = a[1][1], w2 = a[2][2];
    const float lo = fminf(w0, fminf(w1, w2)), hi = fmaxf(w0, fmaxf(w1, w2));
    const float mid = w0 + w1 + w2 - lo - hi;
    if (mid - lo <= EIG_FP64_GAP * (hi - lo)) {
        // nearly degenerate: fp64 decomposition of the fp64 matrix
        double ad[3][3] = {{C[0], C[1], C[2]}, {C[1], C[3], C[4]}, {C[2], C[4], C[5]}};
        double vd[3][3];
        jacobi3_dev<double>(ad, vd);
        const int c = smallest_col_dev<double>(ad[0][0], ad[1][1], ad[2][2]);
        double x = vd[0][c], y = vd[1][c], z = vd[2][c];
        const double l = sqrt(x * x + y * y + z * z);
        n0 = (float)(x / l);
        n1 = (float)(y / l);
        n2 = (float)(z / l);
        return;
    }
    const int c = smallest_col_dev<float>(w0, w1, w2);
    const float x = v[0][c], y = v[1][c], z = v[2][c];
    const float l = sqrtf((x * x + y * y) + z * z);
    n0 = x / l;
    n1 = y / l;
    n2 = z / l;
}

// In-place solve of A x = b (3x3, partial pivoting); x returned in b.  Singular pivots
// give 0 for that unknown.
__device__ inline void solve3(double A[3][3], double b[3])
{
#pragma unroll
    for (int c = 0; c < 3; c++) {
        int piv = c;
        for (int r = c + 1; r < 3; r++)
            if (fabs(A[r][c]) > fabs(A[piv][c])) piv = r;
        if (piv != c) {
            for (int k = 0; k < 3; k++) {
                double t = A[c][k];
                A[c][k] = A[piv][k];
                A[piv][k] = t;
            }
            double t = b[c];
            b[c] = b[piv];
            b[piv] = t;
        }
        if (A[c][c] == 0.0) continue;
        for (int r = c + 1; r < 3; r++) {
            const double f = A[r][c] / A[c][c];
            for (int k = c; k < 3; k++) A[r][k] -= f * A[c][k];
            b[r] -= f * b[c];
        }
    }
    double x[3];
    for (int r = 2; r >= 0; r--) {
        double s = b[r];
        for (int k = r + 1; k < 3; k++) s -= A[r][k] * x[k];
        x[r] = A[r][r] != 0.0 ? s / A[r][r] : 0.0;
    }
    b[0] = x[0];
    b[1] = x[1];
    b[2] = x[2];
}

}  // namespace lvrhip
