/*
 * ORACLE — test infrastructure only (see lvr2_oracle.h).
 *
 * CPU restatement of the lvr2 AdaptiveKSearchSurface path.  Each function cites the
 * reference file:line it follows.  Eigen is not available here, so two solvers are
 * substituted: the LS plane fit of method 0 solves the 3x3 normal equations in FP64
 * (reference: JacobiSVD<MatrixXf>, AdaptiveKSearchSurface.tcc:494) and the eigen
 * problem of method 3 uses cyclic Jacobi (reference: EigenSolver<Matrix3f>, tcc:642).
 * Both agree with the reference to FP32 tolerance, not to the last bit.
 *
 * Built with -ffp-contract=off so every FP32 expression rounds exactly as written.
 */
#include "lvr2_oracle.h"
#include "../lvr2_amd/csrc/mc_table.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <queue>
#include <unordered_map>
#include <vector>
#include <omp.h>

namespace {

struct IdxPoint {
    float p[3];
    uint32_t index;
    float operator[](unsigned i) const { return p[i]; }
};

/* ---- kNN keys ------------------------------------------------------------------ */

// Contract key: fp32 squared distance evaluated as (dx*dx + dy*dy) + dz*dz.
inline float d2_f32(const float* q, const float* p)
{
    float dx = q[0] - p[0];
    float dy = q[1] - p[1];
    float dz = q[2] - p[2];
    return (dx * dx + dy * dy) + dz * dz;
}

// Reference metric: KDTree.tcc:233-236 accumulates pow(point[i] - p[i], 2) in double
// with the query converted to double (KDTree.hpp:170-180).
inline double d2_f64(const double* q, const float* p)
{
    double acc = 0.0;
    for (int i = 0; i < 3; i++) {
        double d = q[i] - (double)p[i];
        acc += d * d;
    }
    return acc;
}

struct KeyC {  // contract: lexicographic (d2, index)
    float d2;
    uint32_t idx;
    bool operator<(const KeyC& o) const { return d2 < o.d2 || (d2 == o.d2 && idx < o.idx); }
};

struct KeyR {  // reference: KDTree.hpp:157-165, ordering on distance only
    double d2;
    uint32_t idx;
    bool operator<(const KeyR& o) const { return d2 < o.d2; }
};

/* ---- KD tree: KDTree.tcc:47-76 (build) and 186-248 (search) ------------------------ */

class KDTree {
public:
    KDTree(const float* xyz, int64_t n, size_t leaf = 20) : m_n(n)
    {
        m_pts.resize(n);
        for (int64_t i = 0; i < n; i++) {  // SearchKDTree ctor, KDTree.tcc:316-329
            m_pts[i].p[0] = xyz[3 * i];
            m_pts[i].p[1] = xyz[3 * i + 1];
            m_pts[i].p[2] = xyz[3 * i + 2];
            m_pts[i].index = (uint32_t)i;
        }
        double mn[3], mx[3];
        for (int j = 0; j < 3; j++) {
            mn[j] = std::numeric_limits<double>::max();
            mx[j] = -std::numeric_limits<double>::max();
        }
        for (int64_t i = 0; i < n; i++)
            for (int j = 0; j < 3; j++) {
                mn[j] = std::min(mn[j], (double)m_pts[i].p[j]);
                mx[j] = std::max(mx[j], (double)m_pts[i].p[j]);
            }
        m_nodes.reserve(n > 0 ? (size_t)(4 * n / leaf + 4) : 4);
        if (n > 0) {
            m_root = build(0, (size_t)n, mn, mx, leaf);
        }
    }

    // Contract search: exact k smallest (d2_f32, index).
    void knn_contract(const float* q, int k, std::vector<KeyC>& out) const
    {
        std::priority_queue<KeyC> heap;
        if (m_n > 0 && k > 0) search_c(m_root, q, (size_t)k, heap);
        out.resize(heap.size());
        for (size_t i = out.size(); i-- > 0;) {
            out[i] = heap.top();
            heap.pop();
        }
    }

    // Contract radius search (SearchKDTree::radiusSearch, KDTree.tcc:368-386): the k smallest
    // keys among the points with d2_f32 < r2.  The reference starts its worst distance at r^2
    // (KDTree.tcc:101-110) and admits a point only below it (KDTree.tcc:236); here the heap
    // admits keys below (r2, -) and the far side is pruned by the smaller of r2 and the worst.
    void radius_contract(const float* q, int k, float r2, std::vector<KeyC>& out) const
    {
        std::priority_queue<KeyC> heap;
        if (m_n > 0 && k > 0) search_cr(m_root, q, (size_t)k, r2, heap);
        out.resize(heap.size());
        for (size_t i = out.size(); i-- > 0;) {
            out[i] = heap.top();
            heap.pop();
        }
    }

    // Reference search: same traversal, KDTree.tcc:99-126 and 186-248.
    void knn_reference(const float* qf, int k, std::vector<KeyR>& out) const
    {
        double q[3] = {qf[0], qf[1], qf[2]};
        std::priority_queue<KeyR> heap;
        double worst = std::numeric_limits<double>::infinity();
        if (m_n > 0 && k > 0) search_r(m_root, q, (size_t)k, heap, worst);
        out.resize(heap.size());
        for (size_t i = out.size(); i-- > 0;) {
            out[i] = heap.top();
            heap.pop();
        }
    }

private:
    struct Node {
        int32_t lesser, greater;  // -1 for a leaf
        uint32_t axis;
        double split;
        size_t start, end;
    };

    int32_t build(size_t start, size_t end, const double* mn, const double* mx, size_t leaf)
    {
        size_t n = end - start;
        int32_t id = (int32_t)m_nodes.size();
        m_nodes.push_back(Node{-1, -1, 0, 0.0, start, end});
        if (n <= leaf) return id;
        unsigned axis = 0;
        double len = mx[0] - mn[0];
        for (unsigned i = 1; i < 3; i++)
            if (mx[i] - mn[i] > len) {
                axis = i;
                len = mx[i] - mn[i];
            }
        size_t mid = start + n / 2;
        std::nth_element(m_pts.begin() + start, m_pts.begin() + mid, m_pts.begin() + end,
                         [axis](const IdxPoint& a, const IdxPoint& b) { return a[axis] < b[axis]; });
        double split = m_pts[mid][axis];
        double lmx[3] = {mx[0], mx[1], mx[2]}, gmn[3] = {mn[0], mn[1], mn[2]};
        lmx[axis] = split;
        gmn[axis] = split;
        int32_t l = build(start, mid, mn, lmx, leaf);
        int32_t g = build(mid, end, gmn, mx, leaf);
        m_nodes[id].lesser = l;
        m_nodes[id].greater = g;
        m_nodes[id].axis = axis;
        m_nodes[id].split = split;
        return id;
    }

    void search_c(int32_t id, const float* q, size_t k, std::priority_queue<KeyC>& heap) const
    {
        const Node& nd = m_nodes[id];
        if (nd.lesser < 0) {
            for (size_t i = nd.start; i < nd.end; i++) {
                KeyC key{d2_f32(q, m_pts[i].p), m_pts[i].index};
                if (heap.size() < k) {
                    heap.push(key);
                } else if (key < heap.top()) {
                    heap.pop();
                    heap.push(key);
                }
            }
            return;
        }
        // child order as in KDNode::childOrder (KDTree.tcc:171-174)
        bool lfirst = (double)q[nd.axis] < nd.split;
        int32_t first = lfirst ? nd.lesser : nd.greater;
        int32_t second = lfirst ? nd.greater : nd.lesser;
        search_c(first, q, k, heap);
        // Every point of the far side lies on the other side of the split plane, so its
        // fp32 key is >= fl(fl(q - split)^2): rounding is monotone and the key only adds
        // non-negative terms.  Visit unless that bound strictly exceeds the current worst.
        float s = (float)nd.split;
        float dd = q[nd.axis] - s;
        float bound = dd * dd;
        if (heap.size() < k || !(bound > heap.top().d2)) search_c(second, q, k, heap);
    }

    void search_cr(int32_t id, const float* q, size_t k, float r2, std::priority_queue<KeyC>& heap) const
    {
        const Node& nd = m_nodes[id];
        if (nd.lesser < 0) {
            for (size_t i = nd.start; i < nd.end; i++) {
                KeyC key{d2_f32(q, m_pts[i].p), m_pts[i].index};
                if (!(key.d2 < r2)) continue;
                if (heap.size() < k) {
                    heap.push(key);
                } else if (key < heap.top()) {
                    heap.pop();
                    heap.push(key);
                }
            }
            return;
        }
        bool lfirst = (double)q[nd.axis] < nd.split;
        int32_t first = lfirst ? nd.lesser : nd.greater;
        int32_t second = lfirst ? nd.greater : nd.lesser;
        search_cr(first, q, k, r2, heap);
        float s = (float)nd.split;
        float dd = q[nd.axis] - s;
        float bound = dd * dd;
        if (bound < r2 && (heap.size() < k || !(bound > heap.top().d2))) search_cr(second, q, k, r2, heap);
    }

    void search_r(int32_t id, const double* q, size_t k, std::priority_queue<KeyR>& heap, double& worst) const
    {
        const Node& nd = m_nodes[id];
        if (nd.lesser < 0) {  // KDLeaf::knnInternal, KDTree.tcc:227-248
            for (size_t i = nd.start; i < nd.end; i++) {
                KeyR key{d2_f64(q, m_pts[i].p), m_pts[i].index};
                if (key.d2 < worst) {
                    heap.push(key);
                    if (heap.size() > k) {
                        heap.pop();
                        worst = heap.top().d2;
                    }
                }
            }
            return;
        }
        // KDNode::knnInternal, KDTree.tcc:186-195
        double cmp = (q[nd.axis] - nd.split) * (q[nd.axis] - nd.split);
        bool lfirst = q[nd.axis] < nd.split;
        search_r(lfirst ? nd.lesser : nd.greater, q, k, heap, worst);
        if (cmp < worst) search_r(lfirst ? nd.greater : nd.lesser, q, k, heap, worst);
    }

    int64_t m_n;
    std::vector<IdxPoint> m_pts;
    std::vector<Node> m_nodes;
    int32_t m_root = 0;
};

/* ---- small vector helpers (geometry/BaseVector.tcc:46-85, Normal.tcc:42-95) --------- */

struct V3 {
    float x, y, z;
};
inline float dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float len3(V3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
// BaseVector::normalize (BaseVector.hpp:130-143): the zero vector is left unchanged.
inline V3 normalized(V3 a)
{
    if (!(a.x == 0 && a.y == 0 && a.z == 0)) {
        float l = len3(a);
        a.x /= l;
        a.y /= l;
        a.z /= l;
    }
    return a;
}

/* ---- symmetric 3x3 eigen-decomposition (cyclic Jacobi) ------------------------------ */

template <typename T>
void jacobi3(const T* A, T* w, T* V)
{
    T a[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) a[i][j] = A[3 * i + j];
    T v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    const int pq[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    for (int sweep = 0; sweep < 64; sweep++) {
        T off = std::fabs(a[0][1]) + std::fabs(a[0][2]) + std::fabs(a[1][2]);
        if (off == T(0)) break;
        for (int r = 0; r < 3; r++) {
            int p = pq[r][0], q = pq[r][1];
            T apq = a[p][q];
            if (apq == T(0)) continue;
            // skip rotations that can no longer change the diagonal
            T g = T(100) * std::fabs(apq);
            if (sweep > 3 && std::fabs(a[p][p]) + g == std::fabs(a[p][p]) &&
                std::fabs(a[q][q]) + g == std::fabs(a[q][q])) {
                a[p][q] = a[q][p] = T(0);
                continue;
            }
            T theta = (a[q][q] - a[p][p]) / (T(2) * apq);
            T t;
            if (std::fabs(theta) > T(1e30))
                t = T(1) / (T(2) * theta);
            else
                t = (theta >= T(0) ? T(1) : T(-1)) / (std::fabs(theta) + std::sqrt(theta * theta + T(1)));
            T c = T(1) / std::sqrt(t * t + T(1));
            T s = t * c;
            // A' = J^T A J with J the Givens rotation in the (p,q) plane
            for (int k = 0; k < 3; k++) {
                T akp = a[k][p], akq = a[k][q];
                a[k][p] = c * akp - s * akq;
                a[k][q] = s * akp + c * akq;
            }
            for (int k = 0; k < 3; k++) {
                T apk = a[p][k], aqk = a[q][k];
                a[p][k] = c * apk - s * aqk;
                a[q][k] = s * apk + c * aqk;
            }
            a[p][q] = a[q][p] = T(0);
            for (int k = 0; k < 3; k++) {
                T vkp = v[k][p], vkq = v[k][q];
                v[k][p] = c * vkp - s * vkq;
                v[k][q] = s * vkp + c * vkq;
            }
        }
    }
    for (int i = 0; i < 3; i++) {
        w[i] = a[i][i];
        for (int j = 0; j < 3; j++) V[3 * i + j] = v[i][j];
    }
}

// Column choice of calcPlaneIPCAExact (AdaptiveKSearchSurface.tcc:647-659).
template <typename T>
int smallest_col(const T* w)
{
    if (w[0] < w[1] && w[0] < w[2]) return 0;
    if (w[1] < w[2]) return 1;
    return 2;
}

/* ---- plane fits --------------------------------------------------------------------- */

// calcPlaneIPCAExact, FP32 (AdaptiveKSearchSurface.tcc:609-673).  Eigen's float*double
// products promote the double weight to float, which is reproduced here.
V3 plane_ipca_f32(const float* pts, V3 q, const std::vector<uint32_t>& ids)
{
    float cov[9] = {0};
    float mean[3] = {q.x, q.y, q.z};
    double n_meas = 1.0;
    for (uint32_t id : ids) {
        const float* p = pts + 3 * (size_t)id;
        const double n_new = n_meas + 1.0;
        const float w1 = (float)(n_meas / n_new);
        const float w2 = (float)(1.0 / n_new);
        float mn[3], a[3], b[3];
        for (int i = 0; i < 3; i++) mn[i] = mean[i] * w1 + p[i] * w2;
        for (int i = 0; i < 3; i++) {
            a[i] = mean[i] - mn[i];
            b[i] = p[i] - mn[i];
        }
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                float P1 = a[i] * a[j];
                float P2 = b[i] * b[j];
                cov[3 * i + j] = (cov[3 * i + j] + P1) * w1 + P2 * w2;
            }
        n_meas = n_new;
        for (int i = 0; i < 3; i++) mean[i] = mn[i];
    }
    float w[3], V[9];
    jacobi3<float>(cov, w, V);
    int c = smallest_col(w);
    float vx = V[c], vy = V[3 + c], vz = V[6 + c];
    // Eigen normalises eigenvectors: v / sqrt((x^2 + y^2) + z^2)
    float nrm = std::sqrt((vx * vx + vy * vy) + vz * vz);
    return V3{vx / nrm, vy / nrm, vz / nrm};
}

// Same neighbourhood, FP64: the covariance of the multiset {q, p_ids...} (the IPCA seed
// counts q once more), which is what the IPCA recursion computes in exact arithmetic.
void plane_ipca_f64(const float* pts, V3 q, const std::vector<uint32_t>& ids, double* n, double* gap = nullptr)
{
    const double qd[3] = {q.x, q.y, q.z};
    double m[3] = {0, 0, 0};
    const double cnt = (double)ids.size() + 1.0;
    for (uint32_t id : ids)
        for (int i = 0; i < 3; i++) m[i] += (double)pts[3 * (size_t)id + i] - qd[i];
    for (int i = 0; i < 3; i++) m[i] /= cnt;  // mean offset from q (q itself adds 0)
    double cov[9] = {0};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) cov[3 * i + j] = m[i] * m[j];  // the seed sample q
    for (uint32_t id : ids) {
        double r[3];
        for (int i = 0; i < 3; i++) r[i] = (double)pts[3 * (size_t)id + i] - qd[i] - m[i];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) cov[3 * i + j] += r[i] * r[j];
    }
    for (int i = 0; i < 9; i++) cov[i] /= cnt;
    double w[3], V[9];
    jacobi3<double>(cov, w, V);
    int c = smallest_col(w);
    double vx = V[c], vy = V[3 + c], vz = V[6 + c];
    double nrm = std::sqrt(vx * vx + vy * vy + vz * vz);
    n[0] = vx / nrm;
    n[1] = vy / nrm;
    n[2] = vz / nrm;
    if (gap) {  // (lambda_mid - lambda_min) / lambda_max
        double s[3] = {w[0], w[1], w[2]};
        std::sort(s, s + 3);
        *gap = s[2] > 0 ? (s[1] - s[0]) / s[2] : 0.0;
    }
}

// calcPlaneIterative (AdaptiveKSearchSurface.tcc:522-607), templated on precision.
template <typename T>
void plane_iterative(const float* pts, V3 q, const std::vector<uint32_t>& ids, T* n)
{
    T xx = 0, xy = 0, xz = 0, yy = 0, yz = 0, zz = 0;
    for (uint32_t id : ids) {
        const float* p = pts + 3 * (size_t)id;
        T rx, ry, rz;
        if (sizeof(T) == 4) {
            rx = (T)(p[0] - q.x);
            ry = (T)(p[1] - q.y);
            rz = (T)(p[2] - q.z);
        } else {
            rx = (T)p[0] - (T)q.x;
            ry = (T)p[1] - (T)q.y;
            rz = (T)p[2] - (T)q.z;
        }
        xx += rx * rx;
        xy += rx * ry;
        xz += rx * rz;
        yy += ry * ry;
        yz += ry * rz;
        zz += rz * rz;
    }
    T det_x = yy * zz - yz * yz;
    T det_y = xx * zz - xz * xz;
    T det_z = xx * yy - xy * xy;
    T dx, dy, dz;
    if (det_x >= det_y && det_x >= det_z) {
        dx = 1;
        dy = (xz * yz - xy * zz) / det_x;
        dz = (xy * yz - xz * yy) / det_x;
    } else if (det_y >= det_x && det_y >= det_z) {
        dx = (yz * xz - xy * zz) / det_y;
        dy = 1;
        dz = (xy * xz - yz * xx) / det_y;
    } else {
        dx = (yz * xy - xz * yy) / det_z;
        dy = (xz * xy - yz * xx) / det_z;
        dz = 1;
    }
    T inv = T(1) / std::sqrt(dx * dx + dy * dy + dz * dz);
    n[0] = dx * inv;
    n[1] = dy * inv;
    n[2] = dz * inv;
}

// calcPlane, method 0 (AdaptiveKSearchSurface.tcc:467-520): least squares
// y = c0 + c1*x + c2*z, then the normal of the two tangents spanned 100 units away
// from the query.  The 3x3 normal equations are solved in FP64 (substitute for
// JacobiSVD); the tangent/cross/normalise tail runs in T.
template <typename T>
void plane_ls(const float* pts, V3 q, const std::vector<uint32_t>& ids, T* n)
{
    // centred at q for conditioning; c0 is re-expressed in absolute coordinates below
    double s1 = 0, sx = 0, sz = 0, sxx = 0, sxz = 0, szz = 0, sy = 0, sxy = 0, szy = 0;
    for (uint32_t id : ids) {
        const float* p = pts + 3 * (size_t)id;
        double x = (double)p[0] - q.x, y = (double)p[1] - q.y, z = (double)p[2] - q.z;
        s1 += 1;
        sx += x;
        sz += z;
        sxx += x * x;
        sxz += x * z;
        szz += z * z;
        sy += y;
        sxy += x * y;
        szy += z * y;
    }
    double A[3][3] = {{s1, sx, sz}, {sx, sxx, sxz}, {sz, sxz, szz}};
    double b[3] = {sy, sxy, szy};
    // Gaussian elimination with partial pivoting
    int perm[3] = {0, 1, 2};
    for (int c = 0; c < 3; c++) {
        int piv = c;
        for (int r = c + 1; r < 3; r++)
            if (std::fabs(A[r][c]) > std::fabs(A[piv][c])) piv = r;
        std::swap(A[c], A[piv]);
        std::swap(b[c], b[piv]);
        std::swap(perm[c], perm[piv]);
        if (A[c][c] == 0) continue;
        for (int r = c + 1; r < 3; r++) {
            double f = A[r][c] / A[c][c];
            for (int k = c; k < 3; k++) A[r][k] -= f * A[c][k];
            b[r] -= f * b[c];
        }
    }
    double x[3] = {0, 0, 0};
    for (int r = 2; r >= 0; r--) {
        double s = b[r];
        for (int k = r + 1; k < 3; k++) s -= A[r][k] * x[k];
        x[r] = A[r][r] != 0 ? s / A[r][r] : 0.0;
    }
    // back to absolute coordinates: y = qy + c0' + c1 (x - qx) + c2 (z - qz)
    const double c1d = x[1], c2d = x[2];
    const double c0d = (double)q.y + x[0] - c1d * q.x - c2d * q.z;
    const T eps = 100;
    T C0 = (T)c0d, C1 = (T)c1d, C2 = (T)c2d;
    T qx = q.x, qy = q.y, qz = q.z;
    T z1 = C0 + C1 * (qx + eps) + C2 * qz;
    T z2 = C0 + C1 * qx + C2 * (qz + eps);
    T d1[3] = {(qx + eps) - qx, z1 - qy, qz - qz};
    T d2[3] = {qx - qx, z2 - qy, (qz + eps) - qz};
    T cx = d1[1] * d2[2] - d1[2] * d2[1];
    T cy = d1[2] * d2[0] - d1[0] * d2[2];
    T cz = d1[0] * d2[1] - d1[1] * d2[0];
    if (!(cx == 0 && cy == 0 && cz == 0)) {
        T l = std::sqrt(cx * cx + cy * cy + cz * cz);
        cx /= l;
        cy /= l;
        cz /= l;
    }
    n[0] = cx;
    n[1] = cy;
    n[2] = cz;
}

// boundingBoxOK (AdaptiveKSearchSurface.tcc:337-350) over the FP32 bbox of the ids.
bool bbox_ok(const float* pts, const std::vector<uint32_t>& ids)
{
    float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
    float mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (uint32_t id : ids)
        for (int i = 0; i < 3; i++) {
            mn[i] = std::min(pts[3 * (size_t)id + i], mn[i]);
            mx[i] = std::max(pts[3 * (size_t)id + i], mx[i]);
        }
    float dx = mx[0] - mn[0], dy = mx[1] - mn[1], dz = mx[2] - mn[2];
    float e = 0.05f;
    return dx >= e * dy && dx >= e * dz && dy >= e * dx && dy >= e * dz && dz >= e * dx && dz >= e * dy;
}

}  // namespace

struct orc_surface {
    std::vector<float> pts;
    int64_t n;
    int metric;
    float bbox[6];
    std::unique_ptr<KDTree> tree;

    // kSearch returning ids only, in ascending order
    void ksearch(const float* q, int k, std::vector<uint32_t>& ids) const
    {
        ids.clear();
        if (metric == ORC_METRIC_CONTRACT) {
            std::vector<KeyC> keys;
            tree->knn_contract(q, k, keys);
            for (auto& c : keys) ids.push_back(c.idx);
        } else {
            std::vector<KeyR> keys;
            tree->knn_reference(q, k, keys);
            for (auto& c : keys) ids.push_back(c.idx);
        }
    }
};

extern "C" {

orc_surface* orc_surface_create(const float* xyz, int64_t n, int metric)
{
    auto* s = new orc_surface;
    s->pts.assign(xyz, xyz + 3 * n);
    s->n = n;
    s->metric = metric;
    // PointsetSurface ctor: BoundingBox::expand over all points (PointsetSurface.tcc:52-56)
    float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int64_t i = 0; i < n; i++)
        for (int j = 0; j < 3; j++) {
            mn[j] = std::min(xyz[3 * i + j], mn[j]);
            mx[j] = std::max(xyz[3 * i + j], mx[j]);
        }
    for (int j = 0; j < 3; j++) {
        s->bbox[j] = mn[j];
        s->bbox[3 + j] = mx[j];
    }
    s->tree.reset(new KDTree(s->pts.data(), n));
    return s;
}

void orc_surface_destroy(orc_surface* s) { delete s; }

void orc_surface_bbox(const orc_surface* s, float* out6) { std::memcpy(out6, s->bbox, sizeof(s->bbox)); }

int orc_knn(const orc_surface* s, const float* q, int64_t m, int k, int32_t* out_idx, float* out_d2)
{
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < m; i++) {
        const float* qi = q + 3 * i;
        size_t got = 0;
        if (s->metric == ORC_METRIC_CONTRACT) {
            std::vector<KeyC> keys;
            s->tree->knn_contract(qi, k, keys);
            got = keys.size();
            for (size_t j = 0; j < got; j++) {
                out_idx[i * k + j] = (int32_t)keys[j].idx;
                if (out_d2) out_d2[i * k + j] = keys[j].d2;
            }
        } else {
            std::vector<KeyR> keys;
            s->tree->knn_reference(qi, k, keys);
            got = keys.size();
            for (size_t j = 0; j < got; j++) {
                out_idx[i * k + j] = (int32_t)keys[j].idx;
                if (out_d2) out_d2[i * k + j] = d2_f32(qi, &s->pts[3 * (size_t)keys[j].idx]);
            }
        }
        for (size_t j = got; j < (size_t)k; j++) {
            out_idx[i * k + j] = -1;
            if (out_d2) out_d2[i * k + j] = INFINITY;
        }
    }
    return 0;
}

int orc_radius(const orc_surface* s, const float* q, int64_t m, int k, float r, int32_t* out_idx, float* out_d2,
               int32_t* out_count)
{
    const float r2 = r * r;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < m; i++) {
        std::vector<KeyC> keys;
        s->tree->radius_contract(q + 3 * i, k, r2, keys);
        const size_t got = keys.size();
        for (size_t j = 0; j < (size_t)k; j++) {
            out_idx[i * k + j] = j < got ? (int32_t)keys[j].idx : -1;
            if (out_d2) out_d2[i * k + j] = j < got ? keys[j].d2 : INFINITY;
        }
        if (out_count) out_count[i] = (int32_t)got;
    }
    return 0;
}

int orc_knn_brute(const float* xyz, int64_t n, const float* q, int64_t m, int k, int32_t* out_idx, float* out_d2)
{
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < m; i++) {
        std::vector<KeyC> keys((size_t)n);
        for (int64_t j = 0; j < n; j++) keys[j] = KeyC{d2_f32(q + 3 * i, xyz + 3 * j), (uint32_t)j};
        size_t kk = std::min<size_t>((size_t)k, (size_t)n);
        std::partial_sort(keys.begin(), keys.begin() + kk, keys.end());
        for (size_t j = 0; j < (size_t)k; j++) {
            out_idx[i * k + j] = j < kk ? (int32_t)keys[j].idx : -1;
            if (out_d2) out_d2[i * k + j] = j < kk ? keys[j].d2 : INFINITY;
        }
    }
    return 0;
}

int orc_normals(const orc_surface* s, int kn, int method, const float flip[3], const int64_t* subset, int64_t nsub,
                float* out_n32, double* out_n64, int32_t* out_k, double* out_gap)
{
    const float* pts = s->pts.data();
    const int64_t cnt = subset ? nsub : s->n;
    const V3 fp{flip[0], flip[1], flip[2]};
    // OMP dynamic loop over points (AdaptiveKSearchSurface.tcc:178)
#pragma omp parallel for schedule(dynamic)
    for (int64_t t = 0; t < cnt; t++) {
        const int64_t i = subset ? subset[t] : t;
        const float* qp = pts + 3 * i;
        const V3 q{qp[0], qp[1], qp[2]};
        std::vector<uint32_t> ids;
        int n = 0;
        size_t k = (size_t)kn;
        while (n < 5) {  // adaptive neighbourhood, tcc:187-216
            n++;
            k = k * 2;
            s->ksearch(qp, (int)k, ids);
            if (bbox_ok(pts, ids)) break;
        }
        V3 nrm;
        if (method == 2) {
            float v[3];
            plane_iterative<float>(pts, q, ids, v);
            nrm = V3{v[0], v[1], v[2]};
        } else if (method == 3) {
            nrm = plane_ipca_f32(pts, q, ids);
        } else {
            float v[3];
            plane_ls<float>(pts, q, ids, v);
            nrm = V3{v[0], v[1], v[2]};
        }
        // flip towards the flip point (tcc:268-274); Normal::operator- renormalises
        V3 d{fp.x - q.x, fp.y - q.y, fp.z - q.z};
        if (dot3(nrm, d) < 0) nrm = normalized(V3{-nrm.x, -nrm.y, -nrm.z});
        out_n32[3 * t] = nrm.x;
        out_n32[3 * t + 1] = nrm.y;
        out_n32[3 * t + 2] = nrm.z;
        if (out_n64) {
            double v[3];
            if (method == 2)
                plane_iterative<double>(pts, q, ids, v);
            else if (method == 3)
                plane_ipca_f64(pts, q, ids, v, out_gap ? out_gap + t : nullptr);
            else
                plane_ls<double>(pts, q, ids, v);
            double dd = v[0] * ((double)fp.x - q.x) + v[1] * ((double)fp.y - q.y) + v[2] * ((double)fp.z - q.z);
            double sgn = dd < 0 ? -1.0 : 1.0;
            for (int j = 0; j < 3; j++) out_n64[3 * t + j] = sgn * v[j];
        }
        if (out_k) out_k[t] = (int32_t)k;
    }
    return 0;
}

int orc_interpolate(const orc_surface* s, const float* nin, int ki, const int64_t* subset, int64_t nsub, float* out)
{
    const float* pts = s->pts.data();
    const int64_t cnt = subset ? nsub : s->n;
    // AdaptiveKSearchSurface::interpolateSurfaceNormals, tcc:312-327
#pragma omp parallel for schedule(dynamic)
    for (int64_t t = 0; t < cnt; t++) {
        const int64_t i = subset ? subset[t] : t;
        std::vector<uint32_t> ids;
        s->ksearch(pts + 3 * i, ki, ids);
        V3 mean{nin[3 * i], nin[3 * i + 1], nin[3 * i + 2]};
        for (uint32_t j : ids) {
            mean.x += nin[3 * (size_t)j];
            mean.y += nin[3 * (size_t)j + 1];
            mean.z += nin[3 * (size_t)j + 2];
        }
        V3 r = normalized(mean);
        out[3 * t] = r.x;
        out[3 * t + 1] = r.y;
        out[3 * t + 2] = r.z;
    }
    return 0;
}

int orc_interpolate_f64(const orc_surface* s, const double* nin, int ki, const int64_t* subset, int64_t nsub,
                        double* out)
{
    const float* pts = s->pts.data();
    const int64_t cnt = subset ? nsub : s->n;
    // interpolateSurfaceNormals (tcc:312-327) on the same neighbourhoods, every sum in fp64
#pragma omp parallel for schedule(dynamic)
    for (int64_t t = 0; t < cnt; t++) {
        const int64_t i = subset ? subset[t] : t;
        std::vector<uint32_t> ids;
        s->ksearch(pts + 3 * i, ki, ids);
        double m[3] = {nin[3 * i], nin[3 * i + 1], nin[3 * i + 2]};
        for (uint32_t j : ids)
            for (int c = 0; c < 3; c++) m[c] += nin[3 * (size_t)j + c];
        const double l = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
        for (int c = 0; c < 3; c++) out[3 * t + c] = l > 0 ? m[c] / l : m[c];
    }
    return 0;
}

int orc_distance_f64(const orc_surface* s, const double* normals, int kd, const float* q, int64_t m,
                     double* out_proj, double* out_eucl)
{
    const float* pts = s->pts.data();
    // AdaptiveKSearchSurface::distance (tcc:414-457) on the same neighbourhoods, in fp64
#pragma omp parallel for schedule(dynamic)
    for (int64_t i = 0; i < m; i++) {
        std::vector<uint32_t> ids;
        s->ksearch(q + 3 * i, kd, ids);
        if (ids.empty()) {
            out_proj[i] = out_eucl[i] = FLT_MAX;
            continue;
        }
        double c[3] = {0, 0, 0}, a[3] = {0, 0, 0};
        for (uint32_t j : ids)
            for (int k = 0; k < 3; k++) {
                c[k] += pts[3 * (size_t)j + k];
                a[k] += normals[3 * (size_t)j + k];
            }
        const double cnt = (double)ids.size();
        for (int k = 0; k < 3; k++) {
            c[k] /= cnt;
            a[k] /= cnt;
        }
        const double l = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
        double d[3], pr = 0, e = 0;
        for (int k = 0; k < 3; k++) {
            d[k] = (double)q[3 * i + k] - c[k];
            pr += d[k] * (l > 0 ? a[k] / l : a[k]);
            e += d[k] * d[k];
        }
        out_proj[i] = pr;
        out_eucl[i] = std::sqrt(e);
    }
    return 0;
}

int orc_interpolated_normal(const orc_surface* s, const float* normals, int ki, const float* q, int64_t m,
                            float* out)
{
    // PointsetSurface::getInterpolatedNormal (PointsetSurface.tcc:61-74): result is a
    // default Normal (0, 1, 0) (Normal.hpp:62); each neighbour's normal is read through
    // Normal(x, y, z), which normalises (Normal.tcc:49-54); BaseVector::operator+= and /= add
    // and divide component-wise; the returned copy is not normalised again.
#pragma omp parallel for schedule(dynamic)
    for (int64_t i = 0; i < m; i++) {
        std::vector<uint32_t> ids;
        s->ksearch(q + 3 * i, ki, ids);
        V3 r{0.0f, 1.0f, 0.0f};
        for (uint32_t j : ids) {
            const V3 n = normalized(V3{normals[3 * (size_t)j], normals[3 * (size_t)j + 1], normals[3 * (size_t)j + 2]});
            r.x += n.x;
            r.y += n.y;
            r.z += n.z;
        }
        const float c = (float)ki;
        out[3 * i] = r.x / c;
        out[3 * i + 1] = r.y / c;
        out[3 * i + 2] = r.z / c;
    }
    return 0;
}

int orc_distance(const orc_surface* s, const float* normals, int kd, const float* q, int64_t m, float* out_proj,
                 float* out_eucl)
{
    const float* pts = s->pts.data();
    // AdaptiveKSearchSurface::distance, tcc:414-457 (called from the OMP loop of
    // PointsetGrid::calcDistanceValues, PointsetGrid.tcc:123)
#pragma omp parallel for schedule(dynamic)
    for (int64_t i = 0; i < m; i++) {
        std::vector<uint32_t> ids;
        s->ksearch(q + 3 * i, kd, ids);
        if (ids.empty()) {
            out_proj[i] = out_eucl[i] = FLT_MAX;
            continue;
        }
        V3 nearest{0, 0, 0}, avg{0, 0, 0};
        for (uint32_t j : ids) {
            nearest.x += pts[3 * (size_t)j];
            nearest.y += pts[3 * (size_t)j + 1];
            nearest.z += pts[3 * (size_t)j + 2];
            avg.x += normals[3 * (size_t)j];
            avg.y += normals[3 * (size_t)j + 1];
            avg.z += normals[3 * (size_t)j + 2];
        }
        const float c = (float)ids.size();
        avg.x /= c;
        avg.y /= c;
        avg.z /= c;
        nearest.x /= c;
        nearest.y /= c;
        nearest.z /= c;
        V3 nrm = normalized(avg);
        V3 d{q[3 * i] - nearest.x, q[3 * i + 1] - nearest.y, q[3 * i + 2] - nearest.z};
        out_proj[i] = dot3(d, nrm);
        out_eucl[i] = len3(d);
    }
    return 0;
}

int orc_grid(const float* xyz, int64_t n, const float bb[6], float voxel, int extrude, int64_t* nv, int64_t* nc,
             float* out_vtx, int32_t* out_cells, int32_t* out_cell_idx)
{
    struct I3 {
        int x, y, z;
    };
    auto contains = [&](float cx, float cy, float cz) {  // BoundingBox::contains, BoundingBox.tcc:166-171
        return bb[0] <= cx && bb[3] >= cx && bb[1] <= cy && bb[4] >= cy && bb[2] <= cz && bb[5] >= cz;
    };
    auto center = [&](int i) { return (i + 0.5f) * voxel; };  // HashGrid.hpp:283-287
    std::vector<int64_t> keys;
    auto pack = [](int x, int y, int z) {
        return ((int64_t)(x + (1 << 20)) << 42) | ((int64_t)(y + (1 << 20)) << 21) | (int64_t)(z + (1 << 20));
    };
    auto unpack = [](int64_t k) {
        return I3{(int)(k >> 42) - (1 << 20), (int)((k >> 21) & ((1 << 21) - 1)) - (1 << 20),
                  (int)(k & ((1 << 21) - 1)) - (1 << 20)};
    };
    // cells touched by points (PointsetGrid.tcc:61-75); calcIndex = floor(p / voxel)
    for (int64_t i = 0; i < n; i++) {
        int ix = (int)std::floor(xyz[3 * i] / voxel);
        int iy = (int)std::floor(xyz[3 * i + 1] / voxel);
        int iz = (int)std::floor(xyz[3 * i + 2] / voxel);
        if (contains(center(ix), center(iy), center(iz))) keys.push_back(pack(ix, iy, iz));
    }
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    if (extrude) {  // PointsetGrid.tcc:77-100
        std::vector<int64_t> inner = keys;
        for (int64_t k : inner) {
            I3 c = unpack(k);
            for (int dx = -1; dx <= 1; dx++)
                for (int dy = -1; dy <= 1; dy++)
                    for (int dz = -1; dz <= 1; dz++)
                        if (contains(center(c.x + dx), center(c.y + dy), center(c.z + dz)))
                            keys.push_back(pack(c.x + dx, c.y + dy, c.z + dz));
        }
        std::sort(keys.begin(), keys.end());
        keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    }
    // addLatticePoints -> addBox (HashGrid.tcc:235-259, 333-345): vertices shared
    // between cells are created once, at the lattice corner; box_creation_table
    // (FastReconstructionTables.hpp:86-95) gives the corner offsets.
    static const int bct[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                  {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
    std::unordered_map<int64_t, int32_t> vmap;
    vmap.reserve(keys.size() * 2);
    std::vector<float> vtx;
    std::vector<int32_t> cells;
    const float half = voxel / 2.0f;
    for (int64_t k : keys) {
        I3 c = unpack(k);
        float cx = center(c.x), cy = center(c.y), cz = center(c.z);
        for (int v = 0; v < 8; v++) {
            int lx = c.x + (bct[v][0] + 1) / 2, ly = c.y + (bct[v][1] + 1) / 2, lz = c.z + (bct[v][2] + 1) / 2;
            int64_t lk = pack(lx, ly, lz);
            auto it = vmap.find(lk);
            int32_t id;
            if (it == vmap.end()) {
                id = (int32_t)(vtx.size() / 3);
                vmap.emplace(lk, id);
                vtx.push_back(cx + (float)bct[v][0] * half);
                vtx.push_back(cy + (float)bct[v][1] * half);
                vtx.push_back(cz + (float)bct[v][2] * half);
            } else {
                id = it->second;
            }
            cells.push_back(id);
        }
    }
    *nv = (int64_t)(vtx.size() / 3);
    *nc = (int64_t)keys.size();
    if (out_vtx) std::memcpy(out_vtx, vtx.data(), vtx.size() * sizeof(float));
    if (out_cells) std::memcpy(out_cells, cells.data(), cells.size() * sizeof(int32_t));
    if (out_cell_idx)
        for (size_t i = 0; i < keys.size(); i++) {
            I3 c = unpack(keys[i]);
            out_cell_idx[3 * i] = c.x;
            out_cell_idx[3 * i + 1] = c.y;
            out_cell_idx[3 * i + 2] = c.z;
        }
    return 0;
}

int orc_mesh(const float* vtx, int64_t nv, const int32_t* cells, int64_t nc, const float* dist,
             const uint8_t* invalid, int pmc, int64_t cap_v, int64_t cap_f, int64_t* out_nv, int64_t* out_nf,
             float* out_vtx, int32_t* out_faces, int32_t* out_face_cell)
{
    // FastReconstruction::getMesh over FastBox cells (FastReconstruction.tcc:47-65), the cells in
    // increasing index order (the lattice's order; lvr2 iterates an unordered map), after the
    // erase of the cells with an invalid corner (PointsetGrid.tcc:153-184).
    static const uint64_t table[256] = {LVRHIP_MC_TABLE};
    // lvr2's edge numbering (FastBox::getIntersections, FastBox.tcc:159-202): the two corners
    // along the edge, low then high, and the corner the other two coordinates are read from
    static const int edge_lo[12] = {0, 1, 3, 0, 4, 5, 7, 4, 0, 1, 3, 2};
    static const int edge_hi[12] = {1, 2, 2, 3, 5, 6, 6, 7, 4, 5, 7, 6};
    static const int edge_axis[12] = {0, 1, 0, 1, 0, 1, 0, 1, 2, 2, 2, 2};
    static const int edge_src[12] = {0, 1, 2, 3, 4, 5, 6, 7, 0, 1, 3, 2};
    // a vertex created by one cell is handed to the cells that share its lattice edge
    // (FastBox.tcc:236-253 through neighbor_table / neighbor_vertex_table): keyed here by the
    // edge's low lattice vertex and axis
    std::unordered_map<uint64_t, int32_t> shared;
    int64_t nverts = 0, nfaces = 0;
    for (int64_t c = 0; c < nc; c++) {
        const int32_t* cv = cells + 8 * c;
        bool bad = false;
        for (int v = 0; v < 8; v++) bad = bad || (invalid && invalid[cv[v]]);
        if (bad) continue;  // erased (PointsetGrid.tcc:153-184) / FastBox::isInvalid
        float corner[8][3], d[8];
        for (int v = 0; v < 8; v++) {
            for (int a = 0; a < 3; a++) corner[v][a] = vtx[3 * (size_t)cv[v] + a];
            d[v] = dist[cv[v]];
        }
        int index = 0;  // FastBox::getIndex
        for (int v = 0; v < 8; v++)
            if (d[v] > 0) index |= 1 << v;
        // the cell's twelve edge positions (FastBox::getIntersections)
        float pos[12][3];
        for (int e = 0; e < 12; e++) {
            const int lo = edge_lo[e], hi = edge_hi[e], ax = edge_axis[e];
            // FastBox::calcIntersection (FastBox.tcc:140-160)
            const float x1 = corner[lo][ax], x2 = corner[hi][ax], d1 = d[lo], d2 = d[hi];
            float x;
            if (d1 == 0)
                x = x1;
            else if (d2 == 0)
                x = x2;
            else if ((d1 < 0 && d2 > 0) || (d2 < 0 && d1 > 0))
                x = x2 - d2 * (x1 - x2) / (d1 - d2);
            else
                x = 0;
            for (int a = 0; a < 3; a++) pos[e][a] = corner[edge_src[e]][a];
            pos[e][ax] = x;
        }
        const uint64_t w = table[index];
        const int ntri = (int)(w >> 60);
        for (int t = 0; t < ntri; t++) {
            int es[3];
            for (int b = 0; b < 3; b++) es[b] = (int)((w >> (4 * (3 * t + b))) & 0xF);
            if (pmc) {
                // BilinearFastBox::getSurface (BilinearFastBox.tcc:83-97): a face with two equal
                // corner positions is skipped before any of its vertices is created
                auto same = [&](int a, int b) {
                    return pos[es[a]][0] == pos[es[b]][0] && pos[es[a]][1] == pos[es[b]][1] &&
                           pos[es[a]][2] == pos[es[b]][2];
                };
                if (same(0, 1) || same(0, 2) || same(2, 1)) continue;
            }
            int32_t tri[3];
            for (int b = 0; b < 3; b++) {
                const int e = es[b];
                const uint64_t key = (uint64_t)(uint32_t)cv[edge_lo[e]] * 3 + (uint64_t)edge_axis[e];
                auto it = shared.find(key);
                if (it == shared.end()) {
                    if (nverts < cap_v)
                        for (int a = 0; a < 3; a++) out_vtx[3 * nverts + a] = pos[e][a];
                    it = shared.emplace(key, (int32_t)nverts++).first;
                }
                tri[b] = it->second;
            }
            if (nfaces < cap_f) {
                for (int b = 0; b < 3; b++) out_faces[3 * nfaces + b] = tri[b];
                if (out_face_cell) out_face_cell[nfaces] = (int32_t)c;
            }
            nfaces++;
        }
    }
    *out_nv = nverts;
    *out_nf = nfaces;
    return (nverts <= cap_v && nfaces <= cap_f) ? 0 : 1;
}

int orc_pmc_optimize(const orc_surface* s, int kc, float* vtx, int64_t nv, const int32_t* faces,
                     const int32_t* face_cell, int64_t nf, int64_t* out_moves, int64_t* out_nonmanifold)
{
    // FastReconstruction::getMesh's BilinearFastBox tail (FastReconstruction.tcc:155-165) calling
    // BilinearFastBox::optimizePlanarFaces(mesh, 5) (BilinearFastBox.tcc:140-212) on every cell,
    // in cell order, literally: a cell whose own faces have one or two border edges (edges with a
    // face on one side only, HalfEdgeMesh::isBorderEdge, HalfEdgeMesh.tcc:732-746) moves both
    // end points of each such edge, one after the other, to the centroid of the kc points nearest
    // to the end point's current position.
    auto ukey = [](int32_t a, int32_t b) {
        return a < b ? ((uint64_t)(uint32_t)a << 32) | (uint32_t)b : ((uint64_t)(uint32_t)b << 32) | (uint32_t)a;
    };
    std::unordered_map<uint64_t, int> faces_of, half_edges;
    for (int64_t f = 0; f < nf; f++)
        for (int b = 0; b < 3; b++) {
            const int32_t u = faces[3 * f + b], v = faces[3 * f + (b + 1) % 3];
            faces_of[ukey(u, v)]++;
            half_edges[((uint64_t)(uint32_t)u << 32) | (uint32_t)v]++;
        }
    // configurations on which HalfEdgeMesh::addFace panics (an edge with more than two faces, or
    // a half-edge used twice): counted so that a test can require there are none
    int64_t bad = 0;
    for (auto& e : faces_of) bad += e.second > 2;
    for (auto& e : half_edges) bad += e.second > 1;
    if (out_nonmanifold) *out_nonmanifold = bad;
    int64_t moves = 0;
    std::vector<uint32_t> ids;
    for (int64_t f0 = 0; f0 < nf;) {
        int64_t f1 = f0;
        while (f1 < nf && face_cell[f1] == face_cell[f0]) f1++;
        std::vector<std::pair<int32_t, int32_t>> border;
        for (int64_t f = f0; f < f1; f++)
            for (int b = 0; b < 3; b++) {
                const int32_t u = faces[3 * f + b], v = faces[3 * f + (b + 1) % 3];
                if (faces_of[ukey(u, v)] == 1) border.emplace_back(u, v);
            }
        if (border.size() == 1 || border.size() == 2) {
            for (auto& e : border)
                for (int32_t vid : {e.first, e.second}) {
                    float* p = vtx + 3 * (size_t)vid;
                    s->ksearch(p, kc, ids);
                    const size_t nk = std::min((size_t)kc, ids.size());
                    if (nk == 0) continue;
                    float c[3] = {0.0f, 0.0f, 0.0f};
                    for (uint32_t j : ids)
                        for (int a = 0; a < 3; a++) c[a] += s->pts[3 * (size_t)j + a];
                    for (int a = 0; a < 3; a++) p[a] = c[a] / (float)nk;
                    moves++;
                }
        }
        f0 = f1;
    }
    if (out_moves) *out_moves = moves;
    (void)nv;
    return 0;
}

int orc_calc_distance_values(const float* eucl, int64_t nv, float voxel, const int32_t* cells, int64_t nc,
                             uint8_t* out_invalid, uint8_t* out_keep)
{
    // PointsetGrid.tcc:138-143: the comparison is done in double (1.75 is a double)
    for (int64_t i = 0; i < nv; i++) out_invalid[i] = ((double)eucl[i] > 1.75 * (double)voxel) ? 1 : 0;
    // PointsetGrid.tcc:153-184: erase cells with any invalid corner
    for (int64_t c = 0; c < nc; c++) {
        bool bad = false;
        for (int v = 0; v < 8; v++) bad = bad || out_invalid[cells[8 * c + v]];
        out_keep[c] = bad ? 0 : 1;
    }
    return 0;
}

void orc_eig3_f32(const float* cov, float* evals, float* evecs) { jacobi3<float>(cov, evals, evecs); }
void orc_eig3_f64(const double* cov, double* evals, double* evecs) { jacobi3<double>(cov, evals, evecs); }

int orc_num_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
