"""Writes the golden fixtures that pin the oracle (run from the repo root:
``python tests/golden/make_golden.py``).

Everything here is computed independently of the oracle, with numpy:
  knn_golden.npz    brute-force kNN under the contract ordering: fp32 d^2 evaluated as
                    (dx*dx + dy*dy) + dz*dz (numpy float32 arithmetic rounds every
                    operation, no FMA), ties broken by point index; clouds with exact
                    duplicates and integer lattices so ties occur.
  planes_golden.npz tiny clouds (21 points) whose whole extent is one neighbourhood, with
                    the FP64 plane normals of lvr2's methods 3 (PCA of {q} + all points,
                    numpy.linalg.eigh), 2 (second-moment determinant solve) and 0 (least
                    squares y = c0 + c1 x + c2 z, numpy.linalg.lstsq), oriented to a flip point.
"""
from __future__ import annotations

import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def knn_contract(points: np.ndarray, q: np.ndarray, k: int):
    p = points.astype(np.float32)
    out_i = np.full((q.shape[0], k), -1, np.int32)
    out_d = np.full((q.shape[0], k), np.inf, np.float32)
    for r in range(q.shape[0]):
        qq = q[r].astype(np.float32)
        dx = qq[0] - p[:, 0]
        dy = qq[1] - p[:, 1]
        dz = qq[2] - p[:, 2]
        d2 = (dx * dx + dy * dy) + dz * dz
        order = np.lexsort((np.arange(p.shape[0]), d2))[:k]
        out_i[r, : order.size] = order
        out_d[r, : order.size] = d2[order]
    return out_i, out_d


def make_knn():
    rng = np.random.default_rng(20240501)
    clouds = {}
    a = rng.normal(size=(700, 3)).astype(np.float32)
    a[600:] = a[:100]  # exact duplicates
    clouds["gauss_dups"] = a
    g = np.stack(np.meshgrid(np.arange(9), np.arange(9), np.arange(4), indexing="ij"), -1).reshape(-1, 3)
    clouds["lattice"] = g.astype(np.float32)
    clouds["tiny"] = rng.random((7, 3)).astype(np.float32)
    out = {}
    for name, pts in clouds.items():
        q = np.concatenate([pts[:: max(1, pts.shape[0] // 40)], rng.normal(size=(10, 3)).astype(np.float32) * 3])
        k = 16
        i, d = knn_contract(pts, q, k)
        out[f"{name}_points"] = pts
        out[f"{name}_queries"] = q
        out[f"{name}_idx"] = i
        out[f"{name}_d2"] = d
    np.savez_compressed(os.path.join(HERE, "knn_golden.npz"), **out)


def orient(n, q, flip):
    return -n if np.dot(n, flip - q) < 0 else n


def make_planes():
    rng = np.random.default_rng(777)
    cases = []
    for c in range(10):
        n_true = rng.normal(size=3)
        n_true /= np.linalg.norm(n_true)
        u = np.cross(n_true, [1.0, 0.0, 0.0])
        if np.linalg.norm(u) < 0.1:
            u = np.cross(n_true, [0.0, 1.0, 0.0])
        u /= np.linalg.norm(u)
        v = np.cross(n_true, u)
        center = rng.uniform(-50, 50, size=3)
        s, t = rng.uniform(-1, 1, size=(2, 20))
        noise = rng.normal(scale=0.05 * (c % 3), size=20)
        pts = center + s[:, None] * u * (1 + c % 4) + t[:, None] * v + noise[:, None] * n_true
        if c == 9:  # a corner-like, strongly non-planar case
            pts = center + rng.normal(size=(20, 3))
        cases.append(pts.astype(np.float32))
    P = np.stack(cases)  # (10, 20, 3)
    flips = rng.uniform(-100, 100, size=(10, 3)).astype(np.float32)
    n3 = np.zeros((10, 3))
    n2 = np.zeros((10, 3))
    n0 = np.zeros((10, 3))
    for c in range(10):
        pts = P[c].astype(np.float64)
        q = pts[0]
        fl = flips[c].astype(np.float64)
        # method 3: covariance of {q} + all 20 points (q appears twice)
        m = np.vstack([q[None], pts])
        C = np.cov(m.T, bias=True)
        w, V = np.linalg.eigh(C)
        n3[c] = orient(V[:, 0], q, fl)
        # method 2: second moments about q, largest-determinant axis solve
        r = pts - q
        xx, xy, xz = (r[:, 0] * r[:, 0]).sum(), (r[:, 0] * r[:, 1]).sum(), (r[:, 0] * r[:, 2]).sum()
        yy, yz, zz = (r[:, 1] * r[:, 1]).sum(), (r[:, 1] * r[:, 2]).sum(), (r[:, 2] * r[:, 2]).sum()
        dets = [yy * zz - yz * yz, xx * zz - xz * xz, xx * yy - xy * xy]
        if dets[0] >= dets[1] and dets[0] >= dets[2]:
            d = np.array([1.0, (xz * yz - xy * zz) / dets[0], (xy * yz - xz * yy) / dets[0]])
        elif dets[1] >= dets[0] and dets[1] >= dets[2]:
            d = np.array([(yz * xz - xy * zz) / dets[1], 1.0, (xy * xz - yz * xx) / dets[1]])
        else:
            d = np.array([(yz * xy - xz * yy) / dets[2], (xz * xy - yz * xx) / dets[2], 1.0])
        n2[c] = orient(d / np.linalg.norm(d), q, fl)
        # method 0: y = c0 + c1 x + c2 z by least squares; lvr2 then takes the cross product
        # of two tangents that start at the query point itself, 100 units along x and z
        # (AdaptiveKSearchSurface.tcc:497-504), so q's residual tilts the normal slightly.
        B = np.stack([np.ones(20), pts[:, 0], pts[:, 2]], 1)
        coef, *_ = np.linalg.lstsq(B, pts[:, 1], rcond=None)
        eps = 100.0
        z1 = coef[0] + coef[1] * (q[0] + eps) + coef[2] * q[2]
        z2 = coef[0] + coef[1] * q[0] + coef[2] * (q[2] + eps)
        nn = np.cross([eps, z1 - q[1], 0.0], [0.0, z2 - q[1], eps])
        n0[c] = orient(nn / np.linalg.norm(nn), q, fl)
    np.savez_compressed(os.path.join(HERE, "planes_golden.npz"), points=P, flips=flips, n3=n3, n2=n2, n0=n0)


if __name__ == "__main__":
    make_knn()
    make_planes()
    print("wrote", sorted(f for f in os.listdir(HERE) if f.endswith(".npz")))
