"""FP64 finite-difference check of the SDF adjoint (test infrastructure).

The loss is L = sum_v proj_v^2 with, per AdaptiveKSearchSurface::distance (tcc:414-457),
proj_v = (v - mean_{N(v)} p) . normalize(mean_{N(v)} n) and, per calcPlaneIPCAExact
(tcc:609-673), n_j = s_j * (eigenvector of the smallest eigenvalue of the covariance of the
k_j + 1 samples {p_j} + kNN_{k_j}(p_j)).  Neighbour lists come from the oracle (the same
exact sets as the GPU), adaptive k from the oracle's adaptive loop, signs s_j from the flip.

For one coordinate (point i, axis a) only the normals whose list holds i and the vertices
whose list holds one of those points change, so L(p_i[a] +- eps) is evaluated in FP64 on
that local part.  A coordinate is *certified* when no neighbour set, adaptive k or flip sign
can change under the +-eps move:
  - no point j's list gains or loses i: |p_j - p_i|^2 + delta < d_{k+1}(j)^2 when i is in it,
    |p_j - p_i|^2 - delta > d_k(j)^2 when not (delta = 2 |p_j - p_i| eps + eps^2), likewise for
    every vertex with its kd-th / (kd+1)-th neighbour, and i's own k-th / (k+1)-th stay apart,
  - the bounding-box test (tcc:337-350) of every affected point keeps a margin above 2 eps
    in every round up to its own,
  - |u0 . (f - p_j)| of every affected normal stays above 1e-6 |f - p_j|.
"""
from __future__ import annotations

import numpy as np


def _eig_normals(P, seeds, lists, ref):
    """FP64 normals for points `seeds` with fixed neighbour lists, oriented like ref[seed]."""
    out = np.zeros((len(seeds), 3))
    for t, (j, row) in enumerate(zip(seeds, lists)):
        X = np.concatenate([P[j][None], P[row]])
        mu = X.mean(0)
        C = (X - mu).T @ (X - mu) / len(X)
        w, V = np.linalg.eigh(C)
        u = V[:, 0]
        out[t] = u if np.dot(u, ref[j]) >= 0 else -u
    return out


def _all_normals(P, lists_by_k, signs):
    N = np.zeros_like(P)
    for k, (idx, L) in lists_by_k.items():
        X = np.concatenate([P[idx][:, None, :], P[L]], axis=1)  # (m, k+1, 3)
        mu = X.mean(1, keepdims=True)
        D = X - mu
        C = np.matmul(D.transpose(0, 2, 1), D) / X.shape[1]
        w, V = np.linalg.eigh(C)
        N[idx] = signs[idx][:, None] * V[:, :, 0]
    return N


def _proj(P, N, V, vl):
    c = P[vl].mean(1)
    mt = N[vl].mean(1)
    m = mt / np.linalg.norm(mt, axis=1, keepdims=True)
    return np.sum((V - c) * m, 1)


def _bbox_margin(X):
    """Smallest slack of the six boundingBoxOK inequalities (e = 0.05) of the sample box."""
    e = 0.05
    d = X.max(0) - X.min(0)
    sl = [d[0] - e * d[1], d[0] - e * d[2], d[1] - e * d[0], d[1] - e * d[2], d[2] - e * d[0], d[2] - e * d[1]]
    return min(sl), all(s >= 0 for s in sl)


class FDCheck:
    def __init__(self, oracle_mod, points, queries, kn, kd, flip):
        self.o = oracle_mod.Surface(points)
        self.ov = oracle_mod.Surface(queries)
        self.P = points.astype(np.float64)
        self.V = queries.astype(np.float64)
        self.kn, self.kd = kn, kd
        self.flip = np.asarray(flip, np.float64)
        _, n64, k = self.o.normals(kn, 3, tuple(flip), fp64=True)
        self.k = k.astype(np.int64)
        n = len(points)
        self.lists = {}
        self.dk = np.zeros(n)
        self.dk1 = np.full(n, np.inf)
        rows = [None] * n
        own, mem = [], []
        for kv in np.unique(self.k):
            idx = np.where(self.k == kv)[0]
            kk = int(min(kv + 1, n))
            li, ld = self.o.knn(points[idx], kk)
            L = li[:, :kv].astype(np.int64)
            self.lists[int(kv)] = (idx, L)
            self.dk[idx] = ld[:, kv - 1]
            if kk > kv:
                self.dk1[idx] = ld[:, kv]
            for j, r in zip(idx.tolist(), list(L)):
                rows[j] = r
            own.append(np.repeat(idx, kv))
            mem.append(L.ravel())
        self.rows = rows
        # signs: the oracle's oriented fp64 normals fix s_j
        N0 = _all_normals(self.P, self.lists, np.ones(n))
        self.signs = np.where(np.sum(N0 * n64, 1) < 0, -1.0, 1.0)
        self.N = N0 * self.signs[:, None]
        vi, vd = self.o.knn(queries, kd + 1)
        self.vl = vi[:, :kd].astype(np.int64)
        self.vdk = vd[:, kd - 1]
        self.vdk1 = vd[:, kd]
        self.proj = _proj(self.P, self.N, self.V, self.vl)
        # reverse maps: point -> points whose list holds it; point -> vertices whose list holds it
        owners = np.concatenate(own)
        members = np.concatenate(mem)
        o_ = np.argsort(members, kind="stable")
        self.rp_ptr = np.searchsorted(members[o_], np.arange(n + 1))
        self.rp = owners[o_]
        vm = self.vl.ravel()
        vo = np.repeat(np.arange(len(queries)), kd)
        o_ = np.argsort(vm, kind="stable")
        self.rv_ptr = np.searchsorted(vm[o_], np.arange(n + 1))
        self.rv = vo[o_]
        self.kmax = int(self.k.max())

    def loss(self):
        return float(np.sum(self.proj ** 2))

    def coord(self, i, a, eps):
        """(fd, certified) for d L / d p_i[a]."""
        A = np.unique(self.rp[self.rp_ptr[i]:self.rp_ptr[i + 1]])
        Vs = np.unique(np.concatenate([self.rv[self.rv_ptr[j]:self.rv_ptr[j + 1]] for j in A] + [np.zeros(0, int)]))
        ok = self._certify(i, a, eps, A, Vs)
        Ls = []
        for sgn in (1.0, -1.0):
            P = self.P.copy()
            P[i, a] += sgn * eps
            N = self.N.copy()
            N[A] = _eig_normals(P, A, [self.rows[j] for j in A], self.N)
            Ls.append(np.sum(_proj(P, N, self.V[Vs], self.vl[Vs]) ** 2) if len(Vs) else 0.0)
        return (Ls[0] - Ls[1]) / (2 * eps), ok

    def _certify(self, i, a, eps, A, Vs):
        n = len(self.P)
        inA = set(int(x) for x in A)
        kc = int(min(n, 2 * self.kmax + 2))
        ci, cd = self.o.knn(self.P[i][None].astype(np.float32), kc)
        for j, d2 in zip(ci[0], cd[0]):
            if j < 0:
                continue
            delta = 2 * np.sqrt(d2) * eps + eps * eps
            if j == i:
                # the moving point's own list: its k-th and (k+1)-th neighbours keep their order
                d1 = 2 * np.sqrt(self.dk[i]) * eps + eps * eps
                d2_ = 2 * np.sqrt(self.dk1[i]) * eps + eps * eps if np.isfinite(self.dk1[i]) else 0.0
                if self.dk[i] + d1 >= self.dk1[i] - d2_:
                    return False
                continue
            # only p_i moves: j's other neighbours keep their distances
            if j in inA:
                if d2 + delta >= self.dk1[j]:
                    return False
            elif d2 - delta <= self.dk[j]:
                return False
        in_v = set(int(x) for x in self.rv[self.rv_ptr[i]:self.rv_ptr[i + 1]])
        vi, vd = self.ov.knn(self.P[i][None].astype(np.float32), int(min(len(self.V), 64)))
        for v, d2 in zip(vi[0], vd[0]):
            if v < 0:
                continue
            delta = 2 * np.sqrt(d2) * eps + eps * eps
            if v in in_v:
                if d2 + delta >= self.vdk1[v]:
                    return False
            elif d2 - delta <= self.vdk[v]:
                return False
        for j in A:
            row = self.rows[j]
            kj = len(row)
            for sgn in (1.0, -1.0):
                P = self.P[np.concatenate([[j], row])].copy()
                hit = np.where(np.concatenate([[j], row]) == i)[0]
                P[hit, a] += sgn * eps
                r = 1
                while (self.kn << r) <= kj:
                    kr = self.kn << r
                    margin, passed = _bbox_margin(P[1: kr + 1])
                    if abs(margin) < 2 * eps:
                        return False
                    if passed != (kr == kj) and kr < kj:
                        return False
                    r += 1
            fv = self.flip - self.P[j]
            if abs(np.dot(self.N[j], fv)) < 1e-6 * np.linalg.norm(fv):
                return False
        return True
