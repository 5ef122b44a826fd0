"""ORACLE — test infrastructure only.

ctypes binding of the CPU restatement of lvr2's AdaptiveKSearchSurface path
(lvr2_oracle.cpp).  Imported only by tests/, __graft_entry__.smoke() and the cpu_baseline
leg of bench.py, always as the checker or the timed CPU baseline, never as the product.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblvr2_oracle.so")

CONTRACT = 0
REFERENCE = 1
_lib = None


def build() -> str:
    subprocess.run(["make", "-C", HERE, "-s"], check=True)
    return LIB_PATH


def _l():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} not built; run `make -C oracle`")
        l = C.CDLL(LIB_PATH)
        vp, i64, i32, f32 = C.c_void_p, C.c_int64, C.c_int, C.c_float
        l.orc_surface_create.restype = vp
        l.orc_surface_create.argtypes = [vp, i64, i32]
        l.orc_surface_destroy.argtypes = [vp]
        l.orc_surface_bbox.argtypes = [vp, vp]
        l.orc_knn.argtypes = [vp, vp, i64, i32, vp, vp]
        l.orc_knn_brute.argtypes = [vp, i64, vp, i64, i32, vp, vp]
        l.orc_radius.argtypes = [vp, vp, i64, i32, f32, vp, vp, vp]
        l.orc_interpolated_normal.argtypes = [vp, vp, i32, vp, i64, vp]
        l.orc_interpolate_f64.argtypes = [vp, vp, i32, vp, i64, vp]
        l.orc_distance_f64.argtypes = [vp, vp, i32, vp, i64, vp, vp]
        l.orc_normals.argtypes = [vp, i32, i32, vp, vp, i64, vp, vp, vp, vp]
        l.orc_interpolate.argtypes = [vp, vp, i32, vp, i64, vp]
        l.orc_distance.argtypes = [vp, vp, i32, vp, i64, vp, vp]
        l.orc_grid.argtypes = [vp, i64, vp, f32, i32, C.POINTER(i64), C.POINTER(i64), vp, vp, vp]
        l.orc_calc_distance_values.argtypes = [vp, i64, f32, vp, i64, vp, vp]
        l.orc_mesh.argtypes = [vp, i64, vp, i64, vp, vp, i32, i64, i64, C.POINTER(i64), C.POINTER(i64), vp, vp, vp]
        l.orc_pmc_optimize.argtypes = [vp, i32, vp, i64, vp, vp, i64, C.POINTER(i64), C.POINTER(i64)]
        l.orc_eig3_f32.argtypes = [vp, vp, vp]
        l.orc_eig3_f64.argtypes = [vp, vp, vp]
        l.orc_num_threads.restype = i32
        _lib = l
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def num_threads() -> int:
    return int(_l().orc_num_threads())


class Surface:
    """AdaptiveKSearchSurface over the in-tree SearchKDTree ("lvr2")."""

    def __init__(self, points, metric: int = CONTRACT):
        self.points = _f32(points).reshape(-1, 3)
        self.n = self.points.shape[0]
        self._s = _l().orc_surface_create(_p(self.points), self.n, metric)

    def __del__(self):
        if getattr(self, "_s", None):
            _l().orc_surface_destroy(self._s)
            self._s = None

    def bbox(self) -> np.ndarray:
        b = np.empty(6, np.float32)
        _l().orc_surface_bbox(self._s, _p(b))
        return b

    def knn(self, q, k: int):
        q = _f32(q).reshape(-1, 3)
        idx = np.empty((q.shape[0], k), np.int32)
        d2 = np.empty((q.shape[0], k), np.float32)
        _l().orc_knn(self._s, _p(q), q.shape[0], k, _p(idx), _p(d2))
        return idx, d2

    def pmc_optimize(self, vertices, faces, face_cell, kc: int = 5):
        """BilinearFastBox::optimizePlanarFaces over every cell, sequentially: returns (moved
        vertices, end-point moves, non-manifold edge count)."""
        v = _f32(vertices).reshape(-1, 3).copy()
        f = np.ascontiguousarray(faces, dtype=np.int32)
        fc = np.ascontiguousarray(face_cell, dtype=np.int32)
        mv, bad = C.c_int64(), C.c_int64()
        _l().orc_pmc_optimize(self._s, kc, _p(v), v.shape[0], _p(f), _p(fc), f.shape[0], C.byref(mv), C.byref(bad))
        return v, mv.value, bad.value

    def normals(self, kn: int, method: int = 3, flip=(0.0, 0.0, 0.0), subset=None, fp64: bool = False,
                gap: bool = False):
        """Returns (n32 (m,3) float32, n64 (m,3) float64 or None, k_used (m,) int32), plus the
        fp64 eigen gap (lambda_mid - lambda_min) / lambda_max of method 3 when gap=True."""
        fl = _f32(flip)
        sub = None if subset is None else np.ascontiguousarray(subset, dtype=np.int64)
        m = self.n if sub is None else sub.shape[0]
        n32 = np.empty((m, 3), np.float32)
        n64 = np.empty((m, 3), np.float64) if (fp64 or gap) else None
        k = np.empty(m, np.int32)
        g = np.zeros(m, np.float64) if gap else None
        _l().orc_normals(self._s, kn, method, _p(fl), _p(sub), 0 if sub is None else m, _p(n32), _p(n64), _p(k),
                         _p(g))
        if gap:
            return n32, n64, k, g
        return n32, n64, k

    def interpolate(self, normals, ki: int, subset=None):
        nin = _f32(normals).reshape(-1, 3)
        sub = None if subset is None else np.ascontiguousarray(subset, dtype=np.int64)
        m = self.n if sub is None else sub.shape[0]
        out = np.empty((m, 3), np.float32)
        _l().orc_interpolate(self._s, _p(nin), ki, _p(sub), 0 if sub is None else m, _p(out))
        return out

    def radius(self, q, k: int, r: float):
        """SearchKDTree::radiusSearch: (idx (m,k), d2 (m,k), count (m,))."""
        q = _f32(q).reshape(-1, 3)
        idx = np.empty((q.shape[0], k), np.int32)
        d2 = np.empty((q.shape[0], k), np.float32)
        cnt = np.empty(q.shape[0], np.int32)
        _l().orc_radius(self._s, _p(q), q.shape[0], k, float(r), _p(idx), _p(d2), _p(cnt))
        return idx, d2, cnt

    def interpolated_normal(self, normals, ki: int, q):
        """PointsetSurface::getInterpolatedNormal at the rows of q: (m, 3) float32."""
        nin = _f32(normals).reshape(-1, 3)
        q = _f32(q).reshape(-1, 3)
        out = np.empty((q.shape[0], 3), np.float32)
        _l().orc_interpolated_normal(self._s, _p(nin), ki, _p(q), q.shape[0], _p(out))
        return out

    def interpolate_f64(self, normals64, ki: int, subset=None):
        """interpolateSurfaceNormals in fp64 on the same neighbourhoods: (m, 3) float64."""
        nin = np.ascontiguousarray(normals64, dtype=np.float64).reshape(-1, 3)
        sub = None if subset is None else np.ascontiguousarray(subset, dtype=np.int64)
        m = self.n if sub is None else sub.shape[0]
        out = np.empty((m, 3), np.float64)
        _l().orc_interpolate_f64(self._s, _p(nin), ki, _p(sub), 0 if sub is None else m, _p(out))
        return out

    def distance_f64(self, normals64, kd: int, q):
        """AdaptiveKSearchSurface::distance in fp64: (proj, eucl) float64."""
        nin = np.ascontiguousarray(normals64, dtype=np.float64).reshape(-1, 3)
        q = _f32(q).reshape(-1, 3)
        proj = np.empty(q.shape[0], np.float64)
        eucl = np.empty(q.shape[0], np.float64)
        _l().orc_distance_f64(self._s, _p(nin), kd, _p(q), q.shape[0], _p(proj), _p(eucl))
        return proj, eucl

    def distance(self, normals, kd: int, q):
        nin = _f32(normals).reshape(-1, 3)
        q = _f32(q).reshape(-1, 3)
        proj = np.empty(q.shape[0], np.float32)
        eucl = np.empty(q.shape[0], np.float32)
        _l().orc_distance(self._s, _p(nin), kd, _p(q), q.shape[0], _p(proj), _p(eucl))
        return proj, eucl


def knn_brute(points, q, k: int):
    p = _f32(points).reshape(-1, 3)
    q = _f32(q).reshape(-1, 3)
    idx = np.empty((q.shape[0], k), np.int32)
    d2 = np.empty((q.shape[0], k), np.float32)
    _l().orc_knn_brute(_p(p), p.shape[0], _p(q), q.shape[0], k, _p(idx), _p(d2))
    return idx, d2


def grid(points, bbox, voxel: float, extrude: bool = True):
    """PointsetGrid vertex set: (vertices (nv,3) float32, cells (nc,8) int32, cell index (nc,3))."""
    p = _f32(points).reshape(-1, 3)
    bb = _f32(bbox)
    nv, nc = C.c_int64(), C.c_int64()
    _l().orc_grid(_p(p), p.shape[0], _p(bb), voxel, int(extrude), C.byref(nv), C.byref(nc), None, None, None)
    v = np.empty((nv.value, 3), np.float32)
    c = np.empty((nc.value, 8), np.int32)
    ci = np.empty((nc.value, 3), np.int32)
    _l().orc_grid(_p(p), p.shape[0], _p(bb), voxel, int(extrude), C.byref(nv), C.byref(nc), _p(v), _p(c), _p(ci))
    return v, c, ci


def calc_distance_values(eucl, voxel: float, cells):
    e = _f32(eucl)
    c = np.ascontiguousarray(cells, dtype=np.int32)
    inv = np.empty(e.shape[0], np.uint8)
    keep = np.empty(c.shape[0], np.uint8)
    _l().orc_calc_distance_values(_p(e), e.shape[0], voxel, _p(c), c.shape[0], _p(inv), _p(keep))
    return inv, keep


def mesh(vertices, cells, dist, invalid=None, pmc: bool = False, face_cells: bool = False):
    """FastReconstruction::getMesh over a lattice with FastBox cells (pmc=False) or BilinearFastBox
    cells (pmc=True, degenerate faces skipped): (vertices (V, 3) float32, faces (F, 3) int32), plus
    the cell of every face when face_cells=True."""
    v = _f32(vertices).reshape(-1, 3)
    c = np.ascontiguousarray(cells, dtype=np.int32)
    d = _f32(dist).ravel()
    inv = None if invalid is None else np.ascontiguousarray(invalid, dtype=np.uint8)
    nv_, nf_ = C.c_int64(), C.c_int64()
    cap_v, cap_f = 3 * v.shape[0] + 1, 4 * c.shape[0] + 1
    ov = np.empty((cap_v, 3), np.float32)
    of = np.empty((cap_f, 3), np.int32)
    oc = np.empty(cap_f, np.int32)
    _l().orc_mesh(_p(v), v.shape[0], _p(c), c.shape[0], _p(d), _p(inv), int(pmc), cap_v, cap_f, C.byref(nv_),
                  C.byref(nf_), _p(ov), _p(of), _p(oc))
    out = (ov[: nv_.value].copy(), of[: nf_.value].copy())
    return out + (oc[: nf_.value].copy(),) if face_cells else out


def eig3(cov, fp64: bool = False):
    if fp64:
        a = np.ascontiguousarray(cov, dtype=np.float64)
        w = np.empty(3, np.float64)
        v = np.empty((3, 3), np.float64)
        _l().orc_eig3_f64(_p(a), _p(w), _p(v))
    else:
        a = np.ascontiguousarray(cov, dtype=np.float32)
        w = np.empty(3, np.float32)
        v = np.empty((3, 3), np.float32)
        _l().orc_eig3_f32(_p(a), _p(w), _p(v))
    return w, v
