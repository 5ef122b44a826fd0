"""HipSurface: the Python mirror of lvr2's CudaSurface / PointsetSurface call surface.

Method names, argument meaning and defaults follow ``lvr2::CudaSurface``
(include/lvr2/reconstruction/cuda/CudaSurface.hpp:70-158; defaults CudaSurface.cu:1007-1023)
and the setters of ``lvr2::PointsetSurface`` (PointsetSurface.hpp:112-131).  Every call goes
through the C ABI of liblvrhip.so.  Points may be given as a host array (N x 3 float32) or
as a torch tensor on a ROCm device (the device-resident path, no host copy).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib

METHODS = {"LS": 0, "PLANE": 0, "ITERATIVE": 2, "PCA": 3}


def _as_host_points(a) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    if a.ndim != 2 or a.shape[1] != 3:
        raise ValueError("points must have shape (N, 3)")
    return a


def _is_device_tensor(a) -> bool:
    return hasattr(a, "is_cuda") and bool(getattr(a, "is_cuda"))


def _device_points(t, what: str):
    """A (M, 3) float32 contiguous view of a device tensor (converted if needed)."""
    import torch

    if t.ndim != 2 or t.shape[1] != 3:
        raise ValueError(f"{what} must have shape (M, 3)")
    return t.detach().to(torch.float32).contiguous()


def _check_out(t, shape, dtype, device, what: str) -> None:
    """An output tensor the library writes through its data pointer: exact shape and dtype,
    contiguous, on the surface's device."""
    if not _is_device_tensor(t):
        raise ValueError(f"{what} must be a device tensor")
    if tuple(t.shape) != tuple(shape) or t.dtype != dtype or not t.is_contiguous() or t.device != device:
        raise ValueError(f"{what} must be a contiguous {dtype} tensor of shape {tuple(shape)} on {device}")


def _scan_table(sizes) -> np.ndarray:
    sz = np.asarray(sizes, dtype=np.int64).ravel()
    if sz.size < 1 or (sz < 0).any():
        raise ValueError("scan sizes must be a non-empty list of non-negative counts")
    return np.concatenate([[0], np.cumsum(sz)]).astype(np.int64)


class HipSurface:
    """Drop-in for ``CudaSurface(floatArr& points, size_t num_points, int device)``.

    ``scan_sizes`` makes it the batched form: the points are the concatenation of scans of
    those sizes, each searched on its own (estimateProjectNormals, ScanProjectUtils.cpp:361-410).
    With a device tensor the tensor must stay alive for ``rebuild`` and graph replays; the
    surface keeps a reference to it.
    """

    def __init__(self, points, device: int = 0, stream=None, scan_sizes=None):
        self._h = C.c_void_p()
        self._keep = None
        if _is_device_tensor(points):
            import torch

            t = points.detach()
            if t.dtype != torch.float32 or not t.is_contiguous():
                t = t.to(torch.float32).contiguous()
            if t.ndim != 2 or t.shape[1] != 3:
                raise ValueError("points must have shape (N, 3)")
            self._n = int(t.shape[0])
            self._keep = t
            device = t.device.index if t.device.index is not None else device
            st = stream if stream is not None else torch.cuda.current_stream(t.device).cuda_stream
            off = _scan_table([self._n] if scan_sizes is None else scan_sizes)
            if off[-1] != self._n:
                raise ValueError("scan sizes must add up to the number of points")
            check(lib().lvrhip_create_scans_device(C.c_void_p(t.data_ptr()), off.ctypes.data_as(C.c_void_p),
                                                   len(off) - 1, device, C.c_void_p(st), C.byref(self._h)))
            torch.cuda.current_stream(t.device).synchronize()
        else:
            p = _as_host_points(points)
            self._n = int(p.shape[0])
            off = _scan_table([self._n] if scan_sizes is None else scan_sizes)
            if off[-1] != self._n:
                raise ValueError("scan sizes must add up to the number of points")
            check(lib().lvrhip_create_scans(p.ctypes.data_as(C.c_void_p), off.ctypes.data_as(C.c_void_p),
                                            len(off) - 1, device, C.byref(self._h)))
            if stream is not None:
                self.setStream(stream)
        self.nscans = len(off) - 1
        self.device = device
        self._kn, self._ki, self._kd = 10, 10, 5

    # -- lifecycle ---------------------------------------------------------------------------
    def freeGPU(self) -> None:
        if self._h:
            check(lib().lvrhip_destroy(self._h))
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.freeGPU()
        except Exception:
            pass

    def __len__(self) -> int:
        return self._n

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    # -- parameters (CudaSurface.hpp:116-158) ---------------------------------------------------
    def setKn(self, kn: int) -> None:
        self._kn = int(kn)
        self._push_k()

    def setKi(self, ki: int) -> None:
        self._ki = int(ki)
        self._push_k()

    def setKd(self, kd: int) -> None:
        self._kd = int(kd)
        self._push_k()

    def _push_k(self) -> None:
        check(lib().lvrhip_set_k(self._h, self._kn, self._ki, self._kd))

    def setFlippoint(self, x: float, y: float, z: float) -> None:
        check(lib().lvrhip_set_flip_point(self._h, float(x), float(y), float(z)))

    def setScanFlippoints(self, flips) -> None:
        """One flip point per scan, (nscans, 3)."""
        f = _as_host_points(flips)
        if f.shape[0] != self.nscans:
            raise ValueError("one flip point per scan")
        check(lib().lvrhip_set_scan_flip_points(self._h, f.ctypes.data_as(C.c_void_p)))

    def setMethod(self, method) -> None:
        """"PCA" (calcPlaneIPCAExact), "ITERATIVE" (calcPlaneIterative), "LS" (calcPlane),
        or the AdaptiveKSearchSurface method number 3 / 2 / 0."""
        m = METHODS[method.upper()] if isinstance(method, str) else int(method)
        check(lib().lvrhip_set_method(self._h, m))

    def setStream(self, stream) -> None:
        v = stream.cuda_stream if hasattr(stream, "cuda_stream") else stream
        check(lib().lvrhip_set_stream(self._h, C.c_void_p(v or 0)))

    # -- normals ---------------------------------------------------------------------------------
    def calculateNormals(self) -> None:
        check(lib().lvrhip_calculate_normals(self._h))

    def interpolateNormals(self) -> None:
        check(lib().lvrhip_interpolate_normals(self._h))

    def setNormals(self, normals) -> None:
        """Use given normals (input order) instead of calculateNormals."""
        nrm = _as_host_points(normals)
        if nrm.shape[0] != self._n:
            raise ValueError("normals must have one row per point")
        check(lib().lvrhip_set_normals(self._h, nrm.ctypes.data_as(C.c_void_p)))

    def getNormals(self, out=None):
        """Normals in input order.  ``out`` may be a host array or a device tensor (N, 3)."""
        if out is not None and _is_device_tensor(out):
            import torch

            _check_out(out, (self._n, 3), torch.float32, out.device, "out")
            if out.device.index is not None and out.device.index != self.device:
                raise ValueError("out must be on the surface's device")
            check(lib().lvrhip_get_normals_device(self._h, C.c_void_p(out.data_ptr())))
            return out
        if out is None:
            res = np.empty((self._n, 3), np.float32)
        else:
            res = out
            if not (isinstance(res, np.ndarray) and res.dtype == np.float32 and res.shape == (self._n, 3)
                    and res.flags["C_CONTIGUOUS"] and res.flags["WRITEABLE"]):
                raise ValueError("out must be a writeable C-contiguous float32 array of shape (N, 3)")
        check(lib().lvrhip_get_normals(self._h, res.ctypes.data_as(C.c_void_p)))
        return res

    def setProfiling(self, on: bool = True) -> None:
        check(lib().lvrhip_set_profiling(self._h, int(bool(on))))

    def timings(self, reset: bool = False) -> dict:
        """{timer: (milliseconds, launches)} accumulated since the last reset."""
        from ._lib import TIMERS

        ms = np.zeros(len(TIMERS), np.float64)
        cnt = np.zeros(len(TIMERS), np.int64)
        check(lib().lvrhip_get_timings(self._h, ms.ctypes.data_as(C.c_void_p), cnt.ctypes.data_as(C.c_void_p),
                                       int(bool(reset))))
        return {name: (float(ms[i]), int(cnt[i])) for i, name in enumerate(TIMERS)}

    def kernelStats(self, reset: bool = False) -> dict:
        """Work counters of the kNN tile kernels, per group (collected while profiling)."""
        from ._lib import STAT_COUNTERS, STAT_GROUPS

        a = np.zeros(16 * len(STAT_GROUPS), np.int64)
        check(lib().lvrhip_get_kernel_stats(self._h, a.ctypes.data_as(C.c_void_p), int(bool(reset))))
        return {g: dict(zip(STAT_COUNTERS, (int(x) for x in a[16 * i: 16 * i + 16]))) for i, g in enumerate(STAT_GROUPS)}

    def setTuning(self, lanes_per_query: int = 0, tiles_per_wave: int = 0, r2_margin: float = 0.0) -> None:
        check(lib().lvrhip_set_tuning(self._h, int(lanes_per_query), int(tiles_per_wave), float(r2_margin)))

    def setSdfTuning(self, tiles_per_wave: int = 0, r2_margin: float = 0.0) -> None:
        check(lib().lvrhip_set_sdf_tuning(self._h, int(tiles_per_wave), float(r2_margin)))

    def setRetryMode(self, merged: bool) -> None:
        """Adaptive rounds 2..5 in one four-target search (True) or round by round (False)."""
        check(lib().lvrhip_set_retry_mode(self._h, int(bool(merged))))

    def autotune(self) -> dict:
        """Pick lanes per query / tiles per wave / radius margin / retry mode by timing a sample."""
        p, t, m = C.c_int(), C.c_int(), C.c_float()
        check(lib().lvrhip_autotune(self._h, C.byref(p), C.byref(t), C.byref(m)))
        return self.tuning()

    def tuning(self) -> dict:
        p, t, m, r = C.c_int(), C.c_int(), C.c_float(), C.c_int()
        check(lib().lvrhip_get_tuning(self._h, C.byref(p), C.byref(t), C.byref(m)))
        check(lib().lvrhip_get_retry_mode(self._h, C.byref(r)))
        return {"lanes_per_query": p.value, "tiles_per_wave": t.value, "r2_margin": m.value,
                "merged_retry": bool(r.value)}

    def distancesBackward(self, queries, grad_proj, grad_points=None):
        """Reverse mode of ``distances``' proj w.r.t. the points (device tensors only):
        returns grad_points (N, 3) += d(sum grad_proj * proj)/d(points)."""
        import torch

        if not (_is_device_tensor(queries) and _is_device_tensor(grad_proj)):
            raise ValueError("queries and grad_proj must be device tensors")
        q = _device_points(queries, "queries")
        w = grad_proj.detach().to(torch.float32).contiguous()
        if tuple(w.shape) != (q.shape[0],):
            raise ValueError("one weight per query")
        if w.device != q.device:
            raise ValueError("queries and grad_proj must be on the same device")
        if grad_points is None:
            grad_points = torch.zeros((self._n, 3), dtype=torch.float32, device=q.device)
        else:
            _check_out(grad_points, (self._n, 3), torch.float32, q.device, "grad_points")
        check(lib().lvrhip_distances_backward_device(self._h, C.c_void_p(q.data_ptr()), int(q.shape[0]),
                                                     C.c_void_p(w.data_ptr()), C.c_void_p(grad_points.data_ptr())))
        return grad_points

    # -- rebuild and graphs -----------------------------------------------------------------------
    def rebuild(self) -> None:
        """Rebuild the LBVH from the point buffer (after the points were changed in place)."""
        check(lib().lvrhip_rebuild(self._h))

    def captureGraph(self, out, rebuild: bool = True) -> None:
        """Capture [build +] normals + interpolation + copy to ``out`` (device (N, 3) float32)."""
        if not _is_device_tensor(out) or tuple(out.shape) != (self._n, 3) or not out.is_contiguous():
            raise ValueError("out must be a contiguous device tensor of shape (N, 3)")
        self._graph_out = out
        check(lib().lvrhip_graph_capture(self._h, C.c_void_p(out.data_ptr()), int(bool(rebuild))))

    def launchGraph(self) -> None:
        check(lib().lvrhip_graph_launch(self._h))

    def synchronize(self) -> None:
        check(lib().lvrhip_synchronize(self._h))

    def grid(self, voxel: float, extrude: bool = True, device: bool = False, box=None):
        """The PointsetGrid lattice: (vertices (nv, 3) float32, cells (nc, 8) int32 vertex ids,
        cell index (nc, 3) int32); torch device tensors with device=True.  box: the grid's bounding
        box (min xyz, max xyz), default the cloud's."""
        nv, nc = C.c_size_t(), C.c_size_t()
        if box is None:
            check(lib().lvrhip_grid_build(self._h, float(voxel), int(bool(extrude)), C.byref(nv), C.byref(nc)))
        else:
            b = np.ascontiguousarray(box, dtype=np.float32).ravel()
            if b.shape != (6,):
                raise ValueError("box must hold 6 values: min x, y, z, max x, y, z")
            check(lib().lvrhip_grid_build_box(self._h, float(voxel), int(bool(extrude)), b.ctypes.data_as(C.c_void_p),
                                              C.byref(nv), C.byref(nc)))
        nv, nc = nv.value, nc.value
        if device:
            import torch

            dev = torch.device("cuda", self.device)
            v = torch.empty((nv, 3), dtype=torch.float32, device=dev)
            c = torch.empty((nc, 8), dtype=torch.int32, device=dev)
            ci = torch.empty((nc, 3), dtype=torch.int32, device=dev)
            check(lib().lvrhip_grid_get_device(self._h, C.c_void_p(v.data_ptr()), C.c_void_p(c.data_ptr()),
                                               C.c_void_p(ci.data_ptr())))
            torch.cuda.synchronize(dev)
            return v, c, ci
        v = np.empty((nv, 3), np.float32)
        c = np.empty((nc, 8), np.int32)
        ci = np.empty((nc, 3), np.int32)
        check(lib().lvrhip_grid_get(self._h, v.ctypes.data_as(C.c_void_p), c.ctypes.data_as(C.c_void_p),
                                    ci.ctypes.data_as(C.c_void_p)))
        return v, c, ci

    def mesh(self, dist, invalid=None, decomposition: str = "PMC", device: bool = False):
        """FastReconstruction::getMesh over the last lattice (``grid``) given the signed distance of
        every lattice vertex (device tensor, nv float32) and their invalid flags (device tensor, nv
        uint8, or None).  decomposition "PMC" (lvr2_reconstruct's default: BilinearFastBox cells,
        degenerate faces skipped, border vertices moved by optimizePlanarFaces with this surface's
        points) or "MC" (FastBox cells).  Returns (vertices (V, 3) float32, faces (F, 3) int32),
        torch device tensors with device=True."""
        import torch

        modes = {"MC": 0, "PMC": 1}
        if decomposition not in modes:
            raise ValueError(f"decomposition must be one of {sorted(modes)}")
        if not _is_device_tensor(dist) or dist.dtype != torch.float32 or not dist.is_contiguous() or dist.ndim != 1:
            raise ValueError("dist must be a contiguous 1-D float32 device tensor (one value per lattice vertex)")
        if invalid is not None:
            if (not _is_device_tensor(invalid) or invalid.dtype != torch.uint8 or not invalid.is_contiguous()
                    or tuple(invalid.shape) != tuple(dist.shape)):
                raise ValueError("invalid must be a contiguous uint8 device tensor shaped like dist")
        nv, nf = C.c_size_t(), C.c_size_t()
        check(lib().lvrhip_mesh_build(self._h, C.c_void_p(dist.data_ptr()),
                                      C.c_void_p(invalid.data_ptr() if invalid is not None else 0),
                                      int(dist.shape[0]), modes[decomposition], C.byref(nv), C.byref(nf)))
        nv, nf = nv.value, nf.value
        if device:
            v = torch.empty((nv, 3), dtype=torch.float32, device=dist.device)
            f = torch.empty((nf, 3), dtype=torch.int32, device=dist.device)
            check(lib().lvrhip_mesh_get_device(self._h, C.c_void_p(v.data_ptr()), C.c_void_p(f.data_ptr())))
            torch.cuda.synchronize(dist.device)
            return v, f
        v = np.empty((nv, 3), np.float32)
        f = np.empty((nf, 3), np.int32)
        check(lib().lvrhip_mesh_get(self._h, v.ctypes.data_as(C.c_void_p), f.ctypes.data_as(C.c_void_p)))
        return v, f

    def reconstruct(self, voxel: float, extrude: bool = True, decomposition: str = "PMC", device: bool = False):
        """The reconstruction after the normals, on the device: the PointsetGrid lattice, the signed
        distances at its vertices with the invalid rule, and the mesh over the cells without an
        invalid corner (PointsetGrid.tcc:41-185, FastReconstruction.tcc:47-165)."""
        v, _, _ = self.grid(voxel, extrude, device=True)
        proj, _, inv = self.distances(v, voxel)
        return self.mesh(proj, inv, decomposition, device=device)

    def adaptiveK(self) -> np.ndarray:
        """Per point (input order): the k its normal was fitted on (the adaptive loop's final k)."""
        out = np.empty(self._n, np.int32)
        check(lib().lvrhip_get_adaptive_k(self._h, out.ctypes.data_as(C.c_void_p)))
        return out

    def normalsStats(self) -> np.ndarray:
        """Points searched in each adaptive round (k = kn*2, kn*4, ..., kn*32)."""
        a = (C.c_int64 * 5)()
        check(lib().lvrhip_normals_stats(self._h, a))
        return np.array(list(a), np.int64)

    # -- queries ----------------------------------------------------------------------------------
    def kSearch(self, queries, k: int):
        """SearchTree::kSearch for many queries: (idx int32 (m,k), d2 float32 (m,k))."""
        q = _as_host_points(queries)
        m = q.shape[0]
        idx = np.empty((m, k), np.int32)
        d2 = np.empty((m, k), np.float32)
        check(lib().lvrhip_knn(self._h, q.ctypes.data_as(C.c_void_p), m, int(k),
                               idx.ctypes.data_as(C.c_void_p), d2.ctypes.data_as(C.c_void_p)))
        return idx, d2

    def radiusSearch(self, queries, k: int, r: float):
        """SearchTree::radiusSearch for many queries: the k nearest with d^2 < r^2:
        (idx int32 (m,k), d2 float32 (m,k), count int32 (m,))."""
        q = _as_host_points(queries)
        m = q.shape[0]
        idx = np.empty((m, k), np.int32)
        d2 = np.empty((m, k), np.float32)
        cnt = np.empty(m, np.int32)
        check(lib().lvrhip_radius_search(self._h, q.ctypes.data_as(C.c_void_p), m, int(k), float(r),
                                         idx.ctypes.data_as(C.c_void_p), d2.ctypes.data_as(C.c_void_p),
                                         cnt.ctypes.data_as(C.c_void_p)))
        return idx, d2, cnt

    def getInterpolatedNormals(self, positions) -> np.ndarray:
        """PointsetSurface::getInterpolatedNormal at each row of ``positions``: (m, 3) float32."""
        q = _as_host_points(positions)
        out = np.empty((q.shape[0], 3), np.float32)
        check(lib().lvrhip_interpolated_normals(self._h, q.ctypes.data_as(C.c_void_p), q.shape[0],
                                                out.ctypes.data_as(C.c_void_p)))
        return out

    def setSearchMode(self, mode) -> None:
        """"auto" (best-first for searches of at most 32 neighbours), "tile" (always the tile kernel)
        or "wave" (the wave-per-query best-first search for every query)."""
        m = {"auto": 0, "tile": 1, "wave": 2}[mode] if isinstance(mode, str) else int(mode)
        check(lib().lvrhip_set_search_mode(self._h, m))

    def distances(self, query_points, voxel_size: float, out=None):
        """CudaSurface::distances semantics: returns (proj, eucl, invalid) per query point.

        ``query_points`` is (M, 3) float32 (host), or a device tensor; with a device tensor
        ``out`` may be a tuple of device tensors (proj, eucl, invalid) to fill in place.
        """
        if _is_device_tensor(query_points):
            import torch

            qt = _device_points(query_points, "query_points")
            m = int(qt.shape[0])
            if out is None:
                out = (torch.empty(m, dtype=torch.float32, device=qt.device),
                       torch.empty(m, dtype=torch.float32, device=qt.device),
                       torch.empty(m, dtype=torch.uint8, device=qt.device))
            else:
                if len(out) != 3:
                    raise ValueError("out must be a tuple (proj, eucl, invalid)")
                _check_out(out[0], (m,), torch.float32, qt.device, "out[0] (proj)")
                _check_out(out[1], (m,), torch.float32, qt.device, "out[1] (eucl)")
                _check_out(out[2], (m,), torch.uint8, qt.device, "out[2] (invalid)")
            check(lib().lvrhip_distances_device(self._h, C.c_void_p(qt.data_ptr()), m, float(voxel_size),
                                                C.c_void_p(out[0].data_ptr()), C.c_void_p(out[1].data_ptr()),
                                                C.c_void_p(out[2].data_ptr())))
            return out
        q = _as_host_points(query_points)
        m = q.shape[0]
        proj = np.empty(m, np.float32)
        eucl = np.empty(m, np.float32)
        inv = np.empty(m, np.uint8)
        check(lib().lvrhip_distances(self._h, q.ctypes.data_as(C.c_void_p), m, float(voxel_size),
                                     proj.ctypes.data_as(C.c_void_p), eucl.ctypes.data_as(C.c_void_p),
                                     inv.ctypes.data_as(C.c_void_p)))
        return proj, eucl, inv
