"""Seeded synthetic point clouds (bindings of csrc/scenes.cpp) and the workload configs.

Config names follow the workloads the path is measured on:
  plain_small      300,000 points, plane + 3 boxes + sphere, sigma 0.5; kn10 ki10 kd5, voxel 10
  room_4m          4,000,000 rays into a 20x15x4 m room with 40 objects, sigma 2 mm;
                   kn50 ki20 kd10, flip = scanner origin, voxel 0.05 m
  sphere_torus_1m  1,000,000 points on a noisy sphere + torus (sigma 1 % of R), kn30
  large_outdoor    32,000,000 points ray-cast from the centre of a 200 x 200 x 30 m outdoor area
                   (Perlin-noise terrain, 500 box buildings), sigma 3 mm; kn100 ki30 kd10,
                   flip = scanner, voxel 0.1 m
  scan_project_outdoor  64 scans of that area from random scanner poses, sizes log-uniform in
                   [50,000, 2,000,000]; kn50 ki20 kd10, per-scan flip point = scanner
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SCENES_PATH = os.path.join(HERE, "libscenes.so")
_lib = None


def _l():
    global _lib
    if _lib is None:
        if not os.path.exists(SCENES_PATH):
            raise RuntimeError(f"{SCENES_PATH} is not built; run `python -m lvr2_amd.build scenes`")
        l = C.CDLL(SCENES_PATH)
        l.scene_plane_boxes_sphere.argtypes = [C.c_uint64, C.c_int64, C.c_double, C.c_void_p]
        l.scene_room_scan.argtypes = [C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_double, C.c_void_p]
        l.scene_sphere_torus.argtypes = [C.c_uint64, C.c_int64, C.c_double, C.c_double, C.c_void_p]
        l.scene_room_pose.argtypes = [C.c_uint64, C.c_int, C.c_uint64, C.c_void_p]
        l.scene_room_scan_at.argtypes = [C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_double, C.c_void_p,
                                         C.c_uint64, C.c_void_p]
        l.scene_outdoor_scan.argtypes = [C.c_uint64, C.c_int, C.c_int64, C.c_double, C.c_void_p, C.c_void_p]
        l.scene_outdoor_pose.argtypes = [C.c_uint64, C.c_int, C.c_uint64, C.c_void_p]
        l.scene_outdoor_scan_at.argtypes = [C.c_uint64, C.c_int, C.c_void_p, C.c_int64, C.c_double, C.c_uint64,
                                            C.c_void_p]
        _lib = l
    return _lib


def plane_boxes_sphere(seed: int, n: int, sigma: float = 0.5) -> np.ndarray:
    out = np.empty((n, 3), np.float32)
    _l().scene_plane_boxes_sphere(seed, n, sigma, out.ctypes.data_as(C.c_void_p))
    return out


def room_scan(seed: int, n_az: int, n_el: int, n_obj: int = 40, sigma: float = 0.002) -> np.ndarray:
    out = np.empty((n_az * n_el, 3), np.float32)
    _l().scene_room_scan(seed, n_az, n_el, n_obj, sigma, out.ctypes.data_as(C.c_void_p))
    return out


def room_pose(seed: int, pose_seed: int, n_obj: int = 40) -> np.ndarray:
    out = np.empty(3, np.float64)
    if _l().scene_room_pose(seed, n_obj, pose_seed, out.ctypes.data_as(C.c_void_p)) != 0:
        raise RuntimeError("no free scanner position found")
    return out


def room_scan_at(seed: int, n_az: int, n_el: int, pos, noise_seed: int, n_obj: int = 40,
                 sigma: float = 0.002) -> np.ndarray:
    out = np.empty((n_az * n_el, 3), np.float32)
    p = np.ascontiguousarray(pos, np.float64)
    _l().scene_room_scan_at(seed, n_az, n_el, n_obj, sigma, p.ctypes.data_as(C.c_void_p), noise_seed,
                            out.ctypes.data_as(C.c_void_p))
    return out


def outdoor_scan(seed: int, n: int, n_bld: int = 500, sigma: float = 0.003):
    """Exactly n points ray-cast from the centre of the outdoor area; returns (points, scanner)."""
    out = np.empty((n, 3), np.float32)
    sc = np.empty(3, np.float64)
    if _l().scene_outdoor_scan(seed, n_bld, n, sigma, out.ctypes.data_as(C.c_void_p),
                               sc.ctypes.data_as(C.c_void_p)) != 0:
        raise RuntimeError("outdoor scan: too few hits")
    return out, sc


def outdoor_pose(seed: int, pose_seed: int, n_bld: int = 500) -> np.ndarray:
    out = np.empty(3, np.float64)
    if _l().scene_outdoor_pose(seed, n_bld, pose_seed, out.ctypes.data_as(C.c_void_p)) != 0:
        raise RuntimeError("no free scanner position found")
    return out


def outdoor_scan_at(seed: int, pos, n: int, noise_seed: int, n_bld: int = 500, sigma: float = 0.003) -> np.ndarray:
    out = np.empty((n, 3), np.float32)
    p = np.ascontiguousarray(pos, np.float64)
    if _l().scene_outdoor_scan_at(seed, n_bld, p.ctypes.data_as(C.c_void_p), n, sigma, noise_seed,
                                  out.ctypes.data_as(C.c_void_p)) != 0:
        raise RuntimeError("outdoor scan: too few hits")
    return out


def sphere_torus(seed: int, n: int, radius: float = 1.0, sigma_rel: float = 0.01) -> np.ndarray:
    out = np.empty((n, 3), np.float32)
    _l().scene_sphere_torus(seed, n, radius, sigma_rel, out.ctypes.data_as(C.c_void_p))
    return out


@dataclass
class Workload:
    name: str
    points: np.ndarray
    kn: int
    ki: int
    kd: int
    flip: tuple
    voxel: float


def plain_small(n: int = 300_000, seed: int = 1) -> Workload:
    return Workload("plain_small", plane_boxes_sphere(seed, n, 0.5), 10, 10, 5, (0.0, 0.0, 0.0), 10.0)


def room(n_az: int = 3200, n_el: int = 1250, seed: int = 2) -> Workload:
    """Terrestrial-scan-like room; the default grid gives 4,000,000 rays."""
    return Workload(f"room_{n_az * n_el}", room_scan(seed, n_az, n_el, 40, 0.002), 50, 20, 10, (0.0, 0.0, 0.0),
                    0.05)


def large_outdoor(n: int = 32_000_000, seed: int = 3) -> Workload:
    """The large single cloud: n points from one scanner at the centre of the outdoor area."""
    pts, sc = outdoor_scan(seed, n, 500, 0.003)
    return Workload(f"outdoor_{n}", pts, 100, 30, 10, tuple(float(c) for c in sc), 0.1)


def sphere_torus_1m(n: int = 1_000_000, seed: int = 5) -> Workload:
    return Workload("sphere_torus", sphere_torus(seed, n, 1.0, 0.01), 30, 0, 10, (0.0, 0.0, 0.0), 0.05)


@dataclass
class ScanProject:
    name: str
    points: np.ndarray  # all scans concatenated
    sizes: np.ndarray   # points per scan
    flips: np.ndarray   # scanner position per scan (the flip point)
    kn: int
    ki: int
    kd: int


def scan_project(nscans: int = 64, lo: int = 50_000, hi: int = 2_000_000, seed: int = 4, kn: int = 50,
                 ki: int = 20, kd: int = 10) -> ScanProject:
    """Many scans of one room from random scanner positions, sizes log-uniform in [lo, hi]
    (each an n_az x n_el ray grid with the room scan's 2.56 : 1 aspect)."""
    rng = np.random.default_rng(seed)
    target = np.exp(rng.uniform(np.log(lo), np.log(hi), nscans))
    pts, sizes, flips = [], [], []
    for s_, t in enumerate(target):
        n_el = max(4, int(round(np.sqrt(t / 2.56))))
        n_az = max(4, int(round(t / n_el)))
        pos = room_pose(seed, 1000 + s_)
        pts.append(room_scan_at(seed, n_az, n_el, pos, 7000 + s_))
        sizes.append(n_az * n_el)
        flips.append(pos)
    return ScanProject(f"scan_project_{nscans}", np.concatenate(pts), np.asarray(sizes, np.int64),
                       np.asarray(flips, np.float32), kn, ki, kd)


def scan_project_outdoor(nscans: int = 64, lo: int = 50_000, hi: int = 2_000_000, seed: int = 4, kn: int = 50,
                         ki: int = 20, kd: int = 10) -> ScanProject:
    """Many scans of the outdoor area (seed selects the scene), each from its own scanner pose,
    sizes log-uniform in [lo, hi]; the flip point of each scan is its scanner."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(lo), np.log(hi), nscans)).astype(np.int64)
    pts, flips = [], []
    for s_, n in enumerate(sizes):
        pos = outdoor_pose(seed, 1000 + s_)
        pts.append(outdoor_scan_at(seed, pos, int(n), 7000 + s_))
        flips.append(pos)
    return ScanProject(f"outdoor_project_{nscans}", np.concatenate(pts), sizes, np.asarray(flips, np.float32), kn,
                       ki, kd)


def band_queries(points: np.ndarray, m: int, width: float, seed: int = 6) -> np.ndarray:
    """m query positions within `width` of the cloud: a random point of the cloud plus a
    uniform random offset in the ball of radius `width`."""
    rng = np.random.default_rng(seed)
    base = points[rng.integers(0, len(points), m)].astype(np.float64)
    d = rng.normal(size=(m, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = width * rng.uniform(0, 1, m) ** (1.0 / 3.0)
    return (base + d * r[:, None]).astype(np.float32)


def sdf_denoise(n: int = 1_000_000, m: int = 500_000, seed: int = 5):
    """The differentiation workload: the noisy sphere + torus (sigma 1 % of R), kn 30, kd 10,
    voxel 0.05 R, m query positions in a 2-voxel band."""
    w = sphere_torus_1m(n, seed)
    return w, band_queries(w.points, m, 2 * w.voxel, seed + 1)
