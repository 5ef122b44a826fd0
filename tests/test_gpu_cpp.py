"""The C++ drop-in (include/lvr2_amd/HipSurface.hpp) run where lvr2 runs CudaSurface: the
--useGPU block of lvr2_reconstruct (Main.cpp:356-373) and LargeScaleReconstruction's GPU block
(LargeScaleReconstruction.tcc:640-688), replayed by tests/cpp/hipsurface_check (built by
__graft_entry__.build() through tests/cpp/Makefile).

Bars: the normals calculateNormals() returns are the interpolated normals of
calculateSurfaceNormals (angle to the oracle's interpolation of its fp32 normals <= 1e-4 rad
for >= 99.99 % of the points whose interpolated normal is not near-orthogonal to the flip
direction); m_distance and m_invalid are bit-exact against the oracle given those normals."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, axis_angles

pytestmark = pytest.mark.gpu

EXE = os.path.join(ROOT, "tests", "cpp", "hipsurface_check")


def _exe():
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built (run __graft_entry__.build())")
    return EXE


def test_lvr2_call_sites_through_cpp_dropin(tmp_path, oracle_mod):
    from lvr2_amd import scenes

    w = scenes.plain_small()
    o = oracle_mod.Surface(w.points)
    v, cells, _ = oracle_mod.grid(w.points, o.bbox(), w.voxel, extrude=True)
    w.points.tofile(tmp_path / "points.f32")
    v.astype(np.float32).tofile(tmp_path / "queries.f32")
    args = [_exe(), "run", str(tmp_path / "points.f32"), str(len(w.points)), str(w.kn), str(w.ki), str(w.kd),
            *(repr(float(c)) for c in w.flip), repr(float(w.voxel)), str(tmp_path / "queries.f32"), str(len(v)),
            str(tmp_path)]
    print("running", " ".join(args[:2]), len(w.points), "points", len(v), "vertices", flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    n_main = np.fromfile(tmp_path / "normals_main.f32", np.float32).reshape(-1, 3)
    n_lsr = np.fromfile(tmp_path / "normals_lsr.f32", np.float32).reshape(-1, 3)
    np.testing.assert_array_equal(n_main, n_lsr)
    # calculateNormals = estimate + interpolate (CudaSurface::GPU_NN), against the oracle
    n32, _, _ = o.normals(w.kn, 3, w.flip)
    oi = o.interpolate(n32, w.ki)
    flipdir = np.asarray(w.flip, np.float64) - w.points.astype(np.float64)
    stable = np.abs(np.sum(oi * flipdir, 1)) > 1e-3 * np.linalg.norm(flipdir, axis=1)
    a = axis_angles(n_main[stable], oi[stable])
    print(f"interpolated normals vs oracle: p99.99 {np.quantile(a, 0.9999):.2e} max {a.max():.2e} rad")
    assert np.quantile(a, 0.9999) <= 1e-4
    # not the un-interpolated estimate: interpolation moves the normals far more than that error
    assert np.median(axis_angles(n_main, n32)) > 100 * np.median(a)
    # m_distance / m_invalid given those normals, bit-exact
    dist = np.fromfile(tmp_path / "distance.f32", np.float32)
    inv = np.fromfile(tmp_path / "invalid.u8", np.uint8)
    oproj, oeucl = o.distance(n_lsr, w.kd, v)
    np.testing.assert_array_equal(dist, oproj)
    oinv, _ = oracle_mod.calc_distance_values(oeucl, w.voxel, cells)
    np.testing.assert_array_equal(inv, oinv)


def test_dropin_host_code_under_sanitizers(tmp_path, oracle_mod):
    """hipsurface_check built with AddressSanitizer and UBSan on its host code (the C++ drop-in
    header, the QueryPoint and floatArr handling), run on the GPU like the plain build."""
    from lvr2_amd import scenes

    exe = os.path.join(ROOT, "tests", "cpp", "hipsurface_check_asan")
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built (run __graft_entry__.build())")
    w = scenes.plain_small(60_000)
    o = oracle_mod.Surface(w.points)
    v, _, _ = oracle_mod.grid(w.points, o.bbox(), w.voxel, extrude=True)
    w.points.tofile(tmp_path / "points.f32")
    v.astype(np.float32).tofile(tmp_path / "queries.f32")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    args = [exe, "run", str(tmp_path / "points.f32"), str(len(w.points)), str(w.kn), str(w.ki), str(w.kd),
            *(repr(float(c)) for c in w.flip), repr(float(w.voxel)), str(tmp_path / "queries.f32"), str(len(v)),
            str(tmp_path)]
    print("running hipsurface_check_asan", flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=300, env=env)
    if "Shadow memory range interleaves" in r.stderr:
        pytest.skip("AddressSanitizer's shadow memory collides with the GPU runtime's mappings on this box")
    assert r.returncode == 0, (r.returncode, r.stdout[-2000:], r.stderr[-4000:])
    assert "runtime error" not in r.stderr


def test_out_of_memory_contract(tmp_path):
    """A refused device allocation (LVRHIP_FAIL_ALLOC_AFTER, refused on the host) surfaces as
    LVRHIP_ENOMEM with "out of memory" in the C ABI and as the std::runtime_error that
    LargeScaleReconstruction.tcc:656-662 catches to shrink its chunks; the handle then works."""
    rng = np.random.default_rng(0)
    p = rng.normal(size=(20000, 3)).astype(np.float32)
    p.tofile(tmp_path / "points.f32")
    env = {k: v for k, v in os.environ.items() if k != "LVRHIP_FAIL_ALLOC_AFTER"}
    r = subprocess.run([_exe(), "oom", str(tmp_path / "points.f32"), str(len(p))], capture_output=True, text=True,
                       timeout=120, env=env)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    assert "out of memory" in r.stdout


KS_EXE = os.path.join(ROOT, "tests", "cpp", "hipksearch_check")


def test_pointset_surface_dropin(tmp_path, oracle_mod):
    """lvr2::HipKSearchSurface (the PointsetSurface-side class) run by tests/cpp/hipksearch_check
    the way PointsetGrid and FastReconstruction drive a surface.  Bars: the flip point is the fp32
    box centroid; normals bit-equal to the Python binding's with that flip point (the same
    library call) and the CudaKSearchSurface-shaped constructor bit-equal to those; distance(v),
    calcDistanceValuesHip (m_distance, m_invalid, the cells left) and the search tree's kSearch /
    radiusSearch bit-exact against the oracle."""
    from lvr2_amd import HipSurface, scenes

    if not os.path.exists(KS_EXE):
        pytest.fail(f"{KS_EXE} is not built (run __graft_entry__.build())")
    w = scenes.plain_small()
    o = oracle_mod.Surface(w.points)
    v, cells, ci = oracle_mod.grid(w.points, o.bbox(), w.voxel, extrude=True)
    rng = np.random.default_rng(11)
    near = w.points[::997]
    q = np.concatenate([v[:: max(1, len(v) // 3000)], near + rng.normal(scale=5, size=near.shape)])
    q = q.astype(np.float32)
    k, r = 12, 9.0
    w.points.tofile(tmp_path / "points.f32")
    v.astype(np.float32).tofile(tmp_path / "lattice.f32")
    cells.astype(np.int32).tofile(tmp_path / "cells.i32")
    ci.astype(np.int32).tofile(tmp_path / "cell_index.i32")
    q.tofile(tmp_path / "queries.f32")
    args = [KS_EXE, "run", str(tmp_path / "points.f32"), str(len(w.points)), str(w.kn), str(w.ki), str(w.kd),
            repr(float(w.voxel)), str(tmp_path / "lattice.f32"), str(len(v)), str(tmp_path / "cells.i32"),
            str(tmp_path / "cell_index.i32"), str(len(cells)), str(tmp_path / "queries.f32"), str(len(q)), str(k),
            repr(r), str(tmp_path)]
    print("running hipksearch_check:", len(w.points), "points", len(v), "vertices", len(q), "queries", flush=True)
    res = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, (res.returncode, res.stdout, res.stderr)
    # init(): the bounding-box centroid, in fp32
    lo, hi = w.points.min(0), w.points.max(0)
    np.testing.assert_array_equal(np.fromfile(tmp_path / "flip.f32", np.float32), lo + np.float32(0.5) * (hi - lo))
    flip = np.fromfile(tmp_path / "flip.f32", np.float32)
    # calculateSurfaceNormals = estimate + interpolate, stored in the point buffer
    nrm = np.fromfile(tmp_path / "normals.f32", np.float32).reshape(-1, 3)
    s = HipSurface(w.points)
    s.setKn(w.kn)
    s.setKi(w.ki)
    s.setFlippoint(*(float(c) for c in flip))
    s.calculateNormals()
    s.interpolateNormals()
    np.testing.assert_array_equal(nrm, s.getNormals())
    np.testing.assert_array_equal(np.fromfile(tmp_path / "cuda_normals.f32", np.float32).reshape(-1, 3), nrm)
    # distance(v): proj, eucl bit-exact given those normals
    d = np.fromfile(tmp_path / "dist_single.f32", np.float32).reshape(-1, 2)
    op, oe = o.distance(nrm, w.kd, q)
    np.testing.assert_array_equal(d[:, 0], op)
    np.testing.assert_array_equal(d[:, 1], oe)
    # calcDistanceValuesHip
    gp, ge = o.distance(nrm, w.kd, v)
    oinv, keep = oracle_mod.calc_distance_values(ge, w.voxel, cells)
    np.testing.assert_array_equal(np.fromfile(tmp_path / "grid_dist.f32", np.float32), gp)
    np.testing.assert_array_equal(np.fromfile(tmp_path / "grid_inv.u8", np.uint8), oinv)
    np.testing.assert_array_equal(np.fromfile(tmp_path / "kept.i32", np.int32), np.nonzero(keep)[0])
    assert 0 < keep.sum() < len(keep)
    # the search tree
    oi, od = o.knn(q, k)
    np.testing.assert_array_equal(np.fromfile(tmp_path / "knn.i32", np.int32).reshape(-1, k), oi)
    np.testing.assert_array_equal(np.fromfile(tmp_path / "knn_d2.f32", np.float32).reshape(-1, k), od)
    ri, rd, rc = o.radius(q, k, r)
    np.testing.assert_array_equal(np.fromfile(tmp_path / "rad.i32", np.int32).reshape(-1, k), ri)
    np.testing.assert_array_equal(np.fromfile(tmp_path / "rad_d2.f32", np.float32).reshape(-1, k), rd)
    assert rc.max() == k and ((rc > 0) & (rc < k)).any()
