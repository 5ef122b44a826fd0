"""GPU parity of the mesh step: the device marching cubes (FastBox "MC" and BilinearFastBox "PMC"
with optimizePlanarFaces' border-edge moves) against the oracle on the same lattice, signed
distances and invalid flags.

Bar: bit-exact (lattice, distances, invalid flags, mesh vertices and faces), given the same
normals."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _oracle_mesh(oracle_mod, o, v, cells, proj, inv, decomposition):
    pmc = decomposition == "PMC"
    mv, mf, fc = oracle_mod.mesh(v, cells, proj, inv, pmc=pmc, face_cells=True)
    bad = 0
    if pmc:
        mv, _, bad = o.pmc_optimize(mv, mf, fc)
    return mv, mf, bad


def _check(s, o, oracle_mod, points, kd, voxel, decomposition):
    import torch

    vd, cd, _ = s.grid(voxel, True, device=True)
    proj, eucl, inv = s.distances(vd, voxel)
    gv, gf = s.mesh(proj, inv, decomposition)
    v, cells = vd.cpu().numpy(), cd.cpu().numpy()
    ov, oc, _ = oracle_mod.grid(points, o.bbox(), voxel, extrude=True)
    np.testing.assert_array_equal(v, ov)
    np.testing.assert_array_equal(cells, oc)
    n = s.getNormals()
    op, oe = o.distance(n, kd, v)
    np.testing.assert_array_equal(proj.cpu().numpy(), op)
    oinv, _ = oracle_mod.calc_distance_values(oe, voxel, cells)
    np.testing.assert_array_equal(inv.cpu().numpy(), oinv)
    mv, mf, bad = _oracle_mesh(oracle_mod, o, v, cells, op, oinv, decomposition)
    assert bad == 0  # a mesh HalfEdgeMesh accepts
    assert len(mf) > 0
    np.testing.assert_array_equal(gf, mf)
    np.testing.assert_array_equal(gv, mv)
    # the device copy is the same mesh
    dv, df = s.mesh(proj, inv, decomposition, device=True)
    assert isinstance(dv, torch.Tensor) and dv.is_cuda
    np.testing.assert_array_equal(dv.cpu().numpy(), gv)
    np.testing.assert_array_equal(df.cpu().numpy(), gf)
    return gv, gf


@pytest.fixture(scope="module")
def c1(oracle_mod):
    from lvr2_amd import scenes

    w = scenes.plain_small()
    o = oracle_mod.Surface(w.points)
    n32, _, _ = o.normals(w.kn, 3, w.flip)
    return w, o, n32


@pytest.mark.parametrize("decomposition", ["MC", "PMC"])
def test_mesh_equals_oracle_plain(c1, oracle_mod, decomposition):
    from lvr2_amd import HipSurface

    w, o, n32 = c1
    s = HipSurface(w.points)
    s.setKd(w.kd)
    s.setNormals(n32)
    gv, gf = _check(s, o, oracle_mod, w.points, w.kd, w.voxel, decomposition)
    print(f"plain 300k {decomposition}: {len(gv)} vertices, {len(gf)} faces", flush=True)


@pytest.mark.parametrize("decomposition", ["MC", "PMC"])
def test_mesh_equals_oracle_room(oracle_mod, decomposition):
    """A 1M-ray room scan with the device's own normals (handed to the oracle as its normals)."""
    from lvr2_amd import HipSurface, scenes

    w = scenes.room(1600, 625)
    o = oracle_mod.Surface(w.points)
    s = HipSurface(w.points)
    s.setKn(w.kn)
    s.setKi(w.ki)
    s.setKd(w.kd)
    s.setFlippoint(*w.flip)
    s.calculateNormals()
    gv, gf = _check(s, o, oracle_mod, w.points, w.kd, w.voxel, decomposition)
    print(f"room 1M {decomposition}: {len(gv)} vertices, {len(gf)} faces", flush=True)


@pytest.mark.parametrize("radius", [3.0, 2.99])
def test_mesh_of_a_given_field_equals_oracle(oracle_mod, radius):
    """The mesh of a given field on the lattice of a sphere cloud: radius 3 puts lattice vertices
    (multiples of 0.125) exactly on the surface, so FastBox makes zero-area faces, PMC drops them
    and moves the rims of the holes this opens; radius 2.99 gives a closed surface that PMC does
    not change."""
    import torch

    from lvr2_amd import HipSurface

    rng = np.random.default_rng(2)
    d = rng.normal(size=(40000, 3))
    pts = (3.0 * d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    o = oracle_mod.Surface(pts)
    s = HipSurface(pts)
    vd, cd, _ = s.grid(0.25, True, device=True)
    v, cells = vd.cpu().numpy(), cd.cpu().numpy()
    dist = (radius - np.linalg.norm(v.astype(np.float64), axis=1)).astype(np.float32)
    dd = torch.from_numpy(dist).to(vd.device)
    mv, mf = s.mesh(dd, None, "MC")
    pv, pf = s.mesh(dd, None, "PMC")
    for (gv, gf), dec in (((mv, mf), "MC"), ((pv, pf), "PMC")):
        ov, of, bad = _oracle_mesh(oracle_mod, o, v, cells, dist, None, dec)
        assert bad == 0
        np.testing.assert_array_equal(gf, of, err_msg=dec)
        np.testing.assert_array_equal(gv, ov, err_msg=dec)
    p = mv[mf]
    degenerate = (np.all(p[:, 0] == p[:, 1], axis=1) | np.all(p[:, 0] == p[:, 2], axis=1)
                  | np.all(p[:, 1] == p[:, 2], axis=1))
    assert len(pf) == len(mf) - degenerate.sum()
    if radius == 3.0:
        assert degenerate.sum() > 0 and not np.array_equal(pv, mv[np.unique(mf[~degenerate])])
    else:
        assert degenerate.sum() == 0
        np.testing.assert_array_equal(pv, mv)
        np.testing.assert_array_equal(pf, mf)


def test_reconstruct_is_grid_distances_mesh(c1):
    from lvr2_amd import HipSurface

    w, _, n32 = c1
    s = HipSurface(w.points)
    s.setKd(w.kd)
    s.setNormals(n32)
    rv, rf = s.reconstruct(w.voxel)
    vd, _, _ = s.grid(w.voxel, True, device=True)
    proj, _, inv = s.distances(vd, w.voxel)
    mv, mf = s.mesh(proj, inv)
    np.testing.assert_array_equal(rv, mv)
    np.testing.assert_array_equal(rf, mf)


def test_mesh_rejects_bad_arguments(c1):
    import torch

    from lvr2_amd import HipSurface

    w, _, n32 = c1
    s = HipSurface(w.points)
    s.setNormals(n32)
    vd, _, _ = s.grid(w.voxel, True, device=True)
    proj, _, inv = s.distances(vd, w.voxel)
    with pytest.raises(ValueError):
        s.mesh(proj, inv, "DMC")
    with pytest.raises(ValueError):
        s.mesh(proj.cpu(), inv)
    with pytest.raises(ValueError):
        s.mesh(proj, inv[:-1])
    with pytest.raises(ValueError):
        s.mesh(proj.to(torch.float64), inv)
    # the C ABI checks the count against the lattice
    from lvr2_amd._lib import LvrHipError

    with pytest.raises(LvrHipError, match="one value per lattice vertex"):
        s.mesh(proj[:-1].contiguous())
