"""Degenerate and adversarial clouds, and the limits the library does or does not have.

Bars: kNN index sets and d^2 bit-exact against the oracle (ties by index); adaptive k bit-exact;
signed distances bit-exact given the same normals; interpolated normals for ki above 64
bit-exact given the same normals; normals finite and of unit length where the plane is
undefined (identical or collinear points)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _surface(points, **kw):
    from lvr2_amd import HipSurface

    return HipSurface(points, **kw)


@pytest.mark.parametrize("ki", [65, 100, 150])
def test_interpolation_with_ki_above_64(oracle_mod, ki):
    from lvr2_amd import scenes

    w = scenes.plain_small(60_000)
    o = oracle_mod.Surface(w.points)
    n32, _, _ = o.normals(w.kn, 3, w.flip)
    s = _surface(w.points)
    s.setKi(ki)
    s.setNormals(n32)
    s.interpolateNormals()
    np.testing.assert_array_equal(s.getNormals(), o.interpolate(n32, ki))


def test_ki_above_64_limits():
    """A captured graph and a multi-scan handle keep the limit of the kNN pass's ki list."""
    import torch

    from lvr2_amd import scenes
    from lvr2_amd._lib import LvrHipError

    w = scenes.plain_small(40_000)
    s = _surface(w.points)
    s.setKn(10)
    s.setKi(80)
    out = torch.empty((len(w.points), 3), dtype=torch.float32, device="cuda")
    with pytest.raises(LvrHipError, match="ki <= 64"):
        s.captureGraph(out)
    b = _surface(w.points, scan_sizes=[20_000, 20_000])
    b.setKn(10)
    b.setKi(80)
    b.calculateNormals()
    with pytest.raises(LvrHipError, match="single-scan"):
        b.interpolateNormals()


def test_identical_points(oracle_mod):
    """3,000 copies of one point: every neighbour list is the lowest indices at d^2 = 0."""
    p = np.tile(np.array([[1.5, -2.25, 3.0]], np.float32), (3000, 1))
    o = oracle_mod.Surface(p)
    s = _surface(p)
    q = np.concatenate([p[:50], np.array([[1.5, -2.25, 3.5], [0.0, 0.0, 0.0]], np.float32)])
    for k in (5, 32, 50, 200):
        gi, gd = s.kSearch(q, k)
        np.testing.assert_array_equal(gi, np.tile(np.arange(k, dtype=np.int32), (len(q), 1)))
        oi, od = o.knn(q, k)
        np.testing.assert_array_equal(gi, oi)
        np.testing.assert_array_equal(gd, od)
    s.setKn(10)
    s.setKi(10)
    s.calculateNormals()
    np.testing.assert_array_equal(s.adaptiveK(), o.normals(10, 3, (0.0, 0.0, 0.0))[2])
    n = s.getNormals()
    assert np.all(np.isfinite(n))
    # signed distances given the oracle's normals
    n32, _, _ = o.normals(10, 3, (0.0, 0.0, 0.0))
    s.setNormals(n32)
    s.setKd(10)
    gp, ge, _ = s.distances(q, 0.5)
    op, oe = o.distance(n32, 10, q)
    np.testing.assert_array_equal(gp, op)
    np.testing.assert_array_equal(ge, oe)


def test_tied_piles_streamed(oracle_mod):
    """Ten points at x = 0 next to two piles of 10,000 copies at x = 1 and x = 2 (collinear): the
    ten points' k-sets (k = 20 ... 320) take the nearest copies at x = 1 by index, their boxes
    fail every round, and each search's tile list (the 625 leaves of a pile, all at one d^2)
    outgrows a tile's 448 leaves, so the one-target rounds stream and count whole-word bins (a
    pile's own points pass at k = 20, their boxes being empty)."""
    rng = np.random.default_rng(5)
    p = np.concatenate([np.zeros((10, 3)), np.tile([[1.0, 0.0, 0.0]], (10_000, 1)),
                        np.tile([[2.0, 0.0, 0.0]], (10_000, 1))]).astype(np.float32)
    p = p[rng.permutation(len(p))]
    o = oracle_mod.Surface(p)
    s = _surface(p)
    lone = np.flatnonzero(p[:, 0] == 0.0)
    for k in (64, 320):
        gi, gd = s.kSearch(p[lone], k)
        oi, od = o.knn(p[lone], k)
        np.testing.assert_array_equal(gi, oi)
        np.testing.assert_array_equal(gd, od)
    s.setKn(10)
    s.setKi(10)
    s.calculateNormals()
    ak = s.adaptiveK()
    np.testing.assert_array_equal(ak, o.normals(10, 3, (0.0, 0.0, 0.0))[2])
    assert np.all(ak[lone] == 320) and np.all(np.delete(ak, lone) == 20)
    assert np.all(np.isfinite(s.getNormals()))


def test_duplicated_positions(oracle_mod):
    """1,000 positions, each 20 times, shuffled: ties broken by index in every search."""
    rng = np.random.default_rng(3)
    base = rng.uniform(-10, 10, size=(1000, 3)).astype(np.float32)
    p = base[rng.permutation(np.repeat(np.arange(1000), 20))]
    o = oracle_mod.Surface(p)
    s = _surface(p)
    q = np.concatenate([p[::37], base[:40] + np.float32(0.01)])
    for k in (8, 20, 21, 32, 100):
        gi, gd = s.kSearch(q, k)
        oi, od = o.knn(q, k)
        np.testing.assert_array_equal(gi, oi, err_msg=str(k))
        np.testing.assert_array_equal(gd, od, err_msg=str(k))
    gi, gd, gc = s.radiusSearch(q, 40, 1.0)
    oi, od, oc = o.radius(q, 40, 1.0)
    np.testing.assert_array_equal(gc, oc)
    np.testing.assert_array_equal(gi, oi)
    s.setKn(10)
    s.calculateNormals()
    np.testing.assert_array_equal(s.adaptiveK(), o.normals(10, 3, (0.0, 0.0, 0.0))[2])


def test_collinear_points(oracle_mod):
    t = np.linspace(-100, 100, 50_000, dtype=np.float64)
    p = np.column_stack([t, 2 * t, -t]).astype(np.float32)
    o = oracle_mod.Surface(p)
    s = _surface(p)
    q = np.concatenate([p[::499], p[::997] + np.float32(0.3)])
    for k in (10, 50):
        gi, gd = s.kSearch(q, k)
        oi, od = o.knn(q, k)
        np.testing.assert_array_equal(gi, oi)
        np.testing.assert_array_equal(gd, od)
    s.setKn(10)
    s.setKi(10)
    s.calculateNormals()
    np.testing.assert_array_equal(s.adaptiveK(), o.normals(10, 3, (0.0, 0.0, 0.0))[2])
    n = s.getNormals()
    assert np.all(np.isfinite(n))
    np.testing.assert_allclose(np.linalg.norm(n, axis=1), 1.0, atol=1e-5)


def test_extreme_clustering(oracle_mod):
    """Half the points in a 1 mm cube, half spread over 1 km: searches that start in one regime
    and end in the other."""
    rng = np.random.default_rng(4)
    p = np.concatenate([rng.uniform(0, 1e-3, size=(50_000, 3)), rng.uniform(-500, 500, size=(50_000, 3))])
    p = p[rng.permutation(len(p))].astype(np.float32)
    o = oracle_mod.Surface(p)
    s = _surface(p)
    q = np.concatenate([p[::211], rng.uniform(-1e-3, 2e-3, size=(200, 3)).astype(np.float32),
                        rng.uniform(-600, 600, size=(200, 3)).astype(np.float32)])
    for k in (16, 64, 300):
        gi, gd = s.kSearch(q, k)
        oi, od = o.knn(q, k)
        np.testing.assert_array_equal(gi, oi, err_msg=str(k))
        np.testing.assert_array_equal(gd, od, err_msg=str(k))
    s.setKn(20)
    s.calculateNormals()
    np.testing.assert_array_equal(s.adaptiveK(), o.normals(20, 3, (0.0, 0.0, 0.0))[2])
    n32, _, _ = o.normals(20, 3, (0.0, 0.0, 0.0))
    s.setNormals(n32)
    s.setKd(10)
    gp, ge, _ = s.distances(q, 1.0)
    op, oe = o.distance(n32, 10, q)
    np.testing.assert_array_equal(gp, op)
    np.testing.assert_array_equal(ge, oe)


def test_batched_scan_below_the_minimum():
    """A multi-scan handle needs 32 kn points in every non-empty scan (its adaptive rounds
    search each scan alone up to k = 32 kn): a clear error, not a wrong result."""
    from lvr2_amd import scenes
    from lvr2_amd._lib import LvrHipError

    w = scenes.plain_small(20_000)
    b = _surface(w.points, scan_sizes=[19_700, 300])
    b.setKn(10)
    with pytest.raises(LvrHipError, match="at least"):
        b.calculateNormals()
