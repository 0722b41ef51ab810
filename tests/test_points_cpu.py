"""Device-side point grids (tools/cuda/PointsToGrid.cuh pointsToGrid, float world points): the
oracle's restatement (CPU only), pinned by the reference's known answer (ThreePointsToGrid in
unittest/TestNanoVDB.cu), by tools::isValid, and by the reference's own PointAccessor reading the
oracle's buffers back: every voxel holds exactly the input points that round into it, and each
encoding decodes to them within its resolution."""
import numpy as np
import pytest

from pyoracle import have_reference

pytestmark = pytest.mark.skipif(not have_reference(), reason="reference sources not built")

TYPES = {"PointID": 1, "World64": 2, "World32": 3, "Grid64": 4, "Grid32": 5, "Voxel32": 6, "Voxel16": 7,
         "Voxel8": 8, "Default": 9}


def _u(b, off, dt, n=1):
    return np.frombuffer(b[off:off + np.dtype(dt).itemsize * n].tobytes(), dt)


def points(seed=3, n=50000, scale=40.0):
    r = np.random.default_rng(seed)
    p = (r.normal(size=(n, 3)) * scale).astype(np.float32)
    return np.concatenate([p, p[:500]])  # coincident points too


def blind(buf):
    meta = int(_u(buf, 640, np.int64)[0])
    count, size = int(_u(buf, meta + 8, np.uint64)[0]), int(_u(buf, meta + 16, np.uint32)[0])
    return buf[meta + 288:meta + 288 + count * size], size


def test_reference_kat_three_points(oracle, ref):
    pts = np.array([[1, 0, 0], [1, 2, 3], [1, 2, 3]], np.float32)
    g = oracle.points_to_grid(pts, 1.0, TYPES["Default"])
    assert len(g) == 672 + 64 + 96 + 32 + 270400 + 33856 + 1120 + 288 + 3 * 12
    assert ref.validate(g, full=True)
    assert list(_u(g, 632, np.uint32, 2)) == [5, 23]  # PointData, PointIndex
    leaf = 672 + int(_u(g, 672, np.int64)[0])
    vals = _u(g, leaf + 96, np.uint16, 512)
    n = np.arange(512)
    xyz = np.stack([n >> 6, (n >> 3) & 7, n & 7], 1)
    lex = lambda c: (c[:, 0] * 64 + c[:, 1] * 8 + c[:, 2])
    expect = np.where(lex(xyz) < 64, 0, np.where(lex(xyz) < 64 + 16 + 3, 1, 3))
    np.testing.assert_array_equal(vals, expect)
    count, first = ref.point_voxels(g, np.array([[0, 0, 0], [1, 0, 0], [1, 2, 3]]))
    assert list(count) == [0, 1, 2] and list(first[1:]) == [0, 1]
    data, size = blind(g)
    np.testing.assert_array_equal(np.frombuffer(data.tobytes(), np.float32).reshape(-1, 3), pts)


@pytest.mark.parametrize("name", sorted(TYPES))
def test_point_accessor_reads_back_every_point(oracle, ref, name):
    pts = points()
    dx = 0.5
    g = oracle.points_to_grid(pts, dx, TYPES[name])
    assert ref.validate(g, full=True)
    ijk = np.floor(pts * np.float32(1.0 / dx) + np.float32(0.5)).astype(np.int32)
    uniq, inv = np.unique(ijk, axis=0, return_inverse=True)
    inv = inv.reshape(-1)
    count, first = ref.point_voxels(g, uniq)
    np.testing.assert_array_equal(count.astype(np.int64), np.bincount(inv))
    data, size = blind(g)
    order = np.argsort(inv, kind="stable")  # input points grouped by voxel, input order inside a voxel
    pos = np.repeat(first.astype(np.int64), count.astype(np.int64)) + \
        (np.arange(len(pts)) - np.repeat(np.cumsum(count.astype(np.int64)) - count.astype(np.int64), count.astype(np.int64)))
    rec = np.frombuffer(data.tobytes(), np.uint8).reshape(len(pts), size)[pos]
    want = pts[order]
    vox = uniq[inv[order]].astype(np.float64)
    if name == "PointID":
        np.testing.assert_array_equal(rec.copy().view(np.uint32).reshape(-1), order)
    elif name in ("World32", "Default"):
        np.testing.assert_array_equal(rec.copy().view(np.float32), want)
    elif name == "World64":
        np.testing.assert_array_equal(rec.copy().view(np.float64), want.astype(np.float64))
    elif name in ("Grid32", "Grid64"):
        grid = rec.copy().view(np.float32 if name == "Grid32" else np.float64).astype(np.float64)
        np.testing.assert_allclose(grid, want.astype(np.float64) / dx, rtol=2e-7, atol=1e-6)
        assert np.all(np.floor(grid + 0.5) == vox)
    else:
        off = {"Voxel32": lambda r: r.copy().view(np.float32).astype(np.float64),
               "Voxel16": lambda r: r.copy().view(np.uint16).astype(np.float64) / 65535.0 - 0.5,
               "Voxel8": lambda r: r.astype(np.float64) / 255.0 - 0.5}[name](rec)
        tol = {"Voxel32": 1e-5, "Voxel16": 2.0 / 65535.0, "Voxel8": 2.0 / 255.0}[name]
        np.testing.assert_allclose((vox + off) * dx, want.astype(np.float64), atol=tol * dx + 1e-5)


@pytest.mark.parametrize("name", sorted(TYPES))
def test_double_points(oracle, ref, name):
    """Vec3d input (PtrT = Vec3d*): voxels rounded in FP64, the reference's PointAccessor reads every
    point back, and each encoding decodes to the double input within its resolution."""
    r = np.random.default_rng(11)
    pts = np.concatenate([r.normal(size=(40000, 3)) * 30.0, r.integers(-512, 512, (4000, 3)).astype(np.float64)])
    dx = 0.75
    g = oracle.points_to_grid(pts, dx, TYPES[name])
    assert ref.validate(g, full=True)
    ijk = np.floor(pts * (1.0 / dx) + 0.5).astype(np.int32)
    uniq, inv = np.unique(ijk, axis=0, return_inverse=True)
    inv = inv.reshape(-1)
    count, first = ref.point_voxels(g, uniq)
    np.testing.assert_array_equal(count.astype(np.int64), np.bincount(inv))
    data, size = blind(g)
    c = count.astype(np.int64)
    order = np.argsort(inv, kind="stable")
    pos = np.repeat(first.astype(np.int64), c) + (np.arange(len(pts)) - np.repeat(np.cumsum(c) - c, c))
    rec = np.frombuffer(data.tobytes(), np.uint8).reshape(len(pts), size)[pos]
    want = pts[order]
    vox = uniq[inv[order]].astype(np.float64)
    if name == "PointID":
        np.testing.assert_array_equal(rec.copy().view(np.uint32).reshape(-1), order)
    elif name in ("World64", "Default"):
        assert size == 24
        np.testing.assert_array_equal(rec.copy().view(np.float64), want)
    elif name == "World32":
        np.testing.assert_array_equal(rec.copy().view(np.float32), want.astype(np.float32))
    elif name == "Grid64":
        np.testing.assert_array_equal(rec.copy().view(np.float64), want * (1.0 / dx))
    elif name == "Grid32":
        grid = rec.copy().view(np.float32).astype(np.float64)
        np.testing.assert_allclose(grid, want / dx, rtol=2e-7, atol=1e-6)
    else:
        off = {"Voxel32": lambda r: r.copy().view(np.float32).astype(np.float64),
               "Voxel16": lambda r: r.copy().view(np.uint16).astype(np.float64) / 65535.0 - 0.5,
               "Voxel8": lambda r: r.astype(np.float64) / 255.0 - 0.5}[name](rec)
        tol = {"Voxel32": 1e-6, "Voxel16": 2.0 / 65535.0, "Voxel8": 2.0 / 255.0}[name]
        np.testing.assert_allclose((vox + off) * dx, want, atol=tol * dx + 1e-9)


def test_double_and_float_integer_points_share_the_tree(oracle):
    """Integer points are exact in both precisions: the float and double builds differ only in
    the blind data, and PointID blind data (indices) not at all."""
    pts = np.random.default_rng(98765).integers(-512, 512, (1 << 14, 3))
    a = oracle.points_to_grid(pts.astype(np.float32), 8.0, TYPES["PointID"])
    b = oracle.points_to_grid(pts.astype(np.float64), 8.0, TYPES["PointID"])
    np.testing.assert_array_equal(a, b)


def test_density_search_reference_case(oracle, ref):
    """Large_CudaPointsToGrid_World64_density (unittest/TestNanoVDB.cu:1714-1831): 2^20 Vec3d points
    with integer coordinates in [-512, 512)^3, 60 points per voxel wanted, tolerance 1.  The
    reference expects the densest voxel within the tolerance of 60 and at most 60 * 512 points per
    leaf.  The reported count is recomputed here from the chosen voxel size with numpy, and the
    grid equals the fixed-voxel-size build at that size."""
    pts = np.random.default_rng(98765).integers(-512, 512, (1 << 20, 3)).astype(np.float64)
    g, dx, most = oracle.points_to_grid_density(pts, 60, 1, 10, TYPES["World64"])
    assert abs(most - 60) <= 1
    assert ref.validate(g, full=True)
    ijk = np.floor(pts * (1.0 / dx) + 0.5).astype(np.int64)
    _, counts = np.unique(ijk, axis=0, return_counts=True)
    assert counts.max() == most
    _, leaf_counts = np.unique(ijk >> 3, axis=0, return_counts=True)
    assert leaf_counts.max() <= 60 * 512
    np.testing.assert_array_equal(g, oracle.points_to_grid(pts, dx, TYPES["World64"]))


def test_density_search_stops(oracle):
    """A single point stops at once (voxel size 1); max_iterations bounds the counts; a target
    above the point count is clamped to it."""
    pts = (np.random.default_rng(5).normal(size=(20000, 3)) * 50).astype(np.float32)
    _, dx, most = oracle.points_to_grid_density(pts[:1], 20)
    assert (dx, most) == (1.0, 1)
    _, dx1, most1 = oracle.points_to_grid_density(pts, 20, 0, 1)
    assert dx1 == 1.0  # one count, no update
    ijk = np.floor(pts * np.float32(1.0) + np.float32(0.5)).astype(np.int64)
    assert most1 == np.unique(ijk, axis=0, return_counts=True)[1].max()
    _, dx2, most2 = oracle.points_to_grid_density(pts[:50], 1000, 0, 10)
    assert most2 <= 50


def test_density_search_degenerate_target(oracle):
    """A target of one point per voxel with coincident points drives the interpolated voxel size
    to 0: an error, not a grid."""
    pts = np.zeros((4, 3), np.float32)
    with pytest.raises(ValueError):
        oracle.points_to_grid_density(pts, 1, 0, 10)


def test_too_many_points_in_a_leaf(oracle):
    """More than 65535 points in one leaf: the reference throws; the oracle reports -3."""
    pts = np.zeros((65536, 3), np.float32)
    with pytest.raises(ValueError):
        oracle.points_to_grid(pts, 1.0, TYPES["World32"])
