"""Device-side point grids (nvq_points_to_grid / _f64 = pointsToGrid for float and double world
points) on the GPU:
the buffer equals the oracle's byte for byte for every point encoding, the reference's validator
accepts it, a leaf with more points than its uint16 values count is an error, and the query entry
points refuse a point grid."""
import numpy as np
import pytest

from conftest import rng
from pyoracle import have_reference

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

TYPES = {"PointID": 1, "World64": 2, "World32": 3, "Grid64": 4, "Grid32": 5, "Voxel32": 6, "Voxel16": 7,
         "Voxel8": 8, "Default": 9}


def _nvq():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import openvdb_amd
    return openvdb_amd


def _points(seed, n, scale):
    p = (rng(seed).normal(size=(n, 3)) * scale).astype(np.float32)
    return np.concatenate([p, p[: n // 50]])  # coincident points


CASES = {
    "three": (np.array([[1, 0, 0], [1, 2, 3], [1, 2, 3]], np.float32), 1.0),
    "cloud": (_points(7, 60000, 30.0), 0.5),
    "wide": (_points(8, 20000, 3000.0), 2.0),
    "one_leaf": (rng(9).uniform(0.0, 7.99, (5000, 3)).astype(np.float32), 1.0),
    # 2^20 integer points in [-512, 512)^3 at voxel size 8: Large_CudaPointsToGrid_World64
    # (unittest/TestNanoVDB.cu:1594-1712), whose points are exactly representable in float
    "large": (rng(98765).integers(-512, 512, (1 << 20, 3)).astype(np.float32), 8.0),
    # double (Vec3d) points: the same reference test's own input type, and a Gaussian cloud
    "large64": (rng(98765).integers(-512, 512, (1 << 20, 3)).astype(np.float64), 8.0),
    "cloud64": (_points(17, 60000, 30.0).astype(np.float64) + rng(18).uniform(-1e-3, 1e-3, (61200, 3)), 0.75),
}


@pytest.mark.parametrize("point_type", sorted(TYPES))
@pytest.mark.parametrize("case", sorted(CASES))
def test_points_match_oracle(oracle, case, point_type):
    nvq = _nvq()
    pts, dx = CASES[case]
    print(f"[points_to_grid] case={case} type={point_type} n={len(pts)}", flush=True)
    g = nvq.DeviceGrid.from_points(torch.from_numpy(pts).cuda(), dx, point_type, name="points")
    got = g.download()
    want = oracle.points_to_grid(pts, dx, TYPES[point_type], name="points")
    assert got.nbytes == want.nbytes
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, f"{bad.size} bytes differ, first at {bad[:8]}"
    assert g.grid_type == "PointIndex"
    if case.startswith("large"):  # maxPointsPerLeaf < 1024, as the reference test expects
        leaf0 = 672 + int(np.frombuffer(got[672:680].tobytes(), np.int64)[0])
        counts = np.frombuffer(got[leaf0:leaf0 + 1120 * g.leaf_count].tobytes(), np.uint64).reshape(-1, 140)[:, 11]
        assert counts.sum() == len(pts) and 0 < counts.max() < 1024
    if have_reference():
        from pyoracle import Reference
        assert Reference().validate(got, full=True)
    g.close()


def test_too_many_points_in_a_leaf():
    """65536 points in one voxel: an error, and the next build on the same device succeeds."""
    nvq = _nvq()
    print("[points_to_grid] overfull leaf", flush=True)
    pts = torch.zeros((65536, 3), dtype=torch.float32, device="cuda")
    with pytest.raises(RuntimeError):
        nvq.DeviceGrid.from_points(pts, 1.0, "World32")
    g = nvq.DeviceGrid.from_points(pts[:65535], 1.0, "World32")
    assert list(g.info.node_count) == [1, 1, 1]
    g.close()


def test_queries_refuse_point_grids():
    nvq = _nvq()
    from openvdb_amd import ops
    print("[points_to_grid] queries refused", flush=True)
    pts, dx = CASES["three"]
    g = nvq.DeviceGrid.from_points(torch.from_numpy(pts).cuda(), dx, "Default")
    ijk = torch.zeros((4, 3), dtype=torch.int32, device="cuda")
    with pytest.raises(RuntimeError):
        ops.get_value(g, ijk)
    with pytest.raises(RuntimeError):
        ops.sample(g, ijk.float())
    g.close()


DENSITY = {
    # Large_CudaPointsToGrid_World64_density (unittest/TestNanoVDB.cu:1714-1831)
    "reference": (rng(98765).integers(-512, 512, (1 << 20, 3)).astype(np.float64), 60, 1, 10, "World64"),
    "gauss32": (_points(21, 200000, 50.0), 20, 1, 10, "Default"),
    "few_iterations": (_points(22, 50000, 20.0), 40, 0, 3, "Voxel16"),
}


@pytest.mark.parametrize("case", sorted(DENSITY))
def test_density_matches_oracle(oracle, case):
    """The voxel-size search on the device takes the oracle's voxel sizes (bit for bit) and builds
    the same grid."""
    nvq = _nvq()
    pts, target, tol, iters, ptype = DENSITY[case]
    print(f"[points_to_grid_density] case={case} n={len(pts)} target={target}", flush=True)
    g = nvq.DeviceGrid.from_points_density(torch.from_numpy(pts).cuda(), target, tol, iters, ptype)
    want, dx, most = oracle.points_to_grid_density(pts, target, tol, iters, TYPES[ptype])
    assert g.points_per_voxel == most
    assert g.voxel_size[0] == dx
    got = g.download()
    assert got.nbytes == want.nbytes
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, f"{bad.size} bytes differ, first at {bad[:8]}"
    g.close()


def test_density_degenerate_target_is_an_error():
    nvq = _nvq()
    print("[points_to_grid_density] degenerate target", flush=True)
    pts = torch.zeros((4, 3), dtype=torch.float32, device="cuda")
    with pytest.raises(RuntimeError):
        nvq.DeviceGrid.from_points_density(pts, 1, 0, 10)
