"""Device-side construction (nvq_voxels_to_grid = voxelsToGrid<float | ValueOnIndex>) on the GPU:
the buffer equals the oracle's byte for byte, the reference's validator accepts it, and an index
grid built from a float grid's active voxels answers every query as that float grid does."""
import numpy as np
import pytest

from conftest import rng, voxel_cases
from pyoracle import have_reference

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

CASES = voxel_cases()


def _nvq():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import openvdb_amd
    return openvdb_amd


def _build(nvq, ijk, grid_type, voxel_size=0.5, name=None):
    d = torch.from_numpy(np.ascontiguousarray(ijk, np.int32)).cuda()
    if grid_type in ("OnIndex", "Index"):
        g, idx = nvq.DeviceGrid.from_voxels(d, grid_type, voxel_size, name=name, return_index=True)
        return g, idx.cpu().numpy().astype(np.uint64)
    return nvq.DeviceGrid.from_voxels(d, grid_type, voxel_size, name=name), None


@pytest.mark.parametrize("grid_type", ["OnIndex", "Index", "float"])
@pytest.mark.parametrize("case", sorted(CASES))
def test_build_matches_oracle(oracle, case, grid_type):
    nvq = _nvq()
    ijk = CASES[case]
    gt = {"OnIndex": 20, "Index": 19, "float": 1}[grid_type]
    print(f"[voxels_to_grid] case={case} type={grid_type} n={len(ijk)}", flush=True)
    g, idx = _build(nvq, ijk, grid_type)
    got = g.download()
    ob, oidx = oracle.voxels_to_grid(ijk, gt, 0.5, want_index=gt != 1)
    assert got.nbytes == ob.nbytes
    bad = np.nonzero(got != ob)[0]
    assert bad.size == 0, f"{bad.size} bytes differ, first at {bad[:8]}"
    if idx is not None:
        np.testing.assert_array_equal(idx, oidx)
    if have_reference():
        from pyoracle import Reference
        assert Reference().validate(got, full=True)
    g.close()


def test_large_random_index_grid(oracle):
    """2^20 random voxels in [-512, 512)^3 (the size of Large_CudaPointsToGrid_DeviceBuffer in
    unittest/TestNanoVDB.cu), two builds in a row on one stream."""
    nvq = _nvq()
    ijk = rng(98765).integers(-512, 512, (1 << 20, 3))
    print("[voxels_to_grid] large n=2^20", flush=True)
    ob, oidx = oracle.voxels_to_grid(ijk, 20, 1.0, want_index=True)
    for _ in range(2):
        g, idx = _build(nvq, ijk, "OnIndex", 1.0)
        got = g.download()
        assert np.array_equal(got, ob)
        np.testing.assert_array_equal(idx, oidx)
        assert g.value_count == int(oidx.max()) + 1
        g.close()


def test_name_is_written(oracle):
    nvq = _nvq()
    g, _ = _build(nvq, CASES["three"], "OnIndex", 0.1, name="points")
    ob, _ = oracle.voxels_to_grid(CASES["three"], 20, 0.1, name="points")
    assert np.array_equal(g.download(), ob)


def test_span_too_wide_is_an_error():
    """More than 2^28 upper-node tiles in the input's bbox is refused (no device fault, no grid)."""
    nvq = _nvq()
    ijk = torch.tensor([[-2**31, -2**31, -2**31], [2**31 - 1, 2**31 - 1, 2**31 - 1]], dtype=torch.int32).cuda()
    with pytest.raises(nvq.NvqError):
        nvq.DeviceGrid.from_voxels(ijk, "OnIndex")


def _active_voxels(buf):
    """(ijk[n,3], value[n]) of the active leaf voxels of a float grid buffer."""
    tree = 672
    off = int(np.frombuffer(buf[tree:tree + 8].tobytes(), np.int64)[0])
    count = int(np.frombuffer(buf[tree + 32:tree + 36].tobytes(), np.uint32)[0])
    leaves = np.frombuffer(buf[tree + off:tree + off + 2144 * count].tobytes(), np.uint8).reshape(count, 2144)
    origin = leaves[:, 0:12].copy().view(np.int32) & ~7
    mask = np.unpackbits(leaves[:, 16:80], axis=1, bitorder="little").astype(bool)
    values = leaves[:, 96:].copy().view(np.float32)
    n = np.arange(512)
    local = np.stack([n >> 6, (n >> 3) & 7, n & 7], 1)
    li, vi = np.nonzero(mask)
    return origin[li] + local[vi], values[li, vi]


def test_index_grid_from_float_voxels_answers_like_the_float_grid():
    """Build an index grid on the GPU from a fog's active voxels (shuffled, with duplicates), fill
    its channel through the returned indices, and query both grids: values, activity and
    trilinear samples agree bit for bit wherever the float grid's voxels are active."""
    nvq = _nvq()
    fog = nvq.tools.create_fog_fbm([0, 0, 0, 95, 95, 95], seed=5)
    ijk, val = _active_voxels(fog)
    r = rng(3)
    perm = np.concatenate([r.permutation(len(ijk)), r.integers(0, len(ijk), 1000)])
    fg = nvq.DeviceGrid(fog)
    ig, idx = nvq.DeviceGrid.from_voxels(torch.from_numpy(ijk[perm].astype(np.int32)).cuda(), "OnIndex",
                                         fg.voxel_size[0], return_index=True)
    channel = torch.zeros(ig.value_count, dtype=torch.float32, device="cuda")
    channel[idx] = torch.from_numpy(val[perm]).cuda()
    ig.set_channel(channel)
    q = torch.from_numpy(r.integers(-4, 100, (200000, 3)).astype(np.int32)).cuda()
    fv, fa, fd = nvq.ops.get_value(fg, q, want_dim=True)
    iv, ia = nvq.ops.get_value(ig, q)
    leaf_active = (fa == 1) & (fd == 1)
    assert bool(torch.all(ia[leaf_active] == 1))
    assert torch.equal(iv[leaf_active], fv[leaf_active])
    assert int((ia == 1).sum()) == int(leaf_active.sum())
    # trilinear samples whose 8 corners are all active leaf voxels
    xyz = torch.from_numpy(r.uniform(0, 95, (200000, 3)).astype(np.float32)).cuda()
    _, fs, _ = nvq.ops.sample(fg, xyz, nn=False, grad_mode=0)
    _, is_, _ = nvq.ops.sample(ig, xyz, nn=False, grad_mode=0)
    base = torch.floor(xyz).to(torch.int32)
    ok = torch.ones(len(xyz), dtype=torch.bool, device="cuda")
    for c in range(8):
        corner = base + torch.tensor([c >> 2, (c >> 1) & 1, c & 1], dtype=torch.int32, device="cuda")
        _, a, d = nvq.ops.get_value(fg, corner, want_dim=True)
        ok &= (a == 1) & (d == 1)
    assert int(ok.sum()) > 1000
    assert torch.equal(is_[ok], fs[ok])


def test_device_built_index_grid_marches_and_differentiates_like_the_fog():
    """voxelsToGrid<ValueOnIndex> from a fog's active voxels, channel filled through the returned
    indices: the fog march (radiance, transmittance) and its adjoint on the index grid equal those
    on the fog bit for bit (the fog has no tiles, no empty leaves and inactive values equal to its
    background, so both grids hold the same values everywhere).  The adjoint's fixed-point sums
    are compared per active voxel: leaf slot of the fog against channel index of the index grid."""
    nvq = _nvq()
    fog = nvq.tools.create_fog_fbm([0, 0, 0, 95, 95, 95], seed=5)
    ijk, val = nvq.tools.active_voxels(fog)
    fg = nvq.DeviceGrid(fog)
    ig, idx = nvq.DeviceGrid.from_voxels(torch.from_numpy(ijk).cuda(), "OnIndex", fg.voxel_size[0], return_index=True)
    channel = torch.zeros(ig.value_count, dtype=torch.float32, device="cuda")
    channel[idx] = torch.from_numpy(val).cuda()
    ig.set_channel(channel)
    eye, dir = nvq.tools.camera_rays(96, 96, (150.0, 130.0, 190.0), (48.0, 48.0, 48.0), (0.0, 1.0, 0.0), 45.0)
    E, D = torch.from_numpy(eye).cuda(), torch.from_numpy(dir).cuda()
    Lf, Tf = nvq.ops.fog_march(fg, E, D, 0.5, 0.05, image_width=96)
    Li, Ti = nvq.ops.fog_march(ig, E, D, 0.5, 0.05, image_width=96)
    assert float(Lf.max()) > 0.0
    assert torch.equal(Lf, Li) and torch.equal(Tf, Ti)
    gL = torch.from_numpy(rng(21).normal(0, 1, len(eye)).astype(np.float32)).cuda()
    af = nvq.ops.fog_march_backward(fg, E, D, Lf, gL, image_width=96)
    ai = nvq.ops.fog_march_backward(ig, E, D, Li, gL, image_width=96)
    offs = np.frombuffer(fog[672:704].tobytes(), np.int64)
    count = int(np.frombuffer(fog[704:708].tobytes(), np.uint32)[0])
    leaves = fog[672 + offs[0]:672 + offs[0] + 2144 * count].reshape(count, 2144)
    li, vi = np.nonzero(np.unpackbits(leaves[:, 16:80], axis=1, bitorder="little").astype(bool))
    slots = torch.from_numpy(li.astype(np.int64) * 512 + vi).cuda()
    assert int((af[slots] != 0).sum()) > 1000
    assert torch.equal(af[slots], ai[idx])


@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_floating_point_voxel_coordinates(oracle, dtype):
    """Vec3f / Vec3d index-space input (PtrT = Vec3f* or Vec3d*): each coordinate rounds to its
    nearest voxel, Floor(x + 0.5) in the input's precision; the grid and the per-voxel indices equal
    the integer build of the rounded coordinates."""
    nvq = _nvq()
    r = rng(4321)
    x = (r.normal(size=(30000, 3)) * 40.0).astype(dtype)
    x[:6] = np.array([[0.5, -0.5, 1.5], [-1.5, 2.5, -2.5], [0.49999997, -0.50000006, 0.0], [7.5, 8.5, -8.5],
                      [1e-8, -1e-8, 3.4999999], [-4095.5, 4095.5, 0.5]], dtype)
    half = np.asarray(0.5, dtype)
    ijk = np.floor(x + half).astype(np.int32)
    print(f"[voxels_to_grid] floating input {dtype}", flush=True)
    g, idx = nvq.DeviceGrid.from_voxels(torch.from_numpy(x).cuda(), "OnIndex", 0.5, return_index=True)
    ob, oidx = oracle.voxels_to_grid(ijk, 20, 0.5, want_index=True)
    got = g.download()
    assert got.nbytes == ob.nbytes and bool((got == ob).all())
    np.testing.assert_array_equal(idx.cpu().numpy().astype(np.uint64), oidx)
    g.close()


@pytest.mark.parametrize("grid_type", ["float", "OnIndex", "Index"])
def test_build_into_reused_pool_memory(oracle, grid_type):
    """Only the node region of a new grid is cleared; the leaf kernels write every leaf byte.  A
    build into stream-ordered pool memory just released by a grid full of non-zero bytes (a point
    grid of the same extent) must still equal the oracle byte for byte."""
    nvq = _nvq()
    print(f"[voxels_to_grid] reused pool memory type={grid_type}", flush=True)
    ijk = rng(777).integers(-300, 300, (150000, 3))
    pts = torch.from_numpy(ijk.astype(np.float32)).cuda()
    for _ in range(2):
        p = nvq.DeviceGrid.from_points(pts.repeat(3, 1), 1.0, "Voxel8")
        p.close()
        g, idx = _build(nvq, ijk, grid_type)
        gt = {"OnIndex": 20, "Index": 19, "float": 1}[grid_type]
        ob, _ = oracle.voxels_to_grid(ijk, gt, 0.5)
        got = g.download()
        assert got.nbytes == ob.nbytes and bool((got == ob).all())
        g.close()
