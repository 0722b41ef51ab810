"""GPU parity: the HIP kernels (through the C ABI) against the oracle on the same inputs.

Bar: voxel lookups, active states, getDim, hit flags and hit voxels bit-exact;
samplers, gradients, t-values and normals bit-exact as well (same FP32
operation order, -ffp-contract=off on both sides); the fog march in FP32
within 1e-5 relative of the FP64 oracle (absolute floor 1e-6).
"""
import numpy as np
import pytest

from conftest import grid_bbox, points_near, rng

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import openvdb_amd
    return openvdb_amd


def _t(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    return t.to(device="cuda", dtype=dtype) if dtype is not None else t.to("cuda")


def _bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def _check_bits(name, gpu, ref):
    gpu = gpu.cpu().numpy() if hasattr(gpu, "cpu") else gpu
    if not _bits_equal(gpu.astype(ref.dtype, copy=False), ref):
        bad = np.nonzero((gpu.reshape(len(ref), -1) != ref.reshape(len(ref), -1)).any(axis=1))[0]
        raise AssertionError(f"{name}: {len(bad)} mismatches, first at {bad[:5]}: gpu={gpu[bad[:3]]} ref={ref[bad[:3]]}")


@pytest.fixture(scope="module")
def nvq():
    return _gpu()


@pytest.fixture(scope="module")
def d_sphere(nvq, sphere):
    return nvq.DeviceGrid(sphere)


@pytest.fixture(scope="module")
def d_torus(nvq, torus):
    return nvq.DeviceGrid(torus)


def test_upload_info(nvq, d_sphere, sphere):
    print("grid info", d_sphere.grid_type, d_sphere.grid_class, d_sphere.leaf_count, d_sphere.nbytes, flush=True)
    assert d_sphere.grid_type == "float"
    assert d_sphere.grid_class == "LevelSet"
    assert d_sphere.leaf_count == 4025
    assert d_sphere.nbytes == sphere.nbytes == 11064704


@pytest.mark.parametrize("n", [1, 63, 65, 4097, 200003])
def test_get_value_sphere(nvq, d_sphere, sphere, oracle, n):
    r = rng(n)
    ijk = r.integers(-112, 112, size=(n, 3)).astype(np.int32)
    val, act, dim = nvq.ops.get_value(d_sphere, _t(ijk), want_dim=True)
    ov, oa, od = oracle.get_value(sphere, ijk)
    _check_bits("value", val, ov)
    _check_bits("active", act, oa)
    _check_bits("dim", dim.cpu().numpy().astype(np.uint32), od)


def test_get_value_far_and_empty(nvq, d_torus, torus, oracle):
    r = rng(5)
    far = r.integers(-(2 ** 30), 2 ** 30, size=(50000, 3)).astype(np.int32)
    edge = np.array([[0, 0, 0], [-1, -1, -1], [4095, 4096, -4097], [-(2 ** 31), 2 ** 31 - 1, 0]], np.int32)
    ijk = np.concatenate([far, edge])
    val, act, dim = nvq.ops.get_value(d_torus, _t(ijk), want_dim=True)
    ov, oa, od = oracle.get_value(torus, ijk)
    _check_bits("value", val, ov)
    _check_bits("active", act, oa)
    _check_bits("dim", dim.cpu().numpy().astype(np.uint32), od)
    # empty batch is a no-op
    e = torch.empty((0, 3), dtype=torch.int32, device="cuda")
    v0, a0 = nvq.ops.get_value(d_torus, e)
    assert v0.numel() == 0 and a0.numel() == 0


@pytest.mark.parametrize("grad_mode", [1, 2, 3])
def test_sample_torus(nvq, d_torus, torus, oracle, grad_mode):
    xyz = points_near(oracle, torus, 300000, seed=11 + grad_mode)
    nn, val, grad = nvq.ops.sample(d_torus, _t(xyz), grad_mode=grad_mode)
    onn, oval, ograd = oracle.sample(torus, xyz, grad_mode=grad_mode)
    _check_bits("nn", nn, onn)
    _check_bits("trilinear", val, oval)
    _check_bits(f"grad{grad_mode}", grad, ograd)


def test_sample_world_and_fd6_edges(nvq, d_torus, torus, oracle):
    r = rng(21)
    xyz = points_near(oracle, torus, 100000, seed=21)
    # values whose +-1 offsets round onto integer cell boundaries
    edge = np.array([[0.99999994, 0.5, 0.5], [-1e-8, 3.0, -2.0], [7.9999995, 8.0, -0.0],
                     [299.99997, 0.0, 79.99999], [-300.0, 1e-30, -1e-30]], np.float32)
    xyz = np.concatenate([xyz, edge, r.normal(0, 1e4, size=(1000, 3)).astype(np.float32)])
    for world in (0, 1):
        nn, val, grad = nvq.ops.sample(d_torus, _t(xyz), world=bool(world), grad_mode=1)
        onn, oval, ograd = oracle.sample(torus, xyz, world=bool(world), grad_mode=1)
        _check_bits(f"nn w{world}", nn, onn)
        _check_bits(f"trilinear w{world}", val, oval)
        _check_bits(f"fd6 w{world}", grad, ograd)


def test_sample_from_voxels_golden(nvq, ref):
    """python/test/TestNanoVDB.py:1330-1390: sphere R=100, dx=0.5, half-width 5."""
    grid = ref.level_set_sphere(100.0, (0, 0, 0), 0.5, 5.0)
    d = nvq.DeviceGrid(grid)
    pts = _t(np.array([[10, 0, 0], [100, 0, 0], [0, 0, 99], [0, 101, 0], [110, 0, 0]], np.float32))
    values = torch.zeros(5, dtype=torch.float32, device="cuda")
    grads = torch.zeros((5, 3), dtype=torch.float32, device="cuda")
    nvq.ops.sample_from_voxels(pts, d, values, grads)
    assert values.cpu().tolist() == [-2.5, 0.0, -1.0, 1.0, 2.5]
    assert grads.cpu().tolist() == [[0, 0, 0], [1, 0, 0], [0, 0, 1], [0, 1, 0], [0, 0, 0]]


def test_grad_stencil(nvq, d_sphere, sphere, oracle):
    r = rng(3)
    ijk = r.integers(-105, 105, size=(100000, 3)).astype(np.int32)
    g, ns, lap, zc = nvq.ops.grad_stencil(d_sphere, _t(ijk))
    og, ons, olap, ozc = oracle.grad_stencil(sphere, ijk)
    _check_bits("grad", g, og)
    _check_bits("normSqGrad", ns, ons)
    _check_bits("laplacian", lap, olap)
    _check_bits("zeroCrossing", zc, ozc)


@pytest.mark.parametrize("tiled", [0, 1])
def test_zero_crossing_sphere_512(nvq, d_sphere, sphere, oracle, ref, tiled):
    """C1: 512x512 pinhole camera at (0,0,400), FOV 45 (RayGenOp), HDDA + zeroCrossing."""
    eye, dir = ref.ray_gen(512, 512, 400.0)
    hit, ijk, t, v, nrm = nvq.ops.zero_crossing(d_sphere, _t(eye), _t(dir), image_width=512 if tiled else 0)
    oh, oijk, ot, ov, onrm = oracle.zero_crossing(sphere, eye, dir)
    assert int(oh.sum()) == 81104
    _check_bits("hit", hit, oh)
    _check_bits("ijk", ijk, oijk)
    _check_bits("t", t, ot)
    _check_bits("v", v, ov)
    _check_bits("normal", nrm, onrm)
    c = 256 * 512 + 256
    print("centre-ray hit", ijk[c].tolist(), float(t[c]), flush=True)


def test_zero_crossing_random_rays(nvq, d_torus, torus, oracle):
    r = rng(8)
    n = 20000
    eye = (r.normal(0, 1, (n, 3)) * 600).astype(np.float32)
    target = r.uniform(-350, 350, (n, 3)).astype(np.float32)
    target[:, 2] *= 0.2
    d = target - eye
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    dir = d.astype(np.float32)
    dir[:7] = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [-1, 0, 0], [0, -1, 0], [0, 0, -1], [0.0, -0.0, 1.0]],
                       np.float32)
    hit, ijk, t, v, nrm = nvq.ops.zero_crossing(d_torus, _t(eye), _t(dir))
    oh, oijk, ot, ov, onrm = oracle.zero_crossing(torus, eye, dir)
    _check_bits("hit", hit, oh)
    _check_bits("ijk", ijk, oijk)
    _check_bits("t", t, ot)
    _check_bits("normal", nrm, onrm)


@pytest.mark.parametrize("kind", [1, 2, 3, 4])
def test_quantised_lookup_and_sample(nvq, ref, oracle, kind):
    """Fp4 / Fp8 / Fp16 / FpN(1e-3) leaves decoded in-register, bit-exact with the reference decode."""
    grid = ref.level_set_sphere(40.0, (1.5, -2.0, 3.0), 0.5, 3.0, kind=kind, tol=1e-3)
    d = nvq.DeviceGrid(grid)
    ijk = rng(kind).integers(-100, 100, size=(100000, 3)).astype(np.int32)
    val, act = nvq.ops.get_value(d, _t(ijk))
    ov, oa, _ = oracle.get_value(grid, ijk)
    _check_bits("value", val, ov)
    _check_bits("active", act, oa)
    xyz = points_near(oracle, grid, 100000, seed=40 + kind)
    nn, sv, g = nvq.ops.sample(d, _t(xyz), grad_mode=1)
    onn, osv, og = oracle.sample(grid, xyz, grad_mode=1)
    _check_bits("nn", nn, onn)
    _check_bits("trilinear", sv, osv)
    _check_bits("fd6", g, og)


def test_fog_march_sphere(nvq, fog_sphere, oracle):
    d = nvq.DeviceGrid(fog_sphere)
    assert d.grid_class == "FogVolume"
    r = rng(17)
    n = 4096
    eye = np.tile(np.array([[3.0, -2.0, 300.0]], np.float32), (n, 1))
    target = np.concatenate([r.uniform(-70, 70, (n, 2)), np.zeros((n, 1))], 1) + np.array([3.0, -2.0, 5.0])
    dd = target - eye
    dir = (dd / np.linalg.norm(dd, axis=1, keepdims=True)).astype(np.float32)
    L, T = nvq.ops.fog_march(d, _t(eye), _t(dir), 0.5, 0.05)
    oL, oT, steps = oracle.fog_march(fog_sphere, eye, dir, 0.5, 0.05)
    L = L.cpu().numpy().astype(np.float64)
    T = T.cpu().numpy().astype(np.float64)
    assert steps.max() > 100
    np.testing.assert_allclose(L, oL, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(T, oT, rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("grad_mode", [1, 2, 3])
def test_sample_binned_large_batch(nvq, d_torus, torus, oracle, grad_mode):
    """Batches of >= 2^20 points take the binned LDS-brick path; results stay bit-exact,
    including points outside the bbox (border bins) and far away (accessor fallback)."""
    r = rng(100 + grad_mode)
    xyz = points_near(oracle, torus, (1 << 20) + 12345, seed=100 + grad_mode)
    far = r.normal(0, 3e3, size=(5000, 3)).astype(np.float32)
    xyz = np.concatenate([xyz, far])
    nn, val, grad = nvq.ops.sample(d_torus, _t(xyz), grad_mode=grad_mode)
    onn, oval, ograd = oracle.sample(torus, xyz, grad_mode=grad_mode)
    _check_bits("nn", nn, onn)
    _check_bits("trilinear", val, oval)
    _check_bits(f"grad{grad_mode}", grad, ograd)


@pytest.mark.parametrize("kind", [3, 4])
def test_sample_binned_quantised_world(nvq, ref, oracle, kind):
    grid = ref.level_set_sphere(60.0, (5.0, -7.0, 2.0), 0.5, 3.0, kind=kind, tol=1e-3)
    d = nvq.DeviceGrid(grid)
    xyz = points_near(oracle, grid, 1 << 20, seed=7 + kind) * np.float32(0.5)  # world space (voxel 0.5)
    nn, val, grad = nvq.ops.sample(d, _t(xyz), world=True, grad_mode=1)
    onn, oval, ograd = oracle.sample(grid, xyz, world=True, grad_mode=1)
    _check_bits("nn", nn, onn)
    _check_bits("trilinear", val, oval)
    _check_bits("fd6", grad, ograd)
