"""The device API (include/nvq_device.h) inside client kernels: tests/device_api/api_kernels.hip is
written against that header alone, the way a NanoVDB user writes kernels against NanoVDB.h
(grid.getAccessor(), createSampler<0|1>, GradStencil, BoxStencil, Ray::worldToIndexF /
applyInverseMap, zeroCrossing, firstActive, TreeMarcher, DDA).  Every output is compared with the
oracle bit for bit; the oracle itself is pinned to the compiled reference by test_spans_cpu.py and
test_oracle_cpu.py.

Reference call shapes: NanoVDB.h:2210, :5370-5447; math/SampleFromVoxels.h:49-53, :325-476;
math/Stencils.h:109-152, :306-432, :455-576; math/Ray.h:223-268; math/HDDA.h:189-214, :258-536.
"""
import ctypes
import os

import numpy as np
import pytest

from conftest import ROOT, points_near, rng
from test_spans_cpu import large_coordinate_rays, large_coordinate_sphere, rays_for

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

API_LIB = os.path.join(ROOT, "tests", "device_api", "libnvq_api_test.so")


@pytest.fixture(scope="module")
def api():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import openvdb_amd
    openvdb_amd.lib()
    if not os.path.exists(API_LIB):
        raise RuntimeError("tests/device_api/libnvq_api_test.so is not built (openvdb_amd.build.build())")
    L = ctypes.CDLL(API_LIB)
    vp, u64, i = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int
    L.api_accessor.argtypes = [vp, vp, u64, vp, vp, vp, vp, vp, vp]
    L.api_sampler.argtypes = [vp, vp, u64, i, vp, vp, vp, vp, vp]
    L.api_grad_stencil.argtypes = [vp, vp, vp, u64, vp, vp, vp, vp, vp, vp, vp]
    L.api_box_stencil.argtypes = [vp, vp, u64, vp, vp, vp]
    L.api_zero_crossing.argtypes = [vp, vp, vp, u64, i, vp, vp, vp, vp, vp, vp]
    L.api_first_active.argtypes = [vp, vp, vp, u64, vp, vp, vp, vp]
    L.api_tree_march.argtypes = [vp, vp, vp, u64, i, ctypes.c_uint32, vp, vp, vp, vp, vp]
    L.api_dda_trace.argtypes = [vp, vp, vp, ctypes.c_uint32, i, i, vp, vp, vp, vp, vp]
    return L


def _d(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _out(shape, dtype):
    return torch.zeros(shape, dtype=dtype, device="cuda")


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _same(name, gpu, ref):
    g = np.ascontiguousarray(gpu.cpu().numpy())
    r = np.ascontiguousarray(ref)
    assert g.shape == r.shape, (name, g.shape, r.shape)
    if not np.array_equal(g.view(np.uint8), r.view(np.uint8)):
        bad = np.nonzero((g.reshape(len(r), -1) != r.reshape(len(r), -1)).any(axis=1))[0]
        raise AssertionError(f"{name}: {len(bad)} mismatches, first {bad[:4]}: gpu={g[bad[:2]]} oracle={r[bad[:2]]}")


@pytest.fixture(scope="module")
def grids(ref):
    from openvdb_amd import tools
    sphere = ref.level_set_sphere(40.0, (1.5, -2.0, 0.5), 1.0, 3.0)
    out = {
        "sphere": sphere,
        "torus": ref.level_set_torus(60.0, 20.0, (0, 0, 0), 1.0, 3.0),
        "fog_sphere": ref.fog_volume_sphere(60.0, (3.0, -2.0, 5.0), 1.0, 3.0),
        "fbm_tiles": tools.create_fog_fbm((0, 0, 0, 159, 159, 159), seed=3, amplitude=8.0),
        "scaled": ref.level_set_sphere(10.0, (6.0, 8.0, 10.0), 0.5, 3.0),
        "index": ref.index_grid(sphere, 1),
    }
    for kind, name in ((ref.FP4, "fp4"), (ref.FP8, "fp8"), (ref.FP16, "fp16"), (ref.FPN, "fpn")):
        out[name] = ref.quantize(sphere, kind)
    return out


ALL = ["sphere", "torus", "fog_sphere", "fbm_tiles", "scaled", "index", "fp4", "fp8", "fp16", "fpn"]
FLOAT = ["sphere", "torus", "fog_sphere", "fbm_tiles", "scaled"]


def _device(buf):
    import openvdb_amd
    return openvdb_amd.DeviceGrid(buf)


@pytest.mark.parametrize("name", ALL)
def test_accessor(api, oracle, grids, name):
    buf = grids[name]
    print(f"[device api] accessor {name}", flush=True)
    ijk = np.floor(points_near(oracle, buf, 60000, 1, bbox_pad=20)).astype(np.int32)
    ijk[:16] = [[2 ** 30, 0, 0], [-(2 ** 31), 5, 5], [0, 0, 0], [-1, -1, -1]] * 4
    d = _device(buf)
    n = len(ijk)
    val, act, pval, pact, dim = (_out(n, torch.float32), _out(n, torch.uint8), _out(n, torch.float32),
                                 _out(n, torch.uint8), _out(n, torch.int32))
    assert api.api_accessor(d.handle, _ptr(_d(ijk)), n, _ptr(val), _ptr(act), _ptr(pval), _ptr(pact), _ptr(dim),
                            None) == 0
    torch.cuda.synchronize()
    oval, oact, odim = oracle.get_value(buf, ijk)
    _same("getValue", val, oval)
    _same("isActive", act, oact)
    _same("probeValue value", pval, oval)
    _same("probeValue state", pact, oact)
    _same("getDim", dim, odim.astype(np.int32))


@pytest.mark.parametrize("name", ALL)
@pytest.mark.parametrize("cached", [0, 1])
def test_samplers(api, oracle, grids, name, cached):
    buf = grids[name]
    print(f"[device api] createSampler {name} cached={cached}", flush=True)
    xyz = points_near(oracle, buf, 50000, 2)
    d = _device(buf)
    n = len(xyz)
    nn, val, grad, zc = _out(n, torch.float32), _out(n, torch.float32), _out((n, 3), torch.float32), _out(n, torch.uint8)
    assert api.api_sampler(d.handle, _ptr(_d(xyz)), n, cached, _ptr(nn), _ptr(val), _ptr(grad), _ptr(zc), None) == 0
    torch.cuda.synchronize()
    onn, oval, og = oracle.sample(buf, xyz, grad_mode=3)
    _same("sampler<0>", nn, onn)
    _same("sampler<1>", val, oval)
    _same("sampler<1>.gradient", grad, og)
    _same("sampler<1>.zeroCrossing", zc, oracle.sample_zero_crossing(buf, xyz))


@pytest.mark.parametrize("name", FLOAT + ["fp8", "fpn"])
def test_stencils(api, oracle, grids, name):
    buf = grids[name]
    print(f"[device api] GradStencil / BoxStencil {name}", flush=True)
    xyz = points_near(oracle, buf, 50000, 3)
    ijk = np.floor(xyz).astype(np.int32)
    V = rng(4).normal(size=ijk.shape).astype(np.float32)
    V[::5] = 0.0
    d = _device(buf)
    n = len(ijk)
    outs = [_out((n, 3), torch.float32), _out(n, torch.float32), _out(n, torch.float32), _out(n, torch.uint8),
            _out((n, 3), torch.float32), _out((n, 3), torch.float32)]
    assert api.api_grad_stencil(d.handle, _ptr(_d(ijk)), _ptr(_d(V)), n, *[_ptr(o) for o in outs], None) == 0
    interp, bgrad = _out(n, torch.float32), _out((n, 3), torch.float32)
    assert api.api_box_stencil(d.handle, _ptr(_d(xyz)), n, _ptr(interp), _ptr(bgrad), None) == 0
    torch.cuda.synchronize()
    g, ns, lap, zc = oracle.grad_stencil(buf, ijk)
    up, cpt = oracle.grad_stencil_ext(buf, ijk, V)
    for nm, a, b in zip(["gradient", "normSqGrad", "laplacian", "zeroCrossing", "gradient(V)", "cpt"], outs,
                        [g, ns, lap, zc, up, cpt]):
        _same("GradStencil." + nm, a, b)
    oi, og = oracle.box_stencil(buf, xyz)
    _same("BoxStencil.interpolation", interp, oi)
    _same("BoxStencil.gradient", bgrad, og)


def _ray_outputs(n):
    return (_out(n, torch.uint8), _out((n, 3), torch.int32), _out(n, torch.float32), _out(n, torch.float32),
            _out((n, 3), torch.float32))


@pytest.mark.parametrize("name", ["sphere", "torus", "scaled", "index", "fp16", "fpn"])
def test_zero_crossing_and_first_active(api, oracle, ref, grids, name):
    buf = grids[name]
    print(f"[device api] zeroCrossing / firstActive {name}", flush=True)
    eye, dirs = rays_for(buf, 8192, 5)
    if name == "sphere":
        e2, d2 = ref.ray_gen(64, 64, 160.0)
        eye, dirs = np.concatenate([eye, e2]), np.concatenate([dirs, d2])
    d = _device(buf)
    n = len(eye)
    out = _ray_outputs(n)
    assert api.api_zero_crossing(d.handle, _ptr(_d(eye)), _ptr(_d(dirs)), n, 0, *[_ptr(o) for o in out], None) == 0
    fa = (_out(n, torch.uint8), _out((n, 3), torch.int32), _out(n, torch.float32))
    assert api.api_first_active(d.handle, _ptr(_d(eye)), _ptr(_d(dirs)), n, *[_ptr(o) for o in fa], None) == 0
    torch.cuda.synchronize()
    o = oracle.zero_crossing(buf, eye, dirs)
    assert o[0].sum() > 100
    for nm, a, b in zip(["hit", "ijk", "t", "v", "normal"], out, o):
        _same("zeroCrossing " + nm, a, b)
    of = oracle.first_active(buf, eye, dirs)
    for nm, a, b in zip(["hit", "ijk", "t"], fa, of):
        _same("firstActive " + nm, a, b)


def test_zero_crossing_fp64_ray_map(api, oracle, ref):
    """Ray::applyInverseMap (FP64 map) in a client kernel and through nvq_zero_crossing with
    nvq_grid_set_ray_map(NVQ_RAY_MAP_F64), on a grid far from the origin."""
    import openvdb_amd as nvq
    buf, center = large_coordinate_sphere(ref)
    eye, dirs = large_coordinate_rays(center, 8192, 22)
    d = _device(buf)
    n = len(eye)
    o64 = oracle.zero_crossing(buf, eye, dirs, f64map=True)
    o32 = oracle.zero_crossing(buf, eye, dirs, f64map=False)
    for f64, o in ((1, o64), (0, o32)):
        print(f"[device api] zeroCrossing f64map={f64}", flush=True)
        out = _ray_outputs(n)
        assert api.api_zero_crossing(d.handle, _ptr(_d(eye)), _ptr(_d(dirs)), n, f64, *[_ptr(x) for x in out],
                                     None) == 0
        torch.cuda.synchronize()
        for nm, a, b in zip(["hit", "ijk", "t", "v", "normal"], out, o):
            _same(f"zeroCrossing(f64map={f64}) {nm}", a, b)
    d.set_ray_map("f64")
    hit, ijk, t, v, nrm = nvq.ops.zero_crossing(d, _d(eye), _d(dirs))
    for nm, a, b in zip(["hit", "ijk", "t", "v", "normal"], (hit, ijk, t, v, nrm), o64):
        _same(f"nvq_zero_crossing(F64 map) {nm}", a, b)
    d.set_ray_map("f32")
    hit, ijk, t, v, nrm = nvq.ops.zero_crossing(d, _d(eye), _d(dirs))
    _same("nvq_zero_crossing(F32 map) ijk", ijk, o32[1])
    assert ((o64[1] != o32[1]).any(axis=1) | (o64[2] != o32[2])).any()


@pytest.mark.parametrize("name", ["sphere", "fog_sphere", "fbm_tiles", "fp8", "fpn"])
@pytest.mark.parametrize("level", [0, 1, 2])
def test_tree_march(api, oracle, grids, name, level):
    buf = grids[name]
    print(f"[device api] TreeMarcher level={level} {name}", flush=True)
    eye, dirs = rays_for(buf, 4096, 6 + level)
    d = _device(buf)
    n, m = len(eye), 128
    cnt, t0, t1, node = _out(n, torch.int32), _out((n, m), torch.float32), _out((n, m), torch.float32), \
        _out((n, m), torch.int64)
    assert api.api_tree_march(d.handle, _ptr(_d(eye)), _ptr(_d(dirs)), n, level, m, _ptr(cnt), _ptr(t0), _ptr(t1),
                              _ptr(node), None) == 0
    torch.cuda.synchronize()
    oc, o0, o1, on = oracle.tree_march(buf, eye, dirs, level, m)
    assert oc.max() < m and oc.sum() > 0
    _same("span count", cnt, oc.astype(np.int32))
    _same("t0", t0, o0)
    _same("t1", t1, o1)
    _same("node", node, on)


@pytest.mark.parametrize("dim", [1, 8, 128])
def test_dda(api, oracle, dim):
    r = rng(30 + dim)
    nr, steps = 256, 40
    eye = (r.normal(size=(nr, 3)) * 50).astype(np.float32)
    v = r.normal(size=(nr, 3))
    v[:16, 1:] = 0.0
    dirs = (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)
    span = np.tile(np.array([1e-5, 300.0], np.float32), (nr, 1))
    outs = [_out((nr, steps + 1), torch.float32), _out((nr, steps + 1), torch.float32),
            _out((nr, steps + 1, 3), torch.int32), _out((nr, steps + 1), torch.uint8)]
    assert api.api_dda_trace(_ptr(_d(eye)), _ptr(_d(dirs)), _ptr(_d(span)), nr, dim, steps, *[_ptr(o) for o in outs],
                             None) == 0
    torch.cuda.synchronize()
    for k in range(nr):
        o = oracle.dda_trace(eye[k], dirs[k], 1e-5, 300.0, dim, steps)
        for nm, a, b in zip(["time", "next", "voxel", "ok"], outs, o):
            _same(f"DDA<{dim}> ray {k} {nm}", a[k], b)
