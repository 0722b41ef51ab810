"""GPU isoCrossing and the iso-surface renderer (tiled and persistent) against the oracle, bit-exact."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _nvq():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import openvdb_amd
    return openvdb_amd


def _bits(name, gpu, ref):
    gpu = np.ascontiguousarray(gpu.cpu().numpy() if hasattr(gpu, "cpu") else gpu).reshape(ref.shape)
    if not np.array_equal(gpu.astype(ref.dtype, copy=False).view(np.uint8), np.ascontiguousarray(ref).view(np.uint8)):
        bad = np.nonzero((gpu.reshape(len(ref), -1) != ref.reshape(len(ref), -1)).any(axis=1))[0]
        raise AssertionError(f"{name}: {len(bad)} mismatches, first {bad[:5]}")


@pytest.fixture(scope="module")
def example_sphere():
    nvq = _nvq()
    return nvq.tools.create_level_set_sphere(100.0, (-20.0, 0.0, 0.0), 1.0, 3.0)


@pytest.mark.parametrize("codec,iso", [("float", 1.0), ("float", -2.25), ("float", 0.0), ("fp16", 0.5), ("fpn", 1.0)])
def test_iso_crossing_matches_oracle(oracle, example_sphere, codec, iso):
    nvq = _nvq()
    g = example_sphere if codec == "float" else nvq.tools.quantize(example_sphere, codec, 1e-3)
    d = nvq.DeviceGrid(g)
    e, dr = nvq.tools.camera_rays(200, 136, (10.0, -30.0, 380.0), (-20.0, 0.0, 0.0), (0.0, 1.0, 0.0), 45.0)
    hit, ijk, t, v, nrm = nvq.ops.iso_crossing(d, torch.from_numpy(e).cuda(), torch.from_numpy(dr).cuda(), iso,
                                               image_width=200)
    oh, oijk, ot, ov, onrm = oracle.iso_crossing(g, e, dr, iso)
    assert int(oh.sum()) > 5000
    for name, x, y in (("hit", hit, oh), ("ijk", ijk, oijk), ("t", t, ot), ("v", v, ov), ("normal", nrm, onrm)):
        _bits(f"{codec} iso {iso} {name}", x, y)


@pytest.mark.parametrize("persistent", [True, False])
def test_render_iso_surface_matches_oracle(oracle, example_sphere, persistent):
    nvq = _nvq()
    d = nvq.DeviceGrid(example_sphere)
    cam = nvq.tools.iso_camera(example_sphere, 300, 201)  # partial 8x8 tiles on both edges
    image, hit, t = nvq.ops.render_iso_surface(d, cam, persistent=persistent)
    oimg, ohit, ot = oracle.iso_render(example_sphere, cam)
    assert 1000 < int(ohit.sum()) < 300 * 201
    _bits("image", image.reshape(-1), oimg)
    _bits("hit", hit.reshape(-1), ohit)
    _bits("t", t.reshape(-1), ot)


def test_render_iso_surface_persistent_equals_tiled_large(example_sphere):
    nvq = _nvq()
    d = nvq.DeviceGrid(example_sphere)
    cam = nvq.tools.iso_camera(example_sphere, 1024, 1024)
    a = nvq.ops.render_iso_surface(d, cam, persistent=True)
    b = nvq.ops.render_iso_surface(d, cam, persistent=False)
    for x, y in zip(a, b):
        assert torch.equal(x, y)
    assert int(a[1].sum()) > 100000
