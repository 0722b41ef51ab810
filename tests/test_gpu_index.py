"""Index grids (ValueOnIndex + float channel) on the GPU against the oracle, bit-exact, including
an external channel set with nvq_grid_set_channel."""
import numpy as np
import pytest

from conftest import points_near, rng

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


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.fixture(scope="module")
def sphere_index(ref):
    _nvq()
    return ref.index_grid(ref.level_set_sphere(40.0, (1.5, -2.0, 3.0), 0.5, 3.0), 1, True, True)


@pytest.fixture(scope="module")
def fog_index(ref):
    nvq = _nvq()
    return ref.index_grid(nvq.tools.create_fog_fbm([0, 0, 0, 63, 63, 63], seed=3), 1, False, False)


def test_get_value_and_samples(oracle, sphere_index):
    nvq = _nvq()
    d = nvq.DeviceGrid(sphere_index)
    assert d.grid_type == "OnIndex" and d.grid_class == "IndexGrid"
    ijk = rng(4).integers(-110, 110, (100000, 3)).astype(np.int32)
    v, a, dim = nvq.ops.get_value(d, _t(ijk), want_dim=True)
    ov, oa, odim = oracle.get_value(sphere_index, ijk)
    _bits("value", v, ov)
    _bits("active", a, oa)
    _bits("dim", dim, odim.astype(np.int32))
    for n, seed in ((100000, 5), ((1 << 20) + 777, 6)):  # lane-per-point and binned paths
        xyz = points_near(oracle, sphere_index, n, seed=seed)
        for mode in (1, 2, 3):
            nn, val, g = nvq.ops.sample(d, _t(xyz), grad_mode=mode)
            onn, oval, og = oracle.sample(sphere_index, xyz, grad_mode=mode)
            _bits(f"n={n} nn", nn, onn)
            _bits(f"n={n} val", val, oval)
            _bits(f"n={n} grad{mode}", g, og)


def test_rays_on_index_grids(oracle, sphere_index, fog_index):
    nvq = _nvq()
    d = nvq.DeviceGrid(sphere_index)
    e, dr = nvq.tools.camera_rays(96, 64, (20.0, 35.0, 250.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 40.0)
    for iso in (0.0, 0.5):
        got = nvq.ops.iso_crossing(d, _t(e), _t(dr), iso, image_width=96)
        exp = oracle.iso_crossing(sphere_index, e, dr, iso)
        assert int(exp[0].sum()) > 300
        for name, x, y in zip(("hit", "ijk", "t", "v", "normal"), got, exp):
            _bits(f"iso {iso} {name}", x, y)
    f = nvq.DeviceGrid(fog_index)
    e2, d2 = nvq.tools.camera_rays(48, 48, (32.0, 32.0, 200.0), (32.0, 32.0, 32.0), (0.0, 1.0, 0.0), 40.0)
    L, T = nvq.ops.fog_march(f, _t(e2), _t(d2), 0.5, 0.05, image_width=48)
    oL, oT, steps = oracle.fog_march(fog_index, e2, d2, 0.5, 0.05)
    assert steps.max() > 50
    np.testing.assert_allclose(L.cpu().numpy(), oL, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(T.cpu().numpy(), oT, rtol=1e-5, atol=1e-6)
    # both index grids in one persistent launch: identical to the single-grid kernels
    n1 = len(e)
    hit, ijk, a, b, nrm = nvq.ops.multi_query([d, f], _t(np.concatenate([e, e2])), _t(np.concatenate([dr, d2])),
                                              [0, n1, n1 + len(e2)], [96, 48],
                                              ops=[nvq.ops.OP_ZERO_CROSSING, nvq.ops.OP_FOG_MARCH])
    z = nvq.ops.zero_crossing(d, _t(e), _t(dr), image_width=96)
    _bits("multi hit", hit[:n1], z[0].cpu().numpy())
    _bits("multi t", a[:n1], z[2].cpu().numpy())
    _bits("multi fog L", a[n1:], L.cpu().numpy())
    _bits("multi fog T", b[n1:], T.cpu().numpy())
    with pytest.raises(nvq.NvqError):  # an index grid's class does not choose the query
        nvq.ops.multi_query([d, f], _t(np.concatenate([e, e2])), _t(np.concatenate([dr, d2])),
                            [0, n1, n1 + len(e2)], [96, 48])


def test_iso_render_on_an_index_grid(oracle, ref):
    nvq = _nvq()
    src = ref.level_set_sphere(100.0, (-20.0, 0.0, 0.0), 1.0, 3.0)
    ix = ref.index_grid(src, 1, True, True)
    cam = nvq.tools.iso_camera(ix, 256, 160)
    image, hit, t = nvq.ops.render_iso_surface(nvq.DeviceGrid(ix), cam, persistent=True)
    oimg, ohit, ot = oracle.iso_render(ix, cam)
    assert int(ohit.sum()) > 1000
    _bits("image", image.reshape(-1), oimg)
    _bits("t", t.reshape(-1), ot)


def _channel_view(buf):
    meta = int(np.frombuffer(buf[640:648].tobytes(), np.int64)[0])
    data = meta + int(np.frombuffer(buf[meta:meta + 8].tobytes(), np.int64)[0])
    count = int(np.frombuffer(buf[meta + 8:meta + 16].tobytes(), np.uint64)[0])
    return data, count


def test_external_channel(oracle, ref, sphere_index):
    nvq = _nvq()
    d = nvq.DeviceGrid(sphere_index)
    data, count = _channel_view(sphere_index)
    assert count >= d.value_count
    own = np.frombuffer(sphere_index[data:data + 4 * count].tobytes(), np.float32)
    doubled = sphere_index.copy()
    doubled[data:data + 4 * count] = np.frombuffer((own * np.float32(2)).tobytes(), np.uint8)
    xyz = points_near(oracle, sphere_index, 50000, seed=9)
    d.set_channel(torch.from_numpy(own * np.float32(2)).cuda())
    nn, val, g = nvq.ops.sample(d, _t(xyz), grad_mode=1)
    onn, oval, og = oracle.sample(doubled, xyz, grad_mode=1)
    _bits("doubled val", val, oval)
    _bits("doubled grad", g, og)
    d.set_channel(None)
    nn, val, g = nvq.ops.sample(d, _t(xyz), grad_mode=1)
    _bits("own val", val, oracle.sample(sphere_index, xyz, grad_mode=1)[1])
    with pytest.raises(ValueError):
        d.set_channel(torch.zeros(4, dtype=torch.float64, device="cuda"))
    with pytest.raises(nvq.NvqError):
        d.set_channel(torch.zeros(4, device="cuda"))  # fewer floats than the grid indexes


def test_index_grid_without_a_channel(oracle, ref):
    nvq = _nvq()
    src = ref.level_set_sphere(30.0, (0.0, 0.0, 0.0), 1.0, 3.0)
    bare = ref.index_grid(src, 0, True, True)   # no channel in the buffer
    with_ch = ref.index_grid(src, 1, True, True)
    d = nvq.DeviceGrid(bare)
    xyz = points_near(oracle, src, 10000, seed=2)
    with pytest.raises(nvq.NvqError):
        nvq.ops.sample(d, _t(xyz))
    data, count = _channel_view(with_ch)
    ch = np.frombuffer(with_ch[data:data + 4 * count].tobytes(), np.float32).copy()
    d.set_channel(torch.from_numpy(ch).cuda())
    _, val, _ = nvq.ops.sample(d, _t(xyz), grad_mode=1)
    _bits("bare grid + channel", val, oracle.sample(with_ch, xyz, grad_mode=1)[1])
