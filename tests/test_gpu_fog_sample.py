"""Point queries on fog volumes (the fog half of the headline point-query workload): nvq_sample on
Float fog grids against the oracle, bit-exact, in all three gradient modes, on the lane-per-query
path (batches below 2^20) and the binned LDS-brick path (2^20 points and more).

Grids: the headline 512^3 fBm fog itself (seed 7, amplitude 1) with points of the headline
distribution; fogs whose constant regions are pruned into active tiles (createFogVolumeSphere, and
fBm at amplitude 8, clamped to 1 over whole leaves); a small fBm fog in world space.  The reference
op these follow is python/cuda/PySampleFromVoxels.cu:31-52 (FD6) on top of
math/SampleFromVoxels.h:423-476.
"""
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


def _check_bits(name, gpu, ref):
    gpu = np.ascontiguousarray(gpu.cpu().numpy() if hasattr(gpu, "cpu") else gpu)
    ref = np.ascontiguousarray(ref)
    if gpu.shape != ref.shape or not np.array_equal(gpu.view(np.uint8), ref.view(np.uint8)):
        rows = np.nonzero((gpu.reshape(len(ref), -1) != ref.reshape(len(ref), -1)).any(axis=1))[0]
        raise AssertionError(f"{name}: {len(rows)} mismatches, first at {rows[:5]}: gpu={gpu[rows[:3]]} "
                             f"oracle={ref[rows[:3]]}")


def _compare(nvq, oracle, grid_np, d, xyz, mode, world=False, tag=""):
    print(f"[fog sample] {tag} n={len(xyz)} grad_mode={mode} world={world}", flush=True)
    nn, val, grad = nvq.ops.sample(d, torch.from_numpy(np.ascontiguousarray(xyz)).cuda(), world=world,
                                   grad_mode=mode)
    onn, oval, ograd = oracle.sample(grid_np, xyz, world=world, grad_mode=mode)
    _check_bits(f"{tag} nn", nn, onn)
    _check_bits(f"{tag} trilinear", val, oval)
    _check_bits(f"{tag} grad{mode}", grad, ograd)
    return oval


@pytest.fixture(scope="module")
def headline():
    """The 512^3 fog and 2^21 points of the headline batch: its first 2^20 (uniform in the bbox)
    and its last 2^20 (Gaussian, sigma 2, around the density's 0.15 contour)."""
    nvq = _nvq()
    from bench import c2_fog_grid, c2_points
    fog_np = c2_fog_grid(nvq, 512)
    torus = nvq.DeviceGrid(nvq.tools.create_level_set_torus(300.0, 80.0, (0, 0, 0), 1.0, 3.0))
    fog = nvq.DeviceGrid(fog_np)
    n = 1 << 26
    p_torus, p_fog, _ = c2_points(torch, nvq, torus, fog, n)
    del p_torus
    half = 1 << 20
    pts = torch.cat([p_fog[:half], p_fog[n - half:]]).cpu().numpy()
    del p_fog
    torus.close()
    torch.cuda.empty_cache()
    return fog_np, fog, pts


@pytest.mark.parametrize("mode", [1, 2, 3])
def test_headline_fog_binned(oracle, headline, mode):
    nvq = _nvq()
    fog_np, fog, pts = headline
    assert fog.leaf_count == 141950
    oval = _compare(nvq, oracle, fog_np, fog, pts, mode, tag="fbm512 binned")
    assert 0.05 < float((oval > 0).mean()) < 0.95  # both empty space and fog are sampled


def test_headline_fog_direct(oracle, headline):
    """The same points in a batch below the binned threshold (lane-per-query kernel)."""
    nvq = _nvq()
    fog_np, fog, pts = headline
    sub = np.ascontiguousarray(pts[::5][: (1 << 20) - 1])
    for mode in (1, 2, 3):
        _compare(nvq, oracle, fog_np, fog, sub, mode, tag="fbm512 direct")


@pytest.fixture(scope="module")
def tiled_fogs(ref):
    nvq = _nvq()
    return {
        # interior pruned into active tiles of value 1 at the leaf and lower-node levels
        "fog_sphere": ref.fog_volume_sphere(70.0, (3.0, -2.0, 5.0), 1.0, 3.0),
        # fBm clamped to 1 over whole leaves: active tiles inside the noise
        "fbm_amp8": nvq.tools.create_fog_fbm([0, 0, 0, 191, 191, 191], seed=5, amplitude=8.0),
        "fbm_amp1": nvq.tools.create_fog_fbm([-40, -8, 16, 87, 119, 143], seed=9, amplitude=1.0),
    }


def _tile_count(buf):
    """Active tiles at the leaf (lower-node table) level, from the buffer's TreeData::mTileCount."""
    return int(np.frombuffer(buf[672 + 44:672 + 48].tobytes(), np.uint32)[0])


@pytest.mark.parametrize("name", ["fog_sphere", "fbm_amp8", "fbm_amp1"])
@pytest.mark.parametrize("mode", [1, 2, 3])
def test_tiled_fog_direct_and_binned(oracle, tiled_fogs, name, mode):
    nvq = _nvq()
    g = tiled_fogs[name]
    if name != "fbm_amp1":
        assert _tile_count(g) > 0, "the grid should hold active tiles"
    d = nvq.DeviceGrid(g)
    small = points_near(oracle, g, 200003, seed=70 + mode)
    _compare(nvq, oracle, g, d, small, mode, tag=f"{name} direct")
    big = points_near(oracle, g, (1 << 20) + 4099, seed=80 + mode)
    far = rng(90 + mode).normal(0, 2e3, size=(3000, 3)).astype(np.float32)
    _compare(nvq, oracle, g, d, np.concatenate([big, far]), mode, tag=f"{name} binned")
    d.close()


def test_fog_world_space(oracle, tiled_fogs):
    """World-space points (FP64 map, as the reference op) on a fog, both paths."""
    nvq = _nvq()
    g = tiled_fogs["fog_sphere"]
    d = nvq.DeviceGrid(g)
    for n in (65537, (1 << 20) + 17):
        xyz = points_near(oracle, g, n, seed=n % 97)
        _compare(nvq, oracle, g, d, xyz, 1, world=True, tag="fog_sphere world")
    d.close()
