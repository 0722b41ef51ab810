import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU; run on the GPU box")


@pytest.fixture(scope="session")
def oracle():
    from pyoracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def ref():
    from pyoracle import Reference, have_reference
    if not have_reference():
        pytest.skip("oracle/_ref/libnanoref.so not built (needs the reference sources)")
    return Reference()


@pytest.fixture(scope="session")
def sphere(ref):
    """createLevelSetSphere<float>(100, 0, 1, 3): the level-set ray-hit grid."""
    return ref.level_set_sphere(100.0, (0, 0, 0), 1.0, 3.0)


@pytest.fixture(scope="session")
def torus(ref):
    """createLevelSetTorus<float>(300, 80, 0, 1, 3): the point-query grid."""
    return ref.level_set_torus(300.0, 80.0, (0, 0, 0), 1.0, 3.0)


@pytest.fixture(scope="session")
def fog_sphere(ref):
    """createFogVolumeSphere<float>(60, (3,-2,5), 1, 3): fog with active interior tiles."""
    return ref.fog_volume_sphere(60.0, (3.0, -2.0, 5.0), 1.0, 3.0)


def rng(seed):
    return np.random.default_rng(seed)


def points_near(oracle, grid, n, seed, bbox_pad=4, sigma=2.0):
    """Half uniform in the (padded) index bbox, half Gaussian around active voxels."""
    from pyoracle import Oracle  # noqa: F401
    r = rng(seed)
    lo, hi = grid_bbox(grid)
    lo = lo - bbox_pad
    hi = hi + bbox_pad
    a = r.random((n // 2, 3)) * (hi - lo + 1) + lo
    cand = r.integers(lo, hi + 1, size=(n * 4, 3)).astype(np.int32)
    _, act, _ = oracle.get_value(grid, cand)
    act_pts = cand[act.astype(bool)]
    if len(act_pts) == 0:
        act_pts = cand
    pick = act_pts[r.integers(0, len(act_pts), size=n - n // 2)]
    b = pick + r.normal(0.0, sigma, size=pick.shape)
    return np.concatenate([a, b]).astype(np.float32)


def grid_bbox(buf):
    tree = 672
    off_root = int(np.frombuffer(buf[tree + 24:tree + 32].tobytes(), np.int64)[0])
    root = tree + off_root
    bb = np.frombuffer(buf[root:root + 24].tobytes(), np.int32)
    return bb[:3].astype(np.int64), bb[3:].astype(np.int64)


def voxel_cases():
    """Voxel lists for the device builder (voxelsToGrid): the reference's own test inputs
    (unittest/TestNanoVDB.cu: one voxel (1,2,3); three voxels; the 125..145 cube; random voxels in
    [-512, 512)) plus duplicates, negative coordinates and inputs spanning several upper tiles."""
    r = rng(11)
    cube = np.stack(np.meshgrid(*[np.arange(125, 146)] * 3, indexing="ij"), -1).reshape(-1, 3)
    rand = r.integers(-512, 512, (20000, 3))
    dups = np.concatenate([rand[:3000], rand[:3000][::-1], rand[:500]])
    wide = r.integers(-9000, 9000, (6000, 3))
    shell = r.normal(size=(30000, 3))
    shell = np.rint(shell / np.linalg.norm(shell, axis=1, keepdims=True) * r.uniform(198, 202, (30000, 1)) +
                    np.array([-37.0, 11.0, 5.0]))
    return {
        "one": np.array([[1, 2, 3]]),
        "three": np.array([[1, 2, 3], [1, 2, 4], [8, 2, 3]]),
        "cube": cube,
        "random": rand,
        "duplicates": dups,
        "wide": wide,
        "shell": shell,
    }


def stats_grids(ref, oracle):
    """Float grids for the statistics tests: reference-built level set and fog with active tiles,
    our builder's torus and fBm fog, and a device-builder grid (statistics zero)."""
    import openvdb_amd as nvq
    v2g, _ = oracle.voxels_to_grid(voxel_cases()["shell"], 1, 0.5)
    return {
        "sphere": ref.level_set_sphere(40.0, (1.5, -2.0, 3.0), 0.5, 3.0),
        "fog_sphere_tiles": ref.fog_volume_sphere(30.0, (0, 0, 0), 1.0, 3.0),
        "torus": nvq.tools.create_level_set_torus(60.0, 15.0, (0.0, 0.0, 0.0), 1.0, 3.0),
        "fbm": nvq.tools.create_fog_fbm([0, 0, 0, 63, 63, 63], seed=3),
        "voxels_to_grid": v2g,
    }
