import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


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
    [-512, 512)) plus duplicates, negative coordinates, inputs spanning several upper tiles, and a
    large input inside one upper tile (the builder's 32-bit sort keys)."""
    r = rng(11)
    cube = np.stack(np.meshgrid(*[np.arange(125, 146)] * 3, indexing="ij"), -1).reshape(-1, 3)
    rand = r.integers(-512, 512, (20000, 3))
    dups = np.concatenate([rand[:3000], rand[:3000][::-1], rand[:500]])
    wide = r.integers(-9000, 9000, (6000, 3))
    shell = r.normal(size=(30000, 3))
    shell = np.rint(shell / np.linalg.norm(shell, axis=1, keepdims=True) * r.uniform(198, 202, (30000, 1)) +
                    np.array([-37.0, 11.0, 5.0]))
    single = r.integers(0, 1500, (200000, 3))
    single = np.concatenate([single, single[:5000]])
    return {
        "single_tile": single,
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


def _u(b, off, dt, n=1):
    return np.frombuffer(b[off:off + np.dtype(dt).itemsize * n].tobytes(), dt)


def damage_inactive(g, value=None):
    """Copy of a float grid whose inactive leaf values and non-child node entries are all set to
    `value` (default: +background), as a level set looks before its signed flood fill."""
    d = g.copy()
    offs, cnt = _u(g, 672, np.int64, 4), _u(g, 704, np.uint32, 3)
    v = np.float32(_u(g, 672 + offs[3] + 28, np.float32)[0] if value is None else value)
    lf = 672 + offs[0]
    leaves = d[lf:lf + 2144 * int(cnt[0])].reshape(-1, 2144)
    mask = np.unpackbits(leaves[:, 16:80], axis=1, bitorder="little").astype(bool)
    vals = leaves[:, 96:].copy().view(np.float32)
    vals[~mask] = v
    leaves[:, 96:] = vals.view(np.uint8)
    bits = np.frombuffer(v.tobytes(), np.uint32)[0]
    for lvl, size, cm, tab, n in ((1, 33856, 544, 1088, 4096), (2, 270400, 4128, 8256, 32768)):
        base = 672 + offs[lvl]
        for k in range(int(cnt[lvl])):
            node = d[base + size * k:base + size * (k + 1)]
            child = np.unpackbits(node[cm:cm + n // 8], bitorder="little").astype(bool)
            table = node[tab:tab + 8 * n].copy().view(np.uint64)
            table[~child] = np.uint64(bits)
            node[tab:tab + 8 * n] = table.view(np.uint8)
    return d


def root_gap_grid(oracle, missing=False, background=3.0):
    """A float level set (all active values -1, background 3) with two root children on the
    z-scanline x = y = 0 (tiles z = 0 and z = 12288) and the inactive root tiles between them
    (z = 4096, 8192; only z = 4096 when `missing`) holding +background: the case
    SignedFloodFill's root pass repairs (tools/cuda/SignedFloodFill.cuh:91-145)."""
    cube = np.stack(np.meshgrid(np.arange(0, 20), np.arange(0, 20), np.arange(0, 20), indexing="ij"), -1).reshape(-1, 3)
    ijk = np.concatenate([cube + [0, 0, 4070], cube + [0, 0, 12288]])
    g, _ = oracle.voxels_to_grid(ijk, 1, 1.0)
    offs, cnt = _u(g, 672, np.int64, 4), _u(g, 704, np.uint32, 3)
    lf = 672 + offs[0]
    leaves = g[lf:lf + 2144 * int(cnt[0])].reshape(-1, 2144)
    mask = np.unpackbits(leaves[:, 16:80], axis=1, bitorder="little").astype(bool)
    vals = leaves[:, 96:].copy().view(np.float32)
    vals[mask] = -1.0
    leaves[:, 96:] = vals.view(np.uint8)
    root = 672 + int(offs[3])
    table = root + 64 + 32 * int(_u(g, root + 24, np.uint32)[0])
    zs = [4096] if missing else [4096, 8192]
    add = np.zeros(32 * len(zs), np.uint8)
    for i, z in enumerate(zs):
        key = np.uint64(z >> 12)  # x = y = 0
        add[32 * i:32 * i + 8] = np.frombuffer(key.tobytes(), np.uint8)
        add[32 * i + 20:32 * i + 24] = np.frombuffer(np.float32(background).tobytes(), np.uint8)
    out = np.concatenate([g[:table], add, g[table:]])
    shift = len(add)
    raw = np.empty(len(out) + 32, np.uint8)
    o = (-raw.ctypes.data) % 32
    buf = raw[o:o + len(out)]
    buf[:] = out

    def put(off, val):
        buf[off:off + val.nbytes] = np.frombuffer(val.tobytes(), np.uint8)
    put(32, np.uint64(len(buf)))
    put(640, np.int64(len(buf)))
    for i in range(3):
        put(672 + 8 * i, np.int64(offs[i] + shift))
    put(root + 24, np.uint32(int(_u(g, root + 24, np.uint32)[0]) + len(zs)))
    put(root + 28, np.float32(background))
    for t in range(int(_u(buf, root + 24, np.uint32)[0])):
        child = int(_u(buf, root + 64 + 32 * t + 8, np.int64)[0])
        if child:
            put(root + 64 + 32 * t + 8, np.int64(child + shift))
    return buf
