"""Device-side statistics (tools/cuda/GridStats.cuh updateGridStats<float>): the oracle's
restatement of the device merge order against the reference host tools::updateGridStats
(CPU only).  Extrema, bboxes, flags and the world bbox do not depend on the merge order and must
be bit-identical; an average may differ in its last bit, and only where the two orders differ."""
import numpy as np
import pytest

from conftest import stats_grids
from pyoracle import have_reference

pytestmark = pytest.mark.skipif(not have_reference(), reason="reference sources not built")


def _u(b, off, dt, n=1):
    return np.frombuffer(b[off:off + np.dtype(dt).itemsize * n].tobytes(), dt)


def moment_fields(buf):
    """Byte offsets of every average and standard deviation field (leaves, lower and upper nodes,
    root) with the offset of the same node's minimum."""
    offs, cnt = _u(buf, 672, np.int64, 4), _u(buf, 704, np.uint32, 3)
    out = {}
    root = 672 + offs[3]
    for f in (40, 44):
        out[root + f] = root + 32
    for base, size, st, c in ((672 + offs[2], 270400, 8224, cnt[2]), (672 + offs[1], 33856, 1056, cnt[1]),
                              (672 + offs[0], 2144, 80, cnt[0])):
        for i in range(int(c)):
            for f in (8, 12):
                out[base + size * i + st + f] = base + size * i + st
    return out


@pytest.fixture(scope="module")
def grids(ref, oracle):
    return stats_grids(ref, oracle)


@pytest.mark.parametrize("name", ["sphere", "fog_sphere_tiles", "torus", "fbm", "voxels_to_grid"])
def test_oracle_vs_host_update_grid_stats(grids, oracle, ref, name):
    g = grids[name]
    o = oracle.update_grid_stats(g)
    r = ref.update_grid_stats(g)
    fields = moment_fields(g)
    diff_fields = {int(x) - (int(x) - 672) % 4 for x in np.nonzero(o != r)[0]}
    assert diff_fields <= set(int(k) for k in fields), "an extremum, bbox, flag or world bbox differs"
    for f in sorted(diff_fields):
        fo, fr = _u(o, f, np.float32)[0], _u(r, f, np.float32)[0]
        lo, hi = _u(o, fields[f], np.float32, 2)
        scale = max(abs(float(lo)), abs(float(hi)), 1e-30)
        ulps = abs(int(np.float32(fo).view(np.int32)) - int(np.float32(fr).view(np.int32)))
        assert ulps <= 1 or abs(float(fo) - float(fr)) <= 1e-6 * scale, (f, fo, fr)


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("name", ["sphere", "fog_sphere_tiles", "torus", "fbm", "voxels_to_grid"])
def test_bbox_and_minmax_modes_match_host(grids, oracle, ref, name, mode):
    """StatsMode::BBox and StatsMode::MinMax involve no averages, so the device merge order cannot
    matter: the oracle equals the host tools::updateGridStats byte for byte, starting from grids
    whose statistics were first scrambled by a pass in the other mode."""
    g = oracle.update_grid_stats(grids[name], 3 if mode == 1 else 1)
    o = oracle.update_grid_stats(g, mode)
    r = ref.update_grid_stats(g, mode)
    bad = np.nonzero(o != r)[0]
    assert bad.size == 0, f"{bad.size} bytes differ, first at {bad[:8]}"
    flags = int(_u(o, 20, np.uint32)[0])
    assert flags & 0x1c == (0 if mode == 1 else 4)  # HasMinMax only in MinMax mode


def test_modes_write_only_their_fields(grids, oracle):
    """BBox leaves every statistic as it was; MinMax rewrites the extrema and nothing else;
    Disable (0) and unknown modes are errors."""
    g = grids["fbm"]
    full = oracle.update_grid_stats(g, 3)
    fields = moment_fields(g)
    avg_std = np.zeros(len(g), bool)
    for f in fields:
        avg_std[f:f + 4] = True
    mm = oracle.update_grid_stats(full, 2)
    assert bool((mm[avg_std] == full[avg_std]).all())
    # extrema: equal to the full pass's extrema
    ext = np.zeros(len(g), bool)
    for f, m in fields.items():
        ext[m:m + 8] = True
    assert bool((mm[ext] == full[ext]).all())
    zero = full.copy()
    zero[ext | avg_std] = 0
    bb = oracle.update_grid_stats(zero, 1)
    assert bool((bb[ext | avg_std] == 0).all())
    for bad_mode in (0, 4):
        with pytest.raises(ValueError):
            oracle.update_grid_stats(g, bad_mode)


def test_stats_do_not_depend_on_previous_stats(grids, oracle):
    """Zeroed statistics and statistics flags are recomputed to the same bytes."""
    g = grids["fog_sphere_tiles"]
    d = g.copy()
    offs, cnt = _u(g, 672, np.int64, 4), _u(g, 704, np.uint32, 3)
    root = 672 + offs[3]
    d[root + 32:root + 48] = 0
    for base, size, lo, c in ((672 + offs[2], 270400, 8224, cnt[2]), (672 + offs[1], 33856, 1056, cnt[1]),
                              (672 + offs[0], 2144, 80, cnt[0])):
        for i in range(int(c)):
            d[base + size * i + lo:base + size * i + lo + 16] = 0
    d[20] &= ~0x1e & 0xff  # HasBBox, HasMinMax, HasAverage, HasStdDeviation (IsBreadthFirst stays)
    np.testing.assert_array_equal(oracle.update_grid_stats(d), oracle.update_grid_stats(g))


def test_leaf_average_matches_exact_mean(grids, oracle):
    """Leaf averages and standard deviations against long-double sums over the active values."""
    g = oracle.update_grid_stats(grids["torus"])
    offs, cnt = _u(g, 672, np.int64, 4), _u(g, 704, np.uint32, 3)
    leaves = np.frombuffer(g[672 + offs[0]:672 + offs[0] + 2144 * int(cnt[0])].tobytes(), np.uint8).reshape(-1, 2144)
    mask = np.unpackbits(leaves[:, 16:80], axis=1, bitorder="little").astype(bool)
    vals = leaves[:, 96:].copy().view(np.float32)
    st = leaves[:, 80:96].copy().view(np.float32)
    for i in range(0, len(leaves), max(1, len(leaves) // 200)):
        v = vals[i][mask[i]].astype(np.longdouble)
        mean = v.mean()
        std = np.sqrt(((v - mean) ** 2).mean())
        assert st[i, 0] == v.min() and st[i, 1] == v.max()
        assert abs(float(st[i, 2]) - float(mean)) <= 1e-6 * max(1.0, abs(float(mean)))
        assert abs(float(st[i, 3]) - float(std)) <= 1e-5 * max(1.0, float(std))
