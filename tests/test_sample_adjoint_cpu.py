"""Pins of the point-query adjoint oracle (oracle/oracle.c: or_sample_bwd_fixed, or_sample_bwd_f64).

The reference sampler has no backward, so nothing in the reference can be compared with: the
adjoint is pinned by calculus and invariants instead.
- central differences of an FP64 forward (or_sample_loss_f64) with one value perturbed, every
  gradient mode and the nearest and trilinear outputs, Float and index grids;
- Euler's identity for a loss linear in the values: sum_k dL/dv_k * v_k = L when every value a
  query reads is a differentiated one (index grids always; Float grids for queries whose whole
  stencil lies in leaves);
- GradStencil and nearest known answers (one query, one voxel pair);
- the FP32 / fixed-point restatement against the FP64 one, and its independence of query order
  and batch splits (integer sums).
"""
import numpy as np
import pytest

from conftest import points_near, rng

MODES = [1, 2, 3]


def _leaf_values(buf):
    """Float grid: [leafCount*512] leaf values in buffer order (leaf size 2144 B, values at +96)."""
    leaf0 = 672 + int(np.frombuffer(buf[672:680].tobytes(), np.int64)[0])
    count = int(np.frombuffer(buf[672 + 32:672 + 36].tobytes(), np.uint32)[0])
    out = np.empty(count * 512, np.float64)
    for i in range(count):
        o = leaf0 + 2144 * i + 96
        out[512 * i:512 * (i + 1)] = np.frombuffer(buf[o:o + 2048].tobytes(), np.float32)
    return out


def _channel(buf):
    """Index grid: its own float channel (blind data #0)."""
    meta = int(np.frombuffer(buf[640:648].tobytes(), np.int64)[0])
    off = int(np.frombuffer(buf[meta:meta + 8].tobytes(), np.int64)[0])
    count = int(np.frombuffer(buf[656:664].tobytes(), np.uint64)[0])
    return np.frombuffer(buf[meta + off:meta + off + 4 * count].tobytes(), np.float32).astype(np.float64)


def _upstream(n, seed, mode):
    r = rng(seed)
    gn = r.normal(0, 1, n)
    gv = r.normal(0, 1, n)
    gg = r.normal(0, 1, (n, 3)) if mode else None
    return gn, gv, gg


@pytest.fixture(scope="module")
def small_fog(ref):
    return ref.fog_volume_sphere(14.0, (1.0, -2.0, 3.0), 1.0, 3.0)


@pytest.fixture(scope="module")
def small_index(ref, small_fog):
    return ref.index_grid(small_fog, channels=1)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("kind", ["float", "index"])
def test_fp64_adjoint_matches_central_differences(oracle, ref, small_fog, small_index, mode, kind):
    grid = small_fog if kind == "float" else small_index
    src = small_fog
    xyz = points_near(oracle, src, 300, seed=10 + mode)
    gn, gv, gg = _upstream(len(xyz), 20 + mode, mode)
    g = oracle.sample_bwd_f64(grid, xyz, gn, gv, gg, mode)
    nz = np.flatnonzero(g)
    assert len(nz) > 200
    r = rng(30 + mode)
    keys = np.concatenate([nz[np.argsort(-np.abs(g[nz]))[:25]], r.choice(nz, 25, replace=False),
                           r.choice(np.flatnonzero(g == 0), 5, replace=False)])
    h = 1e-3
    scale = np.abs(g).max()
    for k in keys:
        lp = oracle.sample_loss_f64(grid, xyz, gn, gv, gg, mode, int(k), h)
        lm = oracle.sample_loss_f64(grid, xyz, gn, gv, gg, mode, int(k), -h)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-9 * max(scale, 1.0), (kind, mode, int(k), fd, g[k])


@pytest.mark.parametrize("mode", MODES)
def test_euler_identity_index_grid(oracle, small_fog, small_index, mode):
    """Every value an index-grid query reads is a channel entry: sum_k g_k c_k = L exactly."""
    xyz = points_near(oracle, small_fog, 400, seed=40 + mode, bbox_pad=6)
    gn, gv, gg = _upstream(len(xyz), 50 + mode, mode)
    g = oracle.sample_bwd_f64(small_index, xyz, gn, gv, gg, mode)
    c = _channel(small_index)
    L = oracle.sample_loss_f64(small_index, xyz, gn, gv, gg, mode)
    assert abs(g @ c - L) <= 1e-10 * (np.abs(g) @ np.abs(c))


@pytest.mark.parametrize("mode", MODES)
def test_euler_identity_float_grid_inside_leaves(oracle, small_fog, mode):
    """Queries whose 4^3 block lies in leaves (getDim 1) read only leaf values: sum_k g_k v_k = L."""
    xyz = points_near(oracle, small_fog, 3000, seed=60 + mode, sigma=1.0)
    base = np.floor(xyz).astype(np.int32)
    offs = np.stack(np.meshgrid(*[np.arange(-1, 3)] * 3, indexing="ij"), -1).reshape(-1, 3).astype(np.int32)
    ijk = (base[:, None, :] + offs[None]).reshape(-1, 3)
    _, _, dim = oracle.get_value(small_fog, ijk)
    inside = (dim.reshape(len(xyz), 64) == 1).all(1)
    xyz = xyz[inside][:300]
    assert len(xyz) >= 100
    gn, gv, gg = _upstream(len(xyz), 70 + mode, mode)
    g = oracle.sample_bwd_f64(small_fog, xyz, gn, gv, gg, mode)
    v = _leaf_values(small_fog)
    L = oracle.sample_loss_f64(small_fog, xyz, gn, gv, gg, mode)
    assert abs(g @ v - L) <= 1e-10 * (np.abs(g) @ np.abs(v))


def test_known_answers_grad_stencil_and_nearest(oracle, small_fog):
    """GradStencil at RoundDown(p) differentiates v(i+1) - v(i-1) with 0.5/dx; the nearest value
    is the voxel at Round(p); nothing else receives gradient."""
    p = np.array([[1.3, -2.6, 15.2]], np.float32)  # cell (1,-3,15) in a leaf; Round -> (1,-3,15)
    g = oracle.sample_bwd_f64(small_fog, p, None, None, np.array([[2.0, 0.0, 0.0]]), 2)
    ijk = np.array([[2, -3, 15], [0, -3, 15]], np.int32)
    keys = _keys(oracle, small_fog, ijk)
    assert np.count_nonzero(g) == 2
    assert g[keys[0]] == pytest.approx(1.0) and g[keys[1]] == pytest.approx(-1.0)
    g = oracle.sample_bwd_f64(small_fog, p, np.array([3.0]), None, None, 0)
    assert np.count_nonzero(g) == 1 and g[_keys(oracle, small_fog, np.array([[1, -3, 15]], np.int32))[0]] == 3.0
    fx, _ = oracle.sample_bwd_fixed(small_fog, p, None, None, np.array([[2.0, 0.0, 0.0]], np.float32), 2, shift=40)
    assert fx[keys[0]] == 2 ** 40 and fx[keys[1]] == -2 ** 40 and np.count_nonzero(fx) == 2


def _keys(oracle, buf, ijk):
    """Leaf-value keys (leafIndex*512 + offset) of voxels known to lie in leaves, found by
    probing the adjoint of the nearest value at each voxel."""
    out = []
    for c in ijk:
        g = oracle.sample_bwd_f64(buf, c[None].astype(np.float32), np.array([1.0]), None, None, 0)
        nz = np.flatnonzero(g)
        assert len(nz) == 1
        out.append(int(nz[0]))
    return out


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
@pytest.mark.parametrize("kind", ["float", "index"])
def test_fixed_point_restatement_matches_fp64(oracle, small_fog, small_index, mode, kind):
    grid = small_fog if kind == "float" else small_index
    xyz = points_near(oracle, small_fog, 2000, seed=80 + mode)
    gn, gv, gg = _upstream(len(xyz), 90 + mode, mode)
    f32 = lambda a: None if a is None else a.astype(np.float32)
    acc, over = oracle.sample_bwd_fixed(grid, xyz, f32(gn), f32(gv), f32(gg), mode, shift=40)
    assert not over
    g = acc.astype(np.float64) * 2.0 ** -40
    ref64 = oracle.sample_bwd_f64(grid, xyz, f32(gn).astype(np.float64), f32(gv).astype(np.float64),
                                  None if gg is None else f32(gg).astype(np.float64), mode)
    scale = np.abs(ref64).max()
    assert np.abs(g - ref64).max() <= 1e-5 * scale
    assert np.linalg.norm(g - ref64) <= 1e-5 * np.linalg.norm(ref64)


def test_fixed_point_is_order_and_split_independent(oracle, small_fog):
    xyz = points_near(oracle, small_fog, 3000, seed=99)
    gn, gv, gg = (a.astype(np.float32) for a in _upstream(len(xyz), 98, 1))
    a, _ = oracle.sample_bwd_fixed(small_fog, xyz, gn, gv, gg, 1)
    perm = rng(5).permutation(len(xyz))
    b, _ = oracle.sample_bwd_fixed(small_fog, xyz[perm], gn[perm], gv[perm], gg[perm], 1)
    assert np.array_equal(a, b)
    c, _ = oracle.sample_bwd_fixed(small_fog, xyz[:1000], gn[:1000], gv[:1000], gg[:1000], 1)
    c, _ = oracle.sample_bwd_fixed(small_fog, xyz[1000:], gn[1000:], gv[1000:], gg[1000:], 1, accum=c)
    assert np.array_equal(a, c)


def test_fixed_point_saturation_flag(oracle, small_fog):
    xyz = np.array([[1.5, -2.5, 15.5]], np.float32)
    acc, over = oracle.sample_bwd_fixed(small_fog, xyz, None, np.array([1e30], np.float32), None, 0, shift=40)
    assert over
    assert acc.max() == 2 ** 62
    acc, over = oracle.sample_bwd_fixed(small_fog, xyz, None, np.array([1.0], np.float32), None, 0, shift=40)
    assert not over


def test_fd6_rounding_cells(oracle, small_fog):
    """p.x + 1 rounding up to the next integer moves the FD6 +x sample one cell further; the
    restatement follows the forward's cells (the +x sample then reads cell + 2 with u = 0)."""
    x = np.float32(1.0) - np.float32(2.0 ** -24)  # 1 + x rounds to 2.0 in float
    p = np.array([[x, -2.25, 3.5]], np.float32)
    assert np.floor(np.float32(p[0, 0] + np.float32(1.0))) == 2.0
    acc, _ = oracle.sample_bwd_fixed(small_fog, p, None, None, np.array([[1.0, 0.0, 0.0]], np.float32), 1)
    # the FP64 derivative (cells floor(p) +- 1, exact fractions) differs from it by one float ulp of u
    g = oracle.sample_bwd_f64(small_fog, p, None, None, np.array([[1.0, 0.0, 0.0]]), 1)
    assert np.count_nonzero(acc) > 0
    assert np.abs(acc.astype(np.float64) * 2.0 ** -40 - g).max() <= 1e-5 * np.abs(g).max()
