"""The oracle restatement, pinned against the reference's own known answers and against the
reference host code compiled from its sources (oracle/_ref)."""
import json
import os

import numpy as np
import pytest

from conftest import ROOT, grid_bbox, points_near, rng

GOLDEN = os.path.join(ROOT, "tests", "golden")


def _eqbits(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint8), np.ascontiguousarray(b).view(np.uint8))


# ----------------------------------------------------------- known answers (unit tests of the reference)
def test_ray_clip_known_answers(oracle):
    # unittest/TestNanoVDB.cc:1426-1460 (RayBasic): bbox (0,0,0)-(0,0,0) -> t0=1, t1=2
    assert oracle.ray_clip([-1, 0.5, 0.5], [1, 0, 0], 0.0, 10000.0, [0, 0, 0, 0, 0, 0]) == (True, 1.0, 2.0)
    # :1462-1480 (Ray): dir(-1,2,3) eye(2,1,1) bbox (0,2,2)-(2,4,6) -> t0=0.5, t1=2.0
    assert oracle.ray_clip([2, 1, 1], [-1, 2, 3], 0.1, 12589.0, [0, 2, 2, 2, 4, 6]) == (True, 0.5, 2.0)
    # :1492-1500 no intersection keeps the times
    hit, t0, t1 = oracle.ray_clip([2, 1, 1], [-1, 2, 3], 0.1, 12589.0, [4, 2, 2, 6, 4, 6])
    assert not hit
    # :1402-1424 empty bbox
    assert not oracle.ray_clip([-1, 0.5, 0.5], [1, 0, 0], 0.0, 10000.0, [1, 1, 1, 0, 0, 0])[0]


def test_hdda_known_answers(oracle):
    # unittest/TestNanoVDB.cc:1501-1603
    fmax = np.finfo(np.float32).max
    times, nexts, _, _ = oracle.hdda_trace([-1, 0, 0], [1, 0, 0], 1e-5, fmax, 4096, 1)
    assert times[0] == np.float32(1e-5) and nexts[0] == 1.0
    assert times[1] == 1.0 and nexts[1] == 4097.0
    # +-0 direction handling: all four traces agree
    tr = [oracle.hdda_trace([2, 0, 0], d, 1e-5, fmax, 8, 1) for d in
          ([1, 0, 0], [1, -0.0, -0.0], [1, -1e-9, -1e-9])]
    for t in tr[1:]:
        assert t[0][1] == tr[0][0][1] and t[1][1] == tr[0][1][1]
    for s in (-1, 1):
        for a in range(3):
            d = [0, 0, 0]
            d[a] = s
            times, _, _, ok = oracle.hdda_trace([0, 0, 0], d, 1e-5, fmax, 1, 10)
            assert ok[1:].all() and list(times[1:]) == list(range(1, 11))
            times, _, _, ok = oracle.hdda_trace([0, 0, 0], d, 1e-5, fmax, 8, 10)
            assert list(times[1:]) == [8 * i for i in range(1, 11)]
            d2 = [2.0 * x for x in d]
            times, nexts, _, _ = oracle.hdda_trace([0, 0, 0], d2, 1e-5, fmax, 8, 10)
            assert list(times[1:]) == [4 * i for i in range(1, 11)]
            assert list(times[2:]) == list(nexts[1:-1])


def test_sphere_values_closed_form(oracle, ref):
    """createLevelSetSphere: every voxel of the narrow band equals the closed form of initSphere
    (tools/CreatePrimitives.h:631-657; unittest/TestNanoVDB.cc:4535-4594), active <=> |d| < halfWidth."""
    grid = ref.level_set_sphere(20.0, (0, 0, 0), 1.0, 3.0)
    r = np.arange(-25, 26)
    ijk = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3).astype(np.int32)
    val, act, _ = oracle.get_value(grid, ijk)
    f = ijk.astype(np.float32)
    x2 = (f[:, 0] - np.float32(0)) ** 2
    d = np.sqrt((f[:, 1] ** 2 + x2) + f[:, 2] ** 2, dtype=np.float32) - np.float32(20.0)
    band = np.abs(d) < 3.0
    assert np.array_equal(act.astype(bool), band)
    assert _eqbits(val[band], (np.float32(1.0) * d[band]).astype(np.float32))
    assert np.all(val[~band & (d < 0)] == -3.0) and np.all(val[~band & (d > 0)] == 3.0)


def test_sampler_closed_form_linear_field(oracle, ref):
    """unittest/TestNanoVDB.cc:5458-5514: trilinear of a linear field within 1e-5, gradient 2e-5."""
    dx = 0.5
    f = lambda i, j, k: np.float32(0.34) + np.float32(1.6 * dx) * i + np.float32(6.7 * dx) * j - np.float32(3.5 * dx) * k
    grid = ref.grid_from_function(lambda i, j, k: float(f(i, j, k)), [0, 0, 0, 40, 40, 40], background=1.0,
                                  grid_class=0)
    xyz = np.array([[13.4, 24.67, 5.23], [3.1, 7.9, 30.2], [20.5, 20.5, 20.5]], np.float32)
    nn, val, grad = oracle.sample(grid, xyz, grad_mode=3)
    exact = 0.34 + 1.6 * dx * xyz[:, 0] + 6.7 * dx * xyz[:, 1] - 3.5 * dx * xyz[:, 2]
    np.testing.assert_allclose(val, exact, atol=1e-5 * 4)
    np.testing.assert_allclose(grad / dx, np.tile([1.6, 6.7, -3.5], (3, 1)), atol=2e-5)
    ni = np.rint(xyz).astype(np.int32)
    assert _eqbits(nn, oracle.get_value(grid, ni)[0])


def test_sample_from_voxels_golden(oracle, ref):
    """python/test/TestNanoVDB.py:1330-1390 (reference CUDA op), world-space points, exact."""
    grid = ref.level_set_sphere(100.0, (0, 0, 0), 0.5, 5.0)
    pts = np.array([[10, 0, 0], [100, 0, 0], [0, 0, 99], [0, 101, 0], [110, 0, 0]], np.float32)
    _, val, grad = oracle.sample(grid, pts, world=True, grad_mode=1)
    assert val.tolist() == [-2.5, 0.0, -1.0, 1.0, 2.5]
    assert grad.tolist() == [[0, 0, 0], [1, 0, 0], [0, 0, 1], [0, 1, 0], [0, 0, 0]]


def test_grad_stencil_sphere_known_answers(oracle, ref):
    """unittest/TestNanoVDB.cc:6077-6123: sphere R=10, dx=0.5 -> |grad|^2 ~ 1, grad (1,0,0), laplacian ~ 2/R."""
    grid = ref.level_set_sphere(10.0, (0, 0, 0), 0.5, 3.0)
    g, ns, lap, _ = oracle.grad_stencil(grid, np.array([[20, 0, 0]], np.int32))
    np.testing.assert_allclose(g[0], [1, 0, 0], atol=1e-6)
    np.testing.assert_allclose(ns[0], 1.0, atol=2e-3)
    np.testing.assert_allclose(lap[0], 2.0 / 10.0, atol=1e-3 * 2)


def test_golden_fixtures(oracle, ref):
    """tests/golden/*.json: vectors written by tests/golden/make_golden.py from the compiled reference."""
    path = os.path.join(GOLDEN, "c1_sphere_zero_crossing_64x64.json")
    with open(path) as fh:
        fx = json.load(fh)
    grid = ref.level_set_sphere(*fx["grid"]["args"])
    eye = np.array(fx["eye"], np.float32)
    dir = np.array(fx["dir"], np.float32)
    hit, ijk, t, v, _ = oracle.zero_crossing(grid, eye, dir)
    assert hit.tolist() == fx["hit"]
    assert ijk.tolist() == fx["ijk"]
    assert _eqbits(t, np.array(fx["t_bits"], np.uint32).view(np.float32))
    assert _eqbits(v, np.array(fx["v_bits"], np.uint32).view(np.float32))


# ----------------------------------------------------------- oracle vs compiled reference (bit-exact)
@pytest.mark.parametrize("kind", [0, 1, 2, 3, 4])
def test_oracle_matches_reference_lookups(oracle, ref, kind):
    grid = ref.level_set_sphere(30.0, (0.5, -1.0, 2.0), 0.5, 3.0, kind=kind, tol=1e-3)
    ijk = rng(kind).integers(-80, 80, size=(100000, 3)).astype(np.int32)
    for a, b in zip(oracle.get_value(grid, ijk), ref.get_value(grid, ijk)):
        assert _eqbits(a, b)
    xyz = points_near(oracle, grid, 50000, seed=kind)
    for gm in (1, 2, 3):
        for w in (False, True):
            for a, b in zip(oracle.sample(grid, xyz, world=w, grad_mode=gm), ref.sample(grid, xyz, world=w, grad_mode=gm)):
                assert _eqbits(a, b), (gm, w)


def test_oracle_matches_reference_zero_crossing(oracle, ref, sphere):
    eye, dir = ref.ray_gen(512, 512, 400.0)
    a = oracle.zero_crossing(sphere, eye, dir)
    b = ref.zero_crossing(sphere, eye, dir, nthreads=8)
    for x, y in zip(a, b):
        assert _eqbits(x, y)
    assert int(a[0].sum()) == 81104
    c = 256 * 512 + 256
    assert a[1][c].tolist() == [0, 0, 99]


def test_oracle_matches_reference_grad_stencil(oracle, ref, sphere):
    ijk = rng(9).integers(-105, 105, size=(50000, 3)).astype(np.int32)
    for a, b in zip(oracle.grad_stencil(sphere, ijk), ref.grad_stencil(sphere, ijk)):
        assert _eqbits(a, b)


def test_oracle_matches_reference_fog_march(oracle, ref, fog_sphere):
    r = rng(17)
    n = 2000
    eye = np.tile(np.array([[3.0, -2.0, 300.0]], np.float32), (n, 1))
    target = np.concatenate([r.uniform(-70, 70, (n, 2)), np.zeros((n, 1))], 1) + np.array([3.0, -2.0, 5.0])
    dd = target - eye
    dir = (dd / np.linalg.norm(dd, axis=1, keepdims=True)).astype(np.float32)
    oL, oT, osteps = oracle.fog_march(fog_sphere, eye, dir, 0.5, 0.05)
    L, T, L64, T64, steps = ref.fog_march(fog_sphere, eye, dir, 0.5, 0.05)
    assert np.array_equal(osteps, steps)
    assert _eqbits(oL, L64) and _eqbits(oT, T64)
    assert steps.max() > 100
    np.testing.assert_allclose(L, L64, rtol=2e-5, atol=1e-6)


def test_fog_march_constant_density_closed_form(oracle, ref):
    """Constant density rho on [0,63]^3 (pruned into active tiles): along an axis-aligned ray the
    trilinear density is rho inside and rho*(64-z) on the last half cell, so
    T = exp(-sigma*dt*sum rho_k) and L = sum T_k (1-a_k) rho_k in closed form."""
    rho = 0.625
    grid = ref.grid_from_function(lambda i, j, k: rho, [0, 0, 0, 63, 63, 63], background=0.0, grid_class=2)
    eye = np.array([[10.25, 20.5, -50.0]], np.float32)
    dir = np.array([[0.0, 0.0, 1.0]], np.float32)
    L, T, steps = oracle.fog_march(grid, eye, dir, 0.5, 0.05)
    t0 = np.float32(50.0) + np.float32(1e-6)
    t1 = np.float32(114.0) - np.float32(1e-6)
    ts = []
    kf = np.ceil(np.float32(t0 / np.float32(0.5)))
    while np.float32(kf * 0.5) < t1:
        ts.append(float(np.float32(kf * 0.5)))
        kf += 1
    z = np.array(ts) - 50.0
    dens = rho * np.clip(64.0 - z, 0.0, 1.0)
    a = np.exp(-float(np.float32(0.05)) * 0.5 * dens)
    Tk = np.concatenate([[1.0], np.cumprod(a)[:-1]])
    assert steps[0] == len(ts) == 128
    np.testing.assert_allclose(T[0], np.prod(a), rtol=1e-14)
    np.testing.assert_allclose(L[0], np.sum(Tk * (1 - a) * dens), rtol=1e-14)


def _leaf_value_offsets(buf, oracle):
    """Byte offset of every leaf value (leaf order, n = x<<6|y<<3|z) of a Float grid."""
    leaf0 = 672 + int(np.frombuffer(buf[672:680].tobytes(), np.int64)[0])
    cnt = oracle.leaf_count(buf)
    return leaf0 + 96 + (np.arange(cnt)[:, None] * 2144 + np.arange(512)[None, :] * 4)


def test_fog_adjoint_matches_central_differences(oracle, ref):
    """The FP64 oracle adjoint (dLoss/dvalue) against central differences of the FP64 oracle
    march, h = 1e-3, on voxels that the rays actually sample.  Pins the adjoint by calculus:
    no reference implementation exists."""
    grid = ref.fog_volume_sphere(9.0, (0.5, -0.25, 0.75), 1.0, 3.0)
    r = rng(5)
    n = 64
    eye = np.tile(np.array([[0.0, 0.0, 40.0]], np.float32), (n, 1)) + r.normal(0, 1.5, (n, 3)).astype(np.float32)
    tgt = r.uniform(-6, 6, (n, 3))
    d = tgt - eye
    dir = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    wL = r.normal(0, 1, n)
    wT = r.normal(0, 1, n)
    g = oracle.fog_march_adjoint(grid, eye, dir, 0.5, 0.05, wL, wT)
    offs = _leaf_value_offsets(grid, oracle).reshape(-1)
    cand = np.nonzero(np.abs(g) > 1e-3 * np.abs(g).max())[0]
    pick = r.choice(cand, size=min(40, len(cand)), replace=False)
    h = 1e-3

    def loss(buf):
        L, T, _ = oracle.fog_march(buf, eye, dir, 0.5, 0.05)
        return float(np.dot(wL, L) + np.dot(wT, T))

    for k in pick:
        b = grid.copy()
        v = np.frombuffer(grid[offs[k]:offs[k] + 4].tobytes(), np.float32)[0]
        b[offs[k]:offs[k] + 4] = np.frombuffer(np.float32(v + h).tobytes(), np.uint8)
        lp = loss(b)
        b[offs[k]:offs[k] + 4] = np.frombuffer(np.float32(v - h).tobytes(), np.uint8)
        lm = loss(b)
        dh = float(np.float32(v + h)) - float(np.float32(v - h))
        fd = (lp - lm) / dh
        assert abs(fd - g[k]) <= 2e-5 * max(1.0, abs(g[k])) + 1e-6 * np.abs(g).max(), (k, fd, g[k])
