"""The differentiable-march configuration's gradient check: the GPU adjoint (FP32 march, int64
fixed-point reduction) on one full 256x256 camera of the 256^3 fBm fog (seed 7, target seed 11),
against

* the FP64 oracle adjoint over every leaf value, and
* central differences (h = 1e-3) of the FP64 oracle march on 1000 seeded voxels the rays touch,

both at 1e-4 relative (max-norm and 2-norm).  The loss is the L2 image loss, linearised at the
current values: loss = sum_r w_r L_r with w = 2 (L - target), so every check differentiates the same
function.  The FP64 oracle adjoint is itself pinned to the central differences at 1e-6.
"""
import numpy as np
import pytest

from conftest import rng

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

H = 1e-3


def _nvq():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import openvdb_amd
    return openvdb_amd


def _leaf_origin_and_offset(buf, keys):
    """Voxel coordinates and byte offsets of leaf values key = leaf * 512 + n (Float grid)."""
    leaf_off = 672 + int(np.frombuffer(buf[672:680].tobytes(), np.int64)[0])
    k, n = keys // 512, keys % 512
    rows = np.frombuffer(buf, np.uint8)[leaf_off + 2144 * k[:, None] + np.arange(12)].copy()
    origin = rows.view(np.int32).reshape(-1, 3) & ~7  # LeafData::mBBoxMin & ~7 is the leaf origin
    ijk = origin + np.stack([n >> 6, (n >> 3) & 7, n & 7], 1)
    return ijk, leaf_off + 2144 * k + 96 + 4 * n


def _rays_touching(E, D, v):
    """Rays that pass through the open box (v-1, v+1)^3: the samples x whose trilinear cell has
    corner v (floor(x) in {v-1, v} on every axis).  The grid's map is the identity."""
    lo, hi = v - 1.0 - 1e-3, v + 1.0 + 1e-3
    with np.errstate(divide="ignore", invalid="ignore"):
        a = (lo - E) / D
        b = (hi - E) / D
    t0 = np.where(np.isnan(np.minimum(a, b)), -np.inf, np.minimum(a, b)).max(axis=1)
    t1 = np.where(np.isnan(np.maximum(a, b)), np.inf, np.maximum(a, b)).min(axis=1)
    return np.nonzero((t1 >= t0) & (t1 > 0))[0]


def central_differences(oracle, grid, eye, dir, w, keys, dt=0.5, sigma=0.05, workers=8):
    """d(sum_r w_r L_r)/d value for each leaf-value key by central differences of the FP64 oracle
    march, re-marching only the rays that reach the value (the others do not depend on it)."""
    from concurrent.futures import ThreadPoolExecutor
    ijk, offs = _leaf_origin_and_offset(grid, keys)
    E, D = eye.astype(np.float64), dir.astype(np.float64)

    def one(chunk):
        buf = grid.copy()
        out = []
        for i in chunk:
            R = _rays_touching(E, D, ijk[i].astype(np.float64))
            off = int(offs[i])
            x0 = np.frombuffer(buf[off:off + 4].tobytes(), np.float32)[0]
            xp, xm = np.float32(x0 + np.float32(H)), np.float32(x0 - np.float32(H))
            buf[off:off + 4] = np.frombuffer(xp.tobytes(), np.uint8)
            Lp, _, _ = oracle.fog_march(buf, eye[R], dir[R], dt, sigma)
            buf[off:off + 4] = np.frombuffer(xm.tobytes(), np.uint8)
            Lm, _, _ = oracle.fog_march(buf, eye[R], dir[R], dt, sigma)
            buf[off:off + 4] = np.frombuffer(x0.tobytes(), np.uint8)
            out.append(float(np.dot(w[R], Lp - Lm)) / (float(xp) - float(xm)))
        return out

    chunks = np.array_split(np.arange(len(keys)), workers)
    with ThreadPoolExecutor(workers) as ex:
        parts = list(ex.map(one, chunks))
    return np.array([x for p in parts for x in p])


def c4_camera(nvq):
    """The differentiable-march scene: 256^3 fBm fogs (seeds 7 and 11) and the first of its 64
    cameras on the radius-600 shell (256^2 rays)."""
    from bench import shell_cameras
    box = [0, 0, 0, 255, 255, 255]
    gp = nvq.tools.create_fog_fbm(box, seed=7)
    gt = nvq.tools.create_fog_fbm(box, seed=11)
    e, d = shell_cameras(nvq, 1, 256, (128.0, 128.0, 128.0), 600.0, seed=2024)
    return gp, gt, e, d


def _errors(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max()), float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_c4_adjoint_against_oracle_and_central_differences(oracle):
    nvq = _nvq()
    gp, gt, e, d = c4_camera(nvq)
    dp, dtg = nvq.DeviceGrid(gp), nvq.DeviceGrid(gt)
    E, D = torch.from_numpy(e).cuda(), torch.from_numpy(d).cuda()
    L, T = nvq.ops.fog_march(dp, E, D, 0.5, 0.05, image_width=256)
    target, _ = nvq.ops.fog_march(dtg, E, D, 0.5, 0.05, image_width=256)
    gL = (2.0 * (L - target)).contiguous()
    acc = nvq.ops.fog_march_backward(dp, E, D, L, gL, dt=0.5, sigma_t=0.05, image_width=256)
    g_gpu = nvq.ops.fixed_to_float(acc).cpu().numpy().astype(np.float64)
    w = gL.cpu().numpy().astype(np.float64)
    g_or = oracle.fog_march_adjoint(gp, e, d, 0.5, 0.05, w)
    mx, nr = _errors(g_gpu, g_or)
    print(f"[c4 camera 0] GPU vs FP64 oracle adjoint over {g_or.size} leaf values "
          f"({int((g_or != 0).sum())} non-zero): max err {mx:.2e} of max|g|, norm rel {nr:.2e}", flush=True)
    assert mx <= 1e-4 and nr <= 1e-4

    keys = np.sort(rng(1000).choice(np.nonzero(g_or)[0], 1000, replace=False))
    fd = central_differences(oracle, gp, e, d, w, keys)
    mo, no = _errors(g_or[keys], fd)
    mg, ng = _errors(g_gpu[keys], fd)
    print(f"[c4 camera 0] central differences (h={H}) on 1000 voxels: FP64 oracle max {mo:.2e} norm {no:.2e}; "
          f"GPU max {mg:.2e} norm {ng:.2e}", flush=True)
    assert mo <= 1e-6 and no <= 1e-6  # pins the oracle adjoint
    assert mg <= 1e-4 and ng <= 1e-4


def test_adjoint_overflow_is_reported_not_faulted():
    """Contributions beyond the fixed-point range saturate and return NVQ_ERR_OVERFLOW."""
    nvq = _nvq()
    from openvdb_amd import NvqError
    g = nvq.tools.create_fog_fbm([0, 0, 0, 31, 31, 31], seed=3)
    d = nvq.DeviceGrid(g)
    eye, dir = nvq.tools.camera_rays(16, 16, (16.0, 16.0, 90.0), (16.0, 16.0, 16.0), (0.0, 1.0, 0.0), 40.0)
    E, D = torch.from_numpy(eye).cuda(), torch.from_numpy(dir).cuda()
    L, T = nvq.ops.fog_march(d, E, D, 0.5, 0.05)
    assert float(L.max()) > 0
    huge = torch.full_like(L, 1e30)
    with pytest.raises(NvqError) as ei:
        nvq.ops.fog_march_backward(d, E, D, L, huge)
    assert ei.value.status == -9
    nan = torch.full_like(L, float("nan"))
    with pytest.raises(NvqError) as ei:
        nvq.ops.fog_march_backward(d, E, D, L, nan)
    assert ei.value.status == -9
    ok = nvq.ops.fog_march_backward(d, E, D, L, torch.ones_like(L))  # in range: no error
    assert int((ok != 0).sum()) > 0
