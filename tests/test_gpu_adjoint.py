"""GPU adjoint of the fog march against the FP64 oracle adjoint, determinism, autograd."""
import numpy as np
import pytest

from conftest import rng

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _nvq():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import openvdb_amd
    return openvdb_amd


def _scene(nvq, size=48, res=48, seed=7):
    grid = nvq.tools.create_fog_fbm([0, 0, 0, size - 1, size - 1, size - 1], seed=seed, amplitude=1.0)
    c = size / 2.0
    eye, dir = nvq.tools.camera_rays(res, res, (c + 0.3 * size, c - 0.2 * size, c + 2.2 * size), (c, c, c),
                                     (0.0, 1.0, 0.0), 40.0)
    return grid, eye, dir


def test_adjoint_matches_fp64_oracle(oracle):
    nvq = _nvq()
    grid, eye, dir = _scene(nvq)
    d = nvq.DeviceGrid(grid)
    r = rng(3)
    wL = r.normal(0, 1, len(eye)).astype(np.float32)
    wT = r.normal(0, 1, len(eye)).astype(np.float32)
    E, D = torch.from_numpy(eye).cuda(), torch.from_numpy(dir).cuda()
    L, T = nvq.ops.fog_march(d, E, D, 0.5, 0.05, image_width=48)
    oL, oT, steps = oracle.fog_march(grid, eye, dir, 0.5, 0.05)
    assert steps.max() > 50
    np.testing.assert_allclose(L.cpu().numpy(), oL, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(T.cpu().numpy(), oT, rtol=1e-5, atol=1e-6)
    acc = nvq.ops.fog_march_backward(d, E, D, L, torch.from_numpy(wL).cuda(), T, torch.from_numpy(wT).cuda(), 0.5,
                                     0.05, image_width=48)
    g = nvq.ops.fixed_to_float(acc).cpu().numpy().astype(np.float64)
    og = oracle.fog_march_adjoint(grid, eye, dir, 0.5, 0.05, wL.astype(np.float64), wT.astype(np.float64))
    scale = np.abs(og).max()
    err = np.abs(g - og)
    print(f"adjoint: max|g| {scale:.3e}, max err {err.max():.3e} ({err.max()/scale:.2e} of max), "
          f"norm rel {np.linalg.norm(g-og)/np.linalg.norm(og):.2e}", flush=True)
    assert err.max() <= 1e-4 * scale
    assert np.linalg.norm(g - og) <= 1e-4 * np.linalg.norm(og)


def test_adjoint_is_deterministic():
    nvq = _nvq()
    grid, eye, dir = _scene(nvq, 64, 64, seed=11)
    d = nvq.DeviceGrid(grid)
    E, D = torch.from_numpy(eye).cuda(), torch.from_numpy(dir).cuda()
    L, T = nvq.ops.fog_march(d, E, D, 0.5, 0.05)
    gL = torch.randn(len(eye), device="cuda")
    a1 = nvq.ops.fog_march_backward(d, E, D, L, gL)
    a2 = nvq.ops.fog_march_backward(d, E, D, L, gL)
    E2, D2 = E.flip(0).contiguous(), D.flip(0).contiguous()  # same rays, different schedule
    L2, _ = nvq.ops.fog_march(d, E2, D2, 0.5, 0.05)
    a3 = nvq.ops.fog_march_backward(d, E2, D2, L2, gL.flip(0).contiguous())
    assert torch.equal(a1, a2) and torch.equal(a1, a3)
    assert int((a1 != 0).sum()) > 1000


def test_autograd_l2_loss_and_optimizer_step():
    nvq = _nvq()
    grid, eye, dir = _scene(nvq, 48, 32, seed=7)
    target_grid, _, _ = _scene(nvq, 48, 32, seed=11)
    d = nvq.DeviceGrid(grid)
    t = nvq.DeviceGrid(target_grid)
    E, D = torch.from_numpy(eye).cuda(), torch.from_numpy(dir).cuda()
    with torch.no_grad():
        target, _ = nvq.ops.fog_march(t, E, D, 0.5, 0.05)
    values = d.leaf_values()
    values.requires_grad_(True)
    L, T = nvq.ops.FogMarch.apply(values, d, E, D, 0.5, 0.05, 0, 40)
    loss = ((L - target) ** 2).sum()
    loss.backward()
    g = values.grad.detach().clone()
    acc = nvq.ops.fog_march_backward(d, E, D, L.detach(), 2 * (L.detach() - target))
    assert torch.equal(g.reshape(-1), nvq.ops.fixed_to_float(acc))
    # one gradient step through the in-place view lowers the loss
    with torch.no_grad():
        values -= 0.5 * g / g.abs().max()
        L2, _ = nvq.ops.fog_march(d, E, D, 0.5, 0.05)
        loss2 = ((L2 - target) ** 2).sum()
    assert float(loss2) < float(loss)


def _index_scene(nvq, ref, size=48, res=48, seed=7):
    grid, eye, dir = _scene(nvq, size, res, seed)
    return ref.index_grid(grid, 1, True, True), eye, dir


def test_index_grid_adjoint_matches_fp64_oracle(oracle, ref):
    """Index grids: dLoss/dchannel (a dense scatter by channel index) against the FP64 oracle."""
    nvq = _nvq()
    grid, eye, dir = _index_scene(nvq, ref)
    d = nvq.DeviceGrid(grid)
    r = rng(13)
    wL = r.normal(0, 1, len(eye)).astype(np.float32)
    wT = r.normal(0, 1, len(eye)).astype(np.float32)
    E, D = torch.from_numpy(eye).cuda(), torch.from_numpy(dir).cuda()
    L, T = nvq.ops.fog_march(d, E, D, 0.5, 0.05, image_width=48)
    acc = nvq.ops.fog_march_backward(d, E, D, L, torch.from_numpy(wL).cuda(), T, torch.from_numpy(wT).cuda(), 0.5,
                                     0.05, image_width=48)
    assert acc.numel() == d.value_count
    g = nvq.ops.fixed_to_float(acc).cpu().numpy().astype(np.float64)
    og = oracle.fog_march_adjoint(grid, eye, dir, 0.5, 0.05, wL.astype(np.float64), wT.astype(np.float64))
    scale = np.abs(og).max()
    err = np.abs(g - og)
    print(f"index adjoint: max|g| {scale:.3e}, max err {err.max()/scale:.2e} of max, nonzero {int((og != 0).sum())}",
          flush=True)
    assert int((og != 0).sum()) > 1000
    assert err.max() <= 1e-4 * scale
    assert np.linalg.norm(g - og) <= 1e-4 * np.linalg.norm(og)
    # same sum for another schedule of the same rays
    E2, D2 = E.flip(0).contiguous(), D.flip(0).contiguous()
    L2, T2 = nvq.ops.fog_march(d, E2, D2, 0.5, 0.05)
    acc2 = nvq.ops.fog_march_backward(d, E2, D2, L2, torch.from_numpy(wL[::-1].copy()).cuda(), T2,
                                      torch.from_numpy(wT[::-1].copy()).cuda(), 0.5, 0.05)
    assert torch.equal(acc, acc2)


def test_index_grid_autograd_with_an_external_channel(ref):
    nvq = _nvq()
    grid, eye, dir = _index_scene(nvq, ref, 48, 32, seed=7)
    target_grid, _, _ = _index_scene(nvq, ref, 48, 32, seed=11)
    d = nvq.DeviceGrid(grid)
    t = nvq.DeviceGrid(target_grid)
    E, D = torch.from_numpy(eye).cuda(), torch.from_numpy(dir).cuda()
    with torch.no_grad():
        target, _ = nvq.ops.fog_march(t, E, D, 0.5, 0.05)
    values = d.channel_values().clone().requires_grad_(True)  # a sidecar owned by the caller
    d.set_channel(values.detach())
    L, _ = nvq.ops.FogMarch.apply(values, d, E, D, 0.5, 0.05, 0, 40)
    loss = ((L - target) ** 2).sum()
    loss.backward()
    assert values.grad is not None and values.grad.shape == values.shape
    assert int((values.grad != 0).sum()) > 1000
    with torch.no_grad():
        values -= 0.05 * values.grad / values.grad.abs().max()
        L2, _ = nvq.ops.fog_march(d, E, D, 0.5, 0.05)  # the grid reads the updated sidecar in place
        loss2 = ((L2 - target) ** 2).sum()
    assert float(loss2) < float(loss.detach())
