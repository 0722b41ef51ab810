"""Torch-facing batched queries over a ``DeviceGrid`` (thin wrappers of the C ABI).

Each function mirrors one reference call shape, applied to a batch:

* ``get_value``        ReadAccessor::getValue / isActive / getDim   (NanoVDB.h:5370-5447)
* ``sample_from_voxels`` nanovdb.tools.cuda.sampleFromVoxels(points, d_grid, values[, gradients])
                       (python/cuda/PySampleFromVoxels.cu:58-89)
* ``sample``           SampleFromVoxels<Acc,0|1> + gradient modes   (math/SampleFromVoxels.h)
* ``grad_stencil``     GradStencil                                  (math/Stencils.h:455-576)
* ``zero_crossing``    Ray::worldToIndexF + math::zeroCrossing      (math/HDDA.h:189-214)
* ``iso_crossing``     Ray::worldToIndexF + math::isoCrossing       (math/HDDA.h:225-248)
* ``render_iso_surface`` the iso-surface example's RenderOp (examples/ex_raytrace_iso_surface/common.h)
* ``fog_march``        emission-absorption march with HDDA node skipping
* ``multi_query``      one persistent launch over many grids (zeroCrossing / fog march per grid)
* ``sample_backward``  adjoint of ``sample`` (dLoss/dvalue, deterministic int64 fixed point);
                       ``SampleFromVoxels`` / ``sample_differentiable`` wrap it for autograd

All tensors must be CUDA tensors on the grid's device; work is queued on the
current torch stream.
"""
import ctypes

import torch

from ._lib import check, lib

GRAD_FD6, GRAD_STENCIL, GRAD_TRILINEAR = 1, 2, 3


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _on_grid(x, grid, name):
    """x must be a CUDA tensor on the grid's device: the kernels dereference its pointer there."""
    if not isinstance(x, torch.Tensor) or not x.is_cuda:
        raise ValueError(f"{name} must be a CUDA tensor")
    dev = getattr(grid, "device", None)
    if dev is not None and x.device != dev:
        raise ValueError(f"{name} is on {x.device}, the grid on {dev}")


def _points(x, dtype, name, grid=None):
    _on_grid(x, grid, name)
    if x.dtype != dtype:
        raise ValueError(f"{name} must be {dtype}")
    if x.dim() != 2 or x.shape[1] != 3:
        raise ValueError(f"{name} must have shape [N, 3]")
    return x.contiguous()


def _out(x, grid, name, shape, dtype=torch.float32):
    """An output the kernels write in place: a contiguous tensor of exactly this dtype, shape and
    device (a narrower dtype would be written past its end)."""
    if not isinstance(x, torch.Tensor) or x.dtype != dtype:
        raise ValueError(f"{name} must be a {dtype} tensor, not {getattr(x, 'dtype', type(x))}")
    if tuple(x.shape) != tuple(shape) or not x.is_contiguous():
        raise ValueError(f"{name} must be a contiguous tensor of shape {list(shape)}")
    _on_grid(x, grid, name)


def _rays(grid, eye, dir):
    eye = _points(eye, torch.float32, "eye", grid)
    dir = _points(dir, torch.float32, "dir", grid)
    if eye.shape[0] != dir.shape[0]:
        raise ValueError("eye and dir must hold the same number of rays")
    return eye, dir


def _per_ray(x, grid, name, n):
    """A per-ray float32 input, kept alive by the caller's reference for the length of the call."""
    _on_grid(x, grid, name)
    if x.numel() != n:
        raise ValueError(f"{name} must hold {n} values")
    return x.reshape(-1).contiguous().float()


def get_value(grid, ijk, want_dim=False):
    ijk = _points(ijk, torch.int32, "ijk", grid)
    n = ijk.shape[0]
    val = torch.empty(n, dtype=torch.float32, device=ijk.device)
    act = torch.empty(n, dtype=torch.uint8, device=ijk.device)
    dim = torch.empty(n, dtype=torch.int32, device=ijk.device) if want_dim else None
    check(lib().nvq_get_value(grid.handle, _ptr(ijk), n, _ptr(val), _ptr(act), _ptr(dim), _stream()), "nvq_get_value")
    return (val, act, dim) if want_dim else (val, act)


def sample(grid, xyz, world=False, nn=True, val=True, grad_mode=GRAD_FD6):
    xyz = _points(xyz, torch.float32, "xyz", grid)
    n = xyz.shape[0]
    dev = xyz.device
    o_nn = torch.empty(n, dtype=torch.float32, device=dev) if nn else None
    o_val = torch.empty(n, dtype=torch.float32, device=dev) if val else None
    o_grad = torch.empty((n, 3), dtype=torch.float32, device=dev) if grad_mode else None
    check(lib().nvq_sample(grid.handle, _ptr(xyz), n, int(world), _ptr(o_nn), _ptr(o_val), _ptr(o_grad),
                           int(grad_mode or 0), _stream()), "nvq_sample")
    return o_nn, o_val, o_grad


def sample_from_voxels(points, grid, values, gradients=None):
    """Same arguments and semantics as the reference op: world-space ``points[N,3]``,
    trilinear ``values[N]`` and optional FD6 ``gradients[N,3]`` written in place."""
    points = _points(points, torch.float32, "points", grid)
    n = points.shape[0]
    _out(values, grid, "values", (n,))
    if gradients is not None:
        _out(gradients, grid, "gradients", (n, 3))
    check(lib().nvq_sample(grid.handle, _ptr(points), n, 1, None, _ptr(values), _ptr(gradients), GRAD_FD6,
                           _stream()), "nvq_sample")


def grad_stencil(grid, ijk):
    ijk = _points(ijk, torch.int32, "ijk", grid)
    n = ijk.shape[0]
    dev = ijk.device
    g = torch.empty((n, 3), dtype=torch.float32, device=dev)
    ns = torch.empty(n, dtype=torch.float32, device=dev)
    lap = torch.empty(n, dtype=torch.float32, device=dev)
    zc = torch.empty(n, dtype=torch.uint8, device=dev)
    check(lib().nvq_grad_stencil(grid.handle, _ptr(ijk), n, _ptr(g), _ptr(ns), _ptr(lap), _ptr(zc), _stream()),
          "nvq_grad_stencil")
    return g, ns, lap, zc


def zero_crossing(grid, eye, dir, image_width=0):
    eye, dir = _rays(grid, eye, dir)
    n = eye.shape[0]
    dev = eye.device
    hit = torch.empty(n, dtype=torch.uint8, device=dev)
    ijk = torch.empty((n, 3), dtype=torch.int32, device=dev)
    t = torch.empty(n, dtype=torch.float32, device=dev)
    v = torch.empty(n, dtype=torch.float32, device=dev)
    nrm = torch.empty((n, 3), dtype=torch.float32, device=dev)
    check(lib().nvq_zero_crossing(grid.handle, _ptr(eye), _ptr(dir), n, image_width, _ptr(hit), _ptr(ijk), _ptr(t),
                                  _ptr(v), _ptr(nrm), _stream()), "nvq_zero_crossing")
    return hit, ijk, t, v, nrm


def iso_crossing(grid, eye, dir, iso, image_width=0):
    eye, dir = _rays(grid, eye, dir)
    n = eye.shape[0]
    dev = eye.device
    hit = torch.empty(n, dtype=torch.uint8, device=dev)
    ijk = torch.empty((n, 3), dtype=torch.int32, device=dev)
    t = torch.empty(n, dtype=torch.float32, device=dev)
    v = torch.empty(n, dtype=torch.float32, device=dev)
    nrm = torch.empty((n, 3), dtype=torch.float32, device=dev)
    check(lib().nvq_iso_crossing(grid.handle, _ptr(eye), _ptr(dir), n, image_width, float(iso), _ptr(hit), _ptr(ijk),
                                 _ptr(t), _ptr(v), _ptr(nrm), _stream()), "nvq_iso_crossing")
    return hit, ijk, t, v, nrm


def render_iso_surface(grid, camera, persistent=True, device="cuda"):
    """Render the iso-surface example's image for ``camera`` (a dict from ``tools.iso_camera``).
    Returns (image [h, w], hit [h, w], t [h, w])."""
    from ._lib import IsoCamera
    w, h = int(camera["width"]), int(camera["height"])
    c = IsoCamera(w, h, camera["iso"], camera["voxel_size"], camera["wbbox_dim_z"], camera["tan_half_fov"],
                  (ctypes.c_float * 3)(*camera["origin"]))
    image = torch.empty((h, w), dtype=torch.float32, device=device)
    hit = torch.empty((h, w), dtype=torch.uint8, device=device)
    t = torch.empty((h, w), dtype=torch.float32, device=device)
    check(lib().nvq_render_iso_surface(grid.handle, ctypes.byref(c), int(bool(persistent)), _ptr(image), _ptr(hit),
                                       _ptr(t), _stream()), "nvq_render_iso_surface")
    return image, hit, t


def fog_march(grid, eye, dir, dt=0.5, sigma_t=0.05, image_width=0):
    eye, dir = _rays(grid, eye, dir)
    n = eye.shape[0]
    L = torch.empty(n, dtype=torch.float32, device=eye.device)
    T = torch.empty(n, dtype=torch.float32, device=eye.device)
    check(lib().nvq_fog_march(grid.handle, _ptr(eye), _ptr(dir), n, image_width, dt, sigma_t, _ptr(L), _ptr(T),
                              _stream()), "nvq_fog_march")
    return L, T


OP_ZERO_CROSSING, OP_FOG_MARCH = 0, 1


def multi_query(grids, eye, dir, offsets, image_widths=None, dt=0.5, sigma_t=0.05, ops=None, noalloc=False):
    """One persistent launch over ``grids`` (<= 32, any codecs).  The rays of grid j are
    ``eye/dir[offsets[j]:offsets[j+1]]``; ``image_widths[j]`` (0 = linear batch) makes them a
    row-major image.  ``ops[j]`` is OP_ZERO_CROSSING or OP_FOG_MARCH (None: level sets run
    zeroCrossing, fog volumes the march; index grids need ``ops``).  Returns
    (hit, ijk, a, b, normal): zeroCrossing gives hit/ijk/t/v/normal, the march gives radiance in
    ``a`` and transmittance in ``b``.  ``noalloc`` calls nvq_multi_query_noalloc, whose work
    queue lives in the library's per-device array (the library allocates nothing)."""
    eye, dir = _rays(grids[0] if len(grids) else None, eye, dir)
    for g in grids[1:]:
        _on_grid(eye, g, "eye")
    k = len(grids)
    if len(offsets) != k + 1:
        raise ValueError("offsets must have len(grids) + 1 entries")
    n = eye.shape[0]
    dev = eye.device
    hs = (ctypes.c_void_p * max(k, 1))(*[g.handle.value for g in grids])
    offs = (ctypes.c_uint64 * (k + 1))(*[int(o) for o in offsets])
    widths = None if image_widths is None else (ctypes.c_uint32 * max(k, 1))(*[int(w) for w in image_widths])
    opa = None if ops is None else (ctypes.c_uint32 * max(k, 1))(*[int(x) for x in ops])
    hit = torch.zeros(n, dtype=torch.uint8, device=dev)
    ijk = torch.zeros((n, 3), dtype=torch.int32, device=dev)
    a = torch.zeros(n, dtype=torch.float32, device=dev)
    b = torch.zeros(n, dtype=torch.float32, device=dev)
    nrm = torch.zeros((n, 3), dtype=torch.float32, device=dev)
    fn = lib().nvq_multi_query_noalloc if noalloc else lib().nvq_multi_query
    check(fn(hs, k, offs, widths, opa, _ptr(eye), _ptr(dir), dt, sigma_t, _ptr(hit), _ptr(ijk), _ptr(a), _ptr(b),
             _ptr(nrm), _stream()), "nvq_multi_query_noalloc" if noalloc else "nvq_multi_query")
    return hit, ijk, a, b, nrm


def fog_march_backward(grid, eye, dir, radiance, grad_radiance, transmittance=None, grad_transmittance=None, dt=0.5,
                       sigma_t=0.05, image_width=0, accum=None, fixed_shift=40):
    """Adjoint of ``fog_march``: adds dLoss/dvalue (int64 fixed point, 2^-fixed_shift) into
    ``accum`` [leafCount*512] and returns it (allocated and zeroed when None)."""
    eye, dir = _rays(grid, eye, dir)
    n = eye.shape[0]
    if accum is None:
        accum = torch.zeros(grid.grad_size, dtype=torch.int64, device=eye.device)
    else:
        _out(accum.view(-1) if accum.is_contiguous() else accum, grid, "accum", (grid.grad_size,), torch.int64)
    # float32 copies bound to locals, so each stays alive (and distinct) until the call returns
    L = _per_ray(radiance, grid, "radiance", n)
    T = None if transmittance is None else _per_ray(transmittance, grid, "transmittance", n)
    gr = None if grad_radiance is None else _per_ray(grad_radiance, grid, "grad_radiance", n)
    gt = None if grad_transmittance is None else _per_ray(grad_transmittance, grid, "grad_transmittance", n)
    check(lib().nvq_fog_march_bwd(grid.handle, _ptr(eye), _ptr(dir), n, image_width, dt, sigma_t, _ptr(L), _ptr(T),
                                  _ptr(gr), _ptr(gt), _ptr(accum), fixed_shift, _stream()), "nvq_fog_march_bwd")
    return accum


def fixed_to_float(accum, fixed_shift=40, out=None):
    if not isinstance(accum, torch.Tensor) or not accum.is_cuda or accum.dtype != torch.int64 or \
            not accum.is_contiguous():
        raise ValueError("accum must be a contiguous int64 CUDA tensor")
    if out is None:
        out = torch.empty(accum.shape, dtype=torch.float32, device=accum.device)
    elif not out.is_cuda or out.device != accum.device or out.numel() != accum.numel() or \
            out.dtype != torch.float32 or not out.is_contiguous():
        raise ValueError("out must be a contiguous float32 tensor with accum's element count")
    check(lib().nvq_grad_fixed_to_float(_ptr(accum), accum.numel(), fixed_shift, _ptr(out), _stream()),
          "nvq_grad_fixed_to_float")
    return out


def sample_backward(grid, xyz, grad_nn=None, grad_val=None, grad_grad=None, grad_mode=GRAD_FD6, world=False,
                    accum=None, fixed_shift=40):
    """Adjoint of ``sample``: adds dLoss/dvalue (int64 fixed point, 2^-fixed_shift) for the given
    dLoss/d(nearest) [N], dLoss/d(value) [N] and dLoss/d(gradient) [N, 3] into ``accum``
    [grid.grad_size] and returns it (allocated and zeroed when None).  Float grids receive
    gradients on leaf values only; index grids on every channel entry a query reads."""
    xyz = _points(xyz, torch.float32, "xyz", grid)
    n = xyz.shape[0]
    if accum is None:
        accum = torch.zeros(grid.grad_size, dtype=torch.int64, device=xyz.device)
    else:
        _out(accum, grid, "accum", (grid.grad_size,), torch.int64)
    gn = None if grad_nn is None else _per_ray(grad_nn, grid, "grad_nn", n)
    gv = None if grad_val is None else _per_ray(grad_val, grid, "grad_val", n)
    gg = None if grad_grad is None else _per_ray(grad_grad, grid, "grad_grad", 3 * n)
    if gg is not None and grad_mode not in (GRAD_FD6, GRAD_STENCIL, GRAD_TRILINEAR):
        raise ValueError("grad_grad needs grad_mode 1, 2 or 3")
    check(lib().nvq_sample_bwd(grid.handle, _ptr(xyz), n, int(world), _ptr(gn), _ptr(gv), _ptr(gg),
                               int(grad_mode or 0), _ptr(accum), fixed_shift, _stream()), "nvq_sample_bwd")
    return accum


class SampleFromVoxels(torch.autograd.Function):
    """Differentiable point queries: (nearest, value, gradient) = sample(values) at fixed points.

    ``values`` is the tensor the grid reads in place (``grid.leaf_values()`` for a Float grid, the
    channel for an index grid); backward returns dLoss/dvalues of its shape, reduced
    deterministically (int64 fixed point).  Points are not differentiated."""

    @staticmethod
    def forward(ctx, values, grid, xyz, world, grad_mode, fixed_shift):
        nn, val, g = sample(grid, xyz, world=world, nn=True, val=True, grad_mode=grad_mode)
        ctx.save_for_backward(xyz)
        ctx.grid, ctx.world, ctx.mode, ctx.shift, ctx.shape = grid, world, grad_mode, fixed_shift, values.shape
        return nn, val, g

    @staticmethod
    def backward(ctx, gnn, gval, ggrad):
        (xyz,) = ctx.saved_tensors
        accum = sample_backward(ctx.grid, xyz, gnn, gval, ggrad if ctx.mode else None, ctx.mode, ctx.world,
                                fixed_shift=ctx.shift)
        return fixed_to_float(accum, ctx.shift).view(ctx.shape), None, None, None, None, None


def sample_differentiable(grid, xyz, world=False, grad_mode=GRAD_FD6, fixed_shift=40):
    """``sample`` with autograd through the grid's values (``leaf_values()`` or the channel)."""
    values = grid.channel_values() if grid.grid_type == "OnIndex" else grid.leaf_values()
    return SampleFromVoxels.apply(values, grid, xyz, world, grad_mode, fixed_shift)


class FogMarch(torch.autograd.Function):
    """Differentiable emission-absorption render: (radiance, transmittance) = march(values).

    ``values`` must be what the grid reads in place: ``grid.leaf_values()`` for a Float grid, or
    for an index grid its channel (``grid.channel_values()``, or a tensor given to
    ``grid.set_channel``); the backward pass returns dLoss/dvalues of the same shape, reduced
    deterministically.
    """

    @staticmethod
    def forward(ctx, values, grid, eye, dir, dt, sigma_t, image_width, fixed_shift):
        L, T = fog_march(grid, eye, dir, dt, sigma_t, image_width)
        ctx.save_for_backward(eye, dir, L, T)
        ctx.grid, ctx.dt, ctx.sigma, ctx.width, ctx.shift = grid, dt, sigma_t, image_width, fixed_shift
        ctx.shape = values.shape
        return L, T

    @staticmethod
    def backward(ctx, gL, gT):
        eye, dir, L, T = ctx.saved_tensors
        accum = fog_march_backward(ctx.grid, eye, dir, L, gL, T, gT, ctx.dt, ctx.sigma, ctx.width,
                                   fixed_shift=ctx.shift)
        g = fixed_to_float(accum, ctx.shift).view(ctx.shape)
        return g, None, None, None, None, None, None, None


def fog_render(grid, eye, dir, dt=0.5, sigma_t=0.05, image_width=0, fixed_shift=40):
    """Differentiable fog march over ``grid.leaf_values()`` (Float grids) or the channel
    ``grid.channel_values()`` (index grids); returns (radiance, transmittance)."""
    values = grid.channel_values() if grid.grid_type == "OnIndex" else grid.leaf_values()
    return FogMarch.apply(values, grid, eye, dir, dt, sigma_t, image_width, fixed_shift)
