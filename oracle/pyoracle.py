"""ctypes front-end for the two checkers in this folder (test infrastructure only).

* ``Oracle`` wraps ``liboracle.so``: the C restatement (oracle.c) of NanoVDB's
  query path.
* ``Reference`` wraps ``_ref/libnanoref.so``: the reference NanoVDB host code
  compiled from the reference's own headers (ref_wrap.cpp).

Only tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of bench.py
import this module; the product package never does.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(_HERE, "liboracle.so")
REF_SO = os.path.join(_HERE, "_ref", "libnanoref.so")

_vp = ctypes.c_void_p
_u64 = ctypes.c_uint64
_i = ctypes.c_int
_f = ctypes.c_float
_d = ctypes.c_double


def _p(a):
    return None if a is None else a.ctypes.data_as(_vp)


def aligned_empty(nbytes):
    """Uninitialised uint8 array whose data is 32-byte aligned (NanoVDB.h:69-74)."""
    raw = np.empty(nbytes + 32, np.uint8)
    off = (-raw.ctypes.data) % 32
    return raw[off:off + nbytes]


class _SpanQueries:
    """Node spans, first hits, DDA traces and the stencil extras, bound to the same C signature in
    liboracle.so (prefix "or_") and libnanoref.so (prefix "ref_")."""

    def _bind_span_queries(self, prefix):
        L = self.L
        self._pre = prefix
        f = lambda name: getattr(L, prefix + name)  # noqa: E731
        f("tree_march").argtypes = [_vp, _vp, _vp, _u64, _i, ctypes.c_uint32, _vp, _vp, _vp, _vp]
        f("first_active").argtypes = [_vp, _vp, _vp, _u64, _vp, _vp, _vp]
        f("dda_trace").argtypes = [_vp, _vp, _f, _f, _i, _i, _vp, _vp, _vp, _vp]
        f("box_stencil").argtypes = [_vp, _vp, _u64, _vp, _vp]
        f("grad_stencil_ext").argtypes = [_vp, _vp, _u64, _vp, _vp, _vp]
        f("sample_zero_crossing").argtypes = [_vp, _vp, _u64, _vp]

    def _fn(self, name):
        return getattr(self.L, self._pre + name)

    def tree_march(self, buf, eye, dir, level, max_spans=256):
        """TreeMarcher spans of world rays: (count[n], t0[n,max], t1[n,max], node byte offset[n,max])."""
        eye = np.ascontiguousarray(eye, np.float32).reshape(-1, 3)
        dir = np.ascontiguousarray(dir, np.float32).reshape(-1, 3)
        n = eye.shape[0]
        cnt = np.zeros(n, np.uint32)
        t0 = np.zeros((n, max_spans), np.float32)
        t1 = np.zeros((n, max_spans), np.float32)
        node = np.zeros((n, max_spans), np.int64)
        self._fn("tree_march")(_p(buf), _p(eye), _p(dir), n, level, max_spans, _p(cnt), _p(t0), _p(t1), _p(node))
        return cnt, t0, t1, node

    def first_active(self, buf, eye, dir):
        eye = np.ascontiguousarray(eye, np.float32).reshape(-1, 3)
        dir = np.ascontiguousarray(dir, np.float32).reshape(-1, 3)
        n = eye.shape[0]
        hit = np.empty(n, np.uint8)
        ijk = np.empty((n, 3), np.int32)
        t = np.empty(n, np.float32)
        self._fn("first_active")(_p(buf), _p(eye), _p(dir), n, _p(hit), _p(ijk), _p(t))
        return hit, ijk, t

    def dda_trace(self, eye, dir, t0, t1, dim, nsteps):
        eye = np.asarray(eye, np.float32)
        dir = np.asarray(dir, np.float32)
        times = np.empty(nsteps + 1, np.float32)
        nexts = np.empty(nsteps + 1, np.float32)
        vox = np.empty((nsteps + 1, 3), np.int32)
        ok = np.empty(nsteps + 1, np.uint8)
        self._fn("dda_trace")(_p(eye), _p(dir), t0, t1, dim, nsteps, _p(times), _p(nexts), _p(vox), _p(ok))
        return times, nexts, vox, ok

    def box_stencil(self, buf, xyz):
        xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
        n = xyz.shape[0]
        interp = np.empty(n, np.float32)
        grad = np.empty((n, 3), np.float32)
        self._fn("box_stencil")(_p(buf), _p(xyz), n, _p(interp), _p(grad))
        return interp, grad

    def grad_stencil_ext(self, buf, ijk, V):
        ijk = np.ascontiguousarray(ijk, np.int32).reshape(-1, 3)
        V = np.ascontiguousarray(V, np.float32).reshape(-1, 3)
        n = ijk.shape[0]
        up = np.empty((n, 3), np.float32)
        cpt = np.empty((n, 3), np.float32)
        self._fn("grad_stencil_ext")(_p(buf), _p(ijk), n, _p(V), _p(up), _p(cpt))
        return up, cpt

    def sample_zero_crossing(self, buf, xyz):
        xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
        zc = np.empty(xyz.shape[0], np.uint8)
        self._fn("sample_zero_crossing")(_p(buf), _p(xyz), xyz.shape[0], _p(zc))
        return zc


class Oracle(_SpanQueries):
    def __init__(self, path=ORACLE_SO):
        L = ctypes.CDLL(path)
        self.L = L
        L.or_validate.argtypes = [_vp, _u64]
        L.or_grid_type.argtypes = [_vp]
        L.or_grid_type.restype = ctypes.c_uint32
        L.or_leaf_count.argtypes = [_vp]
        L.or_leaf_count.restype = ctypes.c_uint32
        L.or_get_value.argtypes = [_vp, _vp, _u64, _vp, _vp, _vp]
        L.or_sample.argtypes = [_vp, _vp, _u64, _i, _vp, _vp, _vp, _i]
        L.or_sample_f64.argtypes = [_vp, _vp, _u64, _vp, _vp]
        L.or_grad_stencil.argtypes = [_vp, _vp, _u64, _vp, _vp, _vp, _vp]
        L.or_zero_crossing.argtypes = [_vp, _vp, _vp, _u64, _vp, _vp, _vp, _vp, _vp]
        L.or_value_count.argtypes = [_vp]
        L.or_value_count.restype = _u64
        L.or_iso_crossing.argtypes = [_vp, _vp, _vp, _u64, _f, _vp, _vp, _vp, _vp, _vp]
        L.or_iso_crossing_map.argtypes = [_vp, _vp, _vp, _u64, _f, _i, _vp, _vp, _vp, _vp, _vp]
        L.or_iso_render.argtypes = [_vp, ctypes.c_int32, ctypes.c_int32, _f, _f, _f, _f, _vp, _vp, _vp, _vp]
        L.or_fog_march.argtypes = [_vp, _vp, _vp, _u64, _f, _f, _vp, _vp, _vp]
        L.or_fog_march_adjoint.argtypes = [_vp, _vp, _vp, _u64, _f, _f, _vp, _vp, _vp]
        L.or_sample_bwd_fixed.argtypes = [_vp, _vp, _u64, _i, _vp, _vp, _vp, _i, _i, _vp, _vp]
        L.or_sample_bwd_f64.argtypes = [_vp, _vp, _u64, _vp, _vp, _vp, _i, _vp]
        L.or_sample_loss_f64.argtypes = [_vp, _vp, _u64, _vp, _vp, _vp, _i, ctypes.c_int64, _d]
        L.or_sample_loss_f64.restype = _d
        L.or_hdda_trace.argtypes = [_vp, _vp, _f, _f, _i, _i, _vp, _vp, _vp, _vp]
        L.or_ray_clip.argtypes = [_vp, _vp, _f, _f, _vp, _vp]
        L.or_voxels_to_grid.argtypes = [_vp, _u64, ctypes.c_uint32, _d, ctypes.c_char_p, ctypes.POINTER(_vp),
                                        ctypes.POINTER(_u64), _vp]
        L.or_free.argtypes = [_vp]
        L.or_update_grid_stats.argtypes = [_vp]
        L.or_update_grid_stats_mode.argtypes = [_vp, ctypes.c_int]
        L.or_signed_flood_fill.argtypes = [_vp]
        L.or_points_to_grid.argtypes = [_vp, _u64, _d, ctypes.c_uint32, ctypes.c_char_p, ctypes.POINTER(_vp),
                                        ctypes.POINTER(_u64)]
        L.or_points_to_grid_f64.argtypes = L.or_points_to_grid.argtypes
        L.or_points_to_grid_density.argtypes = [_vp, ctypes.c_int, _u64, ctypes.c_uint32, ctypes.c_int, ctypes.c_int,
                                                ctypes.c_uint32, ctypes.c_char_p, ctypes.POINTER(_vp),
                                                ctypes.POINTER(_u64), ctypes.POINTER(_d), ctypes.POINTER(ctypes.c_uint32)]
        self._bind_span_queries("or_")

    def validate(self, buf):
        return self.L.or_validate(_p(buf), buf.nbytes)

    def leaf_count(self, buf):
        return self.L.or_leaf_count(_p(buf))

    def get_value(self, buf, ijk):
        ijk = np.ascontiguousarray(ijk, np.int32).reshape(-1, 3)
        n = ijk.shape[0]
        val = np.empty(n, np.float32)
        act = np.empty(n, np.uint8)
        dim = np.empty(n, np.uint32)
        self.L.or_get_value(_p(buf), _p(ijk), n, _p(val), _p(act), _p(dim))
        return val, act, dim

    def sample(self, buf, xyz, world=False, grad_mode=1):
        xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
        n = xyz.shape[0]
        nn = np.empty(n, np.float32)
        val = np.empty(n, np.float32)
        grad = np.empty((n, 3), np.float32)
        self.L.or_sample(_p(buf), _p(xyz), n, int(world), _p(nn), _p(val), _p(grad), grad_mode)
        return nn, val, grad

    def sample_f64(self, buf, xyz):
        xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
        n = xyz.shape[0]
        val = np.empty(n, np.float64)
        grad = np.empty((n, 3), np.float64)
        self.L.or_sample_f64(_p(buf), _p(xyz), n, _p(val), _p(grad))
        return val, grad

    def grad_stencil(self, buf, ijk):
        ijk = np.ascontiguousarray(ijk, np.int32).reshape(-1, 3)
        n = ijk.shape[0]
        g = np.empty((n, 3), np.float32)
        ns = np.empty(n, np.float32)
        lap = np.empty(n, np.float32)
        zc = np.empty(n, np.uint8)
        self.L.or_grad_stencil(_p(buf), _p(ijk), n, _p(g), _p(ns), _p(lap), _p(zc))
        return g, ns, lap, zc

    def zero_crossing(self, buf, eye, dir, f64map=False):
        """f64map: the index-space ray comes from the FP64 Ray::applyInverseMap instead of worldToIndexF."""
        eye = np.ascontiguousarray(eye, np.float32).reshape(-1, 3)
        dir = np.ascontiguousarray(dir, np.float32).reshape(-1, 3)
        n = eye.shape[0]
        hit = np.empty(n, np.uint8)
        ijk = np.empty((n, 3), np.int32)
        t = np.empty(n, np.float32)
        v = np.empty(n, np.float32)
        nrm = np.empty((n, 3), np.float32)
        self.L.or_iso_crossing_map(_p(buf), _p(eye), _p(dir), n, 0.0, int(f64map), _p(hit), _p(ijk), _p(t), _p(v),
                                   _p(nrm))
        return hit, ijk, t, v, nrm

    def iso_crossing(self, buf, eye, dir, iso):
        eye = np.ascontiguousarray(eye, np.float32).reshape(-1, 3)
        dir = np.ascontiguousarray(dir, np.float32).reshape(-1, 3)
        n = eye.shape[0]
        hit = np.empty(n, np.uint8)
        ijk = np.empty((n, 3), np.int32)
        t = np.empty(n, np.float32)
        v = np.empty(n, np.float32)
        nrm = np.empty((n, 3), np.float32)
        self.L.or_iso_crossing(_p(buf), _p(eye), _p(dir), n, iso, _p(hit), _p(ijk), _p(t), _p(v), _p(nrm))
        return hit, ijk, t, v, nrm

    def iso_render(self, buf, cam):
        """cam: dict with width, height, iso, voxel_size, wbbox_dim_z, tan_half_fov, origin."""
        w, h = int(cam["width"]), int(cam["height"])
        image = np.empty(w * h, np.float32)
        hit = np.empty(w * h, np.uint8)
        t = np.empty(w * h, np.float32)
        origin = np.ascontiguousarray(cam["origin"], np.float32)
        self.L.or_iso_render(_p(buf), w, h, cam["iso"], cam["voxel_size"], cam["wbbox_dim_z"], cam["tan_half_fov"],
                             _p(origin), _p(image), _p(hit), _p(t))
        return image, hit, t

    def fog_march(self, buf, eye, dir, dt, sigma):
        eye = np.ascontiguousarray(eye, np.float32).reshape(-1, 3)
        dir = np.ascontiguousarray(dir, np.float32).reshape(-1, 3)
        n = eye.shape[0]
        L = np.empty(n, np.float64)
        T = np.empty(n, np.float64)
        steps = np.empty(n, np.uint32)
        self.L.or_fog_march(_p(buf), _p(eye), _p(dir), n, dt, sigma, _p(L), _p(T), _p(steps))
        return L, T, steps

    def fog_march_adjoint(self, buf, eye, dir, dt, sigma, dL, dT=None):
        eye = np.ascontiguousarray(eye, np.float32).reshape(-1, 3)
        dir = np.ascontiguousarray(dir, np.float32).reshape(-1, 3)
        n = eye.shape[0]
        dL = np.ascontiguousarray(dL, np.float64)
        dT = None if dT is None else np.ascontiguousarray(dT, np.float64)
        is_index = int(np.frombuffer(buf[636:640].tobytes(), np.uint32)[0]) == 20
        grad = np.zeros(int(self.L.or_value_count(_p(buf))) if is_index else self.leaf_count(buf) * 512, np.float64)
        self.L.or_fog_march_adjoint(_p(buf), _p(eye), _p(dir), n, dt, sigma, _p(dL), _p(dT), _p(grad))
        return grad

    def grad_size(self, buf):
        """Length of an adjoint accumulator: value count (index grids) or leafCount*512."""
        is_index = int(np.frombuffer(buf[636:640].tobytes(), np.uint32)[0]) == 20
        return int(self.L.or_value_count(_p(buf))) if is_index else self.leaf_count(buf) * 512

    def sample_bwd_fixed(self, buf, xyz, gnn=None, gval=None, ggrad=None, grad_mode=1, world=False, shift=40,
                         accum=None):
        """FP32 / int64 fixed-point adjoint of sample(): (accum int64[grad_size], overflow flag)."""
        xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
        n = xyz.shape[0]
        f = lambda a, k: None if a is None else np.ascontiguousarray(a, np.float32).reshape(n * k)
        gnn, gval, ggrad = f(gnn, 1), f(gval, 1), f(ggrad, 3)
        if accum is None:
            accum = np.zeros(self.grad_size(buf), np.int64)
        over = ctypes.c_int(0)
        self.L.or_sample_bwd_fixed(_p(buf), _p(xyz), n, int(world), _p(gnn), _p(gval), _p(ggrad), int(grad_mode),
                                   int(shift), _p(accum), ctypes.byref(over))
        return accum, bool(over.value)

    def sample_bwd_f64(self, buf, xyz, gnn=None, gval=None, ggrad=None, grad_mode=1):
        """Exact FP64 adjoint of sample() (index-space points): float64[grad_size]."""
        xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
        n = xyz.shape[0]
        f = lambda a, k: None if a is None else np.ascontiguousarray(a, np.float64).reshape(n * k)
        grad = np.zeros(self.grad_size(buf), np.float64)
        self.L.or_sample_bwd_f64(_p(buf), _p(xyz), n, _p(f(gnn, 1)), _p(f(gval, 1)), _p(f(ggrad, 3)), int(grad_mode),
                                 _p(grad))
        return grad

    def sample_loss_f64(self, buf, xyz, gnn=None, gval=None, ggrad=None, grad_mode=1, key=-1, delta=0.0):
        """FP64 loss sum(gnn*nn + gval*val + ggrad.grad) with value `key` raised by delta."""
        xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
        n = xyz.shape[0]
        f = lambda a, k: None if a is None else np.ascontiguousarray(a, np.float64).reshape(n * k)
        a, b, c = f(gnn, 1), f(gval, 1), f(ggrad, 3)
        return float(self.L.or_sample_loss_f64(_p(buf), _p(xyz), n, _p(a), _p(b), _p(c), int(grad_mode), int(key),
                                               float(delta)))

    def voxels_to_grid(self, ijk, grid_type=20, voxel_size=1.0, name=None, want_index=False):
        """voxelsToGrid<float | ValueOnIndex | ValueIndex> restated (grid_type 1, 20, 19): (buffer,
        per-input value index or None)."""
        ijk = np.ascontiguousarray(ijk, dtype=np.int32).reshape(-1, 3)
        n = len(ijk)
        index = np.zeros(n, np.uint64) if want_index else None
        p, nb = _vp(), _u64()
        st = self.L.or_voxels_to_grid(_p(ijk), n, grid_type, voxel_size, None if name is None else name.encode(),
                                      ctypes.byref(p), ctypes.byref(nb), _p(index))
        if st != 0:
            raise ValueError("or_voxels_to_grid failed: %d" % st)
        buf = aligned_empty(nb.value)
        ctypes.memmove(buf.ctypes.data, p.value, nb.value)
        self.L.or_free(p)
        return buf, index

    def update_grid_stats(self, buf, mode=3):
        """cuda::updateGridStats<float>(grid, mode) restated (mode 1 BBox, 2 MinMax, 3 All = Default):
        a copy of the float grid buffer with its bboxes, statistics and flags recomputed in the
        device code's merge order."""
        out = aligned_empty(buf.nbytes)
        out[:] = buf
        st = self.L.or_update_grid_stats_mode(_p(out), mode)
        if st != 0:
            raise ValueError("or_update_grid_stats failed: %d" % st)
        return out

    def points_to_grid(self, xyz, voxel_size=1.0, point_type=9, name=None):
        """pointsToGrid restated for float or double (Vec3d) world points: the Point grid buffer."""
        f64 = np.asarray(xyz).dtype == np.float64
        xyz = np.ascontiguousarray(xyz, dtype=np.float64 if f64 else np.float32).reshape(-1, 3)
        p, nb = _vp(), _u64()
        fn = self.L.or_points_to_grid_f64 if f64 else self.L.or_points_to_grid
        st = fn(_p(xyz), len(xyz), voxel_size, point_type, None if name is None else name.encode(), ctypes.byref(p),
                ctypes.byref(nb))
        if st != 0:
            raise ValueError("or_points_to_grid failed: %d" % st)
        buf = aligned_empty(nb.value)
        ctypes.memmove(buf.ctypes.data, p.value, nb.value)
        self.L.or_free(p)
        return buf

    def points_to_grid_density(self, xyz, max_points_per_voxel, tolerance=1, max_iterations=10, point_type=9,
                               name=None):
        """PointsToGrid(maxPointsPerVoxel, tolerance, maxIterations) restated: (buffer, voxel size,
        largest points per voxel)."""
        f64 = np.asarray(xyz).dtype == np.float64
        xyz = np.ascontiguousarray(xyz, dtype=np.float64 if f64 else np.float32).reshape(-1, 3)
        p, nb, dx, dens = _vp(), _u64(), _d(), ctypes.c_uint32()
        st = self.L.or_points_to_grid_density(_p(xyz), int(f64), len(xyz), max_points_per_voxel, tolerance,
                                              max_iterations, point_type, None if name is None else name.encode(),
                                              ctypes.byref(p), ctypes.byref(nb), ctypes.byref(dx), ctypes.byref(dens))
        if st != 0:
            raise ValueError("or_points_to_grid_density failed: %d" % st)
        buf = aligned_empty(nb.value)
        ctypes.memmove(buf.ctypes.data, p.value, nb.value)
        self.L.or_free(p)
        return buf, dx.value, dens.value

    def signed_flood_fill(self, buf):
        """cuda::signedFloodFill<float> restated, on a copy of the float grid buffer."""
        out = aligned_empty(buf.nbytes)
        out[:] = buf
        st = self.L.or_signed_flood_fill(_p(out))
        if st != 0:
            raise ValueError("or_signed_flood_fill failed: %d" % st)
        return out

    def hdda_trace(self, eye, dir, t0, t1, dim, nsteps):
        eye = np.asarray(eye, np.float32)
        dir = np.asarray(dir, np.float32)
        times = np.empty(nsteps + 1, np.float32)
        nexts = np.empty(nsteps + 1, np.float32)
        vox = np.empty((nsteps + 1, 3), np.int32)
        ok = np.empty(nsteps + 1, np.uint8)
        self.L.or_hdda_trace(_p(eye), _p(dir), t0, t1, dim, nsteps, _p(times), _p(nexts), _p(vox), _p(ok))
        return times, nexts, vox, ok

    def ray_clip(self, eye, dir, t0, t1, bbox):
        out = np.empty(2, np.float32)
        hit = self.L.or_ray_clip(_p(np.asarray(eye, np.float32)), _p(np.asarray(dir, np.float32)), t0, t1,
                                 _p(np.asarray(bbox, np.int32)), _p(out))
        return bool(hit), float(out[0]), float(out[1])


class Reference(_SpanQueries):
    """The reference NanoVDB host code (compiled from the reference headers)."""

    FLOAT, FP4, FP8, FP16, FPN = 0, 1, 2, 3, 4
    VOXEL_FN = ctypes.CFUNCTYPE(ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p)

    def __init__(self, path=REF_SO):
        L = ctypes.CDLL(path)
        self.L = L
        pu = ctypes.POINTER(_u64)
        L.ref_hw_threads.restype = ctypes.c_uint
        L.ref_free.argtypes = [_vp]
        L.ref_level_set_sphere.restype = _vp
        L.ref_level_set_sphere.argtypes = [_i, _d, _d, _d, _d, _d, _d, _f, pu]
        L.ref_level_set_torus.restype = _vp
        L.ref_level_set_torus.argtypes = [_i, _d, _d, _d, _d, _d, _d, _d, _f, pu]
        L.ref_fog_volume_sphere.restype = _vp
        L.ref_fog_volume_sphere.argtypes = [_d, _d, _d, _d, _d, _d, pu]
        L.ref_grid_from_function.restype = _vp
        L.ref_grid_from_function.argtypes = [self.VOXEL_FN, _vp, _vp, _f, _i, pu]
        L.ref_grid_from_dense.restype = _vp
        L.ref_grid_from_dense.argtypes = [_vp, _vp, _f, _i, pu]
        L.ref_quantize.restype = _vp
        L.ref_quantize.argtypes = [_vp, _i, _f, pu]
        L.ref_get_value.argtypes = [_vp, _vp, _u64, _vp, _vp, _vp, _i]
        L.ref_sample.argtypes = [_vp, _vp, _u64, _i, _vp, _vp, _vp, _i, _i]
        L.ref_grad_stencil.argtypes = [_vp, _vp, _u64, _vp, _vp, _vp, _vp]
        L.ref_zero_crossing.argtypes = [_vp, _vp, _vp, _u64, _vp, _vp, _vp, _vp, _vp, _i]
        L.ref_zero_crossing_map.argtypes = [_vp, _vp, _vp, _u64, _i, _vp, _vp, _vp, _vp, _vp, _i]
        L.ref_iso_crossing.argtypes = [_vp, _vp, _vp, _u64, _f, _vp, _vp, _vp, _vp, _i]
        L.ref_write_nvdb.argtypes = [ctypes.c_char_p, _vp, _vp, _i, _i, _i]
        L.ref_index_grid.argtypes = [_vp, ctypes.c_uint32, _i, _i, ctypes.POINTER(_u64)]
        L.ref_index_grid.restype = _vp
        L.ref_read_nvdb.argtypes = [ctypes.c_char_p, _i, ctypes.c_char_p, ctypes.POINTER(_u64)]
        L.ref_read_nvdb.restype = _vp
        L.ref_fog_march.argtypes = [_vp, _vp, _vp, _u64, _f, _f, _vp, _vp, _vp, _vp, _vp, _i]
        L.ref_ray_gen.argtypes = [_i, _i, _f, _f, _f, _f, _vp, _vp]
        L.ref_hdda_trace.argtypes = [_vp, _vp, _f, _f, _i, _i, _vp, _vp, _vp, _vp]
        L.ref_ray_clip.argtypes = [_vp, _vp, _f, _f, _vp, _vp]
        L.ref_validate.argtypes = [_vp, _i]
        L.ref_grid_from_voxels.argtypes = [_vp, _u64, _i, _d, ctypes.POINTER(_u64)]
        L.ref_grid_from_voxels.restype = _vp
        L.ref_index_values.argtypes = [_vp, _vp, _u64, _vp, _vp]
        L.ref_update_grid_stats.argtypes = [_vp, ctypes.c_int]
        L.ref_point_voxels.argtypes = [_vp, _vp, _u64, _vp, _vp]
        self._bind_span_queries("ref_")

    def hw_threads(self):
        return int(self.L.ref_hw_threads())

    def _take(self, ptr, n):
        if not ptr:
            raise RuntimeError("reference grid construction failed")
        res = aligned_empty(n)
        ctypes.memmove(res.ctypes.data, ptr, n)
        self.L.ref_free(ptr)
        return res

    def level_set_sphere(self, radius, center=(0, 0, 0), voxel=1.0, half_width=3.0, kind=0, tol=1e-3):
        n = _u64()
        p = self.L.ref_level_set_sphere(kind, radius, *center, voxel, half_width, tol, ctypes.byref(n))
        return self._take(p, n.value)

    def level_set_torus(self, R, r, center=(0, 0, 0), voxel=1.0, half_width=3.0, kind=0, tol=1e-3):
        n = _u64()
        p = self.L.ref_level_set_torus(kind, R, r, *center, voxel, half_width, tol, ctypes.byref(n))
        return self._take(p, n.value)

    def fog_volume_sphere(self, radius, center=(0, 0, 0), voxel=1.0, half_width=3.0):
        n = _u64()
        p = self.L.ref_fog_volume_sphere(radius, *center, voxel, half_width, ctypes.byref(n))
        return self._take(p, n.value)

    def grid_from_function(self, fn, bbox, background=0.0, grid_class=2):
        """fn(i, j, k) -> float, evaluated by build::Grid::operator()(func, bbox)."""
        cb = self.VOXEL_FN(lambda i, j, k, _ctx: fn(i, j, k))
        n = _u64()
        p = self.L.ref_grid_from_function(cb, None, _p(np.asarray(bbox, np.int32)), background, grid_class,
                                          ctypes.byref(n))
        return self._take(p, n.value)

    def grid_from_dense(self, dense, bbox, background=0.0, grid_class=2):
        """build::Grid::operator()(func, bbox) with func reading dense[x][y][z] over bbox."""
        dense = np.ascontiguousarray(dense, np.float32)
        n = _u64()
        p = self.L.ref_grid_from_dense(_p(dense), _p(np.asarray(bbox, np.int32)), background, grid_class,
                                       ctypes.byref(n))
        return self._take(p, n.value)

    def quantize(self, buf, kind, tol=1e-3):
        n = _u64()
        p = self.L.ref_quantize(_p(buf), kind, tol, ctypes.byref(n))
        return self._take(p, n.value)

    def get_value(self, buf, ijk, nthreads=1):
        ijk = np.ascontiguousarray(ijk, np.int32).reshape(-1, 3)
        n = ijk.shape[0]
        val = np.empty(n, np.float32)
        act = np.empty(n, np.uint8)
        dim = np.empty(n, np.uint32)
        self.L.ref_get_value(_p(buf), _p(ijk), n, _p(val), _p(act), _p(dim), nthreads)
        return val, act, dim

    def sample(self, buf, xyz, world=False, grad_mode=1, nthreads=1, nn=True, val=True, grad=True):
        xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
        n = xyz.shape[0]
        o_nn = np.empty(n, np.float32) if nn else None
        o_val = np.empty(n, np.float32) if val else None
        o_grad = np.empty((n, 3), np.float32) if grad else None
        self.L.ref_sample(_p(buf), _p(xyz), n, int(world), _p(o_nn), _p(o_val), _p(o_grad), grad_mode, nthreads)
        return o_nn, o_val, o_grad

    def grad_stencil(self, buf, ijk):
        ijk = np.ascontiguousarray(ijk, np.int32).reshape(-1, 3)
        n = ijk.shape[0]
        g = np.empty((n, 3), np.float32)
        ns = np.empty(n, np.float32)
        lap = np.empty(n, np.float32)
        zc = np.empty(n, np.uint8)
        self.L.ref_grad_stencil(_p(buf), _p(ijk), n, _p(g), _p(ns), _p(lap), _p(zc))
        return g, ns, lap, zc

    def zero_crossing(self, buf, eye, dir, nthreads=1, f64map=False):
        eye = np.ascontiguousarray(eye, np.float32).reshape(-1, 3)
        dir = np.ascontiguousarray(dir, np.float32).reshape(-1, 3)
        n = eye.shape[0]
        hit = np.empty(n, np.uint8)
        ijk = np.empty((n, 3), np.int32)
        t = np.empty(n, np.float32)
        v = np.empty(n, np.float32)
        nrm = np.empty((n, 3), np.float32)
        self.L.ref_zero_crossing_map(_p(buf), _p(eye), _p(dir), n, int(f64map), _p(hit), _p(ijk), _p(t), _p(v),
                                     _p(nrm), nthreads)
        return hit, ijk, t, v, nrm

    def iso_crossing(self, buf, eye, dir, iso, nthreads=1):
        eye = np.ascontiguousarray(eye, np.float32).reshape(-1, 3)
        dir = np.ascontiguousarray(dir, np.float32).reshape(-1, 3)
        n = eye.shape[0]
        hit = np.empty(n, np.uint8)
        ijk = np.empty((n, 3), np.int32)
        t = np.empty(n, np.float32)
        v = np.empty(n, np.float32)
        self.L.ref_iso_crossing(_p(buf), _p(eye), _p(dir), n, iso, _p(hit), _p(ijk), _p(t), _p(v), nthreads)
        return hit, ijk, t, v

    def index_grid(self, float_buf, channels=1, include_stats=True, include_tiles=True):
        n = ctypes.c_uint64()
        p = self.L.ref_index_grid(_p(float_buf), channels, int(include_stats), int(include_tiles), ctypes.byref(n))
        return self._take(p, n.value)

    def validate(self, buf, full=True):
        """tools::isValid(grid, CheckMode::Full | Half): True when the reference accepts the buffer."""
        st = self.L.ref_validate(_p(buf), int(full))
        if st < 0:
            raise ValueError("grid type not covered by the reference dispatch")
        return st == 1

    def point_voxels(self, buf, ijk):
        """PointAccessor::voxelPoints per voxel: (point count, first position in the blind data)."""
        ijk = np.ascontiguousarray(ijk, dtype=np.int32).reshape(-1, 3)
        count = np.zeros(len(ijk), np.uint64)
        first = np.zeros(len(ijk), np.uint64)
        if self.L.ref_point_voxels(_p(buf), _p(ijk), len(ijk), _p(count), _p(first)) != 0:
            raise ValueError("not a point grid")
        return count, first

    def update_grid_stats(self, buf, mode=3):
        """Host tools::updateGridStats(StatsMode(mode)) on a copy of a float grid buffer."""
        out = aligned_empty(buf.nbytes)
        out[:] = buf
        if self.L.ref_update_grid_stats(_p(out), mode) != 0:
            raise ValueError("not a float grid")
        return out

    def index_values(self, buf, ijk, with_active=False):
        """ReadAccessor<ValueOnIndex | ValueIndex>::getValue per voxel (uint64 value indices), and
        isActive with ``with_active``."""
        ijk = np.ascontiguousarray(ijk, dtype=np.int32).reshape(-1, 3)
        out = np.zeros(len(ijk), np.uint64)
        active = np.zeros(len(ijk), np.uint8)
        if self.L.ref_index_values(_p(buf), _p(ijk), len(ijk), _p(out), _p(active)) != 0:
            raise ValueError("not an index grid")
        return (out, active.astype(bool)) if with_active else out

    def grid_from_voxels(self, ijk, grid_type=1, voxel_size=1.0):
        """build::Grid<float> with value 1 at each voxel, converted by createNanoGrid."""
        ijk = np.ascontiguousarray(ijk, dtype=np.int32).reshape(-1, 3)
        n = ctypes.c_uint64()
        p = self.L.ref_grid_from_voxels(_p(ijk), len(ijk), grid_type, voxel_size, ctypes.byref(n))
        return self._take(p, n.value)

    def write_nvdb(self, path, bufs, merge=False, raw=False):
        bufs = [np.ascontiguousarray(b) for b in bufs]
        ptrs = (ctypes.c_void_p * len(bufs))(*[b.ctypes.data for b in bufs])
        sizes = (ctypes.c_uint64 * len(bufs))(*[b.nbytes for b in bufs])
        st = self.L.ref_write_nvdb(str(path).encode(), ptrs, sizes, len(bufs), int(merge), int(raw))
        if st != 0:
            raise RuntimeError(f"ref_write_nvdb failed ({st})")

    def read_nvdb(self, path, index=0, name=None):
        n = ctypes.c_uint64()
        p = self.L.ref_read_nvdb(str(path).encode(), int(index), None if name is None else name.encode(), ctypes.byref(n))
        if not p:
            raise RuntimeError("ref_read_nvdb failed")
        return self._take(p, n.value)

    def fog_march(self, buf, eye, dir, dt, sigma, nthreads=1, fp32=True, fp64=True):
        eye = np.ascontiguousarray(eye, np.float32).reshape(-1, 3)
        dir = np.ascontiguousarray(dir, np.float32).reshape(-1, 3)
        n = eye.shape[0]
        L = np.empty(n, np.float32) if fp32 else None
        T = np.empty(n, np.float32) if fp32 else None
        L64 = np.empty(n, np.float64) if fp64 else None
        T64 = np.empty(n, np.float64) if fp64 else None
        steps = np.empty(n, np.uint32)
        self.L.ref_fog_march(_p(buf), _p(eye), _p(dir), n, dt, sigma, _p(L), _p(T), _p(L64), _p(T64), _p(steps),
                             nthreads)
        return L, T, L64, T64, steps

    def ray_gen(self, w, h, dimz, center=(0, 0, 0)):
        eye = np.empty((w * h, 3), np.float32)
        dir = np.empty((w * h, 3), np.float32)
        self.L.ref_ray_gen(w, h, dimz, *center, _p(eye), _p(dir))
        return eye, dir

    def hdda_trace(self, eye, dir, t0, t1, dim, nsteps):
        eye = np.asarray(eye, np.float32)
        dir = np.asarray(dir, np.float32)
        times = np.empty(nsteps + 1, np.float32)
        nexts = np.empty(nsteps + 1, np.float32)
        vox = np.empty((nsteps + 1, 3), np.int32)
        ok = np.empty(nsteps + 1, np.uint8)
        self.L.ref_hdda_trace(_p(eye), _p(dir), t0, t1, dim, nsteps, _p(times), _p(nexts), _p(vox), _p(ok))
        return times, nexts, vox, ok

    def ray_clip(self, eye, dir, t0, t1, bbox):
        out = np.empty(2, np.float32)
        hit = self.L.ref_ray_clip(_p(np.asarray(eye, np.float32)), _p(np.asarray(dir, np.float32)), t0, t1,
                                  _p(np.asarray(bbox, np.int32)), _p(out))
        return bool(hit), float(out[0]), float(out[1])


def have_reference():
    return os.path.exists(REF_SO)
