"""Host grid producers and camera rays (libnvq.so's builder, include/nvq_build.h).

Names mirror the reference's tools (tools/CreatePrimitives.h, tools/CreateNanoGrid.h):
each returns a 32-byte aligned uint8 numpy array holding one NanoVDB grid buffer,
ready for ``DeviceGrid``.
"""
import ctypes

import numpy as np

from ._lib import check, lib

GRID_TYPE = {"float": 1, "fp4": 13, "fp8": 14, "fp16": 15, "fpn": 16}


def _gt(grid_type):
    if isinstance(grid_type, int):
        return grid_type
    return GRID_TYPE[grid_type.lower()]


def _take(ptr, nbytes):
    raw = np.empty(nbytes.value + 32, np.uint8)
    off = (-raw.ctypes.data) % 32
    out = raw[off:off + nbytes.value]
    ctypes.memmove(out.ctypes.data, ptr, nbytes.value)
    lib().nvqb_free(ptr)
    return out


def _call(fn, *args):
    p = ctypes.c_void_p()
    n = ctypes.c_uint64()
    check(fn(*args, ctypes.byref(p), ctypes.byref(n)), fn.__name__)
    return _take(p, n)


def _d3(v):
    return (ctypes.c_double * 3)(*[float(x) for x in v])


def _f3(v):
    return (ctypes.c_float * 3)(*[float(x) for x in v])


def create_level_set_sphere(radius=100.0, center=(0.0, 0.0, 0.0), voxel_size=1.0, half_width=3.0,
                            grid_type="float", tol=1e-3):
    """createLevelSetSphere (tools/CreatePrimitives.h:1016-1100)."""
    return _call(lib().nvqb_level_set_sphere, radius, _d3(center), voxel_size, half_width, _gt(grid_type), tol)


def create_level_set_torus(major_radius=100.0, minor_radius=50.0, center=(0.0, 0.0, 0.0), voxel_size=1.0,
                           half_width=3.0, grid_type="float", tol=1e-3):
    """createLevelSetTorus (tools/CreatePrimitives.h:1186-1243)."""
    return _call(lib().nvqb_level_set_torus, major_radius, minor_radius, _d3(center), voxel_size, half_width,
                 _gt(grid_type), tol)


def create_fog_volume_sphere(radius=100.0, center=(0.0, 0.0, 0.0), voxel_size=1.0, half_width=3.0):
    """createFogVolumeSphere<float> (tools/CreatePrimitives.h:1103-1130)."""
    return _call(lib().nvqb_fog_volume_sphere, radius, _d3(center), voxel_size, half_width)


def create_fog_fbm(bbox, octaves=6, base_freq=1.0 / 64.0, seed=7, amplitude=1.0, threshold=0.15, grid_type="float",
                   tol=1e-3):
    """Fog volume of clamped fBm improved-Perlin noise filled like build::Grid::operator()(func, bbox)."""
    bb = (ctypes.c_int32 * 6)(*[int(x) for x in bbox])
    return _call(lib().nvqb_fog_fbm, bb, octaves, base_freq, seed, amplitude, threshold, _gt(grid_type), tol)


def fbm_density(ijk, octaves=6, base_freq=1.0 / 64.0, seed=7, amplitude=1.0, threshold=0.15):
    ijk = np.ascontiguousarray(ijk, np.int32).reshape(-1, 3)
    out = np.empty(len(ijk), np.float32)
    check(lib().nvqb_fbm_density(ijk.ctypes.data_as(ctypes.c_void_p), len(ijk), octaves, base_freq, seed, amplitude,
                                 threshold, out.ctypes.data_as(ctypes.c_void_p)), "nvqb_fbm_density")
    return out


def quantize(float_grid, grid_type, tol=1e-3):
    """createNanoGrid<NanoGrid<float>, Fp4|Fp8|Fp16|FpN> (tools/CreateNanoGrid.h:1331-1474)."""
    buf = np.ascontiguousarray(float_grid)
    return _call(lib().nvqb_quantize, buf.ctypes.data_as(ctypes.c_void_p), buf.nbytes, _gt(grid_type), tol)


def camera_rays(width, height, eye, look_at=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov_deg=45.0):
    """Pinhole rays with the RayGenOp convention (examples/ex_raytrace_level_set/common.h:91-125)."""
    n = width * height
    e = np.empty((n, 3), np.float32)
    d = np.empty((n, 3), np.float32)
    check(lib().nvqb_camera_rays(width, height, _f3(eye), _f3(look_at), _f3(up), fov_deg,
                                 e.ctypes.data_as(ctypes.c_void_p), d.ctypes.data_as(ctypes.c_void_p)),
          "nvqb_camera_rays")
    return e, d


def iso_camera(grid, width, height):
    """The iso-surface example's camera for a host grid buffer (RenderOp's constructor,
    examples/ex_raytrace_iso_surface/common.h:26-37) as a dict of nvq_iso_camera's fields."""
    from ._lib import IsoCamera
    buf = np.ascontiguousarray(grid)
    cam = IsoCamera()
    check(lib().nvqb_iso_camera(buf.ctypes.data_as(ctypes.c_void_p), buf.nbytes, int(width), int(height),
                                ctypes.byref(cam)), "nvqb_iso_camera")
    return {"width": cam.width, "height": cam.height, "iso": cam.iso, "voxel_size": cam.voxel_size,
            "wbbox_dim_z": cam.wbbox_dim_z, "tan_half_fov": cam.tan_half_fov, "origin": list(cam.origin)}


def read_nvdb(path, index=0, name=None):
    """Grid `index` (or the first grid called `name`) of a NanoVDB file or raw grid file, as a
    uint8 host buffer ready for ``DeviceGrid`` (io::readGrid semantics; uncompressed files)."""
    ptr = ctypes.c_void_p()
    n = ctypes.c_uint64()
    check(lib().nvqb_read_nvdb(str(path).encode(), int(index), None if name is None else name.encode(),
                               ctypes.byref(ptr), ctypes.byref(n)), "nvqb_read_nvdb")
    return _take(ptr, n)


def write_nvdb(path, grid):
    """Write one grid buffer as an uncompressed NanoVDB file (io::writeGrid with Codec::NONE)."""
    buf = np.ascontiguousarray(grid)
    check(lib().nvqb_write_nvdb(str(path).encode(), buf.ctypes.data_as(ctypes.c_void_p), buf.nbytes),
          "nvqb_write_nvdb")


def active_voxels(grid):
    """(ijk [n, 3] int32, values [n] float32) of the active voxels inside the leaves of a float grid
    buffer, in leaf order then bit order (LeafNode::beginValueOn, NanoVDB.h).  Active tiles of the
    internal nodes are not expanded."""
    buf = np.ascontiguousarray(grid).reshape(-1).view(np.uint8)
    if int(np.frombuffer(buf[636:640].tobytes(), np.uint32)[0]) != 1:
        raise ValueError("active_voxels needs a float grid")
    tree = 672
    off = int(np.frombuffer(buf[tree:tree + 8].tobytes(), np.int64)[0])
    count = int(np.frombuffer(buf[tree + 32:tree + 36].tobytes(), np.uint32)[0])
    leaves = buf[tree + off:tree + off + 2144 * count].reshape(count, 2144)
    origin = leaves[:, 0:12].copy().view(np.int32) & ~7
    mask = np.unpackbits(leaves[:, 16:80], axis=1, bitorder="little").astype(bool)
    n = np.arange(512)
    local = np.stack([n >> 6, (n >> 3) & 7, n & 7], 1).astype(np.int32)
    li, vi = np.nonzero(mask)
    values = np.ascontiguousarray(leaves[:, 96:]).view(np.float32)[li, vi]
    return (origin[li] + local[vi]).astype(np.int32), values
