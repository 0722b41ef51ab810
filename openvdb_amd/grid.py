"""Device-resident NanoVDB grids (the GridHandle::deviceUpload / deviceGrid<T>() role).

A ``DeviceGrid`` owns one device copy of an unmodified NanoVDB buffer, made by
``nvq_grid_upload`` (one hipMalloc + one hipMemcpyAsync; GridHandle.h:159-194,
cuda/DeviceBuffer.h:422-437).  Errors surface as ``NvqError`` with the C-ABI
status, where the reference's ``deviceGrid<T>()`` would return nullptr.
"""
import ctypes

import numpy as np

from ._lib import GridInfo, check, lib

GRID_TYPES = {1: "float", 13: "Fp4", 14: "Fp8", 15: "Fp16", 16: "FpN", 20: "OnIndex", 19: "Index", 23: "PointIndex"}
GRID_CLASSES = {0: "Unknown", 1: "LevelSet", 2: "FogVolume", 4: "PointIndex", 5: "PointData", 8: "IndexGrid"}


def aligned_buffer(data):
    """Copy bytes into a 32-byte aligned uint8 array (NanoVDB.h:69-74)."""
    src = np.frombuffer(data, np.uint8) if not isinstance(data, np.ndarray) else data.reshape(-1).view(np.uint8)
    raw = np.empty(src.nbytes + 32, np.uint8)
    off = (-raw.ctypes.data) % 32
    out = raw[off:off + src.nbytes]
    out[:] = src
    return out


def _stream_ptr(stream):
    if stream is None:
        import torch
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    if isinstance(stream, int):
        return ctypes.c_void_p(stream)
    return ctypes.c_void_p(stream.cuda_stream)


class DeviceGrid:
    """An uploaded grid.  ``host_buffer`` is a uint8 numpy array holding the NanoVDB buffer."""

    def __init__(self, host_buffer, stream=None):
        import torch
        buf = np.ascontiguousarray(host_buffer).reshape(-1).view(np.uint8)
        if buf.ctypes.data % 32:
            buf = aligned_buffer(buf)
        self._host = buf
        # device memory owned by torch (caching allocator, 256-byte aligned); the grid wraps it
        self.device_bytes = torch.empty(buf.nbytes, dtype=torch.uint8, device="cuda")
        handle = ctypes.c_void_p()
        s = _stream_ptr(stream)
        check(lib().nvq_grid_upload_into(buf.ctypes.data_as(ctypes.c_void_p), buf.nbytes,
                                         ctypes.c_void_p(self.device_bytes.data_ptr()), s, ctypes.byref(handle)),
              "nvq_grid_upload_into")
        self.handle = handle
        self._device = self.device_bytes.device
        torch.cuda.synchronize()  # the copy is asynchronous: keep the host bytes alive until it is done
        info = GridInfo()
        check(lib().nvq_grid_get_info(handle, ctypes.byref(info)), "nvq_grid_get_info")
        self.info = info

    @classmethod
    def from_voxels(cls, ijk, grid_type="OnIndex", voxel_size=1.0, name=None, return_index=False, stream=None):
        """Build a grid on the device from integer voxels: tools::cuda::voxelsToGrid<BuildT>
        (tools/cuda/PointsToGrid.cuh:119-125) for BuildT = ValueOnIndex ("OnIndex") or float.

        ``ijk`` is an [n, 3] int32 CUDA tensor (duplicates allowed), or float32 / float64 index-space
        coordinates that are rounded to the nearest voxel (Vec3f / Vec3d input).  With ``return_index`` (index
        grids) also returns an int64 CUDA tensor holding each input voxel's value index, so that
        ``channel[index] = values`` fills a channel for ``set_channel``.  The library owns the new
        grid's device buffer (``leaf_values`` is not available on it)."""
        import torch
        codes = {"OnIndex": 20, "Index": 19, "float": 1}
        if grid_type not in codes:
            raise ValueError("grid_type must be 'OnIndex', 'Index' or 'float'")
        if not (ijk.is_cuda and ijk.dtype in (torch.int32, torch.float32, torch.float64) and ijk.dim() == 2 and
                ijk.shape[1] == 3):
            raise ValueError("ijk must be an [n, 3] int32, float32 or float64 CUDA tensor")
        ijk = ijk.contiguous()
        n = ijk.shape[0]
        index = torch.empty(n, dtype=torch.int64, device=ijk.device) if return_index else None
        handle = ctypes.c_void_p()
        args = (n, codes[grid_type], float(voxel_size), None if name is None else name.encode(),
                None if index is None else ctypes.c_void_p(index.data_ptr()), _stream_ptr(stream), ctypes.byref(handle))
        if ijk.dtype == torch.int32:
            check(lib().nvq_voxels_to_grid(ctypes.c_void_p(ijk.data_ptr()), *args), "nvq_voxels_to_grid")
        else:  # index-space coordinates rounded to the nearest voxel (Vec3f / Vec3d input)
            check(lib().nvq_voxels_to_grid_fp(ctypes.c_void_p(ijk.data_ptr()), int(ijk.dtype == torch.float64), *args),
                  "nvq_voxels_to_grid_fp")
        g = cls.__new__(cls)
        g.handle = handle
        g.device_bytes = None
        g._device = torch.device("cuda", torch.cuda.current_device())
        info = GridInfo()
        check(lib().nvq_grid_get_info(handle, ctypes.byref(info)), "nvq_grid_get_info")
        g.info = info
        g._host = g.download(stream, limit=1024)  # grid, tree and root headers
        return (g, index) if return_index else g

    POINT_TYPES = {"PointID": 1, "World64": 2, "World32": 3, "Grid64": 4, "Grid32": 5, "Voxel32": 6, "Voxel16": 7,
                   "Voxel8": 8, "Default": 9}

    @classmethod
    def from_points(cls, xyz, voxel_size=1.0, point_type="Default", name=None, stream=None):
        """Build a point grid on the device from float world-space points: tools::cuda::pointsToGrid
        (tools/cuda/PointsToGrid.cuh:57-64).  ``xyz`` is an [n, 3] float32 (Vec3f) or float64 (Vec3d)
        CUDA tensor; the points are stored as blind data encoded by ``point_type`` (a key of
        POINT_TYPES).  The queries of this library do not read point grids; ``download`` returns
        the NanoVDB buffer."""
        import torch
        if point_type not in cls.POINT_TYPES:
            raise ValueError(f"point_type must be one of {sorted(cls.POINT_TYPES)}")
        if not (xyz.is_cuda and xyz.dtype in (torch.float32, torch.float64) and xyz.dim() == 2 and xyz.shape[1] == 3):
            raise ValueError("xyz must be an [n, 3] float32 or float64 CUDA tensor")
        xyz = xyz.contiguous()
        handle = ctypes.c_void_p()
        fn = "nvq_points_to_grid_f64" if xyz.dtype == torch.float64 else "nvq_points_to_grid"
        check(getattr(lib(), fn)(ctypes.c_void_p(xyz.data_ptr()), xyz.shape[0], float(voxel_size),
                                 cls.POINT_TYPES[point_type], None if name is None else name.encode(),
                                 _stream_ptr(stream), ctypes.byref(handle)), fn)
        g = cls.__new__(cls)
        g.handle = handle
        g.device_bytes = None
        g._device = torch.device("cuda", torch.cuda.current_device())
        info = GridInfo()
        check(lib().nvq_grid_get_info(handle, ctypes.byref(info)), "nvq_grid_get_info")
        g.info = info
        g._host = g.download(stream, limit=1024)
        return g

    @classmethod
    def from_points_density(cls, xyz, max_points_per_voxel, tolerance=1, max_iterations=10, point_type="Default",
                            name=None, stream=None):
        """Point grid whose voxel size gives about ``max_points_per_voxel`` points in the densest
        voxel: PointsToGrid(maxPointsPerVoxel, tolerance, maxIterations)
        (tools/cuda/PointsToGrid.cuh:335-342).  Returns the grid; its ``points_per_voxel`` is the
        largest count reached (PointsToGrid::maxPointsPerVoxel) and ``voxel_size`` the chosen size."""
        import torch
        if point_type not in cls.POINT_TYPES:
            raise ValueError(f"point_type must be one of {sorted(cls.POINT_TYPES)}")
        if not (xyz.is_cuda and xyz.dtype in (torch.float32, torch.float64) and xyz.dim() == 2 and xyz.shape[1] == 3):
            raise ValueError("xyz must be an [n, 3] float32 or float64 CUDA tensor")
        xyz = xyz.contiguous()
        handle = ctypes.c_void_p()
        dx, dens = ctypes.c_double(), ctypes.c_uint32()
        check(lib().nvq_points_to_grid_density(ctypes.c_void_p(xyz.data_ptr()), int(xyz.dtype == torch.float64),
                                               xyz.shape[0], int(max_points_per_voxel), int(tolerance),
                                               int(max_iterations), cls.POINT_TYPES[point_type],
                                               None if name is None else name.encode(), _stream_ptr(stream),
                                               ctypes.byref(handle), ctypes.byref(dx), ctypes.byref(dens)),
              "nvq_points_to_grid_density")
        g = cls.__new__(cls)
        g.handle = handle
        g.device_bytes = None
        g._device = torch.device("cuda", torch.cuda.current_device())
        info = GridInfo()
        check(lib().nvq_grid_get_info(handle, ctypes.byref(info)), "nvq_grid_get_info")
        g.info = info
        g._host = g.download(stream, limit=1024)
        g.points_per_voxel = int(dens.value)
        return g

    def download(self, stream=None, limit=None):
        """The grid buffer (or its first ``limit`` bytes) as a 32-byte aligned uint8 numpy array
        (GridHandle::deviceDownload)."""
        n = self.nbytes if limit is None else min(self.nbytes, int(limit))
        out = aligned_buffer(np.zeros(n, np.uint8))
        check(lib().nvq_grid_download(self.handle, out.ctypes.data_as(ctypes.c_void_p), n, _stream_ptr(stream)),
              "nvq_grid_download")
        return out

    STATS_MODES = {"BBox": 1, "MinMax": 2, "All": 3, "Default": 3}

    def update_stats(self, mode="Default", stream=None):
        """Recompute bboxes and statistics on the device (tools::cuda::updateGridStats(grid, mode),
        float grids), e.g. after the leaf values were changed through ``leaf_values``.  ``mode`` is
        a StatsMode name: "BBox", "MinMax" or "All" (= "Default")."""
        if mode not in self.STATS_MODES:
            raise ValueError(f"mode must be one of {sorted(self.STATS_MODES)}")
        check(lib().nvq_update_grid_stats_mode(self.handle, self.STATS_MODES[mode], _stream_ptr(stream)),
              "nvq_update_grid_stats_mode")
        info = GridInfo()
        check(lib().nvq_grid_get_info(self.handle, ctypes.byref(info)), "nvq_grid_get_info")
        self.info = info

    def signed_flood_fill(self, stream=None):
        """Re-sign the inactive values of a float level set on the device
        (tools::cuda::signedFloodFill)."""
        check(lib().nvq_signed_flood_fill(self.handle, _stream_ptr(stream)), "nvq_signed_flood_fill")

    def leaf_values(self):
        """[leafCount, 512] float32 view of the leaf values inside the device buffer (Float grids).
        Writing to it changes the grid; it is the parameter tensor of the differentiable march."""
        if self.info.grid_type != 1:
            raise ValueError("leaf_values() needs a Float grid")
        if self.device_bytes is None:
            raise ValueError("leaf_values() needs a grid uploaded from host memory")
        import torch
        leaf_off = 672 + int(np.frombuffer(self._host[672:680].tobytes(), np.int64)[0])
        f32 = self.device_bytes.view(torch.float32)
        return torch.as_strided(f32, (self.leaf_count, 512), (2144 // 4, 1), (leaf_off + 96) // 4)

    def channel_values(self):
        """Index grids: [value_count] float32 view of the grid's own channel inside the device buffer
        (the dense sidecar; writing to it changes the values the grid returns).  It is the parameter
        tensor of the differentiable march on an index grid."""
        if self.info.grid_type != 20:
            raise ValueError("channel_values() needs an index grid")
        import torch
        meta = int(np.frombuffer(self._host[640:648].tobytes(), np.int64)[0])
        count = int(np.frombuffer(self._host[648:652].tobytes(), np.uint32)[0])
        if count == 0 or int(np.frombuffer(self._host[meta + 28:meta + 32].tobytes(), np.uint32)[0]) != 1:
            raise ValueError("the grid has no float channel")
        data = meta + int(np.frombuffer(self._host[meta:meta + 8].tobytes(), np.int64)[0])
        f32 = self.device_bytes.view(torch.float32)
        return f32[data // 4:data // 4 + self.value_count]

    @property
    def grad_size(self):
        """Length of the adjoint's accumulator: leafCount*512 (Float grids) or value_count (index grids)."""
        return self.value_count if self.info.grid_type == 20 else self.leaf_count * 512

    @property
    def value_count(self):
        """Index grids: number of values the channel holds (GridData::mData1)."""
        return int(np.frombuffer(self._host[656:664].tobytes(), np.uint64)[0])

    def set_channel(self, channel):
        """Index grids: read values from ``channel`` (a float32 CUDA tensor of at least
        ``value_count`` elements, e.g. a trainable sidecar) instead of the grid's own channel;
        None restores the grid's own channel.  The tensor is kept alive while in use."""
        if channel is None:
            check(lib().nvq_grid_set_channel(self.handle, None, 0), "nvq_grid_set_channel")
            self._channel = None
            return
        if not channel.is_cuda or channel.dtype.itemsize != 4 or not channel.is_contiguous():
            raise ValueError("channel must be a contiguous float32 CUDA tensor")
        check(lib().nvq_grid_set_channel(self.handle, ctypes.c_void_p(channel.data_ptr()), channel.numel()),
              "nvq_grid_set_channel")
        self._channel = channel

    def set_ray_map(self, mode):
        """How ray queries map world rays to index space: "f32" (Ray::worldToIndexF, the default)
        or "f64" (Ray::applyInverseMap, the FP64 Map, for eyes far from the origin of grids with a
        large translation or a non-power-of-two voxel size)."""
        modes = {"f32": 0, "f64": 1}
        if mode not in modes:
            raise ValueError("ray map must be 'f32' or 'f64'")
        check(lib().nvq_grid_set_ray_map(self.handle, modes[mode]), "nvq_grid_set_ray_map")

    @property
    def device(self):
        """The torch device whose memory holds the grid; queries take tensors on this device."""
        return self._device

    @property
    def grid_type(self):
        return GRID_TYPES.get(self.info.grid_type, str(self.info.grid_type))

    @property
    def grid_class(self):
        return GRID_CLASSES.get(self.info.grid_class, str(self.info.grid_class))

    @property
    def leaf_count(self):
        return int(self.info.node_count[0])

    @property
    def nbytes(self):
        return int(self.info.bytes)

    @property
    def index_bbox(self):
        return tuple(self.info.index_bbox)

    @property
    def voxel_size(self):
        return tuple(self.info.voxel_size)

    def close(self):
        if self.handle is not None and self.handle.value:
            check(lib().nvq_grid_free(self.handle), "nvq_grid_free")
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
