"""ctypes binding of libnvq.so (the C ABI declared in include/nvq.h).

The shared library is built in-tree by ``openvdb_amd.build.build()``.  There
is no fallback path: if the library is missing or fails to load, every entry
point raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnvq.so")

_vp = ctypes.c_void_p
_u64 = ctypes.c_uint64
_u32 = ctypes.c_uint32
_i = ctypes.c_int
_f = ctypes.c_float
_d = ctypes.c_double


class IsoCamera(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_int32),
        ("height", ctypes.c_int32),
        ("iso", ctypes.c_float),
        ("voxel_size", ctypes.c_float),
        ("wbbox_dim_z", ctypes.c_float),
        ("tan_half_fov", ctypes.c_float),
        ("origin", ctypes.c_float * 3),
    ]


class GridInfo(ctypes.Structure):
    _fields_ = [
        ("grid_type", ctypes.c_uint32),
        ("grid_class", ctypes.c_uint32),
        ("bytes", ctypes.c_uint64),
        ("node_count", ctypes.c_uint32 * 3),
        ("root_tiles", ctypes.c_uint32),
        ("index_bbox", ctypes.c_int32 * 6),
        ("voxel_size", ctypes.c_double * 3),
        ("background", ctypes.c_float),
        ("pad", ctypes.c_uint32),
        ("device_buffer", ctypes.c_void_p),
    ]


# (name, argtypes) of every entry point the header declares
SIGNATURES = {
    "nvq_status_string": ([_i], ctypes.c_char_p),
    "nvq_grid_upload": ([_vp, _u64, _vp, ctypes.POINTER(_vp)], _i),
    "nvq_grid_upload_into": ([_vp, _u64, _vp, _vp, ctypes.POINTER(_vp)], _i),
    "nvq_grid_free": ([_vp], _i),
    "nvq_grid_set_channel": ([_vp, _vp, _u64], _i),
    "nvq_grid_get_info": ([_vp, ctypes.POINTER(GridInfo)], _i),
    "nvq_grid_get_view": ([_vp, _vp, _u64], _i),
    "nvq_grid_set_ray_map": ([_vp, _i], _i),
    "nvq_get_value": ([_vp, _vp, _u64, _vp, _vp, _vp, _vp], _i),
    "nvq_sample": ([_vp, _vp, _u64, _i, _vp, _vp, _vp, _i, _vp], _i),
    "nvq_grad_stencil": ([_vp, _vp, _u64, _vp, _vp, _vp, _vp, _vp], _i),
    "nvq_zero_crossing": ([_vp, _vp, _vp, _u64, _u32, _vp, _vp, _vp, _vp, _vp, _vp], _i),
    "nvq_iso_crossing": ([_vp, _vp, _vp, _u64, _u32, _f, _vp, _vp, _vp, _vp, _vp, _vp], _i),
    "nvq_render_iso_surface": ([_vp, ctypes.POINTER(IsoCamera), _i, _vp, _vp, _vp, _vp], _i),
    "nvq_fog_march": ([_vp, _vp, _vp, _u64, _u32, _f, _f, _vp, _vp, _vp], _i),
    "nvq_multi_query": ([_vp, _u32, _vp, _vp, _vp, _vp, _vp, _f, _f, _vp, _vp, _vp, _vp, _vp, _vp], _i),
    "nvq_multi_query_noalloc": ([_vp, _u32, _vp, _vp, _vp, _vp, _vp, _f, _f, _vp, _vp, _vp, _vp, _vp, _vp], _i),
    "nvq_fog_march_bwd": ([_vp, _vp, _vp, _u64, _u32, _f, _f, _vp, _vp, _vp, _vp, _vp, _i, _vp], _i),
    "nvq_grad_fixed_to_float": ([_vp, _u64, _i, _vp, _vp], _i),
    "nvq_sample_bwd": ([_vp, _vp, _u64, _i, _vp, _vp, _vp, _i, _vp, _i, _vp], _i),
    "nvq_voxels_to_grid": ([_vp, _u64, _u32, _d, ctypes.c_char_p, _vp, _vp, ctypes.POINTER(_vp)], _i),
    "nvq_voxels_to_grid_fp": ([_vp, _i, _u64, _u32, _d, ctypes.c_char_p, _vp, _vp, ctypes.POINTER(_vp)], _i),
    "nvq_points_to_grid": ([_vp, _u64, _d, _u32, ctypes.c_char_p, _vp, ctypes.POINTER(_vp)], _i),
    "nvq_points_to_grid_f64": ([_vp, _u64, _d, _u32, ctypes.c_char_p, _vp, ctypes.POINTER(_vp)], _i),
    "nvq_points_to_grid_density": ([_vp, _i, _u64, _u32, _i, _i, _u32, ctypes.c_char_p, _vp, ctypes.POINTER(_vp),
                                    ctypes.POINTER(_d), ctypes.POINTER(_u32)], _i),
    "nvq_grid_download": ([_vp, _vp, _u64, _vp], _i),
    "nvq_update_grid_stats": ([_vp, _vp], _i),
    "nvq_update_grid_stats_mode": ([_vp, _i, _vp], _i),
    "nvq_signed_flood_fill": ([_vp, _vp], _i),
    "nvq_profile_enable": ([_i], _i),
    "nvq_profile_read": ([ctypes.c_char_p, ctypes.POINTER(_d), ctypes.POINTER(_u64)], _i),
    "nvq_profile_reset": ([], _i),
    # host builder (include/nvq_build.h)
    "nvqb_level_set_sphere": ([_d, _vp, _d, _d, _u32, _f, ctypes.POINTER(_vp), ctypes.POINTER(_u64)], _i),
    "nvqb_level_set_torus": ([_d, _d, _vp, _d, _d, _u32, _f, ctypes.POINTER(_vp), ctypes.POINTER(_u64)], _i),
    "nvqb_fog_volume_sphere": ([_d, _vp, _d, _d, ctypes.POINTER(_vp), ctypes.POINTER(_u64)], _i),
    "nvqb_fog_fbm": ([_vp, _u32, _d, _u64, _f, _f, _u32, _f, ctypes.POINTER(_vp), ctypes.POINTER(_u64)], _i),
    "nvqb_fbm_density": ([_vp, _u64, _u32, _d, _u64, _f, _f, _vp], _i),
    "nvqb_quantize": ([_vp, _u64, _u32, _f, ctypes.POINTER(_vp), ctypes.POINTER(_u64)], _i),
    "nvqb_camera_rays": ([_u32, _u32, _vp, _vp, _vp, _f, _vp, _vp], _i),
    "nvqb_iso_camera": ([_vp, _u64, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(IsoCamera)], _i),
    "nvqb_read_nvdb": ([ctypes.c_char_p, ctypes.c_int32, ctypes.c_char_p, ctypes.POINTER(_vp), ctypes.POINTER(_u64)], _i),
    "nvqb_write_nvdb": ([ctypes.c_char_p, _vp, _u64], _i),
    "nvqb_crc32": ([_vp, _u64], _u32),
    "nvqb_free": ([_vp], None),
}

_lib = None


class NvqError(RuntimeError):
    def __init__(self, status, what=""):
        self.status = status
        msg = lib().nvq_status_string(status).decode() if _lib is not None else str(status)
        super().__init__(f"{what}: {msg} (status {status})" if what else f"{msg} (status {status})")


def lib():
    """Load libnvq.so once; raise if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run openvdb_amd.build.build() (hipcc, gfx950)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def check(status, what=""):
    if status != 0:
        raise NvqError(status, what)
