"""Error paths that must fail cleanly on the host: malformed .nvdb headers, the status codes of
the C ABI, and the argument checks of the torch layer (no GPU needed)."""
import ctypes

import numpy as np
import pytest
import torch


def _lib():
    import openvdb_amd
    return openvdb_amd.lib()


def _read(path, index=0):
    p, n = ctypes.c_void_p(), ctypes.c_uint64()
    st = _lib().nvqb_read_nvdb(str(path).encode(), index, None, ctypes.byref(p), ctypes.byref(n))
    if st == 0:
        _lib().nvqb_free(p)
    return st


@pytest.fixture(scope="module")
def small_grid():
    import openvdb_amd as nvq
    return nvq.tools.create_level_set_sphere(12.0, (0.0, 0.0, 0.0), 1.0, 3.0)


def test_nvdb_reader_rejects_sizes_past_the_file(small_grid, tmp_path):
    lib = _lib()
    good = tmp_path / "good.nvdb"
    assert lib.nvqb_write_nvdb(str(good).encode(), small_grid.ctypes.data_as(ctypes.c_void_p), small_grid.nbytes) == 0
    assert _read(good) == 0
    data = bytearray(good.read_bytes())
    # FileHeader (16 B) then FileMetaData: nameSize at +136, fileSize at +8, gridSize at +0
    for off, val, fmt in ((16 + 136, 0xFFFFFFFF, "<u4"), (16 + 136, 1 << 30, "<u4"), (16 + 8, 1 << 60, "<u8"),
                          (16 + 0, 1 << 40, "<u8"), (16 + 0, 0, "<u8")):
        bad = bytearray(data)
        raw = np.array([val], dtype=fmt).tobytes()
        bad[off:off + len(raw)] = raw
        path = tmp_path / "bad.nvdb"
        path.write_bytes(bytes(bad))
        # asking for grid 1 walks past grid 0 with the corrupted sizes
        assert _read(path, 0) in (0, -6) and _read(path, 1) in (-1, -6), (off, val)
        if off == 16 + 136:
            assert _read(path, 0) == -6


def test_raw_grid_file_with_zero_size_is_refused(small_grid, tmp_path):
    raw = bytearray(small_grid.tobytes())
    path = tmp_path / "raw.nvdb"
    path.write_bytes(bytes(raw))
    assert _read(path, 0) == 0
    assert _read(path, 1) == -1  # no second grid
    raw[32:40] = np.uint64(0).tobytes()  # GridData::mGridSize = 0 would never advance
    path.write_bytes(bytes(raw))
    assert _read(path, 0) == -6
    assert _read(path, 3) == -6
    raw[32:40] = np.uint64(len(raw) + 4096).tobytes()  # past the end of the file
    path.write_bytes(bytes(raw))
    assert _read(path, 0) == -6


def test_status_mapping():
    lib = _lib()
    from openvdb_amd import NvqError
    assert b"fixed-point" in lib.nvq_status_string(-9)
    assert lib.nvq_status_string(-10) == b"unknown status"
    e = NvqError(-9, "nvq_fog_march_bwd")
    assert e.status == -9 and "fixed-point" in str(e)


class _StubGrid:
    """Stands in for a DeviceGrid: the argument checks run before any handle is used."""
    handle = ctypes.c_void_p(0)
    device = torch.device("cuda", 0)
    grad_size = 1024


def test_sample_from_voxels_rejects_bad_outputs():
    from openvdb_amd import ops
    g = _StubGrid()
    pts = torch.zeros((4, 3), dtype=torch.float32)
    with pytest.raises(ValueError, match="CUDA"):  # a host pointer would reach the kernel
        ops.sample_from_voxels(pts, g, torch.zeros(4))
    # the output checks come first, so they are reachable without a GPU
    for bad in (torch.zeros(4, dtype=torch.float16), torch.zeros(4, dtype=torch.bfloat16), torch.zeros(4, dtype=torch.float64),
                torch.zeros(3), torch.zeros((4, 2))[:, 0]):
        with pytest.raises(ValueError):
            ops._out(bad, g, "values", (4,))
    with pytest.raises(ValueError, match="CUDA"):
        ops._out(torch.zeros(4), g, "values", (4,))
    with pytest.raises(ValueError):
        ops._out(torch.zeros((4, 3), dtype=torch.float16), g, "gradients", (4, 3))


def test_backward_rejects_bad_accumulators():
    from openvdb_amd import ops
    g = _StubGrid()
    for bad in (torch.zeros(1024, dtype=torch.float32), torch.zeros(1023, dtype=torch.int64),
                torch.zeros(2048, dtype=torch.int64)[::2], torch.zeros(1024, dtype=torch.int32)):
        with pytest.raises(ValueError):
            ops._out(bad, g, "accum", (g.grad_size,), torch.int64)
    with pytest.raises(ValueError, match="CUDA"):
        ops._out(torch.zeros(1024, dtype=torch.int64), g, "accum", (g.grad_size,), torch.int64)
    with pytest.raises(ValueError):
        ops.fixed_to_float(torch.zeros(8, dtype=torch.int64))  # host tensor
    with pytest.raises(ValueError):
        ops._per_ray(torch.zeros(5), g, "radiance", 5)  # host tensor
    with pytest.raises(ValueError):
        ops.fog_march_backward(g, torch.zeros((5, 3)), torch.zeros((5, 3)), torch.zeros(5), torch.zeros(5))
