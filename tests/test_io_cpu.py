"""NanoVDB files (uncompressed) and raw grid files: our reader and writer against the reference's
io::readGrid / io::writeGrid(s) and GridHandle::write, byte for byte (CPU only)."""
import zlib

import numpy as np
import pytest

from pyoracle import have_reference

pytestmark = pytest.mark.skipif(not have_reference(), reason="reference sources not built")


def _same(a, b):
    return a.nbytes == b.nbytes and np.array_equal(np.frombuffer(a, np.uint8), np.frombuffer(b, np.uint8))


@pytest.fixture(scope="module")
def grids(ref):
    import openvdb_amd as nvq
    return [
        ref.level_set_sphere(30.0, (1.0, 2.0, 3.0), 0.5, 3.0),               # reference builder: checksum on
        nvq.tools.create_level_set_torus(40.0, 10.0, (0.0, 0.0, 0.0), 1.0, 3.0),  # ours: checksum disabled
        nvq.tools.quantize(nvq.tools.create_fog_fbm([0, 0, 0, 47, 47, 47], seed=3), "fp8"),  # Fp8
        nvq.tools.create_fog_volume_sphere(20.0, (0.0, 0.0, 0.0), 1.0, 3.0),
    ]


def test_crc32_is_the_ieee_crc():
    import ctypes
    import openvdb_amd as nvq
    lib = nvq.lib()
    check = b"123456789"
    assert lib.nvqb_crc32(ctypes.c_char_p(check), len(check)) == 0xCBF43926
    r = np.random.default_rng(1)
    for n in (0, 1, 7, 736, 100003):
        d = r.integers(0, 256, n, dtype=np.uint8)
        assert lib.nvqb_crc32(d.ctypes.data_as(ctypes.c_void_p), n) == zlib.crc32(d.tobytes())


def test_read_files_written_by_the_reference(ref, grids, tmp_path):
    import openvdb_amd as nvq
    path = tmp_path / "four.nvdb"
    ref.write_nvdb(path, grids)  # one segment per grid
    for i, g in enumerate(grids):
        ours = nvq.tools.read_nvdb(path, i)
        assert _same(ours, ref.read_nvdb(path, i))
        assert _same(ours, g)
    name = bytes(grids[1][40:296]).split(b"\0")[0].decode()
    assert _same(nvq.tools.read_nvdb(path, name=name), ref.read_nvdb(path, name=name))
    with pytest.raises(nvq.NvqError):
        nvq.tools.read_nvdb(path, 4)
    with pytest.raises(nvq.NvqError):
        nvq.tools.read_nvdb(path, name="no such grid")


def test_writer_matches_the_reference_writer(ref, grids, tmp_path):
    import openvdb_amd as nvq
    for i, g in enumerate(grids):
        a, b = tmp_path / f"ours{i}.nvdb", tmp_path / f"ref{i}.nvdb"
        nvq.tools.write_nvdb(a, g)
        ref.write_nvdb(b, [g])
        assert a.read_bytes() == b.read_bytes()
        assert _same(nvq.tools.read_nvdb(a, 0), g)


def _name(g):
    return bytes(g[40:296]).split(b"\0")[0].decode()


@pytest.mark.parametrize("raw", [False, True])
def test_multi_grid_buffers_renumber_and_refresh_the_header_crc(ref, grids, tmp_path, raw):
    """A merged handle holds grid i of 4 with its header CRC; reading one grid renumbers it 0 of 1
    and refreshes that CRC.  For a segment of several grids the reference's by-index read
    io::readGrid(is, n) reads from the segment's first grid whatever n is (io/IO.h:529-540 adds the
    preceding grids to `seek` but reads before seeking), so by-index reads are compared there with
    the reference's by-name read, which seeks (io/IO.h:579-596)."""
    import openvdb_amd as nvq
    path = tmp_path / ("merged.raw" if raw else "merged.nvdb")
    ref.write_nvdb(path, grids, merge=True, raw=raw)
    names = [_name(g) for g in grids]
    assert len(set(names)) == len(names)
    for i in range(len(grids)):
        ours = nvq.tools.read_nvdb(path, i)
        by_name = ref.read_nvdb(path, name=names[i])
        assert _same(ours, by_name)
        assert _same(nvq.tools.read_nvdb(path, name=names[i]), by_name)
        if raw or i == 0:
            assert _same(ours, ref.read_nvdb(path, i))
        hdr = np.frombuffer(ours[24:32].tobytes(), np.uint32)
        assert hdr.tolist() == [0, 1]
        # everything past the renumbered header fields and the CRC is the grid as built
        assert _same(ours[32:], grids[i][32:])
    if not raw:  # the reference quirk described above
        wrong = ref.read_nvdb(path, 1)
        assert _name(wrong) == names[0]
