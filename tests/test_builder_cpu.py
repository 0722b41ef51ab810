"""The product's host grid builder (libnvq.so, include/nvq_build.h) against the reference host
tools compiled from their sources: buffers must be byte-identical except the checksum field
(written disabled, all bits set) and, for FpN, the unused tail the reference appends."""
import numpy as np
import pytest

from conftest import rng


def _same(mine, theirs, tail=0):
    assert mine.nbytes == theirs.nbytes, (mine.nbytes, theirs.nbytes)
    end = mine.nbytes - tail
    diff = np.nonzero(mine[:end] != theirs[:end])[0]
    diff = diff[(diff < 8) | (diff >= 16)]  # bytes 8..15: GridData::mChecksum
    assert len(diff) == 0, f"{len(diff)} differing bytes, first at {diff[:8]}"
    assert np.all(mine[8:16] == 0xFF)


def _leaf_offset(buf):
    return 672 + int(np.frombuffer(buf[672:680].tobytes(), np.int64)[0])


@pytest.fixture(scope="module")
def nvq():
    import openvdb_amd
    return openvdb_amd


def test_level_set_sphere_c1(nvq, sphere):
    mine = nvq.tools.create_level_set_sphere(100.0, (0, 0, 0), 1.0, 3.0)
    _same(mine, sphere)
    assert mine.nbytes == 11064704


def test_level_set_torus_c2(nvq, torus):
    _same(nvq.tools.create_level_set_torus(300.0, 80.0, (0, 0, 0), 1.0, 3.0), torus)


@pytest.mark.parametrize("args", [(7.5, (0.3, -0.2, 0.1), 0.25, 2.0), (20.0, (-100.0, 50.0, 3000.0), 1.0, 3.0),
                                  (100.0, (0, 0, 0), 0.5, 5.0)])
def test_level_set_sphere_variants(nvq, ref, args):
    _same(nvq.tools.create_level_set_sphere(*args), ref.level_set_sphere(*args))


def test_fog_volume_sphere(nvq, fog_sphere):
    _same(nvq.tools.create_fog_volume_sphere(60.0, (3.0, -2.0, 5.0), 1.0, 3.0), fog_sphere)


@pytest.mark.parametrize("gt,kind", [("fp4", 1), ("fp8", 2), ("fp16", 3), ("fpn", 4)])
def test_quantised_codecs(nvq, ref, gt, kind):
    args = (40.0, (1.5, -2.0, 3.0), 0.5, 3.0)
    mine = nvq.tools.create_level_set_sphere(*args, grid_type=gt, tol=1e-3)
    theirs = ref.level_set_sphere(*args, kind=kind, tol=1e-3)
    tail = _leaf_offset(mine) if gt == "fpn" else 0
    _same(mine, theirs, tail)
    src = nvq.tools.create_level_set_sphere(*args)
    _same(nvq.tools.quantize(src, gt, 1e-3), ref.quantize(src, kind, 1e-3), tail)
    torus = nvq.tools.create_level_set_torus(60.0, 20.0, (0, 0, 0), 0.5, 3.0, grid_type=gt, tol=1e-3)
    tsrc = nvq.tools.create_level_set_torus(60.0, 20.0, (0, 0, 0), 0.5, 3.0)
    _same(torus, ref.quantize(tsrc, kind, 1e-3), _leaf_offset(torus) if gt == "fpn" else 0)


def test_fbm_fog_fill_matches_grid_builder(nvq, ref):
    """The fBm fog through build::Grid<float>::operator()(func, bbox) of the reference."""
    bbox = [-20, -5, 3, 60, 70, 50]
    r = [np.arange(bbox[a], bbox[3 + a] + 1) for a in range(3)]
    ijk = np.stack(np.meshgrid(*r, indexing="ij"), -1).reshape(-1, 3).astype(np.int32)
    for amp in (1.0, 2.5):
        dens = nvq.tools.fbm_density(ijk, amplitude=amp).reshape(len(r[0]), len(r[1]), len(r[2]))
        theirs = ref.grid_from_dense(dens, bbox, 0.0, 2)
        mine = nvq.tools.create_fog_fbm(bbox, amplitude=amp)
        theirs[40:40 + 256] = 0  # grid name
        mine = mine.copy()
        mine[40:40 + 256] = 0
        _same(mine, theirs)


def test_fbm_constant_regions_become_tiles(nvq, oracle):
    """amplitude large enough that the density saturates at 1: constant leaves are pruned to
    active lower-node tiles (tools/GridBuilder.h:1974-2018)."""
    buf = nvq.tools.create_fog_fbm([0, 0, 0, 127, 127, 127], amplitude=8.0)
    tiles = int(np.frombuffer(buf[672 + 44:672 + 48].tobytes(), np.uint32)[0])
    assert tiles > 0
    ijk = rng(1).integers(0, 128, size=(20000, 3)).astype(np.int32)
    val, act, dim = oracle.get_value(buf, ijk)
    tile = (dim == 8) & (act == 1)
    assert tile.sum() > 0 and np.all(val[tile] == 1.0)
    assert np.all(val[(dim == 8) & (act == 0)] == 0.0)
    expect = nvq.tools.fbm_density(ijk, amplitude=8.0)
    assert np.array_equal(val, expect)


def test_camera_rays_match_raygenop(nvq, ref):
    e, d = nvq.tools.camera_rays(512, 512, (0.0, 0.0, 400.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 45.0)
    re_, rd = ref.ray_gen(512, 512, 400.0)
    assert np.array_equal(e, re_)
    assert np.array_equal(d.view(np.uint32), rd.view(np.uint32))


def test_builder_rejects_bad_arguments(nvq):
    with pytest.raises(Exception):
        nvq.tools.create_level_set_sphere(-1.0)
    with pytest.raises(Exception):
        nvq.tools.create_level_set_torus(10.0, 20.0)
    with pytest.raises(Exception):
        nvq.tools.create_fog_fbm([0, 0, 0, -1, 5, 5])
