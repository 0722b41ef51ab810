"""Device-side signed flood fill (tools/cuda/SignedFloodFill.cuh signedFloodFill<float>): the
oracle's restatement (CPU only), pinned by the reference's own known answers
(CudaSignedFloodFill, NonBlockingStreamSignedFloodFill in unittest/TestNanoVDB.cu), by the host
build::Grid fill wherever the two algorithms must agree (leaves whose active voxels share one
sign), and by invariants (active values unchanged, inactive values +-background, no dependence on
the previous inactive values, idempotence)."""
import numpy as np
import pytest

from conftest import _u, damage_inactive, root_gap_grid
from pyoracle import have_reference

pytestmark = pytest.mark.skipif(not have_reference(), reason="reference sources not built")


def _leaf_value_index(g, ijk):
    """(byte offset of the value, active) of voxel ijk in a float grid buffer."""
    offs, cnt = _u(g, 672, np.int64, 4), _u(g, 704, np.uint32, 3)
    lf = 672 + int(offs[0])
    origins = np.frombuffer(g[lf:lf + 2144 * int(cnt[0])].tobytes(), np.uint8).reshape(-1, 2144)[:, :12].copy().view(np.int32) & ~7
    i = int(np.nonzero((origins == (np.array(ijk) & ~7)).all(1))[0][0])
    n = ((ijk[0] & 7) << 6) | ((ijk[1] & 7) << 3) | (ijk[2] & 7)
    base = lf + 2144 * i
    return base + 96 + 4 * n, bool((int(_u(g, base + 16 + 8 * (n >> 6), np.uint64)[0]) >> (n & 63)) & 1)


def kat_sphere(ref, radius):
    """createLevelSetSphere<float>(radius) with voxels (radius+3,0,0) and (radius-3,0,0) flipped
    to -1 and +1, as the reference's CudaSignedFloodFill test corrupts it."""
    g = ref.level_set_sphere(float(radius), (0, 0, 0), 1.0, 3.0)
    for ijk, v in (((radius + 3, 0, 0), -1.0), ((radius - 3, 0, 0), 1.0)):
        off, active = _leaf_value_index(g, ijk)
        assert not active
        g[off:off + 4] = np.frombuffer(np.float32(v).tobytes(), np.uint8)
    return g


def test_reference_kat_sphere_100(oracle, ref):
    g = kat_sphere(ref, 100)
    f = oracle.signed_flood_fill(g)
    for ijk, want in (((103, 0, 0), 3.0), ((100, 0, 0), 0.0), ((97, 0, 0), -3.0)):
        off, _ = _leaf_value_index(f, ijk)
        assert _u(f, off, np.float32)[0] == want


def _invariants(g, f):
    offs, cnt = _u(g, 672, np.int64, 4), _u(g, 704, np.uint32, 3)
    bg = float(_u(g, 672 + offs[3] + 28, np.float32)[0])
    lf = 672 + int(offs[0])
    G = np.frombuffer(g[lf:lf + 2144 * int(cnt[0])].tobytes(), np.uint8).reshape(-1, 2144)
    F = np.frombuffer(f[lf:lf + 2144 * int(cnt[0])].tobytes(), np.uint8).reshape(-1, 2144)
    mask = np.unpackbits(G[:, 16:80], axis=1, bitorder="little").astype(bool)
    gv, fv = G[:, 96:].copy().view(np.float32), F[:, 96:].copy().view(np.float32)
    assert np.array_equal(gv[mask].view(np.uint32), fv[mask].view(np.uint32))
    assert np.all(np.abs(fv[~mask]) == np.float32(bg))
    assert np.array_equal(G[:, :96], F[:, :96])  # headers, masks, statistics untouched
    return mask, gv, fv


@pytest.mark.parametrize("name", ["ref_sphere", "ref_torus", "our_torus"])
def test_invariants_and_host_agreement(oracle, ref, name):
    import openvdb_amd as nvq
    g = {"ref_sphere": lambda: ref.level_set_sphere(40.0, (1.5, -2.0, 3.0), 0.5, 3.0),
         "ref_torus": lambda: ref.level_set_torus(60.0, 15.0, (0, 0, 0), 1.0, 3.0),
         "our_torus": lambda: nvq.tools.create_level_set_torus(300.0, 80.0, (0, 0, 0), 1.0, 3.0)}[name]()
    f = oracle.signed_flood_fill(g)
    mask, gv, fv = _invariants(g, f)
    # the host build::Grid fill (which produced g) and the device fill agree on every leaf whose
    # active voxels share one sign
    neg = np.where(mask, gv < 0, False).any(1)
    pos = np.where(mask, gv >= 0, False).any(1)
    uniform = ~(neg & pos)
    assert uniform.sum() > 0
    assert np.array_equal(gv[uniform].view(np.uint32), fv[uniform].view(np.uint32))
    # only inactive values are rewritten, and they depend on nothing but the active ones
    np.testing.assert_array_equal(oracle.signed_flood_fill(damage_inactive(g)), f)
    np.testing.assert_array_equal(oracle.signed_flood_fill(damage_inactive(g, -7.0)), f)
    np.testing.assert_array_equal(oracle.signed_flood_fill(f), f)


def test_root_scanline_fill(oracle, ref):
    g = root_gap_grid(oracle)
    assert ref.validate(g, full=True)
    f = oracle.signed_flood_fill(g)
    root = 672 + int(_u(f, 696, np.int64)[0])
    tiles = {int(_u(f, root + 64 + 32 * t, np.uint64)[0]): t for t in range(int(_u(f, root + 24, np.uint32)[0]))}
    for z in (4096, 8192):
        t = tiles[z >> 12]
        assert _u(f, root + 64 + 32 * t + 20, np.float32)[0] == -3.0
        assert _u(f, root + 64 + 32 * t + 16, np.uint32)[0] == 0
    # every inactive leaf value became -background (all active values are negative)
    _, _, fv = _invariants(g, f)
    assert np.all(fv <= 0)


def test_root_missing_tile_is_an_error(oracle):
    with pytest.raises(ValueError):
        oracle.signed_flood_fill(root_gap_grid(oracle, missing=True))
