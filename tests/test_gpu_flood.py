"""Device-side signed flood fill (nvq_signed_flood_fill = cuda::signedFloodFill<float>) on the GPU:
the buffer after the fill equals the oracle's byte for byte, and the reference's own known answers
(CudaSignedFloodFill in unittest/TestNanoVDB.cu) hold."""
import numpy as np
import pytest

from conftest import _u, damage_inactive, root_gap_grid
from test_flood_cpu import _leaf_value_index, kat_sphere

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _nvq():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import openvdb_amd
    return openvdb_amd


def _fill_on_gpu(nvq, g):
    dg = nvq.DeviceGrid(g)
    dg.signed_flood_fill()
    return dg.download()


@pytest.mark.parametrize("name", ["ref_sphere", "our_torus", "root_gap"])
def test_fill_matches_oracle(oracle, ref, name):
    nvq = _nvq()
    g = {"ref_sphere": lambda: damage_inactive(ref.level_set_sphere(40.0, (1.5, -2.0, 3.0), 0.5, 3.0)),
         "our_torus": lambda: damage_inactive(nvq.tools.create_level_set_torus(300.0, 80.0, (0, 0, 0), 1.0, 3.0), -5.0),
         "root_gap": lambda: root_gap_grid(oracle)}[name]()
    print(f"[signed_flood_fill] grid={name} bytes={g.nbytes}", flush=True)
    got = _fill_on_gpu(nvq, g)
    want = oracle.signed_flood_fill(g)
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, f"{bad.size} bytes differ, first at {bad[:8]}"


@pytest.mark.parametrize("radius", [100, 1000])
def test_reference_kat(oracle, ref, radius):
    nvq = _nvq()
    g = kat_sphere(ref, radius)
    print(f"[signed_flood_fill] KAT sphere radius={radius} bytes={g.nbytes}", flush=True)
    got = _fill_on_gpu(nvq, g)
    for ijk, want in (((radius + 3, 0, 0), 3.0), ((radius, 0, 0), 0.0), ((radius - 3, 0, 0), -3.0)):
        off, _ = _leaf_value_index(got, ijk)
        assert _u(got, off, np.float32)[0] == want
    if radius == 100:
        np.testing.assert_array_equal(got, oracle.signed_flood_fill(g))


def test_missing_root_tile_is_an_error(oracle):
    nvq = _nvq()
    dg = nvq.DeviceGrid(root_gap_grid(oracle, missing=True))
    with pytest.raises(nvq.NvqError):
        dg.signed_flood_fill()
