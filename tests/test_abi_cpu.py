"""CPU checks of the C-ABI library: it loads, exports every declared symbol, and rejects
bad buffers before touching any device (the failure-detection half of GridData::isValid,
NanoVDB.h:2002-2013)."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT


def _declared():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            text = open(os.path.join(ROOT, "include", h)).read()
            names |= set(re.findall(r"^\s*(?:const\s+)?[\w\s\*]*?\b(nvq\w*)\s*\(", text, re.M))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    import openvdb_amd
    lib = openvdb_amd.lib()
    names = _declared()
    assert "nvq_grid_upload" in names and "nvq_sample" in names
    for n in names:
        assert hasattr(lib, n), n


def test_status_strings():
    import openvdb_amd
    lib = openvdb_amd.lib()
    for s in range(-8, 1):
        assert lib.nvq_status_string(s)


def _upload(buf, nbytes=None):
    import openvdb_amd
    lib = openvdb_amd.lib()
    h = ctypes.c_void_p()
    st = lib.nvq_grid_upload(buf.ctypes.data_as(ctypes.c_void_p), buf.nbytes if nbytes is None else nbytes, None,
                             ctypes.byref(h))
    return st, h


def test_upload_rejects_bad_buffers(ref):
    from openvdb_amd import aligned_buffer
    good = ref.level_set_sphere(20.0)
    bad = aligned_buffer(good.copy())
    bad[0:8] = np.frombuffer(b"NotAGrid", np.uint8)
    assert _upload(bad)[0] == -2
    bad = aligned_buffer(good.copy())
    bad[16:20] = np.frombuffer(np.uint32(31 << 21).tobytes(), np.uint8)
    assert _upload(bad)[0] == -3
    bad = aligned_buffer(good.copy())
    bad[636:640] = np.frombuffer(np.uint32(2).tobytes(), np.uint8)  # Double
    assert _upload(bad)[0] == -4
    raw = np.zeros(good.nbytes + 64, np.uint8)
    off = ((-raw.ctypes.data) % 32) + 4
    mis = raw[off:off + good.nbytes]
    mis[:] = good
    assert _upload(mis)[0] == -5
    assert _upload(good, nbytes=good.nbytes // 2)[0] == -6
    assert _upload(good, nbytes=100)[0] == -6
    import openvdb_amd
    lib = openvdb_amd.lib()
    assert lib.nvq_grid_upload(None, 0, None, None) == -1
    # queries on a NULL grid are rejected without a launch
    assert lib.nvq_sample(None, None, 0, 0, None, None, None, 1, None) == -1
    assert lib.nvq_get_value(None, None, 1, None, None, None, None) == -1
