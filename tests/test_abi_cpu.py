"""The C ABI: liblvrhip.so loads and exports every entry point include/lvrhip.h declares;
argument checking needs no GPU."""
import ctypes as C
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "lvrhip.h")


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(lvrhip_\w+)\s*\(", txt, re.M)))


def test_header_lists_expected_entry_points():
    names = declared()
    for must in ("lvrhip_create", "lvrhip_calculate_normals", "lvrhip_interpolate_normals", "lvrhip_distances",
                 "lvrhip_knn", "lvrhip_get_normals", "lvrhip_destroy", "lvrhip_set_k", "lvrhip_set_flip_point"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from lvr2_amd import _lib

    lib = _lib.lib()
    for name in declared():
        assert hasattr(lib, name), name
    assert sorted(_lib.EXPORTED) == declared()


def test_null_handle_is_rejected_without_gpu():
    from lvr2_amd import _lib

    lib = _lib.lib()
    assert lib.lvrhip_set_k(None, 10, 10, 5) == 1
    assert b"NULL handle" in lib.lvrhip_last_error()
    assert lib.lvrhip_set_method(None, 3) == 1
    assert lib.lvrhip_calculate_normals(None) == 1
    out = C.c_void_p()
    assert lib.lvrhip_create(None, 10, 0, C.byref(out)) == 1


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_create_without_gpu_raises_cleanly():
    import numpy as np

    from lvr2_amd import HipSurface, LvrHipError

    with pytest.raises(LvrHipError):
        HipSurface(np.zeros((4, 3), np.float32))


def test_scan_table_is_checked_without_gpu():
    import numpy as np

    from lvr2_amd import _lib

    lib = _lib.lib()
    out = C.c_void_p()
    pts = np.zeros((10, 3), np.float32)
    bad = [np.array([1, 10], np.int64),          # does not start at 0
           np.array([0, 6, 4, 10], np.int64)]    # decreasing
    for off in bad:
        rc = lib.lvrhip_create_scans_device(pts.ctypes.data_as(C.c_void_p), off.ctypes.data_as(C.c_void_p),
                                            len(off) - 1, 0, None, C.byref(out))
        assert rc == 1, lib.lvrhip_last_error()
    off = np.array([0, 10], np.int64)
    assert lib.lvrhip_create_scans_device(pts.ctypes.data_as(C.c_void_p), off.ctypes.data_as(C.c_void_p), 0, 0,
                                          None, C.byref(out)) == 1
    assert lib.lvrhip_graph_launch(None) == 1
    assert lib.lvrhip_autotune(None, None, None, None) == 1
