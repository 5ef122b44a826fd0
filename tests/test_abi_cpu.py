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
