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


def test_cpp_header_compiles_and_links(tmp_path):
    """include/lvr2_amd/HipSurface.hpp against stand-ins with lvr2's layouts (boost::shared_array
    -like floatArr, LBPointArray, a 32-byte QueryPoint with a vtable), linked with liblvrhip.so."""
    import subprocess

    src = tmp_path / "use_hipsurface.cpp"
    src.write_text(r'''
#include "lvr2_amd/HipSurface.hpp"
#include <memory>
struct Vec { float x, y, z; };
struct QP { Vec m_position; float m_distance; bool m_invalid; virtual ~QP() {} };
struct FloatArr { std::shared_ptr<float> p; float* get() const { return p.get(); } };
template <class T> struct LBPointArray { unsigned int width = 0; unsigned int dim = 1; T* elements = nullptr; };
static_assert(sizeof(QP) == 32, "QueryPoint stand-in layout");
int main(int argc, char**) {
    if (argc < 5) return 0;  // compiled and linked only: running needs a GPU
    FloatArr pts{std::shared_ptr<float>(new float[30], std::default_delete<float[]>())};
    lvr2::HipSurface s(pts, 10);
    LBPointArray<float> lb; lb.width = 10; lb.dim = 3; lb.elements = pts.get();
    lvr2::HipSurface s2(lb);
    s.setKn(5); s.setKi(5); s.setKd(3); s.setFlippoint(0, 0, 0);
    std::string m = "PCA"; s.setMethod(m);
    s.calculateNormals(); s.interpolateNormals(); s.getNormals(pts);
    std::vector<QP> q(4); s.distances(q, 0.1f);
    return 0;
}
''')
    inc = os.path.join(ROOT, "include")
    libdir = os.path.join(ROOT, "lvr2_amd")
    exe = tmp_path / "use_hipsurface"
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I", inc, str(src), "-L", libdir, "-llvrhip",
                        f"-Wl,-rpath,{libdir}", "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
