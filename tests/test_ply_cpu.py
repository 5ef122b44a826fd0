"""include/lvr2_amd/PlyIO.hpp on the CPU: a round trip through a small program built with g++.

Pins: every PLY storage format and scalar type the reader accepts gives back the float32 values
numpy wrote (doubles rounded to float32), the points come from "vertex" (or "point" when no
vertex element has x, y, z), normals only when nx, ny, nz are all present, polygons become fans
of triangles, and the writer's binary output is what numpy reads back."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

PROG = r'''
#include "lvr2_amd/PlyIO.hpp"
#include <cstdio>
int main(int argc, char** argv) {
    if (argc != 3) return 2;
    try {
        auto c = lvr2::hip_ply::readPly(argv[1]);
        lvr2::hip_ply::writePly(argv[2], c.points.data(), c.size(), c.normals.empty() ? nullptr : c.normals.data(),
                                c.faces.empty() ? nullptr : c.faces.data(), c.faces.size() / 3);
        std::printf("%zu %zu %zu\n", c.size(), c.normals.size() / 3, c.faces.size() / 3);
    } catch (const std::exception& e) {
        std::printf("error: %s\n", e.what());
        return 3;
    }
    return 0;
}
'''


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    d = tmp_path_factory.mktemp("ply")
    src = d / "ply_roundtrip.cpp"
    src.write_text(PROG)
    out = d / "ply_roundtrip"
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
                        str(src), "-o", str(out)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return str(out)


def _header(fmt, elements):
    h = f"ply\nformat {fmt} 1.0\ncomment made by a test\n"
    for name, n, props in elements:
        h += f"element {name} {n}\n"
        for p in props:
            h += f"property {p}\n"
    return (h + "end_header\n").encode()


def _read_back(path):
    with open(path, "rb") as f:
        data = f.read()
    end = data.index(b"end_header\n") + len(b"end_header\n")
    head = data[:end].decode().split("\n")
    n = int([ln for ln in head if ln.startswith("element vertex")][0].split()[2])
    nn = 6 if "property float nx" in head else 3
    v = np.frombuffer(data[end:end + 4 * nn * n], "<f4").reshape(n, nn)
    faces = []
    fl = [ln for ln in head if ln.startswith("element face")]
    if fl:
        nf = int(fl[0].split()[2])
        rec = np.frombuffer(data[end + 4 * nn * n:], dtype=[("c", "u1"), ("i", "<i4", 3)], count=nf)
        assert np.all(rec["c"] == 3)
        faces = rec["i"]
    return v, np.asarray(faces)


def _run(exe, tmp_path, body):
    src = tmp_path / "in.ply"
    src.write_bytes(body)
    dst = tmp_path / "out.ply"
    r = subprocess.run([exe, str(src), str(dst)], capture_output=True, text=True)
    return r, dst


@pytest.mark.parametrize("fmt,order", [("binary_little_endian", "<"), ("binary_big_endian", ">")])
@pytest.mark.parametrize("t,np_t", [("float", "f4"), ("double", "f8"), ("int", "i4"), ("short", "i2"),
                                    ("uchar", "u1"), ("float32", "f4"), ("uint16", "u2")])
def test_binary_vertex_types(exe, tmp_path, fmt, order, t, np_t):
    rng = np.random.default_rng(1)
    n = 257
    xyz = rng.uniform(0, 200, size=(n, 3)).astype(np_t)
    rec = np.zeros(n, dtype=[("a", order + "u2"), ("xyz", order + np_t, 3), ("nrm", order + "f4", 3)])
    rec["xyz"] = xyz
    rec["nrm"] = rng.normal(size=(n, 3))
    body = _header(fmt, [("vertex", n, ["ushort tag", f"{t} x", f"{t} y", f"{t} z", "float nx", "float ny",
                                        "float nz"])]) + rec.tobytes()
    r, dst = _run(exe, tmp_path, body)
    assert r.returncode == 0, r.stdout
    assert r.stdout.split() == [str(n), str(n), "0"]
    v, _ = _read_back(dst)
    np.testing.assert_array_equal(v[:, :3], xyz.astype(np.float64).astype(np.float32))
    np.testing.assert_array_equal(v[:, 3:], rec["nrm"].astype(np.float32))


def test_ascii_points_faces_and_partial_normals(exe, tmp_path):
    pts = np.array([[0.1, 0.2, 0.3], [1e-7, -2.5, 3.0], [4.0, 5.0, 6.0], [7.25, 8.5, 9.75]])
    lines = "".join(f"{float(p[0])!r} {float(p[1])!r} {float(p[2])!r} 0.5\n" for p in pts)
    faces = "3 0 1 2\n4 0 1 2 3\n"  # a triangle and a quad (two fan triangles)
    body = _header("ascii", [("vertex", 4, ["double x", "double y", "double z", "float nx"]),
                             ("face", 2, ["list uchar int vertex_indices"])]) + (lines + faces).encode()
    r, dst = _run(exe, tmp_path, body)
    assert r.returncode == 0, r.stdout
    assert r.stdout.split() == ["4", "0", "3"]  # nx alone is not a normal
    v, f = _read_back(dst)
    np.testing.assert_array_equal(v, pts.astype(np.float32))
    np.testing.assert_array_equal(f, [[0, 1, 2], [0, 1, 2], [0, 2, 3]])


def test_point_element_when_no_vertex_has_xyz(exe, tmp_path):
    rec = np.arange(12, dtype="<f4").reshape(4, 3)
    body = _header("binary_little_endian", [("vertex", 0, ["float x", "float y", "float z"]),
                                            ("point", 4, ["float x", "float y", "float z"])]) + rec.tobytes()
    r, dst = _run(exe, tmp_path, body)
    assert r.returncode == 0, r.stdout
    v, _ = _read_back(dst)
    # an empty vertex element with x, y, z is still the vertex element
    assert r.stdout.split() == ["0", "0", "0"] and len(v) == 0
    body = _header("binary_little_endian", [("point", 4, ["float x", "float y", "float z"])]) + rec.tobytes()
    r, dst = _run(exe, tmp_path, body)
    v, _ = _read_back(dst)
    np.testing.assert_array_equal(v, rec)


@pytest.mark.parametrize("body,msg", [
    (b"plx\n", "not a PLY file"),
    (b"ply\nformat binary_little_endian 1.0\nelement vertex 2\nproperty float x\nproperty float y\n"
     b"property float z\nend_header\n" + b"\0" * 20, "truncated"),
    (b"ply\nformat binary_little_endian 1.0\nelement vertex 2\nproperty float x\nend_header\n", "no vertex or point"),
    (b"ply\nformat weird 1.0\nend_header\n", "unknown format"),
    (b"ply\nformat ascii 1.0\nelement vertex 1\nproperty quad x\nend_header\n", "unknown property type"),
])
def test_errors(exe, tmp_path, body, msg):
    r, _ = _run(exe, tmp_path, body)
    assert r.returncode == 3 and msg in r.stdout
