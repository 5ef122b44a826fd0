"""AddressSanitizer + UndefinedBehaviorSanitizer builds of host code, run on the CPU: the oracle
(every entry point, tests/cpp/oracle_sanitize.cpp) and the PLY reader/writer
(include/lvr2_amd/PlyIO.hpp) on well-formed and malformed files.  A finding aborts the program,
which fails the test."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1",
           OMP_NUM_THREADS="4")


def _build(tmp_path, name, sources, extra=()):
    exe = tmp_path / name
    r = subprocess.run(["g++", "-std=c++17", *SAN, *extra, *sources, "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return str(exe)


def test_oracle_under_sanitizers(tmp_path):
    exe = _build(tmp_path, "oracle_sanitize",
                 [os.path.join(ROOT, "tests", "cpp", "oracle_sanitize.cpp"), os.path.join(ROOT, "oracle", "lvr2_oracle.cpp")],
                 ["-fopenmp", "-ffp-contract=off", "-I", os.path.join(ROOT, "oracle")])
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=ENV)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
    assert "oracle sanitize run: ok" in r.stdout
    assert "runtime error" not in r.stderr


PLY_PROG = r'''
#include "lvr2_amd/PlyIO.hpp"
#include <cstdio>
int main(int argc, char** argv) {
    int ok = 0, refused = 0;
    for (int i = 1; i < argc; i++) {
        try {
            auto c = lvr2::hip_ply::readPly(argv[i]);
            std::string out = std::string(argv[i]) + ".out";
            lvr2::hip_ply::writePly(out, c.points.data(), c.size(), c.normals.empty() ? nullptr : c.normals.data(),
                                    c.faces.empty() ? nullptr : c.faces.data(), c.faces.size() / 3);
            lvr2::hip_ply::readPly(out);
            ok++;
        } catch (const std::exception& e) {
            refused++;
        }
    }
    std::printf("read %d refused %d\n", ok, refused);
    return 0;
}
'''


def test_ply_reader_under_sanitizers(tmp_path):
    src = tmp_path / "ply_san.cpp"
    src.write_text(PLY_PROG)
    exe = _build(tmp_path, "ply_san", [str(src)], ["-I", os.path.join(ROOT, "include")])
    rng = np.random.default_rng(0)
    good, bad = [], []
    pts = rng.normal(size=(100, 3)).astype("<f4")
    head = (b"ply\nformat binary_little_endian 1.0\nelement vertex 100\nproperty float x\nproperty float y\n"
            b"property float z\nelement face 2\nproperty list uchar int vertex_indices\nend_header\n")
    faces = b"\x03" + np.array([0, 1, 2], "<i4").tobytes() + b"\x04" + np.array([3, 4, 5, 6], "<i4").tobytes()
    body = head + pts.tobytes() + faces
    good.append(body)
    good.append(b"ply\nformat ascii 1.0\nelement vertex 2\nproperty double x\nproperty double y\nproperty double z\n"
                b"end_header\n1 2 3\n4 5 6")
    # malformed: every truncation of the binary file, header lies, absurd counts and lists
    for cut in range(0, len(body), 7):
        bad.append(body[:cut])
    bad.append(head.replace(b"vertex 100", b"vertex 99999999999"))
    bad.append(head.replace(b"face 2", b"face 4000000000") + pts.tobytes())
    bad.append(head + pts.tobytes() + b"\xff" + b"\x00" * 16)
    bad.append(b"ply\nformat binary_little_endian 1.0\nelement vertex 1\nproperty list int float x\nproperty float y\n"
               b"property float z\nend_header\n" + np.array([2000000000], "<i4").tobytes() + b"\x00" * 8)
    bad.append(b"ply\nformat ascii 1.0\nelement vertex 3\nproperty float x\nproperty float y\nproperty float z\n"
               b"end_header\n1 2 3\n4 5")
    bad.append(b"ply\nelement vertex 1\nend_header\n")
    paths = []
    for i, b in enumerate(good + bad):
        p = tmp_path / f"f{i}.ply"
        p.write_bytes(b)
        paths.append(str(p))
    r = subprocess.run([exe, *paths], capture_output=True, text=True, timeout=300, env=ENV)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "runtime error" not in r.stderr
    n_ok, n_ref = map(int, r.stdout.split()[1::2])
    # every well-formed file reads; the malformed ones are refused (a truncation that ends on a
    # row boundary of an element of the right size can still be a valid file)
    assert n_ok >= len(good) and n_ref >= len(bad) - 3, r.stdout
