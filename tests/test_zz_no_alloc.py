"""The no-allocation signed distances (lvrhip_distances_device_noalloc, through the C ABI) against
the allocating entry point lvrhip_distances_device on the same handle: proj, eucl and invalid
bit-identical on two calls in a row, the second with fewer queries, while every device allocation
of the library is refused (LVRHIP_FAIL_ALLOC_AFTER=0); sizes beyond the maximum are rejected."""
import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(lib, h, q, voxel, outs):
    os.environ["LVRHIP_FAIL_ALLOC_AFTER"] = "0"
    try:
        return lib.lvrhip_distances_device_noalloc(h, C.c_void_p(q.data_ptr()), int(q.shape[0]), float(voxel),
                                                   C.c_void_p(outs[0].data_ptr()), C.c_void_p(outs[1].data_ptr()),
                                                   C.c_void_p(outs[2].data_ptr()))
    finally:
        del os.environ["LVRHIP_FAIL_ALLOC_AFTER"]


def test_noalloc_distances_match_the_allocating_entry_point():
    import torch

    from lvr2_amd import HipSurface, _lib, scenes

    lib = _lib.lib()
    w = scenes.room(1600, 625)  # 1,000,000 rays
    s = HipSurface(w.points)
    s.setKn(w.kn)
    s.setKi(w.ki)
    s.setKd(w.kd)
    s.setFlippoint(*w.flip)
    s.calculateNormals()
    s.interpolateNormals()
    v, _, _ = s.grid(w.voxel, True, device=True)
    rng = np.random.default_rng(1)
    perm = torch.from_numpy(rng.permutation(v.shape[0])[: v.shape[0] // 3]).cuda()
    calls = [v, v[perm].contiguous()]  # the lattice order, then a third of it in random order
    for c, q in enumerate(calls):
        m = int(q.shape[0])
        print(f"call {c}: {m} queries", flush=True)
        torch.cuda.synchronize()
        ref = s.distances(q, w.voxel)
        torch.cuda.synchronize()
        outs = (torch.full((m,), float("nan"), dtype=torch.float32, device="cuda"),
                torch.full((m,), float("nan"), dtype=torch.float32, device="cuda"),
                torch.full((m,), 7, dtype=torch.uint8, device="cuda"))
        torch.cuda.synchronize()
        rc = _noalloc(lib, s.handle, q, w.voxel, outs)
        assert rc == 0, lib.lvrhip_last_error()
        torch.cuda.synchronize()
        for a, b, name in zip(outs, ref, ("proj", "eucl", "invalid")):
            np.testing.assert_array_equal(a.cpu().numpy(), b.cpu().numpy(), err_msg=f"call {c}: {name}")
    # beyond the maximum: rejected before any work
    big = 1048577
    rc = lib.lvrhip_distances_device_noalloc(s.handle, C.c_void_p(calls[0].data_ptr()), big, float(w.voxel),
                                             C.c_void_p(outs[0].data_ptr()), None, None)
    assert rc == 1 and b"at most 1048576 queries" in lib.lvrhip_last_error()
    # kd above 32: rejected too
    s.setKd(33)
    rc = _noalloc(lib, s.handle, calls[1], w.voxel, outs)
    assert rc == 1 and b"kd <= 32" in lib.lvrhip_last_error()
