"""The allocation-free binning entry point, dsplat_gauss_tile2idx_idx2pnt_noalloc
(a static device scratch instead of a workspace argument), against the
workspace entry point dsplat_gauss_tile2idx_idx2pnt_ws on the same inputs, both
through the C ABI: two calls in a row on one stream, the second on a smaller
input (the scratch then holds the first call's contents), lists and K
bit-exact; an overflowing capacity reports K as the _ws variant does; sizes
beyond the scratch are refused before any launch."""
import ctypes

import numpy as np
import pytest
import torch

from _helpers import small_scene, u32
from del_splat_dlpack import _native, scene

pytestmark = pytest.mark.gpu
DEV = "cuda"
SCRATCH = 128 << 20  # DSPLAT_NOALLOC_SCRATCH_BYTES


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def _frame(n, seed, shape, radius):
    import del_splat_dlpack.SplatGauss.torch as SG
    s, m = small_scene(n, seed, shape, radius=radius)
    pxd, _, aabb, _ = SG.project_fwd(*(_t(s[k]) for k in ("xyz", "log_scale", "quat", "opacity",
                                                           "rgb")), _t(m), shape)
    return pxd, aabb


def _bin(kind, shape, ts, pxd, aabb, capacity):
    w, h = shape
    tw, th = (w + ts - 1) // ts, (h + ts - 1) // ts
    n = pxd.shape[0]
    t2i = torch.full((tw * th + 1,), -1, device=DEV, dtype=torch.int32)
    i2p = torch.full((max(capacity, 1),), -1, device=DEV, dtype=torch.int32)
    k = torch.zeros(1, device=DEV, dtype=torch.int32)
    st = torch.cuda.current_stream().cuda_stream
    if kind == "ws":
        nb = _native.lib().dsplat_tile2idx_idx2pnt_workspace_size(n, tw * th, capacity)
        ws = torch.empty(nb, device=DEV, dtype=torch.uint8)
        _native.call("dsplat_gauss_tile2idx_idx2pnt_ws", st, w, h, ts, n, pxd.data_ptr(),
                     aabb.data_ptr(), t2i.data_ptr(), i2p.data_ptr(), capacity, ws.data_ptr(),
                     nb, k.data_ptr())
    else:
        _native.call("dsplat_gauss_tile2idx_idx2pnt_noalloc", st, w, h, ts, n, pxd.data_ptr(),
                     aabb.data_ptr(), t2i.data_ptr(), i2p.data_ptr(), capacity, k.data_ptr())
    torch.cuda.synchronize()
    return u32(t2i), u32(i2p), int(u32(k)[0])


def test_noalloc_matches_workspace_entry_point_two_calls():
    cases = [((1920, 1080), 32, 1_000_000, 2, 2.0),   # the headline frame
             ((320, 240), 16, 20_000, 5, 1.0)]        # then a smaller one
    for shape, ts, n, seed, radius in cases:
        print(f"noalloc: {n} Gaussians {shape[0]}x{shape[1]} ts={ts}", flush=True)
        pxd, aabb = _frame(n, seed, shape, radius)
        _, _, k = _bin("ws", shape, ts, pxd, aabb, 0)  # K only (capacity 0)
        cap = k + k // 8 + 1024
        lib = _native.lib()
        tw, th = (shape[0] + ts - 1) // ts, (shape[1] + ts - 1) // ts
        assert lib.dsplat_tile2idx_idx2pnt_workspace_size(n, tw * th, cap) <= SCRATCH
        ref = _bin("ws", shape, ts, pxd, aabb, cap)
        got = _bin("noalloc", shape, ts, pxd, aabb, cap)
        assert got[2] == ref[2] == k
        np.testing.assert_array_equal(got[0], ref[0])
        np.testing.assert_array_equal(got[1][:k], ref[1][:k])
        # an overflowing capacity: K reported, as the _ws variant
        small = _bin("noalloc", shape, ts, pxd, aabb, k // 2)
        assert small[2] == k


def test_noalloc_refuses_sizes_beyond_the_scratch():
    lib = _native.lib()
    n, ts, shape = 1_000_000, 16, (1920, 1080)
    tw, th = (shape[0] + ts - 1) // ts, (shape[1] + ts - 1) // ts
    cap = 10_000_000
    assert lib.dsplat_tile2idx_idx2pnt_workspace_size(n, tw * th, cap) > SCRATCH
    pxd = torch.zeros((n, 3), device=DEV)
    aabb = torch.zeros((n, 4), device=DEV)
    t2i = torch.zeros(tw * th + 1, device=DEV, dtype=torch.int32)
    i2p = torch.zeros(16, device=DEV, dtype=torch.int32)  # never written: refused first
    k = torch.full((1,), 7, device=DEV, dtype=torch.int32)
    rc = lib.dsplat_gauss_tile2idx_idx2pnt_noalloc(
        ctypes.c_void_p(torch.cuda.current_stream().cuda_stream), shape[0], shape[1], ts, n,
        pxd.data_ptr(), aabb.data_ptr(), t2i.data_ptr(), i2p.data_ptr(), cap, k.data_ptr())
    assert rc == _native.DSPLAT_ERR_CAPACITY
    assert b"static scratch" in lib.dsplat_last_error()
    torch.cuda.synchronize()
    assert int(k.item()) == 7 and not t2i.any()  # nothing ran
