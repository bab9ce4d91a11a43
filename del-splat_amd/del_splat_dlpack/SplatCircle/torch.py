"""Circle splatting, torch front end.

`from_project_spheres` keeps SplatCircle/torch.py:6-40 of the reference;
`tile2idx_idx2pnt` and `splat` mirror del-splat-cudarc/src/splat_circle.rs
(:3-94, :134-177), which the reference exposes only to Rust.
"""
import torch

from .. import _CapsuleAsDLPack, _native, util_torch
from .._native import call
from .._torch_util import bin_phases, ptr, require_gpu, stream_ptr, tile_shape


def from_project_spheres(pnt2xyz: torch.Tensor, radius: float, transform_world2ndc: torch.Tensor,
                         img_shape):
    """-> (pnt2pixxyndcz [N,3], pnt2pixrad [N]); transform_world2ndc (4,4) row-major."""
    device = pnt2xyz.device
    num_pnt = pnt2xyz.shape[0]
    assert transform_world2ndc.device == device
    assert transform_world2ndc.shape == (4, 4)
    assert transform_world2ndc.dtype == torch.float32
    pnt2pixxyndcz = torch.empty((num_pnt, 3), device=device, dtype=torch.float32)
    pnt2pixrad = torch.empty((num_pnt), device=device, dtype=torch.float32)
    transform_world2ndc = transform_world2ndc.transpose(0, 1).flatten().contiguous()
    sp = 0
    if device.type == "cuda":
        torch.cuda.set_device(device)
        sp = torch.cuda.current_stream(device).cuda_stream
    from .. import SplatCircle

    SplatCircle.from_project_spheres(
        util_torch.to_dlpack_safe(pnt2xyz.contiguous(), sp), radius,
        util_torch.to_dlpack_safe(transform_world2ndc, sp), img_shape,
        util_torch.to_dlpack_safe(pnt2pixxyndcz, sp), util_torch.to_dlpack_safe(pnt2pixrad, sp),
        sp)
    return pnt2pixxyndcz, pnt2pixrad


def tile2idx_idx2pnt(tile_shape_, tile_size, pnt2pixxydepth, pnt2pixrad):
    """splat_circle.rs:3-94 (tile size threaded through instead of the
    hard-coded 16) -> (tile2idx [T+1], idx2pnt [K]) int32 tensors of u32."""
    require_gpu(pnt2pixxydepth, "pnt2pixxydepth")
    dev = pnt2pixxydepth.device
    n = pnt2pixxydepth.shape[0]
    tw, th = tile_shape_
    pxd = pnt2pixxydepth.contiguous()
    rad = pnt2pixrad.contiguous()
    # boxes [x-r, y-r, x+r, y+r] counted on the device, K read back once

    def count(s, wsp, wsp_bytes, kbuf):
        call("dsplat_circle_bin_count", s, tw, th, tile_size, n, ptr(pxd), ptr(rad), ptr(wsp),
             wsp_bytes, ptr(kbuf))

    def fill(s, wsp, wsp_bytes, kbuf, t2i, i2p, k, wsq, wsq_bytes):
        call("dsplat_circle_bin_fill", s, tw, th, tile_size, n, ptr(pxd), ptr(rad), ptr(wsp),
             wsp_bytes, ptr(kbuf), ptr(t2i), ptr(i2p), k, ptr(wsq), wsq_bytes)

    return bin_phases(dev, n, tw * th, count, fill)


def splat(img_shape, pnt2pixxydepth, pnt2pixrad, pnt2rgb, tile_size, tile2idx, idx2pnt):
    """splat_circle.rs:134-177 -> pix2rgb [H,W,3] (uncovered pixels 0)."""
    require_gpu(pnt2pixxydepth, "pnt2pixxydepth")
    dev = pnt2pixxydepth.device
    w, h = img_shape
    img = torch.empty((h, w, 3), device=dev, dtype=torch.float32)
    i2p = idx2pnt if idx2pnt.numel() else torch.zeros(1, device=dev, dtype=torch.int32)
    call("dsplat_circle_splat", stream_ptr(dev), w, h, ptr(img), ptr(pnt2pixxydepth.contiguous()),
         ptr(pnt2pixrad.contiguous()), ptr(pnt2rgb.contiguous()), tile_size, ptr(tile2idx),
         ptr(i2p))
    return img


def rasterize_zbuffer(pnt2pixxyndcz, pnt2pixrad, pnt2rgb, img_shape):
    """Circle z-buffer: every pixel centre within the pixel radius of a point
    keeps the nearest point (min packed key) -> (pix2rgb [H,W,3],
    pix2unitdepth [H,W])."""
    require_gpu(pnt2pixxyndcz, "pnt2pixxyndcz")
    dev = pnt2pixxyndcz.device
    n = pnt2pixxyndcz.shape[0]
    w, h = img_shape
    img = torch.empty((h, w, 3), device=dev, dtype=torch.float32)
    dep = torch.empty((h, w), device=dev, dtype=torch.float32)
    keys = torch.empty(h * w, device=dev, dtype=torch.int64)
    call("dsplat_circle_rasterize_zbuffer", stream_ptr(dev), n, ptr(pnt2pixxyndcz.contiguous()),
         ptr(pnt2pixrad.contiguous()), ptr(pnt2rgb.contiguous()), ptr(dep), ptr(img), w, h,
         ptr(keys))
    return img, dep
