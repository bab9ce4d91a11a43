"""Point splatting, torch front end (SplatPoint/torch.py:5-64 of the reference).

Outputs are allocated with zeros for the background (the reference uses
torch.empty and leaves uncovered pixels uninitialised)."""
import torch

from .. import util_torch


def _stream(device):
    if device.type == "cuda":
        torch.cuda.set_device(device)
        return torch.cuda.current_stream(device).cuda_stream
    return 0


def rasterize_pix(pnt2pixxyndcz, pnt2rgb, img_shape):
    """-> pix2rgb [H,W,3]: nearest point per pixel, painter semantics."""
    device = pnt2pixxyndcz.device
    num_pnt = pnt2pixxyndcz.shape[0]
    assert pnt2rgb.shape == (num_pnt, 3)
    assert pnt2rgb.device == device
    assert pnt2rgb.dtype == torch.float32
    pix2rgb = torch.zeros((img_shape[1], img_shape[0], 3), device=device, dtype=torch.float32)
    sp = _stream(device)
    from .. import SplatPoint

    SplatPoint.rasterize_pix(util_torch.to_dlpack_safe(pnt2pixxyndcz.contiguous(), sp),
                             util_torch.to_dlpack_safe(pnt2rgb.contiguous(), sp),
                             util_torch.to_dlpack_safe(pix2rgb, sp), sp)
    return pix2rgb


def rasterize_pix_zbuffer(pnt2pixxyndcz, pnt2rgb, img_shape, return_depth=False):
    """-> pix2rgb [H,W,3] (and pix2unitdepth [H,W] if return_depth)."""
    device = pnt2pixxyndcz.device
    num_pnt = pnt2pixxyndcz.shape[0]
    assert pnt2rgb.shape == (num_pnt, 3)
    assert pnt2rgb.device == device
    assert pnt2rgb.dtype == torch.float32
    pix2rgb = torch.zeros((img_shape[1], img_shape[0], 3), device=device, dtype=torch.float32)
    pix2unitdepth = torch.empty((img_shape[1], img_shape[0]), device=device, dtype=torch.float32)
    sp = _stream(device)
    from .. import SplatPoint

    SplatPoint.rasterize_pix_zbuffer(util_torch.to_dlpack_safe(pnt2pixxyndcz.contiguous(), sp),
                                     util_torch.to_dlpack_safe(pnt2rgb.contiguous(), sp),
                                     util_torch.to_dlpack_safe(pix2rgb, sp),
                                     util_torch.to_dlpack_safe(pix2unitdepth, sp), sp)
    if return_depth:
        return pix2rgb, pix2unitdepth
    return pix2rgb
