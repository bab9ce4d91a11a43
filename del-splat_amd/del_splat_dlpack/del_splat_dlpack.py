"""The low-level module of the reference package (the PyO3 cdylib
`del_splat_dlpack`, del-splat-dlpack/src/lib.rs:8-15), restated over the
HIP library's C ABI.

Every tensor argument is a DLPack capsule or an object with __dlpack__,
checked for dtype float32, shape and device as the reference checks them
(del_dlpack::check_*_tensor). Device types: GPU (kDLCUDA = 2 and kDLROCM = 10,
which is what torch on ROCm exports) run the HIP kernels on `stream_ptr`. CPU
tensors (kDLCPU = 1) are staged through the GPU and back -- the library has
no CPU implementation, so without a GPU they raise.
"""
import numpy as np
import torch

from . import _native, io_ply
from ._native import call


def _tensor(obj):
    if isinstance(obj, torch.Tensor):
        return obj
    if hasattr(obj, "__dlpack__"):
        return torch.from_dlpack(obj)
    return torch.from_dlpack(_Cap(obj))


class _Cap:
    def __init__(self, cap):
        self.cap = cap

    def __dlpack__(self, *, max_version=None, stream=None, dl_device=None, copy=None):
        return self.cap

    def __dlpack_device__(self):
        # read the device from the DLManagedTensor behind the capsule
        import ctypes
        get = ctypes.pythonapi.PyCapsule_GetPointer
        get.restype = ctypes.c_void_p
        get.argtypes = [ctypes.py_object, ctypes.c_char_p]
        p = get(self.cap, b"dltensor")
        # DLTensor: void* data; DLDevice {int32 type, int32 id}
        dev = (ctypes.c_int32 * 2).from_address(p + ctypes.sizeof(ctypes.c_void_p))
        return (dev[0], dev[1])


def _check(t, shape, name):
    if t.dtype != torch.float32:
        raise ValueError(f"{name}: dtype must be float32")
    if len(shape) != t.dim() or any(s != -1 and s != d for s, d in zip(shape, t.shape)):
        raise ValueError(f"{name}: shape {tuple(t.shape)} does not match {shape}")
    if not t.is_contiguous():
        raise ValueError(f"{name}: must be contiguous")


class _OnGpu:
    """Runs a body on GPU copies of CPU tensors and copies outputs back."""

    def __init__(self, tensors, outputs):
        self.tensors = tensors
        self.outputs = outputs

    def __enter__(self):
        dev0 = self.tensors[0].device
        devs = {t.device for t in self.tensors}
        if len(devs) != 1:
            raise ValueError("all tensors must be on the same device")
        self.cpu = dev0.type == "cpu"
        if self.cpu:
            if not torch.cuda.is_available():
                raise _native.DsplatError("no GPU: the HIP library has no CPU path")
            self.g = [t.cuda() for t in self.tensors]
        else:
            _native.lib()
            self.g = list(self.tensors)
        return self.g

    def __exit__(self, exc_type, exc, tb):
        if exc_type is None and self.cpu:
            for i in self.outputs:
                self.tensors[i].copy_(self.g[i].cpu())
        return False


def _stream(t, stream_ptr):
    if t.device.type == "cpu" or not stream_ptr:
        return torch.cuda.current_stream().cuda_stream
    return stream_ptr


def points_load_ply(path):
    """points.rs:10-21: (capsule xyz [N,3], capsule rgb [N,3]) on the CPU."""
    xyz, rgb = io_ply.read_xyzrgb(path)
    return (torch.from_numpy(xyz.reshape(-1, 3).copy()).__dlpack__(),
            torch.from_numpy(rgb.reshape(-1, 3).copy()).__dlpack__())


def point_splat_rasterize_pix(pnt2pixxyndcz, pnt2rgb, pix2rgb, stream_ptr=0):
    """splat_point.rs:11-43 (GPU branch todo!() in the reference): the point
    painter of pnt2pixxyndcz.rs:1-40 into pix2rgb [H,W,3]."""
    pxz, rgb, img = _tensor(pnt2pixxyndcz), _tensor(pnt2rgb), _tensor(pix2rgb)
    n = pxz.shape[0]
    _check(pxz, (n, 3), "pnt2pixxyndcz")
    _check(rgb, (n, 3), "pnt2rgb")
    _check(img, (-1, -1, 3), "pix2rgb")
    h, w = img.shape[0], img.shape[1]
    with _OnGpu([pxz, rgb, img], [2]) as (g_pxz, g_rgb, g_img):
        keys = torch.empty(h * w, device=g_img.device, dtype=torch.int64)
        call("dsplat_point_rasterize_pix", _stream(pxz, stream_ptr), n, g_pxz.data_ptr(),
             g_rgb.data_ptr(), g_img.data_ptr(), w, h, keys.data_ptr())


def point_splat_rasterize_pix_zbuffer(pnt2pixxyndcz, pnt2rgb, pix2rgb, pix2unitdepth,
                                      stream_ptr=0):
    """splat_point.rs:46-106 with the deterministic packed-key z-buffer."""
    pxz, rgb, img, dep = (_tensor(a) for a in (pnt2pixxyndcz, pnt2rgb, pix2rgb, pix2unitdepth))
    n = pxz.shape[0]
    _check(pxz, (n, 3), "pnt2pixxyndcz")
    _check(rgb, (n, 3), "pnt2rgb")
    _check(img, (-1, -1, 3), "pix2rgb")
    h, w = img.shape[0], img.shape[1]
    _check(dep, (h, w), "pix2unitdepth")
    with _OnGpu([pxz, rgb, img, dep], [2, 3]) as (g_pxz, g_rgb, g_img, g_dep):
        keys = torch.empty(h * w, device=g_img.device, dtype=torch.int64)
        call("dsplat_point_rasterize_zbuffer", _stream(pxz, stream_ptr), n, g_pxz.data_ptr(),
             g_rgb.data_ptr(), g_dep.data_ptr(), g_img.data_ptr(), w, h, keys.data_ptr())


def circle_splat_from_project_spheres(pnt2xyz, radius, transform_world2ndc, img_shape,
                                      pnt2pixxyndcz, pnt2pixrad, stream_ptr=0):
    """splat_circle.rs:10-75: sphere centres -> pixel (x, y, ndc z) and pixel
    radius. transform_world2ndc: column-major flat 16."""
    xyz, m, pxz, rad = (_tensor(a) for a in (pnt2xyz, transform_world2ndc, pnt2pixxyndcz,
                                             pnt2pixrad))
    n = xyz.shape[0]
    _check(xyz, (n, 3), "pnt2xyz")
    _check(m, (16,), "transform_world2ndc")
    _check(pxz, (n, 3), "pnt2pixxyndcz")
    _check(rad, (n,), "pnt2pixrad")
    with _OnGpu([xyz, m, pxz, rad], [2, 3]) as (g_xyz, g_m, g_pxz, g_rad):
        call("dsplat_circle_pnt2splat3_to_pnt2splat2", _stream(xyz, stream_ptr), n,
             g_xyz.data_ptr(), g_pxz.data_ptr(), g_rad.data_ptr(), g_m.data_ptr(),
             int(img_shape[0]), int(img_shape[1]), float(radius))


def load_3d_gauss_splat(path):
    """io_ply.rs:128-227 as numpy arrays (xyz, rgb, sh, opacity, scale, quat)."""
    return io_ply.read_3d_gauss_splat(path)


__all__ = ["points_load_ply", "point_splat_rasterize_pix", "point_splat_rasterize_pix_zbuffer",
           "circle_splat_from_project_spheres", "load_3d_gauss_splat", "np"]
