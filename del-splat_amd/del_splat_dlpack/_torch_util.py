"""Small torch helpers shared by the torch front ends: device checks, the
current HIP stream, raw pointers and the world->ndc matrix layout."""
import torch

from . import _native


def stream_ptr(device):
    return torch.cuda.current_stream(device).cuda_stream


def ptr(t):
    return None if t is None else t.data_ptr()


def require_gpu(t, name):
    if t.device.type != "cuda":
        raise ValueError(f"{name} must be a GPU tensor (got {t.device}); the HIP library has no "
                         "CPU path")
    _native.lib()


def check_f32(t, shape, name):
    if t.dtype != torch.float32:
        raise ValueError(f"{name} must be float32 (got {t.dtype})")
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} must have shape {tuple(shape)} (got {tuple(t.shape)})")


def world2ndc_col_major(m, device):
    """(4,4) row-major -> flat 16 column-major (SplatCircle/torch.py:22); a
    flat (16,) input is taken as already column-major."""
    m = torch.as_tensor(m, dtype=torch.float32)
    if tuple(m.shape) == (4, 4):
        m = m.transpose(0, 1).reshape(16)
    elif tuple(m.shape) != (16,):
        raise ValueError(f"transform_world2ndc must be (4,4) or (16,), got {tuple(m.shape)}")
    return m.to(device).contiguous()


def tile_shape(img_shape, tile_size):
    w, h = img_shape
    return (w + tile_size - 1) // tile_size, (h + tile_size - 1) // tile_size
