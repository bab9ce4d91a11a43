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


def bin_phases(device, n, num_tile, count, fill):
    """Binning in three phases with K read back between them (the reference
    reads K before allocating idx2pnt, splat_gauss.rs:137-147): `count(s,
    ws_points, ws_bytes, kbuf)` queues the per-Gaussian counts, K is copied
    to pinned host memory, the depth order is queued behind it, and only then
    does the host wait for K, allocate idx2pnt and queue `fill(s, ws_points,
    ws_bytes, kbuf, tile2idx, idx2pnt, k, ws_pairs, ws_pairs_bytes)`.
    Returns (tile2idx [T+1], idx2pnt [K]) int32 tensors holding u32."""
    L = _native.lib()
    s = stream_ptr(device)
    wsp_bytes = L.dsplat_bin_points_workspace_size(n)
    wsp = torch.empty(max(wsp_bytes, 1), device=device, dtype=torch.uint8)
    kbuf = torch.empty(4, device=device, dtype=torch.int32)
    count(s, wsp, wsp_bytes, kbuf)
    k_host = torch.empty(1, dtype=torch.int32, pin_memory=True)
    k_host.copy_(kbuf[:1], non_blocking=True)
    ready = torch.cuda.Event()
    ready.record(torch.cuda.current_stream(device))
    _native.call("dsplat_bin_order", s, n, ptr(wsp), wsp_bytes)
    ready.synchronize()
    k = int(k_host[0]) & 0xFFFFFFFF
    wsq_bytes = L.dsplat_bin_pairs_workspace_size(k)
    wsq = torch.empty(max(wsq_bytes, 1), device=device, dtype=torch.uint8)
    tile2idx = torch.empty(num_tile + 1, device=device, dtype=torch.int32)
    idx2pnt = torch.empty(max(k, 1), device=device, dtype=torch.int32)
    fill(s, wsp, wsp_bytes, kbuf, tile2idx, idx2pnt, k, wsq, wsq_bytes)
    return tile2idx, idx2pnt[:k]
