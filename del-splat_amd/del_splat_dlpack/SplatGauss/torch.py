"""Gaussian splatting on MI355X, torch front end.

The stage functions mirror del-splat-cudarc's Gaussian host API
(del-splat-cudarc/src/splat_gauss.rs): `pnt2splat3_to_pnt2splat2` (:6-44),
`tile2idx_idx2pnt` (:96-195), `rasterize_pnt2splat2` (:46-94). `render` is
the differentiable renderer (no counterpart in the reference): a
torch.autograd.Function whose backward is the hand-written adjoint kernels.
All tensors live on the GPU; torch only provides memory and the stream.
"""
import torch

from .. import _native
from .._native import call
from .._torch_util import (bin_phases, check_f32, ptr, require_gpu, stream_ptr, tile_shape,
                           world2ndc_col_major)

DEFAULT_TILE = 16


def pnt2splat3_to_pnt2splat2(pnt2xyz, pnt2scale, pnt2quat, transform_world2ndc, img_shape):
    """Projection (splat_gauss.rs:6-44): returns (pnt2pixxydepth [N,3],
    pnt2pixconvinv [N,3], pnt2pixaabb [N,4]). pnt2scale holds sigma."""
    require_gpu(pnt2xyz, "pnt2xyz")
    dev = pnt2xyz.device
    n = pnt2xyz.shape[0]
    check_f32(pnt2xyz, (n, 3), "pnt2xyz")
    check_f32(pnt2scale, (n, 3), "pnt2scale")
    check_f32(pnt2quat, (n, 4), "pnt2quat")
    m = world2ndc_col_major(transform_world2ndc, dev)
    pxd = torch.empty((n, 3), device=dev, dtype=torch.float32)
    conic = torch.empty((n, 3), device=dev, dtype=torch.float32)
    aabb = torch.empty((n, 4), device=dev, dtype=torch.float32)
    call("dsplat_gauss_pnt2splat3_to_pnt2splat2", stream_ptr(dev), n,
         ptr(pnt2xyz.contiguous()), ptr(pnt2scale.contiguous()), ptr(pnt2quat.contiguous()),
         ptr(pxd), ptr(conic), ptr(aabb), ptr(m), img_shape[0], img_shape[1])
    return pxd, conic, aabb


def tile2idx_idx2pnt(img_shape, tile_size, pnt2pixxydepth, pnt2pixaabb):
    """Binning + stable sort (splat_gauss.rs:96-195): returns (tile2idx
    u32[T+1], idx2pnt u32[K]) as int32-typed tensors holding u32 values.
    Reads K back once, as the reference does; the depth order runs on the
    GPU meanwhile."""
    require_gpu(pnt2pixaabb, "pnt2pixaabb")
    dev = pnt2pixaabb.device
    n = pnt2pixaabb.shape[0]
    w, h = img_shape
    tw, th = tile_shape(img_shape, tile_size)
    pxd = pnt2pixxydepth.contiguous()
    bb = pnt2pixaabb.contiguous()

    def count(s, wsp, wsp_bytes, kbuf):
        call("dsplat_gauss_bin_count", s, w, h, tile_size, n, ptr(pxd), ptr(bb), ptr(wsp),
             wsp_bytes, ptr(kbuf))

    def fill(s, wsp, wsp_bytes, kbuf, t2i, i2p, k, wsq, wsq_bytes):
        call("dsplat_gauss_bin_fill", s, w, h, tile_size, n, ptr(bb), ptr(wsp), wsp_bytes,
             ptr(kbuf), ptr(t2i), ptr(i2p), k, ptr(wsq), wsq_bytes)

    return bin_phases(dev, n, tw * th, count, fill)


def bin_gauss_capacity(img_shape, tile_size, pnt2pixxydepth, pnt2pixaabb, capacity,
                       d_num_idx=None):
    """Binning + stable sort with no host synchronisation
    (dsplat_gauss_tile2idx_idx2pnt_ws): returns (tile2idx [T+1], idx2pnt
    [capacity], d_num_idx), where d_num_idx[0] is K on the device. When
    K <= capacity the lists equal tile2idx_idx2pnt's (idx2pnt[:K]); when
    K > capacity every list is empty and K tells the capacity to use."""
    require_gpu(pnt2pixaabb, "pnt2pixaabb")
    dev = pnt2pixaabb.device
    n = pnt2pixaabb.shape[0]
    pxd = pnt2pixxydepth.contiguous()
    bb = pnt2pixaabb.contiguous()
    tw, th = tile_shape(img_shape, tile_size)
    ws_bytes = _native.lib().dsplat_tile2idx_idx2pnt_workspace_size(n, tw * th, capacity)
    ws = torch.empty(ws_bytes, device=dev, dtype=torch.uint8)
    tile2idx = torch.empty(tw * th + 1, device=dev, dtype=torch.int32)
    idx2pnt = torch.empty(max(capacity, 1), device=dev, dtype=torch.int32)
    if d_num_idx is None:
        d_num_idx = torch.empty(1, device=dev, dtype=torch.int32)
    call("dsplat_gauss_tile2idx_idx2pnt_ws", stream_ptr(dev), img_shape[0], img_shape[1],
         tile_size, n, ptr(pxd), ptr(bb), ptr(tile2idx), ptr(idx2pnt), capacity, ptr(ws),
         ws_bytes, ptr(d_num_idx))
    return tile2idx, idx2pnt[:capacity], d_num_idx



def rasterize_pnt2splat2(img_shape, pnt2pixxydepth, pnt2pixconvinv, pnt2pixaabb, pnt2opacity,
                         pnt2rgb, tile_size, tile2idx, idx2pnt):
    """Tiled front-to-back raster (splat_gauss.rs:46-94) -> pix2rgb [H,W,3]."""
    require_gpu(pnt2pixaabb, "pnt2pixaabb")
    dev = pnt2pixaabb.device
    w, h = img_shape
    img = torch.empty((h, w, 3), device=dev, dtype=torch.float32)
    i2p = idx2pnt if idx2pnt.numel() else torch.zeros(1, device=dev, dtype=torch.int32)
    call("dsplat_gauss_rasterize_pnt2splat2", stream_ptr(dev), w, h, ptr(img),
         ptr(pnt2pixxydepth.contiguous()), ptr(pnt2pixconvinv.contiguous()),
         ptr(pnt2pixaabb.contiguous()), ptr(pnt2opacity.contiguous()), ptr(pnt2rgb.contiguous()),
         tile_size, ptr(tile2idx), ptr(i2p))
    return img


# ------------------------------------------------------------------ render --
def project_fwd(xyz, scale, quat, opacity, rgb, m16, img_shape, log_scale=True,
                with_conic=True, grad_acc=None):
    """Projection with the packed raster record. Returns (pixxydepth, conic,
    aabb, splat2 [N,16]); conic is None without with_conic (the records hold
    it). grad_acc: a float32 [9N] tensor the kernel sets to zero (the
    accumulators rasterize_bwd / rasterize_bwd_l2 take as grad_acc)."""
    dev = xyz.device
    n = xyz.shape[0]
    pxd = torch.empty((n, 3), device=dev, dtype=torch.float32)
    conic = torch.empty((n, 3), device=dev, dtype=torch.float32) if with_conic else None
    aabb = torch.empty((n, 4), device=dev, dtype=torch.float32)
    splat2 = torch.empty((n, 16), device=dev, dtype=torch.float32)
    if grad_acc is not None:
        _check_grad_acc(grad_acc, n, dev)
    call("dsplat_gauss_project_fwd", stream_ptr(dev), n, ptr(xyz), ptr(scale), ptr(quat),
         ptr(opacity), ptr(rgb), ptr(m16), img_shape[0], img_shape[1],
         _native.DSPLAT_SCALE_IS_LOG if log_scale else 0, ptr(pxd), ptr(conic), ptr(aabb),
         ptr(splat2), ptr(grad_acc))
    return pxd, conic, aabb, splat2


def project_bin(xyz, scale, quat, opacity, rgb, m16, img_shape, tile_size, capacity,
                log_scale=True, grad_acc=None, d_num_idx=None, boxes=False, workspace=None):
    """Projection and binning in one call without a host wait
    (dsplat_gauss_project_bin_ws): returns (splat2 [N,16], tile2idx [T+1],
    idx2pnt [capacity], d_num_idx, pixxydepth, aabb). The records, lists and
    K equal project_fwd followed by bin_gauss_capacity; the projection kernel
    computes what the binning's count phase would, so pixxydepth and aabb are
    returned (as (N,3) / (N,4) tensors) only with boxes=True, else None.
    workspace (uint8, reused across calls) may be passed to avoid an
    allocation per call."""
    dev = xyz.device
    n = xyz.shape[0]
    w, h = img_shape
    tw, th = tile_shape(img_shape, tile_size)
    splat2 = torch.empty((n, 16), device=dev, dtype=torch.float32)
    pxd = torch.empty((n, 3), device=dev, dtype=torch.float32) if boxes else None
    aabb = torch.empty((n, 4), device=dev, dtype=torch.float32) if boxes else None
    if grad_acc is not None:
        _check_grad_acc(grad_acc, n, dev)
    ws_bytes = _native.lib().dsplat_tile2idx_idx2pnt_workspace_size(n, tw * th, capacity)
    if workspace is None or workspace.numel() < ws_bytes:
        workspace = torch.empty(ws_bytes, device=dev, dtype=torch.uint8)
    tile2idx = torch.empty(tw * th + 1, device=dev, dtype=torch.int32)
    idx2pnt = torch.empty(max(capacity, 1), device=dev, dtype=torch.int32)
    if d_num_idx is None:
        d_num_idx = torch.empty(1, device=dev, dtype=torch.int32)
    call("dsplat_gauss_project_bin_ws", stream_ptr(dev), n, ptr(xyz), ptr(scale), ptr(quat),
         ptr(opacity), ptr(rgb), ptr(m16), w, h,
         _native.DSPLAT_SCALE_IS_LOG if log_scale else 0, ptr(pxd), ptr(aabb), ptr(splat2),
         ptr(grad_acc), tile_size, ptr(tile2idx), ptr(idx2pnt), capacity, ptr(workspace),
         workspace.numel(), ptr(d_num_idx))
    return splat2, tile2idx, idx2pnt[:capacity], d_num_idx, pxd, aabb


def _check_grad_acc(grad_acc, n, dev):
    if (grad_acc.dtype != torch.float32 or grad_acc.numel() != 9 * n
            or not grad_acc.is_contiguous() or grad_acc.device != dev):
        raise ValueError("grad_acc must be a contiguous float32 [9N] tensor on the "
                         "splats' device")


def rasterize_fwd(img_shape, tile_size, splat2, tile2idx, idx2pnt, threads=0, tile_order=None,
                  tile_cost=None, state=True):
    """Returns (rgb [H,W,3], alpha [H,W], nwalk [H,W] int32, tfinal [H,W]).
    tile_order (int32 [T], a permutation) schedules the tiles (results do not
    depend on it); tile_cost (int32 [T]) receives each tile's longest walk.
    state=False: a plain forward, without the adjoint's state (nwalk and
    tfinal are returned as None; the kernel skips tracking them)."""
    dev = splat2.device
    w, h = img_shape
    img = torch.empty((h, w, 3), device=dev, dtype=torch.float32)
    alpha = torch.empty((h, w), device=dev, dtype=torch.float32)
    nwalk = torch.empty((h, w), device=dev, dtype=torch.int32) if state else None
    tfinal = torch.empty((h, w), device=dev, dtype=torch.float32) if state else None
    i2p = idx2pnt if idx2pnt.numel() else torch.zeros(1, device=dev, dtype=torch.int32)
    if tile_order is None and tile_cost is None:
        call("dsplat_gauss_rasterize_fwd", stream_ptr(dev), w, h, tile_size, threads, ptr(splat2),
             ptr(tile2idx), ptr(i2p), ptr(img), ptr(alpha), ptr(nwalk), ptr(tfinal))
    else:
        _check_schedule(tile2idx, tile_order, tile_cost)
        call("dsplat_gauss_rasterize_fwd_ordered", stream_ptr(dev), w, h, tile_size, threads,
             ptr(splat2), ptr(tile2idx), ptr(i2p), ptr(tile_order), ptr(img), ptr(alpha),
             ptr(nwalk), ptr(tfinal), ptr(tile_cost))
    return img, alpha, nwalk, tfinal


def _check_schedule(tile2idx, *arrays):
    for a in arrays:
        if a is not None and (a.dtype != torch.int32 or a.numel() != tile2idx.numel() - 1):
            raise ValueError("tile schedule arrays must be int32 with one entry per tile")


def tile_order(tile_cost, out=None):
    """Tiles by descending cost (int32 [T]): the schedule for rasterize_bwd and,
    as a hint, the next frame's rasterize_fwd. out (int32 [T]) receives it if
    given (a graph-captured frame keeps its schedule in one buffer)."""
    order = torch.empty_like(tile_cost) if out is None else out
    if order.dtype != torch.int32 or order.numel() != tile_cost.numel():
        raise ValueError("tile_order: out must be an int32 tensor of the cost's size")
    call("dsplat_tile_order", stream_ptr(tile_cost.device), tile_cost.numel(), ptr(tile_cost),
         ptr(order))
    return order


def rasterize_bwd(img_shape, tile_size, splat2, tile2idx, idx2pnt, nwalk, tfinal, d_img,
                  d_alpha=None, threads=0, tile_order=None, grad_acc=None):
    """Returns 2D gradients (d_pixxy [N,2], d_conic [N,3], d_opacity [N],
    d_rgb [N,3]). tile_order as for rasterize_fwd (tile_order(tile_cost) of
    the forward is the natural one). The adjoint adds into its outputs:
    grad_acc, if given, is the caller's float32 [9N] buffer, already zeroed,
    that the four outputs are views of; otherwise one is allocated zeroed."""
    dev = splat2.device
    n = splat2.shape[0]
    w, h = img_shape
    if grad_acc is None:
        acc = torch.zeros(9 * n, device=dev, dtype=torch.float32)
    else:
        _check_grad_acc(grad_acc, n, dev)
        acc = grad_acc.view(-1)
    d_pixxy = acc[:2 * n].view(n, 2)
    d_conic = acc[2 * n:5 * n].view(n, 3)
    d_op = acc[5 * n:6 * n]
    d_rgb = acc[6 * n:].view(n, 3)
    i2p = idx2pnt if idx2pnt.numel() else torch.zeros(1, device=dev, dtype=torch.int32)
    da = None if d_alpha is None else d_alpha.contiguous()
    if tile_order is None:
        call("dsplat_gauss_rasterize_bwd", stream_ptr(dev), w, h, tile_size, threads,
             ptr(splat2), ptr(tile2idx), ptr(i2p), ptr(nwalk), ptr(tfinal),
             ptr(d_img.contiguous()), ptr(da), ptr(d_pixxy), ptr(d_conic), ptr(d_op), ptr(d_rgb))
    else:
        _check_schedule(tile2idx, tile_order)
        call("dsplat_gauss_rasterize_bwd_ordered", stream_ptr(dev), w, h, tile_size, threads,
             ptr(splat2), ptr(tile2idx), ptr(i2p), ptr(tile_order), ptr(nwalk), ptr(tfinal),
             ptr(d_img.contiguous()), ptr(da), ptr(d_pixxy), ptr(d_conic), ptr(d_op), ptr(d_rgb))
    return d_pixxy, d_conic, d_op, d_rgb


def project_bwd(xyz, scale, quat, m16, img_shape, d_pixxy, d_conic, log_scale=True):
    """Returns (d_xyz [N,3], d_scale [N,3], d_quat [N,4])."""
    dev = xyz.device
    n = xyz.shape[0]
    d_xyz = torch.empty((n, 3), device=dev, dtype=torch.float32)
    d_scale = torch.empty((n, 3), device=dev, dtype=torch.float32)
    d_quat = torch.empty((n, 4), device=dev, dtype=torch.float32)
    flags = _native.DSPLAT_GRAD_OVERWRITE | (_native.DSPLAT_SCALE_IS_LOG if log_scale else 0)
    call("dsplat_gauss_project_bwd", stream_ptr(dev), n, ptr(xyz), ptr(scale), ptr(quat),
         ptr(m16), img_shape[0], img_shape[1], flags, ptr(d_pixxy), ptr(d_conic), ptr(d_xyz),
         ptr(d_scale), ptr(d_quat))
    return d_xyz, d_scale, d_quat


def _cam_pos(cam_pos, device):
    if torch.is_tensor(cam_pos):
        c = cam_pos.to(device=device, dtype=torch.float32).reshape(-1).contiguous()
    else:
        c = torch.tensor([float(v) for v in cam_pos], device=device, dtype=torch.float32)
    if c.numel() != 3:
        raise ValueError("cam_pos must hold 3 values")
    return c


def _check_sh(xyz, sh, degree):
    require_gpu(xyz, "xyz")
    n = xyz.shape[0]
    check_f32(xyz, (n, 3), "xyz")
    if sh.dim() != 3 or sh.shape[0] != n or sh.shape[2] != 3 or sh.shape[1] not in (1, 4, 9, 16):
        raise ValueError(f"sh must be [N, K, 3] with K in (1, 4, 9, 16); got {tuple(sh.shape)}")
    check_f32(sh, tuple(sh.shape), "sh")
    if not 0 <= degree <= 3 or (degree + 1) ** 2 > sh.shape[1]:
        raise ValueError(f"degree {degree} needs {(degree + 1) ** 2} coefficients, sh has "
                         f"{sh.shape[1]}")
    if not (xyz.is_contiguous() and sh.is_contiguous()):
        raise ValueError("xyz and sh must be contiguous")


def sh_to_rgb_fwd(xyz, sh, cam_pos, degree):
    """View-dependent colour [N,3] of spherical-harmonic coefficients sh
    [N,K,3] seen from cam_pos (no counterpart in the reference)."""
    _check_sh(xyz, sh, degree)
    dev = xyz.device
    rgb = torch.empty((xyz.shape[0], 3), device=dev, dtype=torch.float32)
    call("dsplat_sh_to_rgb", stream_ptr(dev), xyz.shape[0], degree, sh.shape[1],
         ptr(_cam_pos(cam_pos, dev)), ptr(xyz), ptr(sh), ptr(rgb))
    return rgb


def sh_to_rgb_bwd(xyz, sh, cam_pos, degree, d_rgb, d_sh=None, d_xyz=None):
    """(d_sh [N,K,3], d_xyz [N,3]) of the colour's adjoint; added into d_sh
    and d_xyz when both are given, stored in new tensors otherwise."""
    _check_sh(xyz, sh, degree)
    dev = xyz.device
    n = xyz.shape[0]
    check_f32(d_rgb, (n, 3), "d_rgb")
    add = d_sh is not None and d_xyz is not None
    if not add:
        d_sh = torch.empty_like(sh)
        d_xyz = torch.empty((n, 3), device=dev, dtype=torch.float32)
    check_f32(d_sh, tuple(sh.shape), "d_sh")
    check_f32(d_xyz, (n, 3), "d_xyz")
    if not (d_rgb.is_contiguous() and d_sh.is_contiguous() and d_xyz.is_contiguous()):
        raise ValueError("d_rgb, d_sh and d_xyz must be contiguous")
    call("dsplat_sh_to_rgb_bwd", stream_ptr(dev), n, degree, sh.shape[1],
         ptr(_cam_pos(cam_pos, dev)), ptr(xyz), ptr(sh), ptr(d_rgb),
         0 if add else _native.DSPLAT_GRAD_OVERWRITE, ptr(d_sh), ptr(d_xyz))
    return d_sh, d_xyz


class _ShColour(torch.autograd.Function):
    @staticmethod
    def forward(ctx, xyz, sh, cam, degree):
        xyz, sh = xyz.contiguous(), sh.contiguous()
        ctx.save_for_backward(xyz, sh, cam)
        ctx.degree = degree
        return sh_to_rgb_fwd(xyz, sh, cam, degree)

    @staticmethod
    def backward(ctx, d_rgb):
        xyz, sh, cam = ctx.saved_tensors
        d_sh, d_xyz = sh_to_rgb_bwd(xyz, sh, cam, ctx.degree, d_rgb.contiguous())
        return d_xyz, d_sh, None, None


def sh_to_rgb(xyz, sh, cam_pos, degree=3):
    """Differentiable view-dependent colour: rgb [N,3] =
    max(sum_k Y_k(normalize(xyz - cam_pos)) sh[:,k,:] + 1/2, 0), Y_k the real
    spherical harmonics of 3D Gaussian splatting (k < (degree+1)^2). Feed the
    result to `render` as its rgb; gradients reach sh and xyz."""
    _check_sh(xyz, sh, degree)
    return _ShColour.apply(xyz, sh, _cam_pos(cam_pos, xyz.device), int(degree))


def rasterize_bwd_l2(img_shape, tile_size, splat2, tile2idx, idx2pnt, nwalk, tfinal, img, target,
                     threads=0, tile_order=None, grad_acc=None):
    """rasterize_bwd for the loss sum (img - target)^2: dL/drgb = 2 (img -
    target) is formed per pixel inside the adjoint kernel (img is the
    forward's output, target [H,W,3]). Returns the same four 2D gradients."""
    dev = splat2.device
    n = splat2.shape[0]
    w, h = img_shape
    for t, name in ((img, "img"), (target, "target")):
        check_f32(t, (h, w, 3), name)
    if grad_acc is None:
        acc = torch.zeros(9 * n, device=dev, dtype=torch.float32)
    else:
        _check_grad_acc(grad_acc, n, dev)
        acc = grad_acc.view(-1)
    if tile_order is not None:
        _check_schedule(tile2idx, tile_order)
    d_pixxy = acc[:2 * n].view(n, 2)
    d_conic = acc[2 * n:5 * n].view(n, 3)
    d_op = acc[5 * n:6 * n]
    d_rgb = acc[6 * n:].view(n, 3)
    i2p = idx2pnt if idx2pnt.numel() else torch.zeros(1, device=dev, dtype=torch.int32)
    call("dsplat_gauss_rasterize_bwd_l2", stream_ptr(dev), w, h, tile_size, threads, ptr(splat2),
         ptr(tile2idx), ptr(i2p), ptr(tile_order), ptr(nwalk), ptr(tfinal), ptr(img.contiguous()),
         ptr(target.contiguous()), ptr(d_pixxy), ptr(d_conic), ptr(d_op), ptr(d_rgb))
    return d_pixxy, d_conic, d_op, d_rgb


def bin_gauss(img_shape, tile_size, pxd, aabb):
    return tile2idx_idx2pnt(img_shape, tile_size, pxd, aabb)


# last tile order per (device, image shape, tile size): a scheduling hint only
_SCHEDULE = {}


class _Render(torch.autograd.Function):
    @staticmethod
    def forward(ctx, xyz, log_scale, quat, opacity, rgb, m16, img_shape, tile_size, threads,
                threads_bwd):
        xyz, log_scale, quat, opacity, rgb = (t.contiguous() for t in
                                              (xyz, log_scale, quat, opacity, rgb))
        # the projection also clears the adjoint's accumulators
        gacc = torch.empty(9 * xyz.shape[0], device=xyz.device, dtype=torch.float32)
        pxd, _, aabb, splat2 = project_fwd(xyz, log_scale, quat, opacity, rgb, m16, img_shape,
                                           log_scale=True, with_conic=False, grad_acc=gacc)
        tile2idx, idx2pnt = bin_gauss(img_shape, tile_size, pxd, aabb)
        # schedule: the previous frame's order at this image / tile size as a
        # hint for the forward, this frame's walks for the adjoint
        key = (str(xyz.device), tuple(img_shape), tile_size)
        hint = _SCHEDULE.get(key)
        cost = torch.empty(tile2idx.numel() - 1, device=xyz.device, dtype=torch.int32)
        img, alpha, nwalk, tfinal = rasterize_fwd(img_shape, tile_size, splat2, tile2idx,
                                                 idx2pnt, threads, tile_order=hint,
                                                 tile_cost=cost)
        order = tile_order(cost)
        _SCHEDULE[key] = order
        ctx.save_for_backward(xyz, log_scale, quat, m16, splat2, tile2idx, idx2pnt, nwalk, tfinal,
                              order)
        ctx.gacc = gacc  # zeroed; handed to the first backward only
        ctx.img_shape = img_shape
        ctx.tile_size = tile_size
        ctx.threads = threads_bwd
        ctx.mark_non_differentiable(nwalk)
        return img, alpha

    @staticmethod
    def backward(ctx, d_img, d_alpha):
        (xyz, log_scale, quat, m16, splat2, tile2idx, idx2pnt, nwalk, tfinal,
         order) = ctx.saved_tensors
        if d_img is None:
            d_img = torch.zeros((ctx.img_shape[1], ctx.img_shape[0], 3), device=xyz.device)
        gacc, ctx.gacc = ctx.gacc, None
        d_pixxy, d_conic, d_op, d_rgb = rasterize_bwd(ctx.img_shape, ctx.tile_size, splat2,
                                                      tile2idx, idx2pnt, nwalk, tfinal, d_img,
                                                      d_alpha, ctx.threads, tile_order=order,
                                                      grad_acc=gacc)
        d_xyz, d_ls, d_quat = project_bwd(xyz, log_scale, quat, m16, ctx.img_shape, d_pixxy,
                                          d_conic, log_scale=True)
        return d_xyz, d_ls, d_quat, d_op, d_rgb, None, None, None, None, None


def render(xyz, log_scale, quat, opacity, rgb, transform_world2ndc, img_shape,
           tile_size=DEFAULT_TILE, threads=0, threads_bwd=None):
    """Differentiable render of N Gaussians -> (rgb [H,W,3], alpha [H,W]).

    xyz [N,3] means, log_scale [N,3] log sigma, quat [N,4] (x,y,z,w; used as
    given), opacity [N] in (0,1), rgb [N,3]; transform_world2ndc (4,4)
    row-major or (16,) column-major; img_shape (W, H). Gradients flow to all
    five per-Gaussian inputs through the adjoint kernels. threads /
    threads_bwd: workgroup size of the forward / adjoint raster (0: default;
    threads_bwd defaults to threads).
    """
    require_gpu(xyz, "xyz")
    n = xyz.shape[0]
    check_f32(xyz, (n, 3), "xyz")
    check_f32(log_scale, (n, 3), "log_scale")
    check_f32(quat, (n, 4), "quat")
    check_f32(opacity, (n,), "opacity")
    check_f32(rgb, (n, 3), "rgb")
    if tile_size not in (8, 16, 32):
        raise ValueError("tile_size must be 8, 16 or 32")
    m16 = world2ndc_col_major(transform_world2ndc, xyz.device)
    return _Render.apply(xyz, log_scale, quat, opacity, rgb, m16, tuple(img_shape), tile_size,
                         threads, threads if threads_bwd is None else threads_bwd)
