"""ctypes front end of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker: the product library never calls it.
Every function takes and returns numpy arrays. The C++ sources cite the
reference file:line each function restates.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u32 = ctypes.c_uint32
_vp = ctypes.c_void_p


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_gauss_project_f32.restype = _u32
        L.orc_gauss_project_f32.argtypes = [_u32, _f32p, _f32p, _f32p, _f32p, _u32, _u32,
                                            _f32p, _f32p, _f32p]
        L.orc_gauss_project_f64.restype = _u32
        L.orc_gauss_project_f64.argtypes = [_u32, _f64p, _f64p, _f64p, _f64p, _u32, _u32,
                                            _f64p, _f64p, _f64p]
        L.orc_gauss_project_one_f64.restype = None
        L.orc_gauss_project_one_f64.argtypes = [_f64p, _f64p, _f64p, _f64p, _u32, _u32, _f64p]
        L.orc_safe_inverse_spd_f64.restype = ctypes.c_int
        L.orc_safe_inverse_spd_f64.argtypes = [_f64p, ctypes.c_double, _f64p]
        L.orc_tile2pnt.restype = ctypes.c_uint64
        L.orc_tile2pnt.argtypes = [_u32, _f32p, _f32p, _u32, _u32, _u32, _vp, _vp, _vp]
        L.orc_circle_boxes.restype = None
        L.orc_circle_boxes.argtypes = [_u32, _f32p, _f32p, _f32p]
        L.orc_tile_rects.restype = None
        L.orc_tile_rects.argtypes = [_u32, _f32p, _u32, _u32, _u32, _i32p]
        L.orc_gauss_rasterize_f32.restype = None
        L.orc_gauss_rasterize_f32.argtypes = [_u32, _u32, _f32p, _f32p, _f32p, _f32p, _f32p,
                                              _u32, _u32p, _u32p, _f32p, _f32p, _u32p, _f32p, _vp]
        L.orc_gauss_rasterize_exp_ulp_f32.restype = None
        L.orc_gauss_rasterize_exp_ulp_f32.argtypes = [_u32, _u32, _f32p, _f32p, _f32p, _f32p,
                                                      _f32p, _u32, _u32p, _u32p, _f32p, _f32p,
                                                      _u32p, _f32p, ctypes.c_int]
        L.orc_gauss_l2_loss_tiles_f64.restype = ctypes.c_double
        L.orc_gauss_l2_loss_tiles_f64.argtypes = [_u32, _u32, _f64p, _f64p, _f32p, _f64p, _f64p,
                                                  _u32, _u32p, _u32p, _u32p, _f32p, _u32, _u32,
                                                  _u32, _u32]
        L.orc_gauss_rasterize_f64.restype = None
        L.orc_gauss_rasterize_f64.argtypes = [_u32, _u32, _f64p, _f64p, _f32p, _f64p, _f64p,
                                              _u32, _u32p, _u32p, _f64p, _f64p, _u32p, _f64p, _vp]
        L.orc_gauss_rasterize_bwd_f32.restype = None
        L.orc_gauss_rasterize_bwd_f32.argtypes = [_u32, _u32, _f32p, _f32p, _f32p, _f32p, _f32p,
                                                  _u32, _u32p, _u32p, _u32p, _f32p, _f32p, _vp,
                                                  _f32p, _f32p, _f32p, _f32p]
        L.orc_set_threads.restype = ctypes.c_int
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_gauss_rasterize_bwd_mt_f32.restype = None
        L.orc_gauss_rasterize_bwd_mt_f32.argtypes = [_u32, _u32, _u32, _f32p, _f32p, _f32p,
                                                     _f32p, _f32p, _u32, _u32p, _u32p, _u32p,
                                                     _f32p, _f32p, _vp, _f32p, _f32p, _f32p,
                                                     _f32p]
        L.orc_gauss_rasterize_bwd_f64.restype = None
        L.orc_gauss_rasterize_bwd_f64.argtypes = [_u32, _u32, _f64p, _f64p, _f32p, _f64p, _f64p,
                                                  _u32, _u32p, _u32p, _u32p, _f64p, _f64p, _vp,
                                                  _f64p, _f64p, _f64p, _f64p]
        L.orc_gauss_project_bwd_f32.restype = None
        L.orc_gauss_project_bwd_f32.argtypes = [_u32, _f32p, _f32p, _f32p, _f32p, _u32, _u32,
                                                _f32p, _f32p, _f32p, _f32p, _f32p, ctypes.c_int]
        L.orc_gauss_project_bwd_f64.restype = None
        L.orc_gauss_project_bwd_f64.argtypes = [_u32, _f64p, _f64p, _f64p, _f64p, _u32, _u32,
                                                _f64p, _f64p, _f64p, _f64p, _f64p, ctypes.c_int]
        L.orc_gauss_render_sort_depth.restype = None
        L.orc_gauss_render_sort_depth.argtypes = [_u32, _f32p, _f32p, _f32p, _f32p, _f32p,
                                                  _u32, _u32, _f32p]
        L.orc_circle_project.restype = None
        L.orc_circle_project.argtypes = [_u32, _f32p, ctypes.c_float, _f32p, _u32, _u32,
                                         _f32p, _f32p]
        L.orc_circle_rasterize.restype = None
        L.orc_circle_rasterize.argtypes = [_u32, _u32, _f32p, _f32p, _f32p, _u32, _u32p, _u32p,
                                           _f32p]
        L.orc_zbuffer.restype = None
        L.orc_zbuffer.argtypes = [_u32, _f32p, _vp, _f32p, _u32, _u32, _u64p, _f32p]
        L.orc_render_pix_sort_depth.restype = None
        L.orc_render_pix_sort_depth.argtypes = [_u32, _f32p, _f32p, _u32, _u32, _f32p]
        L.orc_transfer.restype = ctypes.c_float
        L.orc_transfer.argtypes = [_f32p, _u32]
        L.orc_dw_transfer.restype = ctypes.c_float
        L.orc_dw_transfer.argtypes = [_f32p, _u32, _u32]
        L.orc_adam.restype = None
        L.orc_adam.argtypes = [_u32, _f32p, _f32p, _f32p, _f32p, ctypes.c_float, ctypes.c_float,
                               ctypes.c_float, ctypes.c_float, _u32]
        L.orc_sh_basis_f64.restype = None
        L.orc_sh_basis_f64.argtypes = [ctypes.c_double] * 3 + [_u32, _f64p, _vp]
        for sfx, fp in (("f32", _f32p), ("f64", _f64p)):
            f = getattr(L, "orc_sh_to_rgb_" + sfx)
            f.restype = None
            f.argtypes = [_u32, _u32, _u32, fp, fp, fp, fp]
            f = getattr(L, "orc_sh_to_rgb_bwd_" + sfx)
            f.restype = None
            f.argtypes = [_u32, _u32, _u32, fp, fp, fp, fp, fp, fp, ctypes.c_int]
        _lib = L
    return _lib


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def tile_shape(img_shape, ts):
    w, h = img_shape
    return (w + ts - 1) // ts, (h + ts - 1) // ts


# ---------------------------------------------------------------- projection
def gauss_project(xyz, quat, scale, world2ndc, img_shape, dtype=np.float32):
    """splat_gauss.cu:10-53 / pnt2xyz.rs:51-96. world2ndc: col-major flat 16."""
    xyz, quat, scale, m = (_c(a, dtype).reshape(-1) for a in (xyz, quat, scale, world2ndc))
    n = xyz.size // 3
    pxd = np.empty(3 * n, dtype)
    conic = np.empty(3 * n, dtype)
    aabb = np.empty(4 * n, dtype)
    f = lib().orc_gauss_project_f32 if dtype == np.float32 else lib().orc_gauss_project_f64
    nclamp = f(n, xyz, quat, scale, m, img_shape[0], img_shape[1], pxd, conic, aabb)
    return pxd.reshape(n, 3), conic.reshape(n, 3), aabb.reshape(n, 4), int(nclamp)


def gauss_project_one(xyz, quat, scale, world2ndc, img_shape):
    out = np.zeros(39, np.float64)
    lib().orc_gauss_project_one_f64(_c(xyz, np.float64), _c(quat, np.float64),
                                    _c(scale, np.float64), _c(world2ndc, np.float64).reshape(-1),
                                    img_shape[0], img_shape[1], out)
    return dict(pix=out[0:2], ndcz=out[2], w=out[3], jac=out[4:13], wpix=out[13:19],
                rot=out[19:28], cov=out[28:31], conic=out[31:34], aabb=out[34:38],
                clamped=int(out[38]))


def safe_inverse_spd(abc, eps):
    out = np.zeros(3, np.float64)
    c = lib().orc_safe_inverse_spd_f64(_c(abc, np.float64), eps, out)
    return out, c


# ------------------------------------------------------------------- binning
def tile2pnt(boxes, pixxydepth, img_shape, ts, with_keys=False):
    """tile_acceleration.rs:3-58: returns (tile2idx[T+1], idx2pnt[K][, keys[K]])."""
    boxes = _c(boxes, np.float32).reshape(-1)
    pxd = _c(pixxydepth, np.float32).reshape(-1)
    n = boxes.size // 4
    tw, th = tile_shape(img_shape, ts)
    tile2idx = np.zeros(tw * th + 1, np.uint32)
    L = lib()
    K = L.orc_tile2pnt(n, boxes, pxd, img_shape[0], img_shape[1], ts, _ptr(tile2idx), None, None)
    idx2pnt = np.zeros(max(K, 1), np.uint32)
    keys = np.zeros(max(K, 1), np.uint64)
    L.orc_tile2pnt(n, boxes, pxd, img_shape[0], img_shape[1], ts, _ptr(tile2idx),
                   _ptr(idx2pnt), _ptr(keys))
    if with_keys:
        return tile2idx, idx2pnt[:K], keys[:K]
    return tile2idx, idx2pnt[:K]


def tile2pnt_gauss(pixxydepth, aabb, img_shape, ts, with_keys=False):
    """tile_acceleration.rs:60-77."""
    return tile2pnt(aabb, pixxydepth, img_shape, ts, with_keys)


def circle_boxes(pixxydepth, pixrad):
    pxd = _c(pixxydepth, np.float32).reshape(-1)
    rad = _c(pixrad, np.float32).reshape(-1)
    n = rad.size
    boxes = np.empty(4 * n, np.float32)
    lib().orc_circle_boxes(n, pxd, rad, boxes)
    return boxes.reshape(n, 4)


def tile2pnt_circle(pixxydepth, pixrad, img_shape, ts, with_keys=False):
    """tile_acceleration.rs:79-97."""
    return tile2pnt(circle_boxes(pixxydepth, pixrad), pixxydepth, img_shape, ts, with_keys)


def tile_rects(boxes, ts, tw, th):
    boxes = _c(boxes, np.float32).reshape(-1)
    n = boxes.size // 4
    out = np.zeros(4 * n, np.int32)
    lib().orc_tile_rects(n, boxes, ts, tw, th, out)
    return out.reshape(n, 4)


# -------------------------------------------------------------------- raster
def gauss_rasterize(img_shape, pixxydepth, conic, aabb, opacity, rgb, ts, tile2idx, idx2pnt,
                    dtype=np.float32, frozen_nwalk=None):
    """splat_gauss.cu:56-109 plus alpha / nwalk / tfinal."""
    w, h = img_shape
    pxd, cn, op, col = (_c(a, dtype).reshape(-1) for a in (pixxydepth, conic, opacity, rgb))
    bb = _c(aabb, np.float32).reshape(-1)
    t2i = _c(tile2idx, np.uint32)
    i2p = _c(idx2pnt, np.uint32)
    if i2p.size == 0:
        i2p = np.zeros(1, np.uint32)
    img = np.zeros(h * w * 3, dtype)
    alpha = np.zeros(h * w, dtype)
    nwalk = np.zeros(h * w, np.uint32)
    tfinal = np.zeros(h * w, dtype)
    fz = None if frozen_nwalk is None else _c(frozen_nwalk, np.uint32).reshape(-1)
    f = lib().orc_gauss_rasterize_f32 if dtype == np.float32 else lib().orc_gauss_rasterize_f64
    f(w, h, pxd, cn, bb, op, col, ts, t2i, i2p, img, alpha, nwalk, tfinal, _ptr(fz))
    return (img.reshape(h, w, 3), alpha.reshape(h, w), nwalk.reshape(h, w), tfinal.reshape(h, w))


def gauss_rasterize_exp_ulp(img_shape, pixxydepth, conic, aabb, opacity, rgb, ts, tile2idx,
                            idx2pnt, exp_ulp):
    """gauss_rasterize (float) with every exp result moved exp_ulp ulps (a
    planted perturbation: tests count the pixels whose termination it flips)."""
    w, h = img_shape
    pxd, cn, op, col = (_c(a, np.float32).reshape(-1) for a in (pixxydepth, conic, opacity, rgb))
    bb = _c(aabb, np.float32).reshape(-1)
    i2p = _c(idx2pnt, np.uint32)
    if i2p.size == 0:
        i2p = np.zeros(1, np.uint32)
    img = np.zeros(h * w * 3, np.float32)
    alpha = np.zeros(h * w, np.float32)
    nwalk = np.zeros(h * w, np.uint32)
    tfinal = np.zeros(h * w, np.float32)
    lib().orc_gauss_rasterize_exp_ulp_f32(w, h, pxd, cn, bb, op, col, ts,
                                          _c(tile2idx, np.uint32), i2p, img, alpha, nwalk, tfinal,
                                          int(exp_ulp))
    return (img.reshape(h, w, 3), alpha.reshape(h, w), nwalk.reshape(h, w), tfinal.reshape(h, w))


def gauss_l2_loss_tiles(img_shape, pixxydepth, conic, aabb, opacity, rgb, ts, tile2idx, idx2pnt,
                        frozen_nwalk, target, tile_rect):
    """float64: sum (rgb - target)^2 over the tiles tile_rect = (tx0, tx1, ty0,
    ty1) (inclusive) of the tiled raster with every walk frozen at
    frozen_nwalk. The arrays are passed as they are (no copies), so a caller
    can change one Gaussian in place between calls: pixxydepth, conic,
    opacity, rgb must be C-contiguous float64, aabb float32, the lists and
    frozen_nwalk uint32, target float32."""
    for a, dt in ((pixxydepth, np.float64), (conic, np.float64), (opacity, np.float64),
                  (rgb, np.float64), (aabb, np.float32), (tile2idx, np.uint32),
                  (idx2pnt, np.uint32), (frozen_nwalk, np.uint32), (target, np.float32)):
        assert a.dtype == dt and a.flags.c_contiguous
    w, h = img_shape
    tx0, tx1, ty0, ty1 = (int(v) for v in tile_rect)
    return float(lib().orc_gauss_l2_loss_tiles_f64(
        w, h, pixxydepth.reshape(-1), conic.reshape(-1), aabb.reshape(-1), opacity.reshape(-1),
        rgb.reshape(-1), ts, tile2idx, idx2pnt, frozen_nwalk.reshape(-1), target.reshape(-1),
        tx0, tx1, ty0, ty1))


def self_flips(img_shape, pixxydepth, conic, aabb, opacity, rgb, ts, tile2idx, idx2pnt, nwalk):
    """Pixels whose walk length changes when every exp result of the oracle
    moves one ulp up or one ulp down, against the oracle's own nwalk (the
    union of both sets): the termination flips that the last bit of the
    exponential decides."""
    out = np.zeros(np.asarray(nwalk).shape, bool)
    for u in (1, -1):
        nw = gauss_rasterize_exp_ulp(img_shape, pixxydepth, conic, aabb, opacity, rgb, ts,
                                     tile2idx, idx2pnt, u)[2]
        out |= nw != nwalk
    return out


def gauss_rasterize_bwd(img_shape, pixxydepth, conic, aabb, opacity, rgb, ts, tile2idx, idx2pnt,
                        nwalk, tfinal, d_img, d_alpha=None, dtype=np.float32):
    w, h = img_shape
    pxd, cn, op, col, tl, dimg = (_c(a, dtype).reshape(-1)
                                  for a in (pixxydepth, conic, opacity, rgb, tfinal, d_img))
    bb = _c(aabb, np.float32).reshape(-1)
    n = op.size
    da = None if d_alpha is None else _c(d_alpha, dtype).reshape(-1)
    i2p = _c(idx2pnt, np.uint32)
    if i2p.size == 0:
        i2p = np.zeros(1, np.uint32)
    d_pixxy = np.zeros(2 * n, dtype)
    d_conic = np.zeros(3 * n, dtype)
    d_op = np.zeros(n, dtype)
    d_rgb = np.zeros(3 * n, dtype)
    f = (lib().orc_gauss_rasterize_bwd_f32 if dtype == np.float32
         else lib().orc_gauss_rasterize_bwd_f64)
    f(w, h, pxd, cn, bb, op, col, ts, _c(tile2idx, np.uint32), i2p,
      _c(nwalk, np.uint32).reshape(-1), tl, dimg, _ptr(da), d_pixxy, d_conic, d_op, d_rgb)
    return d_pixxy.reshape(n, 2), d_conic.reshape(n, 3), d_op, d_rgb.reshape(n, 3)


def set_threads(n):
    """Threads of the oracle's loops: 1 (the default) is the sequential parity
    oracle; more give the all-core CPU baseline (same results, except the
    adjoint's float sums in another order). Returns the number set."""
    return int(lib().orc_set_threads(int(n)))


def gauss_rasterize_bwd_mt(img_shape, pixxydepth, conic, aabb, opacity, rgb, ts, tile2idx,
                           idx2pnt, nwalk, tfinal, d_img, d_alpha=None):
    """gauss_rasterize_bwd (float) over the threads set by set_threads (row
    bands into per-thread gradient buffers)."""
    w, h = img_shape
    pxd, cn, op, col, tl, dimg = (_c(a, np.float32).reshape(-1)
                                  for a in (pixxydepth, conic, opacity, rgb, tfinal, d_img))
    bb = _c(aabb, np.float32).reshape(-1)
    n = op.size
    da = None if d_alpha is None else _c(d_alpha, np.float32).reshape(-1)
    i2p = _c(idx2pnt, np.uint32)
    if i2p.size == 0:
        i2p = np.zeros(1, np.uint32)
    out = [np.zeros(k * n, np.float32) for k in (2, 3, 1, 3)]
    lib().orc_gauss_rasterize_bwd_mt_f32(n, w, h, pxd, cn, bb, op, col, ts,
                                         _c(tile2idx, np.uint32), i2p,
                                         _c(nwalk, np.uint32).reshape(-1), tl, dimg, _ptr(da),
                                         *out)
    return out[0].reshape(n, 2), out[1].reshape(n, 3), out[2], out[3].reshape(n, 3)


def gauss_project_bwd(xyz, quat, scale, world2ndc, img_shape, d_pixxy, d_conic, log_scale=False,
                      dtype=np.float32):
    xyz, quat, scale, m, dp, dc = (_c(a, dtype).reshape(-1)
                                   for a in (xyz, quat, scale, world2ndc, d_pixxy, d_conic))
    n = xyz.size // 3
    d_xyz = np.zeros(3 * n, dtype)
    d_scale = np.zeros(3 * n, dtype)
    d_quat = np.zeros(4 * n, dtype)
    f = lib().orc_gauss_project_bwd_f32 if dtype == np.float32 else lib().orc_gauss_project_bwd_f64
    f(n, xyz, quat, scale, m, img_shape[0], img_shape[1], dp, dc, d_xyz, d_scale, d_quat,
      int(log_scale))
    return d_xyz.reshape(n, 3), d_scale.reshape(n, 3), d_quat.reshape(n, 4)


def gauss_render_sort_depth(pixxydepth, conic, aabb, opacity, rgb, img_shape):
    """pnt2pixxyndcz.rs:42-100 (naive, non-tiled)."""
    w, h = img_shape
    pxd, cn, bb, op, col = (_c(a, np.float32).reshape(-1)
                            for a in (pixxydepth, conic, aabb, opacity, rgb))
    img = np.zeros(h * w * 3, np.float32)
    lib().orc_gauss_render_sort_depth(op.size, pxd, cn, bb, op, col, w, h, img)
    return img.reshape(h, w, 3)


# -------------------------------------------------------------- circle/point
def circle_project(xyz, radius, world2ndc, img_shape):
    """splat_circle.cu:7-41 / splat_sphere.rs:51-87."""
    xyz = _c(xyz, np.float32).reshape(-1)
    n = xyz.size // 3
    pxd = np.empty(3 * n, np.float32)
    rad = np.empty(n, np.float32)
    lib().orc_circle_project(n, xyz, float(radius), _c(world2ndc, np.float32).reshape(-1),
                             img_shape[0], img_shape[1], pxd, rad)
    return pxd.reshape(n, 3), rad


def circle_rasterize(img_shape, pixxydepth, pixrad, rgb, ts, tile2idx, idx2pnt):
    """splat_circle.cu:93-132."""
    w, h = img_shape
    img = np.zeros(h * w * 3, np.float32)
    i2p = _c(idx2pnt, np.uint32)
    if i2p.size == 0:
        i2p = np.zeros(1, np.uint32)
    lib().orc_circle_rasterize(w, h, _c(pixxydepth, np.float32).reshape(-1),
                               _c(pixrad, np.float32).reshape(-1),
                               _c(rgb, np.float32).reshape(-1), ts, _c(tile2idx, np.uint32), i2p,
                               img)
    return img.reshape(h, w, 3)


def zbuffer(pixxydepth, rgb, img_shape, pixrad=None):
    """splat_point.cu:49-74 with the packed u64 key (min key wins)."""
    w, h = img_shape
    pxd = _c(pixxydepth, np.float32).reshape(-1)
    n = pxd.size // 3
    rad = None if pixrad is None else _c(pixrad, np.float32).reshape(-1)
    keys = np.zeros(h * w, np.uint64)
    img = np.zeros(h * w * 3, np.float32)
    lib().orc_zbuffer(n, pxd, _ptr(rad), _c(rgb, np.float32).reshape(-1), w, h, keys, img)
    return keys.reshape(h, w), img.reshape(h, w, 3)


def render_pix_sort_depth(pixxyndcz, rgb, img_shape):
    """pnt2pixxyndcz.rs:1-40 (point painter, background 0)."""
    w, h = img_shape
    pxd = _c(pixxyndcz, np.float32).reshape(-1)
    img = np.zeros(h * w * 3, np.float32)
    lib().orc_render_pix_sort_depth(pxd.size // 3, pxd, _c(rgb, np.float32).reshape(-1), w, h, img)
    return img.reshape(h, w, 3)


def transfer(alphas, i):
    return lib().orc_transfer(_c(alphas, np.float32), i)


def dw_transfer(alphas, i, j):
    return lib().orc_dw_transfer(_c(alphas, np.float32), i, j)


def adam(param, grad, m1, m2, lr, b1, b2, eps, step):
    """In place on param/m1/m2 (float32 arrays)."""
    lib().orc_adam(param.size, param, _c(grad, np.float32), m1, m2, lr, b1, b2, eps, step)


# ------------------------------------------------------------- SH colour
def sh_basis(d, degree=3, with_grad=False):
    """Real SH basis values (16) and optionally their gradients (16x3) at
    the unit direction d, float64."""
    b = np.zeros(16, np.float64)
    g = np.zeros(48, np.float64) if with_grad else None
    lib().orc_sh_basis_f64(float(d[0]), float(d[1]), float(d[2]), degree, b, _ptr(g))
    return (b, g.reshape(16, 3)) if with_grad else b


def sh_to_rgb(xyz, sh, cam, degree, dtype=np.float32):
    """rgb [N,3] = max(sum_k Y_k(normalize(xyz - cam)) sh[:,k,:] + 1/2, 0);
    sh is [N, K, 3]."""
    sh = _c(sh, dtype)
    n, K = sh.shape[0], sh.shape[1]
    xyz, cam = _c(xyz, dtype).reshape(-1), _c(cam, dtype).reshape(-1)
    rgb = np.empty(3 * n, dtype)
    f = lib().orc_sh_to_rgb_f32 if dtype == np.float32 else lib().orc_sh_to_rgb_f64
    f(n, degree, K, cam, xyz, sh.reshape(-1), rgb)
    return rgb.reshape(n, 3)


def sh_to_rgb_bwd(xyz, sh, cam, degree, d_rgb, dtype=np.float32, d_sh=None, d_xyz=None):
    """(d_sh [N,K,3], d_xyz [N,3]); added into d_sh / d_xyz when given,
    stored otherwise."""
    sh = _c(sh, dtype)
    n, K = sh.shape[0], sh.shape[1]
    overwrite = d_sh is None
    d_sh = np.zeros(n * K * 3, dtype) if d_sh is None else _c(d_sh, dtype).reshape(-1).copy()
    d_xyz = np.zeros(3 * n, dtype) if d_xyz is None else _c(d_xyz, dtype).reshape(-1).copy()
    f = lib().orc_sh_to_rgb_bwd_f32 if dtype == np.float32 else lib().orc_sh_to_rgb_bwd_f64
    f(n, degree, K, _c(cam, dtype).reshape(-1), _c(xyz, dtype).reshape(-1), sh.reshape(-1),
      _c(d_rgb, dtype).reshape(-1), d_sh, d_xyz, int(overwrite))
    return d_sh.reshape(n, K, 3), d_xyz.reshape(n, 3)
