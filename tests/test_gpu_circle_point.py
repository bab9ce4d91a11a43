"""GPU parity of the circle and point paths against the oracle.

All of these are bit-exact: projection (same op order, correctly rounded
divide/sqrt on both sides), binning, painter raster, z-buffer winners and
the gathered colours. The z-buffer case runs at the full 2M-point 1920x1080
size of the point workload.
"""
import numpy as np
import pytest
import torch

from _helpers import u32
from del_splat_dlpack import scene

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def _points(n, seed, shape):
    p = scene.points(n, seed)
    m = scene.front_camera(shape)
    return p, m


def test_circle_projection_bit_exact(oracle):
    import del_splat_dlpack.SplatCircle.torch as SC
    shape = (1280, 720)
    p, m = _points(100_000, 30, shape)
    M = torch.from_numpy(scene.camera.world2ndc((0, 0, 3.0), (0, 0, 0), shape).astype(np.float32))
    pxd, rad = SC.from_project_spheres(_t(p["xyz"]), 0.004, M.to(DEV), shape)
    o_pxd, o_rad = oracle.circle_project(p["xyz"], 0.004, m, shape)
    np.testing.assert_array_equal(pxd.cpu().numpy(), o_pxd)
    np.testing.assert_array_equal(rad.cpu().numpy(), o_rad)


def test_circle_projection_dlpack_cpu_staging(oracle):
    """CPU tensors go through the GPU and back (the library has no CPU path)."""
    import del_splat_dlpack.SplatCircle.torch as SC
    shape = (320, 240)
    p, m = _points(1000, 31, shape)
    M = torch.from_numpy(scene.camera.world2ndc((0, 0, 3.0), (0, 0, 0), shape).astype(np.float32))
    pxd, rad = SC.from_project_spheres(torch.from_numpy(p["xyz"]), 0.01, M, shape)
    assert pxd.device.type == "cpu"
    o_pxd, o_rad = oracle.circle_project(p["xyz"], 0.01, m, shape)
    np.testing.assert_array_equal(pxd.numpy(), o_pxd)
    np.testing.assert_array_equal(rad.numpy(), o_rad)


@pytest.mark.parametrize("ts", [8, 16, 32])
def test_circle_binning_and_painter(oracle, ts):
    import del_splat_dlpack.SplatCircle.torch as SC
    shape = (333, 211)
    p, m = _points(20000, 32 + ts, shape)
    o_pxd, _ = oracle.circle_project(p["xyz"], 0.0, m, shape)
    rad = p["pixrad"]
    tw, th = (shape[0] + ts - 1) // ts, (shape[1] + ts - 1) // ts
    t2i, i2p = SC.tile2idx_idx2pnt((tw, th), ts, _t(o_pxd), _t(rad))
    o_t2i, o_i2p = oracle.tile2pnt_circle(o_pxd, rad, shape, ts)
    np.testing.assert_array_equal(u32(t2i), o_t2i)
    np.testing.assert_array_equal(u32(i2p), o_i2p)
    img = SC.splat(shape, _t(o_pxd), _t(rad), _t(p["rgb"]), ts, t2i, i2p).cpu().numpy()
    o_img = oracle.circle_rasterize(shape, o_pxd, rad, p["rgb"], ts, o_t2i, o_i2p)
    np.testing.assert_array_equal(img, o_img)


def test_circle_zbuffer_full_size(oracle):
    """Point workload: 2M points in the unit cube, pixel radius U[0.5,3],
    1920x1080: winners (packed keys) and colours bit-exact."""
    import del_splat_dlpack.SplatCircle.torch as SC
    shape = (1920, 1080)
    p, m = _points(2_000_000, 1, shape)
    o_pxd, _ = oracle.circle_project(p["xyz"], 0.0, m, shape)
    img, dep = SC.rasterize_zbuffer(_t(o_pxd), _t(p["pixrad"]), _t(p["rgb"]), shape)
    keys, o_img = oracle.zbuffer(o_pxd, p["rgb"], shape, pixrad=p["pixrad"])
    np.testing.assert_array_equal(img.cpu().numpy(), o_img)
    win = keys != np.iinfo(np.uint64).max
    o_dep = np.where(win, (keys >> np.uint64(32)).astype(np.uint32).view(np.float32),
                     np.float32(np.finfo(np.float32).max))
    np.testing.assert_array_equal(dep.cpu().numpy(), o_dep)
    assert win.mean() > 0.05  # the unit cube covers ~7% of the frame


def test_point_zbuffer_and_painter(oracle):
    import del_splat_dlpack.SplatPoint.torch as SP
    shape = (640, 480)
    p, m = _points(300_000, 33, shape)
    o_pxd, _ = oracle.circle_project(p["xyz"], 0.0, m, shape)
    o_pxd[::97, 2] = 1.5    # clipped
    o_pxd[1::89, :2] = -3.0  # off-screen top-left
    img, dep = SP.rasterize_pix_zbuffer(_t(o_pxd), _t(p["rgb"]), shape, return_depth=True)
    keys, o_img = oracle.zbuffer(o_pxd, p["rgb"], shape)
    np.testing.assert_array_equal(img.cpu().numpy(), o_img)
    # painter: render_pix_sort_depth semantics (nearest wins, later index on ties)
    pimg = SP.rasterize_pix(_t(o_pxd), _t(p["rgb"]), shape).cpu().numpy()
    np.testing.assert_array_equal(pimg, oracle.render_pix_sort_depth(o_pxd, p["rgb"], shape))
    # and through CPU tensors (staged via the GPU)
    cimg = SP.rasterize_pix(torch.from_numpy(o_pxd), torch.from_numpy(p["rgb"]), shape)
    np.testing.assert_array_equal(cimg.numpy(), pimg)


@pytest.mark.parametrize("with_radius", [True, False])
def test_zbuffer_edge_cases(oracle, with_radius):
    """Edge rows (NaN centres, radius 0 on a pixel centre, negative, infinite
    and huge radii, centres far off the image, clipped depths, equal depths)
    among random points: winners and colours bit-exact, for the circle and
    the point z-buffer, at two sizes (the second not a multiple of 64)."""
    import del_splat_dlpack.SplatCircle.torch as SC
    import del_splat_dlpack.SplatPoint.torch as SP
    from _helpers import zbuffer_edge_points
    shape = (203, 117)
    rng = np.random.default_rng(61)
    pxz, rad = zbuffer_edge_points(rng, shape, 5000)
    rgb = rng.random((pxz.shape[0], 3)).astype(np.float32)
    for n in (5000, 1337):
        print(f"zbuffer edge cases n={n} radius={with_radius}", flush=True)
        if with_radius:
            img, dep = SC.rasterize_zbuffer(_t(pxz[:n]), _t(rad[:n]), _t(rgb[:n]), shape)
            keys, o_img = oracle.zbuffer(pxz[:n], rgb[:n], shape, pixrad=rad[:n])
        else:
            img, dep = SP.rasterize_pix_zbuffer(_t(pxz[:n]), _t(rgb[:n]), shape,
                                                return_depth=True)
            keys, o_img = oracle.zbuffer(pxz[:n], rgb[:n], shape)
        np.testing.assert_array_equal(img.cpu().numpy(), o_img)
        win = keys != np.iinfo(np.uint64).max
        o_dep = np.where(win, (keys >> np.uint64(32)).astype(np.uint32).view(np.float32),
                         np.float32(np.finfo(np.float32).max))
        np.testing.assert_array_equal(dep.cpu().numpy(), o_dep)
