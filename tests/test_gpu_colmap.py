"""A COLMAP reconstruction driving the renderer on the GPU: the readers'
cameras and images become the batched plan's uneven views (three pinhole
cameras of different sizes and focal lengths), the points become the
initial Gaussians, and every view is checked against the oracle's frame of
the same inputs (binning bit for bit, pixels as in the single-view tests).
"""
import numpy as np
import pytest
import torch

from _colmap_files import synthetic_reconstruction
from _helpers import check_pixels, u32
from del_splat_dlpack import colmap

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def _load(directory, n_points=6000, n_images=6, seed=11):
    synthetic_reconstruction(str(directory), n_points=n_points, n_images=n_images, seed=seed)
    return (colmap.read_cameras(directory / "cameras.bin"),
            colmap.read_images(directory / "images.bin"),
            colmap.read_points3d(directory / "points3D.bin"))


def _gauss(g):
    return [_t(g[k]) for k in ("xyz", "scale", "quat", "opacity", "rgb")]


def test_colmap_projection_lands_on_keypoints(tmp_path):
    """The kernel's projection of each point through the bridged matrix is
    the keypoint COLMAP recorded for it (written as the exact pinhole
    projection), to float32 rounding."""
    import del_splat_dlpack.SplatGauss.torch as SG
    cams, images, pts = _load(tmp_path)
    g = colmap.gaussians_from_points(pts)
    xyz, sc, q, op, rgb = _gauss(g)
    row = pts.index_of()
    for (m16, shape), im in zip(colmap.views(cams, images), images):
        pxd, _, _, _ = SG.project_fwd(xyz, sc, q, op, rgb, _t(m16), shape, log_scale=False)
        sel = np.nonzero(im.vtx2id >= 0)[0]
        rows = [row[int(p)] for p in im.vtx2id[sel]]
        print(f"image {im.name} {shape}: {len(sel)} keypoints", flush=True)
        np.testing.assert_allclose(pxd.cpu().numpy()[rows, :2].astype(np.float64),
                                   im.vtx2xy.reshape(-1, 2)[sel], rtol=0, atol=5e-4)


@pytest.mark.parametrize("ts,threads", [(16, 0), (32, 512)])
def test_colmap_multiview_forward_matches_oracle(tmp_path, oracle, ts, threads):
    from del_splat_dlpack.SplatGauss.multiview import MultiViewPlan
    cams, images, pts = _load(tmp_path)
    g = colmap.gaussians_from_points(pts)
    n = len(pts)
    views = colmap.views(cams, images, 0.1, 20.0)
    assert len({shape for _, shape in views}) == 3
    plan = MultiViewPlan(n, views, tile_size=ts, threads=threads)
    out_rgb, _ = plan.forward(*_gauss(g), flags=0)
    t2i, i2p = plan.lists()
    alpha_all, nwalk_all, _ = plan.state()
    g_t2i = u32(t2i).astype(np.int64)
    g_i2p = u32(i2p).astype(np.int64)
    imgs = plan.split(out_rgb, 3)
    for v, (m16, shape) in enumerate(views):
        print(f"view {v} {shape} ts {ts} threads {threads}", flush=True)
        o_pxd, o_con, o_aabb, _ = oracle.gauss_project(g["xyz"], g["quat"], g["scale"], m16,
                                                       shape)
        o_t2i, o_i2p = oracle.tile2pnt_gauss(o_pxd, o_aabb, shape, ts)
        a, b = plan.tile_off[v], plan.tile_off[v + 1]
        np.testing.assert_array_equal(g_t2i[a:b + 1] - g_t2i[a], o_t2i)
        np.testing.assert_array_equal(g_i2p[g_t2i[a]:g_t2i[b]] - v * n, o_i2p)
        o_img, o_alpha, o_nwalk, _ = oracle.gauss_rasterize(shape, o_pxd, o_con, o_aabb,
                                                            g["opacity"], g["rgb"], ts, o_t2i,
                                                            o_i2p)
        w, h = shape
        pa, pb = plan.pix_off[v], plan.pix_off[v + 1]
        check_pixels(imgs[v].cpu().numpy(), alpha_all[pa:pb].cpu().numpy().reshape(h, w),
                     nwalk_all[pa:pb].cpu().numpy().reshape(h, w), o_img, o_alpha, o_nwalk)
        assert o_alpha.max() > 0.5  # the point cloud covers the view
    plan.close()
