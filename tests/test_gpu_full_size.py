"""Parity at the full sizes of the workloads the framework is measured on
(the 1M-Gaussian 1080p frame and the 2M-point z-buffer have their own
full-size tests next to their kernels):

* the plain forward of 100k Gaussians at 512x512 with 16x16 tiles, RGB and
  alpha, against the oracle (projection and lists bit-exact, pixels as in the
  single-view raster tests);
* 2M Gaussians seen by 16 uneven views (640x480, 1280x720, 1920x1080,
  800x800) in one batched forward, per view bit-exact against the
  single-view path (lists and images), and a graph replay equal to it;
* 3M Gaussians with 8 uneven views: the batched forward + L2 adjoint
  against the sum of the per-view single-view adjoints;
* the one-call projection and binning the headline frame times (1M
  Gaussians, 1920x1080, 32x32 tiles) and the 100k-Gaussian forward,
  bit-exact against the projection followed by the binning.

The single-view path is itself checked against the oracle; at these sizes
the oracle would take minutes per view, so the batched cases compare
against it."""
import numpy as np
import pytest
import torch

from _helpers import check_grad, check_pixels, u32
from del_splat_dlpack import scene

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def test_plain_forward_100k_512(oracle):
    import del_splat_dlpack.SplatGauss.torch as SG
    shape, ts = (512, 512), 16
    s = scene.gaussians(100_000, 0)
    m = scene.front_camera(shape)
    xyz, sc, q, op, rgb = (_t(s[k]) for k in ("xyz", "scale", "quat", "opacity", "rgb"))
    pxd, conic, aabb, splat2 = SG.project_fwd(xyz, sc, q, op, rgb, _t(m), shape, log_scale=False)
    t2i, i2p = SG.tile2idx_idx2pnt(shape, ts, pxd, aabb)
    img, alpha, nwalk, _ = SG.rasterize_fwd(shape, ts, splat2, t2i, i2p)
    o_pxd, o_conic, o_aabb, _ = oracle.gauss_project(s["xyz"], s["quat"], s["scale"], m, shape)
    np.testing.assert_array_equal(pxd.cpu().numpy(), o_pxd)
    np.testing.assert_array_equal(conic.cpu().numpy(), o_conic)
    np.testing.assert_array_equal(aabb.cpu().numpy(), o_aabb)
    o_t2i, o_i2p = oracle.tile2pnt_gauss(o_pxd, o_aabb, shape, ts)
    np.testing.assert_array_equal(u32(t2i), o_t2i)
    np.testing.assert_array_equal(u32(i2p), o_i2p)
    o_img, o_alpha, o_nwalk, _ = oracle.gauss_rasterize(shape, o_pxd, o_conic, o_aabb,
                                                        s["opacity"], s["rgb"], ts, o_t2i, o_i2p)
    own = oracle.self_flips(shape, o_pxd, o_conic, o_aabb, s["opacity"], s["rgb"], ts, o_t2i,
                            o_i2p, o_nwalk)
    flips = check_pixels(img.cpu().numpy(), alpha.cpu().numpy(), nwalk.cpu().numpy(), o_img,
                         o_alpha, o_nwalk, self_flip=own)
    print(f"pairs {len(o_i2p)}, termination flips {flips}", flush=True)
    assert o_alpha.max() > 0.99
    # the plain forward the workload times (RGB + alpha, no adjoint state), at
    # every workgroup shape of 16x16 tiles: the same pixels, bit for bit
    for thr in (64, 128, 256):
        p_img, p_alpha, _, _ = SG.rasterize_fwd(shape, ts, splat2, t2i, i2p, thr, state=False)
        np.testing.assert_array_equal(p_img.cpu().numpy(), img.cpu().numpy())
        np.testing.assert_array_equal(p_alpha.cpu().numpy(), alpha.cpu().numpy())


def test_batched_forward_2m_16_views():
    import del_splat_dlpack.SplatGauss.torch as SG
    from del_splat_dlpack.SplatGauss.multiview import MultiViewPlan, ensure_capacity
    s = scene.gaussians(2_000_000, 3)
    views = scene.sphere_cameras(16, 3)
    assert len({shape for _, shape in views}) >= 3
    n = 2_000_000
    args = [_t(s[k]) for k in ("xyz", "log_scale", "quat", "opacity", "rgb")]
    for ts, th in ((32, 512), (16, 0)):
        print(f"tile {ts} threads {th}", flush=True)
        plan, (out_rgb, _) = ensure_capacity(
            lambda cap: MultiViewPlan(n, views, ts, th, pair_capacity=cap),
            lambda p: p.forward(*args))
        t2i, i2p = plan.lists()
        g_t2i = u32(t2i).astype(np.int64)
        g_i2p = u32(i2p).astype(np.int64)
        imgs = [x.cpu().numpy() for x in plan.split(out_rgb, 3)]
        for v, (m, shape) in enumerate(views):
            pxd, _, aabb, splat2 = SG.project_fwd(*args, _t(m), shape)
            s_t2i, s_i2p = SG.tile2idx_idx2pnt(shape, ts, pxd, aabb)
            a, b = plan.tile_off[v], plan.tile_off[v + 1]
            np.testing.assert_array_equal(g_t2i[a:b + 1] - g_t2i[a], u32(s_t2i))
            np.testing.assert_array_equal(g_i2p[g_t2i[a]:g_t2i[b]] - v * n, u32(s_i2p))
            img = SG.rasterize_fwd(shape, ts, splat2, s_t2i, s_i2p, th)[0]
            np.testing.assert_array_equal(imgs[v], img.cpu().numpy())
        # the captured graph replays the same frame
        g_rgb, g_a = plan.new_image(3), plan.new_image(1)
        plan.capture_forward(*args, g_rgb, g_a)
        plan.replay()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(g_rgb.cpu().numpy(), out_rgb.cpu().numpy())
        print(f"pairs {len(g_i2p)}", flush=True)
        plan.close()
        del plan, out_rgb, g_rgb, g_a
        torch.cuda.empty_cache()


def test_batched_adjoint_3m_8_views():
    import del_splat_dlpack.SplatGauss.torch as SG
    from del_splat_dlpack.SplatGauss.multiview import MultiViewPlan, ensure_capacity
    s = scene.gaussians(3_000_000, 4)
    views = scene.sphere_cameras(8, 4)
    n = 3_000_000
    base = {k: _t(s[k]) for k in ("xyz", "log_scale", "quat", "opacity", "rgb")}
    targets = [_t(scene.target_image(shape, 90 + v)) for v, (_, shape) in enumerate(views)]
    tflat = torch.cat([t.reshape(-1) for t in targets])
    args = [base[k] for k in ("xyz", "log_scale", "quat", "opacity", "rgb")]
    plan, (out_rgb, _) = ensure_capacity(
        lambda cap: MultiViewPlan(n, views, 32, 512, pair_capacity=cap),
        lambda p: p.forward(*args))
    g = plan.backward(*args[:4], out_rgb=out_rgb, target=tflat)
    plan.close()
    leaf = {k: v.clone().requires_grad_(True) for k, v in base.items()}
    for v, (m, shape) in enumerate(views):
        img, _ = SG.render(leaf["xyz"], leaf["log_scale"], leaf["quat"], leaf["opacity"],
                           leaf["rgb"], _t(m), shape, tile_size=32, threads=512)
        ((img - targets[v]) ** 2).sum().backward()
    names = dict(xyz="xyz", scale="log_scale", quat="quat", opacity="opacity", rgb="rgb")
    for k, lk in names.items():
        check_grad(g[k].cpu().numpy(), leaf[lk].grad.cpu().numpy(), k, max_rel_tol=1e-3)


@pytest.mark.parametrize("n,shape,ts,seed,radius", [(1_000_000, (1920, 1080), 32, 2, 2.0),
                                                    (100_000, (512, 512), 16, 0, 1.0)])
def test_project_bin_full_size(n, shape, ts, seed, radius):
    import del_splat_dlpack.SplatGauss.torch as SG
    s = scene.gaussians(n, seed, radius=radius)
    m16 = _t(scene.front_camera(shape))
    xyz, ls, q, op, rgb = (_t(s[k]) for k in ("xyz", "log_scale", "quat", "opacity", "rgb"))
    pxd, _, aabb, splat2 = SG.project_fwd(xyz, ls, q, op, rgb, m16, shape, with_conic=False)
    k = SG.bin_gauss(shape, ts, pxd, aabb)[1].numel()
    cap = k + k // 8 + 65536
    k_ref = torch.zeros(1, device=DEV, dtype=torch.int32)
    t2i, i2p, _ = SG.bin_gauss_capacity(shape, ts, pxd, aabb, cap, k_ref)
    rec, t2i_b, i2p_b, k_b, _, _ = SG.project_bin(xyz, ls, q, op, rgb, m16, shape, ts, cap)
    torch.cuda.synchronize()
    assert int(k_b.item()) == int(k_ref.item()) == k
    assert torch.equal(rec, splat2)
    assert torch.equal(t2i_b, t2i)
    assert torch.equal(i2p_b[:k], i2p[:k])
