"""The batched plan with spherical-harmonic colour (dsplat_mv_set_sh): per
view, the lists and pixels against the oracle's frame whose colours are the
oracle's SH colours from that view's camera centre; the summed adjoint
against the per-view autograd path (sh_to_rgb + render); the training step
in the SH parameter layout, graph against eager; mode switches and errors."""
import numpy as np
import pytest
import torch

from _helpers import check_grad, check_pixels, u32
from del_splat_dlpack import camera, scene

pytestmark = pytest.mark.gpu
DEV = "cuda"
SHAPES = [(160, 120), (200, 200), (256, 144), (120, 90)]


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def _setup(n, seed, nview=4, K=16):
    s = scene.gaussians(n, seed)
    views = scene.sphere_cameras(nview, seed, resolutions=SHAPES)
    sh = np.random.default_rng(seed).normal(0, 0.3, (n, K, 3)).astype(np.float32)
    return s, views, sh


@pytest.mark.parametrize("K,deg,ts,threads", [(16, 3, 16, 0), (9, 2, 32, 512), (4, 1, 8, 0),
                                              (16, 0, 16, 0)])
def test_sh_forward_matches_oracle(oracle, K, deg, ts, threads):
    from del_splat_dlpack.SplatGauss.multiview import MultiViewPlan
    s, views, sh = _setup(8000, 50 + K + deg, K=K)
    n = s["xyz"].shape[0]
    plan = MultiViewPlan(n, views, tile_size=ts, threads=threads)
    plan.set_sh(deg, K)
    # linear scale and opacity: the oracle's projection sees the same floats
    out_rgb, _ = plan.forward(_t(s["xyz"]), _t(s["scale"]), _t(s["quat"]), _t(s["opacity"]),
                              _t(sh), flags=0)
    t2i, i2p = plan.lists()
    alpha_all, nwalk_all, _ = plan.state()
    g_t2i = u32(t2i).astype(np.int64)
    g_i2p = u32(i2p).astype(np.int64)
    imgs = plan.split(out_rgb, 3)
    for v, (m, shape) in enumerate(views):
        print(f"view {v} {shape} K {K} deg {deg} ts {ts}", flush=True)
        cam = camera.camera_center(m).astype(np.float32)
        col = oracle.sh_to_rgb(s["xyz"], sh, cam, deg)
        o_pxd, o_con, o_aabb, _ = oracle.gauss_project(s["xyz"], s["quat"], s["scale"], m, shape)
        o_t2i, o_i2p = oracle.tile2pnt_gauss(o_pxd, o_aabb, shape, ts)
        a, b = plan.tile_off[v], plan.tile_off[v + 1]
        np.testing.assert_array_equal(g_t2i[a:b + 1] - g_t2i[a], o_t2i)
        np.testing.assert_array_equal(g_i2p[g_t2i[a]:g_t2i[b]] - v * n, o_i2p)
        o_img, o_alpha, o_nwalk, _ = oracle.gauss_rasterize(shape, o_pxd, o_con, o_aabb,
                                                            s["opacity"], col, ts, o_t2i, o_i2p)
        w, h = shape
        pa, pb = plan.pix_off[v], plan.pix_off[v + 1]
        check_pixels(imgs[v].cpu().numpy(), alpha_all[pa:pb].cpu().numpy().reshape(h, w),
                     nwalk_all[pa:pb].cpu().numpy().reshape(h, w), o_img, o_alpha, o_nwalk)
    plan.close()


@pytest.mark.parametrize("ts,threads", [(16, 0), (32, 512)])
def test_sh_backward_matches_sum_of_per_view(ts, threads):
    import del_splat_dlpack.SplatGauss.torch as SG
    from del_splat_dlpack.SplatGauss.multiview import MultiViewPlan
    s, views, sh = _setup(8000, 61)
    n = s["xyz"].shape[0]
    base = {k: _t(s[k]) for k in ("xyz", "log_scale", "quat", "opacity")}
    plan = MultiViewPlan(n, views, tile_size=ts, threads=threads)
    plan.set_sh(3, 16)
    targets = [_t(scene.target_image(shape, 70 + v)) for v, (_, shape) in enumerate(views)]
    target_flat = torch.cat([t.reshape(-1) for t in targets])
    out_rgb, _ = plan.forward(base["xyz"], base["log_scale"], base["quat"], base["opacity"],
                              _t(sh))
    g = plan.backward(base["xyz"], base["log_scale"], base["quat"], base["opacity"],
                      out_rgb=out_rgb, target=target_flat)
    assert g["rgb"].shape == (n, 16, 3)
    leaf = {k: v.clone().requires_grad_(True) for k, v in base.items()}
    sh_leaf = _t(sh).requires_grad_(True)
    total = 0.0
    for v, (m, shape) in enumerate(views):
        cam = camera.camera_center(m).astype(np.float32)
        rgb_v = SG.sh_to_rgb(leaf["xyz"], sh_leaf, cam, 3)
        img, _ = SG.render(leaf["xyz"], leaf["log_scale"], leaf["quat"], leaf["opacity"], rgb_v,
                           _t(m), shape)
        total = total + ((img - targets[v]) ** 2).sum()
    total.backward()
    pairs = dict(xyz=leaf["xyz"].grad, scale=leaf["log_scale"].grad, quat=leaf["quat"].grad,
                 opacity=leaf["opacity"].grad, rgb=sh_leaf.grad)
    for k, ref in pairs.items():
        check_grad(g[k].cpu().numpy(), ref.cpu().numpy(), k, abs_frac=1e-4, max_rel_tol=1e-3)
    plan.close()


def test_sh_train_step_graph_matches_eager():
    from del_splat_dlpack.SplatGauss.multiview import MultiViewPlan
    n = 5998  # even (the quaternion block is 16-byte aligned), 11N = 2 mod 4: padded
    s, views, sh = _setup(n, 62, nview=3)
    logit = np.log(s["opacity"] / (1 - s["opacity"])).astype(np.float32)
    targets = torch.cat([_t(scene.target_image(shape, 80 + v)).reshape(-1)
                         for v, (_, shape) in enumerate(views)])

    def run(graph):
        plan = MultiViewPlan(n, views, tile_size=16)
        plan.set_sh(3, 16)
        lay = plan.param_layout()
        assert lay["colour"][0] % 4 == 0 and lay["colour"][0] >= 11 * n
        params = plan.pack_params(s["xyz"], s["log_scale"], s["quat"], logit, sh)
        # NaN everywhere the adjoint must write, and in the padding, which
        # the step itself zeroes
        grads = torch.full_like(params, float("nan"))
        m1 = torch.zeros_like(params)
        m2 = torch.zeros_like(params)
        step = torch.zeros(1, device=DEV, dtype=torch.int32)
        out = plan.new_image(3)
        if graph:
            plan.capture_train_step(params, grads, m1, m2, step, targets, out, lr=1e-3)
        losses = []
        for _ in range(4):
            if graph:
                plan.replay()
            else:
                plan.train_step(params, grads, m1, m2, step, targets, out, lr=1e-3)
            losses.append(float(((out - targets) ** 2).sum()))
        torch.cuda.synchronize()
        plan.close()
        c0 = lay["colour"][0]
        assert not grads[11 * n:c0].any() and not m1[11 * n:c0].any()
        assert not m2[11 * n:c0].any()
        assert torch.isfinite(grads).all() and torch.isfinite(params).all()
        return params.cpu().numpy(), losses, int(step.item()), lay

    pe, le, se, lay = run(False)
    pg, lg, sg, _ = run(True)
    assert se == sg == 4
    assert le[-1] < le[0]
    np.testing.assert_allclose(lg, le, rtol=1e-4)
    assert np.abs(pe - pg).max() <= 1e-4 * max(1.0, np.abs(pe).max())
    c0, c1 = lay["colour"]
    assert not pe[11 * n:c0].any()  # padding untouched
    assert np.abs(pe[c0:c1] - sh.reshape(-1)).max() > 0  # the coefficients trained


def test_sh_mode_switch_and_errors():
    from del_splat_dlpack import _native
    from del_splat_dlpack.SplatGauss.multiview import MultiViewPlan
    s, views, sh = _setup(3000, 63, nview=2)
    n = 3000
    args = [_t(s[k]) for k in ("xyz", "log_scale", "quat", "opacity")]
    plan = MultiViewPlan(n, views, tile_size=16)
    ref, _ = plan.forward(*args, _t(s["rgb"]))
    ref = ref.clone()
    plan.set_sh(2, 16)
    plan.forward(*args, _t(sh))
    plan.set_sh(0, 0)  # back to one rgb per Gaussian
    again, _ = plan.forward(*args, _t(s["rgb"]))
    np.testing.assert_array_equal(again.cpu().numpy(), ref.cpu().numpy())
    with pytest.raises(ValueError):
        plan.set_sh(3, 9)
    with pytest.raises(ValueError):
        plan.set_sh(1, 5)
    # an orthographic view has no camera centre
    ortho = np.eye(4, dtype=np.float32)
    ortho[2, 2] = -0.1
    plan.set_views([(camera.to_col_major(ortho), (64, 64))] + views[1:])
    plan.set_sh(3, 16)
    with pytest.raises(ValueError, match="perspective"):
        plan.forward(*args, _t(sh))
    # backward without an SH forward under the current mode
    plan.set_views(views)
    plan.set_sh(0, 0)
    out, _ = plan.forward(*args, _t(s["rgb"]))
    plan.set_sh(3, 16)
    with pytest.raises(ValueError, match="no forward over the current views"):
        plan.backward(*args, out_rgb=out, target=torch.zeros_like(out))
    plan.close()
    assert _native.DSPLAT_GRAD_OVERWRITE  # the module's constants stay importable
