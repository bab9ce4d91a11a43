"""Batched multi-view engine against the per-view path and the oracle.

Per view, the batched binning must equal the single-view binning bit for bit
(the global lists restricted to the view's tile range, items mapped back to
Gaussian indices), images must equal the single-view raster (same kernels,
same order) and the summed gradients must agree with the sum of per-view
gradients within the reordering of float atomics. Graph replays are checked
against eager runs, several times, with eager work in between.
"""
import numpy as np
import pytest
import torch

from _helpers import check_grad, u32
from del_splat_dlpack import _native, scene

pytestmark = pytest.mark.gpu
DEV = "cuda"
SHAPES = [(160, 120), (200, 200), (256, 144), (120, 90)]


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def _setup(n=20000, seed=40, nview=4, radius=1.0):
    s = scene.gaussians(n, seed, radius=radius)
    cams = scene.sphere_cameras(nview, seed, resolutions=SHAPES)
    views = [(m, shape) for m, shape in cams]
    return s, views


@pytest.mark.parametrize("ts,threads", [(16, 0), (8, 0), (32, 256), (16, 64)])
def test_multiview_forward_matches_per_view(oracle, ts, threads):
    import del_splat_dlpack.SplatGauss.torch as SG
    from del_splat_dlpack.SplatGauss.multiview import MultiViewPlan
    s, views = _setup()
    n = s["xyz"].shape[0]
    xyz, ls, q, op, rgb = (_t(s[k]) for k in ("xyz", "log_scale", "quat", "opacity", "rgb"))
    plan = MultiViewPlan(n, views, tile_size=ts, threads=threads)
    out_rgb, out_alpha = plan.forward(xyz, ls, q, op, rgb)
    t2i, i2p = plan.lists()
    alpha_all, nwalk_all, tfinal_all = plan.state()
    g_t2i = u32(t2i).astype(np.int64)
    g_i2p = u32(i2p).astype(np.int64)
    imgs = plan.split(out_rgb, 3)
    for v, (m, shape) in enumerate(views):
        a, b = plan.tile_off[v], plan.tile_off[v + 1]
        loc_t2i = g_t2i[a:b + 1] - g_t2i[a]
        loc_i2p = g_i2p[g_t2i[a]:g_t2i[b]] - v * n
        pxd, conic, aabb, splat2 = SG.project_fwd(xyz, ls, q, op, rgb, _t(m), shape)
        s_t2i, s_i2p = SG.tile2idx_idx2pnt(shape, ts, pxd, aabb)
        np.testing.assert_array_equal(loc_t2i, u32(s_t2i))
        np.testing.assert_array_equal(loc_i2p, u32(s_i2p))
        o_t2i, o_i2p = oracle.tile2pnt_gauss(pxd.cpu().numpy(), aabb.cpu().numpy(), shape, ts)
        np.testing.assert_array_equal(loc_i2p, o_i2p)
        img, alpha, nwalk, tfinal = SG.rasterize_fwd(shape, ts, splat2, s_t2i, s_i2p, threads)
        np.testing.assert_array_equal(imgs[v].cpu().numpy(), img.cpu().numpy())
        pa, pb = plan.pix_off[v], plan.pix_off[v + 1]
        np.testing.assert_array_equal(nwalk_all[pa:pb].cpu().numpy(), nwalk.reshape(-1).cpu().numpy())
        np.testing.assert_array_equal(alpha_all[pa:pb].cpu().numpy(), alpha.reshape(-1).cpu().numpy())
        np.testing.assert_array_equal(out_alpha[pa:pb].cpu().numpy(), alpha.reshape(-1).cpu().numpy())
    plan.close()


@pytest.mark.parametrize("n", [1 << 18, 300000])
def test_views_depth_sorted_in_segments_match_per_view(n):
    """From 2^18 Gaussians the views are depth-sorted each in its own
    segment, padded to whole segments (300000: 27680 padding entries per
    view; 2^18: none). The tile lists and images must still equal the
    single-view binning and raster per view, bit for bit."""
    import del_splat_dlpack.SplatGauss.torch as SG
    from del_splat_dlpack.SplatGauss.multiview import MultiViewPlan
    s, views = _setup(n=n, seed=44, nview=3)
    xyz, ls, q, op, rgb = (_t(s[k]) for k in ("xyz", "log_scale", "quat", "opacity", "rgb"))
    ts = 16
    plan = MultiViewPlan(n, views, tile_size=ts, threads=0)
    out_rgb, _ = plan.forward(xyz, ls, q, op, rgb)
    t2i, i2p = plan.lists()
    g_t2i = u32(t2i).astype(np.int64)
    g_i2p = u32(i2p).astype(np.int64)
    imgs = plan.split(out_rgb, 3)
    for v, (m, shape) in enumerate(views):
        a, b = plan.tile_off[v], plan.tile_off[v + 1]
        pxd, conic, aabb, splat2 = SG.project_fwd(xyz, ls, q, op, rgb, _t(m), shape)
        s_t2i, s_i2p = SG.tile2idx_idx2pnt(shape, ts, pxd, aabb)
        assert g_t2i[b] - g_t2i[a] > 0
        np.testing.assert_array_equal(g_t2i[a:b + 1] - g_t2i[a], u32(s_t2i))
        np.testing.assert_array_equal(g_i2p[g_t2i[a]:g_t2i[b]] - v * n, u32(s_i2p))
        img, _, _, _ = SG.rasterize_fwd(shape, ts, splat2, s_t2i, s_i2p, 0)
        np.testing.assert_array_equal(imgs[v].cpu().numpy(), img.cpu().numpy())
    plan.close()


@pytest.mark.parametrize("ts,threads", [(16, 0), (32, 512)])
def test_multiview_backward_matches_sum_of_per_view(ts, threads):
    import del_splat_dlpack.SplatGauss.torch as SG
    from del_splat_dlpack.SplatGauss.multiview import MultiViewPlan
    s, views = _setup(n=8000, seed=41)
    n = s["xyz"].shape[0]
    base = {k: _t(s[k]) for k in ("xyz", "log_scale", "quat", "opacity", "rgb")}
    plan = MultiViewPlan(n, views, tile_size=ts, threads=threads)
    targets = [_t(scene.target_image(shape, 50 + v)) for v, (_, shape) in enumerate(views)]
    target_flat = torch.cat([t.reshape(-1) for t in targets])
    out_rgb, _ = plan.forward(base["xyz"], base["log_scale"], base["quat"], base["opacity"],
                              base["rgb"])
    g = plan.backward(base["xyz"], base["log_scale"], base["quat"], base["opacity"],
                      out_rgb=out_rgb, target=target_flat)
    # per-view autograd path, summed
    leaf = {k: v.clone().requires_grad_(True) for k, v in base.items()}
    total = 0.0
    for v, (m, shape) in enumerate(views):
        img, _ = SG.render(leaf["xyz"], leaf["log_scale"], leaf["quat"], leaf["opacity"],
                           leaf["rgb"], _t(m), shape)
        total = total + ((img - targets[v]) ** 2).sum()
    total.backward()
    names = dict(xyz="xyz", scale="log_scale", quat="quat", opacity="opacity", rgb="rgb")
    for k, lk in names.items():
        check_grad(g[k].cpu().numpy(), leaf[lk].grad.cpu().numpy(), k, abs_frac=1e-4,
                   max_rel_tol=1e-3)
    plan.close()


def test_graph_replay_forward():
    from del_splat_dlpack.SplatGauss.multiview import MultiViewPlan
    s, views = _setup(n=10000, seed=42)
    n = s["xyz"].shape[0]
    xyz, ls, q, op, rgb = (_t(s[k]) for k in ("xyz", "log_scale", "quat", "opacity", "rgb"))
    plan = MultiViewPlan(n, views, tile_size=16)
    eager, eager_a = plan.forward(xyz, ls, q, op, rgb)
    eager, eager_a = eager.clone(), eager_a.clone()
    out_rgb = plan.new_image(3)
    out_a = plan.new_image(1)
    plan.capture_forward(xyz, ls, q, op, rgb, out_rgb, out_a)
    for rep in range(3):
        out_rgb.fill_(-1.0)
        plan.replay()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(out_rgb.cpu().numpy(), eager.cpu().numpy())
        np.testing.assert_array_equal(plan.state()[0].cpu().numpy(), eager_a.cpu().numpy())
        # eager work between replays (a different frame into other buffers)
        plan.forward(xyz * 1.01, ls, q, op, rgb)
        torch.cuda.synchronize()
    plan.close()


def test_set_views_is_stream_ordered():
    """set_views does not wait for the GPU: frames queued before it render the
    old views, frames after it the new ones, with no synchronisation in
    between (a slow frame first, so the table upload really overlaps it)."""
    from del_splat_dlpack.SplatGauss.multiview import MultiViewPlan
    s, views = _setup(n=30000, seed=47)
    n = s["xyz"].shape[0]
    args = [_t(s[k]) for k in ("xyz", "log_scale", "quat", "opacity", "rgb")]
    other = scene.sphere_cameras(3, 48, resolutions=SHAPES)
    refs = []
    for vs in (views, other):
        p = MultiViewPlan(n, vs, tile_size=16)
        refs.append(p.forward(*args)[0].cpu().numpy())
        p.close()
    plan = MultiViewPlan(n, views, tile_size=16)
    torch.cuda.synchronize()
    outs = []
    for vs in (views, other, views, other):
        plan.set_views(vs)
        outs.append(plan.forward(*args)[0])
    torch.cuda.synchronize()
    for k, o in enumerate(outs):
        np.testing.assert_array_equal(o.cpu().numpy(), refs[k % 2])
    plan.close()


def test_changing_views_or_colour_drops_the_graph():
    """A captured graph is sized for its views; after set_views or set_sh a
    replay raises instead of running stale launches, and a new capture
    works."""
    from del_splat_dlpack.SplatGauss.multiview import MultiViewPlan
    s, views = _setup(n=4000, seed=43)
    n = s["xyz"].shape[0]
    args = [_t(s[k]) for k in ("xyz", "log_scale", "quat", "opacity", "rgb")]
    plan = MultiViewPlan(n, views, tile_size=16)
    out_rgb, out_a = plan.new_image(3), plan.new_image(1)
    plan.capture_forward(*args, out_rgb, out_a)
    plan.replay()
    plan.set_views(views[::-1])  # other image offsets than the captured launches
    with pytest.raises(ValueError, match="nothing captured"):
        plan.replay()
    big_rgb, big_a = plan.new_image(3), plan.new_image(1)
    plan.capture_forward(*args, big_rgb, big_a)
    plan.replay()
    torch.cuda.synchronize()
    eager, _ = plan.forward(*args)
    np.testing.assert_array_equal(big_rgb.cpu().numpy(), eager.cpu().numpy())
    plan.set_sh(0, 0)
    with pytest.raises(ValueError, match="nothing captured"):
        plan.replay()
    plan.close()


def test_backward_needs_a_forward_over_the_current_views():
    """The adjoint reads the last forward's lists, walks and transmittances:
    on a new plan, after set_views or set_sh, or after a capture that was not
    replayed, the C ABI refuses it before launching anything (and so does the
    Python plan); a forward, a training step or a replay makes it valid."""
    from del_splat_dlpack.SplatGauss.multiview import MultiViewPlan
    s, views = _setup(n=3000, seed=46)
    n = s["xyz"].shape[0]
    args = [_t(s[k]) for k in ("xyz", "log_scale", "quat", "opacity", "rgb")]
    plan = MultiViewPlan(n, views, tile_size=16)
    tgt = plan.new_image(3).zero_()
    g = [torch.empty((n, c), device=DEV) for c in (3, 3, 4, 1, 3)]

    def c_backward(out):
        _native.call("dsplat_mv_backward", plan._plan, torch.cuda.current_stream().cuda_stream,
                     *[a.data_ptr() for a in args[:4]], out.data_ptr(), None, tgt.data_ptr(),
                     None, *[x.data_ptr() for x in g], _native.DSPLAT_SCALE_IS_LOG)

    out = plan.new_image(3)
    with pytest.raises(ValueError, match="no forward"):
        c_backward(out)
    with pytest.raises(ValueError, match="no forward"):
        plan.backward(*args[:4], out_rgb=out, target=tgt)
    out, _ = plan.forward(*args)
    c_backward(out)
    plan.set_views(views[::-1])
    with pytest.raises(ValueError, match="no forward"):
        c_backward(out)
    o_rgb, o_a = plan.new_image(3), plan.new_image(1)
    plan.capture_forward(*args, o_rgb, o_a)
    with pytest.raises(ValueError, match="no forward"):
        c_backward(o_rgb)  # captured, not run
    plan.replay()
    c_backward(o_rgb)
    plan.set_sh(0, 0)
    with pytest.raises(ValueError, match="no forward"):
        plan.backward(*args[:4], out_rgb=o_rgb, target=tgt)
    torch.cuda.synchronize()
    plan.close()


def test_capacity_overflow_and_growth():
    from del_splat_dlpack.SplatGauss.multiview import MultiViewPlan, ensure_capacity
    s, views = _setup(n=5000, seed=43)
    n = s["xyz"].shape[0]
    args = [_t(s[k]) for k in ("xyz", "log_scale", "quat", "opacity", "rgb")]
    small = MultiViewPlan(n, views, tile_size=16, pair_capacity=100)
    small.forward(*args)
    k, over = small.counters()
    assert over and k > 100
    small.close()
    plan, _ = ensure_capacity(lambda cap: MultiViewPlan(n, views, 16, pair_capacity=cap or 100),
                              lambda p: p.forward(*args))
    k2, over2 = plan.counters()
    assert not over2 and k2 == k
    plan.close()


def test_adam_matches_torch():
    rng = np.random.default_rng(44)
    p0 = rng.standard_normal(100_000).astype(np.float32)
    p = _t(p0)
    m1 = torch.zeros_like(p)
    m2 = torch.zeros_like(p)
    step = torch.zeros(1, device=DEV, dtype=torch.int32)
    ref = torch.tensor(p0, device=DEV, requires_grad=True)
    opt = torch.optim.Adam([ref], lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    for k in range(5):
        gk = _t(rng.standard_normal(100_000).astype(np.float32))
        _native.call("dsplat_adam_step", torch.cuda.current_stream().cuda_stream, p.numel(),
                     p.data_ptr(), gk.data_ptr(), m1.data_ptr(), m2.data_ptr(), 1e-2, 0.9, 0.999,
                     1e-8, step.data_ptr())
        ref.grad = gk.clone()
        opt.step()
    assert int(step.item()) == 5
    np.testing.assert_allclose(p.cpu().numpy(), ref.detach().cpu().numpy(), rtol=1e-5, atol=1e-6)


def test_train_step_graph_matches_eager():
    from del_splat_dlpack.SplatGauss.multiview import MultiViewPlan
    s, views = _setup(n=6000, seed=45, nview=3)
    n = s["xyz"].shape[0]
    logit = np.log(s["opacity"] / (1 - s["opacity"])).astype(np.float32)
    flat = np.concatenate([s["xyz"].reshape(-1), s["log_scale"].reshape(-1),
                           s["quat"].reshape(-1), logit, s["rgb"].reshape(-1)])
    targets = torch.cat([_t(scene.target_image(shape, 60 + v)).reshape(-1)
                         for v, (_, shape) in enumerate(views)])

    def run(graph):
        plan = MultiViewPlan(n, views, tile_size=16)
        params = _t(flat)
        grads = torch.zeros_like(params)
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
        return params.cpu().numpy(), losses, int(step.item())

    pe, le, se = run(False)
    pg, lg, sg = run(True)
    assert se == sg == 4
    assert le[-1] < le[0]  # the loss goes down
    np.testing.assert_allclose(lg, le, rtol=1e-4)
    assert np.abs(pe - pg).max() <= 1e-4 * max(1.0, np.abs(pe).max())
