"""End to end on the widened stack: a COLMAP model's cameras and points
drive the batched plan with spherical-harmonic colour. Target images are
rendered from ground-truth Gaussians whose colours depend on the view
(degree-1 SH); a perturbed copy, started from points3D the way 3D Gaussian
splatting starts (gaussians_from_points), is trained with the fused forward
+ L2 adjoint + Adam step (graph-captured), and the loss over every
registered view falls well below its start."""
import numpy as np
import pytest
import torch

from _colmap_files import synthetic_reconstruction
from del_splat_dlpack import _native, colmap

pytestmark = pytest.mark.gpu
DEV = "cuda"
C0 = 0.28209479177387814


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def test_train_from_colmap_with_sh(tmp_path):
    from del_splat_dlpack.SplatGauss.multiview import MultiViewPlan
    synthetic_reconstruction(str(tmp_path), n_points=3000, n_images=6, seed=21)
    cams = colmap.read_cameras(tmp_path / "cameras.bin")
    images = colmap.read_images(tmp_path / "images.bin")
    pts = colmap.read_points3d(tmp_path / "points3D.bin")
    views = colmap.views(cams, images, 0.1, 20.0)
    g = colmap.gaussians_from_points(pts)
    n = len(pts)
    rng = np.random.default_rng(22)
    # ground truth: point colours as the DC term plus a view-dependent band 1
    sh_gt = np.zeros((n, 4, 3), np.float32)
    sh_gt[:, 0] = (g["rgb"] - 0.5) / C0
    sh_gt[:, 1:] = rng.normal(0, 0.3, (n, 3, 3))
    logit_gt = np.full(n, 1.5, np.float32)
    plan = MultiViewPlan(n, views, tile_size=16)
    plan.set_sh(1, 4)
    flags = _native.DSPLAT_SCALE_IS_LOG | _native.DSPLAT_OPACITY_IS_LOGIT
    targets, _ = plan.forward(_t(g["xyz"]), _t(g["log_scale"]), _t(g["quat"]), _t(logit_gt),
                              _t(sh_gt), flags=flags)
    targets = targets.clone()
    assert float(targets.max()) > 0.5  # the views see the cloud
    # start: grey, view-independent, fainter and larger, slightly moved
    xyz0 = g["xyz"] + rng.normal(0, 0.01, g["xyz"].shape).astype(np.float32)
    params = plan.pack_params(xyz0, g["log_scale"] + 0.3, g["quat"], np.zeros(n, np.float32),
                              np.zeros((n, 4, 3), np.float32))
    grads = torch.zeros_like(params)
    m1 = torch.zeros_like(params)
    m2 = torch.zeros_like(params)
    step = torch.zeros(1, device=DEV, dtype=torch.int32)
    out = plan.new_image(3)
    plan.capture_train_step(params, grads, m1, m2, step, targets, out, lr=2e-2)
    losses = []
    for it in range(60):
        plan.replay()
        if it % 10 == 0 or it == 59:
            losses.append(float(((out - targets) ** 2).sum()))
            print(f"step {it}: loss {losses[-1]:.2f}", flush=True)
    torch.cuda.synchronize()
    assert int(step.item()) == 60
    assert np.isfinite(losses).all()
    assert losses[-1] < 0.25 * losses[0], losses
    # the view-dependent band moved away from zero
    lay = plan.param_layout()
    sh = params[lay["colour"][0]:lay["colour"][1]].reshape(n, 4, 3).cpu().numpy()
    assert np.abs(sh[:, 1:]).max() > 1e-2
    plan.close()
