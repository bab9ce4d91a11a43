"""Spherical-harmonic colour kernels (dsplat_sh_to_rgb / _bwd) against the
oracle, through the C ABI: every (coefficient count, degree) pair, sizes
with partial waves and blocks, both gradient modes, the differentiable
wrapper composed with the renderer, and the reference's asset."""
import ctypes
import os

import numpy as np
import pytest
import torch

from _helpers import check_grad
from del_splat_dlpack import _native, io_ply

pytestmark = pytest.mark.gpu
DEV = "cuda"
PAIRS = [(1, 0), (4, 0), (4, 1), (9, 0), (9, 1), (9, 2), (16, 0), (16, 1), (16, 2), (16, 3)]
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def _scene(n, K, seed, cam=(0.3, -0.2, 4.0)):
    rng = np.random.default_rng(seed)
    xyz = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    sh = rng.normal(0, 0.4, (n, K, 3)).astype(np.float32)
    return xyz, sh, np.float32(cam)


@pytest.mark.parametrize("K,deg", PAIRS)
@pytest.mark.parametrize("n", [1, 63, 64, 1000, 70001])
def test_forward_matches_oracle(oracle, K, deg, n):
    import del_splat_dlpack.SplatGauss.torch as SG
    xyz, sh, cam = _scene(n, K, K * 10 + deg + n)
    print(f"sh fwd K={K} deg={deg} n={n}", flush=True)
    rgb = SG.sh_to_rgb_fwd(_t(xyz), _t(sh), _t(cam), deg).cpu().numpy()
    want = oracle.sh_to_rgb(xyz, sh, cam, deg)
    np.testing.assert_allclose(rgb, want, rtol=0, atol=4e-6)
    assert (want == 0).any() or n < 100 or deg == 0


@pytest.mark.parametrize("K,deg", PAIRS)
@pytest.mark.parametrize("n", [63, 1000, 70001])
def test_adjoint_matches_oracle(oracle, K, deg, n):
    import del_splat_dlpack.SplatGauss.torch as SG
    xyz, sh, cam = _scene(n, K, K * 10 + deg + n + 1)
    d_rgb = np.random.default_rng(n).standard_normal((n, 3)).astype(np.float32)
    print(f"sh bwd K={K} deg={deg} n={n}", flush=True)
    d_sh, d_xyz = SG.sh_to_rgb_bwd(_t(xyz), _t(sh), _t(cam), deg, _t(d_rgb))
    o_sh, o_xyz = oracle.sh_to_rgb_bwd(xyz, sh, cam, deg, d_rgb)
    check_grad(d_sh.cpu().numpy(), o_sh, "d_sh", rel=1e-4, abs_frac=1e-5, max_rel_tol=1e-5)
    if deg > 0:
        check_grad(d_xyz.cpu().numpy(), o_xyz, "d_xyz", abs_frac=1e-4, max_rel_tol=1e-4)
    else:
        assert not d_xyz.cpu().numpy().any()
    # add mode: the same gradients on top of what the buffers hold
    base_sh = np.random.default_rng(1).standard_normal(sh.shape).astype(np.float32)
    base_xyz = np.random.default_rng(2).standard_normal(xyz.shape).astype(np.float32)
    a_sh, a_xyz = _t(base_sh), _t(base_xyz)
    SG.sh_to_rgb_bwd(_t(xyz), _t(sh), _t(cam), deg, _t(d_rgb), a_sh, a_xyz)
    np.testing.assert_allclose(a_sh.cpu().numpy(), base_sh + d_sh.cpu().numpy(), rtol=0,
                               atol=1e-6)
    np.testing.assert_allclose(a_xyz.cpu().numpy(), base_xyz + d_xyz.cpu().numpy(), rtol=0,
                               atol=1e-5)


def test_full_size_million(oracle):
    """1M Gaussians, 16 coefficients, degree 3 (the 3DGS configuration)."""
    import del_splat_dlpack.SplatGauss.torch as SG
    n = 1 << 20
    xyz, sh, cam = _scene(n, 16, 5)
    g_xyz, g_sh, g_cam = _t(xyz), _t(sh), _t(cam)
    rgb = SG.sh_to_rgb_fwd(g_xyz, g_sh, g_cam, 3).cpu().numpy()
    np.testing.assert_allclose(rgb, oracle.sh_to_rgb(xyz, sh, cam, 3), rtol=0, atol=4e-6)
    d_rgb = np.random.default_rng(6).standard_normal((n, 3)).astype(np.float32)
    d_sh, d_xyz = SG.sh_to_rgb_bwd(g_xyz, g_sh, g_cam, 3, _t(d_rgb))
    o_sh, o_xyz = oracle.sh_to_rgb_bwd(xyz, sh, cam, 3, d_rgb)
    check_grad(d_sh.cpu().numpy(), o_sh, "d_sh", rel=1e-4, abs_frac=1e-5, max_rel_tol=1e-5)
    check_grad(d_xyz.cpu().numpy(), o_xyz, "d_xyz", abs_frac=1e-4, max_rel_tol=1e-4)


def test_camera_position_is_read_on_device(oracle):
    """cam_pos is a device buffer: changing its contents changes the next
    call's colours."""
    import del_splat_dlpack.SplatGauss.torch as SG
    xyz, sh, cam = _scene(500, 16, 7)
    g_cam = _t(cam)
    a = SG.sh_to_rgb_fwd(_t(xyz), _t(sh), g_cam, 3).cpu().numpy()
    g_cam.copy_(_t(np.float32([-3.0, 1.0, 0.5])))
    b = SG.sh_to_rgb_fwd(_t(xyz), _t(sh), g_cam, 3).cpu().numpy()
    np.testing.assert_allclose(a, oracle.sh_to_rgb(xyz, sh, cam, 3), rtol=0, atol=4e-6)
    np.testing.assert_allclose(b, oracle.sh_to_rgb(xyz, sh, np.float32([-3, 1, 0.5]), 3), rtol=0,
                               atol=4e-6)


def test_render_with_sh_colour_gradients(oracle):
    """render(sh_to_rgb(...)): sh.grad is the SH adjoint of the renderer's
    colour gradient, and xyz.grad is the projection's part plus the SH
    direction part."""
    import del_splat_dlpack.SplatGauss.torch as SG
    from del_splat_dlpack import scene
    shape = (160, 120)
    s = scene.gaussians(4000, 8)
    m = scene.front_camera(shape)
    eye = scene.camera.camera_center(m)
    np.testing.assert_allclose(eye, [0.0, 0.0, 3.0], atol=1e-5)
    sh_np = np.random.default_rng(9).normal(0, 0.3, (4000, 16, 3)).astype(np.float32)
    base = {k: _t(s[k]) for k in ("xyz", "log_scale", "quat", "opacity")}

    def run(with_sh):
        leaf = {k: v.clone().requires_grad_(True) for k, v in base.items()}
        sh = _t(sh_np).requires_grad_(True)
        rgb = SG.sh_to_rgb(leaf["xyz"], sh, eye.astype(np.float32), 3)
        if not with_sh:
            rgb = rgb.detach().requires_grad_(True)
        rgb.retain_grad()
        img, _ = SG.render(leaf["xyz"], leaf["log_scale"], leaf["quat"], leaf["opacity"], rgb,
                           m, shape)
        (img * img).sum().backward()
        return leaf, sh, rgb

    leaf, sh, rgb = run(True)
    leaf0, _, rgb0 = run(False)
    d_rgb = rgb.grad.cpu().numpy()
    # the two renders' colour gradients agree up to the order of the
    # adjoint's float atomics
    check_grad(d_rgb, rgb0.grad.cpu().numpy(), "d_rgb", rel=1e-4, abs_frac=1e-5,
               max_rel_tol=1e-5)
    o_sh, o_xyz = oracle.sh_to_rgb_bwd(s["xyz"], sh_np, eye, 3, d_rgb)
    check_grad(sh.grad.cpu().numpy(), o_sh, "d_sh", rel=1e-4, abs_frac=1e-5, max_rel_tol=1e-5)
    sh_part = leaf["xyz"].grad.cpu().numpy() - leaf0["xyz"].grad.cpu().numpy()
    # a difference of two adjoint results: absolute noise of the larger one
    check_grad(sh_part, o_xyz, "d_xyz (SH part)", max_rel_tol=1e-3)


def test_reference_asset_colours(oracle):
    """The SH colours of the first 256 Gaussians of the reference's asset,
    seen from two sides, against the oracle; degree 0 is c0 f_dc + 1/2."""
    import del_splat_dlpack.SplatGauss.torch as SG
    path = os.path.join(GOLDEN, "bulba_first256.ply")
    xyz = io_ply.read_3d_gauss_splat(path)[0].reshape(-1, 3)
    sh = io_ply.read_3d_gauss_splat_sh(path)
    for cam in (np.float32([0, 0, 5]), np.float32([3, -2, -4])):
        for deg in (0, 3):
            rgb = SG.sh_to_rgb_fwd(_t(xyz), _t(sh), _t(cam), deg).cpu().numpy()
            np.testing.assert_allclose(rgb, oracle.sh_to_rgb(xyz, sh, cam, deg), rtol=0,
                                       atol=4e-6)
            if deg == 0:
                c0 = 0.5 / np.sqrt(np.pi)
                np.testing.assert_allclose(rgb, np.maximum(c0 * sh[:, 0] + 0.5, 0), atol=1e-6)


def test_argument_errors():
    import del_splat_dlpack.SplatGauss.torch as SG
    xyz, sh, cam = _scene(100, 16, 10)
    g_xyz, g_sh, g_cam = _t(xyz), _t(sh), _t(cam)
    with pytest.raises(ValueError, match="needs 16"):
        SG.sh_to_rgb_fwd(g_xyz, g_sh[:, :9].contiguous(), g_cam, 3)
    with pytest.raises(ValueError, match="contiguous"):
        SG.sh_to_rgb_fwd(g_xyz, g_sh.transpose(0, 1).contiguous().transpose(0, 1), g_cam, 3)
    with pytest.raises(ValueError, match="K in"):
        SG.sh_to_rgb_fwd(g_xyz, torch.zeros((100, 5, 3), device=DEV), g_cam, 1)
    rgb = torch.empty((100, 3), device=DEV)
    lib = _native.lib()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    # C ABI: bad coefficient count, degree beyond the coefficients, a
    # misaligned sh pointer, null buffers
    assert lib.dsplat_sh_to_rgb(stream, 100, 1, 5, p(g_cam), p(g_xyz), p(g_sh), p(rgb)) != 0
    assert lib.dsplat_sh_to_rgb(stream, 100, 3, 9, p(g_cam), p(g_xyz), p(g_sh), p(rgb)) != 0
    assert lib.dsplat_sh_to_rgb(stream, 99, 3, 16, p(g_cam), p(g_xyz),
                                ctypes.c_void_p(g_sh.data_ptr() + 4), p(rgb)) != 0
    assert lib.dsplat_sh_to_rgb(stream, 100, 3, 16, None, p(g_xyz), p(g_sh), p(rgb)) != 0
    assert b"sh_to_rgb" in lib.dsplat_last_error()
    assert lib.dsplat_sh_to_rgb(stream, 0, 3, 16, None, None, None, None) == 0
