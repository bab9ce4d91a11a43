"""The HIP adjoint against central finite differences on the full headline
frame (1M Gaussians, 1920x1080, 32x32 tiles, the benchmarked schedule:
512-thread forward, walk-ordered tiles, fused L2 adjoint at 256 threads).

The recipe of del-splat-cpu/src/splat_gaussian2.rs:245-266 as the
benchmark's configuration states it: 64 random (visible) Gaussians, float64
central differences with eps 1e-3 of the oracle's forward in each of mean,
log-scale, quaternion, opacity and colour, relative tolerance 1e-2. The
differences freeze the GPU frame's boxes, lists and walk lengths (the
function the adjoint differentiates) and re-render only the Gaussian's tiles
(tests/_fd.py). The HIP gradients are the library's own (projection adjoint
included), compared per element: relative 1e-2 above 1% of each
parameter's largest FD magnitude, 1e-3 of it below."""
import numpy as np
import pytest
import torch

from _fd import fd_gradients, visible_pick
from _helpers import check_grad, u32
from del_splat_dlpack import scene

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def test_hip_gradients_match_finite_differences_headline(oracle):
    import del_splat_dlpack.SplatGauss.torch as SG
    shape, ts = (1920, 1080), 32
    s = scene.gaussians(1_000_000, 2, radius=2.0)
    m = scene.front_camera(shape)
    target = scene.target_image(shape, 102)
    xyz, ls, q, op, rgb = (_t(s[k]) for k in ("xyz", "log_scale", "quat", "opacity", "rgb"))
    m16 = _t(m)
    pxd, conic, aabb, splat2 = SG.project_fwd(xyz, ls, q, op, rgb, m16, shape)
    t2i, i2p = SG.bin_gauss(shape, ts, pxd, aabb)
    cost = torch.empty(t2i.numel() - 1, device=DEV, dtype=torch.int32)
    img, alpha, nwalk, tfinal = SG.rasterize_fwd(shape, ts, splat2, t2i, i2p, 512, tile_cost=cost)
    dpix, dcon, dop, drgb = SG.rasterize_bwd_l2(shape, ts, splat2, t2i, i2p, nwalk, tfinal, img,
                                                _t(target), 256, tile_order=SG.tile_order(cost))
    dxyz, dls, dq = SG.project_bwd(xyz, ls, q, m16, shape, dpix, dcon)
    torch.cuda.synchronize()
    hip = dict(xyz=dxyz, log_scale=dls, quat=dq, opacity=dop, rgb=drgb)
    hip = {k: v.cpu().numpy() for k, v in hip.items()}
    pick = visible_pick(hip["rgb"], 64, 2024)
    print(f"pairs {i2p.numel()}, Gaussians {pick[:8]}...", flush=True)
    fd = fd_gradients(oracle, s, m, shape, ts, u32(t2i), u32(i2p), aabb.cpu().numpy(),
                      nwalk.cpu().numpy().view(np.uint32), target, pick, eps=1e-3)
    for name in ("xyz", "log_scale", "quat", "opacity", "rgb"):
        check_grad(hip[name][pick], fd[name], f"HIP vs FD {name}")
