"""Shared test helpers: seeded scenes and GPU pipeline runs."""
import numpy as np

from del_splat_dlpack import scene


def small_scene(n, seed, img_shape=(256, 192), radius=1.0, distance=3.0):
    s = scene.gaussians(n, seed, radius=radius)
    m = scene.camera.to_col_major(
        scene.camera.world2ndc((0.0, 0.0, distance), (0.0, 0.0, 0.0), img_shape))
    return s, m


def max_rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = max(np.abs(b).max(), 1e-30)
    return np.abs(a - b).max() / den


def u32(t):
    """int32 tensor holding u32 values -> numpy uint32."""
    return t.cpu().numpy().view(np.uint32)


def check_pixels(g_img, g_alpha, g_nwalk, o_img, o_alpha, o_nwalk, tol=1e-4):
    """Pixels agree within tol; where the two sides' termination differs
    (the walks end at different Gaussians) the final alpha must sit within
    2e-3 of the 0.999 threshold, and such pixels must be rare."""
    diff = np.abs(g_img - o_img).max(axis=2)
    flip = g_nwalk != o_nwalk
    bad = (diff > tol) & ~flip
    assert not bad.any(), f"{bad.sum()} pixels off by up to {diff[bad].max()}"
    if flip.any():
        # a flip can only happen when the running alpha crosses 0.999 within
        # rounding distance on one side
        near = np.abs(np.maximum(g_alpha, o_alpha) - 0.999) < 2e-3
        assert near[flip].all()
        assert flip.mean() < 1e-3, flip.mean()
    assert np.abs(g_alpha - o_alpha)[~flip].max(initial=0) <= tol
    return int(flip.sum())
