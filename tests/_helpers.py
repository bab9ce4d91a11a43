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
