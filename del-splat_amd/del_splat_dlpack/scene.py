"""Seeded synthetic scenes for the splatting workloads.

Distributions (all float32, numpy.random.default_rng(seed)):
  Gaussians: means uniform in a ball of radius R (direction normalize(N(0,I3)),
  radius R U^(1/3)); log-scale N(-4, 0.5) per axis; quaternion
  normalize(N(0,I4)) stored [x,y,z,w]; opacity sigmoid(N(0,1)); colour U[0,1].
  Points: uniform in the cube [-0.5,0.5]^3, pixel radius U[0.5,3], colour U[0,1].
Cameras look at the origin from distance 3 with a 60 degree vertical field of
view, near 0.01 and far 100.
"""
import numpy as np

from . import camera

RESOLUTIONS = ((640, 480), (1280, 720), (1920, 1080), (800, 800))


def gaussians(n, seed, radius=1.0):
    rng = np.random.default_rng(seed)
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = radius * rng.random(n) ** (1.0 / 3.0)
    xyz = (d * r[:, None]).astype(np.float32)
    log_scale = rng.normal(-4.0, 0.5, (n, 3)).astype(np.float32)
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    quat = q.astype(np.float32)
    opacity = (1.0 / (1.0 + np.exp(-rng.standard_normal(n)))).astype(np.float32)
    rgb = rng.random((n, 3)).astype(np.float32)
    return dict(xyz=xyz, log_scale=log_scale, scale=np.exp(log_scale).astype(np.float32),
                quat=quat, opacity=opacity, rgb=rgb)


def points(n, seed):
    rng = np.random.default_rng(seed)
    xyz = (rng.random((n, 3)) - 0.5).astype(np.float32)
    pixrad = rng.uniform(0.5, 3.0, n).astype(np.float32)
    rgb = rng.random((n, 3)).astype(np.float32)
    return dict(xyz=xyz, pixrad=pixrad, rgb=rgb)


def front_camera(img_shape, distance=3.0):
    """Column-major world->ndc of the camera at (0,0,distance) looking at 0."""
    return camera.to_col_major(camera.world2ndc((0.0, 0.0, distance), (0.0, 0.0, 0.0), img_shape))


def sphere_cameras(num_views, seed, distance=3.0, resolutions=RESOLUTIONS):
    """Cameras on a sphere (Fibonacci directions under a seeded random rotation),
    each with a resolution drawn from `resolutions`. Returns a list of
    (col-major world2ndc float32[16], (w, h))."""
    rng = np.random.default_rng(seed)
    k = np.arange(num_views) + 0.5
    phi = np.arccos(1.0 - 2.0 * k / num_views)
    theta = np.pi * (1.0 + 5.0 ** 0.5) * k
    dirs = np.stack([np.cos(theta) * np.sin(phi), np.sin(theta) * np.sin(phi), np.cos(phi)], 1)
    qm, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    dirs = dirs @ qm.T
    res_idx = rng.integers(0, len(resolutions), num_views)
    out = []
    for v in range(num_views):
        shape = tuple(resolutions[res_idx[v]])
        m = camera.world2ndc(distance * dirs[v], (0.0, 0.0, 0.0), shape)
        out.append((camera.to_col_major(m), shape))
    return out


def target_image(img_shape, seed):
    w, h = img_shape
    return np.random.default_rng(seed).random((h, w, 3)).astype(np.float32)
