"""Camera matrices in the reference's conventions.

The projection follows the numpy restatement the reference's Python test uses
(del-splat-dlpack/tests/test_splat_sphere.py:16-28): an OpenGL-style frustum
whose third row is negated, so ndc z = +1 at the near plane and -1 at the far
plane (nearer = larger ndc z, which is what the depth key and the
front-to-back walk of the tile raster assume). Pixel y grows downwards
(transform_ndc2pix). Matrices are built row-major in float64 and handed to
the kernels as the column-major flat 16 (`to_col_major`), as
SplatCircle/torch.py:22 does with `transpose(0,1).flatten()`.
"""
import numpy as np


def perspective_fov_y(fov_y_deg, aspect, z_near, z_far):
    """Row-major 4x4 projection (test_splat_sphere.py:16-28, fov given directly)."""
    f = 1.0 / np.tan(np.radians(fov_y_deg) / 2.0)
    nf = 1.0 / (z_near - z_far)
    return np.array([
        [f / aspect, 0, 0, 0],
        [0, f, 0, 0],
        [0, 0, -(z_far + z_near) * nf, -(2 * z_far * z_near) * nf],
        [0, 0, -1, 0],
    ], dtype=np.float64)


def perspective_blender(f_mm, aspect, z_near, z_far, sensor_height_mm=24.0):
    """test_splat_sphere.py:16-28 (focal length in mm on a 24 mm sensor)."""
    fov_y_deg = 2.0 * np.degrees(np.arctan(sensor_height_mm / (2.0 * f_mm)))
    return perspective_fov_y(fov_y_deg, aspect, z_near, z_far)


def translation_modelview(t):
    """test_splat_sphere.py:30-36."""
    return np.array([[1, 0, 0, -t[0]], [0, 1, 0, -t[1]], [0, 0, 1, -t[2]], [0, 0, 0, 1]],
                    dtype=np.float64)


def look_at(eye, target, up=(0.0, 1.0, 0.0)):
    """Row-major world->camera matrix; the camera looks down its -z axis."""
    eye = np.asarray(eye, np.float64)
    target = np.asarray(target, np.float64)
    up = np.asarray(up, np.float64)
    fwd = target - eye
    fwd /= np.linalg.norm(fwd)
    if abs(np.dot(fwd, up)) > 0.999:
        up = np.array([0.0, 0.0, 1.0]) if abs(fwd[2]) < 0.9 else np.array([1.0, 0.0, 0.0])
    right = np.cross(fwd, up)
    right /= np.linalg.norm(right)
    true_up = np.cross(right, fwd)
    m = np.eye(4)
    m[0, :3], m[1, :3], m[2, :3] = right, true_up, -fwd
    m[:3, 3] = -m[:3, :3] @ eye
    return m


def world2ndc(eye, target, img_shape, fov_y_deg=60.0, z_near=0.01, z_far=100.0,
              up=(0.0, 1.0, 0.0)):
    """Row-major world->ndc for a pinhole camera looking at `target`."""
    w, h = img_shape
    return perspective_fov_y(fov_y_deg, w / h, z_near, z_far) @ look_at(eye, target, up)


def to_col_major(m):
    """Row-major 4x4 -> column-major flat 16 float32 (m[i + 4 j] = M_ij)."""
    return np.ascontiguousarray(np.asarray(m, np.float64).T.reshape(16), dtype=np.float32)


def camera_center(m):
    """Eye point of a perspective world->ndc matrix (row-major 4x4, or the
    column-major flat 16): the point where clip x, y and w all vanish."""
    m = np.asarray(m, np.float64)
    if m.shape == (16,):
        m = m.reshape(4, 4).T
    rows = m[[0, 1, 3]]
    return np.linalg.solve(rows[:, :3], -rows[:, 3])
