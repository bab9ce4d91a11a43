"""COLMAP binary files for the reader tests, packed with the struct formats of
COLMAP's published sparse-model layout (the format strings of COLMAP's
read_write_model.py): cameras "<iiQQ" then the params as "<d"; images
"<idddddddi", the name, 0x00, "<Q", then "<ddq" per keypoint; points3D
"<QdddBBBd", "<Q", then "<ii" per track element; every file opens with a
"<Q" count.

`synthetic_reconstruction` makes a small consistent model: points in a ball,
pinhole cameras of several resolutions at distance 3 looking at the origin
(COLMAP axes: x right, y down, z forward), keypoints that are the exact
pinhole projections of the visible points plus a few keypoints with no
point (id -1), and tracks that list every observation.
"""
import struct

import numpy as np


def write_cameras(path, cams):
    """cams: [(camera_id, model_id, width, height, params)]."""
    with open(path, "wb") as f:
        f.write(struct.pack("<Q", len(cams)))
        for cid, mid, w, h, params in cams:
            f.write(struct.pack("<iiQQ", cid, mid, w, h))
            f.write(struct.pack("<%dd" % len(params), *params))


def write_images(path, images):
    """images: [(image_id, qvec wxyz, tvec, camera_id, name, xy [n,2], ids [n])]."""
    with open(path, "wb") as f:
        f.write(struct.pack("<Q", len(images)))
        for iid, q, t, cid, name, xy, ids in images:
            f.write(struct.pack("<idddddddi", iid, *q, *t, cid))
            f.write(name.encode("utf-8") + b"\x00")
            f.write(struct.pack("<Q", len(ids)))
            for (x, y), p in zip(xy, ids):
                f.write(struct.pack("<ddq", x, y, p))


def write_points3d(path, points):
    """points: [(id, xyz, rgb, error, [(image_id, point2D_idx)])]."""
    with open(path, "wb") as f:
        f.write(struct.pack("<Q", len(points)))
        for pid, xyz, rgb, err, track in points:
            f.write(struct.pack("<QdddBBBd", pid, *xyz, *rgb, err))
            f.write(struct.pack("<Q", len(track)))
            for a, b in track:
                f.write(struct.pack("<ii", a, b))


def rotation_to_qvec(r):
    """Unit quaternion (w, x, y, z) of a rotation matrix (Shepperd's method:
    divide by the largest of the four diagonal combinations)."""
    r = np.asarray(r, np.float64)
    tr = np.trace(r)
    cand = [tr, r[0, 0], r[1, 1], r[2, 2]]
    k = int(np.argmax(cand))
    if k == 0:
        s = 2.0 * np.sqrt(1.0 + tr)
        q = [0.25 * s, (r[2, 1] - r[1, 2]) / s, (r[0, 2] - r[2, 0]) / s, (r[1, 0] - r[0, 1]) / s]
    elif k == 1:
        s = 2.0 * np.sqrt(1.0 + r[0, 0] - r[1, 1] - r[2, 2])
        q = [(r[2, 1] - r[1, 2]) / s, 0.25 * s, (r[0, 1] + r[1, 0]) / s, (r[0, 2] + r[2, 0]) / s]
    elif k == 2:
        s = 2.0 * np.sqrt(1.0 + r[1, 1] - r[0, 0] - r[2, 2])
        q = [(r[0, 2] - r[2, 0]) / s, (r[0, 1] + r[1, 0]) / s, 0.25 * s, (r[1, 2] + r[2, 1]) / s]
    else:
        s = 2.0 * np.sqrt(1.0 + r[2, 2] - r[0, 0] - r[1, 1])
        q = [(r[1, 0] - r[0, 1]) / s, (r[0, 2] + r[2, 0]) / s, (r[1, 2] + r[2, 1]) / s, 0.25 * s]
    q = np.array(q)
    return q if q[0] >= 0 else -q


def look_at_colmap(centre, rng):
    """World->camera rotation (rows: right, down, forward) of a camera at
    `centre` looking at the origin, rolled by a random angle about its axis."""
    fwd = -np.asarray(centre, np.float64)
    fwd /= np.linalg.norm(fwd)
    up = np.array([0.0, 0.0, 1.0]) if abs(fwd[2]) < 0.9 else np.array([1.0, 0.0, 0.0])
    right = np.cross(fwd, up)
    right /= np.linalg.norm(right)
    down = np.cross(fwd, right)
    roll = rng.uniform(-0.3, 0.3)
    c, s = np.cos(roll), np.sin(roll)
    right, down = c * right + s * down, -s * right + c * down
    return np.stack([right, down, fwd])


# (camera_id, model_id, width, height, params): PINHOLE with fx != fy and an
# off-centre principal point, SIMPLE_PINHOLE, and a second PINHOLE size.
CAMERAS = [
    (1, 1, 160, 120, [150.0, 140.0, 78.5, 62.25]),
    (2, 0, 200, 200, [180.0, 100.0, 100.0]),
    (7, 1, 256, 144, [220.0, 230.0, 131.0, 70.0]),
]


def _intrinsics(cam):
    _, mid, _, _, p = cam
    return (p[0], p[0], p[1], p[2]) if mid == 0 else tuple(p)


def synthetic_reconstruction(directory, n_points=3000, n_images=5, seed=0, distance=3.0):
    """Writes cameras.bin, images.bin, points3D.bin under `directory`;
    returns the ground truth as a dict."""
    rng = np.random.default_rng(seed)
    d = rng.standard_normal((n_points, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    xyz = d * (rng.random(n_points) ** (1.0 / 3.0))[:, None]
    rgb = rng.integers(0, 256, (n_points, 3), dtype=np.uint8)
    err = rng.random(n_points)
    ids = 1000 + 7 * np.arange(n_points)  # sparse ids, as COLMAP leaves them
    tracks = [[] for _ in range(n_points)]
    images = []
    for i in range(n_images):
        cam = CAMERAS[i % len(CAMERAS)]
        fx, fy, cx, cy = _intrinsics(cam)
        c = rng.standard_normal(3)
        c *= distance / np.linalg.norm(c)
        r = look_at_colmap(c, rng)
        t = -r @ c
        pc = xyz @ r.T + t
        uv = np.stack([fx * pc[:, 0] / pc[:, 2] + cx, fy * pc[:, 1] / pc[:, 2] + cy], 1)
        # visible points, minus unmatched ones (a random fifth, and every
        # 50th point is never matched so that some tracks are empty)
        vis = np.nonzero((uv[:, 0] >= 0) & (uv[:, 0] < cam[2]) & (uv[:, 1] >= 0)
                         & (uv[:, 1] < cam[3]) & (rng.random(n_points) < 0.8)
                         & (np.arange(n_points) % 50 != 0))[0]
        vis = vis[rng.permutation(len(vis))]
        n_free = 5
        kp_xy = np.concatenate([uv[vis], rng.random((n_free, 2)) * [cam[2], cam[3]]])
        kp_id = np.concatenate([ids[vis], -np.ones(n_free, np.int64)])
        image_id = 10 + 3 * i
        for k, p in enumerate(vis):
            tracks[p].append((image_id, k))
        images.append((image_id, rotation_to_qvec(r), t, cam[0], f"view_{i:03d}.png", kp_xy,
                       kp_id))
    points = [(int(ids[p]), xyz[p], rgb[p], err[p], tracks[p]) for p in range(n_points)]
    write_cameras(f"{directory}/cameras.bin", CAMERAS)
    write_images(f"{directory}/images.bin", images)
    write_points3d(f"{directory}/points3D.bin", points)
    return dict(xyz=xyz, rgb=rgb, err=err, ids=ids, tracks=tracks, images=images)
