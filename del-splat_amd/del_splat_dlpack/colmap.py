"""COLMAP sparse-model readers (del-splat-cpu/src/colmap.rs) and the bridge
from a COLMAP reconstruction to the renderer's cameras and Gaussians.

Binary layout (little-endian, COLMAP's `cameras.bin`, `images.bin`,
`points3D.bin`; the reference reads the same fields in the same order):

  cameras.bin   u64 count, then per camera            colmap.rs:16-49
                i32 camera_id, u32 model_id, u64 width, u64 height,
                f64 params[n(model)]
  images.bin    u64 count, then per image             colmap.rs:72-115
                i32 image_id, f64 qvec[4] (w,x,y,z), f64 tvec[3],
                i32 camera_id, name (bytes up to a 0x00),
                u64 n, then n x (f64 x, f64 y, i64 point3D_id)
  points3D.bin  u64 count, then per point             colmap.rs:125-162
                u64 id, f64 xyz[3], u8 rgb[3], f64 error,
                u64 track length L, then L x (i32 image_id, i32 point2D_idx)

Differences from the reference, on purpose:
  * the reference accepts only model 1 (PINHOLE) and panics on any other id;
    here every model of COLMAP's table (ids 0-10) is read, so a file with a
    distorted camera still loads, and an unknown id raises ValueError. Only
    the undistorted models (SIMPLE_PINHOLE, PINHOLE) turn into a projection
    (`intrinsics`); the others raise ValueError there.
  * the reference keeps the terminating 0x00 inside `Image.name`
    (colmap.rs:52-62 pushes the byte before testing it); here it is dropped.
  * a truncated file raises ValueError instead of an I/O error.

The bridge builds, for image i, the column-major world->ndc matrix whose
pixel map (x_pix = (ndc_x + 1) W/2, y_pix = (1 - ndc_y) H/2, pixel centres
at i + 0.5) reproduces COLMAP's pinhole projection u = fx X/Z + cx,
v = fy Y/Z + cy of the camera-frame point (X, Y, Z) = R(q) x + t, the same
chain as the reference example (examples/colmap.rs:41-60: quaternion
(w,x,y,z) -> (x,y,z,w), M = T(tvec) R(q), then K). Depth follows the
renderer's convention: ndc z = +1 at z_near and -1 at z_far.
"""
import struct
from dataclasses import dataclass

import numpy as np

# COLMAP's camera model table: id -> (name, number of parameters).
CAMERA_MODELS = {
    0: ("SIMPLE_PINHOLE", 3),
    1: ("PINHOLE", 4),
    2: ("SIMPLE_RADIAL", 4),
    3: ("RADIAL", 5),
    4: ("OPENCV", 8),
    5: ("OPENCV_FISHEYE", 8),
    6: ("FULL_OPENCV", 12),
    7: ("FOV", 5),
    8: ("SIMPLE_RADIAL_FISHEYE", 4),
    9: ("RADIAL_FISHEYE", 5),
    10: ("THIN_PRISM_FISHEYE", 12),
}
MODEL_IDS = {name: mid for mid, (name, _) in CAMERA_MODELS.items()}


@dataclass
class Camera:
    """colmap.rs:9-14 (`model` is the model name, `params` its f64 values)."""
    camera_id: int
    width: int
    height: int
    model: str
    params: np.ndarray

    def intrinsics(self):
        """(fx, fy, cx, cy) of an undistorted pinhole camera."""
        p = self.params
        if self.model == "PINHOLE":
            return float(p[0]), float(p[1]), float(p[2]), float(p[3])
        if self.model == "SIMPLE_PINHOLE":
            return float(p[0]), float(p[0]), float(p[1]), float(p[2])
        raise ValueError(f"camera {self.camera_id}: model {self.model} has lens distortion; "
                         "undistort the images to PINHOLE first")


@dataclass
class Image:
    """colmap.rs:64-71; `vtx2xy` is the flat f64 [2n] keypoint array and
    `vtx2id` the i64 [n] point3D ids (-1 for keypoints without a point)."""
    image_id: int
    qvec: np.ndarray
    tvec: np.ndarray
    camera_id: int
    name: str
    vtx2xy: np.ndarray
    vtx2id: np.ndarray


@dataclass
class Point3D:
    """colmap.rs:117-123; `img` is an i32 [L, 2] array of (image_id, point2D_idx)."""
    id: int
    xyz: np.ndarray
    rgb: np.ndarray
    err: float
    img: np.ndarray


class Points3D:
    """The points of `points3D.bin` as arrays (ids u64 [N], xyz f64 [N,3],
    rgb u8 [N,3], err f64 [N], track_offsets i64 [N+1], track i32 [T,2]).
    Indexing and iteration give `Point3D` records, as the reference's Vec."""

    def __init__(self, ids, xyz, rgb, err, track_offsets, track):
        self.ids, self.xyz, self.rgb, self.err = ids, xyz, rgb, err
        self.track_offsets, self.track = track_offsets, track

    def __len__(self):
        return len(self.ids)

    def __getitem__(self, i):
        if not -len(self) <= i < len(self):
            raise IndexError(i)
        i %= len(self)
        t0, t1 = self.track_offsets[i], self.track_offsets[i + 1]
        return Point3D(int(self.ids[i]), self.xyz[i], self.rgb[i], float(self.err[i]),
                       self.track[t0:t1])

    def __iter__(self):
        return (self[i] for i in range(len(self)))

    def index_of(self):
        """point3D id -> row, as a dict (the reference example's pntid2pntidx)."""
        return {int(p): k for k, p in enumerate(self.ids)}


def _read(path):
    with open(path, "rb") as f:
        return f.read()


def _need(buf, off, n, what):
    if off + n > len(buf):
        raise ValueError(f"truncated COLMAP file while reading {what} at byte {off}")


def read_cameras(path):
    """colmap.rs:16-49 -> list of Camera."""
    buf = _read(path)
    _need(buf, 0, 8, "camera count")
    (num,) = struct.unpack_from("<Q", buf, 0)
    off = 8
    cams = []
    for _ in range(num):
        _need(buf, off, 24, "camera header")
        cid, mid, w, h = struct.unpack_from("<iIQQ", buf, off)
        off += 24
        if mid not in CAMERA_MODELS:
            raise ValueError(f"camera {cid}: unknown COLMAP model id {mid}")
        name, npar = CAMERA_MODELS[mid]
        _need(buf, off, 8 * npar, "camera parameters")
        params = np.frombuffer(buf, "<f8", npar, off).copy()
        off += 8 * npar
        cams.append(Camera(cid, w, h, name, params))
    return cams


_KEYPOINT = np.dtype([("x", "<f8"), ("y", "<f8"), ("id", "<i8")])


def read_images(path):
    """colmap.rs:72-115 -> list of Image."""
    buf = _read(path)
    _need(buf, 0, 8, "image count")
    (num,) = struct.unpack_from("<Q", buf, 0)
    off = 8
    images = []
    for _ in range(num):
        _need(buf, off, 64, "image header")
        head = struct.unpack_from("<i7di", buf, off)
        off += 64
        end = buf.find(b"\x00", off)
        if end < 0:
            raise ValueError(f"truncated COLMAP file: image name at byte {off} has no 0x00")
        name = buf[off:end].decode("utf-8")
        off = end + 1
        _need(buf, off, 8, "keypoint count")
        (npt,) = struct.unpack_from("<Q", buf, off)
        off += 8
        _need(buf, off, 24 * npt, "keypoints")
        kp = np.frombuffer(buf, _KEYPOINT, npt, off)
        off += 24 * npt
        vtx2xy = np.stack([kp["x"], kp["y"]], 1).reshape(-1)
        images.append(Image(head[0], np.array(head[1:5]), np.array(head[5:8]), head[8], name,
                            vtx2xy, kp["id"].copy()))
    return images


_POINT_HEAD = np.dtype([("id", "<u8"), ("xyz", "<f8", (3,)), ("rgb", "u1", (3,)),
                        ("err", "<f8"), ("len", "<u8")])  # 51 bytes, packed


def read_points3d(path):
    """colmap.rs:125-162 -> Points3D."""
    buf = _read(path)
    _need(buf, 0, 8, "point count")
    (num,) = struct.unpack_from("<Q", buf, 0)
    hsize = _POINT_HEAD.itemsize
    starts = np.empty(num, np.int64)
    lens = np.empty(num, np.int64)
    off = 8
    unpack = struct.Struct("<Q").unpack_from
    for i in range(num):
        _need(buf, off, hsize, "point")
        starts[i] = off
        (n,) = unpack(buf, off + hsize - 8)
        lens[i] = n
        off += hsize + 8 * n
    if off > len(buf):
        raise ValueError(f"truncated COLMAP file: point tracks end at byte {off}")
    # The records are [head | track] back to back: a byte mask of the heads
    # splits the payload into all heads and all tracks, each contiguous.
    payload = np.frombuffer(buf, np.uint8, off - 8, 8)
    edge = np.zeros(len(payload) + 1, np.int8)
    edge[starts - 8] += 1
    edge[starts - 8 + hsize] -= 1
    is_head = np.cumsum(edge[:-1], dtype=np.int8).view(np.bool_)
    heads = payload[is_head].view(_POINT_HEAD)
    track = payload[~is_head].view("<i4").reshape(-1, 2)
    offsets = np.zeros(num + 1, np.int64)
    np.cumsum(lens, out=offsets[1:])
    return Points3D(heads["id"].copy(), heads["xyz"].copy(), heads["rgb"].copy(),
                    heads["err"].copy(), offsets, track.astype(np.int32))


# ------------------------------------------------------------------ bridge --

def qvec_to_rotation(qvec):
    """Row-major 3x3 of the unit quaternion (w, x, y, z), COLMAP's world->camera
    rotation (the Hamilton formula; the quaternion is used as stored)."""
    w, x, y, z = (float(v) for v in qvec)
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
        [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
        [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)],
    ])


def world2cam(image):
    """Row-major 4x4 [R(q) | t] (examples/colmap.rs:52-54)."""
    m = np.eye(4)
    m[:3, :3] = qvec_to_rotation(image.qvec)
    m[:3, 3] = np.asarray(image.tvec, np.float64)
    return m


def pinhole_projection(camera, z_near=0.01, z_far=100.0):
    """Row-major 4x4 camera-frame -> clip matrix of a pinhole camera (camera
    looks down +Z, image y down): clip w = Z, and the renderer's ndc -> pixel
    map turns clip/w into u = fx X/Z + cx, v = fy Y/Z + cy."""
    if not 0.0 < z_near < z_far:
        raise ValueError("need 0 < z_near < z_far")
    fx, fy, cx, cy = camera.intrinsics()
    w, h = float(camera.width), float(camera.height)
    a = -(z_far + z_near) / (z_far - z_near)
    b = 2.0 * z_far * z_near / (z_far - z_near)
    return np.array([
        [2.0 * fx / w, 0.0, 2.0 * cx / w - 1.0, 0.0],
        [0.0, -2.0 * fy / h, 1.0 - 2.0 * cy / h, 0.0],
        [0.0, 0.0, a, b],
        [0.0, 0.0, 1.0, 0.0],
    ])


def world2ndc(camera, image, z_near=0.01, z_far=100.0):
    """Column-major float32 [16] world->ndc of one registered image."""
    m = pinhole_projection(camera, z_near, z_far) @ world2cam(image)
    return np.ascontiguousarray(m.T.reshape(16), dtype=np.float32)


def views(cameras, images, z_near=0.01, z_far=100.0):
    """[(world2ndc float32[16], (width, height))] of every image, in the order
    of `images`: the view list of SplatGauss.multiview.MultiViewPlan."""
    by_id = {c.camera_id: c for c in cameras}
    out = []
    for im in images:
        if im.camera_id not in by_id:
            raise ValueError(f"image {im.image_id} names camera {im.camera_id}, not in cameras")
        cam = by_id[im.camera_id]
        out.append((world2ndc(cam, im, z_near, z_far), (int(cam.width), int(cam.height))))
    return out


def project_pinhole(camera, image, xyz):
    """Pixel coordinates (f64 [N,2]) and camera depth Z (f64 [N]) of world
    points through COLMAP's pinhole model (examples/colmap.rs:63-67)."""
    fx, fy, cx, cy = camera.intrinsics()
    m = world2cam(image)
    pc = np.asarray(xyz, np.float64) @ m[:3, :3].T + m[:3, 3]
    z = pc[:, 2]
    return np.stack([fx * pc[:, 0] / z + cx, fy * pc[:, 1] / z + cy], 1), z


def reprojection_errors(cameras, images, points):
    """Per observation (keypoint with a point3D id) the pixel distance between
    the keypoint and its point's pinhole projection, one f64 array per image
    (the check of examples/colmap.rs:56-70)."""
    by_id = {c.camera_id: c for c in cameras}
    row = points.index_of()
    out = []
    for im in images:
        sel = np.nonzero(im.vtx2id >= 0)[0]
        rows = np.fromiter((row[int(p)] for p in im.vtx2id[sel]), np.int64, len(sel))
        uv, _ = project_pinhole(by_id[im.camera_id], im, points.xyz[rows])
        kp = im.vtx2xy.reshape(-1, 2)[sel]
        out.append(np.linalg.norm(uv - kp, axis=1))
    return out


def gaussians_from_points(points, k=3, opacity=0.1, min_scale=1e-7):
    """Initial Gaussians of a COLMAP point cloud, the standard 3D Gaussian
    splatting start: one isotropic Gaussian per point at its position with
    its colour (rgb / 255), log-scale log(sqrt(mean squared distance to the
    k nearest other points)), identity rotation (x,y,z,w = 0,0,0,1) and the
    given opacity. Returns float32 arrays in the scene-dictionary layout
    (xyz, log_scale, scale, quat, opacity, rgb)."""
    from scipy.spatial import cKDTree

    xyz = np.asarray(points.xyz, np.float64)
    n = len(xyz)
    if n == 0:
        raise ValueError("empty point cloud")
    kk = min(k, n - 1)
    if kk > 0:
        d, _ = cKDTree(xyz).query(xyz, kk + 1)
        d2 = np.mean(np.square(d[:, 1:]), axis=1)
    else:
        d2 = np.ones(n)
    log_scale = 0.5 * np.log(np.maximum(d2, min_scale ** 2))
    log_scale = np.repeat(log_scale[:, None], 3, 1).astype(np.float32)
    quat = np.zeros((n, 4), np.float32)
    quat[:, 3] = 1.0
    return dict(xyz=xyz.astype(np.float32), log_scale=log_scale,
                scale=np.exp(log_scale).astype(np.float32), quat=quat,
                opacity=np.full(n, opacity, np.float32),
                rgb=(np.asarray(points.rgb, np.float32) / 255.0))
