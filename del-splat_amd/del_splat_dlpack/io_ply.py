"""PLY readers (del-splat-cpu/src/io_ply.rs), numpy implementation.

read_xyzrgb          io_ply.rs:30-105  binary little-endian, per vertex
                     f64 x,y,z + u8 r,g,b; colours scaled to [0,1] float32.
read_3d_gauss_splat  io_ply.rs:128-227 the 3DGS layout of exactly 62 float
                     properties (x,y,z, nx,ny,nz, f_dc 3, f_rest 45,
                     opacity, scale 3, rot 4) with the reference's
                     activations: rgb = (f_dc + 0.5) * 0.5 (its sh_c0 is 0.5),
                     quaternion (r0,r1,r2,r3) -> (r1,r2,r3,r0) normalised,
                     scale = exp, opacity = sigmoid.
read_3d_gauss_splat_sh  the same file's raw SH coefficients [N,16,3] for the
                     view-dependent colour (new).
All expect the header shape the reference asserts (magic, format, one
comment line, element vertex, the properties, end_header).
"""
import numpy as np


def _read_header(f, num_props_expected=None):
    lines = []
    line = f.readline()
    if line != b"ply\n":
        raise ValueError("not a PLY file")
    fmt = f.readline().split()
    if fmt[0] != b"format" or fmt[1] not in (b"binary_little_endian", b"binary_big_endian",
                                             b"ascii"):
        raise ValueError("bad PLY format line")
    comment = f.readline().split()
    if not comment or comment[0] != b"comment":
        raise ValueError("expected one comment line")
    elem = f.readline().split()
    if elem[0] != b"element":
        raise ValueError("expected element line")
    num = int(elem[2])
    while True:
        line = f.readline()
        if line == b"end_header\n":
            break
        if not line:
            raise ValueError("missing end_header")
        lines.append(line.split())
    if num_props_expected is not None and len(lines) != num_props_expected:
        raise ValueError(f"expected {num_props_expected} properties, found {len(lines)}")
    return num, lines, fmt[1].decode()


def read_xyzrgb(path):
    """-> (pnt2xyz float32 [N*3], pnt2rgb float32 [N*3]) flat, as the reference."""
    with open(path, "rb") as f:
        num, props, _ = _read_header(f, 6)
        buf = f.read()
    dt = np.dtype([("x", "<f8"), ("y", "<f8"), ("z", "<f8"), ("r", "u1"), ("g", "u1"),
                   ("b", "u1")])
    a = np.frombuffer(buf, dtype=dt, count=num)
    xyz = np.stack([a["x"], a["y"], a["z"]], 1).astype(np.float32).reshape(-1)
    rgb = (np.stack([a["r"], a["g"], a["b"]], 1).astype(np.float32) / 255.0).reshape(-1)
    return xyz, rgb


def _read_gauss_records(path):
    with open(path, "rb") as f:
        num, props, _ = _read_header(f, 62)
        for p in props:
            if p[0] != b"property" or p[1] != b"float":
                raise ValueError("all 62 properties must be float")
        buf = f.read()
    if len(buf) != 62 * num * 4:
        raise ValueError("payload size mismatch")
    return np.frombuffer(buf, dtype="<f4").reshape(num, 62)


def read_3d_gauss_splat(path):
    """-> (xyz [N*3], rgb [N*3], sh [N*45], opacity [N], scale [N*3], quat [N*4])."""
    a = _read_gauss_records(path)
    sh_c0 = np.float32(0.5)
    xyz = np.ascontiguousarray(a[:, 0:3])
    rgb = (a[:, 6:9] + np.float32(0.5)) * sh_c0
    sh = np.ascontiguousarray(a[:, 9:54])
    op = a[:, 54]
    op = (np.float32(1.0) / (np.float32(1.0) + np.exp(-op))).astype(np.float32)
    scale = np.exp(a[:, 55:58]).astype(np.float32)
    q = a[:, 58:62]
    q = np.stack([q[:, 1], q[:, 2], q[:, 3], q[:, 0]], 1)
    q = (q / np.sqrt((q * q).sum(1, keepdims=True))).astype(np.float32)
    return (xyz.reshape(-1), rgb.astype(np.float32).reshape(-1), sh.reshape(-1), op,
            scale.reshape(-1), q.reshape(-1))


def read_3d_gauss_splat_sh(path):
    """Raw spherical-harmonic coefficients of the same layout as float32
    [N, 16, 3] (coefficient-major, channel inner; the layout of the kernels'
    SH colour): coefficient 0 is f_dc, coefficients 1..15 of channel c are
    f_rest[15 c .. 15 c + 14] (the file stores the higher orders channel by
    channel). No counterpart in the reference, whose reader folds f_dc into
    its rgb and returns f_rest flat (io_ply.rs:196-204)."""
    a = _read_gauss_records(path)
    n = a.shape[0]
    sh = np.empty((n, 16, 3), np.float32)
    sh[:, 0, :] = a[:, 6:9]
    sh[:, 1:, :] = a[:, 9:54].reshape(n, 3, 15).transpose(0, 2, 1)
    return sh
