"""PLY loaders (del-splat-cpu/src/io_ply.rs). Files written with distinct
known values per property pin every record offset; hand-computed values pin
the activations the reference applies (io_ply.rs:182-216): rgb =
(f_dc + 0.5) * 0.5, quaternion (r0,r1,r2,r3) -> (r1,r2,r3,r0) normalised,
scale = exp, opacity = sigmoid. A 256-Gaussian cut of the reference's own
asset (tests/golden/bulba_first256.ply) checks the real header shape."""
import os

import numpy as np
import pytest

from _ply_files import GAUSS_PROPS, write_gauss, write_xyzrgb
from del_splat_dlpack import io_ply

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _records(n, seed=0):
    rng = np.random.default_rng(seed)
    r = rng.standard_normal((n, 62)).astype(np.float32)
    r[:, 9:54] = np.arange(45, dtype=np.float32)[None, :] + 100.0 * np.arange(n)[:, None]
    return r


def test_gauss_layout_and_activations(tmp_path):
    r = _records(5)
    # record 0: hand-computed activations
    r[0, 0:3] = [1.0, -2.0, 3.5]
    r[0, 6:9] = [0.5, -0.5, 1.5]        # rgb = (f_dc + 0.5) * 0.5 = 0.5, 0, 1
    r[0, 54] = 0.0                      # opacity sigmoid(0) = 0.5
    r[0, 55:58] = [0.0, np.log(2.0), -np.log(4.0)]  # scale 1, 2, 0.25
    r[0, 58:62] = [2.0, 0.0, 0.0, 0.0]  # (w,x,y,z) = (2,0,0,0) -> (0,0,0,1)
    write_gauss(tmp_path / "g.ply", r)
    xyz, rgb, sh, op, scale, quat = io_ply.read_3d_gauss_splat(tmp_path / "g.ply")
    assert xyz.shape == (15,) and rgb.shape == (15,) and sh.shape == (225,)
    assert op.shape == (5,) and scale.shape == (15,) and quat.shape == (20,)
    np.testing.assert_array_equal(xyz.reshape(5, 3), r[:, 0:3])
    np.testing.assert_array_equal(rgb[:3], [0.5, 0.0, 1.0])
    np.testing.assert_array_equal(sh.reshape(5, 45), r[:, 9:54])
    assert op[0] == 0.5
    np.testing.assert_allclose(scale[:3], [1.0, 2.0, 0.25], rtol=1e-6)
    np.testing.assert_array_equal(quat[:4], [0.0, 0.0, 0.0, 1.0])
    # the rest: the same formulas in float64, to float32 rounding
    r64 = r.astype(np.float64)
    np.testing.assert_allclose(rgb.reshape(5, 3), (r64[:, 6:9] + 0.5) * 0.5, rtol=1e-6)
    np.testing.assert_allclose(op, 1 / (1 + np.exp(-r64[:, 54])), rtol=1e-6)
    np.testing.assert_allclose(scale.reshape(5, 3), np.exp(r64[:, 55:58]), rtol=1e-6)
    q = r64[:, [59, 60, 61, 58]]
    np.testing.assert_allclose(quat.reshape(5, 4), q / np.linalg.norm(q, axis=1, keepdims=True),
                               rtol=0, atol=1e-6)
    assert all(a.dtype == np.float32 for a in (xyz, rgb, sh, op, scale, quat))


def test_gauss_raw_sh_layout(tmp_path):
    """Coefficient 0 is f_dc; coefficient k >= 1 of channel c is
    f_rest[15 c + k - 1] (the file stores the higher orders channel by
    channel)."""
    r = _records(3, 1)
    write_gauss(tmp_path / "g.ply", r)
    sh = io_ply.read_3d_gauss_splat_sh(tmp_path / "g.ply")
    assert sh.shape == (3, 16, 3) and sh.dtype == np.float32
    for i in range(3):
        for c in range(3):
            assert sh[i, 0, c] == r[i, 6 + c]
            for k in range(1, 16):
                assert sh[i, k, c] == r[i, 9 + 15 * c + k - 1]


def test_gauss_header_checks(tmp_path):
    r = _records(2)
    write_gauss(tmp_path / "g.ply", r)
    raw = (tmp_path / "g.ply").read_bytes()
    (tmp_path / "short.ply").write_bytes(raw[:-4])
    with pytest.raises(ValueError, match="payload"):
        io_ply.read_3d_gauss_splat(tmp_path / "short.ply")
    (tmp_path / "magic.ply").write_bytes(b"plx" + raw[3:])
    with pytest.raises(ValueError, match="not a PLY"):
        io_ply.read_3d_gauss_splat(tmp_path / "magic.ply")
    (tmp_path / "dbl.ply").write_bytes(raw.replace(b"property float opacity",
                                                   b"property double opacity"))
    with pytest.raises(ValueError, match="float"):
        io_ply.read_3d_gauss_splat(tmp_path / "dbl.ply")
    (tmp_path / "nocomment.ply").write_bytes(raw.replace(b"comment test file\n", b""))
    with pytest.raises(ValueError):
        io_ply.read_3d_gauss_splat(tmp_path / "nocomment.ply")
    (tmp_path / "props.ply").write_bytes(raw.replace(b"property float rot_3\n", b""))
    with pytest.raises(ValueError, match="62 properties"):
        io_ply.read_3d_gauss_splat(tmp_path / "props.ply")


def test_xyzrgb(tmp_path):
    xyz = np.array([[0.1, -2.0, 1e10], [3.0, 4.0, 5.0]])
    rgb = np.array([[0, 255, 51], [1, 2, 3]], np.uint8)
    write_xyzrgb(tmp_path / "p.ply", xyz, rgb)
    p, c = io_ply.read_xyzrgb(tmp_path / "p.ply")
    np.testing.assert_array_equal(p, xyz.astype(np.float32).reshape(-1))  # f64 -> f32 rounding
    np.testing.assert_array_equal(c[:3], np.float32([0.0, 1.0, 0.2]))
    np.testing.assert_array_equal(c, (rgb.astype(np.float32) / np.float32(255)).reshape(-1))


def test_reference_asset_cut():
    """The first 256 Gaussians of the reference's asset/bulba.ply: the
    reader accepts its header, quaternions come out unit, opacities in
    (0, 1), scales positive, and the raw SH agree with the flat f_rest."""
    path = os.path.join(GOLDEN, "bulba_first256.ply")
    xyz, rgb, sh45, op, scale, quat = io_ply.read_3d_gauss_splat(path)
    n = 256
    assert xyz.shape == (3 * n,)
    np.testing.assert_allclose(np.linalg.norm(quat.reshape(n, 4), axis=1), 1.0, atol=1e-6)
    assert ((op > 0) & (op < 1)).all() and (scale > 0).all()
    sh = io_ply.read_3d_gauss_splat_sh(path)
    np.testing.assert_array_equal(sh[:, 1:, :].transpose(0, 2, 1).reshape(n, 45),
                                  sh45.reshape(n, 45))
    np.testing.assert_array_equal(((sh[:, 0, :] + np.float32(0.5)) * np.float32(0.5)).reshape(-1),
                                  rgb)
    assert np.abs(sh[:, 1:, :]).max() > 0  # the asset carries view-dependent terms


def test_fixture_header_matches_property_names():
    with open(os.path.join(GOLDEN, "bulba_first256.ply"), "rb") as f:
        head = f.read(4096).split(b"end_header\n")[0].decode().splitlines()
    props = [l.split()[2] for l in head if l.startswith("property")]
    assert props == GAUSS_PROPS
