"""COLMAP readers and the camera bridge (del-splat-cpu/src/colmap.rs,
examples/colmap.rs). The reference holds no COLMAP data, so the readers are
pinned by files packed with COLMAP's published struct formats, the bridge by
hand-computed pinhole projections, the ndc depth convention and the oracle's
projection of the same points."""
import struct

import numpy as np
import pytest

from _colmap_files import (CAMERAS, rotation_to_qvec, synthetic_reconstruction, write_cameras,
                           write_images, write_points3d)
from del_splat_dlpack import colmap


@pytest.fixture(scope="module")
def recon(tmp_path_factory):
    d = tmp_path_factory.mktemp("colmap")
    truth = synthetic_reconstruction(str(d), n_points=2000, n_images=5, seed=3)
    return (colmap.read_cameras(d / "cameras.bin"), colmap.read_images(d / "images.bin"),
            colmap.read_points3d(d / "points3D.bin"), truth)


def test_read_cameras_all_fields(recon):
    cams = recon[0]
    assert len(cams) == len(CAMERAS)
    for c, (cid, mid, w, h, params) in zip(cams, CAMERAS):
        assert (c.camera_id, c.width, c.height) == (cid, w, h)
        assert c.model == colmap.CAMERA_MODELS[mid][0]
        np.testing.assert_array_equal(c.params, params)
    assert cams[1].intrinsics() == (180.0, 180.0, 100.0, 100.0)  # SIMPLE_PINHOLE: f, cx, cy


def test_read_images_all_fields(recon):
    images, truth = recon[1], recon[3]
    assert len(images) == len(truth["images"])
    for im, (iid, q, t, cid, name, xy, ids) in zip(images, truth["images"]):
        assert (im.image_id, im.camera_id, im.name) == (iid, cid, name)
        np.testing.assert_array_equal(im.qvec, q)
        np.testing.assert_array_equal(im.tvec, t)
        np.testing.assert_array_equal(im.vtx2xy, np.asarray(xy).reshape(-1))
        np.testing.assert_array_equal(im.vtx2id, ids)
        assert im.vtx2xy.dtype == np.float64 and im.vtx2id.dtype == np.int64


def test_read_points3d_all_fields(recon):
    pts, truth = recon[2], recon[3]
    assert len(pts) == len(truth["ids"])
    np.testing.assert_array_equal(pts.ids, truth["ids"])
    np.testing.assert_array_equal(pts.xyz, truth["xyz"])
    np.testing.assert_array_equal(pts.rgb, truth["rgb"])
    np.testing.assert_array_equal(pts.err, truth["err"])
    lens = [len(t) for t in truth["tracks"]]
    assert min(lens) == 0 < max(lens)  # empty and non-empty tracks both present
    for i in (0, 1, len(pts) // 2, len(pts) - 1, -1):
        p = pts[i]
        assert p.id == truth["ids"][i]
        np.testing.assert_array_equal(p.img, np.asarray(truth["tracks"][i], np.int32).reshape(-1, 2))
    assert sum(1 for _ in pts) == len(pts)
    with pytest.raises(IndexError):
        pts[len(pts)]


def test_tracks_and_keypoints_agree(recon):
    """Every track element names an image keypoint that carries the point's id."""
    images, pts = recon[1], recon[2]
    by_id = {im.image_id: im for im in images}
    n_obs = 0
    for p in pts:
        for img_id, k in p.img:
            assert by_id[int(img_id)].vtx2id[k] == p.id
            n_obs += 1
    assert n_obs == sum(int((im.vtx2id >= 0).sum()) for im in images)


def test_empty_files(tmp_path):
    write_cameras(tmp_path / "c.bin", [])
    write_images(tmp_path / "i.bin", [])
    write_points3d(tmp_path / "p.bin", [])
    assert colmap.read_cameras(tmp_path / "c.bin") == []
    assert colmap.read_images(tmp_path / "i.bin") == []
    p = colmap.read_points3d(tmp_path / "p.bin")
    assert len(p) == 0 and p.xyz.shape == (0, 3) and p.track.shape == (0, 2)


def test_image_without_keypoints_and_point_without_track(tmp_path):
    write_images(tmp_path / "i.bin", [(5, [1, 0, 0, 0], [0, 0, 0], 1, "a", np.zeros((0, 2)), [])])
    write_points3d(tmp_path / "p.bin", [(9, [1.0, 2.0, 3.0], [1, 2, 3], 0.5, [])])
    (im,) = colmap.read_images(tmp_path / "i.bin")
    assert im.name == "a" and im.vtx2xy.shape == (0,) and im.vtx2id.shape == (0,)
    p = colmap.read_points3d(tmp_path / "p.bin")[0]
    assert p.id == 9 and p.err == 0.5 and p.img.shape == (0, 2)
    np.testing.assert_array_equal(p.rgb, [1, 2, 3])


def test_all_camera_models_read_and_distorted_ones_refuse_projection(tmp_path):
    cams = [(i + 1, mid, 64, 48, list(np.arange(n) + 1.0))
            for i, (mid, (_, n)) in enumerate(sorted(colmap.CAMERA_MODELS.items()))]
    write_cameras(tmp_path / "c.bin", cams)
    got = colmap.read_cameras(tmp_path / "c.bin")
    assert [c.model for c in got] == [colmap.CAMERA_MODELS[m][0] for _, m, *_ in cams]
    for c, (*_, params) in zip(got, cams):
        np.testing.assert_array_equal(c.params, params)
        if c.model in ("PINHOLE", "SIMPLE_PINHOLE"):
            c.intrinsics()
        else:
            with pytest.raises(ValueError, match="distortion"):
                c.intrinsics()


def test_unknown_model_and_truncation_raise(tmp_path):
    write_cameras(tmp_path / "c.bin", [(1, 42, 8, 8, [])])
    with pytest.raises(ValueError, match="unknown COLMAP model id 42"):
        colmap.read_cameras(tmp_path / "c.bin")
    write_points3d(tmp_path / "p.bin", [(1, [0.0, 0.0, 0.0], [0, 0, 0], 0.0, [(1, 2), (3, 4)])])
    raw = (tmp_path / "p.bin").read_bytes()
    (tmp_path / "t.bin").write_bytes(raw[:-3])
    with pytest.raises(ValueError, match="truncated"):
        colmap.read_points3d(tmp_path / "t.bin")
    write_images(tmp_path / "i.bin", [(5, [1, 0, 0, 0], [0, 0, 0], 1, "abc", np.zeros((1, 2)), [3])])
    raw = (tmp_path / "i.bin").read_bytes()
    for cut in (10, 8 + 64 + 2, len(raw) - 1):
        (tmp_path / "t.bin").write_bytes(raw[:cut])
        with pytest.raises(ValueError, match="truncated"):
            colmap.read_images(tmp_path / "t.bin")


def test_name_is_utf8_without_terminator(tmp_path):
    write_images(tmp_path / "i.bin", [(1, [1, 0, 0, 0], [0, 0, 0], 1, "dir/bär 01.jpg",
                                       np.zeros((0, 2)), [])])
    assert colmap.read_images(tmp_path / "i.bin")[0].name == "dir/bär 01.jpg"


def test_point_record_is_51_bytes():
    assert struct.calcsize("<QdddBBBdQ") == colmap._POINT_HEAD.itemsize == 51


# ------------------------------------------------------------------ bridge --

def _ndc(m16, x):
    m = np.asarray(m16, np.float64).reshape(4, 4).T  # col-major -> row-major
    y = m @ np.append(np.asarray(x, np.float64), 1.0)
    return y[:3] / y[3], y[3]


def test_world2ndc_known_answer():
    """fx 100, fy 120, cx 50, cy 40, camera at the origin looking down +Z
    shifted by t = (0,0,5): the world point (1,2,0) is at camera (1,2,5), so
    u = 100/5 + 50 = 70 and v = 120*2/5 + 40 = 88."""
    cam = colmap.Camera(1, 100, 80, "PINHOLE", np.array([100.0, 120.0, 50.0, 40.0]))
    im = colmap.Image(1, np.array([1.0, 0, 0, 0]), np.array([0.0, 0, 5]), 1, "x",
                      np.zeros(0), np.zeros(0, np.int64))
    m16 = colmap.pinhole_projection(cam, 0.5, 50.0) @ colmap.world2cam(im)
    ndc, w = _ndc(m16.T.reshape(16), [1.0, 2.0, 0.0])
    assert w == 5.0
    np.testing.assert_allclose([(ndc[0] + 1) * 50, (1 - ndc[1]) * 40], [70.0, 88.0], rtol=0,
                               atol=1e-12)
    # depth: ndc z = +1 at the near plane, -1 at the far plane, decreasing
    for z, want in ((0.5, 1.0), (50.0, -1.0)):
        nz, _ = _ndc(m16.T.reshape(16), [0.0, 0.0, z - 5.0])
        assert abs(nz[2] - want) < 1e-12
    z1, _ = _ndc(m16.T.reshape(16), [0.0, 0.0, 1.0])
    z2, _ = _ndc(m16.T.reshape(16), [0.0, 0.0, 2.0])
    assert 1.0 > z1[2] > z2[2] > -1.0
    with pytest.raises(ValueError):
        colmap.pinhole_projection(cam, 1.0, 0.5)


def test_quaternion_convention():
    """A rotation of +90 degrees about z, qvec (cos 45, 0, 0, sin 45), maps the
    x axis to the y axis (Hamilton convention, as COLMAP's qvec2rotmat)."""
    h = np.sqrt(0.5)
    r = colmap.qvec_to_rotation([h, 0.0, 0.0, h])
    np.testing.assert_allclose(r @ [1.0, 0, 0], [0.0, 1.0, 0.0], atol=1e-15)
    rng = np.random.default_rng(0)
    for _ in range(20):
        q = rng.standard_normal(4)
        q /= np.linalg.norm(q)
        r = colmap.qvec_to_rotation(q)
        np.testing.assert_allclose(r @ r.T, np.eye(3), atol=1e-14)
        assert abs(np.linalg.det(r) - 1.0) < 1e-14
        np.testing.assert_allclose(colmap.qvec_to_rotation(rotation_to_qvec(r)), r, atol=1e-14)


def test_bridge_reproduces_keypoints(recon):
    """Keypoints were written as exact pinhole projections: the renderer's
    world->ndc matrix plus its ndc->pixel map lands on them, and the camera
    centre looks at the origin (which projects to the principal point)."""
    cams, images, pts, _ = recon
    by_id = {c.camera_id: c for c in cams}
    row = pts.index_of()
    for (m16, (w, h)), im in zip(colmap.views(cams, images), images):
        cam = by_id[im.camera_id]
        assert (w, h) == (cam.width, cam.height)
        sel = np.nonzero(im.vtx2id >= 0)[0]
        xyz = pts.xyz[[row[int(p)] for p in im.vtx2id[sel]]]
        m = np.asarray(m16, np.float64).reshape(4, 4).T
        y = np.c_[xyz, np.ones(len(xyz))] @ m.T
        ndc = y[:, :3] / y[:, 3:]
        pix = np.stack([(ndc[:, 0] + 1) * 0.5 * w, (1 - ndc[:, 1]) * 0.5 * h], 1)
        # float32 matrix entries: ~1e-7 relative error on coordinates of ~1e2
        np.testing.assert_allclose(pix, im.vtx2xy.reshape(-1, 2)[sel], rtol=0, atol=2e-4)
        fx, fy, cx, cy = cam.intrinsics()
        c_ndc, _ = _ndc(m16, [0.0, 0.0, 0.0])
        np.testing.assert_allclose([(c_ndc[0] + 1) * 0.5 * w, (1 - c_ndc[1]) * 0.5 * h],
                                   [cx, cy], atol=1e-4)


def test_reprojection_errors(recon):
    cams, images, pts, _ = recon
    errs = colmap.reprojection_errors(cams, images, pts)
    assert len(errs) == len(images)
    for e, im in zip(errs, images):
        assert len(e) == int((im.vtx2id >= 0).sum()) > 100
        assert e.max() < 1e-9
    # a known offset on one keypoint shows up as that distance
    im = images[0]
    k = int(np.nonzero(im.vtx2id >= 0)[0][0])
    moved = colmap.Image(im.image_id, im.qvec, im.tvec, im.camera_id, im.name,
                         im.vtx2xy.copy(), im.vtx2id)
    moved.vtx2xy[2 * k:2 * k + 2] += [3.0, 4.0]
    e = colmap.reprojection_errors(cams, [moved], pts)[0]
    assert abs(e.max() - 5.0) < 1e-9 and np.count_nonzero(e > 1e-9) == 1


def test_oracle_projection_through_bridge(recon, oracle):
    """The oracle's Gaussian projection (the kernel's CPU restatement) of the
    points through the bridged matrix puts their centres on the pinhole
    projection, and every point in front of the camera has ndc depth in
    (-1, 1)."""
    cams, images, pts, _ = recon
    by_id = {c.camera_id: c for c in cams}
    n = len(pts)
    quat = np.tile(np.float32([0, 0, 0, 1]), (n, 1))
    scale = np.full((n, 3), 0.01, np.float32)
    for (m16, shape), im in zip(colmap.views(cams, images, 0.1, 20.0), images):
        pxd, _, _, _ = oracle.gauss_project(pts.xyz.astype(np.float32), quat, scale, m16, shape)
        uv, z = colmap.project_pinhole(by_id[im.camera_id], im, pts.xyz)
        assert (z > 0.1).all()
        np.testing.assert_allclose(pxd[:, :2], uv, rtol=0, atol=5e-4)
        assert (np.abs(pxd[:, 2]) < 1).all()
        order = np.argsort(z)
        assert (np.diff(pxd[order, 2].astype(np.float64)) <= 1e-6).all()  # nearer = larger


def test_views_unknown_camera_raises(recon):
    cams, images, _, _ = recon
    with pytest.raises(ValueError, match="names camera"):
        colmap.views(cams[:1], images)


def test_gaussians_from_points_lattice():
    """On a lattice of spacing s every interior point has its 3 nearest
    neighbours at distance s, so its log-scale is log s."""
    s = 0.25
    g = np.stack(np.meshgrid(*[np.arange(6) * s] * 3, indexing="ij"), -1).reshape(-1, 3)
    pts = colmap.Points3D(np.arange(len(g), dtype=np.uint64), g,
                          np.tile(np.uint8([255, 0, 51]), (len(g), 1)), np.zeros(len(g)),
                          np.zeros(len(g) + 1, np.int64), np.zeros((0, 2), np.int32))
    out = colmap.gaussians_from_points(pts)
    np.testing.assert_allclose(out["log_scale"], np.log(s), rtol=1e-6)
    np.testing.assert_allclose(out["scale"], s, rtol=1e-6)
    np.testing.assert_array_equal(out["quat"], np.tile([0, 0, 0, 1], (len(g), 1)))
    np.testing.assert_allclose(out["rgb"], np.tile([1.0, 0.0, 0.2], (len(g), 1)), rtol=1e-6)
    np.testing.assert_array_equal(out["opacity"], np.float32(0.1))
    assert all(v.dtype == np.float32 for v in out.values())
    # duplicate points: distances 0 are clamped to the minimum scale
    dup = colmap.Points3D(np.arange(4, dtype=np.uint64), np.zeros((4, 3)),
                          np.zeros((4, 3), np.uint8), np.zeros(4), np.zeros(5, np.int64),
                          np.zeros((0, 2), np.int32))
    assert np.isfinite(colmap.gaussians_from_points(dup)["log_scale"]).all()


def test_camera_center_is_the_eye(recon):
    """camera.camera_center recovers the eye of look-at cameras and COLMAP's
    centre -R^T t of registered images from their world->ndc matrices."""
    from del_splat_dlpack import camera
    rng = np.random.default_rng(12)
    for _ in range(10):
        eye = rng.standard_normal(3) * 3
        m = camera.world2ndc(eye, (0.0, 0.0, 0.0), (320, 240))
        np.testing.assert_allclose(camera.camera_center(m), eye, rtol=0, atol=1e-9)
        np.testing.assert_allclose(camera.camera_center(camera.to_col_major(m)), eye, rtol=0,
                                   atol=5e-5)
    cams, images, _, _ = recon
    for (m16, _), im in zip(colmap.views(cams, images), images):
        r = colmap.qvec_to_rotation(im.qvec)
        np.testing.assert_allclose(camera.camera_center(m16), -r.T @ im.tvec, rtol=0, atol=5e-5)
