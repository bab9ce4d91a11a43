"""Pins of the CPU oracle (no GPU). Each check is chosen so that a plausible
slip in the restatement -- a dropped term, a wrong sign or index, a transposed
operand -- fails: closed forms, brute force on tiny inputs, fp64 finite
differences, and the reference's own (commented) known-answer test."""
import numpy as np
import pytest

from _helpers import check_grad, small_scene, zbuffer_edge_points
from del_splat_dlpack import camera, scene


def _m(shape, eye=(0.0, 0.0, 3.0), target=(0.0, 0.0, 0.0), fov=60.0, near=0.01, far=100.0):
    return camera.to_col_major(camera.world2ndc(eye, target, shape, fov, near, far))


def _m64(shape, **kw):
    eye = kw.pop("eye", (0.0, 0.0, 3.0))
    M = camera.world2ndc(eye, (0.0, 0.0, 0.0), shape, **kw)
    return np.ascontiguousarray(M.T.reshape(16))


def _rand_quat(rng):
    q = rng.standard_normal(4)
    return q / np.linalg.norm(q)


# ---------------------------------------------------------------- del-geo ---
def test_homogeneous_and_jacobian_fd(oracle):
    rng = np.random.default_rng(0)
    shape = (640, 480)
    m = _m64(shape)
    M = m.reshape(4, 4).T
    for _ in range(20):
        x = rng.uniform(-1, 1, 3)
        p = oracle.gauss_project_one(x, [0, 0, 0, 1.0], [0.01] * 3, m, shape)
        y = M @ np.append(x, 1.0)
        assert np.isclose(p["w"], y[3])
        q = y[:3] / y[3]
        assert np.isclose(p["ndcz"], q[2])
        # pixel = (W/2 (qx + 1), H/2 (1 - qy))
        np.testing.assert_allclose(p["pix"], [shape[0] / 2 * (q[0] + 1), shape[1] / 2 * (1 - q[1])],
                                   rtol=1e-12)
        # Jacobian (column-major) against central differences of q(x)
        jac = p["jac"].reshape(3, 3).T
        fd = np.zeros((3, 3))
        for k in range(3):
            e = np.zeros(3)
            e[k] = 1e-6
            yp, ym = M @ np.append(x + e, 1), M @ np.append(x - e, 1)
            fd[:, k] = (yp[:3] / yp[3] - ym[:3] / ym[3]) / 2e-6
        np.testing.assert_allclose(jac, fd, rtol=1e-6, atol=1e-9)
        # world->pix Jacobian: A J with A's third column multiplying the ndc-z row
        A = np.array([[shape[0] / 2, 0, shape[0] / 2], [0, -shape[1] / 2, shape[1] / 2]])
        np.testing.assert_allclose(p["wpix"].reshape(3, 2).T, A @ jac, rtol=1e-12)


def test_quaternion_rotation(oracle):
    rng = np.random.default_rng(1)
    m = _m64((64, 64))
    p = oracle.gauss_project_one([0, 0, 0], [0, 0, 0, 1.0], [1, 1, 1], m, (64, 64))
    np.testing.assert_allclose(p["rot"], np.eye(3).reshape(-1), atol=0)
    # 90 degrees about z (Hamilton, (x,y,z,w)): x axis -> y axis
    s = np.sqrt(0.5)
    R = oracle.gauss_project_one([0, 0, 0], [0, 0, s, s], [1, 1, 1], m, (64, 64))["rot"]
    R = R.reshape(3, 3).T
    np.testing.assert_allclose(R @ [1, 0, 0], [0, 1, 0], atol=1e-15)
    np.testing.assert_allclose(R @ [0, 1, 0], [-1, 0, 0], atol=1e-15)
    for _ in range(50):
        R = oracle.gauss_project_one([0, 0, 0], _rand_quat(rng), [1, 1, 1], m, (64, 64))["rot"]
        R = R.reshape(3, 3).T
        np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-12)
        assert np.isclose(np.linalg.det(R), 1.0)


def test_projected_covariance_and_conic(oracle):
    rng = np.random.default_rng(2)
    shape = (800, 600)
    m = _m64(shape)
    for _ in range(50):
        x = rng.uniform(-0.5, 0.5, 3)
        q = _rand_quat(rng)
        s = np.exp(rng.normal(-3, 0.5, 3))
        p = oracle.gauss_project_one(x, q, s, m, shape)
        W = p["wpix"].reshape(3, 2).T
        R = p["rot"].reshape(3, 3).T
        cov = W @ R @ np.diag(s ** 2) @ R.T @ W.T
        np.testing.assert_allclose(p["cov"], [cov[0, 0], cov[0, 1], cov[1, 1]], rtol=1e-10)
        assert np.all(np.linalg.eigvalsh(cov) > 0)
        C = np.array([[p["conic"][0], p["conic"][1]], [p["conic"][1], p["conic"][2]]])
        np.testing.assert_allclose(C @ cov, np.eye(2), atol=1e-9)
        assert p["clamped"] == 0


def test_isotropic_on_axis_closed_form(oracle):
    """Isotropic sigma at the origin seen from (0,0,d): closed-form covariance
    including the ndc-z row the reference's world->pix Jacobian carries."""
    W_, H_ = 500, 400
    d, fov, n, f = 3.0, 60.0, 0.1, 50.0
    m = _m64((W_, H_), eye=(0.0, 0.0, d), fov_y_deg=fov, z_near=n, z_far=f)
    sig = 0.02
    p = oracle.gauss_project_one([0, 0, 0], [0, 0, 0, 1.0], [sig] * 3, m, (W_, H_))
    fy = 1.0 / np.tan(np.radians(fov) / 2)
    fx = fy / (W_ / H_)
    # ndc z(z_cam = -depth) = ((f+n) - 2fn/depth)/(n-f); world z = 3 - depth
    dzdx3 = -(2 * f * n / d ** 2) / (n - f)  # d ndc_z / d world_z
    row0 = np.array([W_ / 2 * fx / d, 0.0, W_ / 2 * dzdx3])
    row1 = np.array([0.0, -H_ / 2 * fy / d, H_ / 2 * dzdx3])
    cov = sig ** 2 * np.array([row0 @ row0, row0 @ row1, row1 @ row1])
    np.testing.assert_allclose(p["cov"], cov, rtol=1e-9)
    np.testing.assert_allclose(p["pix"], [W_ / 2, H_ / 2], atol=1e-9)


def test_safe_inverse_clamps_small_eigenvalues(oracle):
    rng = np.random.default_rng(3)
    for _ in range(30):
        U, _ = np.linalg.qr(rng.standard_normal((2, 2)))
        lam = np.array([rng.uniform(1e-7, 9e-6), rng.uniform(1e-3, 1.0)])
        A = U @ np.diag(lam) @ U.T
        out, c = oracle.safe_inverse_spd([A[0, 0], A[0, 1], A[1, 1]], 1e-5)
        assert c == 1
        ref = U @ np.diag(1 / np.maximum(lam, 1e-5)) @ U.T
        np.testing.assert_allclose(out, [ref[0, 0], ref[0, 1], ref[1, 1]], rtol=1e-9)
        out, c = oracle.safe_inverse_spd([A[0, 0] + 1, A[0, 1], A[1, 1] + 1], 1e-5)
        assert c == 0


def test_aabb_brute_force(oracle):
    """The box is the 3-sigma extent of the ellipse {d : d^T C d <= 9}."""
    rng = np.random.default_rng(4)
    m = _m64((600, 600))
    th = np.linspace(0, 2 * np.pi, 20001)
    for _ in range(20):
        p = oracle.gauss_project_one(rng.uniform(-0.5, 0.5, 3), _rand_quat(rng),
                                     np.exp(rng.normal(-3, 0.5, 3)), m, (600, 600))
        C = np.array([[p["conic"][0], p["conic"][1]], [p["conic"][1], p["conic"][2]]])
        L = np.linalg.cholesky(np.linalg.inv(C))
        pts = 3.0 * (L @ np.stack([np.cos(th), np.sin(th)]))
        ext = [pts[0].min(), pts[1].min(), pts[0].max(), pts[1].max()]
        np.testing.assert_allclose(np.array(p["aabb"]) - np.tile(p["pix"], 2), ext,
                                   rtol=1e-6, atol=1e-9)


def test_projection_f32_close_to_f64(oracle):
    s, m = small_scene(5000, 5, (512, 512))
    a = oracle.gauss_project(s["xyz"], s["quat"], s["scale"], m, (512, 512))
    b = oracle.gauss_project(s["xyz"], s["quat"], s["scale"], m, (512, 512), dtype=np.float64)
    for x, y in zip(a[:3], b[:3]):
        assert np.abs(x - y).max() / np.abs(y).max() < 1e-5


# ---------------------------------------------------------------- binning ---
def _brute_lists(boxes, depth, shape, ts):
    tw, th = (shape[0] + ts - 1) // ts, (shape[1] + ts - 1) // ts
    out = []
    key = np.maximum(depth.astype(np.float32) + np.float32(1), np.float32(0)).view(np.uint32)
    for t in range(tw * th):
        tx, ty = t % tw, t // tw
        x0, y0, x1, y1 = tx * ts, ty * ts, (tx + 1) * ts, (ty + 1) * ts
        # box overlaps the half-open tile [x0, x1) x [y0, y1)
        sel = [i for i, b in enumerate(boxes)
               if b[0] <= b[2] and b[1] <= b[3] and b[0] < x1 and b[2] >= x0 and b[1] < y1
               and b[3] >= y0]
        sel.sort(key=lambda i: (key[i], i))
        out.append(sel)
    return out


@pytest.mark.parametrize("ts", [8, 16, 32])
def test_binning_brute_force(oracle, ts):
    shape = (64, 48)
    s, m = small_scene(200, 6 + ts, shape)
    pxd, conic, aabb, _ = oracle.gauss_project(s["xyz"], s["quat"], s["scale"], m, shape)
    t2i, i2p, keys = oracle.tile2pnt_gauss(pxd, aabb, shape, ts, with_keys=True)
    lists = _brute_lists(aabb, pxd[:, 2], shape, ts)
    assert t2i[-1] == len(i2p) == sum(len(x) for x in lists)
    assert np.all(np.diff(t2i.astype(np.int64)) >= 0)
    for t, lst in enumerate(lists):
        assert list(i2p[t2i[t]:t2i[t + 1]]) == lst
    # keys: tile << 32 | bits(max(z+1, 0)), ascending
    assert np.all(np.diff(keys.astype(np.float64)) >= 0)
    tiles = (keys >> np.uint64(32)).astype(np.int64)
    for t in range(len(t2i) - 1):
        assert np.all(tiles[t2i[t]:t2i[t + 1]] == t)
    dk = np.maximum(pxd[:, 2] + np.float32(1), 0).astype(np.float32).view(np.uint32)
    np.testing.assert_array_equal((keys & np.uint64(0xFFFFFFFF)).astype(np.uint32), dk[i2p])


def test_tile_rect_edges(oracle):
    boxes = np.array([[15.999, 0, 16.0, 0.0], [16, 16, 32, 32], [-5, -5, -1, -1],
                      [np.inf, np.inf, -np.inf, -np.inf], [np.nan, 0, 1, 1],
                      [-1e30, -1e30, 1e30, 1e30], [9, 5, 3, 2], [20, 5, 10, 8]], np.float32)
    r = oracle.tile_rects(boxes, 16, 4, 3)
    np.testing.assert_array_equal(r[0], [0, 2, 0, 1])   # x1 = 16 touches tile 1
    np.testing.assert_array_equal(r[1], [1, 3, 1, 3])
    assert (r[2, 1] - r[2, 0]) * (r[2, 3] - r[2, 2]) == 0
    assert (r[3, 1] - r[3, 0]) * (r[3, 3] - r[3, 2]) == 0  # culled: inverted to infinity
    assert (r[4, 1] - r[4, 0]) * (r[4, 3] - r[4, 2]) == 0
    np.testing.assert_array_equal(r[5], [0, 4, 0, 3])
    # inverted inside one tile: the reference's floor loops still visit it
    np.testing.assert_array_equal(r[6], [0, 1, 0, 1])
    assert (r[7, 1] - r[7, 0]) * (r[7, 3] - r[7, 2]) == 0  # inverted across a tile boundary


def test_tile_rect_matches_reference_loops(oracle):
    """Against the loops of tile_acceleration.h:27-44 written out (ix from
    floor(x0 / ts) up to floor(x1 / ts), tiles outside the grid skipped), on
    random boxes of every orientation, partly off the grid."""
    rng = np.random.default_rng(3)
    ts, tw, th = 16, 7, 5
    n = 4000
    c = rng.uniform(-40, 150, (n, 2))
    hw = rng.uniform(-30, 60, (n, 2))  # negative half-widths: inverted boxes
    boxes = np.concatenate([c - hw, c + hw], 1).astype(np.float32)
    boxes[::7, 2] = boxes[::7, 0] - rng.uniform(0, 3, len(boxes[::7]))  # inverted in x only
    r = oracle.tile_rects(boxes, ts, tw, th)
    for i, b in enumerate(boxes):
        ix0, iy0 = int(np.floor(b[0] / np.float32(ts))), int(np.floor(b[1] / np.float32(ts)))
        ix1, iy1 = int(np.floor(b[2] / np.float32(ts))) + 1, int(np.floor(b[3] / np.float32(ts))) + 1
        want = {(ix, iy) for ix in range(ix0, ix1) if 0 <= ix < tw
                for iy in range(iy0, iy1) if 0 <= iy < th}
        got = {(ix, iy) for ix in range(r[i, 0], r[i, 1]) for iy in range(r[i, 2], r[i, 3])}
        assert got == want, (i, b, r[i])


# ----------------------------------------------------------------- raster ---
def test_threshold_equivalence():
    """The reference tests float alpha_sum > 0.999 (double); the kernels test
    alpha_sum >= 0.999f. Equivalent because 0.999f is the least float above
    the double 0.999."""
    f = np.float32(0.999)
    assert float(f) > 0.999
    assert float(np.nextafter(f, np.float32(0))) < 0.999


def test_single_and_stacked_splats_closed_form(oracle):
    shape = (16, 16)
    pxd = np.array([[8.5, 8.5, 0.5], [8.5, 8.5, 0.2]], np.float32)  # first is nearer
    conic = np.array([[0.1, 0.0, 0.1]] * 2, np.float32)
    aabb = np.array([[0, 0, 16, 16]] * 2, np.float32)
    op = np.array([0.6, 0.5], np.float32)
    rgb = np.array([[1, 0, 0], [0, 1, 0]], np.float32)
    t2i, i2p = oracle.tile2pnt_gauss(pxd, aabb, shape, 16)
    img, alpha, nwalk, tfinal = oracle.gauss_rasterize(shape, pxd, conic, aabb, op, rgb, 16, t2i,
                                                      i2p)
    # pixel centre (8.5, 8.5) is both means: e = opacity
    np.testing.assert_allclose(img[8, 8], [0.6, 0.5 * 0.4, 0.0], rtol=1e-6)
    np.testing.assert_allclose(alpha[8, 8], 0.6 + 0.5 * 0.4, rtol=1e-6)
    np.testing.assert_allclose(tfinal[8, 8], 0.4 * 0.5, rtol=1e-6)  # after both splats
    # one pixel to the right: q = 0.1 * 1^2
    e = np.exp(-0.05)
    np.testing.assert_allclose(img[8, 9], [0.6 * e, 0.5 * e * (1 - 0.6 * e), 0], rtol=1e-6)


def _py_raster(shape, pxd, conic, aabb, op, rgb, ts, t2i, i2p):
    """Independent pure-Python restatement of splat_gauss.cu:56-109."""
    w, h = shape
    tw = (w + ts - 1) // ts
    img = np.zeros((h, w, 3), np.float64)
    for iy in range(h):
        for ix in range(w):
            t = (iy // ts) * tw + ix // ts
            px, py = ix + 0.5, iy + 0.5
            a_sum, occ = 0.0, 1.0
            for idx in range(int(t2i[t + 1]) - 1, int(t2i[t]) - 1, -1):
                p = i2p[idx]
                b = aabb[p]
                if not (b[0] <= px <= b[2] and b[1] <= py <= b[3]):
                    continue
                dx, dy = px - pxd[p, 0], py - pxd[p, 1]
                e = np.exp(-0.5 * (conic[p, 0] * dx * dx + 2 * conic[p, 1] * dx * dy
                                   + conic[p, 2] * dy * dy)) * op[p]
                eo = occ * e
                img[iy, ix] += rgb[p] * eo
                occ *= 1 - e
                a_sum += eo
                if a_sum > 0.999:
                    break
    return img


def test_raster_against_python_and_naive(oracle):
    shape = (40, 32)
    s, m = small_scene(60, 7, shape, distance=2.5)
    pxd, conic, aabb, _ = oracle.gauss_project(s["xyz"], s["quat"], s["scale"], m, shape)
    t2i, i2p = oracle.tile2pnt_gauss(pxd, aabb, shape, 8)
    img = oracle.gauss_rasterize(shape, pxd, conic, aabb, s["opacity"], s["rgb"], 8, t2i, i2p)[0]
    ref = _py_raster(shape, pxd, conic, aabb, s["opacity"], s["rgb"], 8, t2i, i2p)
    assert np.abs(img - ref).max() < 1e-5
    assert np.all((pxd[:, 2] > -1) & (pxd[:, 2] < 1))
    naive = oracle.gauss_render_sort_depth(pxd, conic, aabb, s["opacity"], s["rgb"], shape)
    np.testing.assert_array_equal(img, naive)


# -------------------------------------------------------------- backward ---
def test_transfer_known_answer(oracle):
    """splat_gaussian2.rs:245-266 (commented in the reference), revived."""
    a = np.array([0.4, 0.7, 0.5, 0.3], np.float32)
    eps = 1e-4
    for i in range(4):
        t0 = oracle.transfer(a, i)
        for j in range(4):
            a1 = a.copy()
            a1[j] += eps
            fd = (oracle.transfer(a1, i) - t0) / eps
            assert abs(fd - oracle.dw_transfer(a, i, j)) < 2e-4


def _raster_case(oracle, n=40, shape=(32, 24), seed=8, ts=8):
    s, m = small_scene(n, seed, shape, distance=2.2)
    pxd, conic, aabb, _ = oracle.gauss_project(s["xyz"], s["quat"], s["scale"], m, shape)
    t2i, i2p = oracle.tile2pnt_gauss(pxd, aabb, shape, ts)
    return s, m, pxd, conic, aabb, t2i, i2p


def test_raster_adjoint_fd_fp64(oracle):
    """Raster adjoint vs central differences in fp64, with the walk lengths and
    boxes frozen (the function the adjoint differentiates)."""
    shape, ts = (32, 24), 8
    s, m, pxd, conic, aabb, t2i, i2p = _raster_case(oracle, shape=shape, ts=ts)
    f64 = np.float64
    P = dict(pxd=pxd.astype(f64), conic=conic.astype(f64), op=s["opacity"].astype(f64),
             rgb=s["rgb"].astype(f64))
    _, _, nwalk, _ = oracle.gauss_rasterize(shape, pxd, conic, aabb, s["opacity"], s["rgb"], ts,
                                            t2i, i2p)
    rng = np.random.default_rng(9)
    wr = rng.standard_normal((shape[1], shape[0], 3))
    wa = rng.standard_normal((shape[1], shape[0]))

    def loss(P):
        img, alpha, _, _ = oracle.gauss_rasterize(shape, P["pxd"], P["conic"], aabb, P["op"],
                                                  P["rgb"], ts, t2i, i2p, dtype=f64,
                                                  frozen_nwalk=nwalk)
        return (img * wr).sum() + (alpha * wa).sum()

    img, alpha, nw64, tl64 = oracle.gauss_rasterize(shape, P["pxd"], P["conic"], aabb, P["op"],
                                                    P["rgb"], ts, t2i, i2p, dtype=f64,
                                                    frozen_nwalk=nwalk)
    dpix, dcon, dop, drgb = oracle.gauss_rasterize_bwd(shape, P["pxd"], P["conic"], aabb, P["op"],
                                                       P["rgb"], ts, t2i, i2p, nwalk, tl64, wr,
                                                       wa, dtype=f64)
    eps = 1e-6
    checks = [("pxd", 0, dpix[:, 0]), ("pxd", 1, dpix[:, 1]), ("conic", 0, dcon[:, 0]),
              ("conic", 1, dcon[:, 1]), ("conic", 2, dcon[:, 2]), ("op", None, dop),
              ("rgb", 0, drgb[:, 0]), ("rgb", 2, drgb[:, 2])]
    for name, c, an in checks:
        for i in range(0, len(P["op"]), 3):
            Pp = {k: v.copy() for k, v in P.items()}
            Pm = {k: v.copy() for k, v in P.items()}
            if c is None:
                Pp[name][i] += eps
                Pm[name][i] -= eps
            else:
                Pp[name][i, c] += eps
                Pm[name][i, c] -= eps
            fd = (loss(Pp) - loss(Pm)) / (2 * eps)
            assert abs(fd - an[i]) <= 1e-5 * max(1.0, abs(fd)), (name, c, i, fd, an[i])


def test_projection_adjoint_fd_fp64(oracle):
    rng = np.random.default_rng(10)
    shape = (320, 240)
    n = 30
    s, m = small_scene(n, 11, shape)
    m64 = m.astype(np.float64)
    x = s["xyz"].astype(np.float64)
    q = s["quat"].astype(np.float64)
    ls = s["log_scale"].astype(np.float64)
    gp = rng.standard_normal((n, 2))
    gc = rng.standard_normal((n, 3)) * 100

    def loss(x, q, ls):
        pxd, conic, _, _ = oracle.gauss_project(x, q, np.exp(ls), m64, shape, dtype=np.float64)
        return (pxd[:, :2] * gp).sum() + (conic * gc).sum()

    dx, dls, dq = oracle.gauss_project_bwd(x, q, np.exp(ls), m64, shape, gp, gc, log_scale=True,
                                           dtype=np.float64)
    eps = 1e-7
    for arr, an, name in ((x, dx, "xyz"), (ls, dls, "log_scale"), (q, dq, "quat")):
        for i in range(n):
            for c in range(arr.shape[1]):
                a_p, a_m = arr.copy(), arr.copy()
                a_p[i, c] += eps
                a_m[i, c] -= eps
                args_p = {"xyz": (a_p, q, ls), "quat": (x, a_p, ls), "log_scale": (x, q, a_p)}[name]
                args_m = {"xyz": (a_m, q, ls), "quat": (x, a_m, ls), "log_scale": (x, q, a_m)}[name]
                fd = (loss(*args_p) - loss(*args_m)) / (2 * eps)
                assert abs(fd - an[i, c]) <= 1e-5 * max(1.0, abs(fd)), (name, i, c, fd, an[i, c])


def test_end_to_end_fd_64_gaussians(oracle):
    """North-star recipe: loss = sum (rgb - target)^2, gradients w.r.t. means,
    log-scales, quaternions, opacities and colours of 64 random Gaussians by
    central differences (eps 1e-3, rel tol 1e-2) in fp64 against the fp64
    adjoint chain; plus the fp32 chain against the fp64 one."""
    shape, ts = (96, 72), 16
    n = 1500
    s, m = small_scene(n, 12, shape)
    f64 = np.float64
    m64 = m.astype(f64)
    P0 = dict(xyz=s["xyz"].astype(f64), ls=s["log_scale"].astype(f64), q=s["quat"].astype(f64),
              op=s["opacity"].astype(f64), rgb=s["rgb"].astype(f64))
    pxd32, conic32, aabb, _ = oracle.gauss_project(s["xyz"], s["quat"], s["scale"], m, shape)
    t2i, i2p = oracle.tile2pnt_gauss(pxd32, aabb, shape, ts)
    _, _, nwalk, _ = oracle.gauss_rasterize(shape, pxd32, conic32, aabb, s["opacity"], s["rgb"],
                                            ts, t2i, i2p)
    target = scene.target_image(shape, 13).astype(f64)

    def fwd(P):
        pxd, conic, _, _ = oracle.gauss_project(P["xyz"], P["q"], np.exp(P["ls"]), m64, shape,
                                                dtype=f64)
        return pxd, conic, oracle.gauss_rasterize(shape, pxd, conic, aabb, P["op"], P["rgb"], ts,
                                                  t2i, i2p, dtype=f64, frozen_nwalk=nwalk)

    def loss(P):
        return ((fwd(P)[2][0] - target) ** 2).sum()

    pxd, conic, (img, alpha, _, tfinal) = fwd(P0)
    d_img = 2 * (img - target)
    dpix, dcon, dop, drgb = oracle.gauss_rasterize_bwd(shape, pxd, conic, aabb, P0["op"],
                                                       P0["rgb"], ts, t2i, i2p, nwalk, tfinal,
                                                       d_img, dtype=f64)
    dxyz, dls, dq = oracle.gauss_project_bwd(P0["xyz"], P0["q"], np.exp(P0["ls"]), m64, shape,
                                             dpix, dcon, log_scale=True, dtype=f64)
    grads = dict(xyz=dxyz, ls=dls, q=dq, op=dop, rgb=drgb)
    # 64 Gaussians that touch the image
    touched = np.unique(i2p)
    pick = np.random.default_rng(14).choice(touched, 64, replace=False)
    eps = 1e-3
    for name, g in grads.items():
        fd = np.zeros((64,) + g.shape[1:])
        for r, i in enumerate(pick):
            for c in range(1 if g.ndim == 1 else g.shape[1]):
                Pp = {k: v.copy() for k, v in P0.items()}
                Pm = {k: v.copy() for k, v in P0.items()}
                if g.ndim == 1:
                    Pp[name][i] += eps
                    Pm[name][i] -= eps
                    fd[r] = (loss(Pp) - loss(Pm)) / (2 * eps)
                else:
                    Pp[name][i, c] += eps
                    Pm[name][i, c] -= eps
                    fd[r, c] = (loss(Pp) - loss(Pm)) / (2 * eps)
        an = g[pick]
        err = np.abs(fd - an).max() / np.abs(an).max()
        assert err <= 1e-2, (name, err)
    # fp32 adjoint chain vs fp64
    img32, alpha32, nw32, tl32 = oracle.gauss_rasterize(shape, pxd32, conic32, aabb, s["opacity"],
                                                        s["rgb"], ts, t2i, i2p)
    d32 = oracle.gauss_rasterize_bwd(shape, pxd32, conic32, aabb, s["opacity"], s["rgb"], ts, t2i,
                                     i2p, nw32, tl32, (2 * (img32 - target)).astype(np.float32))
    g32 = oracle.gauss_project_bwd(s["xyz"], s["quat"], s["scale"], m, shape, d32[0], d32[1],
                                   log_scale=True)
    for a, b in zip(list(g32) + [d32[2], d32[3]], [dxyz, dls, dq, dop, drgb]):
        assert np.abs(a - b).max() / np.abs(b).max() < 1e-3


# --------------------------------------------------------- points/circles ---
def test_zbuffer_brute_force(oracle):
    rng = np.random.default_rng(15)
    shape = (24, 16)
    n = 300
    pxz = np.stack([rng.uniform(-2, 26, n), rng.uniform(-2, 18, n),
                    rng.uniform(-1.2, 1.2, n)], 1).astype(np.float32)
    pxz[5:10, 2] = pxz[4, 2]  # depth ties -> smallest index
    pxz[5:10, :2] = pxz[4, :2]
    rad = rng.uniform(0.5, 3.0, n).astype(np.float32)
    rgb = rng.random((n, 3)).astype(np.float32)
    for r in (None, rad):
        keys, img = oracle.zbuffer(pxz, rgb, shape, pixrad=r)
        best = np.full((shape[1], shape[0]), np.iinfo(np.uint64).max, np.uint64)
        for i in range(n):
            z = pxz[i, 2]
            if z < -1 or z > 1:
                continue
            u = np.float32(1.0 - (float(z) + 1.0) * 0.5)
            key = (np.uint64(u.view(np.uint32)) << np.uint64(32)) | np.uint64(i)
            for iy in range(shape[1]):
                for ix in range(shape[0]):
                    if r is None:
                        hit = np.floor(pxz[i, 0]) == ix and np.floor(pxz[i, 1]) == iy
                    else:
                        dx = np.float32(pxz[i, 0] - np.float32(ix + 0.5))
                        dy = np.float32(pxz[i, 1] - np.float32(iy + 0.5))
                        hit = np.sqrt(np.float32(dx * dx + dy * dy)) <= r[i]
                    if hit and key < best[iy, ix]:
                        best[iy, ix] = key
        np.testing.assert_array_equal(keys, best)
        win = best != np.iinfo(np.uint64).max
        idx = (best & np.uint64(0xFFFFFFFF)).astype(np.int64)
        np.testing.assert_array_equal(img[win], rgb[idx[win]])
        assert np.all(img[~win] == 0)


def test_zbuffer_edge_cases_brute_force(oracle):
    """Circle z-buffer on the edge rows of zbuffer_edge_points against the
    per-pixel brute force (float32 distance test, min key)."""
    rng = np.random.default_rng(16)
    shape = (20, 14)
    pxz, rad = zbuffer_edge_points(rng, shape, 80)
    rgb = rng.random((pxz.shape[0], 3)).astype(np.float32)
    with np.errstate(invalid="ignore", over="ignore"):
        keys, _ = oracle.zbuffer(pxz, rgb, shape, pixrad=rad)
        best = np.full((shape[1], shape[0]), np.iinfo(np.uint64).max, np.uint64)
        for i in range(pxz.shape[0]):
            z = pxz[i, 2]
            if not (-1 <= z <= 1) or not (rad[i] >= 0):
                continue
            u = np.float32(1.0 - (float(z) + 1.0) * 0.5)
            key = (np.uint64(u.view(np.uint32)) << np.uint64(32)) | np.uint64(i)
            ix, iy = np.meshgrid(np.arange(shape[0]), np.arange(shape[1]))
            dx = pxz[i, 0] - (ix.astype(np.float32) + np.float32(0.5))
            dy = pxz[i, 1] - (iy.astype(np.float32) + np.float32(0.5))
            hit = np.sqrt(dx * dx + dy * dy) <= rad[i]
            best = np.where(hit & (key < best), key, best)
    np.testing.assert_array_equal(keys, best)
    assert (keys != np.iinfo(np.uint64).max).all()  # the infinite radius covers every pixel


def test_point_painter_matches_python(oracle):
    rng = np.random.default_rng(16)
    shape = (20, 10)
    n = 500
    pxz = np.stack([rng.uniform(-3, 23, n), rng.uniform(-3, 13, n),
                    rng.uniform(-1.1, 1.1, n)], 1).astype(np.float32)
    pxz[::7, 2] = pxz[0, 2]
    rgb = rng.random((n, 3)).astype(np.float32)
    img = oracle.render_pix_sort_depth(pxz, rgb, shape)
    ref = np.zeros((shape[1], shape[0], 3), np.float32)
    order = sorted(range(n), key=lambda i: np.float32(pxz[i, 2] + np.float32(1)))  # stable
    for i in order:
        z = pxz[i, 2]
        if z <= -1 or z >= 1:
            continue
        ix = int(pxz[i, 0]) if pxz[i, 0] > 0 else 0
        iy = int(pxz[i, 1]) if pxz[i, 1] > 0 else 0
        if ix >= shape[0] or iy >= shape[1]:
            continue
        ref[iy, ix] = rgb[i]
    np.testing.assert_array_equal(img, ref)


def test_circle_radius_on_axis(oracle):
    """A sphere of radius r at the origin seen from distance d: pixel radius
    f H r / (2 d) (both axes agree when the aspect is W/H)."""
    W_, H_, d = 640, 480, 3.0
    m = _m((W_, H_), eye=(0.0, 0.0, d))
    pxd, rad = oracle.circle_project(np.zeros((1, 3), np.float32), 0.01, m, (W_, H_))
    fy = 1.0 / np.tan(np.radians(60.0) / 2)
    np.testing.assert_allclose(rad[0], fy * H_ * 0.01 / (2 * d), rtol=1e-4)
    np.testing.assert_allclose(pxd[0, :2], [W_ / 2, H_ / 2], atol=1e-4)


def test_circle_painter_against_python(oracle):
    rng = np.random.default_rng(17)
    shape, ts = (40, 24), 8
    n = 80
    pxd = np.stack([rng.uniform(0, 40, n), rng.uniform(0, 24, n), rng.uniform(-1, 1, n)],
                   1).astype(np.float32)
    rad = rng.uniform(0.5, 4, n).astype(np.float32)
    rgb = rng.random((n, 3)).astype(np.float32)
    t2i, i2p = oracle.tile2pnt_circle(pxd, rad, shape, ts)
    img = oracle.circle_rasterize(shape, pxd, rad, rgb, ts, t2i, i2p)
    tw = (shape[0] + ts - 1) // ts
    for iy in range(shape[1]):
        for ix in range(shape[0]):
            t = (iy // ts) * tw + ix // ts
            c = np.zeros(3, np.float32)
            for k in range(t2i[t], t2i[t + 1]):
                p = i2p[k]
                dx = np.float32(pxd[p, 0] - np.float32(ix + 0.5))
                dy = np.float32(pxd[p, 1] - np.float32(iy + 0.5))
                if np.sqrt(np.float32(dx * dx + dy * dy)) > rad[p]:
                    continue
                c = rgb[p]
            np.testing.assert_array_equal(img[iy, ix], c)


def test_adam_closed_form_and_torch(oracle):
    import torch
    rng = np.random.default_rng(18)
    p0 = rng.standard_normal(1000).astype(np.float32)
    p = p0.copy()
    m1 = np.zeros_like(p)
    m2 = np.zeros_like(p)
    g = rng.standard_normal(1000).astype(np.float32)
    oracle.adam(p, g, m1, m2, 1e-3, 0.9, 0.999, 1e-8, 1)
    # first step: m_hat = g, v_hat = g^2 -> p0 - lr g / (|g| + eps)
    np.testing.assert_allclose(p, p0 - 1e-3 * g / (np.abs(g) + 1e-8), rtol=1e-5, atol=1e-7)
    tp = torch.tensor(p0, requires_grad=True)
    opt = torch.optim.Adam([tp], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    p = p0.copy()
    m1[:] = 0
    m2[:] = 0
    for step in range(1, 6):
        gs = (g * step).astype(np.float32)
        tp.grad = torch.tensor(gs)
        opt.step()
        oracle.adam(p, gs, m1, m2, 1e-3, 0.9, 0.999, 1e-8, step)
    np.testing.assert_allclose(p, tp.detach().numpy(), rtol=1e-5, atol=1e-6)


# ------------------------------------------- planted exp perturbation, FD ---
def test_exp_ulp_perturbation_and_self_flips(oracle):
    """exp_ulp = 0 is the oracle's raster; moving every exp one ulp flips the
    termination of a few pixels only, each at the 0.999 threshold, and moves
    those pixels by less than the ~1e-3 of transmittance left there."""
    shape, ts = (256, 192), 16
    s, m = small_scene(20000, 70, shape)
    pxd, conic, aabb, _ = oracle.gauss_project(s["xyz"], s["quat"], s["scale"], m, shape)
    t2i, i2p = oracle.tile2pnt_gauss(pxd, aabb, shape, ts)
    img, alpha, nwalk, tfinal = oracle.gauss_rasterize(shape, pxd, conic, aabb, s["opacity"],
                                                      s["rgb"], ts, t2i, i2p)
    assert (alpha >= 0.999).mean() > 0.05
    z = oracle.gauss_rasterize_exp_ulp(shape, pxd, conic, aabb, s["opacity"], s["rgb"], ts, t2i,
                                       i2p, 0)
    for a, b in zip(z, (img, alpha, nwalk, tfinal)):
        np.testing.assert_array_equal(a, b)
    for u in (1, -1, 3):
        img2, alpha2, nw2, _ = oracle.gauss_rasterize_exp_ulp(shape, pxd, conic, aabb,
                                                              s["opacity"], s["rgb"], ts, t2i,
                                                              i2p, u)
        flip = nw2 != nwalk
        assert flip.mean() < 1e-3
        assert (np.abs(np.maximum(alpha, alpha2)[flip] - 0.999) < 2e-3).all()
        assert np.abs(img2 - img).max() <= 1.1e-3
        assert np.abs(img2 - img)[~flip].max() <= 1e-5
        if u > 0:  # larger weights: at least as much alpha wherever the walk is the same
            assert (alpha2[~flip] >= alpha[~flip] - 1e-6).all()
    own = oracle.self_flips(shape, pxd, conic, aabb, s["opacity"], s["rgb"], ts, t2i, i2p, nwalk)
    assert own.shape == nwalk.shape and own.sum() < 1e-3 * own.size


def test_tile_loss_and_local_fd_match_the_adjoint(oracle):
    """The tile-rectangle L2 loss over all tiles equals the full frozen-walk
    frame's loss, and central differences through it (the full-frame FD
    recipe of the GPU gradient test, _fd.fd_gradients) reproduce the oracle's
    own fp64 adjoint chain for 12 Gaussians."""
    from _fd import fd_gradients, visible_pick
    f64 = np.float64
    shape, ts = (200, 150), 16
    s, m = small_scene(3000, 71, shape)
    pxd32, conic32, aabb, _ = oracle.gauss_project(s["xyz"], s["quat"], s["scale"], m, shape)
    t2i, i2p = oracle.tile2pnt_gauss(pxd32, aabb, shape, ts)
    _, _, nwalk, _ = oracle.gauss_rasterize(shape, pxd32, conic32, aabb, s["opacity"], s["rgb"],
                                            ts, t2i, i2p)
    target = scene.target_image(shape, 72)
    pxd, conic, _, _ = oracle.gauss_project(s["xyz"].astype(f64), s["quat"].astype(f64),
                                            np.exp(s["log_scale"].astype(f64)), m.astype(f64),
                                            shape, dtype=f64)
    img, alpha, _, tfinal = oracle.gauss_rasterize(shape, pxd, conic, aabb, s["opacity"], s["rgb"],
                                                  ts, t2i, i2p, dtype=f64, frozen_nwalk=nwalk)
    full = ((img - target.astype(f64)) ** 2).sum()
    tw, th = (shape[0] + ts - 1) // ts, (shape[1] + ts - 1) // ts
    tiles = oracle.gauss_l2_loss_tiles(shape, np.ascontiguousarray(pxd),
                                       np.ascontiguousarray(conic), aabb,
                                       s["opacity"].astype(f64), s["rgb"].astype(f64), ts, t2i,
                                       i2p, nwalk.astype(np.uint32), target, (0, tw - 1, 0, th - 1))
    assert abs(tiles - full) <= 1e-12 * full
    dpix, dcon, dop, drgb = oracle.gauss_rasterize_bwd(shape, pxd, conic, aabb, s["opacity"],
                                                       s["rgb"], ts, t2i, i2p, nwalk, tfinal,
                                                       2 * (img - target), dtype=f64)
    dxyz, dls, dq = oracle.gauss_project_bwd(s["xyz"], s["quat"], np.exp(s["log_scale"]), m,
                                             shape, dpix, dcon, log_scale=True, dtype=f64)
    pick = visible_pick(drgb, 12, 73)
    fd = fd_gradients(oracle, s, m, shape, ts, t2i, i2p, aabb, nwalk, target, pick)
    for name, an in (("xyz", dxyz), ("log_scale", dls), ("quat", dq), ("opacity", dop),
                     ("rgb", drgb)):
        check_grad(an[pick], fd[name], f"fd {name}")


def test_multithreaded_oracle_matches_sequential(oracle):
    """The all-core CPU baseline (oracle.set_threads > 1) computes the same
    frame as the sequential parity oracle: projection, binning and raster
    bit-exact (the binning as a per-thread stable counting sort by tile, then
    per-tile depth sorts), adjoints equal up to the order of float sums."""
    import os
    shape, ts = (240, 176), 16
    s, m = small_scene(20000, 74, shape)
    ref = {}
    for threads in (1, min(4, os.cpu_count() or 1)):
        assert oracle.set_threads(threads) >= 1
        try:
            pxd, conic, aabb, _ = oracle.gauss_project(s["xyz"], s["quat"], s["scale"], m, shape)
            t2i, i2p, keys = oracle.tile2pnt_gauss(pxd, aabb, shape, ts, with_keys=True)
            img, alpha, nwalk, tfinal = oracle.gauss_rasterize(shape, pxd, conic, aabb,
                                                              s["opacity"], s["rgb"], ts, t2i, i2p)
            d_img = (2 * img - 1).astype(np.float32)
            fn = oracle.gauss_rasterize_bwd if threads == 1 else oracle.gauss_rasterize_bwd_mt
            d = fn(shape, pxd, conic, aabb, s["opacity"], s["rgb"], ts, t2i, i2p, nwalk, tfinal,
                   d_img)
            g = oracle.gauss_project_bwd(s["xyz"], s["quat"], s["scale"], m, shape, d[0], d[1])
        finally:
            oracle.set_threads(1)
        out = [pxd, conic, aabb, t2i, i2p, keys, img, alpha, nwalk, tfinal] + list(d) + list(g)
        if threads == 1:
            ref = out
            continue
        for a, b in zip(out[:10], ref[:10]):
            np.testing.assert_array_equal(a, b)
        for a, b in zip(out[10:], ref[10:]):
            check_grad(a, b, "multithreaded adjoint", rel=1e-4, abs_frac=1e-6)


def test_staged_opacity_bound_keeps_one_minus_e_above_2_pow_minus_24():
    """The HIP walks need no per-pixel floor on 1 - e: they clamp the splat
    weight g to [0, 1] and stage the opacity at most 1 - 2^-24 (the largest
    float32 below 1). float32 products round to nearest, monotonically, so
    e = g * o never exceeds 1 - 2^-24 and 1 - e >= 2^-24, exact for
    e >= 1/2. Checked for every float32 g in [1 - 2^-12, 1] (where e comes
    closest to 1) and a spread of smaller ones, against the float64 value:
    the subtraction is exact, so 1 - e matches the float64 difference."""
    f32 = np.float32
    o = f32(1.0) - f32(2.0 ** -24)
    assert o == np.nextafter(f32(1.0), f32(0.0))
    lo = np.float32(1.0 - 2.0 ** -12).view(np.uint32)
    hi = np.float32(1.0).view(np.uint32)
    g = np.arange(lo, hi + 1, dtype=np.uint32).view(np.float32)
    g = np.concatenate([g, np.linspace(0, 1, 100001, dtype=np.float32), [f32(0.5), f32(0.75)]])
    e = g * o
    assert e.dtype == np.float32
    assert (e <= o).all()
    ome = f32(1.0) - e
    assert (ome >= f32(2.0 ** -24)).all()
    big = e >= f32(0.5)
    np.testing.assert_array_equal(ome[big].astype(np.float64), 1.0 - e[big].astype(np.float64))
    # the closest case: g = 1 gives 1 - e = 2^-24 exactly, the oracle's floor
    assert f32(1.0) - f32(1.0) * o == f32(2.0 ** -24)


def test_opaque_splat_on_a_pixel_centre_gives_finite_gradients(oracle):
    """An opacity-1 splat whose mean is a pixel centre has e = 1 there: the
    walk ends on it, the stored transmittance is floored at 2^-24 of the one
    in front of it (never 0) and the adjoint divides it back exactly, so no
    gradient is 0 * inf. A second splat behind it never contributes there."""
    shape = (16, 16)
    pxd = np.array([[8.5, 8.5, 0.5], [8.5, 8.5, 0.2]], np.float32)  # first is nearer
    conic = np.array([[0.1, 0.0, 0.1]] * 2, np.float32)
    aabb = np.array([[0, 0, 16, 16]] * 2, np.float32)
    op = np.array([1.0, 0.5], np.float32)
    rgb = np.array([[0.2, 0.3, 0.4], [0, 1, 0]], np.float32)
    t2i, i2p = oracle.tile2pnt_gauss(pxd, aabb, shape, 16)
    img, alpha, nwalk, tfinal = oracle.gauss_rasterize(shape, pxd, conic, aabb, op, rgb, 16, t2i,
                                                       i2p)
    np.testing.assert_array_equal(img[8, 8], rgb[0])  # e = 1: only the first splat
    assert alpha[8, 8] == 1.0 and nwalk[8, 8] == 1
    assert tfinal[8, 8] == np.float32(2.0 ** -24)
    d_img = np.ones_like(img)
    for dt in (np.float32, np.float64):
        g = oracle.gauss_rasterize_bwd(shape, pxd, conic, aabb, op, rgb, 16, t2i, i2p, nwalk,
                                       tfinal, d_img, dtype=dt)
        for a in g:
            assert np.isfinite(a).all()
        # at the opaque splat's centre pixel dL/drgb = T e = 1 per channel
        assert (g[3][:3] >= 1.0).all()
