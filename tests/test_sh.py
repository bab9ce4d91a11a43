"""Spherical-harmonic colour: the oracle pinned by mathematics (the
reference reads SH coefficients, io_ply.rs:196, but never evaluates them,
so there is no reference output to compare with).

* basis: every one of the 16 functions equals the real form built from
  scipy's complex harmonics (Condon-Shortley phase): sqrt(2) Re Y_l^|m| for
  m > 0, sqrt(2) Im Y_l^|m| for m < 0, Y_l^0 for m = 0;
* the addition theorem sum_m Y_lm(d)^2 = (2l+1)/(4 pi) for every direction;
* the basis gradients against central differences of the basis;
* colour: the oracle's sum + 1/2 with the clamp at 0, against a numpy sum
  over the scipy basis, for every degree and coefficient stride;
* adjoint: the float64 twin against central differences in sh and xyz, and
  the clamp zeroing the gradient.
"""
import numpy as np
import pytest
import scipy.special as sp


def _scipy_real_basis(d):
    th, ph = np.arccos(np.clip(d[2], -1, 1)), np.arctan2(d[1], d[0])
    out = np.zeros(16)
    for l in range(4):
        for m in range(-l, l + 1):
            y = sp.sph_harm_y(l, abs(m), th, ph)
            out[l * l + l + m] = (np.sqrt(2) * y.real if m > 0 else
                                  np.sqrt(2) * y.imag if m < 0 else y.real)
    return out


def _dirs(n, seed):
    d = np.random.default_rng(seed).standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    # the poles and the axes, where the azimuth is degenerate
    return np.concatenate([d, np.eye(3), -np.eye(3)])


def test_basis_matches_scipy_real_harmonics(oracle):
    for d in _dirs(300, 0):
        np.testing.assert_allclose(oracle.sh_basis(d), _scipy_real_basis(d), rtol=0, atol=1e-13)


def test_basis_addition_theorem(oracle):
    for d in _dirs(100, 1):
        b = oracle.sh_basis(d)
        for l in range(4):
            band = b[l * l:(l + 1) * (l + 1)]
            assert abs(np.dot(band, band) - (2 * l + 1) / (4 * np.pi)) < 1e-13


def test_basis_lower_degree_is_a_prefix(oracle):
    d = _dirs(1, 2)[0]
    full = oracle.sh_basis(d, 3)
    for deg in range(3):
        nk = (deg + 1) ** 2
        np.testing.assert_array_equal(oracle.sh_basis(d, deg)[:nk], full[:nk])


def test_basis_gradient_central_differences(oracle):
    h = 1e-6
    for d in _dirs(20, 3):
        _, g = oracle.sh_basis(d, 3, with_grad=True)
        for a in range(3):
            e = np.zeros(3)
            e[a] = h
            fd = (oracle.sh_basis(d + e) - oracle.sh_basis(d - e)) / (2 * h)
            np.testing.assert_allclose(g[:, a], fd, rtol=0, atol=1e-8)


def _scene(n, K, seed, cam=(0.3, -0.2, 4.0)):
    rng = np.random.default_rng(seed)
    xyz = rng.uniform(-1, 1, (n, 3))
    sh = rng.normal(0, 0.4, (n, K, 3))
    return xyz, sh, np.array(cam)


@pytest.mark.parametrize("degree,K", [(0, 1), (1, 4), (2, 9), (3, 16), (1, 16), (2, 16)])
def test_colour_matches_scipy_sum(oracle, degree, K):
    xyz, sh, cam = _scene(200, K, 4 + degree)
    nk = (degree + 1) ** 2
    want = np.empty((len(xyz), 3))
    for i in range(len(xyz)):
        d = xyz[i] - cam
        b = _scipy_real_basis(d / np.linalg.norm(d))[:nk]
        want[i] = np.maximum(b @ sh[i, :nk] + 0.5, 0.0)
    np.testing.assert_allclose(oracle.sh_to_rgb(xyz, sh, cam, degree, np.float64), want,
                               rtol=0, atol=1e-13)
    np.testing.assert_allclose(oracle.sh_to_rgb(xyz, sh, cam, degree), want, rtol=0, atol=2e-6)
    assert (want == 0).any() or degree == 0  # some channels clamp at these magnitudes


def test_degree0_is_c0_dc_plus_half(oracle):
    xyz, sh, cam = _scene(50, 16, 5)
    c0 = 0.5 / np.sqrt(np.pi)
    np.testing.assert_allclose(oracle.sh_to_rgb(xyz, sh, cam, 0, np.float64),
                               np.maximum(c0 * sh[:, 0] + 0.5, 0), rtol=0, atol=1e-15)


def test_colour_depends_on_direction_only(oracle):
    """Moving a mean along the viewing ray keeps its colour."""
    xyz, sh, cam = _scene(50, 16, 6)
    far = cam + 2.5 * (xyz - cam)
    np.testing.assert_allclose(oracle.sh_to_rgb(far, sh, cam, 3, np.float64),
                               oracle.sh_to_rgb(xyz, sh, cam, 3, np.float64), rtol=0, atol=1e-13)


@pytest.mark.parametrize("degree", [1, 2, 3])
def test_adjoint_central_differences(oracle, degree):
    xyz, sh, cam = _scene(40, 16, 7 + degree)
    sh *= 0.5  # mostly unclamped; the clamp has its own test
    rng = np.random.default_rng(8)
    w = rng.standard_normal((len(xyz), 3))

    def loss(x, s):
        return float(np.sum(w * oracle.sh_to_rgb(x, s, cam, degree, np.float64)))

    d_sh, d_xyz = oracle.sh_to_rgb_bwd(xyz, sh, cam, degree, w, np.float64)
    nk = (degree + 1) ** 2
    assert not d_sh[:, nk:].any()
    h = 1e-6
    for i in range(0, len(xyz), 7):
        for a in range(3):
            xp, xm = xyz.copy(), xyz.copy()
            xp[i, a] += h
            xm[i, a] -= h
            fd = (loss(xp, sh) - loss(xm, sh)) / (2 * h)
            assert abs(d_xyz[i, a] - fd) <= 1e-7 * max(1.0, abs(fd)), (i, a, d_xyz[i, a], fd)
        for k in range(nk):
            for c in range(3):
                sp_, sm = sh.copy(), sh.copy()
                sp_[i, k, c] += h
                sm[i, k, c] -= h
                fd = (loss(xyz, sp_) - loss(xyz, sm)) / (2 * h)
                assert abs(d_sh[i, k, c] - fd) <= 1e-7 * max(1.0, abs(fd))


def test_adjoint_clamped_channel_has_no_gradient(oracle):
    xyz, sh, cam = _scene(1, 16, 9)
    sh[0, :, 0] = 0.0
    sh[0, 0, 0] = -5.0  # red channel: c0 * -5 + 0.5 < 0
    rgb = oracle.sh_to_rgb(xyz, sh, cam, 3, np.float64)
    assert rgb[0, 0] == 0.0
    d_sh, d_xyz = oracle.sh_to_rgb_bwd(xyz, sh, cam, 3, np.array([[1.0, 0.0, 0.0]]), np.float64)
    assert not d_sh.any() and not d_xyz.any()


def test_adjoint_add_mode_adds(oracle):
    xyz, sh, cam = _scene(30, 16, 10)
    w = np.random.default_rng(11).standard_normal((30, 3))
    d_sh, d_xyz = oracle.sh_to_rgb_bwd(xyz, sh, cam, 3, w, np.float64)
    a_sh, a_xyz = oracle.sh_to_rgb_bwd(xyz, sh, cam, 3, w, np.float64, d_sh=np.ones_like(sh),
                                       d_xyz=np.ones_like(xyz))
    np.testing.assert_allclose(a_sh, d_sh + 1.0, rtol=0, atol=1e-15)
    np.testing.assert_allclose(a_xyz, d_xyz + 1.0, rtol=0, atol=1e-15)
