"""Shared test helpers: seeded scenes and GPU pipeline runs."""
import numpy as np

from del_splat_dlpack import scene


def small_scene(n, seed, img_shape=(256, 192), radius=1.0, distance=3.0):
    s = scene.gaussians(n, seed, radius=radius)
    m = scene.camera.to_col_major(
        scene.camera.world2ndc((0.0, 0.0, distance), (0.0, 0.0, 0.0), img_shape))
    return s, m


def max_rel(a, b):
    """Largest absolute difference over the largest magnitude of b (for
    values that should agree to rounding everywhere; gradients use
    check_grad)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = max(np.abs(b).max(), 1e-30)
    return np.abs(a - b).max() / den


def check_grad(got, ref, name="", rel=1e-2, band=1e-2, abs_frac=1e-3, max_rel_tol=None):
    """Gradient agreement per element. Entries with |ref| above band * max|ref|
    must agree to rel of their own magnitude; the others to abs_frac *
    max|ref| (fp32 sums in another order leave absolute, not relative, noise
    on entries that are sums of cancelling terms). With max_rel_tol, every
    entry must also lie within max_rel_tol * max|ref|. Prints how many
    entries each band holds and the worst error in each; returns them."""
    g = np.asarray(got, np.float64).reshape(-1)
    r = np.asarray(ref, np.float64).reshape(-1)
    assert g.shape == r.shape, (name, g.shape, r.shape)
    assert np.isfinite(g).all(), f"{name}: non-finite gradient"
    top = max(np.abs(r).max(initial=0.0), 1e-30)
    d = np.abs(g - r)
    big = np.abs(r) > band * top
    worst_rel = float((d[big] / np.abs(r[big])).max(initial=0.0))
    worst_abs = float(d[~big].max(initial=0.0) / top)
    print(f"{name}: {int(big.sum())} entries > {band:g} max (worst rel {worst_rel:.2e}), "
          f"{int((~big).sum())} below (worst {worst_abs:.2e} of max)", flush=True)
    assert worst_rel <= rel, (name, "relative", worst_rel)
    assert worst_abs <= abs_frac, (name, "absolute", worst_abs)
    if max_rel_tol is not None:
        assert d.max(initial=0.0) / top <= max_rel_tol, (name, "max_rel", d.max() / top)
    return worst_rel, worst_abs


def check_grad_vs_truth(got, ref32, ref64, name="", rel=1e-2, abs_frac=1e-4, factor=2.0):
    """Accuracy no worse than the oracle's own float32 arithmetic, against the
    oracle's float64 twin ref64 on the same frame: with the per-element error
    e = |x - ref64| / (|ref64| + abs_frac max|ref64|) of the result (got) and
    of the float32 oracle (ref32), the result may have at most factor times
    as many entries with e > rel (plus 4), and its worst and its RMS e may be
    at most factor times the oracle's (plus rel). For gradients whose float32
    evaluation is ill-conditioned (the inverse of a nearly singular
    covariance), where two float32 results differ by rounding noise of the
    same size and comparing them element by element pins nothing."""
    g = np.asarray(got, np.float64).reshape(-1)
    r32 = np.asarray(ref32, np.float64).reshape(-1)
    r64 = np.asarray(ref64, np.float64).reshape(-1)
    assert np.isfinite(g).all(), f"{name}: non-finite gradient"
    top = max(np.abs(r64).max(initial=0.0), 1e-30)
    den = np.abs(r64) + abs_frac * top
    eh, eo = np.abs(g - r64) / den, np.abs(r32 - r64) / den
    nh, no = int((eh > rel).sum()), int((eo > rel).sum())
    rh, ro = float(np.sqrt((eh ** 2).mean())), float(np.sqrt((eo ** 2).mean()))
    print(f"{name}: vs float64, entries beyond {rel:g}: {nh} (float32 oracle {no}); worst "
          f"{eh.max():.2e} ({eo.max():.2e}); rms {rh:.2e} ({ro:.2e})", flush=True)
    assert nh <= factor * no + 4, (name, nh, no)
    assert eh.max() <= factor * eo.max() + rel, (name, eh.max(), eo.max())
    assert rh <= factor * ro + rel, (name, rh, ro)


def u32(t):
    """int32 tensor holding u32 values -> numpy uint32."""
    return t.cpu().numpy().view(np.uint32)


def check_pixels(g_img, g_alpha, g_nwalk, o_img, o_alpha, o_nwalk, tol=1e-4, flip_tol=1.1e-3,
                 self_flip=None):
    """Pixels agree within tol. Where the two sides' termination differs (the
    walks end at different Gaussians: a "flip"), the final alpha must sit
    within 2e-3 of the 0.999 threshold, such pixels must be rare, and they
    may differ by at most flip_tol: once alpha_sum >= 0.999 the
    transmittance left is about 1e-3, which bounds what the splats one side
    still adds (colours up to 1). With self_flip (the pixels the oracle
    flips against itself when its exp moves by one ulp, oracle.self_flips),
    the flips must not outnumber those. Returns the number of flips."""
    diff = np.abs(g_img - o_img).max(axis=2)
    flip = g_nwalk != o_nwalk
    bad = (diff > tol) & ~flip
    assert not bad.any(), f"{bad.sum()} pixels off by up to {diff[bad].max()}"
    assert np.abs(g_alpha - o_alpha)[~flip].max(initial=0) <= tol
    if flip.any():
        # a flip can only happen when the running alpha crosses 0.999 within
        # rounding distance on one side
        near = np.abs(np.maximum(g_alpha, o_alpha) - 0.999) < 2e-3
        assert near[flip].all()
        assert flip.mean() < 1e-3, flip.mean()
        fd = max(diff[flip].max(), np.abs(g_alpha - o_alpha)[flip].max())
        print(f"termination flips: {int(flip.sum())} pixels, largest difference {fd:.2e} "
              f"(bound {flip_tol:g})", flush=True)
        assert fd <= flip_tol, (fd, flip_tol)
    if self_flip is not None:
        ns = int(np.asarray(self_flip).sum())
        print(f"flips {int(flip.sum())} vs the oracle's own at +-1 ulp of exp: {ns}", flush=True)
        assert flip.sum() <= ns, (int(flip.sum()), ns)
    return int(flip.sum())


def zbuffer_edge_points(rng, shape, n):
    """Random circles plus the edge rows: NaN centres, radius 0 on a pixel
    centre, negative / infinite / huge radii, centres far off the image,
    depths outside [-1, 1], equal depths."""
    w, h = shape
    pxz = np.stack([rng.uniform(-4, w + 4, n), rng.uniform(-4, h + 4, n),
                    rng.uniform(-1.1, 1.1, n)], 1).astype(np.float32)
    rad = rng.uniform(0.0, 4.0, n).astype(np.float32)
    edge = [((np.nan, 3.0, 0.1), 2.0), ((3.0, np.nan, 0.1), 2.0), ((3.5, 4.5, -0.3), 0.0),
            ((5.0, 5.0, 0.2), -1.0), ((7.2, 6.1, 0.9), np.inf), ((1e30, 5.0, 0.0), 3.0),
            ((-1e30, 5.0, 0.0), 3.0), ((4.0, 1e30, 0.0), 1e30), ((6.0, 7.0, 0.95), 1e30),
            ((np.inf, 5.0, 0.0), np.inf), ((2.0, 2.0, 1.5), 3.0), ((9.5, 3.5, 0.25), 2.5),
            ((9.5, 3.5, 0.25), 2.5)]
    for k, (c, r) in enumerate(edge):
        pxz[k] = c
        rad[k] = r
    return pxz, rad
