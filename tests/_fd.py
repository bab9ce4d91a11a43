"""Central finite differences of the L2 frame loss for single Gaussians on
large frames (the recipe of del-splat-cpu/src/splat_gaussian2.rs:245-266:
eps 1e-3, relative tolerance 1e-2), in float64 on the oracle.

The walk lengths and the boxes are frozen (the function the adjoint
differentiates: termination and box edges are steps), so a change of
Gaussian i can only move the pixels of the tiles its box covers; each
difference re-renders just those tiles (oracle.gauss_l2_loss_tiles) with
Gaussian i re-projected in float64 and spliced into the float64 frame."""
import numpy as np

PARAMS = ("xyz", "log_scale", "quat", "opacity", "rgb")


def fd_gradients(oracle, s, m, img_shape, ts, tile2idx, idx2pnt, aabb, nwalk, target, pick,
                 eps=1e-3):
    """dL/dparam of L = sum over pixels (rgb - target)^2 for the Gaussians
    `pick` of the scene s (xyz, log_scale, quat, opacity, rgb), by central
    differences of size eps in each component. aabb (float32 [N,4]), the
    lists and nwalk fix the frame's boxes, binning and walk lengths.
    Returns {name: [len(pick), components]} (opacity: [len(pick)])."""
    f64 = np.float64
    P = {k: np.array(s[k], f64) for k in PARAMS}
    m64 = np.asarray(m, f64)
    pxd, conic, _, _ = oracle.gauss_project(P["xyz"], P["quat"], np.exp(P["log_scale"]), m64,
                                            img_shape, dtype=f64)
    pxd, conic = np.ascontiguousarray(pxd), np.ascontiguousarray(conic)
    # the frame's float64 arrays, changed in place one Gaussian at a time
    # (copies: the parameters P stay as given)
    op, rgb = P["opacity"].copy(), P["rgb"].copy()
    bb = np.ascontiguousarray(aabb, np.float32)
    t2i = np.ascontiguousarray(tile2idx, np.uint32)
    i2p = np.ascontiguousarray(idx2pnt, np.uint32)
    nw = np.ascontiguousarray(nwalk, np.uint32)
    tgt = np.ascontiguousarray(target, np.float32)
    tw = (img_shape[0] + ts - 1) // ts
    th = (img_shape[1] + ts - 1) // ts
    # half-open tile ranges [x0, x1) x [y0, y1) of each box (the binning's)
    rects = oracle.tile_rects(bb[pick], ts, tw, th)
    out = {k: np.zeros((len(pick),) + P[k].shape[1:]) for k in PARAMS}

    def loss(rect):
        return oracle.gauss_l2_loss_tiles(img_shape, pxd, conic, bb, op, rgb, ts, t2i, i2p, nw,
                                          tgt, rect)

    for r, i in enumerate(pick):
        x0, x1, y0, y1 = (int(v) for v in rects[r])
        if x1 <= x0 or y1 <= y0:
            continue  # no tile: the Gaussian moves no pixel
        rect = (x0, x1 - 1, y0, y1 - 1)
        saved = (pxd[i].copy(), conic[i].copy(), op[i], rgb[i].copy())
        for name in PARAMS:
            ncomp = 1 if P[name].ndim == 1 else P[name].shape[1]
            for c in range(ncomp):
                vals = []
                for sign in (1.0, -1.0):
                    q = {k: P[k][i].copy() for k in PARAMS}
                    if ncomp == 1:
                        q[name] = q[name] + sign * eps
                    else:
                        q[name][c] += sign * eps
                    one = oracle.gauss_project_one(q["xyz"], q["quat"], np.exp(q["log_scale"]),
                                                   m64, img_shape)
                    pxd[i, :2], pxd[i, 2], conic[i] = one["pix"], one["ndcz"], one["conic"]
                    op[i] = q["opacity"]
                    rgb[i] = q["rgb"]
                    vals.append(loss(rect))
                g = (vals[0] - vals[1]) / (2.0 * eps)
                if ncomp == 1:
                    out[name][r] = g
                else:
                    out[name][r, c] = g
                pxd[i], conic[i], op[i], rgb[i] = saved
    return out


def visible_pick(d_rgb, count, seed):
    """`count` Gaussians drawn at random among the clearly visible ones: those
    whose colour gradient is above the median of the non-zero ones."""
    mag = np.abs(np.asarray(d_rgb, np.float64)).reshape(len(d_rgb), -1).sum(1)
    nz = mag[mag > 0]
    cand = np.nonzero(mag > np.median(nz))[0]
    return np.sort(np.random.default_rng(seed).choice(cand, count, replace=False))
