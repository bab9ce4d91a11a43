"""Autotuner for the raster schedule: tile size (8, 16, 32) and workgroup
size per tile (waves per tile), measured on the actual workload with HIP
events and cached as JSON per (Gaussian count, view set, mode) signature."""
import json
import os
import time

import torch

CANDIDATES = ((8, 0), (16, 64), (16, 128), (16, 256), (32, 256), (32, 512), (32, 1024))
CANDIDATES_BWD = ((8, 0), (16, 64), (16, 128), (16, 256), (32, 256))
CACHE = os.environ.get("DSPLAT_AUTOTUNE_CACHE",
                       os.path.join(os.path.expanduser("~"), ".cache", "dsplat_autotune.json"))


def _signature(num_pnt, views, mode):
    shapes = ",".join(f"{w}x{h}" for _, (w, h) in views)
    return f"{mode}:{num_pnt}:{shapes}"


def _load():
    try:
        with open(CACHE) as f:
            return json.load(f)
    except Exception:
        return {}


def _store(db):
    try:
        os.makedirs(os.path.dirname(CACHE), exist_ok=True)
        with open(CACHE, "w") as f:
            json.dump(db, f, indent=1)
    except OSError:
        pass


def tune_single(num_pnt, m16, img_shape, tensors, target, reps=2, candidates=None,
                use_cache=True):
    """Single-view forward + L2 adjoint (the differentiable render path).
    Returns ((tile_size, threads), {config: ms})."""
    from .SplatGauss import torch as SG

    sig = _signature(num_pnt, [(None, img_shape)], "single_fwd_bwd")
    db = _load() if use_cache else {}
    if sig in db:
        e = db[sig]
        return (e["best"][0], e["best"][1]), dict(e["ms"])
    xyz, ls, q, op, rgb = tensors
    res = {}
    for ts, th in candidates or CANDIDATES_BWD:
        def step():
            pxd, conic, aabb, splat2 = SG.project_fwd(xyz, ls, q, op, rgb, m16, img_shape)
            t2i, i2p = SG.bin_gauss(img_shape, ts, pxd, aabb)
            img, alpha, nwalk, tlast = SG.rasterize_fwd(img_shape, ts, splat2, t2i, i2p, th)
            d = SG.rasterize_bwd(img_shape, ts, splat2, t2i, i2p, nwalk, tlast,
                                 2.0 * (img - target), None, th)
            SG.project_bwd(xyz, ls, q, m16, img_shape, d[0], d[1])

        step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            step()
        torch.cuda.synchronize()
        res[f"{ts}/{th}"] = (time.perf_counter() - t0) * 1000.0 / reps
    best = min(res, key=res.get)
    ts, th = (int(x) for x in best.split("/"))
    db[sig] = {"best": [ts, th], "ms": res}
    if use_cache:
        _store(db)
    return (ts, th), res


def tune_multiview(num_pnt, views, tensors, mode="forward", target=None, reps=3,
                   candidates=None, use_cache=True):
    """Returns ((tile_size, threads), {config: ms}). `tensors` are (xyz,
    log_scale, quat, opacity, rgb) on the GPU; mode "forward" times
    dsplat_mv_forward, "fwd_bwd" forward + L2 adjoint against `target`."""
    from .SplatGauss.multiview import MultiViewPlan, ensure_capacity

    sig = _signature(num_pnt, views, mode)
    db = _load() if use_cache else {}
    if sig in db:
        e = db[sig]
        return (e["best"][0], e["best"][1]), {k: v for k, v in e["ms"].items()}
    cands = candidates or (CANDIDATES if mode == "forward" else CANDIDATES_BWD)
    xyz, ls, q, op, rgb = tensors
    res = {}
    for ts, th in cands:
        def step(p):
            out, _ = p.forward(xyz, ls, q, op, rgb)
            if mode != "forward":
                p.backward(xyz, ls, q, op, out_rgb=out, target=target)
            return out

        plan, _ = ensure_capacity(
            lambda cap: MultiViewPlan(num_pnt, views, ts, th, pair_capacity=cap), step)
        step(plan)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            step(plan)
        torch.cuda.synchronize()
        res[f"{ts}/{th}"] = (time.perf_counter() - t0) * 1000.0 / reps
        plan.close()
        torch.cuda.empty_cache()
    best = min(res, key=res.get)
    ts, th = (int(x) for x in best.split("/"))
    db[sig] = {"best": [ts, th], "ms": res}
    if use_cache:
        _store(db)
    return (ts, th), res
