"""Autotuner for the raster schedule: tile size (8, 16, 32) and workgroup
size per tile (waves per tile), measured on the actual workload with HIP
events and cached as JSON per (Gaussian count, view set, mode) signature."""
import json
import os
import time

import torch

# (tile size, workgroup threads; 0 = the default for the tile size). The
# batched plan runs its forward and its adjoint at the same size, and both
# kernels exist for every entry, so one list serves both modes.
CANDIDATES = ((8, 0), (16, 64), (16, 128), (16, 256), (32, 256), (32, 512), (32, 1024))
CANDIDATES_BWD = CANDIDATES
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
    """Single-view forward + L2 adjoint (the differentiable render path). Per
    tile size the forward and the adjoint workgroup sizes are timed
    separately (the kernels are independent; the forward result does not
    depend on its workgroup size). Returns ((tile_size, threads_fwd,
    threads_bwd), {config: ms})."""
    from .SplatGauss import torch as SG

    sig = _signature(num_pnt, [(None, img_shape)], "single_fwd_bwd_v4")
    db = _load() if use_cache else {}
    if sig in db:
        e = db[sig]
        return tuple(e["best"]), dict(e["ms"])
    xyz, ls, q, op, rgb = tensors
    fwd_opts = {8: (64,), 16: (64, 128, 256), 32: (128, 256, 512, 1024)}
    bwd_opts = {8: (64,), 16: (64, 128, 256), 32: (128, 256, 512, 1024)}

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1000.0 / reps

    res, best = {}, None
    for ts in (candidates or (8, 16, 32)):
        pxd, conic, aabb, splat2 = SG.project_fwd(xyz, ls, q, op, rgb, m16, img_shape)
        t_front = timed(lambda: SG.bin_gauss(img_shape, ts, *SG.project_fwd(
            xyz, ls, q, op, rgb, m16, img_shape)[::2]))
        t2i, i2p = SG.bin_gauss(img_shape, ts, pxd, aabb)
        # timed with the walk-ordered tile schedule the render path uses
        cost = torch.empty(t2i.numel() - 1, device=t2i.device, dtype=torch.int32)
        img, alpha, nwalk, tfinal = SG.rasterize_fwd(img_shape, ts, splat2, t2i, i2p,
                                                    tile_cost=cost)
        order = SG.tile_order(cost)
        tf = {th: timed(lambda th=th: SG.rasterize_fwd(img_shape, ts, splat2, t2i, i2p, th,
                                                       tile_order=order, tile_cost=cost))
              for th in fwd_opts[ts]}
        thf = min(tf, key=tf.get)
        d_img = 2.0 * (img - target)
        tb = {th: timed(lambda th=th: SG.rasterize_bwd(img_shape, ts, splat2, t2i, i2p, nwalk,
                                                      tfinal, d_img, None, th, tile_order=order))
              for th in bwd_opts[ts]}
        thb = min(tb, key=tb.get)
        total = t_front + tf[thf] + tb[thb]
        for th, v in tf.items():
            res[f"{ts}/fwd{th}"] = v
        for th, v in tb.items():
            res[f"{ts}/bwd{th}"] = v
        res[f"{ts}/total"] = total
        if best is None or total < best[0]:
            best = (total, ts, thf, thb)
        del pxd, conic, aabb, splat2, t2i, i2p, img, alpha, nwalk, tfinal, d_img
    _, ts, thf, thb = best
    db[sig] = {"best": [ts, thf, thb], "ms": res}
    if use_cache:
        _store(db)
    return (ts, thf, thb), res


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

        try:
            plan, _ = ensure_capacity(
                lambda cap: MultiViewPlan(num_pnt, views, ts, th, pair_capacity=cap), step)
        except RuntimeError:
            continue  # K beyond what one plan can hold at this tile size
        step(plan)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            step(plan)
        torch.cuda.synchronize()
        res[f"{ts}/{th}"] = (time.perf_counter() - t0) * 1000.0 / reps
        plan.close()
        torch.cuda.empty_cache()
    if not res:
        raise RuntimeError("no tile size fits the pair limit of a plan")
    best = min(res, key=res.get)
    ts, th = (int(x) for x in best.split("/"))
    db[sig] = {"best": [ts, th], "ms": res}
    if use_cache:
        _store(db)
    return (ts, th), res
