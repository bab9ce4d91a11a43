"""Batched multi-view rendering and training over one shared Gaussian set
(no counterpart in the reference, which renders one view per call).

`MultiViewPlan` wraps the C ABI's dsplat_mv_* plan: device memory for N
Gaussians and a set of uneven views is allocated once; a forward, a backward
or a whole training step (forward, L2 adjoint, Adam) is then a fixed kernel
sequence that can be captured into a hipGraph and replayed.
"""
import ctypes

import numpy as np
import torch

from .. import _native
from .._native import ViewDesc, call
from .._torch_util import ptr, stream_ptr


def _view_array(views):
    arr = (ViewDesc * len(views))()
    for k, (m16, shape) in enumerate(views):
        m = np.ascontiguousarray(np.asarray(m16, np.float32).reshape(16))
        arr[k].world2ndc[:] = m.tolist()
        arr[k].img_w, arr[k].img_h = int(shape[0]), int(shape[1])
    return arr


MAX_PAIRS = (1 << 32) - 1  # pair counts and capacities are 32-bit


def pair_estimate(num_pnt, views, tile_size, mean_box_px=40.0):
    """A generous guess of the pair count K for the plan's capacity."""
    tiles_per = (mean_box_px / tile_size + 1.0) ** 2
    return int(num_pnt * len(views) * tiles_per) + (1 << 20)


class MultiViewPlan:
    def __init__(self, num_pnt, views, tile_size=16, threads=0, pair_capacity=None,
                 max_pixels=None, device="cuda"):
        self.num_pnt = int(num_pnt)
        self.tile_size = int(tile_size)
        self.threads = int(threads)
        self.device = torch.device(device)
        self.max_view = max(len(views), 1)
        pix = sum(w * h for _, (w, h) in views)
        self.max_pixels = int(max_pixels or pix)
        self.capacity = min(int(pair_capacity or pair_estimate(num_pnt, views, tile_size)),
                            MAX_PAIRS)
        self._plan = ctypes.c_void_p()
        call("dsplat_mv_plan_create", self.num_pnt, self.max_view, self.max_pixels,
             self.tile_size, self.threads, self.capacity, ctypes.byref(self._plan))
        self.graph_kind = None
        # where the last forward wrote alpha: a caller tensor, or None for the
        # plan's own buffer (forwards given no out_alpha, training steps)
        self._last_alpha = None
        self._graph_alpha = None
        self.sh_degree, self.sh_coef = 0, 0  # plain rgb until set_sh
        # a forward over the current views and colour mode has run (the C
        # ABI keeps the same state and refuses a backward without it)
        self._fwd_valid = False
        self.set_views(views)

    # ------------------------------------------------------------- views --
    def set_views(self, views):
        self.views = [(np.asarray(m, np.float32).reshape(16), tuple(s)) for m, s in views]
        arr = _view_array(self.views)
        call("dsplat_mv_set_views", self._plan, stream_ptr(self.device), len(views), arr)
        nv = ctypes.c_uint32()
        tp = ctypes.c_uint64()
        tt = ctypes.c_uint32()
        po = (ctypes.c_uint64 * (len(views) + 1))()
        to = (ctypes.c_uint32 * (len(views) + 1))()
        call("dsplat_mv_plan_info", self._plan, ctypes.byref(nv), ctypes.byref(tp),
             ctypes.byref(tt), po, to)
        self.total_pix = int(tp.value)
        self.total_tiles = int(tt.value)
        self.pix_off = [int(x) for x in po]
        self.tile_off = [int(x) for x in to]
        self.graph_kind = None
        self._fwd_valid = False

    def set_sh(self, degree, num_coef=16):
        """Colour from spherical-harmonic coefficients [N, num_coef, 3] seen
        from each view's camera centre (num_coef = 0: one rgb per Gaussian).
        The forward's colour argument and the backward's colour gradient then
        have that shape."""
        call("dsplat_mv_set_sh", self._plan, int(degree), int(num_coef))
        self.sh_degree, self.sh_coef = (int(degree), int(num_coef)) if num_coef else (0, 0)
        self.graph_kind = None
        self._fwd_valid = False

    def param_layout(self):
        """Offsets of the flat training block: dict of name -> (start, stop)
        for xyz, log_scale, quat, opacity (logit) and colour (rgb, or the SH
        coefficients, which start at 11N rounded up to a multiple of 4)."""
        n = self.num_pnt
        c0 = 11 * n
        if self.sh_coef:
            c0 = (c0 + 3) // 4 * 4
        k = self.sh_coef or 1
        return dict(xyz=(0, 3 * n), log_scale=(3 * n, 6 * n), quat=(6 * n, 10 * n),
                    opacity=(10 * n, 11 * n), colour=(c0, c0 + 3 * k * n))

    def pack_params(self, xyz, log_scale, quat, opacity_logit, colour):
        """Flat training block (float32 tensor on the plan's device; padding 0)."""
        lay = self.param_layout()
        flat = torch.zeros(lay["colour"][1], device=self.device, dtype=torch.float32)
        for name, t in (("xyz", xyz), ("log_scale", log_scale), ("quat", quat),
                        ("opacity", opacity_logit), ("colour", colour)):
            a, b = lay[name]
            flat[a:b] = torch.as_tensor(t, dtype=torch.float32).reshape(-1).to(self.device)
        return flat

    def split(self, flat, channels):
        """Concatenated [total_pix, C] buffer -> list of per-view [H, W, C]."""
        out = []
        for v, (_, (w, h)) in enumerate(self.views):
            a, b = self.pix_off[v], self.pix_off[v + 1]
            t = flat.reshape(-1)[a * channels:b * channels]
            out.append(t.reshape(h, w, channels) if channels > 1 else t.reshape(h, w))
        return out

    def new_image(self, channels=3):
        return torch.empty(self.total_pix * channels, device=self.device, dtype=torch.float32)

    # ---------------------------------------------------------- compute --
    def forward(self, xyz, scale, quat, opacity, rgb, flags=_native.DSPLAT_SCALE_IS_LOG,
                out_rgb=None, out_alpha=None):
        out_rgb = self.new_image(3) if out_rgb is None else out_rgb
        out_alpha = self.new_image(1) if out_alpha is None else out_alpha
        call("dsplat_mv_forward", self._plan, stream_ptr(self.device), ptr(xyz), ptr(scale),
             ptr(quat), ptr(opacity), ptr(rgb), flags, ptr(out_rgb), ptr(out_alpha))
        self._last_alpha = out_alpha
        self._fwd_valid = True
        return out_rgb, out_alpha

    def backward(self, xyz, scale, quat, opacity, out_rgb=None, d_img=None, target=None,
                 d_alpha=None, flags=_native.DSPLAT_SCALE_IS_LOG):
        if not self._fwd_valid:
            raise ValueError("MultiViewPlan.backward: no forward over the current views and "
                             "colour mode (call forward, train_step or replay first)")
        n = self.num_pnt
        dev = self.device
        colour_shape = (n, self.sh_coef, 3) if self.sh_coef else (n, 3)
        g = dict(xyz=torch.empty((n, 3), device=dev), scale=torch.empty((n, 3), device=dev),
                 quat=torch.empty((n, 4), device=dev), opacity=torch.empty(n, device=dev),
                 rgb=torch.empty(colour_shape, device=dev))
        call("dsplat_mv_backward", self._plan, stream_ptr(dev), ptr(xyz), ptr(scale), ptr(quat),
             ptr(opacity), ptr(out_rgb), ptr(d_img), ptr(target), ptr(d_alpha), ptr(g["xyz"]),
             ptr(g["scale"]), ptr(g["quat"]), ptr(g["opacity"]), ptr(g["rgb"]), flags)
        return g

    def train_step(self, params, grads, m1, m2, d_step, targets, out_rgb, lr=1e-3,
                   betas=(0.9, 0.999), eps=1e-8):
        call("dsplat_mv_train_step", self._plan, stream_ptr(self.device), ptr(params), ptr(grads),
             ptr(m1), ptr(m2), ptr(d_step), ptr(targets), ptr(out_rgb), lr, betas[0], betas[1],
             eps)
        self._last_alpha = None
        self._fwd_valid = True

    # ------------------------------------------------------------ graphs --
    def capture_forward(self, xyz, scale, quat, opacity, rgb, out_rgb, out_alpha,
                        flags=_native.DSPLAT_SCALE_IS_LOG):
        call("dsplat_mv_graph_capture", self._plan, 0, ptr(xyz), ptr(scale), ptr(quat),
             ptr(opacity), ptr(rgb), flags, ptr(out_rgb), ptr(out_alpha), None, None, None, None,
             None, None, 0.0, 0.0, 0.0, 0.0)
        self.graph_kind = "forward"
        self._graph_alpha = out_alpha

    def capture_fwd_bwd(self, xyz, scale, quat, opacity, rgb, out_rgb, targets, grads,
                        flags=_native.DSPLAT_SCALE_IS_LOG):
        """Graph of forward + adjoint of sum (out_rgb - targets)^2; grads [14N]
        receive (xyz 3N | scale 3N | quat 4N | opacity N | rgb 3N)."""
        call("dsplat_mv_graph_capture", self._plan, 2, ptr(xyz), ptr(scale), ptr(quat),
             ptr(opacity), ptr(rgb), flags, ptr(out_rgb), None, None, ptr(grads), None, None,
             None, ptr(targets), 0.0, 0.0, 0.0, 0.0)
        self.graph_kind = "fwd_bwd"
        self._graph_alpha = None

    def capture_train_step(self, params, grads, m1, m2, d_step, targets, out_rgb, lr=1e-3,
                           betas=(0.9, 0.999), eps=1e-8):
        call("dsplat_mv_graph_capture", self._plan, 1, None, None, None, None, None, 0,
             ptr(out_rgb), None, ptr(params), ptr(grads), ptr(m1), ptr(m2), ptr(d_step),
             ptr(targets), lr, betas[0], betas[1], eps)
        self.graph_kind = "train"
        self._graph_alpha = None

    def replay(self):
        call("dsplat_mv_graph_launch", self._plan, stream_ptr(self.device))
        self._last_alpha = self._graph_alpha
        self._fwd_valid = True

    # ------------------------------------------------------------ state --
    def _outputs(self):
        p = [ctypes.c_void_p() for _ in range(7)]
        call("dsplat_mv_outputs", self._plan, *[ctypes.byref(x) for x in p])
        return [x.value for x in p]

    def counters(self):
        """SYNC: (K, overflowed)."""
        c = self._outputs()[6]
        buf = torch.empty(8, device=self.device, dtype=torch.int32)
        call("dsplat_memcpy_d2d", stream_ptr(self.device), ptr(buf), c, 20)
        host = buf.cpu().numpy().view(np.uint32)
        return int(host[0]), bool(host[4])

    def _copy(self, src, count, dtype):
        t = torch.empty(max(count, 1), device=self.device, dtype=dtype)
        call("dsplat_memcpy_d2d", stream_ptr(self.device), ptr(t), src,
             count * t.element_size())
        return t[:count]

    def lists(self):
        """(tile2idx [tiles+1], idx2pnt [K]) over global tiles, as int32
        tensors of u32 values; items are view*N + Gaussian."""
        o = self._outputs()
        k, over = self.counters()
        if over:
            raise RuntimeError(f"pair capacity {self.capacity} < K = {k}")
        return (self._copy(o[4], self.total_tiles + 1, torch.int32),
                self._copy(o[5], k, torch.int32))

    def state(self):
        """(alpha, nwalk, tfinal) of the last forward, concatenated per pixel
        (alpha read from where that forward wrote it)."""
        o = self._outputs()
        if self._last_alpha is not None:
            alpha = self._last_alpha.reshape(-1)[:self.total_pix].clone()
        else:
            alpha = self._copy(o[1], self.total_pix, torch.float32)
        return (alpha,
                self._copy(o[2], self.total_pix, torch.int32),
                self._copy(o[3], self.total_pix, torch.float32))

    def close(self):
        if self._plan:
            _native.lib().dsplat_mv_plan_destroy(self._plan)
            self._plan = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ensure_capacity(plan_factory, run, max_tries=3):
    """Runs `run(plan)` and grows the plan when the pair capacity overflowed.
    plan_factory(capacity or None) -> plan. Returns (plan, result). Raises
    RuntimeError when K exceeds the largest capacity a plan can hold."""
    plan = plan_factory(None)
    for _ in range(max_tries):
        res = run(plan)
        k, over = plan.counters()
        if not over:
            return plan, res
        plan.close()
        if k > MAX_PAIRS:
            raise RuntimeError(f"K = {k} pairs exceeds the plan limit {MAX_PAIRS}")
        plan = plan_factory(min(int(k * 1.25) + 1024, MAX_PAIRS))
    raise RuntimeError("pair capacity could not be satisfied")
