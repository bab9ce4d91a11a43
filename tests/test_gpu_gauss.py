"""GPU parity of the Gaussian path against the oracle (same seeded inputs).

Bars: projection bit-exact (both sides round every op in the same order;
tolerance conic 1e-5 relative as the fallback bar in the assert message);
binning / sort / tile ranges bit-exact, checked as the reference checks them
(del-splat-cudarc/examples/03_splat_gauss.rs:124-153: the CPU binning is fed
the GPU's projected boxes and depths); pixels within 1e-4 except at pixels
where the two sides' 0.999 termination decision differs by rounding (counted,
bounded, and each one explained); gradients within 2e-3 of the largest
magnitude.
"""
import numpy as np
import pytest
import torch

from _helpers import check_grad, check_pixels as _check_pixels, max_rel, small_scene, u32

pytestmark = pytest.mark.gpu

DEV = "cuda"


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def _project(s, m, shape):
    import del_splat_dlpack.SplatGauss.torch as SG
    return SG.pnt2splat3_to_pnt2splat2(_t(s["xyz"]), _t(s["scale"]), _t(s["quat"]),
                                       _t(m), shape)


def _check_bin(shape, ts, pxd, aabb, o_t2i, o_i2p):
    """Both binning paths against the oracle's lists: the phased entry points
    (K read back; depth sort and tile sort) and the one-call workspace entry
    point sized with the tile count (tile buckets and per-tile sorts), at
    capacity K and K + 1000."""
    import del_splat_dlpack.SplatGauss.torch as SG
    t2i, i2p = SG.tile2idx_idx2pnt(shape, ts, pxd, aabb)
    np.testing.assert_array_equal(u32(t2i), o_t2i)
    np.testing.assert_array_equal(u32(i2p), o_i2p)
    k = len(o_i2p)
    for cap in (k, k + 1000):
        kbuf = torch.full((1,), -1, device=DEV, dtype=torch.int32)
        t2i, i2p, _ = SG.bin_gauss_capacity(shape, ts, pxd, aabb, cap, kbuf)
        assert int(kbuf[0].item()) == k
        np.testing.assert_array_equal(u32(t2i), o_t2i)
        np.testing.assert_array_equal(u32(i2p)[:k], o_i2p)


@pytest.mark.parametrize("n,shape,seed", [(5000, (256, 192), 0), (20000, (333, 201), 1)])
def test_projection_parity(oracle, n, shape, seed):
    s, m = small_scene(n, seed, shape)
    pxd, conic, aabb = _project(s, m, shape)
    o_pxd, o_conic, o_aabb, nclamp = oracle.gauss_project(s["xyz"], s["quat"], s["scale"], m,
                                                          shape)
    g_pxd, g_conic, g_aabb = (t.cpu().numpy() for t in (pxd, conic, aabb))
    assert nclamp == 0
    assert max_rel(g_conic, o_conic) <= 1e-5
    np.testing.assert_array_equal(g_pxd, o_pxd)
    np.testing.assert_array_equal(g_conic, o_conic)
    np.testing.assert_array_equal(g_aabb, o_aabb)


@pytest.mark.parametrize("ts", [8, 16, 32])
@pytest.mark.parametrize("n,shape,seed", [(3000, (256, 192), 2), (20000, (301, 177), 3),
                                          (0, (64, 48), 4), (1, (64, 48), 5)])
def test_binning_bit_exact(oracle, ts, n, shape, seed):
    import del_splat_dlpack.SplatGauss.torch as SG
    s, m = small_scene(max(n, 1), seed, shape)
    if n == 0:
        s = {k: v[:0] for k, v in s.items()}
    pxd, conic, aabb = _project(s, m, shape)
    o_t2i, o_i2p = oracle.tile2pnt_gauss(pxd.cpu().numpy(), aabb.cpu().numpy(), shape, ts)
    _check_bin(shape, ts, pxd, aabb, o_t2i, o_i2p)


def test_binning_edge_boxes(oracle):
    """Boxes off-screen, straddling every border, covering the whole image,
    culled (inverted to infinity), finite inverted inside one tile (binned
    to it, as the reference's loops do) and across tiles (none), NaN and
    exactly on tile boundaries."""
    import del_splat_dlpack.SplatGauss.torch as SG
    shape, ts = (100, 70), 16
    boxes = np.array([
        [-50, -50, -10, -10], [110, 10, 200, 20], [-5, -5, 5, 5], [95, 65, 130, 90],
        [-1e9, -1e9, 1e9, 1e9], [np.inf, np.inf, -np.inf, -np.inf], [np.nan, 0, 10, 10],
        [16, 16, 32, 32], [15.999, 0, 16.0, 0.0], [0, 0, 0, 0], [99.9, 69.9, 99.95, 69.95],
        [40, 30, 35, 20], [40, 30, 20, 20], [30, 60, 34, 50],
    ], np.float32)
    n = len(boxes)
    pxd = np.zeros((n, 3), np.float32)
    pxd[:, 2] = np.linspace(-1.5, 1.5, n).astype(np.float32)
    pxd[3, 2] = pxd[4, 2]  # equal depths: ties by index
    o_t2i, o_i2p = oracle.tile2pnt(boxes, pxd, shape, ts)
    _check_bin(shape, ts, _t(pxd), _t(boxes), o_t2i, o_i2p)


def test_binning_allocating_variant(oracle):
    """dsplat_gauss_tile2idx_idx2pnt (library-owned idx2pnt, SYNC) == _ws."""
    import ctypes
    from del_splat_dlpack import _native
    s, m = small_scene(4000, 6, (200, 150))
    pxd, conic, aabb = _project(s, m, (200, 150))
    tw, th = (200 + 15) // 16, (150 + 15) // 16
    t2i = torch.empty(tw * th + 1, device=DEV, dtype=torch.int32)
    out = ctypes.c_void_p()
    k = ctypes.c_uint32()
    stream = torch.cuda.current_stream().cuda_stream
    _native.call("dsplat_gauss_tile2idx_idx2pnt", stream, 200, 150, 16, 4000, pxd.data_ptr(),
                 aabb.data_ptr(), t2i.data_ptr(), ctypes.byref(out), ctypes.byref(k))
    torch.cuda.synchronize()
    buf = torch.empty(k.value, device=DEV, dtype=torch.int32)
    _copy_device(buf.data_ptr(), out.value, 4 * k.value)
    _native.call("dsplat_free", stream, out.value)
    torch.cuda.synchronize()
    o_t2i, o_i2p = oracle.tile2pnt_gauss(pxd.cpu().numpy(), aabb.cpu().numpy(), (200, 150), 16)
    assert k.value == len(o_i2p)
    np.testing.assert_array_equal(u32(t2i), o_t2i)
    np.testing.assert_array_equal(u32(buf), o_i2p)


def test_binning_one_call_and_count_only(oracle):
    """The one-call _ws entry at capacity K and above, and the count-only
    entry, against the oracle's lists; empty input through the phased path."""
    import del_splat_dlpack.SplatGauss.torch as SG
    from del_splat_dlpack import _native
    shape = (200, 150)
    s, m = small_scene(4000, 16, shape)
    pxd, conic, aabb = _project(s, m, shape)
    o_t2i, o_i2p = oracle.tile2pnt_gauss(pxd.cpu().numpy(), aabb.cpu().numpy(), shape, 16)
    kbuf = torch.zeros(4, device=DEV, dtype=torch.int32)
    _native.call("dsplat_gauss_count_pairs", torch.cuda.current_stream().cuda_stream, shape[0],
                 shape[1], 16, 4000, aabb.data_ptr(), kbuf.data_ptr())
    assert int(kbuf[0].item()) == len(o_i2p)
    for cap in (len(o_i2p), len(o_i2p) + 1000):
        t2i, i2p, _ = SG.bin_gauss_capacity(shape, 16, pxd, aabb, cap, kbuf)
        assert int(kbuf[0].item()) == len(o_i2p)
        np.testing.assert_array_equal(u32(t2i), o_t2i)
        np.testing.assert_array_equal(u32(i2p)[:len(o_i2p)], o_i2p)
    e3, e4 = torch.zeros((0, 3), device=DEV), torch.zeros((0, 4), device=DEV)
    t2i, i2p = SG.tile2idx_idx2pnt(shape, 16, e3, e4)
    assert i2p.numel() == 0 and int(t2i.abs().sum().item()) == 0


@pytest.mark.parametrize("n,shape,ties", [(5000, (64, 64), True), (20000, (64, 64), True),
                                          (20000, (40, 30), True), (12000, (64, 64), False)])
def test_binning_dense_tiles_and_equal_depths(oracle, n, shape, ties):
    """Every box covers the whole image, so each tile's list holds (nearly)
    all n Gaussians (beyond the 8192 pairs a tile sort keeps in LDS from n =
    12000 on); with ties, a third of the depths repeat three values, so long
    runs of equal depth keys must keep ascending index order."""
    rng = np.random.default_rng(n + shape[0])
    boxes = np.tile(np.array([[-1.0, -1.0, shape[0] + 1.0, shape[1] + 1.0]], np.float32), (n, 1))
    boxes[::7, 0] = 10.0  # some boxes start one tile in
    pxd = np.zeros((n, 3), np.float32)
    if ties:
        pxd[:, 2] = rng.uniform(-0.9, 0.9, n).astype(np.float32)
        pxd[::3, 2] = np.float32([0.1, 0.2, 0.3])[rng.integers(0, 3, len(pxd[::3]))]
    else:
        pxd[:, 2] = rng.permutation(np.linspace(-0.9, 0.9, n)).astype(np.float32)
        assert len(np.unique(pxd[:, 2])) == n
    for ts in (16, 32):
        o_t2i, o_i2p = oracle.tile2pnt(boxes, pxd, shape, ts)
        _check_bin(shape, ts, _t(pxd), _t(boxes), o_t2i, o_i2p)


def test_binning_large_tile_grid(oracle):
    """A 1024x1024 image at 8-pixel tiles: 16384 tiles, so the tile sort
    runs on 14 key bits (two 7-bit passes); same lists as the oracle."""
    import del_splat_dlpack.SplatGauss.torch as SG
    shape = (1024, 1024)
    s, m = small_scene(20000, 8, shape)
    pxd, conic, aabb = _project(s, m, shape)
    o_t2i, o_i2p = oracle.tile2pnt_gauss(pxd.cpu().numpy(), aabb.cpu().numpy(), shape, 8)
    _check_bin(shape, 8, pxd, aabb, o_t2i, o_i2p)


def test_binning_many_gaussians(oracle):
    """4.5M Gaussians: more than 16384 chunk sums, so the pair offsets take
    the two-level scan; same lists as the oracle."""
    import del_splat_dlpack.SplatGauss.torch as SG
    shape = (320, 240)
    s, m = small_scene(4_500_000, 9, shape)
    pxd, conic, aabb = _project(s, m, shape)
    o_t2i, o_i2p = oracle.tile2pnt_gauss(pxd.cpu().numpy(), aabb.cpu().numpy(), shape, 16)
    print(f"pairs {len(o_i2p)}", flush=True)
    _check_bin(shape, 16, pxd, aabb, o_t2i, o_i2p)


def _copy_device(dst, src, nbytes):
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    assert hip.hipMemcpy(dst, src, nbytes, 3) == 0  # hipMemcpyDeviceToDevice


def test_capacity_overflow_reports_k(oracle):
    import del_splat_dlpack.SplatGauss.torch as SG
    s, m = small_scene(2000, 7, (128, 128))
    pxd, conic, aabb = _project(s, m, (128, 128))
    kbuf = torch.zeros(4, device=DEV, dtype=torch.int32)
    t2i, i2p, _ = SG.bin_gauss_capacity((128, 128), 16, pxd, aabb, 10, kbuf)
    o_t2i, o_i2p = oracle.tile2pnt_gauss(pxd.cpu().numpy(), aabb.cpu().numpy(), (128, 128), 16)
    assert int(kbuf[0].item()) == len(o_i2p) > 10
    assert int(t2i.abs().sum().item()) == 0  # nothing binned on overflow


@pytest.mark.parametrize("n", [1, 1000, 2048, 4096, 4097, 100_003])
@pytest.mark.parametrize("bits", [1, 13, 37, 64])
def test_radix_sort_stable(n, bits):
    from del_splat_dlpack import _native
    rng = np.random.default_rng(n + bits)
    keys = rng.integers(0, 2**63, n, dtype=np.uint64)
    if bits == 64:
        keys[: n // 2] |= np.uint64(1 << 63)
    else:
        keys &= np.uint64((1 << bits) - 1)
    keys[rng.integers(0, n, n // 3 + 1)] = keys[0]  # many ties
    vals = np.arange(n, dtype=np.uint32)
    k = torch.from_numpy(keys.view(np.int64).copy()).to(DEV)
    v = torch.from_numpy(vals.view(np.int32).copy()).to(DEV)
    L = _native.lib()
    wsb = L.dsplat_radix_sort_workspace_size(n)
    ws = torch.empty(wsb, device=DEV, dtype=torch.uint8)
    _native.call("dsplat_radix_sort_by_key_u64_u32", torch.cuda.current_stream().cuda_stream, n,
                 k.data_ptr(), v.data_ptr(), bits, ws.data_ptr(), wsb)
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(k.cpu().numpy().view(np.uint64), keys[order])
    np.testing.assert_array_equal(v.cpu().numpy().view(np.uint32), vals[order])


def test_radix_sort_large_repeated():
    """Thousands of tiles per pass (long look-back chains), the same
    workspace reused for three sorts of different sizes: every sort must
    clear what the previous one left."""
    from del_splat_dlpack import _native
    L = _native.lib()
    nmax = 6_000_011
    wsb = L.dsplat_radix_sort_workspace_size(nmax)
    ws = torch.full((wsb,), 0xA5, device=DEV, dtype=torch.uint8)  # garbage in the workspace
    for n, bits in ((nmax, 40), (1_234_567, 31), (nmax - 5, 17)):
        rng = np.random.default_rng(n)
        keys = rng.integers(0, 1 << bits, n, dtype=np.uint64)
        k = torch.from_numpy(keys.view(np.int64).copy()).to(DEV)
        v = torch.arange(n, device=DEV, dtype=torch.int32)
        _native.call("dsplat_radix_sort_by_key_u64_u32", torch.cuda.current_stream().cuda_stream,
                     n, k.data_ptr(), v.data_ptr(), bits, ws.data_ptr(), wsb)
        order = np.argsort(keys, kind="stable")
        np.testing.assert_array_equal(k.cpu().numpy().view(np.uint64), keys[order])
        np.testing.assert_array_equal(v.cpu().numpy().view(np.uint32), order.astype(np.uint32))


@pytest.mark.parametrize("n,bits", [(9_000_017, 37), (8 << 20, 15)])
def test_radix_sort_big_tiles(n, bits):
    """Sorts of at least 8M elements take 4096-element tiles: stable against
    numpy, with ties, in a dirty workspace, at the size boundary too."""
    from del_splat_dlpack import _native
    L = _native.lib()
    wsb = L.dsplat_radix_sort_workspace_size(n)
    ws = torch.full((wsb,), 0x5A, device=DEV, dtype=torch.uint8)
    rng = np.random.default_rng(n + bits)
    keys = rng.integers(0, 1 << bits, n, dtype=np.uint64)
    keys[rng.integers(0, n, n // 4)] = keys[7]
    k = torch.from_numpy(keys.view(np.int64).copy()).to(DEV)
    v = torch.arange(n, device=DEV, dtype=torch.int32)
    _native.call("dsplat_radix_sort_by_key_u64_u32", torch.cuda.current_stream().cuda_stream, n,
                 k.data_ptr(), v.data_ptr(), bits, ws.data_ptr(), wsb)
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(k.cpu().numpy().view(np.uint64), keys[order])
    np.testing.assert_array_equal(v.cpu().numpy().view(np.uint32), order.astype(np.uint32))


@pytest.mark.parametrize("n", [0, 1, 2047, 2048, 2049, 1_000_001])
def test_exclusive_scan(n):
    from del_splat_dlpack import _native
    rng = np.random.default_rng(n)
    a = rng.integers(0, 100, n).astype(np.uint32)
    t = torch.from_numpy(a.view(np.int32).copy()).to(DEV) if n else torch.zeros(1, device=DEV, dtype=torch.int32)
    out = torch.empty_like(t)
    L = _native.lib()
    wsb = L.dsplat_exclusive_scan_workspace_size(n)
    ws = torch.empty(wsb, device=DEV, dtype=torch.uint8)
    _native.call("dsplat_exclusive_scan_u32", torch.cuda.current_stream().cuda_stream, n,
                 t.data_ptr(), out.data_ptr(), ws.data_ptr(), wsb)
    if n:
        ref = np.concatenate([[0], np.cumsum(a, dtype=np.uint64)[:-1]]).astype(np.uint32)
        np.testing.assert_array_equal(out.cpu().numpy().view(np.uint32), ref)


@pytest.mark.parametrize("ts", [8, 16, 32])
def test_raster_parity(oracle, ts):
    import del_splat_dlpack.SplatGauss.torch as SG
    shape = (320, 240)
    s, m = small_scene(30000, 10 + ts, shape)
    pxd, conic, aabb = _project(s, m, shape)
    t2i, i2p = SG.tile2idx_idx2pnt(shape, ts, pxd, aabb)
    op, rgb = _t(s["opacity"]), _t(s["rgb"])
    img = SG.rasterize_pnt2splat2(shape, pxd, conic, aabb, op, rgb, ts, t2i, i2p).cpu().numpy()
    o_img, o_alpha, o_nwalk, o_tfinal = oracle.gauss_rasterize(
        shape, pxd.cpu().numpy(), conic.cpu().numpy(), aabb.cpu().numpy(), s["opacity"],
        s["rgb"], ts, u32(t2i), u32(i2p))
    # packed-record path with the adjoint's state
    _, _, _, splat2 = SG.project_fwd(_t(s["xyz"]), _t(s["scale"]), _t(s["quat"]), op, rgb,
                                     _t(m), shape, log_scale=False)
    for thr in ([0, 64, 128, 256] if ts == 16 else [0, 128, 256, 512, 1024] if ts == 32 else [0]):
        g_img, g_alpha, g_nwalk, g_tfinal = (t.cpu().numpy() for t in SG.rasterize_fwd(
            shape, ts, splat2, t2i, i2p, thr))
        flips = _check_pixels(g_img, g_alpha, g_nwalk.view(np.uint32), o_img, o_alpha, o_nwalk)
        print(f"ts={ts} threads={thr} threshold flips={flips}")
        np.testing.assert_allclose(g_img, img, atol=1e-6)
        # the plain forward (no adjoint state tracked): the same pixels, bit for bit
        p_img, p_alpha, none_w, none_t = SG.rasterize_fwd(shape, ts, splat2, t2i, i2p, thr,
                                                          state=False)
        assert none_w is None and none_t is None
        np.testing.assert_array_equal(p_img.cpu().numpy(), g_img)
        np.testing.assert_array_equal(p_alpha.cpu().numpy(), g_alpha)
    assert (o_alpha > 0.999).mean() > 0.05  # the scene does saturate pixels


def test_raster_matches_naive_renderer(oracle):
    """Tiled GPU raster == the reference's naive non-tiled renderer on a tiny
    scene whose depths stay inside (-1, 1) (pnt2pixxyndcz.rs:42-100)."""
    import del_splat_dlpack.SplatGauss.torch as SG
    shape = (64, 48)
    s, m = small_scene(200, 11, shape)
    pxd, conic, aabb = _project(s, m, shape)
    t2i, i2p = SG.tile2idx_idx2pnt(shape, 16, pxd, aabb)
    img = SG.rasterize_pnt2splat2(shape, pxd, conic, aabb, _t(s["opacity"]), _t(s["rgb"]), 16,
                                  t2i, i2p).cpu().numpy()
    naive = oracle.gauss_render_sort_depth(pxd.cpu().numpy(), conic.cpu().numpy(),
                                           aabb.cpu().numpy(), s["opacity"], s["rgb"], shape)
    assert np.abs(img - naive).max() <= 1e-4


def _grad_case(oracle, n, shape, seed, ts, with_alpha):
    import del_splat_dlpack.SplatGauss.torch as SG
    s, m = small_scene(n, seed, shape)
    xyz, ls, q = _t(s["xyz"]), _t(s["log_scale"]), _t(s["quat"])
    op, rgb = _t(s["opacity"]), _t(s["rgb"])
    for t in (xyz, ls, q, op, rgb):
        t.requires_grad_(True)
    img, alpha = SG.render(xyz, ls, q, op, rgb, _t(m), shape, tile_size=ts)
    rng = np.random.default_rng(seed + 100)
    target = rng.random((shape[1], shape[0], 3)).astype(np.float32)
    w_alpha = rng.random((shape[1], shape[0])).astype(np.float32) if with_alpha else None
    loss = ((img - _t(target)) ** 2).sum()
    if with_alpha:
        loss = loss + (alpha * _t(w_alpha)).sum()
    loss.backward()
    # oracle chain on the same projection / lists
    scale = np.exp(s["log_scale"]).astype(np.float32)
    o_pxd, o_conic, o_aabb, _ = oracle.gauss_project(s["xyz"], s["quat"], scale, m, shape)
    o_t2i, o_i2p = oracle.tile2pnt_gauss(o_pxd, o_aabb, shape, ts)
    o_img, o_alpha, o_nwalk, o_tfinal = oracle.gauss_rasterize(
        shape, o_pxd, o_conic, o_aabb, s["opacity"], s["rgb"], ts, o_t2i, o_i2p)
    d_img = 2.0 * (img.detach().cpu().numpy() - target)
    dpix, dcon, dop, drgb = oracle.gauss_rasterize_bwd(
        shape, o_pxd, o_conic, o_aabb, s["opacity"], s["rgb"], ts, o_t2i, o_i2p, o_nwalk,
        o_tfinal, d_img, w_alpha)
    dxyz, dls, dq = oracle.gauss_project_bwd(s["xyz"], s["quat"], scale, m, shape, dpix, dcon,
                                             log_scale=True)
    return dict(xyz=(xyz.grad, dxyz), log_scale=(ls.grad, dls), quat=(q.grad, dq),
                opacity=(op.grad, dop), rgb=(rgb.grad, drgb))


@pytest.mark.parametrize("ts,with_alpha", [(16, False), (16, True), (8, False), (32, True)])
def test_backward_matches_oracle_adjoint(oracle, ts, with_alpha):
    res = _grad_case(oracle, 3000, (160, 120), 20 + ts, ts, with_alpha)
    for name, (g, o) in res.items():
        check_grad(g.cpu().numpy(), o, name, abs_frac=1e-4)


@pytest.mark.parametrize("ts,threads", [(16, 64), (16, 128), (16, 256), (32, 64), (32, 128),
                                        (32, 256), (32, 512), (32, 1024), (8, 64)])
@pytest.mark.parametrize("with_alpha", [False, True])
def test_raster_bwd_every_workgroup_shape(oracle, ts, threads, with_alpha):
    """The adjoint raster at every workgroup shape (pixel groups per wave,
    per-wave partial slots) against the oracle adjoint on the same lists; the
    image is not a multiple of the tile size."""
    import del_splat_dlpack.SplatGauss.torch as SG
    shape = (171, 133)
    s, m = small_scene(4000, 40 + ts, shape)
    op, rgb = _t(s["opacity"]), _t(s["rgb"])
    pxd, conic, aabb, splat2 = SG.project_fwd(_t(s["xyz"]), _t(s["scale"]), _t(s["quat"]), op,
                                              rgb, _t(m), shape, log_scale=False)
    t2i, i2p = SG.tile2idx_idx2pnt(shape, ts, pxd, aabb)
    img, alpha, nwalk, tfinal = SG.rasterize_fwd(shape, ts, splat2, t2i, i2p)
    rng = np.random.default_rng(ts + threads)
    d_img = rng.standard_normal((shape[1], shape[0], 3)).astype(np.float32)
    d_alpha = rng.standard_normal((shape[1], shape[0])).astype(np.float32) if with_alpha else None
    g = SG.rasterize_bwd(shape, ts, splat2, t2i, i2p, nwalk, tfinal, _t(d_img),
                         None if d_alpha is None else _t(d_alpha), threads)
    o = oracle.gauss_rasterize_bwd(shape, pxd.cpu().numpy(), conic.cpu().numpy(),
                                   aabb.cpu().numpy(), s["opacity"], s["rgb"], ts, u32(t2i),
                                   u32(i2p), nwalk.cpu().numpy().view(np.uint32),
                                   tfinal.cpu().numpy(), d_img, d_alpha)
    for name, a, b in zip(("d_pixxy", "d_conic", "d_opacity", "d_rgb"), g, o):
        check_grad(a.cpu().numpy(), b, f"ts={ts} threads={threads} alpha={with_alpha} {name}",
                   abs_frac=1e-4)


@pytest.mark.parametrize("ts,threads", [(16, 0), (32, 512)])
def test_raster_bwd_fused_l2(ts, threads):
    """The fused-L2 adjoint (dL/drgb = 2 (img - target) formed in the kernel)
    against the adjoint fed that gradient explicitly, with a tile schedule."""
    import del_splat_dlpack.SplatGauss.torch as SG
    from del_splat_dlpack import scene
    shape = (333, 201)
    s, m = small_scene(20000, 21, shape)
    xyz, ls, q, op, rgb = (_t(s[k]) for k in ("xyz", "log_scale", "quat", "opacity", "rgb"))
    pxd, conic, aabb, splat2 = SG.project_fwd(xyz, ls, q, op, rgb, _t(m), shape)
    t2i, i2p = SG.tile2idx_idx2pnt(shape, ts, pxd, aabb)
    cost = torch.empty(t2i.numel() - 1, device=DEV, dtype=torch.int32)
    img, alpha, nwalk, tfinal = SG.rasterize_fwd(shape, ts, splat2, t2i, i2p, threads,
                                                tile_cost=cost)
    order = SG.tile_order(cost)
    target = _t(scene.target_image(shape, 77))
    ref = SG.rasterize_bwd(shape, ts, splat2, t2i, i2p, nwalk, tfinal, 2.0 * (img - target), None,
                           threads, tile_order=order)
    got = SG.rasterize_bwd_l2(shape, ts, splat2, t2i, i2p, nwalk, tfinal, img, target, threads,
                              tile_order=order)
    for a, b in zip(got, ref):
        # same terms, float atomics in another order
        assert max_rel(a.cpu().numpy(), b.cpu().numpy()) <= 1e-5


@pytest.mark.parametrize("n,shape,ts,cap", [(20000, (333, 201), 8, 0), (20000, (333, 201), 16, 0),
                                             (20000, (640, 480), 32, 0), (1, (64, 48), 16, 0),
                                             (0, (64, 48), 16, 0), (2000, (128, 128), 16, 10),
                                             (500, (1, 1), 8, 0), (3000, (96, 64), 16, -1)])
@pytest.mark.parametrize("boxes", [False, True])
def test_project_bin_matches_project_then_bin(n, shape, ts, cap, boxes):
    """dsplat_gauss_project_bin_ws (the projection writes the binning's count
    outputs itself) against dsplat_gauss_project_fwd followed by
    dsplat_gauss_tile2idx_idx2pnt_ws on the same inputs: records, K, tile
    ranges and lists bit-exact, the optional pixel/depth and box outputs too,
    the cleared accumulators zero; a capacity below K (the last case) leaves
    every list empty in both."""
    import del_splat_dlpack.SplatGauss.torch as SG
    from del_splat_dlpack import camera
    s, m = small_scene(max(n, 1), 11, shape)
    if n == 0:
        s = {k: v[:0] for k, v in s.items()}
    if cap == -1:
        # a camera looking away from the scene: every Gaussian culled (w <= 0)
        m = camera.to_col_major(camera.world2ndc((0.0, 0.0, 3.0), (0.0, 0.0, 6.0), shape))
        cap = 0
    xyz, ls, q, op, rgb = (_t(s[k]) for k in ("xyz", "log_scale", "quat", "opacity", "rgb"))
    m16 = _t(m)
    pxd, _, aabb, splat2 = SG.project_fwd(xyz, ls, q, op, rgb, m16, shape, with_conic=False)
    k_ref = torch.zeros(1, device=DEV, dtype=torch.int32)
    capacity = cap
    if capacity == 0:
        kb = SG.bin_gauss(shape, ts, pxd, aabb)[1].numel()
        capacity = kb + 17
    t2i, i2p, _ = SG.bin_gauss_capacity(shape, ts, pxd, aabb, capacity, k_ref)
    gacc = torch.full((9 * n,), 3.0, device=DEV, dtype=torch.float32)
    rec, t2i_b, i2p_b, k_b, pxd_b, aabb_b = SG.project_bin(xyz, ls, q, op, rgb, m16, shape, ts,
                                                           capacity, grad_acc=gacc, boxes=boxes)
    torch.cuda.synchronize()
    k = int(k_ref.item())
    assert int(k_b.item()) == k
    assert torch.equal(rec, splat2)
    assert torch.equal(t2i_b, t2i)
    if k <= capacity:
        assert torch.equal(i2p_b[:k], i2p[:k])
    else:
        assert int(t2i_b.abs().sum().item()) == 0  # overflow: nothing binned
    if boxes:
        assert torch.equal(pxd_b, pxd) and torch.equal(aabb_b, aabb)
    else:
        assert pxd_b is None and aabb_b is None
    assert int((gacc != 0).sum().item()) == 0


@pytest.mark.parametrize("n", [20000, 1337])
def test_project_fwd_clears_accumulators_record_only(n):
    """project_fwd without the conic array (with_conic=False) writes the same
    records, depths and boxes, and with grad_acc it sets the [9N]
    accumulators to zero (filled with garbage first, n not a multiple of 64
    included): the fused-L2 adjoint into them equals the adjoint into fresh
    zeros up to the order of its float atomics (1e-6 of the largest)."""
    import del_splat_dlpack.SplatGauss.torch as SG
    from del_splat_dlpack import scene
    shape = (333, 201)
    s, m = small_scene(n, 23, shape)
    xyz, ls, q, op, rgb = (_t(s[k]) for k in ("xyz", "log_scale", "quat", "opacity", "rgb"))
    pxd, conic, aabb, splat2 = SG.project_fwd(xyz, ls, q, op, rgb, _t(m), shape)
    gacc = torch.full((9 * n + 64,), float("nan"), device=DEV)
    pxd2, none, aabb2, splat22 = SG.project_fwd(xyz, ls, q, op, rgb, _t(m), shape,
                                                with_conic=False, grad_acc=gacc[:9 * n])
    assert none is None
    for a, b in ((pxd, pxd2), (aabb, aabb2), (splat2, splat22)):
        assert torch.equal(a, b)
    np.testing.assert_array_equal(splat2.cpu().numpy()[:, 2:5], conic.cpu().numpy())
    assert (gacc[:9 * n] == 0).all() and gacc[9 * n:].isnan().all()
    t2i, i2p = SG.tile2idx_idx2pnt(shape, 16, pxd, aabb)
    img, alpha, nwalk, tfinal = SG.rasterize_fwd(shape, 16, splat2, t2i, i2p)
    target = _t(scene.target_image(shape, 78))
    ref = SG.rasterize_bwd_l2(shape, 16, splat2, t2i, i2p, nwalk, tfinal, img, target)
    got = SG.rasterize_bwd_l2(shape, 16, splat2, t2i, i2p, nwalk, tfinal, img, target,
                              grad_acc=gacc[:9 * n])
    for a, b in zip(got, ref):
        d = (a - b).abs().max().item()
        assert d <= 1e-6 * max(b.abs().max().item(), 1e-30), d


def test_project_bwd_add_and_overwrite():
    """dsplat_gauss_project_bwd adds into the gradient buffers by default and
    stores into them with DSPLAT_GRAD_OVERWRITE: add(into ones) == 1 + store."""
    import del_splat_dlpack.SplatGauss.torch as SG
    from del_splat_dlpack import _native
    shape = (160, 120)
    s, m = small_scene(3000, 77, shape)
    xyz, ls, q = _t(s["xyz"]), _t(s["log_scale"]), _t(s["quat"])
    n = xyz.shape[0]
    rng = np.random.default_rng(5)
    dpix = _t(rng.standard_normal((n, 2)).astype(np.float32))
    dcon = _t(rng.standard_normal((n, 3)).astype(np.float32))
    store = SG.project_bwd(xyz, ls, q, _t(m), shape, dpix, dcon)
    add = [torch.ones((n, k), device=DEV) for k in (3, 3, 4)]
    _native.call("dsplat_gauss_project_bwd", torch.cuda.current_stream().cuda_stream, n,
                 xyz.data_ptr(), ls.data_ptr(), q.data_ptr(), _t(m).data_ptr(), shape[0],
                 shape[1], _native.DSPLAT_SCALE_IS_LOG, dpix.data_ptr(), dcon.data_ptr(),
                 *[a.data_ptr() for a in add])
    for a, b in zip(add, store):
        np.testing.assert_allclose(a.cpu().numpy(), 1.0 + b.cpu().numpy(), rtol=1e-6, atol=1e-6)


def test_tile_schedule_orders_and_results():
    """dsplat_tile_order is a permutation by descending (capped) cost; the
    forward in any tile order is bit-identical to raster order and reports
    each tile's longest walk; the adjoint in any order agrees to float
    atomic reordering."""
    import del_splat_dlpack.SplatGauss.torch as SG
    rng = np.random.default_rng(3)
    cost = rng.integers(0, 5000, 9000).astype(np.int32)
    order = SG.tile_order(_t(cost)).cpu().numpy()
    assert sorted(order.tolist()) == list(range(9000))
    capped = np.minimum(cost[order], 2047)
    assert (np.diff(capped) <= 0).all()
    shape, ts = (171, 133), 16
    s, m = small_scene(4000, 91, shape)
    op, rgb = _t(s["opacity"]), _t(s["rgb"])
    pxd, conic, aabb, splat2 = SG.project_fwd(_t(s["xyz"]), _t(s["scale"]), _t(s["quat"]), op,
                                              rgb, _t(m), shape, log_scale=False)
    t2i, i2p = SG.tile2idx_idx2pnt(shape, ts, pxd, aabb)
    T = t2i.numel() - 1
    ref = SG.rasterize_fwd(shape, ts, splat2, t2i, i2p)
    perm = _t(rng.permutation(T).astype(np.int32))
    tc = torch.zeros(T, device=DEV, dtype=torch.int32)
    got = SG.rasterize_fwd(shape, ts, splat2, t2i, i2p, tile_order=perm, tile_cost=tc)
    for a, b in zip(ref, got):
        np.testing.assert_array_equal(a.cpu().numpy(), b.cpu().numpy())
    nw = ref[2].cpu().numpy()
    tw, th = (shape[0] + ts - 1) // ts, (shape[1] + ts - 1) // ts
    pad = np.zeros((th * ts, tw * ts), np.int64)
    pad[:shape[1], :shape[0]] = nw
    np.testing.assert_array_equal(tc.cpu().numpy(), pad.reshape(th, ts, tw, ts).max(axis=(1, 3)).ravel())
    d_img = _t(rng.standard_normal((shape[1], shape[0], 3)).astype(np.float32))
    g0 = SG.rasterize_bwd(shape, ts, splat2, t2i, i2p, ref[2], ref[3], d_img)
    g1 = SG.rasterize_bwd(shape, ts, splat2, t2i, i2p, ref[2], ref[3], d_img,
                          tile_order=SG.tile_order(tc))
    for a, b in zip(g0, g1):
        assert max_rel(b.cpu().numpy(), a.cpu().numpy()) <= 1e-5
    # order entries beyond the grid are skipped: every other tile renders as in
    # raster order
    bad = perm.clone()
    skipped = bad[:10].cpu().numpy().copy()
    bad[:10] = torch.arange(T + 3, T + 13, device=DEV, dtype=torch.int32)
    got = SG.rasterize_fwd(shape, ts, splat2, t2i, i2p, tile_order=bad)
    tile_of_pix = ((np.arange(shape[1])[:, None] // ts) * tw + np.arange(shape[0])[None, :] // ts)
    keep = ~np.isin(tile_of_pix, skipped)
    np.testing.assert_array_equal(got[0].cpu().numpy()[keep], ref[0].cpu().numpy()[keep])
    np.testing.assert_array_equal(got[2].cpu().numpy()[keep], ref[2].cpu().numpy()[keep])


def test_headline_frame_matches_oracle(oracle):
    """The headline workload at full size (1M Gaussians, 1920x1080, 32x32
    tiles, 512-thread forward, walk-ordered tiles, fused L2 adjoint at the
    benchmarked 256 threads and at 512): projection and lists bit-exact,
    pixels within 1e-4 with termination flips bounded in size and no more
    numerous than the oracle's own flips when its exp moves by one ulp,
    every gradient per element (1e-2 relative above 1% of the largest
    magnitude, 1e-4 of it below), all against the oracle's frame on the same
    inputs."""
    import del_splat_dlpack.SplatGauss.torch as SG
    from del_splat_dlpack import scene
    shape, ts, th = (1920, 1080), 32, 512
    s = scene.gaussians(1_000_000, 2, radius=2.0)
    m = scene.front_camera(shape)
    target = scene.target_image(shape, 102)
    # linear scales on both sides (the same float inputs for a bit-exact
    # projection)
    xyz, sc, q, op, rgb = (_t(s[k]) for k in ("xyz", "scale", "quat", "opacity", "rgb"))
    m16 = _t(m)
    pxd, conic, aabb, splat2 = SG.project_fwd(xyz, sc, q, op, rgb, m16, shape, log_scale=False)
    t2i, i2p = SG.bin_gauss(shape, ts, pxd, aabb)
    cost = torch.empty(t2i.numel() - 1, device=DEV, dtype=torch.int32)
    img, alpha, nwalk, tfinal = SG.rasterize_fwd(shape, ts, splat2, t2i, i2p, th, tile_cost=cost)
    order = SG.tile_order(cost)
    grads = {}
    for thb in (256, 512):
        dpix, dcon, dop, drgb = SG.rasterize_bwd_l2(shape, ts, splat2, t2i, i2p, nwalk, tfinal,
                                                    img, _t(target), thb, tile_order=order)
        dxyz, dsc, dq = SG.project_bwd(xyz, sc, q, m16, shape, dpix, dcon, log_scale=False)
        grads[thb] = [t.cpu().numpy() for t in (dxyz, dsc, dq, dop, drgb)]
    # the oracle's frame
    o_pxd, o_conic, o_aabb, _ = oracle.gauss_project(s["xyz"], s["quat"], s["scale"], m, shape)
    np.testing.assert_array_equal(pxd.cpu().numpy(), o_pxd)
    np.testing.assert_array_equal(conic.cpu().numpy(), o_conic)
    np.testing.assert_array_equal(aabb.cpu().numpy(), o_aabb)
    o_t2i, o_i2p = oracle.tile2pnt_gauss(o_pxd, o_aabb, shape, ts)
    np.testing.assert_array_equal(u32(t2i), o_t2i)
    np.testing.assert_array_equal(u32(i2p), o_i2p)
    o_img, o_alpha, o_nwalk, o_tfinal = oracle.gauss_rasterize(
        shape, o_pxd, o_conic, o_aabb, s["opacity"], s["rgb"], ts, o_t2i, o_i2p)
    own = oracle.self_flips(shape, o_pxd, o_conic, o_aabb, s["opacity"], s["rgb"], ts, o_t2i,
                            o_i2p, o_nwalk)
    flips = _check_pixels(img.cpu().numpy(), alpha.cpu().numpy(),
                          nwalk.cpu().numpy().view(np.uint32), o_img, o_alpha, o_nwalk,
                          self_flip=own)
    print(f"threshold flips: {flips}")
    o_dpix, o_dcon, o_dop, o_drgb = oracle.gauss_rasterize_bwd(
        shape, o_pxd, o_conic, o_aabb, s["opacity"], s["rgb"], ts, o_t2i, o_i2p, o_nwalk,
        o_tfinal, (2.0 * (o_img - target)).astype(np.float32))
    o_dxyz, o_dsc, o_dq = oracle.gauss_project_bwd(s["xyz"], s["quat"], s["scale"], m, shape,
                                                   o_dpix, o_dcon)
    for thb, gl in grads.items():
        for name, g, o in zip(("xyz", "scale", "quat", "opacity", "rgb"), gl,
                              (o_dxyz, o_dsc, o_dq, o_dop, o_drgb)):
            check_grad(g, o, f"adjoint {thb} threads {name}", abs_frac=1e-4)


def _crafted_boxes(n, shape, seed):
    """Splats whose boxes put their edges exactly on pixel centres, one ulp
    either side of them, at random positions, or at NaN / +-inf / +-1e30, some
    inverted; every tile lists every splat (in depth order), so the box test
    alone decides which pixels a splat reaches."""
    w, h = shape
    rng = np.random.default_rng(seed)
    cx = rng.uniform(-10, w + 10, n)
    cy = rng.uniform(-10, h + 10, n)
    sig = rng.uniform(2.0, 12.0, (n, 2))
    th = rng.uniform(0, np.pi, n)
    c, s_ = np.cos(th), np.sin(th)
    a = c * c / sig[:, 0] ** 2 + s_ * s_ / sig[:, 1] ** 2
    b = c * s_ * (1 / sig[:, 0] ** 2 - 1 / sig[:, 1] ** 2)
    cc = s_ * s_ / sig[:, 0] ** 2 + c * c / sig[:, 1] ** 2
    r = 3.0 * sig.max(axis=1)
    box = np.stack([cx - r, cy - r, cx + r, cy + r], 1).astype(np.float32)
    special = np.float32([np.nan, np.inf, -np.inf, 1e30, -1e30])
    for i in range(n):
        for e in range(4):
            mode = rng.integers(0, 7)
            v = box[i, e]
            if mode in (1, 2, 3):
                v = np.float32(np.floor(v) + 0.5)  # a pixel centre
                if mode == 2:
                    v = np.nextafter(v, np.float32(np.inf))
                elif mode == 3:
                    v = np.nextafter(v, np.float32(-np.inf))
            elif mode == 4 and rng.random() < 0.15:
                v = special[rng.integers(0, len(special))]
            box[i, e] = v
        if rng.random() < 0.05:
            box[i, [0, 2]] = box[i, [2, 0]]  # inverted
    pxd = np.stack([cx, cy, rng.uniform(0.5, 5.0, n)], 1).astype(np.float32)
    conic = np.stack([a, b, cc], 1).astype(np.float32)
    op = rng.uniform(0.2, 0.9, n).astype(np.float32)
    rgb = rng.random((n, 3)).astype(np.float32)
    splat2 = np.zeros((n, 16), np.float32)
    splat2[:, 0:2] = pxd[:, :2]
    splat2[:, 2:5] = conic
    splat2[:, 5] = op
    splat2[:, 6:9] = rgb
    splat2[:, 9:13] = box
    return pxd, conic, box, op, rgb, splat2


@pytest.mark.parametrize("ts,threads", [(8, 64), (16, 64), (16, 128), (16, 256), (32, 256),
                                        (32, 512), (32, 1024)])
def test_box_edges_on_pixel_centres(oracle, ts, threads):
    """The inclusive box test at its edges (the raster forms it per tile row /
    column and per 8x8 pixel region, or per pixel, by workgroup shape):
    forward pixels and the adjoint against the oracle's per-pixel float
    comparisons on boxes crafted to sit on, and one ulp off, pixel centres."""
    import del_splat_dlpack.SplatGauss.torch as SG
    shape = (100, 70)
    n = 300
    pxd, conic, box, op, rgb, splat2 = _crafted_boxes(n, shape, 7 + ts + threads)
    ntile = ((shape[0] + ts - 1) // ts) * ((shape[1] + ts - 1) // ts)
    order = np.argsort(pxd[:, 2], kind="stable").astype(np.uint32)
    t2i = (np.arange(ntile + 1) * n).astype(np.uint32)
    i2p = np.tile(order, ntile)
    o_img, o_alpha, o_nwalk, o_tfinal = oracle.gauss_rasterize(shape, pxd, conic, box, op, rgb,
                                                              ts, t2i, i2p)
    g_t2i, g_i2p = _t(t2i.view(np.int32)), _t(i2p.view(np.int32))
    img, alpha, nwalk, tfinal = SG.rasterize_fwd(shape, ts, _t(splat2), g_t2i, g_i2p, threads)
    _check_pixels(img.cpu().numpy(), alpha.cpu().numpy(), nwalk.cpu().numpy().view(np.uint32),
                  o_img, o_alpha, o_nwalk)
    d_img = np.random.default_rng(ts).standard_normal((shape[1], shape[0], 3)).astype(np.float32)
    g = SG.rasterize_bwd(shape, ts, _t(splat2), g_t2i, g_i2p, nwalk, tfinal, _t(d_img), None,
                         threads)
    o = oracle.gauss_rasterize_bwd(shape, pxd, conic, box, op, rgb, ts, t2i, i2p,
                                   nwalk.cpu().numpy().view(np.uint32), tfinal.cpu().numpy(),
                                   d_img)
    for name, a, b in zip(("d_pixxy", "d_conic", "d_opacity", "d_rgb"), g, o):
        check_grad(a.cpu().numpy(), b, f"ts={ts} threads={threads} {name}", abs_frac=1e-4)
    # the crafted edges do decide pixels: moving every edge one ulp outward
    # changes the image
    wide = splat2.copy()
    for e, d in ((9, -np.inf), (10, -np.inf), (11, np.inf), (12, np.inf)):
        wide[:, e] = np.nextafter(wide[:, e], np.float32(d))
    img2 = SG.rasterize_fwd(shape, ts, _t(wide), g_t2i, g_i2p, threads)[0]
    assert (img2 != img).any()


@pytest.mark.parametrize("ts,threads", [(16, 64), (16, 256), (32, 256), (32, 512)])
def test_opaque_splat_on_a_pixel_centre(oracle, ts, threads):
    """An opacity-1 splat whose mean is a pixel centre (e = 1 there, the walk
    ends on it): the stored transmittance is 2^-24 of the one in front, never
    0 (the kernels stage opacity at most 1 - 2^-24, the oracle floors 1 - e at
    2^-24, which agree here), and the adjoint's gradients are finite and
    match the oracle's."""
    import del_splat_dlpack.SplatGauss.torch as SG
    shape = (32, 32)
    pxd = np.array([[8.5, 8.5, 0.5], [8.5, 8.5, 0.2], [20.3, 11.7, 0.4]], np.float32)
    conic = np.array([[0.1, 0.0, 0.1], [0.1, 0.0, 0.1], [0.05, 0.01, 0.08]], np.float32)
    box = np.array([[0, 0, 32, 32]] * 3, np.float32)
    op = np.array([1.0, 0.5, 1.0], np.float32)
    rgb = np.array([[0.2, 0.3, 0.4], [0, 1, 0], [0.9, 0.1, 0.5]], np.float32)
    splat2 = np.zeros((3, 16), np.float32)
    splat2[:, 0:2] = pxd[:, :2]
    splat2[:, 2:5] = conic
    splat2[:, 5] = op
    splat2[:, 6:9] = rgb
    splat2[:, 9:13] = box
    t2i, i2p = oracle.tile2pnt_gauss(pxd, box, shape, ts)
    o_img, o_alpha, o_nwalk, o_tfinal = oracle.gauss_rasterize(shape, pxd, conic, box, op, rgb,
                                                              ts, t2i, i2p)
    g_t2i, g_i2p = _t(t2i.view(np.int32)), _t(i2p.view(np.int32))
    img, alpha, nwalk, tfinal = SG.rasterize_fwd(shape, ts, _t(splat2), g_t2i, g_i2p, threads)
    np.testing.assert_allclose(img.cpu().numpy(), o_img, atol=1e-5)
    np.testing.assert_array_equal(nwalk.cpu().numpy().view(np.uint32), o_nwalk)
    assert tfinal.cpu().numpy()[8, 8] == np.float32(2.0 ** -24)
    d_img = np.ones((shape[1], shape[0], 3), np.float32)
    g = SG.rasterize_bwd(shape, ts, _t(splat2), g_t2i, g_i2p, nwalk, tfinal, _t(d_img), None,
                         threads)
    o = oracle.gauss_rasterize_bwd(shape, pxd, conic, box, op, rgb, ts, t2i, i2p,
                                   nwalk.cpu().numpy().view(np.uint32), tfinal.cpu().numpy(),
                                   d_img)
    for name, a, b in zip(("d_pixxy", "d_conic", "d_opacity", "d_rgb"), g, o):
        a = a.cpu().numpy()
        assert np.isfinite(a).all(), name
        check_grad(a, b, f"ts={ts} threads={threads} {name}")
