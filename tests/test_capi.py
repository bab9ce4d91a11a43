"""The C-ABI library loads and exports every symbol include/dsplat.h declares
(no device work: runs without a GPU)."""
import ctypes
import os
import re

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "dsplat.h")
LIB = os.path.join(ROOT, "del-splat_amd", "libdsplat.so")


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(dsplat_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = declared()
    assert "dsplat_gauss_pnt2splat3_to_pnt2splat2" in names
    assert "dsplat_gauss_tile2idx_idx2pnt" in names
    assert "dsplat_gauss_rasterize_pnt2splat2" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    from del_splat_dlpack import _native
    assert sorted(_native.symbols()) == declared()


def test_host_only_queries():
    from del_splat_dlpack import _native
    L = _native.lib()
    assert L.dsplat_version() >= 100
    assert L.dsplat_tile2idx_idx2pnt_workspace_size(1000, 100, 5000) > 5000 * 4
    assert L.dsplat_radix_sort_workspace_size(1 << 20) >= (1 << 20) * 12
    assert L.dsplat_exclusive_scan_workspace_size(0) > 0
    assert L.dsplat_last_error() == b""


def test_project_bin_argument_checks():
    """dsplat_gauss_project_bin_ws refuses bad arguments before any device
    work: a tile size other than 8, 16 or 32, null buffers, a grid of more
    than 65535 tiles along an axis, a workspace below the binning's size (the
    pointers below are never dereferenced: every check precedes the launch)."""
    from del_splat_dlpack import _native
    L = _native.lib()
    f = L.dsplat_gauss_project_bin_ws
    fake = 4096  # a 16-byte aligned stand-in address
    n, cap = 1000, 5000
    need = L.dsplat_tile2idx_idx2pnt_workspace_size(n, 0, cap)
    ok = dict(xyz=fake, scale=fake, quat=fake, op=fake, rgb=fake, m=fake, splat=fake,
              t2i=fake, i2p=fake, ws=fake, k=fake)

    def call(w=640, h=480, ts=16, ws_bytes=need, **over):
        a = dict(ok, **over)
        return f(None, n, a["xyz"], a["scale"], a["quat"], a["op"], a["rgb"], a["m"], w, h, 0,
                 None, None, a["splat"], None, ts, a["t2i"], a["i2p"], cap, a["ws"], ws_bytes,
                 a["k"])

    assert call(ts=12) == _native.DSPLAT_ERR_ARG
    assert call(xyz=None) == _native.DSPLAT_ERR_ARG
    assert call(splat=None) == _native.DSPLAT_ERR_ARG
    assert call(k=None) == _native.DSPLAT_ERR_ARG
    assert call(i2p=None) == _native.DSPLAT_ERR_ARG
    assert call(quat=fake + 4) == _native.DSPLAT_ERR_ARG  # not 16-byte aligned
    assert call(w=65536 * 8, ts=8) == _native.DSPLAT_ERR_ARG
    assert call(ws_bytes=need - 1) == _native.DSPLAT_ERR_CAPACITY
    assert b"workspace" in L.dsplat_last_error()
