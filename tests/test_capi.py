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
