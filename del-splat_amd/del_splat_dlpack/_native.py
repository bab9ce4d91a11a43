"""ctypes binding of libdsplat.so (the HIP library's C ABI, include/dsplat.h).

The library is built in-tree (`make -C del-splat_amd`); importing this module
never compiles anything. There is no CPU implementation behind these calls:
if the library or a GPU is missing, `lib()` raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.normpath(os.path.join(_HERE, "..", "libdsplat.so"))
_lib = None

_vp = ctypes.c_void_p
_u32 = ctypes.c_uint32
_f32 = ctypes.c_float
_sz = ctypes.c_size_t
_int = ctypes.c_int
_u64 = ctypes.c_uint64

DSPLAT_SCALE_IS_LOG = 1
DSPLAT_OPACITY_IS_LOGIT = 2
DSPLAT_GRAD_OVERWRITE = 4
DSPLAT_ERR_ARG = -1  # status: invalid argument
DSPLAT_ERR_CAPACITY = -3  # status: workspace, scratch or capacity too small


class ViewDesc(ctypes.Structure):
    """dsplat_view_desc: column-major world->ndc 4x4 and the image size."""
    _fields_ = [("world2ndc", ctypes.c_float * 16), ("img_w", ctypes.c_uint32),
                ("img_h", ctypes.c_uint32)]

# name -> (restype, argtypes)
_SIGS = {
    "dsplat_last_error": (ctypes.c_char_p, []),
    "dsplat_version": (_int, []),
    "dsplat_gauss_pnt2splat3_to_pnt2splat2": (_int, [_vp, _u32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _u32, _u32]),
    "dsplat_gauss_count_pairs": (_int, [_vp, _u32, _u32, _u32, _u32, _vp, _vp]),
    "dsplat_tile2idx_idx2pnt_workspace_size": (_sz, [_u32, _u32, _u32]),
    "dsplat_gauss_tile2idx_idx2pnt_ws": (_int, [_vp, _u32, _u32, _u32, _u32, _vp, _vp, _vp, _vp, _u32, _vp, _sz, _vp]),
    "dsplat_gauss_tile2idx_idx2pnt_noalloc": (_int, [_vp, _u32, _u32, _u32, _u32, _vp, _vp, _vp, _vp, _u32, _vp]),
    "dsplat_gauss_tile2idx_idx2pnt": (_int, [_vp, _u32, _u32, _u32, _u32, _vp, _vp, _vp, ctypes.POINTER(_vp), ctypes.POINTER(_u32)]),
    "dsplat_bin_points_workspace_size": (_sz, [_u32]),
    "dsplat_bin_pairs_workspace_size": (_sz, [_u32]),
    "dsplat_gauss_bin_count": (_int, [_vp, _u32, _u32, _u32, _u32, _vp, _vp, _vp, _sz, _vp]),
    "dsplat_bin_order": (_int, [_vp, _u32, _vp, _sz]),
    "dsplat_gauss_bin_fill": (_int, [_vp, _u32, _u32, _u32, _u32, _vp, _vp, _sz, _vp, _vp, _vp, _u32, _vp, _sz]),
    "dsplat_circle_bin_count": (_int, [_vp, _u32, _u32, _u32, _u32, _vp, _vp, _vp, _sz, _vp]),
    "dsplat_circle_bin_fill": (_int, [_vp, _u32, _u32, _u32, _u32, _vp, _vp, _vp, _sz, _vp, _vp, _vp, _u32, _vp, _sz]),
    "dsplat_free": (_int, [_vp, _vp]),
    "dsplat_gauss_rasterize_pnt2splat2": (_int, [_vp, _u32, _u32, _vp, _vp, _vp, _vp, _vp, _vp, _u32, _vp, _vp]),
    "dsplat_gauss_project_fwd": (_int, [_vp, _u32, _vp, _vp, _vp, _vp, _vp, _vp, _u32, _u32, _u32, _vp, _vp, _vp, _vp, _vp]),
    "dsplat_gauss_project_bin_ws": (_int, [_vp, _u32, _vp, _vp, _vp, _vp, _vp, _vp, _u32, _u32, _u32, _vp, _vp, _vp, _vp, _u32, _vp, _vp, _u32, _vp, _sz, _vp]),
    "dsplat_gauss_rasterize_fwd": (_int, [_vp, _u32, _u32, _u32, _u32, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "dsplat_gauss_rasterize_bwd": (_int, [_vp, _u32, _u32, _u32, _u32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "dsplat_gauss_rasterize_fwd_ordered": (_int, [_vp, _u32, _u32, _u32, _u32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "dsplat_gauss_rasterize_bwd_ordered": (_int, [_vp, _u32, _u32, _u32, _u32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "dsplat_tile_order": (_int, [_vp, _u32, _vp, _vp]),
    "dsplat_gauss_rasterize_bwd_l2": (_int, [_vp, _u32, _u32, _u32, _u32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "dsplat_gauss_project_bwd": (_int, [_vp, _u32, _vp, _vp, _vp, _vp, _u32, _u32, _u32, _vp, _vp, _vp, _vp, _vp]),
    "dsplat_sh_to_rgb": (_int, [_vp, _u32, _u32, _u32, _vp, _vp, _vp, _vp]),
    "dsplat_mv_set_sh": (_int, [_vp, _u32, _u32]),
    "dsplat_sh_to_rgb_bwd": (_int, [_vp, _u32, _u32, _u32, _vp, _vp, _vp, _vp, _u32, _vp, _vp]),
    "dsplat_circle_pnt2splat3_to_pnt2splat2": (_int, [_vp, _u32, _vp, _vp, _vp, _vp, _u32, _u32, _f32]),
    "dsplat_circle_tile2idx_idx2pnt_ws": (_int, [_vp, _u32, _u32, _u32, _u32, _vp, _vp, _vp, _vp, _u32, _vp, _sz, _vp]),
    "dsplat_circle_tile2idx_idx2pnt": (_int, [_vp, _u32, _u32, _u32, _u32, _vp, _vp, _vp, ctypes.POINTER(_vp), ctypes.POINTER(_u32)]),
    "dsplat_circle_splat": (_int, [_vp, _u32, _u32, _vp, _vp, _vp, _vp, _u32, _vp, _vp]),
    "dsplat_point_rasterize_zbuffer": (_int, [_vp, _u32, _vp, _vp, _vp, _vp, _u32, _u32, _vp]),
    "dsplat_circle_rasterize_zbuffer": (_int, [_vp, _u32, _vp, _vp, _vp, _vp, _vp, _u32, _u32, _vp]),
    "dsplat_point_rasterize_pix": (_int, [_vp, _u32, _vp, _vp, _vp, _u32, _u32, _vp]),
    "dsplat_mv_plan_create": (_int, [_u32, _u32, _u64, _u32, _u32, _u32, ctypes.POINTER(_vp)]),
    "dsplat_mv_plan_destroy": (_int, [_vp]),
    "dsplat_mv_set_views": (_int, [_vp, _vp, _u32, _vp]),
    "dsplat_mv_plan_info": (_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "dsplat_mv_outputs": (_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "dsplat_mv_forward": (_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _u32, _vp, _vp]),
    "dsplat_mv_backward": (_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _u32]),
    "dsplat_adam_step": (_int, [_vp, _u64, _vp, _vp, _vp, _vp, _f32, _f32, _f32, _f32, _vp]),
    "dsplat_mv_train_step": (_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _f32, _f32, _f32, _f32]),
    "dsplat_mv_graph_capture": (_int, [_vp, _u32, _vp, _vp, _vp, _vp, _vp, _u32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _f32, _f32, _f32, _f32]),
    "dsplat_mv_graph_launch": (_int, [_vp, _vp]),
    "dsplat_memcpy_d2d": (_int, [_vp, _vp, _vp, _u64]),
    "dsplat_exclusive_scan_workspace_size": (_sz, [_u32]),
    "dsplat_exclusive_scan_u32": (_int, [_vp, _u32, _vp, _vp, _vp, _sz]),
    "dsplat_radix_sort_workspace_size": (_sz, [_u32]),
    "dsplat_radix_sort_by_key_u64_u32": (_int, [_vp, _u32, _vp, _vp, _u32, _vp, _sz]),
    "dsplat_fill_f32": (_int, [_vp, _u32, _vp, _f32]),
}


class DsplatError(RuntimeError):
    pass


def lib():
    """The loaded library; raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise DsplatError(
                f"{LIB_PATH} is missing: build the HIP library first "
                "(python -c 'import __graft_entry__ as g; g.build()' or make -C del-splat_amd)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def symbols():
    return list(_SIGS)


def check(rc, what=""):
    """Raise on a non-zero status of the C ABI."""
    if rc != 0:
        msg = lib().dsplat_last_error().decode(errors="replace")
        if rc == -1:
            raise ValueError(f"{what}: {msg}")
        raise DsplatError(f"{what}: status {rc}: {msg}")


def call(name, *args):
    check(getattr(lib(), name)(*args), name)
