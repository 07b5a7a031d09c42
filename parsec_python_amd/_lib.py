"""ctypes binding of libparsec_hip.so (the C-ABI declared in include/parsec_hip.h).

The library is built in-tree by ``parsec_python_amd.build_lib``.  There is no
fallback: if the shared object is missing or a call fails, a
``ParsecHipError`` is raised.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libparsec_hip.so")


class ParsecHipError(RuntimeError):
    pass


_vp = C.c_void_p
_i = C.c_int
_ll = C.c_longlong
_d = C.c_double
_f = C.c_float
_sz = C.c_size_t
_u64 = C.c_ulonglong
_pi = C.POINTER(C.c_int)
_pd = C.POINTER(C.c_double)
_pll = C.POINTER(C.c_longlong)
_psz = C.POINTER(C.c_size_t)

# name -> argtypes (all return int status unless listed in _RESTYPES)
PROTOTYPES = {
    "pc_last_error": [],
    "pc_version": [],
    "pc_grid_create": [_i, _i, _i, _pi, _pi, _i, _pd, C.POINTER(_vp), _pll],
    "pc_grid_destroy": [_vp],
    "pc_stencil_f64": [_vp, _vp, _vp, _vp, _vp, _i, _ll, _d, _d, _d, _d, _vp, _vp],
    "pc_stencil_f32": [_vp, _vp, _vp, _vp, _vp, _i, _ll, _d, _d, _d, _d, _vp, _vp],
    "pc_stencil_set_variant": [_i],
    "pc_stencil_direct_f64": [_vp, _vp, _vp, _vp, _vp, _i, _ll, _d, _d, _d, _d, _vp, _vp],
    "pc_stencil_direct_f32": [_vp, _vp, _vp, _vp, _vp, _i, _ll, _d, _d, _d, _d, _vp, _vp],
    "pc_stencil_num_partials": [_vp, _i, _i, _pi, _pi],
    "pc_proj_create": [_i, _pi, _pi, _pd, _pd, _ll, C.POINTER(_vp)],
    "pc_proj_destroy": [_vp],
    "pc_proj_ncols": [_vp, _pi],
    "pc_proj_gather_f64": [_vp, _vp, _ll, _i, _vp, _vp],
    "pc_proj_gather_f32": [_vp, _vp, _ll, _i, _vp, _vp],
    "pc_proj_scatter_f64": [_vp, _vp, _vp, _ll, _i, _d, _vp],
    "pc_proj_scatter_f32": [_vp, _vp, _vp, _ll, _i, _d, _vp],
    "pc_gram_workspace": [_ll, _i, _i, _i, _psz],
    "pc_gram_f64": [_vp, _ll, _vp, _ll, _ll, _i, _i, _vp, _ll, _i, _vp, _sz, _vp],
    "pc_gram_f32": [_vp, _ll, _vp, _ll, _ll, _i, _i, _vp, _ll, _i, _vp, _sz, _vp],
    "pc_rotate_f64": [_vp, _ll, _vp, _i, _vp, _ll, _ll, _i, _i, _vp, _sz, _vp],
    "pc_rotate_f32": [_vp, _ll, _vp, _i, _vp, _ll, _ll, _i, _i, _vp, _sz, _vp],
    "pc_rotate_workspace": [_i, _i, _i, _psz],
    "pc_rotate_set_split_bf16": [_i],
    "pc_dense_create": [_vp, _i, C.POINTER(_vp)],
    "pc_dense_destroy": [_vp],
    "pc_dense_rr": [_vp, _i, _vp, _i, _vp, _i, _vp, _i, _vp, _pd],
    "pc_dense_syevd": [_vp, _i, _vp, _i, _pd],
    "pc_dense_trd_control": [_vp, _i, _i, C.POINTER(_i)],
    "pc_dense_set_method": [_vp, _i],
    "pc_dense_tri_fallbacks": [_vp, _pi],
    "pc_dense_set_triangular": [_vp, _i],
    "pc_dense_set_backtransform": [_vp, _i],
    "pc_dense_chol_async": [_vp, _i, _vp, _i],
    "pc_dense_gram_chol_async": [_vp, _i, _vp, _ll, _ll, _i, _vp, _i, _vp, _sz],
    "pc_dense_shift_count": [_vp, C.POINTER(_i)],
    "pc_density_f64": [_vp, _ll, _ll, _i, _vp, _vp, _vp],
    "pc_density_f32": [_vp, _ll, _ll, _i, _vp, _vp, _vp],
    "pc_num_blocks": [_ll, _pi],
    "pc_scf_post": [_ll, _vp, _d, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "pc_xc": [_ll, _vp, _vp, _vp, _vp, _vp],
    "pc_hartree_rhs": [_ll, _vp, _vp, _d, _vp, _vp],
    "pc_reduce_sets": [_vp, _i, _i, _vp, _vp],
    "pc_dots_nchunk": [_ll, _pi],
    "pc_dots": [_ll, _vp, _ll, _i, _vp, _vp, _vp, _vp],
    "pc_lincomb": [_ll, _vp, _ll, _i, _pd, _d, _vp, _d, _vp, _vp],
    "pc_lincomb_dev": [_ll, _vp, _ll, _i, _vp, _d, _d, _vp, _d, _vp, _d, _vp, _d, _vp, _vp],
    "pc_lanczos_nblocks": [_ll, _pi],
    "pc_lanczos_start": [_ll, _vp, _vp, _vp],
    "pc_lanczos_step": [_ll, _vp, _vp, _vp, _vp, _i, _vp, _vp],
    "pc_convert_f64_to_f32": [_ll, _i, _vp, _ll, _vp, _ll, _vp],
    "pc_convert_f32_to_f64": [_ll, _i, _vp, _ll, _vp, _ll, _vp],
    "pc_rand_f64": [_ll, _i, _ll, _u64, _i, _vp, _vp],
    "pc_rand_f32": [_ll, _i, _ll, _u64, _i, _vp, _vp],
    "pc_pcg_workspace": [_vp, _psz],
    "pc_pcg_solve": [_vp, _d, _vp, _vp, _i, _d, _i, _i, _vp, _sz, _pi, _vp, _vp],
    "pc_pcg_solve_noalloc": [_vp, _d, _vp, _vp, _i, _d, _i, _i, _pi, _vp, _vp],
    "pc_pcg_noalloc_max_points": [_pll],
    "pc_mg_create": [_vp, _d, C.POINTER(_vp)],
    "pc_mg_destroy": [_vp],
    "pc_mg_workspace": [_vp, _psz],
    "pc_pcg_solve_mg": [_vp, _vp, _vp, _i, _d, _i, _vp, _sz, _pi, _vp, _vp],
    "pc_mg_set_first_check": [_vp, _i],
    "pc_sad_eval": [_vp, _d, _d, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "pc_spline_eval": [_vp, _vp, _vp, _vp, _i, _vp, _ll, _vp, _vp],
    # learned initialiser (3D U-Net) kernels
    "pc_unet_conv3d_fwd": [_vp, _i, _i, _vp, _vp, _i, _i, _i, _i, _i, _f, _vp, _i, _vp],
    "pc_unet_conv3d_wflip": [_vp, _i, _i, _vp, _vp],
    "pc_unet_wgrad_workspace": [_i, _i, _psz],
    "pc_unet_conv3d_wgrad": [_vp, _i, _i, _vp, _i, _i, _i, _i, _i, _vp, _vp, _sz, _vp],
    "pc_unet_act_bwd_workspace": [_i, _ll, _psz],
    "pc_unet_act_bwd": [_vp, _i, _vp, _i, _i, _i, _i, _i, _f, _vp, _i, _vp, _vp, _sz, _vp],
    "pc_unet_pool2_fwd": [_vp, _i, _i, _i, _i, _i, _vp, _i, _vp],
    "pc_unet_pool2_bwd": [_vp, _i, _vp, _i, _i, _i, _i, _i, _vp, _i, _vp],
    "pc_unet_up2_fwd": [_vp, _i, _i, _i, _i, _i, _vp, _i, _vp],
    "pc_unet_up2_bwd": [_vp, _i, _i, _i, _i, _i, _vp, _i, _vp],
    "pc_unet_cast_bf16": [_vp, _ll, _vp, _vp],
    "pc_unet_adam": [_vp, _vp, _vp, _vp, _ll, _d, _d, _d, _d, _d, _i, _vp, _vp],
    "pc_unet_mse_workspace": [_psz],
    "pc_unet_mse": [_vp, _vp, _ll, _vp, _vp, _vp, _sz, _vp],
}

_RESTYPES = {"pc_last_error": C.c_char_p}

_lib = None


def lib():
    """Load the library once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ParsecHipError(
            f"{LIB_PATH} not found: build it with `python -m parsec_python_amd.build_lib` "
            "(there is no CPU fallback for the HIP path)")
    handle = C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL)
    for name, argtypes in PROTOTYPES.items():
        try:
            fn = getattr(handle, name)
        except AttributeError:
            continue
        fn.argtypes = argtypes
        fn.restype = _RESTYPES.get(name, C.c_int)
    _lib = handle
    return _lib


def check(status: int, what: str = "") -> None:
    if status != 0:
        msg = lib().pc_last_error().decode(errors="replace")
        raise ParsecHipError(f"{what or 'parsec_hip'} failed (status {status}): {msg}")


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args), name)


def exported_symbols() -> list[str]:
    """Names the C header declares; used to check the shared object exports them."""
    return list(PROTOTYPES.keys())
