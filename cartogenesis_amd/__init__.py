"""MI355X-native (gfx950) ocean-circulation multigrid solver.

The product is ``libcgocean.so``: hand-written HIP kernels and a C ABI declared in
``include/cgocean.h``. This package loads it with ctypes and mirrors the reference's Kotlin
interface for the solve (OceanAccelerator, OceanStencil, OceanCirculation) in ``ocean.py``.
Importing the package never falls back to anything: if the library is missing, ``lib()`` raises.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libcgocean.so"

# Status codes and methods, as in include/cgocean.h.
CG_OK, CG_DECLINED, CG_CANCELLED = 0, 1, 2
CG_ERR_ARG, CG_ERR_DEVICE, CG_ERR_NOMEM, CG_ERR_STATE = -1, -2, -3, -4
CG_METHOD_REFERENCE, CG_METHOD_RESCUE, CG_METHOD_REFINED, CG_METHOD_FGMRES, CG_METHOD_KRYLOV = 0, 1, 2, 3, 4
CG_METHOD_LEARNED, CG_METHOD_LEARNED_FGMRES = 5, 6
CG_OPT_MIN_CELLS, CG_OPT_TILE_ROWS, CG_OPT_USE_GRAPH = 1, 2, 3
CG_OPT_FGMRES_RESTART, CG_OPT_TIME_KERNELS = 4, 6
CG_OPT_FUSED, CG_OPT_LEVEL_PASSES, CG_OPT_FUSE_OUTER, CG_OPT_DIRECT_BOTTOM = 7, 8, 9, 10
CG_OPT_STOP_PERMILLE, CG_OPT_ERROR_PICO, CG_OPT_NET_FALLBACK, CG_OPT_GROUP_MERGED = 11, 12, 13, 14


class CgSolution(ctypes.Structure):
    _fields_ = [("cycles", ctypes.c_int), ("relative_residual", ctypes.c_double),
                ("refinements", ctypes.c_int), ("status", ctypes.c_int)]


_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_vp = ctypes.c_void_p
_int, _dbl, _long, _u64 = ctypes.c_int, ctypes.c_double, ctypes.c_long, ctypes.c_uint64
_flt, _size = ctypes.c_float, ctypes.c_size_t
# int (*cg_precond_fn)(void *user, const float *v, float *z, void *stream)
PRECOND_FN = ctypes.CFUNCTYPE(_int, _vp, _vp, _vp, _vp)

# Every entry point of include/cgocean.h with its ctypes signature: (restype, argtypes).
SIGNATURES = {
    "cg_ocean_open": (_int, [_int, ctypes.POINTER(_vp), ctypes.c_char_p, ctypes.c_size_t]),
    "cg_ocean_close": (None, [_vp]),
    "cg_ocean_name": (_int, [_vp, ctypes.c_char_p, ctypes.c_size_t]),
    "cg_ocean_last_error": (ctypes.c_char_p, [_vp]),
    "cg_ocean_set_option": (_int, [_vp, _int, _long]),
    "cg_ocean_get_option": (_long, [_vp, _int]),
    "cg_ocean_stencil_upload": (_int, [_vp, _int, _int, _u8p, _f32p, _f32p, _f32p, _f32p, _vp,
                                       ctypes.POINTER(_u64)]),
    "cg_ocean_stencil_release": (_int, [_vp, _u64]),
    "cg_ocean_relax": (_int, [_vp, _u64, _f32p, _f32p, _f32p, _int, _vp]),
    "cg_ocean_relax_noalloc": (_int, [_vp, _u64, _vp, _vp, _vp, _int, _vp]),
    "cg_ocean_solve": (_int, [_vp, ctypes.POINTER(_u64), _int, _f32p, _vp, _dbl, _int, _int, _vp, _vp,
                              ctypes.POINTER(CgSolution), _vp]),
    "cg_ocean_plan_create": (_int, [_vp, ctypes.POINTER(_u64), _int, _int, ctypes.POINTER(_vp)]),
    "cg_ocean_plan_destroy": (_int, [_vp]),
    "cg_ocean_plan_create_batch": (_int, [_vp, ctypes.POINTER(_u64), _int, _int, _int, ctypes.POINTER(_vp)]),
    "cg_ocean_plan_set_forcing_at": (_int, [_vp, _int, _f32p]),
    "cg_ocean_plan_solve_all": (_int, [_vp, _dbl, _int, ctypes.POINTER(CgSolution)]),
    "cg_ocean_plan_read_at": (_int, [_vp, _int, _vp, _vp]),
    "cg_ocean_plan_set_forcing": (_int, [_vp, _f32p]),
    "cg_ocean_plan_solve": (_int, [_vp, _dbl, _int, _int, ctypes.POINTER(CgSolution)]),
    "cg_ocean_plan_solve_batch": (_int, [ctypes.POINTER(_vp), _int, _dbl, _int, ctypes.POINTER(CgSolution)]),
    "cg_ocean_group_create": (_int, [ctypes.POINTER(_vp), _int, ctypes.POINTER(_vp)]),
    "cg_ocean_group_solve": (_int, [_vp, _dbl, _int, ctypes.POINTER(CgSolution)]),
    "cg_ocean_group_destroy": (_int, [_vp]),
    "cg_ocean_plan_set_bodies": (_int, [_vp, _vp, _int]),
    "cg_ocean_plan_correct_bodies": (_int, [_vp, _f32p, _f32p]),
    "cg_ocean_plan_set_start": (_int, [_vp, _f32p]),
    "cg_ocean_plan_read": (_int, [_vp, _vp, _vp]),
    "cg_ocean_kernel_time_us": (_dbl, [_vp, _int, ctypes.POINTER(_long)]),
    "cg_ocean_plan_time_kernel": (_dbl, [_vp, _int, _int]),
    "cg_ocean_plan_autotune": (_int, [_vp, _int, _vp]),
    "cg_ocean_plan_set_tile_rows": (_int, [_vp, _vp]),
    "cg_ocean_water_on": (_int, [_vp, _u8p, _int, _int, _int, _int, _u8p]),
    "cg_ocean_curl_forcing": (_int, [_vp, _int, _int, _dbl, _dbl, _f64p, _f64p, _vp, _f64p]),
    "cg_ocean_plan_set_stress": (_int, [_vp, _dbl, _dbl, _f64p, _f64p, _vp]),
    "cg_ocean_plan_velocities": (_int, [_vp, _dbl, _dbl, _vp, _vp]),
    "cg_ocean_plan_set_preconditioner": (_int, [_vp, PRECOND_FN, _vp]),
    "cg_ocean_plan_set_coarse_net": (_int, [_vp, _int, _int, _vp, _size, _vp, _dbl]),
    "cg_unet_weight_count": (_int, [_int]),
    "cg_ocean_plan_apply_coarse_net": (_int, [_vp, _f32p, _f32p]),
    "cg_ocean_plan_capture_coarse": (_int, [_vp, _int, _dbl, _int, _vp]),
    "cg_ocean_plan_net_fallbacks": (_int, [_vp]),
    "cg_conv3x3": (_int, [_vp, _int, _vp, _int, _int, _int, _int, _vp, _int, _int, _vp, _int, _vp, _vp, _vp]),
    "cg_conv3x3_in3": (_int, [_vp, _int, _int, _int, _vp, _int, _int, _vp, _vp]),
    "cg_conv3x3_up": (_int, [_vp, _int, _vp, _int, _int, _int, _int, _vp, _int, _int, _vp, _int, _vp, _vp, _vp]),
    "cg_conv3x3_adjoint": (_int, [_vp, _int, _int, _int, _int, _vp, _int, _vp, _int, _vp, _vp, _vp]),
    "cg_conv3x3_flip_weights": (_int, [_vp, _int, _int, _vp, _vp]),
    "cg_conv3x3_wgrad_slices": (_int, [_int, _int, _int]),
    "cg_conv3x3_wgrad": (_int, [_vp, _int, _vp, _int, _vp, _int, _int, _int, _int, _vp, _size, _vp, _vp]),
    "cg_pool2": (_int, [_vp, _int, _int, _int, _int, _flt, _vp, _vp]),
    "cg_upsample2": (_int, [_vp, _int, _int, _int, _int, _flt, _vp, _vp]),
    "cg_relu_mask": (_int, [_vp, _vp, _long, _vp]),
    "cg_add_bf16": (_int, [_vp, _vp, _long, _vp, _vp]),
    "cg_unet_pack_input": (_int, [_vp, _vp, _flt, _long, _vp, _vp]),
    "cg_unet_unpack_output": (_int, [_vp, _vp, _flt, _long, _vp, _vp]),
    "cg_bottom_drag": (_dbl, []),
    "cg_beta_rows": (_int, [_int, _dbl, _f64p]),
    "cg_circulation_stencil": (_int, [_int, _int, _dbl, _dbl, _u8p, _f64p, _dbl, _vp, _f32p, _f32p, _f32p,
                                      _f32p, _f64p, _f32p]),
    "cg_coarse_water": (_int, [_u8p, _int, _int, _int, _int, _int, _u8p]),
    "cg_bottom_inverse": (_int, [_int, _int, _u8p, _f32p, _f32p, _f32p, _f32p, _int, _int, _i32p, _f64p]),
    "cg_level_shapes": (_int, [_int, _int, _dbl, _dbl, _int, _i32p, _i32p, _f64p, _f64p]),
    "cg_heat_diffusivity": (_dbl, [_dbl, _dbl]),
    "cg_heat_stencil": (_int, [_int, _int, _dbl, _dbl, _u8p, _f32p, _vp, _dbl, _int, _dbl, _dbl, _vp, _vp,
                               _f32p, _f32p, _f32p, _f32p, _f64p, _f32p]),
    "cg_water_bodies": (_int, [_u8p, _int, _int, _i32p]),
    "cg_resample": (_int, [_f32p, _int, _int, _int, _int, _f32p]),
    "cg_coarse_mean": (_int, [_f32p, _int, _int, _int, _int, _f32p]),
    "cg_problem_generate": (_int, [_int, _int, _u64, _dbl, _int, _vp, _vp, _vp]),
    "cg_problem_stress": (_int, [_int, _int, _u64, _int, _vp, _vp]),
    # include/cgsave.h: the ocean's fields in the save format, host only.
    "cg_save_crc32": (ctypes.c_uint32, [_vp, _size, ctypes.c_uint32]),
    "cg_save_chunk_checksum": (ctypes.c_uint32, [ctypes.c_uint32, _int, _vp, _size]),
    "cg_save_directory": (_int, [_int, _int, _vp, _int, ctypes.POINTER(ctypes.c_int64)]),
    "cg_ocean_records_write": (_int, [_int, _f32p, _f32p, _f32p, _f32p, _vp, _size, ctypes.POINTER(_size)]),
    "cg_ocean_records_read": (_int, [_vp, _size, _int, _f32p, _f32p, _f32p, _f32p, ctypes.c_char_p, _size]),
    "cg_save_inspect": (_int, [_vp, _size, ctypes.POINTER(_int), ctypes.POINTER(_int), ctypes.POINTER(ctypes.c_int64),
                               ctypes.POINTER(ctypes.c_int64), ctypes.c_char_p, _size]),
    "cg_save_read_ocean": (_int, [_vp, _size, _f32p, _f32p, _f32p, _f32p, ctypes.c_char_p, _size]),
    "cg_save_replace_ocean": (_int, [_vp, _size, _f32p, _f32p, _f32p, _f32p, _int, _vp, _size, ctypes.POINTER(_size),
                                     ctypes.c_char_p, _size]),
}

_lib = None


def lib() -> ctypes.CDLL:
    """The loaded HIP library. Raises if it has not been built: there is no fallback."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m cartogenesis_amd.build` "
                           "(or __graft_entry__.build()); there is no CPU fallback")
    L = ctypes.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


class CgError(RuntimeError):
    pass


def check(status: int, ctx=None, what: str = "") -> int:
    if status < 0:
        msg = lib().cg_ocean_last_error(ctx).decode() if ctx else ""
        raise CgError(f"{what} failed with status {status}: {msg}")
    return status


from . import ocean, problem  # noqa: E402,F401
