"""Loader of libexosht.so and ctypes prototypes of the C ABI
(include/exo_sht.h)."""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_double, c_float, c_int, c_long, c_void_p, c_char_p

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libexosht.so")

_dp = POINTER(c_double)
_fp = POINTER(c_float)

# name -> (restype, argtypes); every symbol of include/exo_sht.h
PROTOTYPES = {
    "exo_strerror": (c_char_p, [c_int]),
    "exo_device_count": (c_int, []),
    "exo_inigau": (c_int, [c_int, _dp, _dp]),
    "exo_legini": (c_int, [c_int, c_int, c_int, _dp, _dp, _dp, _dp, c_double]),
    "exo_legini_tables": (c_int, [_dp]),
    "exo_sp2fc_f64": (c_int, [_dp, _dp, c_int]),
    "exo_sp2fc_f32": (c_int, [_fp, _fp, c_int]),
    "exo_sp2fcdmu_f64": (c_int, [_dp, _dp, c_int]),
    "exo_sp2fcdmu_f32": (c_int, [_fp, _fp, c_int]),
    "exo_fc2sp_f64": (c_int, [_dp, _dp, c_int]),
    "exo_fc2sp_f32": (c_int, [_fp, _fp, c_int]),
    "exo_dv2uv_f64": (c_int, [_dp, _dp, _dp, _dp]),
    "exo_dv2uv_f32": (c_int, [_fp, _fp, _fp, _fp]),
    "exo_uv2dv_f64": (c_int, [_dp, _dp, _dp, _dp]),
    "exo_uv2dv_f32": (c_int, [_fp, _fp, _fp, _fp]),
    "exo_qtend_f64": (c_int, [_dp, _dp, _dp, _dp]),
    "exo_qtend_f32": (c_int, [_fp, _fp, _fp, _fp]),
    "exo_mktend_f64": (c_int, [_dp] * 9),
    "exo_mktend_f32": (c_int, [_fp] * 9),
    "exo_gp2fc_f64": (c_int, [_dp, c_int, c_int]),
    "exo_gp2fc_f32": (c_int, [_fp, c_int, c_int]),
    "exo_fc2gp_f64": (c_int, [_dp, c_int, c_int]),
    "exo_fc2gp_f32": (c_int, [_fp, c_int, c_int]),
    "exo_sht_plan_create": (c_int, [POINTER(c_void_p), c_int, _dp, _dp, _dp, _dp, c_double]),
    "exo_sht_plan_destroy": (c_int, [c_void_p]),
    "exo_sht_plan_dims": (c_int, [c_void_p, POINTER(c_int)]),
    "exo_sht_plan_grid": (c_int, [c_void_p, _dp, _dp]),
    "exo_sht_plan_tables": (c_int, [c_void_p, _dp]),
    "exo_sht_inverse": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_long, c_int,
                                c_void_p, c_long, c_int, c_int, c_void_p]),
    "exo_sht_inverse_uv": (c_int, [c_void_p, c_int, c_int, c_void_p, c_long, c_int, c_int,
                                   c_void_p, c_long, c_int, c_int, c_int, c_void_p]),
    "exo_sht_direct": (c_int, [c_void_p, c_int, c_int, c_void_p, c_long, c_int, c_void_p,
                               c_long, c_int, c_int, c_void_p]),
    "exo_sht_roundtrip": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p,
                                  c_void_p, c_void_p]),
    "exo_sht_inverse_all": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p]),
    "exo_sht_roundtrip_noalloc": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p,
                                          c_void_p, c_void_p]),
    "exo_sht_sp2fc": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_long, c_int,
                              c_void_p, c_long, c_int, c_int, c_void_p]),
    "exo_sht_dv2uv": (c_int, [c_void_p, c_int, c_int, c_void_p, c_long, c_int, c_int,
                              c_void_p, c_long, c_int, c_int, c_int, c_void_p]),
    "exo_sht_fc2sp": (c_int, [c_void_p, c_int, c_int, c_void_p, c_long, c_int, c_void_p,
                              c_long, c_int, c_int, c_void_p]),
    "exo_sht_fft": (c_int, [c_void_p, c_int, c_int, c_void_p, c_long, c_void_p]),
    "exo_sht_uv2dv": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_long,
                              c_void_p, c_void_p, c_long, c_void_p]),
    "exo_sht_qtend": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_long, c_void_p,
                              c_void_p, c_void_p, c_long, c_void_p]),
    "exo_sht_mktend": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p,
                               c_long] + [c_void_p] * 6 + [c_long, c_void_p]),
    "exo_sht_set_tiles": (c_int, [c_void_p, c_int, c_int, c_int, c_int]),
    "exo_sht_get_tiles": (c_int, [c_void_p, c_int, POINTER(c_int)]),
    "exo_sht_autotune": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, POINTER(c_int)]),
    "exo_sht_ens_create": (c_int, [c_void_p, c_int, c_int, c_void_p]),
    "exo_sht_ens_destroy": (c_int, [c_void_p]),
    "exo_sht_ens_run": (c_int, [c_void_p, c_void_p]),
    "exo_sht_ens_run_eager": (c_int, [c_void_p, c_void_p]),
    "exo_sht_ens_run_groups": (c_int, [c_void_p, c_void_p]),
    "exo_sht_ens_run_members": (c_int, [c_void_p, c_void_p]),
    "exo_sht_ens_get_tiles": (c_int, [c_void_p, POINTER(c_int)]),
    "exo_sht_ens_set_tiles": (c_int, [c_void_p, POINTER(c_int)]),
    "exo_sht_ens_autotune": (c_int, [c_void_p, c_void_p, POINTER(c_int)]),
    "exo_sht_ens_get_schedule": (c_int, [c_void_p, POINTER(c_int)]),
    "exo_sht_ens_set_schedule": (c_int, [c_void_p, c_int]),
    "exo_si_create": (c_int, [POINTER(c_void_p), c_int, c_int, _dp, _dp, _dp, _dp]),
    "exo_si_destroy": (c_int, [c_void_p]),
    "exo_si_makebm": (c_int, [c_void_p, c_double, c_void_p]),
    "exo_si_bm1": (c_int, [c_void_p, _dp]),
    "exo_si_apply": (c_int, [c_void_p, c_int, c_double] + [c_void_p] * 8),
}

_lib = None


class ExoError(RuntimeError):
    pass


def lib() -> ctypes.CDLL:
    """The loaded library; raises if it has not been built (there is no
    fallback implementation)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ExoError(f"{LIB_PATH} is not built; run exoplasim_amd.build.build_library()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = lib().exo_strerror(rc).decode()
        raise ExoError(f"{what}: {msg} ({rc})")
