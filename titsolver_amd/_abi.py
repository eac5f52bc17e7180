"""ctypes declarations of the C ABI in include/tit_hip.h.

The shared library is loaded from this package directory; a missing library
raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libtit_hip.so"

TIT_OK, TIT_E_ARG, TIT_E_HIP, TIT_E_OOM, TIT_E_NUMERIC, TIT_E_STATE = range(6)
STATUS_NAMES = {0: "TIT_OK", 1: "TIT_E_ARG", 2: "TIT_E_HIP", 3: "TIT_E_OOM",
                4: "TIT_E_NUMERIC", 5: "TIT_E_STATE"}

KERNEL_WENDLAND_C2 = 0
KERNEL_WENDLAND_C4 = 1
FLAVOUR_REF = 0
FLAVOUR_CFG = 1


class Params(C.Structure):
    _fields_ = [
        ("dim", C.c_int), ("kernel", C.c_int), ("flavour", C.c_int),
        ("ssprk_order", C.c_int),
        ("h", C.c_double), ("m", C.c_double), ("rho0", C.c_double),
        ("cs0", C.c_double), ("xi", C.c_double), ("g", C.c_double),
        ("mu", C.c_double), ("alpha", C.c_double), ("delta", C.c_double),
        ("cfl", C.c_double), ("c_force", C.c_double), ("c_visc", C.c_double),
        ("skin", C.c_double),
        ("post_integrate", C.c_int), ("reserved", C.c_int),
    ]


class Tank(C.Structure):
    _fields_ = [("lo", C.c_double * 3), ("hi", C.c_double * 3)]


class HostView(C.Structure):
    _fields_ = [
        ("r", C.POINTER(C.c_double)), ("v", C.POINTER(C.c_double)),
        ("rho", C.POINTER(C.c_double)), ("dv_dt", C.POINTER(C.c_double)),
        ("normal", C.POINTER(C.c_double)),
        ("n_fluid", C.c_size_t), ("n_fixed", C.c_size_t),
        ("m", C.POINTER(C.c_double)),
    ]


class Walls(C.Structure):
    _fields_ = [("verts", C.POINTER(C.c_double)), ("n_verts", C.c_uint32),
                ("n_faces", C.c_uint32), ("faces", C.POINTER(C.c_uint32)),
                ("cverts", C.POINTER(C.c_double)), ("n_cverts", C.c_uint32),
                ("n_cfaces", C.c_uint32), ("cfaces", C.POINTER(C.c_uint32))]


class Schedule(C.Structure):
    _fields_ = [("chunk_targets", C.c_int), ("lds_capacity", C.c_int),
                ("occupancy", C.c_double), ("cached", C.c_int), ("reserved", C.c_int)]


class LearnedReport(C.Structure):
    _fields_ = [("learned_steps", C.c_int), ("judged", C.c_int),
                ("rejected", C.c_int), ("plain_steps", C.c_int),
                ("max_rho_err", C.c_double), ("last_rho_err", C.c_double),
                ("max_front_err", C.c_double), ("max_dv_err", C.c_double),
                ("last_dv_err", C.c_double)]


class KernelStat(C.Structure):
    _fields_ = [("name", C.c_char * 48), ("launches", C.c_uint64),
                ("total_ms", C.c_double)]


# name -> (restype, argtypes); mirrors include/tit_hip.h one to one.
_DP = C.POINTER(C.c_double)
_FP = C.POINTER(C.c_float)
SIGNATURES = {
    "tit_hip_create": (C.c_int, [C.c_int, C.POINTER(Params), C.POINTER(C.c_void_p)]),
    "tit_hip_upload": (C.c_int, [C.c_void_p, C.POINTER(HostView), C.POINTER(Tank)]),
    "tit_hip_upload_batch": (C.c_int, [C.c_void_p, C.POINTER(HostView),
                                       C.POINTER(Tank), C.c_size_t]),
    "tit_hip_step": (C.c_int, [C.c_void_p, C.c_int, _DP]),
    "tit_hip_download": (C.c_int, [C.c_void_p, _DP, _DP, _DP, _DP]),
    "tit_hip_tank_clock": (C.c_int, [C.c_void_p, _DP, _DP]),
    "tit_hip_prepare": (C.c_int, [C.c_void_p]),
    "tit_hip_compute_time_step": (C.c_int, [C.c_void_p, _DP]),
    "tit_hip_compute_rhs": (C.c_int, [C.c_void_p, _DP, _DP]),
    "tit_hip_post_integrate": (C.c_int, [C.c_void_p, _DP]),
    "tit_hip_set_walls": (C.c_int, [C.c_void_p, C.POINTER(Walls)]),
    "tit_hip_neighbors": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64),
                                    C.POINTER(C.c_uint32), C.c_size_t,
                                    C.POINTER(C.c_size_t)]),
    "tit_hip_neighbors_noalloc": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64),
                                            C.POINTER(C.c_uint32), C.c_size_t,
                                            C.POINTER(C.c_size_t)]),
    "tit_hip_pair_counts": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32)]),
    "tit_hip_set_profiling": (C.c_int, [C.c_void_p, C.c_int]),
    "tit_hip_kernel_stats": (C.c_int, [C.c_void_p, C.POINTER(KernelStat), C.c_int,
                                       C.POINTER(C.c_int)]),
    "tit_hip_get_schedule": (C.c_int, [C.c_void_p, C.POINTER(Schedule)]),
    "tit_hip_set_schedule": (C.c_int, [C.c_void_p, C.POINTER(Schedule)]),
    "tit_hip_autotune": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(Schedule), _DP]),
    "tit_hip_autotune_clear_cache": (None, []),
    "tit_hip_counts": (C.c_int, [C.c_void_p, C.POINTER(C.c_size_t),
                                 C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "tit_hip_last_error": (C.c_char_p, [C.c_void_p]),
    "tit_hip_destroy": (None, [C.c_void_p]),
    "tit_hip_learned_init": (C.c_int, [C.c_void_p, C.c_uint64]),
    "tit_hip_learned_params": (C.c_int, [C.c_void_p, _FP, C.c_size_t,
                                         C.POINTER(C.c_size_t)]),
    "tit_hip_learned_set_params": (C.c_int, [C.c_void_p, _FP, C.c_size_t]),
    "tit_hip_learned_scales": (C.c_int, [C.c_void_p, _DP]),
    "tit_hip_learned_set_scales": (C.c_int, [C.c_void_p, _DP]),
    "tit_hip_learned_record": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int)]),
    "tit_hip_learned_clear_samples": (C.c_int, [C.c_void_p]),
    "tit_hip_learned_sample": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_size_t),
                                         C.POINTER(C.c_uint32), _FP, _FP, _FP,
                                         C.POINTER(C.c_uint8)]),
    "tit_hip_learned_loss_grad": (C.c_int, [C.c_void_p, C.c_int, _DP, _FP]),
    "tit_hip_learned_train": (C.c_int, [C.c_void_p, C.c_int, C.c_double, _DP]),
    "tit_hip_learned_graph": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), _FP, _FP]),
    "tit_hip_learned_forward": (C.c_int, [C.c_void_p, _FP]),
    "tit_hip_learned_set_fused": (C.c_int, [C.c_void_p, C.c_int]),
    "tit_hip_learned_run": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_double,
                                      C.c_double, C.POINTER(LearnedReport)]),
    "tit_hip_nn_gemm": (C.c_int, [C.c_int, C.c_int, C.c_int, _FP, C.c_int, _FP,
                                  C.c_int, _FP, C.c_int, C.c_int, _FP]),
    "tit_dam_break_count": (C.c_int, [C.c_int, C.c_double, _DP, _DP, C.c_int,
                                      C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "tit_dam_break_fill": (C.c_int, [C.c_int, C.c_double, _DP, _DP, C.c_int,
                                     C.c_double, C.c_double, C.c_double,
                                     _DP, _DP, _DP]),
}

_lib = None


def load() -> C.CDLL:
    """Loads libtit_hip.so from the package directory (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is not built; run __graft_entry__.build() or "
            "`make -C titsolver_amd/csrc`")
    lib = C.CDLL(os.fspath(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib
