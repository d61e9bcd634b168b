"""ctypes declarations of include/uvlm_fmm.h and the loader of libuvlm_fmm.so.

The library is built in-tree (uvlm_amd/lib/libuvlm_fmm.so) by uvlm_amd/csrc/Makefile.
Loading fails loudly when it is missing: the package has no CPU fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libuvlm_fmm.so")

STATUS = {0: "OK", 1: "ERR_ARG", 2: "ERR_DEPTH", 3: "ERR_STATE", 4: "ERR_HIP", 5: "ERR_NOMEM", 6: "ERR_INTERNAL"}

# Every entry point declared in include/uvlm_fmm.h.
DIRECT_NOALLOC_MAX = 16384  # UVLM_DIRECT_NOALLOC_MAX
EXPORTS = [
    "uvlm_last_error", "uvlm_abi_version", "uvlm_create", "uvlm_destroy", "uvlm_set_bodies_aos",
    "uvlm_get_bodies_aos", "uvlm_set_bodies_soa", "uvlm_get_bodies_soa", "uvlm_get_permutation",
    "uvlm_build", "uvlm_eval", "uvlm_stage", "uvlm_advance", "uvlm_num_cells", "uvlm_get_tree",
    "uvlm_get_lists", "uvlm_get_expansions", "uvlm_sensors_aos", "uvlm_direct", "uvlm_probe_adjoint",
    "uvlm_update_bodies_soa", "uvlm_relax", "uvlm_diagnostics", "uvlm_timings", "uvlm_sync", "uvlm_autotune",
    "uvlm_set_viscosity", "uvlm_viscosity", "uvlm_vorticity_aos", "uvlm_rbf_aos", "uvlm_eval_vjp",
    "uvlm_step_vjp", "uvlm_trajectory_forward", "uvlm_trajectory_vjp", "uvlm_append_bodies_soa",
    "uvlm_wing_create", "uvlm_wing_destroy", "uvlm_wing_get", "uvlm_wing_translate", "uvlm_wing_solve",
    "uvlm_wing_shed", "uvlm_wing_step", "uvlm_wing_lift_adjoint", "uvlm_num_bodies", "uvlm_direct_noalloc", "uvlm_get_list",
    "uvlm_wing_window_adjoint", "uvlm_set_near_field",
]


class UVLMError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


class Params(ctypes.Structure):
    _fields_ = [("P", ctypes.c_int), ("theta", ctypes.c_double), ("ncrit", ctypes.c_int64),
                ("precision", ctypes.c_int), ("stretching", ctypes.c_int), ("formulation", ctypes.c_int),
                ("scheme", ctypes.c_int), ("device", ctypes.c_int)]


class Timing(ctypes.Structure):
    _fields_ = [("tree_ms", ctypes.c_double), ("lists_ms", ctypes.c_double), ("upward_ms", ctypes.c_double),
                ("m2l_ms", ctypes.c_double), ("p2p_ms", ctypes.c_double), ("downward_ms", ctypes.c_double),
                ("eval_ms", ctypes.c_double), ("n_bodies", ctypes.c_int64), ("n_cells", ctypes.c_int64),
                ("n_leaves", ctypes.c_int64), ("n_p2p_pairs", ctypes.c_int64), ("n_m2l_pairs", ctypes.c_int64),
                ("n_levels", ctypes.c_int64), ("n_interactions", ctypes.c_int64), ("sensor_ms", ctypes.c_double),
                ("adjoint_ms", ctypes.c_double), ("ncrit", ctypes.c_int64), ("m2l_slots", ctypes.c_int64),
                ("vjp_ms", ctypes.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                              "or `make -C uvlm_amd/csrc`")
        L = ctypes.CDLL(LIB_PATH)
        I, I64, D, VP = ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p
        PD = ctypes.POINTER(ctypes.c_double)
        PI64 = ctypes.POINTER(ctypes.c_int64)
        sig = {
            "uvlm_last_error": (ctypes.c_char_p, []),
            "uvlm_abi_version": (I, []),
            "uvlm_create": (I, [ctypes.POINTER(Params), ctypes.POINTER(VP)]),
            "uvlm_destroy": (I, [VP]),
            "uvlm_set_bodies_aos": (I, [VP, VP, I64, I64]),
            "uvlm_get_bodies_aos": (I, [VP, VP, I64, I64]),
            "uvlm_set_bodies_soa": (I, [VP, I64, PD, PD, PD]),
            "uvlm_get_bodies_soa": (I, [VP, I64, PD, PD, PD, PD, PD, PD, PD]),
            "uvlm_get_permutation": (I, [VP, I64, PI64]),
            "uvlm_build": (I, [VP]),
            "uvlm_eval": (I, [VP]),
            "uvlm_stage": (I, [VP, I, D]),
            "uvlm_advance": (I, [VP, D]),
            "uvlm_num_cells": (I, [VP, PI64]),
            "uvlm_get_tree": (I, [VP, I64, PI64, PD]),
            "uvlm_get_lists": (I, [VP, PI64, PI64, PI64, PI64]),
            "uvlm_get_expansions": (I, [VP, PD, PD]),
            "uvlm_sensors_aos": (I, [VP, VP, I64, I64]),
            "uvlm_direct": (I, [VP, I64, PD, PD, PD, PD, PD]),
            "uvlm_direct_noalloc": (I, [VP, I64, PD, PD, PD, PD, PD]),
            "uvlm_get_list": (I, [I, D, I64, PI64, PD, I64, PI64, PD, I64, I64, PI64, PI64, PI64, PI64]),
            "uvlm_probe_adjoint": (I, [VP, I64, PD, PD, PD, PD]),
            "uvlm_update_bodies_soa": (I, [VP, I64, PD, PD]),
            "uvlm_relax": (I, [VP, D]),
            "uvlm_diagnostics": (I, [VP, PD]),
            "uvlm_timings": (I, [VP, ctypes.POINTER(Timing)]),
            "uvlm_sync": (I, [VP]),
            "uvlm_autotune": (I, [VP, PI64, I, I, PI64, PD]),
            "uvlm_set_viscosity": (I, [VP, I, D]),
            "uvlm_set_near_field": (I, [VP, D]),
            "uvlm_viscosity": (I, [VP, D]),
            "uvlm_vorticity_aos": (I, [VP, VP, I64, I64]),
            "uvlm_rbf_aos": (I, [VP, VP, I64, I64, PI64, ctypes.POINTER(ctypes.c_int), PD]),
            "uvlm_eval_vjp": (I, [VP, I64, PD, PD, PD, PD]),
            "uvlm_step_vjp": (I, [VP, D, I64, PD, PD]),
            "uvlm_trajectory_forward": (I, [VP, I64, D, I64]),
            "uvlm_trajectory_vjp": (I, [VP, I64, PD, PD]),
            "uvlm_append_bodies_soa": (I, [VP, I64, PD, PD, PD]),
            "uvlm_num_bodies": (I, [VP, PI64]),
            "uvlm_wing_create": (I, [I, I64, I64, PD, ctypes.POINTER(VP)]),
            "uvlm_wing_destroy": (I, [VP]),
            "uvlm_wing_get": (I, [VP, PD, PD, PD, PD]),
            "uvlm_wing_translate": (I, [VP, PD]),
            "uvlm_wing_solve": (I, [VP, VP, PD, PD, PD]),
            "uvlm_wing_shed": (I, [VP, VP, PD, D, I]),
            "uvlm_wing_step": (I, [VP, VP, PD, D, I, PD, PD]),
            "uvlm_wing_lift_adjoint": (I, [VP, VP, PD, D, I64, I, PD, PD, I64, PD, PD]),
            "uvlm_wing_window_adjoint": (I, [VP, VP, PD, D, I64, I64, I, PD, PD, I64, PD, PD, PD]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(code):
    if code != 0:
        raise UVLMError(code, lib().uvlm_last_error().decode(errors="replace"))
