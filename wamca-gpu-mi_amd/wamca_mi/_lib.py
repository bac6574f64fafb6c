"""Loader for the in-tree HIP library ``wamca-gpu-mi_amd/lib/libwamca_mi.so``.

The library is the product: there is no CPU fallback.  If it is missing the
import fails loudly with the command that builds it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libwamca_mi.so")

I8P = np.ctypeslib.ndpointer(np.int8, flags="C_CONTIGUOUS")
I32P = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
I64P = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
U32P = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
U64P = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")


class MLPError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"status {status}: {msg}")
        self.status = status


class MoveRec(C.Structure):
    _fields_ = [("cost", C.c_int64), ("id", C.c_int32), ("i", C.c_int32), ("j", C.c_int32),
                ("row", C.c_int32)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_int64), ("max_active", C.c_int32), ("rowmap", C.c_int32),
                ("rounds_per_step", C.c_int32), ("max_rounds", C.c_int32), ("big_n", C.c_int32),
                ("max_bypass", C.c_int32)]


class JobResult(C.Structure):
    _fields_ = [("job", C.c_int64), ("tag", C.c_int64), ("latency", C.c_int64),
                ("rounds", C.c_int32), ("n", C.c_int32), ("t_submit", C.c_double),
                ("t_admit", C.c_double), ("t_finish", C.c_double),
                ("tour", C.POINTER(C.c_int32)), ("status", C.c_int32)]


_lib = None

# name -> (restype, argtypes)
_SIGS = {
    "mlp_ctx_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "mlp_ctx_destroy": (C.c_int, [C.c_void_p]),
    "mlp_ctx_error": (C.c_char_p, [C.c_void_p]),
    "mlp_strerror": (C.c_char_p, [C.c_int]),
    "mlp_sync": (C.c_int, [C.c_void_p]),
    "mlp_query": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32)]),
    "mlp_ctx_device": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32)]),
    "mlp_device_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int64),
                                  C.c_char_p, C.c_int32]),
    "mlp_instance_create": (C.c_int, [C.c_void_p, I32P, I32P, C.c_int32, C.c_int32,
                                      C.POINTER(C.c_int32)]),
    "mlp_instance_create_async": (C.c_int, [C.c_void_p, I32P, I32P, C.c_int32, C.c_int32,
                                            C.POINTER(C.c_int32)]),
    "mlp_instance_destroy": (C.c_int, [C.c_void_p, C.c_int32]),
    "mlp_check_job": (C.c_int, [C.c_void_p, I32P, I32P, C.c_int32, C.c_int32, I32P]),
    "mlp_mem_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "mlp_launch_probe": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
    "mlp_dvnd_fused": (C.c_int, [C.c_void_p, C.c_int32]),
    "mlp_dvnd_groups": (C.c_int, [C.c_void_p, C.c_int32]),
    "mlp_dev_alloc": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(C.c_void_p)]),
    "mlp_dev_free": (C.c_int, [C.c_void_p, C.c_void_p]),
    "mlp_copy_to_host": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]),
    "mlp_copy_to_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]),
    "mlp_eval_static": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_uint32, C.c_int32,
                                  C.c_void_p, C.c_void_p, C.c_void_p]),
    "mlp_pool_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "mlp_server_create": (C.c_int, [C.c_void_p, C.POINTER(ServerConfig), C.POINTER(C.c_void_p)]),
    "mlp_server_destroy": (C.c_int, [C.c_void_p]),
    "mlp_server_submit": (C.c_int, [C.c_void_p, I32P, I32P, C.c_int32, C.c_int32, I32P,
                                    C.c_int64, C.POINTER(C.c_int64)]),
    "mlp_server_step": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "mlp_server_poll": (C.c_int, [C.c_void_p, C.POINTER(JobResult), C.c_int32,
                                  C.POINTER(C.c_int32)]),
    "mlp_server_job_bytes": (C.c_int64, [C.c_int32]),
    "mlp_server_now": (C.c_double, [C.c_void_p]),
    "mlp_instance_weights": (C.c_int, [C.c_void_p, C.c_int32, I32P]),
    "mlp_slot_footprint": (C.c_int64, [C.c_int32]),
    "mlp_slot_create": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_int32)]),
    "mlp_slot_destroy": (C.c_int, [C.c_void_p, C.c_int32]),
    "mlp_set_tours": (C.c_int, [C.c_void_p, I32P, C.c_int32, I32P]),
    "mlp_get_tours": (C.c_int, [C.c_void_p, I32P, C.c_int32, I32P]),
    "mlp_get_latency": (C.c_int, [C.c_void_p, I32P, C.c_int32, I64P]),
    "mlp_eval": (C.c_int, [C.c_void_p, I32P, C.c_int32, C.c_uint32, C.c_int32, C.c_int32]),
    "mlp_get_rows": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, I64P, I32P, I32P,
                               C.POINTER(C.c_int32)]),
    "mlp_get_moves64": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, U64P,
                                  C.POINTER(C.c_int32)]),
    "mlp_merge_apply": (C.c_int, [C.c_void_p, I32P, C.c_int32, C.c_int32]),
    "mlp_eval_total": (C.c_int, [C.c_void_p, I32P, C.c_int32, C.c_int32, C.c_int32]),
    "mlp_get_total": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                C.c_int64, C.POINTER(C.c_int64)]),
    "mlp_merge_total": (C.c_int, [C.c_void_p, I32P, C.c_int32]),
    "mlp_get_applied": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(MoveRec),
                                  C.c_int32, C.POINTER(C.c_int32)]),
    "mlp_dvnd": (C.c_int, [C.c_void_p, I32P, C.c_int32, C.c_int32, C.c_int32, I32P]),
    "mlp_dvnd_start": (C.c_int, [C.c_void_p, I32P, C.c_int32]),
    "mlp_dvnd_rounds": (C.c_int, [C.c_void_p, I32P, C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
    "mlp_dvnd_status": (C.c_int, [C.c_void_p, I32P, C.c_int32, I32P, I32P]),
    "mlp_dvnd_config": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32]),
    "mlp_ads_materialize": (C.c_int, [C.c_void_p, C.c_int32, U32P, C.c_int64]),
    "mlp_dist_table": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, I32P]),
    "mlp_dist_list": (C.c_int, [C.c_void_p, I32P, C.c_int32, C.c_int32, I32P]),
    "mlp_region_begin": (C.c_int, [C.c_void_p]),
    "mlp_region_end": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "mlp_eval_each": (C.c_int, [C.c_void_p, I32P, C.c_int32, C.c_uint32, C.c_int32]),
    "mlp_copy_probe": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32]),
    "mlp_sqrt_selfcheck": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    "mlp_set_tile": (C.c_int, [C.c_void_p, C.c_int32]),
    "mlp_merge_force_global": (C.c_int, [C.c_void_p, C.c_int32]),
    "mlp_set_neighborhoods": (C.c_int, [C.c_void_p, C.c_int32]),
    "mlp_dvnd_neighborhoods": (C.c_int, [C.c_void_p, C.c_int32]),
    "mlp_slot_bytes": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_int64)]),
    "mlp_graph_begin": (C.c_int, [C.c_void_p]),
    "mlp_graph_end": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32)]),
    "mlp_graph_launch": (C.c_int, [C.c_void_p, C.c_int32]),
    "mlp_graph_destroy": (C.c_int, [C.c_void_p, C.c_int32]),
    "mlp_autotune": (C.c_int, [C.c_void_p, I32P, C.c_int32, C.c_uint32, C.c_int32,
                               C.POINTER(C.c_int32), C.POINTER(C.c_float)]),
    "mlp_tune_get": (C.c_int, [C.c_void_p, I8P]),
    "mlp_tune_set": (C.c_int, [C.c_void_p, I8P]),
    "mlp_timing_begin": (C.c_int, [C.c_void_p, C.c_int32]),
    "mlp_timing_end": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int32)]),
    # host-side inputs (include/mlp_host.h)
    "mlp_mt_first": (C.c_uint32, [C.c_uint32, C.c_int32]),
    "mlp_gen_uniform": (C.c_int, [C.c_int32, C.c_uint32, C.c_int32, I32P, I32P]),
    "mlp_gen_clustered": (C.c_int, [C.c_int32, C.c_uint32, C.c_int32, C.c_int32, C.c_double,
                                    I32P, I32P]),
    "mlp_gen_loguniform_n": (C.c_int32, [C.c_uint32, C.c_int32, C.c_int32, C.c_int32]),
    "mlp_tsplib_load": (C.c_int, [C.c_char_p, I32P, I32P, C.c_int32, C.POINTER(C.c_int32)]),
    "mlp_shift_nonnegative": (C.c_int, [C.c_int32, I32P, I32P]),
    "mlp_tour_perm": (C.c_int, [C.c_uint32, C.c_int32, I32P]),
    "mlp_tour_random": (C.c_int, [I32P, I32P, C.c_int32, C.c_int32, C.c_uint32, C.c_int32,
                                  C.c_float, I32P]),
    "mlp_tour_random_fast": (C.c_int, [I32P, I32P, C.c_int32, C.c_int32, C.c_uint32, C.c_float,
                                       I32P]),
}


def lib():
    """Load the library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with "
                "`make -C wamca-gpu-mi_amd/csrc` (or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    """Every name the C ABI headers declare, for the export check."""
    return sorted(_SIGS)
