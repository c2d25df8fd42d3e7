"""ctypes binding of liteqg_amd/liblqg.so (the C ABI in include/liteqg.h).

The library is the only compute path: if it is missing this module raises,
there is no fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# LQG_LIB names another build of the same sources (the sanitizer build of
# `python -m liteqg_amd.build --sanitize`, liteqg_amd/_san/liblqg.so)
LIB_PATH = os.environ.get("LQG_LIB") or os.path.join(HERE, "liblqg.so")

LQG_OK = 0
ORDER = {"reference": 0, "nearest": 1}
MODE = {"sq8": 0, "fastscan": 1}
ROTATION = {"fht": 0, "gemm": 1}
BUILD_CANDIDATES = {"sq8": 0, "fastscan": 1}


class LqgError(RuntimeError):
    """A failure reported by the C ABI (status code + message)."""

    def __init__(self, status: int, msg: str):
        super().__init__(f"[lqg status {status}] {msg}")
        self.status = status


class Launch(C.Structure):
    _fields_ = [("waves_per_group", C.c_int32), ("cache_bits", C.c_int32),
                ("occupancy", C.c_int32), ("groups", C.c_int32), ("prefetch", C.c_int32),
                ("split", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("queries", C.c_uint64), ("expansions", C.c_uint64),
                ("evaluations", C.c_uint64), ("device_ms", C.c_double)]


class ServeConfig(C.Structure):
    _fields_ = [("lanes", C.c_uint32), ("max_batch", C.c_uint32), ("slots", C.c_uint32),
                ("ef_max", C.c_uint32), ("kmax", C.c_uint32), ("order", C.c_int)]


class ServeStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("batches", C.c_uint64),
                ("max_lanes_busy", C.c_uint32), ("max_batch_seen", C.c_uint32)]


LQG_ERR_BUSY = 7

_lib = None

# name -> (restype, argtypes); every function returns lqg_status (int)
_P = C.c_void_p
_SIGS = {
    "lqg_create": [C.c_char_p, C.c_char_p, C.c_uint64, C.c_uint64, C.c_uint64, C.POINTER(_P)],
    "lqg_destroy": [_P],
    "lqg_load": [_P, C.c_char_p],
    "lqg_save": [_P, C.c_char_p],
    "lqg_set_ef": [_P, C.c_uint32],
    "lqg_set_order": [_P, C.c_int],
    "lqg_build_index": [_P, _P, C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32],
    "lqg_set_build_candidates": [_P, C.c_int],
    "lqg_search": [_P, _P, C.c_uint32, _P],
    "lqg_search_batch": [_P, _P, C.c_uint64, _P, _P, C.c_uint32, C.c_int, _P, _P, _P],
    "lqg_search_batch_device": [_P, _P, C.c_uint64, _P, _P, C.c_uint32, C.c_uint32, C.c_int, _P,
                                _P, _P, _P],
    "lqg_search_batch_device_noalloc": [_P, _P, C.c_uint64, _P, _P, C.c_uint32, C.c_uint32,
                                        C.c_int, _P, _P, _P, _P],
    "lqg_set_launch": [_P, C.c_uint32, _P],
    "lqg_get_launch": [_P, C.c_uint32, _P],
    "lqg_autotune": [_P, _P, C.c_uint64, C.c_uint32, C.c_uint32, _P],
    "lqg_count_visits": [_P, _P, C.c_uint64, C.c_uint32, C.c_uint32, _P],
    "lqg_info": [_P, _P, _P, _P, _P, _P],
    "lqg_get_graph": [_P, _P],
    "lqg_set_sq8_integer": [_P, C.c_int],
    "lqg_set_mode": [_P, C.c_int],
    "lqg_get_mode": [_P, _P],
    "lqg_h2_build_sidecar": [_P, _P, C.c_uint64],
    "lqg_h2_set_sidecar": [_P, _P, _P, C.c_int, _P],
    "lqg_h2_get_sidecar": [_P, _P, _P, _P],
    "lqg_h2_save_sidecar": [_P, C.c_char_p],
    "lqg_h2_load_sidecar": [_P, C.c_char_p],
    "lqg_h2_set_rotation": [_P, C.c_int],
    "lqg_h2_get_rotation": [_P, _P],
    "lqg_h2_rotate": [_P, _P, C.c_uint64, C.c_int, _P],
    "lqg_h2_probe": [_P, _P, C.c_uint64, _P, _P, _P, _P, _P, _P],
    "lqg_h2_expansions": [_P, _P, C.c_uint64, C.c_uint32, C.c_uint32, _P],
    "lqg_h2_set_overflow_bits": [_P, C.c_uint32],
    "lqg_h2_set_generic": [_P, C.c_int],
    "lqg_graph_create": [_P, _P, C.c_uint64, _P, _P, C.c_uint32, C.c_uint32, C.c_int, _P, _P, _P,
                         C.POINTER(_P)],
    "lqg_graph_launch": [_P, _P],
    "lqg_graph_destroy": [_P],
    "lqg_serve_start": [_P, _P, C.POINTER(_P)],
    "lqg_serve_submit": [_P, _P, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint64)],
    "lqg_serve_poll": [_P, C.c_uint64, _P, _P, C.c_uint32],
    "lqg_serve_wait": [_P, C.c_uint64, _P, _P, C.c_uint32],
    "lqg_serve_get_stats": [_P, _P],
    "lqg_serve_stop": [_P],
    "lqg_serve_replay": [_P, _P, C.c_uint64, _P, _P, _P, _P, _P],
}
EXPORTED = tuple(_SIGS) + ("lqg_last_error",)


def lib():
    """Load liblqg.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m liteqg_amd.build`")
        # One HIP runtime serves the whole process: the first libamdhip64 loaded.
        # PyTorch-ROCm ships its own and cannot initialise its devices on another
        # one loaded before it, while this library runs on either.  So when torch
        # is installed it is loaded first, and device tensors can be shared.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_int
        L.lqg_last_error.argtypes = []
        L.lqg_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def check(status: int) -> None:
    if status != LQG_OK:
        msg = lib().lqg_last_error()
        raise LqgError(status, msg.decode() if msg else "")


def ptr(a):
    return None if a is None else C.c_void_p(a.ctypes.data)
