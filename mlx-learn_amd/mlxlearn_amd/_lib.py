"""ctypes binding of ``libmlxlearn_amd.so`` (the C ABI of ``include/mlxlearn_amd.h``).

The library is loaded from the package's ``_lib`` directory only. There is no
CPU fallback: without the library, or without a HIP device, every compute
call raises :class:`~mlxlearn_amd.exceptions.DeviceError`.
"""
from __future__ import annotations

import ctypes
import os

from .exceptions import BackendExecutionError, DeviceError

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "libmlxlearn_amd.so")

# flags (include/mlxlearn_amd.h)
EXCLUDE_SELF = 0x1
PREC_FP32 = 0x00
PREC_BF16 = 0x10
PREC_FP16 = 0x20
SORT_RESULTS = 0x100
NEED_DISTANCE = 0x200
SQUARED_DIST = 0x400
GRAPH_REPLAY = 0x800
WEIGHTS_UNIFORM = 0
WEIGHTS_DISTANCE = 1
HOOK_FORCE_FALLBACK = 0x1
HOOK_CAPACITY = 0x2
HOOK_FAIL = 0x4
HOOK_FORCE_UNCERTIFIED = 0x8
HOOK_SORT_LIMIT = 0x10
HOOK_FORCE_ORDER = 0x20
HOOK_NO_ORDER = 0x40
TUNE_RADIUS = 0x1000
MAX_ROUNDS = 8

PRECISIONS = {"float32": PREC_FP32, "bfloat16": PREC_BF16, "float16": PREC_FP16}

# Every exported entry point with its ctypes signature.
_vp, _i64, _i32, _u32, _dbl = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_uint32, ctypes.c_double
_i64p = ctypes.POINTER(ctypes.c_int64)


class Stats(ctypes.Structure):
    _fields_ = [("queries", _i64), ("candidates", _i64), ("reranked", _i64), ("fallback_queries", _i64),
                ("rounds", _i64), ("splits", _i64), ("capacity", _i64), ("screen_precision", _i64),
                ("screen_ms", _dbl), ("total_ms", _dbl)]


class Plan(ctypes.Structure):
    _fields_ = [("splits", _i32), ("capacity", _i32), ("query_tiles", _i32), ("stages", _i32), ("n_rounds", _i32),
                ("rounds", _dbl * MAX_ROUNDS), ("tuned", _i32), ("trials", _i32), ("tuned_ms", _dbl),
                ("default_ms", _dbl)]

    def to_dict(self) -> dict:
        return {"splits": self.splits, "capacity": self.capacity, "query_tiles": self.query_tiles,
                "stages": self.stages, "rounds": [self.rounds[i] for i in range(self.n_rounds)],
                "tuned": bool(self.tuned), "trials": self.trials, "tuned_ms": self.tuned_ms,
                "default_ms": self.default_ms}


SIGNATURES = {
    "mlk_last_error": ([], ctypes.c_char_p),
    "mlk_version": ([], ctypes.c_int),
    "mlk_get_stats": ([ctypes.POINTER(Stats)], ctypes.c_int),
    "mlk_set_test_hooks": ([_u32, _i64], ctypes.c_int),
    "mlk_set_forced_rows": ([_i64p, _i64], ctypes.c_int),
    "mlk_set_tuning": ([_i32, _i32, _i32, _i32, _i32, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "mlk_index_create": ([_vp, _i64, _i32, _u32, _vp, ctypes.POINTER(_vp)], ctypes.c_int),
    "mlk_index_destroy": ([_vp], ctypes.c_int),
    "mlk_index_info": ([_vp, _i64p, ctypes.POINTER(_i32), ctypes.POINTER(_i32)], ctypes.c_int),
    "mlk_knn_query": ([_vp, _vp, _i64, _i32, _i64, _u32, _vp, _vp, _vp], ctypes.c_int),
    "mlk_knn_query_noalloc": ([_vp, _vp, _i64, _i32, _i64, _u32, _vp, _vp, _vp], ctypes.c_int),
    "mlk_knn_query_batched": ([_vp, _vp, _i64p, _i32, _i32, _u32, _vp, _vp, _vp], ctypes.c_int),
    "mlk_radius_query": ([_vp, _vp, _i64, _dbl, _i64, _u32, _vp, _i64p, ctypes.POINTER(_vp), _vp], ctypes.c_int),
    "mlk_radius_fill": ([_vp, _vp, _vp, _vp], ctypes.c_int),
    "mlk_radius_free": ([_vp], ctypes.c_int),
    "mlk_radius_votes": ([_vp, _vp, _vp, _i32, _i32, _vp, _vp, _vp], ctypes.c_int),
    "mlk_radius_mean": ([_vp, _vp, _vp, _i32, _vp, _vp, _vp], ctypes.c_int),
    "mlk_kmeans_assign": ([_vp, _i64, _i32, _vp, _i64, _u32, _vp, _vp, _vp], ctypes.c_int),
    "mlk_kmeans_assign_reuse": ([_vp, _i64, _i32, _vp, _i64, _vp, _u32, _vp, _vp, _i64p, _vp], ctypes.c_int),
    "mlk_kmeans_update": ([_vp, _i64, _i32, _vp, _i64, _vp, _vp, _vp, _vp, _vp], ctypes.c_int),
    "mlk_knn_votes": ([_vp, _vp, _i64, _i32, _vp, _i32, _i32, _vp, _vp, _vp], ctypes.c_int),
    "mlk_autotune": ([_vp, _vp, _i64, _i32, _dbl, _u32, _i32, ctypes.POINTER(Plan), _vp], ctypes.c_int),
    "mlk_get_plan": ([_vp, _i64, _i32, _u32, ctypes.POINTER(Plan)], ctypes.c_int),
    "mlk_clear_plans": ([], ctypes.c_int),
    "mlk_export_plans": ([ctypes.c_char_p, _i64, _i64p], ctypes.c_int),
    "mlk_import_plans": ([ctypes.c_char_p, _i64], ctypes.c_int),
    "mlk_mfma_probe": ([_i32, _i32, ctypes.POINTER(_dbl), _vp], ctypes.c_int),
}

_LIB = None


def load(path: str | None = None):
    """Load the library and bind every entry point (raises DeviceError if absent)."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise DeviceError(f"native library {path} is missing; build it with mlx-learn_amd/build_lib.py")
    lib = ctypes.CDLL(path)
    for name, (args, res) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    if path == LIB_PATH:
        _LIB = lib
        from . import _config
        try:
            _config.apply_tuning()
        except Exception:
            _LIB = None  # a rejected MLXLEARN_* value fails every load until it is fixed
            raise
    return lib


def check(rc: int, what: str = "call") -> None:
    """Map a status code to the exception family of the reference surface."""
    if rc == 0:
        return
    msg = (_LIB.mlk_last_error() or b"").decode(errors="replace") if _LIB is not None else ""
    if rc > 0:
        raise ValueError(msg or f"{what}: invalid input")
    if rc == -1:
        raise MemoryError(msg or f"{what}: device allocation failed")
    raise BackendExecutionError(f"{what} failed: {msg}")


def stats() -> dict:
    out = Stats()
    check(load().mlk_get_stats(ctypes.byref(out)), "mlk_get_stats")
    return {name: getattr(out, name) for name, _ in Stats._fields_}


def set_test_hooks(hooks: int, param: int = 0) -> None:
    check(load().mlk_set_test_hooks(int(hooks), int(param)), "mlk_set_test_hooks")


def set_forced_rows(rows) -> None:
    import numpy as np
    arr = np.ascontiguousarray(rows, dtype=np.int64)
    check(load().mlk_set_forced_rows(arr.ctypes.data_as(_i64p), arr.size), "mlk_set_forced_rows")


def set_tuning(splits: int = 0, capacity: int = 0, query_tiles: int = 0, stages: int = 0, rounds=None) -> None:
    rounds = list(rounds or [])
    arr = (ctypes.c_double * max(len(rounds), 1))(*rounds)
    check(load().mlk_set_tuning(int(splits), int(capacity), int(query_tiles), int(stages), len(rounds), arr),
          "mlk_set_tuning")


def clear_plans() -> None:
    """Forget every autotuned plan of this process."""
    check(load().mlk_clear_plans(), "mlk_clear_plans")


def export_plans() -> str:
    """The autotuned plans of this process as text (see import_plans)."""
    lib = load()
    n = ctypes.c_int64()
    check(lib.mlk_export_plans(None, 0, ctypes.byref(n)), "mlk_export_plans")
    buf = ctypes.create_string_buffer(max(int(n.value), 1))
    check(lib.mlk_export_plans(buf, int(n.value), ctypes.byref(n)), "mlk_export_plans")
    return buf.raw[:n.value].decode()


def import_plans(text: str) -> None:
    """Add plans exported by export_plans (in this or another process)."""
    data = text.encode()
    check(load().mlk_import_plans(data, len(data)), "mlk_import_plans")
