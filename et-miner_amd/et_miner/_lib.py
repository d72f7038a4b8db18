"""ctypes binding of libetm.so (the C ABI declared in include/etm.h).

``lib()`` loads the in-tree library once and declares every entry point's
signature; ``call(name, *args)`` invokes one and maps a nonzero status to the
exception the reference raises for the same condition (ValueError for bad
arguments and sizes, MemoryError for allocation failures, RuntimeError for
HIP errors). There is no fallback: without the library every GPU entry point
raises.
"""

from __future__ import annotations

import ctypes
import os
import threading
from ctypes import POINTER, c_char_p, c_int, c_int32, c_int64, c_size_t, c_void_p
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libetm.so"
ABI_VERSION = 10

ETM_OK, ETM_EINVAL, ETM_ENOMEM, ETM_EHIP, ETM_EOVERFLOW, ETM_ENODEV, ETM_ECANCELED = 0, 1, 2, 3, 4, 5, 6
CAND_COUNT, CAND_SKIP, CAND_INFER = 0, 1, 2
SUBSET_PRUNE, SUBSET_INFER = 1, 2
DEDUCTION_AUTO, DEDUCTION_OFF = 0, 1
ROW_ORDERS = {"auto": 0, "keep": 1, "lex": 2}
#: etm_mine_opts.k3_schedule (ETM_SCHEDULE_*).
K3_SCHEDULES = {"auto": 0, "all": 1, "live": 2}
#: Accepted etm_mine_opts.small_max values (0 = the library's default, 8).
SMALL_MAXES = (0, 4, 8, 16)
#: Words per tile of the occupancy bitmap (ETM_OCC_TILE_WORDS).
OCC_TILE_WORDS = 64
#: Largest candidate the deduction bounds settle (etm_level_deduce).
BOUNDS_MAX_K = 24


def deduction_setting() -> int:
    """DEDUCTION_AUTO unless ``ET_MINER_DEDUCTION`` is "0" (the reference's switch, ``_env.py:285-291``): the
    bounds run in the inference mode. Any value other than "", "0" or "1" raises."""
    v = os.environ.get("ET_MINER_DEDUCTION", "").strip()
    if v not in ("", "0", "1"):
        raise ValueError(f"ET_MINER_DEDUCTION must be 0 or 1, got {v!r}")
    return DEDUCTION_OFF if v == "0" else DEDUCTION_AUTO
LAYOUT_LEX, LAYOUT_TRI = 0, 1
#: etm_mine_opts.emit_levels (ETM_EMIT_*).
EMIT_NONE, EMIT_HOST, EMIT_DEVICE = 0, 1, 2
KERNEL_AUTO, KERNEL_VALU, KERNEL_MFMA, KERNEL_ROWS = 0, 1, 2, 3
K2_KERNELS = {"auto": KERNEL_AUTO, "valu": KERNEL_VALU, "mfma": KERNEL_MFMA, "rows": KERNEL_ROWS}
K3_KERNELS = {"auto": KERNEL_AUTO, "valu": KERNEL_VALU, "rows": KERNEL_ROWS}
KERNEL_NAMES = {KERNEL_AUTO: "", KERNEL_VALU: "valu", KERNEL_MFMA: "mfma", KERNEL_ROWS: "rows"}

P = c_void_p  # every device pointer crosses the ABI as void*

#: name -> (restype, argtypes), mirroring include/etm.h.
SIGNATURES: dict = {
    "etm_abi_version": (c_int, []),
    "etm_last_error": (c_int, [c_char_p, c_size_t]),
    "etm_device_count": (c_int, [POINTER(c_int)]),
    "etm_upload": (c_int, [P, c_int64, P, P]),
    "etm_download": (c_int, [P, c_int64, P, P]),
    "etm_support": (c_int, [P, c_int64, c_int64, P, P]),
    "etm_ids_minmax": (c_int, [P, c_int, c_int64, P, P]),
    "etm_ids_hist": (c_int, [P, c_int, c_int64, c_int64, c_int64, P, P]),
    "etm_bitvec_build": (c_int, [P, P, c_int, c_int64, c_int64, c_int64, P, c_int32, c_int64, P, P, P]),
    "etm_row_order_ws": (c_size_t, [c_int64]),
    "etm_row_order": (c_int, [P, P, c_int, c_int64, c_int64, c_int64, P, P, c_int32, P, P, c_size_t, P]),
    "etm_tile_occupancy": (c_int, [P, c_int64, c_int64, c_int32, P, c_int64, P]),
    "etm_column_popcounts": (c_int, [P, c_int32, c_int64, c_int64, P, P]),
    "etm_level_groups_ws": (c_size_t, [c_int64]),
    "etm_level_groups": (c_int, [P, c_int64, c_int32, c_int32, P, P, P, P, P, P, c_size_t, P]),
    "etm_level_hash": (c_int, [P, c_int64, c_int32, P, c_int64, P]),
    "etm_level_classify": (c_int, [P, c_int64, c_int32, P, P, P, c_int64, P, P, c_int64, c_int64, P, P, c_int32, P,
                                   P, P, P, P]),
    "etm_level_deduce": (c_int, [P, c_int32, P, P, c_int64, P, P, P, c_int64, P, P, P, c_int64, c_int64, c_int64,
                                 P, P, c_int64, c_int64, P, P, P, P, P]),
    "etm_count_level": (c_int, [P, c_int64, c_int64, P, c_int32, P, P, P, c_int64, c_int64, P, P, P, P, c_int32,
                                c_int32, c_int32, P, c_int64, P, c_size_t, P]),
    "etm_count_level_ws": (c_size_t, [c_int64, c_int64]),
    "etm_level_compact_ws": (c_size_t, [c_int64]),
    "etm_level_compact": (c_int, [P, c_int32, P, P, c_int64, c_int64, P, P, P, c_int64, P, P, P, P, P, c_size_t, P]),
    "etm_level_compact_noalloc_max": (c_int64, []),
    "etm_level_compact_noalloc": (c_int, [P, c_int32, P, P, c_int64, c_int64, P, P, P, c_int64, P, P, P, c_int64, P,
                                          P]),
    "etm_level_free": (c_int, [P, c_int64, c_int32, P, P, P, P, c_int64, P, P]),
    "etm_select_rows": (c_int, [P, c_int64, c_int32, P, P, P, P, P, P, c_size_t, P]),
    "etm_count_itemsets": (c_int, [P, c_int64, c_int64, P, P, c_int64, P, c_int32, P]),
    "etm_rows_ws": (c_size_t, [c_int64]),
    "etm_rows_build": (c_int, [P, P, c_int, c_int64, c_int64, c_int64, P, P, P, P, P, P, c_size_t, P]),
    "etm_count_k2_rows": (c_int, [P, P, c_int64, c_int64, P, P]),
    "etm_k3_rows_prepare": (c_int, [P, c_int64, P, c_int64, P, c_int64, P, P, P, P]),
    "etm_count_k3_rows": (c_int, [P, P, c_int64, P, P, P, c_int64, P, P, P, P, P]),
    "etm_union_insert": (c_int, [P, P, c_int64, c_int32, P, c_int64, P, c_int64, P, P, P, P, c_int64, P, P]),
    "etm_mine_bitvecs": (c_int, [P, c_int64, c_int64, c_int64, P, c_int32, P, P, P]),
    "etm_mine_csr": (c_int, [P, P, c_int, c_int64, c_int64, c_int64, P, P, P]),
    "etm_mine_batch": (c_int, [P, c_int32, c_int64, P, P, P]),
    "etm_levels_count": (c_int, [P, P]),
    "etm_levels_shape": (c_int, [P, c_int32, P, P]),
    "etm_levels_copy": (c_int, [P, c_int32, P, P]),
    "etm_levels_columns": (c_int, [P, P, P, P]),
    "etm_levels_stats": (c_int, [P, c_int32, P]),
    "etm_levels_free": (c_int, [P]),
    "etm_levels_flat": (c_int, [P, P, P, P, P]),
    "etm_levels_frame": (c_int, [P, P, P, P, P, P, c_int64]),
    "etm_levels_frame_view": (c_int, [P, P, P, c_int32]),
}


class LevelStat(ctypes.Structure):
    """etm_level_stat."""

    _fields_ = [("k", c_int32), ("kernel", c_int32), ("n_candidates", c_int64), ("n_counted", c_int64),
                ("n_inferred", c_int64), ("n_skipped", c_int64), ("n_frequent", c_int64), ("bytes_alg", c_int64),
                ("bytes_live", c_int64), ("count_ms", ctypes.c_float), ("wall_ms", ctypes.c_float)]


class LevelInfo(ctypes.Structure):
    """etm_level_info: what the level callback receives as each level completes."""

    _fields_ = [("stat", LevelStat), ("h_rows", POINTER(c_int32)), ("h_counts", POINTER(c_int32)),
                ("h_col_to_item", POINTER(c_int64)), ("n_cols", c_int64), ("device_bytes", c_int64), ("host_bytes", c_int64), ("next_level", c_int32),
                ("reserved", c_int32), ("d_rows", c_void_p), ("d_counts", c_void_p)]


#: etm_level_cb: int (*)(const etm_level_info*, void*)
LEVEL_CB = ctypes.CFUNCTYPE(c_int, POINTER(LevelInfo), c_void_p)


class BatchJob(ctypes.Structure):
    """etm_batch_job: one job of etm_mine_batch."""

    _fields_ = [("d_indptr", c_void_p), ("d_ids", c_void_p), ("n_rows", c_int64), ("nnz", c_int64),
                ("min_count", c_int64), ("id_bytes", c_int32), ("max_length", c_int32)]


class MineOpts(ctypes.Structure):
    """etm_mine_opts."""

    _fields_ = [("min_count", c_int64), ("max_length", c_int32), ("prune_non_free", c_int32),
                ("infer_counts", c_int32), ("time_kernels", c_int32), ("k2_kernel", c_int32),
                ("k3_kernel", c_int32), ("deduction", c_int32), ("row_order", c_int32), ("live_bytes", c_int32),
                ("emit_levels", c_int32), ("k3_schedule", c_int32), ("small_max", c_int32), ("level_cb", LEVEL_CB),
                ("cb_user", c_void_p)]

_lock = threading.Lock()
_lib = None


class Cancelled(RuntimeError):
    """A level callback stopped a mining call (the callback's own exception is re-raised by the caller)."""


class LibraryMissing(RuntimeError):
    """libetm.so is not built: the GPU path cannot run (there is no fallback)."""


def lib() -> ctypes.CDLL:
    """The loaded library with every signature declared (loaded once per process)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise LibraryMissing(
                    f"{LIB_PATH} is not built: run `python __graft_entry__.py build` (or et_miner._build.build()) "
                    "first; the GPU counting path has no CPU fallback"
                )
            handle = ctypes.CDLL(str(LIB_PATH))
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(handle, name)
                fn.restype = res
                fn.argtypes = args
            if handle.etm_abi_version() != ABI_VERSION:
                raise LibraryMissing(f"{LIB_PATH} has ABI {handle.etm_abi_version()}, expected {ABI_VERSION}")
            _lib = handle
    return _lib


def last_error() -> str:
    buf = ctypes.create_string_buffer(1024)
    lib().etm_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def call(name: str, *args) -> int:
    """Calls an entry point; raises on a nonzero status."""
    rc = getattr(lib(), name)(*args)
    if rc == ETM_OK:
        return rc
    msg = f"{name}: {last_error()}"
    if rc == ETM_ECANCELED:
        raise Cancelled(msg)
    if rc in (ETM_EINVAL, ETM_EOVERFLOW):
        raise ValueError(msg)
    if rc == ETM_ENOMEM:
        raise MemoryError(msg)
    raise RuntimeError(msg)


def ws_bytes(name: str, n: int) -> int:
    return int(getattr(lib(), name)(int(n)))


def device_count() -> int:
    n = c_int(0)
    call("etm_device_count", ctypes.byref(n))
    return n.value
