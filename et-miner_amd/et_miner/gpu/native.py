"""The native level loop (``etm_mine_csr`` / ``etm_mine_bitvecs`` in libetm.so), bound for Python.

One C call mines a whole lattice on the device; the result comes back as host arrays per level with
per-level statistics (candidates counted / inferred / skipped, compulsory bytes, event-timed counting).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from .. import _lib
from . import device as dev
from .row_split import LevelStats

#: Largest id span counted with a direct histogram on the device.
SPAN_LIMIT = 1 << 26


@dataclass
class NativeResult:
    levels: list = field(default_factory=list)  # (rows int32 (n, k) of column ids, counts int64)
    stats: list = field(default_factory=list)  # LevelStats per level
    col_to_item: np.ndarray | None = None
    col_counts: np.ndarray | None = None
    flat: tuple | None = None  # (item ids int64, offsets int64, counts int64) over every level

    def table(self, n_rows: int, with_counts: bool = False):
        """The result as a pyarrow Table (itemset, support[, count]) built from the flat arrays."""
        import pyarrow as pa

        values, offsets, counts = self.flat
        cols = {"itemset": pa.LargeListArray.from_arrays(pa.array(offsets), pa.array(values)),
                "support": pa.array(counts / float(n_rows))}
        if with_counts:
            cols["count"] = pa.array(counts)
        return pa.table(cols)


def _opts(min_count, max_length, prune, infer, time_kernels, k2_kernel="auto", k3_kernel="auto",
          row_order="auto") -> _lib.MineOpts:
    if k2_kernel not in _lib.K2_KERNELS:
        raise ValueError(f"k2_kernel must be one of {sorted(_lib.K2_KERNELS)}, got {k2_kernel!r}")
    if k3_kernel not in _lib.K3_KERNELS:
        raise ValueError(f"k3_kernel must be one of {sorted(_lib.K3_KERNELS)}, got {k3_kernel!r}")
    if row_order not in _lib.ROW_ORDERS:
        raise ValueError(f"row_order must be one of {sorted(_lib.ROW_ORDERS)}, got {row_order!r}")
    return _lib.MineOpts(int(max(1, min_count)), int(max_length or 0), int(bool(prune)), int(bool(infer)),
                         int(bool(time_kernels)), _lib.K2_KERNELS[k2_kernel], _lib.K3_KERNELS[k3_kernel],
                         _lib.deduction_setting(), _lib.ROW_ORDERS[row_order])


def _collect(handle, levels: bool = True) -> NativeResult:
    """Reads a result handle (and frees it): the flat arrays always, the per-level arrays when ``levels``."""
    L = _lib.lib()
    res = NativeResult()
    try:
        tot = np.zeros(2, dtype=np.int64)
        _lib.call("etm_levels_flat", handle, tot.ctypes.data, None, None, None)
        values = np.empty(tot[1], dtype=np.int64)
        offsets = np.empty(tot[0] + 1, dtype=np.int64)
        counts = np.empty(tot[0], dtype=np.int64)
        _lib.call("etm_levels_flat", handle, tot.ctypes.data, values.ctypes.data, offsets.ctypes.data,
                  counts.ctypes.data)
        res.flat = (values, offsets, counts)
        n_levels = ctypes.c_int32()
        _lib.call("etm_levels_count", handle, ctypes.byref(n_levels))
        for lv in range(n_levels.value):
            k, n = ctypes.c_int32(), ctypes.c_int64()
            _lib.call("etm_levels_shape", handle, lv, ctypes.byref(k), ctypes.byref(n))
            if levels:
                rows = np.empty((n.value, k.value), dtype=np.int32)
                lv_counts = np.empty(n.value, dtype=np.int64)
                _lib.call("etm_levels_copy", handle, lv, rows.ctypes.data, lv_counts.ctypes.data)
                res.levels.append((rows, lv_counts))
            st = _lib.LevelStat()
            _lib.call("etm_levels_stats", handle, lv, ctypes.byref(st))
            res.stats.append(LevelStats(st.k, st.n_candidates, st.n_frequent, st.n_counted, st.n_inferred,
                                        st.n_skipped, st.bytes_alg, 0.0, st.count_ms,
                                        _lib.KERNEL_NAMES.get(st.kernel, "")))
        n_cols = ctypes.c_int32()
        _lib.call("etm_levels_columns", handle, ctypes.byref(n_cols), None, None)
        c2i = np.empty(n_cols.value, dtype=np.int64)
        cnt = np.empty(n_cols.value, dtype=np.int64)
        _lib.call("etm_levels_columns", handle, ctypes.byref(n_cols), c2i.ctypes.data, cnt.ctypes.data)
        res.col_to_item, res.col_counts = c2i, cnt
    finally:
        L.etm_levels_free(handle)
    return res


def mine_csr_device(indptr_d, ids_d, n_rows: int, min_count: int, max_length=None, prune_non_free=False,
                    infer_counts=False, time_kernels=False, levels: bool = True,
                    k2_kernel: str = "auto", k3_kernel: str = "auto", row_order: str = "auto") -> NativeResult:
    """Every frequent itemset of a device CSR of item ids (int32 or int64), mined by the native loop.

    ``k2_kernel``: "auto", "valu", "mfma" or "rows"; ``k3_kernel``: "auto", "valu" (bitvector kernels) or
    "rows" (row-wise, which needs rows of strictly ascending distinct ids); ``row_order``: "lex" (rows sorted by
    their most frequent items, with the tile skip), "keep", or "auto" (keeps the rows)."""
    torch = dev.torch()
    id_bytes = 8 if ids_d.dtype == torch.int64 else 4
    if ids_d.dtype not in (torch.int64, torch.int32):
        raise ValueError(f"item ids must be int32 or int64, got {ids_d.dtype}")
    handle = ctypes.c_void_p()
    opts = _opts(min_count, max_length, prune_non_free, infer_counts, time_kernels, k2_kernel, k3_kernel, row_order)
    _lib.call("etm_mine_csr", indptr_d.data_ptr(), ids_d.data_ptr() if ids_d.numel() else None, id_bytes, n_rows,
              ids_d.numel(), SPAN_LIMIT, ctypes.byref(opts), ctypes.byref(handle), dev.stream_ptr())
    return _collect(handle, levels)


def mine_bitvecs_device(bv_ptr: int, pitch: int, n_words: int, n_rows: int, col_counts: np.ndarray, min_count: int,
                        max_length=None, prune_non_free=False, infer_counts=False,
                        time_kernels=False, k2_kernel: str = "auto") -> NativeResult:
    """Every frequent itemset of device bitvectors (read only), from the columns' K=1 counts."""
    counts = np.ascontiguousarray(col_counts, dtype=np.int64)
    handle = ctypes.c_void_p()
    opts = _opts(min_count, max_length, prune_non_free, infer_counts, time_kernels, k2_kernel)
    _lib.call("etm_mine_bitvecs", bv_ptr, pitch, n_words, n_rows, counts.ctypes.data, len(counts),
              ctypes.byref(opts), ctypes.byref(handle), dev.stream_ptr())
    return _collect(handle)
