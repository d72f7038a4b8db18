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
    support: np.ndarray | None = None  # count / support_rows per itemset (float64), when the export wrote it
    support_rows: int = 0

    def table(self, n_rows: int, with_counts: bool = False):
        """The result as a pyarrow Table (itemset, support[, count]) built from the flat arrays."""
        import pyarrow as pa

        values, offsets, counts = self.flat
        support = self.support if self.support is not None and self.support_rows == n_rows else counts / n_rows
        cols = {"itemset": pa.LargeListArray.from_arrays(pa.array(offsets), pa.array(values)),
                "support": pa.array(support)}
        if with_counts:
            cols["count"] = pa.array(counts)
        return pa.table(cols)


def _stats_of(st) -> LevelStats:
    return LevelStats(st.k, st.n_candidates, st.n_frequent, st.n_counted, st.n_inferred, st.n_skipped, st.bytes_alg,
                      float(st.wall_ms), float(st.count_ms), _lib.KERNEL_NAMES.get(st.kernel, ""), st.bytes_live)


@dataclass
class LevelReport:
    """A level as the native loop hands it over on completion: its statistics, its rows (column ids int32 (n, k))
    and counts (int64) when the levels are emitted, and the memory the call holds."""

    stats: LevelStats
    rows: np.ndarray | None
    counts: np.ndarray | None
    device_bytes: int
    host_bytes: int
    next_level: bool
    col_to_item: np.ndarray | None = None  # the item id of each column (etm_mine_csr)
    d_rows: int | None = None  # with emit_levels="device", levels K >= 2: device pointers, valid during the call
    d_counts: int | None = None


class _Callback:
    """The ctypes level callback around a Python function of a LevelReport. An exception it raises stops the
    mining call (the callback returns nonzero) and is re-raised by ``run``."""

    def __init__(self, fn, emit):
        self.mode = (_lib.EMIT_DEVICE if emit == "device" else _lib.EMIT_HOST if emit else _lib.EMIT_NONE)
        self.fn, self.emit, self.error = fn, self.mode == _lib.EMIT_HOST, None
        self.c = _lib.LEVEL_CB(self._call)

    def _call(self, info_p, _user) -> int:
        try:
            info = info_p.contents
            st = _stats_of(info.stat)
            rows = counts = None
            if self.mode == _lib.EMIT_DEVICE and st.k == 1 and st.n_frequent:
                n = st.n_frequent
                rows = np.ctypeslib.as_array(info.h_rows, shape=(n,)).reshape(n, 1).copy()
                counts = np.ctypeslib.as_array(info.h_counts, shape=(n,)).astype(np.int64)
            elif self.emit and st.n_frequent:
                n, k = st.n_frequent, st.k
                rows = np.ctypeslib.as_array(info.h_rows, shape=(n * k,)).reshape(n, k).copy()
                counts = np.ctypeslib.as_array(info.h_counts, shape=(n,)).astype(np.int64)
            elif self.emit:
                rows, counts = np.empty((0, st.k), np.int32), np.empty(0, np.int64)
            c2i = None
            if info.h_col_to_item and info.n_cols:
                c2i = np.ctypeslib.as_array(info.h_col_to_item, shape=(info.n_cols,)).copy()
            self.fn(LevelReport(st, rows, counts, int(info.device_bytes), int(info.host_bytes),
                                bool(info.next_level), c2i, info.d_rows, info.d_counts))
            return 0
        except BaseException as e:  # noqa: BLE001 - handed back to the caller after the C call returns
            self.error = e
            return 1

    def run(self, name, *args):
        try:
            _lib.call(name, *args)
        except _lib.Cancelled:
            if self.error is not None:
                raise self.error
            raise


def _opts(min_count, max_length, prune, infer, time_kernels, k2_kernel="auto", k3_kernel="auto",
          row_order="auto", live_bytes=False, callback: _Callback | None = None, k3_schedule="auto",
          small_max=0) -> _lib.MineOpts:
    if k2_kernel not in _lib.K2_KERNELS:
        raise ValueError(f"k2_kernel must be one of {sorted(_lib.K2_KERNELS)}, got {k2_kernel!r}")
    if k3_kernel not in _lib.K3_KERNELS:
        raise ValueError(f"k3_kernel must be one of {sorted(_lib.K3_KERNELS)}, got {k3_kernel!r}")
    if row_order not in _lib.ROW_ORDERS:
        raise ValueError(f"row_order must be one of {sorted(_lib.ROW_ORDERS)}, got {row_order!r}")
    if k3_schedule not in _lib.K3_SCHEDULES:
        raise ValueError(f"k3_schedule must be one of {sorted(_lib.K3_SCHEDULES)}, got {k3_schedule!r}")
    if small_max not in _lib.SMALL_MAXES:
        raise ValueError(f"small_max must be one of {_lib.SMALL_MAXES}, got {small_max!r}")
    opts = _lib.MineOpts(int(max(1, min_count)), int(max_length or 0), int(bool(prune)), int(bool(infer)),
                         int(bool(time_kernels)), _lib.K2_KERNELS[k2_kernel], _lib.K3_KERNELS[k3_kernel],
                         _lib.deduction_setting(), _lib.ROW_ORDERS[row_order], int(bool(live_bytes)), 0,
                         _lib.K3_SCHEDULES[k3_schedule], int(small_max))
    if callback is not None:
        opts.emit_levels = callback.mode
        opts.level_cb = callback.c
    return opts


class _ResultOwner:
    """Owns a result handle whose frame arrays are viewed in place: freed with the last array viewing them."""

    def __init__(self, handle):
        self.handle = handle

    def __del__(self):
        if self.handle is None:
            return
        try:
            _lib.lib().etm_levels_free(self.handle)
        except Exception:  # interpreter shutdown: the library may be gone already
            pass


class _FrameView:
    """One of a result's frame arrays, as numpy sees it (``np.asarray`` keeps this object, hence the owner)."""

    def __init__(self, owner: _ResultOwner, ptr: int, n: int, typestr: str):
        self.owner = owner
        self.__array_interface__ = {"shape": (n,), "typestr": typestr, "data": (ptr, False), "version": 3}


def _frame_view(handle):
    """The frame the mining call built while it ran (large lattices), as numpy arrays over the library's memory,
    or None; the result's per-level rows are released."""
    tot = np.zeros(2, dtype=np.int64)
    ptrs = (ctypes.c_void_p * 4)()
    _lib.call("etm_levels_frame_view", handle, tot.ctypes.data, ptrs, 1)
    if tot[0] < 0:
        return None
    owner = _ResultOwner(handle)
    n_sets, n_vals = int(tot[0]), int(tot[1])

    def view(i, n, typestr):
        return np.asarray(_FrameView(owner, ptrs[i], n, typestr)) if n > 0 else np.empty(0, dtype=np.dtype(typestr))

    try:
        return view(0, n_vals, "<i8"), view(1, n_sets + 1, "<i8"), view(2, n_sets, "<i8"), view(3, n_sets, "<f8")
    except BaseException:
        owner.handle = None  # the caller still frees the handle
        raise


def _collect(handle, levels: bool = True, n_rows: int = 0) -> NativeResult:
    """Reads a result handle (and frees it, or hands it to the arrays viewing its frame): the flat arrays
    always, the per-level arrays when ``levels``, and with ``n_rows`` > 0 each itemset's support (count /
    n_rows), written by the library's export threads."""
    L = _lib.lib()
    res = NativeResult()
    try:
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
            res.stats.append(_stats_of(st))
        n_cols = ctypes.c_int32()
        _lib.call("etm_levels_columns", handle, ctypes.byref(n_cols), None, None)
        c2i = np.empty(n_cols.value, dtype=np.int64)
        cnt = np.empty(n_cols.value, dtype=np.int64)
        _lib.call("etm_levels_columns", handle, ctypes.byref(n_cols), c2i.ctypes.data, cnt.ctypes.data)
        res.col_to_item, res.col_counts = c2i, cnt
        view = _frame_view(handle)
        if view is not None:
            handle = None  # the arrays own it now
            values, offsets, counts, support = view
            res.flat = (values, offsets, counts)
            res.support, res.support_rows = (support, int(n_rows)) if n_rows > 0 else (None, 0)
            return res
        tot = np.zeros(2, dtype=np.int64)
        _lib.call("etm_levels_frame", handle, tot.ctypes.data, None, None, None, None, 0)
        values = np.empty(tot[1], dtype=np.int64)
        offsets = np.empty(tot[0] + 1, dtype=np.int64)
        counts = np.empty(tot[0], dtype=np.int64)
        support = np.empty(tot[0], dtype=np.float64) if n_rows > 0 else None
        _lib.call("etm_levels_frame", handle, tot.ctypes.data, values.ctypes.data, offsets.ctypes.data,
                  counts.ctypes.data, None if support is None else support.ctypes.data, int(n_rows))
        res.flat = (values, offsets, counts)
        res.support, res.support_rows = support, int(n_rows)
    finally:
        if handle is not None:
            L.etm_levels_free(handle)
    return res


def mine_csr_device(indptr_d, ids_d, n_rows: int, min_count: int, max_length=None, prune_non_free=False,
                    infer_counts=False, time_kernels=False, levels: bool = True,
                    k2_kernel: str = "auto", k3_kernel: str = "auto", row_order: str = "auto",
                    on_level=None, emit_levels: bool = False, live_bytes: bool = False, k3_schedule: str = "auto",
                    small_max: int = 0) -> NativeResult:
    """Every frequent itemset of a device CSR of item ids (int32 or int64), mined by the native loop.

    ``k2_kernel``: "auto", "valu", "mfma" or "rows"; ``k3_kernel``: "auto", "valu" (bitvector kernels) or
    "rows" (row-wise, which needs rows of strictly ascending distinct ids); ``row_order``: "lex" (rows sorted by
    their most frequent items, with the tile skip), "keep", or "auto" (keeps the rows). ``on_level(report)`` is
    called with a LevelReport as each level completes (K=1 included); with ``emit_levels`` the report carries the
    level's rows and counts and the result keeps no itemsets (host memory holds one level at a time). An
    exception raised by ``on_level`` stops mining and propagates. With ``emit_levels="device"`` the levels K >= 2
    stay in device memory: the report carries ``d_rows`` / ``d_counts`` pointers, valid during the callback (work
    it enqueues on the current stream is ordered before any reuse), and no level is copied to the host; K=1 comes
    as host arrays. ``live_bytes`` fills ``LevelStats.bytes_live``.
    ``k3_schedule`` ("auto", "all", "live") and ``small_max`` (0, 4, 8, 16) choose how levels K >= 3 launch their
    counting (etm.h ETM_SCHEDULE_*, etm_mine_opts.small_max); neither changes a count."""
    torch = dev.torch()
    id_bytes = 8 if ids_d.dtype == torch.int64 else 4
    if ids_d.dtype not in (torch.int64, torch.int32):
        raise ValueError(f"item ids must be int32 or int64, got {ids_d.dtype}")
    if emit_levels and on_level is None:
        raise ValueError("emit_levels needs on_level")
    handle = ctypes.c_void_p()
    cb = _Callback(on_level, emit_levels) if on_level is not None else None
    opts = _opts(min_count, max_length, prune_non_free, infer_counts, time_kernels, k2_kernel, k3_kernel, row_order,
                 live_bytes, cb, k3_schedule, small_max)
    args = (indptr_d.data_ptr(), ids_d.data_ptr() if ids_d.numel() else None, id_bytes, n_rows, ids_d.numel(),
            SPAN_LIMIT, ctypes.byref(opts), ctypes.byref(handle), dev.stream_ptr())
    if cb is None:
        _lib.call("etm_mine_csr", *args)
    else:
        cb.run("etm_mine_csr", *args)
    return _collect(handle, levels, n_rows)


def mine_batch_device(jobs, k2_kernel: str = "auto", k3_kernel: str = "auto", levels: bool = True) -> list:
    """Many independent jobs mined in one call (``etm_mine_batch``): ``jobs`` holds (indptr_d, ids_d, n_rows,
    min_count, max_length) per job, device CSRs as for ``mine_csr_device``. Returns one NativeResult per job, the
    same itemsets and counts as ``mine_csr_device`` gives that job alone. Jobs whose K=1 level has at most 256
    columns advance together, one level per round; the rest (and any whose levels outgrow the rounds) are mined
    by the per-job loop with ``k2_kernel`` / ``k3_kernel``."""
    torch = dev.torch()
    n = len(jobs)
    arr = (_lib.BatchJob * max(n, 1))()
    keep = []
    for i, (indptr_d, ids_d, n_rows, min_count, max_length) in enumerate(jobs):
        if ids_d.dtype not in (torch.int64, torch.int32):
            raise ValueError(f"item ids must be int32 or int64, got {ids_d.dtype}")
        keep.append((indptr_d, ids_d))
        arr[i] = _lib.BatchJob(indptr_d.data_ptr(), ids_d.data_ptr() if ids_d.numel() else None, int(n_rows),
                               int(ids_d.numel()), int(max(1, min_count)), 8 if ids_d.dtype == torch.int64 else 4,
                               int(max_length or 0))
    opts = _opts(1, None, False, False, False, k2_kernel, k3_kernel)
    handles = (ctypes.c_void_p * max(n, 1))()
    _lib.call("etm_mine_batch", ctypes.byref(arr), n, SPAN_LIMIT, ctypes.byref(opts), ctypes.byref(handles),
              dev.stream_ptr())
    out = []
    try:
        for i in range(n):
            h, handles[i] = handles[i], None
            out.append(_collect(ctypes.c_void_p(h), levels))
    finally:
        for i in range(n):
            if handles[i]:
                _lib.lib().etm_levels_free(ctypes.c_void_p(handles[i]))
    return out


def mine_bitvecs_device(bv_ptr: int, pitch: int, n_words: int, n_rows: int, col_counts: np.ndarray, min_count: int,
                        max_length=None, prune_non_free=False, infer_counts=False,
                        time_kernels=False, k2_kernel: str = "auto", on_level=None,
                        emit_levels: bool = False, live_bytes: bool = False) -> NativeResult:
    """Every frequent itemset of device bitvectors (read only), from the columns' K=1 counts; ``on_level``,
    ``emit_levels`` and ``live_bytes`` as in ``mine_csr_device``."""
    counts = np.ascontiguousarray(col_counts, dtype=np.int64)
    if emit_levels and on_level is None:
        raise ValueError("emit_levels needs on_level")
    handle = ctypes.c_void_p()
    cb = _Callback(on_level, emit_levels) if on_level is not None else None
    opts = _opts(min_count, max_length, prune_non_free, infer_counts, time_kernels, k2_kernel, live_bytes=live_bytes,
                 callback=cb)
    args = (bv_ptr, pitch, n_words, n_rows, counts.ctypes.data, len(counts), ctypes.byref(opts),
            ctypes.byref(handle), dev.stream_ptr())
    if cb is None:
        _lib.call("etm_mine_bitvecs", *args)
    else:
        cb.run("etm_mine_bitvecs", *args)
    return _collect(handle, True, n_rows)
