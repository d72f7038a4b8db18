"""apriori(): the reference's public entry point, with the GPU counting path behind it.

Signature, parameter validation and route checks follow the reference's ``core/apriori.py:64-251`` and
``:303-754``. Routes served here:

* ``bitvecs=(bitvecs_gpu, col_to_item, n_transactions)`` -> the level loop on the caller's bitvectors
  (read-only), from K=1;
* ``use_gpu=True`` -> the device input layer (transactions -> CSR in HBM -> bitvectors), then the level loop;
* ``streaming=True, use_gpu=True`` -> SON's two passes, each chunk on the GPU.

The CPU route (``use_gpu=False`` without streaming or bitvecs) is not part of this build: it raises
ValueError rather than falling back. One deliberate deviation: ``streaming=True`` accepts
``use_generator_pruning=True`` (the reference raises, ``core/apriori.py:160-170``); Pascal inference is exact,
so pass 1 applies it per chunk and the result is the complete lattice either way.
"""

from __future__ import annotations

import math
import numbers
import warnings
from math import comb

import numpy as np
import pyarrow as pa

from .profiling import ProfilingSession
from .result import _min_count, _mining_min_count, empty_result, itemset_table, to_frame


def _warn_complexity(n_frequent_items: int, min_support: float) -> None:
    n_pairs = comb(n_frequent_items, 2) if n_frequent_items > 1 else 0
    if n_pairs > 10_000_000:
        warnings.warn(
            f"Very large computation ahead!\n  Frequent items: {n_frequent_items:,}\n  Candidate pairs: {n_pairs:,}\n"
            f"  Consider increasing min_support (currently {min_support})", UserWarning, stacklevel=3)
    elif n_pairs > 1_000_000:
        warnings.warn(f"Large computation: {n_pairs:,} candidate pairs.", UserWarning, stacklevel=3)


def _validate_parameters(min_support, max_length, batch_size, sparse_from_k=None, prune_apriori=True) -> None:
    """Parameter checks with the reference's messages and exception types (``core/apriori.py:64-106``)."""
    if isinstance(min_support, bool) or not isinstance(min_support, (int, float)):
        raise TypeError(f"min_support must be numeric, got {type(min_support).__name__}")
    if not 0.0 <= min_support <= 1.0:
        raise ValueError(f"min_support must be in range [0.0, 1.0], got {min_support}\n"
                         f"  Hint: Support is a proportion (e.g., 0.01 = 1%)")
    if max_length is not None:
        if isinstance(max_length, bool) or not isinstance(max_length, int):
            raise TypeError(f"max_length must be int or None, got {type(max_length).__name__}")
        if max_length < 1:
            raise ValueError(f"max_length must be >= 1, got {max_length}")
    if batch_size is not None:
        if isinstance(batch_size, bool) or not isinstance(batch_size, int):
            raise TypeError(f"batch_size must be int or None, got {type(batch_size).__name__}")
        if batch_size < 1:
            raise ValueError(f"batch_size must be >= 1, got {batch_size}")
    if sparse_from_k is not None:
        raise ValueError("sparse_from_k was removed: the GPU miner keeps dense bitvectors at every level. "
                         "Drop the argument.")
    if not isinstance(prune_apriori, bool):
        raise TypeError(f"prune_apriori must be bool, got {type(prune_apriori).__name__}")
    if not prune_apriori:
        raise ValueError("prune_apriori=False was removed: every route tests each candidate's (k-1)-subsets. "
                         "Drop the argument.")


def _validate_route_support(*, streaming, use_gpu, has_bitvecs, n_gpus, profile, gpu_resident, prune_equal_support,
                            anchor_items, output_dir, resume_from_k, memory_budget_gb) -> None:
    """Rejects route/parameter combinations a route cannot honour (``core/apriori.py:109-251``)."""
    if gpu_resident:
        raise ValueError("gpu_resident was removed: use_gpu=True (or bitvecs=) mines on the GPU level loop, "
                         "which keeps candidates and counts on the GPU. Drop the argument.")
    if streaming and prune_equal_support:
        raise ValueError("streaming=True cannot be combined with prune_equal_support: the SON streaming path "
                         "does not implement the free-set gate and would return the complete frequent lattice "
                         "instead of the free-sets. Drop one of the two, or mine without streaming.")
    if anchor_items is not None:
        raise ValueError("anchor_items (two-phase anchored mining) is not part of this build's GPU counting path")
    if output_dir is not None or resume_from_k is not None:
        which = " and ".join(n for n, v in (("output_dir", output_dir), ("resume_from_k", resume_from_k))
                             if v is not None)
        if streaming and not has_bitvecs:
            raise ValueError(f"{which} requires the level-loop miner, reached with bitvecs= or with use_gpu=True "
                             "and streaming=False. On this route the per-K parquet flush does not run: output_dir "
                             "would stay empty and resume_from_k would silently re-mine from K=1.")
    if memory_budget_gb is not None and not streaming:
        raise ValueError("memory_budget_gb requires streaming=True: it overrides chunk_size for the SON paths "
                         "and is read nowhere else.")
    if n_gpus is not None and n_gpus > 1:
        raise ValueError(f"n_gpus={n_gpus}: this build counts on one GPU (row-split over several GPUs is not "
                         "built)")
    if not (use_gpu or has_bitvecs):
        raise ValueError("this build implements the GPU counting path only: pass use_gpu=True (optionally with "
                         "streaming=True) or bitvecs=; the CPU route is not provided")


def _encode_values(values: pa.Array) -> tuple:
    """(int64 ids, items to map back or None): integer items pass through; others get ids in sorted order."""
    if pa.types.is_integer(values.type):
        return np.asarray(values.cast(pa.int64()).to_numpy(zero_copy_only=False)), None
    uniq = pa.compute.unique(values)
    order = pa.compute.sort_indices(uniq)
    items = uniq.take(order)
    ids = pa.compute.index_in(values, value_set=items)
    return np.asarray(ids.cast(pa.int64()).to_numpy(zero_copy_only=False)), items


def transactions_to_csr(transactions, item_col: str = "items") -> tuple:
    """(indptr int64, ids int64, items or None) of a transaction list column.

    Accepts a polars DataFrame/LazyFrame, a pyarrow Table / (Large)ListArray / ChunkedArray, a pandas
    DataFrame, a list of lists, or a CSR tuple (indptr, ids). Nulls inside lists are dropped; a null list
    is an empty transaction. ``items`` is the sorted distinct item values when they are not integers.
    """
    if isinstance(transactions, tuple) and len(transactions) == 2:
        indptr = np.asarray(transactions[0], dtype=np.int64)
        ids = np.asarray(transactions[1])
        if ids.dtype.kind not in "iu":
            raise ValueError("a CSR tuple needs integer item ids")
        return indptr, ids.astype(np.int64, copy=False), None
    col = None
    mod = type(transactions).__module__
    if mod.startswith("polars"):
        if hasattr(transactions, "collect") and not hasattr(transactions, "to_arrow"):
            transactions = transactions.collect()
        col = transactions.get_column(item_col).to_arrow()
    elif isinstance(transactions, pa.Table):
        col = transactions.column(item_col)
    elif isinstance(transactions, (pa.Array, pa.ChunkedArray)):
        col = transactions
    elif mod.startswith("pandas"):
        col = pa.array(transactions[item_col].tolist())
    elif isinstance(transactions, list):
        col = pa.array(transactions)
    else:
        raise TypeError(f"unsupported transactions type {type(transactions).__name__}")
    if isinstance(col, pa.ChunkedArray):
        col = col.combine_chunks() if col.num_chunks != 1 else col.chunk(0)
    if not (pa.types.is_list(col.type) or pa.types.is_large_list(col.type)):
        raise TypeError(f"column {item_col!r} must be a list column, got {col.type}")
    col = col.cast(pa.large_list(col.type.value_type))
    n = len(col)
    offsets = np.asarray(col.offsets.to_numpy(zero_copy_only=False), dtype=np.int64)
    values = col.values.slice(offsets[0], offsets[-1] - offsets[0]) if n else col.values.slice(0, 0)
    offsets = offsets - offsets[0]
    if col.null_count:
        lens = np.diff(offsets)
        lens[np.asarray(col.is_null().to_numpy(zero_copy_only=False))] = 0
        keep = np.repeat(lens > 0, np.diff(offsets))
        values = values.filter(pa.array(keep))
        offsets = np.zeros(n + 1, dtype=np.int64)
        np.cumsum(lens, out=offsets[1:])
    if values.null_count:
        valid = np.asarray(values.is_valid().to_numpy(zero_copy_only=False))
        rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(offsets))
        values = values.filter(pa.array(valid))
        offsets = np.zeros(n + 1, dtype=np.int64)
        np.cumsum(np.bincount(rows[valid], minlength=n), out=offsets[1:])
    ids, items = _encode_values(values)
    return offsets, ids, items


def _frame(levels, col_to_item, n_rows, items):
    table = itemset_table(levels, col_to_item, n_rows)
    if items is not None:
        flat = table.column("itemset").combine_chunks()
        mapped = items.take(flat.values)
        table = pa.table({"itemset": pa.LargeListArray.from_arrays(flat.offsets, mapped),
                          "support": table.column("support")})
    return to_frame(table)


def apriori(
    transactions=None,
    min_support: float = 0.5,
    max_length: int | None = None,
    item_col: str = "items",
    use_gpu: bool = False,
    batch_size: int | None = 10_000,
    profile: bool = False,
    show_progress: bool = False,
    warn_complexity: bool = True,
    prune_equal_support: bool = False,
    use_generator_pruning: bool = False,
    prune_apriori: bool = True,
    sparse: bool | None = None,
    n_jobs: int = 1,
    enable_length_filter: bool = True,
    streaming: bool = False,
    chunk_size: int = 10_000_000,
    n_gpus: int = 1,
    memory_budget_gb: float | None = None,
    progress_callback=None,
    level_callback=None,
    gpu_resident: bool = False,
    bitvecs: tuple | None = None,
    output_dir: str | None = None,
    resume_from_k: int | None = None,
    max_ram_gb: float = 800.0,
    max_vram_gb: float = 70.0,
    sparse_from_k=None,
    anchor_items: set | None = None,
):
    """Frequent itemsets (or free-sets) of a transaction list column, counted on the GPU.

    Returns a frame with ``itemset`` (int64 item ids ascending within each itemset) and ``support`` (float64 =
    count / N); ``(frame, ProfilingSession)`` with ``profile=True``. The parameters mean what they mean in the
    reference (``core/apriori.py:303-420``); see the module docstring for the routes served.
    """
    _validate_parameters(min_support, max_length, batch_size, sparse_from_k, prune_apriori)
    if sparse is not None:
        raise ValueError("sparse= was removed: no route reads it. Drop the argument.")
    _validate_route_support(
        streaming=streaming, use_gpu=use_gpu, has_bitvecs=bitvecs is not None, n_gpus=n_gpus, profile=profile,
        gpu_resident=gpu_resident, prune_equal_support=prune_equal_support, anchor_items=anchor_items,
        output_dir=output_dir, resume_from_k=resume_from_k, memory_budget_gb=memory_budget_gb,
    )
    session = ProfilingSession() if profile else None

    def done(df):
        return (df, session) if profile else df

    if bitvecs is not None:
        if transactions is not None:
            raise ValueError("Cannot provide both 'transactions' and 'bitvecs'. Use 'bitvecs' for pre-built GPU "
                             "bitvectors or 'transactions' for DataFrame input.")
        return done(_mine_bitvecs(bitvecs, min_support, max_length, prune_equal_support, use_generator_pruning,
                                  level_callback, session, warn_complexity, output_dir, resume_from_k, max_ram_gb,
                                  max_vram_gb))
    if transactions is None:
        raise ValueError("Either 'transactions' or 'bitvecs' must be provided")
    if streaming:
        from ..streaming.son import apriori_streaming

        return apriori_streaming(transactions, min_support=min_support, max_length=max_length, item_col=item_col,
                                 chunk_size=chunk_size, memory_budget_gb=memory_budget_gb, use_gpu=True,
                                 profile=profile, progress_callback=progress_callback,
                                 use_generator_pruning=use_generator_pruning)
    if session:
        session.start_phase("matrix_build")
    indptr, ids, items = transactions_to_csr(transactions, item_col)
    if session:
        session.end_phase(n_transactions=len(indptr) - 1)
    return done(mine_csr(indptr, ids, min_support, max_length, prune_equal_support, use_generator_pruning,
                         level_callback, session, items=items, warn_complexity=warn_complexity,
                         output_dir=output_dir, resume_from_k=resume_from_k, max_ram_gb=max_ram_gb,
                         max_vram_gb=max_vram_gb))


def mine_csr(indptr, ids, min_support, max_length=None, prune_equal_support=False, use_generator_pruning=False,
             level_callback=None, session=None, items=None, warn_complexity=False, device_csr=None,
             with_counts=False, timer=None, stats_out=None, engine: str = "native", time_kernels=False,
             k2_kernel: str = "auto", k3_kernel: str = "auto", row_order: str = "auto", output_dir=None,
             resume_from_k=None, max_ram_gb=None, max_vram_gb=None, live_bytes=False, k3_schedule: str = "auto",
             small_max: int = 0):
    """The use_gpu route from a host CSR (or ``device_csr=(indptr_d, ids_d)`` already in HBM).

    ``engine="native"`` runs the whole level loop inside libetm (one C call); ``engine="python"`` drives the
    same kernels level by level from Python through the operator wrappers (``timer`` hooks apply there).
    ``stats_out`` (a list) receives the per-level statistics. ``k2_kernel`` ("auto", "valu", "mfma",
    "rows") picks the K=2 pair space's kernel: the fp4 MFMA Gram, the VALU AND+popcount tiles or the row-wise
    kernel; ``k3_kernel`` ("auto", "valu", "rows") picks K=3's; ``row_order`` ("auto", "keep", "lex") the
    bitvectors' row order (native engine only; counts never depend on it).

    ``level_callback(k, n_candidates, n_frequent, duration_ms)`` is called as each level completes, with the
    level's wall time; ``max_ram_gb`` / ``max_vram_gb`` raise MemoryError between levels as the reference's
    memory guard does (``gpu/row_split.py:467-487``). ``output_dir`` writes each level to
    ``frequent_k{k}.parquet`` there as it completes (host memory holds one level at a time) and returns an
    empty result, as the reference's row-split miner does (``gpu/row_split.py:512-537``, ``:1172-1178``); with
    ``resume_from_k`` >= 2 as well, level ``resume_from_k`` is read back from that directory and mining
    continues above it (the Python level loop; ``gpu/row_split.py:539-606``). ``live_bytes`` fills the native
    loop's ``LevelStats.bytes_live``; ``k3_schedule`` and ``small_max`` pick the native loop's launch schedule of
    levels K >= 3 (see ``gpu.native.mine_csr_device``).
    """
    from ..gpu import device as dev

    dev.require_gpu()
    n_rows = (len(indptr) - 1) if device_csr is None else int(device_csr[0].numel()) - 1
    if n_rows <= 0:
        return itemset_table([], np.zeros(0, np.int64), 1, with_counts=True) if with_counts else empty_result()
    min_count = _mining_min_count(min_support, n_rows)
    if output_dir is not None and resume_from_k is not None and resume_from_k >= 2:
        return _resume_csr(indptr, ids, device_csr, n_rows, min_count, max_length, prune_equal_support,
                           use_generator_pruning, level_callback, session, items, output_dir, resume_from_k,
                           stats_out, max_ram_gb, max_vram_gb)
    if session:
        session.start_phase("input_layer")
    if device_csr is None:
        from ..gpu.input_layer import host_csr_to_device

        device_csr = host_csr_to_device(indptr, ids)
    indptr_d, ids_d = device_csr
    hooks = _LevelHooks(level_callback, session, max_ram_gb, max_vram_gb, output_dir, n_rows, items)
    if engine == "native":
        from ..gpu.native import mine_csr_device

        if session:
            session.end_phase()
        res = mine_csr_device(indptr_d, ids_d, n_rows, min_count, max_length, prune_equal_support,
                              use_generator_pruning, time_kernels=time_kernels, levels=False, k2_kernel=k2_kernel,
                              k3_kernel=k3_kernel, row_order=row_order, on_level=hooks.native(),
                              emit_levels=hooks.emit, live_bytes=live_bytes, k3_schedule=k3_schedule,
                              small_max=small_max)
        if warn_complexity and res.stats:
            _warn_complexity(res.stats[0].n_frequent, min_support)
        if stats_out is not None:
            stats_out.extend(res.stats)
        if hooks.emit:
            return empty_result()
        table = res.table(n_rows, with_counts)
        if items is not None:
            flat = table.column("itemset").combine_chunks()
            table = pa.table({"itemset": pa.LargeListArray.from_arrays(flat.offsets, items.take(flat.values)),
                              "support": table.column("support")})
        return table if with_counts else to_frame(table)
    from ..gpu.input_layer import build_device_input
    from ..gpu.row_split import mine_levels

    built = build_device_input(indptr_d, ids_d, n_rows, min_count)
    if session:
        session.end_phase(n_columns=0 if built is None else len(built.col_to_item))
    if built is None:
        levels, stats, col_to_item = [], [], np.zeros(0, np.int64)
    else:
        if warn_complexity:
            _warn_complexity(int((built.col_counts >= min_count).sum()), min_support)
        hooks.col_to_item = np.asarray(built.col_to_item, dtype=np.int64)
        res = mine_levels(built.bitvecs, built.n_words, n_rows, built.col_counts, min_count, max_length,
                          prune_equal_support, use_generator_pruning, hooks.python(), timer=timer,
                          guard=hooks.guard_python, level_sink=hooks.sink() if hooks.emit else None)
        levels, stats, col_to_item = res.levels, res.stats, built.col_to_item
    if stats_out is not None:
        stats_out.extend(stats)
    if hooks.emit:
        return empty_result()
    levels = [(r, c) for r, c in levels if len(r)]
    if with_counts:
        return itemset_table(levels, col_to_item, n_rows, with_counts=True)
    if not levels:
        return empty_result()
    return _frame(levels, col_to_item, n_rows, items)


class _LevelHooks:
    """The per-level work of a mining call, run as each level completes: the reference's ``level_callback``
    (``k, n_candidates, n_frequent, duration_ms``; ``gpu/row_split.py:855-856, 1140-1141``), the profile
    session's phase per level, the memory guards (``gpu/row_split.py:467-487``: checked when another level
    follows) and the per-level parquet flush of ``output_dir`` (``gpu/row_split.py:512-537``).

    ``items`` (the sorted distinct values of non-integer items, or None) and ``col_to_item`` (set by the route,
    or taken from the native loop's report) map column ids to what the files hold."""

    def __init__(self, level_callback, session, max_ram_gb, max_vram_gb, output_dir, n_rows, items=None,
                 col_to_item=None, sort_rows=False):
        self.level_callback = level_callback
        self.session = session
        self.max_ram_gb, self.max_vram_gb = max_ram_gb, max_vram_gb
        self.output_dir = output_dir
        self.n_rows = n_rows
        self.items = items
        self.col_to_item = col_to_item
        self.sort_rows = sort_rows
        self.emit = output_dir is not None
        if self.emit:
            from ..io import flush

            flush._check_local(output_dir)

    @property
    def active(self) -> bool:
        return bool(self.level_callback or self.session or self.emit or self.max_ram_gb is not None
                    or self.max_vram_gb is not None)

    def _record(self, k, n_cand, n_freq, ms):
        if self.session:
            self.session.record_phase(f"k{k}", ms, n_candidates=n_cand, n_frequent=n_freq)
        if self.level_callback:
            self.level_callback(k, n_cand, n_freq, ms)

    def _write(self, k, rows, counts):
        from ..io import flush

        if len(rows) == 0:  # the reference writes no file for an empty level (gpu/row_split.py:1131-1133)
            return
        item_rows = np.asarray(self.col_to_item, dtype=np.int64)[rows] if len(rows) else np.empty((0, k), np.int64)
        if self.sort_rows:
            item_rows = np.sort(item_rows, axis=1)
        if self.items is not None:
            item_rows = np.asarray(self.items.take(pa.array(item_rows.reshape(-1))).to_numpy(zero_copy_only=False),
                                   dtype=object).reshape(item_rows.shape)
        flush.write_level(self.output_dir, k, item_rows, np.asarray(counts, dtype=np.float64) / float(self.n_rows))

    def guard(self, k, device_bytes, host_bytes) -> None:
        if self.max_ram_gb is None and self.max_vram_gb is None:
            return
        import psutil

        ram = (psutil.Process().memory_info().rss + host_bytes) / float(1 << 30)
        vram = (device_bytes + _torch_allocated()) / float(1 << 30)
        for used, limit, what, remedy in (
            (ram, self.max_ram_gb, "RAM", "pass output_dir so each level is flushed as it completes"),
            (vram, self.max_vram_gb, "VRAM", "raise min_support, or mine with streaming=True to count in chunks"),
        ):
            if limit is not None and used > limit:
                raise MemoryError(
                    f"Memory guard tripped after K={k}: {what}={used:.1f}GB exceeds max_{what.lower()}_gb={limit}GB. "
                    "The lattice is INCOMPLETE at this point, so it is not returned. "
                    f"Raise max_{what.lower()}_gb, lower max_length, or {remedy}.")

    def native(self):
        """The native loop's on_level (None when nothing is asked for)."""
        if not self.active:
            return None

        def on_level(rep):
            st = rep.stats
            if self.col_to_item is None and rep.col_to_item is not None:
                self.col_to_item = rep.col_to_item
            if self.emit:
                self._write(st.k, rep.rows, rep.counts)
            self._record(st.k, st.n_candidates, st.n_frequent, st.ms)
            if rep.next_level:
                self.guard(st.k, rep.device_bytes, rep.host_bytes)

        return on_level

    def python(self):
        return self._record if (self.level_callback or self.session) else None

    def guard_python(self, k) -> None:
        self.guard(k, 0, 0)

    def sink(self):
        return self._write


def _torch_allocated() -> int:
    """Device bytes torch holds in this process (the resident CSR, the caller's bitvectors)."""
    from ..gpu import device as dev

    t = dev.torch()
    return int(t.cuda.memory_allocated()) if t.cuda.is_available() else 0


def _resume_csr(indptr, ids, device_csr, n_rows, min_count, max_length, prune_equal_support, use_generator_pruning,
                level_callback, session, items, output_dir, resume_from_k, stats_out, max_ram_gb, max_vram_gb):
    """output_dir + resume_from_k: level ``resume_from_k`` read back, the levels above it mined and written there
    one at a time; an empty result back."""
    if items is not None:
        raise ValueError("resume_from_k needs integer item ids (the flushed levels hold ids)")
    from ..gpu import device as dev
    from ..gpu.input_layer import build_device_input, host_csr_to_device

    if device_csr is None:
        device_csr = host_csr_to_device(indptr, ids)
    indptr_d, ids_d = device_csr
    built = build_device_input(indptr_d, ids_d, n_rows, min_count)
    if built is None:
        return empty_result()
    hooks = _LevelHooks(level_callback, session, max_ram_gb, max_vram_gb, output_dir, n_rows,
                        col_to_item=np.asarray(built.col_to_item, dtype=np.int64))
    res = _resume_levels(built.bitvecs, built.n_words, n_rows, built.col_counts, min_count, max_length,
                         prune_equal_support, use_generator_pruning, hooks, resume_from_k, None)
    if stats_out is not None:
        stats_out.extend(res.stats)
    dev.synchronize()
    return empty_result()


def _resume_levels(bv, n_words, n_rows, col_counts, min_count, max_length, prune, infer, hooks, resume_from_k, pitch):
    """Level ``resume_from_k`` back from ``hooks.output_dir``; the levels above it mined on the Python level loop
    and written there as each completes (``gpu/row_split.py:539-606``)."""
    from ..gpu.row_split import mine_levels
    from ..io import flush

    items_rows, counts = flush.read_level(hooks.output_dir, resume_from_k, n_rows)
    cols, counts = flush.lexsorted(flush.columns_of(items_rows, np.asarray(hooks.col_to_item, dtype=np.int64)),
                                   counts)
    hooks.sort_rows = True
    return mine_levels(bv, n_words, n_rows, col_counts, min_count, max_length, prune, infer, hooks.python(),
                       pitch=pitch, start=(cols, counts), guard=hooks.guard_python, level_sink=hooks.sink())


def _mine_bitvecs(bitvecs, min_support, max_length, prune_equal_support, use_generator_pruning, level_callback,
                  session, warn_complexity, output_dir=None, resume_from_k=None, max_ram_gb=None, max_vram_gb=None):
    """The bitvecs= route: validation as the reference (``core/apriori.py:477-527``), then the level loop."""
    from ..gpu import device as dev
    from ..gpu import kernels as K

    if not isinstance(bitvecs, tuple) or len(bitvecs) != 3:
        raise ValueError("bitvecs must be a tuple of (bitvecs_gpu, col_to_item, n_transactions)")
    bitvecs_gpu, col_to_item, n_trans = bitvecs
    arr = dev.describe(bitvecs_gpu)
    if arr.ndim != 2:
        raise ValueError(f"bitvecs_gpu must be 2D (n_cols, n_u64s), got {arr.ndim}D")
    expected = math.ceil(n_trans / 64)
    if arr.shape[1] != expected:
        raise ValueError(f"bitvecs_gpu.shape[1] ({arr.shape[1]}) must equal ceil(n_transactions/64) = "
                         f"ceil({n_trans}/64) = {expected}")
    if len(col_to_item) != arr.shape[0]:
        raise ValueError(f"col_to_item has {len(col_to_item)} keys but bitvecs_gpu has {arr.shape[0]} columns")
    if not all(isinstance(v, numbers.Integral) for v in col_to_item.values()):
        raise ValueError("col_to_item must map every column to an integer item id: the bitvecs route holds item "
                         "ids as 64-bit integers. Map the items to integer ids first.")
    if arr.dtype is None or arr.dtype.itemsize != 8 or arr.dtype.kind not in "iu" or not arr.c_contiguous:
        raise ValueError(f"bitvecs_gpu must be a C-contiguous uint64 array, got {arr.typestr}, "
                         f"c_contiguous={arr.c_contiguous}")
    dev.require_gpu()
    n_cols = arr.shape[0]
    mapping = np.zeros(n_cols, dtype=np.int64)
    for c, item in col_to_item.items():
        mapping[int(c)] = int(item)
    order_ok = n_cols < 2 or bool(np.all(np.diff(mapping) > 0))
    min_count = _min_count(min_support, n_trans)
    view = _RawBitvecs(arr)
    counts = dev.to_host(K.column_popcounts(view, arr.shape[1])).astype(np.int64)
    if warn_complexity:
        _warn_complexity(int((counts >= max(1, min_count)).sum()), min_support)
    hooks = _LevelHooks(level_callback, session, max_ram_gb, max_vram_gb, output_dir, n_trans, col_to_item=mapping,
                        sort_rows=not order_ok)
    if output_dir is not None and resume_from_k is not None and resume_from_k >= 2:
        _resume_levels(view, arr.shape[1], n_trans, counts, max(1, min_count), max_length, prune_equal_support,
                       use_generator_pruning, hooks, resume_from_k, arr.shape[1])
        dev.synchronize()
        return empty_result()
    from ..gpu.native import mine_bitvecs_device

    res = mine_bitvecs_device(arr.ptr, arr.shape[1], arr.shape[1], n_trans, counts, max(1, min_count), max_length,
                              prune_equal_support, use_generator_pruning, on_level=hooks.native(),
                              emit_levels=hooks.emit)
    if hooks.emit:
        return empty_result()
    levels = [(r, c) for r, c in res.levels if len(r)]
    if not order_ok:
        # Item ids that do not ascend with the columns: map each level, then sort every itemset's items.
        levels = [(np.sort(mapping[rows], axis=1), cnt) for rows, cnt in levels]
        mapping = None
    if mapping is None:
        return to_frame(itemset_table(levels, None, n_trans))
    return _frame(levels, mapping, n_trans, None)


class _RawBitvecs:
    """A borrowed (n_cols, n_words) uint64 device array, exposing what the kernel wrappers read."""

    def __init__(self, arr):
        self._arr = arr
        self.shape = arr.shape

    def data_ptr(self) -> int:
        return self._arr.ptr
