"""Per-level parquet output and resume: ``frequent_k{k}.parquet`` in ``output_dir``.

The on-disk format is the reference's (``io/flush.py:63-200``, ``gpu/row_split.py:491-606``):

* one file per itemset length k, ``frequent_k{k}.parquet``, or a directory ``frequent_k{k}/`` of
  ``part_*.parquet`` files (read, not written, here);
* columns ``itemset`` (``large_list<int32>``: the item ids of each itemset, ascending; ``large_list<int64>``
  when an id does not fit int32) and ``support`` (float64 = count / N).

Resume reads level k back as the generation base of level k + 1: item ids map to the current run's columns,
counts are ``rint(support * N)`` (exact through float64 for any int32 count), and rows are lexsorted. Only
local directories are served (the reference's ``gs://`` upload is outside the counting path).
"""

from __future__ import annotations

import os
from pathlib import Path

import numpy as np
import pyarrow as pa
import pyarrow.parquet as pq


def _check_local(output_dir: str) -> None:
    if "://" in str(output_dir):
        raise ValueError(f"output_dir={output_dir!r}: only local directories are served (no remote upload)")


def write_level(output_dir: str, k: int, item_rows: np.ndarray, support: np.ndarray) -> str:
    """Writes one level, ``frequent_k{k}.parquet``: ``item_rows`` (n, k) item ids ascending within each row,
    ``support`` float64 (n,). Returns the path. The reference's per-level flush (``gpu/row_split.py:512-537``)."""
    _check_local(output_dir)
    os.makedirs(output_dir, exist_ok=True)
    item_rows = np.asarray(item_rows)
    n = len(item_rows)
    flat = item_rows.reshape(-1)
    if flat.dtype.kind in "iu":
        narrow = len(flat) == 0 or (flat.min() >= -(2**31) and flat.max() < 2**31)
        flat = flat.astype(np.int32 if narrow else np.int64)
    lst = pa.LargeListArray.from_arrays(pa.array(np.arange(n + 1, dtype=np.int64) * k), pa.array(flat))
    path = os.path.join(output_dir, f"frequent_k{k}.parquet")
    pq.write_table(pa.table({"itemset": lst, "support": pa.array(np.asarray(support, dtype=np.float64))}), path)
    return path


def write_levels(table: pa.Table, output_dir: str, n_rows: int, min_k: int = 1) -> list:
    """Writes the itemsets of ``table`` (``itemset`` lists, ``count`` or ``support``) with at least ``min_k``
    items, one file per length; returns the paths written. Levels are written in length order."""
    _check_local(output_dir)
    os.makedirs(output_dir, exist_ok=True)
    lists = table.column("itemset").combine_chunks()
    if isinstance(lists, pa.ChunkedArray):
        lists = lists.combine_chunks()
    offsets = np.asarray(lists.offsets.to_numpy(zero_copy_only=False), dtype=np.int64)
    values = lists.values.to_numpy(zero_copy_only=False)
    lengths = np.diff(offsets)
    if "count" in table.column_names:
        support = np.asarray(table.column("count").to_numpy(), dtype=np.float64) / float(n_rows)
    else:
        support = np.asarray(table.column("support").to_numpy(), dtype=np.float64)
    narrow = values.dtype.kind in "iu" and (len(values) == 0 or (values.min() >= -(2**31) and values.max() < 2**31))
    written = []
    for k in sorted(set(lengths.tolist())):
        if k < min_k:
            continue
        idx = np.flatnonzero(lengths == k)
        starts = offsets[idx]
        flat = values[(starts[:, None] + np.arange(k)[None, :]).ravel()]
        if values.dtype.kind in "iu":
            flat = flat.astype(np.int32 if narrow else np.int64)
        item_arr = pa.array(flat)
        lst = pa.LargeListArray.from_arrays(pa.array(np.arange(len(idx) + 1, dtype=np.int64) * k), item_arr)
        path = os.path.join(output_dir, f"frequent_k{k}.parquet")
        pq.write_table(pa.table({"itemset": lst, "support": pa.array(support[idx])}), path)
        written.append(path)
    return written


def resolve_level(output_dir: str, k: int) -> list:
    """The files holding level k: ``frequent_k{k}.parquet``, or the parts of ``frequent_k{k}/``
    (``io/gcs.py:79-110`` ``resolve_k_parquet``)."""
    _check_local(output_dir)
    single = Path(output_dir) / f"frequent_k{k}.parquet"
    if single.is_file():
        return [str(single)]
    parts = sorted((Path(output_dir) / f"frequent_k{k}").glob("part_*.parquet"))
    if parts:
        return [str(p) for p in parts]
    raise FileNotFoundError(f"resume_from_k={k}: neither {single} nor {single.with_suffix('')}/part_*.parquet "
                            "exists")


def read_level(output_dir: str, k: int, n_rows: int) -> tuple:
    """(item ids int64 (n, k), counts int64) of level k as written by a previous run."""
    tables = [pq.read_table(p) for p in resolve_level(output_dir, k)]
    table = pa.concat_tables(tables) if len(tables) > 1 else tables[0]
    lists = table.column("itemset").combine_chunks()
    values = np.asarray(lists.values.to_numpy(zero_copy_only=False), dtype=np.int64)
    n = len(lists)
    if len(values) != n * k:
        raise ValueError(f"frequent_k{k}: {len(values)} items for {n} itemsets of {k}")
    rows = values.reshape(n, k)
    if "support" in table.column_names:
        counts = np.rint(np.asarray(table.column("support").to_numpy(), dtype=np.float64) * n_rows).astype(np.int64)
    else:
        raise ValueError(f"frequent_k{k}: no support column, so the level's counts cannot be rebuilt")
    return rows, counts


def columns_of(rows_items: np.ndarray, col_to_item: np.ndarray) -> np.ndarray:
    """Item-id rows mapped to this run's column ids, each row ascending, rows lexsorted. Raises when an item is
    not a frequent column of this run (the file does not belong to this input and support)."""
    order = np.argsort(col_to_item, kind="stable")
    sorted_items = col_to_item[order]
    flat = rows_items.ravel()
    pos = np.searchsorted(sorted_items, flat)
    pos = np.minimum(pos, max(0, len(sorted_items) - 1))
    if len(flat) and (len(sorted_items) == 0 or np.any(sorted_items[pos] != flat)):
        raise ValueError("the resumed level holds items that are not frequent items of this input: it was "
                         "written for other transactions or another min_support")
    cols = order[pos].astype(np.int32).reshape(rows_items.shape)
    return np.sort(cols, axis=1)


def lexsorted(rows: np.ndarray, counts: np.ndarray) -> tuple:
    if len(rows) == 0:
        return rows, counts
    order = np.lexsort(rows[:, ::-1].T)
    return np.ascontiguousarray(rows[order]), counts[order]
