"""The threshold contract and the result frame.

``_min_count`` follows the reference's ``core/result.py:57-79``: min_support denotes the shortest decimal that
round-trips to the float, so the threshold is ceil(Fraction(str(s)) * N) exactly (0.07 * 10,000 -> 700, where
the binary product would give 701).

The result has two columns: ``itemset`` (list of int64 item ids, ascending within each itemset) and
``support`` (float64 = count / N), as the reference returns (``core/apriori.py:423-440``). Row order is not
part of the contract. It is a polars DataFrame when polars is importable, else a pyarrow Table of the same
schema.
"""

from __future__ import annotations

import math
from fractions import Fraction

import numpy as np
import pyarrow as pa


def _min_count(min_support: float, n_transactions: int) -> int:
    """ceil(s * N) with s read as the shortest round-tripping decimal of the float."""
    return math.ceil(Fraction(str(min_support)) * n_transactions)


def _mining_min_count(min_support: float, n_transactions: int) -> int:
    """``_min_count``, at least 1 (reference ``core/cpu_miner.py:967-970``)."""
    return max(1, _min_count(min_support, n_transactions))


_POLARS = None


def _have_polars() -> bool:
    global _POLARS
    if _POLARS is None:
        try:
            import polars  # noqa: F401

            _POLARS = True
        except ImportError:
            _POLARS = False
    return _POLARS


def itemset_table(levels: list, col_to_item: np.ndarray, n_rows: int, with_counts: bool = False) -> pa.Table:
    """The result as a pyarrow Table from levels of (column-id rows int32 (n, k), counts int64).

    ``col_to_item`` maps column ids to item ids; None means the rows already hold item ids."""
    values, offsets, supports, counts = [], [np.zeros(1, dtype=np.int64)], [], []
    at = 0
    for rows, cnt in levels:
        n = len(rows)
        if n == 0:
            continue
        k = rows.shape[1]
        values.append((rows.ravel() if col_to_item is None else col_to_item[rows.ravel()]).astype(np.int64))
        offsets.append(at + k * np.arange(1, n + 1, dtype=np.int64))
        at += n * k
        supports.append(np.asarray(cnt, dtype=np.int64) / float(n_rows))
        counts.append(np.asarray(cnt, dtype=np.int64))
    vals = np.concatenate(values) if values else np.zeros(0, dtype=np.int64)
    offs = np.concatenate(offsets)
    itemsets = pa.LargeListArray.from_arrays(pa.array(offs, type=pa.int64()), pa.array(vals, type=pa.int64()))
    cols = {
        "itemset": itemsets,
        "support": pa.array(np.concatenate(supports) if supports else np.zeros(0), type=pa.float64()),
    }
    if with_counts:
        cols["count"] = pa.array(np.concatenate(counts) if counts else np.zeros(0, np.int64), type=pa.int64())
    return pa.table(cols)


def to_frame(table: pa.Table):
    """The result in the caller's frame type: polars when importable, else the pyarrow Table."""
    if _have_polars():
        import polars as pl

        df = pl.from_arrow(table)
        return df.with_columns(pl.col("itemset").cast(pl.List(pl.Int64)))
    return table


def empty_result():
    return to_frame(itemset_table([], np.zeros(0, dtype=np.int64), 1))


def as_dict(result) -> dict:
    """{tuple(itemset): count} of a result frame with counts, or {tuple: support} without (test helper)."""
    table = result if isinstance(result, pa.Table) else result.to_arrow()
    sets = [tuple(x) for x in table.column("itemset").to_pylist()]
    key = "count" if "count" in table.column_names else "support"
    return dict(zip(sets, table.column(key).to_pylist()))
