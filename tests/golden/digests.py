"""Per-level digests of a frequent-itemset lattice (test infrastructure).

A level's digest is the SHA-256 of its (itemset, count) pairs laid out as one little-endian int64 array of shape
(n, k + 1): items ascending within each itemset, itemsets in lexicographic order, the count last. Two lattices
have equal digests at every level exactly when they hold the same itemsets with the same counts, whatever order
either side produced them in.
"""

from __future__ import annotations

import hashlib

import numpy as np


def _level_digest(rows: np.ndarray, counts: np.ndarray) -> str:
    rows = np.sort(np.asarray(rows, dtype=np.int64), axis=1)
    counts = np.asarray(counts, dtype=np.int64)
    order = np.lexsort(rows.T[::-1]) if len(rows) else np.zeros(0, dtype=np.int64)
    arr = np.concatenate([rows[order], counts[order][:, None]], axis=1).astype("<i8")
    return hashlib.sha256(np.ascontiguousarray(arr).tobytes()).hexdigest()


def level_digests(values: np.ndarray, offsets: np.ndarray, counts: np.ndarray) -> list:
    """[{"k", "n", "sha256"}] per itemset length of a flat lattice (item ids, offsets, counts), ascending k."""
    values = np.asarray(values, dtype=np.int64)
    offsets = np.asarray(offsets, dtype=np.int64)
    counts = np.asarray(counts, dtype=np.int64)
    lens = np.diff(offsets)
    out = []
    for k in np.unique(lens).tolist():
        sel = np.flatnonzero(lens == k)
        rows = values[offsets[sel][:, None] + np.arange(k)]
        out.append({"k": int(k), "n": int(len(sel)), "sha256": _level_digest(rows, counts[sel])})
    return out


def table_digests(table) -> list:
    """level_digests of a pyarrow table with ``itemset`` (list of item ids) and ``count`` columns."""
    flat = table.column("itemset").combine_chunks()
    return level_digests(np.asarray(flat.values), np.asarray(flat.offsets), np.asarray(table.column("count")))


def lattice_digests(levels: list, items: np.ndarray) -> list:
    """level_digests of levels given as (column-id rows (n, k), counts), mapped to item ids by ``items``."""
    out = []
    for sets, counts in levels:
        if len(counts) == 0:
            continue
        rows = np.asarray(items, dtype=np.int64)[np.asarray(sets)]
        out.append({"k": int(rows.shape[1]), "n": int(len(counts)), "sha256": _level_digest(rows, counts)})
    return sorted(out, key=lambda d: d["k"])
