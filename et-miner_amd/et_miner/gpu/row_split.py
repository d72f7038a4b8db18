"""The GPU level loop on one device: K=1 through exhaustion, every level resident in HBM.

Counterpart of the reference's ``_apriori_row_split_multi_gpu`` (``gpu/row_split.py:229-1185``) for a single
GPU. Per level k >= 2 the loop runs, all on the device:

1. the prefix groups of level k-1 (``etm_level_groups``; for k = 2 the K=1 level is one group);
2. for k >= 3, the (k-1)-subset test of every candidate against the hash of level k-1, with Pascal inference
   when ``infer_counts`` (``etm_level_classify``), then the deduction bounds over levels k-2 and k-3
   (``etm_level_deduce``) unless ``ET_MINER_DEDUCTION=0``;
3. AND + popcount of the counted candidates (``etm_count_level``);
4. compaction of the survivors (count >= min_count) in lexicographic order (``etm_level_compact``) -- the
   candidate numbering is lexicographic, so no level-end sort is needed;
5. free flags when tracked (``etm_level_free``), and the free-set prune (``etm_select_rows``).

The host reads two scalars per level (the candidate total and the survivor total) to size the next buffers;
the itemsets and counts of every level come back once, at the end.

Free-set mode (``prune_non_free``) keeps one population per level, as the reference does
(``gpu/row_split.py:292-309``): the free-sets, which are emitted, generated from, and tested against.
"""

from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np

from .. import _lib
from . import device as dev
from . import kernels as K


@dataclass
class LevelStats:
    """One level: candidates generated by the prefix join, and of them counted / inferred / skipped by the
    subset test; ``bytes_alg`` is the compulsory HBM traffic of the prefix-group decomposition over every group,
    ``bytes_live`` the same over what the counting kernels must read for the candidates left counted (native
    loop, when asked for); ``ms`` is the level's wall time."""

    k: int
    n_candidates: int
    n_frequent: int
    n_counted: int = 0
    n_inferred: int = 0
    n_skipped: int = 0
    bytes_alg: int = 0
    ms: float = 0.0
    count_ms: float = 0.0  # event-timed duration of the level's counting launches (native loop, when timed)
    kernel: str = ""  # the native loop's counting kernel family for the level: "valu", "mfma" or "rows"
    bytes_live: int = 0

    @property
    def n_evaluated(self) -> int:
        return self.n_counted + self.n_inferred


@dataclass
class MineResult:
    """Levels as (rows of bitvector column ids int32 (n, k), counts int64), plus per-level statistics."""

    levels: list = field(default_factory=list)
    stats: list = field(default_factory=list)


@dataclass
class _Level:
    rows: object  # device int32 (n, m)
    counts: object  # device int32 (n,)
    free: object = None  # device uint8 (n,) or None
    table: object = None
    table_size: int = 0

    @property
    def n(self) -> int:
        return int(self.rows.shape[0])

    @property
    def m(self) -> int:
        return int(self.rows.shape[1])

    def hashed(self):
        if self.table is None:
            self.table, self.table_size = K.level_hash(self.rows, self.m)
        return self.table, self.table_size


def mine_levels(
    bv,
    n_words: int,
    n_rows: int,
    col_counts: np.ndarray,
    min_count: int,
    max_length: int | None = None,
    prune_non_free: bool = False,
    infer_counts: bool = False,
    level_callback=None,
    pitch: int | None = None,
    keep_device: bool = False,
    timer=None,
    k2_kernel: str = "auto",
    start: tuple | None = None,
    guard=None,
    level_sink=None,
) -> MineResult:
    """Mines every frequent itemset (or free-set) of device bitvectors ``bv`` (n_cols, pitch).

    ``col_counts`` are the K=1 counts of the columns (host int64). Counts are exact integers throughout.
    With ``keep_device`` the levels stay device tensors (rows int32, counts int32) instead of host arrays.
    ``timer`` (optional) brackets each level's counting launches: ``timer.begin(k)`` / ``timer.end(k, stats)``
    are called on the host around them, on the stream they run on. ``start`` = (rows int32 (n, k0) of column
    ids, lexsorted; counts int64) resumes from a level a previous run wrote: mining continues at k0 + 1 and the
    result holds the levels above k0 only. The loaded level's free flags are unknown, so its first successor
    runs without inference (``gpu/row_split.py:594-601``). ``guard(k)`` is called after level k when another
    level follows (the reference's memory guard, ``gpu/row_split.py:1143-1144``); ``level_sink(k, rows, counts)``
    receives each level (host arrays) as it completes, and the result then keeps no levels.
    """
    if n_rows > np.iinfo(np.int32).max:
        raise ValueError(f"n_transactions={n_rows:,} exceeds int32 max: the count arrays are int32")
    out = MineResult()
    n_cols = len(col_counts)
    if start is not None:
        return _mine_from(bv, n_words, n_rows, n_cols, min_count, max_length, prune_non_free, infer_counts,
                          level_callback, pitch, keep_device, timer, k2_kernel, start, out, guard, level_sink)
    t0 = time.perf_counter()
    freq = np.flatnonzero(col_counts >= min_count).astype(np.int32)
    free1 = col_counts[freq] < n_rows
    gen_cols = freq[free1] if prune_non_free else freq
    k1_rows = gen_cols.reshape(-1, 1)
    k1_counts = col_counts[gen_cols].astype(np.int64)
    if level_sink is None:
        out.levels.append((k1_rows, k1_counts))
    else:
        level_sink(1, k1_rows, k1_counts)
    out.stats.append(LevelStats(1, n_cols, len(gen_cols), n_counted=n_cols, ms=(time.perf_counter() - t0) * 1000))
    if level_callback:
        level_callback(1, n_cols, len(gen_cols), out.stats[-1].ms)
    if len(gen_cols) == 0:
        return out
    eml = min(max_length if max_length else math.inf, n_cols)
    if guard is not None and eml >= 2 and len(gen_cols) >= 2:
        guard(1)
    track_free = prune_non_free or infer_counts
    prev = _Level(dev.to_device(k1_rows), dev.to_device(k1_counts.astype(np.int32)))
    if track_free:
        prev.free = dev.to_device((k1_counts < n_rows).astype(np.uint8))
    return _level_loop(bv, n_words, n_rows, min_count, eml, prune_non_free, infer_counts, level_callback,
                       pitch, keep_device, timer, k2_kernel, prev, 2, [], out,
                       first_host=out.levels[0] if out.levels else None, guard=guard, level_sink=level_sink)


def _mine_from(bv, n_words, n_rows, n_cols, min_count, max_length, prune_non_free, infer_counts, level_callback,
               pitch, keep_device, timer, k2_kernel, start, out, guard=None, level_sink=None) -> MineResult:
    rows0, counts0 = start
    rows0 = np.ascontiguousarray(rows0, dtype=np.int32)
    if rows0.ndim != 2 or rows0.shape[1] < 1 or len(rows0) != len(counts0):
        raise ValueError("start must be (rows (n, k0) of column ids, counts (n,))")
    k0 = rows0.shape[1]
    eml = min(max_length if max_length else math.inf, n_cols)
    if len(rows0) == 0 or k0 + 1 > eml:
        return out
    prev = _Level(dev.to_device(rows0), dev.to_device(np.asarray(counts0, dtype=np.int32)))
    return _level_loop(bv, n_words, n_rows, min_count, eml, prune_non_free, infer_counts, level_callback, pitch,
                       keep_device, timer, k2_kernel, prev, k0 + 1, [], out, first_host=None, guard=guard,
                       level_sink=level_sink)


def _level_loop(bv, n_words, n_rows, min_count, eml, prune_non_free, infer_counts, level_callback, pitch,
                keep_device, timer, k2_kernel, prev, k, older, out, first_host, guard=None,
                level_sink=None) -> MineResult:
    """Levels k, k+1, ... from ``prev`` (level k-1 on the device); ``older`` holds levels k-3, k-2 when known."""
    track_free = prune_non_free or infer_counts
    dev_levels = []
    deduction = _lib.deduction_setting() == _lib.DEDUCTION_AUTO
    while k <= eml and prev.n >= k:
        t_k = time.perf_counter()
        m = k - 1
        groups = K.level_groups(prev.rows, m)
        if groups.n_cands == 0:
            break
        tally = gid = None
        if k >= 3:
            table, tsize = prev.hashed()
            infer = infer_counts and not prune_non_free and prev.free is not None
            mode = K.SUBSET_PRUNE | (K.SUBSET_INFER if infer else 0)
            tally = dev.zeros(3, "int64")
            # Both classification forms are exercised by this operator-level loop: group-centric on odd levels,
            # one thread per candidate (binary search for the group) on even ones.
            status, counts, gid = K.level_classify(prev.rows, m, prev.counts, prev.free if infer else None, table,
                                                   tsize, groups, mode, tally, group_centric=(k % 2 == 1))
            if infer and deduction and k <= _lib.BOUNDS_MAX_K and len(older) >= (1 if k == 3 else 2):
                older3 = None
                if k > 3:
                    older3 = (older[-2].rows, older[-2].counts) + older[-2].hashed()
                K.level_deduce((prev.rows, prev.counts, table, tsize),
                               (older[-1].rows, older[-1].counts) + older[-1].hashed(), older3, n_rows, min_count,
                               groups, status, counts, gid=gid, tally=tally)
        else:
            status, counts = None, dev.zeros(groups.n_cands, "int32")
        st = LevelStats(k, groups.n_cands, 0, groups.n_cands, 0, 0,
                        ((m - 1) * groups.n_groups2 + groups.rows2) * n_words * 8 + 4 * groups.n_cands)
        if timer is not None:
            timer.begin(k)
        K.count_level(bv, n_words, prev.rows, m, groups, status, counts, pitch=pitch, kernel=k2_kernel)
        if timer is not None:
            timer.end(k, st)
        rows_k, counts_k = K.level_compact(prev.rows, m, groups, status, counts, min_count, gid=gid)
        if tally is not None:
            st.n_counted, st.n_inferred, st.n_skipped = (int(x) for x in dev.to_host(tally))
        cur = _Level(rows_k, counts_k)
        if track_free and cur.n:
            table, tsize = prev.hashed()
            cur.free = K.level_free(cur.rows, k, cur.counts, prev.rows, prev.counts, table, tsize)
            if prune_non_free:
                cur.rows, cur.counts = K.select_rows(cur.rows, k, cur.counts, cur.free)
                cur.free = dev.empty(cur.n, "uint8").fill_(1) if cur.n else None
        n_freq = cur.n
        if level_sink is None:
            dev_levels.append((cur.rows, cur.counts))
        else:
            level_sink(k, dev.to_host(cur.rows), dev.to_host(cur.counts).astype(np.int64))
        dt = (time.perf_counter() - t_k) * 1000
        st.n_frequent, st.ms = n_freq, dt
        out.stats.append(st)
        if level_callback:
            level_callback(k, groups.n_cands, n_freq, dt)
        if n_freq == 0:
            break
        if guard is not None and k + 1 <= eml and n_freq >= k + 1:
            guard(k)
        older = (older + [prev])[-2:]
        prev = cur
        k += 1
    if keep_device:
        head = [(dev.to_device(first_host[0]), dev.to_device(first_host[1]))] if first_host is not None else []
        out.levels = head + dev_levels
        return out
    for rows_d, counts_d in dev_levels:
        if rows_d.shape[0]:
            out.levels.append((dev.to_host(rows_d), dev.to_host(counts_d).astype(np.int64)))
    return out
