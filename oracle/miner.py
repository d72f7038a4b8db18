"""CPU oracle for the exact frequent-itemset counting path (TEST INFRASTRUCTURE ONLY).

This module restates, in numpy/scipy, the reference's CPU route -- the
answer every route of the reference must reproduce bit for bit. Only the
test suite, the smoke check and the benchmark's CPU-baseline leg may import
it, and only as the checker: the product never calls it.

Functions and the reference code each one follows (paths relative to the
reference's ``src/et_miner``):

* ``min_count`` / ``mining_min_count``: ``core/result.py:57-79`` and
  ``core/cpu_miner.py:967-970`` -- ceil(Fraction(str(s)) * N), at least 1 when
  mining.
* ``transaction_csr``: ``core/cpu_miner.py:222-308`` -- an item is frequent when
  the number of distinct rows holding it reaches min_count; columns are the
  frequent items in ascending id order; rows are deduplicated and sorted.
* ``count_pairs``: ``core/cpu_miner.py:332-378`` -- K=2 from the Gram M^T M.
* ``generate_candidates``: ``core/cpu_miner.py:448-498`` -- prefix join of the
  lexsorted level k-1 plus the test of the k-2 subsets that drop a prefix item.
* ``count_candidates``: ``core/cpu_miner.py:706-760`` -- per prefix group: AND of
  the prefix columns, then AND + popcount per suffix pair.
* ``mine``: ``core/cpu_miner.py:984-1161`` (``mine_cpu``, ``_k1_seed``,
  ``_mine_levels``): levels to exhaustion, Pascal inference
  (``:1119-1124``), free flags (``:943-945``, ``:1139-1151``), free-set mode.
* ``count_itemsets``: ``core/cpu_miner.py:838-896`` -- counts of given itemsets.
* ``mine_son``: ``streaming/son.py:324-421``, ``:570-693`` -- SON's two passes,
  the candidate union with known counts and slack, and the bound.
* ``deduction_outcome``: ``gpu/kernels/_src/_subset_index.cu:95-178`` -- Pascal
  inference, then the depth-3 deduction bounds of one candidate.

Pinned by (tests/test_oracle.py): the reference's committed lattices
(``bench/results/2026-10-17-moonshots-baseline/*.lattice.parquet``) on the
regenerated preset datasets, the reference's ``min_count_cases.json``, the
README's known answer, the boundary case s=0.07, planted-motif lower bounds,
and brute-force subset enumeration on small random inputs; ``deduction_outcome``
by brute-force counts (l <= count <= u for every candidate, an inferred value
equal to the count, a skipped candidate infrequent) and hand-worked cases.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from fractions import Fraction

import numpy as np
from scipy.sparse import csr_matrix

LOCAL_SUPPORT_FACTOR = 0.95  # streaming/son.py:82


def min_count(min_support: float, n_rows: int) -> int:
    """ceil(s * N) with s read as the shortest decimal that round-trips to the float (core/result.py:57-79)."""
    return math.ceil(Fraction(str(min_support)) * n_rows)


def mining_min_count(min_support: float, n_rows: int) -> int:
    """min_count, at least 1 (core/cpu_miner.py:967-970)."""
    return max(1, min_count(min_support, n_rows))


@dataclass
class Lattice:
    """Mined levels: ``items`` are the column ids' item ids; each level is (lexsorted int32 (n, k), int64 counts)."""

    items: np.ndarray
    n_rows: int
    levels: list = field(default_factory=list)

    def as_dict(self) -> dict:
        out = {}
        for sets, counts in self.levels:
            for row, c in zip(self.items[sets].tolist(), counts.tolist()):
                out[tuple(row)] = c
        return out

    def n_itemsets(self) -> int:
        return int(sum(len(c) for _, c in self.levels))


def transaction_csr(indptr: np.ndarray, ids: np.ndarray, mc: int):
    """(col_indptr int64, cols int32, items int64, counts int64) of the frequent items (core/cpu_miner.py:222-308)."""
    n_rows = len(indptr) - 1
    ids = np.asarray(ids).astype(np.int64, copy=False)
    rows = np.repeat(np.arange(n_rows, dtype=np.int64), np.diff(indptr))
    if len(ids) == 0:
        return np.zeros(n_rows + 1, np.int64), np.zeros(0, np.int32), np.zeros(0, np.int64), np.zeros(0, np.int64)
    lo = int(ids.min())
    span = int(ids.max()) - lo + 1
    keys = np.unique(rows * span + (ids - lo))
    r, it = keys // span, keys % span
    counts = np.bincount(it, minlength=span)
    freq = np.flatnonzero(counts >= mc)
    lut = np.full(span, -1, dtype=np.int64)
    lut[freq] = np.arange(len(freq))
    c = lut[it]
    keep = c >= 0
    r, c = r[keep], c[keep]
    col_indptr = np.zeros(n_rows + 1, dtype=np.int64)
    np.cumsum(np.bincount(r, minlength=n_rows), out=col_indptr[1:])
    return col_indptr, c.astype(np.int32), (freq + lo).astype(np.int64), counts[freq].astype(np.int64)


def bitvectors(indptr: np.ndarray, cols: np.ndarray, n_cols: int) -> np.ndarray:
    """(n_cols, ceil(N/64)) uint64: bit r%64 of word r//64 of column c is set when row r holds c."""
    n_rows = len(indptr) - 1
    w = (n_rows + 63) // 64
    out = np.zeros(n_cols * w, dtype=np.uint64)
    rows = np.repeat(np.arange(n_rows, dtype=np.int64), np.diff(indptr))
    word = cols.astype(np.int64) * w + (rows >> 6)
    bit = (rows & 63).astype(np.uint64)
    # Entries of one (word, bit) are unique, so each bit plane is a plain fancy-index OR.
    for b in range(64):
        sel = bit == b
        if sel.any():
            out[word[sel]] |= np.uint64(1) << np.uint64(b)
    return out.reshape(n_cols, w)


def count_pairs(indptr: np.ndarray, cols: np.ndarray, n_cols: int, gen: np.ndarray, mc: int):
    """Pairs of generating columns with count >= mc from G = M^T M (core/cpu_miner.py:332-378).

    Returns (lexsorted int32 (n, 2) with i < j, int64 counts)."""
    if n_cols < 2:
        return np.empty((0, 2), np.int32), np.empty(0, np.int64)
    m = csr_matrix((np.ones(len(cols), dtype=np.int32), cols, indptr), shape=(len(indptr) - 1, n_cols))
    g = (m.T.tocsr() @ m).tocsr()
    rows = np.repeat(np.arange(n_cols, dtype=np.int32), np.diff(g.indptr))
    c = g.indices
    keep = (c > rows) & (g.data >= mc) & gen[rows] & gen[c]
    i, j, cnt = rows[keep], c[keep].astype(np.int32), g.data[keep].astype(np.int64)
    order = np.lexsort((j, i))
    return np.stack([i[order], j[order]], axis=1), cnt[order]


class LevelIndex:
    """Positions of rows in a lexsorted int32 level (packed keys when they fit 62 bits, else void rows)."""

    def __init__(self, level: np.ndarray, base: int):
        self.base = max(base, 2)
        self.k = level.shape[1]
        self.packed = self.k * math.log2(self.base) < 62
        if self.packed:
            self.keys = self._pack(level)
        else:
            self.void = np.dtype((np.void, 4 * self.k))
            v = np.ascontiguousarray(level, dtype=np.int32).view(self.void).ravel()
            self.order = np.argsort(v, kind="stable")
            self.sorted = v[self.order]

    def _pack(self, rows: np.ndarray) -> np.ndarray:
        key = np.zeros(len(rows), dtype=np.int64)
        for c in range(rows.shape[1]):
            key = key * self.base + rows[:, c].astype(np.int64)
        return key

    def find(self, sub: np.ndarray) -> np.ndarray:
        if len(sub) == 0:
            return np.empty(0, dtype=np.int64)
        if self.packed:
            if len(self.keys) == 0:
                return np.full(len(sub), -1, dtype=np.int64)
            key = self._pack(sub)
            pos = np.minimum(np.searchsorted(self.keys, key), len(self.keys) - 1)
            return np.where(self.keys[pos] == key, pos, -1)
        if len(self.sorted) == 0:
            return np.full(len(sub), -1, dtype=np.int64)
        q = np.ascontiguousarray(sub, dtype=np.int32).view(self.void).ravel()
        pos = np.minimum(np.searchsorted(self.sorted, q), len(self.sorted) - 1)
        return np.where(self.sorted[pos] == q, self.order[pos], -1)


def _group_starts(sets: np.ndarray, plen: int) -> np.ndarray:
    n = len(sets)
    if plen == 0:
        return np.array([0], dtype=np.int64) if n else np.empty(0, np.int64)
    change = np.any(sets[1:, :plen] != sets[:-1, :plen], axis=1)
    return np.flatnonzero(np.r_[True, change])


def generate_candidates(prev: np.ndarray, k: int, n_cols: int, chunk: int = 2_000_000):
    """Lexsorted k-candidates from the lexsorted level ``prev`` (core/cpu_miner.py:448-498), in chunks."""
    n = len(prev)
    if n < 2:
        return
    starts = _group_starts(prev, k - 2)
    sizes = np.diff(np.r_[starts, n])
    pos = np.arange(n) - np.repeat(starts, sizes)
    partners = np.repeat(sizes, sizes) - 1 - pos
    cum = np.cumsum(partners)
    index = LevelIndex(prev, n_cols)
    e0 = 0
    while e0 < n:
        before = int(cum[e0 - 1]) if e0 else 0
        e1 = min(max(int(np.searchsorted(cum, before + chunk, side="right")), e0 + 1), n)
        p = partners[e0:e1]
        total = int(p.sum())
        if total:
            left = np.repeat(np.arange(e0, e1, dtype=np.int64), p)
            right = left + 1 + (np.arange(total, dtype=np.int64) - np.repeat(np.cumsum(p) - p, p))
            cands = np.empty((total, k), dtype=np.int32)
            cands[:, : k - 1] = prev[left]
            cands[:, k - 1] = prev[right, k - 2]
            keep = np.ones(total, dtype=bool)
            for drop in range(k - 2):
                idx = np.flatnonzero(keep)
                if len(idx) == 0:
                    break
                keep[idx] = index.find(cands[idx][:, [c for c in range(k) if c != drop]]) >= 0
            out = cands[keep]
            if len(out):
                yield out
        e0 = e1


def popcount_rows(a: np.ndarray) -> np.ndarray:
    return np.bitwise_count(a).sum(axis=1, dtype=np.int64)


def count_candidates(cands: np.ndarray, k: int, bv: np.ndarray, n_jobs: int = 1) -> np.ndarray:
    """Counts of lexsorted k-candidates, one prefix group at a time (core/cpu_miner.py:706-760).

    Per group: AND the prefix's k-2 columns, keep its nonzero words, AND each suffix column with it, then
    popcount each pair's AND. ``n_jobs`` > 1 counts disjoint runs of groups on that many threads (numpy releases
    the GIL inside these array operations), as the reference's CPU route spreads groups over its workers."""
    out = np.empty(len(cands), dtype=np.int64)
    starts = _group_starts(cands, k - 2)
    ends = np.r_[starts[1:], len(cands)]
    if n_jobs > 1 and len(starts) > 1:
        from concurrent.futures import ThreadPoolExecutor

        # Runs of whole groups with about equal candidate counts, one task each (several per thread).
        bounds = np.searchsorted(starts, np.linspace(0, len(cands), 4 * n_jobs + 1)[1:-1].astype(np.int64))
        runs = [r for r in np.split(np.arange(len(starts)), np.unique(bounds)) if len(r)]
        with ThreadPoolExecutor(n_jobs) as pool:
            list(pool.map(lambda r: _count_groups(cands, k, bv, starts[r], ends[r], out), runs))
        return out
    _count_groups(cands, k, bv, starts, ends, out)
    return out


def _count_groups(cands, k, bv, starts, ends, out) -> None:
    for s, e in zip(np.asarray(starts).tolist(), np.asarray(ends).tolist()):
        a, b = cands[s:e, k - 2], cands[s:e, k - 1]
        suffix = np.union1d(a, b)
        ia, ib = np.searchsorted(suffix, a), np.searchsorted(suffix, b)
        if k == 2:
            sub = bv[suffix]
        else:
            pre = bv[cands[s, 0]].copy()
            for it in cands[s, 1 : k - 2]:
                pre &= bv[it]
            nz = np.flatnonzero(pre)
            if len(nz) == 0:
                out[s:e] = 0
                continue
            sub = bv[suffix[:, None], nz] & pre[nz]
        step = max(1, (4 << 20) // (8 * max(1, sub.shape[1])))
        for c0 in range(0, len(ia), step):
            c1 = min(c0 + step, len(ia))
            out[s + c0 : s + c1] = popcount_rows(sub[ia[c0:c1]] & sub[ib[c0:c1]])


def count_per_candidate(cands: np.ndarray, bv: np.ndarray) -> np.ndarray:
    """Counts of itemsets as the popcount of the AND of their columns (core/cpu_miner.py:763-781)."""
    out = np.empty(len(cands), dtype=np.int64)
    step = max(1, (4 << 20) // (8 * max(1, bv.shape[1])))
    for c0 in range(0, len(cands), step):
        c = cands[c0 : c0 + step]
        acc = bv[c[:, 0]].copy()
        for j in range(1, c.shape[1]):
            acc &= bv[c[:, j]]
        out[c0 : c0 + len(c)] = popcount_rows(acc)
    return out


@dataclass
class _Level:
    sets: np.ndarray
    counts: np.ndarray
    free: np.ndarray | None
    index: LevelIndex | None = None

    def subset_positions(self, cands: np.ndarray, base: int, n_jobs: int = 1) -> np.ndarray:
        if self.index is None:
            self.index = LevelIndex(self.sets, base)
        k = cands.shape[1]
        out = np.empty((len(cands), k), dtype=np.int64)

        def fill(lo, hi):
            for j in range(k):
                out[lo:hi, j] = self.index.find(cands[lo:hi][:, [c for c in range(k) if c != j]])

        if n_jobs > 1 and len(cands) >= 4096:
            from concurrent.futures import ThreadPoolExecutor

            cuts = np.linspace(0, len(cands), 4 * n_jobs + 1).astype(np.int64)
            with ThreadPoolExecutor(n_jobs) as pool:
                list(pool.map(lambda i: fill(cuts[i], cuts[i + 1]), range(len(cuts) - 1)))
        else:
            fill(0, len(cands))
        return out


def mine(
    indptr: np.ndarray,
    ids: np.ndarray,
    min_support: float,
    max_length: int | None = None,
    prune_equal_support: bool = False,
    use_generator_pruning: bool = False,
    level_callback=None,
    n_jobs: int = 1,
) -> Lattice:
    """Every frequent itemset (or free-set) of a CSR of item ids, with exact counts (core/cpu_miner.py:984-1161).

    ``indptr`` (N+1) and ``ids`` are the transactions; rows may repeat items. Levels come back lexsorted over
    column ids, which ascend with item ids. ``n_jobs`` threads count the prefix groups of each level (the result
    does not depend on it).
    """
    indptr = np.asarray(indptr, dtype=np.int64)
    n_rows = len(indptr) - 1
    mc = mining_min_count(min_support, n_rows)
    col_ptr, cols, items, counts1 = transaction_csr(indptr, ids, mc)
    lat = Lattice(items, n_rows)
    n_cols = len(items)
    if n_cols == 0:
        return lat
    max_tx = int(np.diff(indptr).max()) if n_rows else 0
    eml = min(max_length if max_length else math.inf, max_tx, n_cols)
    ones = np.arange(n_cols, dtype=np.int32)[:, None]
    free1 = counts1 < n_rows
    gen = ones[free1] if prune_equal_support else ones
    lat.levels.append((gen, counts1[free1] if prune_equal_support else counts1))
    if level_callback:
        level_callback(1, n_cols, len(gen))
    prev = _Level(ones, counts1, free1)
    track_free = prune_equal_support or use_generator_pruning
    bv = None
    k = 2
    while k <= eml and len(gen) >= k:
        if k == 2:
            gmask = np.zeros(n_cols, dtype=bool)
            gmask[gen[:, 0]] = True
            sets, counts = count_pairs(col_ptr, cols, n_cols, gmask, mc)
            n_cand = len(gen) * (len(gen) - 1) // 2
        else:
            if bv is None:
                bv = bitvectors(col_ptr, cols, n_cols)
            infer = use_generator_pruning and not prune_equal_support
            parts_s, parts_c = [], []
            n_cand = 0
            for cands in generate_candidates(gen, k, n_cols):
                n_cand += len(cands)
                cc = np.empty(len(cands), dtype=np.int64)
                needs = np.ones(len(cands), dtype=bool)
                if infer:
                    pos = prev.subset_positions(cands, n_cols, n_jobs)
                    licensed = np.any(~prev.free[pos], axis=1)
                    cc[licensed] = prev.counts[pos[licensed]].min(axis=1)
                    needs = ~licensed
                if needs.any():
                    cc[needs] = count_candidates(cands[needs], k, bv, n_jobs)
                keep = cc >= mc
                parts_s.append(cands[keep])
                parts_c.append(cc[keep])
            if n_cand == 0:
                break
            sets = np.concatenate(parts_s) if parts_s else np.empty((0, k), np.int32)
            counts = np.concatenate(parts_c) if parts_c else np.empty(0, np.int64)
        level_free = None
        if track_free:
            if len(sets):
                pos = prev.subset_positions(sets, n_cols, n_jobs)
                level_free = ~np.any(prev.counts[pos] == counts[:, None], axis=1)
            else:
                level_free = np.zeros(0, dtype=bool)
        if prune_equal_support:
            gen = sets[level_free]
            lat.levels.append((gen, counts[level_free]))
        else:
            gen = sets
            lat.levels.append((sets, counts))
        if level_callback:
            level_callback(k, n_cand, len(gen))
        prev = _Level(sets, counts, level_free)
        if len(gen) == 0:
            break
        k += 1
    return lat


def count_itemsets(indptr: np.ndarray, cols: np.ndarray, n_cols: int, itemsets: dict) -> dict:
    """Counts of given itemsets (column ids, per length) over every row of a CSR (core/cpu_miner.py:838-896)."""
    bv = bitvectors(np.asarray(indptr, np.int64), np.asarray(cols), n_cols)
    out = {}
    for k, c in itemsets.items():
        if len(c) == 0:
            out[k] = np.zeros(0, np.int64)
        elif k == 1:
            out[k] = popcount_rows(bv[c[:, 0]])
        else:
            out[k] = count_per_candidate(c, bv)
    return out


def sum_rows(sets: np.ndarray, weights: list) -> tuple:
    """Lexsorted distinct rows with each weight summed per distinct row (core/cpu_miner.py:396-414)."""
    if len(sets) == 0:
        return sets.astype(np.int32), [w[:0] for w in weights]
    order = np.lexsort(sets.T[::-1])
    s = sets[order]
    new = np.r_[True, (s[1:] != s[:-1]).any(axis=1)]
    starts = np.flatnonzero(new)
    return s[starts].astype(np.int32), [np.add.reduceat(w[order], starts) for w in weights]


def mine_son(indptr: np.ndarray, ids: np.ndarray, min_support: float, chunk_size: int,
             max_length: int | None = None, local_support_factor: float = LOCAL_SUPPORT_FACTOR) -> Lattice:
    """SON's two passes over row chunks (streaming/son.py:570-693); the result equals ``mine``'s.

    Pass 1 mines each chunk at ``local_support_factor * min_support``; the union sums each itemset's local
    counts (``known``) and the slack (local min_count - 1) of the chunks that emitted it. The bound
    known + total_slack - slack drops candidates that cannot reach min_count and is exact where no slack is
    left; pass 2 counts the rest over every chunk.
    """
    indptr = np.asarray(indptr, dtype=np.int64)
    ids = np.asarray(ids).astype(np.int64, copy=False)
    n_total = len(indptr) - 1
    if n_total <= chunk_size:
        return mine(indptr, ids, min_support, max_length)
    local_s = min_support * local_support_factor
    bounds = [(a, min(a + chunk_size, n_total)) for a in range(0, n_total, chunk_size)]
    slacks = [mining_min_count(local_s, b - a) - 1 for a, b in bounds]
    mc = min_count(min_support, n_total)
    if sum(slacks) >= mc:
        raise RuntimeError(f"SON's slack {sum(slacks)} reaches min_count {mc}")
    rows_by_k: dict = {}
    for (a, b), slack in zip(bounds, slacks):
        sub_ptr = indptr[a : b + 1] - indptr[a]
        lat = mine(sub_ptr, ids[indptr[a] : indptr[b]], local_s, max_length)
        for sets, counts in lat.levels:
            if len(sets):
                rows_by_k.setdefault(sets.shape[1], []).append(
                    (np.sort(lat.items[sets], axis=1), counts.astype(np.int64), slack))
    if not rows_by_k:
        return Lattice(np.zeros(0, np.int64), n_total)
    all_items = np.unique(np.concatenate([np.concatenate([p[0].ravel() for p in v]) for v in rows_by_k.values()]))
    total_slack = sum(slacks)
    bounded = {}
    for k, parts in sorted(rows_by_k.items()):
        rows = np.searchsorted(all_items, np.concatenate([p[0] for p in parts])).astype(np.int32)
        known = np.concatenate([p[1] for p in parts])
        sl = np.concatenate([np.full(len(p[0]), p[2], dtype=np.int64) for p in parts])
        rows, (known, sl) = sum_rows(rows, [known, sl])
        left = total_slack - sl
        keep = known + left >= mc
        need = left[keep] > 0
        bounded[k] = (rows[keep], np.where(need, 0, known[keep]).astype(np.int64), need)
    cands = {k: r[need] for k, (r, _, need) in bounded.items() if need.any()}
    totals = {k: np.zeros(len(c), np.int64) for k, c in cands.items()}
    if cands:
        lut_lo = int(all_items.min())
        lut = np.full(int(all_items.max()) - lut_lo + 1, -1, np.int64)
        lut[all_items - lut_lo] = np.arange(len(all_items))
        for a, b in bounds:
            seg = ids[indptr[a] : indptr[b]]
            r = np.repeat(np.arange(b - a, dtype=np.int64), np.diff(indptr[a : b + 1]))
            inside = (seg >= lut_lo) & (seg <= int(all_items.max()))
            c = np.full(len(seg), -1, np.int64)
            c[inside] = lut[seg[inside] - lut_lo]
            keep = c >= 0
            keys = np.unique(r[keep] * len(all_items) + c[keep])
            ptr = np.zeros(b - a + 1, np.int64)
            np.cumsum(np.bincount(keys // len(all_items), minlength=b - a), out=ptr[1:])
            got = count_itemsets(ptr, (keys % len(all_items)).astype(np.int32), len(all_items), cands)
            for k, v in got.items():
                totals[k] += v
    lat = Lattice(all_items.astype(np.int64), n_total)
    for k, (rows, counts, need) in bounded.items():
        if k in totals:
            counts[need] = totals[k]
        keep = counts >= mc
        lat.levels.append((rows[keep], counts[keep]))
    return lat


def brute_force(indptr: np.ndarray, ids: np.ndarray, min_support: float, max_length: int | None = None) -> dict:
    """Every frequent itemset by enumerating the subsets of each row's frequent items (small inputs only)."""
    from itertools import combinations

    n_rows = len(indptr) - 1
    mc = mining_min_count(min_support, n_rows)
    rows = [sorted(set(ids[indptr[r] : indptr[r + 1]].tolist())) for r in range(n_rows)]
    single: dict = {}
    for row in rows:
        for it in row:
            single[it] = single.get(it, 0) + 1
    freq = {it for it, c in single.items() if c >= mc}
    counts: dict = {}
    top = max_length or 10**9
    for row in rows:
        fr = [it for it in row if it in freq]
        for r in range(1, min(len(fr), top) + 1):
            for sub in combinations(fr, r):
                counts[sub] = counts.get(sub, 0) + 1
    return {s: c for s, c in counts.items() if c >= mc}


def _drop(x: tuple, drop: set, support: dict, n_total: int) -> int:
    sub = tuple(v for t, v in enumerate(x) if t not in drop)
    return support[sub] if sub else n_total


def deduction_bounds(x: tuple, support: dict, n_total: int) -> tuple:
    """(l, u): the depth-3 bounds of ``gpu/kernels/_src/_subset_index.cu:95-136`` on the count of ``x``.

    u = min(min_a s(x-a), min_{a<b<c} s(x-a)+s(x-b)+s(x-c)-s(x-ab)-s(x-ac)-s(x-bc)+s(x-abc)),
    l = max(0, max_{a<b} s(x-a)+s(x-b)-s(x-ab)); ``support`` holds the subsets of lengths k-1..k-3.
    """
    from itertools import combinations

    k = len(x)
    c1 = [_drop(x, {t}, support, n_total) for t in range(k)]
    upper = min(c1)
    lower = 0
    for a, b in combinations(range(k), 2):
        lower = max(lower, c1[a] + c1[b] - _drop(x, {a, b}, support, n_total))
    for a, b, c in combinations(range(k), 3):
        s2 = (_drop(x, {a, b}, support, n_total) + _drop(x, {a, c}, support, n_total)
              + _drop(x, {b, c}, support, n_total))
        upper = min(upper, c1[a] + c1[b] + c1[c] - s2 + _drop(x, {a, b, c}, support, n_total))
    return lower, upper


def deduction_outcome(x: tuple, support: dict, n_total: int, mc: int) -> tuple:
    """('pascal' | 'skip' | 'infer' | 'count', value) of candidate ``x`` (ascending, every (k-1)-subset frequent).

    ``support`` maps the subsets of x of lengths k-1, k-2 and k-3 to their counts (the empty set is ``n_total``).
    Follows ``gpu/kernels/_src/_subset_index.cu:139-178``: a non-free (k-1)-subset gives Pascal inference (the
    minimum of the subsets' counts); otherwise the bounds decide -- u < mc skips, l == u infers u.
    """
    k = len(x)
    c1 = [_drop(x, {t}, support, n_total) for t in range(k)]
    for t in range(k):
        # a (k-1)-subset is free when its count differs from each of its own (k-2)-subsets'
        if any(c1[t] == _drop(x, {t, u}, support, n_total) for u in range(k) if u != t):
            return "pascal", min(c1)
    lower, upper = deduction_bounds(x, support, n_total)
    if upper < mc:
        return "skip", 0
    if lower == upper:
        return "infer", upper
    return "count", 0
