"""Deduction bounds (GPU): etm_level_deduce against the oracle's per-candidate outcome, and mining with the
bounds on and off against the oracle.

The per-candidate test follows the reference's tests/test_deduction.py:93-146: levels k-1, k-2, k-3 of a dense
seven-item lattice, every candidate of level k classified (subset test + Pascal inference), then bounded. Every
candidate's status and initial count must equal the oracle's ``deduction_outcome`` (Pascal -> inferred
minimum; bounds -> skipped or inferred; otherwise counted). The mining test follows :149-186 (ties at min_count,
empty rows, an item in every row) on both level loops.
"""

from __future__ import annotations

import itertools
from collections import Counter

import numpy as np
import pytest
from conftest import dense_rows, lattice_counts

from oracle import miner

pytestmark = pytest.mark.gpu


def _level(lat, m):
    items = sorted(s for s in lat if len(s) == m)
    return (np.array(items, dtype=np.int32).reshape(-1, m), np.array([lat[s] for s in items], dtype=np.int32))


def _expected(prev_rows, lat, n_total, mc):
    """Candidates of level k in the library's numbering (prefix groups in order, suffix pairs i-major), with
    the oracle's (status, count)."""
    m = prev_rows.shape[1]
    prev_set = {tuple(r) for r in prev_rows.tolist()}
    out = []
    start = 0
    n = len(prev_rows)
    while start < n:
        end = start + 1
        while end < n and np.array_equal(prev_rows[end, : m - 1], prev_rows[start, : m - 1]):
            end += 1
        pre = tuple(prev_rows[start, : m - 1].tolist())
        suf = prev_rows[start:end, m - 1].tolist()
        for i in range(len(suf)):
            for j in range(i + 1, len(suf)):
                x = pre + (suf[i], suf[j])
                if not all(sub in prev_set for sub in itertools.combinations(x, len(x) - 1)):
                    out.append((x, 1, -1, "subset"))  # ETM_CAND_SKIP: a subset is infrequent
                    continue
                kind, value = miner.deduction_outcome(x, lat, n_total, mc)
                if kind in ("pascal", "infer"):
                    out.append((x, 2, value, kind))  # ETM_CAND_INFER
                elif kind == "skip":
                    out.append((x, 1, -1, kind))
                else:
                    out.append((x, 0, 0, kind))  # ETM_CAND_COUNT
        start = end
    return out


CASES = [(3, 5, 40), (3, 10, 40), (3, 53, 40), (4, 10, 40), (4, 17, 40), (4, 53, 40), (5, 53, 15), (6, 17, 10)]


@pytest.mark.parametrize("group_centric", [True, False], ids=["gid", "search"])
@pytest.mark.parametrize("k,seed,mc", CASES, ids=[f"k{c[0]}-s{c[1]}" for c in CASES])
def test_deduce_matches_the_oracle(k, seed, mc, group_centric):
    from et_miner.gpu import device as dev
    from et_miner.gpu import kernels as K

    rows = dense_rows(seed, 300)
    n_total = len(rows)
    lat = lattice_counts(rows, mc, k)
    prev, prev_counts = _level(lat, k - 1)
    free = np.array([all(lat[tuple(r)] != lat.get(tuple(v for v in r if v != b), n_total) for b in r)
                     for r in prev.tolist()], dtype=np.uint8)
    d_prev, d_pc, d_free = dev.to_device(prev), dev.to_device(prev_counts), dev.to_device(free)
    groups = K.level_groups(d_prev, k - 1)
    table, tsize = K.level_hash(d_prev, k - 1)
    tally = dev.zeros(3, "int64")
    status, counts, gid = K.level_classify(d_prev, k - 1, d_pc, d_free, table, tsize, groups,
                                           K.SUBSET_PRUNE | K.SUBSET_INFER, tally, group_centric=group_centric)
    r2, c2 = _level(lat, k - 2)
    d_r2, d_c2 = dev.to_device(r2), dev.to_device(c2)
    older2 = (d_r2, d_c2) + K.level_hash(d_r2, k - 2)
    older3 = None
    if k > 3:
        r3, c3 = _level(lat, k - 3)
        d_r3, d_c3 = dev.to_device(r3), dev.to_device(c3)
        older3 = (d_r3, d_c3) + K.level_hash(d_r3, k - 3)
    K.level_deduce((d_prev, d_pc, table, tsize), older2, older3, n_total, mc, groups, status, counts, gid=gid,
                   tally=tally)
    got_s, got_c = dev.to_host(status), dev.to_host(counts)
    want = _expected(prev, lat, n_total, mc)
    assert len(want) == groups.n_cands
    by_status, by_kind = Counter(), Counter()
    for c, (x, st, val, kind) in enumerate(want):
        assert (int(got_s[c]), int(got_c[c])) == (st, val), (x, kind, int(got_s[c]), int(got_c[c]), st, val)
        by_status[st] += 1
        by_kind[kind] += 1
    assert dev.to_host(tally).tolist() == [by_status[0], by_status[2], by_status[1]]
    # the bounds settled candidates that the subset test and Pascal inference leave counted
    assert by_kind["skip"] + by_kind["infer"] > 0, by_kind


def _mine(rows, s, engine, deduction, monkeypatch):
    from et_miner.core.apriori import mine_csr
    from et_miner.core.result import as_dict

    monkeypatch.setenv("ET_MINER_DEDUCTION", deduction)
    indptr = np.zeros(len(rows) + 1, dtype=np.int64)
    indptr[1:] = np.cumsum([len(r) for r in rows])
    ids = np.array([x for r in rows for x in r], dtype=np.int64)
    stats = []
    got = as_dict(mine_csr(indptr, ids, s, None, use_generator_pruning=True, with_counts=True, engine=engine,
                           stats_out=stats))
    return got, stats, indptr, ids


@pytest.mark.parametrize("engine", ["native", "python"])
@pytest.mark.parametrize("seed", [10, 53])
def test_mining_with_the_bounds_matches_the_oracle(seed, engine, monkeypatch):
    """Ties at min_count (the median 3-itemset count), every ninth row empty, item 99 in every non-empty row (all
    ones in its bitvector words)."""
    from et_miner.core.result import _min_count

    rows = [list(r) + [99] for r in dense_rows(seed, 1_000 + 37 * seed)]
    for i in range(0, len(rows), 9):
        rows[i] = []
    counts = sorted(Counter(p for r in rows for p in itertools.combinations(sorted(r), 3)).values())
    c = counts[len(counts) // 2]
    s = c / len(rows)
    while _min_count(s, len(rows)) > c:
        s = float(np.nextafter(s, 0.0))
    on, st_on, indptr, ids = _mine(rows, s, engine, "1", monkeypatch)
    off, st_off, _, _ = _mine(rows, s, engine, "0", monkeypatch)
    want = miner.mine(indptr, ids, s, None, use_generator_pruning=True).as_dict()
    assert on == off == want
    counted_on = sum(st.n_counted for st in st_on if st.k >= 3)
    counted_off = sum(st.n_counted for st in st_off if st.k >= 3)
    assert counted_on < counted_off, (counted_on, counted_off)
    for a, b in zip(st_on, st_off):
        assert a.n_candidates == b.n_candidates
        assert a.n_counted + a.n_inferred + a.n_skipped == b.n_counted + b.n_inferred + b.n_skipped
