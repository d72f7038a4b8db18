"""The allocation-free compaction (etm_level_compact_noalloc) against etm_level_compact, through the C ABI.

Both entry points filter a level's candidates by status and count and write the survivors in candidate order;
the allocation-free one keeps its scan in a static device array instead of a workspace argument. Its output
must equal the two-call entry point's bit for bit, on two calls in a row (the second on a smaller level), with
a capacity below the survivor count, and it must reject sizes beyond its scratch's limit before any launch.
"""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _level(rng, n_rows: int, m: int, n_cols: int) -> np.ndarray:
    rows = set()
    while len(rows) < n_rows:
        rows.add(tuple(sorted(rng.choice(n_cols, size=m, replace=False).tolist())))
    return np.array(sorted(rows), dtype=np.int32).reshape(-1, m)


def _inputs(seed: int, n_rows: int, m: int, n_cols: int):
    from et_miner.gpu import device as dev
    from et_miner.gpu import kernels as K

    rng = np.random.default_rng(seed)
    rows = dev.to_device(_level(rng, n_rows, m, n_cols))
    groups = K.level_groups(rows, m)
    status = dev.to_device(rng.integers(0, 3, size=groups.n_cands).astype(np.uint8))
    counts = dev.to_device(rng.integers(0, 100, size=groups.n_cands).astype(np.int32))
    return rows, groups, status, counts


def _noalloc(rows, m, groups, status, counts, min_count, cap):
    from et_miner import _lib
    from et_miner.gpu import device as dev

    out_rows = dev.empty((max(cap, 1), m + 1), "int32")
    out_counts = dev.empty(max(cap, 1), "int32")
    out_cand = dev.empty(max(cap, 1), "int64")
    n_out = dev.empty(1, "int64")
    _lib.call("etm_level_compact_noalloc", rows.data_ptr(), m, groups.group_start.data_ptr(),
              groups.cum_pairs.data_ptr(), groups.n_groups, groups.n_cands, status.data_ptr(), counts.data_ptr(),
              None, min_count, out_rows.data_ptr(), out_counts.data_ptr(), out_cand.data_ptr(), cap,
              n_out.data_ptr(), dev.stream_ptr())
    n = int(dev.to_host(n_out)[0])
    k = min(n, cap)
    return n, dev.to_host(out_rows)[:k], dev.to_host(out_counts)[:k], dev.to_host(out_cand)[:k]


def _check(seed, n_rows, m, n_cols, min_count=50):
    from et_miner.gpu import device as dev
    from et_miner.gpu import kernels as K

    rows, groups, status, counts = _inputs(seed, n_rows, m, n_cols)
    want_rows, want_counts, want_cand = (dev.to_host(x) for x in
                                         K.level_compact(rows, m, groups, status, counts, min_count, with_cand=True))
    print(f"noalloc compaction: seed {seed}, {n_rows} rows of {m}, {groups.n_cands} candidates, "
          f"{len(want_counts)} survivors", flush=True)
    assert len(want_counts) > 0
    n, got_rows, got_counts, got_cand = _noalloc(rows, m, groups, status, counts, min_count, groups.n_cands)
    assert n == len(want_counts)
    assert np.array_equal(got_rows, want_rows)
    assert np.array_equal(got_counts, want_counts)
    assert np.array_equal(got_cand, want_cand)
    return rows, groups, status, counts, (want_rows, want_counts, want_cand)


def test_noalloc_compaction_equals_workspace_compaction():
    # Two calls in a row on one GPU: a level of 6,000 pairs, then a smaller level of 400 triples (the static
    # scratch keeps the first call's block sums; the second must not read them).
    _check(seed=11, n_rows=6000, m=2, n_cols=300)
    rows, groups, status, counts, (w_rows, w_counts, w_cand) = _check(seed=12, n_rows=400, m=3, n_cols=40)
    # A capacity below the survivor count: the true count comes back, and only the first survivors are written.
    cap = len(w_counts) // 2
    n, g_rows, g_counts, g_cand = _noalloc(rows, 3, groups, status, counts, 50, cap)
    assert n == len(w_counts)
    assert np.array_equal(g_rows, w_rows[:cap])
    assert np.array_equal(g_counts, w_counts[:cap])
    assert np.array_equal(g_cand, w_cand[:cap])
    # Counting only.
    n0, *_ = _noalloc(rows, 3, groups, status, counts, 50, 0)
    assert n0 == len(w_counts)


def test_noalloc_compaction_rejects_sizes_beyond_its_scratch():
    from et_miner import _lib
    from et_miner.gpu import device as dev

    limit = int(_lib.lib().etm_level_compact_noalloc_max())
    assert limit == ((1 << 20) // 8 - 1) * 1024
    dummy = dev.empty(16, "int64")
    p = dummy.data_ptr()
    with pytest.raises(ValueError, match="static scratch"):
        _lib.call("etm_level_compact_noalloc", p, 2, p, p, 1, limit + 1, None, p, None, 1, p, p, None, 1, p,
                  dev.stream_ptr())
