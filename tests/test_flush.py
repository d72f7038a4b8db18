"""Per-level parquet output and resume (host side): the reference's on-disk format (``io/flush.py``,
``gpu/row_split.py:539-606``), round trips, partitioned levels, and the errors a foreign file raises."""

from __future__ import annotations

import numpy as np
import pyarrow as pa
import pyarrow.parquet as pq
import pytest


def _table(sets, counts):
    offsets = np.concatenate([[0], np.cumsum([len(s) for s in sets])]).astype(np.int64)
    values = np.array([x for s in sets for x in s], dtype=np.int64)
    return pa.table({"itemset": pa.LargeListArray.from_arrays(pa.array(offsets), pa.array(values)),
                     "count": pa.array(np.array(counts, dtype=np.int64))})


def test_write_levels_format_and_round_trip(tmp_path):
    from et_miner.io import flush

    sets = [(3,), (7,), (3, 7), (3, 9), (3, 7, 9)]
    counts = [9, 8, 6, 5, 4]
    paths = flush.write_levels(_table(sets, counts), str(tmp_path), n_rows=10)
    assert [p.split("/")[-1] for p in paths] == ["frequent_k1.parquet", "frequent_k2.parquet",
                                                  "frequent_k3.parquet"]
    t2 = pq.read_table(tmp_path / "frequent_k2.parquet")
    assert t2.schema.field("itemset").type == pa.large_list(pa.int32())
    assert t2.schema.field("support").type == pa.float64()
    assert t2.column("itemset").to_pylist() == [[3, 7], [3, 9]]
    assert t2.column("support").to_pylist() == [0.6, 0.5]
    rows, got = flush.read_level(str(tmp_path), 2, n_rows=10)
    assert rows.tolist() == [[3, 7], [3, 9]] and got.tolist() == [6, 5]
    # min_k skips the levels a resumed run already holds
    (tmp_path / "again").mkdir()
    assert [p.split("/")[-1] for p in flush.write_levels(_table(sets, counts), str(tmp_path / "again"), 10,
                                                         min_k=3)] == ["frequent_k3.parquet"]


def test_counts_survive_float64_support_exactly(tmp_path):
    """count -> count / N -> rint(support * N) is exact for int32-range counts (spot checks at the edges)."""
    from et_miner.io import flush

    n = 2_000_000_000
    counts = [1, 3, n // 3, n - 1, 2**31 - 1 if 2**31 - 1 <= n else n]
    sets = [(i, i + 1) for i in range(len(counts))]
    flush.write_levels(_table(sets, counts), str(tmp_path), n_rows=n)
    _, got = flush.read_level(str(tmp_path), 2, n_rows=n)
    assert got.tolist() == counts


def test_wide_ids_are_written_as_int64(tmp_path):
    from et_miner.io import flush

    flush.write_levels(_table([(1, 2**40)], [3]), str(tmp_path), n_rows=4)
    t = pq.read_table(tmp_path / "frequent_k2.parquet")
    assert t.schema.field("itemset").type == pa.large_list(pa.int64())
    assert flush.read_level(str(tmp_path), 2, 4)[0].tolist() == [[1, 2**40]]


def test_partitioned_level_and_missing_level(tmp_path):
    from et_miner.io import flush

    part = tmp_path / "frequent_k2"
    part.mkdir()
    for i, (sets, counts) in enumerate([([(1, 2)], [4]), ([(1, 3), (2, 3)], [3, 2])]):
        t = _table(sets, counts)
        pq.write_table(pa.table({"itemset": t.column("itemset"),
                                 "support": pa.array(np.array(counts) / 8.0)}), part / f"part_{i:03d}.parquet")
    rows, counts = flush.read_level(str(tmp_path), 2, n_rows=8)
    assert rows.tolist() == [[1, 2], [1, 3], [2, 3]] and counts.tolist() == [4, 3, 2]
    with pytest.raises(FileNotFoundError, match="resume_from_k=3"):
        flush.read_level(str(tmp_path), 3, n_rows=8)
    with pytest.raises(ValueError, match="remote"):
        flush.write_levels(_table([(1,)], [1]), "gs://bucket/x", 1)


def test_columns_of_maps_sorts_and_rejects_foreign_items():
    from et_miner.io import flush

    col_to_item = np.array([10, 30, 20], dtype=np.int64)  # column 1 holds item 30
    cols = flush.columns_of(np.array([[10, 30], [20, 30]], dtype=np.int64), col_to_item)
    assert cols.tolist() == [[0, 1], [1, 2]]
    rows, counts = flush.lexsorted(np.array([[1, 2], [0, 1]], dtype=np.int32), np.array([5, 6]))
    assert rows.tolist() == [[0, 1], [1, 2]] and counts.tolist() == [6, 5]
    with pytest.raises(ValueError, match="not frequent items"):
        flush.columns_of(np.array([[10, 40]], dtype=np.int64), col_to_item)


def test_output_dir_route_validation():
    from et_miner import apriori

    with pytest.raises(ValueError, match="requires the level-loop miner"):
        apriori([[1, 2]], 0.5, use_gpu=True, streaming=True, output_dir="/tmp/x")


def test_level_hooks_write_then_report_and_guard(tmp_path):
    """The per-level hook of the native loop, driven with host reports: each level is written before the level
    callback runs, an empty level writes no file, and the memory guards raise the reference's MemoryError only
    when another level follows."""
    from et_miner.core.apriori import _LevelHooks
    from et_miner.gpu.native import LevelReport
    from et_miner.gpu.row_split import LevelStats

    seen = []
    hooks = _LevelHooks(lambda k, c, f, ms: seen.append((k, f, ms, sorted(p.name for p in tmp_path.iterdir()))),
                        None, None, None, str(tmp_path), n_rows=10, col_to_item=np.array([7, 8, 9], np.int64))
    on_level = hooks.native()
    on_level(LevelReport(LevelStats(1, 3, 3, ms=1.5), np.array([[0], [1], [2]], np.int32),
                         np.array([5, 6, 7], np.int64), 0, 0, True))
    on_level(LevelReport(LevelStats(2, 3, 1, ms=2.5), np.array([[0, 2]], np.int32), np.array([4], np.int64), 0, 0,
                         False))
    on_level(LevelReport(LevelStats(3, 0, 0, ms=0.5), np.empty((0, 3), np.int32), np.empty(0, np.int64), 0, 0,
                         False))
    assert seen == [(1, 3, 1.5, ["frequent_k1.parquet"]),
                    (2, 1, 2.5, ["frequent_k1.parquet", "frequent_k2.parquet"]),
                    (3, 0, 0.5, ["frequent_k1.parquet", "frequent_k2.parquet"])]
    t = pq.read_table(tmp_path / "frequent_k2.parquet").to_pydict()
    assert t["itemset"] == [[7, 9]] and t["support"] == [0.4]

    tight = _LevelHooks(None, None, 1e-9, None, None, n_rows=10)
    report = LevelReport(LevelStats(2, 3, 3), None, None, 0, 0, True)
    with pytest.raises(MemoryError, match=r"Memory guard tripped after K=2: RAM=.*max_ram_gb=1e-09GB"):
        tight.native()(report)
    tight.native()(LevelReport(LevelStats(2, 3, 3), None, None, 0, 0, False))  # the last level: no guard
    vram = _LevelHooks(None, None, None, 1.0, None, n_rows=10)
    vram.native()(LevelReport(LevelStats(2, 3, 3), None, None, 1 << 29, 0, True))
    with pytest.raises(MemoryError, match="VRAM=2.0GB exceeds max_vram_gb=1.0GB"):
        vram.native()(LevelReport(LevelStats(2, 3, 3), None, None, 2 << 30, 0, True))
    assert _LevelHooks(None, None, None, None, None, n_rows=10).native() is None
