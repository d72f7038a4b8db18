"""The committed workload digests on the CPU: the digest helper's invariants, and that the oracle reproduces
the committed entries for the workloads small enough to mine here in seconds (plain_60k and a few served jobs).
"""

from __future__ import annotations

import json

import numpy as np
from conftest import GOLDEN

from golden.digests import lattice_digests, level_digests
from oracle.miner import mine

DIGESTS = json.loads((GOLDEN / "workload_digests.json").read_text())


def _flat(sets):
    values = np.array([x for s, _ in sets for x in s], dtype=np.int64)
    offsets = np.zeros(len(sets) + 1, dtype=np.int64)
    offsets[1:] = np.cumsum([len(s) for s, _ in sets])
    counts = np.array([c for _, c in sets], dtype=np.int64)
    return values, offsets, counts


def test_digest_is_order_free_and_sensitive():
    sets = [((1, 5), 7), ((2,), 9), ((1, 2), 4), ((5,), 11), ((1, 2, 5), 3)]
    a = level_digests(*_flat(sets))
    # Another order of itemsets and of the items inside them: the same digests.
    b = level_digests(*_flat([((5, 1), 7), ((5,), 11), ((2, 1, 5), 3), ((2, 1), 4), ((2,), 9)]))
    assert a == b
    assert [d["k"] for d in a] == [1, 2, 3] and [d["n"] for d in a] == [2, 2, 1]
    # One count off by one, or one item changed: that level's digest changes, the others do not.
    c = level_digests(*_flat([((1, 5), 8)] + sets[1:]))
    assert c[1]["sha256"] != a[1]["sha256"] and c[0] == a[0] and c[2] == a[2]
    d = level_digests(*_flat([((1, 6), 7)] + sets[1:]))
    assert d[1]["sha256"] != a[1]["sha256"]


def test_committed_entries_reproduce():
    from workloads import served_jobs, workload

    w = workload("plain_60k")
    data = w.generate()
    lat = mine(data.indptr, data.indices, w.min_support, w.max_length)
    entry = DIGESTS["workloads"]["plain_60k"]
    assert lattice_digests(lat.levels, lat.items) == entry["levels"]
    specs = served_jobs(DIGESTS["served_jobs"]["n_jobs"], seed=DIGESTS["served_jobs"]["seed"])
    for j in (96, 304, 448):  # three of the sampled jobs of 10,000-21,000 rows
        wj = specs[j].workload
        d = wj.generate()
        lat = mine(d.indptr, d.indices, wj.min_support, wj.max_length)
        assert lattice_digests(lat.levels, lat.items) == DIGESTS["served_jobs"]["jobs"][str(j)]["levels"], j


def test_committed_entries_cover_every_benchmarked_workload():
    assert set(DIGESTS["workloads"]) == {"plain_60k", "deep_k_1m", "wide_vocab_2m", "deep_sparse_20m"}
    assert DIGESTS["workloads"]["deep_sparse_20m"]["n_rows"] == 20_000_000
    assert sum(d["n"] for d in DIGESTS["workloads"]["deep_sparse_20m"]["levels"]) == 2_000_582
    assert len(DIGESTS["served_jobs"]["jobs"]) == 32
    rows = [v["n_rows"] for v in DIGESTS["served_jobs"]["jobs"].values()]
    assert min(rows) >= 10_000 and max(rows) <= 2_000_000 and max(rows) > 1_000_000
