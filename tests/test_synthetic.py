"""The data generators: determinism and the laws the workloads are defined by (CPU only)."""

from __future__ import annotations

import numpy as np

from workloads import PRESETS, generate_csr, served_jobs, workload


def _rows_ok(data):
    lens = np.diff(data.indptr)
    assert lens.min() >= 0 and data.indptr[-1] == len(data.indices)
    r = np.repeat(np.arange(data.n_rows), lens)
    # unique and ascending within every row
    same = r[1:] == r[:-1]
    assert np.all(data.indices[1:][same] > data.indices[:-1][same])
    assert data.indices.min() >= 0 and data.indices.max() < data.n_cols


def test_preset_smoke_matches_reference_shape():
    d = generate_csr(PRESETS["smoke"])
    _rows_ok(d)
    assert d.n_rows == 60_000 and len(d.indices) == 498_549
    assert [p[1] for p in d.planted] == [1200, 1200, 1200]


def test_workload_deterministic_and_lawful():
    a = workload("plain_60k", n_rows=5_000).generate()
    b = workload("plain_60k", n_rows=5_000).generate()
    assert np.array_equal(a.indptr, b.indptr) and np.array_equal(a.indices, b.indices)
    _rows_ok(a)
    lens = np.diff(a.indptr)
    # 1 + Poisson(9), drawn without replacement: lengths are exactly the drawn basket sizes.
    assert lens.min() >= 1 and abs(lens.mean() - 10.0) < 0.3


def test_planted_workload_counts():
    w = workload("deep_k_1m", n_rows=50_000)
    d = w.generate()
    _rows_ok(d)
    assert len(d.planted) == 20
    for items, full in d.planted:
        assert 8 <= len(items) <= 12
        assert full == 1500 - 150  # 3% of rows, 10% of them missing one member
        rows_with_all = np.ones(d.n_rows, bool)
        r = np.repeat(np.arange(d.n_rows), np.diff(d.indptr))
        for it in items:
            has = np.zeros(d.n_rows, bool)
            has[r[d.indices == it]] = True
            rows_with_all &= has
        assert rows_with_all.sum() >= full


def test_lognormal_baskets_clipped():
    d = workload("wide_vocab_2m", n_rows=20_000).generate()
    _rows_ok(d)
    lens = np.diff(d.indptr)
    assert lens.min() >= 1 and lens.max() <= 200


def test_served_jobs_stream():
    jobs = served_jobs(64, rows_range=(10_000, 200_000))
    assert len(jobs) == 64
    arr = [j.arrival_s for j in jobs]
    assert arr == sorted(arr) and arr[0] == 0.0
    for j in jobs:
        w = j.workload
        assert 10_000 <= w.n_rows <= 200_000 and 500 <= w.vocab <= 20_000
        assert 0.001 <= w.min_support <= 0.05 and 3 <= w.max_length <= 10
