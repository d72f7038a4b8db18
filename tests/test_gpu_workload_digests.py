"""The benchmarked workloads at their full sizes against the oracle, through committed per-level digests.

tests/golden/make_workload_digests.py mined each workload with the oracle (CPU, once) and stored, per itemset
length, the number of frequent itemsets and a SHA-256 over every (itemset, count) pair
(tests/golden/digests.py). Here the GPU mines the same seeded data and must reproduce every level's number and
digest, so each itemset and each count is checked without running the oracle on the GPU host:

* plain_60k (60,000 rows), deep_k_1m (1,000,000), wide_vocab_2m (2,000,000) with each K=2 kernel (the VALU
  AND+popcount tiles and the fp4 MFMA Gram), deep_sparse_20m (20,000,000 rows, max_length 16, generator
  pruning) resident and by SON in 2,000,000-row chunks;
* a fixed sample of 32 jobs of the served stream (jobs 0, 16, ..., 496 of the seed-4 stream: 10,000-2,000,000
  transactions each), every job alone on the plain path and in the mix of the served form.
"""

from __future__ import annotations

import json

import numpy as np
import pytest
from conftest import GOLDEN

from golden.digests import level_digests, table_digests

pytestmark = pytest.mark.gpu

DIGESTS = json.loads((GOLDEN / "workload_digests.json").read_text())


def _want(entry) -> list:
    return [{"k": d["k"], "n": d["n"], "sha256": d["sha256"]} for d in entry["levels"]]


def _mine(data, w, **kw):
    from et_miner.core.apriori import mine_csr

    return table_digests(mine_csr(data.indptr, data.indices.astype(np.int64), w.min_support, w.max_length,
                                  use_generator_pruning=w.use_generator_pruning, with_counts=True, **kw))


@pytest.mark.parametrize("name,k2_kernel", [("plain_60k", "auto"), ("deep_k_1m", "auto"), ("wide_vocab_2m", "valu"),
                                            ("wide_vocab_2m", "mfma")])
def test_full_size_workload_equals_oracle(name, k2_kernel):
    from workloads import workload

    w = workload(name)
    entry = DIGESTS["workloads"][name]
    assert (w.n_rows, w.min_support, w.max_length) == (entry["n_rows"], entry["min_support"], entry["max_length"])
    data = w.generate()
    print(f"{name}: {data.n_rows:,} rows, k2_kernel={k2_kernel}", flush=True)
    assert len(data.indices) == entry["nnz"]
    assert _mine(data, w, k2_kernel=k2_kernel) == _want(entry)


def test_deep_sparse_20m_resident_and_son_equal_oracle():
    """20,000,000 rows: the resident run and SON in 2,000,000-row chunks both give the oracle's lattice."""
    from et_miner.streaming.son import son_csr
    from workloads import workload

    w = workload("deep_sparse_20m")
    entry = DIGESTS["workloads"]["deep_sparse_20m"]
    assert (w.n_rows, w.min_support, w.max_length) == (entry["n_rows"], entry["min_support"], entry["max_length"])
    data = w.generate()
    print(f"deep_sparse_20m: {data.n_rows:,} rows, {len(data.indices):,} entries", flush=True)
    assert len(data.indices) == entry["nnz"]
    want = _want(entry)
    assert sum(d["n"] for d in want) == 2_000_582
    assert _mine(data, w) == want
    ids = data.indices.astype(np.int64)
    son = son_csr(data.indptr, ids, w.min_support, w.max_length, w.chunk_size, use_generator_pruning=True,
                  with_counts=True)
    assert table_digests(son) == want


def test_served_sample_alone_and_mixed_equal_oracle():
    """32 jobs of the served stream at their true sizes: each job's lattice equals the oracle's alone and in
    the mix."""
    from et_miner.serve import run_plain, serve
    from et_miner.serve.server import jobs_from_specs
    from workloads import served_jobs

    sample = DIGESTS["served_jobs"]
    specs = served_jobs(sample["n_jobs"], seed=sample["seed"])
    jobs = jobs_from_specs([specs[j] for j in sample["sample"]], workers=8)
    print(f"served sample: {len(jobs)} jobs, {sum(j.n_rows for j in jobs):,} rows", flush=True)
    alone, _ = run_plain(jobs, time_scale=0.0)
    mixed, summary = serve(jobs, workers=8, time_scale=0.0)
    assert summary["peak_concurrent_jobs"] > 1
    for a, b in zip(alone, mixed):
        want = _want(sample["jobs"][str(a.job_id)])
        assert level_digests(*a.flat) == want, a.job_id
        assert level_digests(*b.flat) == want, b.job_id
