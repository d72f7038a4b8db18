"""The batched entry point (``etm_mine_batch``): many jobs in one call, each with the result it has alone.

Jobs of every kind go into one call: narrow ones that advance in the lockstep rounds (K=1 of at most 256
columns), wide ones and ones whose levels outgrow the rounds' 32,768-candidate limit (both mined by the per-job
loop), rows with repeated items, jobs with no frequent column, and empty jobs. Each job's itemsets and counts must
equal the oracle's and those of ``etm_mine_csr`` on that job alone; the served stream's fixed 32-job sample must
reproduce the oracle's committed per-level digests.
"""

from __future__ import annotations

import json

import numpy as np
import pytest
from conftest import GOLDEN

from golden.digests import level_digests

pytestmark = pytest.mark.gpu


def _flat_dict(flat) -> dict:
    values, offsets, counts = flat
    return {tuple(sorted(values[offsets[i]:offsets[i + 1]].tolist())): int(counts[i]) for i in range(len(counts))}


def _random_job(rng, n_rows, vocab, mean_len, zipf_s, dup=False):
    """A CSR of Zipf-popular item ids (int64): 1 + Poisson(mean_len) draws per row, distinct and ascending within a
    row; with ``dup`` about 30% of the rows repeat their first item."""
    w = 1.0 / np.arange(1, vocab + 1) ** zipf_s
    w /= w.sum()
    lens = 1 + rng.poisson(mean_len, n_rows)
    row = np.repeat(np.arange(n_rows, dtype=np.int64), lens)
    key = np.unique(row * vocab + rng.choice(vocab, size=len(row), p=w))
    row, item = key // vocab, key % vocab
    if dup:
        first = np.flatnonzero(np.r_[True, row[1:] != row[:-1]])
        again = first[rng.random(len(first)) < 0.3]
        order = np.argsort(np.r_[row, row[again]], kind="stable")
        row, item = np.r_[row, row[again]][order], np.r_[item, item[again]][order]
    indptr = np.zeros(n_rows + 1, dtype=np.int64)
    indptr[1:] = np.cumsum(np.bincount(row, minlength=n_rows))
    return indptr, (item + 1000).astype(np.int64)


def _dense_job(rng, n_rows, n_items, p):
    """Every item independently in a row with probability p: all pairs and triples frequent at low support, so
    K=3 has more candidates than the rounds take (C(n_items, 3) > 32,768 for n_items >= 60)."""
    rows = [np.flatnonzero(rng.random(n_items) < p) for _ in range(n_rows)]
    indptr = np.zeros(n_rows + 1, dtype=np.int64)
    indptr[1:] = np.cumsum([len(r) for r in rows])
    return indptr, np.concatenate(rows).astype(np.int64)


def _jobs():
    rng = np.random.default_rng(20261204)
    jobs = []
    for t in range(24):
        n_rows = int(rng.integers(500, 40_000))
        vocab = int(rng.integers(40, 3000))
        indptr, ids = _random_job(rng, n_rows, vocab, float(rng.choice([3.0, 5.0, 8.0])),
                                  float(rng.uniform(0.8, 1.3)), dup=t % 5 == 0)
        ms = float(np.exp(rng.uniform(np.log(0.003), np.log(0.05))))
        ml = [None, 2, 3, 5, 8][t % 5]
        jobs.append((indptr, ids, ms, ml))
    # Wide: about 300-600 frequent columns (the per-job loop).
    indptr, ids = _random_job(rng, 20_000, 800, 12.0, 0.3)
    jobs.append((indptr, ids, 0.004, 3))
    # Overflow: K=3 holds C(64, 3) = 41,664 candidates.
    indptr, ids = _dense_job(rng, 1500, 64, 0.45)
    jobs.append((indptr, ids, 0.08, 4))
    # No frequent column; one frequent column; empty.
    indptr, ids = _random_job(rng, 2000, 5000, 2.0, 0.0)
    jobs.append((indptr, ids, 0.2, None))
    indptr, ids = _random_job(rng, 3000, 50, 1.0, 3.0)
    jobs.append((indptr, ids, 0.3, None))
    jobs.append((np.zeros(1, np.int64), np.zeros(0, np.int64), 0.1, None))
    return jobs


def _device(jobs):
    import torch

    from et_miner.core.result import _mining_min_count

    out = []
    for indptr, ids, ms, ml in jobs:
        n = len(indptr) - 1
        out.append((torch.from_numpy(indptr).cuda(), torch.from_numpy(ids).cuda(), n,
                    _mining_min_count(ms, max(1, n)), ml))
    return out


def test_batch_equals_alone_and_oracle():
    from et_miner.gpu.native import mine_batch_device, mine_csr_device
    from oracle.miner import mine

    jobs = _jobs()
    dev_jobs = _device(jobs)
    print(f"batch of {len(jobs)} jobs", flush=True)
    batched = mine_batch_device(dev_jobs, levels=False)
    assert len(batched) == len(jobs)
    for t, ((indptr, ids, ms, ml), (d_ptr, d_ids, n, mc, _), got) in enumerate(zip(jobs, dev_jobs, batched)):
        alone = mine_csr_device(d_ptr, d_ids, n, mc, ml, levels=False)
        want = mine(indptr, ids, ms, ml).as_dict() if n else {}
        assert _flat_dict(got.flat) == want, t
        assert _flat_dict(alone.flat) == want, t
        assert [s.n_frequent for s in got.stats] == [s.n_frequent for s in alone.stats], t
        assert [s.n_candidates for s in got.stats] == [s.n_candidates for s in alone.stats], t


def test_batch_twice_in_a_row_and_shorter():
    """Two calls in a row on one thread (the second reuses the captured round graph and the cached arena), the
    second with fewer jobs, give the same results as the first on the shared jobs."""
    from et_miner.gpu.native import mine_batch_device

    dev_jobs = _device(_jobs()[:20])
    first = mine_batch_device(dev_jobs, levels=False)
    second = mine_batch_device(dev_jobs[:7], levels=False)
    for a, b in zip(first[:7], second):
        assert _flat_dict(a.flat) == _flat_dict(b.flat)


def test_batch_served_sample_equals_oracle():
    """The served stream's fixed 32-job sample (10,000-2,000,000 transactions each) in one batched call: every
    job's levels equal the oracle's committed digests."""
    import torch

    from et_miner.core.result import _mining_min_count
    from et_miner.gpu.native import mine_batch_device
    from et_miner.serve.server import jobs_from_specs
    from workloads import served_jobs

    digests = json.loads((GOLDEN / "workload_digests.json").read_text())["served_jobs"]
    specs = served_jobs(digests["n_jobs"], seed=digests["seed"])
    jobs = jobs_from_specs([specs[j] for j in digests["sample"]], workers=8)
    dev_jobs = [(torch.from_numpy(j.indptr).cuda(), torch.from_numpy(j.ids).cuda(), j.n_rows,
                 _mining_min_count(j.min_support, j.n_rows), j.max_length) for j in jobs]
    print(f"served sample: {len(jobs)} jobs in one batched call", flush=True)
    got = mine_batch_device(dev_jobs, levels=False)
    for j, r in zip(jobs, got):
        want = [{"k": d["k"], "n": d["n"], "sha256": d["sha256"]} for d in digests["jobs"][str(j.job_id)]["levels"]]
        assert level_digests(*r.flat) == want, j.job_id


def test_batch_rejects_the_modes_it_does_not_batch():
    import ctypes

    from et_miner import _lib

    opts = _lib.MineOpts(1, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0)
    handles = (ctypes.c_void_p * 1)()
    arr = (_lib.BatchJob * 1)()
    with pytest.raises(ValueError, match="per job"):
        _lib.call("etm_mine_batch", ctypes.byref(arr), 1, 1 << 26, ctypes.byref(opts), ctypes.byref(handles), None)
