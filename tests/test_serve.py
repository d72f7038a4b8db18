"""The served form: many independent jobs mined concurrently on one GPU.

Host tests drive the request loop with a stand-in executor (admission order, the memory budget, per-job
failures, the result digest). GPU tests check that every job's result is identical alone and in the mix, and
equal to the oracle's on small jobs.
"""

from __future__ import annotations

import threading
import time

import numpy as np
import pytest

from oracle import miner


def _tiny_job(job_id, n_rows=50, seed=0, min_support=0.1, dataset=None):
    from et_miner.serve import Job

    rng = np.random.default_rng(seed)
    lens = 1 + rng.poisson(3, size=n_rows)
    rows = [np.unique(rng.integers(0, 12, size=L)) for L in lens]
    indptr = np.zeros(n_rows + 1, dtype=np.int64)
    indptr[1:] = np.cumsum([len(r) for r in rows])
    return Job(job_id, indptr, np.concatenate(rows).astype(np.int64), min_support, 4, dataset=dataset)


def _oracle_flat(job):
    lat = miner.mine(job.indptr, job.ids, job.min_support, job.max_length,
                     use_generator_pruning=job.use_generator_pruning)
    d = lat.as_dict()
    values, offsets, counts = [], [0], []
    for items, c in d.items():
        values.extend(sorted(items))
        offsets.append(len(values))
        counts.append(c)
    return (np.array(values, dtype=np.int64), np.array(offsets, dtype=np.int64), np.array(counts, dtype=np.int64))


def test_digest_is_order_independent_and_count_sensitive():
    from et_miner.serve import result_digest

    v = np.array([1, 1, 2, 3], dtype=np.int64)
    o = np.array([0, 1, 3, 4], dtype=np.int64)
    c = np.array([5, 4, 7], dtype=np.int64)
    # same itemsets and counts, another order
    v2 = np.array([3, 1, 2, 1], dtype=np.int64)
    o2 = np.array([0, 1, 3, 4], dtype=np.int64)
    c2 = np.array([7, 4, 5], dtype=np.int64)
    assert result_digest(v, o, c) == result_digest(v2, o2, c2)
    # the items inside an itemset in another order
    assert result_digest(v, o, c) == result_digest(np.array([1, 2, 1, 3], dtype=np.int64), o, c)
    assert result_digest(v, o, c) != result_digest(v, o, np.array([5, 4, 6], dtype=np.int64))
    # {1},{1,2} vs {1,1},{2}: the itemset boundaries matter
    assert result_digest(v, o, c) != result_digest(v, np.array([0, 2, 3, 4], dtype=np.int64), c)


def test_scheduler_fifo_budget_and_failures():
    """Jobs start in arrival order, the admitted footprint never exceeds the budget (a job larger than the
    budget runs alone), and a failing job reports its error without stopping the others."""
    from et_miner.serve import Server
    from et_miner.serve.server import footprint_bytes

    jobs = [_tiny_job(i, n_rows=30 + 10 * i, seed=i) for i in range(12)]
    need = {j.job_id: footprint_bytes(j) for j in jobs}
    budget = int(2.5 * max(need.values()))
    state = {"used": 0, "peak": 0, "running": 0, "peak_running": 0}
    lock = threading.Lock()

    def executor(job):
        with lock:
            state["used"] += min(need[job.job_id], budget)
            state["running"] += 1
            state["peak"] = max(state["peak"], state["used"])
            state["peak_running"] = max(state["peak_running"], state["running"])
        time.sleep(0.01 * (1 + job.job_id % 3))
        with lock:
            state["used"] -= min(need[job.job_id], budget)
            state["running"] -= 1
        if job.job_id == 5:
            raise ValueError("job 5 is broken")
        return _oracle_flat(job) + (1, 0)

    srv = Server(workers=4, memory_budget_bytes=budget, executor=executor)
    futs = [srv.submit(j, arrival_s=0.0) for j in jobs]
    out = {}
    for j, f in zip(jobs, futs):
        if j.job_id == 5:
            with pytest.raises(ValueError, match="job 5"):
                f.result()
        else:
            out[j.job_id] = f.result()
    srv.close()
    assert srv.order == [j.job_id for j in jobs]
    assert state["peak"] <= budget
    assert 1 < state["peak_running"] <= 4
    from et_miner.serve import result_digest

    for j in jobs:
        if j.job_id != 5:
            assert out[j.job_id].digest == result_digest(*_oracle_flat(j))


def test_oversized_job_runs_alone():
    from et_miner.serve import Server

    jobs = [_tiny_job(i, seed=i) for i in range(6)]
    live = {"n": 0, "max_with_big": 0}
    lock = threading.Lock()

    def executor(job):
        with lock:
            live["n"] += 1
            if job.job_id == 2:
                live["max_with_big"] = live["n"]
        time.sleep(0.02)
        with lock:
            if job.job_id == 2:
                live["max_with_big"] = max(live["max_with_big"], live["n"])
            live["n"] -= 1
        return _oracle_flat(job) + (1, 0)

    # a budget below one job's footprint: every job is oversized and runs alone
    srv = Server(workers=3, memory_budget_bytes=1, executor=executor)
    futs = [srv.submit(j, arrival_s=0.0) for j in jobs]
    for f in futs:
        f.result()
    srv.close()
    assert live["max_with_big"] == 1
    assert srv.peak_running == 1


def test_batch_workers_take_narrow_jobs_in_groups():
    """Narrow jobs (at most 256 frequent columns, no generator pruning) go to the batch workers, several per
    call, within the memory budget; the others to the per-job workers. Every job gets its own result, and a
    failing batch reports its error on each of its jobs."""
    from et_miner.serve import Server
    from et_miner.serve.server import batchable, occurrence_profile

    narrow = [_tiny_job(i, n_rows=40, seed=i) for i in range(20)]
    wide_job = _tiny_job(100, n_rows=60, seed=100, min_support=0.1)
    wide_job.use_generator_pruning = True  # generator pruning is mined per job
    assert all(batchable(j, occurrence_profile(j.ids)) for j in narrow)
    assert not batchable(wide_job, occurrence_profile(wide_job.ids))
    calls, per_job = [], []
    gate = threading.Event()

    def batch_executor(jobs):
        gate.wait(5)
        calls.append([j.job_id for j in jobs])
        if any(j.job_id == 13 for j in jobs):
            raise RuntimeError("batch with job 13 failed")
        return [_oracle_flat(j) + (1, 0) for j in jobs]

    def executor(job):
        per_job.append(job.job_id)
        return _oracle_flat(job) + (1, 0)

    srv = Server(workers=1, memory_budget_bytes=1 << 40, executor=executor, batch_workers=2, max_batch=8,
                 batch_executor=batch_executor)
    futs = [srv.submit(j, arrival_s=0.0) for j in narrow + [wide_job]]
    gate.set()
    results = {}
    for j, f in zip(narrow + [wide_job], futs):
        exc = f.exception(10)
        results[j.job_id] = exc if exc is not None else f.result()
    srv.close()
    from et_miner.serve import result_digest

    assert per_job == [100]
    assert sorted(i for c in calls for i in c) == list(range(20))
    assert all(len(c) <= 8 for c in calls) and max(len(c) for c in calls) > 1
    failed = next(c for c in calls if 13 in c)
    for j in narrow + [wide_job]:
        r = results[j.job_id]
        if j.job_id in failed:
            assert isinstance(r, RuntimeError) and "job 13" in str(r)
        else:
            assert r.digest == result_digest(*_oracle_flat(j))


def test_footprint_counts_frequent_columns():
    from et_miner.serve import Job
    from et_miner.serve.server import footprint_bytes

    # 3 ids reach min_count 2 (ids 0, 1, 2); id 9 occurs once
    indptr = np.array([0, 3, 6, 7], dtype=np.int64)
    ids = np.array([0, 1, 2, 0, 1, 2, 9], dtype=np.int64)
    j = Job(0, indptr, ids, 0.5)
    pitch = 64
    assert footprint_bytes(j) == 8 * 4 + 8 * 7 + 8 * 3 * pitch + 25 * 3 + (64 << 20)


def _served_sample(n_jobs, rows_range, seed=4):
    from et_miner.serve.server import jobs_from_specs
    from workloads import served_jobs

    return jobs_from_specs(served_jobs(n_jobs, seed=seed, rows_range=rows_range), workers=1)


@pytest.mark.gpu
def test_served_equals_alone_and_oracle():
    """Each job's result is the same alone (plain path) and in the mix (served, all jobs at once)."""
    from et_miner.serve import result_digest, run_plain, serve

    jobs = _served_sample(16, (2_000, 40_000))
    alone, _ = run_plain(jobs, time_scale=0.0)
    mixed, summary = serve(jobs, workers=6, time_scale=0.0)
    assert [r.job_id for r in alone] == [r.job_id for r in mixed]
    for a, b in zip(alone, mixed):
        assert (a.digest, a.n_itemsets) == (b.digest, b.n_itemsets), a.job_id
    assert summary["jobs"] == 16 and summary["peak_concurrent_jobs"] > 1
    for j in [j for j in jobs if j.n_rows <= 6_000][:3]:
        assert mixed[j.job_id].digest == result_digest(*_oracle_flat(j)), j.job_id


@pytest.mark.gpu
def test_served_under_tight_budget_and_arrivals():
    """A budget that admits about two jobs at a time, Poisson arrivals, a dataset shared by repeated jobs, and
    the autotuner (K=2 kernel, K>=3 schedule and class split per shape bucket): results unchanged."""
    import dataclasses

    from et_miner.serve import footprint_bytes, run_plain, serve

    jobs = _served_sample(8, (2_000, 30_000), seed=9)
    jobs += [dataclasses.replace(j, job_id=100 + j.job_id, min_support=j.min_support * 2) for j in jobs[:4]]
    budget = 2 * max(footprint_bytes(j) for j in jobs)
    alone, _ = run_plain(jobs, time_scale=0.0)
    from et_miner.serve import autotune

    autotune.reset()
    mixed, summary = serve(jobs, workers=4, time_scale=0.2, memory_budget_bytes=budget, autotune=True)
    for a, b in zip(alone, mixed):
        assert a.digest == b.digest, a.job_id
    # Every tuned bucket holds a valid setting of each tuned choice.
    tuned = autotune.decisions()
    assert tuned
    for s in tuned.values():
        assert s.k2_kernel in ("valu", "mfma", "rows")
        assert s.k3_schedule in ("auto", "all", "live") and s.small_max in (0, 4, 8, 16)
    assert summary["latency_p99_s"] >= summary["latency_p50_s"] > 0
