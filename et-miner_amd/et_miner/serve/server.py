"""A request loop that mines many independent jobs concurrently on one GPU.

Jobs arrive over time and wait in arrival order. A fixed pool of workers, each with its own HIP stream (default
priority) and its own cached device arena inside libetm, takes the head of the queue when its estimated device
footprint (resident CSR + K=1 bitvectors + the pair level's per-candidate buffers + the level hash) fits the
memory budget beside the jobs already running; a job larger than the whole budget runs alone. A worker runs the
whole level loop of its job inside libetm (``etm_mine_csr``; the C call releases the GIL), so the levels of
several jobs -- their kernels and their per-level host syncs -- interleave on the device: a job that arrives
while others are on the GPU starts as soon as a worker and memory free up. Datasets stay resident in HBM in a
byte-bounded store (least recently used out first), so jobs on a registered dataset skip the upload.

Jobs narrow enough for the batched rounds (plain mode, at most 256 frequent columns: most small jobs) wait in a
second queue. Batch workers take every such job queued when they come free (up to ``max_batch``, within the
memory budget) and mine them in one call (``etm_mine_batch``): the input layers share their host syncs, and the
jobs' levels advance together, one round of shared launches (work-queue-fed counting kernels, a workgroup per job
for group build and compaction, replayed from a HIP graph) per level, instead of a dozen launches and a host sync
per job and level.

The plain form (``run_plain``) runs the same jobs one after another on one stream, as one caller of
``apriori(use_gpu=True)`` per job would. Both report jobs/s and latency percentiles (completion - arrival), and
every job's result is the same in either form (``JobResult.digest`` over every itemset and count).
"""

from __future__ import annotations

import hashlib
import threading
import time
from collections import OrderedDict, deque
from concurrent.futures import Future
from dataclasses import dataclass

import numpy as np

from ..core.result import _mining_min_count


@dataclass
class Job:
    job_id: int
    indptr: np.ndarray  # int64 (N+1)
    ids: np.ndarray  # item ids (int32 or int64)
    min_support: float
    max_length: int | None = None
    use_generator_pruning: bool = False
    arrival_s: float = 0.0
    dataset: str | None = None  # jobs naming one dataset share its resident copy

    @property
    def n_rows(self) -> int:
        return len(self.indptr) - 1


@dataclass
class JobResult:
    """A finished job: its frequent itemsets (flat: item ids, offsets, counts) and its timeline."""

    job_id: int
    flat: tuple
    arrival_s: float
    start_s: float
    end_s: float
    levels: int = 0
    candidates: int = 0

    @property
    def n_itemsets(self) -> int:
        return len(self.flat[2])

    @property
    def digest(self) -> str:
        return result_digest(*self.flat)

    @property
    def latency_s(self) -> float:
        return self.end_s - self.arrival_s


def _mix64(x: np.ndarray) -> np.ndarray:
    """splitmix64's finaliser (Steele, Lea, Flood 2014) on uint64, elementwise."""
    with np.errstate(over="ignore"):
        x = x + np.uint64(0x9E3779B97F4A7C15)
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return x ^ (x >> np.uint64(31))


def result_digest(values: np.ndarray, offsets: np.ndarray, counts: np.ndarray) -> str:
    """Digest of a lattice's (itemset, count) pairs, independent of the order of itemsets and of the items in
    each: per itemset the sum of its items' 64-bit mixes, mixed with its count; then the sorted hashes, hashed."""
    n = len(counts)
    if n == 0:
        return hashlib.sha256(b"").hexdigest()[:16]
    lens = np.diff(offsets)
    if np.any(lens <= 0):
        raise ValueError("empty itemset in a result")
    with np.errstate(over="ignore"):
        h = np.add.reduceat(_mix64(np.asarray(values).astype(np.uint64)), np.asarray(offsets[:-1]))
        h = _mix64(h ^ _mix64(np.asarray(counts).astype(np.uint64) + (lens.astype(np.uint64) << np.uint64(48))))
    return hashlib.sha256(np.sort(h).tobytes()).hexdigest()[:16]


def occurrence_profile(ids: np.ndarray) -> np.ndarray:
    """Occurrence counts of a dataset's ids, ascending (the footprint of any threshold is one search away)."""
    ids = np.asarray(ids)
    if len(ids) == 0:
        return np.zeros(0, dtype=np.int64)
    c = np.bincount(ids - int(ids.min()))
    return np.sort(c[c > 0])


def footprint_bytes(job: Job, profile: np.ndarray | None = None) -> int:
    """Estimated device bytes of a job: CSR + F bitvector columns + the pair level's per-candidate buffers.

    F = the ids with at least min_count occurrences (the K=1 survivors; ``profile`` is the dataset's
    ``occurrence_profile``); P = F(F-1)/2 pair candidates, each with a status byte, an int32 count and an int32
    group id; survivors and the level hash are bounded by P as well (4 B per survivor item + 8 B of hash per
    survivor). A fixed 64 MiB covers the arena's first block."""
    n = job.n_rows
    nnz = len(job.ids)
    mc = _mining_min_count(job.min_support, max(1, n))
    prof = occurrence_profile(job.ids) if profile is None else profile
    f = int(len(prof) - np.searchsorted(prof, mc, side="left"))
    pitch = ((n + 63) // 64 + 63) // 64 * 64
    pairs = f * (f - 1) // 2
    return int(8 * (n + 1) + 8 * nnz + 8 * f * pitch + 25 * pairs + (64 << 20))


class DatasetStore:
    """Device-resident CSRs keyed by dataset name, evicted least recently used first beyond ``budget`` bytes."""

    def __init__(self, budget_bytes: int):
        self.budget = budget_bytes
        self.items: OrderedDict = OrderedDict()
        self.bytes = 0
        self.lock = threading.Lock()
        self.hits = 0
        self.misses = 0
        self.profiles: dict = {}

    def profile(self, key, ids: np.ndarray) -> np.ndarray:
        """The dataset's occurrence profile, computed once per dataset name."""
        if key is None:
            return occurrence_profile(ids)
        with self.lock:
            p = self.profiles.get(key)
        if p is None:
            p = occurrence_profile(ids)
            with self.lock:
                self.profiles[key] = p
        return p

    def register(self, key, indptr: np.ndarray, ids: np.ndarray, device):
        """Uploads a dataset (synchronously) and keeps it resident; returns (indptr, ids) device tensors."""
        import torch

        with self.lock:
            if key is not None and key in self.items:
                self.items.move_to_end(key)
                self.hits += 1
                return self.items[key][0]
            self.misses += 1
        d_ptr = torch.from_numpy(np.ascontiguousarray(indptr, np.int64)).to(device)
        d_ids = torch.from_numpy(np.ascontiguousarray(ids)).to(device)
        torch.cuda.current_stream(device).synchronize()
        size = d_ptr.numel() * 8 + d_ids.numel() * d_ids.element_size()
        if key is not None:
            with self.lock:
                if key not in self.items:
                    self.items[key] = ((d_ptr, d_ids), size)
                    self.bytes += size
                while self.bytes > self.budget and len(self.items) > 1:
                    _, (_, sz) = self.items.popitem(last=False)
                    self.bytes -= sz
        return d_ptr, d_ids


#: K=1 columns of a job the batched rounds take (etm_mine_batch): K=2 then has at most 32,640 candidates.
BATCH_MAX_COLS = 256
#: Device bytes per batched job beside its CSR and bitvectors: the round buffers (hash, per-candidate status,
#: counts and group ids, two group buffers) and the 8 MiB level store.
BATCH_JOB_BYTES = 11 << 20


def frequent_columns(job: Job, profile: np.ndarray) -> int:
    """Item ids of a job occurring at least min_count times (its K=1 columns before duplicate rows are merged)."""
    mc = _mining_min_count(job.min_support, max(1, job.n_rows))
    return int(len(profile) - np.searchsorted(profile, mc, side="left"))


def batchable(job: Job, profile: np.ndarray) -> bool:
    """Whether the batched rounds take the job: the plain mode (no generator pruning) and at most
    BATCH_MAX_COLS frequent columns."""
    return not job.use_generator_pruning and frequent_columns(job, profile) <= BATCH_MAX_COLS


class Server:
    """Continuous batching of mining jobs on one GPU (see the module docstring)."""

    def __init__(self, workers: int = 8, memory_budget_bytes: int | None = None,
                 dataset_budget_bytes: int | None = None, autotune: bool = False, executor=None,
                 batch_workers: int = 0, max_batch: int = 64, batch_executor=None):
        """``executor(job) -> (values, offsets, counts, n_levels, n_candidates)`` replaces the device run; it
        exists to test the scheduling on a host without a GPU and then needs ``memory_budget_bytes``.

        ``batch_workers`` threads take the batchable jobs (``batchable``), up to ``max_batch`` of those queued at
        a time, and mine each group in one batched call (``mine_batch_device``): their levels advance together,
        one round of shared launches per level. ``batch_executor(jobs) -> [executor results]`` replaces that
        call in host-only tests."""
        self.executor = executor
        self.batch_executor = batch_executor
        self.autotune = autotune
        self.max_batch = max(1, int(max_batch))
        if executor is None:
            import torch

            from ..gpu import device as dev

            dev.require_gpu()
            self.torch = torch
            self.device = dev.device()
            free, total = torch.cuda.mem_get_info(self.device)
            self.budget = memory_budget_bytes or int(0.6 * free)
            self.store = DatasetStore(dataset_budget_bytes or int(0.25 * total))
        else:
            if memory_budget_bytes is None:
                raise ValueError("a custom executor needs memory_budget_bytes")
            self.torch = None
            self.budget = memory_budget_bytes
            self.store = DatasetStore(0)
        self.order: list = []  # job ids in start order
        self.cv = threading.Condition()
        self.queue: deque = deque()  # (job, future, arrival_s, footprint)
        self.bqueue: deque = deque()  # the same, for the batch workers
        self.batches = 0
        self.batched_jobs = 0
        self.used = 0
        self.running = 0
        self.peak_running = 0
        self.closed = False
        self.t0 = time.perf_counter()
        self.workers = [threading.Thread(target=self._worker, daemon=True, name=f"etm-serve-{i}")
                        for i in range(workers)]
        self.workers += [threading.Thread(target=self._batch_worker, daemon=True, name=f"etm-batch-{i}")
                         for i in range(batch_workers)]
        self.n_batch_workers = batch_workers
        for t in self.workers:
            t.start()

    def now(self) -> float:
        return time.perf_counter() - self.t0

    def _take(self):
        with self.cv:
            while True:
                if self.queue:
                    job, fut, arrival, need = self.queue[0]
                    if self.running == 0 or self.used + need <= self.budget:
                        self.queue.popleft()
                        self.order.append(job.job_id)
                        self.running += 1
                        self.peak_running = max(self.peak_running, self.running)
                        self.used += need
                        return job, fut, arrival, need
                elif self.closed:
                    return None
                self.cv.wait()

    def _take_batch(self):
        """The batch queue's head and the jobs behind it (up to max_batch) that fit the memory budget; the head
        alone when nothing else runs and it exceeds the budget."""
        with self.cv:
            while True:
                if self.bqueue:
                    taken = []
                    need = 0
                    while self.bqueue and len(taken) < self.max_batch:
                        job, fut, arrival, n = self.bqueue[0]
                        if self.used + need + n > self.budget and (taken or self.running > 0):
                            break
                        self.bqueue.popleft()
                        taken.append((job, fut, arrival, n))
                        need += n
                    if taken:
                        for job, _, _, _ in taken:
                            self.order.append(job.job_id)
                        self.running += len(taken)
                        self.peak_running = max(self.peak_running, self.running)
                        self.used += need
                        return taken
                elif self.closed:
                    return None
                self.cv.wait()

    def _batch_worker(self) -> None:
        torch = self.torch
        stream = torch.cuda.Stream(device=self.device) if torch is not None else None
        while True:
            items = self._take_batch()
            if items is None:
                return
            jobs = [it[0] for it in items]
            try:
                start = self.now()
                if stream is None:
                    outs = self.batch_executor(jobs)
                else:
                    with torch.cuda.stream(stream):
                        res = self._execute_batch(jobs)
                    stream.synchronize()
                    outs = [(*r.flat, len(r.stats), sum(st.n_evaluated for st in r.stats)) for r in res]
                end = self.now()
                for (job, fut, arrival, _), (values, offsets, counts, n_levels, n_cands) in zip(items, outs):
                    fut.set_result(JobResult(job.job_id, (values, offsets, counts), arrival, start, end, n_levels,
                                             n_cands))
                with self.cv:
                    self.batches += 1
                    self.batched_jobs += len(items)
            except BaseException as e:  # every job of the failing batch reports it; the server keeps serving
                for _, fut, _, _ in items:
                    if not fut.done():
                        fut.set_exception(e)
            finally:
                with self.cv:
                    self.running -= len(items)
                    self.used -= sum(it[3] for it in items)
                    self.cv.notify_all()

    def _execute_batch(self, jobs: list):
        from ..gpu.native import mine_batch_device

        dev_jobs = []
        for job in jobs:
            d_ptr, d_ids = self.store.register(job.dataset, job.indptr, job.ids, self.device)
            dev_jobs.append((d_ptr, d_ids, job.n_rows, _mining_min_count(job.min_support, job.n_rows),
                             job.max_length))
        return mine_batch_device(dev_jobs, levels=False)

    def _worker(self) -> None:
        torch = self.torch
        stream = torch.cuda.Stream(device=self.device) if torch is not None else None
        while True:
            item = self._take()
            if item is None:
                return
            job, fut, arrival, need = item
            try:
                start = self.now()
                if stream is None:
                    values, offsets, counts, n_levels, n_cands = self.executor(job)
                else:
                    with torch.cuda.stream(stream):
                        res = self._execute(job)
                    stream.synchronize()
                    values, offsets, counts = res.flat
                    n_levels, n_cands = len(res.stats), sum(s.n_evaluated for s in res.stats)
                end = self.now()
                fut.set_result(JobResult(job.job_id, (values, offsets, counts), arrival, start, end, n_levels,
                                         n_cands))
            except BaseException as e:  # the failing job reports it; the server keeps serving
                fut.set_exception(e)
            finally:
                with self.cv:
                    self.running -= 1
                    self.used -= need
                    self.cv.notify_all()

    def _execute(self, job: Job):
        from ..gpu.native import mine_csr_device

        d_ptr, d_ids = self.store.register(job.dataset, job.indptr, job.ids, self.device)
        mc = _mining_min_count(job.min_support, job.n_rows)
        kw = {}
        if self.autotune:
            from .autotune import setting_for

            kw = setting_for(job, mc, d_ptr, d_ids, self.store.profile(job.dataset, job.ids)).kwargs()
        return mine_csr_device(d_ptr, d_ids, job.n_rows, mc, job.max_length,
                               infer_counts=job.use_generator_pruning, levels=False, **kw)

    def tune(self, jobs: list) -> dict:
        """Tunes the kernel settings of the jobs' shape buckets serially, before they are served (needs
        ``autotune``; the workers must be idle)."""
        if self.executor is not None or not self.autotune:
            return {}
        from .autotune import tune

        with self.torch.cuda.device(self.device):
            return tune(jobs, self.store, self.device)

    def submit(self, job: Job, arrival_s: float | None = None) -> Future:
        """Queues a job; its future resolves to a JobResult (or raises the job's error)."""
        fut: Future = Future()
        prof = self.store.profile(job.dataset, job.ids) if self.store is not None else None
        if prof is None:
            prof = occurrence_profile(job.ids)
        to_batch = self.n_batch_workers > 0 and batchable(job, prof)
        need = footprint_bytes(job, prof) + (BATCH_JOB_BYTES if to_batch else 0)
        need = min(need, self.budget)
        with self.cv:
            if self.closed:
                raise RuntimeError("server is closed")
            (self.bqueue if to_batch else self.queue).append(
                (job, fut, self.now() if arrival_s is None else arrival_s, need))
            self.cv.notify_all()
        return fut

    def close(self) -> None:
        """Stops accepting jobs; workers exit once the queue is drained."""
        with self.cv:
            self.closed = True
            self.cv.notify_all()
        for t in self.workers:
            t.join()


def summarize(results: list, makespan: float) -> dict:
    lat = np.array([r.latency_s for r in results]) if results else np.zeros(1)
    return {
        "jobs": len(results),
        "makespan_s": makespan,
        "jobs_per_s": len(results) / makespan if makespan > 0 else 0.0,
        "latency_p50_s": float(np.percentile(lat, 50)),
        "latency_p99_s": float(np.percentile(lat, 99)),
        "latency_mean_s": float(lat.mean()),
    }


def _pace(t0: float, due: float) -> None:
    while True:
        left = due - (time.perf_counter() - t0)
        if left <= 0:
            return
        time.sleep(min(0.002, left))


def serve(jobs: list, workers: int = 8, time_scale: float = 1.0, resident: bool = True, server=None,
          **server_kw) -> tuple:
    """Replays ``jobs`` at their arrival times (``arrival_s * time_scale``; 0 = all at once) through a Server
    (``server``, kept open, or a new one closed at the end).

    With ``resident`` every job's dataset is registered in HBM before the clock starts. Returns (results sorted
    by job id, summary)."""
    own = server is None
    srv = Server(workers=workers, **server_kw) if own else server
    for j in jobs:
        if j.dataset is not None:
            srv.store.profile(j.dataset, j.ids)
            if resident and srv.executor is None:
                srv.store.register(j.dataset, j.indptr, j.ids, srv.device)
    futs = []
    t0 = time.perf_counter()
    srv.t0 = t0
    srv.peak_running = srv.running
    batches0, batched0 = srv.batches, srv.batched_jobs
    for job in sorted(jobs, key=lambda j: j.arrival_s):
        due = job.arrival_s * time_scale
        _pace(t0, due)
        futs.append(srv.submit(job, arrival_s=due))
    results = [f.result() for f in futs]
    if own:
        srv.close()
    makespan = max(r.end_s for r in results) if results else 0.0
    out = summarize(results, makespan)
    out["workers"] = len(srv.workers) - srv.n_batch_workers
    out["batch_workers"] = srv.n_batch_workers
    out["peak_concurrent_jobs"] = srv.peak_running
    out["batches"] = srv.batches - batches0
    out["batched_jobs"] = srv.batched_jobs - batched0
    return sorted(results, key=lambda r: r.job_id), out


def run_plain(jobs: list, time_scale: float = 1.0, resident: bool = True) -> tuple:
    """The plain per-job path: each job alone, in arrival order, on the current stream, once it has arrived."""
    import torch

    from ..gpu import device as dev
    from ..gpu.native import mine_csr_device

    dev.require_gpu()
    store = DatasetStore(1 << 62)
    if resident:
        for j in jobs:
            store.register(j.dataset, j.indptr, j.ids, dev.device())
    results = []
    t0 = time.perf_counter()
    for job in sorted(jobs, key=lambda j: j.arrival_s):
        due = job.arrival_s * time_scale
        _pace(t0, due)
        start = time.perf_counter() - t0
        d_ptr, d_ids = store.register(job.dataset, job.indptr, job.ids, dev.device())
        mc = _mining_min_count(job.min_support, job.n_rows)
        res = mine_csr_device(d_ptr, d_ids, job.n_rows, mc, job.max_length,
                              infer_counts=job.use_generator_pruning, levels=False)
        torch.cuda.synchronize()
        end = time.perf_counter() - t0
        results.append(JobResult(job.job_id, res.flat, due, start, end, len(res.stats),
                                 sum(s.n_evaluated for s in res.stats)))
    makespan = max(r.end_s for r in results) if results else 0.0
    return sorted(results, key=lambda r: r.job_id), summarize(results, makespan)


def jobs_from_specs(specs: list, workers: int = 8) -> list:
    """The data of served-job specs (each with ``job_id``, ``arrival_s`` and a ``workload`` whose ``generate()``
    returns CSR arrays), generated in a process pool.

    The pool's workers are spawned, not forked, so a caller that has already initialised the GPU hands no device
    state to them."""
    import multiprocessing
    from concurrent.futures import ProcessPoolExecutor

    if workers <= 1:
        datas = [s.workload.generate() for s in specs]
    else:
        with ProcessPoolExecutor(max_workers=workers, mp_context=multiprocessing.get_context("spawn")) as pool:
            datas = list(pool.map(_generate, [s.workload for s in specs]))
    return [Job(s.job_id, d.indptr, d.indices, s.workload.min_support, s.workload.max_length,
                arrival_s=s.arrival_s, dataset=s.workload.name) for s, d in zip(specs, datas)]


def _generate(w):
    return w.generate()
