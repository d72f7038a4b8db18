"""Kernel choice per shape bucket for the served jobs, tuned by timing on the bucket's first job.

A job's bucket is (frequent columns F, words W), each rounded up to a power of two. The first job of a bucket is
mined once per candidate setting with the level counting launches event-timed; later jobs of the bucket reuse the
fastest (``tune`` does this for a list of jobs with the GPU otherwise idle, before serving them). Every setting is
exact, so the choice never changes a result, only its time. Three choices are tuned, one after another (each
with the earlier ones fixed at their best; the library's automatic choice stays unless another is more than 3%
faster):

* the K=2 kernel: the VALU AND+popcount tiles or the fp4 MFMA Gram, timed on the pair level
  alone; the row-wise kernel joins the candidates where the rows are sparse enough for it to have a chance
  (sum over rows of C(len, 2) below 1% of F(F-1)/2 x W, the reference's crossover times 2.5);
* the launch schedule of levels K >= 3: every group and tile-pair, or only those holding a counted candidate
  (``k3_schedule``), timed over the whole lattice;
* the split between the wave-per-group kernels and the tiled kernel (``small_max``: groups of at most 4, 8 or
  16 suffixes on the wave-per-group kernels), timed over the whole lattice.

The reference tunes the same decisions by fixed crossovers measured per K (``gpu/row_split_chunks.py:52-67``,
``:96-121``, ``:130-159``); here they are measured on the GPU that serves the jobs.
"""

from __future__ import annotations

import threading
from dataclasses import dataclass

import numpy as np

_DECISIONS: dict = {}
_LOCK = threading.Lock()

#: Row-wise K=2 joins the candidates below this sum C(len, 2) / (pairs x words).
ROWS_TRY_MAX_R = 1.0e-2
#: A candidate replaces the current choice only when it is faster by more than this fraction (timing noise).
MIN_GAIN = 0.03


def _pick(times: dict, current):
    """The fastest entry of ``times``, unless it beats ``current`` (also a key) by no more than MIN_GAIN."""
    best = min(times, key=times.get)
    return best if times[best] < (1.0 - MIN_GAIN) * times[current] else current


@dataclass(frozen=True)
class Setting:
    """The tuned kernel settings of a bucket (the defaults are the library's automatic choices)."""

    k2_kernel: str = "auto"
    k3_schedule: str = "auto"
    small_max: int = 0

    def kwargs(self) -> dict:
        return {"k2_kernel": self.k2_kernel, "k3_schedule": self.k3_schedule, "small_max": self.small_max}


def bucket(n_cols: int, n_words: int) -> tuple:
    return (1 << max(0, int(n_cols - 1).bit_length()), 1 << max(0, int(n_words - 1).bit_length()))


def _frequent_cols(profile: np.ndarray, min_count: int) -> int:
    return int(len(profile) - np.searchsorted(profile, min_count, side="left"))


def _row_pairs_ratio(job, min_count: int, n_cols: int) -> float:
    """sum over rows of C(len, 2) of the frequent items, over F(F-1)/2 pairs x words."""
    ids = np.asarray(job.ids)
    if n_cols < 2 or len(ids) == 0 or job.n_rows == 0:
        return float("inf")
    lo = int(ids.min())
    freq = np.bincount(ids - lo) >= min_count
    rows = np.repeat(np.arange(job.n_rows, dtype=np.int64), np.diff(np.asarray(job.indptr)))
    lens = np.bincount(rows[freq[ids - lo]], minlength=job.n_rows).astype(np.int64)
    pairs = float((lens * (lens - 1) // 2).sum())
    return pairs / (n_cols * (n_cols - 1) / 2 * max(1, (job.n_rows + 63) // 64))


def _timed(job, min_count, d_ptr, d_ids, max_length, **kw) -> float:
    """Event-timed counting of one mining run (sum over levels of the counting launches' duration, ms)."""
    from ..gpu.native import mine_csr_device

    res = mine_csr_device(d_ptr, d_ids, job.n_rows, min_count, max_length, time_kernels=True, levels=False,
                          infer_counts=job.use_generator_pruning, **kw)
    return float(sum(s.count_ms for s in res.stats))


def _wall(job, min_count, d_ptr, d_ids, **kw) -> float:
    """Wall time (ms) of one whole mining run, after the device is idle."""
    import time

    import torch

    from ..gpu.native import mine_csr_device

    torch.cuda.current_stream().synchronize()
    t = time.perf_counter()
    mine_csr_device(d_ptr, d_ids, job.n_rows, min_count, job.max_length, levels=False,
                    infer_counts=job.use_generator_pruning, **kw)
    torch.cuda.current_stream().synchronize()
    return (time.perf_counter() - t) * 1000.0


def setting_for(job, min_count: int, d_ptr, d_ids, profile: np.ndarray) -> Setting:
    """The tuned setting of a job's bucket, tuning the bucket on first sight. ``profile`` is the dataset's
    ascending occurrence counts, so the frequent-column count is one search."""
    n_cols = _frequent_cols(profile, min_count)
    if n_cols < 2:
        return Setting()
    key = bucket(n_cols, (job.n_rows + 63) // 64)
    with _LOCK:
        if key in _DECISIONS:
            return _DECISIONS[key]
    # K=2 kernel, on the pair level alone.
    k2 = ["valu"] + (["mfma"] if n_cols >= 128 else [])
    if _row_pairs_ratio(job, min_count, n_cols) < ROWS_TRY_MAX_R:
        k2.append("rows")
    best_k2 = k2[0]
    if len(k2) > 1:
        t2 = {}
        for k in k2:
            try:
                t2[k] = min(_timed(job, min_count, d_ptr, d_ids, 2, k2_kernel=k) for _ in range(2))
            except ValueError:  # the row-wise kernel refuses rows whose ids do not ascend strictly
                pass
        best_k2 = min(t2, key=t2.get)
    best = Setting(k2_kernel=best_k2)
    defaults = {"k3_schedule": "auto", "small_max": 0}
    if job.max_length is None or job.max_length >= 3:
        # The schedule, then the class split, over the whole lattice (wall time: the schedules differ in
        # launches outside the counting events too).
        for field, values in (("k3_schedule", ("auto", "all", "live")), ("small_max", (0, 4, 16))):
            times = {}
            for v in values:
                cand = Setting(**{**best.__dict__, field: v})
                _wall(job, min_count, d_ptr, d_ids, **cand.kwargs())  # warm: arena sized, code loaded
                times[v] = min(_wall(job, min_count, d_ptr, d_ids, **cand.kwargs()) for _ in range(3))
            best = Setting(**{**best.__dict__, field: _pick(times, defaults[field])})
    with _LOCK:
        return _DECISIONS.setdefault(key, best)


def tune(jobs, store, device) -> dict:
    """Tunes the bucket of each job in turn on the calling thread, with nothing else on the GPU (timings taken
    beside other jobs' kernels are too noisy to choose by); returns the decisions."""
    from ..core.result import _mining_min_count

    for job in jobs:
        d_ptr, d_ids = store.register(job.dataset, job.indptr, job.ids, device)
        setting_for(job, _mining_min_count(job.min_support, job.n_rows), d_ptr, d_ids,
                    store.profile(job.dataset, job.ids))
    return decisions()


def k2_kernel_for(job, min_count: int, d_ptr, d_ids, profile: np.ndarray) -> str:
    """The tuned K=2 kernel of a job's bucket (see ``setting_for``)."""
    return setting_for(job, min_count, d_ptr, d_ids, profile).k2_kernel


def decisions() -> dict:
    with _LOCK:
        return dict(_DECISIONS)


def reset() -> None:
    """Forgets every tuned bucket."""
    with _LOCK:
        _DECISIONS.clear()
