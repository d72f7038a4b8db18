"""Kernel choice for the K=2 pair space per shape bucket: int8 MFMA Gram or VALU AND+popcount tiles.

The first job of a bucket (frequent columns F and words W, each rounded up to a power of two) runs its pair
level once per kernel with event timing; later jobs of the bucket use the faster one. Both kernels are exact,
so the choice never changes a result, only its time.
"""

from __future__ import annotations

import threading

import numpy as np

_DECISIONS: dict = {}
_LOCK = threading.Lock()


def bucket(n_cols: int, n_words: int) -> tuple:
    return (1 << max(0, int(n_cols - 1).bit_length()), 1 << max(0, int(n_words - 1).bit_length()))


def k2_kernel_for(job, min_count: int, d_ptr, d_ids, profile: np.ndarray) -> str:
    """The K=2 kernel for a job, tuning its bucket on first sight. ``profile`` is the dataset's ascending
    occurrence counts, so the frequent-column count is one search."""
    from ..gpu.native import mine_csr_device

    n_cols = int(len(profile) - np.searchsorted(profile, min_count, side="left"))
    if n_cols < 128:  # below one MFMA tile the VALU tiles win by construction
        return "valu"
    key = bucket(n_cols, (job.n_rows + 63) // 64)
    with _LOCK:
        if key in _DECISIONS:
            return _DECISIONS[key]
    timing = {}
    for kern in ("valu", "mfma"):
        res = mine_csr_device(d_ptr, d_ids, job.n_rows, min_count, 2, time_kernels=True, levels=False,
                              k2_kernel=kern)
        timing[kern] = sum(s.count_ms for s in res.stats if s.k == 2)
    choice = min(timing, key=timing.get)
    with _LOCK:
        return _DECISIONS.setdefault(key, choice)


def decisions() -> dict:
    with _LOCK:
        return dict(_DECISIONS)
