"""Test configuration: the `gpu` marker, import paths, and shared fixtures.

Tests marked `gpu` need a ROCm device and the built libetm.so; they skip when no device is visible. Every
other test runs on the CPU (the oracle, host logic, and that the C-ABI library loads and exports its symbols).
"""

from __future__ import annotations

import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT, ROOT / "et-miner_amd"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built libetm.so")
    config.addinivalue_line("markers", "slow: long-running CPU check (opt-in with ETM_SLOW=1)")


def _have_gpu() -> bool:
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    import os

    gpu = None
    for item in items:
        if "gpu" in item.keywords:
            if gpu is None:
                gpu = _have_gpu()
            if not gpu:
                item.add_marker(pytest.mark.skip(reason="no ROCm device visible"))
        if "slow" in item.keywords and os.environ.get("ETM_SLOW") != "1":
            item.add_marker(pytest.mark.skip(reason="slow check; set ETM_SLOW=1"))


def load_lattice(name: str) -> tuple:
    """({itemset tuple: count}, params) of a committed golden lattice."""
    import json

    import pyarrow.parquet as pq

    t = pq.read_table(GOLDEN / f"{name}.lattice.parquet").to_pydict()
    params = json.loads((GOLDEN / f"{name}.lattice.json").read_text())
    return {tuple(s): c for s, c in zip(t["itemset"], t["count"])}, params


def dense_rows(seed: int, n_rows: int) -> list:
    """Rows over seven items present independently (40-80 %), where a row holding item 0 also holds item 1 or
    item 2: a dense lattice whose depth-3 upper bounds fall below min_count, and whose cover of item 0 by items 1
    and 2 makes the bounds meet (the data shape of the reference's tests/test_deduction.py:72-88)."""
    import numpy as np

    rng = np.random.default_rng(seed)
    p = np.array([0.8, 0.5, 0.5, 0.6, 0.6, 0.4, 0.4])
    out = []
    for _ in range(n_rows):
        r = set(np.flatnonzero(rng.random(len(p)) < p).tolist())
        if 0 in r and not {1, 2} & r:
            r.add(1 if rng.random() < 0.5 else 2)
        out.append(tuple(sorted(r)))
    return out


def lattice_counts(rows, min_count: int, max_k: int) -> dict:
    """{itemset: count} of every itemset of at most max_k items with count >= min_count (brute force)."""
    from collections import Counter
    from itertools import combinations

    counts: Counter = Counter()
    for r in rows:
        for k in range(1, max_k + 1):
            counts.update(combinations(r, k))
    return {s: c for s, c in counts.items() if c >= min_count}
