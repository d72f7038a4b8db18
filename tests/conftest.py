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
