"""The drop-in boundary on the CPU: the C-ABI library and the apriori() validation contract.

No compute entry point is called here (there is no GPU); the library must load and export every function
include/etm.h declares, with the ctypes signatures matching the header's parameter counts.
"""

from __future__ import annotations

import re

import numpy as np
import pytest
from conftest import ROOT

from et_miner import _lib, apriori
from et_miner.core.apriori import transactions_to_csr
from et_miner.core.result import _min_count


def _header_functions() -> dict:
    text = (ROOT / "include" / "etm.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = {}
    for m in re.finditer(r"\b(?:int|int64_t|size_t)\s+(etm_\w+)\s*\(([^)]*)\)\s*;", text):
        params = [p for p in m.group(2).split(",") if p.strip() and p.strip() != "void"]
        out[m.group(1)] = len(params)
    return out


def test_library_exports_every_declared_symbol():
    declared = _header_functions()
    assert len(declared) >= 15
    lib = _lib.lib()
    for name in declared:
        assert hasattr(lib, name), f"libetm.so does not export {name}"


def test_ctypes_signatures_match_header():
    declared = _header_functions()
    assert set(declared) == set(_lib.SIGNATURES)
    for name, n_params in declared.items():
        assert len(_lib.SIGNATURES[name][1]) == n_params, name


def test_abi_version_and_errors():
    lib = _lib.lib()
    assert lib.etm_abi_version() == _lib.ABI_VERSION
    assert _lib.device_count() >= 0
    # A guard failure maps to ValueError with the library's message, before any launch.
    with pytest.raises(ValueError, match="id_bytes"):
        _lib.call("etm_ids_minmax", 1, 3, 10, 1, None)
    with pytest.raises(ValueError, match="power of two"):
        _lib.call("etm_level_hash", 1, 4, 2, 1, 12, None)


def test_min_count_contract():
    assert _min_count(0.07, 10_000) == 700
    assert _min_count(0.07, 100) == 7
    assert _min_count(1e-07, 76_890_945) == 8


@pytest.mark.parametrize(
    "kwargs,exc",
    [
        (dict(min_support="0.1", use_gpu=True), TypeError),
        (dict(min_support=1.5, use_gpu=True), ValueError),
        (dict(min_support=0.1, max_length=0, use_gpu=True), ValueError),
        (dict(min_support=0.1, max_length=2.0, use_gpu=True), TypeError),
        (dict(min_support=0.1, batch_size=0, use_gpu=True), ValueError),
        (dict(min_support=0.1, prune_apriori=False, use_gpu=True), ValueError),
        (dict(min_support=0.1, sparse=True, use_gpu=True), ValueError),
        (dict(min_support=0.1, sparse_from_k=3, use_gpu=True), ValueError),
        (dict(min_support=0.1, gpu_resident=True, use_gpu=True), ValueError),
        (dict(min_support=0.1, streaming=True, prune_equal_support=True, use_gpu=True), ValueError),
        (dict(min_support=0.1, memory_budget_gb=1.0, use_gpu=True), ValueError),
        (dict(min_support=0.1, n_gpus=2, use_gpu=True), ValueError),
        (dict(min_support=0.1, output_dir="/tmp/x", use_gpu=True, streaming=True), ValueError),
        (dict(min_support=0.1, resume_from_k=2, use_gpu=True, streaming=True), ValueError),
        (dict(min_support=0.1, anchor_items={1}, use_gpu=True), ValueError),
        (dict(min_support=0.1), ValueError),  # the CPU route is not part of this build
    ],
)
def test_validation_matches_reference(kwargs, exc):
    with pytest.raises(exc):
        apriori([[1, 2], [2, 3]], **kwargs)


def test_bitvecs_route_validation():
    class FakeDev:
        __cuda_array_interface__ = {"data": (0, False), "shape": (2, 3), "typestr": "<u8", "version": 3}

    with pytest.raises(ValueError, match="both"):
        apriori([[1]], bitvecs=(FakeDev(), {0: 1, 1: 2}, 130), use_gpu=True)
    with pytest.raises(ValueError, match="tuple"):
        apriori(bitvecs=(FakeDev(), {}), use_gpu=True)
    with pytest.raises(ValueError, match="ceil"):
        apriori(bitvecs=(FakeDev(), {0: 1, 1: 2}, 64))
    with pytest.raises(ValueError, match="keys"):
        apriori(bitvecs=(FakeDev(), {0: 1}, 130))
    with pytest.raises(ValueError, match="integer"):
        apriori(bitvecs=(FakeDev(), {0: 1.5, 1: 2}, 130))
    with pytest.raises(TypeError):
        apriori(bitvecs=(np.zeros((2, 3), np.uint64), {0: 1, 1: 2}, 130))


def test_transactions_to_csr_inputs():
    import pyarrow as pa

    rows = [[3, 1, 2], [], None, [2, None, 5]]
    indptr, ids, items = transactions_to_csr(rows)
    assert items is None
    assert indptr.tolist() == [0, 3, 3, 3, 5] and ids.tolist() == [3, 1, 2, 2, 5]
    t = pa.table({"basket": pa.array([[1, 2], [2]])})
    indptr, ids, _ = transactions_to_csr(t, "basket")
    assert indptr.tolist() == [0, 2, 3] and ids.tolist() == [1, 2, 2]
    indptr, ids, items = transactions_to_csr([["b", "a"], ["a"]])
    assert items.to_pylist() == ["a", "b"] and ids.tolist() == [1, 0, 0]
    import pandas as pd

    indptr, ids, _ = transactions_to_csr(pd.DataFrame({"items": [[4, 5], [5]]}))
    assert ids.tolist() == [4, 5, 5]


def test_deduction_switch(monkeypatch):
    """ET_MINER_DEDUCTION: unset or "1" keeps the bounds, "0" turns them off, anything else raises (the
    reference's _env.py:285-291)."""
    from et_miner import _lib

    monkeypatch.delenv("ET_MINER_DEDUCTION", raising=False)
    assert _lib.deduction_setting() == _lib.DEDUCTION_AUTO
    monkeypatch.setenv("ET_MINER_DEDUCTION", "1")
    assert _lib.deduction_setting() == _lib.DEDUCTION_AUTO
    monkeypatch.setenv("ET_MINER_DEDUCTION", "0")
    assert _lib.deduction_setting() == _lib.DEDUCTION_OFF
    monkeypatch.setenv("ET_MINER_DEDUCTION", "2")
    with pytest.raises(ValueError, match="ET_MINER_DEDUCTION"):
        _lib.deduction_setting()


def test_frequent_first_ranks():
    """Ranks 0.. go to the columns of most occurrences, ties by column; past n_ranked, -1."""
    from et_miner.gpu.kernels import frequent_first_ranks

    r = frequent_first_ranks(np.array([5, 1, 9, 1, 3]), n_ranked=3)
    assert r.tolist() == [1, -1, 0, -1, 2]
    assert frequent_first_ranks(np.array([2, 2]), n_ranked=64).tolist() == [0, 1]
