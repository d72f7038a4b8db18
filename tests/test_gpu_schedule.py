"""The launch schedules of the counting levels give the oracle's lattice: every group and tile-pair launched or
only the live ones (k3_schedule), and each split between the wave-per-group kernels and the tiled kernel
(small_max), on full-size workloads checked against the oracle's committed per-level digests."""

from __future__ import annotations

import json

import numpy as np
import pytest
from conftest import GOLDEN

from golden.digests import table_digests

pytestmark = pytest.mark.gpu

DIGESTS = json.loads((GOLDEN / "workload_digests.json").read_text())


def _want(name):
    return [{"k": d["k"], "n": d["n"], "sha256": d["sha256"]} for d in DIGESTS["workloads"][name]["levels"]]


@pytest.fixture(scope="module")
def data():
    from workloads import workload

    return {n: workload(n).generate() for n in ("plain_60k", "deep_k_1m")}


@pytest.mark.parametrize("name", ["plain_60k", "deep_k_1m"])
@pytest.mark.parametrize("k3_schedule", ["all", "live"])
@pytest.mark.parametrize("small_max", [4, 8, 16])
@pytest.mark.parametrize("infer", [False, True])
def test_schedules_match_oracle(data, name, k3_schedule, small_max, infer):
    from et_miner.core.apriori import mine_csr
    from workloads import workload

    w = workload(name)
    d = data[name]
    print(f"{name}: k3_schedule={k3_schedule} small_max={small_max} infer={infer}", flush=True)
    got = mine_csr(d.indptr, d.indices.astype(np.int64), w.min_support, w.max_length, use_generator_pruning=infer,
                   with_counts=True, k3_schedule=k3_schedule, small_max=small_max)
    assert table_digests(got) == _want(name)


def test_bad_schedule_options_raise():
    from et_miner.core.apriori import mine_csr

    indptr = np.array([0, 2, 4], dtype=np.int64)
    ids = np.array([1, 2, 1, 2], dtype=np.int64)
    with pytest.raises(ValueError, match="k3_schedule"):
        mine_csr(indptr, ids, 0.5, k3_schedule="sometimes")
    with pytest.raises(ValueError, match="small_max"):
        mine_csr(indptr, ids, 0.5, small_max=5)
