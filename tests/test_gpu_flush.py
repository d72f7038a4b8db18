"""output_dir / resume_from_k on the GPU routes (the reference's tests/test_row_split_e2e.py:88-91,
tests/test_routing_contracts.py:150-199, tests/test_free_set_semantics.py:411-431).

The flushed levels equal the oracle's lattice; a run resumed from level 2 or 3 of an earlier run's files writes
exactly the levels the uninterrupted run wrote (plain, free-set and inference modes); the bitvecs route flushes
and resumes the same way.
"""

from __future__ import annotations

import shutil
from pathlib import Path

import numpy as np
import pyarrow.parquet as pq
import pytest

from oracle import miner

pytestmark = pytest.mark.gpu


def _rows(seed, n=6000, vocab=40, motifs=2, motif_len=6, frac=0.05):
    rng = np.random.default_rng(seed)
    w = 1.0 / np.arange(1, vocab + 1) ** 1.0
    w /= w.sum()
    planted = [rng.choice(vocab, size=motif_len, replace=False) for _ in range(motifs)]
    out = []
    for _ in range(n):
        r = set(rng.choice(vocab, size=1 + rng.poisson(4), replace=False, p=w).tolist())
        for m in planted:
            if rng.random() < frac:
                r |= set(m.tolist())
        out.append(sorted(r))
    return out


def _csr(rows):
    indptr = np.zeros(len(rows) + 1, dtype=np.int64)
    indptr[1:] = np.cumsum([len(r) for r in rows])
    return indptr, np.array([x for r in rows for x in r], dtype=np.int64)


def _levels(d: Path, n_rows: int) -> dict:
    out = {}
    for p in sorted(d.glob("frequent_k*.parquet")):
        k = int(p.stem.split("frequent_k")[1])
        t = pq.read_table(p)
        sets = [tuple(x) for x in t.column("itemset").to_pylist()]
        counts = np.rint(np.asarray(t.column("support").to_numpy()) * n_rows).astype(np.int64).tolist()
        out[k] = dict(zip(sets, counts))
    return out


def _flat(levels: dict) -> dict:
    return {s: c for lv in levels.values() for s, c in lv.items()}


MODES = {"plain": {}, "free": {"prune_equal_support": True}, "infer": {"use_generator_pruning": True}}


@pytest.mark.parametrize("mode", list(MODES))
def test_output_dir_holds_the_oracle_lattice(mode, tmp_path):
    from et_miner import apriori

    rows = _rows(1)
    got = apriori(rows, min_support=0.01, use_gpu=True, output_dir=str(tmp_path), **MODES[mode])
    assert len(got) == 0, "the output_dir route returns an empty frame; the levels are on disk"
    indptr, ids = _csr(rows)
    want = miner.mine(indptr, ids, 0.01, None, prune_equal_support=mode == "free",
                      use_generator_pruning=mode == "infer").as_dict()
    levels = _levels(tmp_path, len(rows))
    assert _flat(levels) == want
    assert all(len(s) == k for k, lv in levels.items() for s in lv)


@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("resume_from_k", [2, 3])
def test_resume_reproduces_the_uninterrupted_run(mode, resume_from_k, tmp_path):
    from et_miner import apriori

    rows = _rows(2)
    full, resumed = tmp_path / "full", tmp_path / "resumed"
    apriori(rows, min_support=0.01, use_gpu=True, output_dir=str(full), **MODES[mode])
    a = _levels(full, len(rows))
    assert len(a) > resume_from_k + 1, f"the fixture must reach past K={resume_from_k + 1}: {sorted(a)}"
    resumed.mkdir()
    for p in full.glob("frequent_k*.parquet"):
        if int(p.stem.split("frequent_k")[1]) <= resume_from_k:
            shutil.copy(p, resumed / p.name)
    apriori(rows, min_support=0.01, use_gpu=True, output_dir=str(resumed), resume_from_k=resume_from_k,
            **MODES[mode])
    b = _levels(resumed, len(rows))
    assert sorted(a) == sorted(b)
    for k in a:
        assert a[k] == b[k], f"K={k} differs between the full and the resumed run"


def test_bitvecs_route_flushes_and_resumes(tmp_path):
    """The caller's bitvectors (item ids 100 + column): output_dir equals the in-memory result; resume from
    level 2 reproduces the levels above it."""
    from et_miner import apriori
    from et_miner.core.result import as_dict
    from et_miner.gpu import device as dev
    from et_miner.gpu import kernels as K

    rows = _rows(3, n=5000)
    indptr, ids = _csr(rows)
    n = len(rows)
    cols = np.unique(ids)
    lut = np.full(int(ids.max()) + 1, -1, np.int32)
    lut[cols] = np.arange(len(cols), dtype=np.int32)
    bv = K.build_bitvecs(dev.to_device(indptr), dev.to_device(ids), n, 0, int(ids.max()) + 1, dev.to_device(lut),
                         len(cols), pitch=K.words_for(n))
    col_to_item = {c: int(100 + it) for c, it in enumerate(cols)}
    in_memory = apriori(bitvecs=(bv, col_to_item, n), min_support=0.02)
    want = {k: round(v * n) for k, v in as_dict(in_memory).items()}
    full = tmp_path / "full"
    assert len(apriori(bitvecs=(bv, col_to_item, n), min_support=0.02, output_dir=str(full))) == 0
    a = _levels(full, n)
    assert _flat(a) == want
    resumed = tmp_path / "resumed"
    resumed.mkdir()
    for k in (1, 2):
        shutil.copy(full / f"frequent_k{k}.parquet", resumed / f"frequent_k{k}.parquet")
    apriori(bitvecs=(bv, col_to_item, n), min_support=0.02, output_dir=str(resumed), resume_from_k=2)
    assert _levels(resumed, n) == a


def test_resuming_a_foreign_level_raises(tmp_path):
    from et_miner import apriori

    apriori(_rows(4), min_support=0.01, use_gpu=True, output_dir=str(tmp_path))
    with pytest.raises(ValueError, match="not frequent items"):
        apriori([[x + 1000 for x in r] for r in _rows(4)], min_support=0.01, use_gpu=True,
                output_dir=str(tmp_path), resume_from_k=2)
