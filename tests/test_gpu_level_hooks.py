"""Per-level behaviour of the native loop behind apriori(): the level callback as each level completes with its
wall time, the memory guards, the per-level parquet flush, cancellation from the callback, and the bytes the
counting kernels must read (``bytes_live``), against the reference's contract (``core/apriori.py:394-408``,
``gpu/row_split.py:467-537, 855-856, 1136-1144``).
"""

from __future__ import annotations

import time

import numpy as np
import pyarrow.parquet as pq
import pytest

from oracle import miner

pytestmark = pytest.mark.gpu


def _rows(seed=3, n=6000, vocab=40):
    rng = np.random.default_rng(seed)
    w = 1.0 / np.arange(1, vocab + 1) ** 0.8
    w /= w.sum()
    return [sorted(set(rng.choice(vocab, size=1 + rng.poisson(6), p=w).tolist())) for _ in range(n)]


def _csr(rows):
    indptr = np.zeros(len(rows) + 1, dtype=np.int64)
    indptr[1:] = np.cumsum([len(r) for r in rows])
    return indptr, np.array([x for r in rows for x in r], dtype=np.int64)


@pytest.mark.parametrize("bitvecs_route", [False, True])
def test_level_callback_fires_per_level_with_wall_time(bitvecs_route):
    from et_miner import apriori
    from et_miner.core.result import as_dict

    rows = _rows()
    calls = []
    t0 = time.perf_counter()

    def cb(k, n_cand, n_freq, ms):
        calls.append((k, n_cand, n_freq, ms, time.perf_counter()))

    if bitvecs_route:
        bv, c2i, n = _bitvecs(rows)
        res = apriori(bitvecs=(bv, c2i, n), min_support=0.02, level_callback=cb)
    else:
        res = apriori(rows, min_support=0.02, use_gpu=True, level_callback=cb)
    t1 = time.perf_counter()
    got = as_dict(res) if not bitvecs_route else None
    ks = [c[0] for c in calls]
    assert ks == list(range(1, len(ks) + 1)) and len(ks) >= 4
    assert all(c[3] > 0 for c in calls), "every level reports a positive wall time"
    stamps = [c[4] for c in calls]
    assert t0 < stamps[0] <= stamps[-1] < t1 and stamps == sorted(stamps)
    # The level of k items holds n_frequent itemsets.
    want = miner.mine(*_csr(rows), 0.02).as_dict()
    for k, _, n_freq, _, _ in calls:
        assert n_freq == sum(1 for s in want if len(s) == k), k
    if got is not None:
        assert set(got) == set(want)


def _bitvecs(rows):
    import torch

    indptr, ids = _csr(rows)
    n = len(rows)
    cols = np.unique(ids)
    bv = miner.bitvectors(indptr, np.searchsorted(cols, ids).astype(np.int32), len(cols))
    d = torch.from_numpy(bv.view(np.int64)).cuda()
    return d, {int(c): int(v) for c, v in enumerate(cols)}, n


def test_profile_session_records_every_level():
    from et_miner import apriori

    _, session = apriori(_rows(), min_support=0.02, use_gpu=True, profile=True)
    phases = [p for p in session.phases if p.name.startswith("k")]
    assert [p.name for p in phases] == [f"k{i}" for i in range(1, len(phases) + 1)]
    assert all(p.duration_ms > 0 for p in phases)


def test_callback_exception_stops_mining():
    from et_miner import apriori

    seen = []

    class Stop(Exception):
        pass

    def cb(k, *_):
        seen.append(k)
        if k == 2:
            raise Stop("enough")

    with pytest.raises(Stop, match="enough"):
        apriori(_rows(), min_support=0.02, use_gpu=True, level_callback=cb)
    assert seen == [1, 2]
    # The library is still usable afterwards.
    assert len(apriori(_rows(), min_support=0.02, use_gpu=True)) > 0


@pytest.mark.parametrize("which", ["ram", "vram"])
def test_memory_guards_raise_between_levels(which):
    from et_miner import apriori

    kw = {"max_ram_gb": 1e-6} if which == "ram" else {"max_vram_gb": 1e-6}
    with pytest.raises(MemoryError, match=f"Memory guard tripped after K=1: {which.upper()}="):
        apriori(_rows(), min_support=0.02, use_gpu=True, **kw)
    # Generous limits (and None) let the call finish.
    assert len(apriori(_rows(), min_support=0.02, use_gpu=True, max_ram_gb=None, max_vram_gb=None)) > 0
    assert len(apriori(_rows(), min_support=0.02, use_gpu=True, max_ram_gb=4096.0, max_vram_gb=280.0)) > 0


def test_output_dir_writes_each_level_as_it_completes(tmp_path):
    from et_miner import apriori

    rows = _rows()
    on_disk = []

    def cb(k, n_cand, n_freq, ms):
        on_disk.append((n_freq, sorted(p.name for p in tmp_path.glob("frequent_k*.parquet"))))

    assert len(apriori(rows, min_support=0.02, use_gpu=True, output_dir=str(tmp_path), level_callback=cb)) == 0
    # When level k's callback runs, the files of levels 1..k are written (an empty level has none) and none above.
    for k, (_, names) in enumerate(on_disk, start=1):
        want_names = [f"frequent_k{i}.parquet" for i in range(1, k + 1) if on_disk[i - 1][0] > 0]
        assert names == sorted(want_names), (k, names)
    assert len(on_disk) >= 4
    want = miner.mine(*_csr(rows), 0.02).as_dict()
    got = {}
    for p in tmp_path.glob("frequent_k*.parquet"):
        t = pq.read_table(p).to_pydict()
        for s, sup in zip(t["itemset"], t["support"]):
            got[tuple(s)] = int(round(sup * len(rows)))
    assert got == want


def test_bytes_live_equals_the_counted_groups_bytes():
    """bytes_live of each level K >= 3 in the prune-only mode, against an independent count from the oracle's
    candidate generation (which applies the same subset test): (prefix items x groups with a candidate + distinct
    suffix rows of the candidates) x words x 8 + 4 per candidate."""
    from et_miner.core.apriori import mine_csr

    rows = _rows(seed=5, n=20000, vocab=60)
    indptr, ids = _csr(rows)
    stats = []
    mine_csr(indptr, ids, 0.01, stats_out=stats, live_bytes=True, with_counts=True)
    lat = miner.mine(indptr, ids, 0.01)
    n_cols = len(lat.items)
    words = (len(rows) + 63) // 64
    checked = 0
    for st in stats:
        if st.k < 3:
            assert st.bytes_live == st.bytes_alg
            continue
        prev = lat.levels[st.k - 2][0]
        cands = np.concatenate(list(miner.generate_candidates(prev, st.k, n_cols)) or [np.empty((0, st.k), np.int32)])
        assert st.n_counted == len(cands), st.k
        prefixes = {tuple(c[: st.k - 2]) for c in cands.tolist()}
        suffix_rows = {tuple(c[: st.k - 2]) + (c[j],) for c in cands.tolist() for j in (st.k - 2, st.k - 1)}
        want = ((st.k - 2) * len(prefixes) + len(suffix_rows)) * words * 8 + 4 * len(cands)
        assert st.bytes_live == want, (st.k, st.bytes_live, want)
        assert st.bytes_live <= st.bytes_alg
        checked += 1
    assert checked >= 2
