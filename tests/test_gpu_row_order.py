"""Row order and the tile-occupancy skip (GPU).

The row order permutes bit positions only, so every count must be unchanged: the ordered bitvectors equal the
oracle's bitvectors of the permuted rows, the order is a stable sort of the host-computed keys, the occupancy
equals a host reduction of the bitvectors, group counts with the skip equal counts without it, and whole
lattices under "lex", "keep" and "auto" equal the oracle's. The reference holds the same contract
(gpu/input_layer.py:469-480: "Counts do not depend on it").
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import miner

pytestmark = pytest.mark.gpu


def _csr(rows):
    indptr = np.zeros(len(rows) + 1, dtype=np.int64)
    indptr[1:] = np.cumsum([len(r) for r in rows])
    ids = np.array([x for r in rows for x in r], dtype=np.int64)
    return indptr, ids


def _planted_rows(seed, n, vocab=60, mean=5, motifs=3, motif_len=7, frac=0.04, dup_every=0):
    """Zipf baskets plus motifs planted whole in a fraction of rows; every dup_every-th row repeats an id."""
    rng = np.random.default_rng(seed)
    w = 1.0 / np.arange(1, vocab + 1) ** 1.0
    w /= w.sum()
    planted = [rng.choice(vocab, size=motif_len, replace=False) for _ in range(motifs)]
    rows = []
    for r in range(n):
        row = set(rng.choice(vocab, size=min(vocab, 1 + rng.poisson(mean)), replace=False, p=w).tolist())
        for m in planted:
            if rng.random() < frac:
                row |= set(m.tolist())
        row = sorted(row)
        if dup_every and r % dup_every == 0 and row:
            row = row + [row[0]]
        rows.append(row)
    return rows


def _device_input(rows):
    from et_miner.gpu import device as dev
    from et_miner.gpu import kernels as K

    indptr, ids = _csr(rows)
    n = len(rows)
    d_ptr, d_ids = dev.to_device(indptr), dev.to_device(ids)
    lo, hi = K.ids_minmax(d_ids)
    span = hi - lo + 1
    occ = np.bincount(ids - lo, minlength=span)
    cols = np.flatnonzero(occ > 0)
    lut = np.full(span, -1, dtype=np.int32)
    lut[cols] = np.arange(len(cols), dtype=np.int32)
    rank = K.frequent_first_ranks(occ[cols])
    return indptr, ids, n, d_ptr, d_ids, lo, span, lut, cols, rank


def _host_keys(indptr, ids, lo, lut, rank, n_ranked):
    n = len(indptr) - 1
    keys = np.zeros(n, dtype=np.uint64)
    for r in range(n):
        for v in ids[indptr[r] : indptr[r + 1]]:
            rk = rank[lut[v - lo]]
            if 0 <= rk < n_ranked:
                keys[r] |= np.uint64(1) << np.uint64(63 - rk)
    return keys


@pytest.mark.parametrize("n_ranked", [64, 20, 3])
@pytest.mark.parametrize("n_rows", [1, 63, 64, 1000, 20_011])
def test_order_is_a_stable_key_sort_and_bitvectors_follow_it(n_rows, n_ranked):
    """Ranks in 64, 24 or 8 key bits: the sort covers whole 8-bit digits from the top (60 of 64 columns ranked
    when n_ranked is 64)."""
    from et_miner.gpu import device as dev
    from et_miner.gpu import kernels as K

    rows = _planted_rows(3, n_rows, dup_every=7)
    indptr, ids, n, d_ptr, d_ids, lo, span, lut, cols, rank = _device_input(rows)
    d_lut, d_rank = dev.to_device(lut), dev.to_device(rank)
    n_ranked = int(min(n_ranked, len(cols)))
    order = dev.to_host(K.row_order(d_ptr, d_ids, n, lo, span, d_lut, d_rank, n_ranked))
    keys = _host_keys(indptr, ids, lo, lut, rank, n_ranked)
    assert np.array_equal(order, np.argsort(keys, kind="stable"))
    bv = dev.to_host(K.build_bitvecs(d_ptr, d_ids, n, lo, span, d_lut, len(cols),
                                     order=dev.to_device(order.astype(np.int32))))
    perm_rows = [rows[i] for i in order]
    p_indptr, p_ids = _csr(perm_rows)
    want = miner.bitvectors(p_indptr, lut[p_ids - lo], len(cols))
    assert np.array_equal(bv[:, : want.shape[1]].view(np.uint64), want)
    assert not bv[:, want.shape[1]:].any()


def test_ordered_build_of_a_wide_column_set():
    """About 1,500 columns (the LDS-tile builder, two tiles) through an order: the oracle's bitvectors of the
    permuted rows."""
    from et_miner.gpu import device as dev
    from et_miner.gpu import kernels as K

    rows = _planted_rows(4, 9_001, vocab=3000, mean=8, dup_every=3)
    indptr, ids, n, d_ptr, d_ids, lo, span, lut, cols, rank = _device_input(rows)
    assert len(cols) > 1024
    d_lut = dev.to_device(lut)
    order = dev.to_host(K.row_order(d_ptr, d_ids, n, lo, span, d_lut, dev.to_device(rank), 64))
    bv = dev.to_host(K.build_bitvecs(d_ptr, d_ids, n, lo, span, d_lut, len(cols),
                                     order=dev.to_device(order.astype(np.int32))))
    p_indptr, p_ids = _csr([rows[i] for i in order])
    want = miner.bitvectors(p_indptr, lut[p_ids - lo], len(cols))
    assert np.array_equal(bv[:, : want.shape[1]].view(np.uint64), want)
    assert not bv[:, want.shape[1]:].any()


def test_occupancy_and_skip_keep_every_count():
    from et_miner.gpu import device as dev
    from et_miner.gpu import kernels as K

    rows = _planted_rows(5, 300_000, vocab=40, mean=3, motifs=2, motif_len=8, frac=0.02)
    indptr, ids, n, d_ptr, d_ids, lo, span, lut, cols, rank = _device_input(rows)
    d_lut, d_rank = dev.to_device(lut), dev.to_device(rank)
    order = K.row_order(d_ptr, d_ids, n, lo, span, d_lut, d_rank, int(min(64, len(cols))))
    bv = K.build_bitvecs(d_ptr, d_ids, n, lo, span, d_lut, len(cols), order=order)
    n_words = K.words_for(n)
    occ = K.tile_occupancy(bv, n_words)
    host_bv = dev.to_host(bv).view(np.uint64)[:, :n_words]
    tiles = -(-n_words // 64)
    padded = np.zeros((len(cols), tiles * 64), dtype=np.uint64)
    padded[:, :n_words] = host_bv
    nz = (padded.reshape(len(cols), tiles, 64) != 0).any(axis=2)
    want = np.zeros((len(cols), occ.pitch), dtype=np.uint64)
    for t in range(tiles):
        want[nz[:, t], t // 64] |= np.uint64(1) << np.uint64(t % 64)
    assert np.array_equal(dev.to_host(occ.bits).view(np.uint64), want)
    # A level of 4-itemsets (every 4-subset of the columns' frequent pairs' closure) counted with and without it.
    lat = miner.mine(indptr, ids, 0.004, 4)
    rows4 = lat.levels[3][0] if len(lat.levels) > 3 else None
    assert rows4 is not None and len(rows4) > 0
    col_of = {int(it): c for c, it in enumerate(cols + lo)}
    items = lat.items
    lvl = np.vectorize(lambda x: col_of[int(items[x])])(rows4).astype(np.int32)
    lvl = lvl[np.lexsort(lvl.T[::-1])]
    d_lvl = dev.to_device(lvl)
    groups = K.level_groups(d_lvl, 4)
    assert groups.n_small > 0
    got = []
    for o in (None, occ):
        counts = dev.zeros(max(1, groups.n_cands), "int32")
        K.count_level(bv, n_words, d_lvl, 4, groups, None, counts, occ=o)
        got.append(dev.to_host(counts)[: groups.n_cands])
    assert np.array_equal(got[0], got[1])
    assert got[0].sum() > 0


@pytest.mark.parametrize("row_order", ["lex", "keep", "auto"])
@pytest.mark.parametrize("case", ["planted", "dups_and_empty", "small"])
def test_mining_under_each_order_matches_the_oracle(case, row_order):
    from et_miner.core.apriori import mine_csr
    from et_miner.core.result import as_dict

    if case == "planted":
        rows, s, ml = _planted_rows(7, 40_000), 0.01, None
    elif case == "dups_and_empty":
        rows = _planted_rows(8, 30_001, dup_every=5)
        rows = [[] if i % 11 == 0 else r for i, r in enumerate(rows)]
        s, ml = 0.012, None
    else:
        rows, s, ml = _planted_rows(9, 700), 0.03, 6
    indptr, ids = _csr(rows)
    want = miner.mine(indptr, ids, s, ml).as_dict()
    got = as_dict(mine_csr(indptr, ids, s, ml, with_counts=True, row_order=row_order))
    assert got == want
    inferred = as_dict(mine_csr(indptr, ids, s, ml, with_counts=True, row_order=row_order,
                                use_generator_pruning=True))
    assert inferred == miner.mine(indptr, ids, s, ml, use_generator_pruning=True).as_dict()
