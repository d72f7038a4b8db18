"""Row-wise K=2 / K=3 counting (GPU): bit-exact with the oracle and with the bitvector kernels.

The row-wise kernels count from the transactions (one atomic add per contained candidate) instead of ANDing
bitvectors; the reference routes sparse wide levels the same way (gpu/kernels/k2.py:79-135,
gpu/kernels/k3_rows.py:126-178). Cases: short and long rows (over 48 and over 64 positions take their own
kernels), max_length 3 and unbounded, Pascal inference, the free-set mode, and rows that do not ascend
(refused when forced, counted by the bitvectors when automatic).
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


def _rand_rows(rng, n, vocab, mean, zipf=1.1, long_every=0, long_len=0):
    w = 1.0 / np.arange(1, vocab + 1) ** zipf
    w /= w.sum()
    out = []
    for r in range(n):
        L = long_len if long_every and r % long_every == 0 else 1 + rng.poisson(mean)
        out.append(sorted(set(rng.choice(vocab, size=min(L, vocab), replace=False, p=w).tolist())))
    return out


def _mine(indptr, ids, s, ml=None, **kw):
    from et_miner.core.apriori import mine_csr
    from et_miner.core.result import as_dict

    return as_dict(mine_csr(indptr, ids, s, ml, with_counts=True, **kw))


CASES = [
    # (seed, rows, vocab, mean basket, min_support, max_length, long row every / length)
    (1, 3000, 60, 6, 0.01, None, 0, 0),
    (2, 5000, 300, 8, 0.004, 3, 0, 0),
    # Long rows share most of their ids, so their subsets explode combinatorially: bound these cases to K=3.
    (3, 2000, 120, 5, 0.01, 3, 97, 70),  # rows of 70 positions: the long-row K=2 and K=3 kernels
    (4, 1500, 80, 4, 0.02, 3, 13, 55),  # rows of 55 positions: long for K=2, short for K=3
]


@pytest.mark.parametrize("case", CASES, ids=[f"seed{c[0]}" for c in CASES])
def test_rows_forced_equals_oracle(case):
    seed, n, vocab, mean, s, ml, every, long_len = case
    rows = _rand_rows(np.random.default_rng(seed), n, vocab, mean, long_every=every, long_len=long_len)
    indptr, ids = _csr(rows)
    want = miner.mine(indptr, ids, s, ml).as_dict()
    stats = []
    got = _mine(indptr, ids, s, ml, k2_kernel="rows", k3_kernel="rows", stats_out=stats)
    assert got == want
    kinds = {st.k: st.kernel for st in stats}
    assert kinds.get(2) == "rows"
    if 3 in kinds:
        assert kinds[3] == "rows"


@pytest.mark.parametrize("infer", [False, True])
def test_rows_with_inference_and_free_sets(infer):
    rows = _rand_rows(np.random.default_rng(11), 4000, 90, 7)
    indptr, ids = _csr(rows)
    from et_miner.core.apriori import mine_csr
    from et_miner.core.result import as_dict

    want = miner.mine(indptr, ids, 0.01, None, use_generator_pruning=infer).as_dict()
    got = as_dict(mine_csr(indptr, ids, 0.01, None, use_generator_pruning=infer, with_counts=True,
                           k2_kernel="rows", k3_kernel="rows"))
    assert got == want
    want_free = miner.mine(indptr, ids, 0.01, None, prune_equal_support=True).as_dict()
    got_free = as_dict(mine_csr(indptr, ids, 0.01, None, prune_equal_support=True, with_counts=True,
                                k2_kernel="rows", k3_kernel="rows"))
    assert got_free == want_free


def test_rows_match_bitvector_kernels_on_a_wide_sparse_level():
    """A wide vocabulary at a low threshold: the automatic choice takes the rows; every count equals the
    bitvector kernels' (VALU tiles / MFMA Gram)."""
    from workloads import workload

    w = workload("wide_vocab_2m", 60_000)
    d = w.generate()
    auto_stats = []
    auto = _mine(d.indptr, d.indices.astype(np.int64), w.min_support, 3, stats_out=auto_stats)
    bitvec = _mine(d.indptr, d.indices.astype(np.int64), w.min_support, 3, k2_kernel="valu", k3_kernel="valu")
    rows = _mine(d.indptr, d.indices.astype(np.int64), w.min_support, 3, k2_kernel="rows", k3_kernel="rows")
    assert rows == bitvec
    assert auto == bitvec


def test_unsorted_rows():
    """Rows with repeated or unordered ids: forced row-wise counting refuses them, the automatic path counts
    them from the bitvectors (exact)."""
    rng = np.random.default_rng(5)
    rows = [list(rng.choice(40, size=6, replace=True)) for _ in range(800)]
    indptr, ids = _csr(rows)
    with pytest.raises(ValueError, match="ascending"):
        _mine(indptr, ids, 0.02, 3, k2_kernel="rows")
    want = miner.mine(indptr, ids, 0.02, 3).as_dict()
    assert _mine(indptr, ids, 0.02, 3) == want


def test_operator_wrappers_match_tiles():
    """count_pairs_k2_rows equals count_pairs_k2_shared, and count_k3_rows equals count_level, pair by pair."""
    import torch

    from et_miner.gpu import device as dev
    from et_miner.gpu import kernels as K

    rows = _rand_rows(np.random.default_rng(8), 3000, 70, 6, long_every=50, long_len=66)
    indptr, ids = _csr(rows)
    n_rows = len(rows)
    d_ptr, d_ids = dev.to_device(indptr), dev.to_device(ids)
    lo, hi = K.ids_minmax(d_ids)
    span = hi - lo + 1
    n_cols = span
    lut = dev.to_device(np.arange(span, dtype=np.int32))
    bv = K.build_bitvecs(d_ptr, d_ids, n_rows, lo, span, lut, n_cols)
    n_words = K.words_for(n_rows)
    colpos = dev.to_device(np.arange(n_cols, dtype=np.int32))
    drows = K.upload_rows(d_ptr, d_ids, n_rows, lo, span, lut, colpos)
    assert drows.ascending and drows.longest == 66
    k2_rows = dev.to_host(K.count_pairs_k2_rows(drows, n_cols))
    k2_tiles = dev.to_host(K.count_pairs_k2_shared(bv, np.arange(n_cols), n_words, layout=K.LAYOUT_LEX))
    assert np.array_equal(k2_rows[: len(k2_tiles)], k2_tiles)
    # K=3 over every pair of the K=2 level with a nonzero count (a level of n2 lexsorted pairs).
    iu, ju = np.triu_indices(n_cols, 1)
    keep = k2_tiles > 0
    rows2 = np.stack([iu[keep], ju[keep]], axis=1).astype(np.int32)
    d_rows2 = dev.to_device(rows2)
    groups = K.level_groups(d_rows2, 2)
    graph = K.build_k3_graph(d_rows2, colpos, n_cols, groups)
    k3_rows = dev.to_host(K.count_k3_rows(drows, graph, groups))
    status = dev.zeros(max(1, groups.n_cands), "uint8")
    counts = dev.zeros(max(1, groups.n_cands), "int32")
    K.count_level(bv, n_words, d_rows2, 2, groups, status, counts)
    torch.cuda.synchronize()
    assert np.array_equal(k3_rows[: groups.n_cands], dev.to_host(counts)[: groups.n_cands])
