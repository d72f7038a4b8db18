"""Parity of the HIP counting path with the CPU oracle (GPU only, through the C ABI).

Integer, byte and index work must be bit-exact: every itemset and every count equals the oracle's.
Known answers follow the reference's own preflight and unit tests (bench/selfcheck.py:136-181,
tests/test_k2_counting.py:102-136).
"""

from __future__ import annotations

import numpy as np
import pytest
from conftest import load_lattice

from oracle import miner

pytestmark = pytest.mark.gpu


def _csr(rows):
    indptr = np.zeros(len(rows) + 1, dtype=np.int64)
    indptr[1:] = np.cumsum([len(r) for r in rows])
    ids = np.array([x for r in rows for x in r], dtype=np.int64)
    return indptr, ids


ENGINES = ["native", "python"]


def _gpu_mine(indptr, ids, s, ml=None, engine="native", **kw):
    from et_miner.core.apriori import mine_csr
    from et_miner.core.result import as_dict

    return as_dict(mine_csr(indptr, ids, s, ml, with_counts=True, engine=engine, **kw))


def _rand_rows(rng, n, vocab, mean, zipf=1.1, dup=False):
    w = 1.0 / np.arange(1, vocab + 1) ** zipf
    w /= w.sum()
    lens = 1 + rng.poisson(mean, size=n)
    out = []
    for L in lens:
        r = rng.choice(vocab, size=L, replace=True, p=w).tolist()
        out.append(r if dup else sorted(set(r)))
    return out


def test_native_library_is_loaded():
    from et_miner import _lib

    assert _lib.device_count() >= 1
    assert _lib.lib()._name.endswith("libetm.so")


def test_bitvec_known_answer():
    """rows {0,2},{1},{0,1,2},{} -> column words 0b0101, 0b0110, 0b0101 (reference bench/selfcheck.py:136)."""
    from et_miner.gpu import device as dev
    from et_miner.gpu import kernels as K

    indptr, ids = _csr([[0, 2], [1], [0, 1, 2], []])
    d_ptr, d_ids = dev.to_device(indptr), dev.to_device(ids)
    lut = dev.to_device(np.arange(3, dtype=np.int32))
    bv = dev.to_host(K.build_bitvecs(d_ptr, d_ids, 4, 0, 3, lut, 3)).view(np.uint64)
    assert bv[:, 0].tolist() == [0b0101, 0b0110, 0b0101]
    assert not bv[:, 1:].any()


@pytest.mark.parametrize("n_rows,vocab", [(1, 300), (63, 300), (64, 300), (65, 300), (4095, 300), (4096, 300),
                                          (4097, 300), (20_000, 300), (1, 700), (513, 700), (20_000, 700),
                                          (65, 5000), (4097, 5000), (20_000, 5000)])
def test_bitvec_build_and_popcount_match_oracle(n_rows, vocab):
    """Every other item of the vocabulary gets a column: about 150 or 350 columns (one LDS tile of 1,024
    columns), or about 2,000 (two LDS tiles, the last partial)."""
    from et_miner.gpu import device as dev
    from et_miner.gpu import kernels as K

    rng = np.random.default_rng(n_rows + vocab)
    rows = _rand_rows(rng, n_rows, vocab, 6, zipf=0.8, dup=True)
    indptr, ids = _csr(rows)
    ids = ids + 1000  # a nonzero id base
    lo, span = int(ids.min()), int(ids.max() - ids.min() + 1)
    lut = np.full(span, -1, np.int32)
    keep = np.unique(ids)[::2]  # every other item gets a column
    lut[keep - lo] = np.arange(len(keep), dtype=np.int32)
    for id_dtype in (np.int64, np.int32):
        bv = K.build_bitvecs(dev.to_device(indptr), dev.to_device(ids.astype(id_dtype)), n_rows, lo, span,
                             dev.to_device(lut), len(keep))
        got = dev.to_host(bv).view(np.uint64)
        col = lut[ids - lo]
        sel = col >= 0
        rr = np.repeat(np.arange(n_rows), np.diff(indptr))[sel]
        keys = np.unique(rr * len(keep) + col[sel])
        ptr = np.zeros(n_rows + 1, np.int64)
        np.cumsum(np.bincount(keys // len(keep), minlength=n_rows), out=ptr[1:])
        want = miner.bitvectors(ptr, (keys % len(keep)).astype(np.int32), len(keep))
        w = K.words_for(n_rows)
        assert np.array_equal(got[:, :w], want)
        assert not got[:, w:].any(), "padding words must be zero"
        pc = dev.to_host(K.column_popcounts(bv, w))
        assert np.array_equal(pc, miner.popcount_rows(want))


def _numpy_pairs(bv: np.ndarray, cols: np.ndarray) -> dict:
    out = {}
    for a in range(len(cols)):
        for b in range(a + 1, len(cols)):
            out[(a, b)] = int(np.bitwise_count(bv[cols[a]] & bv[cols[b]]).sum())
    return out


@pytest.mark.parametrize("kernel", ["valu", "mfma"])
@pytest.mark.parametrize("n_cols,n_words", [(2, 1), (5, 7), (16, 130), (17, 64), (70, 33), (130, 100), (300, 21)])
def test_k2_counts_both_layouts(n_cols, n_words, kernel):
    from et_miner.gpu import device as dev
    from et_miner.gpu import kernels as K

    rng = np.random.default_rng(n_cols * 1000 + n_words)
    bv = rng.integers(0, 2**63, size=(n_cols + 3, n_words), dtype=np.int64).view(np.uint64)
    bv &= rng.integers(0, 2**63, size=bv.shape, dtype=np.int64).view(np.uint64)
    cols = np.sort(rng.choice(n_cols + 3, size=n_cols, replace=False)).astype(np.int32)
    d_bv = dev.to_device(bv)
    want = _numpy_pairs(bv, cols)
    tri = dev.to_host(K.count_pairs_k2_shared(d_bv, cols, n_words, layout=K.LAYOUT_TRI, kernel=kernel))
    lex = dev.to_host(K.count_pairs_k2_shared(d_bv, cols, n_words, layout=K.LAYOUT_LEX, kernel=kernel))
    s = len(cols)
    for (a, b), c in want.items():
        assert tri[b * (b - 1) // 2 + a] == c
        assert lex[a * (2 * s - a - 1) // 2 + (b - a - 1)] == c


def test_k2_known_answer_128_rows():
    """128 rows: item 0 in all, 1 in the first 64, 2 in rows 64-95 (reference tests/test_k2_counting.py:102)."""
    rows = [[0] + ([1] if r < 64 else []) + ([2] if 64 <= r < 96 else []) for r in range(128)]
    got = _gpu_mine(*_csr(rows), 0.1, 2)
    assert got[(0, 1)] == 64 and got[(0, 2)] == 32 and (1, 2) not in got


def test_readme_and_boundary():
    got = _gpu_mine(*_csr([[1, 2, 3], [2, 3, 4], [1, 3, 5], [2, 3]]), 0.5)
    assert got == {(1,): 2, (2,): 3, (3,): 4, (1, 3): 2, (2, 3): 3}
    rows = [[0, 1, 2]] * 700 + [[1, 2]] * 9300
    got = _gpu_mine(*_csr(rows), 0.07)
    for s in [(0,), (0, 1), (0, 2), (0, 1, 2)]:
        assert got[s] == 700


@pytest.mark.parametrize("engine", ENGINES)
def test_support_column_is_count_over_rows(engine):
    """The default frame's support (written by the library's export threads on the native route) is each
    itemset's oracle count / n_rows as float64, bit for bit, in the reference frame's division."""
    import pyarrow as pa

    from et_miner.core.apriori import mine_csr

    rng = np.random.default_rng(77)
    indptr, ids = _csr(_rand_rows(rng, 7_001, 40, 4, dup=True))
    n = len(indptr) - 1
    want = miner.mine(indptr, ids, 0.01, None).as_dict()
    frame = mine_csr(indptr, ids, 0.01, None, engine=engine)
    table = frame if isinstance(frame, pa.Table) else frame.to_arrow()
    sets = [tuple(sorted(x)) for x in table.column("itemset").to_pylist()]
    got = dict(zip(sets, table.column("support").to_numpy()))
    assert got.keys() == want.keys()
    wk = list(want)
    exact = np.array([want[k] for k in wk], dtype=np.int64) / n
    assert np.array_equal(np.array([got[k] for k in wk]), exact)


def test_early_exported_levels_match_oracle():
    """Levels of at least 256K ints are copied to the host as soon as they are compacted and written into the
    result frame by host threads while later levels run: a planted block of 40 items (91,390 itemsets at K=4,
    457K ints) goes that way, K=1..3 through the final gather. The frame, the per-level arrays read before it, and
    the default frame's supports all equal the oracle."""
    import pyarrow as pa

    from et_miner.core.apriori import mine_csr
    from et_miner.core.result import _mining_min_count
    from et_miner.gpu import device as dev
    from et_miner.gpu.native import mine_csr_device

    rng = np.random.default_rng(4040)
    rows = []
    for _ in range(3_000):
        r = set(rng.choice(400, size=rng.integers(1, 6), replace=False).tolist())
        if rng.random() < 0.3:
            r |= set(range(1000, 1040))
        rows.append(sorted(r))
    indptr, ids = _csr(rows)
    n = len(rows)
    want = miner.mine(indptr, ids, 0.1, 4).as_dict()
    assert sum(1 for s in want if len(s) == 4) == 91_390
    # The frame (flat arrays viewed in place) and the per-level arrays.
    d_ptr, d_ids = dev.to_device(indptr), dev.to_device(ids)
    res = mine_csr_device(d_ptr, d_ids, n, _mining_min_count(0.1, n), 4, levels=True)
    values, offsets, counts = res.flat
    assert res.support is not None and res.support_rows == n
    got = {tuple(values[offsets[i]:offsets[i + 1]].tolist()): int(counts[i]) for i in range(len(counts))}
    assert got == want
    assert np.array_equal(res.support, counts / n)
    per_level = {}
    for lv_rows, lv_counts in res.levels:
        for r, c in zip(lv_rows, lv_counts):
            per_level[tuple(res.col_to_item[r].tolist())] = int(c)
    assert per_level == want
    # The default frame through the API, twice in a row (the second call reuses the pooled pinned blocks).
    for _ in range(2):
        frame = mine_csr(indptr, ids, 0.1, 4)
        table = frame if isinstance(frame, pa.Table) else frame.to_arrow()
        sets = [tuple(x) for x in table.column("itemset").to_pylist()]
        sup = table.column("support").to_numpy()
        assert dict(zip(sets, sup)) == {s: c / n for s, c in want.items()}


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("seed", range(8))
def test_random_datasets_match_oracle(seed, engine):
    rng = np.random.default_rng(seed)
    rows = _rand_rows(rng, 2_000 + 1_000 * seed, 20 + 10 * seed, 3 + seed % 4, zipf=0.8 + 0.1 * (seed % 3),
                      dup=seed % 2 == 1)
    indptr, ids = _csr(rows)
    for s, ml in [(0.01, None), (0.03, 3)]:
        want = miner.mine(indptr, ids, s, ml).as_dict()
        assert _gpu_mine(indptr, ids, s, ml, engine) == want
        assert _gpu_mine(indptr, ids, s, ml, engine, use_generator_pruning=True) == want


@pytest.mark.parametrize("engine", ENGINES)
def test_free_set_mode_matches_oracle(engine):
    from workloads import PRESETS, generate_csr

    data = generate_csr(PRESETS["smoke"])
    want = miner.mine(data.indptr, data.indices, 0.01, prune_equal_support=True).as_dict()
    assert _gpu_mine(data.indptr, data.indices, 0.01, engine=engine, prune_equal_support=True) == want
    rng = np.random.default_rng(5)
    rows = [sorted(set(r) | {0}) for r in _rand_rows(rng, 3000, 12, 4)]
    for r in rows[:1000]:
        r.extend(x for x in (7, 8) if x not in r)
        r.sort()
    indptr, ids = _csr(rows)
    want = miner.mine(indptr, ids, 0.02, prune_equal_support=True).as_dict()
    assert _gpu_mine(indptr, ids, 0.02, engine=engine, prune_equal_support=True) == want


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("name", ["smoke-C1-base", "deepk-C1-base", "skew-C1-base"])
def test_reference_lattices_on_gpu(name, engine):
    from workloads import PRESETS, generate_csr

    gold, params = load_lattice(name)
    data = generate_csr(PRESETS[params["dataset"]])
    got = _gpu_mine(data.indptr, data.indices, params["min_support"], params["max_length"], engine)
    assert got == gold
    got = _gpu_mine(data.indptr, data.indices, params["min_support"], params["max_length"], engine,
                    use_generator_pruning=True)
    assert got == gold


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("name,n_rows", [("plain_60k", None), ("deep_k_1m", 200_000), ("wide_vocab_2m", 100_000),
                                         ("deep_sparse_20m", 400_000)])
def test_workloads_match_oracle(name, n_rows, engine):
    from workloads import workload

    w = workload(name, n_rows)
    data = w.generate()
    want = miner.mine(data.indptr, data.indices, w.min_support, w.max_length).as_dict()
    got = _gpu_mine(data.indptr, data.indices, w.min_support, w.max_length, engine,
                    use_generator_pruning=w.use_generator_pruning)
    assert got == want


@pytest.mark.parametrize("k2_kernel", ["valu", "mfma"])
def test_k2_kernels_mine_identically(k2_kernel):
    """The fp4 MFMA Gram and the VALU tiles give the same lattice (wide vocabulary: hundreds of columns)."""
    from workloads import workload

    w = workload("wide_vocab_2m", 150_000)
    data = w.generate()
    want = miner.mine(data.indptr, data.indices, w.min_support, 2).as_dict()
    assert _gpu_mine(data.indptr, data.indices, w.min_support, 2, k2_kernel=k2_kernel) == want


def test_son_matches_in_core():
    from et_miner.core.result import as_dict
    from et_miner.streaming.son import son_csr
    from workloads import workload

    for name, n_rows, chunk in [("plain_60k", None, 7_000), ("deep_k_1m", 100_000, 30_000)]:
        w = workload(name, n_rows)
        data = w.generate()
        core = _gpu_mine(data.indptr, data.indices, w.min_support, w.max_length)
        son = as_dict(son_csr(data.indptr, data.indices, w.min_support, w.max_length, chunk, with_counts=True))
        assert son == core
        assert son == miner.mine(data.indptr, data.indices, w.min_support, w.max_length).as_dict()


@pytest.mark.parametrize("infer", [False, True])
def test_son_with_uneven_chunks(infer):
    """Chunks drawn from two different item popularities: many itemsets are frequent in some chunks only, so
    the union's bound leaves slack and pass 2 counts them over every chunk."""
    from et_miner.core.result import as_dict
    from et_miner.streaming.son import son_csr

    rng = np.random.default_rng(21)
    first = _rand_rows(rng, 6000, 40, 5)
    perm = rng.permutation(40)
    second = [sorted(int(perm[x]) for x in r) for r in _rand_rows(rng, 6000, 40, 5)]
    indptr, ids = _csr(first + second)
    want = miner.mine(indptr, ids, 0.02, None, use_generator_pruning=infer).as_dict()
    table = son_csr(indptr, ids, 0.02, None, 2500, use_generator_pruning=infer, with_counts=True)
    assert as_dict(table) == want
    # The frame is assembled on the device: its supports are the counts / n_rows of numpy's division, bit for bit.
    counts = table.column("count").to_numpy()
    assert np.array_equal(table.column("support").to_numpy(), counts / (len(indptr) - 1))


def test_bitvecs_route_is_read_only_and_exact():
    import torch

    from et_miner import apriori
    from et_miner.core.result import as_dict
    from workloads import PRESETS, generate_csr

    data = generate_csr(PRESETS["smoke"])
    gold, _ = load_lattice("smoke-C1-base")
    items = np.unique(data.indices)
    lut = {int(it): c for c, it in enumerate(items)}
    cols = np.array([lut[int(x)] for x in data.indices], np.int32)
    bv = miner.bitvectors(data.indptr, cols, len(items))
    t = torch.from_numpy(bv.view(np.int64)).cuda().view(torch.uint64)
    before = t.view(torch.int64).clone()
    res = apriori(bitvecs=(t, {c: int(it) for c, it in enumerate(items)}, data.n_rows), min_support=0.01)
    assert torch.equal(before, t.view(torch.int64)), "the caller's bitvectors must not be written"
    sup = as_dict(res)
    assert {k: round(v * data.n_rows) for k, v in sup.items()} == gold


def test_count_itemsets_flat_matches_oracle():
    from et_miner.gpu import device as dev
    from et_miner.gpu import kernels as K

    rng = np.random.default_rng(11)
    rows = _rand_rows(rng, 5_000, 30, 5)
    indptr, ids = _csr(rows)
    bv = miner.bitvectors(indptr, ids.astype(np.int32), 30)
    sets = [rng.choice(30, size=k, replace=False) for k in [1, 2, 3, 4, 6] for _ in range(20)]
    flat = np.concatenate(sets).astype(np.int32)
    offsets = np.zeros(len(sets) + 1, np.int64)
    np.cumsum([len(s) for s in sets], out=offsets[1:])
    got = K.count_itemsets_flat(dev.to_device(bv), flat, offsets)
    want = [int(np.bitwise_count(np.bitwise_and.reduce(bv[s], axis=0)).sum()) for s in sets]
    assert got.tolist() == want


@pytest.mark.parametrize("engine", ENGINES)
def test_edge_cases(engine):
    assert _gpu_mine(*_csr([[5]]), 0.5, None, engine) == {(5,): 1}
    assert _gpu_mine(*_csr([[]] * 10), 0.1, None, engine) == {}
    rows = [[1, 2, 3]] * 50
    assert _gpu_mine(*_csr(rows), 1.0, None, engine) == {(1,): 50, (2,): 50, (3,): 50, (1, 2): 50, (1, 3): 50,
                                                         (2, 3): 50, (1, 2, 3): 50}
    assert _gpu_mine(*_csr(rows), 1.0, 1, engine) == {(1,): 50, (2,): 50, (3,): 50}
    # Levels larger than the single-workgroup fast path (> 32,768 rows / candidates) take the multi-block scans.
    rng = np.random.default_rng(3)
    big = [sorted(set(rng.choice(400, size=12, replace=False).tolist())) for _ in range(20_000)]
    indptr, ids = _csr(big)
    assert _gpu_mine(indptr, ids, 0.0005, 3, engine) == miner.mine(indptr, ids, 0.0005, 3).as_dict()
