"""Pins the CPU oracle before anything is checked against it (CPU only).

Each check is independent of the oracle's own formulas: the reference's committed lattices on the regenerated
preset datasets, the reference's threshold table, the README's known answer, the boundary case, brute-force
subset enumeration, the free-set definition, planted-motif lower bounds and SON == in-core.
"""

from __future__ import annotations

import json
from itertools import combinations

import numpy as np
import pytest
from conftest import GOLDEN, load_lattice

from oracle import miner


def _csr(rows):
    indptr = np.zeros(len(rows) + 1, dtype=np.int64)
    indptr[1:] = np.cumsum([len(r) for r in rows])
    ids = np.array([x for r in rows for x in r], dtype=np.int64)
    return indptr, ids


def test_min_count_matches_reference_table():
    cases = json.loads((GOLDEN / "min_count_cases.json").read_text())["cases"]
    assert len(cases) >= 12
    for c in cases:
        assert miner.min_count(c["min_support"], c["n_rows"]) == c["expected"], c


def test_readme_known_answer():
    indptr, ids = _csr([[1, 2, 3], [2, 3, 4], [1, 3, 5], [2, 3]])
    got = miner.mine(indptr, ids, 0.5).as_dict()
    assert got == {(1,): 2, (2,): 3, (3,): 4, (1, 3): 2, (2, 3): 3}


def test_boundary_count_is_kept():
    threshold = 700  # ceil(0.07 * 10,000) computed from the decimal, not from the code under test
    rows = [[0, 1, 2]] * threshold + [[1, 2]] * (10_000 - threshold)
    got = miner.mine(*_csr(rows), 0.07).as_dict()
    for s in [(0,), (0, 1), (0, 2), (0, 1, 2)]:
        assert got[s] == threshold


def _random_rows(rng, n, vocab, mean, dup=False):
    rows = []
    w = 1.0 / np.arange(1, vocab + 1) ** 1.1
    w /= w.sum()
    for _ in range(n):
        L = 1 + rng.poisson(mean)
        r = rng.choice(vocab, size=L, replace=True, p=w).tolist()
        rows.append(r if dup else sorted(set(r)))
    return rows


@pytest.mark.parametrize("seed", range(6))
def test_mine_equals_brute_force(seed):
    rng = np.random.default_rng(seed)
    rows = _random_rows(rng, 300 + 50 * seed, 12 + seed, 3 + seed % 3, dup=seed % 2 == 1)
    indptr, ids = _csr(rows)
    for s, ml in [(0.05, None), (0.1, 3), (0.02, None)]:
        want = miner.brute_force(indptr, ids, s, ml)
        assert miner.mine(indptr, ids, s, ml).as_dict() == want
        assert miner.mine(indptr, ids, s, ml, use_generator_pruning=True).as_dict() == want


def _free_sets(freq: dict, n_rows: int) -> dict:
    out = {}
    for s, c in freq.items():
        if len(s) == 1:
            ok = c < n_rows
        else:
            ok = all(freq[sub] != c for sub in combinations(s, len(s) - 1))
        if ok:
            out[s] = c
    return out


@pytest.mark.parametrize("seed", range(4))
def test_free_set_mode_equals_definition(seed):
    rng = np.random.default_rng(100 + seed)
    rows = _random_rows(rng, 400, 10, 4)
    # An item in every row is never free; planted pairs make equal-count subsets.
    rows = [sorted(set(r) | {0}) for r in rows]
    for r in rows[: len(rows) // 3]:
        r.extend(x for x in (7, 8) if x not in r)
        r.sort()
    indptr, ids = _csr(rows)
    freq = miner.brute_force(indptr, ids, 0.03)
    got = miner.mine(indptr, ids, 0.03, prune_equal_support=True).as_dict()
    assert got == _free_sets(freq, len(rows))


def test_count_itemsets_equals_brute_force():
    rng = np.random.default_rng(7)
    rows = _random_rows(rng, 500, 15, 4)
    indptr, ids = _csr(rows)
    sets = {2: np.array([[0, 1], [2, 5], [3, 9]], np.int32), 3: np.array([[0, 1, 2], [1, 4, 6]], np.int32)}
    got = miner.count_itemsets(indptr, ids.astype(np.int32), 15, sets)
    for k, arr in sets.items():
        for row, c in zip(arr.tolist(), got[k].tolist()):
            assert c == sum(1 for r in rows if set(row) <= set(r))


@pytest.mark.parametrize("name", ["smoke-C1-base", "deepk-C1-base", "skew-C1-base"])
def test_reference_lattices(name):
    from workloads import PRESETS, generate_csr

    gold, params = load_lattice(name)
    data = generate_csr(PRESETS[params["dataset"]])
    assert data.n_rows == params["n_rows"]
    assert miner.min_count(params["min_support"], data.n_rows) == params["min_count"]
    got = miner.mine(data.indptr, data.indices, params["min_support"], params["max_length"],
                     prune_equal_support=params["prune_equal_support"]).as_dict()
    assert got == gold


def test_planted_motifs_recovered():
    from workloads import PRESETS, generate_csr

    data = generate_csr(PRESETS["smoke"])
    got = miner.mine(data.indptr, data.indices, 0.01).as_dict()
    for motif, planted in data.planted:
        for r in range(1, len(motif) + 1):
            for sub in combinations(motif, r):
                assert got[sub] >= planted


def test_son_equals_in_core():
    from workloads import workload

    data = workload("plain_60k", n_rows=30_000).generate()
    core = miner.mine(data.indptr, data.indices, 0.01).as_dict()
    son = miner.mine_son(data.indptr, data.indices, 0.01, chunk_size=7_000).as_dict()
    assert son == core


def test_son_slack_guard_raises():
    # With local factor <= 1 the slack sum stays below s*N, so only min_support = 0 (min_count 0, local
    # min_count 1) reaches the guard -- as in the reference (streaming/son.py:611-616).
    indptr, ids = _csr([[0, 1]] * 100)
    with pytest.raises(RuntimeError):
        miner.mine_son(indptr, ids, 0.0, chunk_size=10)


def test_deduction_hand_case():
    """Rows {0,1} {0,2} {0,1,2} {1} {2} {}: s(0)=s(1)=s(2)=3, s(01)=s(02)=2, s(12)=1, s(012)=1, N=6.
    Lower bound over the pair dropping items 1 and 2: s(02)+s(01)-s(0) = 1; upper: min(s(12)=1, triple
    2+2+1-3-3-3+6 = 2) = 1. Bounds meet: inferred 1 at min_count 1, skipped at min_count 2."""
    rows = [(0, 1), (0, 2), (0, 1, 2), (1,), (2,), ()]
    from conftest import lattice_counts

    lat = lattice_counts(rows, 1, 3)
    assert miner.deduction_bounds((0, 1, 2), lat, 6) == (1, 1)
    assert miner.deduction_outcome((0, 1, 2), lat, 6, 1) == ("infer", 1)
    assert miner.deduction_outcome((0, 1, 2), lat, 6, 2) == ("skip", 0)
    # {0,1,3} with 3 in every row holding 0: s(03) = s(0), so {0,3} is not free and Pascal applies
    rows = [(0, 1, 3), (0, 3), (1,), (1, 3), (0, 1, 3)]
    lat = lattice_counts(rows, 1, 3)
    assert miner.deduction_outcome((0, 1, 3), lat, 5, 1) == ("pascal", 2)


@pytest.mark.parametrize("k,seed,mc", [(3, 5, 40), (3, 53, 40), (4, 10, 40), (4, 17, 40), (5, 53, 15), (6, 17, 10)])
def test_deduction_bounds_bracket_every_count(k, seed, mc):
    """On every candidate of level k (all (k-1)-subsets frequent): l <= count <= u, Pascal's value and an
    inferred value equal the count, a skipped candidate is infrequent -- and the bounds settle some."""
    from conftest import dense_rows, lattice_counts

    rows = dense_rows(seed, 300)
    n = len(rows)
    exact = lattice_counts(rows, 0, k)
    lat = {s: c for s, c in exact.items() if c >= mc}
    sets = [frozenset(r) for r in rows]
    seen = {"pascal": 0, "skip": 0, "infer": 0, "count": 0}
    for x in combinations(range(7), k):
        if not all(sub in lat for sub in combinations(x, k - 1)):
            continue
        true = exact.get(x, 0)
        lower, upper = miner.deduction_bounds(x, lat, n)
        assert lower <= true <= upper, (x, lower, true, upper)
        # Inclusion-exclusion, counted from the rows: the pair term is s(x) minus the rows holding x-ab but
        # neither a nor b; the triple term is s(x) plus the rows holding x-abc but none of a, b, c.
        def none_of(drop):
            rest = set(x) - set(drop)
            return sum(1 for r in sets if rest <= r and not set(drop) & r)

        want_u = min([exact.get(tuple(v for v in x if v != a), 0) for a in x]
                     + [true + none_of(t) for t in combinations(x, 3)])
        want_l = max([0] + [true - none_of(t) for t in combinations(x, 2)])
        assert (lower, upper) == (want_l, want_u), (x, lower, upper, want_l, want_u)
        kind, value = miner.deduction_outcome(x, lat, n, mc)
        seen[kind] += 1
        if kind in ("pascal", "infer"):
            assert value == true, (x, kind, value, true)
        if kind == "skip":
            assert true < mc, (x, true)
    assert seen["skip"] + seen["infer"] > 0, seen


def test_thread_count_does_not_change_the_lattice():
    """The oracle's n_jobs (threads over prefix groups and subset lookups) leaves every itemset and count as
    with one thread, in each mining mode."""
    from workloads import workload

    w = workload("deep_k_1m", 60_000)
    d = w.generate()
    for kw in ({}, {"use_generator_pruning": True}, {"prune_equal_support": True}):
        one = miner.mine(d.indptr, d.indices, w.min_support, None, **kw).as_dict()
        many = miner.mine(d.indptr, d.indices, w.min_support, None, n_jobs=6, **kw).as_dict()
        assert one == many and len(one) > 1000, kw
