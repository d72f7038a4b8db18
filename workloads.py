"""Seeded synthetic transaction data, as CSR arrays of item ids: the inputs of the benchmark and of the tests.

This module is measurement and test infrastructure, not part of the et_miner package: nothing in the package
imports it. Two families live here:

* ``PRESETS`` / ``generate_csr``: the reference's own presets
  (et_miner/synthetic.py:105-195, :335-449 in the reference). Their lattices
  were mined by the reference and committed as golden results, so the
  datasets have to come out bit-identical: the parameters and the order of
  the random draws follow the reference's generator exactly (Poisson lengths,
  inverse-CDF Zipf draws with replacement, planted motifs on rows chosen
  without replacement, then one global unique over ``row * vocab + item``).
* The named mining workloads (``plain_60k``, ``deep_k_1m``, ``wide_vocab_2m``,
  ``deep_sparse_20m``, ``served_jobs``): baskets drawn WITHOUT replacement
  from Zipf popularity, item ids permuted by a seeded permutation so that
  id order is not popularity order, planted itemsets unioned into their rows.

Without-replacement draws: a basket of length L takes the first L distinct
values of an i.i.d. stream drawn with probability proportional to the Zipf
weights. Each new distinct value is then distributed proportionally to the
weights of the items not yet drawn, which is exactly successive weighted
sampling without replacement.

Every generator returns ``GeneratedData(indptr int64, indices int32 or int64,
n_rows, n_cols, planted)``; rows hold unique, ascending item ids.
"""

from __future__ import annotations

import dataclasses
import math
from dataclasses import dataclass
from fractions import Fraction
from typing import NamedTuple

import numpy as np

__all__ = [
    "SynthSpec",
    "GeneratedData",
    "PRESETS",
    "generate_csr",
    "Workload",
    "WORKLOADS",
    "workload",
    "served_jobs",
    "JobSpec",
]


class GeneratedData(NamedTuple):
    indptr: np.ndarray  # int64, n_rows + 1
    indices: np.ndarray  # item ids, unique and ascending within each row
    n_rows: int
    n_cols: int  # vocabulary size (ids are 0 .. n_cols - 1)
    planted: list  # [(itemset tuple, rows it was planted in), ...]


# ── the reference's presets ───────────────────────────────────────────────


@dataclass(frozen=True)
class SynthSpec:
    """One preset of the reference generator (et_miner/synthetic.py:45-96 in the reference)."""

    name: str
    n_rows: int
    vocab_size: int
    zipf_a: float
    row_len_mean: float
    row_len_min: int = 1
    row_len_max: int = 64
    skew_frac: float = 0.0
    skew_mult: float = 1.0
    motif_count: int = 0
    motif_size: int = 0
    motif_penetration: float = 0.0
    motif_tiers: tuple = ()
    min_support: float = 0.01
    seed: int = 42
    block_rows: int = 0

    @property
    def min_count(self) -> int:
        return math.ceil(Fraction(str(self.min_support)) * self.n_rows)


def _zipf_weights(n: int, a: float) -> np.ndarray:
    w = 1.0 / np.power(np.arange(1, n + 1, dtype=np.float64), a)
    return w / w.sum()


def _preset_motifs(spec: SynthSpec) -> list:
    """Disjoint motifs from the vocabulary's mid-tail (reference synthetic.py:109-121)."""
    if spec.motif_count <= 0 or spec.motif_size <= 0:
        return []
    lo = spec.vocab_size // 10
    hi = max(lo + spec.motif_count * spec.motif_size, spec.vocab_size // 2)
    if hi > spec.vocab_size:
        raise ValueError(f"{spec.name}: vocabulary too small for the motifs")
    picked = np.random.default_rng(spec.seed + 1).choice(
        np.arange(lo, hi, dtype=np.int64), size=spec.motif_count * spec.motif_size, replace=False
    )
    m = spec.motif_size
    return [np.sort(picked[i * m : (i + 1) * m]) for i in range(spec.motif_count)]


def _preset_block(spec: SynthSpec, row_seed: int) -> GeneratedData:
    """One block of a preset (reference synthetic.py:139-195): draws in the reference's order."""
    rng = np.random.default_rng(row_seed)
    n, vocab = spec.n_rows, spec.vocab_size
    lens = rng.poisson(spec.row_len_mean, size=n)
    if spec.skew_frac > 0 and spec.skew_mult != 1.0:
        n_skew = int(n * spec.skew_frac)
        lens[:n_skew] = np.round(lens[:n_skew] * spec.skew_mult).astype(lens.dtype)
    lens = np.clip(lens, spec.row_len_min, spec.row_len_max)
    rows = np.repeat(np.arange(n, dtype=np.int64), lens)
    cdf = np.cumsum(_zipf_weights(vocab, spec.zipf_a))
    cdf[-1] = 1.0
    items = np.searchsorted(cdf, rng.random(int(lens.sum())), side="right").astype(np.int64)

    planted = []
    extra_r, extra_i = [rows], [items]
    motifs = _preset_motifs(spec)
    for m_idx, motif in enumerate(motifs):
        k = int(round(spec.motif_penetration * n))
        at = np.random.default_rng(row_seed + 100 + m_idx).choice(n, size=k, replace=False).astype(np.int64)
        extra_r.append(np.repeat(at, len(motif)))
        extra_i.append(np.tile(motif, k))
        planted.append((tuple(int(x) for x in motif), k))
    for t_idx, (plen, pen) in enumerate(spec.motif_tiers):
        for m_idx, motif in enumerate(motifs):
            pre = motif[:plen]
            k = int(round(pen * n))
            at = np.random.default_rng(row_seed + 1000 + 100 * t_idx + m_idx).choice(n, size=k, replace=False)
            at = at.astype(np.int64)
            extra_r.append(np.repeat(at, len(pre)))
            extra_i.append(np.tile(pre, k))
            planted.append((tuple(int(x) for x in pre), k))
    keys = np.unique(np.concatenate(extra_r) * vocab + np.concatenate(extra_i))
    out_rows = keys // vocab
    indptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(out_rows, minlength=n), out=indptr[1:])
    return GeneratedData(indptr, (keys % vocab).astype(np.int32), n, vocab, planted)


def generate_csr(spec: SynthSpec) -> GeneratedData:
    """A preset as CSR arrays; blocks of ``block_rows`` draw from ``seed + 7919 * b``."""
    if 0 < spec.block_rows < spec.n_rows:
        parts = []
        for b, start in enumerate(range(0, spec.n_rows, spec.block_rows)):
            sub = dataclasses.replace(spec, n_rows=min(spec.block_rows, spec.n_rows - start), block_rows=0)
            parts.append(_preset_block(sub, spec.seed + 7919 * b))
        indptr = np.zeros(spec.n_rows + 1, dtype=np.int64)
        np.cumsum(np.concatenate([np.diff(p.indptr) for p in parts]), out=indptr[1:])
        planted = [(m, sum(p.planted[i][1] for p in parts)) for i, (m, _) in enumerate(parts[0].planted)]
        return GeneratedData(indptr, np.concatenate([p.indices for p in parts]), spec.n_rows, spec.vocab_size, planted)
    return _preset_block(spec, spec.seed)


#: The reference's presets (reference synthetic.py:335-449); the parameters define the golden datasets.
PRESETS: dict = {
    "smoke": SynthSpec("smoke", 60_000, 2_000, 1.05, 9, row_len_max=40, motif_count=3, motif_size=5,
                       motif_penetration=0.02, min_support=0.01, seed=42),
    "stress_k2": SynthSpec("stress_k2", 2_000_000, 35_000, 0.75, 24, row_len_max=64, min_support=0.000015,
                           seed=1337),
    "deep_k": SynthSpec("deep_k", 1_000_000, 4_000, 1.1, 14, row_len_max=48, motif_count=4, motif_size=10,
                        motif_penetration=0.03, min_support=0.02, seed=777),
    "skewed_rows": SynthSpec("skewed_rows", 1_000_000, 4_000, 1.1, 12, row_len_max=64, skew_frac=0.2,
                             skew_mult=5.0, motif_count=2, motif_size=8, motif_penetration=0.03,
                             min_support=0.02, seed=778),
    "oom_regression": SynthSpec("oom_regression", 500_000, 30_000, 0.7, 20, row_len_max=64,
                                min_support=0.00003, seed=99),
    "deep_sparse_large": SynthSpec("deep_sparse_large", 20_000_000, 1_000, 1.3, 6, row_len_max=48,
                                   motif_count=4, motif_size=12, motif_penetration=0.018,
                                   motif_tiers=((5, 0.22), (8, 0.09)), min_support=0.015, seed=4242),
    "wide_vocab": SynthSpec("wide_vocab", 100_000, 3_000, 0.6, 25, row_len_max=80, min_support=0.004, seed=31),
    "deep_k20": SynthSpec("deep_k20", 1_000_000, 4_000, 1.1, 10, row_len_max=48, motif_count=2, motif_size=20,
                          motif_penetration=0.011, min_support=0.01, seed=2020),
}


# ── the mining workloads ──────────────────────────────────────────────────


def _distinct_draws(rng: np.random.Generator, lens: np.ndarray, cdf: np.ndarray, vocab: int):
    """(row ids, item ranks): per row the first ``lens[r]`` distinct values of an i.i.d. stream drawn by ``cdf``.

    Rows are drawn in rounds: each round draws ``need + margin`` more values for every row still short,
    appends them to the row's stream, and keeps each value's first occurrence in stream order. Rows that are
    complete leave the working set, so later rounds sort only the rows still short.
    """
    n = len(lens)
    lens = np.minimum(lens, vocab).astype(np.int64)
    got = np.zeros(n, dtype=np.int64)
    stream_len = np.zeros(n, dtype=np.int64)
    short = np.flatnonzero(lens > 0)
    done: list = []
    # Distinct (row, item) keys of the rows still short, sorted, with their stream positions.
    seen_keys = np.empty(0, dtype=np.int64)
    seen_pos = np.empty(0, dtype=np.int64)
    while len(short):
        need = lens[short] - got[short]
        extra = need + np.maximum(2, need // 2)
        r = np.repeat(short, extra)
        first = np.repeat(stream_len[short], extra)
        pos = first + (np.arange(len(r), dtype=np.int64) - np.repeat(np.cumsum(extra) - extra, extra))
        it = np.searchsorted(cdf, rng.random(len(r)), side="right").astype(np.int64)
        stream_len[short] += extra
        keys = np.concatenate([seen_keys, r * vocab + it])
        allpos = np.concatenate([seen_pos, pos])
        span = int(stream_len.max()) + 1
        if n * vocab * span * vocab < (1 << 62):  # one int64 key per (row, item, position): plain sorts
            c = np.sort(keys * span + allpos)
            keys, allpos = c // span, c % span
            firsts = np.r_[True, keys[1:] != keys[:-1]]
            keys, allpos = keys[firsts], allpos[firsts]
            # Keep, per row, the lens[r] earliest distinct values.
            rr = keys // vocab
            c = np.sort((rr * span + allpos) * vocab + keys % vocab)
            rr_o = c // (span * vocab)
            head = np.r_[True, rr_o[1:] != rr_o[:-1]]
            first_of_row = np.maximum.accumulate(np.where(head, np.arange(len(rr_o)), 0))
            keep = (np.arange(len(rr_o)) - first_of_row) < lens[rr_o]
            c = c[keep]
            kp = np.sort((rr_o[keep] * vocab + c % vocab) * span + (c // vocab) % span)  # back to key order
            keys, allpos = kp // span, kp % span
        else:
            order = np.lexsort((allpos, keys))
            keys, allpos = keys[order], allpos[order]
            firsts = np.r_[True, keys[1:] != keys[:-1]]
            keys, allpos = keys[firsts], allpos[firsts]
            rr = keys // vocab
            order = np.lexsort((allpos, rr))
            rr_o = rr[order]
            head = np.r_[True, rr_o[1:] != rr_o[:-1]]
            first_of_row = np.maximum.accumulate(np.where(head, np.arange(len(rr_o)), 0))
            keep = (np.arange(len(rr_o)) - first_of_row) < lens[rr_o]
            sel = np.sort(order[keep])
            keys, allpos = keys[sel], allpos[sel]
        rr = keys // vocab
        got[short] = np.bincount(rr, minlength=n)[short]
        fin = got[rr] >= lens[rr]
        done.append(keys[fin])
        seen_keys, seen_pos = keys[~fin], allpos[~fin]
        short = short[got[short] < lens[short]]
    keys = np.sort(np.concatenate(done)) if done else np.empty(0, dtype=np.int64)
    return keys // vocab, keys % vocab


def _rows_to_csr(rows: np.ndarray, items: np.ndarray, n: int, vocab: int) -> tuple:
    keys = np.unique(rows.astype(np.int64) * vocab + items.astype(np.int64))
    indptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(keys // vocab, minlength=n), out=indptr[1:])
    return indptr, keys % vocab


@dataclass(frozen=True)
class Workload:
    """A mining workload: its data law, threshold and options (see each constructor's docstring)."""

    name: str
    seed: int
    n_rows: int
    vocab: int
    zipf_s: float
    basket: str  # "poisson" (1 + Poisson(mean)) or "lognormal" (round(lognormal(mu, sigma)) clipped)
    basket_mean: float = 0.0
    ln_mu: float = 0.0
    ln_sigma: float = 0.0
    ln_clip: tuple = (1, 200)
    plant_count: int = 0
    plant_len: tuple = (0, 0)
    plant_frac: float = 0.0
    plant_drop: float = 0.0
    min_support: float = 0.01
    max_length: int | None = None
    use_generator_pruning: bool = False
    chunk_size: int | None = None

    @property
    def min_count(self) -> int:
        return max(1, math.ceil(Fraction(str(self.min_support)) * self.n_rows))

    def generate(self, block_rows: int = 2_000_000) -> GeneratedData:
        return _generate_workload(self, block_rows)


def _generate_workload(w: Workload, block_rows: int) -> GeneratedData:
    """Rows of ``w``: baskets without replacement, planted itemsets unioned in, ids permuted.

    Draw order (all from ``default_rng(seed)``): the id permutation; the planted itemsets (length, members,
    rows, dropped members); then the baskets block by block (lengths, then the item stream).
    """
    rng = np.random.default_rng(w.seed)
    perm = rng.permutation(w.vocab).astype(np.int64)  # popularity rank -> item id
    planted_sets, plant_r, plant_i = [], [], []
    n = w.n_rows
    for _ in range(w.plant_count):
        length = int(rng.integers(w.plant_len[0], w.plant_len[1] + 1))
        members = np.sort(rng.choice(w.vocab, size=length, replace=False).astype(np.int64))
        at = np.sort(rng.choice(n, size=int(round(w.plant_frac * n)), replace=False).astype(np.int64))
        n_drop = int(round(w.plant_drop * len(at)))
        dropped_rows = rng.choice(len(at), size=n_drop, replace=False) if n_drop else np.empty(0, dtype=np.int64)
        drop_member = rng.integers(0, length, size=n_drop) if n_drop else np.empty(0, dtype=np.int64)
        mask = np.ones((len(at), length), dtype=bool)
        mask[dropped_rows, drop_member] = False
        rr = np.repeat(at, length).reshape(len(at), length)
        ii = np.tile(members, (len(at), 1))
        plant_r.append(rr[mask])
        plant_i.append(ii[mask])
        planted_sets.append((tuple(int(x) for x in members), len(at) - n_drop))
    cdf = np.cumsum(_zipf_weights(w.vocab, w.zipf_s))
    cdf[-1] = 1.0
    rows_parts, items_parts = [], []
    for start in range(0, n, block_rows):
        m = min(block_rows, n - start)
        if w.basket == "poisson":
            lens = 1 + rng.poisson(w.basket_mean, size=m)
        else:
            lens = np.clip(np.round(rng.lognormal(w.ln_mu, w.ln_sigma, size=m)), *w.ln_clip).astype(np.int64)
        r, rank = _distinct_draws(rng, lens, cdf, w.vocab)
        rows_parts.append(r + start)
        items_parts.append(perm[rank])
    rows = np.concatenate(rows_parts + plant_r)
    items = np.concatenate(items_parts + plant_i)
    indptr, indices = _rows_to_csr(rows, items, n, w.vocab)
    return GeneratedData(indptr, indices, n, w.vocab, planted_sets)


#: The named workloads.
WORKLOADS: dict = {
    # 60,000 baskets of 1 + Poisson(9) distinct Zipf(1.1) items over 1,000 ids; s = 0.01, mined to exhaustion.
    "plain_60k": Workload("plain_60k", 0, 60_000, 1_000, 1.1, "poisson", basket_mean=9, min_support=0.01),
    # 1,000,000 baskets of 1 + Poisson(6) Zipf(0.9) items over 200 ids, 20 planted itemsets of 8-12 items, each in
    # 3% of rows with 10% of insertions missing one random member; s = 0.02, no max_length.
    "deep_k_1m": Workload("deep_k_1m", 1, 1_000_000, 200, 0.9, "poisson", basket_mean=6, plant_count=20,
                          plant_len=(8, 12), plant_frac=0.03, plant_drop=0.10, min_support=0.02),
    # 2,000,000 baskets of round(lognormal(2.3, 0.6)) clipped to 1-200 Zipf(1.05) items over 20,000 ids;
    # s = 0.0005, max_length = 3.
    "wide_vocab_2m": Workload("wide_vocab_2m", 2, 2_000_000, 20_000, 1.05, "lognormal", ln_mu=2.3, ln_sigma=0.6,
                              min_support=0.0005, max_length=3),
    # 20,000,000 baskets of 1 + Poisson(4) Zipf(1.2) items over 5,000 ids, 50 planted itemsets of 6-16 items in
    # 0.5% of rows each; s = 0.002, max_length = 16, generator pruning; streamed in 2,000,000-row chunks.
    "deep_sparse_20m": Workload("deep_sparse_20m", 3, 20_000_000, 5_000, 1.2, "poisson", basket_mean=4,
                                plant_count=50, plant_len=(6, 16), plant_frac=0.005, min_support=0.002,
                                max_length=16, use_generator_pruning=True, chunk_size=2_000_000),
}


def workload(name: str, n_rows: int | None = None) -> Workload:
    """A named workload, optionally with another row count (same law; used for reduced-size parity cases)."""
    w = WORKLOADS[name]
    return w if n_rows is None else dataclasses.replace(w, n_rows=n_rows)


# ── the served-jobs stream ───────────────────────────────────────────────


@dataclass(frozen=True)
class JobSpec:
    """One independent mining job of the served stream."""

    job_id: int
    arrival_s: float
    workload: Workload


def served_jobs(n_jobs: int = 512, seed: int = 4, rate_per_s: float = 50.0,
                rows_range: tuple = (10_000, 2_000_000)) -> list:
    """``n_jobs`` independent jobs arriving by a Poisson process of ``rate_per_s`` jobs/s.

    Per job, from ``default_rng(seed)``: transactions log-uniform in ``rows_range``; vocabulary uniform in
    500-20,000; Zipf s uniform in 0.9-1.3; the basket law a fair coin between 1 + Poisson(lambda), lambda
    uniform in {4, 6, 9}, and round(lognormal(2.3, 0.6)) clipped to 1-200; min_support log-uniform in
    0.001-0.05; max_length uniform in 3-10. Each job's data seed is ``seed * 1_000_003 + job_id``.
    """
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate_per_s, size=n_jobs)
    arrivals = np.cumsum(gaps) - gaps[0]
    jobs = []
    lo, hi = math.log(rows_range[0]), math.log(rows_range[1])
    for j in range(n_jobs):
        n_rows = int(round(math.exp(rng.uniform(lo, hi))))
        vocab = int(rng.integers(500, 20_001))
        s = float(rng.uniform(0.9, 1.3))
        poisson = bool(rng.integers(0, 2))
        lam = float(rng.choice([4.0, 6.0, 9.0]))
        ms = float(math.exp(rng.uniform(math.log(0.001), math.log(0.05))))
        ms = float(f"{ms:.6g}")
        max_len = int(rng.integers(3, 11))
        wl = Workload(
            f"job{j}", seed * 1_000_003 + j, n_rows, vocab, round(s, 4),
            "poisson" if poisson else "lognormal", basket_mean=lam, ln_mu=2.3, ln_sigma=0.6,
            min_support=ms, max_length=max_len,
        )
        jobs.append(JobSpec(j, float(arrivals[j]), wl))
    return jobs
