"""Benchmark of the GPU counting path: candidate counts/s on one MI355X.

A step is one complete mining pass of the hot path over the deep_k_1m workload (1,000,000 transactions over 200
item ids, baskets of 1 + Poisson(6) Zipf(0.9) items, 20 planted itemsets of 8-12 items in 3% of rows with 10%
of insertions missing one member, min_support 0.02, no max_length -> levels to K=12). The transactions' CSR is
resident in HBM before timing; a step runs the device input layer (id range, histogram, bitvector build, K=1
popcount) and the level loop (prefix groups, subset test, AND + popcount counting, compaction) to exhaustion,
and copies every frequent itemset and count back to the host.

value = itemset candidates evaluated per second (counted + inferred, all levels; K=1 columns included),
whole-job, one GPU. The roofline object covers the counting launches of levels K >= 3 (the dominant kernel
family); their algorithmic bytes are the prefix-group decomposition's compulsory traffic,
sum over groups of (prefix items + suffixes) x words x 8 B + 4 B per candidate, and their duration is measured
with HIP events recorded on the stream the kernels run on, over the timed steps.

cpu_baseline: the oracle (numpy/scipy restatement of the reference's CPU route, one thread) on the same
workload, in the same run, on this host; its output is checked equal to the GPU's.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "et-miner_amd"))
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
VALU_CW_PEAK = 9.8e12  # candidate-words/s: 256 CUs x 64 lanes x 2.4 GHz / 4 VALU ops per word


class EventTimer:
    """HIP events around each level's counting launches, on the current (kernel) stream."""

    def __init__(self, torch):
        self.torch = torch
        self.records = []
        self.open = None
        self.enabled = True

    def begin(self, k):
        if not self.enabled:
            return
        ev = self.torch.cuda.Event(enable_timing=True)
        ev.record()
        self.open = (k, ev)

    def end(self, k, stats):
        if not self.enabled or self.open is None:
            return
        ev = self.torch.cuda.Event(enable_timing=True)
        ev.record()
        self.records.append((k, self.open[1], ev, stats))
        self.open = None

    def summary(self, n_words):
        out = {"k2": [0.0, 0, 0], "k3p": [0.0, 0, 0]}  # ms, bytes, candidate-words
        for k, a, b, st in self.records:
            key = "k2" if k == 2 else "k3p"
            out[key][0] += a.elapsed_time(b)
            out[key][1] += st.bytes_alg
            out[key][2] += st.n_counted * n_words
        return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="deep_k_1m")
    ap.add_argument("--rows", type=int, default=None, help="override the workload's row count")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--engine", default="native", choices=["native", "python"],
                    help="native: the level loop inside libetm; python: the same kernels driven level by level")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(json.dumps({"error": f"--gpus {args.gpus}: this build runs on one GPU; multi-GPU counting "
                                   "(row-split with RCCL) is not built"}))
        return 2

    import numpy as np
    import torch

    from et_miner.core.apriori import mine_csr
    from et_miner.core.result import as_dict, _mining_min_count
    from et_miner.gpu import device as dev
    from et_miner.gpu import kernels as K
    from et_miner.synthetic import workload

    t_gen = time.perf_counter()
    w = workload(args.workload, args.rows)
    data = w.generate()
    t_gen = time.perf_counter() - t_gen
    print(f"[bench] generated {w.name}: {data.n_rows:,} rows, {len(data.indices):,} entries in {t_gen:.1f}s",
          file=sys.stderr, flush=True)

    torch.cuda.set_device(0)
    d_csr = (dev.to_device(data.indptr), dev.to_device(data.indices.astype(np.int64)))
    torch.cuda.synchronize()
    n_words = K.words_for(data.n_rows)
    timer = EventTimer(torch)

    def step(collect=False):
        stats = []
        table = mine_csr(None, None, w.min_support, w.max_length, use_generator_pruning=w.use_generator_pruning,
                         device_csr=d_csr, with_counts=collect, timer=timer, stats_out=stats, engine=args.engine,
                         time_kernels=timer.enabled)
        return table, stats

    timer.enabled = False
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    timer.enabled = True
    stats_all = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        _, stats = step()
        stats_all.append(stats)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed * 1000 / args.steps
    per_step = stats_all[-1]
    n_eval = sum(s.n_evaluated for s in per_step)
    n_gen = sum(s.n_candidates for s in per_step)
    n_itemsets = sum(s.n_frequent for s in per_step)
    value = n_eval * args.steps / elapsed
    if args.engine == "python":
        ksum = timer.summary(n_words)
    else:  # the native loop records the events itself, on the stream its kernels run on
        ksum = {"k2": [0.0, 0, 0], "k3p": [0.0, 0, 0]}
        for stats in stats_all:
            for st in stats:
                if st.k >= 2:
                    key = "k2" if st.k == 2 else "k3p"
                    ksum[key][0] += st.count_ms
                    ksum[key][1] += st.bytes_alg
                    ksum[key][2] += st.n_counted * n_words
    k3_ms, k3_bytes = ksum["k3p"][0] / args.steps, ksum["k3p"][1] / args.steps
    k2_ms, k2_cw = ksum["k2"][0] / args.steps, ksum["k2"][2] / args.steps
    k3_launches = sum(1 for s in per_step if s.k >= 3)
    achieved = (k3_bytes / (k3_ms / 1000) / 1e9) if k3_ms > 0 else 0.0
    print(f"[bench] {args.steps} steps, {ms_per_step:.2f} ms/step, {n_eval:,} candidates/step, "
          f"K>=3 counting {k3_ms:.3f} ms/step over {k3_launches} launches", file=sys.stderr, flush=True)

    # Parity of the timed configuration with the oracle, and the CPU baseline from the same oracle run.
    gpu_res = as_dict(step(collect=True)[0])
    cpu = None
    parity = None
    if not args.no_cpu_baseline:
        from oracle.miner import mine

        cpu_stats = []
        t_c = time.perf_counter()
        lat = mine(data.indptr, data.indices, w.min_support, w.max_length,
                   use_generator_pruning=w.use_generator_pruning,
                   level_callback=lambda k, c, f: cpu_stats.append((k, c, f)))
        t_c = time.perf_counter() - t_c
        parity = lat.as_dict() == gpu_res
        cpu_cands = sum(c for _, c, _ in cpu_stats)
        cpu = {"value": cpu_cands / t_c, "unit": "candidates/s", "cores": 1, "kind": "port",
               "sample": f"the full {w.name} workload ({data.n_rows:,} rows), oracle mine() single-threaded "
                         f"numpy/scipy, {t_c:.2f} s, {cpu_cands:,} candidates, host {os.cpu_count()} cpus"}
        print(f"[bench] cpu oracle {t_c:.2f}s, parity={parity}", file=sys.stderr, flush=True)

    line = {
        "metric": "itemset-candidate counts/s",
        "value": value,
        "unit": "candidates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {
            "workload": w.name,
            "engine": args.engine,
            "n_transactions": data.n_rows,
            "n_items": w.vocab,
            "min_support": w.min_support,
            "min_count": _mining_min_count(w.min_support, data.n_rows),
            "max_length": w.max_length,
            "candidates_evaluated_per_step": n_eval,
            "candidates_generated_per_step": n_gen,
            "frequent_itemsets": n_itemsets,
            "levels": max(s.k for s in per_step),
            "parity_vs_oracle": parity,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_count_small/k_count_tiled, levels K>=3",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "bytes_alg_per_step": k3_bytes,
            "kernel_ms_per_step": k3_ms,
            "launches_per_step": k3_launches,
        },
        "k2_tiled": {"bound": "valu", "kernel_ms_per_step": k2_ms, "candidate_words_per_step": k2_cw,
                     "achieved_cw_per_s": (k2_cw / (k2_ms / 1000)) if k2_ms > 0 else 0.0,
                     "peak_cw_per_s": VALU_CW_PEAK},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
