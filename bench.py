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

served: the served_jobs stream (512 independent jobs, seed 4: transactions log-uniform in 10K-2M, vocabulary
500-20K, Zipf 0.9-1.3, a coin between 1 + Poisson(4|6|9) and lognormal(2.3, 0.6) baskets, min_support
log-uniform in 0.001-0.05, max_length 3-10), every dataset resident in HBM before timing. The plain form mines
the jobs one after another on one stream; the served form runs them through the request loop (8 workers, one
stream each, footprint admission, K=2 kernel autotuned per shape bucket) after one warm-up pass. Throughput is
jobs / makespan with every job queued at t=0; latency (completion - arrival) is measured by replaying Poisson
arrivals at twice the plain form's measured jobs/s. Every job's itemsets and counts are checked identical in
both forms.

streaming: deep_sparse_20m (20,000,000 transactions over 5,000 ids, s = 0.002, max_length 16, generator pruning)
mined from host arrays resident and by SON in 2,000,000-row chunks; wall seconds and identical results.
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
TRAFFIC_PROFILE = ROOT / "profiles" / "deep_k_1m_pmc_traffic.md"


def measured_traffic(workload_name: str):
    """HBM bytes per step of the K>=3 counting dispatches, from the committed rocprofv3 PMC summary (FETCH_SIZE +
    WRITE_SIZE, each scaled by a plain copy's reading); None when no summary covers the workload."""
    import re

    if workload_name != "deep_k_1m" or not TRAFFIC_PROFILE.exists():
        return None
    m = re.search(r"^traffic_bytes_per_step = (\d+)$", TRAFFIC_PROFILE.read_text(), re.M)
    return int(m.group(1)) if m else None
VALU_CW_PEAK = 9.8e12  # candidate-words/s: 256 CUs x 64 lanes x 2.4 GHz / 4 VALU ops per word
# candidate-words/s of the int8 MFMA Gram: 5.0 POP/s dense int8 / 128 ops per word (64 bit products, each a
# multiply and an add)
MFMA_I8_CW_PEAK = 5.0e15 / 128


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
    ap.add_argument("--row-order", default="auto", choices=["auto", "keep", "lex"],
                    help="bitvector row order of the native loop (counts never depend on it)")
    ap.add_argument("--served-jobs", type=int, default=512, help="jobs of the served stream (0: skip that leg)")
    ap.add_argument("--served-workers", type=int, default=8)
    ap.add_argument("--no-streaming", action="store_true",
                    help="skip the streaming leg (deep_sparse_20m resident vs SON in 2,000,000-row chunks)")
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

    served_jobs_data = None
    if args.served_jobs > 0:  # generated before the GPU is initialised: the generator forks worker processes
        from et_miner.serve.server import jobs_from_specs
        from et_miner.synthetic import served_jobs

        t_gen = time.perf_counter()
        served_jobs_data = jobs_from_specs(served_jobs(512)[:args.served_jobs], workers=min(16, os.cpu_count() or 1))
        print(f"[bench] generated {len(served_jobs_data)} served jobs, "
              f"{sum(j.n_rows for j in served_jobs_data):,} rows in {time.perf_counter() - t_gen:.1f}s",
              file=sys.stderr, flush=True)

    stream_data = None
    if not args.no_streaming:
        t_gen = time.perf_counter()
        stream_w = workload("deep_sparse_20m")
        stream_data = stream_w.generate()
        print(f"[bench] generated {stream_w.name}: {stream_data.n_rows:,} rows in {time.perf_counter() - t_gen:.1f}s",
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
                         time_kernels=timer.enabled, row_order=args.row_order)
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
    # the K=2 kernel family the level ran on ("valu" tiles, int8 "mfma" Gram or the CSR "rows" kernel); the
    # python engine drives the VALU tiles
    k2_kernel = next((s.kernel for s in per_step if s.k == 2 and s.kernel), "valu")
    k2_bound, k2_peak = {"valu": ("valu", VALU_CW_PEAK), "mfma": ("mfma", MFMA_I8_CW_PEAK)}.get(k2_kernel,
                                                                                               ("hbm", None))
    achieved = (k3_bytes / (k3_ms / 1000) / 1e9) if k3_ms > 0 else 0.0
    print(f"[bench] {args.steps} steps, {ms_per_step:.2f} ms/step, {n_eval:,} candidates/step, "
          f"K>=3 counting {k3_ms:.3f} ms/step over {k3_launches} launches", file=sys.stderr, flush=True)

    # The same step when the caller hands over host arrays: the CSR crosses PCIe (pinned) before mining.
    host_ptr = torch.from_numpy(data.indptr).pin_memory()
    host_ids = torch.from_numpy(data.indices.astype(np.int64)).pin_memory()
    up_ms = []
    for _ in range(3):
        torch.cuda.synchronize()
        t_up = time.perf_counter()
        up = (host_ptr.to("cuda", non_blocking=True), host_ids.to("cuda", non_blocking=True))
        torch.cuda.synchronize()
        up_ms.append((time.perf_counter() - t_up) * 1000)
        del up
    upload_ms = sorted(up_ms)[1]
    csr_bytes = host_ptr.numel() * 8 + host_ids.numel() * 8
    pcie = {"upload_ms": upload_ms, "csr_bytes": csr_bytes, "upload_gb_per_s": csr_bytes / upload_ms / 1e6,
            "value": n_eval / ((ms_per_step + upload_ms) / 1000), "unit": "candidates/s",
            "note": "host CSR uploaded from pinned memory each step, then the same resident step"}

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

    served = _served_leg(served_jobs_data, args.served_workers) if served_jobs_data else None
    streaming = _streaming_leg(stream_w, stream_data) if stream_data is not None else None

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
            "traffic": measured_traffic(w.name) if args.rows is None else None,
            "traffic_unit": "HBM bytes per step (rocprofv3 FETCH_SIZE + WRITE_SIZE over the same dispatches)",
            "bytes_alg_per_step": k3_bytes,
            "kernel_ms_per_step": k3_ms,
            "launches_per_step": k3_launches,
        },
        "k2_count": {"kernel": k2_kernel, "bound": k2_bound, "kernel_ms_per_step": k2_ms,
                     "candidate_words_per_step": k2_cw,
                     "achieved_cw_per_s": (k2_cw / (k2_ms / 1000)) if k2_ms > 0 else 0.0,
                     "peak_cw_per_s": k2_peak},
        "cpu_baseline": cpu,
        "pcie_inclusive": pcie,
        "served": served,
        "streaming": streaming,
    }
    print(json.dumps(line), flush=True)
    return 0


def _streaming_leg(w, data):
    """deep_sparse_20m (use_generator_pruning) mined resident and by SON in 2,000,000-row chunks, from host
    arrays (the CSR crosses PCIe in both); wall seconds of one warm run each, results compared by digest."""
    import numpy as np

    from et_miner.core.apriori import mine_csr
    from et_miner.serve.server import result_digest
    from et_miner.streaming.son import son_csr

    ids = data.indices.astype(np.int64)

    def digest(tb):
        flat = tb.column("itemset").combine_chunks()
        return result_digest(np.asarray(flat.values), np.asarray(flat.offsets), np.asarray(tb.column("count")))

    out = {"workload": f"{w.name}: {data.n_rows:,} rows, {len(ids):,} entries, s={w.min_support}, "
                       f"max_length={w.max_length}, generator pruning", "chunk_rows": w.chunk_size}
    for name, run in (("resident", lambda: mine_csr(data.indptr, ids, w.min_support, w.max_length,
                                                    use_generator_pruning=True, with_counts=True)),
                      ("son", lambda: son_csr(data.indptr, ids, w.min_support, w.max_length, w.chunk_size,
                                              use_generator_pruning=True, with_counts=True))):
        run()  # warm: per-thread arenas sized, code objects loaded
        t = time.perf_counter()
        tb = run()
        out[name + "_s"] = time.perf_counter() - t
        out[name + "_digest"] = digest(tb)
        out["itemsets"] = tb.num_rows
    out["identical_results"] = out.pop("resident_digest") == out.pop("son_digest")
    print(f"[bench] streaming: resident {out['resident_s']:.2f}s, son {out['son_s']:.2f}s, "
          f"identical={out['identical_results']}", file=sys.stderr, flush=True)
    return out


def _served_leg(jobs, workers):
    """Plain vs served jobs/s on the served stream, latency under Poisson arrivals, identical results."""
    import dataclasses

    from et_miner.serve.server import Server, run_plain, serve

    zero = [dataclasses.replace(j, arrival_s=0.0) for j in jobs]
    run_plain(zero, time_scale=0.0)  # warm-up: uploads cached, per-thread arenas sized
    plain, ps = run_plain(zero, time_scale=0.0)
    print(f"[bench] plain: {ps['jobs_per_s']:.1f} jobs/s", file=sys.stderr, flush=True)
    srv = Server(workers=workers, autotune=True)
    try:
        serve(zero, server=srv, time_scale=0.0)  # warm-up: the workers' arenas sized, K=2 buckets tuned
        mixed, ss = serve(zero, server=srv, time_scale=0.0)
        print(f"[bench] served: {ss['jobs_per_s']:.1f} jobs/s", file=sys.stderr, flush=True)
        rate = 2.0 * ps["jobs_per_s"]
        gaps = __import__("numpy").random.default_rng(4).exponential(1.0 / rate, size=len(jobs))
        arrivals = gaps.cumsum() - gaps[0]
        timed = [dataclasses.replace(j, arrival_s=float(a)) for j, a in zip(jobs, arrivals)]
        mixed_p, sp = serve(timed, server=srv, time_scale=1.0)
    finally:
        srv.close()
    plain_p, pp = run_plain(timed, time_scale=1.0)
    same = all(a.digest == b.digest == c.digest == d.digest
               for a, b, c, d in zip(plain, mixed, mixed_p, plain_p))
    keep = ("jobs_per_s", "makespan_s", "latency_p50_s", "latency_p99_s")
    return {
        "workload": f"served_jobs: {len(jobs)} jobs, {sum(j.n_rows for j in jobs):,} transactions, seed 4",
        "unit": "jobs/s",
        "plain": {k: ps[k] for k in keep},
        "served": {k: ss[k] for k in keep} | {"workers": workers, "peak_concurrent_jobs": ss["peak_concurrent_jobs"]},
        "served_over_plain": ss["jobs_per_s"] / ps["jobs_per_s"],
        "poisson_rate_per_s": rate,
        "poisson_served": {k: sp[k] for k in keep},
        "poisson_plain": {k: pp[k] for k in keep},
        "identical_results": same,
    }


if __name__ == "__main__":
    sys.exit(main())
