"""Writes tests/golden/workload_digests.json: the oracle's per-level digests of the mining workloads at their
full sizes and of a fixed 32-job sample of the served stream (CPU only; run once, in the build container).

    python tests/golden/make_workload_digests.py [--only NAME ...] [--procs P]

The GPU parity tests regenerate the same data (seeded numpy generators), mine it on the device and compare
every level's (n, digest) with these. The oracle (``oracle.miner.mine``) restates the reference's CPU route; it
is pinned against the reference's committed lattices and brute force by tests/test_oracle.py.

Entries:

* plain_60k, deep_k_1m, wide_vocab_2m, deep_sparse_20m: each workload at its full row count, its own
  min_support and max_length (deep_sparse_20m: 20,000,000 rows, max_length 16; the oracle applies Pascal
  inference, which leaves the complete lattice unchanged).
* served_jobs: jobs 0, 16, 32, ..., 496 of the seed-4 stream of 512 jobs (transactions log-uniform in
  10,000-2,000,000), each with its own threshold and max_length.

Existing entries are kept unless named with --only; the file records the numpy version the data was drawn
with.
"""

from __future__ import annotations

import argparse
import json
import sys
import time
from concurrent.futures import ProcessPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
for p in (ROOT, ROOT / "et-miner_amd", HERE.parent):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

OUT = HERE / "workload_digests.json"
WORKLOADS = ["plain_60k", "deep_k_1m", "wide_vocab_2m", "deep_sparse_20m"]
SERVED_SAMPLE = list(range(0, 512, 16))


def _mine_workload(w) -> dict:
    from golden.digests import lattice_digests

    from oracle.miner import mine

    t0 = time.perf_counter()
    data = w.generate()
    t_gen = time.perf_counter() - t0
    t0 = time.perf_counter()
    lat = mine(data.indptr, data.indices, w.min_support, w.max_length,
               use_generator_pruning=w.use_generator_pruning)
    t_mine = time.perf_counter() - t0
    levels = lattice_digests(lat.levels, lat.items)
    return {"n_rows": w.n_rows, "vocab": w.vocab, "min_support": w.min_support, "max_length": w.max_length,
            "nnz": int(len(data.indices)), "n_itemsets": int(sum(d["n"] for d in levels)), "levels": levels,
            "oracle_s": round(t_mine, 1), "generate_s": round(t_gen, 1)}


def _job(spec) -> tuple:
    out = _mine_workload(spec.workload)
    return spec.job_id, out


def main(argv=None) -> int:
    import numpy as np

    from workloads import served_jobs, workload

    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--only", nargs="*", default=None, help="workload names and/or 'served_jobs'")
    ap.add_argument("--procs", type=int, default=4)
    args = ap.parse_args(argv)
    doc = json.loads(OUT.read_text()) if OUT.exists() else {}
    doc["numpy"] = np.__version__
    doc.setdefault("workloads", {})
    names = args.only if args.only is not None else WORKLOADS + ["served_jobs"]

    def save():
        OUT.write_text(json.dumps(doc, indent=1, sort_keys=True) + "\n")

    todo = [n for n in names if n in WORKLOADS]
    specs = []
    if "served_jobs" in names:
        all_specs = served_jobs(512)
        specs = [all_specs[j] for j in SERVED_SAMPLE]
        doc["served_jobs"] = {"seed": 4, "n_jobs": 512, "sample": SERVED_SAMPLE, "jobs": {}}
    with ProcessPoolExecutor(max_workers=max(1, args.procs)) as pool:
        futs = {pool.submit(_mine_workload, workload(n)): n for n in todo}
        jfuts = [pool.submit(_job, s) for s in specs]
        for f, n in futs.items():
            doc["workloads"][n] = f.result()
            print(f"{n}: {doc['workloads'][n]['n_itemsets']:,} itemsets, oracle {doc['workloads'][n]['oracle_s']} s",
                  flush=True)
            save()
        for f in jfuts:
            jid, out = f.result()
            doc["served_jobs"]["jobs"][str(jid)] = out
            print(f"job {jid}: {out['n_rows']:,} rows, {out['n_itemsets']:,} itemsets, oracle {out['oracle_s']} s",
                  flush=True)
            save()
    save()
    return 0


if __name__ == "__main__":
    sys.exit(main())
