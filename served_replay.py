"""Replays the served-jobs stream (``workloads.served_jobs``) on one GPU, plain or served.

    python served_replay.py --mode plain --jobs 512
    python served_replay.py --mode served --workers 8 --jobs 512
    python served_replay.py --mode served --workers 6 --batch-workers 2 --max-batch 64 --jobs 512
    python served_replay.py --mode batch --max-batch 64

``--mode batch`` keeps the stream's narrow jobs only (at most 256 frequent columns) and mines them through one
batch worker, in batched calls of up to ``--max-batch`` jobs.

Every job's dataset is generated on the host (before the GPU is touched) and made resident in HBM. The
stream runs ``--passes`` times with every job queued at t=0 (the first pass warms the per-thread arenas); the
last pass's summary (jobs/s, makespan, latency percentiles) is printed as one JSON line.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "et-miner_amd"))
sys.path.insert(0, str(ROOT))


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--mode", choices=["plain", "served", "batch"], default="served")
    ap.add_argument("--jobs", type=int, default=512)
    ap.add_argument("--workers", type=int, default=8)
    ap.add_argument("--batch-workers", type=int, default=0, help="threads mining narrow jobs in batched calls")
    ap.add_argument("--max-batch", type=int, default=64, help="most jobs per batched call")
    ap.add_argument("--passes", type=int, default=2)
    args = ap.parse_args(argv)

    from workloads import served_jobs
    from et_miner.serve.server import Server, batchable, jobs_from_specs, occurrence_profile, run_plain, serve

    jobs = jobs_from_specs(served_jobs(512)[: args.jobs], workers=min(16, os.cpu_count() or 1))
    if args.mode == "batch":
        jobs = [j for j in jobs if batchable(j, occurrence_profile(j.ids))]
        args.workers, args.batch_workers = 1, 1
    import torch

    torch.cuda.set_device(0)
    summary = None
    if args.mode == "plain":
        for _ in range(args.passes):
            _, summary = run_plain(jobs, time_scale=0.0)
    else:
        srv = Server(workers=args.workers, batch_workers=args.batch_workers, max_batch=args.max_batch)
        try:
            for _ in range(args.passes):
                _, summary = serve(jobs, server=srv, time_scale=0.0)
        finally:
            srv.close()
    print(json.dumps({"mode": args.mode, "jobs": len(jobs), **summary}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
