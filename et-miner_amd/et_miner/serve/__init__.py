"""Batched and served mining on one GPU: continuous batching of independent jobs over HBM-resident datasets."""

from .server import Job, JobResult, Server, footprint_bytes, result_digest, run_plain, serve

__all__ = ["Job", "JobResult", "Server", "footprint_bytes", "result_digest", "run_plain", "serve"]
