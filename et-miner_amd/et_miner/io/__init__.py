"""Output formats of the mining results (per-level parquet and resume)."""
