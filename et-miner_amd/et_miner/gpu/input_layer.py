"""Device input layer: a CSR of item ids in HBM -> column bitvectors and K=1 counts.

Counterpart of the reference's ``build_device_input`` (``gpu/input_layer.py:469-544``):

1. the id range (``etm_ids_minmax``) and the occurrence histogram (``etm_ids_hist``) give the K=1 candidate
   columns: ids occurring at least min_count times, in ascending id order;
2. a lookup table id -> column drives the coalesced bitvector builder (``etm_bitvec_build``);
3. the column popcount (``etm_column_popcounts``) gives each column's exact row count; the level loop keeps
   the columns reaching min_count.

An item repeated inside a row counts once per occurrence in step 1 and once per row in step 3, as in the
reference (``core/cpu_miner.py:275-308``).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import device as dev
from . import kernels as K

#: Largest id span (max - min + 1) counted with a direct histogram; wider ids are coded on the host first.
SPAN_LIMIT = 1 << 26


@dataclass
class DeviceInput:
    bitvecs: object  # device int64 (n_cols, pitch) holding uint64 words
    n_rows: int
    n_words: int
    pitch: int
    col_to_item: np.ndarray  # int64 item id of each column (ascending)
    col_counts: np.ndarray  # int64 rows holding each column


def host_csr_to_device(indptr: np.ndarray, ids: np.ndarray):
    """(indptr int64, ids) copied to the device through pinned buffers."""
    ids = np.asarray(ids)
    if ids.dtype not in (np.int32, np.int64):
        ids = ids.astype(np.int64)
    return dev.to_device(np.asarray(indptr, dtype=np.int64)), dev.to_device(ids)


def build_device_input(indptr_d, ids_d, n_rows: int, min_count: int) -> DeviceInput | None:
    """Bitvectors of the K=1 candidate columns of a device CSR, or None when no item occurs min_count times."""
    if ids_d.numel() == 0:
        return None
    lo, hi = K.ids_minmax(ids_d)
    span = hi - lo + 1
    if span > SPAN_LIMIT:
        raise ValueError(f"item ids span {span:,} values (> {SPAN_LIMIT:,}): code them to dense ids first")
    hist = dev.to_host(K.ids_hist(ids_d, lo, span))
    cand = np.flatnonzero(hist >= min_count)
    if len(cand) == 0:
        return None
    lut = np.full(span, -1, dtype=np.int32)
    lut[cand] = np.arange(len(cand), dtype=np.int32)
    lut_d = dev.to_device(lut)
    pitch = K.padded_pitch(n_rows)
    bv = K.build_bitvecs(indptr_d, ids_d, n_rows, lo, span, lut_d, len(cand), pitch)
    n_words = K.words_for(n_rows)
    counts = dev.to_host(K.column_popcounts(bv, n_words)).astype(np.int64)
    return DeviceInput(bv, n_rows, n_words, pitch, (cand + lo).astype(np.int64), counts)
