"""The operator API of the GPU counting path: one wrapper per C-ABI entry point.

Each wrapper checks its guards before launching (raising ValueError as the
reference's wrappers do, ``gpu/kernels/loader.py:40-138``), allocates its
outputs on the current device, enqueues the HIP kernels on the current stream
and returns device tensors. Wrappers that hand results to the host
synchronize. Names follow the reference's ``gpu/kernels/__init__.py:8-80``
where a counterpart exists, so a spy-based equivalence test can pin a kernel
by wrapper name.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from .. import _lib
from .._lib import CAND_COUNT, CAND_INFER, CAND_SKIP, LAYOUT_LEX, LAYOUT_TRI, SUBSET_INFER, SUBSET_PRUNE, call
from . import device as dev

__all__ = [
    "CAND_COUNT", "CAND_SKIP", "CAND_INFER", "SUBSET_PRUNE", "SUBSET_INFER", "LAYOUT_LEX", "LAYOUT_TRI",
    "SMALL_MAX", "MAX_K", "words_for", "padded_pitch", "ids_minmax", "ids_hist", "build_bitvecs",
    "column_popcounts", "Groups", "level_groups", "level_hash", "level_classify", "count_level", "level_compact",
    "level_free", "select_rows", "count_pairs_k2_shared", "count_itemsets_flat", "threshold_filter", "UNTOUCHED",
    "DeviceRows", "upload_rows", "count_pairs_k2_rows", "K3Graph", "build_k3_graph", "count_k3_rows",
]

#: Groups with at most this many suffixes run the wave-per-group kernel; larger ones the tiled kernel. The native
#: loop uses 8, measured faster on MI355X; this driver keeps 16, so the 9-16 suffix kernel stays exercised.
SMALL_MAX = 16
#: Largest itemset length (the reference's cap, gpu/kernels/loader.py:25-37).
MAX_K = 62
#: Marker of an unwritten count (gpu/kernels/filter.py:60).
UNTOUCHED = -1


def words_for(n_rows: int) -> int:
    return (int(n_rows) + 63) // 64


def padded_pitch(n_rows: int) -> int:
    """Words per column of the builder's output: ceil(N/64) rounded up to 64 words (512 bytes)."""
    w = words_for(n_rows)
    return max(64, (w + 63) // 64 * 64)


def _s():
    return dev.stream_ptr()


def _id_bytes(t) -> int:
    torch = dev.torch()
    if t.dtype == torch.int64:
        return 8
    if t.dtype == torch.int32:
        return 4
    raise ValueError(f"item ids must be int32 or int64, got {t.dtype}")


def ids_minmax(ids) -> tuple:
    """(min, max) of a device id array (synchronizes)."""
    if ids.numel() == 0:
        raise ValueError("ids_minmax: no ids")
    out = dev.empty(2, "int64")
    call("etm_ids_minmax", ids.data_ptr(), _id_bytes(ids), ids.numel(), out.data_ptr(), _s())
    lo, hi = dev.to_host(out).tolist()
    return int(lo), int(hi)


def ids_hist(ids, lo: int, span: int):
    """Device int32 occurrence counts of ids lo .. lo+span-1."""
    out = dev.empty(span, "int32")
    call("etm_ids_hist", ids.data_ptr() if ids.numel() else None, _id_bytes(ids), ids.numel(), lo, span,
         out.data_ptr(), _s())
    return out


def build_bitvecs(indptr, ids, n_rows: int, lo: int, span: int, lut, n_cols: int, pitch: int | None = None,
                  order=None):
    """(n_cols, pitch) bitvectors (int64 tensor holding the uint64 words) of a device CSR of item ids; with
    ``order`` (device int32, a permutation from ``row_order``) bit position p holds row ``order[p]``."""
    if indptr.numel() != n_rows + 1:
        raise ValueError(f"indptr holds {indptr.numel()} entries, expected n_rows + 1 = {n_rows + 1}")
    pitch = padded_pitch(n_rows) if pitch is None else int(pitch)
    if pitch < words_for(n_rows):
        raise ValueError(f"pitch {pitch} < ceil(n_rows/64) = {words_for(n_rows)}")
    if order is not None and order.numel() != n_rows:
        raise ValueError(f"order holds {order.numel()} entries, expected n_rows = {n_rows}")
    out = dev.empty((n_cols, pitch), "int64")
    call("etm_bitvec_build", indptr.data_ptr(), ids.data_ptr() if ids.numel() else indptr.data_ptr(),
         _id_bytes(ids), n_rows, lo, span, lut.data_ptr(), n_cols, pitch, dev.ptr(order), out.data_ptr(), _s())
    return out


def row_order(indptr, ids, n_rows: int, lo: int, span: int, lut, rank, n_ranked: int):
    """Device int32 permutation: the row at each bit position, rows sorted by the key of their ranked columns
    (bit 63 - r for a column of rank r < ``n_ranked`` <= 64). Replaces ``gpu/input_layer.py:408-431``."""
    order = dev.empty(max(n_rows, 1), "int32")
    ws_n = _lib.ws_bytes("etm_row_order_ws", n_rows)
    ws = dev.empty((ws_n + 7) // 8, "int64")
    call("etm_row_order", indptr.data_ptr(), ids.data_ptr() if ids.numel() else indptr.data_ptr(), _id_bytes(ids),
         n_rows, lo, span, lut.data_ptr(), rank.data_ptr(), n_ranked, order.data_ptr(), ws.data_ptr(), ws_n, _s())
    return order[:n_rows]


def frequent_first_ranks(col_counts: np.ndarray, n_ranked: int = 64) -> np.ndarray:
    """Host int32 rank of each column for ``row_order``: the ``n_ranked`` columns of most occurrences get ranks
    0.. in descending count order (ties by column), the others -1 (``gpu/input_layer.py:516-517``)."""
    order = np.argsort(-np.asarray(col_counts, dtype=np.int64), kind="stable")[:n_ranked]
    rank = np.full(len(col_counts), -1, dtype=np.int32)
    rank[order] = np.arange(len(order), dtype=np.int32)
    return rank


@dataclass
class Occupancy:
    """Tile occupancy: bit t % 64 of ``bits[c, t // 64]`` is set when column c holds a row in tile t."""

    bits: object  # device int64 (n_cols, pitch) holding uint64 words
    pitch: int


def tile_occupancy(bv, n_words: int, pitch: int | None = None) -> Occupancy:
    """The 64-word tile occupancy of (n_cols, pitch) bitvectors (reference ``gpu/kernels/occupancy.py:37-58``)."""
    n_cols = int(bv.shape[0])
    pitch = int(bv.shape[1]) if pitch is None else int(pitch)
    tiles = -(-int(n_words) // _lib.OCC_TILE_WORDS)
    occ_pitch = -(-tiles // 64) + 1
    bits = dev.empty((n_cols, occ_pitch), "int64")
    call("etm_tile_occupancy", bv.data_ptr(), pitch, n_words, n_cols, bits.data_ptr(), occ_pitch, _s())
    return Occupancy(bits, occ_pitch)


def column_popcounts(bv, n_words: int | None = None, pitch: int | None = None):
    """Device int64 popcount of every column (K=1), over the first n_words words (reference loader.py:218-257)."""
    n_cols = int(bv.shape[0])
    pitch = int(bv.shape[1]) if pitch is None else int(pitch)
    n_words = pitch if n_words is None else int(n_words)
    out = dev.empty(n_cols, "int64")
    call("etm_column_popcounts", bv.data_ptr(), n_cols, pitch, n_words, out.data_ptr(), _s())
    return out


@dataclass
class DeviceRows:
    """Transactions as ascending K=1 positions on the device (the row-wise kernels' input)."""

    row_ptr: object  # int64 (n_rows + 1)
    pos: object  # int32, row_ptr[-1] entries used
    n_rows: int
    row_pairs: int  # sum over rows of C(len, 2)
    longest: int
    ascending: bool  # every row's ids ascend strictly (the row-wise kernels need it)


def upload_rows(indptr, ids, n_rows: int, lo: int, span: int, lut, colpos) -> DeviceRows:
    """The rows of a device CSR as K=1 positions: id v -> colpos[lut[v - lo]] (-1 drops it). Synchronizes
    for the row statistics (reference DeviceRows.upload, gpu/input_layer.py:160-197)."""
    if indptr.numel() != n_rows + 1:
        raise ValueError(f"indptr holds {indptr.numel()} entries, expected n_rows + 1 = {n_rows + 1}")
    row_ptr = dev.empty(n_rows + 1, "int64")
    pos = dev.empty(max(1, ids.numel()), "int32")
    stats = dev.empty(4, "int64")
    ws_n = _lib.ws_bytes("etm_rows_ws", n_rows)
    ws = dev.empty((ws_n + 7) // 8, "int64")
    call("etm_rows_build", indptr.data_ptr(), ids.data_ptr() if ids.numel() else indptr.data_ptr(), _id_bytes(ids),
         n_rows, lo, span, lut.data_ptr(), colpos.data_ptr(), row_ptr.data_ptr(), pos.data_ptr(), stats.data_ptr(),
         ws.data_ptr(), ws_n, _s())
    st = dev.to_host(stats).tolist()
    return DeviceRows(row_ptr, pos, n_rows, int(st[0]), int(st[1]), st[2] == 0)


def count_pairs_k2_rows(rows: DeviceRows, n1: int):
    """Row-wise K=2 over n1 K=1 positions: device int32 counts in lexicographic pair order (reference
    gpu/kernels/k2.py:79-135)."""
    if not rows.ascending:
        raise ValueError("count_pairs_k2_rows: rows must hold strictly ascending, distinct ids")
    out = dev.zeros(max(1, n1 * (n1 - 1) // 2), "int32")
    call("etm_count_k2_rows", rows.row_ptr.data_ptr(), rows.pos.data_ptr(), rows.n_rows, n1, out.data_ptr(), _s())
    return out


@dataclass
class K3Graph:
    """The K=2 level as a pair bitmap over K=1 positions (reference K3Graph, gpu/kernels/k3_rows.py:76-90)."""

    bits: object
    rank: object
    gof: object
    n1: int


def build_k3_graph(rows2, colpos, n1: int, groups) -> K3Graph:
    """Pair bitmap, word ranks and prefix groups of the lexsorted K=2 level ``rows2`` (n2 x 2 column ids)."""
    if n1 > 65536:
        raise ValueError(f"build_k3_graph: {n1} K=1 positions exceed 65536")
    words = (n1 + 63) // 64
    bits = dev.empty(n1 * words, "int64")
    rank = dev.empty(n1 * words, "int32")
    gof = dev.empty(n1, "int32")
    call("etm_k3_rows_prepare", rows2.data_ptr(), int(rows2.shape[0]), colpos.data_ptr(), n1,
         groups.group_start.data_ptr(), groups.n_groups, bits.data_ptr(), rank.data_ptr(), gof.data_ptr(), _s())
    return K3Graph(bits, rank, gof, n1)


def count_k3_rows(rows: DeviceRows, graph: K3Graph, groups, status=None, counts=None):
    """Row-wise K=3 into the K=2 level's candidate numbering: adds 1 per containing row to each candidate whose
    status is COUNT (every candidate when ``status`` is None); returns the int32 counts (reference
    gpu/kernels/k3_rows.py:126-178)."""
    if not rows.ascending:
        raise ValueError("count_k3_rows: rows must hold strictly ascending, distinct ids")
    if counts is None:
        counts = dev.zeros(max(1, groups.n_cands), "int32")
    call("etm_count_k3_rows", rows.row_ptr.data_ptr(), rows.pos.data_ptr(), rows.n_rows, graph.bits.data_ptr(),
         graph.rank.data_ptr(), graph.gof.data_ptr(), graph.n1, groups.group_start.data_ptr(),
         groups.cum_pairs.data_ptr(), status.data_ptr() if status is not None else None, counts.data_ptr(), _s())
    return counts


@dataclass
class Groups:
    """Prefix groups of a level (device arrays) and their totals (host ints)."""

    group_start: object
    cum_pairs: object
    cum_tiles: object
    small_list: object
    n_groups: int
    n_cands: int
    n_tile_pairs: int
    n_small: int
    max_suffixes: int
    n_groups2: int  # groups with >= 2 rows (the groups that yield candidates)
    rows2: int  # rows in those groups (their suffixes)
    class_counts: tuple  # groups per size class (2, 3-4, 5-8, 9-16, > 16 suffixes)


def level_groups(rows, m: int) -> Groups:
    """Prefix groups of a lexsorted level of ``m`` items (synchronizes to read the totals)."""
    n = int(rows.shape[0])
    if n == 0:
        raise ValueError("level_groups: empty level")
    gs, cp, ct = dev.empty(n + 1, "int64"), dev.empty(n + 1, "int64"), dev.empty(n + 1, "int64")
    small = dev.empty(n, "int32")
    totals = dev.empty(24, "int64")
    ws_n = _lib.ws_bytes("etm_level_groups_ws", n)
    ws = dev.empty((ws_n + 7) // 8, "int64")
    call("etm_level_groups", rows.data_ptr(), n, m, SMALL_MAX, gs.data_ptr(), cp.data_ptr(), ct.data_ptr(),
         small.data_ptr(), totals.data_ptr(), ws.data_ptr(), ws_n, _s())
    t = dev.to_host(totals).tolist()
    return Groups(gs, cp, ct, small, *(int(x) for x in t[:7]), tuple(int(x) for x in t[8:13]))


def table_size_for(n: int) -> int:
    size = 1
    while size < 2 * max(1, n):
        size <<= 1
    return size


def level_hash(rows, m: int):
    """(table, size): the open-addressing hash of a level's rows."""
    n = int(rows.shape[0])
    size = table_size_for(n)
    table = dev.empty(size, "int32")
    call("etm_level_hash", rows.data_ptr(), n, m, table.data_ptr(), size, _s())
    return table, size


def level_classify(rows, m: int, prev_counts, prev_free, table, table_size: int, groups: Groups, mode: int,
                   tally=None, group_centric: bool = True):
    """(status uint8, counts int32, group id int32 or None) of every candidate: subset test, optional Pascal
    inference. ``tally`` (device int64[3], optional) accumulates the counted, inferred and skipped candidates;
    ``group_centric`` picks the wave-per-group form (else a thread per candidate with a binary search)."""
    status = dev.empty(groups.n_cands, "uint8")
    counts = dev.empty(groups.n_cands, "int32")
    gid = dev.empty(groups.n_cands, "int32") if group_centric else None
    cls = (ctypes.c_int64 * 6)(*groups.class_counts, groups.max_suffixes) if group_centric else None
    call("etm_level_classify", rows.data_ptr(), int(rows.shape[0]), m, dev.ptr(prev_counts), dev.ptr(prev_free),
         table.data_ptr(), table_size, groups.group_start.data_ptr(), groups.cum_pairs.data_ptr(), groups.n_groups,
         groups.n_cands, groups.small_list.data_ptr() if group_centric else None, cls, mode, status.data_ptr(),
         counts.data_ptr(), dev.ptr(gid), dev.ptr(tally), _s())
    return status, counts, gid


def level_deduce(prev, older2, older3, n_total: int, min_count: int, groups: Groups, status, counts, gid=None,
                 tally=None) -> None:
    """Deduction bounds of the candidates ``level_classify`` left counted (its inference mode), in place.

    ``prev``, ``older2``, ``older3`` are levels k-1, k-2, k-3 as ``(rows, counts, table, table_size)`` (``older3``
    is None at k = 3, where the empty set's count is ``n_total``). Replaces the SUBSET_BOUNDS mode of the
    reference's subset index (``gpu/kernels/_src/_subset_index.cu:95-136``)."""
    rows, pc, table, tsize = prev
    rows2, counts2, table2, tsize2 = older2
    rows3, counts3, table3, tsize3 = older3 if older3 is not None else (None, None, None, 0)
    call("etm_level_deduce", rows.data_ptr(), int(rows.shape[1]), pc.data_ptr(), table.data_ptr(), tsize,
         rows2.data_ptr(), counts2.data_ptr(), table2.data_ptr(), tsize2, dev.ptr(rows3), dev.ptr(counts3),
         dev.ptr(table3), tsize3, int(n_total), int(min_count), groups.group_start.data_ptr(),
         groups.cum_pairs.data_ptr(), groups.n_groups, groups.n_cands, dev.ptr(gid), status.data_ptr(),
         counts.data_ptr(), dev.ptr(tally), _s())


def count_level(bv, n_words: int, rows, m: int, groups: Groups, status, counts, layout: int = LAYOUT_LEX,
                word_chunk: int = 0, pitch: int | None = None, kernel: str = "auto",
                occ: Occupancy | None = None, live: bool = True) -> None:
    """Adds the AND + popcount counts of the counted candidates into ``counts``.

    ``kernel`` ("auto", "valu", "mfma") picks the K=2 pair space's kernel (ignored for other levels). With
    ``occ`` (``tile_occupancy`` of ``bv``) the group kernels skip the tiles where their AND is zero. With a
    ``status`` array and ``live`` (and neither ``occ`` nor ``word_chunk``), only the groups and tile-pairs
    holding a counted candidate are launched, by the persistent kernels."""
    if kernel not in ("auto", "valu", "mfma"):
        raise ValueError(f"count_level: kernel must be 'auto', 'valu' or 'mfma', got {kernel!r}")
    pitch = int(bv.shape[1]) if pitch is None else int(pitch)
    cls = (ctypes.c_int64 * 4)(*groups.class_counts[:4])
    ws, ws_n = None, 0
    if live and status is not None and occ is None and word_chunk <= 0:
        ws_n = int(_lib.lib().etm_count_level_ws(sum(groups.class_counts[:4]), groups.n_tile_pairs))
        ws = dev.empty((ws_n + 7) // 8, "int64")
    call("etm_count_level", bv.data_ptr(), pitch, n_words, rows.data_ptr(), m, groups.group_start.data_ptr(),
         groups.cum_pairs.data_ptr(), groups.cum_tiles.data_ptr(), groups.n_groups, groups.n_tile_pairs,
         groups.small_list.data_ptr(), cls, dev.ptr(status), counts.data_ptr(), layout, word_chunk,
         _lib.K2_KERNELS[kernel], dev.ptr(occ.bits) if occ is not None else None, occ.pitch if occ else 0,
         dev.ptr(ws), ws_n, _s())


def level_compact(rows, m: int, groups: Groups, status, counts, min_count: int, with_cand: bool = False, gid=None):
    """(rows (n, m+1) int32, counts int32[, candidate ids int64]) of the survivors, in lexicographic order."""
    n_c = groups.n_cands
    ws_n = _lib.ws_bytes("etm_level_compact_ws", n_c)
    ws = dev.empty((ws_n + 7) // 8, "int64")
    n_out = dev.empty(1, "int64")
    args = (rows.data_ptr(), m, groups.group_start.data_ptr(), groups.cum_pairs.data_ptr(), groups.n_groups, n_c,
            dev.ptr(status), counts.data_ptr(), dev.ptr(gid), int(min_count))
    call("etm_level_compact", *args, None, None, None, n_out.data_ptr(), ws.data_ptr(), ws_n, _s())
    n = int(dev.to_host(n_out)[0])
    out_rows = dev.empty((n, m + 1), "int32")
    out_counts = dev.empty(n, "int32")
    out_cand = dev.empty(n, "int64") if with_cand else None
    if n:
        call("etm_level_compact", *args, out_rows.data_ptr(), out_counts.data_ptr(), dev.ptr(out_cand),
             n_out.data_ptr(), ws.data_ptr(), ws_n, _s())
    return (out_rows, out_counts, out_cand) if with_cand else (out_rows, out_counts)


def level_free(rows, k: int, counts, prev_rows, prev_counts, prev_table, table_size: int):
    """uint8 free flags of a level of k-itemsets against its complete previous level's hash."""
    n = int(rows.shape[0])
    out = dev.empty(n, "uint8")
    if n:
        call("etm_level_free", rows.data_ptr(), n, k, counts.data_ptr(), prev_rows.data_ptr(),
             prev_counts.data_ptr(), prev_table.data_ptr(), table_size, out.data_ptr(), _s())
    return out


def select_rows(rows, m: int, counts, flag):
    """The rows (and counts) whose flag is set, in order."""
    n = int(rows.shape[0])
    ws_n = _lib.ws_bytes("etm_level_compact_ws", n)
    ws = dev.empty((ws_n + 7) // 8, "int64")
    n_out = dev.empty(1, "int64")
    call("etm_select_rows", rows.data_ptr(), n, m, counts.data_ptr(), flag.data_ptr(), None, None, n_out.data_ptr(),
         ws.data_ptr(), ws_n, _s())
    k = int(dev.to_host(n_out)[0])
    out_rows, out_counts = dev.empty((k, m), "int32"), dev.empty(k, "int32")
    if k:
        call("etm_select_rows", rows.data_ptr(), n, m, counts.data_ptr(), flag.data_ptr(), out_rows.data_ptr(),
             out_counts.data_ptr(), n_out.data_ptr(), ws.data_ptr(), ws_n, _s())
    return out_rows, out_counts


def count_pairs_k2_shared(bv, freq_cols, n_words: int, layout: int = LAYOUT_TRI, kernel: str = "auto"):
    """Dense int32 counts of every pair of ``freq_cols`` (reference ``shared_tiled.py:145-150``).

    Pair (a < b) of the column list lives at b(b-1)/2 + a (LAYOUT_TRI, the reference's layout) or at its
    lexicographic index (LAYOUT_LEX)."""
    cols = np.asarray(freq_cols, dtype=np.int32).reshape(-1, 1)
    if len(cols) < 2:
        return dev.zeros(0, "int32")
    if np.any(np.diff(cols[:, 0]) <= 0):
        raise ValueError("freq_cols must be strictly ascending")
    rows = dev.to_device(cols)
    g = level_groups(rows, 1)
    counts = dev.zeros(g.n_cands, "int32")
    count_level(bv, n_words, rows, 1, g, None, counts, layout=layout, kernel=kernel)
    return counts


def count_itemsets_flat(bv, flat, offsets, n_words: int | None = None) -> np.ndarray:
    """Counts (host int64) of itemsets given as flat column ids and offsets (reference ``batch.py:32-117``)."""
    flat = np.asarray(flat, dtype=np.int32)
    offsets = np.asarray(offsets, dtype=np.int64)
    n = len(offsets) - 1
    if n <= 0:
        return np.zeros(0, dtype=np.int64)
    if offsets[0] != 0 or offsets[-1] != len(flat) or np.any(np.diff(offsets) < 1):
        raise ValueError("offsets must start at 0, end at len(flat) and hold non-empty itemsets")
    if len(flat) and (flat.min() < 0 or flat.max() >= bv.shape[0]):
        raise ValueError("itemset column ids out of range")
    pitch = int(bv.shape[1])
    n_words = pitch if n_words is None else int(n_words)
    out = dev.empty(n, "int64")
    d_flat, d_off = dev.to_device(flat), dev.to_device(offsets)
    call("etm_count_itemsets", bv.data_ptr(), pitch, n_words, d_flat.data_ptr(), d_off.data_ptr(), n,
         out.data_ptr(), 0, _s())
    return dev.to_host(out)


def count_itemsets_device(bv, n_words: int, d_flat, d_offsets, n_sets: int, out, accumulate: bool = True,
                          pitch: int | None = None) -> None:
    """Counts of device-resident itemsets (flat column ids, offsets) stored into or added to ``out``."""
    pitch = int(bv.shape[1]) if pitch is None else int(pitch)
    call("etm_count_itemsets", bv.data_ptr(), pitch, n_words, d_flat.data_ptr(), d_offsets.data_ptr(), n_sets,
         out.data_ptr(), 1 if accumulate else 0, _s())


def threshold_filter(counts, min_count: int) -> tuple:
    """(indices int64, counts int64) on the host of the entries >= min_count (reference ``filter.py:85-160``)."""
    torch = dev.torch()
    keep = torch.nonzero(counts >= int(min_count)).flatten()
    return dev.to_host(keep).astype(np.int64), dev.to_host(counts[keep]).astype(np.int64)
