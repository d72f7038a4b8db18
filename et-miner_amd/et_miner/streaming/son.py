"""SON two-pass streaming on one GPU (the reference's ``streaming/son.py:131-693``).

Pass 1 mines every row chunk at ``local_support_factor * min_support`` on the GPU level loop; the union of
the chunks' local levels, a device hash map per itemset length (``etm_union_insert``), sums per itemset the
local counts (``known``) and the slack (local min_count - 1) of the chunks that emitted it. The bound ``known + total_slack - slack`` drops candidates that cannot reach
the global min_count and is the exact count where no slack is left (``son.py:402-421``). Pass 2 counts the
remaining candidates over every chunk with the batched itemset kernel, summing the chunks on the device.

Transfers: each chunk's CSR is staged through pinned host buffers and copied with ``hipMemcpyAsync`` on a side
stream (``etm_upload``); the copy of chunk i+1 is issued before chunk i is counted, so transfer overlaps
counting. Only the CSR travels; bitvectors are built on the device, and pass 1's local levels stay on the device
from the mining call to the union.
"""

from __future__ import annotations

import math

import numpy as np

from ..core.apriori import transactions_to_csr
from ..core.profiling import ProfilingSession
from ..core.result import _min_count, _mining_min_count, empty_result, itemset_table, to_frame

#: The default local support factor (reference ``streaming/son.py:82``).
LOCAL_SUPPORT_FACTOR = 0.95
#: Chunks of pass 1 mined at a time (one stream and host thread each; 3 measured fastest).
PASS1_STREAMS = 3


def _estimate_chunk_size_from_memory(memory_budget_gb: float, n_items_estimate: int = 1000) -> int:
    """Rows per chunk for a memory budget (reference ``son.py:85-110``)."""
    bytes_per_transaction = (n_items_estimate / 8) * 2
    chunk_size = int(memory_budget_gb * 1024**3 / max(bytes_per_transaction, 1))
    return max(100_000, min(chunk_size, 100_000_000))


class _Uploader:
    """CSR chunks copied to the device on a side stream by a helper thread.

    ``issue(i)`` hands chunk i to the helper, which copies the chunk's rows with ``etm_upload`` (32 MiB pieces
    staged through the helper's ring of pinned buffers by parallel host copies, each piece's DMA overlapping the
    staging of the next) and rebases its offsets on the device. ``take(i)`` makes the current stream wait for
    that copy and issues chunk i+1, so the staging and the copy of the next chunk run while the current one is
    counted.
    """

    def __init__(self, indptr: np.ndarray, ids: np.ndarray, bounds: list):
        from concurrent.futures import ThreadPoolExecutor

        from ..gpu import device as dev

        self.torch = dev.torch()
        self.dev = dev
        self.device = dev.device()
        self.indptr, self.ids, self.bounds = indptr, ids, bounds
        self.stream = self.torch.cuda.Stream(device=self.device)
        self.pool = ThreadPoolExecutor(1)
        self.pending: dict = {}

    def _stage(self, i: int):
        from .. import _lib

        t = self.torch
        a, b = self.bounds[i]
        lo, hi_e = int(self.indptr[a]), int(self.indptr[b])
        n, nnz = b - a, hi_e - lo
        with t.cuda.device(self.device), t.cuda.stream(self.stream):
            d_ptr = t.empty(n + 1, dtype=t.int64, device=self.device)
            d_ids = t.empty(max(1, nnz), dtype=t.int64, device=self.device)
            h = self.stream.cuda_stream
            _lib.call("etm_upload", self.indptr[a:].ctypes.data, (n + 1) * 8, d_ptr.data_ptr(), h)
            _lib.call("etm_upload", self.ids[lo:].ctypes.data, nnz * 8, d_ids.data_ptr(), h)
            d_ptr -= lo
            ev = t.cuda.Event()
            ev.record(self.stream)
        return d_ptr, d_ids[:nnz], ev, n

    def issue(self, i: int) -> None:
        if 0 <= i < len(self.bounds) and i not in self.pending:
            self.pending[i] = self.pool.submit(self._stage, i)

    def take(self, i: int):
        """Chunk i's device CSR (indptr, ids, rows), ordered after its copy on the current stream."""
        self.issue(i)
        d_ptr, d_ids, ev, n = self.pending.pop(i).result()
        cur = self.torch.cuda.current_stream()
        cur.wait_event(ev)
        d_ptr.record_stream(cur)
        d_ids.record_stream(cur)
        self.issue(i + 1)
        return d_ptr, d_ids, n

    def close(self) -> None:
        self.pool.shutdown(wait=True)


def _local_levels(d_ptr, d_ids, n_rows: int, local_support: float, max_length, infer: bool):
    """One chunk mined at the local threshold on the native level loop: (its column items on the device, levels
    [(k, n, device column-id rows (n * k int32), device counts (n int32))]). Levels K >= 2 never leave the
    device: each is copied out of the call's arena as it completes."""
    from ..gpu import device as dev
    from ..gpu.native import mine_csr_device

    levels = []

    def keep(rep):
        n, k = rep.stats.n_frequent, rep.stats.k
        if n == 0:
            return
        if k == 1:
            levels.append((1, n, dev.to_device(np.ascontiguousarray(rep.rows.ravel(), dtype=np.int32)),
                           dev.to_device(rep.counts.astype(np.int32))))
        else:
            levels.append((k, n, dev.copy_from_pointer(rep.d_rows, n * k), dev.copy_from_pointer(rep.d_counts, n)))

    res = mine_csr_device(d_ptr, d_ids, n_rows, _mining_min_count(local_support, n_rows), max_length,
                          infer_counts=infer, on_level=keep, emit_levels="device", levels=False)
    if not levels or len(res.col_to_item) == 0:
        return None
    return res.col_to_item, levels


def apriori_streaming(
    transactions,
    min_support: float = 0.5,
    max_length: int | None = None,
    item_col: str = "items",
    chunk_size: int = 40_000_000,
    memory_budget_gb: float | None = None,
    local_support_factor: float = LOCAL_SUPPORT_FACTOR,
    use_gpu: bool = True,
    gpu_resident: bool = False,
    batch_size: int | None = 10_000,
    profile: bool = False,
    show_progress: bool = False,
    sparse=None,
    n_jobs: int = 1,
    progress_callback=None,
    use_generator_pruning: bool = False,
):
    """Frequent itemsets by SON's two passes over row chunks, each chunk counted on the GPU.

    Same result as the in-core ``apriori(use_gpu=True)``; see the module docstring.
    """
    if gpu_resident:
        raise ValueError("gpu_resident was removed: use_gpu=True keeps the candidates on the GPU. Drop the argument.")
    if not 0 < local_support_factor <= 1:
        raise ValueError(f"local_support_factor must be in (0, 1], got {local_support_factor}: above 1 pass 1 mines "
                         "each chunk above the global threshold and can miss globally frequent itemsets")
    if memory_budget_gb is None and not chunk_size >= 1:
        raise ValueError(f"chunk_size must be >= 1, got {chunk_size}")
    if sparse is not None:
        raise ValueError("sparse= was removed: no route reads it. Drop the argument.")
    if not use_gpu:
        raise ValueError("this build streams on the GPU only: pass use_gpu=True")
    session = ProfilingSession() if profile else None
    indptr, ids, items = transactions_to_csr(transactions, item_col)
    res = son_csr(indptr, ids, min_support, max_length,
                  _estimate_chunk_size_from_memory(memory_budget_gb) if memory_budget_gb is not None else chunk_size,
                  local_support_factor, session, progress_callback, use_generator_pruning)
    if items is not None:
        import pyarrow as pa

        flat = res.column("itemset").combine_chunks()
        res = pa.table({"itemset": pa.LargeListArray.from_arrays(flat.offsets, items.take(flat.values)),
                        "support": res.column("support")})
    df = to_frame(res)
    return (df, session) if profile else df


def son_csr(indptr, ids, min_support, max_length, chunk_size, local_support_factor=LOCAL_SUPPORT_FACTOR,
            session=None, progress_callback=None, use_generator_pruning=False, with_counts=False):
    """SON over a host CSR of item ids; returns a pyarrow Table (``itemset``, ``support``[, ``count``])."""
    from ..core.apriori import mine_csr
    from ..gpu import device as dev
    from ..gpu import kernels as K

    dev.require_gpu()
    indptr = np.asarray(indptr, dtype=np.int64)
    ids = np.asarray(ids, dtype=np.int64)
    n_total = len(indptr) - 1
    if n_total == 0:
        return itemset_table([], np.zeros(0, np.int64), 1, with_counts)
    if n_total <= chunk_size:
        table = mine_csr(indptr, ids, min_support, max_length, use_generator_pruning=use_generator_pruning,
                         with_counts=True)
        return table if with_counts else _drop_counts(table)
    bounds = [(a, min(a + chunk_size, n_total)) for a in range(0, n_total, chunk_size)]
    local_s = min_support * local_support_factor
    slacks = [_mining_min_count(local_s, b - a) - 1 for a, b in bounds]
    min_count = _min_count(min_support, n_total)
    if sum(slacks) >= min_count:
        raise RuntimeError(f"SON's slack {sum(slacks)} reaches min_count {min_count}: an itemset no chunk emits "
                           "could be frequent")
    up = _Uploader(indptr, ids, bounds)

    # Pass 1: local levels per chunk (column-id rows, counts, the chunk's column -> item map).
    if session:
        session.start_phase("pass1_local_mining")
    # PASS1_STREAMS chunks are mined at a time, each on its own stream and host thread (a mining call waits on the
    # host once per level; the other chunks' kernels fill those gaps). Chunk i goes to stream i % PASS1_STREAMS,
    # after chunk i - PASS1_STREAMS. Measured on deep_sparse_20m in 2M-row chunks, the whole call: 1, 2, 3, 4, 6
    # streams 108, 84, 79, 79, 75 ms (six streams hold twice the chunk workspaces for 5%).
    from concurrent.futures import ThreadPoolExecutor

    torch = dev.torch()
    streams = [torch.cuda.Stream(device=dev.device()) for _ in range(PASS1_STREAMS)]

    def mine_chunk(ci, d_ptr, d_ids, n):
        with torch.cuda.device(dev.device()), torch.cuda.stream(streams[ci % PASS1_STREAMS]):
            local = _local_levels(d_ptr, d_ids, n, local_s, max_length, use_generator_pruning)
            if local is None:
                return None
            col_items, levels = local
            return dev.to_device(np.ascontiguousarray(col_items, dtype=np.int64)), levels

    pending: dict = {}
    chunk_items: dict = {}
    futs: list = []
    with ThreadPoolExecutor(PASS1_STREAMS) as pool:
        for ci in range(len(bounds)):
            if ci >= PASS1_STREAMS:
                futs[ci - PASS1_STREAMS].result()
            with torch.cuda.stream(streams[ci % PASS1_STREAMS]):
                d_ptr, d_ids, n = up.take(ci)
            futs.append(pool.submit(mine_chunk, ci, d_ptr, d_ids, n))
        for ci, fut in enumerate(futs):
            local = fut.result()
            if local is not None:
                chunk_items[ci], levels = local
                for k, n_k, d_rows, d_counts in levels:
                    pending.setdefault(k, []).append((ci, n_k, d_rows, d_counts))
            if progress_callback:
                progress_callback("pass1", ci, len(bounds), {"candidates": sum(p[1] for v in pending.values()
                                                                               for p in v)})
    up.close()
    # The union runs on the current stream: it follows the chunks' streams.
    for st in streams:
        torch.cuda.current_stream().wait_stream(st)
    if not pending:
        return itemset_table([], np.zeros(0, np.int64), n_total, with_counts)
    # The union of the chunks' levels, per length, in a device hash map: the local counts (known) and the slack
    # of the emitting chunks summed per itemset; the bound ``known + total_slack - slack`` drops what cannot
    # reach min_count and is exact where no slack is left (reference son.py:402-421). The bound is applied on the
    # device: only the exact survivors and the candidates pass 2 must count come back to the host.
    torch = dev.torch()
    total_slack = sum(slacks)
    unions = {k: _device_union(parts, k, slacks, chunk_items) for k, parts in sorted(pending.items())}
    sizes = dev.to_host(torch.stack([torch.stack([u[3][0], u[4][0].to(torch.int64)]) for u in unions.values()]))
    masks = {}
    for (k, (keys, known, sl, _, _)), (u, overflow) in zip(unions.items(), sizes.tolist()):
        cap = len(known)
        if overflow or u > cap:
            raise RuntimeError(f"SON union of length {k}: {u} itemsets exceed the capacity {cap}")
        known_u, left = known[:u], total_slack - sl[:u]
        keep = known_u + left >= min_count
        need = keep & (left > 0)
        masks[k] = (u, keep & ~need & (known_u >= min_count), need)
    # How many itemsets each length keeps exactly and sends to pass 2, read back at once; the exact ones are then
    # gathered on the device with no further sync (positions by a prefix sum, the rest into a spare row).
    n_sel = dev.to_host(torch.stack([torch.stack([m[1].sum(), m[2].sum()]) for m in masks.values()]))
    exact, cands = {}, {}
    for (k, (u, done, need)), (n_done, n_need) in zip(masks.items(), n_sel.tolist()):
        keys, known = unions[k][0][: u * k].view(u, k), unions[k][1][:u]
        if n_done:
            pos = torch.cumsum(done, 0) - 1
            tgt = torch.where(done, pos, torch.full_like(pos, n_done))
            rows = torch.empty((n_done + 1, k), dtype=keys.dtype, device=keys.device).index_copy_(0, tgt, keys)
            cnts = torch.empty(n_done + 1, dtype=known.dtype, device=known.device).index_copy_(0, tgt, known)
            exact[k] = (rows[:n_done], cnts[:n_done])
        if n_need:
            cands[k] = dev.to_host(keys[need])
    if session:
        session.end_phase(n_candidates=sum(int(m[0]) for m in n_sel.tolist()) + sum(len(c) for c in cands.values()))

    # Pass 2: every candidate with slack left, counted over every chunk; totals summed on the device.
    if session:
        session.start_phase("pass2_global_counting")
    counted = {}
    if cands:
        order = sorted(cands)
        all_items = np.unique(np.concatenate([cands[k].ravel() for k in order]))
        flat = np.searchsorted(all_items, np.concatenate([cands[k].ravel() for k in order])).astype(np.int32)
        sizes = np.concatenate([np.full(len(cands[k]), k, dtype=np.int64) for k in order])
        offsets = np.zeros(len(sizes) + 1, dtype=np.int64)
        np.cumsum(sizes, out=offsets[1:])
        d_flat, d_off = dev.to_device(flat), dev.to_device(offsets)
        totals = dev.zeros(len(sizes), "int64")
        lo, span = int(all_items[0]), int(all_items[-1] - all_items[0] + 1)
        lut = np.full(span, -1, dtype=np.int32)
        lut[all_items - lo] = np.arange(len(all_items), dtype=np.int32)
        d_lut = dev.to_device(lut)
        up2 = _Uploader(indptr, ids, bounds)
        for ci in range(len(bounds)):
            d_ptr, d_ids, n = up2.take(ci)
            bv = K.build_bitvecs(d_ptr, d_ids, n, lo, span, d_lut, len(all_items))
            K.count_itemsets_device(bv, K.words_for(n), d_flat, d_off, len(sizes), totals, accumulate=True)
            if progress_callback:
                progress_callback("pass2", ci, len(bounds), {"counted": ci + 1})
        up2.close()
        got = dev.to_host(totals)
        at = 0
        for k in order:
            counted[k] = got[at : at + len(cands[k])]
            at += len(cands[k])
    if session:
        session.end_phase(n_counted=sum(len(c) for c in cands.values()))
    # The frame, assembled on the device (pass 2's survivors appended to their lengths) and copied back once.
    levels = []
    for k in sorted(set(exact) | set(counted)):
        rows, cnts = exact.get(k, (None, None))
        if k in counted:
            ok = counted[k] >= min_count
            more_rows = dev.to_device(np.ascontiguousarray(cands[k][ok], dtype=np.int64))
            more_cnts = dev.to_device(np.ascontiguousarray(counted[k][ok], dtype=np.int64))
            rows = more_rows.view(-1, k) if rows is None else torch.cat([rows, more_rows.view(-1, k)])
            cnts = more_cnts if cnts is None else torch.cat([cnts, more_cnts])
        if len(cnts):
            levels.append((k, rows, cnts))
    return _frame_table(levels, n_total, with_counts)


def _frame_table(levels: list, n_rows: int, with_counts: bool):
    """The result Table (itemset, support[, count]) of device levels [(k, item-id rows (n, k) int64, counts (n)
    int64)]: item ids, offsets, counts and supports (count / n_rows, IEEE double division) laid out in one device
    buffer and brought back with one ``etm_download``; the host arrays are views of it."""
    import pyarrow as pa

    from .. import _lib
    from ..gpu import device as dev

    torch = dev.torch()
    if not levels:
        return itemset_table([], np.zeros(0, np.int64), n_rows, with_counts)
    n_sets = sum(len(c) for _, _, c in levels)
    n_vals = sum(len(c) * k for k, _, c in levels)
    offs, base = [torch.zeros(1, dtype=torch.int64, device=dev.device())], 0
    for k, _, c in levels:
        offs.append(base + k * torch.arange(1, len(c) + 1, dtype=torch.int64, device=dev.device()))
        base += k * len(c)
    cnts = torch.cat([c.to(torch.int64) for _, _, c in levels])
    support = torch.empty(n_sets, dtype=torch.float64, device=dev.device())
    _lib.call("etm_support", cnts.data_ptr(), n_sets, int(n_rows), support.data_ptr(), dev.stream_ptr())
    buf = torch.cat([torch.cat([r.reshape(-1) for _, r, _ in levels]), torch.cat(offs), cnts, support.view(torch.int64)])
    host = dev.download(buf)
    values, offsets = host[:n_vals], host[n_vals : n_vals + n_sets + 1]
    counts = host[n_vals + n_sets + 1 : n_vals + 2 * n_sets + 1]
    support = host[n_vals + 2 * n_sets + 1 :].view(np.float64)
    cols = {"itemset": pa.LargeListArray.from_arrays(pa.array(offsets), pa.array(values)),
            "support": pa.array(support)}
    if with_counts:
        cols["count"] = pa.array(counts)
    return pa.table(cols)


def _device_union(parts: list, k: int, slacks: list, chunk_items: dict) -> tuple:
    """The distinct itemsets of length k over the chunks' device levels ``parts`` = [(chunk, rows, column-id rows
    int32, counts int32)], inserted through ``etm_union_insert`` with nothing read back: device tensors of their
    item-id rows (flat int64, cap x k), summed local counts and summed slack (the first u entries hold the u
    itemsets), then u and the overflow flag (one element each)."""
    from .. import _lib
    from ..gpu import device as dev

    torch = dev.torch()
    cap = sum(p[1] for p in parts)
    table_size = 1 << max(1, int(2 * cap - 1).bit_length())
    table = torch.full((table_size,), -1, dtype=torch.int32, device=dev.device())
    keys = dev.empty(cap * k, "int64")
    known = dev.zeros(cap, "int64")
    slack = dev.zeros(cap, "int64")
    n_unique = dev.zeros(1, "int64")
    overflow = dev.zeros(1, "int32")
    for ci, n, d_rows, d_counts in parts:
        _lib.call("etm_union_insert", d_rows.data_ptr(), d_counts.data_ptr(), n, k,
                  chunk_items[ci].data_ptr(), int(slacks[ci]), table.data_ptr(), table_size, keys.data_ptr(),
                  known.data_ptr(), slack.data_ptr(), n_unique.data_ptr(), cap, overflow.data_ptr(), dev.stream_ptr())
    return keys, known, slack, n_unique, overflow


def _drop_counts(table):
    return table.drop(["count"]) if "count" in table.column_names else table
