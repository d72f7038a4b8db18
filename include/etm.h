/*
 * etm.h -- C ABI of libetm.so, the MI355X (gfx950) counting path of exact
 * frequent-itemset mining.
 *
 * Conventions
 * -----------
 * - Every entry point returns an int status: ETM_OK (0) or an ETM_E* code.
 *   etm_last_error() copies the calling thread's last message. No C++
 *   exception crosses the ABI.
 * - Pointers named d_* are device (HBM) pointers; h_* are host pointers.
 *   Sizes are element counts unless the name ends in _bytes.
 * - `stream` is a hipStream_t passed as void* (NULL = the null stream).
 *   Entry points only enqueue work on it: they return before the GPU is
 *   done. The caller synchronizes before reading results on the host.
 * - Bitvectors: column c of a (n_cols, pitch) uint64 array holds bit r%64 of
 *   word r/64 set when transaction r holds c. pitch >= n_words =
 *   ceil(n_rows/64); words in [n_words, pitch) are never read.
 * - A "level" is a lexicographically sorted int32 array of itemsets of m
 *   column ids each (row-major, ascending inside a row), with an int32 count
 *   per row and optionally a uint8 free flag per row (1 = no (m-1)-subset
 *   has the same count).
 * - Prefix groups of a level: maximal runs of rows sharing their first m-1
 *   items. A group of s rows yields s(s-1)/2 candidates of m+1 items
 *   (prefix, suffix_i, suffix_j), i < j, numbered in lexicographic order:
 *   candidate (i, j) of group g has index cum_pairs[g] + i*(2s-i-1)/2 + (j-i-1).
 *   The candidates of a level are therefore lexsorted by index.
 *
 * Each declaration cites the reference interface it replaces (paths relative
 * to the reference's src/et_miner).
 */
#ifndef ETM_H
#define ETM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ETM_OK = 0,
    ETM_EINVAL = 1,    /* bad argument: maps to ValueError */
    ETM_ENOMEM = 2,    /* device allocation failed: MemoryError */
    ETM_EHIP = 3,      /* HIP runtime or launch error: RuntimeError */
    ETM_EOVERFLOW = 4, /* a size exceeds what the layout holds: ValueError */
    ETM_ENODEV = 5,    /* no usable gfx950 device: RuntimeError */
    ETM_ECANCELED = 6  /* a level callback stopped the call: the callback's own error */
};

/* Candidate status values written by etm_level_classify. */
enum { ETM_CAND_COUNT = 0, ETM_CAND_SKIP = 1, ETM_CAND_INFER = 2 };
/* etm_level_classify modes (bit set), as SUBSET_PRUNE / SUBSET_INFER in gpu/kernels/_src/_subset_index.cu:29-31. */
enum { ETM_SUBSET_PRUNE = 1, ETM_SUBSET_INFER = 2 };
/* Pair layouts of the K=2 count array: lexicographic (i-major), or the reference's b(b-1)/2 + a. */
enum { ETM_LAYOUT_LEX = 0, ETM_LAYOUT_TRI = 1 };
/* Kernel choice for the pair space of a single-group level (K=2): automatic (the fp4 MFMA Gram, except below 128
 * columns over at least 16,384 words: the VALU AND+popcount tiles), VALU tiles only, the MFMA Gram regardless
 * of size, or (etm_mine_* only) the row-wise kernel. */
enum { ETM_KERNEL_AUTO = 0, ETM_KERNEL_VALU = 1, ETM_KERNEL_MFMA = 2, ETM_KERNEL_ROWS = 3 };

/* etm_mine_opts.deduction: the bounds are on by default in the inference mode, as the reference's
 * ET_MINER_DEDUCTION unset or "1" (gpu/row_split.py:627); OFF is its "0". */
enum { ETM_DEDUCTION_AUTO = 0, ETM_DEDUCTION_OFF = 1 };
/* etm_mine_opts.k3_schedule: every group and tile-pair launched (those with nothing counted return at once), only
 * the ones holding a counted candidate (listed on the device first; see etm_count_level_ws), or automatic (the
 * live units under Pascal inference or over columns of at least 65,536 words, where most units of deep levels
 * have nothing to count; all units otherwise). */
enum { ETM_SCHEDULE_AUTO = 0, ETM_SCHEDULE_ALL = 1, ETM_SCHEDULE_LIVE = 2 };
/* etm_mine_opts.row_order: the reference's row_order ("keep" / "lex", gpu/input_layer.py:469-529). */
enum { ETM_ROW_ORDER_AUTO = 0, ETM_ROW_ORDER_KEEP = 1, ETM_ROW_ORDER_LEX = 2 };

/* etm_mine_opts.emit_levels. */
enum { ETM_EMIT_NONE = 0, ETM_EMIT_HOST = 1, ETM_EMIT_DEVICE = 2 };

/* ABI version of this header (bumped on any signature or struct-meaning change). */
#define ETM_ABI_VERSION 10
int etm_abi_version(void);

/* Copies the calling thread's last error message (NUL-terminated, truncated to n). */
int etm_last_error(char* h_buf, size_t n);

/* Number of visible HIP devices (0 when none). */
int etm_device_count(int* h_out);

/* Host -> device copy of `bytes` bytes of pageable host memory, cut into 32 MiB chunks staged through the
 * calling thread's ring of three pinned buffers: each chunk is copied into its buffer by up to 8 host threads
 * while the previous chunk's DMA runs, and a buffer is refilled only after its DMA completed. Enqueued on
 * `stream`; returns once the last chunk is enqueued, after which h_src may be reused. Replaces the reference's
 * double-buffered pinned upload of the Arrow value buffers (gpu/input_layer.py:328-393). */
int etm_upload(const void* h_src, int64_t bytes, void* d_dst, void* stream);

/* Device -> host copy into pageable host memory through the same ring: up to three 32 MiB chunks' DMAs in
 * flight on `stream` (after the work already queued there), each copied out by up to 8 host threads as it lands.
 * Returns when the data is in h_dst. No counterpart in the reference, whose results leave the device through
 * CuPy's synchronous device-to-host copies (.get()); this is the streaming path's result copy. */
int etm_download(const void* d_src, int64_t bytes, void* h_dst, void* stream);

/* Fills d_out[0] = min, d_out[1] = max of n item ids (int32 or int64 per id_bytes); n >= 1.
 * Replaces the id range scan of gpu/input_layer.py:396-405 (the bincount's span). */
int etm_ids_minmax(const void* d_ids, int id_bytes, int64_t n, int64_t* d_out, void* stream);

/* Occurrence histogram: d_hist[v - lo] = number of entries equal to v, for lo <= v < lo + span.
 * Zero-fills d_hist first. Replaces cp.bincount in gpu/input_layer.py:396-405 (_occurrences). */
int etm_ids_hist(const void* d_ids, int id_bytes, int64_t n, int64_t lo, int64_t span, int32_t* d_hist,
                 void* stream);

/* Column bitvectors from a CSR of item ids: entry e of row r sets column d_lut[id - lo] (skipped when the
 * id is outside [lo, lo+span) or the lut holds -1). Bit position p holds row d_order[p] (a permutation of the
 * rows, from etm_row_order), or row p when d_order is NULL. Writes every word of the (n_cols, pitch) output,
 * padding included. No atomics on HBM: each workgroup owns an 8-word x 1,024-column tile in LDS.
 * Replaces csr_to_bitvec / ids_to_bitvec (gpu/kernels/_src/csr_to_bitvec.cu:8-73, with new_pos) and
 * build_bitvecs_gpu (gpu/csr_bitvec.py:63-155). */
int etm_bitvec_build(const int64_t* d_indptr, const void* d_ids, int id_bytes, int64_t n_rows, int64_t lo,
                     int64_t span, const int32_t* d_lut, int32_t n_cols, int64_t pitch, const int32_t* d_order,
                     uint64_t* d_out, void* stream);

/* Words of tile occupancy: a tile is ETM_OCC_TILE_WORDS words (4,096 rows). */
#define ETM_OCC_TILE_WORDS 64

/* Workspace bytes etm_row_order needs for n_rows rows. */
size_t etm_row_order_ws(int64_t n_rows);

/* Row order: d_order[p] = the row placed at bit position p. Each row's key has bit 63 - r set for every column
 * of rank r < n_ranked it holds (d_rank[column], -1 for unranked columns; ids through d_lut as in
 * etm_bitvec_build); rows are sorted by ascending key, stably. Counts do not depend on the order; it gathers the
 * rows that share ranked columns, which the tile skip of etm_count_level exploits.
 * Replaces row_order_keys and the argsorts of gpu/input_layer.py:408-431 (kernels/_src/csr_to_bitvec.cu:79-101). */
int etm_row_order(const int64_t* d_indptr, const void* d_ids, int id_bytes, int64_t n_rows, int64_t lo,
                  int64_t span, const int32_t* d_lut, const int32_t* d_rank, int32_t n_ranked, int32_t* d_order,
                  void* d_ws, size_t ws_bytes, void* stream);

/* Tile occupancy of (n_cols, pitch) bitvectors: bit t % 64 of d_occ[c * occ_pitch + t / 64] is set when column
 * c has a nonzero word in tile t (words [64t, 64t + 64)). occ_pitch >= ceil(tiles / 64) + 1; the words past
 * the last tile are written as zero. Replaces tile_occupancy (gpu/kernels/occupancy.py:37-58), with 64-word
 * tiles: one wave64 pass per tile. */
int etm_tile_occupancy(const uint64_t* d_bv, int64_t pitch, int64_t n_words, int32_t n_cols, uint64_t* d_occ,
                       int64_t occ_pitch, void* stream);

/* K=1: d_out[c] = popcount of column c over words [0, n_words).
 * Replaces column_popcounts (gpu/kernels/loader.py:218-257). */
int etm_column_popcounts(const uint64_t* d_bv, int32_t n_cols, int64_t pitch, int64_t n_words, int64_t* d_out,
                         void* stream);

/* Workspace bytes etm_level_groups needs for a level of n rows. */
size_t etm_level_groups_ws(int64_t n);

/* Prefix groups of a lexsorted level (n rows of m items, m >= 1; m == 1 gives one group of every row, the
 * K=2 pair space). Writes, for n_groups groups:
 *   d_group_start[n_groups + 1]  first row of each group (last entry n),
 *   d_cum_pairs[n_groups + 1]    exclusive prefix sum of s(s-1)/2,
 *   d_cum_tiles[n_groups + 1]    exclusive prefix sum of the tiled kernel's tile-pairs (groups with more
 *                                than small_max suffixes; 0 for the others),
 *   d_small_list[n_groups]       ids of the groups with >= 2 suffixes, by size class: first the groups of
 *                                2 suffixes, then 3-4, 5-8, 9-16 (the wave-per-group classes up to small_max),
 *                                then those with more than small_max (the tiled kernel's); ascending within
 *                                a class,
 *   d_totals[24]                 {n_groups, n_candidates, n_tile_pairs, groups in the first four classes,
 *                                 max_suffixes, groups with >= 2 rows, rows in those groups, rows,
 *                                 class sizes (2, 3-4, 5-8, 9-16, > small_max), 0, 0, 0, 8 internal slots}.
 * Capacities: n + 1 for the first three, n for d_small_list.
 * Replaces build_k3plus_groups_from_flat (gpu/kernels/k3plus.py:243-376, rust_ext/src/core/groups.rs:59-205)
 * and the kernel split of group_kernels (gpu/row_split_chunks.py:96-121). */
int etm_level_groups(const int32_t* d_rows, int64_t n, int32_t m, int32_t small_max, int64_t* d_group_start,
                     int64_t* d_cum_pairs, int64_t* d_cum_tiles, int32_t* d_small_list, int64_t* d_totals,
                     void* d_ws, size_t ws_bytes, void* stream);

/* Open-addressing hash of a level's rows: d_table[table_size] (a power of two >= 2n) of row ids, -1 empty.
 * Replaces the sorted subset index upload (gpu/kernels/subset_index.py:63-140). */
int etm_level_hash(const int32_t* d_rows, int64_t n, int32_t m, int32_t* d_table, int64_t table_size,
                   void* stream);

/* Subset test and Pascal inference of every candidate of a level's prefix groups, against the hash of the
 * same level (the candidates' (k-1)-subsets). Writes d_status[c] (ETM_CAND_*) and initialises d_counts[c]
 * (0 when counted, the inferred count when inferred, -1 when skipped). mode: ETM_SUBSET_PRUNE, optionally |
 * ETM_SUBSET_INFER (then d_prev_free is required). With d_group_list and h_class_counts[6] (etm_level_groups'
 * five class sizes, then the largest group's size, d_totals[4]) the work is group-centric (one wave per small group, lanes over candidate x subset;
 * one workgroup per big group) and d_gid (optional) receives every candidate's group; without them, one
 * thread per candidate finds its group by binary search. When d_tally is not NULL, adds the numbers of
 * counted, inferred and skipped candidates to d_tally[0..2].
 * Replaces _classify_candidate (gpu/kernels/_src/_subset_index.cu:139-178); the deduction bounds are
 * etm_level_deduce. */
int etm_level_classify(const int32_t* d_rows, int64_t n, int32_t m, const int32_t* d_prev_counts,
                       const uint8_t* d_prev_free, const int32_t* d_table, int64_t table_size,
                       const int64_t* d_group_start, const int64_t* d_cum_pairs, int64_t n_groups,
                       int64_t n_cands, const int32_t* d_group_list, const int64_t* h_class_counts, int32_t mode,
                       uint8_t* d_status, int32_t* d_counts, int32_t* d_gid, int64_t* d_tally, void* stream);

/* Deduction bounds of depth 3 for the candidates of k = m + 1 items (3 <= k <= 24) that etm_level_classify left
 * ETM_CAND_COUNT in its inference mode (their (k-1)-subsets all free). With s() the support:
 *   u = min(min_a s(X-a), min_{a<b<c} s(X-a)+s(X-b)+s(X-c)-s(X-ab)-s(X-ac)-s(X-bc)+s(X-abc)),
 *   l = max(0, max_{a<b} s(X-a)+s(X-b)-s(X-ab)).
 * u < min_count makes the candidate ETM_CAND_SKIP (count -1); l == u makes it ETM_CAND_INFER with count u;
 * otherwise it stays counted. Level k-1 is d_rows (n x m) with its counts and hash; level k-2 is d_rows2
 * (m-1 columns), level k-3 d_rows3 (m-2 columns; unused at k = 3, where s(empty set) = n_total), each with its
 * counts and its etm_level_hash table. Levels k-1..k-3 must hold every frequent itemset of their length.
 * d_gid as etm_level_classify writes it (NULL: binary search). d_tally (optional) moves the settled candidates
 * from d_tally[0] to d_tally[1] (inferred) and d_tally[2] (skipped).
 * Replaces _deduce and the SUBSET_BOUNDS mode (gpu/kernels/_src/_subset_index.cu:95-136,
 * gpu/kernels/subset_index.py:63-140 with deep=). */
int etm_level_deduce(const int32_t* d_rows, int32_t m, const int32_t* d_prev_counts, const int32_t* d_table,
                     int64_t table_size, const int32_t* d_rows2, const int32_t* d_counts2, const int32_t* d_table2,
                     int64_t table_size2, const int32_t* d_rows3, const int32_t* d_counts3, const int32_t* d_table3,
                     int64_t table_size3, int64_t n_total, int64_t min_count, const int64_t* d_group_start,
                     const int64_t* d_cum_pairs, int64_t n_groups, int64_t n_cands, const int32_t* d_gid,
                     uint8_t* d_status, int32_t* d_counts, int64_t* d_tally, void* stream);

/* AND + popcount of every candidate whose status is ETM_CAND_COUNT (every candidate when d_status is
 * NULL), ADDED into d_counts (which the caller initialises: etm_level_classify does). Groups of 2..small_max
 * suffixes run one wave per (group, word chunk) with lanes over words, one kernel instance per size class
 * (h_class_counts[4]: the class sizes etm_level_groups reported in d_totals[8..11]); larger groups run
 * LDS-tiled tile-pairs. layout is ETM_LAYOUT_LEX, or ETM_LAYOUT_TRI for a single-group level (K=2 in the reference's
 * b(b-1)/2 + a order). word_chunk <= 0 picks a chunk from the level size. With d_occ (etm_tile_occupancy of
 * the same bitvectors) the group kernels skip the tiles where a prefix column, or all but one suffix column,
 * is empty; chunks are then at most 4,096 words.
 * Replaces count_k3plus_dense, count_group_pairs and count_shared_tiled_dense
 * (gpu/kernels/_src/k3plus_dense.cu:2-106 with its occupancy skip :78-83, group_pairs.cu:31-217,
 * shared_tiled.cu:217-283) and count_pairs_k2_dense (pairs_k2_dense.cu:2-45). */
int etm_count_level(const uint64_t* d_bv, int64_t pitch, int64_t n_words, const int32_t* d_rows, int32_t m,
                    const int64_t* d_group_start, const int64_t* d_cum_pairs, const int64_t* d_cum_tiles,
                    int64_t n_groups, int64_t n_tile_pairs, const int32_t* d_small_list,
                    const int64_t* h_class_counts, const uint8_t* d_status, int32_t* d_counts, int32_t layout,
                    int32_t word_chunk, int32_t kernel, const uint64_t* d_occ, int64_t occ_pitch, void* d_ws,
                    size_t ws_bytes, void* stream);

/* Workspace bytes of etm_count_level's live-unit path for a level of n_small small-class groups (the four class
 * sizes summed) and n_tile_pairs tile-pairs. With that workspace, a status array, the LEX layout, no occupancy
 * bitmap and word_chunk <= 0, etm_count_level first lists on the device the groups and tile-pairs holding at
 * least one ETM_CAND_COUNT candidate (tile-pairs with at most 32 of them, and at most 4 per column in them, with
 * their pair lists), then counts only those, the chunk length sized on the device from the number of live units:
 * the small groups and sparse tile-pairs one wave per (unit, chunk), the other tile-pairs one workgroup each.
 * Without it, every group and tile-pair is launched and those with nothing to count return at once. */
size_t etm_count_level_ws(int64_t n_small, int64_t n_tile_pairs);

/* Workspace bytes etm_level_compact needs for n_cands candidates. */
size_t etm_level_compact_ws(int64_t n_cands);

/* Survivor compaction of a counted level, in candidate (lexicographic) order. A survivor has status !=
 * ETM_CAND_SKIP (or d_status NULL) and count >= min_count. d_gid (optional) gives each candidate's group,
 * as etm_level_classify writes it; without it the group is found by binary search. Call twice with the same workspace: first with
 * d_out_rows == NULL, which writes the survivor count to d_out_n[0]; then with the outputs, capacity >= that
 * count, which writes d_out_rows (n_out x (m+1)), d_out_counts and, when not NULL, d_out_cand (candidate
 * index of each survivor). Replaces threshold_filter (gpu/kernels/filter.py:85-160) plus
 * decode_k3plus_flat / decode_k2_pairs_flat (gpu/kernels/decode.py:9-73) and the level-end lexsort
 * (gpu/row_split.py:1064-1067), which the lexicographic candidate order makes unnecessary. */
int etm_level_compact(const int32_t* d_rows, int32_t m, const int64_t* d_group_start, const int64_t* d_cum_pairs,
                      int64_t n_groups, int64_t n_cands, const uint8_t* d_status, const int32_t* d_counts,
                      const int32_t* d_gid, int64_t min_count, int32_t* d_out_rows, int32_t* d_out_counts, int64_t* d_out_cand,
                      int64_t* d_out_n, void* d_ws, size_t ws_bytes, void* stream);

/* etm_level_compact in one call that allocates nothing: no workspace argument, no device or host allocation.
 * No counterpart in the reference: it is for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). The scan's block sums live in a 1 MiB static device array,
 * one per GPU, shared by every call on that GPU: calls that overlap on one GPU (other streams, other threads)
 * race, and must be ordered by the caller. Writes the survivor count to d_out_n[0] and the first
 * min(count, out_cap) survivors, in candidate order, to d_out_rows (out_cap x (m+1)), d_out_counts and, when
 * not NULL, d_out_cand; a count above out_cap means the outputs were too small (nothing past them is written).
 * out_cap == 0 only counts. n_cands above etm_level_compact_noalloc_max() (134,216,704) is rejected with
 * ETM_EINVAL. */
int64_t etm_level_compact_noalloc_max(void);
int etm_level_compact_noalloc(const int32_t* d_rows, int32_t m, const int64_t* d_group_start,
                              const int64_t* d_cum_pairs, int64_t n_groups, int64_t n_cands, const uint8_t* d_status,
                              const int32_t* d_counts, const int32_t* d_gid, int64_t min_count, int32_t* d_out_rows,
                              int32_t* d_out_counts, int64_t* d_out_cand, int64_t out_cap, int64_t* d_out_n,
                              void* stream);

/* Free flags of a level of k-itemsets against the hash of its complete previous level: d_free[x] = 1 when
 * no (k-1)-subset of x has x's count. Every subset must be present in the previous level.
 * Replaces _prune_non_free_mask (gpu/mining.py:81-121, rust_ext/src/core/groups.rs:250-308). */
int etm_level_free(const int32_t* d_rows, int64_t n, int32_t k, const int32_t* d_counts,
                   const int32_t* d_prev_rows, const int32_t* d_prev_counts, const int32_t* d_prev_table,
                   int64_t table_size, uint8_t* d_free, void* stream);

/* Keeps the rows whose d_flag is nonzero, in order (two-call protocol as etm_level_compact, workspace
 * etm_level_compact_ws(n)). Replaces the host masks of the free-set prune (gpu/row_split.py:1081-1100). */
int etm_select_rows(const int32_t* d_rows, int64_t n, int32_t m, const int32_t* d_counts, const uint8_t* d_flag,
                    int32_t* d_out_rows, int32_t* d_out_counts, int64_t* d_out_n, void* d_ws, size_t ws_bytes,
                    void* stream);

/* Counts of arbitrary itemsets: popcount(AND of the columns d_items[d_offsets[i] .. d_offsets[i+1])), int64,
 * stored into d_out[i] (accumulate == 0) or added to it (accumulate != 0, e.g. summing the chunks of SON's
 * pass 2 on the device). Replaces count_itemsets_batch
 * (gpu/kernels/_src/itemset_count.cu:4-41) and count_itemsets_flat (gpu/kernels/batch.py:32-117). */
int etm_count_itemsets(const uint64_t* d_bv, int64_t pitch, int64_t n_words, const int32_t* d_items,
                       const int64_t* d_offsets, int64_t n_sets, int64_t* d_out, int32_t accumulate, void* stream);

/* ---- Row-wise counting of K=2 and K=3 ---------------------------------------------------------------- */

/* The transactions as ascending K=1 positions: position of id v = d_colpos[d_lut[v - lo]] (-1 drops it).
 * d_row_ptr (n_rows + 1) and d_pos (capacity nnz) receive the rows; d_stats[0..3] = sum over rows of C(len, 2),
 * the longest row, 1 when some row's ids do not ascend strictly (the row-wise kernels then do not apply), and
 * the number of positions. d_ws holds etm_rows_ws(n_rows) bytes. Replaces DeviceRows and _canonicalize
 * (gpu/input_layer.py:101-242) on one GPU. */
size_t etm_rows_ws(int64_t n_rows);
int etm_rows_build(const int64_t* d_indptr, const void* d_ids, int id_bytes, int64_t n_rows, int64_t lo, int64_t span,
                   const int32_t* d_lut, const int32_t* d_colpos, int64_t* d_row_ptr, int32_t* d_pos,
                   int64_t* d_stats, void* d_ws, size_t ws_bytes, void* stream);

/* Row-wise K=2: every row adds 1 to d_counts[lex_pair_index(i, j, n1)] for each pair i < j of its positions
 * (d_counts zeroed by the caller). Rows must ascend strictly. Replaces count_pairs_k2_rows
 * (gpu/kernels/k2.py:79-135, gpu/kernels/_src/pairs_k2_rows.cu:53-69). */
int etm_count_k2_rows(const int64_t* d_row_ptr, const int32_t* d_pos, int64_t n_rows, int64_t n1, int32_t* d_counts,
                      void* stream);

/* The K=2 level (d_rows2: n2 lexsorted pairs of column ids, its prefix groups d_group_start / n_groups) as a
 * pair bitmap over the n1 <= 65536 K=1 positions: d_bits and d_rank are n1 x ceil(n1/64) (uint64 bits, int32
 * set bits before each word), d_gof (n1) the group of each prefix or -1. Replaces k3_layout, k3_pair_bits and
 * build_k3_graph (gpu/kernels/k3_rows.py:78-125). */
int etm_k3_rows_prepare(const int32_t* d_rows2, int64_t n2, const int32_t* d_colpos, int64_t n1,
                        const int64_t* d_group_start, int64_t n_groups, uint64_t* d_bits, int32_t* d_rank,
                        int32_t* d_gof, void* stream);

/* Row-wise K=3: every row adds 1 to each candidate (a, b, c) of its positions with (a, b) and (a, c) in the
 * K=2 level whose status is ETM_CAND_COUNT (every one when d_status is NULL); candidates are numbered by the
 * K=2 level's prefix groups (d_cum_pairs). Rows must ascend strictly. Replaces count_k3_rows and
 * decode_k3_slots (gpu/kernels/k3_rows.py:126-178, gpu/kernels/_src/k3_rows.cu:49-140). */
int etm_count_k3_rows(const int64_t* d_row_ptr, const int32_t* d_pos, int64_t n_rows, const uint64_t* d_bits,
                      const int32_t* d_rank, const int32_t* d_gof, int64_t n1, const int64_t* d_group_start,
                      const int64_t* d_cum_pairs, const uint8_t* d_status, int32_t* d_counts, void* stream);

/* ---- SON pass-1 union --------------------------------------------------------------------------------- */

/* Adds a chunk's local level (n rows of k column ids, lexsorted within rows, mapped to item ids by
 * d_col_to_item, and their local counts) to the union of all chunks' levels of length k: d_table (table_size,
 * a power of two >= 2 x cap, filled with -1 before the first chunk) indexes d_keys (cap x k item ids), whose
 * d_known sums the local counts and d_slack the emitting chunks' slack (each chunk adds `slack`). *d_n_unique
 * counts the distinct itemsets; more than cap sets *d_overflow. A chunk's rows must be distinct. Replaces the
 * host-side union of the local lattices (streaming/son.py:366-421). */
int etm_union_insert(const int32_t* d_rows, const int32_t* d_counts, int64_t n, int32_t k,
                     const int64_t* d_col_to_item, int64_t slack, int32_t* d_table, int64_t table_size,
                     int64_t* d_keys, int64_t* d_known, int64_t* d_slack, int64_t* d_n_unique, int64_t cap,
                     int32_t* d_overflow, void* stream);

/* Each itemset's support on the device: d_out[i] = d_counts[i] / n_rows as an IEEE double division, the result
 * frame's `support = counts / n_rows` (core/cpu_miner.py:963) for a frame assembled on the device (the SON
 * result, streaming/son.py's final table). */
int etm_support(const int64_t* d_counts, int64_t n, int64_t n_rows, double* d_out, void* stream);

/* ---- The native level loop --------------------------------------------------------------------------- */

/* Per-level statistics: candidates generated by the prefix join and, of them, counted / inferred / skipped by
 * the subset test; survivors; bytes; the counting launches' event-timed duration (when time_kernels); the
 * level's wall time on the host, from the end of the previous level (K=1: from the call's start) to the sync
 * that ends this one.
 *   bytes_alg:  the prefix-group decomposition's compulsory bytes over every group of >= 2 rows,
 *               ((prefix items) x groups + suffix rows) x n_words x 8 + 4 per candidate;
 *   bytes_live: the same over what the counting kernels must read for the candidates left counted by the subset
 *               test, Pascal inference and the deduction bounds: (prefix items x groups with a counted pair +
 *               suffix rows in a counted pair) x n_words x 8 + 4 per counted candidate (K=2: = bytes_alg). */
typedef struct etm_level_stat {
    int32_t k;
    int32_t kernel; /* ETM_KERNEL_* that counted the level (K >= 2; VALU also covers the group kernels) */
    int64_t n_candidates;
    int64_t n_counted;
    int64_t n_inferred;
    int64_t n_skipped;
    int64_t n_frequent;
    int64_t bytes_alg;
    int64_t bytes_live;
    float count_ms;
    float wall_ms;
} etm_level_stat;

/* What a level callback receives as each level completes (pointers valid until the callback returns). */
typedef struct etm_level_info {
    etm_level_stat stat;
    const int32_t* h_rows;   /* with emit_levels: the level's n_frequent x k column ids, else NULL */
    const int32_t* h_counts; /* with emit_levels: their counts, else NULL */
    const int64_t* h_col_to_item; /* etm_mine_csr: the item id of each column (n_cols), else NULL */
    int64_t n_cols;
    int64_t device_bytes;    /* device memory the call holds (its arena, bitvectors included for etm_mine_csr) */
    int64_t host_bytes;      /* host memory of the levels kept for the result so far */
    int32_t next_level;      /* 1 when another level follows */
    int32_t reserved;
    const int32_t* d_rows;   /* levels K >= 2: the level's n_frequent x k column ids in device memory, valid until
                                the callback returns (copies enqueued on the call's stream are ordered before any
                                reuse), else NULL */
    const int32_t* d_counts; /* their counts, the same way */
} etm_level_info;

/* A nonzero return stops the mining call, which then fails with ETM_ECANCELED. Called on the calling thread. */
typedef int (*etm_level_cb)(const etm_level_info* info, void* user);

/* Options of a mining call. min_count is the absolute threshold (>= 1 is enforced); max_length 0 = none. */
typedef struct etm_mine_opts {
    int64_t min_count;
    int32_t max_length;
    int32_t prune_non_free; /* free-set mode (prune_equal_support) */
    int32_t infer_counts;   /* Pascal inference (use_generator_pruning) */
    int32_t time_kernels;   /* record HIP events around each level's counting launches */
    int32_t k2_kernel;      /* ETM_KERNEL_* for the K=2 pair space (ROWS: row-wise) */
    int32_t k3_kernel;      /* ETM_KERNEL_AUTO, ETM_KERNEL_VALU (bitvector kernels) or ETM_KERNEL_ROWS for K=3 */
    int32_t deduction;      /* ETM_DEDUCTION_AUTO: the deduction bounds with infer_counts (k <= 24); ETM_DEDUCTION_OFF */
    int32_t row_order;      /* etm_mine_csr: ETM_ROW_ORDER_LEX (rows sorted by their 64 most frequent columns,
                               and the tile skip), ETM_ROW_ORDER_KEEP, or ETM_ROW_ORDER_AUTO (keeps the rows) */
    int32_t live_bytes;     /* compute etm_level_stat.bytes_live (one extra pass over each level's candidates) */
    int32_t emit_levels;    /* ETM_EMIT_HOST: hand each level to level_cb as it completes (host copies in h_rows,
                               h_counts) and keep none for the result (host memory then holds one level at a time),
                               as the reference's per-level flush; ETM_EMIT_DEVICE: the same with the level left in
                               device memory only (d_rows, d_counts; no host copy of levels K >= 2) */
    int32_t k3_schedule;    /* ETM_SCHEDULE_*: how levels K >= 3 launch their counting units */
    int32_t small_max;      /* largest group (suffixes) counted by the wave-per-group kernels: 4, 8 or 16; 0 = 8 */
    etm_level_cb level_cb;  /* called after every level (K=1 included) when not NULL */
    void* cb_user;
} etm_mine_opts;

/* Host-side result of a mining call: levels of (rows of column ids int32 (n, k), counts), the columns. */
typedef struct etm_levels etm_levels;

/* Mines every frequent itemset (free-set with prune_non_free) of device bitvectors, from K=1 (the columns'
 * counts h_col_counts) to exhaustion, every level resident in HBM; one host sync per level. The caller's
 * bitvectors are only read. Replaces _apriori_row_split_multi_gpu's level loop (gpu/row_split.py:866-1185)
 * on one GPU. Free the result with etm_levels_free. */
int etm_mine_bitvecs(const uint64_t* d_bv, int64_t pitch, int64_t n_words, int64_t n_rows, const int64_t* h_col_counts,
                     int32_t n_cols, const etm_mine_opts* opts, etm_levels** out, void* stream);

/* The same from a device CSR of item ids (nnz entries), through the device input layer: id range,
 * occurrence histogram (K=1 candidates: ids occurring >= min_count times), bitvectors, K=1 popcounts.
 * Ids spanning more than span_limit values raise ETM_EINVAL. Replaces build_device_input
 * (gpu/input_layer.py:469-544) followed by the level loop. */
int etm_mine_csr(const int64_t* d_indptr, const void* d_ids, int id_bytes, int64_t n_rows, int64_t nnz,
                 int64_t span_limit, const etm_mine_opts* opts, etm_levels** out, void* stream);

/* One job of etm_mine_batch: a device CSR of item ids (as etm_mine_csr) with its own threshold and length cap
 * (max_length 0 = none). */
typedef struct etm_batch_job {
    const int64_t* d_indptr;
    const void* d_ids;
    int64_t n_rows;
    int64_t nnz;
    int64_t min_count;
    int32_t id_bytes;
    int32_t max_length;
} etm_batch_job;

/* Mines n_jobs independent jobs in one call, with the same result per job as etm_mine_csr (outs[i], freed with
 * etm_levels_free). The input layers of all jobs share their host syncs; then every job whose K=1 level has at
 * most 256 columns advances in lockstep rounds, one level per round, where each round is one fixed sequence of
 * launches covering every job (a work-queue-fed persistent counting kernel, workgroup-per-job group build and
 * compaction), replayed from a captured HIP graph with no host read between its launches. A job whose level
 * outgrows the rounds' limits (32,768 candidates), or whose K=1 level is wider, is mined by the per-job loop
 * of etm_mine_csr on the same bitvectors. opts gives the kernel settings of that loop; its min_count and
 * max_length are ignored (each job has its own). The free-set and inference modes and the level callback are not
 * supported here (ETM_EINVAL): mine those jobs with etm_mine_csr.
 * No reference counterpart: the reference mines one job per call (core/apriori.py:303-341); this is the batched
 * form of the served path, for callers holding many small jobs at once. */
int etm_mine_batch(const etm_batch_job* jobs, int32_t n_jobs, int64_t span_limit, const etm_mine_opts* opts,
                   etm_levels** outs, void* stream);

/* Accessors of a result: number of levels; level shape; copy of a level (h_rows n*k int32, h_counts n int64,
 * either may be NULL); the columns (count, item id of each column -- empty for etm_mine_bitvecs -- and K=1
 * counts); per-level statistics; release. */
int etm_levels_count(const etm_levels* lv, int32_t* h_n_levels);
int etm_levels_shape(const etm_levels* lv, int32_t level, int32_t* h_k, int64_t* h_n);
int etm_levels_copy(const etm_levels* lv, int32_t level, int32_t* h_rows, int64_t* h_counts);
int etm_levels_columns(const etm_levels* lv, int32_t* h_n_cols, int64_t* h_col_to_item, int64_t* h_col_counts);
int etm_levels_stats(const etm_levels* lv, int32_t level, etm_level_stat* h_out);
int etm_levels_free(etm_levels* lv);

/* Every level at once, flattened for a list column: h_totals[0] = itemsets, h_totals[1] = items over all
 * itemsets (call with NULL outputs to size them); then h_values (item ids through the result's column->item
 * map, or column ids when the result has none), h_offsets (itemsets + 1) and h_counts. */
int etm_levels_flat(const etm_levels* lv, int64_t* h_totals, int64_t* h_values, int64_t* h_offsets,
                    int64_t* h_counts);

/* The same export for a result frame: item ids and offsets (both or neither), and, each optional, counts and
 * supports (count / n_rows as an IEEE double division, the reference frame's `support = counts / n_rows`,
 * core/cpu_miner.py:963), written by the exporting threads so that the caller divides nothing. */
int etm_levels_frame(const etm_levels* lv, int64_t* h_totals, int64_t* h_values, int64_t* h_offsets,
                     int64_t* h_counts, double* h_support, int64_t n_rows);

/* The frame a mining call built while it ran (etm_mine_csr / etm_mine_bitvecs with large levels: each level of at
 * least 256K ints was copied to the host as soon as it was compacted and written out by host threads while the
 * next levels were mined). h_totals[0] = itemsets, h_totals[1] = items, or both -1 when the result holds no such
 * frame (then export with etm_levels_frame). h_ptrs[0..3] = item ids (int64), offsets (int64, itemsets + 1),
 * counts (int64) and supports (double), the same values etm_levels_frame writes, owned by the result until
 * etm_levels_free. With release_rows, the result's per-level rows go back to the pinned pool at once (later
 * etm_levels_copy of levels K >= 2, etm_levels_flat and etm_levels_frame then fail). No counterpart in the
 * reference: its frame is built on the host after the level loop (gpu/row_split.py:1172-1178). */
int etm_levels_frame_view(etm_levels* lv, int64_t* h_totals, void** h_ptrs, int32_t release_rows);

#ifdef __cplusplus
}
#endif

#endif /* ETM_H */
