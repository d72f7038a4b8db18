// Launchers shared between libetm's translation units (not part of the C ABI). Each enqueues kernels on `s`
// and returns the launch status. A non-null d_n means "the count lives on the device at d_n, at most n_cap".
#pragma once

#include <functional>

#include "batch.h"
#include "common.h"

namespace etm {
namespace detail {

// Largest level (rows) / candidate space the single-workgroup fast paths take, and the most items (rows or
// candidates x their items) they handle: one workgroup writes or compares every item, so wide levels go to the
// multi-block paths sooner (deep_sparse_20m's K=13, 24,938 candidates of 13 items: 310 us compacted by one
// workgroup).
constexpr int64_t kSingleMax = 32768;
constexpr int64_t kSingleItems = 65536;
inline bool single_fits(int64_t n, int64_t items) { return n <= kSingleMax && n * items <= kSingleItems; }

// The automatic K=2 kernel choice for a pair space of `cols` columns (an upper bound, rounded up to 64) over
// `n_words` words: the fp4 MFMA Gram wherever it measured faster than the VALU AND+popcount tiles on MI355X (64 to
// 512 columns over 256 to 31,250 words: everywhere except below 128 columns over 16,384 words or more, where the
// two are within 10%).
inline bool gram_preferred(int64_t cols, int64_t n_words) { return cols >= 128 || n_words < 16384; }

// The process's pool of host threads (uploads, result exports): fn(0 .. parts - 1), part 0 on the caller, the
// rest on the pool; returns when all are done. parallel_width() is the pool's threads plus the caller.
int parallel_width();
void parallel_for(int parts, const std::function<void(int)>& fn);

hipError_t launch_groups_single(const int32_t* rows, const int64_t* d_n, int64_t n_cap, int32_t m, int32_t small_max,
                                int64_t* group_start, int64_t* cum_pairs, int64_t* cum_tiles, int32_t* small_list,
                                int64_t* totals, hipStream_t s);
hipError_t launch_compact_single(const int32_t* rows, int32_t m, const int64_t* group_start, const int64_t* cum_pairs,
                                 int64_t n_groups, int64_t n_cands, const uint8_t* status, const int32_t* counts,
                                 const int32_t* gid, int64_t min_count, int32_t* out_rows, int32_t* out_counts,
                                 int64_t* out_cand, int64_t* d_out_n, hipStream_t s);
hipError_t launch_select_single(const int32_t* rows, const int64_t* d_n, int64_t n_cap, int32_t m,
                                const int32_t* counts, const uint8_t* flag, int32_t* out_rows, int32_t* out_counts,
                                int64_t* d_out_n, hipStream_t s);
hipError_t launch_hash(const int32_t* rows, int64_t n_cap, const int64_t* d_n, int32_t m, int32_t* table,
                       int64_t table_size, hipStream_t s);
hipError_t launch_free(const int32_t* rows, int64_t n_cap, const int64_t* d_n, int32_t k, const int32_t* counts,
                       const int32_t* prev_rows, const int32_t* prev_counts, const int32_t* prev_table,
                       int64_t table_size, uint8_t* free_out, hipStream_t s, bool hinted = false);
// The survivors of a level already counted by etm_level_compact's first call (workspace ws) written out, with
// each survivor's free flag pre-set in free_hint (0: inferred count, hence not free; 1: to be computed by
// launch_free(..., hinted = true)).
hipError_t level_compact_hinted(const int32_t* rows, int32_t m, const int64_t* group_start, const int64_t* cum_pairs,
                                int64_t n_groups, int64_t n_cands, const uint8_t* status, const int32_t* counts,
                                const int32_t* gid, int64_t min_count, int32_t* out_rows, int32_t* out_counts,
                                uint8_t* free_hint, const int64_t* ws, hipStream_t s);
hipError_t launch_fill_u8(uint8_t* p, int64_t n, uint8_t v, hipStream_t s);

// Level 2 (rows2: n2 lexsorted pairs of column ids < n_cols) as a pair index for the K=3 subset test: a bitmap
// `bits` (n_cols x words uint64, words = ceil(n_cols / 64)) and `rank` (n_cols x words int32, the row of the first
// pair in each word). pair_index_bytes is their size.
int64_t pair_index_bytes(int64_t n_cols);
hipError_t launch_pair_index(const int32_t* rows2, int64_t n2, int64_t n_cols, uint64_t* bits, int32_t* rank,
                             hipStream_t s);
// etm_level_classify with an optional pair index (used when m = 2; null pointers: the hash).
int level_classify(const int32_t* d_rows, int64_t n, int32_t m, const int32_t* d_prev_counts,
                   const uint8_t* d_prev_free, const int32_t* d_table, int64_t table_size,
                   const int64_t* d_group_start, const int64_t* d_cum_pairs, int64_t n_groups, int64_t n_cands,
                   const int32_t* d_group_list, const int64_t* h_class_counts, int32_t mode, uint8_t* d_status,
                   int32_t* d_counts, int32_t* d_gid, int64_t* d_tally, const uint64_t* d_pair_bits,
                   const int32_t* d_pair_rank, int64_t pair_words, void* stream);
hipError_t launch_zero_i64(int64_t* p, int n, hipStream_t s);
// etm_count_level's live-unit path (LEX layout, a status array) with the caller's live counters d_counters[0..15],
// zero on entry (so no zeroing launch is needed); d_ws holds etm_count_level_ws(...) bytes. With n_tiled > 0 (the
// tiled class's groups, listed after the four small classes in d_small_list), the tiled groups whose counted
// pairs are few go to the group-listed kernel; d_gl_ws holds count_glist_ws(...) bytes.
int count_level_live(const uint64_t* d_bv, int64_t pitch, int64_t n_words, const int32_t* d_rows, int32_t m,
                     const int64_t* d_group_start, const int64_t* d_cum_pairs, const int64_t* d_cum_tiles,
                     int64_t n_groups, int64_t n_tile_pairs, const int32_t* d_small_list,
                     const int64_t* h_class_counts, const uint8_t* d_status, int32_t* d_counts, void* d_ws,
                     int64_t* d_counters, int64_t n_tiled, int64_t n_cands, int64_t n_rows, void* d_gl_ws,
                     void* stream);
// Workspace bytes of count_level_live's group-listed mode for a level of n_rows rows in n_groups groups, n_tiled
// of them in the tiled class, with n_cands candidates (0 without tiled groups).
size_t count_glist_ws(int64_t n_groups, int64_t n_tiled, int64_t n_cands, int64_t n_rows);
// What the counting kernels read for the candidates left ETM_CAND_COUNT (status of n_cands candidates of a level
// of n_rows rows in n_groups prefix groups; gid optional): adds the suffix rows in a counted pair to out[0], the
// groups with a counted pair to out[1] and the counted candidates to out[2]. row_bits and group_bits are scratch
// of ceil(n_rows / 64) and ceil(n_groups / 64) words.
hipError_t launch_live_bytes(const uint8_t* status, const int32_t* gid, const int64_t* group_start,
                             const int64_t* cum_pairs, int64_t n_groups, int64_t n_cands, int64_t n_rows,
                             uint64_t* row_bits, uint64_t* group_bits, int64_t* out, hipStream_t s);

// Rounds of a batch of jobs (batch.h): each launch covers n_jobs job records, reads every size from them, and
// skips the jobs that are not active. The plan launch (count.hip) must precede the hash, classification and
// counting launches of a round; compaction and the next groups end it.
hipError_t launch_batch_plan(etm::batch::Job* jobs, int n_jobs, etm::batch::Queue* q, hipStream_t s);
hipError_t launch_batch_hash(etm::batch::Job* jobs, int n_jobs, hipStream_t s);
hipError_t launch_batch_classify(etm::batch::Job* jobs, int n_jobs, hipStream_t s);
hipError_t launch_batch_count(const etm::batch::Job* jobs, int n_jobs, etm::batch::Queue* q, hipStream_t s);
hipError_t launch_batch_compact(etm::batch::Job* jobs, int n_jobs, hipStream_t s);
hipError_t launch_batch_groups(etm::batch::Job* jobs, int n_jobs, etm::batch::Queue* q, bool init, hipStream_t s);

// The batched input layer (bitvec.hip): per phase (0 id ranges, 1 histograms, 2 bitvectors, 3 K=1 counts) one
// launch over n_jobs InJob records; pre holds the exclusive prefix of batch_input_blocks over the jobs and
// n_blocks its total. batch_input_shape sets a job's column slab width (phase 2 needs max_slab_cols, the largest).
void batch_input_shape(etm::batch::InJob& j);
int64_t batch_input_blocks(int phase, const etm::batch::InJob& j);
hipError_t launch_batch_input(int phase, etm::batch::InJob* jobs, int n_jobs, const int64_t* pre, int64_t n_blocks,
                              int64_t max_slab_cols, hipStream_t s);
hipError_t launch_fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t s);

// etm_bitvec_build, and with d_part (bitvec_part_len(pitch, n_cols) uint16) and d_counts (n_cols) the column
// popcounts as well, from per-tile partials the builder writes beside the bitvectors (no second pass over them).
int64_t bitvec_part_len(int64_t pitch, int32_t n_cols);
int bitvec_build(const int64_t* d_indptr, const void* d_ids, int id_bytes, int64_t n_rows, int64_t lo, int64_t span,
                 const int32_t* d_lut, int32_t n_cols, int64_t pitch, const int32_t* d_order, uint64_t* d_out,
                 uint16_t* d_part, int64_t* d_counts, void* stream);

}  // namespace detail
}  // namespace etm
