// Device state of a batch of independent mining jobs advanced in lockstep rounds (etm_mine_batch).
//
// Every job of a batch keeps one Job record in device memory. A round advances every active job by one level
// with a fixed sequence of launches that read all their sizes from these records (no host read between them):
//
//   k_batch_plan      one workgroup: per job the counting units of the round and its hash size; prefix sums of
//                     the units over the jobs (the work queue's index space); zeroed queue heads and tallies
//   k_batch_hash_*    the hash of each job's previous level (the subset test's index)
//   k_batch_classify  per candidate: the (k-1)-subset test (status and count slot)
//   k_batch_count_*   persistent counting kernels fed by the device work queue: each wave (small groups) or
//                     workgroup (tile-pairs) takes the next unit index with one atomic add and maps it to
//                     (job, group or tile-pair, word chunk) through the prefix sums
//   k_batch_compact   a workgroup per job: survivors in lexicographic order into the job's level store
//   k_batch_groups    a workgroup per job: the next level's prefix groups, then the job's record advances
//                     (or the job ends: done, or overflow when its next level would pass the caps)
//
// The round is therefore a static launch sequence, captured once as a HIP graph and replayed per round.
#pragma once

#include <cstdint>

namespace etm {
namespace batch {

constexpr int kMaxJobs = 256;          // jobs per batch call (one plan thread each)
constexpr int kMaxLevels = 64;         // levels K >= 2 recorded per job
constexpr int64_t kCapCands = 32768;   // candidates of a level (the single-workgroup compaction's limit)
constexpr int64_t kTableSize = 65536;  // hash slots per job: a power of two >= 2 x the rows of a level
constexpr int kMaxCols = 256;          // K=1 columns of an eligible job: K=2 has at most 32,640 candidates

enum : int32_t { kActive = 0, kDone = 1, kOverflow = 2 };

// A finished level K >= 2 of a job: its survivors (rows of k column ids at rows_off, counts at counts_off, in
// the job's level store, in ints) and its statistics.
struct Level {
    int64_t n, rows_off, counts_off, cands, counted, skipped, bytes_alg, t_end;
};

struct Job {
    // Fixed for the call.
    const uint64_t* bv;
    int64_t pitch, n_words, min_count;
    int32_t max_len, small_max;
    int32_t* table;
    uint8_t* status;
    int32_t* counts;
    int32_t* gid;
    int64_t* gs[2];
    int64_t* cp[2];
    int64_t* ct[2];
    int32_t* list[2];
    int32_t* out;  // level store: level 1 first, then each level's rows and counts
    int64_t out_cap;
    // State, advanced on the device.
    int32_t k;         // items of the round's candidates (the previous level holds k - 1)
    int32_t state;     // kActive, kDone, kOverflow
    int32_t buf;       // group buffer of the previous level
    int32_t n_levels;  // entries of lv
    int64_t out_used;
    const int32_t* prev_rows;
    const int32_t* prev_counts;
    int64_t prev_n, table_mask;
    int64_t tot[2][24];  // group totals per buffer: [0] groups, [1] candidates, [2] tile-pairs, [5] groups of
                         // >= 2 rows, [6] their rows, [8 + c] groups of size class c
    // Round scratch.
    int64_t tally[4];  // counted, inferred, skipped
    int64_t s_chunk, s_nch, t_chunk, t_nch;
    int64_t new_n;
    int64_t t_init;  // device wall clock when the rounds began (the first level's start)
    Level lv[kMaxLevels];
};

// The round's work queues, one per unit kind: small groups of size classes 0, 1, 2 (2, 3-4, 5-8 suffixes; a batch
// job's small_max is 8, so class 3 stays empty) and tile-pairs (kTiledQueue). A queue's index space is the
// exclusive prefix of its units (groups or tile-pairs x word chunks) over the jobs; its head is the next index.
constexpr int kSmallQueues = 3;
constexpr int kTiledQueue = 3;
constexpr int kQueues = 4;
struct Queue {
    int64_t pre[kQueues][kMaxJobs + 1];
    unsigned long long head[kQueues];
    int64_t active;        // jobs active at the round's start
    int64_t active_after;  // jobs still active at its end (the host stops the rounds at 0)
};

// One job of the batched input layer: id range, occurrence histogram, bitvectors and K=1 counts of every job in
// one launch per phase. Each phase's launch flattens the jobs' workgroups; pre (n_jobs + 1) is the exclusive
// prefix of the workgroups over the jobs.
struct InJob {
    const int64_t* indptr;
    const void* ids;
    int64_t n_rows, nnz, lo, span;
    int32_t id_bytes, n_cols;
    int64_t pitch, n_words, slab_cols;
    int32_t* hist;       // span bins (zeroed before the histogram launch)
    const int32_t* lut;  // id - lo -> column, or -1
    uint64_t* bv;        // (n_cols, pitch)
    int64_t* mm;         // id range: [0] min, [1] max
    int64_t* pc;         // K=1 counts (n_cols)
};

}  // namespace batch
}  // namespace etm
