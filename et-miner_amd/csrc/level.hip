// Level management on the device: prefix groups of a lexsorted level, the
// open-addressing hash of a level, candidate classification (the (k-1)-subset
// test and Pascal inference), survivor compaction in lexicographic order, and
// free flags. Nothing here returns to the host except the scalar totals the
// caller reads to size the next buffers.
#include <algorithm>
#include <mutex>

#include "batch.h"
#include "common.h"
#include "internal.h"
#include "scan.h"

namespace {

using etm::item_hash;

constexpr int kTileT = 64;  // suffixes per tile of the tiled counting kernel (count.hip)

__device__ __forceinline__ int32_t row_item(const int32_t* __restrict__ rows, int64_t r, int32_t m, int t) {
    return rows[r * (int64_t)m + t];
}

// Row r starts a prefix group: r == 0, or its first m-1 items differ from row r-1's. Branch-free over the
// items, four pairs of loads at a time: an early exit per item would make each load wait on the previous
// comparison (m-1 dependent round trips on deep levels).
struct IsGroupStart {
    const int32_t* rows;
    int32_t m;
    __device__ int64_t operator()(int64_t r) const {
        if (r == 0) return 1;
        const int32_t* a = rows + r * (int64_t)m;
        const int32_t* b = a - m;
        int32_t d = 0;
        int t = 0;
        for (; t + 4 <= m - 1; t += 4) {
            const int32_t a0 = a[t], a1 = a[t + 1], a2 = a[t + 2], a3 = a[t + 3];
            const int32_t b0 = b[t], b1 = b[t + 1], b2 = b[t + 2], b3 = b[t + 3];
            d |= (a0 ^ b0) | (a1 ^ b1) | (a2 ^ b2) | (a3 ^ b3);
        }
        for (; t < m - 1; ++t) d |= a[t] ^ b[t];
        return d != 0;
    }
};

struct WriteGroupStart {
    int64_t* group_start;
    __device__ void operator()(int64_t r, int64_t v, int64_t prefix) const {
        if (v) group_start[prefix] = r;
    }
};

struct SetGroupEnd {
    int64_t* group_start;
    int64_t n;
    int64_t* totals;
    __device__ void operator()(int64_t n_groups) const {
        group_start[n_groups] = n;
        totals[0] = n_groups;
    }
};

// Per group g (g < totals[0]; 0 beyond): its candidate pairs, or its tile-pairs for the tiled kernel.
struct GroupPairs {
    const int64_t* group_start;
    const int64_t* totals;
    __device__ int64_t operator()(int64_t g) const {
        if (g >= totals[0]) return 0;
        const int64_t s = group_start[g + 1] - group_start[g];
        return s * (s - 1) / 2;
    }
};

struct GroupTiles {
    const int64_t* group_start;
    const int64_t* totals;
    int32_t small_max;
    __device__ int64_t operator()(int64_t g) const {
        if (g >= totals[0]) return 0;
        const int64_t s = group_start[g + 1] - group_start[g];
        if (s <= small_max) return 0;
        const int64_t nt = (s + kTileT - 1) / kTileT;
        return nt * (nt + 1) / 2;
    }
};

struct WritePrefix {
    int64_t* out;
    __device__ void operator()(int64_t i, int64_t, int64_t prefix) const { out[i] = prefix; }
};

struct StoreTotal {
    int64_t* slot;
    __device__ void operator()(int64_t t) const { *slot = t; }
};

__global__ void k_zero_i64(int64_t* p, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

// Size class of a group of s suffixes: 0 (2), 1 (3-4), 2 (5-8), 3 (9-16) for the wave-per-group kernels,
// 4 (more than small_max) for the tiled kernel, -1 (fewer than 2: no candidates).
constexpr int kClasses = 5;
__device__ __forceinline__ int size_class(int64_t s, int32_t small_max) {
    if (s < 2) return -1;
    if (s > small_max) return 4;
    return s <= 2 ? 0 : s <= 4 ? 1 : s <= 8 ? 2 : 3;
}

// Pass 1 (classes): class sizes and the totals; pass 2 (place): each group at its class's offset, ascending
// within a class. Both passes cut the groups into the same gridDim.x contiguous ranges, one per workgroup. Pass 1
// writes each range's tallies to a partial row (no contended atomics: thousands of device-scope atomics on a
// few counters serialise) and the last workgroup to finish sums the rows into the totals and
// turns the class columns into exclusive offsets per range; pass 2 then places a range's groups from its
// offsets with a workgroup-wide ballot scan per 256 groups.
constexpr int kGroupBlocks = 256;  // most workgroups of the two passes
constexpr int kPartCols = 3 + kClasses;  // n2, rows2, smax, then the class sizes
constexpr int kGroupDone = 23;           // totals[] slot of pass 1's completion counter

__device__ __forceinline__ void group_range(int64_t n_groups, int64_t* lo, int64_t* hi) {
    const int64_t per = (n_groups + gridDim.x - 1) / gridDim.x;
    *lo = min(n_groups, (int64_t)blockIdx.x * per);
    *hi = min(n_groups, *lo + per);
}

__global__ __launch_bounds__(256) void k_group_classes(const int64_t* __restrict__ group_start, int64_t n,
                                                        int32_t small_max, int64_t* __restrict__ totals,
                                                        int64_t* __restrict__ part) {
    const int64_t n_groups = min(totals[0], n);
    int64_t lo, hi;
    group_range(n_groups, &lo, &hi);
    int64_t n2 = 0, rows2 = 0, smax = 0, cls[kClasses] = {0, 0, 0, 0, 0};
    for (int64_t g = lo + threadIdx.x; g < hi; g += blockDim.x) {
        const int64_t s = group_start[g + 1] - group_start[g];
        const int c = size_class(s, small_max);
        if (c >= 0) cls[c] += 1;
        if (s >= 2) {
            n2 += 1;
            rows2 += s;
        }
        smax = s > smax ? s : smax;
    }
    __shared__ int64_t lds[5];
    __shared__ int last;
    int64_t v[kPartCols] = {n2, rows2, 0, cls[0], cls[1], cls[2], cls[3], cls[4]};
    for (int c = 0; c < kPartCols; ++c)
        if (c != 2) v[c] = etm::block_sum<256>(v[c], lds);
    smax = etm::wave_max(smax);
    __shared__ int64_t wmax[4];
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = smax;
    __syncthreads();
    if (threadIdx.x == 0) {
        v[2] = max(max(wmax[0], wmax[1]), max(wmax[2], wmax[3]));
        for (int c = 0; c < kPartCols; ++c) part[(int64_t)blockIdx.x * kPartCols + c] = v[c];
        __threadfence();
        const unsigned long long done = __hip_atomic_fetch_add((unsigned long long*)&totals[kGroupDone], 1ull,
                                                               __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        last = done == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    // The last workgroup, a thread per range: column sums into the totals, then the class columns as exclusive
    // offsets per range (class c's list starts after the classes before it), by a scan over the ranges in LDS.
    __shared__ int64_t scan[kClasses][kGroupBlocks];
    __shared__ int64_t sums[kClasses];
    const unsigned b = threadIdx.x;
    const bool have = b < gridDim.x;
    int64_t row[kPartCols];
    for (int c = 0; c < kPartCols; ++c) row[c] = have ? part[(int64_t)b * kPartCols + c] : 0;
    const int64_t t_n2 = etm::block_sum<256>(row[0], lds);
    const int64_t t_rows2 = etm::block_sum<256>(row[1], lds);
    const int64_t w_max = etm::wave_max(row[2]);
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = w_max;
    for (int c = 0; c < kClasses; ++c) scan[c][b] = row[3 + c];
    __syncthreads();
    for (int off = 1; off < kGroupBlocks; off <<= 1) {  // inclusive scans of the five class columns
        int64_t add[kClasses];
        for (int c = 0; c < kClasses; ++c) add[c] = b >= (unsigned)off ? scan[c][b - off] : 0;
        __syncthreads();
        for (int c = 0; c < kClasses; ++c) scan[c][b] += add[c];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        int64_t at = 0;
        for (int c = 0; c < kClasses; ++c) {
            sums[c] = at;  // the class's first slot
            const int64_t total = scan[c][kGroupBlocks - 1];
            totals[8 + c] = total;
            at += total;
            if (c == kClasses - 2) totals[3] = at;  // the groups of the four wave-per-group classes
        }
        totals[4] = max(max(wmax[0], wmax[1]), max(wmax[2], wmax[3]));
        totals[5] = t_n2;
        totals[6] = t_rows2;
    }
    __syncthreads();
    if (have)
        for (int c = 0; c < kClasses; ++c) part[(int64_t)b * kPartCols + 3 + c] = sums[c] + scan[c][b] - row[3 + c];
    if (threadIdx.x == 0) totals[kGroupDone] = 0;
}

__global__ __launch_bounds__(256) void k_group_place(const int64_t* __restrict__ group_start, int64_t n,
                                                      int32_t small_max, int32_t* __restrict__ small_list,
                                                      const int64_t* __restrict__ totals,
                                                      const int64_t* __restrict__ part) {
    const int64_t n_groups = min(totals[0], n);
    int64_t lo, hi;
    group_range(n_groups, &lo, &hi);
    __shared__ int64_t at[kClasses];
    __shared__ int wcount[4][kClasses];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
    if (threadIdx.x < kClasses) at[threadIdx.x] = part[(int64_t)blockIdx.x * kPartCols + 3 + threadIdx.x];
    for (int64_t g0 = lo; g0 < hi; g0 += blockDim.x) {
        const int64_t g = g0 + threadIdx.x;
        const int c = g < hi ? size_class(group_start[g + 1] - group_start[g], small_max) : -1;
        uint64_t mine = 0;
#pragma unroll
        for (int q = 0; q < kClasses; ++q) {
            const uint64_t mask = __ballot(c == q);
            if (c == q) mine = mask;
            if (lane == 0) wcount[wave][q] = __popcll(mask);
        }
        __syncthreads();
        if (c >= 0) {
            int64_t pos = at[c] + __popcll(mine & below);
            for (int w = 0; w < wave; ++w) pos += wcount[w][c];
            small_list[pos] = (int32_t)g;
        }
        __syncthreads();
        if (threadIdx.x < kClasses)
            at[threadIdx.x] += wcount[0][threadIdx.x] + wcount[1][threadIdx.x] + wcount[2][threadIdx.x] +
                               wcount[3][threadIdx.x];
        __syncthreads();
    }
}

// ── hash of a level ─────────────────────────────────────────────────────

__global__ void k_fill_table(int32_t* t, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        t[i] = -1;
}

__global__ __launch_bounds__(256) void k_hash_insert(const int32_t* __restrict__ rows, int64_t n, int32_t m,
                                                      int32_t* __restrict__ table, int64_t mask,
                                                      const int64_t* d_n = nullptr) {
    if (d_n) n = min(n, *d_n);
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        uint64_t h = 0;
        for (int t = 0; t < m; ++t) h += item_hash(row_item(rows, r, m, t));
        int64_t slot = (int64_t)(h & (uint64_t)mask);
        while (atomicCAS(&table[slot], -1, (int32_t)r) != -1) slot = (slot + 1) & mask;
    }
}

// The itemset made of `pre` (np items, dropping position `skip`; skip >= np keeps all), then a, then b
// (use_b false: only a). Item t of it:
struct Query {
    const int32_t* pre;
    int np, skip;
    int32_t a, b;
    bool use_b;
    __device__ int size() const { return np - (skip < np ? 1 : 0) + 1 + (use_b ? 1 : 0); }
    __device__ int32_t item(int t) const {
        const int kept = np - (skip < np ? 1 : 0);
        if (t < kept) return pre[(skip < np && t >= skip) ? t + 1 : t];
        return t == kept ? a : b;
    }
    // Whether a level row of size() items equals the query: the kept prefix items against `pre`, then a and b
    // as values (an item() loop would select between a global and a private address per item).
    __device__ __forceinline__ bool match(const int32_t* row, int32_t) const {
        const int kept = np - (skip < np ? 1 : 0);
        for (int t = 0; t < kept; ++t)
            if (row[t] != pre[t + (t >= skip ? 1 : 0)]) return false;
        if (row[kept] != a) return false;
        return !use_b || row[kept + 1] == b;
    }
};

// Row id of the level row equal to q (q.item(0..m)), or -1. Inlined with q by value, so the query stays in
// registers (a query passed by reference to an outlined call lives in scratch memory).
template <class Q>
__device__ __forceinline__ int64_t hash_find(const int32_t* __restrict__ rows, int32_t m,
                                             const int32_t* __restrict__ table, int64_t mask, uint64_t h, const Q q) {
    int64_t slot = (int64_t)(h & (uint64_t)mask);
    for (;;) {
        const int32_t r = table[slot];
        if (r < 0) return -1;
        if (q.match(rows + (int64_t)r * m, m)) return r;
        slot = (slot + 1) & mask;
    }
}

// ── classification ──────────────────────────────────────────────────────

__global__ __launch_bounds__(256) void k_classify(const int32_t* __restrict__ rows, int32_t m,
                                                   const int32_t* __restrict__ prev_counts,
                                                   const uint8_t* __restrict__ prev_free,
                                                   const int32_t* __restrict__ table, int64_t mask,
                                                   const int64_t* __restrict__ group_start,
                                                   const int64_t* __restrict__ cum_pairs, int64_t n_groups,
                                                   int64_t n_cands, int32_t mode, uint8_t* __restrict__ status,
                                                   int32_t* __restrict__ counts, int64_t* __restrict__ tally) {
    __shared__ int64_t lds[5];
    const bool infer = (mode & ETM_SUBSET_INFER) != 0;
    const bool prune = (mode & ETM_SUBSET_PRUNE) != 0;
    int64_t n_count = 0, n_infer = 0, n_skip = 0;
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n_cands;
         c += (int64_t)gridDim.x * blockDim.x) {
        const int64_t g = etm::upper_index(cum_pairs, n_groups, c);
        const int64_t gs = group_start[g];
        const int64_t s = group_start[g + 1] - gs;
        int64_t i, j;
        etm::lex_pair_decode(c - cum_pairs[g], s, &i, &j);
        const int32_t* pre = rows + gs * m;  // the group's prefix: the first m-1 items of its first row
        const int np = m - 1;
        const int64_t ri = gs + i, rj = gs + j;  // the two prefix-parents (rows of the level)
        const int32_t a = rows[ri * m + np], b = rows[rj * m + np];
        uint64_t h = item_hash(a) + item_hash(b);
        for (int t = 0; t < np; ++t) h += item_hash(pre[t]);
        int32_t lowest = prev_counts ? min(prev_counts[ri], prev_counts[rj]) : 0;
        bool not_free = infer && (prev_free[ri] == 0 || prev_free[rj] == 0);
        uint8_t st = ETM_CAND_COUNT;
        if (prune || infer) {
            for (int d = 0; d < np; ++d) {
                Query q{pre, np, d, a, b, true};
                const int64_t r = hash_find(rows, m, table, mask, h - item_hash(pre[d]), q);
                if (r < 0) {
                    if (prune) {
                        st = ETM_CAND_SKIP;
                        break;
                    }
                    not_free = false;  // without the subset Pascal cannot be applied
                    lowest = -1;
                    break;
                }
                if (infer) {
                    lowest = min(lowest, prev_counts[r]);
                    not_free = not_free || prev_free[r] == 0;
                }
            }
        }
        if (st == ETM_CAND_COUNT && infer && not_free && lowest >= 0) st = ETM_CAND_INFER;
        status[c] = st;
        counts[c] = st == ETM_CAND_COUNT ? 0 : (st == ETM_CAND_INFER ? lowest : -1);
        n_count += st == ETM_CAND_COUNT;
        n_infer += st == ETM_CAND_INFER;
        n_skip += st == ETM_CAND_SKIP;
    }
    if (tally) {
        n_count = etm::block_sum<256>(n_count, lds);
        n_infer = etm::block_sum<256>(n_infer, lds);
        n_skip = etm::block_sum<256>(n_skip, lds);
        if (threadIdx.x == 0) {
            if (n_count) atomicAdd((unsigned long long*)&tally[0], (unsigned long long)n_count);
            if (n_infer) atomicAdd((unsigned long long*)&tally[1], (unsigned long long)n_infer);
            if (n_skip) atomicAdd((unsigned long long*)&tally[2], (unsigned long long)n_skip);
        }
    }
}

// Group-centric classification, a thread per candidate: small groups packed several to a wave by size class,
// big groups split into ranges of kBigPairs candidates. Both write the group id of every candidate (compaction
// then needs no search).

struct Classifier {
    const int32_t* rows;
    int32_t m;
    const int32_t* prev_counts;
    const uint8_t* prev_free;
    const int32_t* table;
    int64_t mask;
    int32_t mode;
    // Level 2 as a pair index (set only when the level holds pairs, m = 2): bit b of bitmap row a is set for
    // each pair (a, b) of column ids, and pair_rank[a * pair_words + w] is that pair's row when it is the first
    // of its word, so a K=3 candidate's one subset test is a bitmap word and a rank instead of a hash probe.
    const uint64_t* pair_bits;
    const int32_t* pair_rank;
    int64_t pair_words;

    __device__ int64_t pair_row(int32_t a, int32_t b) const {
        const int64_t at = (int64_t)a * pair_words + (b >> 6);
        const uint64_t w = pair_bits[at], bit = 1ull << (b & 63);
        return (w & bit) ? (int64_t)pair_rank[at] + __popcll(w & (bit - 1)) : -1;
    }

    // Row of the subset of candidate (i, j) of the group at row gs that drops prefix item d, or -1.
    __device__ int64_t probe(int64_t gs, int64_t i, int64_t j, int d) const {
        const int32_t* pre = rows + gs * m;
        const int np = m - 1;
        const int32_t a = rows[(gs + i) * m + np], b = rows[(gs + j) * m + np];
        if (pair_bits) return pair_row(a, b);
        uint64_t h = item_hash(a) + item_hash(b);
        for (int t = 0; t < np; ++t)
            if (t != d) h += item_hash(pre[t]);
        Query q{pre, np, d, a, b, true};
        return hash_find(rows, m, table, mask, h, q);
    }

    // Status and count of candidate (i, j) of the group starting at row gs: the hash of all its items is
    // formed once, and each (k-1)-subset dropping a prefix item is one subtraction and one probe away.
    __device__ void classify(int64_t gs, int64_t i, int64_t j, uint8_t* st_out, int32_t* cnt_out) const {
        const bool infer = (mode & ETM_SUBSET_INFER) != 0;
        const bool prune = (mode & ETM_SUBSET_PRUNE) != 0;
        const int32_t* pre = rows + gs * m;
        const int np = m - 1;
        const int64_t ri = gs + i, rj = gs + j;
        const int32_t a = rows[ri * m + np], b = rows[rj * m + np];
        int32_t lowest = prev_counts ? min(prev_counts[ri], prev_counts[rj]) : 0;
        bool nf = infer && (prev_free[ri] == 0 || prev_free[rj] == 0);
        bool missing = false;
        if (pair_bits && (prune || infer)) {
            const int64_t r = pair_row(a, b);
            if (r < 0) {
                missing = true;
            } else if (infer) {
                lowest = min(lowest, prev_counts[r]);
                nf = nf || prev_free[r] == 0;
            }
        } else if (prune || infer) {
            uint64_t h = item_hash(a) + item_hash(b);
            for (int t = 0; t < np; ++t) h += item_hash(pre[t]);
            for (int d = 0; d < np && !missing; ++d) {
                Query q{pre, np, d, a, b, true};
                const int64_t r = hash_find(rows, m, table, mask, h - item_hash(pre[d]), q);
                if (r < 0) {
                    missing = true;
                } else if (infer) {
                    lowest = min(lowest, prev_counts[r]);
                    nf = nf || prev_free[r] == 0;
                }
            }
        }
        uint8_t st = ETM_CAND_COUNT;
        if (missing && prune) st = ETM_CAND_SKIP;
        else if (!missing && infer && nf) st = ETM_CAND_INFER;
        *st_out = st;
        *cnt_out = st == ETM_CAND_COUNT ? 0 : (st == ETM_CAND_INFER ? lowest : -1);
    }
};

// LDS written by one wave's lanes, made visible to the same wave's other lanes.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Small groups packed into waves, a lane per (candidate, dropped prefix item): a candidate's k-2 subset probes
// run in parallel lanes (their latency does not add up on deep levels) and several groups share a wave when
// they have few (candidate, probe) pairs. Size class c (groups of at most 2, 4, 8, 16 rows: at most
// P = 1, 6, 28, 120 candidates) puts G = max(1, 64 / (P x probes)) groups in a wave; each candidate's outcome
// gathers in LDS (at most 120 candidates per wave). `n` holds the class sizes; the class lists follow one
// another in `list`.
constexpr int kPackedSlots = 120;

struct PackedClasses {
    int64_t n[4];
    int32_t probes;  // k-2 (at least 1)
    __device__ __host__ static int per_group(int c) { return c == 0 ? 1 : c == 1 ? 6 : c == 2 ? 28 : 120; }
    __device__ __host__ int groups(int c) const {
        const int g = 64 / (per_group(c) * probes);
        return g < 1 ? 1 : g;
    }
};

int64_t packed_waves(const PackedClasses& pc) {
    int64_t w = 0;
    for (int c = 0; c < 4; ++c) w += (pc.n[c] + pc.groups(c) - 1) / pc.groups(c);
    return w;
}

__global__ __launch_bounds__(256) void k_classify_packed(Classifier C, const int64_t* __restrict__ group_start,
                                                          const int64_t* __restrict__ cum_pairs,
                                                          const int32_t* __restrict__ list, PackedClasses pc,
                                                          uint8_t* __restrict__ status, int32_t* __restrict__ counts,
                                                          int32_t* __restrict__ gid, int64_t* __restrict__ tally) {
    // Per wave: its G groups' bounds and prefix items (with their hashes), and each candidate slot's two suffix
    // items, staged in LDS in three rounds of independent loads; a probe lane then reads only the hash table and
    // the row it points to (a lane re-deriving all of this from the group list costs some eight dependent global
    // loads per probe).
    __shared__ int32_t s_min[4][kPackedSlots];
    __shared__ int32_t s_flags[4][kPackedSlots];  // bit 0: a subset is missing, bit 1: a subset is not free
    __shared__ int32_t s_a[4][kPackedSlots], s_b[4][kPackedSlots];
    __shared__ int64_t s_gs[4][64], s_cb[4][64];
    __shared__ int32_t s_sz[4][64], s_g[4][64];
    __shared__ int32_t s_pre[4][64];
    __shared__ uint64_t s_hpre[4][64];
    __shared__ int64_t lds[5];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const bool infer = (C.mode & ETM_SUBSET_INFER) != 0;
    const bool prune = (C.mode & ETM_SUBSET_PRUNE) != 0;
    int64_t wave = (int64_t)blockIdx.x * 4 + wv;
    int c = 0;
    int64_t base = 0;
    for (; c < 4; ++c) {
        const int64_t wc = (pc.n[c] + pc.groups(c) - 1) / pc.groups(c);
        if (wave < wc) break;
        wave -= wc;
        base += pc.n[c];
    }
    int64_t n_count = 0, n_infer = 0, n_skip = 0;
    if (c < 4) {
        const int P = PackedClasses::per_group(c), G = pc.groups(c);
        const int m = C.m, np = m - 1;
        const int slots = G * P;
        // Round 1: the groups.
        if (lane < G) {
            const int64_t li = wave * G + lane;
            int32_t g = -1;
            int64_t gs = 0, sz = 0, cb = 0;
            if (li < pc.n[c]) {
                g = list[base + li];
                gs = group_start[g];
                sz = group_start[g + 1] - gs;
                cb = cum_pairs[g];
            }
            s_g[wv][lane] = g;
            s_gs[wv][lane] = gs;
            s_sz[wv][lane] = (int32_t)sz;
            s_cb[wv][lane] = cb;
        }
        wave_lds_sync();
        // Round 2: each group's prefix items (G x np <= 64) and each slot's suffix items and parent counts.
        for (int t = lane; t < G * np; t += 64) {
            const int gi = t / np, q = t % np;
            const int32_t v = s_sz[wv][gi] >= 2 ? C.rows[s_gs[wv][gi] * m + q] : 0;
            s_pre[wv][t] = v;
            s_hpre[wv][t] = item_hash(v);
        }
        for (int sl = lane; sl < slots; sl += 64) {
            const int gi = sl / P, p = sl % P;
            const int64_t sz = s_sz[wv][gi];
            int32_t mn = 0, fl = 4, a = 0, b = 0;  // fl bit 2: no such candidate
            if (p < sz * (sz - 1) / 2) {
                int64_t i, j;
                etm::lex_pair_decode(p, sz, &i, &j);
                const int64_t ri = s_gs[wv][gi] + i, rj = s_gs[wv][gi] + j;
                a = C.rows[ri * m + np];
                b = C.rows[rj * m + np];
                mn = C.prev_counts ? min(C.prev_counts[ri], C.prev_counts[rj]) : 0;
                fl = (infer && (C.prev_free[ri] == 0 || C.prev_free[rj] == 0)) ? 2 : 0;
            }
            s_a[wv][sl] = a;
            s_b[wv][sl] = b;
            s_min[wv][sl] = mn;
            s_flags[wv][sl] = fl;
        }
        wave_lds_sync();
        // Round 3: the probes, a lane per (slot, dropped prefix item).
        if (prune || infer) {
            for (int t = lane; t < slots * np; t += 64) {
                const int sl = t / np, d = t % np;
                if (s_flags[wv][sl] & 4) continue;
                const int gi = sl / P;
                const int32_t a = s_a[wv][sl], b = s_b[wv][sl];
                int64_t r;
                if (C.pair_bits) {
                    r = C.pair_row(a, b);
                } else {
                    const int32_t* pre = &s_pre[wv][gi * np];
                    const uint64_t* hp = &s_hpre[wv][gi * np];
                    uint64_t h = item_hash(a) + item_hash(b);
                    for (int q = 0; q < np; ++q)
                        if (q != d) h += hp[q];
                    r = hash_find(C.rows, m, C.table, C.mask, h, Query{pre, np, d, a, b, true});
                }
                if (r < 0) {
                    atomicOr(&s_flags[wv][sl], 1);
                } else if (infer) {
                    atomicMin(&s_min[wv][sl], C.prev_counts[r]);
                    if (C.prev_free[r] == 0) atomicOr(&s_flags[wv][sl], 2);
                }
            }
        }
        wave_lds_sync();
        for (int sl = lane; sl < slots; sl += 64) {
            const int f = s_flags[wv][sl];
            if (f & 4) continue;
            const int gi = sl / P;
            uint8_t st = ETM_CAND_COUNT;
            if ((f & 1) && prune) st = ETM_CAND_SKIP;
            else if (!(f & 1) && infer && (f & 2)) st = ETM_CAND_INFER;
            const int64_t cand = s_cb[wv][gi] + sl % P;
            status[cand] = st;
            counts[cand] = st == ETM_CAND_COUNT ? 0 : (st == ETM_CAND_INFER ? s_min[wv][sl] : -1);
            if (gid) gid[cand] = s_g[wv][gi];
            n_count += st == ETM_CAND_COUNT;
            n_infer += st == ETM_CAND_INFER;
            n_skip += st == ETM_CAND_SKIP;
        }
    }
    if (tally) {
        n_count = etm::block_sum<256>(n_count, lds);
        n_infer = etm::block_sum<256>(n_infer, lds);
        n_skip = etm::block_sum<256>(n_skip, lds);
        if (threadIdx.x == 0) {
            if (n_count) atomicAdd((unsigned long long*)&tally[0], (unsigned long long)n_count);
            if (n_infer) atomicAdd((unsigned long long*)&tally[1], (unsigned long long)n_infer);
            if (n_skip) atomicAdd((unsigned long long*)&tally[2], (unsigned long long)n_skip);
        }
    }
}

// Big groups: blockIdx.y picks the group, blockIdx.x a range of kBigPairs of its candidates.
constexpr int64_t kBigPairs = 4096;

__global__ __launch_bounds__(256) void k_classify_big(Classifier C, const int64_t* __restrict__ group_start,
                                                       const int64_t* __restrict__ cum_pairs,
                                                       const int32_t* __restrict__ list,
                                                       uint8_t* __restrict__ status, int32_t* __restrict__ counts,
                                                       int32_t* __restrict__ gid, int64_t* __restrict__ tally) {
    __shared__ int64_t lds[5];
    const int64_t g = list[blockIdx.y];
    const int64_t gs = group_start[g];
    const int64_t s = group_start[g + 1] - gs;
    const int64_t npairs = s * (s - 1) / 2;
    const int64_t p0 = (int64_t)blockIdx.x * kBigPairs;
    if (p0 >= npairs) return;
    const int64_t p1 = min(npairs, p0 + kBigPairs);
    const int64_t cbase = cum_pairs[g];
    int64_t n_count = 0, n_infer = 0, n_skip = 0;
    for (int64_t p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
        int64_t i, j;
        etm::lex_pair_decode(p, s, &i, &j);
        uint8_t st;
        int32_t cnt;
        C.classify(gs, i, j, &st, &cnt);
        const int64_t c = cbase + p;
        status[c] = st;
        counts[c] = cnt;
        if (gid) gid[c] = (int32_t)g;
        n_count += st == ETM_CAND_COUNT;
        n_infer += st == ETM_CAND_INFER;
        n_skip += st == ETM_CAND_SKIP;
    }
    if (tally) {
        n_count = etm::block_sum<256>(n_count, lds);
        n_infer = etm::block_sum<256>(n_infer, lds);
        n_skip = etm::block_sum<256>(n_skip, lds);
        if (threadIdx.x == 0) {
            if (n_count) atomicAdd((unsigned long long*)&tally[0], (unsigned long long)n_count);
            if (n_infer) atomicAdd((unsigned long long*)&tally[1], (unsigned long long)n_infer);
            if (n_skip) atomicAdd((unsigned long long*)&tally[2], (unsigned long long)n_skip);
        }
    }
}

// ── deduction bounds ────────────────────────────────────────────────────
//
// For a candidate X of k items whose (k-1)-subsets are all frequent and free, with s(Y) the support of Y:
//   upper u = min( min_a s(X-a), min_{a<b<c} [s(X-a)+s(X-b)+s(X-c) - s(X-ab)-s(X-ac)-s(X-bc) + s(X-abc)] )
//   lower l = max( 0, max_{a<b} [s(X-a) + s(X-b) - s(X-ab)] )
// (inclusion-exclusion over the rows missing a, b, c). u < min_count proves X infrequent; l = u gives its count.

constexpr int kBoundsMaxK = 24;

// Candidate x[0..k) without positions d0 < d1 < d2 (an unused position is kNoPos). Item t of it:
struct DropQuery {
    static constexpr int kNoPos = 1 << 30;
    const int32_t* x;
    int d0, d1, d2;
    __device__ int32_t item(int t) const {
        t += t >= d0;
        t += t >= d1;
        t += t >= d2;
        return x[t];
    }
    __device__ __forceinline__ bool match(const int32_t* row, int32_t m) const {
        for (int t = 0; t < m; ++t)
            if (row[t] != item(t)) return false;
        return true;
    }
};

// A level the bounds read: rows (n, m), counts, and its hash.
struct DeepLevel {
    const int32_t* rows;
    const int32_t* counts;
    const int32_t* table;
    int64_t mask;
};

// Positions (i < j) of pair p in colex order p = j(j-1)/2 + i.
__device__ __forceinline__ void colex_pair(int p, int* i, int* j) {
    int b = 1;
    while ((b + 1) * b / 2 <= p) ++b;
    *j = b;
    *i = p - b * (b - 1) / 2;
}

// Positions (i < j < l) of triple q in colex order q = C(l,3) + C(j,2) + i.
__device__ __forceinline__ void colex_triple(int q, int* i, int* j, int* l) {
    int c = 2;
    while ((c + 1) * c * (c - 1) / 6 <= q) ++c;
    *l = c;
    colex_pair(q - c * (c - 1) * (c - 2) / 6, i, j);
}

__device__ __forceinline__ int64_t sub_max(int64_t v, int L) {
    for (int off = L >> 1; off > 0; off >>= 1) v = max(v, (int64_t)__shfl_xor((long long)v, off, 64));
    return v;
}

__device__ __forceinline__ int64_t sub_min(int64_t v, int L) {
    for (int off = L >> 1; off > 0; off >>= 1) v = min(v, (int64_t)__shfl_xor((long long)v, off, 64));
    return v;
}

// A sub-wave of L lanes (a power of two <= 64) per candidate still ETM_CAND_COUNT: lanes over the k (k-1)-
// subsets, then the C(k,2) pair drops (probes of level k-2), then the C(k,3) triple drops (probes of level k-3;
// the empty set's count n_total at k = 3). Every lane of a wave runs the same iterations (the reductions are
// wave-wide shuffles); each sub-wave keeps its candidate's items, their hashes, s(X-a) and s(X-ab) in LDS.
__global__ __launch_bounds__(256) void k_deduce(Classifier C, DeepLevel d2, DeepLevel d3, int64_t n_total,
                                                 int64_t min_count, const int64_t* __restrict__ group_start,
                                                 const int64_t* __restrict__ cum_pairs, int64_t n_groups,
                                                 const int32_t* __restrict__ gid, int64_t n_cands, int L,
                                                 int stride8, uint8_t* __restrict__ status,
                                                 int32_t* __restrict__ counts, int64_t* __restrict__ tally) {
    extern __shared__ uint64_t dyn_lds[];
    __shared__ int64_t lds[5];
    const int k = C.m + 1, m = C.m;
    const int P2 = k * (k - 1) / 2, P3 = k * (k - 1) * (k - 2) / 6;
    const int per_block = blockDim.x / L;
    const int sub = threadIdx.x / L, sl = threadIdx.x % L;
    uint64_t* hx = dyn_lds + (int64_t)sub * stride8;
    int32_t* x = (int32_t*)(hx + k);
    int32_t* c1 = x + k;
    int32_t* s2 = c1 + k;
    int64_t n_inf = 0, n_skip = 0;
    // Each wave scans 64 candidates at a time and runs the bounds only for those still ETM_CAND_COUNT, 64 / L
    // of them per round (sub-wave w takes the w-th left): on deep levels of inferred candidates most waves
    // find none and move on after one coalesced status load.
    const int lane = threadIdx.x & 63, per_wave = 64 / L, sub_w = lane / L;
    const int64_t wave_step = (int64_t)gridDim.x * (blockDim.x >> 6) * 64;
    for (int64_t base = ((int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 64; base < n_cands;
         base += wave_step) {
      uint64_t left = __ballot(base + lane < n_cands && status[base + lane] == ETM_CAND_COUNT);
      while (left) {
        uint64_t mine = left;
        for (int t = 0; t < sub_w && mine; ++t) mine &= mine - 1;
        for (int t = 0; t < per_wave && left; ++t) left &= left - 1;
        const int64_t c = mine ? base + __ffsll((long long)mine) - 1 : -1;
        const bool act = c >= 0;
        int64_t gs = 0, i = 0, j = 0;
        if (act) {
            const int64_t g = gid ? (int64_t)gid[c] : etm::upper_index(cum_pairs, n_groups, c);
            gs = group_start[g];
            etm::lex_pair_decode(c - cum_pairs[g], group_start[g + 1] - gs, &i, &j);
            for (int t = sl; t < k; t += L) {
                const int32_t v = t < m - 1 ? C.rows[gs * m + t] : C.rows[(gs + (t == m - 1 ? i : j)) * m + m - 1];
                x[t] = v;
                hx[t] = item_hash(v);
            }
        }
        wave_lds_sync();
        uint64_t h = 0;
        int64_t miss = 0;
        if (act) {
            for (int t = 0; t < k; ++t) h += hx[t];
            // s(X - x[t]): dropping a or b is a parent row of the group; a prefix item is a probe.
            for (int t = sl; t < k; t += L) {
                int64_t r;
                if (t == k - 2) r = gs + j;
                else if (t == k - 1) r = gs + i;
                else r = hash_find(C.rows, m, C.table, C.mask, h - hx[t], DropQuery{x, t, DropQuery::kNoPos,
                                                                                       DropQuery::kNoPos});
                if (r < 0) miss = 1;
                c1[t] = r < 0 ? 0 : C.prev_counts[r];
            }
        }
        wave_lds_sync();
        int64_t lower = 0;
        if (act) {
            for (int p = sl; p < P2; p += L) {
                int a, b;
                colex_pair(p, &a, &b);
                const int64_t r = hash_find(d2.rows, m - 1, d2.table, d2.mask, h - hx[a] - hx[b],
                                            DropQuery{x, a, b, DropQuery::kNoPos});
                int32_t v = 0;
                if (r < 0) miss = 1;
                else v = d2.counts[r];
                s2[p] = v;
                lower = max(lower, (int64_t)c1[a] + c1[b] - v);
            }
        }
        wave_lds_sync();
        int64_t upper = INT64_MAX;
        if (act) {
            for (int t = 0; t < k; ++t) upper = min(upper, (int64_t)c1[t]);
            for (int q = sl; q < P3; q += L) {
                int a, b, e;
                colex_triple(q, &a, &b, &e);
                int64_t v3 = n_total;
                if (k > 3) {
                    const int64_t r = hash_find(d3.rows, m - 2, d3.table, d3.mask, h - hx[a] - hx[b] - hx[e],
                                                DropQuery{x, a, b, e});
                    if (r < 0) miss = 1;
                    else v3 = d3.counts[r];
                }
                const int64_t u = (int64_t)c1[a] + c1[b] + c1[e] - s2[b * (b - 1) / 2 + a] -
                                  s2[e * (e - 1) / 2 + a] - s2[e * (e - 1) / 2 + b] + v3;
                upper = min(upper, u);
            }
        }
        lower = sub_max(lower, L);
        upper = sub_min(upper, L);
        miss = sub_max(miss, L);
        if (act && sl == 0 && !miss) {
            if (upper < min_count) {
                status[c] = ETM_CAND_SKIP;
                counts[c] = -1;
                ++n_skip;
            } else if (lower == upper) {
                status[c] = ETM_CAND_INFER;
                counts[c] = (int32_t)upper;
                ++n_inf;
            }
        }
        wave_lds_sync();  // the next candidate's items overwrite this one's
      }
    }
    if (tally) {
        n_inf = etm::block_sum<256>(n_inf, lds);
        n_skip = etm::block_sum<256>(n_skip, lds);
        if (threadIdx.x == 0 && (n_inf || n_skip)) {
            atomicAdd((unsigned long long*)&tally[0], (unsigned long long)(-(n_inf + n_skip)));
            if (n_inf) atomicAdd((unsigned long long*)&tally[1], (unsigned long long)n_inf);
            if (n_skip) atomicAdd((unsigned long long*)&tally[2], (unsigned long long)n_skip);
        }
    }
}

// ── compaction ──────────────────────────────────────────────────────────

struct Survives {
    const uint8_t* status;
    const int32_t* counts;
    int64_t min_count;
    __device__ int64_t operator()(int64_t c) const {
        if (status && status[c] == ETM_CAND_SKIP) return 0;
        return (int64_t)counts[c] >= min_count ? 1 : 0;
    }
};

struct WriteSurvivor {
    const int32_t* rows;
    int32_t m;
    const int64_t* group_start;
    const int64_t* cum_pairs;
    int64_t n_groups;
    const int32_t* counts;
    int32_t* out_rows;
    int32_t* out_counts;
    int64_t* out_cand;
    const int32_t* gid;  // group of every candidate, or null (then found by binary search)
    // With free_hint: each survivor's free flag before the free-flag pass, 0 where the survivor's count was
    // inferred (its count is the smallest of its subsets', so one subset has it: not free), 1 where it is to be
    // computed.
    const uint8_t* status = nullptr;
    uint8_t* free_hint = nullptr;
    __device__ void operator()(int64_t c, int64_t v, int64_t p) const {
        if (!v) return;
        if (free_hint) free_hint[p] = (status != nullptr && status[c] == ETM_CAND_INFER) ? 0 : 1;
        const int64_t g = gid ? gid[c] : (n_groups == 1 ? 0 : etm::upper_index(cum_pairs, n_groups, c));
        const int64_t gs = group_start[g];
        const int64_t s = group_start[g + 1] - gs;
        int64_t i, j;
        etm::lex_pair_decode(c - cum_pairs[g], s, &i, &j);
        const int k = m + 1;
        int32_t* o = out_rows + p * k;
        const int32_t* first = rows + gs * m;
        for (int t = 0; t < m - 1; ++t) o[t] = first[t];
        o[m - 1] = rows[(gs + i) * m + m - 1];
        o[m] = rows[(gs + j) * m + m - 1];
        out_counts[p] = counts[c];
        if (out_cand) out_cand[p] = c;
    }
};

struct Flagged {
    const uint8_t* flag;
    __device__ int64_t operator()(int64_t i) const { return flag[i] ? 1 : 0; }
};

struct WriteSelected {
    const int32_t* rows;
    int32_t m;
    const int32_t* counts;
    int32_t* out_rows;
    int32_t* out_counts;
    __device__ void operator()(int64_t i, int64_t v, int64_t p) const {
        if (!v) return;
        for (int t = 0; t < m; ++t) out_rows[p * m + t] = rows[i * (int64_t)m + t];
        out_counts[p] = counts[i];
    }
};

// ── free flags ──────────────────────────────────────────────────────────

// With hinted, free_out already holds 0 for the rows known not to be free (inferred counts), which are skipped.
__global__ __launch_bounds__(256) void k_free(const int32_t* __restrict__ rows, int64_t n, int32_t k,
                                               const int32_t* __restrict__ counts,
                                               const int32_t* __restrict__ prev_rows,
                                               const int32_t* __restrict__ prev_counts,
                                               const int32_t* __restrict__ table, int64_t mask,
                                               uint8_t* __restrict__ free_out, const int64_t* d_n, bool hinted) {
    if (d_n) n = min(n, *d_n);
    for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < n; x += (int64_t)gridDim.x * blockDim.x) {
        if (hinted && free_out[x] == 0) continue;
        const int32_t* row = rows + x * k;
        uint64_t h = 0;
        for (int t = 0; t < k; ++t) h += item_hash(row[t]);
        uint8_t is_free = 1;
        for (int d = 0; d < k && is_free; ++d) {
            // The subset dropping row[d]: row[0..k) without position d (k-1 items).
            int64_t slot = (int64_t)((h - item_hash(row[d])) & (uint64_t)mask);
            for (;;) {
                const int32_t r = table[slot];
                if (r < 0) break;
                const int32_t* pr = prev_rows + (int64_t)r * (k - 1);
                bool eq = true;
                for (int t = 0; t < k - 1 && eq; ++t) eq = pr[t] == row[t < d ? t : t + 1];
                if (eq) {
                    if (prev_counts[r] == counts[x]) is_free = 0;
                    break;
                }
                slot = (slot + 1) & mask;
            }
        }
        free_out[x] = is_free;
    }
}

// ── single-workgroup fast paths (small levels: one launch, sizes read on the device) ─────────

constexpr int kOne = 1024;
constexpr int kOneWaves = kOne / 64;

// One pass of a workgroup-wide exclusive scan of K int32 values per item over items [base, base + kOne*S).
// Wave v owns the contiguous run [base + 64*S*v, base + 64*S*(v+1)), lane l its items 64*j + l (coalesced).
// Every value is evaluated once, all of a lane's loads are issued before any scan; the S rows of a wave are
// scanned on the DPP crossbar, and one barrier publishes the wave totals, which each wave scans itself.
// f(i, v) fills v[K] for item i < n; w(i, v, prefix) is called for every item i < n with its exclusive
// prefixes (carry included). carry[] advances by the pass totals. Values and totals must fit int32.
template <int K, int S, class F, class W>
__device__ __forceinline__ void stripe_pass(int64_t base, int64_t n, const F& f, const W& w, int32_t (&carry)[K],
                                            int32_t* lds) {
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const int64_t wbase = base + (int64_t)wv * 64 * S;
    int32_t v[S][K], ex[S][K];
#pragma unroll
    for (int j = 0; j < S; ++j) {
        const int64_t i = wbase + j * 64 + lane;
        if (i < n) {
            f(i, v[j]);
        } else {
#pragma unroll
            for (int q = 0; q < K; ++q) v[j][q] = 0;
        }
    }
    int32_t run[K];
#pragma unroll
    for (int q = 0; q < K; ++q) run[q] = 0;
#pragma unroll
    for (int j = 0; j < S; ++j) {
#pragma unroll
        for (int q = 0; q < K; ++q) {
            const int32_t inc = etm::wave_scan_i32(v[j][q]);
            ex[j][q] = run[q] + inc - v[j][q];
            run[q] += __builtin_amdgcn_readlane(inc, 63);
        }
    }
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < K; ++q) lds[q * kOneWaves + wv] = run[q];
    }
    __syncthreads();
    int32_t off[K];
#pragma unroll
    for (int q = 0; q < K; ++q) {
        const int32_t t = lane < kOneWaves ? lds[q * kOneWaves + lane] : 0;
        const int32_t tinc = etm::wave_scan_i32(t);
        off[q] = carry[q] + __shfl(tinc - t, wv, 64);
        carry[q] += __builtin_amdgcn_readlane(tinc, kOneWaves - 1);
    }
#pragma unroll
    for (int j = 0; j < S; ++j) {
        const int64_t i = wbase + j * 64 + lane;
        if (i < n) {
            int32_t pre[K];
#pragma unroll
            for (int q = 0; q < K; ++q) pre[q] = off[q] + ex[j][q];
            w(i, v[j], pre);
        }
    }
    __syncthreads();
}

// A single-valued sweep over items [0, n): passes of kOne*S items. Returns the total.
template <int S, class F, class W>
__device__ __forceinline__ int64_t stripe_scan(int64_t n, const F& f, const W& w, int32_t* lds) {
    int32_t carry[1] = {0};
    auto f1 = [&](int64_t i, int32_t(&v)[1]) { v[0] = (int32_t)f(i); };
    auto w1 = [&](int64_t i, const int32_t(&v)[1], const int32_t(&p)[1]) { w(i, (int64_t)v[0], (int64_t)p[0]); };
    for (int64_t base = 0; base < n; base += (int64_t)kOne * S) stripe_pass<1, S>(base, n, f1, w1, carry, lds);
    return carry[0];
}

struct GroupSize {
    const int64_t* group_start;
    __device__ int64_t operator()(int64_t g) const { return group_start[g + 1] - group_start[g]; }
};

// The whole group build of a level of *d_n rows (at most n_cap <= 32768) in one workgroup: group starts,
// cumulative pairs and tile-pairs, size-class lists (groups in ascending order within each class) and the
// totals. Three sweeps: rows (group starts), groups (pairs, tile-pairs and class counts, classes 0-3 packed
// two 16-bit counts per int32: a class of groups of >= 2 rows holds at most 16384 of them), groups again (class
// lists at their class's base). Groups of one row (no candidates) are in no class.
__device__ __forceinline__ void groups_single_body(const int32_t* __restrict__ rows, int64_t n, int32_t m,
                                                   int32_t small_max, int64_t* __restrict__ group_start,
                                                   int64_t* __restrict__ cum_pairs, int64_t* __restrict__ cum_tiles,
                                                   int32_t* __restrict__ small_list, int64_t* __restrict__ totals,
                                                   int32_t* lds, int64_t* red) {
    const int tid = threadIdx.x;
    IsGroupStart is_start{rows, m};
    auto put_start = [&](int64_t r, int64_t v, int64_t prefix) {
        if (v && prefix < n) group_start[prefix] = r;
    };
    const int64_t n_groups = stripe_scan<8>(n, is_start, put_start, lds);
    if (tid == 0) group_start[n_groups] = n;
    __syncthreads();
    GroupSize size{group_start};
    // Sweep 2: pairs, tile-pairs, class counts.
    int64_t n2 = 0, rows2 = 0, smax = 0;
    auto vals = [&](int64_t g, int32_t(&v)[5]) {
        const int64_t z = size(g);
        const int64_t nt = (z + kTileT - 1) / kTileT;
        const int c = size_class(z, small_max);
        v[0] = (int32_t)(z * (z - 1) / 2);
        v[1] = z > small_max ? (int32_t)(nt * (nt + 1) / 2) : 0;
        v[2] = c == 0 ? 1 : (c == 1 ? (1 << 16) : 0);
        v[3] = c == 2 ? 1 : (c == 3 ? (1 << 16) : 0);
        v[4] = c == 4 ? 1 : 0;
    };
    auto put = [&](int64_t g, const int32_t(&v)[5], const int32_t(&p)[5]) {
        cum_pairs[g] = p[0];
        cum_tiles[g] = p[1];
        const int64_t z = size(g);
        if (z >= 2) {
            n2 += 1;
            rows2 += z;
        }
        smax = z > smax ? z : smax;
    };
    int32_t c4[5] = {0, 0, 0, 0, 0};
    for (int64_t base = 0; base < n_groups; base += (int64_t)kOne * 2) stripe_pass<5, 2>(base, n_groups, vals, put, c4, lds);
    const int64_t cp = c4[0], ct = c4[1];
    int64_t cls[kClasses];
    cls[0] = c4[2] & 0xffff;
    cls[1] = (uint32_t)c4[2] >> 16;
    cls[2] = c4[3] & 0xffff;
    cls[3] = (uint32_t)c4[3] >> 16;
    cls[4] = c4[4];
    if (tid == 0) {
        cum_pairs[n_groups] = cp;
        cum_tiles[n_groups] = ct;
    }
    n2 = etm::block_sum<kOne>(n2, red);
    rows2 = etm::block_sum<kOne>(rows2, red);
    smax = etm::wave_max(smax);
    if ((tid & 63) == 0) red[tid / 64] = smax;
    __syncthreads();
    for (int w = 0; w < kOneWaves; ++w) smax = red[w] > smax ? red[w] : smax;
    __syncthreads();
    // Sweep 3: the class lists, ascending within each class, class c at base[c].
    int64_t cbase[kClasses];
    cbase[0] = 0;
    for (int c = 1; c < kClasses; ++c) cbase[c] = cbase[c - 1] + cls[c - 1];
    auto cvals = [&](int64_t g, int32_t(&v)[3]) {
        const int c = size_class(size(g), small_max);
        v[0] = c == 0 ? 1 : (c == 1 ? (1 << 16) : 0);
        v[1] = c == 2 ? 1 : (c == 3 ? (1 << 16) : 0);
        v[2] = c == 4 ? 1 : 0;
    };
    auto cput = [&](int64_t g, const int32_t(&v)[3], const int32_t(&p)[3]) {
        int64_t at;
        if (v[0] & 0xffff) at = cbase[0] + (p[0] & 0xffff);
        else if (v[0]) at = cbase[1] + ((uint32_t)p[0] >> 16);
        else if (v[1] & 0xffff) at = cbase[2] + (p[1] & 0xffff);
        else if (v[1]) at = cbase[3] + ((uint32_t)p[1] >> 16);
        else if (v[2]) at = cbase[4] + p[2];
        else return;
        if (at < n_groups) small_list[at] = (int32_t)g;
    };
    int32_t c3[3] = {0, 0, 0};
    for (int64_t base = 0; base < n_groups; base += (int64_t)kOne * 2) stripe_pass<3, 2>(base, n_groups, cvals, cput, c3, lds);
    if (tid == 0) {
        totals[0] = n_groups;
        totals[1] = cp;
        totals[2] = ct;
        totals[3] = cls[0] + cls[1] + cls[2] + cls[3];
        totals[4] = smax;
        totals[5] = n2;
        totals[6] = rows2;
        totals[7] = n;
        for (int c = 0; c < kClasses; ++c) totals[8 + c] = cls[c];
        for (int c = kClasses; c < 8; ++c) totals[8 + c] = 0;
        for (int c = 0; c < 8; ++c) totals[16 + c] = 0;
    }
}

__global__ __launch_bounds__(kOne) void k_groups_single(const int32_t* __restrict__ rows, const int64_t* d_n,
                                                        int64_t n_cap, int32_t m, int32_t small_max,
                                                        int64_t* __restrict__ group_start,
                                                        int64_t* __restrict__ cum_pairs,
                                                        int64_t* __restrict__ cum_tiles,
                                                        int32_t* __restrict__ small_list,
                                                        int64_t* __restrict__ totals) {
    __shared__ int32_t lds[5 * kOneWaves];
    __shared__ int64_t red[kOneWaves + 1];
    const int64_t n = d_n ? min(*d_n, n_cap) : n_cap;
    groups_single_body(rows, n, m, small_max, group_start, cum_pairs, cum_tiles, small_list, totals, lds, red);
}

// Survivor compaction of up to n_cands (<= 32768) candidates in one workgroup, in lexicographic order; the
// survivor count goes to *d_out_n. Passes of kOne x 4 candidates, lane-striped per wave as in stripe_pass. The
// arguments are restrict-qualified and each pass issues all its loads (status and count, then group id, then
// group bounds, then the two suffix items) before its first store, so the dependent loads of a lane's four
// candidates overlap instead of queueing behind one another's stores.
__device__ __forceinline__ int64_t compact_single_body(const uint8_t* __restrict__ status,
                                                       const int32_t* __restrict__ counts, int64_t min_count,
                                                       const int32_t* __restrict__ gid,
                                                       const int32_t* __restrict__ rows, int32_t m,
                                                       const int64_t* __restrict__ group_start,
                                                       const int64_t* __restrict__ cum_pairs, int64_t n_groups,
                                                       int64_t n_cands, int32_t* __restrict__ out_rows,
                                                       int32_t* __restrict__ out_counts,
                                                       int64_t* __restrict__ out_cand, int32_t* lds) {
    constexpr int S = 4;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int64_t carry = 0;
    for (int64_t base = 0; base < n_cands; base += (int64_t)kOne * S) {
        const int64_t wbase = base + (int64_t)wv * 64 * S;
        int64_t c[S];
        int32_t v[S], cnt[S], ex[S];
#pragma unroll
        for (int j = 0; j < S; ++j) {
            c[j] = wbase + j * 64 + lane;
            v[j] = 0;
            cnt[j] = 0;
            if (c[j] < n_cands) {
                cnt[j] = counts[c[j]];
                const bool skip = status != nullptr && status[c[j]] == ETM_CAND_SKIP;
                v[j] = (!skip && (int64_t)cnt[j] >= min_count) ? 1 : 0;
            }
        }
        int32_t run = 0;
#pragma unroll
        for (int j = 0; j < S; ++j) {
            const int32_t inc = etm::wave_scan_i32(v[j]);
            ex[j] = run + inc - v[j];
            run += __builtin_amdgcn_readlane(inc, 63);
        }
        if (lane == 0) lds[wv] = run;
        __syncthreads();
        const int32_t t = lane < kOneWaves ? lds[lane] : 0;
        const int32_t tinc = etm::wave_scan_i32(t);
        const int64_t off = carry + __shfl(tinc - t, wv, 64);
        const int64_t total = __builtin_amdgcn_readlane(tinc, kOneWaves - 1);
        // Loads of every surviving candidate of the pass, in stages.
        int64_t g[S], gs[S], sz[S], cb[S], ii[S], jj[S];
        int32_t a[S], b[S];
#pragma unroll
        for (int j = 0; j < S; ++j)
            g[j] = v[j] ? (gid ? (int64_t)gid[c[j]] : (n_groups == 1 ? 0 : etm::upper_index(cum_pairs, n_groups, c[j])))
                        : 0;
#pragma unroll
        for (int j = 0; j < S; ++j) {
            gs[j] = group_start[g[j]];
            sz[j] = group_start[g[j] + 1] - gs[j];
            cb[j] = cum_pairs[g[j]];
        }
#pragma unroll
        for (int j = 0; j < S; ++j) {
            ii[j] = 0;
            jj[j] = 1;
            if (v[j]) etm::lex_pair_decode(c[j] - cb[j], sz[j], &ii[j], &jj[j]);
        }
#pragma unroll
        for (int j = 0; j < S; ++j) {
            a[j] = v[j] ? rows[(gs[j] + ii[j]) * m + m - 1] : 0;
            b[j] = v[j] ? rows[(gs[j] + jj[j]) * m + m - 1] : 0;
        }
#pragma unroll
        for (int j = 0; j < S; ++j) {
            const int64_t p = off + ex[j];
            if (!v[j] || p >= n_cands) continue;
            int32_t* o = out_rows + p * (m + 1);
            const int32_t* first = rows + gs[j] * m;
            for (int q = 0; q < m - 1; ++q) o[q] = first[q];
            o[m - 1] = a[j];
            o[m] = b[j];
            out_counts[p] = cnt[j];
            if (out_cand) out_cand[p] = c[j];
        }
        carry += total;
        __syncthreads();
    }
    return carry;
}

__global__ __launch_bounds__(kOne) void k_compact_single(const uint8_t* __restrict__ status,
                                                         const int32_t* __restrict__ counts, int64_t min_count,
                                                         const int32_t* __restrict__ gid,
                                                         const int32_t* __restrict__ rows, int32_t m,
                                                         const int64_t* __restrict__ group_start,
                                                         const int64_t* __restrict__ cum_pairs, int64_t n_groups,
                                                         int64_t n_cands, int32_t* __restrict__ out_rows,
                                                         int32_t* __restrict__ out_counts,
                                                         int64_t* __restrict__ out_cand,
                                                         int64_t* __restrict__ d_out_n) {
    __shared__ int32_t lds[kOneWaves];
    const int64_t n = compact_single_body(status, counts, min_count, gid, rows, m, group_start, cum_pairs, n_groups,
                                          n_cands, out_rows, out_counts, out_cand, lds);
    if (threadIdx.x == 0) *d_out_n = n;
}

// Rows whose flag is set, in order, in one workgroup (n read from *d_n, at most n_cap).
__global__ __launch_bounds__(kOne) void k_select_single(const int32_t* __restrict__ rows, const int64_t* d_n,
                                                        int64_t n_cap, int32_t m, const int32_t* __restrict__ counts,
                                                        const uint8_t* __restrict__ flag, int32_t* __restrict__ out_rows,
                                                        int32_t* __restrict__ out_counts, int64_t* __restrict__ d_out_n) {
    __shared__ int32_t lds[kOneWaves];
    const int64_t n = d_n ? min(*d_n, n_cap) : n_cap;
    WriteSelected w{rows, m, counts, out_rows, out_counts};
    Flagged f{flag};
    auto put = [&](int64_t i, int64_t v, int64_t p) {
        if (p < n) w(i, v, p);
    };
    const int64_t total = stripe_scan<8>(n, f, put, lds);
    if (threadIdx.x == 0) *d_out_n = total;
}

// The pair index of a level of pairs (rows2, lexsorted): bits first (one atomic OR per pair), then a wave per
// bitmap row scans its words' popcounts from the row's first pair in rows2 (a binary search on the first item).
__global__ __launch_bounds__(256) void k_pair_index_bits(const int32_t* __restrict__ rows2, int64_t n2,
                                                          int64_t words, uint64_t* __restrict__ bits) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n2; p += (int64_t)gridDim.x * blockDim.x) {
        const int32_t a = rows2[2 * p], b = rows2[2 * p + 1];
        atomicOr((unsigned long long*)&bits[(int64_t)a * words + (b >> 6)], 1ull << (b & 63));
    }
}

__global__ __launch_bounds__(256) void k_pair_index_rank(const int32_t* __restrict__ rows2, int64_t n2,
                                                          int64_t n_cols, int64_t words,
                                                          const uint64_t* __restrict__ bits,
                                                          int32_t* __restrict__ rank) {
    const int lane = threadIdx.x & 63;
    const int64_t n_waves = (int64_t)gridDim.x * (blockDim.x / 64);
    for (int64_t a = (int64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6); a < n_cols; a += n_waves) {
        int64_t lo = 0, hi = n2;  // first row whose first item is >= a
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (rows2[2 * mid] < a) lo = mid + 1;
            else hi = mid;
        }
        int32_t run = (int32_t)lo;
        for (int64_t w0 = 0; w0 < words; w0 += 64) {
            const int64_t w = w0 + lane;
            const int32_t v = w < words ? __popcll(bits[a * words + w]) : 0;
            const int32_t inc = etm::wave_scan_i32(v);
            if (w < words) rank[a * words + w] = run + inc - v;
            run += __builtin_amdgcn_readlane(inc, 63);
        }
    }
}

int grid_for(int64_t n, int cap = 16384) {
    int64_t g = (n + 255) / 256;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

bool is_pow2(int64_t x) { return x > 0 && (x & (x - 1)) == 0; }

__global__ void k_fill_u8(uint8_t* p, int64_t n, uint8_t v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// ── what the counting kernels must read ────────────────────────────────

// Marks, for every candidate left ETM_CAND_COUNT, its group (bit g of group_bits) and its two suffix rows (bits
// gs + i and gs + j of row_bits), and counts the counted candidates into out[2].
__global__ __launch_bounds__(256) void k_live_mark(const uint8_t* __restrict__ status, const int32_t* __restrict__ gid,
                                                   const int64_t* __restrict__ group_start,
                                                   const int64_t* __restrict__ cum_pairs, int64_t n_groups,
                                                   int64_t n_cands, unsigned long long* __restrict__ row_bits,
                                                   unsigned long long* __restrict__ group_bits, int64_t* out) {
    __shared__ int64_t lds[256 / etm::kWave + 1];
    int64_t n_live = 0;
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n_cands; c += (int64_t)gridDim.x * blockDim.x) {
        if (status[c] != ETM_CAND_COUNT) continue;
        ++n_live;
        const int64_t g = gid ? gid[c] : etm::upper_index(cum_pairs, n_groups, c);
        const int64_t gs = group_start[g];
        int64_t i, j;
        etm::lex_pair_decode(c - cum_pairs[g], group_start[g + 1] - gs, &i, &j);
        atomicOr(&group_bits[g >> 6], 1ull << (g & 63));
        atomicOr(&row_bits[(gs + i) >> 6], 1ull << ((gs + i) & 63));
        atomicOr(&row_bits[(gs + j) >> 6], 1ull << ((gs + j) & 63));
    }
    n_live = etm::block_sum<256>(n_live, lds);
    if (threadIdx.x == 0 && n_live) atomicAdd((unsigned long long*)&out[2], (unsigned long long)n_live);
}

// out[0] += set bits of row_bits (n_row_words words), out[1] += set bits of group_bits (n_group_words words).
__global__ __launch_bounds__(256) void k_live_sum(const uint64_t* __restrict__ row_bits, int64_t n_row_words,
                                                  const uint64_t* __restrict__ group_bits, int64_t n_group_words,
                                                  int64_t* out) {
    __shared__ int64_t lds[256 / etm::kWave + 1];
    int64_t a = 0, b = 0;
    for (int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; w < n_row_words || w < n_group_words;
         w += (int64_t)gridDim.x * blockDim.x) {
        if (w < n_row_words) a += __popcll(row_bits[w]);
        if (w < n_group_words) b += __popcll(group_bits[w]);
    }
    a = etm::block_sum<256>(a, lds);
    b = etm::block_sum<256>(b, lds);
    if (threadIdx.x == 0) {
        if (a) atomicAdd((unsigned long long*)&out[0], (unsigned long long)a);
        if (b) atomicAdd((unsigned long long*)&out[1], (unsigned long long)b);
    }
}

// ── batches of jobs in lockstep rounds (batch.h) ──────────────────────────
//
// blockIdx.y (or blockIdx.x for the workgroup-per-job kernels) is the job; a job that is not active returns at
// once, so a round's launches keep one shape whatever the jobs' states.

using etm::batch::Job;

constexpr int kBatchFillBlocks = 16;    // workgroups per job clearing its hash slots
constexpr int kBatchInsertBlocks = 32;  // workgroups per job inserting its previous level (32 x 1024 rows)
constexpr int kBatchClassifyBlocks = etm::batch::kCapCands / 256;  // a thread per candidate slot

__global__ __launch_bounds__(256) void k_batch_hash_fill(const Job* __restrict__ jobs) {
    const Job& J = jobs[blockIdx.y];
    if (J.state != etm::batch::kActive || J.k < 3) return;
    int32_t* t = J.table;
    const int64_t size = J.table_mask + 1;
    for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < size; i += (int64_t)gridDim.x * 256) t[i] = -1;
}

__global__ __launch_bounds__(256) void k_batch_hash_insert(const Job* __restrict__ jobs) {
    const Job& J = jobs[blockIdx.y];
    if (J.state != etm::batch::kActive || J.k < 3) return;
    const int32_t m = J.k - 1;
    const int64_t mask = J.table_mask;
    const int32_t* rows = J.prev_rows;
    int32_t* table = J.table;
    for (int64_t r = blockIdx.x * 256 + threadIdx.x; r < J.prev_n; r += (int64_t)gridDim.x * 256) {
        uint64_t h = 0;
        for (int t = 0; t < m; ++t) h += item_hash(row_item(rows, r, m, t));
        int64_t slot = (int64_t)(h & (uint64_t)mask);
        while (atomicCAS(&table[slot], -1, (int32_t)r) != -1) slot = (slot + 1) & mask;
    }
}

// Status, count slot and group of every candidate of the round: at K=2 every pair is counted; from K=3 a
// candidate with a (k-1)-subset missing from the previous level is skipped (the subset test of the plain loop,
// by hash probes).
__global__ __launch_bounds__(256) void k_batch_classify(Job* __restrict__ jobs) {
    __shared__ int64_t lds[5];
    Job& J = jobs[blockIdx.y];
    if (J.state != etm::batch::kActive) return;
    const int b = J.buf;
    const int32_t m = J.k - 1;
    const int64_t n_cands = J.tot[b][1], n_groups = J.tot[b][0];
    const int64_t* gs = J.gs[b];
    const int64_t* cp = J.cp[b];
    const Classifier C{J.prev_rows, m, nullptr, nullptr, J.table, J.table_mask, ETM_SUBSET_PRUNE, nullptr, nullptr, 0};
    int64_t n_count = 0, n_skip = 0;
    for (int64_t c = blockIdx.x * 256 + threadIdx.x; c < n_cands; c += (int64_t)gridDim.x * 256) {
        const int64_t g = n_groups == 1 ? 0 : etm::upper_index(cp, n_groups, c);
        uint8_t st = ETM_CAND_COUNT;
        int32_t cnt = 0;
        if (m >= 2) {
            const int64_t g0 = gs[g];
            int64_t i, j;
            etm::lex_pair_decode(c - cp[g], gs[g + 1] - g0, &i, &j);
            C.classify(g0, i, j, &st, &cnt);
        }
        J.status[c] = st;
        J.counts[c] = cnt;
        J.gid[c] = (int32_t)g;
        n_count += st == ETM_CAND_COUNT;
        n_skip += st == ETM_CAND_SKIP;
    }
    n_count = etm::block_sum<256>(n_count, lds);
    n_skip = etm::block_sum<256>(n_skip, lds);
    if (threadIdx.x == 0) {
        if (n_count) atomicAdd((unsigned long long*)&J.tally[0], (unsigned long long)n_count);
        if (n_skip) atomicAdd((unsigned long long*)&J.tally[2], (unsigned long long)n_skip);
    }
}

// A workgroup per job: the round's survivors (count >= min_count, not skipped) in lexicographic order into the
// job's level store, rows then counts, recorded as the job's next level.
__global__ __launch_bounds__(kOne) void k_batch_compact(Job* __restrict__ jobs) {
    __shared__ int32_t lds[kOneWaves];
    Job& J = jobs[blockIdx.x];
    if (J.state != etm::batch::kActive) return;
    const int b = J.buf;
    const int32_t k = J.k, m = k - 1;
    const int64_t n_cands = J.tot[b][1], n_groups = J.tot[b][0];
    const int64_t rows_off = J.out_used, counts_off = rows_off + n_cands * k;
    if (counts_off + n_cands > J.out_cap || J.n_levels >= etm::batch::kMaxLevels) {
        if (threadIdx.x == 0) J.state = etm::batch::kOverflow;
        return;
    }
    const int64_t n = compact_single_body(J.status, J.counts, J.min_count, J.gid, J.prev_rows, m, J.gs[b], J.cp[b],
                                          n_groups, n_cands, J.out + rows_off, J.out + counts_off, nullptr, lds);
    if (threadIdx.x == 0) {
        etm::batch::Level& L = J.lv[J.n_levels];
        L.n = n;
        L.rows_off = rows_off;
        L.counts_off = counts_off;
        L.cands = n_cands;
        L.counted = J.tally[0];
        L.skipped = J.tally[2];
        L.bytes_alg = ((int64_t)(m - 1) * J.tot[b][5] + J.tot[b][6]) * J.n_words * 8 + 4 * n_cands;
        L.t_end = 0;
        J.new_n = n;
        J.out_used = counts_off + n;
        J.n_levels += 1;
    }
}

// A workgroup per job. With `init`: the prefix groups of the job's previous level (level 1 before the first
// round). Otherwise the groups of the level the round just recorded, after which the job advances to it, or
// ends: done when no level can follow (no survivors, max_len reached, fewer rows than the next size, or no
// candidates), overflow when the next level's candidates pass kCapCands.
__global__ __launch_bounds__(kOne) void k_batch_groups(Job* __restrict__ jobs, etm::batch::Queue* __restrict__ q,
                                                       int init) {
    __shared__ int32_t lds[5 * kOneWaves];
    __shared__ int64_t red[kOneWaves + 1];
    Job& J = jobs[blockIdx.x];
    if (J.state != etm::batch::kActive) return;
    const int32_t k = J.k;
    const int32_t* rows;
    const int32_t* counts = nullptr;
    int64_t n;
    int32_t m;
    int nb;
    if (init) {
        rows = J.prev_rows;
        n = J.prev_n;
        m = k - 1;
        nb = J.buf;
    } else {
        const etm::batch::Level& L = J.lv[J.n_levels - 1];
        rows = J.out + L.rows_off;
        counts = J.out + L.counts_off;
        n = L.n;
        m = k;
        nb = J.buf ^ 1;
        if (n == 0 || k + 1 > J.max_len || n < k + 1) {
            if (threadIdx.x == 0) {
                J.lv[J.n_levels - 1].t_end = wall_clock64();
                J.state = etm::batch::kDone;
            }
            return;
        }
    }
    groups_single_body(rows, n, m, J.small_max, J.gs[nb], J.cp[nb], J.ct[nb], J.list[nb], J.tot[nb], lds, red);
    if (threadIdx.x == 0) {
        const int64_t cands = J.tot[nb][1];
        const int32_t st = cands == 0 ? etm::batch::kDone
                                      : (cands > etm::batch::kCapCands ? etm::batch::kOverflow : etm::batch::kActive);
        if (init) {
            J.t_init = wall_clock64();
        } else {
            J.lv[J.n_levels - 1].t_end = wall_clock64();
            if (st == etm::batch::kActive) atomicAdd((unsigned long long*)&q->active_after, 1ull);
            if (st == etm::batch::kActive) {
                J.prev_rows = rows;
                J.prev_counts = counts;
                J.prev_n = n;
                J.buf = nb;
                J.k = k + 1;
            }
        }
        J.state = st;
    }
}

}  // namespace

namespace etm {
namespace detail {

hipError_t launch_groups_single(const int32_t* rows, const int64_t* d_n, int64_t n_cap, int32_t m, int32_t small_max,
                                int64_t* group_start, int64_t* cum_pairs, int64_t* cum_tiles, int32_t* small_list,
                                int64_t* totals, hipStream_t s) {
    k_groups_single<<<1, kOne, 0, s>>>(rows, d_n, n_cap, m, small_max, group_start, cum_pairs, cum_tiles,
                                       small_list, totals);
    return hipGetLastError();
}

hipError_t launch_compact_single(const int32_t* rows, int32_t m, const int64_t* group_start, const int64_t* cum_pairs,
                                 int64_t n_groups, int64_t n_cands, const uint8_t* status, const int32_t* counts,
                                 const int32_t* gid, int64_t min_count, int32_t* out_rows, int32_t* out_counts,
                                 int64_t* out_cand, int64_t* d_out_n, hipStream_t s) {
    k_compact_single<<<1, kOne, 0, s>>>(status, counts, min_count, gid, rows, m, group_start, cum_pairs, n_groups,
                                        n_cands, out_rows, out_counts, out_cand, d_out_n);
    return hipGetLastError();
}

hipError_t launch_select_single(const int32_t* rows, const int64_t* d_n, int64_t n_cap, int32_t m,
                                const int32_t* counts, const uint8_t* flag, int32_t* out_rows, int32_t* out_counts,
                                int64_t* d_out_n, hipStream_t s) {
    k_select_single<<<1, kOne, 0, s>>>(rows, d_n, n_cap, m, counts, flag, out_rows, out_counts, d_out_n);
    return hipGetLastError();
}

hipError_t launch_hash(const int32_t* rows, int64_t n_cap, const int64_t* d_n, int32_t m, int32_t* table,
                       int64_t table_size, hipStream_t s) {
    k_fill_table<<<grid_for(table_size), 256, 0, s>>>(table, table_size);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || n_cap == 0) return e;
    k_hash_insert<<<grid_for(n_cap), 256, 0, s>>>(rows, n_cap, m, table, table_size - 1, d_n);
    return hipGetLastError();
}

hipError_t launch_free(const int32_t* rows, int64_t n_cap, const int64_t* d_n, int32_t k, const int32_t* counts,
                       const int32_t* prev_rows, const int32_t* prev_counts, const int32_t* prev_table,
                       int64_t table_size, uint8_t* free_out, hipStream_t s, bool hinted) {
    if (n_cap == 0) return hipSuccess;
    k_free<<<grid_for(n_cap), 256, 0, s>>>(rows, n_cap, k, counts, prev_rows, prev_counts, prev_table,
                                           table_size - 1, free_out, d_n, hinted);
    return hipGetLastError();
}

hipError_t level_compact_hinted(const int32_t* rows, int32_t m, const int64_t* group_start, const int64_t* cum_pairs,
                                int64_t n_groups, int64_t n_cands, const uint8_t* status, const int32_t* counts,
                                const int32_t* gid, int64_t min_count, int32_t* out_rows, int32_t* out_counts,
                                uint8_t* free_hint, const int64_t* ws, hipStream_t s) {
    Survives f{status, counts, min_count};
    WriteSurvivor w{rows, m, group_start, cum_pairs, n_groups, counts, out_rows, out_counts, nullptr, gid, status,
                    free_hint};
    return etm::scan_finish(f, w, n_cands, ws, s);
}

hipError_t launch_fill_u8(uint8_t* p, int64_t n, uint8_t v, hipStream_t s) {
    if (n == 0) return hipSuccess;
    k_fill_u8<<<grid_for(n), 256, 0, s>>>(p, n, v);
    return hipGetLastError();
}

hipError_t launch_batch_hash(etm::batch::Job* jobs, int n_jobs, hipStream_t s) {
    k_batch_hash_fill<<<dim3(kBatchFillBlocks, n_jobs), 256, 0, s>>>(jobs);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_batch_hash_insert<<<dim3(kBatchInsertBlocks, n_jobs), 256, 0, s>>>(jobs);
    return hipGetLastError();
}

hipError_t launch_batch_classify(etm::batch::Job* jobs, int n_jobs, hipStream_t s) {
    k_batch_classify<<<dim3(kBatchClassifyBlocks, n_jobs), 256, 0, s>>>(jobs);
    return hipGetLastError();
}

hipError_t launch_batch_compact(etm::batch::Job* jobs, int n_jobs, hipStream_t s) {
    k_batch_compact<<<n_jobs, kOne, 0, s>>>(jobs);
    return hipGetLastError();
}

hipError_t launch_batch_groups(etm::batch::Job* jobs, int n_jobs, etm::batch::Queue* q, bool init, hipStream_t s) {
    k_batch_groups<<<n_jobs, kOne, 0, s>>>(jobs, q, init ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_zero_i64(int64_t* p, int n, hipStream_t s) {
    k_zero_i64<<<std::max(1, std::min((n + 255) / 256, 1024)), 256, 0, s>>>(p, n);
    return hipGetLastError();
}

hipError_t launch_live_bytes(const uint8_t* status, const int32_t* gid, const int64_t* group_start,
                             const int64_t* cum_pairs, int64_t n_groups, int64_t n_cands, int64_t n_rows,
                             uint64_t* row_bits, uint64_t* group_bits, int64_t* out, hipStream_t s) {
    const int64_t rw = (n_rows + 63) / 64, gw = (n_groups + 63) / 64;
    if (rw + gw > 0x7FFFFFFF) return hipErrorInvalidValue;
    hipError_t e = launch_zero_i64((int64_t*)row_bits, (int)rw, s);
    if (e == hipSuccess) e = launch_zero_i64((int64_t*)group_bits, (int)gw, s);
    if (e != hipSuccess || n_cands == 0) return e;
    k_live_mark<<<grid_for(n_cands), 256, 0, s>>>(status, gid, group_start, cum_pairs, n_groups, n_cands,
                                                  (unsigned long long*)row_bits, (unsigned long long*)group_bits,
                                                  out);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_live_sum<<<grid_for(std::max(rw, gw), 1024), 256, 0, s>>>(row_bits, rw, group_bits, gw, out);
    return hipGetLastError();
}

}  // namespace detail
}  // namespace etm


namespace {
// The class passes' partial rows follow the scans' workspace.
size_t group_part_offset(int64_t n) { return (etm::scan_ws_bytes(n + 1) + 255) / 256 * 256; }
}  // namespace

extern "C" size_t etm_level_groups_ws(int64_t n) {
    return group_part_offset(n) + (size_t)kGroupBlocks * kPartCols * sizeof(int64_t);
}

extern "C" int etm_level_groups(const int32_t* d_rows, int64_t n, int32_t m, int32_t small_max,
                                int64_t* d_group_start, int64_t* d_cum_pairs, int64_t* d_cum_tiles,
                                int32_t* d_small_list, int64_t* d_totals, void* d_ws, size_t ws_bytes,
                                void* stream) {
    ETM_ARG(d_rows && d_group_start && d_cum_pairs && d_cum_tiles && d_small_list && d_totals && d_ws,
            "etm_level_groups: null pointer");
    ETM_ARG(n >= 1 && m >= 1 && m <= 62, "etm_level_groups: need n >= 1 and 1 <= m <= 62 (n=%lld, m=%d)",
            (long long)n, m);
    ETM_ARG(n < 0x7FFFFFFF, "etm_level_groups: %lld rows exceed int32 row ids", (long long)n);
    ETM_ARG(small_max == 2 || small_max == 4 || small_max == 8 || small_max == 16,
            "etm_level_groups: small_max must be 2, 4, 8 or 16, got %d", small_max);
    ETM_ARG(ws_bytes >= etm_level_groups_ws(n), "etm_level_groups: workspace %zu < %zu bytes", ws_bytes,
            etm_level_groups_ws(n));
    hipStream_t s = ETM_STREAM(stream);
    int64_t* ws = (int64_t*)d_ws;
    k_zero_i64<<<1, 64, 0, s>>>(d_totals, 24);
    ETM_LAUNCHED("k_zero_i64");
    IsGroupStart is_start{d_rows, m};
    ETM_HIP(etm::scan_prepare(is_start, n, ws, nullptr, s, SetGroupEnd{d_group_start, n, d_totals}));
    ETM_HIP(etm::scan_finish(is_start, WriteGroupStart{d_group_start}, n, ws, s));
    GroupPairs pairs{d_group_start, d_totals};
    ETM_HIP(etm::scan_prepare(pairs, n + 1, ws, &d_totals[1], s));
    ETM_HIP(etm::scan_finish(pairs, WritePrefix{d_cum_pairs}, n + 1, ws, s));
    GroupTiles tiles{d_group_start, d_totals, small_max};
    ETM_HIP(etm::scan_prepare(tiles, n + 1, ws, &d_totals[2], s));
    ETM_HIP(etm::scan_finish(tiles, WritePrefix{d_cum_tiles}, n + 1, ws, s));
    // Both passes take the same ranges: one launch size for both, from the row count (groups <= rows).
    const int blocks = grid_for(n, kGroupBlocks);
    int64_t* part = (int64_t*)((char*)d_ws + group_part_offset(n));
    k_group_classes<<<blocks, 256, 0, s>>>(d_group_start, n, small_max, d_totals, part);
    ETM_LAUNCHED("k_group_classes");
    k_group_place<<<blocks, 256, 0, s>>>(d_group_start, n, small_max, d_small_list, d_totals, part);
    ETM_LAUNCHED("k_group_place");
    return ETM_OK;
}

extern "C" int etm_level_hash(const int32_t* d_rows, int64_t n, int32_t m, int32_t* d_table, int64_t table_size,
                              void* stream) {
    ETM_ARG(d_rows && d_table, "etm_level_hash: null pointer");
    ETM_ARG(m >= 1 && m <= 62, "etm_level_hash: 1 <= m <= 62 required, got %d", m);
    ETM_ARG(is_pow2(table_size) && table_size >= 2 * n, "etm_level_hash: table_size %lld must be a power of two >= 2n",
            (long long)table_size);
    ETM_ARG(n < 0x7FFFFFFF, "etm_level_hash: %lld rows exceed int32 row ids", (long long)n);
    hipStream_t s = ETM_STREAM(stream);
    k_fill_table<<<grid_for(table_size), 256, 0, s>>>(d_table, table_size);
    ETM_LAUNCHED("k_fill_table");
    if (n > 0) {
        k_hash_insert<<<grid_for(n), 256, 0, s>>>(d_rows, n, m, d_table, table_size - 1);
        ETM_LAUNCHED("k_hash_insert");
    }
    return ETM_OK;
}

namespace etm {
namespace detail {

int64_t pair_index_bytes(int64_t n_cols) {
    const int64_t words = (n_cols + 63) / 64;
    return n_cols * words * 12;
}

hipError_t launch_pair_index(const int32_t* rows2, int64_t n2, int64_t n_cols, uint64_t* bits, int32_t* rank,
                             hipStream_t s) {
    const int64_t words = (n_cols + 63) / 64;
    hipError_t e = hipMemsetAsync(bits, 0, (size_t)(n_cols * words) * 8, s);
    if (e != hipSuccess) return e;
    if (n2 > 0) k_pair_index_bits<<<grid_for(n2), 256, 0, s>>>(rows2, n2, words, bits);
    e = hipGetLastError();
    if (e != hipSuccess || n_cols == 0) return e;
    k_pair_index_rank<<<grid_for(n_cols * 64), 256, 0, s>>>(rows2, n2, n_cols, words, bits, rank);
    return hipGetLastError();
}

}  // namespace detail
}  // namespace etm

extern "C" int etm_level_classify(const int32_t* d_rows, int64_t n, int32_t m, const int32_t* d_prev_counts,
                                  const uint8_t* d_prev_free, const int32_t* d_table, int64_t table_size,
                                  const int64_t* d_group_start, const int64_t* d_cum_pairs, int64_t n_groups,
                                  int64_t n_cands, const int32_t* d_group_list, const int64_t* h_class_counts,
                                  int32_t mode, uint8_t* d_status, int32_t* d_counts, int32_t* d_gid,
                                  int64_t* d_tally, void* stream) {
    return etm::detail::level_classify(d_rows, n, m, d_prev_counts, d_prev_free, d_table, table_size, d_group_start,
                                       d_cum_pairs, n_groups, n_cands, d_group_list, h_class_counts, mode, d_status,
                                       d_counts, d_gid, d_tally, nullptr, nullptr, 0, stream);
}

int etm::detail::level_classify(const int32_t* d_rows, int64_t n, int32_t m, const int32_t* d_prev_counts,
                                const uint8_t* d_prev_free, const int32_t* d_table, int64_t table_size,
                                const int64_t* d_group_start, const int64_t* d_cum_pairs, int64_t n_groups,
                                int64_t n_cands, const int32_t* d_group_list, const int64_t* h_class_counts,
                                int32_t mode, uint8_t* d_status, int32_t* d_counts, int32_t* d_gid,
                                int64_t* d_tally, const uint64_t* d_pair_bits, const int32_t* d_pair_rank,
                                int64_t pair_words, void* stream) {
    ETM_ARG(d_rows && d_table && d_group_start && d_cum_pairs && d_status && d_counts,
            "etm_level_classify: null pointer");
    ETM_ARG(m >= 2 && m <= 62, "etm_level_classify: 2 <= m <= 62 required, got %d", m);
    ETM_ARG(is_pow2(table_size), "etm_level_classify: table_size must be a power of two");
    ETM_ARG(!(mode & ETM_SUBSET_INFER) || (d_prev_free && d_prev_counts),
            "etm_level_classify: inference needs the previous level's counts and free flags");
    ETM_ARG(n_groups >= 1 || n_cands == 0, "etm_level_classify: no groups");
    (void)n;
    if (n_cands == 0) return ETM_OK;
    hipStream_t s = ETM_STREAM(stream);
    if (d_group_list && h_class_counts) {
        int64_t small = 0;
        for (int c = 0; c < 4; ++c) small += h_class_counts[c];
        const int64_t big = h_class_counts[4];
        Classifier C{d_rows, m, d_prev_counts, d_prev_free, d_table, table_size - 1, mode,
                     m == 2 ? d_pair_bits : nullptr, d_pair_rank, pair_words};
        if (small > 0) {
            PackedClasses pc{{h_class_counts[0], h_class_counts[1], h_class_counts[2], h_class_counts[3]},
                             std::max(1, m - 1)};
            const int64_t waves = packed_waves(pc);
            ETM_REQUIRE((waves + 3) / 4 <= 0x7FFFFFFF, ETM_EOVERFLOW, "etm_level_classify: grid too large");
            k_classify_packed<<<(unsigned)((waves + 3) / 4), 256, 0, s>>>(C, d_group_start, d_cum_pairs, d_group_list,
                                                                          pc, d_status, d_counts, d_gid, d_tally);
            ETM_LAUNCHED("k_classify_packed");
        }
        if (big > 0) {
            // Pair chunks per group: enough for the largest group (h_class_counts[5], when given, is its size).
            const int64_t max_s = h_class_counts[5] > 0 ? h_class_counts[5] : 0;
            const int64_t max_pairs = max_s > 0 ? max_s * (max_s - 1) / 2 : n_cands;
            const int64_t gx = std::max<int64_t>(1, (max_pairs + kBigPairs - 1) / kBigPairs);
            ETM_REQUIRE(gx <= 0x7FFFFFFF && big <= 65535, ETM_EOVERFLOW, "etm_level_classify: grid too large");
            k_classify_big<<<dim3((unsigned)gx, (unsigned)big), 256, 0, s>>>(C, d_group_start, d_cum_pairs,
                                                                             d_group_list + small, d_status, d_counts,
                                                                             d_gid, d_tally);
            ETM_LAUNCHED("k_classify_big");
        }
        return ETM_OK;
    }
    k_classify<<<grid_for(n_cands, 65536), 256, 0, s>>>(d_rows, m, d_prev_counts, d_prev_free, d_table,
                                                        table_size - 1, d_group_start, d_cum_pairs, n_groups,
                                                        n_cands, mode, d_status, d_counts, d_tally);
    ETM_LAUNCHED("k_classify");
    return ETM_OK;
}

extern "C" int etm_level_deduce(const int32_t* d_rows, int32_t m, const int32_t* d_prev_counts, const int32_t* d_table,
                                int64_t table_size, const int32_t* d_rows2, const int32_t* d_counts2,
                                const int32_t* d_table2, int64_t table_size2, const int32_t* d_rows3,
                                const int32_t* d_counts3, const int32_t* d_table3, int64_t table_size3,
                                int64_t n_total, int64_t min_count, const int64_t* d_group_start,
                                const int64_t* d_cum_pairs, int64_t n_groups, int64_t n_cands, const int32_t* d_gid,
                                uint8_t* d_status, int32_t* d_counts, int64_t* d_tally, void* stream) {
    const int k = m + 1;
    ETM_ARG(k >= 3 && k <= kBoundsMaxK, "etm_level_deduce: candidates of 3 to %d items, got %d", kBoundsMaxK, k);
    ETM_ARG(d_rows && d_prev_counts && d_table && d_rows2 && d_counts2 && d_table2 && d_group_start &&
                d_cum_pairs && d_status && d_counts,
            "etm_level_deduce: null pointer");
    ETM_ARG(k == 3 || (d_rows3 && d_counts3 && d_table3), "etm_level_deduce: k > 3 needs level k-3");
    ETM_ARG(is_pow2(table_size) && is_pow2(table_size2) && (k == 3 || is_pow2(table_size3)),
            "etm_level_deduce: table sizes must be powers of two");
    ETM_ARG(n_groups >= 1 || n_cands == 0, "etm_level_deduce: no groups");
    if (n_cands == 0) return ETM_OK;
    hipStream_t s = ETM_STREAM(stream);
    const int need = std::max(k, std::max(k * (k - 1) / 2, k * (k - 1) * (k - 2) / 6));
    int L = 1;
    while (L < need && L < 64) L <<= 1;
    const int per_block = 256 / L;
    const int stride8 = k + (2 * k + k * (k - 1) / 2 + 1) / 2;
    const size_t dyn = (size_t)per_block * stride8 * 8;
    Classifier C{d_rows, m, d_prev_counts, nullptr, d_table, table_size - 1, ETM_SUBSET_INFER};
    DeepLevel d2{d_rows2, d_counts2, d_table2, table_size2 - 1};
    DeepLevel d3{d_rows3, d_counts3, d_table3, k == 3 ? 0 : table_size3 - 1};
    const int64_t blocks = std::min<int64_t>((n_cands + 255) / 256, 8192);
    k_deduce<<<(unsigned)blocks, 256, dyn, s>>>(C, d2, d3, n_total, min_count, d_group_start, d_cum_pairs, n_groups,
                                                d_gid, n_cands, L, stride8, d_status, d_counts, d_tally);
    ETM_LAUNCHED("k_deduce");
    return ETM_OK;
}

extern "C" size_t etm_level_compact_ws(int64_t n_cands) { return etm::scan_ws_bytes(n_cands); }

extern "C" int etm_level_compact(const int32_t* d_rows, int32_t m, const int64_t* d_group_start,
                                 const int64_t* d_cum_pairs, int64_t n_groups, int64_t n_cands,
                                 const uint8_t* d_status, const int32_t* d_counts, const int32_t* d_gid,
                                 int64_t min_count,
                                 int32_t* d_out_rows, int32_t* d_out_counts, int64_t* d_out_cand, int64_t* d_out_n,
                                 void* d_ws, size_t ws_bytes, void* stream) {
    ETM_ARG(d_rows && d_group_start && d_cum_pairs && d_counts && d_out_n && d_ws,
            "etm_level_compact: null pointer");
    ETM_ARG(m >= 1 && m < 62, "etm_level_compact: 1 <= m < 62 required, got %d", m);
    ETM_ARG(ws_bytes >= etm_level_compact_ws(n_cands), "etm_level_compact: workspace %zu < %zu bytes", ws_bytes,
            etm_level_compact_ws(n_cands));
    hipStream_t s = ETM_STREAM(stream);
    Survives f{d_status, d_counts, min_count};
    if (d_out_rows == nullptr) {
        ETM_HIP(etm::scan_prepare(f, n_cands, (int64_t*)d_ws, d_out_n, s));
        return ETM_OK;
    }
    ETM_ARG(d_out_counts != nullptr, "etm_level_compact: null output counts");
    WriteSurvivor w{d_rows,       m,        d_group_start, d_cum_pairs, n_groups, d_counts,
                    d_out_rows,   d_out_counts, d_out_cand, d_gid};
    ETM_HIP(etm::scan_finish(f, w, n_cands, (const int64_t*)d_ws, s));
    return ETM_OK;
}

// ── compaction without allocation ───────────────────────────────────────
//
// etm_level_compact_noalloc keeps the scan's block sums in `scratch`, a file-scope device array: each device
// that runs the library holds its own copy until the process ends, at its own address. Host code never passes
// the array's name to a kernel (on the host it is a stand-in, not a device address): the first call on a device
// resolves that device's address with hipGetSymbolAddress, and every later use goes through it.

namespace {

constexpr int64_t kNoAllocScratchWords = (1 << 20) / 8;  // 1 MiB of int64 block sums
__device__ __align__(256) int64_t scratch[kNoAllocScratchWords];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
int64_t* g_scratch_ptr[kMaxDevices] = {};

// Survivor writer that drops every survivor at or past the output capacity.
struct WriteSurvivorCapped {
    WriteSurvivor w;
    int64_t cap;
    __device__ void operator()(int64_t c, int64_t v, int64_t p) const {
        if (v && p < cap) w(c, v, p);
    }
};

}  // namespace

extern "C" int64_t etm_level_compact_noalloc_max(void) {
    return (kNoAllocScratchWords - 1) * etm::kScanItems;
}

extern "C" int etm_level_compact_noalloc(const int32_t* d_rows, int32_t m, const int64_t* d_group_start,
                                         const int64_t* d_cum_pairs, int64_t n_groups, int64_t n_cands,
                                         const uint8_t* d_status, const int32_t* d_counts, const int32_t* d_gid,
                                         int64_t min_count, int32_t* d_out_rows, int32_t* d_out_counts,
                                         int64_t* d_out_cand, int64_t out_cap, int64_t* d_out_n, void* stream) {
    ETM_ARG(d_rows && d_group_start && d_cum_pairs && d_counts && d_out_n,
            "etm_level_compact_noalloc: null pointer");
    ETM_ARG(out_cap >= 0 && (out_cap == 0 || (d_out_rows && d_out_counts)),
            "etm_level_compact_noalloc: outputs must hold out_cap >= 0 survivors");
    ETM_ARG(m >= 1 && m < 62, "etm_level_compact_noalloc: 1 <= m < 62 required, got %d", m);
    ETM_ARG(n_cands >= 0 && n_cands <= etm_level_compact_noalloc_max(),
            "etm_level_compact_noalloc: %lld candidates exceed the static scratch's limit of %lld",
            (long long)n_cands, (long long)etm_level_compact_noalloc_max());
    int dev = 0;
    ETM_HIP(hipGetDevice(&dev));
    ETM_ARG(dev >= 0 && dev < kMaxDevices, "etm_level_compact_noalloc: device %d out of range", dev);
    int64_t* ptr = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (g_scratch_ptr[dev] == nullptr) {
            void* p = nullptr;
            ETM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
            g_scratch_ptr[dev] = (int64_t*)p;
        }
        ptr = g_scratch_ptr[dev];
    }
    hipStream_t s = ETM_STREAM(stream);
    // Every block sum the scan reads is written first by this call's reduce (and the total by its block scan),
    // so nothing left by an earlier call is read.
    Survives f{d_status, d_counts, min_count};
    ETM_HIP(etm::scan_prepare(f, n_cands, ptr, d_out_n, s));
    if (out_cap == 0) return ETM_OK;
    WriteSurvivorCapped w{WriteSurvivor{d_rows, m, d_group_start, d_cum_pairs, n_groups, d_counts, d_out_rows,
                                        d_out_counts, d_out_cand, d_gid},
                          out_cap};
    ETM_HIP(etm::scan_finish(f, w, n_cands, ptr, s));
    return ETM_OK;
}

extern "C" int etm_level_free(const int32_t* d_rows, int64_t n, int32_t k, const int32_t* d_counts,
                              const int32_t* d_prev_rows, const int32_t* d_prev_counts, const int32_t* d_prev_table,
                              int64_t table_size, uint8_t* d_free, void* stream) {
    ETM_ARG(d_rows && d_counts && d_prev_rows && d_prev_counts && d_prev_table && d_free,
            "etm_level_free: null pointer");
    ETM_ARG(k >= 2 && k <= 63, "etm_level_free: 2 <= k <= 63 required, got %d", k);
    ETM_ARG(is_pow2(table_size), "etm_level_free: table_size must be a power of two");
    if (n == 0) return ETM_OK;
    k_free<<<grid_for(n), 256, 0, ETM_STREAM(stream)>>>(d_rows, n, k, d_counts, d_prev_rows, d_prev_counts,
                                                        d_prev_table, table_size - 1, d_free, nullptr, false);
    ETM_LAUNCHED("k_free");
    return ETM_OK;
}

extern "C" int etm_select_rows(const int32_t* d_rows, int64_t n, int32_t m, const int32_t* d_counts,
                               const uint8_t* d_flag, int32_t* d_out_rows, int32_t* d_out_counts, int64_t* d_out_n,
                               void* d_ws, size_t ws_bytes, void* stream) {
    ETM_ARG(d_rows && d_counts && d_flag && d_out_n && d_ws, "etm_select_rows: null pointer");
    ETM_ARG(m >= 1 && m <= 63, "etm_select_rows: 1 <= m <= 63 required");
    ETM_ARG(ws_bytes >= etm_level_compact_ws(n), "etm_select_rows: workspace too small");
    hipStream_t s = ETM_STREAM(stream);
    Flagged f{d_flag};
    if (d_out_rows == nullptr) {
        ETM_HIP(etm::scan_prepare(f, n, (int64_t*)d_ws, d_out_n, s));
        return ETM_OK;
    }
    ETM_ARG(d_out_counts != nullptr, "etm_select_rows: null output counts");
    ETM_HIP(etm::scan_finish(f, WriteSelected{d_rows, m, d_counts, d_out_rows, d_out_counts}, n,
                             (const int64_t*)d_ws, s));
    return ETM_OK;
}
