// AND + popcount counting of a level's candidates.
//
// Two kernel families share one work description (prefix groups of a
// lexsorted level, see etm.h):
//
// * k_count_small<S>: groups of 2..S suffixes. One wave64 per (group, word
//   chunk); lane l reads word w0 + l + 64t of every column (or, where every
//   column is 16-byte aligned, words w0 + 2l + 128t and the next one as one
//   16-byte load), so each load instruction fetches one contiguous run per
//   column. Per word the wave ANDs the prefix columns, ANDs each suffix
//   column with that, and accumulates popcounts of the counted suffix pairs
//   in registers. Partial counts are reduced across the wave with shuffles and
//   added to the count array with one integer atomic per pair and chunk
//   (exact and order-independent). Reads (prefix + suffixes) x 8 bytes per
//   word, mostly from L2: bound by load issue there, by HBM otherwise.
//
// * k_count_tiled: groups with more suffixes (and the K=2 pair space, one
//   group with an empty prefix). One workgroup per (tile-pair, word chunk) of
//   64 x 64 suffixes. Dense mode: per 32-word step it stages the prefix AND
//   and the two tiles' suffix columns (already ANDed with the prefix) in LDS,
//   transposed to [word][suffix] so the compute loop reads are conflict-free;
//   each of 256 threads accumulates a 4 x 4 register block of pairs
//   (VALU-bound: 4 instructions, 2 AND + 2 BCNT, per candidate-word). Listed
//   mode (at most kListMax live pairs): only the live pairs' columns are
//   staged, compacted into LDS slots, up to 256 words per step, and a thread
//   counts up to four listed pairs.
//
// A tile-pair or group none of whose candidates has status ETM_CAND_COUNT
// returns before reading any bitvector.
//
// Further sections: the live-unit schedule (k_live_units lists the groups and
// tile-pairs holding a counted candidate; k_count_small_live,
// k_count_tiled_live and the register-only k_count_sparse_live count them),
// the group-listed kernels for tiled groups of few counted pairs (k_gl_scan,
// k_count_glist), the K=2 pair space on the fp4 matrix cores (k_gram_mfma),
// arbitrary itemsets (k_count_itemsets), and the queue-fed counting kernels of
// batched jobs (k_batch_plan, k_batch_count_small, k_batch_count_tiled).
#include <algorithm>
#include <cmath>

#include "common.h"
#include "internal.h"

namespace {

// Work items (unit x word chunk) the launches aim for, and the shortest chunk (words). Launching every unit:
// about 2048 items per small class and per tiled level (measured on MI355X over 2K-64K targets and 256-4096-word
// floors; with 16-byte loads 2048 beat 4096 and 8192 on deep_k_1m). Launching the live units only: 16384 and
// 32768 items with chunks of at least 1024 and 512 words (deep_sparse_20m's levels K >= 3: 23.1 -> 13.4 ms
// against the 2048 targets; shorter floors cost deep_k_1m's levels 10%).
constexpr int64_t kSmallTarget = 2048, kTiledTarget = 2048;
constexpr int64_t kLiveSmallTarget = 16384, kLiveTiledTarget = 32768;
constexpr int64_t kLiveSmallMin = 1024, kLiveTiledMin = 512;

constexpr int kT = 64;   // suffixes per tile
constexpr int kKW = 32;  // words per LDS step
constexpr int kXcds = 8;  // MI355X: 8 XCDs of 32 CUs, each with its own L2
constexpr int kListMax = 1024;  // live pairs from which a tile-pair is counted densely (4096 slots)
constexpr int kListWords = 256;  // most words per step of the listed mode

// Start of each small size class in the class-ordered group list (start[4] = the list's length).
struct ClassBounds {
    int64_t start[5];
};

// A sparse live tile-pair: at most kSparseMax counted pairs (ta * 64 + a[p], tb * 64 + b[p]) of its group, with at
// most kSparseRatio pairs per column in them. One wave counts it from registers (k_count_sparse_live): no LDS
// staging, no barriers, each column read straight from the caches per pair. Measured on deep_sparse_20m's levels
// K >= 3: 32 pairs and 4 per column 10.9 ms; 32 and 2: 11.3 ms; 16 and 2: 11.4 ms; 64 and 4: 11.9 ms; 64 and 8:
// 12.7 ms (12.7 ms without the sparse kernel).
constexpr int kSparseMax = 32;
constexpr int kSparseRatio = 4;
struct SparseTP {
    int32_t tp;
    int32_t n;
    uint8_t a[kSparseMax];
    uint8_t b[kSparseMax];
};

// The (word chunk, item) unit of a workgroup or wave: `xcd` = workgroup index mod 8 (the XCD the hardware
// dispatches it to, round-robin), `local` = its index among that XCD's units. With at least 8 chunks, XCD x takes
// chunks x, x+8, ..., and every item of a chunk in turn, so the waves of one XCD share chunks in its L2; with
// fewer chunks, XCD x takes every 8th unit of the chunk-major order, so all 8 XCDs stay busy. Returns false for
// a unit past the end.
__device__ __forceinline__ bool xcd_unit(int64_t xcd, int64_t local, int64_t n_items, int64_t n_chunks,
                                         int64_t* ch, int64_t* item) {
    if (n_chunks >= kXcds) {
        *ch = xcd + kXcds * (local / n_items);
        *item = local % n_items;
    } else {
        const int64_t u = xcd + kXcds * local;
        *ch = u / n_items;
        *item = u % n_items;
    }
    return *ch < n_chunks;
}

__host__ __device__ __forceinline__ int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Word chunk (a multiple of `align` words) giving about `target` work items over `units` independent units.
// The chunk count rounds up to a multiple of the 8 XCDs where the columns allow it: xcd_unit then keeps each
// XCD on its own chunks, whose words its L2 holds; with fewer chunks every XCD reads every word.
__host__ __device__ inline int64_t pick_chunk(int64_t n_words, int64_t units, int64_t target, int64_t align,
                                              int64_t min_words) {
    if (n_words <= 0) return align;
    int64_t chunks = units > 0 ? ceil_div(target, units) : 1;
    const int64_t floor_words = min_words > align ? min_words : align;
    int64_t max_chunks = n_words / floor_words > 1 ? n_words / floor_words : 1;
    chunks = chunks < max_chunks ? chunks : max_chunks;
    if (chunks < 1) chunks = 1;
    const int64_t aligned = ceil_div(chunks, kXcds) * kXcds;
    if (aligned <= max_chunks) chunks = aligned;
    return ceil_div(ceil_div(n_words, chunks), align) * align;
}

// A wave-uniform 64-bit value (computed by every lane alike) moved to scalar registers: loop bounds derived from
// it stay uniform, which keeps the counting loops' register allocation that of kernel-argument bounds.
__device__ __forceinline__ int64_t uniform64(int64_t x) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)(uint64_t)x);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)x >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

// pick_chunk on the device, for the kernels whose unit count lives on the device.
__device__ __forceinline__ void live_chunks(int64_t n_words, int64_t units, int64_t target, int64_t align,
                                            int64_t min_words, int64_t* chunk, int64_t* n_chunks) {
    *chunk = uniform64(pick_chunk(n_words, units, target, align, min_words));
    *n_chunks = uniform64(ceil_div(n_words, *chunk));
}

// Units per XCD of the mapping above (the grid is 8 x this, rounded up to whole workgroups).
int64_t xcd_units(int64_t n_items, int64_t n_chunks) {
    const int64_t per = (n_chunks + kXcds - 1) / kXcds;
    return n_chunks >= kXcds ? n_items * per : (n_items * n_chunks + kXcds - 1) / kXcds;
}

// acc + popcount(x) as two v_bcnt_u32_b32 with accumulate (one per 32-bit half).
__device__ __forceinline__ int popc_acc(uint64_t x, int acc) {
    return __popc((uint32_t)(x >> 32)) + (__popc((uint32_t)x) + acc);
}

__device__ __forceinline__ int64_t pair_slot(int64_t i, int64_t j, int64_t s, int32_t layout) {
    return layout == ETM_LAYOUT_TRI ? j * (j - 1) / 2 + i : etm::lex_pair_index(i, j, s);
}

// Words a lane handles per iteration in k_count_small<S>: more independent loads in flight for the narrow
// classes, fewer registers for the wide ones.
template <int S>
struct SmallUnroll {
    static constexpr int U = S <= 4 ? 4 : (S <= 8 ? 2 : 1);
};

// One (group, word chunk) unit of a group of 2..S suffixes, counted by the calling wave.
template <int S, bool SKIP, bool VEC>
__device__ __forceinline__ void small_unit(const uint64_t* __restrict__ bv, int64_t pitch, int64_t n_words,
                                           int64_t chunk_words, const int32_t* __restrict__ rows, int32_t m,
                                           const int64_t* __restrict__ group_start,
                                           const int64_t* __restrict__ cum_pairs, const uint8_t* __restrict__ status,
                                           int32_t* __restrict__ counts, int32_t layout,
                                           const uint64_t* __restrict__ occ, int64_t occ_pitch, int64_t g,
                                           int64_t ch) {
    constexpr int NP = S * (S - 1) / 2;
    constexpr int U = SmallUnroll<S>::U;
    const int lane = threadIdx.x & 63;
    const int64_t gs = group_start[g];
    const int s = (int)(group_start[g + 1] - gs);
    if (s < 2 || s > S) return;
    const int np = m - 1;
    const int64_t cbase = cum_pairs[g];
    const int npairs = s * (s - 1) / 2;

    // The group's counted pairs in the unrolled (i, j) order (bit a of live[a / 64]; wave-uniform). Pairs that
    // are not counted (past the group's size, or skipped / inferred by the subset test) cost no VALU work and
    // no reduction below, and a suffix column in no counted pair is not read.
    constexpr int NL = (NP + 63) / 64;
    uint64_t live[NL];
    uint32_t used = 0;
    bool any_live = false;
#pragma unroll
    for (int h = 0; h < NL; ++h) {
        const int a = h * 64 + lane;
        bool lv = false;
        if (a < NP) {
            int i = 0, rem = a;
            while (rem >= S - 1 - i) {
                rem -= S - 1 - i;
                ++i;
            }
            const int j = i + 1 + rem;
            lv = j < s && (status == nullptr || status[cbase + pair_slot(i, j, s, layout)] == ETM_CAND_COUNT);
        }
        live[h] = __ballot(lv);
        any_live = any_live || live[h] != 0;
    }
    if (!any_live) return;
    {
        int a = 0;
#pragma unroll
        for (int i = 0; i < S; ++i)
#pragma unroll
            for (int j = i + 1; j < S; ++j, ++a)
                if ((live[a >> 6] >> (a & 63)) & 1ull) used |= (1u << i) | (1u << j);
    }

    const uint64_t* scol[S];
#pragma unroll
    for (int t = 0; t < S; ++t) scol[t] = bv + (t < s ? (int64_t)rows[(gs + t) * m + np] * pitch : 0);
    const int32_t* prow = rows + gs * m;

    int acc[NP];
#pragma unroll
    for (int a = 0; a < NP; ++a) acc[a] = 0;

    const int64_t w_begin = ch * chunk_words;
    const int64_t w_end = min(n_words, w_begin + chunk_words);
    // The counted pairs' popcounts of U words per lane, the suffix words already ANDed with the prefix.
    auto accumulate = [&](const uint64_t (&sv)[U][S]) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            int a = 0;
#pragma unroll
            for (int i = 0; i < S; ++i) {
#pragma unroll
                for (int j = i + 1; j < S; ++j) {
                    if ((live[a >> 6] >> (a & 63)) & 1ull) acc[a] = popc_acc(sv[u][i] & sv[u][j], acc[a]);
                    ++a;
                }
            }
        }
    };
    // U words per lane: w[u] (a valid word when !in[u]; its result is masked out).
    auto step = [&](const int64_t (&w)[U], const bool (&in)[U]) {
        uint64_t pre[U];
#pragma unroll
        for (int u = 0; u < U; ++u) pre[u] = in[u] ? ~0ull : 0ull;
        // Prefix AND, four prefix columns at a time: 4 x U independent loads in flight.
        for (int t = 0; t < np; t += 4) {
            const uint64_t* c0 = bv + (int64_t)prow[t] * pitch;
            const uint64_t* c1 = bv + (int64_t)prow[t + 1 < np ? t + 1 : t] * pitch;
            const uint64_t* c2 = bv + (int64_t)prow[t + 2 < np ? t + 2 : t] * pitch;
            const uint64_t* c3 = bv + (int64_t)prow[t + 3 < np ? t + 3 : t] * pitch;
#pragma unroll
            for (int u = 0; u < U; ++u) pre[u] &= c0[w[u]] & c1[w[u]] & c2[w[u]] & c3[w[u]];
        }
        uint64_t sv[U][S];
#pragma unroll
        for (int t = 0; t < S; ++t)
#pragma unroll
            for (int u = 0; u < U; ++u) sv[u][t] = ((used >> t) & 1u) ? (scol[t][w[u]] & pre[u]) : 0ull;
        accumulate(sv);
    };
    if constexpr (VEC && U >= 2) {
        // VEC: a lane reads words w2[v] and w2[v] + 1 as one 16-byte load (V = U / 2 of them), half the load
        // instructions of `step`. Needs every column 16-byte aligned (even pitch, aligned base) and an even chunk
        // start; a pair's second word may be the padding word past n_words, which an even pitch keeps in the column.
        constexpr int V = U / 2;
        auto step2 = [&](const int64_t (&w2)[V], const bool (&in)[U]) {
            uint64_t pre[U];
#pragma unroll
            for (int u = 0; u < U; ++u) pre[u] = in[u] ? ~0ull : 0ull;
            for (int t = 0; t < np; t += 4) {
                const uint64_t* c0 = bv + (int64_t)prow[t] * pitch;
                const uint64_t* c1 = bv + (int64_t)prow[t + 1 < np ? t + 1 : t] * pitch;
                const uint64_t* c2 = bv + (int64_t)prow[t + 2 < np ? t + 2 : t] * pitch;
                const uint64_t* c3 = bv + (int64_t)prow[t + 3 < np ? t + 3 : t] * pitch;
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    const ulonglong2 x0 = *reinterpret_cast<const ulonglong2*>(c0 + w2[v]);
                    const ulonglong2 x1 = *reinterpret_cast<const ulonglong2*>(c1 + w2[v]);
                    const ulonglong2 x2 = *reinterpret_cast<const ulonglong2*>(c2 + w2[v]);
                    const ulonglong2 x3 = *reinterpret_cast<const ulonglong2*>(c3 + w2[v]);
                    pre[2 * v] &= x0.x & x1.x & x2.x & x3.x;
                    pre[2 * v + 1] &= x0.y & x1.y & x2.y & x3.y;
                }
            }
            uint64_t sv[U][S];
#pragma unroll
            for (int t = 0; t < S; ++t)
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    ulonglong2 x = make_ulonglong2(0ull, 0ull);
                    if ((used >> t) & 1u) x = *reinterpret_cast<const ulonglong2*>(scol[t] + w2[v]);
                    sv[2 * v][t] = x.x & pre[2 * v];
                    sv[2 * v + 1][t] = x.y & pre[2 * v + 1];
                }
            accumulate(sv);
        };
        for (int64_t w0 = w_begin + 2 * lane; w0 < w_end; w0 += 128 * V) {
            bool in[U];
            int64_t w2[V];
#pragma unroll
            for (int v = 0; v < V; ++v) {
                w2[v] = w0 + 128 * v;
                in[2 * v] = w2[v] < w_end;
                in[2 * v + 1] = w2[v] + 1 < w_end;
                if (!in[2 * v]) w2[v] = w_begin;
            }
            step2(w2, in);
        }
    } else if constexpr (!SKIP) {
        for (int64_t w0 = w_begin + lane; w0 < w_end; w0 += 64 * U) {
            bool in[U];
            int64_t w[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                w[u] = w0 + 64 * u;
                in[u] = w[u] < w_end;
                if (!in[u]) w[u] = w_begin;
            }
            step(w, in);
        }
    } else {
        // Tiles of 64 words the chunk spans (at most 64: the launcher caps the chunk at 4096 words), kept where
        // every prefix column and at least two suffix columns are occupied: elsewhere every pair's AND is zero.
        const int64_t t0 = w_begin >> 6;
        const int nt = (int)((w_end - w_begin + 63) >> 6);
        auto window = [&](int32_t col) {
            const uint64_t* o = occ + (int64_t)col * occ_pitch + (t0 >> 6);
            const int sh = (int)(t0 & 63);
            return sh ? (o[0] >> sh) | (o[1] << (64 - sh)) : o[0];
        };
        uint64_t tiles = nt >= 64 ? ~0ull : ((1ull << nt) - 1);
        for (int t = 0; t < np; ++t) tiles &= window(prow[t]);
        uint64_t one = 0, two = 0;
        for (int t = 0; t < s; ++t) {
            const uint64_t x = window(rows[(gs + t) * m + np]);
            two |= one & x;
            one |= x;
        }
        tiles &= two;
        // The plain loop's shape (U tiles per step, lane-strided words), skipping a step whose U tiles are all
        // empty and masking the empty ones of the rest.
        for (int64_t w0 = w_begin + lane; w0 < w_end; w0 += 64 * U) {
            const int rel = (int)((w0 - lane - w_begin) >> 6);
            const uint64_t sel = (tiles >> rel) & ((1ull << U) - 1);
            if (sel == 0) continue;
            bool in[U];
            int64_t w[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                w[u] = w0 + 64 * u;
                in[u] = w[u] < w_end && ((sel >> u) & 1ull);
                if (!in[u]) w[u] = w_begin;
            }
            step(w, in);
        }
    }

    int a = 0;
#pragma unroll
    for (int i = 0; i < S; ++i) {
#pragma unroll
        for (int j = i + 1; j < S; ++j) {
            if ((live[a >> 6] >> (a & 63)) & 1ull) {
                const int tot = etm::wave_sum(acc[a]);
                if (lane == (a & 63)) atomicAdd(&counts[cbase + pair_slot(i, j, s, layout)], tot);
            }
            ++a;
        }
    }
}

// One wave per (word chunk, group) unit of a size class, spread over the 8 XCDs by xcd_unit.
template <int S, bool SKIP, bool VEC>
__global__ __launch_bounds__(256) void k_count_small(const uint64_t* __restrict__ bv, int64_t pitch, int64_t n_words,
                                                      int64_t chunk_words, int64_t n_chunks,
                                                      const int32_t* __restrict__ rows, int32_t m,
                                                      const int64_t* __restrict__ group_start,
                                                      const int64_t* __restrict__ cum_pairs,
                                                      const int32_t* __restrict__ list, int64_t n_list,
                                                      const uint8_t* __restrict__ status,
                                                      int32_t* __restrict__ counts, int32_t layout,
                                                      const uint64_t* __restrict__ occ, int64_t occ_pitch) {
    int64_t ch, item;
    if (!xcd_unit(blockIdx.x % kXcds, (int64_t)(blockIdx.x / kXcds) * 4 + (threadIdx.x >> 6), n_list, n_chunks, &ch,
                  &item))
        return;
    small_unit<S, SKIP, VEC>(bv, pitch, n_words, chunk_words, rows, m, group_start, cum_pairs, status, counts, layout,
                             occ, occ_pitch, uniform64(list[uniform64(item)]), uniform64(ch));
}

// Over a class's live groups (the groups with at least one counted pair, listed on the device by k_live_units):
// *d_n_live groups, each cut into chunks sized from that count (live_chunks), the (group, chunk) units spread
// over the 8 XCDs by xcd_unit, one wave each at a time. The grid is sized from an upper bound of the units, capped
// (live_grid_per_xcd): each wave takes every stride-th unit of its XCD and returns past the last.
template <int S, bool VEC, bool LOOP>
__global__ __launch_bounds__(256) void k_count_small_live(const uint64_t* __restrict__ bv, int64_t pitch,
                                                           int64_t n_words, const int32_t* __restrict__ rows,
                                                           int32_t m, const int64_t* __restrict__ group_start,
                                                           const int64_t* __restrict__ cum_pairs,
                                                           const int32_t* __restrict__ live,
                                                           const int64_t* __restrict__ d_n_live,
                                                           const uint8_t* __restrict__ status,
                                                           int32_t* __restrict__ counts, int32_t layout,
                                                           int64_t target, int64_t min_words) {
    const int64_t n_live = *d_n_live;
    if (n_live == 0) return;
    int64_t chunk, n_chunks;
    live_chunks(n_words, n_live, target, 64, min_words, &chunk, &n_chunks);
    const int64_t stride = (int64_t)(gridDim.x / kXcds) * 4;
    for (int64_t u = (int64_t)(blockIdx.x / kXcds) * 4 + (threadIdx.x >> 6);; u += stride) {
        int64_t ch, item;
        if (!xcd_unit(blockIdx.x % kXcds, u, n_live, n_chunks, &ch, &item)) return;
        small_unit<S, false, VEC>(bv, pitch, n_words, chunk, rows, m, group_start, cum_pairs, status, counts, layout,
                                  nullptr, 0, uniform64(live[uniform64(item)]), uniform64(ch));
        if (!LOOP) return;
    }
}

// LDS of the tiled kernels. Dense mode: A and B tiles [word][column] of kKW words. Listed mode: the live columns
// compacted into slots, [word][slot] of up to kListWords words, in the same buffer.
struct TiledShared {
    uint64_t tiles[2 * kKW * (kT + 1)];
    uint64_t P[kKW];
    int64_t col[2 * kT];
    uint8_t la[kListMax], lb[kListMax];
    int at[kT];
    unsigned long long use[2];
    int nlist;
};

// One (tile-pair, word chunk) unit, counted by the calling workgroup (every thread calls it with the same unit).
__device__ __forceinline__ void tiled_unit(TiledShared& sh, const uint64_t* __restrict__ bv, int64_t pitch,
                                           int64_t n_words, int64_t chunk_words, const int32_t* __restrict__ rows,
                                           int32_t m, const int64_t* __restrict__ group_start,
                                           const int64_t* __restrict__ cum_pairs,
                                           const int64_t* __restrict__ cum_tiles, int64_t n_groups,
                                           const uint8_t* __restrict__ status, int32_t* __restrict__ counts,
                                           int32_t layout, int64_t tp, int64_t chunk_id) {
    uint64_t* s_tiles = sh.tiles;
    uint64_t* sP = sh.P;
    int64_t* s_col = sh.col;
    uint64_t(*sA)[kT + 1] = reinterpret_cast<uint64_t(*)[kT + 1]>(s_tiles);
    uint64_t(*sB)[kT + 1] = reinterpret_cast<uint64_t(*)[kT + 1]>(s_tiles + kKW * (kT + 1));
    int64_t* colA = s_col;
    int64_t* colB = s_col + kT;
    uint8_t* s_la = sh.la;
    uint8_t* s_lb = sh.lb;
    int& s_nlist = sh.nlist;
    unsigned long long* s_use = sh.use;

    // The thread index passes through an empty asm statement, so that a persistent caller's loop cannot hoist
    // the body's per-thread values out of it and keep them all live across units (255 registers instead of 114).
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const int64_t g = etm::wave_upper_index(cum_tiles, n_groups, tp);
    const int64_t gs = group_start[g];
    const int64_t s = group_start[g + 1] - gs;
    const int64_t nt = (s + kT - 1) / kT;
    int64_t ta, tb1;
    etm::lex_pair_decode(tp - cum_tiles[g], nt + 1, &ta, &tb1);
    const int64_t tb = tb1 - 1;
    const bool diag = ta == tb;
    const int np = m - 1;
    const int64_t cbase = cum_pairs[g];
    const int tx = tid & 15, ty = tid >> 4;

    // Liveness of this thread's 16 pairs; a sparsely live tile-pair lists its live pairs and counts only those.
    int my_live = 0;
    uint32_t live_bits = 0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int64_t i = ta * kT + tx + 16 * r, j = tb * kT + ty + 16 * c;
            if (i < j && j < s &&
                (status == nullptr || status[cbase + pair_slot(i, j, s, layout)] == ETM_CAND_COUNT)) {
                ++my_live;
                live_bits |= 1u << (4 * r + c);
            }
        }
    }
    // Which of the tile's A and B columns take part in a live pair: in listed mode only those are staged.
    if (tid == 0) {
        s_nlist = 0;
        s_use[0] = 0;
        s_use[1] = 0;
    }
    __syncthreads();
    if (my_live) {
        atomicAdd(&s_nlist, my_live);
        unsigned long long ua = 0, ub = 0;
#pragma unroll
        for (int r = 0; r < 4; ++r)
            if ((live_bits >> (4 * r)) & 0xFu) ua |= 1ull << (tx + 16 * r);
#pragma unroll
        for (int c = 0; c < 4; ++c)
            if (live_bits & (0x1111u << c)) ub |= 1ull << (ty + 16 * c);
        atomicOr(&s_use[0], ua);
        atomicOr(&s_use[1], ub);
    }
    __syncthreads();
    const int n_live = s_nlist;
    if (n_live == 0) return;
    const bool listed = n_live <= kListMax;
    __syncthreads();
    if (listed) {
        // The list grouped by the pair's A column (a counting sort over the tile's 64 A columns): the lanes of a
        // wave then mostly read the same A word (an LDS broadcast) instead of scattered ones (with 128 columns
        // staged, arrival order measured 37% of the LDS-active cycles in bank conflicts, grouped 20%).
        int* s_at = sh.at;
        if (tid < kT) s_at[tid] = 0;
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int n_r = __popc((live_bits >> (4 * r)) & 0xFu);
            if (n_r) atomicAdd(&s_at[tx + 16 * r], n_r);
        }
        __syncthreads();
        if (tid < 64) {  // exclusive scan of the 64 counts by wave 0
            const int v = s_at[tid];
            const int inc = etm::wave_scan_i32(v);
            s_at[tid] = inc - v;
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const uint32_t row = (live_bits >> (4 * r)) & 0xFu;
            if (!row) continue;
            const int a = tx + 16 * r;
            int at = atomicAdd(&s_at[a], __popc(row));
#pragma unroll
            for (int c = 0; c < 4; ++c)
                if (row & (1u << c)) {
                    s_la[at] = (uint8_t)a;
                    s_lb[at] = (uint8_t)(ty + 16 * c);
                    ++at;
                }
        }
        __syncthreads();
    }

    const int32_t* prow = rows + gs * m;
    const int64_t w_begin = chunk_id * chunk_words;
    const int64_t w_end = min(n_words, w_begin + chunk_words);

    if (listed) {
        // Only the columns of live pairs are staged, compacted into slots (on a diagonal tile-pair both sides of
        // a pair index the A tile). With fewer slots a step covers more words: the largest power of two up to
        // kListWords whose [word][slot] block fits the tile buffer, so a sparsely live tile-pair pays its global
        // load latencies and barriers over several times more words per step.
        const uint64_t uA = diag ? (s_use[0] | s_use[1]) : s_use[0];
        const uint64_t uB = diag ? 0ull : s_use[1];
        const int nA = __popcll(uA), nS = nA + __popcll(uB);
        const int stride = nS | 1;  // odd: a wave's 64-bit stores down one slot spread over the banks
        int lg = 5;
        while ((1 << (lg + 1)) <= kListWords && (stride << (lg + 1)) <= 2 * kKW * (kT + 1)) ++lg;
        const int kw = 1 << lg;
        if (tid < kT) {
            if ((uA >> tid) & 1ull)
                s_col[__popcll(uA & ((1ull << tid) - 1))] = (int64_t)rows[(gs + ta * kT + tid) * m + np] * pitch;
        } else if (tid < 2 * kT) {
            const int c = tid - kT;
            if ((uB >> c) & 1ull)
                s_col[nA + __popcll(uB & ((1ull << c) - 1))] = (int64_t)rows[(gs + tb * kT + c) * m + np] * pitch;
        }
        // This thread's list entries tid, tid + 256, ... as slots; or, with at most 128 live pairs, tpp threads
        // per pair (pair tid / tpp), each counting every tpp-th word of a step (a short list otherwise leaves most
        // of the workgroup idle while a few threads walk every word).
        int tpp = 1;
        while (tpp < 8 && n_live * (2 * tpp) <= 256) tpp <<= 1;
        const int sub = tid & (tpp - 1);
        int la[kListMax / 256], lb[kListMax / 256], sum[kListMax / 256];
#pragma unroll
        for (int e = 0; e < kListMax / 256; ++e) {
            const int li = tpp > 1 ? (e == 0 ? tid / tpp : n_live) : tid + 256 * e;
            sum[e] = 0;
            la[e] = 0;
            lb[e] = 0;
            if (li < n_live) {
                const int a = s_la[li], b = s_lb[li];
                la[e] = __popcll(uA & ((1ull << a) - 1));
                lb[e] = diag ? __popcll(uA & ((1ull << b) - 1)) : nA + __popcll(uB & ((1ull << b) - 1));
            }
        }
        __syncthreads();
        // kw divides 256, so a thread stages the same word wq of every step, for slots c0, c0 + cstep, ..., and
        // ANDs its own prefix word in: a step waits on one round of global loads. (Loading the next step's words
        // while the current one counts measured slower: the registers held across the count cost a wave per
        // SIMD.)
        const int wq = tid & (kw - 1), c0 = tid >> lg, cstep = 256 >> lg;
        for (int64_t wb = w_begin; wb < w_end; wb += kw) {
            const bool in = wb + wq < w_end;
            uint64_t p = in ? ~0ull : 0ull;
            if (in)
                for (int t = 0; t < np; ++t) p &= bv[(int64_t)prow[t] * pitch + wb + wq];
#pragma unroll 4
            for (int c = c0; c < nS; c += cstep) s_tiles[wq * stride + c] = in ? (bv[s_col[c] + wb + wq] & p) : 0ull;
            __syncthreads();
            if (tpp > 1) {
                if (tid / tpp < n_live) {
                    int acc = sum[0];
                    const uint64_t* pa = s_tiles + la[0];
                    const uint64_t* pb = s_tiles + lb[0];
                    for (int w = sub; w < kw; w += tpp) acc = popc_acc(pa[w * stride] & pb[w * stride], acc);
                    sum[0] = acc;
                }
            } else {
#pragma unroll
                for (int e = 0; e < kListMax / 256; ++e) {
                    if (tid + 256 * e < n_live) {
                        int acc = sum[e];
                        const uint64_t* pa = s_tiles + la[e];
                        const uint64_t* pb = s_tiles + lb[e];
#pragma unroll 8
                        for (int w = 0; w < kw; ++w) acc = popc_acc(pa[w * stride] & pb[w * stride], acc);
                        sum[e] = acc;
                    }
                }
            }
            __syncthreads();
        }
        if (tpp > 1) {
            // The tpp lanes of a pair are adjacent in one wave: sum them, lane sub == 0 adds the pair's count.
            int a = sum[0];
            for (int off = tpp >> 1; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
            const int li = tid / tpp;
            if (sub == 0 && li < n_live && a) {
                const int64_t i = ta * kT + s_la[li], j = tb * kT + s_lb[li];
                atomicAdd(&counts[cbase + pair_slot(i, j, s, layout)], a);
            }
        } else {
#pragma unroll
            for (int e = 0; e < kListMax / 256; ++e) {
                const int li = tid + 256 * e;
                if (li < n_live && sum[e]) {
                    const int64_t i = ta * kT + s_la[li], j = tb * kT + s_lb[li];
                    atomicAdd(&counts[cbase + pair_slot(i, j, s, layout)], sum[e]);
                }
            }
        }
        return;
    }

    if (tid < kT) {
        const int64_t i = ta * kT + tid;
        colA[tid] = i < s ? (int64_t)rows[(gs + i) * m + np] * pitch : -1;
    } else if (tid < 2 * kT) {
        const int64_t j = tb * kT + (tid - kT);
        colB[tid - kT] = j < s ? (int64_t)rows[(gs + j) * m + np] * pitch : -1;
    }

    int acc[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[r][c] = 0;

    __syncthreads();
    for (int64_t wb = w_begin; wb < w_end; wb += kKW) {
        if (tid < kKW) sP[tid] = (wb + tid < w_end) ? ~0ull : 0ull;
        __syncthreads();
        if (np > 0) {
            // 256 threads: word tid % 32, prefix columns tid / 32, tid / 32 + 8, ...
            const int w = tid % kKW;
            if (wb + w < w_end) {
                uint64_t part = ~0ull;
                for (int t = tid / kKW; t < np; t += 256 / kKW) part &= bv[(int64_t)prow[t] * pitch + wb + w];
                if (tid / kKW < np) atomicAnd((unsigned long long*)&sP[w], (unsigned long long)part);
            }
            __syncthreads();
        }
#pragma unroll
        for (int q = 0; q < (kT * kKW) / 256; ++q) {
            const int e = tid + 256 * q;
            const int c = e / kKW, w = e % kKW;
            const int64_t ca = colA[c];
            sA[w][c] = (ca >= 0 && wb + w < w_end) ? (bv[ca + wb + w] & sP[w]) : 0ull;
            if (!diag) {
                const int64_t cb = colB[c];
                sB[w][c] = (cb >= 0 && wb + w < w_end) ? (bv[cb + wb + w] & sP[w]) : 0ull;
            }
        }
        __syncthreads();
        uint64_t(*B)[kT + 1] = diag ? sA : sB;
#pragma unroll 4
        for (int w = 0; w < kKW; ++w) {
            uint64_t av[4], bw[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) av[r] = sA[w][tx + 16 * r];
#pragma unroll
            for (int c = 0; c < 4; ++c) bw[c] = B[w][ty + 16 * c];
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int c = 0; c < 4; ++c) acc[r][c] = popc_acc(av[r] & bw[c], acc[r][c]);
        }
        __syncthreads();
    }

#pragma unroll
    for (int r = 0; r < 4; ++r) {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int64_t i = ta * kT + tx + 16 * r, j = tb * kT + ty + 16 * c;
            if (i < j && j < s) {
                const int64_t slot = cbase + pair_slot(i, j, s, layout);
                if (status == nullptr || status[slot] == ETM_CAND_COUNT) atomicAdd(&counts[slot], acc[r][c]);
            }
        }
    }
}

// One workgroup per (word chunk, tile-pair) unit, spread over the 8 XCDs by xcd_unit.
__global__ __launch_bounds__(256) void k_count_tiled(const uint64_t* __restrict__ bv, int64_t pitch, int64_t n_words,
                                                      int64_t chunk_words, const int32_t* __restrict__ rows,
                                                      int32_t m, const int64_t* __restrict__ group_start,
                                                      const int64_t* __restrict__ cum_pairs,
                                                      const int64_t* __restrict__ cum_tiles, int64_t n_groups,
                                                      int64_t n_tile_pairs, int64_t n_chunks,
                                                      const uint8_t* __restrict__ status,
                                                      int32_t* __restrict__ counts, int32_t layout) {
    __shared__ TiledShared sh;
    int64_t chunk_id, tp;
    if (!xcd_unit(blockIdx.x % kXcds, blockIdx.x / kXcds, n_tile_pairs, n_chunks, &chunk_id, &tp)) return;
    tiled_unit(sh, bv, pitch, n_words, chunk_words, rows, m, group_start, cum_pairs, cum_tiles, n_groups, status,
               counts, layout, tp, chunk_id);
}

// Over the live tile-pairs (at least one counted pair; listed on the device by k_live_units): *d_n_live of
// them, each cut into chunks sized from that count, one workgroup per (tile-pair, chunk) unit spread over the 8
// XCDs by xcd_unit. The grid is sized from a capped upper bound of the units (live_grid_per_xcd); each workgroup
// takes every stride-th unit of its XCD.
template <bool LOOP>
__global__ __launch_bounds__(256) void k_count_tiled_live(const uint64_t* __restrict__ bv, int64_t pitch,
                                                           int64_t n_words, const int32_t* __restrict__ rows,
                                                           int32_t m, const int64_t* __restrict__ group_start,
                                                           const int64_t* __restrict__ cum_pairs,
                                                           const int64_t* __restrict__ cum_tiles, int64_t n_groups,
                                                           const int32_t* __restrict__ live,
                                                           const int64_t* __restrict__ d_n_live,
                                                           const uint8_t* __restrict__ status,
                                                           int32_t* __restrict__ counts, int32_t layout,
                                                           int64_t target, int64_t min_words) {
    __shared__ TiledShared sh;
    const int64_t n_live = *d_n_live;
    if (n_live == 0) return;
    int64_t chunk, n_chunks;
    live_chunks(n_words, n_live, target, kKW, min_words, &chunk, &n_chunks);
    for (int64_t u = blockIdx.x / kXcds;; u += gridDim.x / kXcds) {
        int64_t ch, item;
        if (!xcd_unit(blockIdx.x % kXcds, u, n_live, n_chunks, &ch, &item)) return;
        tiled_unit(sh, bv, pitch, n_words, chunk, rows, m, group_start, cum_pairs, cum_tiles, n_groups, status,
                   counts, layout, uniform64(live[uniform64(item)]), uniform64(ch));
        if (!LOOP) return;
        __syncthreads();  // the next unit's set-up overwrites this one's LDS
    }
}

// One wave per (sparse tile-pair, word chunk): lane p < n holds pair p's two column offsets and its count slot;
// per word, the prefix AND, then for each pair the AND of its two columns with it (the pair's offsets broadcast
// from lane p), popcounts accumulated per lane; a wave sum and one atomic per pair at the end.
template <bool VEC>
__device__ __forceinline__ void sparse_unit(const uint64_t* __restrict__ bv, int64_t pitch, int64_t n_words,
                                            const int32_t* __restrict__ rows, int32_t m,
                                            const int64_t* __restrict__ group_start,
                                            const int64_t* __restrict__ cum_pairs,
                                            const int64_t* __restrict__ cum_tiles, int64_t n_groups,
                                            const SparseTP* __restrict__ e, int64_t chunk, int64_t ch,
                                            int32_t* __restrict__ counts) {
    const int lane = threadIdx.x & 63;
    const int64_t tp = uniform64(e->tp);
    const int n = (int)uniform64(e->n);
    const int64_t g = etm::wave_upper_index(cum_tiles, n_groups, tp);
    const int64_t gs = group_start[g];
    const int64_t s = group_start[g + 1] - gs;
    const int64_t nt = (s + kT - 1) / kT;
    int64_t ta, tb1;
    etm::lex_pair_decode(tp - cum_tiles[g], nt + 1, &ta, &tb1);
    const int64_t tb = tb1 - 1;
    const int np = m - 1;
    int64_t ca = 0, cb = 0, slot = 0;
    if (lane < n) {
        const int64_t i = ta * kT + e->a[lane], j = tb * kT + e->b[lane];
        ca = (int64_t)rows[(gs + i) * m + np] * pitch;
        cb = (int64_t)rows[(gs + j) * m + np] * pitch;
        slot = cum_pairs[g] + etm::lex_pair_index(i, j, s);
    }
    const int32_t* prow = rows + gs * m;
    int acc[kSparseMax];
#pragma unroll
    for (int p = 0; p < kSparseMax; ++p) acc[p] = 0;
    const int64_t w_begin = ch * chunk;
    const int64_t w_end = min(n_words, w_begin + chunk);
    // Two words a lane per iteration: one 16-byte load per column (VEC), or two 8-byte loads 64 words apart.
    for (int64_t w0 = w_begin + (VEC ? 2 * lane : lane); w0 < w_end; w0 += 128) {
        const int64_t w1 = VEC ? w0 + 1 : w0 + 64;
        const bool in0 = true, in1 = w1 < w_end;
        uint64_t pre0 = in0 ? ~0ull : 0ull, pre1 = in1 ? ~0ull : 0ull;
        for (int t = 0; t < np; ++t) {
            const uint64_t* c = bv + (int64_t)prow[t] * pitch;
            if (VEC) {
                const ulonglong2 x = *reinterpret_cast<const ulonglong2*>(c + w0);
                pre0 &= x.x;
                pre1 &= x.y;
            } else {
                pre0 &= c[w0];
                pre1 &= in1 ? c[w1] : 0ull;
            }
        }
        if (__ballot((pre0 | pre1) != 0) == 0) continue;  // the prefix is empty on the wave's words
#pragma unroll
        for (int p = 0; p < kSparseMax; ++p) {
            if (p >= n) break;
            const int64_t oa = (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(ca >> 32), p) << 32) |
                                         (uint32_t)__builtin_amdgcn_readlane((int)ca, p));
            const int64_t ob = (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(cb >> 32), p) << 32) |
                                         (uint32_t)__builtin_amdgcn_readlane((int)cb, p));
            uint64_t x0, x1;
            if (VEC) {
                const ulonglong2 a = *reinterpret_cast<const ulonglong2*>(bv + oa + w0);
                const ulonglong2 b = *reinterpret_cast<const ulonglong2*>(bv + ob + w0);
                x0 = a.x & b.x;
                x1 = a.y & b.y;
            } else {
                x0 = bv[oa + w0] & bv[ob + w0];
                x1 = in1 ? (bv[oa + w1] & bv[ob + w1]) : 0ull;
            }
            acc[p] = popc_acc(x1 & pre1, popc_acc(x0 & pre0, acc[p]));
        }
    }
#pragma unroll
    for (int p = 0; p < kSparseMax; ++p) {
        if (p >= n) break;
        const int tot = etm::wave_sum(acc[p]);
        if (lane == p && tot) atomicAdd(&counts[slot], tot);
    }
}

template <bool VEC, bool LOOP>
__global__ __launch_bounds__(256) void k_count_sparse_live(const uint64_t* __restrict__ bv, int64_t pitch,
                                                            int64_t n_words, const int32_t* __restrict__ rows,
                                                            int32_t m, const int64_t* __restrict__ group_start,
                                                            const int64_t* __restrict__ cum_pairs,
                                                            const int64_t* __restrict__ cum_tiles, int64_t n_groups,
                                                            const SparseTP* __restrict__ sparse,
                                                            const int64_t* __restrict__ d_n_live,
                                                            int32_t* __restrict__ counts, int64_t target,
                                                            int64_t min_words) {
    const int64_t n_live = *d_n_live;
    if (n_live == 0) return;
    int64_t chunk, n_chunks;
    live_chunks(n_words, n_live, target, 64, min_words, &chunk, &n_chunks);
    const int64_t stride = (int64_t)(gridDim.x / kXcds) * 4;
    for (int64_t u = (int64_t)(blockIdx.x / kXcds) * 4 + (threadIdx.x >> 6);; u += stride) {
        int64_t ch, item;
        if (!xcd_unit(blockIdx.x % kXcds, u, n_live, n_chunks, &ch, &item)) return;
        sparse_unit<VEC>(bv, pitch, n_words, rows, m, group_start, cum_pairs, cum_tiles, n_groups,
                         sparse + uniform64(item), chunk, uniform64(ch), counts);
        if (!LOOP) return;
    }
}

// ── live units ─────────────────────────────────────────────────────────
//
// The groups of the small classes and the tile-pairs of the tiled groups with at least one candidate left
// ETM_CAND_COUNT, listed so that the persistent counting kernels run over live work only (after the subset test,
// Pascal inference and the deduction bounds, most units of a deep level have none). Lists are in no particular
// order; every count is an integer atomic, so the result does not depend on it.

// One launch for both lists. Workgroups [0, small_blocks) take the small groups, a thread per group of
// small_list[0 .. n_small): class c's live groups go to live_small[start_c ..], counted in n_live[c]. The other
// workgroups take the tile-pairs, a wave per tile-pair (lane l checks row ta * 64 + l against the rows of tile tb):
// sparse live tile-pairs go to `sparse` with their pair lists, counted in n_live[5], the other live ones to
// live_tp, counted in n_live[4]; tile-pairs of groups flagged in gl_flag (when given) are left to the
// group-listed kernel. n_live[0..5] start at zero.
__global__ __launch_bounds__(256) void k_live_units(const int32_t* __restrict__ small_list, int64_t n_small,
                                                    ClassBounds cb, int64_t small_blocks,
                                                    const int64_t* __restrict__ group_start,
                                                    const int64_t* __restrict__ cum_pairs,
                                                    const int64_t* __restrict__ cum_tiles, int64_t n_groups,
                                                    int64_t n_tile_pairs, const uint8_t* __restrict__ status,
                                                    int32_t* __restrict__ live_small, int32_t* __restrict__ live_tp,
                                                    SparseTP* __restrict__ sparse, int64_t* __restrict__ n_live,
                                                    const uint8_t* __restrict__ gl_flag) {
    if ((int64_t)blockIdx.x < small_blocks) {
        for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n_small;
             e += small_blocks * (int64_t)blockDim.x) {
            const int32_t g = small_list[e];
            const int64_t s = group_start[g + 1] - group_start[g];
            const int64_t base = cum_pairs[g];
            bool any = false;
            for (int64_t q = 0; q < s * (s - 1) / 2 && !any; ++q) any = status[base + q] == ETM_CAND_COUNT;
            if (!any) continue;
            int c = 0;
            while (c < 3 && e >= cb.start[c + 1]) ++c;
            const int64_t at = (int64_t)atomicAdd((unsigned long long*)&n_live[c], 1ull);
            live_small[cb.start[c] + at] = g;
        }
        return;
    }
    const int lane = threadIdx.x & 63;
    const int64_t waves = ((int64_t)gridDim.x - small_blocks) * (blockDim.x >> 6);
    for (int64_t tp = ((int64_t)blockIdx.x - small_blocks) * (blockDim.x >> 6) + (threadIdx.x >> 6); tp < n_tile_pairs;
         tp += waves) {
        const int64_t g = etm::wave_upper_index(cum_tiles, n_groups, tp);
        if (gl_flag != nullptr && gl_flag[g]) continue;  // counted by the group-listed kernel
        const int64_t gs = group_start[g];
        const int64_t s = group_start[g + 1] - gs;
        const int64_t nt = (s + kT - 1) / kT;
        int64_t ta, tb1;
        etm::lex_pair_decode(tp - cum_tiles[g], nt + 1, &ta, &tb1);
        const int64_t tb = tb1 - 1;
        const int64_t i = ta * kT + lane;
        // Bit j - tb * 64 of `mine`: pair (i, j) is counted.
        uint64_t mine = 0;
        if (i < s) {
            const int64_t j0 = max(i + 1, tb * kT), j1 = min(s, tb * kT + kT);
            const int64_t base = cum_pairs[g] + etm::lex_pair_index(i, j0, s);
            for (int64_t j = j0; j < j1; ++j)
                if (status[base + (j - j0)] == ETM_CAND_COUNT) mine |= 1ull << (j - tb * kT);
        }
        const int cnt = __popcll(mine);
        const int n = etm::wave_sum(cnt);
        if (n == 0) continue;
        // Columns in a counted pair: the A rows with a bit, the B columns of any bit (one set on the diagonal).
        uint64_t used_b = mine;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) used_b |= __shfl_xor(used_b, off, 64);
        const uint64_t used_a = __ballot(mine != 0);
        const int u = ta == tb ? __popcll(used_a | used_b) : __popcll(used_a) + __popcll(used_b);
        if (n <= kSparseMax && n <= kSparseRatio * u) {
            int64_t at = 0;
            if (lane == 0) at = (int64_t)atomicAdd((unsigned long long*)&n_live[5], 1ull);
            at = __shfl(at, 0, 64);
            SparseTP* e = sparse + at;
            int off = etm::wave_scan_i32(cnt) - cnt;  // this lane's first slot in the entry's list
            for (uint64_t x = mine; x; x &= x - 1, ++off) {
                e->a[off] = (uint8_t)lane;
                e->b[off] = (uint8_t)__builtin_ctzll(x);
            }
            if (lane == 0) {
                e->tp = (int32_t)tp;
                e->n = n;
            }
        } else if (lane == 0) {
            const int64_t at = (int64_t)atomicAdd((unsigned long long*)&n_live[4], 1ull);
            live_tp[at] = (int32_t)tp;
        }
    }
}

// ── group-listed counting of the tiled groups ───────────────────────────────────────────────────────
//
// A tiled group (more than small_max suffixes) cut into 64 x 64 tile-pairs re-reads each suffix column once per
// tile-pair it appears in, and every (tile-pair, chunk) unit re-derives its live pairs from 4096 statuses. Where
// a group's counted pairs are few enough (kGlSparseMax < pairs <= kGlPairsMax) over few enough columns (at most
// kGlColsMax), k_gl_scan lists them once per level, and k_count_glist counts each (group, word chunk) unit with
// all of the group's used columns staged in LDS: each column is read once per chunk. The other tiled groups keep
// the tile-pair kernels (k_live_units skips the flagged ones).
//
// Groups of more than kGlSparseMax counted pairs are group-listed; fewer go to the tile-pair path (its sparse
// wave kernel). Measured on deep_sparse_20m's K >= 3 counting: 32: 8.70-8.93 ms, 16: 8.43-8.51, 8: 8.94, 4: 9.53.
constexpr int kGlSparseMax = 16;
constexpr int kGlPairsMax = 4096;          // counted pairs of a group-listed group: 8 per thread
constexpr int kGlThreads = 512;            // threads of a k_count_glist workgroup
constexpr int kGlColsMax = 2048;           // its used suffix columns (LDS slots)
constexpr int64_t kGlScanMax = 1 << 22;    // candidates of a group that one workgroup scans
constexpr int kGlPer = 8;                          // words a thread stages per step
constexpr int kGlLdsWords = kGlThreads * kGlPer;   // staging buffer of k_count_glist (32 KiB)
constexpr int kGlMaxWords = 256;                   // most words per staging step (a divisor of kGlThreads)
constexpr int kGlPreMax = 4;               // prefix words loaded ahead with the columns (more: read in step)
constexpr int kGlScanPer = 16;             // candidates per thread per step of k_gl_scan
constexpr int kGlScanThreads = 1024;       // threads of a k_gl_scan workgroup (one per tiled group)

// Words per staging step of a group-listed group of n_used columns: the largest power of two (at most
// kGlMaxWords) whose [word][slot] block (odd stride) fits kGlLdsWords.
__device__ __forceinline__ int gl_step_log2(int n_used) {
    const int stride = n_used | 1;
    int lg = 0;
    while ((1 << (lg + 1)) <= kGlMaxWords && stride * (1 << (lg + 1)) <= kGlLdsWords) ++lg;
    return lg;
}
// Groups whose steps hold at most this many words have long units (many steps per chunk): they are listed first
// and dispatched first, so that none of them starts late and sets the launch's tail.
constexpr int kGlHeavyWords = 32;

struct GlGroup {
    int64_t pair_start, col_start;
    int32_t g, n_pairs, n_used, pad;
};

struct GlistWs {
    uint8_t* flag;         // [n_groups] 1: the group is group-listed
    GlGroup* groups;       // [n_tiled]
    uint32_t* pair_slots;  // [pair_cap] slot(i) | slot(j) << 16
    uint32_t* pair_ij;     // [pair_cap] i | j << 16
    int32_t* cols;         // [col_cap] column ids of the used suffixes, by slot
};

// Counters of the mode: [0] heavy group-listed groups (listed from the front of ws.groups), [1] pair slots
// taken, [2] column slots taken, [3] light ones (listed from the back), [4] k_count_glist's unit head.
__global__ __launch_bounds__(kGlScanThreads) void k_gl_scan(const int32_t* __restrict__ tiled_list, int64_t n_tiled,
                                                 const int32_t* __restrict__ rows, int32_t m,
                                                 const int64_t* __restrict__ group_start,
                                                 const int64_t* __restrict__ cum_pairs,
                                                 const uint8_t* __restrict__ status, GlistWs ws,
                                                 int64_t* __restrict__ gl_cnt) {
    __shared__ uint64_t used[65536 / 64];
    __shared__ int32_t wpre[65536 / 64];
    __shared__ int64_t lds[kGlScanThreads / 64 + 1];
    __shared__ int64_t s_base[3];
    const int tid = threadIdx.x;
    const int32_t g = tiled_list[blockIdx.x];
    const int64_t gs = group_start[g];
    const int64_t s = group_start[g + 1] - gs;
    const int64_t P = s * (s - 1) / 2;
    const int64_t cbase = cum_pairs[g];
    if (s > 65535 || P > kGlScanMax) {
        if (tid == 0) ws.flag[g] = 0;
        return;
    }
    const int nw = (int)((s + 63) / 64);
    for (int w = tid; w < nw; w += kGlScanThreads) used[w] = 0;
    __syncthreads();
    // Thread t takes candidates [t * 16, t * 16 + 16) of each step of 4096, decoded once and then walked in
    // candidate order ((i, j) -> (i, j + 1), or (i + 1, i + 2) at the end of row i).
    constexpr int kStep = kGlScanThreads * kGlScanPer;
    // Pass 1: the counted pairs and the suffixes they use.
    int64_t n_pairs = 0;
    for (int64_t p0 = (int64_t)tid * kGlScanPer; p0 < P; p0 += kStep) {
        uint32_t live = 0;
#pragma unroll
        for (int q = 0; q < kGlScanPer; ++q)
            if (p0 + q < P && status[cbase + p0 + q] == ETM_CAND_COUNT) live |= 1u << q;
        if (!live) continue;
        int64_t i, j;
        etm::lex_pair_decode(p0, s, &i, &j);
        for (int q = 0; q < kGlScanPer; ++q) {
            if ((live >> q) & 1u) {
                atomicOr((unsigned long long*)&used[i >> 6], 1ull << (i & 63));
                atomicOr((unsigned long long*)&used[j >> 6], 1ull << (j & 63));
                ++n_pairs;
            }
            if (++j == s) {
                ++i;
                j = i + 1;
            }
        }
    }
    n_pairs = etm::block_sum<kGlScanThreads>(n_pairs, lds);
    // Slots: exclusive popcount prefix of the used bitmap, four words a thread.
    int32_t cnt4 = 0;
    for (int w = 4 * tid; w < min(nw, 4 * tid + 4); ++w) cnt4 += __popcll(used[w]);
    int64_t n_used = 0;
    const int64_t pre4 = etm::block_exclusive_scan<kGlScanThreads>(cnt4, lds, &n_used);
    {
        int32_t at = (int32_t)pre4;
        for (int w = 4 * tid; w < min(nw, 4 * tid + 4); ++w) {
            wpre[w] = at;
            at += __popcll(used[w]);
        }
    }
    const bool listed = n_pairs > kGlSparseMax && n_pairs <= kGlPairsMax && n_used <= kGlColsMax;
    if (tid == 0) {
        ws.flag[g] = listed ? 1 : 0;
        if (listed) {
            const bool heavy = (1 << gl_step_log2((int)n_used)) <= kGlHeavyWords;
            const int64_t q = heavy ? (int64_t)atomicAdd((unsigned long long*)&gl_cnt[0], 1ull)
                                    : n_tiled - 1 - (int64_t)atomicAdd((unsigned long long*)&gl_cnt[3], 1ull);
            s_base[0] = (int64_t)atomicAdd((unsigned long long*)&gl_cnt[1], (unsigned long long)n_pairs);
            s_base[1] = (int64_t)atomicAdd((unsigned long long*)&gl_cnt[2], (unsigned long long)n_used);
            ws.groups[q] = GlGroup{s_base[0], s_base[1], g, (int32_t)n_pairs, (int32_t)n_used, 0};
        }
    }
    __syncthreads();
    if (!listed) return;
    const int64_t pair_start = s_base[0], col_start = s_base[1];
    const int np = m - 1;
    auto slot_of = [&](int64_t x) {
        const uint64_t w = used[x >> 6];
        return (uint32_t)(wpre[x >> 6] + __popcll(w & ((1ull << (x & 63)) - 1)));
    };
    for (int64_t i = tid; i < s; i += kGlScanThreads)
        if ((used[i >> 6] >> (i & 63)) & 1ull) ws.cols[col_start + slot_of(i)] = rows[(gs + i) * m + np];
    // Pass 2: the pairs in candidate order (their A slots ascend, so a wave's A reads mostly broadcast).
    int64_t run = 0;
    for (int64_t s0 = 0; s0 < P; s0 += kStep) {
        const int64_t p0 = s0 + (int64_t)tid * kGlScanPer;
        uint32_t live = 0;
#pragma unroll
        for (int q = 0; q < kGlScanPer; ++q)
            if (p0 + q < P && status[cbase + p0 + q] == ETM_CAND_COUNT) live |= 1u << q;
        int64_t total = 0;
        int64_t at = run + etm::block_exclusive_scan<kGlScanThreads>(__popc(live), lds, &total);
        if (live) {
            int64_t i, j;
            etm::lex_pair_decode(p0, s, &i, &j);
            for (int q = 0; q < kGlScanPer; ++q) {
                if ((live >> q) & 1u) {
                    ws.pair_slots[pair_start + at] = slot_of(i) | (slot_of(j) << 16);
                    ws.pair_ij[pair_start + at] = (uint32_t)i | ((uint32_t)j << 16);
                    ++at;
                }
                if (++j == s) {
                    ++i;
                    j = i + 1;
                }
            }
        }
        run += total;
    }
}

// One workgroup of kGlThreads per (group-listed group, word chunk) unit. Per step of kw words the group's used
// columns are staged [word][slot] (odd stride) in LDS, ANDed with the group's prefix; the next step's words are
// loaded into registers while the current step counts, so a step costs its counting, not a global load round
// trip. Pairs: with at most kGlThreads / 2 of them, tpp threads share a pair's words (one accumulator each, summed
// over the tpp lanes at the end); with more, each thread takes a run of consecutive pairs over every word, reading
// a shared A column's word once for the run. One atomic per pair at the end. kw is the largest power of two (at
// most kGlMaxWords) whose block fits the buffer. (Eight waves of 107 registers: four waves a SIMD; four waves of
// twice the pairs and staged words took 200 registers, two waves a SIMD.)
// One (group, word chunk) unit of k_count_glist: unit u of the heavy groups' units (chunk-major), then the light
// ones' (listed from the back). Every thread of the workgroup calls it with the same u.
__device__ __forceinline__ void glist_unit(const uint64_t* __restrict__ bv, int64_t pitch, int64_t n_words,
                                           const int32_t* __restrict__ rows, int32_t m,
                                           const int64_t* __restrict__ group_start,
                                           const int64_t* __restrict__ cum_pairs, const GlistWs& ws, int64_t n_tiled,
                                           int64_t n_heavy, int64_t n_light, int64_t chunk, int64_t n_chunks,
                                           int32_t* __restrict__ counts, int64_t u, uint64_t* buf, int32_t* col_id) {
    int64_t ch, idx;
    if (u < n_heavy * n_chunks) {
        ch = u / n_heavy;
        idx = u % n_heavy;
    } else {
        const int64_t v = u - n_heavy * n_chunks;
        ch = v / n_light;
        idx = n_tiled - 1 - v % n_light;
    }
    const GlGroup e = ws.groups[uniform64(idx)];
    // Laundered thread index: the persistent caller's loop keeps no hoisted per-thread values across units.
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const int64_t gs = group_start[e.g];
    const int64_t s = group_start[e.g + 1] - gs;
    const int np = m - 1;
    const int32_t* prow = rows + gs * m;
    const int n_used = e.n_used, n_pairs = e.n_pairs;
    for (int c = tid; c < n_used; c += kGlThreads) col_id[c] = ws.cols[e.col_start + c];
    const int stride = n_used | 1;
    const int lg = gl_step_log2(n_used);
    const int kw = 1 << lg;
    // Pair work: tpp lanes per pair, or (tpp = 1) thread t takes the npt consecutive pairs [t npt, (t + 1) npt):
    // the list is in candidate order, so they mostly share their A column, whose word is then read from LDS
    // once for all of them (bit q of new_a: pair q's A column differs from pair q - 1's). A thread's pairs past
    // the list count slot 0 against itself and are dropped.
    int tpp = 1;
    while (tpp < 4 && n_pairs * (2 * tpp) <= kGlThreads) tpp <<= 1;
    const int sub = tid & (tpp - 1);
    constexpr int kPer = kGlPairsMax / kGlThreads;
    const int npt = (n_pairs + kGlThreads - 1) / kGlThreads;  // at most kPer
    uint32_t slot[kPer];
    int acc[kPer];
    uint32_t new_a = 1u;
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
        const int pi = tpp > 1 ? (q == 0 ? tid / tpp : n_pairs) : (q < npt ? tid * npt + q : n_pairs);
        slot[q] = pi < n_pairs ? ws.pair_slots[e.pair_start + pi] : 0u;
        acc[q] = 0;
        if (q > 0 && (slot[q] & 0xFFFFu) != (slot[q - 1] & 0xFFFFu)) new_a |= 1u << q;
    }
    __syncthreads();
    const int64_t w_begin = uniform64(ch) * chunk;
    const int64_t w_end = min(n_words, w_begin + chunk);
    // kw divides kGlThreads: thread t stages word t % kw of every step, for slots t / kw + q * (kGlThreads / kw),
    // q < kGlPer.
    const int wq = tid & (kw - 1), c0 = tid >> lg, cstep = kGlThreads >> lg;
    uint64_t nxt[kGlPer];
    uint64_t pw[kGlPreMax];
    auto load_step = [&](int64_t wb) {
        const bool in = wb + wq < w_end;
#pragma unroll
        for (int t = 0; t < kGlPreMax; ++t)
            pw[t] = (in && t < np) ? bv[(int64_t)prow[t] * pitch + wb + wq] : ~0ull;
#pragma unroll
        for (int q = 0; q < kGlPer; ++q) {
            const int c = c0 + q * cstep;
            nxt[q] = (in && c < n_used) ? bv[(int64_t)col_id[c] * pitch + wb + wq] : 0ull;
        }
    };
    load_step(w_begin);
    for (int64_t wb = w_begin; wb < w_end; wb += kw) {
        uint64_t pre = wb + wq < w_end ? ~0ull : 0ull;
#pragma unroll
        for (int t = 0; t < kGlPreMax; ++t) pre &= pw[t];
        for (int t = kGlPreMax; t < np; ++t) pre &= bv[(int64_t)prow[t] * pitch + wb + wq];
#pragma unroll
        for (int q = 0; q < kGlPer; ++q) {
            const int c = c0 + q * cstep;
            if (c < n_used) buf[wq * stride + c] = nxt[q] & pre;
        }
        __syncthreads();
        if (wb + kw < w_end) load_step(wb + kw);
        if (tpp > 1) {
            if (tid / tpp < n_pairs) {
                const uint64_t* pa = buf + (slot[0] & 0xFFFFu);
                const uint64_t* pb = buf + (slot[0] >> 16);
                int a = acc[0];
#pragma unroll 8
                for (int w = sub; w < kw; w += tpp) a = popc_acc(pa[w * stride] & pb[w * stride], a);
                acc[0] = a;
            }
        } else {
            for (int w = 0; w < kw; ++w) {
                const uint64_t* row = buf + w * stride;
                uint64_t va = 0;
#pragma unroll
                for (int q = 0; q < kPer; ++q) {
                    if (q < npt) {  // uniform over the workgroup
                        if ((new_a >> q) & 1u) va = row[slot[q] & 0xFFFFu];
                        acc[q] = popc_acc(va & row[slot[q] >> 16], acc[q]);
                    }
                }
            }
        }
        __syncthreads();
    }
    const int64_t cbase = cum_pairs[e.g];
    if (tpp > 1) {
        int a = acc[0];
        for (int off = tpp >> 1; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
        const int pi = tid / tpp;
        if (sub == 0 && pi < n_pairs && a) {
            const uint32_t ij = ws.pair_ij[e.pair_start + pi];
            atomicAdd(&counts[cbase + etm::lex_pair_index(ij & 0xFFFFu, ij >> 16, s)], a);
        }
        return;
    }
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
        const int pi = tid * npt + q;
        if (q < npt && pi < n_pairs && acc[q]) {
            const uint32_t ij = ws.pair_ij[e.pair_start + pi];
            atomicAdd(&counts[cbase + etm::lex_pair_index(ij & 0xFFFFu, ij >> 16, s)], acc[q]);
        }
    }
}

// A persistent grid over the (group-listed group, word chunk) units: each workgroup takes its own index as its
// first unit, then the next from the level's head counter (gl_cnt[4], zero on entry), so a level with few or no
// group-listed groups costs one pass of exits instead of a grid sized for the most units it could hold.
constexpr int64_t kGlGrid = 512;  // workgroups of k_count_glist: two a CU (107 registers, 40 KiB of LDS)

__global__ __launch_bounds__(kGlThreads) void k_count_glist(const uint64_t* __restrict__ bv, int64_t pitch, int64_t n_words,
                                                     const int32_t* __restrict__ rows, int32_t m,
                                                     const int64_t* __restrict__ group_start,
                                                     const int64_t* __restrict__ cum_pairs, GlistWs ws,
                                                     int64_t n_tiled, int64_t* __restrict__ gl_cnt,
                                                     int32_t* __restrict__ counts, int64_t target,
                                                     int64_t min_words) {
    __shared__ uint64_t buf[kGlLdsWords];
    __shared__ int32_t col_id[kGlColsMax];
    __shared__ int64_t s_next;
    const int64_t n_heavy = gl_cnt[0], n_light = gl_cnt[3];
    const int64_t n_gl = n_heavy + n_light;
    if (n_gl == 0) return;
    int64_t chunk, n_chunks;
    live_chunks(n_words, n_gl, target, kGlMaxWords, min_words, &chunk, &n_chunks);
    const int64_t total = uniform64(n_gl * n_chunks);
    int64_t u = blockIdx.x;
    while (u < total) {
        glist_unit(bv, pitch, n_words, rows, m, group_start, cum_pairs, ws, n_tiled, n_heavy, n_light, chunk, n_chunks,
                   counts, u, buf, col_id);
        // Every thread has passed the unit's barriers, so s_next is free; the barrier below also keeps the next
        // unit's writes of col_id and buf behind this unit's last reads.
        if (threadIdx.x == 0) s_next = (int64_t)gridDim.x + (int64_t)atomicAdd((unsigned long long*)&gl_cnt[4], 1ull);
        __syncthreads();
        u = uniform64(s_next);
    }
}

// ── the K=2 pair space (a single-group level) as a Gram on the matrix cores ─────────────────────────────
//
// Each bit of a column becomes one fp4 (e2m1) operand element, 0 or 1.0 (nibble 0x2), and
// v_mfma_scale_f32_32x32x64_f8f6f4 (unit scales) multiplies 32 columns by 32 columns over 64 rows: one bitvector
// word per column per instruction. Operand lane l carries column l % 32 and half l / 32 of the word (32 rows as
// 32 nibbles in four registers; the instruction reads four of the operand's eight for fp4). Which nibble holds
// which row does not matter as long as A and B place rows alike, since the instruction sums products of
// matching elements: register d of the half-word h takes bits d, d + 4, ..., d + 28 of h, each moved to bit 1 of
// its nibble -- one shift and one AND per register. Each product is 0 or 1 and fp32 adds integers below 2^24
// exactly: a word chunk holds fewer than 2^24 rows, and the per-chunk counts leave as int32 with one atomic per
// pair.
typedef float v16f __attribute__((ext_vector_type(16)));
typedef int v8i __attribute__((ext_vector_type(8)));

// The fp4 operand of a half-word h (32 rows): nibble n of register d is 0x2 where bit 4n + d of h is set.
__device__ __forceinline__ v8i fp4_operand(uint32_t h) {
    v8i v;
    v[0] = (int)((h << 1) & 0x22222222u);
    v[1] = (int)(h & 0x22222222u);
    v[2] = (int)((h >> 1) & 0x22222222u);
    v[3] = (int)((h >> 2) & 0x22222222u);
    v[4] = v[5] = v[6] = v[7] = 0;
    return v;
}

// A workgroup of 8 waves covers a tile-pair of kGT x kGT columns over a chunk of words, in steps of kGW words.
// Per step all 512 threads stage the raw words of the tile-pair's 2 kGT columns for the next step in LDS (thread t
// loads column t's kGW consecutive words; [column][word] as 32-bit halves, a column's row 33 dwords long so that
// the 32 columns of an operand read fall in 32 distinct banks), while each wave runs its 2 x 4 block of MFMAs (64 A
// columns x 128 B columns, fp32 accumulators) on the current step: per word it reads its six half-words from LDS
// and expands them in registers. The two LDS buffers alternate; one barrier per step. (Measured against 4 waves
// with 4 x 4 blocks, one wave per SIMD and accumulators in AGPRs: wide_vocab_2m K=2 2.92 against 3.54 ms.)
constexpr int kGT = 256;                      // columns per tile (both sides)
constexpr int kGW = 16;                       // words per step
constexpr int kGRowB = 8 * kGW + 4;           // LDS bytes per column per buffer (132: 33 banks apart)
constexpr int kGBuf = 2 * kGT * kGRowB;       // 67,584 B per buffer
constexpr int kGThreads = 512;
constexpr int64_t kGMaxChunkWords = ((1ll << 24) - 1) / 64 / kGW * kGW;  // rows of a chunk stay below 2^24

// The grid covers the tile-pairs of an upper bound of the columns (n_tp of nt_ub tiles); the exact count s is the
// level's one group's size, read on the device, and the tile-pairs past it return at once.
__global__ __launch_bounds__(kGThreads) void k_gram_mfma(const uint64_t* __restrict__ bv, int64_t pitch,
                                                         int64_t n_words, int64_t chunk_words,
                                                         const int32_t* __restrict__ cols,
                                                         const int64_t* __restrict__ group_start, int64_t nt_ub,
                                                         int64_t n_tp, int64_t n_chunks, int32_t* __restrict__ counts,
                                                         int32_t layout) {
    __shared__ __align__(16) uint8_t ops[2 * kGBuf];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int64_t chunk_id, tp;
    if (!xcd_unit(blockIdx.x % kXcds, blockIdx.x / kXcds, n_tp, n_chunks, &chunk_id, &tp)) return;
    const int64_t s = uniform64(group_start[1] - group_start[0]);
    int64_t ta, tb1;
    etm::lex_pair_decode(uniform64(tp), nt_ub + 1, &ta, &tb1);
    if (uniform64(tb1) - 1 >= ceil_div(s, kGT)) return;
    ta = uniform64(ta);
    const int64_t tb = uniform64(tb1) - 1;
    const int64_t w_begin = uniform64(chunk_id) * chunk_words;
    const int64_t w_end = min(n_words, w_begin + chunk_words);
    const int64_t n_steps = ceil_div(w_end - w_begin, kGW);
    // This thread's staging column: A columns 0..255, then B.
    const int64_t mycol = tid < kGT ? ta * kGT + tid : tb * kGT + (tid - kGT);
    const int64_t base = mycol < s ? (int64_t)cols[mycol] * pitch : -1;
    uint64_t next[kGW];
    auto load = [&](int64_t t) {
        const int64_t w0 = w_begin + t * kGW;
#pragma unroll
        for (int w = 0; w < kGW; ++w) next[w] = (base >= 0 && w0 + w < w_end) ? bv[base + w0 + w] : 0ull;
    };
    auto stage = [&](int64_t t) {
        uint32_t* row = reinterpret_cast<uint32_t*>(ops + (t & 1) * kGBuf + tid * kGRowB);
#pragma unroll
        for (int w = 0; w < kGW; ++w) {
            row[2 * w] = (uint32_t)next[w];
            row[2 * w + 1] = (uint32_t)(next[w] >> 32);
        }
    };
    load(0);
    stage(0);
    // Wave (cm, cn) owns A columns cm * 64 .. + 63 and B columns cn * 128 .. + 127 of the tile-pair.
    const int cm = wave >> 1, cn = wave & 1;
    v16f acc[2][4];
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 4; ++ni)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[mi][ni][r] = 0.0f;
    const uint32_t a_off = (uint32_t)(cm * 64 + (lane & 31)) * kGRowB + 4 * (lane >> 5);
    const uint32_t b_off = (uint32_t)(kGT + cn * 128 + (lane & 31)) * kGRowB + 4 * (lane >> 5);
    __syncthreads();
    for (int64_t t = 0; t < n_steps; ++t) {
        const bool more = t + 1 < n_steps;
        if (more) load(t + 1);
        const uint8_t* buf = ops + (t & 1) * kGBuf;
#pragma unroll 4
        for (int w = 0; w < kGW; ++w) {
            v8i a[2], b[4];
#pragma unroll
            for (int mi = 0; mi < 2; ++mi)
                a[mi] = fp4_operand(*reinterpret_cast<const uint32_t*>(buf + a_off + mi * 32 * kGRowB + 8 * w));
#pragma unroll
            for (int ni = 0; ni < 4; ++ni)
                b[ni] = fp4_operand(*reinterpret_cast<const uint32_t*>(buf + b_off + ni * 32 * kGRowB + 8 * w));
#pragma unroll
            for (int mi = 0; mi < 2; ++mi)
#pragma unroll
                for (int ni = 0; ni < 4; ++ni)
                    acc[mi][ni] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a[mi], b[ni], acc[mi][ni], 4, 4, 0,
                                                                                 127, 0, 127);
        }
        if (more) stage(t + 1);
        __syncthreads();
    }
#pragma unroll
    for (int mi = 0; mi < 2; ++mi) {
#pragma unroll
        for (int ni = 0; ni < 4; ++ni) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int64_t i = ta * kGT + cm * 64 + mi * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                const int64_t j = tb * kGT + cn * 128 + ni * 32 + (lane & 31);
                const int32_t c = (int32_t)acc[mi][ni][r];
                if (i < j && j < s && c) atomicAdd(&counts[pair_slot(i, j, s, layout)], c);
            }
        }
    }
}


// Word chunk of the Gram: the workgroups (tile-pairs x chunks) fill whole rounds of the resident slots (one
// 132 KiB-LDS workgroup per CU) as closely as possible over 2..8 rounds, the chunks per tile-pair rounded down to
// a multiple of the 8 XCDs where at least 8 (xcd_unit then keeps each XCD on its own chunks), no chunk shorter
// than four steps and none of 2^24 rows or more (the fp32 accumulators' exact range). A last round a third full
// costs the launch a third of a round.
int64_t gram_chunk(int64_t n_words, int64_t n_tp) {
    constexpr int64_t kSlots = 256;
    const int64_t max_chunks = std::max<int64_t>(1, n_words / (4 * kGW));
    const int64_t min_chunks = ceil_div(n_words, kGMaxChunkWords);
    int64_t best = 1;
    double best_fill = -1.0;
    for (int64_t rounds = 2; rounds <= 8; ++rounds) {
        int64_t ch = std::max<int64_t>(1, rounds * kSlots / std::max<int64_t>(n_tp, 1));
        if (ch >= kXcds) ch = ch / kXcds * kXcds;
        ch = std::max(std::min(ch, max_chunks), min_chunks);
        const int64_t units = ch * n_tp;
        const double fill = (double)units / (double)(ceil_div(units, kSlots) * kSlots);
        if (fill > best_fill + 0.02) {
            best_fill = fill;
            best = ch;
        }
    }
    return ceil_div(ceil_div(n_words, best), kGW) * kGW;
}

// One wave per (itemset, word chunk): AND of the itemset's columns, popcount, wave sum, one 64-bit add.
__global__ __launch_bounds__(256) void k_count_itemsets(const uint64_t* __restrict__ bv, int64_t pitch,
                                                         int64_t n_words, int64_t chunk_words, int64_t n_chunks,
                                                         const int32_t* __restrict__ items,
                                                         const int64_t* __restrict__ offsets, int64_t n_sets,
                                                         int64_t* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int64_t item = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (item >= n_sets * n_chunks) return;
    const int64_t x = item / n_chunks, ch = item % n_chunks;
    const int64_t a = offsets[x], b = offsets[x + 1];
    const int64_t w_begin = ch * chunk_words;
    const int64_t w_end = min(n_words, w_begin + chunk_words);
    int64_t sum = 0;
    for (int64_t w = w_begin + lane; w < w_end; w += 64) {
        uint64_t v = ~0ull;
        for (int64_t e = a; e < b; ++e) v &= bv[(int64_t)items[e] * pitch + w];
        sum += __popcll(v);
    }
    sum = etm::wave_sum(sum);
    if (lane == 0 && sum) atomicAdd((unsigned long long*)&out[x], (unsigned long long)sum);
}

__global__ void k_zero_i64n(int64_t* p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}


template <int S>
hipError_t launch_small(const uint64_t* bv, int64_t pitch, int64_t n_words, int64_t chunk, const int32_t* rows,
                        int32_t m, const int64_t* gs, const int64_t* cp, const int32_t* list, int64_t n_list,
                        const uint8_t* status, int32_t* counts, int32_t layout, const uint64_t* occ,
                        int64_t occ_pitch, hipStream_t s) {
    const int64_t n_chunks = ceil_div(n_words, chunk);
    const int64_t blocks = kXcds * ceil_div(xcd_units(n_list, n_chunks), 4);
    if (blocks > 0x7FFFFFFF) return hipErrorInvalidConfiguration;
    // 16-byte loads where every column and chunk start is 16-byte aligned and a lane has at least two words.
    const bool vec = SmallUnroll<S>::U >= 2 && pitch % 2 == 0 && chunk % 2 == 0 && ((uintptr_t)bv & 15) == 0;
    if (occ)
        k_count_small<S, true, false><<<(unsigned)blocks, 256, 0, s>>>(bv, pitch, n_words, chunk, n_chunks, rows, m,
                                                                       gs, cp, list, n_list, status, counts, layout,
                                                                       occ, occ_pitch);
    else if constexpr (SmallUnroll<S>::U >= 2) {
        if (vec)
            k_count_small<S, false, true><<<(unsigned)blocks, 256, 0, s>>>(bv, pitch, n_words, chunk, n_chunks, rows,
                                                                           m, gs, cp, list, n_list, status, counts,
                                                                           layout, nullptr, 0);
        else
            k_count_small<S, false, false><<<(unsigned)blocks, 256, 0, s>>>(bv, pitch, n_words, chunk, n_chunks, rows,
                                                                            m, gs, cp, list, n_list, status, counts,
                                                                            layout, nullptr, 0);
    } else
        k_count_small<S, false, false><<<(unsigned)blocks, 256, 0, s>>>(bv, pitch, n_words, chunk, n_chunks, rows, m,
                                                                        gs, cp, list, n_list, status, counts, layout,
                                                                        nullptr, 0);
    return hipGetLastError();
}

constexpr int64_t kLiveCounters = 16;  // int64: [0..3] live groups per small class, [4] live tile-pairs,
                                       // [5] sparse tile-pairs, [8..12] the group-listed mode's counters

// Units one XCD can be handed by xcd_unit when at most `ub` units are live and live_chunks cuts each into
// n_chunks <= max(1, ceil(target / n) rounded up to a multiple of 8, capped) chunks: with n_chunks >= 8 an XCD
// takes n x ceil(n_chunks / 8) <= target / 8 + 2n units, with fewer chunks n x n_chunks / 8 < n.
int64_t live_units_per_xcd(int64_t ub, int64_t target) { return ceil_div(target, kXcds) + 2 * ub + 1; }

// Waves (or workgroups) per XCD of a live launch: the units' upper bound, or, when that bound exceeds 6x the
// target's share (levels of very many groups: most deep levels, which hold few or no live units), 1.5x the
// target's share with each wave taking every stride-th unit of its XCD (*loop). The looping form holds more
// registers (the compiler keeps the unit's set-up live across the loop), so bounds below that use the one-unit
// form.
int64_t live_grid_per_xcd(int64_t ub, int64_t target, bool* loop) {
    const int64_t units = live_units_per_xcd(ub, target), cap = 3 * ceil_div(target, 2 * kXcds);
    *loop = units > 4 * cap;
    return *loop ? cap : units;
}

// ── batches of jobs in lockstep rounds (batch.h) ──────────────────────────

using etm::batch::Job;
using etm::batch::Queue;

// Units a job's round aims for per queue (small groups, tile-pairs), and the persistent grids that drain the
// queues: every wave (workgroup) takes units until the queue's index passes its total, so these fixed shapes
// serve any mix of jobs.
constexpr int64_t kBatchSmallTarget = 2048, kBatchTiledTarget = 2048;
constexpr int kBatchSmallBlocks = 1024, kBatchTiledBlocks = 768;  // the tiled kernel: three workgroups a CU

// One workgroup, a thread per job: zeroes the job's round tallies, sizes its hash (a power of two >= 2 x the
// previous level's rows) and its word chunks, and lays out the round's units per queue (a small class's groups
// or the tile-pairs, x chunks, chunk-major) as exclusive prefix sums over the jobs.
__global__ __launch_bounds__(256) void k_batch_plan(Job* __restrict__ jobs, int n_jobs, Queue* __restrict__ q) {
    __shared__ int64_t lds[256 / etm::kWave + 1];
    const int j = threadIdx.x;
    int64_t units[etm::batch::kQueues] = {0, 0, 0, 0};
    int64_t act = 0;
    if (j < n_jobs) {
        Job& J = jobs[j];
        for (int t = 0; t < 4; ++t) J.tally[t] = 0;
        J.new_n = 0;
        if (J.state == etm::batch::kActive) {
            act = 1;
            const int64_t* T = J.tot[J.buf];
            int64_t size = 64;
            while (size < 2 * J.prev_n) size <<= 1;
            J.table_mask = size - 1;
            const int64_t n_small = T[8] + T[9] + T[10];
            J.s_chunk = pick_chunk(J.n_words, n_small, kBatchSmallTarget, 64, kLiveSmallMin);
            J.s_nch = ceil_div(J.n_words, J.s_chunk);
            J.t_chunk = pick_chunk(J.n_words, T[2], kBatchTiledTarget, kKW, kLiveTiledMin);
            J.t_nch = ceil_div(J.n_words, J.t_chunk);
            for (int c = 0; c < etm::batch::kSmallQueues; ++c) units[c] = T[8 + c] * J.s_nch;
            units[etm::batch::kTiledQueue] = T[2] * J.t_nch;
        }
    }
    for (int c = 0; c < etm::batch::kQueues; ++c) {
        int64_t total = 0;
        const int64_t p = etm::block_exclusive_scan<256>(units[c], lds, &total);
        if (j < n_jobs) q->pre[c][j] = p;
        if (j == 0) {
            q->pre[c][n_jobs] = total;
            q->head[c] = 0;
        }
    }
    const int64_t n_act = etm::block_sum<256>(act, lds);
    if (j == 0) {
        q->active = n_act;
        q->active_after = 0;
    }
}

// The job holding unit u: pre[j] <= u < pre[j + 1] (pre[0] = 0 <= u < pre[n_jobs]; jobs without units are passed
// over).
// The job of unit u (wave-uniform for a wave-uniform u; made so explicitly, so that the job record's fields are
// read by scalar loads into scalar registers).
__device__ __forceinline__ int batch_job_of(const int64_t* __restrict__ pre, int n_jobs, int64_t u) {
    int lo = 0, hi = n_jobs;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (pre[mid] <= u) lo = mid;
        else hi = mid;
    }
    return __builtin_amdgcn_readfirstlane(lo);
}

// The small groups of size class C (S = 2, 4, 8 suffixes at most) of every job: each wave takes the next unit
// index of the class's queue (one atomic add by lane 0) until the index passes the total, and counts that
// (group, word chunk) with the wave-per-group body.
template <int S, int C>
__global__ __launch_bounds__(256) void k_batch_count_small(const Job* __restrict__ jobs, int n_jobs,
                                                            Queue* __restrict__ q) {
    const int lane = threadIdx.x & 63;
    const int64_t* pre = q->pre[C];
    const int64_t total = pre[n_jobs];
    // The first unit of every wave is its own index; the queue's head starts past them (one device-scope atomic
    // per further unit, none at all when the round has no more units than waves).
    const int64_t waves = (int64_t)gridDim.x * 4;
    int64_t u = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    for (;; u = -1) {
        if (u < 0) {
            if (total <= waves) return;
            unsigned long long got = 0;
            if (lane == 0) got = atomicAdd(&q->head[C], 1ull);
            u = uniform64((int64_t)got + waves);
        }
        if (u >= total) return;
        const int j = batch_job_of(pre, n_jobs, u);
        const Job& J = jobs[j];
        const int64_t r = u - pre[j];
        const int b = J.buf;
        const int64_t* T = J.tot[b];
        const int64_t nc = T[8 + C];
        int64_t start = 0;
        for (int t = 0; t < C; ++t) start += T[8 + t];
        small_unit<S, false, true>(J.bv, J.pitch, J.n_words, J.s_chunk, J.prev_rows, J.k - 1, J.gs[b], J.cp[b],
                                   J.status, J.counts, ETM_LAYOUT_LEX, nullptr, 0,
                                   uniform64(J.list[b][start + r % nc]), uniform64(r / nc));
    }
}

// Tile-pair units of every job (groups of more than 8 suffixes, and K=2's pair space): each workgroup takes the
// next unit index from the tiled queue and counts that (tile-pair, word chunk) with the tiled body.
__global__ __launch_bounds__(256) void k_batch_count_tiled(const Job* __restrict__ jobs, int n_jobs,
                                                            Queue* __restrict__ q) {
    __shared__ TiledShared sh;
    __shared__ unsigned long long s_u;
    const int64_t* pre = q->pre[etm::batch::kTiledQueue];
    const int64_t total = pre[n_jobs];
    // As the small-group queues: the workgroup's own index first, the atomic head past the grid after that.
    const int64_t blocks = gridDim.x;
    for (int64_t first = 1;; first = 0) {
        int64_t u = blockIdx.x;
        if (!first) {
            if (total <= blocks) return;
            if (threadIdx.x == 0) s_u = atomicAdd(&q->head[etm::batch::kTiledQueue], 1ull);
            __syncthreads();
            u = uniform64((int64_t)s_u) + blocks;
            __syncthreads();  // also: every thread is past the previous unit before this one's set-up
        }
        if (u >= total) return;
        const int j = batch_job_of(pre, n_jobs, u);
        const Job& J = jobs[j];
        const int64_t r = u - pre[j];
        const int b = J.buf;
        const int64_t n_tp = J.tot[b][2];
        tiled_unit(sh, J.bv, J.pitch, J.n_words, J.t_chunk, J.prev_rows, J.k - 1, J.gs[b], J.cp[b], J.ct[b],
                   J.tot[b][0], J.status, J.counts, ETM_LAYOUT_LEX, uniform64(r % n_tp), uniform64(r / n_tp));
    }
}

}  // namespace

namespace etm {
namespace detail {

hipError_t launch_batch_plan(etm::batch::Job* jobs, int n_jobs, etm::batch::Queue* q, hipStream_t s) {
    k_batch_plan<<<1, 256, 0, s>>>(jobs, n_jobs, q);
    return hipGetLastError();
}

hipError_t launch_batch_count(const etm::batch::Job* jobs, int n_jobs, etm::batch::Queue* q, hipStream_t s) {
    k_batch_count_small<2, 0><<<kBatchSmallBlocks, 256, 0, s>>>(jobs, n_jobs, q);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_batch_count_small<4, 1><<<kBatchSmallBlocks, 256, 0, s>>>(jobs, n_jobs, q);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_batch_count_small<8, 2><<<kBatchSmallBlocks, 256, 0, s>>>(jobs, n_jobs, q);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_batch_count_tiled<<<kBatchTiledBlocks, 256, 0, s>>>(jobs, n_jobs, q);
    return hipGetLastError();
}

}  // namespace detail
}  // namespace etm

// Workspace layout: kLiveCounters int64, the small list (n_small int32), the dense tile-pair list (n_tile_pairs
// int32), then, 8-byte aligned, the sparse tile-pair entries (n_tile_pairs SparseTP).
int64_t ws_sparse_offset(int64_t n_small, int64_t n_tile_pairs) {
    return (kLiveCounters * 8 + (n_small + n_tile_pairs) * 4 + 7) / 8 * 8;
}

extern "C" size_t etm_count_level_ws(int64_t n_small, int64_t n_tile_pairs) {
    n_small = std::max<int64_t>(n_small, 0);
    n_tile_pairs = std::max<int64_t>(n_tile_pairs, 0);
    return (size_t)(ws_sparse_offset(n_small, n_tile_pairs) + n_tile_pairs * (int64_t)sizeof(SparseTP));
}

namespace {

// The counted level through its live units: lists of the small-class groups and tile-pairs with a counted pair
// (k_live_units), then one launch per small class and one for the tiles over those lists. d_counters (8 int64,
// zero on entry) may be given by the caller; else the workspace's are zeroed first.
int live_count_impl(const uint64_t* d_bv, int64_t pitch, int64_t n_words, const int32_t* d_rows, int32_t m,
                    const int64_t* d_group_start, const int64_t* d_cum_pairs, const int64_t* d_cum_tiles,
                    int64_t n_groups, int64_t n_tile_pairs, const int32_t* d_small_list, const int64_t* cls,
                     const uint8_t* d_status, int32_t* d_counts, int32_t layout, void* d_ws, int64_t* d_counters,
                     const GlistWs* gl, int64_t n_tiled, hipStream_t s) {
    int64_t* cnt = d_counters != nullptr ? d_counters : (int64_t*)d_ws;
    int32_t* live_small = (int32_t*)((int64_t*)d_ws + kLiveCounters);
    ClassBounds cb{};
    for (int c = 0; c < 4; ++c) cb.start[c + 1] = cb.start[c] + cls[c];
    const int64_t n_small = cb.start[4];
    int32_t* live_tp = live_small + n_small;
    SparseTP* sparse = (SparseTP*)((char*)d_ws + ws_sparse_offset(n_small, n_tile_pairs));
    if (d_counters == nullptr) {
        k_zero_i64n<<<1, 64, 0, s>>>(cnt, kLiveCounters);
        ETM_LAUNCHED("k_zero_i64n");
    }
    const bool glist = gl != nullptr && n_tiled > 0;
    if (glist) {
        ETM_REQUIRE(n_tiled <= 0x7FFFFFFF, ETM_EOVERFLOW, "etm_count_level: %lld tiled groups exceed the grid",
                    (long long)n_tiled);
        k_gl_scan<<<(unsigned)n_tiled, kGlScanThreads, 0, s>>>(d_small_list + n_small, n_tiled, d_rows, m, d_group_start,
                                                    d_cum_pairs, d_status, *gl, cnt + 8);
        ETM_LAUNCHED("k_gl_scan");
    }
    const int64_t small_blocks = std::min<int64_t>(ceil_div(n_small, 256), 2048);
    const int64_t tile_blocks = std::min<int64_t>(ceil_div(n_tile_pairs, 4), 8192);
    if (small_blocks + tile_blocks > 0) {
        k_live_units<<<(unsigned)(small_blocks + tile_blocks), 256, 0, s>>>(
            d_small_list, n_small, cb, small_blocks, d_group_start, d_cum_pairs, d_cum_tiles, n_groups, n_tile_pairs,
            d_status, live_small, live_tp, sparse, cnt, glist ? gl->flag : nullptr);
        ETM_LAUNCHED("k_live_units");
    }
    const bool vec = pitch % 2 == 0 && ((uintptr_t)d_bv & 15) == 0;
    for (int c = 0; c < 4; ++c) {
        if (cls[c] == 0) continue;
        bool loop = false;
        const int64_t blocks = kXcds * ceil_div(live_grid_per_xcd(cls[c], kLiveSmallTarget, &loop), 4);
        ETM_REQUIRE(blocks <= 0x7FFFFFFF, ETM_EOVERFLOW, "etm_count_level: %lld workgroups exceed the grid",
                    (long long)blocks);
        const unsigned grid = (unsigned)blocks;
        const int32_t* list = live_small + cb.start[c];
#define ETM_SMALL_LIVE_L(S, V, L)                                                                                   \
    k_count_small_live<S, V, L><<<grid, 256, 0, s>>>(d_bv, pitch, n_words, d_rows, m, d_group_start, d_cum_pairs,    \
                                                     list, cnt + c, d_status, d_counts, layout, kLiveSmallTarget,   \
                                                     kLiveSmallMin)
#define ETM_SMALL_LIVE(S, V)                 \
    do {                                         \
        if (loop) ETM_SMALL_LIVE_L(S, V, true);  \
        else ETM_SMALL_LIVE_L(S, V, false);      \
    } while (0)
        switch (c) {
            case 0:
                if (vec) ETM_SMALL_LIVE(2, true);
                else ETM_SMALL_LIVE(2, false);
                break;
            case 1:
                if (vec) ETM_SMALL_LIVE(4, true);
                else ETM_SMALL_LIVE(4, false);
                break;
            case 2:
                if (vec) ETM_SMALL_LIVE(8, true);
                else ETM_SMALL_LIVE(8, false);
                break;
            default: ETM_SMALL_LIVE(16, false); break;
        }
#undef ETM_SMALL_LIVE
#undef ETM_SMALL_LIVE_L
        ETM_LAUNCHED("k_count_small_live");
    }
    if (n_tile_pairs > 0) {
        bool loop = false;
        const int64_t blocks = kXcds * live_grid_per_xcd(n_tile_pairs, kLiveTiledTarget, &loop);
        ETM_REQUIRE(blocks <= 0x7FFFFFFF, ETM_EOVERFLOW, "etm_count_level: %lld workgroups exceed the grid",
                    (long long)blocks);
#define ETM_TILED_LIVE(L)                                                                                          \
    k_count_tiled_live<L><<<(unsigned)blocks, 256, 0, s>>>(d_bv, pitch, n_words, d_rows, m, d_group_start,         \
                                                           d_cum_pairs, d_cum_tiles, n_groups, live_tp, cnt + 4,   \
                                                           d_status, d_counts, layout, kLiveTiledTarget,           \
                                                           kLiveTiledMin)
        if (loop) ETM_TILED_LIVE(true);
        else ETM_TILED_LIVE(false);
#undef ETM_TILED_LIVE
        ETM_LAUNCHED("k_count_tiled_live");
        const int64_t sblocks = kXcds * ceil_div(live_grid_per_xcd(n_tile_pairs, kLiveSmallTarget, &loop), 4);
        ETM_REQUIRE(sblocks <= 0x7FFFFFFF, ETM_EOVERFLOW, "etm_count_level: %lld workgroups exceed the grid",
                    (long long)sblocks);
#define ETM_SPARSE_LIVE(V, L)                                                                                       \
    k_count_sparse_live<V, L><<<(unsigned)sblocks, 256, 0, s>>>(d_bv, pitch, n_words, d_rows, m, d_group_start,     \
                                                                d_cum_pairs, d_cum_tiles, n_groups, sparse, cnt + 5, \
                                                                d_counts, kLiveSmallTarget, kLiveSmallMin)
        if (vec && loop) ETM_SPARSE_LIVE(true, true);
        else if (vec) ETM_SPARSE_LIVE(true, false);
        else if (loop) ETM_SPARSE_LIVE(false, true);
        else ETM_SPARSE_LIVE(false, false);
#undef ETM_SPARSE_LIVE
        ETM_LAUNCHED("k_count_sparse_live");
    }
    if (glist) {
        const int64_t blocks = std::min<int64_t>(kXcds * live_units_per_xcd(n_tiled, kLiveTiledTarget), kGlGrid);
        k_count_glist<<<(unsigned)blocks, kGlThreads, 0, s>>>(d_bv, pitch, n_words, d_rows, m, d_group_start, d_cum_pairs,
                                                       *gl, n_tiled, cnt + 8, d_counts, kLiveTiledTarget,
                                                       kLiveTiledMin);
        ETM_LAUNCHED("k_count_glist");
    }
    return ETM_OK;
}

}  // namespace

extern "C" int etm_count_level(const uint64_t* d_bv, int64_t pitch, int64_t n_words, const int32_t* d_rows, int32_t m,
                               const int64_t* d_group_start, const int64_t* d_cum_pairs, const int64_t* d_cum_tiles,
                               int64_t n_groups, int64_t n_tile_pairs, const int32_t* d_small_list,
                               const int64_t* h_class_counts, const uint8_t* d_status, int32_t* d_counts,
                               int32_t layout, int32_t word_chunk, int32_t kernel, const uint64_t* d_occ,
                               int64_t occ_pitch, void* d_ws, size_t ws_bytes, void* stream) {
    ETM_ARG(d_bv && d_rows && d_group_start && d_cum_pairs && d_counts, "etm_count_level: null pointer");
    ETM_ARG(d_occ == nullptr || occ_pitch >= ((n_words + 63) / 64 + 63) / 64 + 1,
            "etm_count_level: occ_pitch %lld too small for %lld words", (long long)occ_pitch, (long long)n_words);
    ETM_ARG(m >= 1 && m <= 62, "etm_count_level: 1 <= m <= 62 required, got %d", m);
    ETM_ARG(pitch >= n_words && n_words >= 0, "etm_count_level: pitch %lld < n_words %lld", (long long)pitch,
            (long long)n_words);
    ETM_ARG(layout == ETM_LAYOUT_LEX || (layout == ETM_LAYOUT_TRI && n_groups == 1),
            "etm_count_level: the TRI layout needs a single group");
    ETM_ARG(n_tile_pairs == 0 || d_cum_tiles != nullptr, "etm_count_level: null tile offsets");
    int64_t cls[4] = {0, 0, 0, 0};
    if (h_class_counts)
        for (int c = 0; c < 4; ++c) {
            ETM_ARG(h_class_counts[c] >= 0, "etm_count_level: negative class size");
            cls[c] = h_class_counts[c];
        }
    const int64_t n_small = cls[0] + cls[1] + cls[2] + cls[3];
    ETM_ARG(n_small == 0 || d_small_list != nullptr, "etm_count_level: null small-group list");
    ETM_ARG(d_ws == nullptr || ws_bytes >= etm_count_level_ws(n_small, n_tile_pairs),
            "etm_count_level: workspace %zu < %zu bytes", ws_bytes, etm_count_level_ws(n_small, n_tile_pairs));
    ETM_ARG(n_tile_pairs <= 0x7FFFFFFF && n_small <= 0x7FFFFFFF, "etm_count_level: %lld tile-pairs exceed int32",
            (long long)n_tile_pairs);
    if (n_words == 0) return ETM_OK;
    hipStream_t s = ETM_STREAM(stream);
    // With a workspace and a status array (and no occupancy skip) only the units holding a counted pair run.
    if (d_ws != nullptr && d_status != nullptr && d_occ == nullptr && word_chunk <= 0 && layout == ETM_LAYOUT_LEX)
        return live_count_impl(d_bv, pitch, n_words, d_rows, m, d_group_start, d_cum_pairs, d_cum_tiles, n_groups,
                               n_tile_pairs, d_small_list, cls, d_status, d_counts, layout, d_ws, nullptr, nullptr, 0,
                               s);
    int64_t base = 0;
    for (int c = 0; c < 4; ++c) {
        if (cls[c] == 0) continue;
        // About kSmallTarget waves per class and chunks of at least 256 words: classes of many groups keep long
        // chunks (fewer partial sums, chunk-major L2 reuse per XCD), classes of few groups get enough waves.
        int64_t chunk = word_chunk > 0 ? word_chunk : pick_chunk(n_words, cls[c], kSmallTarget, 64, 256);
        // The tile skip reads one 64-bit occupancy window per column: at most 64 tiles of 64 words per chunk.
        if (d_occ) chunk = std::min<int64_t>(ceil_div(chunk, 64) * 64, 4096);
        hipError_t e;
        const int32_t* list = d_small_list + base;
        switch (c) {
            case 0: e = launch_small<2>(d_bv, pitch, n_words, chunk, d_rows, m, d_group_start, d_cum_pairs, list,
                                        cls[c], d_status, d_counts, layout, d_occ, occ_pitch, s); break;
            case 1: e = launch_small<4>(d_bv, pitch, n_words, chunk, d_rows, m, d_group_start, d_cum_pairs, list,
                                        cls[c], d_status, d_counts, layout, d_occ, occ_pitch, s); break;
            case 2: e = launch_small<8>(d_bv, pitch, n_words, chunk, d_rows, m, d_group_start, d_cum_pairs, list,
                                        cls[c], d_status, d_counts, layout, d_occ, occ_pitch, s); break;
            default: e = launch_small<16>(d_bv, pitch, n_words, chunk, d_rows, m, d_group_start, d_cum_pairs, list,
                                          cls[c], d_status, d_counts, layout, d_occ, occ_pitch, s); break;
        }
        ETM_REQUIRE(e == hipSuccess, ETM_EHIP, "launch of k_count_small failed: %s", hipGetErrorString(e));
        base += cls[c];
    }
    // The pair space of a single-group level of one-item rows (K=2) without a status array: the fp4 MFMA Gram
    // when asked for, or by default where it measured faster (gram_preferred), else the VALU tiled kernel.
    const bool gram_level = n_groups == 1 && m == 1 && d_status == nullptr;
    if (n_tile_pairs > 0 && gram_level && kernel != ETM_KERNEL_VALU) {
        // n_tile_pairs counts 64-column tiles, nt64 (nt64 + 1) / 2 of them: the columns are at most 64 nt64.
        int64_t nt64 = (int64_t)((std::sqrt(8.0 * (double)n_tile_pairs + 1.0) - 1.0) / 2.0);
        while (nt64 * (nt64 + 1) / 2 < n_tile_pairs) ++nt64;
        while (nt64 > 0 && (nt64 - 1) * nt64 / 2 >= n_tile_pairs) --nt64;
        const int64_t s_ub = 64 * nt64;
        if (kernel == ETM_KERNEL_MFMA || etm::detail::gram_preferred(s_ub, n_words)) {
            const int64_t nt = ceil_div(s_ub, kGT);
            const int64_t n_tp = nt * (nt + 1) / 2;
            const int64_t chunk = word_chunk > 0 ? std::min(ceil_div(word_chunk, kGW) * kGW, kGMaxChunkWords)
                                                 : gram_chunk(n_words, n_tp);
            const int64_t n_chunks = ceil_div(n_words, chunk);
            const int64_t blocks = kXcds * xcd_units(n_tp, n_chunks);
            ETM_REQUIRE(blocks <= 0x7FFFFFFF, ETM_EOVERFLOW, "etm_count_level: %lld workgroups exceed the grid",
                        (long long)blocks);
            k_gram_mfma<<<(unsigned)blocks, kGThreads, 0, s>>>(d_bv, pitch, n_words, chunk, d_rows, d_group_start, nt,
                                                               n_tp, n_chunks, d_counts, layout);
            ETM_LAUNCHED("k_gram_mfma");
            return ETM_OK;
        }
    }
    if (n_tile_pairs > 0) {
        ETM_REQUIRE(n_tile_pairs <= 0x7FFFFFFF, ETM_EOVERFLOW, "etm_count_level: %lld tile-pairs exceed the grid",
                    (long long)n_tile_pairs);
        const int64_t chunk = word_chunk > 0 ? ceil_div(word_chunk, kKW) * kKW
                                             : pick_chunk(n_words, n_tile_pairs, kTiledTarget, kKW, 128);
        const int64_t n_chunks = ceil_div(n_words, chunk);
        const int64_t blocks = kXcds * xcd_units(n_tile_pairs, n_chunks);
        ETM_REQUIRE(blocks <= 0x7FFFFFFF, ETM_EOVERFLOW, "etm_count_level: %lld workgroups exceed the grid",
                    (long long)blocks);
        k_count_tiled<<<(unsigned)blocks, 256, 0, s>>>(d_bv, pitch, n_words, chunk, d_rows, m, d_group_start,
                                                      d_cum_pairs, d_cum_tiles, n_groups, n_tile_pairs, n_chunks,
                                                      d_status, d_counts, layout);
        ETM_LAUNCHED("k_count_tiled");
    }
    return ETM_OK;
}

namespace {
// The group-listed workspace of a level: flags, group entries, the pair lists and the used columns.
struct GlistSizes {
    int64_t flag, groups, pairs, cols;
    GlistSizes(int64_t n_groups, int64_t n_tiled, int64_t n_cands, int64_t n_rows) {
        const int64_t pair_cap = std::min<int64_t>(n_cands, n_tiled * kGlPairsMax);
        const int64_t col_cap = std::min<int64_t>(n_rows, n_tiled * kGlColsMax);
        flag = (n_groups + 255) / 256 * 256;
        groups = n_tiled * (int64_t)sizeof(GlGroup);
        pairs = (pair_cap * 4 + 255) / 256 * 256;
        cols = col_cap * 4;
    }
    int64_t total() const { return flag + groups + 2 * pairs + cols; }
};
}  // namespace

size_t etm::detail::count_glist_ws(int64_t n_groups, int64_t n_tiled, int64_t n_cands, int64_t n_rows) {
    return n_tiled > 0 ? (size_t)GlistSizes(n_groups, n_tiled, n_cands, n_rows).total() : 0;
}

int etm::detail::count_level_live(const uint64_t* d_bv, int64_t pitch, int64_t n_words, const int32_t* d_rows,
                                  int32_t m, const int64_t* d_group_start, const int64_t* d_cum_pairs,
                                  const int64_t* d_cum_tiles, int64_t n_groups, int64_t n_tile_pairs,
                                  const int32_t* d_small_list, const int64_t* h_class_counts, const uint8_t* d_status,
                                  int32_t* d_counts, void* d_ws, int64_t* d_counters, int64_t n_tiled, int64_t n_cands,
                                  int64_t n_rows, void* d_gl_ws, void* stream) {
    ETM_ARG(d_bv && d_rows && d_group_start && d_cum_pairs && d_counts && d_status && d_ws && d_counters,
            "count_level_live: null pointer");
    ETM_ARG(pitch >= n_words && n_words >= 0 && m >= 1 && m <= 62, "count_level_live: bad sizes");
    ETM_ARG(n_tile_pairs <= 0x7FFFFFFF, "count_level_live: %lld tile-pairs exceed int32", (long long)n_tile_pairs);
    ETM_ARG(n_tiled == 0 || d_gl_ws != nullptr, "count_level_live: tiled groups need the group-listed workspace");
    if (n_words == 0) return ETM_OK;
    GlistWs gl{};
    if (n_tiled > 0) {
        const GlistSizes z(n_groups, n_tiled, n_cands, n_rows);
        char* p = (char*)d_gl_ws;
        gl.flag = (uint8_t*)p;
        gl.groups = (GlGroup*)(p + z.flag);
        gl.pair_slots = (uint32_t*)(p + z.flag + z.groups);
        gl.pair_ij = (uint32_t*)(p + z.flag + z.groups + z.pairs);
        gl.cols = (int32_t*)(p + z.flag + z.groups + 2 * z.pairs);
    }
    return live_count_impl(d_bv, pitch, n_words, d_rows, m, d_group_start, d_cum_pairs, d_cum_tiles, n_groups,
                           n_tile_pairs, d_small_list, h_class_counts, d_status, d_counts, ETM_LAYOUT_LEX, d_ws,
                           d_counters, n_tiled > 0 ? &gl : nullptr, n_tiled, ETM_STREAM(stream));
}

extern "C" int etm_count_itemsets(const uint64_t* d_bv, int64_t pitch, int64_t n_words, const int32_t* d_items,
                                  const int64_t* d_offsets, int64_t n_sets, int64_t* d_out, int32_t accumulate,
                                  void* stream) {
    ETM_ARG(d_bv && d_items && d_offsets && d_out, "etm_count_itemsets: null pointer");
    ETM_ARG(pitch >= n_words && n_words >= 0 && n_sets >= 0, "etm_count_itemsets: bad sizes");
    hipStream_t s = ETM_STREAM(stream);
    if (n_sets == 0) return ETM_OK;
    if (!accumulate) {
        k_zero_i64n<<<(unsigned)std::min<int64_t>(ceil_div(n_sets, 256), 4096), 256, 0, s>>>(d_out, n_sets);
        ETM_LAUNCHED("k_zero_i64n");
    }
    if (n_words == 0) return ETM_OK;
    const int64_t chunk = pick_chunk(n_words, n_sets, 16384, 64, 256);
    const int64_t n_chunks = ceil_div(n_words, chunk);
    const int64_t blocks = ceil_div(n_sets * n_chunks, 4);
    ETM_REQUIRE(blocks <= 0x7FFFFFFF, ETM_EOVERFLOW, "etm_count_itemsets: too many itemsets");
    k_count_itemsets<<<(unsigned)blocks, 256, 0, s>>>(d_bv, pitch, n_words, chunk, n_chunks, d_items, d_offsets,
                                                      n_sets, d_out);
    ETM_LAUNCHED("k_count_itemsets");
    return ETM_OK;
}
