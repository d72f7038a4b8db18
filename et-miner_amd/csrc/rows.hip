// Row-wise counting of the K=2 and K=3 levels: each transaction adds one to every candidate it contains,
// instead of every candidate ANDing its columns over all words. It pays where the pair space is wide and the
// rows are short (work = sum over rows of C(len, 2) or of its partner triples, against P x W candidate-words
// for the bitvector kernels).
//
// Rows input: the transactions as ascending K=1 positions (the index of each id's column in the K=1 level),
// built on the device from the id CSR: a per-row length pass (which also reports sum C(len, 2), the longest
// row, and whether every row's ids ascend strictly), a scan to row pointers, and a write pass. The row-wise
// kernels are exact only on rows of distinct ascending positions, so the caller uses them only when that
// report says every row ascends strictly.
//
// K=2: the level is one group holding every K=1 position; candidate (i, j), i < j, has the lexicographic index
// lex_pair_index(i, j, n1). K=3: candidates are numbered by the prefix groups of the K=2 level (group of prefix
// a, suffix indices i < j); a row counts (a, b, c) when (a, b) and (a, c) are frequent pairs, which the pair
// bitmap answers, and when the candidate was classified COUNT.
#include <algorithm>

#include "common.h"
#include "internal.h"
#include "scan.h"

namespace {

// ── rows as K=1 positions ───────────────────────────────────────────────

// Position of id v in the K=1 level, or -1: lut maps ids to columns, colpos columns to K=1 positions.
struct PosOf {
    int64_t lo, span;
    const int32_t* lut;
    const int32_t* colpos;
    __device__ int32_t operator()(int64_t id) const {
        const int64_t x = id - lo;
        if (x < 0 || x >= span) return -1;
        const int32_t c = lut[x];
        return c < 0 ? -1 : colpos[c];
    }
};

// One thread per row: its number of K=1 positions; per workgroup, the sum of C(len, 2), the longest row and
// whether some row's ids fail to ascend strictly, folded into stats[0..2].
template <typename Id>
__global__ __launch_bounds__(256) void k_rows_len(const int64_t* __restrict__ indptr, const Id* __restrict__ ids,
                                                   int64_t n_rows, PosOf pos_of, int32_t* __restrict__ len,
                                                   int64_t* __restrict__ stats) {
    __shared__ int64_t lds[5];
    int64_t pairs = 0, longest = 0, unsorted = 0;
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n_rows;
         r += (int64_t)gridDim.x * blockDim.x) {
        const int64_t a = indptr[r], b = indptr[r + 1];
        int64_t cnt = 0, prev = INT64_MIN;
        bool up = true;
        auto one = [&](int64_t v) {
            up = up && v > prev;
            prev = v;
            cnt += pos_of(v) >= 0;
        };
        int64_t e = a;
        for (; e + 4 <= b; e += 4) {
            const int64_t v0 = (int64_t)ids[e], v1 = (int64_t)ids[e + 1], v2 = (int64_t)ids[e + 2],
                          v3 = (int64_t)ids[e + 3];
            one(v0);
            one(v1);
            one(v2);
            one(v3);
        }
        for (; e < b; ++e) one((int64_t)ids[e]);
        len[r] = (int32_t)cnt;
        pairs += cnt * (cnt - 1) / 2;
        longest = cnt > longest ? cnt : longest;
        unsorted |= up ? 0 : 1;
    }
    pairs = etm::block_sum<256>(pairs, lds);
    unsorted = etm::block_sum<256>(unsorted, lds);
    longest = etm::wave_max(longest);
    if (threadIdx.x == 0) {
        if (pairs) atomicAdd((unsigned long long*)&stats[0], (unsigned long long)pairs);
        if (unsorted) atomicMax((long long*)&stats[2], 1ll);
    }
    if ((threadIdx.x & 63) == 0 && longest) atomicMax((long long*)&stats[1], (long long)longest);
}

struct RowLen {
    const int32_t* len;
    __device__ int64_t operator()(int64_t r) const { return len[r]; }
};

struct WriteRowPtr {
    int64_t* row_ptr;
    __device__ void operator()(int64_t r, int64_t, int64_t prefix) const { row_ptr[r] = prefix; }
};

// One thread per row: its K=1 positions, in order, at row_ptr[r].
template <typename Id>
__global__ __launch_bounds__(256) void k_rows_write(const int64_t* __restrict__ indptr, const Id* __restrict__ ids,
                                                     int64_t n_rows, PosOf pos_of,
                                                     const int64_t* __restrict__ row_ptr, int32_t* __restrict__ pos) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n_rows;
         r += (int64_t)gridDim.x * blockDim.x) {
        const int64_t a = indptr[r], b = indptr[r + 1];
        int64_t o = row_ptr[r];
        const int64_t end = row_ptr[r + 1];
        for (int64_t e = a; e < b && o < end; ++e) {
            const int32_t p = pos_of((int64_t)ids[e]);
            if (p >= 0) pos[o++] = p;
        }
    }
}

__global__ void k_zero4(int64_t* p) {
    if (threadIdx.x < 4) p[threadIdx.x] = 0;
}

__global__ void k_set_total(int64_t* row_ptr, int64_t n_rows, int64_t* stats) {
    if (threadIdx.x == 0) stats[3] = row_ptr[n_rows];
}

// ── K=2 ─────────────────────────────────────────────────────────────────

constexpr int kShortRow = 48;  // rows up to this many positions: a thread each; longer rows: a wave each

// A thread per short row: for each of its positions i, the pairs (i, j) with the later positions j are the
// consecutive candidates lex_pair_index(i, j, n1) = start(i) + j - i - 1.
__global__ __launch_bounds__(256) void k_count_k2_rows_short(const int64_t* __restrict__ row_ptr,
                                                              const int32_t* __restrict__ pos, int64_t n_rows,
                                                              int64_t n1, int32_t* __restrict__ counts) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n_rows;
         r += (int64_t)gridDim.x * blockDim.x) {
        const int64_t a = row_ptr[r], b = row_ptr[r + 1];
        if (b - a > kShortRow) continue;
        for (int64_t t = a; t + 1 < b; ++t) {
            const int64_t i = pos[t];
            const int64_t base = i * (2 * n1 - i - 1) / 2 - i - 1;  // + j = lex_pair_index(i, j, n1), j > i
            for (int64_t u = t + 1; u < b; ++u) {
                const int64_t j = pos[u];
                if (j > i && j < n1) atomicAdd(&counts[base + j], 1);  // rows that do not ascend add nothing
            }
        }
    }
}

// A wave per long row: lanes take its C(len, 2) pairs in lexicographic order.
__global__ __launch_bounds__(256) void k_count_k2_rows_long(const int64_t* __restrict__ row_ptr,
                                                             const int32_t* __restrict__ pos, int64_t n_rows,
                                                             int64_t n1, int32_t* __restrict__ counts) {
    const int lane = threadIdx.x & 63;
    const int64_t n_waves = (int64_t)gridDim.x * (blockDim.x / 64);
    for (int64_t r = (int64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6); r < n_rows; r += n_waves) {
        const int64_t a = row_ptr[r], L = row_ptr[r + 1] - a;
        if (L <= kShortRow) continue;
        const int64_t np = L * (L - 1) / 2;
        for (int64_t q = lane; q < np; q += 64) {
            int64_t t, u;
            etm::lex_pair_decode(q, L, &t, &u);
            const int64_t i = pos[a + t], j = pos[a + u];
            if (i >= 0 && i < j && j < n1) atomicAdd(&counts[etm::lex_pair_index(i, j, n1)], 1);
        }
    }
}

// ── K=3 ─────────────────────────────────────────────────────────────────

// The K=2 level as a pair bitmap over K=1 positions: bit v of row u is set for each pair (u, v).
__global__ __launch_bounds__(256) void k_pair_bits(const int32_t* __restrict__ rows2, int64_t n2,
                                                    const int32_t* __restrict__ colpos, int64_t n1, int64_t words,
                                                    uint64_t* __restrict__ bits) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n2; p += (int64_t)gridDim.x * blockDim.x) {
        const int64_t u = colpos[rows2[2 * p]], v = colpos[rows2[2 * p + 1]];
        if (u >= 0 && u < n1 && v >= 0 && v < n1)
            atomicOr((unsigned long long*)&bits[u * words + v / 64], 1ull << (v & 63));
    }
}

// rank[u][w] = set bits of bitmap row u before word w (so v's index among u's partners is
// rank[u][v/64] + the set bits below v in its word); gof[u] = the K=2 group whose prefix is u, or -1.
__global__ __launch_bounds__(256) void k_pair_rank(const uint64_t* __restrict__ bits, int64_t n1, int64_t words,
                                                    int32_t* __restrict__ rank, int32_t* __restrict__ gof) {
    for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < n1; u += (int64_t)gridDim.x * blockDim.x) {
        int32_t run = 0;
        for (int64_t w = 0; w < words; ++w) {
            rank[u * words + w] = run;
            run += __popcll(bits[u * words + w]);
        }
        gof[u] = -1;
    }
}

__global__ __launch_bounds__(256) void k_group_of(const int32_t* __restrict__ rows2,
                                                   const int64_t* __restrict__ group_start, int64_t n_groups,
                                                   const int32_t* __restrict__ colpos, int64_t n1,
                                                   int32_t* __restrict__ gof) {
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n_groups;
         g += (int64_t)gridDim.x * blockDim.x) {
        const int64_t u = colpos[rows2[2 * group_start[g]]];
        if (u >= 0 && u < n1) gof[u] = (int32_t)g;
    }
}

constexpr int kK3MaxRow = 64;

// A wave per row of 3 to 64 positions, lane t holding position t. Lane t lists the later positions of the
// row that pair with its own (bitmap), as their indices among its partners, in LDS; each two of them (b, c)
// form candidate (t's group, i_b, i_c), counted when classified COUNT (every candidate when status is null).
__global__ __launch_bounds__(256) void k_count_k3_rows(const int64_t* __restrict__ row_ptr,
                                                        const int32_t* __restrict__ pos, int64_t n_rows,
                                                        const uint64_t* __restrict__ bits,
                                                        const int32_t* __restrict__ rank, int64_t words,
                                                        int64_t n1, const int32_t* __restrict__ gof,
                                                        const int64_t* __restrict__ group_start,
                                                        const int64_t* __restrict__ cum_pairs,
                                                        const uint8_t* __restrict__ status,
                                                        int32_t* __restrict__ counts) {
    __shared__ uint16_t part[4][kK3MaxRow][kK3MaxRow];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t n_waves = (int64_t)gridDim.x * (blockDim.x / 64);
    for (int64_t r = (int64_t)blockIdx.x * (blockDim.x / 64) + wv; r < n_rows; r += n_waves) {
        const int64_t a = row_ptr[r];
        const int64_t len = row_ptr[r + 1] - a;
        if (len < 3 || len > kK3MaxRow) continue;  // longer rows: k_count_k3_rows_long
        const int L = (int)len;
        const int32_t me = lane < L ? pos[a + lane] : 0;
        const int32_t g = (lane < L - 2 && me >= 0 && me < n1) ? gof[me] : -1;
        int d = 0;
        for (int u = 1; u < L; ++u) {
            const int32_t v = __shfl(me, u, 64);
            if (g >= 0 && u > lane && v >= 0 && v < n1) {
                const uint64_t word = bits[(int64_t)me * words + v / 64];
                const uint64_t bit = 1ull << (v & 63);
                if (word & bit)
                    part[wv][lane][d++] =
                        (uint16_t)(rank[(int64_t)me * words + v / 64] + __popcll(word & (bit - 1)));
            }
        }
        // Each lane reads back only its own LDS row: no exchange between lanes.
        if (d >= 2) {
            const int64_t gs = group_start[g];
            const int64_t s = group_start[g + 1] - gs;
            const int64_t cb = cum_pairs[g];
            for (int x = 0; x + 1 < d; ++x) {
                const int64_t i = part[wv][lane][x];
                const int64_t row_base = cb + i * (2 * s - i - 1) / 2 - i - 1;
                for (int y = x + 1; y < d; ++y) {
                    const int64_t j = part[wv][lane][y];
                    if (j <= i) continue;  // rows that do not ascend add nothing
                    const int64_t c = row_base + j;
                    if (status == nullptr || status[c] == ETM_CAND_COUNT) atomicAdd(&counts[c], 1);
                }
            }
        }
    }
}

// Rows of more than 64 positions, a wave each: lane t takes positions t, t+64, ... as the prefix a, and walks
// the later positions for its partners b and, after each b, for the partners c after it.
__global__ __launch_bounds__(256) void k_count_k3_rows_long(const int64_t* __restrict__ row_ptr,
                                                             const int32_t* __restrict__ pos, int64_t n_rows,
                                                             const uint64_t* __restrict__ bits,
                                                             const int32_t* __restrict__ rank, int64_t words,
                                                             int64_t n1, const int32_t* __restrict__ gof,
                                                             const int64_t* __restrict__ group_start,
                                                             const int64_t* __restrict__ cum_pairs,
                                                             const uint8_t* __restrict__ status,
                                                             int32_t* __restrict__ counts) {
    const int lane = threadIdx.x & 63;
    const int64_t n_waves = (int64_t)gridDim.x * (blockDim.x / 64);
    auto partner = [&](int64_t u, int64_t v, int64_t* idx) {
        if (v < 0 || v >= n1) return false;
        const uint64_t word = bits[u * words + v / 64];
        const uint64_t bit = 1ull << (v & 63);
        if (!(word & bit)) return false;
        *idx = rank[u * words + v / 64] + __popcll(word & (bit - 1));
        return true;
    };
    for (int64_t r = (int64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6); r < n_rows; r += n_waves) {
        const int64_t a = row_ptr[r], L = row_ptr[r + 1] - a;
        if (L <= kK3MaxRow) continue;
        for (int64_t t = lane; t + 2 < L; t += 64) {
            const int64_t u = pos[a + t];
            if (u < 0 || u >= n1) continue;
            const int32_t g = gof[u];
            if (g < 0) continue;
            const int64_t gs = group_start[g];
            const int64_t s = group_start[g + 1] - gs;
            const int64_t cb = cum_pairs[g];
            for (int64_t b = t + 1; b + 1 < L; ++b) {
                int64_t i;
                if (!partner(u, pos[a + b], &i)) continue;
                const int64_t row_base = cb + i * (2 * s - i - 1) / 2 - i - 1;
                for (int64_t c = b + 1; c < L; ++c) {
                    int64_t j;
                    if (!partner(u, pos[a + c], &j) || j <= i) continue;
                    const int64_t cand = row_base + j;
                    if (status == nullptr || status[cand] == ETM_CAND_COUNT) atomicAdd(&counts[cand], 1);
                }
            }
        }
    }
}

int grid_for(int64_t n, int cap = 16384) {
    int64_t g = (n + 255) / 256;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

}  // namespace

extern "C" size_t etm_rows_ws(int64_t n_rows) {
    const size_t scan = etm::scan_ws_bytes(std::max<int64_t>(n_rows, 1));
    return ((scan + 255) & ~(size_t)255) + (size_t)std::max<int64_t>(n_rows, 1) * 4;
}

extern "C" int etm_rows_build(const int64_t* d_indptr, const void* d_ids, int id_bytes, int64_t n_rows, int64_t lo,
                              int64_t span, const int32_t* d_lut, const int32_t* d_colpos, int64_t* d_row_ptr,
                              int32_t* d_pos, int64_t* d_stats, void* d_ws, size_t ws_bytes, void* stream) {
    ETM_ARG(d_indptr && d_lut && d_colpos && d_row_ptr && d_pos && d_stats, "etm_rows_build: null pointer");
    ETM_ARG(id_bytes == 4 || id_bytes == 8, "etm_rows_build: id_bytes must be 4 or 8, got %d", id_bytes);
    ETM_ARG(n_rows >= 1 && span >= 1, "etm_rows_build: n_rows and span must be >= 1");
    ETM_ARG(d_ws && ws_bytes >= etm_rows_ws(n_rows), "etm_rows_build: workspace of %zu bytes < %zu", ws_bytes,
            etm_rows_ws(n_rows));
    hipStream_t s = ETM_STREAM(stream);
    const size_t scan = (etm::scan_ws_bytes(n_rows) + 255) & ~(size_t)255;
    int64_t* ws = (int64_t*)d_ws;
    int32_t* len = (int32_t*)((char*)d_ws + scan);
    PosOf pos_of{lo, span, d_lut, d_colpos};
    k_zero4<<<1, 64, 0, s>>>(d_stats);
    ETM_LAUNCHED("k_zero4");
    if (id_bytes == 4)
        k_rows_len<int32_t><<<grid_for(n_rows), 256, 0, s>>>(d_indptr, (const int32_t*)d_ids, n_rows, pos_of, len,
                                                             d_stats);
    else
        k_rows_len<int64_t><<<grid_for(n_rows), 256, 0, s>>>(d_indptr, (const int64_t*)d_ids, n_rows, pos_of, len,
                                                             d_stats);
    ETM_LAUNCHED("k_rows_len");
    RowLen f{len};
    WriteRowPtr w{d_row_ptr};
    ETM_HIP(etm::scan_prepare(f, n_rows, ws, &d_row_ptr[n_rows], s));
    ETM_HIP(etm::scan_finish(f, w, n_rows, ws, s));
    k_set_total<<<1, 64, 0, s>>>(d_row_ptr, n_rows, d_stats);
    ETM_LAUNCHED("k_set_total");
    if (id_bytes == 4)
        k_rows_write<int32_t><<<grid_for(n_rows), 256, 0, s>>>(d_indptr, (const int32_t*)d_ids, n_rows, pos_of,
                                                               d_row_ptr, d_pos);
    else
        k_rows_write<int64_t><<<grid_for(n_rows), 256, 0, s>>>(d_indptr, (const int64_t*)d_ids, n_rows, pos_of,
                                                               d_row_ptr, d_pos);
    ETM_LAUNCHED("k_rows_write");
    return ETM_OK;
}

extern "C" int etm_count_k2_rows(const int64_t* d_row_ptr, const int32_t* d_pos, int64_t n_rows, int64_t n1,
                                 int32_t* d_counts, void* stream) {
    ETM_ARG(d_row_ptr && d_pos && d_counts, "etm_count_k2_rows: null pointer");
    ETM_ARG(n_rows >= 0 && n1 >= 0, "etm_count_k2_rows: negative size");
    if (n_rows == 0 || n1 < 2) return ETM_OK;
    hipStream_t s = ETM_STREAM(stream);
    k_count_k2_rows_short<<<grid_for(n_rows), 256, 0, s>>>(d_row_ptr, d_pos, n_rows, n1, d_counts);
    ETM_LAUNCHED("k_count_k2_rows_short");
    k_count_k2_rows_long<<<grid_for(n_rows * 64, 4096), 256, 0, s>>>(d_row_ptr, d_pos, n_rows, n1, d_counts);
    ETM_LAUNCHED("k_count_k2_rows_long");
    return ETM_OK;
}

extern "C" int etm_k3_rows_prepare(const int32_t* d_rows2, int64_t n2, const int32_t* d_colpos, int64_t n1,
                                   const int64_t* d_group_start, int64_t n_groups, uint64_t* d_bits,
                                   int32_t* d_rank, int32_t* d_gof, void* stream) {
    ETM_ARG(d_rows2 && d_colpos && d_group_start && d_bits && d_rank && d_gof, "etm_k3_rows_prepare: null pointer");
    ETM_ARG(n2 >= 0 && n1 >= 1 && n_groups >= 0, "etm_k3_rows_prepare: bad sizes");
    ETM_ARG(n1 <= 65536, "etm_k3_rows_prepare: %lld K=1 positions exceed the 16-bit suffix index", (long long)n1);
    hipStream_t s = ETM_STREAM(stream);
    const int64_t words = (n1 + 63) / 64;
    ETM_HIP(hipMemsetAsync(d_bits, 0, (size_t)(n1 * words) * 8, s));
    if (n2 > 0) {
        k_pair_bits<<<grid_for(n2), 256, 0, s>>>(d_rows2, n2, d_colpos, n1, words, d_bits);
        ETM_LAUNCHED("k_pair_bits");
    }
    k_pair_rank<<<grid_for(n1), 256, 0, s>>>(d_bits, n1, words, d_rank, d_gof);
    ETM_LAUNCHED("k_pair_rank");
    if (n_groups > 0) {
        k_group_of<<<grid_for(n_groups), 256, 0, s>>>(d_rows2, d_group_start, n_groups, d_colpos, n1, d_gof);
        ETM_LAUNCHED("k_group_of");
    }
    return ETM_OK;
}

extern "C" int etm_count_k3_rows(const int64_t* d_row_ptr, const int32_t* d_pos, int64_t n_rows,
                                 const uint64_t* d_bits, const int32_t* d_rank, const int32_t* d_gof, int64_t n1,
                                 const int64_t* d_group_start, const int64_t* d_cum_pairs, const uint8_t* d_status,
                                 int32_t* d_counts, void* stream) {
    ETM_ARG(d_row_ptr && d_pos && d_bits && d_rank && d_gof && d_group_start && d_cum_pairs && d_counts,
            "etm_count_k3_rows: null pointer");
    ETM_ARG(n1 >= 1 && n1 <= 65536, "etm_count_k3_rows: n1 must be in [1, 65536]");
    if (n_rows == 0) return ETM_OK;
    hipStream_t s = ETM_STREAM(stream);
    const int64_t words = (n1 + 63) / 64;
    k_count_k3_rows<<<grid_for(n_rows * 64, 8192), 256, 0, s>>>(d_row_ptr, d_pos, n_rows, d_bits, d_rank, words,
                                                                 n1, d_gof, d_group_start, d_cum_pairs, d_status,
                                                                 d_counts);
    ETM_LAUNCHED("k_count_k3_rows");
    k_count_k3_rows_long<<<grid_for(n_rows * 64, 4096), 256, 0, s>>>(d_row_ptr, d_pos, n_rows, d_bits, d_rank, words,
                                                                     n1, d_gof, d_group_start, d_cum_pairs, d_status,
                                                                     d_counts);
    ETM_LAUNCHED("k_count_k3_rows_long");
    return ETM_OK;
}

// ── SON pass-1 union ────────────────────────────────────────────────────

namespace {

// Inserts a chunk's local level (n rows of k column ids, mapped to item ids through col_to_item, with their
// local counts) into the union of one itemset length: an open-addressing table of indices into the union's
// key rows (k int64 item ids each), summing the local counts (known) and the chunk's slack per distinct
// itemset. A chunk's rows are distinct, so no two threads of one call insert the same key: a new key takes a
// fresh index (an atomic counter), writes its row, publishes it, then claims an empty slot by CAS; a thread
// that finds an occupied slot compares the published row.
__global__ __launch_bounds__(256) void k_union_insert(const int32_t* __restrict__ rows,
                                                       const int32_t* __restrict__ counts, int64_t n, int32_t k,
                                                       const int64_t* __restrict__ col_to_item, int64_t slack,
                                                       int32_t* table, int64_t mask, int64_t* keys, int64_t* known,
                                                       int64_t* slack_sum, unsigned long long* n_unique,
                                                       int64_t cap, int32_t* overflow) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        uint64_t h = 0;
        for (int t = 0; t < k; ++t) h += etm::item_hash((int32_t)(col_to_item[rows[r * k + t]] & 0x7fffffff)) * (t + 1);
        int64_t slot = (int64_t)(h & (uint64_t)mask);
        int64_t mine = -1;
        for (;;) {
            const int32_t cur = __hip_atomic_load(&table[slot], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
            if (cur >= 0) {
                bool eq = true;
                for (int t = 0; t < k && eq; ++t) eq = keys[(int64_t)cur * k + t] == col_to_item[rows[r * k + t]];
                if (eq) {
                    atomicAdd((unsigned long long*)&known[cur], (unsigned long long)counts[r]);
                    atomicAdd((unsigned long long*)&slack_sum[cur], (unsigned long long)slack);
                    break;
                }
                slot = (slot + 1) & mask;
                continue;
            }
            if (mine < 0) {
                mine = (int64_t)atomicAdd(n_unique, 1ull);
                if (mine >= cap) {
                    atomicOr(overflow, 1);
                    break;
                }
                for (int t = 0; t < k; ++t) keys[mine * k + t] = col_to_item[rows[r * k + t]];
                known[mine] = counts[r];
                slack_sum[mine] = slack;
                __threadfence();
            }
            if (atomicCAS(&table[slot], -1, (int32_t)mine) == -1) break;
            // Another key took the slot meanwhile: look again at the same slot (now occupied).
        }
    }
}

int grid_u(int64_t n) {
    int64_t g = (n + 255) / 256;
    return (int)std::max<int64_t>(1, std::min<int64_t>(g, 16384));
}

}  // namespace

extern "C" int etm_union_insert(const int32_t* d_rows, const int32_t* d_counts, int64_t n, int32_t k,
                                const int64_t* d_col_to_item, int64_t slack, int32_t* d_table, int64_t table_size,
                                int64_t* d_keys, int64_t* d_known, int64_t* d_slack, int64_t* d_n_unique, int64_t cap,
                                int32_t* d_overflow, void* stream) {
    ETM_ARG(d_rows && d_counts && d_col_to_item && d_table && d_keys && d_known && d_slack && d_n_unique && d_overflow,
            "etm_union_insert: null pointer");
    ETM_ARG(k >= 1 && k <= 62, "etm_union_insert: 1 <= k <= 62 required, got %d", k);
    ETM_ARG(table_size > 0 && (table_size & (table_size - 1)) == 0 && table_size >= 2 * cap,
            "etm_union_insert: table_size must be a power of two >= 2 x capacity");
    if (n == 0) return ETM_OK;
    k_union_insert<<<grid_u(n), 256, 0, ETM_STREAM(stream)>>>(d_rows, d_counts, n, k, d_col_to_item, slack, d_table,
                                                              table_size - 1, d_keys, d_known, d_slack,
                                                              (unsigned long long*)d_n_unique, cap, d_overflow);
    ETM_LAUNCHED("k_union_insert");
    return ETM_OK;
}

namespace {

// support[i] = counts[i] / n_rows as an IEEE double division (correctly rounded, as numpy's).
__global__ __launch_bounds__(256) void k_support(const int64_t* __restrict__ counts, int64_t n, double rows,
                                                  double* __restrict__ out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (double)counts[i] / rows;
}

}  // namespace

extern "C" int etm_support(const int64_t* d_counts, int64_t n, int64_t n_rows, double* d_out, void* stream) {
    ETM_ARG(n == 0 || (d_counts && d_out), "etm_support: null pointer");
    ETM_ARG(n_rows > 0, "etm_support: n_rows must be positive");
    if (n == 0) return ETM_OK;
    k_support<<<grid_u(n), 256, 0, ETM_STREAM(stream)>>>(d_counts, n, (double)n_rows, d_out);
    ETM_LAUNCHED("k_support");
    return ETM_OK;
}
