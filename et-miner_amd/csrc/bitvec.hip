// Input side of the counting path: item-id range and histogram, the CSR ->
// column-bitvector builder, and the K=1 column popcount.
//
// Builder layout: a workgroup owns a tile of 8 consecutive words (512
// transactions) times up to 1,024 columns in LDS; each entry sets its bit
// with an LDS atomic OR (no atomic reaches HBM), and each column's 8 words
// leave as one contiguous 64-byte run. Rows may be unsorted or repeat items.
#include <algorithm>

#include <rocprim/device/device_radix_sort.hpp>

#include "batch.h"
#include "common.h"
#include "internal.h"

namespace {

constexpr int kBuildThreads = 512;

// Ids as 16-byte vectors: 2 int64 or 4 int32 per load. The vector body covers the 16-byte aligned prefix
// that holds whole vectors; the remaining ids go through scalar loads.
template <typename Id>
struct IdVec;
template <>
struct IdVec<int64_t> {
    using V = longlong2;
    static constexpr int kPer = 2;
    __device__ static void get(const V& v, int64_t (&o)[2]) {
        o[0] = v.x;
        o[1] = v.y;
    }
};
template <>
struct IdVec<int32_t> {
    using V = int4;
    static constexpr int kPer = 4;
    __device__ static void get(const V& v, int64_t (&o)[4]) {
        o[0] = v.x;
        o[1] = v.y;
        o[2] = v.z;
        o[3] = v.w;
    }
};

// Calls fn(id) for every id of ids[0..n), each thread keeping 4 vector loads in flight, workgroup `blk` of
// `nblk` striding over the ids.
template <typename Id, class Fn>
__device__ __forceinline__ void for_each_id_in(const Id* __restrict__ ids, int64_t n, int64_t blk, int64_t nblk,
                                               Fn fn) {
    using IV = IdVec<Id>;
    constexpr int P = IV::kPer;
    const int64_t nv = ((uintptr_t)ids % 16 == 0) ? n / P : 0;
    const typename IV::V* vec = reinterpret_cast<const typename IV::V*>(ids);
    const int64_t stride = nblk * blockDim.x;
    int64_t i = blk * blockDim.x + threadIdx.x;
    for (; i + 3 * stride < nv; i += 4 * stride) {
        typename IV::V v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = vec[i + u * stride];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            int64_t o[P];
            IV::get(v[u], o);
#pragma unroll
            for (int q = 0; q < P; ++q) fn(o[q]);
        }
    }
    for (; i < nv; i += stride) {
        int64_t o[P];
        IV::get(vec[i], o);
#pragma unroll
        for (int q = 0; q < P; ++q) fn(o[q]);
    }
    for (int64_t j = nv * P + blk * blockDim.x + threadIdx.x; j < n; j += stride) fn((int64_t)ids[j]);
}

// The same over the whole grid.
template <typename Id, class Fn>
__device__ __forceinline__ void for_each_id(const Id* __restrict__ ids, int64_t n, Fn fn) {
    for_each_id_in(ids, n, (int64_t)blockIdx.x, (int64_t)gridDim.x, fn);
}

// Range of the ids: each workgroup reduces its share and updates out[0] / out[1] only where it improves them
// (the plain read first spares almost every atomic once an extreme has been seen).
__device__ __forceinline__ void minmax_publish(int64_t lo, int64_t hi, int64_t* out);

template <typename Id>
__global__ __launch_bounds__(256) void k_ids_minmax(const Id* __restrict__ ids, int64_t n, int64_t* out) {
    int64_t lo = INT64_MAX, hi = INT64_MIN;
    for_each_id(ids, n, [&](int64_t v) {
        lo = v < lo ? v : lo;
        hi = v > hi ? v : hi;
    });
    minmax_publish(lo, hi, out);
}

// The workgroup's extremes into out[0] / out[1] (atomics only where they improve them).
__device__ __forceinline__ void minmax_publish(int64_t lo, int64_t hi, int64_t* out) {
    __shared__ int64_t lds_min[4], lds_max[4];
    lo = etm::wave_min(lo);
    hi = etm::wave_max(hi);
    const int wid = threadIdx.x / 64;
    if ((threadIdx.x & 63) == 0) {
        lds_min[wid] = lo;
        lds_max[wid] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 0; w < 4; ++w) {
            lo = lds_min[w] < lo ? lds_min[w] : lo;
            hi = lds_max[w] > hi ? lds_max[w] : hi;
        }
        if (lo < __atomic_load_n(&out[0], __ATOMIC_RELAXED)) atomicMin((long long*)&out[0], (long long)lo);
        if (hi > __atomic_load_n(&out[1], __ATOMIC_RELAXED)) atomicMax((long long*)&out[1], (long long)hi);
    }
}

__global__ void k_minmax_init(int64_t* out) {
    out[0] = INT64_MAX;
    out[1] = INT64_MIN;
}

__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Histogram through a per-workgroup LDS copy (span <= kLdsBins: 128 KiB of gfx950's 160 KiB LDS), flushed with
// one global add per bin; wider spans take global atomics.
constexpr int kLdsBins = 32768;

template <typename Id>
__global__ __launch_bounds__(256) void k_hist_lds(const Id* __restrict__ ids, int64_t n, int64_t lo, int32_t span,
                                                   int32_t* __restrict__ hist) {
    __shared__ int32_t bins[kLdsBins];
    for (int b = threadIdx.x; b < span; b += blockDim.x) bins[b] = 0;
    __syncthreads();
    for_each_id(ids, n, [&](int64_t id) {
        const int64_t v = id - lo;
        if (v >= 0 && v < span) atomicAdd(&bins[v], 1);
    });
    __syncthreads();
    for (int b = threadIdx.x; b < span; b += blockDim.x)
        if (bins[b]) atomicAdd(&hist[b], bins[b]);
}

template <typename Id>
__global__ __launch_bounds__(256) void k_hist_global(const Id* __restrict__ ids, int64_t n, int64_t lo, int64_t span,
                                                      int32_t* __restrict__ hist) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        int64_t v = (int64_t)ids[i] - lo;
        if (v >= 0 && v < span) atomicAdd(&hist[v], 1);
    }
}

// The builder: a workgroup owns kTileWords words (64 x kTileWords rows) x a slab of up to kTileCols columns in LDS
// (at most 64 KiB); each entry of its rows sets its bit with an LDS atomic OR, so each row's ids are read once per
// kTileCols columns and no work is spent on a column a row does not hold. Each column's kTileWords words leave
// as one contiguous 64-byte run; no atomic reaches HBM. (Measured against a register-mask + wave-ballot
// transpose of 128/256-column slabs: deep_k_1m 2.85 -> 2.77 ms/step, wide_vocab_2m 17.4 -> 16.4 ms/step.)
constexpr int kTileWords = 8;
constexpr int kTileCols = 1024;

// Word tile `wt` x column slab `slab` of the builder, through the LDS tile `tile`. With `part`, the tile's
// popcount of each column (at most 512) goes to part[wt * n_cols + column] as well.
template <typename Id>
__device__ __forceinline__ void build_tile(const int64_t* __restrict__ indptr, const Id* __restrict__ ids,
                                           int64_t n_rows, int64_t lo, int64_t span, const int32_t* __restrict__ lut,
                                           int32_t n_cols, int64_t pitch, const int32_t* __restrict__ order,
                                           int32_t slab_cols, uint64_t* __restrict__ out, int64_t wt, int slab,
                                           unsigned long long* tile, uint16_t* __restrict__ part = nullptr) {
    const int64_t w0 = wt * kTileWords;
    const int c0 = slab * slab_cols;
    const int nc = min(slab_cols, n_cols - c0);
    for (int t = threadIdx.x; t < nc * kTileWords; t += kBuildThreads) tile[t] = 0;
    __syncthreads();
    // Lane l of wave v takes the row at position w0 * 64 + 8 l + v (eight waves, one row a thread: twice the
    // waves of four with two rows each, to hide the dependent indptr -> ids -> LUT loads): the 64 lanes of a
    // wave spread over all kTileWords words, so a column held by many rows meets at most 8 lanes per address.
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int j = wv; j < 8; j += kBuildThreads / 64) {
        const int l = 8 * lane + j;  // 0..511: word l / 64, bit l % 64
        const int64_t pos = w0 * 64 + l;
        if (pos >= n_rows) continue;
        const int64_t r = order != nullptr ? (int64_t)order[pos] : pos;
        const int64_t a = indptr[r], b = indptr[r + 1];
        const unsigned long long bit = 1ull << (l & 63);
        unsigned long long* word = tile + (l >> 6);
        auto set = [&](int64_t id) {
            const int64_t v = id - lo;
            if (v < 0 || v >= span) return;
            const int c = lut[v] - c0;
            if (c < 0 || c >= nc) return;
            atomicOr(word + c * kTileWords, bit);
        };
        int64_t e = a;
        for (; e + 4 <= b; e += 4) {
            const int64_t v0 = (int64_t)ids[e], v1 = (int64_t)ids[e + 1], v2 = (int64_t)ids[e + 2],
                          v3 = (int64_t)ids[e + 3];
            set(v0);
            set(v1);
            set(v2);
            set(v3);
        }
        for (; e < b; ++e) set((int64_t)ids[e]);
    }
    __syncthreads();
    // A column's kTileWords words sit in kTileWords consecutive lanes (all active or all done together: both
    // kBuildThreads and nc * kTileWords are multiples of kTileWords), so its popcount is a three-step butterfly.
    for (int t = threadIdx.x; t < nc * kTileWords; t += kBuildThreads) {
        const int c = t / kTileWords, w = t % kTileWords;
        const unsigned long long x = tile[t];
        if (w0 + w < pitch) out[(int64_t)(c0 + c) * pitch + w0 + w] = x;
        if (part != nullptr) {
            int p = __popcll(x);
            p += __shfl_xor(p, 1);
            p += __shfl_xor(p, 2);
            p += __shfl_xor(p, 4);
            if (w == 0) part[wt * n_cols + c0 + c] = (uint16_t)p;
        }
    }
}

template <typename Id>
__global__ __launch_bounds__(kBuildThreads) void k_bitvec_build_lds(const int64_t* __restrict__ indptr,
                                                                    const Id* __restrict__ ids, int64_t n_rows,
                                                                    int64_t lo, int64_t span,
                                                                    const int32_t* __restrict__ lut, int32_t n_cols,
                                                                    int64_t pitch, const int32_t* __restrict__ order,
                                                                    int32_t slab_cols, uint64_t* __restrict__ out,
                                                                    uint16_t* __restrict__ part) {
    // The slab's tile, sized at launch: slab_cols (at most kTileCols) columns x kTileWords words.
    extern __shared__ unsigned long long tile[];
    build_tile(indptr, ids, n_rows, lo, span, lut, n_cols, pitch, order, slab_cols, out, (int64_t)blockIdx.x,
               (int)blockIdx.y, tile, part);
}

// Column popcounts from the builder's per-tile partials: a thread per column and chunk of kPartTiles tiles
// (consecutive threads read consecutive columns of a tile row), one 64-bit atomic per (column, chunk) into the
// zeroed output. Reads 2 bytes per (column, 8-word tile) where a pass over the bitvectors reads 64.
constexpr int64_t kPartTiles = 128;

__global__ __launch_bounds__(256) void k_colpop_parts(const uint16_t* __restrict__ part, int64_t n_tiles,
                                                       int32_t n_cols, int64_t* __restrict__ out) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n_cols) return;
    const int64_t t0 = (int64_t)blockIdx.y * kPartTiles, t1 = min(n_tiles, t0 + kPartTiles);
    uint32_t s = 0;
    for (int64_t t = t0; t < t1; ++t) s += part[t * n_cols + c];
    atomicAdd((unsigned long long*)&out[c], (unsigned long long)s);
}


// One workgroup per (column, chunk of kColpopChunk words): the sum of popcounts over the chunk's words, four
// loads in flight per thread. A column of one chunk stores its sum; with more chunks each adds its partial sum
// to the zeroed output (one 64-bit atomic per workgroup).
constexpr int64_t kColpopChunk = 16384;

__global__ __launch_bounds__(256) void k_colpop(const uint64_t* __restrict__ bv, int64_t pitch, int64_t n_words,
                                                 int64_t* __restrict__ out) {
    __shared__ int64_t lds[8];
    const uint64_t* col = bv + (int64_t)blockIdx.x * pitch;
    const int64_t w0 = (int64_t)blockIdx.y * kColpopChunk;
    const int64_t w1 = min(n_words, w0 + kColpopChunk);
    int64_t s = 0;
    int64_t w = w0 + threadIdx.x;
    for (; w + 768 < w1; w += 1024)
        s += __popcll(col[w]) + __popcll(col[w + 256]) + __popcll(col[w + 512]) + __popcll(col[w + 768]);
    for (; w < w1; w += 256) s += __popcll(col[w]);
    s = etm::block_sum<256>(s, lds);
    if (threadIdx.x == 0) {
        if (gridDim.y == 1) out[blockIdx.x] = s;
        else atomicAdd((unsigned long long*)&out[blockIdx.x], (unsigned long long)s);
    }
}

__global__ void k_zero_counts(int64_t* __restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// ── the input layer of a batch of jobs (batch.h InJob): one launch per phase over every job ──────────────

using etm::batch::InJob;

// The job of flattened workgroup b: pre[j] <= b < pre[j + 1].
__device__ __forceinline__ int in_job_of(const int64_t* __restrict__ pre, int n_jobs, int64_t b) {
    int lo = 0, hi = n_jobs;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (pre[mid] <= b) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ void k_batch_minmax_init(const InJob* __restrict__ jobs, int n_jobs) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_jobs) return;
    jobs[j].mm[0] = INT64_MAX;
    jobs[j].mm[1] = INT64_MIN;
}

__global__ __launch_bounds__(256) void k_batch_minmax(const InJob* __restrict__ jobs, int n_jobs,
                                                       const int64_t* __restrict__ pre) {
    const int j = in_job_of(pre, n_jobs, blockIdx.x);
    const InJob& J = jobs[j];
    const int64_t blk = blockIdx.x - pre[j], nblk = pre[j + 1] - pre[j];
    int64_t lo = INT64_MAX, hi = INT64_MIN;
    auto fn = [&](int64_t v) {
        lo = v < lo ? v : lo;
        hi = v > hi ? v : hi;
    };
    if (J.id_bytes == 8) for_each_id_in((const int64_t*)J.ids, J.nnz, blk, nblk, fn);
    else for_each_id_in((const int32_t*)J.ids, J.nnz, blk, nblk, fn);
    minmax_publish(lo, hi, J.mm);
}

// Histograms: spans of at most kLdsBins through an LDS copy per workgroup (as k_hist_lds), wider ones with global
// atomics. The bins are zeroed beforehand.
__global__ __launch_bounds__(256) void k_batch_hist(const InJob* __restrict__ jobs, int n_jobs,
                                                     const int64_t* __restrict__ pre) {
    __shared__ int32_t bins[kLdsBins];
    const int j = in_job_of(pre, n_jobs, blockIdx.x);
    const InJob& J = jobs[j];
    const int64_t blk = blockIdx.x - pre[j], nblk = pre[j + 1] - pre[j];
    const int64_t lo = J.lo, span = J.span;
    int32_t* hist = J.hist;
    if (span <= kLdsBins) {
        for (int b = threadIdx.x; b < span; b += blockDim.x) bins[b] = 0;
        __syncthreads();
        auto fn = [&](int64_t id) {
            const int64_t v = id - lo;
            if (v >= 0 && v < span) atomicAdd(&bins[v], 1);
        };
        if (J.id_bytes == 8) for_each_id_in((const int64_t*)J.ids, J.nnz, blk, nblk, fn);
        else for_each_id_in((const int32_t*)J.ids, J.nnz, blk, nblk, fn);
        __syncthreads();
        for (int b = threadIdx.x; b < span; b += blockDim.x)
            if (bins[b]) atomicAdd(&hist[b], bins[b]);
    } else {
        auto fn = [&](int64_t id) {
            const int64_t v = id - lo;
            if (v >= 0 && v < span) atomicAdd(&hist[v], 1);
        };
        if (J.id_bytes == 8) for_each_id_in((const int64_t*)J.ids, J.nnz, blk, nblk, fn);
        else for_each_id_in((const int32_t*)J.ids, J.nnz, blk, nblk, fn);
    }
}

// Bitvectors: a job's workgroups are its (word tile, column slab) pairs, tile-major within a slab.
__global__ __launch_bounds__(kBuildThreads) void k_batch_build(const InJob* __restrict__ jobs, int n_jobs,
                                                                const int64_t* __restrict__ pre) {
    extern __shared__ unsigned long long tile[];
    const int j = in_job_of(pre, n_jobs, blockIdx.x);
    const InJob& J = jobs[j];
    const int64_t local = blockIdx.x - pre[j];
    const int64_t tiles = (J.pitch + kTileWords - 1) / kTileWords;
    const int64_t wt = local % tiles;
    const int slab = (int)(local / tiles);
    if (J.id_bytes == 8)
        build_tile(J.indptr, (const int64_t*)J.ids, J.n_rows, J.lo, J.span, J.lut, J.n_cols, J.pitch, nullptr,
                   (int32_t)J.slab_cols, J.bv, wt, slab, tile);
    else
        build_tile(J.indptr, (const int32_t*)J.ids, J.n_rows, J.lo, J.span, J.lut, J.n_cols, J.pitch, nullptr,
                   (int32_t)J.slab_cols, J.bv, wt, slab, tile);
}

// K=1 counts: a workgroup per column of every job.
__global__ __launch_bounds__(256) void k_batch_colpop(const InJob* __restrict__ jobs, int n_jobs,
                                                       const int64_t* __restrict__ pre) {
    __shared__ int64_t lds[8];
    const int j = in_job_of(pre, n_jobs, blockIdx.x);
    const InJob& J = jobs[j];
    const int64_t c = blockIdx.x - pre[j];
    const uint64_t* col = J.bv + c * J.pitch;
    int64_t s = 0;
    for (int64_t w = threadIdx.x; w < J.n_words; w += blockDim.x) s += __popcll(col[w]);
    s = etm::block_sum<256>(s, lds);
    if (threadIdx.x == 0) J.pc[c] = s;
}

// Row order key: bit 63 - r for each column of rank r < n_ranked the row holds (repeated ids set the same bit).
template <typename Id>
__global__ __launch_bounds__(256) void k_row_keys(const int64_t* __restrict__ indptr, const Id* __restrict__ ids,
                                                   int64_t n_rows, int64_t lo, int64_t span,
                                                   const int32_t* __restrict__ lut, const int32_t* __restrict__ rank,
                                                   int32_t n_ranked, uint64_t* __restrict__ keys,
                                                   int32_t* __restrict__ iota) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n_rows;
         r += (int64_t)gridDim.x * blockDim.x) {
        uint64_t key = 0;
        const int64_t b = indptr[r + 1];
        for (int64_t e = indptr[r]; e < b; ++e) {
            const int64_t v = (int64_t)ids[e] - lo;
            if (v < 0 || v >= span) continue;
            const int32_t c = lut[v];
            if (c < 0) continue;
            const int32_t rk = rank[c];
            if (rk >= 0 && rk < n_ranked) key |= 1ull << (63 - rk);
        }
        keys[r] = key;
        iota[r] = (int32_t)r;
    }
}

// Tile occupancy: bit t % 64 of occ[c * occ_pitch + t / 64] is set when column c has a nonzero word in tile t
// (words [64t, 64t + 64)). One wave per output word: 64 tiles, each one coalesced 512-byte read and a ballot.
// Words of a row past the last tile are written as zero.
__global__ __launch_bounds__(256) void k_occupancy(const uint64_t* __restrict__ bv, int64_t pitch, int64_t n_words,
                                                    int64_t n_cols, uint64_t* __restrict__ occ, int64_t occ_pitch) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (wave >= n_cols * occ_pitch) return;
    const int64_t c = wave / occ_pitch, q = wave % occ_pitch;
    const uint64_t* col = bv + c * pitch;
    uint64_t bits = 0;
    for (int t = 0; t < 64; ++t) {
        const int64_t w = (q * 64 + t) * 64 + lane;
        const bool nz = w < n_words && col[w] != 0;
        if (__ballot(nz) != 0) bits |= 1ull << t;
    }
    if (lane == 0) occ[wave] = bits;
}

int grid_for(int64_t n, int per_block = 256, int cap = 8192) {
    int64_t g = (n + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

}  // namespace

namespace etm {
namespace detail {

void batch_input_shape(etm::batch::InJob& j) {
    const int64_t slabs = std::max<int64_t>(1, (j.n_cols + kTileCols - 1) / kTileCols);
    j.slab_cols = (j.n_cols + slabs - 1) / slabs;
}

int64_t batch_input_blocks(int phase, const etm::batch::InJob& j) {
    switch (phase) {
        case 0:
            // 16K-32K ids a workgroup: the batch's jobs are small, and workgroups of 2K ids spent their time on
            // their set-up and the job search (1 TB/s against the per-job scan's 5 TB/s).
            return j.nnz > 0 ? grid_for(j.nnz, j.id_bytes == 4 ? 256 * 128 : 256 * 64, 1024) : 0;
        case 1:
            if (j.nnz == 0) return 0;
            return j.span <= kLdsBins ? grid_for(j.nnz, (int)std::max<int64_t>(4 * j.span, 8192), j.span <= 2048 ? 1024 : 512)
                                      : grid_for(j.nnz);
        case 2:
            return j.n_cols > 0 ? (j.pitch + kTileWords - 1) / kTileWords * ((j.n_cols + j.slab_cols - 1) / j.slab_cols)
                                : 0;
        default:
            return j.n_cols;
    }
}

hipError_t launch_batch_input(int phase, etm::batch::InJob* jobs, int n_jobs, const int64_t* pre, int64_t n_blocks,
                              int64_t max_slab_cols, hipStream_t s) {
    if (phase == 0) {
        k_batch_minmax_init<<<(n_jobs + 255) / 256, 256, 0, s>>>(jobs, n_jobs);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (n_blocks <= 0) return hipSuccess;
    if (n_blocks > 0x7FFFFFFF) return hipErrorInvalidValue;
    const unsigned g = (unsigned)n_blocks;
    if (phase == 0) k_batch_minmax<<<g, 256, 0, s>>>(jobs, n_jobs, pre);
    else if (phase == 1) k_batch_hist<<<g, 256, 0, s>>>(jobs, n_jobs, pre);
    else if (phase == 2)
        k_batch_build<<<g, kBuildThreads, (size_t)max_slab_cols * kTileWords * sizeof(unsigned long long), s>>>(
            jobs, n_jobs, pre);
    else k_batch_colpop<<<g, 256, 0, s>>>(jobs, n_jobs, pre);
    return hipGetLastError();
}

hipError_t launch_fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    k_fill_i32<<<grid_for(n), 256, 0, s>>>(p, n, v);
    return hipGetLastError();
}

}  // namespace detail
}  // namespace etm

extern "C" int etm_ids_minmax(const void* d_ids, int id_bytes, int64_t n, int64_t* d_out, void* stream) {
    ETM_ARG(d_ids != nullptr && d_out != nullptr, "etm_ids_minmax: null pointer");
    ETM_ARG(id_bytes == 4 || id_bytes == 8, "etm_ids_minmax: id_bytes must be 4 or 8, got %d", id_bytes);
    ETM_ARG(n >= 1, "etm_ids_minmax: n must be >= 1");
    hipStream_t s = ETM_STREAM(stream);
    k_minmax_init<<<1, 1, 0, s>>>(d_out);
    ETM_LAUNCHED("k_minmax_init");
    if (id_bytes == 4)
        k_ids_minmax<int32_t><<<grid_for(n, 256 * 16, 1024), 256, 0, s>>>((const int32_t*)d_ids, n, d_out);
    else
        k_ids_minmax<int64_t><<<grid_for(n, 256 * 8, 1024), 256, 0, s>>>((const int64_t*)d_ids, n, d_out);
    ETM_LAUNCHED("k_ids_minmax");
    return ETM_OK;
}

extern "C" int etm_ids_hist(const void* d_ids, int id_bytes, int64_t n, int64_t lo, int64_t span, int32_t* d_hist,
                            void* stream) {
    ETM_ARG(d_hist != nullptr, "etm_ids_hist: null histogram");
    ETM_ARG(id_bytes == 4 || id_bytes == 8, "etm_ids_hist: id_bytes must be 4 or 8, got %d", id_bytes);
    ETM_ARG(span >= 1 && n >= 0, "etm_ids_hist: span must be >= 1 and n >= 0");
    ETM_ARG(n == 0 || d_ids != nullptr, "etm_ids_hist: null ids");
    hipStream_t s = ETM_STREAM(stream);
    k_fill_i32<<<grid_for(span), 256, 0, s>>>(d_hist, span, 0);
    ETM_LAUNCHED("k_fill_i32");
    if (n == 0) return ETM_OK;
    if (span <= kLdsBins) {
        // Each workgroup flushes every nonzero bin once: give it at least 4 x span ids.
        const int g = grid_for(n, (int)std::max<int64_t>(4 * span, 8192), span <= 2048 ? 1024 : 512);
        if (id_bytes == 4)
            k_hist_lds<int32_t><<<g, 256, 0, s>>>((const int32_t*)d_ids, n, lo, (int32_t)span, d_hist);
        else
            k_hist_lds<int64_t><<<g, 256, 0, s>>>((const int64_t*)d_ids, n, lo, (int32_t)span, d_hist);
        ETM_LAUNCHED("k_hist_lds");
    } else {
        if (id_bytes == 4)
            k_hist_global<int32_t><<<grid_for(n), 256, 0, s>>>((const int32_t*)d_ids, n, lo, span, d_hist);
        else
            k_hist_global<int64_t><<<grid_for(n), 256, 0, s>>>((const int64_t*)d_ids, n, lo, span, d_hist);
        ETM_LAUNCHED("k_hist_global");
    }
    return ETM_OK;
}

namespace {
// Temporary bytes of the key sort (0 when the query fails; the sort itself then reports the error).
size_t sort_tmp_bytes(int64_t n) {
    size_t tmp = 0;
    if (rocprim::radix_sort_pairs(nullptr, tmp, (uint64_t*)nullptr, (uint64_t*)nullptr, (int32_t*)nullptr,
                                  (int32_t*)nullptr, (size_t)n, 0, 64) != hipSuccess)
        return 0;
    return tmp;
}
size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
}  // namespace

extern "C" size_t etm_row_order_ws(int64_t n_rows) {
    const int64_t n = std::max<int64_t>(n_rows, 1);
    return 2 * align256((size_t)n * 8) + align256((size_t)n * 4) + align256(sort_tmp_bytes(n));
}

extern "C" int etm_row_order(const int64_t* d_indptr, const void* d_ids, int id_bytes, int64_t n_rows, int64_t lo,
                             int64_t span, const int32_t* d_lut, const int32_t* d_rank, int32_t n_ranked,
                             int32_t* d_order, void* d_ws, size_t ws_bytes, void* stream) {
    ETM_ARG(d_indptr && d_lut && d_rank && d_order && d_ws, "etm_row_order: null pointer");
    ETM_ARG(id_bytes == 4 || id_bytes == 8, "etm_row_order: id_bytes must be 4 or 8, got %d", id_bytes);
    ETM_ARG(n_rows >= 0 && n_rows < 0x7FFFFFFF && span >= 1, "etm_row_order: bad sizes");
    ETM_ARG(n_ranked >= 1 && n_ranked <= 64, "etm_row_order: 1 <= n_ranked <= 64 required, got %d", n_ranked);
    ETM_ARG(ws_bytes >= etm_row_order_ws(n_rows), "etm_row_order: workspace too small");
    if (n_rows == 0) return ETM_OK;
    ETM_ARG(d_ids != nullptr, "etm_row_order: null ids");
    hipStream_t s = ETM_STREAM(stream);
    char* ws = (char*)d_ws;
    const size_t kb = align256((size_t)n_rows * 8);
    uint64_t* keys = (uint64_t*)ws;
    uint64_t* keys_out = (uint64_t*)(ws + kb);
    int32_t* iota = (int32_t*)(ws + 2 * kb);
    void* tmp = ws + 2 * kb + align256((size_t)n_rows * 4);
    size_t tmp_bytes = sort_tmp_bytes(n_rows);
    if (id_bytes == 4)
        k_row_keys<int32_t><<<grid_for(n_rows), 256, 0, s>>>(d_indptr, (const int32_t*)d_ids, n_rows, lo, span, d_lut,
                                                             d_rank, n_ranked, keys, iota);
    else
        k_row_keys<int64_t><<<grid_for(n_rows), 256, 0, s>>>(d_indptr, (const int64_t*)d_ids, n_rows, lo, span, d_lut,
                                                             d_rank, n_ranked, keys, iota);
    ETM_LAUNCHED("k_row_keys");
    // The whole key is sorted: rocPRIM's multi-block path returned unsorted output here for every begin_bit
    // above 0 tried (4, 24, 40, 56; above 1,000 rows), though only the top n_ranked bits carry ranks.
    const hipError_t e = rocprim::radix_sort_pairs(tmp, tmp_bytes, keys, keys_out, iota, d_order, (size_t)n_rows,
                                                   0, 64, s);
    ETM_REQUIRE(e == hipSuccess, ETM_EHIP, "etm_row_order: radix sort failed: %s", hipGetErrorString(e));
    return ETM_OK;
}

extern "C" int etm_tile_occupancy(const uint64_t* d_bv, int64_t pitch, int64_t n_words, int32_t n_cols,
                                  uint64_t* d_occ, int64_t occ_pitch, void* stream) {
    ETM_ARG(d_bv && d_occ, "etm_tile_occupancy: null pointer");
    ETM_ARG(n_cols >= 0 && n_words >= 0 && pitch >= n_words, "etm_tile_occupancy: bad sizes");
    const int64_t tiles = (n_words + 63) / 64;
    ETM_ARG(occ_pitch >= (tiles + 63) / 64 + 1, "etm_tile_occupancy: occ_pitch %lld < %lld",
            (long long)occ_pitch, (long long)((tiles + 63) / 64 + 1));
    const int64_t waves = (int64_t)n_cols * occ_pitch;
    if (waves == 0) return ETM_OK;
    ETM_REQUIRE((waves + 3) / 4 <= 0x7FFFFFFF, ETM_EOVERFLOW, "etm_tile_occupancy: grid too large");
    k_occupancy<<<(unsigned)((waves + 3) / 4), 256, 0, ETM_STREAM(stream)>>>(d_bv, pitch, n_words, n_cols, d_occ,
                                                                             occ_pitch);
    ETM_LAUNCHED("k_occupancy");
    return ETM_OK;
}

extern "C" int etm_bitvec_build(const int64_t* d_indptr, const void* d_ids, int id_bytes, int64_t n_rows, int64_t lo,
                                int64_t span, const int32_t* d_lut, int32_t n_cols, int64_t pitch,
                                const int32_t* d_order, uint64_t* d_out, void* stream) {
    return etm::detail::bitvec_build(d_indptr, d_ids, id_bytes, n_rows, lo, span, d_lut, n_cols, pitch, d_order, d_out,
                                     nullptr, nullptr, stream);
}

int64_t etm::detail::bitvec_part_len(int64_t pitch, int32_t n_cols) {
    return (pitch + kTileWords - 1) / kTileWords * n_cols;
}

int etm::detail::bitvec_build(const int64_t* d_indptr, const void* d_ids, int id_bytes, int64_t n_rows, int64_t lo,
                              int64_t span, const int32_t* d_lut, int32_t n_cols, int64_t pitch, const int32_t* d_order,
                              uint64_t* d_out, uint16_t* d_part, int64_t* d_counts, void* stream) {
    ETM_ARG(d_indptr != nullptr && d_lut != nullptr && d_out != nullptr, "etm_bitvec_build: null pointer");
    ETM_ARG((d_part == nullptr) == (d_counts == nullptr), "etm_bitvec_build: partials need counts");
    ETM_ARG(id_bytes == 4 || id_bytes == 8, "etm_bitvec_build: id_bytes must be 4 or 8, got %d", id_bytes);
    ETM_ARG(n_rows >= 0 && n_cols >= 0 && span >= 1, "etm_bitvec_build: negative size");
    ETM_ARG(pitch >= (n_rows + 63) / 64, "etm_bitvec_build: pitch %lld < ceil(n_rows/64) = %lld",
            (long long)pitch, (long long)((n_rows + 63) / 64));
    if (n_cols == 0 || pitch == 0) return ETM_OK;
    const int64_t tiles = (pitch + kTileWords - 1) / kTileWords;
    ETM_REQUIRE(tiles <= 0x7FFFFFFF, ETM_EOVERFLOW, "etm_bitvec_build: %lld word tiles exceed the grid",
                (long long)tiles);
    hipStream_t s = ETM_STREAM(stream);
    {
        // Column slabs of equal width (at most kTileCols): the LDS tile is only as large as a slab, so narrow
        // column sets leave room for more workgroups per CU (157 columns: 10 KB instead of 64 KB).
        const int32_t slabs = (n_cols + kTileCols - 1) / kTileCols;
        const int32_t slab_cols = (n_cols + slabs - 1) / slabs;
        const size_t lds = (size_t)slab_cols * kTileWords * sizeof(unsigned long long);
        dim3 grid((unsigned)tiles, (unsigned)slabs);
        if (id_bytes == 4)
            k_bitvec_build_lds<int32_t><<<grid, kBuildThreads, lds, s>>>(d_indptr, (const int32_t*)d_ids, n_rows, lo,
                                                                         span, d_lut, n_cols, pitch, d_order,
                                                                         slab_cols, d_out, d_part);
        else
            k_bitvec_build_lds<int64_t><<<grid, kBuildThreads, lds, s>>>(d_indptr, (const int64_t*)d_ids, n_rows, lo,
                                                                         span, d_lut, n_cols, pitch, d_order,
                                                                         slab_cols, d_out, d_part);
        ETM_LAUNCHED("k_bitvec_build_lds");
    }
    if (d_part != nullptr) {
        const int64_t chunks = (tiles + kPartTiles - 1) / kPartTiles;
        ETM_REQUIRE(chunks <= 65535, ETM_EOVERFLOW, "etm_bitvec_build: %lld tile chunks exceed the grid",
                    (long long)chunks);
        k_zero_counts<<<(unsigned)((n_cols + 255) / 256), 256, 0, s>>>(d_counts, n_cols);
        ETM_LAUNCHED("k_zero_counts");
        k_colpop_parts<<<dim3((unsigned)((n_cols + 255) / 256), (unsigned)chunks), 256, 0, s>>>(d_part, tiles, n_cols,
                                                                                                d_counts);
        ETM_LAUNCHED("k_colpop_parts");
    }
    return ETM_OK;
}

extern "C" int etm_column_popcounts(const uint64_t* d_bv, int32_t n_cols, int64_t pitch, int64_t n_words,
                                    int64_t* d_out, void* stream) {
    ETM_ARG(d_bv != nullptr && d_out != nullptr, "etm_column_popcounts: null pointer");
    ETM_ARG(n_cols >= 0 && n_words >= 0 && pitch >= n_words, "etm_column_popcounts: bad sizes");
    if (n_cols == 0) return ETM_OK;
    hipStream_t s = ETM_STREAM(stream);
    const int64_t chunks = std::max<int64_t>(1, (n_words + kColpopChunk - 1) / kColpopChunk);
    ETM_REQUIRE(chunks <= 65535, ETM_EOVERFLOW, "etm_column_popcounts: %lld word chunks exceed the grid",
                (long long)chunks);
    if (chunks > 1) {
        k_zero_counts<<<(unsigned)((n_cols + 255) / 256), 256, 0, s>>>(d_out, n_cols);
        ETM_LAUNCHED("k_zero_counts");
    }
    k_colpop<<<dim3((unsigned)n_cols, (unsigned)chunks), 256, 0, s>>>(d_bv, pitch, n_words, d_out);
    ETM_LAUNCHED("k_colpop");
    return ETM_OK;
}
