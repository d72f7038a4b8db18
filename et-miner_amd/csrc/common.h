// Shared definitions of libetm's HIP sources: error reporting, argument
// checks, and the wave64 / workgroup reduction and scan helpers every kernel
// family uses. gfx950 only: wavefronts are 64 lanes wide.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstddef>
#include <cstdint>

#include "etm.h"

namespace etm {

// Records the calling thread's last error message (printf format).
void set_error(const char* fmt, ...);

constexpr int kWave = 64;

}  // namespace etm

#define ETM_REQUIRE(cond, code, ...)        \
    do {                                    \
        if (!(cond)) {                      \
            ::etm::set_error(__VA_ARGS__);  \
            return (code);                  \
        }                                   \
    } while (0)

#define ETM_ARG(cond, ...) ETM_REQUIRE(cond, ETM_EINVAL, __VA_ARGS__)

#define ETM_HIP(call)                                                                  \
    do {                                                                               \
        hipError_t etm_e_ = (call);                                                    \
        if (etm_e_ != hipSuccess) {                                                    \
            ::etm::set_error("%s failed: %s", #call, hipGetErrorString(etm_e_));       \
            return ETM_EHIP;                                                           \
        }                                                                              \
    } while (0)

// After a kernel launch: report a launch-configuration error.
#define ETM_LAUNCHED(name)                                                                  \
    do {                                                                                    \
        hipError_t etm_e_ = hipGetLastError();                                              \
        if (etm_e_ != hipSuccess) {                                                         \
            ::etm::set_error("launch of %s failed: %s", name, hipGetErrorString(etm_e_));   \
            return ETM_EHIP;                                                                \
        }                                                                                   \
    } while (0)

#define ETM_STREAM(s) (reinterpret_cast<hipStream_t>(s))

namespace etm {

// Wave64 sum of a 32- or 64-bit integer (butterfly over the 64 lanes).
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_min(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        T o = __shfl_xor(v, off, kWave);
        v = o < v ? o : v;
    }
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        T o = __shfl_xor(v, off, kWave);
        v = o > v ? o : v;
    }
    return v;
}

// Inclusive wave64 scan of int64 (Hillis-Steele over shuffles).
__device__ __forceinline__ int64_t wave_inclusive_scan(int64_t v) {
    const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        int64_t o = __shfl_up(v, off, kWave);
        if (lane >= off) v += o;
    }
    return v;
}

// Inclusive wave64 scan of int32 on the DPP crossbar: row_shr 1, 2, 4, 8 scan each 16-lane row, then
// row_bcast:15 adds row r's last lane to row r+1 (rows 1, 3) and row_bcast:31 adds lane 31 to rows 2 and 3.
// Lanes whose DPP source is outside the row read `old` = 0. No LDS traffic.
__device__ __forceinline__ int32_t wave_scan_i32(int32_t v) {
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);  // row_shr:1
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);  // row_shr:2
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);  // row_shr:4
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);  // row_shr:8
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);  // row_bcast:15
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return v;
}

// Exclusive scan of one int64 per thread over a workgroup of NT threads; returns the thread's exclusive
// prefix and writes the workgroup total to *total. `lds` holds NT/64 + 1 int64. Ends with a barrier.
template <int NT>
__device__ __forceinline__ int64_t block_exclusive_scan(int64_t v, int64_t* lds, int64_t* total) {
    constexpr int NW = NT / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    int64_t inc = wave_inclusive_scan(v);
    if (lane == kWave - 1) lds[wid] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t run = 0;
        for (int w = 0; w < NW; ++w) {
            int64_t t = lds[w];
            lds[w] = run;
            run += t;
        }
        lds[NW] = run;
    }
    __syncthreads();
    int64_t out = lds[wid] + inc - v;
    *total = lds[NW];
    __syncthreads();
    return out;
}

// Workgroup sum of one int64 per thread (NT threads); every thread gets the total. Ends with a barrier.
template <int NT>
__device__ __forceinline__ int64_t block_sum(int64_t v, int64_t* lds) {
    constexpr int NW = NT / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    v = wave_sum(v);
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    int64_t t = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) t += lds[w];
    __syncthreads();
    return t;
}

__device__ __forceinline__ int popc64(uint64_t x) { return __popcll(x); }

// Candidate (i, j), i < j, of a group of s suffixes -> lexicographic index within the group.
__host__ __device__ __forceinline__ int64_t lex_pair_index(int64_t i, int64_t j, int64_t s) {
    return i * (2 * s - i - 1) / 2 + (j - i - 1);
}

// Inverse of lex_pair_index: q in [0, s(s-1)/2) -> (i, j). Floating-point seed, exact integer correction.
__device__ __forceinline__ void lex_pair_decode(int64_t q, int64_t s, int64_t* i_out, int64_t* j_out) {
    // Row i starts at start(i) = i*(2s-i-1)/2; the largest i with start(i) <= q.
    const double b = 2.0 * (double)s - 1.0;
    double disc = b * b - 8.0 * (double)q;
    if (disc < 0) disc = 0;
    int64_t i = (int64_t)((b - sqrt(disc)) * 0.5);
    if (i < 0) i = 0;
    if (i > s - 2) i = s - 2;
    while (i > 0 && i * (2 * s - i - 1) / 2 > q) --i;
    while (i + 1 <= s - 2 && (i + 1) * (2 * s - i - 2) / 2 <= q) ++i;
    *i_out = i;
    *j_out = q - i * (2 * s - i - 1) / 2 + i + 1;
}

// Largest g in [0, n) with cum[g] <= x (cum ascending, cum[0] == 0 <= x).
__device__ __forceinline__ int64_t upper_index(const int64_t* __restrict__ cum, int64_t n, int64_t x) {
    int64_t lo = 0, hi = n - 1;
    while (lo < hi) {
        int64_t mid = (lo + hi + 1) >> 1;
        if (cum[mid] <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// upper_index for a whole wave with the same n and x (every lane gets the answer): each round the 64 lanes read
// 64 evenly spaced entries of the range and a ballot keeps the stretch between the last one <= x and the next,
// so a range of n entries takes ceil(log64 n) dependent loads instead of log2 n.
__device__ __forceinline__ int64_t wave_upper_index(const int64_t* __restrict__ cum, int64_t n, int64_t x) {
    const int lane = threadIdx.x & (kWave - 1);
    int64_t lo = 0, hi = n - 1;  // the answer lies in [lo, hi]; cum[lo] <= x
    while (lo < hi) {
        const int64_t span = hi - lo + 1;
        const int64_t step = span <= kWave ? 1 : (span + kWave - 1) / kWave;
        const int64_t g = lo + (int64_t)lane * step;
        const bool le = g <= hi && cum[g] <= x;
        const int t = __popcll(__ballot(le)) - 1;  // the true lanes are a prefix: cum ascends
        if (step == 1) return lo + t;
        const int64_t nlo = lo + (int64_t)t * step;
        hi = min(hi, nlo + step - 1);
        lo = nlo;
    }
    return lo;
}

// Commutative 64-bit hash of one item (an itemset hashes to the sum of its items' hashes, so dropping an
// item is one subtraction).
__host__ __device__ __forceinline__ uint64_t item_hash(int32_t item) {
    uint64_t z = (uint64_t)(uint32_t)item + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace etm
