// Device-wide exclusive scan over a functor, in three launches:
//   1. k_scan_reduce: one int64 sum per block of kScanItems inputs,
//   2. k_scan_blocks: one workgroup scans the block sums in place (exclusive) and reports the total,
//   3. k_scan_apply:  each block rescans its inputs and hands (index, value, exclusive prefix) to a writer.
// The functor is evaluated twice per input instead of materialising the values.
#pragma once

#include "common.h"

namespace etm {

constexpr int kScanThreads = 256;
constexpr int kScanPer = 4;
constexpr int64_t kScanItems = (int64_t)kScanThreads * kScanPer;

inline int64_t scan_blocks(int64_t n) { return (n + kScanItems - 1) / kScanItems; }
inline size_t scan_ws_bytes(int64_t n) { return (size_t)(scan_blocks(n) + 1) * sizeof(int64_t); }

template <class F>
__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(F f, int64_t n, int64_t* __restrict__ bsum) {
    __shared__ int64_t lds[kScanThreads / kWave + 1];
    const int64_t base = (int64_t)blockIdx.x * kScanItems;
    int64_t s = 0;
#pragma unroll 4
    for (int r = 0; r < kScanPer; ++r) {
        const int64_t i = base + (int64_t)r * kScanThreads + threadIdx.x;
        if (i < n) s += f(i);
    }
    s = block_sum<kScanThreads>(s, lds);
    if (threadIdx.x == 0) bsum[blockIdx.x] = s;
}

// Scans bsum[0..nb) in place (exclusive); writes the total to bsum[nb] and, when `total` is not null, to *total.
template <class Done>
__global__ __launch_bounds__(1024) void k_scan_blocks(int64_t* __restrict__ bsum, int64_t nb, int64_t* total,
                                                      Done done) {
    __shared__ int64_t lds[1024 / kWave + 1];
    int64_t carry = 0;
    for (int64_t base = 0; base < nb; base += 1024) {
        const int64_t i = base + threadIdx.x;
        const int64_t v = i < nb ? bsum[i] : 0;
        int64_t tot;
        const int64_t ex = block_exclusive_scan<1024>(v, lds, &tot);
        if (i < nb) bsum[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) {
        bsum[nb] = carry;
        if (total) *total = carry;
        done(carry);
    }
}

template <class F, class W>
__global__ __launch_bounds__(kScanThreads) void k_scan_apply(F f, W w, int64_t n, const int64_t* __restrict__ bsum) {
    __shared__ int64_t lds[kScanThreads / kWave + 1];
    const int64_t base = (int64_t)blockIdx.x * kScanItems;
    int64_t carry = bsum[blockIdx.x];
    for (int r = 0; r < kScanPer; ++r) {
        const int64_t i = base + (int64_t)r * kScanThreads + threadIdx.x;
        const int64_t v = i < n ? f(i) : 0;
        int64_t tot;
        const int64_t ex = block_exclusive_scan<kScanThreads>(v, lds, &tot);
        if (i < n) w(i, v, carry + ex);
        carry += tot;
    }
}

struct NoDone {
    __device__ void operator()(int64_t) const {}
};

// Launches reduce + block scan. `ws` holds scan_ws_bytes(n). Returns a hipError_t of the launches.
template <class F, class Done = NoDone>
hipError_t scan_prepare(F f, int64_t n, int64_t* ws, int64_t* total, hipStream_t s, Done done = Done()) {
    const int64_t nb = scan_blocks(n);
    if (nb > 0) {
        k_scan_reduce<F><<<(unsigned)nb, kScanThreads, 0, s>>>(f, n, ws);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    k_scan_blocks<Done><<<1, 1024, 0, s>>>(ws, nb, total, done);
    return hipGetLastError();
}

template <class F, class W>
hipError_t scan_finish(F f, W w, int64_t n, const int64_t* ws, hipStream_t s) {
    const int64_t nb = scan_blocks(n);
    if (nb == 0) return hipSuccess;
    k_scan_apply<F, W><<<(unsigned)nb, kScanThreads, 0, s>>>(f, w, n, ws);
    return hipGetLastError();
}

}  // namespace etm
