// Host <-> device copies of pageable host arrays through a ring of pinned staging buffers.
//
// A copy is cut into chunks of kChunk bytes. Each chunk is first copied into the calling thread's next pinned
// staging buffer by up to kCopyThreads host threads, then handed to an asynchronous DMA on the caller's stream,
// and the buffer's event is recorded behind it: while the DMA of chunk i runs, the host threads fill the buffer
// of chunk i + 1. A buffer is refilled only once its previous DMA has completed. The call returns when the last
// chunk is enqueued; stream order puts later work on the stream after the data. (Pinned memory can only be
// allocated here, not registered in place, so every byte passes through a staging buffer once.) Downloads run
// the other way: up to three chunks' DMAs in flight into the ring, each copied out by the host threads as it
// lands, and the call returns with the data in place.
#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#include "common.h"

namespace {

constexpr int64_t kChunk = 32ll << 20;  // bytes per staging buffer
constexpr int kRing = 3;                // staging buffers per calling thread
constexpr int kCopyThreads = 8;         // host threads per chunk copy (the caller included)
constexpr int kPoolThreads = 15;        // threads of the shared pool (uploads and result exports)
constexpr int64_t kParallelMin = 1 << 20;  // below this a chunk is copied by the caller alone

// A fixed pool of host threads for the chunk copies, shared by every calling thread (tasks queue in order).
class CopyPool {
   public:
    static CopyPool& get() {
        static CopyPool* pool = new CopyPool();  // kept for the process: its threads never join at exit
        return *pool;
    }
    int size() const { return (int)threads_.size(); }
    // Runs fn(0..n-1): parts 1..n-1 on the pool, part 0 on the caller; returns when all are done.
    void run(int n, const std::function<void(int)>& fn) {
        struct Join {
            std::mutex mu;
            std::condition_variable cv;
            int left;
        } join;
        join.left = n - 1;
        {
            std::lock_guard<std::mutex> lk(mu_);
            for (int p = 1; p < n; ++p)
                tasks_.push_back([&join, &fn, p] {
                    fn(p);
                    std::lock_guard<std::mutex> jl(join.mu);
                    if (--join.left == 0) join.cv.notify_one();
                });
        }
        cv_.notify_all();
        fn(0);
        std::unique_lock<std::mutex> jl(join.mu);
        join.cv.wait(jl, [&] { return join.left == 0; });
    }

   private:
    CopyPool() {
        const unsigned hw = std::max(2u, std::thread::hardware_concurrency());
        const int n = (int)std::min<unsigned>(kPoolThreads, hw - 1);
        for (int t = 0; t < n; ++t) threads_.emplace_back([this] { loop(); });
        for (auto& t : threads_) t.detach();
    }
    void loop() {
        for (;;) {
            std::function<void()> task;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return !tasks_.empty(); });
                task = std::move(tasks_.front());
                tasks_.erase(tasks_.begin());
            }
            task();
        }
    }
    std::mutex mu_;
    std::condition_variable cv_;
    std::vector<std::function<void()>> tasks_;
    std::vector<std::thread> threads_;
};

// The calling thread's staging ring (pinned buffers and their DMA-done events), created on first use on a
// device and kept for the process.
struct Ring {
    int device = -1;
    char* buf[kRing] = {};
    hipEvent_t done[kRing] = {};
    bool used[kRing] = {};
    int next = 0;
};

Ring& thread_ring() {
    static thread_local Ring ring;
    return ring;
}

hipError_t ring_for_device(Ring& r) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (r.device == dev) return hipSuccess;
    for (int i = 0; i < kRing; ++i) {
        if (r.buf[i] == nullptr) {
            e = hipHostMalloc((void**)&r.buf[i], (size_t)kChunk, hipHostMallocDefault);
            if (e != hipSuccess) return e;
        }
        // Events belong to a device: a ring moving to another device waits for its copies and takes new ones.
        if (r.done[i] != nullptr) {
            if (r.used[i]) (void)hipEventSynchronize(r.done[i]);
            (void)hipEventDestroy(r.done[i]);
        }
        e = hipEventCreateWithFlags(&r.done[i], hipEventDisableTiming);
        if (e != hipSuccess) return e;
        r.used[i] = false;
    }
    r.device = dev;
    return hipSuccess;
}

void copy_parallel(char* dst, const char* src, int64_t n) {
    CopyPool& pool = CopyPool::get();
    const int parts = n < kParallelMin ? 1 : std::min<int>(std::min(pool.size() + 1, kCopyThreads),
                                                           (int)(n / (kParallelMin / 2)));
    if (parts <= 1) {
        std::memcpy(dst, src, (size_t)n);
        return;
    }
    const int64_t per = (n + parts - 1) / parts;
    pool.run(parts, [&](int p) {
        const int64_t a = p * per, b = std::min(n, a + per);
        if (a < b) std::memcpy(dst + a, src + a, (size_t)(b - a));
    });
}

}  // namespace

namespace etm {
namespace detail {

int parallel_width() { return CopyPool::get().size() + 1; }

void parallel_for(int parts, const std::function<void(int)>& fn) {
    if (parts <= 1) {
        if (parts == 1) fn(0);
        return;
    }
    CopyPool::get().run(parts, fn);
}

}  // namespace detail
}  // namespace etm

extern "C" int etm_download(const void* d_src, int64_t bytes, void* h_dst, void* stream) {
    ETM_ARG(bytes >= 0, "etm_download: negative size");
    ETM_ARG(bytes == 0 || (d_src != nullptr && h_dst != nullptr), "etm_download: null pointer");
    if (bytes == 0) return ETM_OK;
    Ring& ring = thread_ring();  // the same thread's ring as etm_upload's
    ETM_HIP(ring_for_device(ring));
    hipStream_t s = ETM_STREAM(stream);
    const char* src = (const char*)d_src;
    char* dst = (char*)h_dst;
    const int64_t n_chunks = (bytes + kChunk - 1) / kChunk;
    // Chunk c lands in buffer c % kRing; up to kRing DMAs run ahead of the host copies out of the buffers.
    auto issue = [&](int64_t c) -> hipError_t {
        const int b = (int)(c % kRing);
        if (ring.used[b]) {
            const hipError_t e = hipEventSynchronize(ring.done[b]);  // an earlier upload still reading it
            if (e != hipSuccess) return e;
        }
        const int64_t off = c * kChunk, n = std::min(kChunk, bytes - off);
        hipError_t e = hipMemcpyAsync(ring.buf[b], src + off, (size_t)n, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipEventRecord(ring.done[b], s);
        ring.used[b] = true;
        return e;
    };
    for (int64_t c = 0; c < std::min<int64_t>(kRing, n_chunks); ++c) ETM_HIP(issue(c));
    for (int64_t c = 0; c < n_chunks; ++c) {
        const int b = (int)(c % kRing);
        ETM_HIP(hipEventSynchronize(ring.done[b]));
        const int64_t off = c * kChunk, n = std::min(kChunk, bytes - off);
        copy_parallel(dst + off, ring.buf[b], n);
        ring.used[b] = false;
        if (c + kRing < n_chunks) ETM_HIP(issue(c + kRing));
    }
    return ETM_OK;
}

extern "C" int etm_upload(const void* h_src, int64_t bytes, void* d_dst, void* stream) {
    ETM_ARG(bytes >= 0, "etm_upload: negative size");
    ETM_ARG(bytes == 0 || (h_src != nullptr && d_dst != nullptr), "etm_upload: null pointer");
    if (bytes == 0) return ETM_OK;
    Ring& ring = thread_ring();
    ETM_HIP(ring_for_device(ring));
    hipStream_t s = ETM_STREAM(stream);
    const char* src = (const char*)h_src;
    char* dst = (char*)d_dst;
    for (int64_t off = 0; off < bytes; off += kChunk) {
        const int64_t n = std::min(kChunk, bytes - off);
        const int b = ring.next;
        ring.next = (ring.next + 1) % kRing;
        if (ring.used[b]) ETM_HIP(hipEventSynchronize(ring.done[b]));
        copy_parallel(ring.buf[b], src + off, n);
        ETM_HIP(hipMemcpyAsync(dst + off, ring.buf[b], (size_t)n, hipMemcpyHostToDevice, s));
        ETM_HIP(hipEventRecord(ring.done[b], s));
        ring.used[b] = true;
    }
    return ETM_OK;
}
