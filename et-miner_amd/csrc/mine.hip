// The native level loop: mines every frequent itemset of device bitvectors (or of a device CSR, through
// the input layer) in one call, with every level resident in HBM.
//
// Per level the host enqueues the group build, the subset test, the counting kernels, the compaction, the
// free flags and the NEXT level's group build, then synchronizes once to read one block of scalars
// (survivors, counted/inferred/skipped tallies, the next level's group totals). Levels of up to
// kSingleMax rows (or candidates) take single-workgroup group-build and compaction kernels that read their
// sizes on the device, so nothing between two syncs waits on the host. Larger levels fall back to the
// multi-block scans with one extra sync. Device memory comes from a per-thread bump allocator over blocks
// kept across calls (DeviceHeap below), so the levels make no allocator calls between their syncs.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include <sys/mman.h>

#include "common.h"
#include "internal.h"

// Mappings of freed result arrays, kept for the next results (at most kMapPoolCount of them and kMapPoolBytes in
// all, process-wide): their pages are already faulted in, so the export threads of the next call write at memory
// speed instead of taking a page fault (and the kernel's zeroing of the page) per first touch.
class MapPool {
   public:
    static constexpr size_t kMapPoolCount = 16;
    static constexpr size_t kMapPoolBytes = size_t(2) << 30;
    static MapPool& get() {
        static MapPool* pool = new MapPool();  // kept for the process
        return *pool;
    }
    // The largest kept mapping of at least `bytes`, or {nullptr, 0}.
    std::pair<void*, size_t> take(size_t bytes) {
        std::lock_guard<std::mutex> lk(mu_);
        size_t best = maps_.size();
        for (size_t i = 0; i < maps_.size(); ++i)
            if (maps_[i].second >= bytes && (best == maps_.size() || maps_[i].second > maps_[best].second)) best = i;
        if (best == maps_.size()) return {nullptr, 0};
        const auto m = maps_[best];
        maps_.erase(maps_.begin() + (std::ptrdiff_t)best);
        held_ -= m.second;
        return m;
    }
    // Keeps a mapping, dropping the smallest kept ones (or this one) past the limits.
    void give(void* p, size_t cap) {
        std::vector<std::pair<void*, size_t>> drop;
        {
            std::lock_guard<std::mutex> lk(mu_);
            maps_.emplace_back(p, cap);
            held_ += cap;
            while (!maps_.empty() && (maps_.size() > kMapPoolCount || held_ > kMapPoolBytes)) {
                size_t small = 0;
                for (size_t i = 1; i < maps_.size(); ++i)
                    if (maps_[i].second < maps_[small].second) small = i;
                drop.push_back(maps_[small]);
                held_ -= maps_[small].second;
                maps_.erase(maps_.begin() + (std::ptrdiff_t)small);
            }
        }
        for (const auto& m : drop) munmap(m.first, m.second);
    }

   private:
    std::mutex mu_;
    std::vector<std::pair<void*, size_t>> maps_;
    size_t held_ = 0;
};

// A host array in an anonymous mapping that grows by remapping (no copy), with huge pages advised: the exported
// result of a large lattice, written by host threads while later levels are still being mined. A new array starts
// in the largest kept mapping that fits, and goes back to the pool when freed.
class etm_host_array {
   public:
    etm_host_array() = default;
    etm_host_array(const etm_host_array&) = delete;
    etm_host_array& operator=(const etm_host_array&) = delete;
    ~etm_host_array() {
        if (p_) MapPool::get().give(p_, cap_);
    }
    char* data() const { return (char*)p_; }
    // Capacity of at least `bytes` (false when the mapping fails); the contents are kept.
    bool reserve(size_t bytes) {
        if (bytes <= cap_) return true;
        constexpr size_t kHuge = 2u << 20;
        size_t want = std::max(bytes, cap_ * 2);
        want = (want + kHuge - 1) / kHuge * kHuge;
        if (p_ == nullptr) {
            const auto m = MapPool::get().take(want);
            if (m.first != nullptr) {
                p_ = m.first;
                cap_ = m.second;
                return true;
            }
        }
        void* q = p_ ? mremap(p_, cap_, want, MREMAP_MAYMOVE)
                     : mmap(nullptr, want, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
        if (q == MAP_FAILED) return false;
        (void)madvise(q, want, MADV_HUGEPAGE);
        p_ = q;
        cap_ = want;
        return true;
    }

   private:
    void* p_ = nullptr;
    size_t cap_ = 0;
};

// The result as a frame (item ids, offsets, counts and supports over every level, in level order), built during
// the call when the lattice has large levels.
struct etm_frame {
    etm_host_array values, offsets, counts, support;
    int64_t n_sets = 0, n_vals = 0;
    bool reserve(int64_t sets, int64_t vals) {
        return values.reserve((size_t)std::max<int64_t>(vals, 1) * 8) && offsets.reserve((size_t)(sets + 1) * 8) &&
               counts.reserve((size_t)std::max<int64_t>(sets, 1) * 8) &&
               support.reserve((size_t)std::max<int64_t>(sets, 1) * 8);
    }
};

// Host result of a mining call. Level 1 lives in host vectors; the levels from K=2 on stay in the pinned block
// their device-to-host copy landed in (taken from a process-wide pool and given back when the result is
// freed), so no further host copy is made before the caller reads them.
struct etm_levels {
    std::vector<int32_t> k;
    std::vector<int64_t> n;
    std::vector<const int32_t*> rows;    // n[l] * k[l] column ids per level
    std::vector<const int32_t*> counts;  // n[l] counts per level
    std::vector<int32_t> k1_rows, k1_counts;
    int32_t* block = nullptr;
    int64_t block_cap = 0;
    std::vector<etm_level_stat> stats;
    std::vector<int64_t> col_to_item;
    std::vector<int64_t> col_counts;
    std::vector<int32_t> owned;  // host copy of levels K >= 2 of a batched job (etm_mine_batch)
    // Large levels copied to the host while later levels were mined: one pinned pool block each.
    std::vector<std::pair<int32_t*, int64_t>> level_blocks;
    etm_frame* frame = nullptr;  // the prebuilt frame, or null
    bool rows_released = false;  // levels K >= 2 no longer readable (etm_levels_frame_view released them)
    ~etm_levels();
};

namespace {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Error(ETM_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

void check_rc(int rc, const char* what) {
    if (rc != ETM_OK) {
        char buf[1024];
        etm_last_error(buf, sizeof buf);
        throw Error(rc, std::string(what) + ": " + buf);
    }
}

// Device memory of mining calls: a per-thread, per-device bump allocator over blocks kept across calls.
// A call takes what it needs by moving a pointer (no allocator API call on the critical path between the
// level syncs); blocks added when a call outgrows the current one are merged into one larger block once the
// call has synchronized. Every call ends with a stream synchronization, so reuse by the next call on the
// same thread is ordered after all of this call's kernels.
struct DeviceHeap {
    int device = -1;
    std::vector<std::pair<char*, size_t>> blocks;  // (base, capacity); blocks[0] is the main block
    size_t used = 0;      // bytes taken from the last block
    size_t demand = 0;    // bytes taken by the current call over all blocks
    int depth = 0;        // nesting of Arena scopes (the input layer's holds the level loop's)
    ~DeviceHeap() {
        for (auto& b : blocks) (void)hipFree(b.first);
    }
};

DeviceHeap& heap() {
    static thread_local DeviceHeap h;
    int dev = 0;
    check(hipGetDevice(&dev), "hipGetDevice");
    if (h.device != dev) {
        for (auto& b : h.blocks) (void)hipFree(b.first);
        h.blocks.clear();
        h.device = dev;
    }
    return h;
}

class Arena {
   public:
    explicit Arena(hipStream_t s) : s_(s), h_(heap()) {
        if (h_.depth++ == 0) {
            h_.used = 0;
            h_.demand = 0;
        }
    }
    // The outermost scope ends after the call's final synchronization: fold overflow blocks into one block. A
    // call that fails part-way synchronizes here first, so that no queued copy or kernel outlives the host
    // buffers and arena memory it uses.
    ~Arena() {
        if (std::uncaught_exceptions() > 0) (void)hipStreamSynchronize(s_);
        if (--h_.depth > 0) return;
        if (h_.blocks.size() > 1) {
            (void)hipStreamSynchronize(s_);
            size_t total = 0;
            for (auto& b : h_.blocks) {
                total += b.second;
                (void)hipFree(b.first);
            }
            h_.blocks.clear();
            char* p = nullptr;
            const size_t want = std::max(total, h_.demand + (h_.demand >> 2));
            if (hipMalloc((void**)&p, want) == hipSuccess) h_.blocks.emplace_back(p, want);
            else (void)hipGetLastError();
        }
        h_.used = 0;
    }
    // Device bytes the current call has taken so far.
    int64_t bytes() const { return (int64_t)h_.demand; }
    template <class T>
    T* get(int64_t n) {
        const size_t bytes = (std::max<size_t>((size_t)std::max<int64_t>(n, 1) * sizeof(T), 256) + 255) & ~(size_t)255;
        h_.demand += bytes;
        if (h_.blocks.empty() || h_.used + bytes > h_.blocks.back().second) {
            const size_t want = std::max<size_t>(bytes, std::max<size_t>(64ull << 20, h_.demand * 2));
            char* p = nullptr;
            hipError_t e = hipMalloc((void**)&p, want);
            if (e == hipErrorOutOfMemory) {
                (void)hipGetLastError();
                throw Error(ETM_ENOMEM, "device allocation of " + std::to_string(want) + " bytes failed");
            }
            check(e, "device allocation");
            h_.blocks.emplace_back(p, want);
            h_.used = 0;
        }
        char* p = h_.blocks.back().first + h_.used;
        h_.used += bytes;
        return (T*)p;
    }

   private:
    hipStream_t s_;
    DeviceHeap& h_;
};

// Pinned host scalars for the per-level readback (one buffer per thread, kept for the process).
struct Pinned {
    int64_t* p = nullptr;
    Pinned() {
        static thread_local int64_t* buf = nullptr;
        if (!buf) check(hipHostMalloc((void**)&buf, 64 * sizeof(int64_t), hipHostMallocDefault), "hipHostMalloc");
        p = buf;
    }
};

// A pinned host buffer of at least n int32 (one per thread, grown on demand, kept for the process).
int32_t* staging(int64_t n) {
    static thread_local int32_t* buf = nullptr;
    static thread_local int64_t cap = 0;
    if (n > cap) {
        if (buf) (void)hipHostFree(buf);
        buf = nullptr;
        const int64_t want = std::max<int64_t>(n, 1 << 20);
        check(hipHostMalloc((void**)&buf, want * 4, hipHostMallocDefault), "hipHostMalloc");
        cap = want;
    }
    return buf;
}

// Waits for the stream by polling it: the per-level readbacks wait tens of microseconds at most, and a
// blocking wait's wake-up latency would add as much again to every level.
void wait_stream(hipStream_t s) {
    for (;;) {
        const hipError_t e = hipStreamQuery(s);
        if (e == hipSuccess) return;
        if (e != hipErrorNotReady) check(e, "stream");
        __builtin_ia32_pause();
    }
}

// Copies of several device int32 segments into one contiguous device buffer (one kernel, then one
// device-to-host copy instead of one per segment).
struct Seg {
    const int32_t* src;
    int64_t n;
    int64_t dst;
};
// Largest group (suffixes) counted by the wave-per-group kernels; bigger groups go to the tiled kernel, whose
// listed mode stages only a group's live columns. Measured on MI355X: 8 beat 16 (the 16-suffix kernel holds
// 120 accumulators a lane, two waves a SIMD) and 4 on deep_k_1m, wide_vocab_2m and the served stream.
constexpr int32_t kSmallMaxDefault = 8;
// Columns of at least this many words take the live-unit schedule automatically (with Pascal inference, every
// length does): there a level's units are long to stream, so launching only the live ones pays off.
constexpr int64_t kLiveMinWords = 65536;

constexpr int kGatherSegs = 32;
struct Segs {
    Seg s[kGatherSegs];
};

__global__ __launch_bounds__(256) void k_gather(Segs segs, int32_t* __restrict__ out) {
    const Seg sg = segs.s[blockIdx.y];
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < sg.n; i += (int64_t)gridDim.x * blockDim.x)
        out[sg.dst + i] = sg.src[i];
}

// Timing events of this thread, created on first use and reused by every later call on the same device
// (event creation costs far more than recording).
hipEvent_t pooled_event(size_t i) {
    static thread_local std::vector<hipEvent_t> pool;
    static thread_local int pool_device = -1;
    int dev = 0;
    check(hipGetDevice(&dev), "hipGetDevice");
    if (dev != pool_device) {
        for (hipEvent_t e : pool) (void)hipEventDestroy(e);
        pool.clear();
        pool_device = dev;
    }
    while (pool.size() <= i) {
        hipEvent_t e = nullptr;
        check(hipEventCreate(&e), "event");
        pool.push_back(e);
    }
    return pool[i];
}

// Process-wide pool of pinned host blocks for results (int32 capacity each). A block leaves the pool with a
// result and comes back when the result is freed; at most kPoolKeep blocks wait in the pool.
std::mutex g_pool_mu;
std::vector<std::pair<int32_t*, int64_t>> g_pool;
constexpr size_t kPoolKeep = 32;

int32_t* pool_take(int64_t n, int64_t* cap) {
    {
        std::lock_guard<std::mutex> lk(g_pool_mu);
        size_t best = g_pool.size();
        for (size_t i = 0; i < g_pool.size(); ++i)
            if (g_pool[i].second >= n && (best == g_pool.size() || g_pool[i].second < g_pool[best].second)) best = i;
        if (best < g_pool.size()) {
            auto b = g_pool[best];
            g_pool.erase(g_pool.begin() + (std::ptrdiff_t)best);
            *cap = b.second;
            return b.first;
        }
    }
    const int64_t want = std::max<int64_t>(n + (n >> 2), 1 << 20);
    int32_t* p = nullptr;
    check(hipHostMalloc((void**)&p, want * 4, hipHostMallocDefault), "hipHostMalloc");
    *cap = want;
    return p;
}

void pool_give(int32_t* p, int64_t cap) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(g_pool_mu);
    g_pool.emplace_back(p, cap);
    if (g_pool.size() > kPoolKeep) {
        size_t small = 0;
        for (size_t i = 1; i < g_pool.size(); ++i)
            if (g_pool[i].second < g_pool[small].second) small = i;
        (void)hipHostFree(g_pool[small].first);
        g_pool.erase(g_pool.begin() + (std::ptrdiff_t)small);
    }
}

// Writes a level (rows: n x k column ids, cnts: n counts) at itemset set0 / item val0 of the frame: item ids
// through c2i (or the column ids themselves), offsets, counts and supports (count / n_rows); ranges of rows go to
// up to 8 of the host pool's threads when the level is large (more leave the launching thread short of cores; 4
// made the export the step's tail: deep_sparse_20m 22.2-23.0 ms with 16, 22.3-22.4 with 8, 23.8 with 4).
void frame_write(etm_frame& f, const int32_t* rows, const int32_t* cnts, int64_t n, int32_t k, int64_t set0,
                 int64_t val0, const int64_t* c2i, double n_rows_d) {
    if (n <= 0) return;
    int64_t* v = (int64_t*)f.values.data();
    int64_t* off = (int64_t*)f.offsets.data();
    int64_t* c = (int64_t*)f.counts.data();
    double* sp = (double*)f.support.data();
    const int parts = (int)std::min<int64_t>(std::min(8, etm::detail::parallel_width()), std::max<int64_t>(1, n * k >> 17));
    const int64_t per = (n + parts - 1) / parts;
    etm::detail::parallel_for(parts, [&](int p) {
        const int64_t lo = p * per, hi = std::min(n, lo + per);
        for (int64_t i = lo * k; i < hi * k; ++i) v[val0 + i] = c2i ? c2i[rows[i]] : rows[i];
        for (int64_t i = lo; i < hi; ++i) {
            off[set0 + i + 1] = val0 + (i + 1) * k;
            c[set0 + i] = cnts[i];
            sp[set0 + i] = (double)cnts[i] / n_rows_d;
        }
    });
}

// The calling thread's stream for result copies (non-blocking, default priority) and its events, created on
// first use on a device and kept for the process.
hipStream_t export_stream() {
    static thread_local hipStream_t xs = nullptr;
    static thread_local int xs_device = -1;
    int dev = 0;
    check(hipGetDevice(&dev), "hipGetDevice");
    if (xs == nullptr || xs_device != dev) {
        check(hipStreamCreateWithFlags(&xs, hipStreamNonBlocking), "export stream");
        xs_device = dev;
    }
    return xs;
}

hipEvent_t export_event(size_t i) {
    static thread_local std::vector<hipEvent_t> pool;
    static thread_local int pool_device = -1;
    int dev = 0;
    check(hipGetDevice(&dev), "hipGetDevice");
    if (dev != pool_device) {
        pool.clear();  // events of another device stay with it
        pool_device = dev;
    }
    while (pool.size() <= i) {
        hipEvent_t e = nullptr;
        check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event");
        pool.push_back(e);
    }
    return pool[i];
}

// A mining call's exporter: a host thread that waits for each large level's copy to land in pinned memory and
// writes it into the frame while later levels are mined. finish() waits for every level pushed and rethrows the
// first failure; the destructor waits as well (copies read device memory the caller frees afterwards).
class Exporter {
   public:
    struct Task {
        hipEvent_t done;
        const int32_t* rows;
        const int32_t* cnts;
        int64_t n;
        int32_t k;
        int64_t set0, val0;
    };
    Exporter(etm_frame& f, const int64_t* c2i, double n_rows_d, int device)
        : f_(f), c2i_(c2i), n_rows_d_(n_rows_d), device_(device) {
        th_ = std::thread([this] { loop(); });
    }
    Exporter(const Exporter&) = delete;
    Exporter& operator=(const Exporter&) = delete;
    ~Exporter() { join(); }
    void push(const Task& t) {
        {
            std::lock_guard<std::mutex> lk(mu_);
            q_.push_back(t);
        }
        cv_.notify_one();
    }
    void finish() {
        join();
        if (err_) std::rethrow_exception(err_);
    }

   private:
    void join() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            closed_ = true;
        }
        cv_.notify_one();
        if (th_.joinable()) th_.join();
    }
    void loop() {
        (void)hipSetDevice(device_);
        for (;;) {
            Task t;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return closed_ || !q_.empty(); });
                if (q_.empty()) return;
                t = q_.front();
                q_.erase(q_.begin());
            }
            try {
                const hipError_t e = hipEventSynchronize(t.done);  // the copy has landed (or failed)
                if (err_) continue;
                check(e, "result copy");
                if (!f_.reserve(t.set0 + t.n, t.val0 + t.n * t.k))
                    throw Error(ETM_ENOMEM, "the host result frame could not grow");
                frame_write(f_, t.rows, t.cnts, t.n, t.k, t.set0, t.val0, c2i_, n_rows_d_);
            } catch (...) {
                if (!err_) err_ = std::current_exception();
            }
        }
    }
    etm_frame& f_;
    const int64_t* c2i_;
    double n_rows_d_;
    int device_;
    std::mutex mu_;
    std::condition_variable cv_;
    std::vector<Task> q_;
    bool closed_ = false;
    std::exception_ptr err_;
    std::thread th_;
};

// Levels of at least this many ints (rows and counts) are copied to the host and exported as soon as they are
// compacted; smaller ones come back in the final gather.
constexpr int64_t kExportEarlyInts = 1 << 18;

using Clock = std::chrono::steady_clock;

double ms_since(Clock::time_point t) {
    return std::chrono::duration<double, std::milli>(Clock::now() - t).count();
}

// Hands a finished level to the caller's callback (when there is one); a nonzero return stops the call.
void report(const etm_mine_opts& o, const etm_levels& out, const etm_level_stat& st, const int32_t* h_rows,
            const int32_t* h_counts, int64_t device_bytes, int64_t host_bytes, bool next_level,
            const int32_t* d_rows = nullptr, const int32_t* d_counts = nullptr) {
    if (o.level_cb == nullptr) return;
    etm_level_info info{};
    info.stat = st;
    info.h_rows = h_rows;
    info.h_counts = h_counts;
    info.d_rows = d_rows;
    info.d_counts = d_counts;
    info.h_col_to_item = out.col_to_item.empty() ? nullptr : out.col_to_item.data();
    info.n_cols = (int64_t)out.col_counts.size();
    info.device_bytes = device_bytes;
    info.host_bytes = host_bytes;
    info.next_level = next_level ? 1 : 0;
    if (o.level_cb(&info, o.cb_user) != 0)
        throw Error(ETM_ECANCELED, "the level callback stopped mining after K=" + std::to_string(st.k));
}

int64_t pow2_at_least(int64_t x) {
    int64_t v = 1;
    while (v < x) v <<= 1;
    return v;
}

struct Level {
    int32_t* rows = nullptr;
    int32_t* counts = nullptr;
    uint8_t* free = nullptr;
    int32_t* table = nullptr;
    int64_t table_size = 0;
    int64_t n = 0;
    int32_t m = 0;
};

// Device scalars of one level: [0] survivors, [1..3] tally, [4] selected (free-set prune), [5..7] live bytes,
// [8..31] totals of the next level's groups, [32..47] live-unit counters of the counting launches.
constexpr int kScal = 48;

struct GroupsDev {
    int64_t *gs, *cp, *ct;
    int32_t* list;
};

GroupsDev alloc_groups(Arena& a, int64_t n) {
    return GroupsDev{a.get<int64_t>(n + 1), a.get<int64_t>(n + 1), a.get<int64_t>(n + 1), a.get<int32_t>(n)};
}

// Group build of a level (single workgroup when small and d_n given; else the multi-block scans, which need n).
void build_groups(Arena& a, const Level& lv, const int64_t* d_n, int64_t n_cap, GroupsDev& g, int64_t* d_totals,
                  int32_t small_max, hipStream_t s) {
    if (etm::detail::single_fits(n_cap, lv.m)) {
        check(etm::detail::launch_groups_single(lv.rows, d_n, n_cap, lv.m, small_max, g.gs, g.cp, g.ct, g.list,
                                                d_totals, s),
              "group build");
        return;
    }
    const size_t ws_n = etm_level_groups_ws(n_cap);
    int64_t* ws = a.get<int64_t>((int64_t)(ws_n + 7) / 8);
    check_rc(etm_level_groups(lv.rows, n_cap, lv.m, small_max, g.gs, g.cp, g.ct, g.list, d_totals, ws, ws_n, s),
             "etm_level_groups");
}

// The id CSR and its id -> column map, when the caller mined from transactions: the row-wise K=2 / K=3
// kernels count from these rows.
struct RowsSrc {
    const int64_t* indptr;
    const void* ids;
    int id_bytes;
    int64_t nnz, lo, span;
    const int32_t* lut;
};

// Crossovers of the row-wise kernels: rows count K=2 when sum C(len, 2) < kK2RowsMaxR x pairs x words, and
// K=3 when it is < kK3RowsMaxR x K=3 candidates x words with no row over kK3RowsLongest positions. Their
// atomics resolve device-wide, so on MI355X they win only on much sparser levels than their row pairs alone
// suggest: wide_vocab_2m's K=2 (r = 9e-4) measured 10.5 ms row-wise against 9.0 ms on the MFMA Gram. The
// rows are built only when a level may take them: sum C(len, 2) >= (E^2/N - E)/2 for E K=1 entries over N rows
// (Cauchy-Schwarz) must leave the rule open.
constexpr double kK2RowsMaxR = 2.0e-4;
constexpr double kK3RowsMaxR = 2.5e-4;
constexpr int64_t kK3RowsLongest = 256;
constexpr int64_t kRowsBitsMax = 512ll << 20;
// Largest K=3 pair index (12 bytes per column x 64 columns: 75 MB at 20,000 columns).
constexpr int64_t kPairIndexMaxBytes = 256ll << 20;

// Levels K=1 .. exhaustion of device bitvectors. t_call: the call's start (K=1's wall time runs from it);
// base_bytes: device bytes the caller's arena already holds for this call (the input layer's).
void mine_levels(const uint64_t* d_bv, int64_t pitch, int64_t n_words, int64_t n_rows, const int64_t* col_counts,
                 int32_t n_cols, const etm_mine_opts& o, etm_levels& out, hipStream_t s, Clock::time_point t_call,
                 const RowsSrc* src = nullptr, bool ordered = false) {
    if (n_rows > 0x7FFFFFFF) throw Error(ETM_EOVERFLOW, "n_rows exceeds int32: the count arrays are int32");
    const int64_t mc = std::max<int64_t>(1, o.min_count);
    const bool prune = o.prune_non_free != 0;
    const bool track_free = prune || o.infer_counts != 0;
    const int64_t max_len = o.max_length > 0 ? std::min<int64_t>(o.max_length, n_cols) : n_cols;
    if (o.emit_levels < ETM_EMIT_NONE || o.emit_levels > ETM_EMIT_DEVICE)
        throw Error(ETM_EINVAL, "emit_levels must be one of ETM_EMIT_*");
    if (o.small_max != 0 && o.small_max != 4 && o.small_max != 8 && o.small_max != 16)
        throw Error(ETM_EINVAL, "small_max must be 0, 4, 8 or 16, got " + std::to_string(o.small_max));
    if (o.k3_schedule < ETM_SCHEDULE_AUTO || o.k3_schedule > ETM_SCHEDULE_LIVE)
        throw Error(ETM_EINVAL, "k3_schedule must be one of ETM_SCHEDULE_*");
    const int32_t small_max = o.small_max ? o.small_max : kSmallMaxDefault;
    const bool live_schedule = o.k3_schedule == ETM_SCHEDULE_LIVE ||
                               (o.k3_schedule == ETM_SCHEDULE_AUTO && (o.infer_counts != 0 || n_words >= kLiveMinWords));

    // K = 1.
    std::vector<int32_t> k1_rows;
    std::vector<int32_t> k1_counts;
    for (int32_t c = 0; c < n_cols; ++c) {
        if (col_counts[c] < mc) continue;
        if (prune && col_counts[c] >= n_rows) continue;
        k1_rows.push_back(c);
        k1_counts.push_back((int32_t)col_counts[c]);
    }
    const bool emit = o.emit_levels != 0;
    out.k1_rows = k1_rows;
    out.k1_counts = k1_counts;
    out.k.push_back(1);
    out.n.push_back(emit ? 0 : (int64_t)out.k1_counts.size());
    out.rows.push_back(out.k1_rows.data());
    out.counts.push_back(out.k1_counts.data());
    etm_level_stat st1{};
    st1.k = 1;
    st1.n_candidates = n_cols;
    st1.n_counted = n_cols;
    st1.n_frequent = (int64_t)k1_rows.size();
    st1.bytes_alg = (int64_t)n_cols * n_words * 8;
    st1.bytes_live = st1.bytes_alg;
    st1.wall_ms = (float)ms_since(t_call);
    out.stats.push_back(st1);
    Arena a(s);
    // Host bytes the kept levels take once copied back (K=1 lives on the host already).
    int64_t host_bytes = emit ? 0 : (int64_t)k1_rows.size() * 8;
    const int64_t n1_all = (int64_t)k1_rows.size();
    report(o, out, st1, emit ? k1_rows.data() : nullptr, emit ? k1_counts.data() : nullptr, a.bytes(), host_bytes,
           n1_all >= 2 && max_len >= 2);
    if (k1_rows.empty() || max_len < 2) return;
    Clock::time_point t_level = Clock::now();

    Pinned host;
    const int64_t n1 = (int64_t)k1_rows.size();
    Level prev, prev2, prev3;  // levels k-1, k-2, k-3 of the level being mined
    prev.m = 1;
    prev.n = n1;
    prev.rows = a.get<int32_t>(n1);
    prev.counts = a.get<int32_t>(n1);
    check(hipMemcpyAsync(prev.rows, k1_rows.data(), n1 * 4, hipMemcpyHostToDevice, s), "upload K=1");
    check(hipMemcpyAsync(prev.counts, k1_counts.data(), n1 * 4, hipMemcpyHostToDevice, s), "upload K=1 counts");
    // Host sources of asynchronous uploads live until the call's end: a copy from pageable memory may still
    // read its source after hipMemcpyAsync returns.
    std::vector<uint8_t> k1_free;
    if (track_free) {
        k1_free.resize((size_t)n1);
        for (int64_t i = 0; i < n1; ++i) k1_free[(size_t)i] = k1_counts[(size_t)i] < n_rows ? 1 : 0;
        prev.free = a.get<uint8_t>(n1);
        check(hipMemcpyAsync(prev.free, k1_free.data(), n1, hipMemcpyHostToDevice, s), "upload K=1 free");
    }
    GroupsDev g = alloc_groups(a, n1);
    int64_t* scal = a.get<int64_t>(kScal);
    // Rows as K=1 positions, built on first need (a forced row-wise level, or one the rule may give them).
    const bool rows_forced = o.k2_kernel == ETM_KERNEL_ROWS || o.k3_kernel == ETM_KERNEL_ROWS;
    if (rows_forced && src == nullptr)
        throw Error(ETM_EINVAL, "the row-wise kernels need the transactions (etm_mine_csr)");
    double k1_entries = 0;
    for (int64_t t = 0; t < n1; ++t) k1_entries += (double)col_counts[k1_rows[(size_t)t]];
    const double row_pairs_lb = std::max(0.0, (k1_entries * k1_entries / (double)n_rows - k1_entries) / 2);
    int64_t* rows_ptr = nullptr;
    int32_t* rows_pos = nullptr;
    int32_t* d_colpos = nullptr;
    std::vector<int32_t> colpos;  // host source of d_colpos, alive until the level loop's syncs
    bool rows_built = false, rows_ok = false;
    int64_t row_pairs = 0, row_longest = 0;
    auto ensure_rows = [&]() {
        if (rows_built) return;
        rows_built = true;
        colpos.assign((size_t)n_cols, -1);
        for (int64_t t = 0; t < n1; ++t) colpos[(size_t)k1_rows[(size_t)t]] = (int32_t)t;
        d_colpos = a.get<int32_t>(n_cols);
        check(hipMemcpyAsync(d_colpos, colpos.data(), (size_t)n_cols * 4, hipMemcpyHostToDevice, s), "upload colpos");
        rows_ptr = a.get<int64_t>(n_rows + 1);
        rows_pos = a.get<int32_t>(std::max<int64_t>(src->nnz, 1));
        int64_t* stats = a.get<int64_t>(4);
        const size_t ws_n = etm_rows_ws(n_rows);
        void* ws = a.get<int64_t>((int64_t)(ws_n + 7) / 8);
        check_rc(etm_rows_build(src->indptr, src->ids, src->id_bytes, n_rows, src->lo, src->span, src->lut, d_colpos,
                                rows_ptr, rows_pos, stats, ws, ws_n, s),
                 "etm_rows_build");
        check(hipMemcpyAsync(host.p + 32, stats, 32, hipMemcpyDeviceToHost, s), "readback");
        wait_stream(s);
        row_pairs = host.p[32];
        row_longest = host.p[33];
        rows_ok = host.p[34] == 0;
        if (rows_forced && !rows_ok)
            throw Error(ETM_EINVAL, "the row-wise kernels need rows of strictly ascending, distinct item ids");
    };
    // Tile occupancy of the columns, built at the first level whose group kernels can skip tiles: under the
    // ordered rows an itemset's rows gather, and the AND of a prefix is zero on most tiles.
    uint64_t* occ = nullptr;
    int64_t occ_pitch = 0;
    auto ensure_occ = [&]() {
        if (!ordered || occ != nullptr) return;
        occ_pitch = ((n_words + 63) / 64 + 63) / 64 + 1;
        occ = a.get<uint64_t>((int64_t)n_cols * occ_pitch);
        check_rc(etm_tile_occupancy(d_bv, pitch, n_words, n_cols, occ, occ_pitch, s), "etm_tile_occupancy");
    };
    build_groups(a, prev, nullptr, n1, g, scal + 8, small_max, s);
    check(hipMemcpyAsync(host.p, scal, kScal * 8, hipMemcpyDeviceToHost, s), "readback");
    wait_stream(s);
    int64_t tot[24];
    std::memcpy(tot, host.p + 8, sizeof tot);


    struct Pending {
        int32_t* rows;
        int32_t* counts;
        int64_t n;
        int32_t k;
        int32_t* host = nullptr;  // the pinned copy of a level exported early
        int64_t set0 = 0, val0 = 0;  // its place in the frame
    };
    std::vector<Pending> levels;
    // The frame: K=1 first, then each level in order; the exporter starts at the first large level.
    int64_t frame_sets = emit ? 0 : n1, frame_vals = emit ? 0 : n1;
    std::unique_ptr<Exporter> exporter;
    hipStream_t xs = nullptr;
    size_t n_export = 0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;

    // Device scalars of every level (k <= 62 + 2), zeroed once.
    constexpr int kMaxLevels = 64;
    int64_t* level_scal = a.get<int64_t>((int64_t)kMaxLevels * kScal);
    check(etm::detail::launch_zero_i64(level_scal, kMaxLevels * kScal, s), "zero");
    for (int32_t k = 2; k <= max_len && prev.n >= k && k - 2 < kMaxLevels; ++k) {
        const int32_t m = k - 1;
        const int64_t n_groups = tot[0], n_cands = tot[1], n_tiles = tot[2];
        if (n_cands == 0) break;
        etm_level_stat st{};
        st.k = k;
        st.n_candidates = n_cands;
        st.bytes_alg = ((int64_t)(m - 1) * tot[5] + tot[6]) * n_words * 8 + 4 * n_cands;
        // Hash of the previous level (subset test from K=3, free flags from K=2).
        if ((k >= 3 || track_free) && prev.table == nullptr) {
            prev.table_size = pow2_at_least(2 * std::max<int64_t>(1, prev.n));
            prev.table = a.get<int32_t>(prev.table_size);
            check(etm::detail::launch_hash(prev.rows, prev.n, nullptr, prev.m, prev.table, prev.table_size, s), "hash");
        }
        int64_t* cur_scal = level_scal + (int64_t)(k - 2) * kScal;
        uint8_t* status = nullptr;
        int32_t* gid = nullptr;
        int32_t* counts = a.get<int32_t>(n_cands);
        const int64_t classes[6] = {tot[8], tot[9], tot[10], tot[11], tot[12], tot[4]};
        if (k >= 3) {
            status = a.get<uint8_t>(n_cands);
            gid = a.get<int32_t>(n_cands);
            const bool infer = o.infer_counts && !prune && prev.free != nullptr;
            const int32_t mode = ETM_SUBSET_PRUNE | (infer ? ETM_SUBSET_INFER : 0);
            // At K=3 the one subset test of a candidate reads a pair index of level 2 (a bitmap word and a rank)
            // instead of probing its hash, where the index stays small.
            uint64_t* pair_bits = nullptr;
            int32_t* pair_rank = nullptr;
            int64_t pair_words = 0;
            if (k == 3 && etm::detail::pair_index_bytes(n_cols) <= kPairIndexMaxBytes) {
                pair_words = ((int64_t)n_cols + 63) / 64;
                pair_bits = a.get<uint64_t>((int64_t)n_cols * pair_words);
                pair_rank = a.get<int32_t>((int64_t)n_cols * pair_words);
                check(etm::detail::launch_pair_index(prev.rows, prev.n, n_cols, pair_bits, pair_rank, s),
                      "pair index");
            }
            check_rc(etm::detail::level_classify(prev.rows, prev.n, m, prev.counts, infer ? prev.free : nullptr,
                                                 prev.table, prev.table_size, g.gs, g.cp, n_groups, n_cands, g.list,
                                                 classes, mode, status, counts, gid, cur_scal + 1, pair_bits,
                                                 pair_rank, pair_words, s),
                     "etm_level_classify");
            // The deduction bounds need levels k-2 and k-3 complete (the inference mode keeps every level whole).
            if (infer && o.deduction != ETM_DEDUCTION_OFF && k <= 24) {
                check_rc(etm_level_deduce(prev.rows, m, prev.counts, prev.table, prev.table_size, prev2.rows,
                                          prev2.counts, prev2.table, prev2.table_size, prev3.rows, prev3.counts,
                                          prev3.table, prev3.table_size, n_rows, mc, g.gs, g.cp, n_groups, n_cands,
                                          gid, status, counts, cur_scal + 1, s),
                         "etm_level_deduce");
            }
        } else {
            check(hipMemsetAsync(counts, 0, n_cands * 4, s), "zero counts");
        }
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (o.time_kernels) {
            e0 = pooled_event(2 * events.size());
            e1 = pooled_event(2 * events.size() + 1);
            check(hipEventRecord(e0, s), "event record");
        }
        const int64_t cls[4] = {tot[8], tot[9], tot[10], tot[11]};
        // At K=2 the single group holds every column: the automatic kernel choice needs no device read.
        int32_t kern = ETM_KERNEL_VALU;
        if (k == 2) {
            const double limit = kK2RowsMaxR * (double)n_cands * (double)n_words;
            if (o.k2_kernel == ETM_KERNEL_ROWS ||
                (o.k2_kernel == ETM_KERNEL_AUTO && src != nullptr && row_pairs_lb < limit)) {
                ensure_rows();
                if (o.k2_kernel == ETM_KERNEL_ROWS || (rows_ok && (double)row_pairs < limit)) kern = ETM_KERNEL_ROWS;
            }
            if (kern != ETM_KERNEL_ROWS)
                kern = o.k2_kernel != ETM_KERNEL_AUTO ? o.k2_kernel
                       : etm::detail::gram_preferred((prev.n + 63) / 64 * 64, n_words) ? ETM_KERNEL_MFMA : ETM_KERNEL_VALU;
        } else if (k == 3) {
            const int64_t words1 = (n1 + 63) / 64;
            const bool fits = n1 <= 65536 && n1 * words1 * 12 <= kRowsBitsMax;
            const double limit = kK3RowsMaxR * (double)n_cands * (double)n_words;
            if (o.k3_kernel == ETM_KERNEL_ROWS ||
                (o.k3_kernel == ETM_KERNEL_AUTO && src != nullptr && fits && row_pairs_lb < limit)) {
                ensure_rows();
                if (o.k3_kernel == ETM_KERNEL_ROWS ||
                    (rows_ok && row_longest <= kK3RowsLongest && (double)row_pairs < limit))
                    kern = ETM_KERNEL_ROWS;
            }
        }
        st.kernel = kern;
        if (kern == ETM_KERNEL_ROWS && k == 2) {
            check_rc(etm_count_k2_rows(rows_ptr, rows_pos, n_rows, n1, counts, s), "etm_count_k2_rows");
        } else if (kern == ETM_KERNEL_ROWS) {
            const int64_t words1 = (n1 + 63) / 64;
            uint64_t* bits = a.get<uint64_t>(n1 * words1);
            int32_t* rank = a.get<int32_t>(n1 * words1);
            int32_t* gof = a.get<int32_t>(n1);
            check_rc(etm_k3_rows_prepare(prev.rows, prev.n, d_colpos, n1, g.gs, n_groups, bits, rank, gof, s),
                     "etm_k3_rows_prepare");
            check_rc(etm_count_k3_rows(rows_ptr, rows_pos, n_rows, bits, rank, gof, n1, g.gs, g.cp, status, counts,
                                       s),
                     "etm_count_k3_rows");
        } else {
            const bool skip_tiles = ordered && k >= 3 && cls[0] + cls[1] + cls[2] + cls[3] > 0;
            if (skip_tiles) ensure_occ();
            if (status != nullptr && !skip_tiles && live_schedule) {
                // From K=3 on, only the groups and tile-pairs holding a counted candidate are launched, and the
                // tiled groups with few counted pairs are counted group-listed; their counters are the level's
                // scalars [32..47], zeroed with the others.
                const size_t ws_n = etm_count_level_ws(cls[0] + cls[1] + cls[2] + cls[3], n_tiles);
                void* ws = a.get<int64_t>((int64_t)(ws_n + 7) / 8);
                const int64_t n_tiled = tot[12];
                const size_t gl_n = etm::detail::count_glist_ws(n_groups, n_tiled, n_cands, prev.n);
                void* gl_ws = gl_n ? a.get<int64_t>((int64_t)(gl_n + 7) / 8) : nullptr;
                check_rc(etm::detail::count_level_live(d_bv, pitch, n_words, prev.rows, m, g.gs, g.cp, g.ct, n_groups,
                                                       n_tiles, g.list, cls, status, counts, ws, cur_scal + 32,
                                                       n_tiled, n_cands, prev.n, gl_ws, s),
                         "etm_count_level");
            } else {
                check_rc(etm_count_level(d_bv, pitch, n_words, prev.rows, m, g.gs, g.cp, g.ct, n_groups, n_tiles,
                                         g.list, cls, status, counts, ETM_LAYOUT_LEX, 0, kern,
                                         skip_tiles ? occ : nullptr, occ_pitch, nullptr, 0, s),
                         "etm_count_level");
            }
        }
        if (o.time_kernels) {
            check(hipEventRecord(e1, s), "event record");
            events.emplace_back(e0, e1);
        }
        // What the counting kernels had to read for the candidates left counted (cur_scal[5..7]).
        if (o.live_bytes && status != nullptr) {
            uint64_t* row_bits = a.get<uint64_t>((prev.n + 63) / 64);
            uint64_t* group_bits = a.get<uint64_t>((n_groups + 63) / 64);
            check(etm::detail::launch_live_bytes(status, gid, g.gs, g.cp, n_groups, n_cands, prev.n, row_bits,
                                                 group_bits, cur_scal + 5, s),
                  "live bytes");
        }
        // Survivors, in lexicographic order.
        Level cur;
        cur.m = k;
        int64_t cap;
        bool cap_exact = false;  // cap is the survivors' count itself, read back (not a bound)
        GroupsDev gn{};
        if (etm::detail::single_fits(n_cands, k)) {
            cap = n_cands;
            cur.rows = a.get<int32_t>(cap * k);
            cur.counts = a.get<int32_t>(cap);
            check(etm::detail::launch_compact_single(prev.rows, m, g.gs, g.cp, n_groups, n_cands, status, counts, gid,
                                                     mc, cur.rows, cur.counts, nullptr, cur_scal + 0, s),
                  "compaction");
        } else {
            const size_t ws_n = etm_level_compact_ws(n_cands);
            int64_t* ws = a.get<int64_t>((int64_t)(ws_n + 7) / 8);
            check_rc(etm_level_compact(prev.rows, m, g.gs, g.cp, n_groups, n_cands, status, counts, gid, mc,
                                       nullptr, nullptr, nullptr, cur_scal + 0, ws, ws_n, s),
                     "etm_level_compact");
            check(hipMemcpyAsync(host.p, cur_scal, 8, hipMemcpyDeviceToHost, s), "readback");
            wait_stream(s);
            cap = host.p[0];
            cap_exact = true;
            cur.rows = a.get<int32_t>(cap * k);
            cur.counts = a.get<int32_t>(cap);
            if (cap && track_free && status != nullptr) {
                // The free flags start as hints: the survivors with inferred counts are not free (no lookup).
                cur.free = a.get<uint8_t>(cap);
                check(etm::detail::level_compact_hinted(prev.rows, m, g.gs, g.cp, n_groups, n_cands, status, counts,
                                                        gid, mc, cur.rows, cur.counts, cur.free, ws, s),
                      "etm_level_compact");
            } else if (cap) {
                check_rc(etm_level_compact(prev.rows, m, g.gs, g.cp, n_groups, n_cands, status, counts, gid, mc,
                                           cur.rows, cur.counts, nullptr, cur_scal + 0, ws, ws_n, s),
                         "etm_level_compact");
            }
        }
        const int64_t* d_n = cur_scal + 0;
        if (track_free) {
            const bool hinted = cur.free != nullptr;
            if (!hinted) cur.free = a.get<uint8_t>(cap);
            check(etm::detail::launch_free(cur.rows, cap, d_n, k, cur.counts, prev.rows, prev.counts, prev.table,
                                           prev.table_size, cur.free, s, hinted),
                  "free flags");
            if (prune) {
                Level sel;
                sel.m = k;
                sel.rows = a.get<int32_t>(cap * k);
                sel.counts = a.get<int32_t>(cap);
                if (etm::detail::single_fits(cap, k)) {
                    check(etm::detail::launch_select_single(cur.rows, d_n, cap, k, cur.counts, cur.free, sel.rows,
                                                            sel.counts, cur_scal + 4, s),
                          "select");
                } else {
                    check(hipMemcpyAsync(host.p, cur_scal, 8, hipMemcpyDeviceToHost, s), "readback");
                    wait_stream(s);
                    const int64_t nn = host.p[0];
                    const size_t ws_n = etm_level_compact_ws(nn);
                    int64_t* ws = a.get<int64_t>((int64_t)(ws_n + 7) / 8);
                    check_rc(etm_select_rows(cur.rows, nn, k, cur.counts, cur.free, nullptr, nullptr, cur_scal + 4, ws,
                                             ws_n, s),
                             "etm_select_rows");
                    check_rc(etm_select_rows(cur.rows, nn, k, cur.counts, cur.free, sel.rows, sel.counts,
                                             cur_scal + 4, ws, ws_n, s),
                             "etm_select_rows");
                }
                sel.free = a.get<uint8_t>(cap);
                check(etm::detail::launch_fill_u8(sel.free, cap, 1, s), "fill");
                cur.rows = sel.rows;
                cur.counts = sel.counts;
                cur.free = sel.free;
                d_n = cur_scal + 4;
            }
        }
        // The next level's groups, sized on the device when the level is small. A large level whose size the
        // host already holds (no free-set selection after the count) builds them here as well, so that the
        // readback below serves the level's scalars and the group totals at once.
        const bool early = etm::detail::single_fits(cap, k) && k + 1 <= max_len;
        const bool prebuilt = !early && cap_exact && !prune && cap > 0 && k + 1 <= max_len;
        if (early) {
            gn = alloc_groups(a, cap);
            check(etm::detail::launch_groups_single(cur.rows, d_n, cap, k, small_max, gn.gs, gn.cp, gn.ct, gn.list,
                                                    cur_scal + 8, s),
                  "group build");
        } else if (prebuilt) {
            gn = alloc_groups(a, cap);
            build_groups(a, cur, nullptr, cap, gn, cur_scal + 8, small_max, s);
        }
        check(hipMemcpyAsync(host.p, cur_scal, kScal * 8, hipMemcpyDeviceToHost, s), "readback");
        wait_stream(s);
        cur.n = host.p[prune ? 4 : 0];
        st.n_counted = k >= 3 ? host.p[1] : n_cands;
        st.n_inferred = k >= 3 ? host.p[2] : 0;
        st.n_skipped = k >= 3 ? host.p[3] : 0;
        st.n_frequent = cur.n;
        st.bytes_live = (o.live_bytes && status != nullptr)
                            ? ((int64_t)(m - 1) * host.p[6] + host.p[5]) * n_words * 8 + 4 * host.p[7]
                            : (status == nullptr ? st.bytes_alg : 0);
        const int32_t* h_rows = nullptr;
        const int32_t* h_counts = nullptr;
        if (o.emit_levels == ETM_EMIT_HOST && cur.n > 0 && o.level_cb) {
            int32_t* buf = staging(cur.n * (k + 1));
            check(hipMemcpyAsync(buf, cur.rows, (size_t)(cur.n * k) * 4, hipMemcpyDeviceToHost, s), "level out");
            check(hipMemcpyAsync(buf + cur.n * k, cur.counts, (size_t)cur.n * 4, hipMemcpyDeviceToHost, s),
                  "level out");
            wait_stream(s);
            h_rows = buf;
            h_counts = buf + cur.n * k;
        }
        st.wall_ms = (float)ms_since(t_level);
        out.stats.push_back(st);
        if (emit) {
            levels.push_back(Pending{nullptr, nullptr, 0, k});
        } else {
            Pending pd{cur.rows, cur.counts, cur.n, k};
            pd.set0 = frame_sets;
            pd.val0 = frame_vals;
            frame_sets += cur.n;
            frame_vals += cur.n * k;
            if (cur.n * (int64_t)(k + 1) >= kExportEarlyInts) {
                // The level is final (the readback above waited for its compaction): copy it to the host on the
                // export stream and let the exporter write it into the frame while the next levels run.
                if (!exporter) {
                    out.frame = new etm_frame();
                    int dev = 0;
                    check(hipGetDevice(&dev), "hipGetDevice");
                    exporter = std::make_unique<Exporter>(*out.frame,
                                                          out.col_to_item.empty() ? nullptr : out.col_to_item.data(),
                                                          (double)n_rows, dev);
                    xs = export_stream();
                }
                int64_t bcap = 0;
                pd.host = pool_take(cur.n * (k + 1), &bcap);
                out.level_blocks.emplace_back(pd.host, bcap);
                check(hipMemcpyAsync(pd.host, cur.rows, (size_t)(cur.n * k) * 4, hipMemcpyDeviceToHost, xs),
                      "level copy");
                check(hipMemcpyAsync(pd.host + cur.n * k, cur.counts, (size_t)cur.n * 4, hipMemcpyDeviceToHost, xs),
                      "level copy");
                const hipEvent_t done = export_event(n_export++);
                check(hipEventRecord(done, xs), "event record");
                exporter->push(Exporter::Task{done, pd.host, pd.host + cur.n * k, cur.n, k, pd.set0, pd.val0});
            }
            levels.push_back(pd);
        }
        if (!emit) host_bytes += cur.n * (int64_t)(k + 1) * 4;
        report(o, out, st, h_rows, h_counts, a.bytes(), host_bytes, k + 1 <= max_len && cur.n >= k + 1,
               cur.n > 0 ? cur.rows : nullptr, cur.n > 0 ? cur.counts : nullptr);
        t_level = Clock::now();
        if (cur.n == 0) break;
        if (early || prebuilt) {
            std::memcpy(tot, host.p + 8, sizeof tot);
            g = gn;
        } else if (k + 1 <= max_len) {
            g = alloc_groups(a, cur.n);
            build_groups(a, cur, nullptr, cur.n, g, cur_scal + 8, small_max, s);
            check(hipMemcpyAsync(host.p, cur_scal, kScal * 8, hipMemcpyDeviceToHost, s), "readback");
            wait_stream(s);
            std::memcpy(tot, host.p + 8, sizeof tot);
        }
        prev3 = prev2;
        prev2 = prev;
        prev = cur;
    }
    // Every level not yet copied back to the host in one pass, through one pinned staging buffer.
    int64_t total_ints = 0;
    for (const Pending& p : levels)
        if (!p.host) total_ints += p.n * (p.k + 1);
    int64_t cap_ints = 0;
    int32_t* stage = pool_take(std::max<int64_t>(total_ints, 1), &cap_ints);
    out.block = stage;
    out.block_cap = cap_ints;
    int64_t at = 0;
    if (total_ints > 0) {
        int32_t* dev_stage = a.get<int32_t>(total_ints);
        Segs segs{};
        int ns = 0;
        int64_t longest = 0;
        auto flush = [&]() {
            if (ns == 0) return;
            const int64_t bx = std::min<int64_t>(std::max<int64_t>((longest + 255) / 256, 1), 1024);
            k_gather<<<dim3((unsigned)bx, (unsigned)ns), 256, 0, s>>>(segs, dev_stage);
            check(hipGetLastError(), "gather");
            segs = Segs{};
            ns = 0;
            longest = 0;
        };
        for (const Pending& p : levels) {
            if (p.host) continue;
            if (p.n) {
                segs.s[ns++] = Seg{p.rows, p.n * p.k, at};
                segs.s[ns++] = Seg{p.counts, p.n, at + p.n * p.k};
                longest = std::max(longest, p.n * p.k);
                if (ns + 2 > kGatherSegs) flush();
            }
            at += p.n * (p.k + 1);
        }
        flush();
        check(hipMemcpyAsync(stage, dev_stage, (size_t)total_ints * 4, hipMemcpyDeviceToHost, s), "copy out");
    }
    wait_stream(s);
    at = 0;
    for (const Pending& p : levels) {
        out.k.push_back(p.k);
        out.n.push_back(p.n);
        if (p.host) {
            out.rows.push_back(p.host);
            out.counts.push_back(p.host + p.n * p.k);
            continue;
        }
        out.rows.push_back(stage + at);
        out.counts.push_back(stage + at + p.n * p.k);
        at += p.n * (p.k + 1);
    }
    if (exporter) {
        // The frame: the exporter's levels, then K=1 and the levels of the final gather.
        exporter->finish();
        etm_frame& f = *out.frame;
        if (!f.reserve(frame_sets, frame_vals)) throw Error(ETM_ENOMEM, "the host result frame could not grow");
        ((int64_t*)f.offsets.data())[0] = 0;
        const int64_t* c2i = out.col_to_item.empty() ? nullptr : out.col_to_item.data();
        frame_write(f, out.k1_rows.data(), out.k1_counts.data(), n1, 1, 0, 0, c2i, (double)n_rows);
        for (size_t l = 0; l < levels.size(); ++l) {
            const Pending& p = levels[l];
            if (!p.host) frame_write(f, out.rows[l + 1], out.counts[l + 1], p.n, p.k, p.set0, p.val0, c2i, (double)n_rows);
        }
        f.n_sets = frame_sets;
        f.n_vals = frame_vals;
    }
    for (size_t i = 0; i < events.size(); ++i) {
        float ms = 0;
        check(hipEventElapsedTime(&ms, events[i].first, events[i].second), "event time");
        out.stats[i + 1].count_ms = ms;
    }
}

// ── batches of jobs (etm_mine_batch) ──────────────────────────────────────
//
// The input layers of all jobs run phase by phase (id ranges, histograms, bitvectors and K=1 counts), each phase
// enqueued for every job before one host sync. The jobs whose K=1 level is narrow enough then advance in lockstep
// rounds (batch.h): one round is a fixed launch sequence over the job records, replayed from a HIP graph
// captured once per (records, queue, batch width) on the thread, with up to two rounds in flight and one
// scalar read back per round. The other jobs, and any that outgrow the rounds, take the per-job loop.

namespace bt = etm::batch;

// A thread's pinned host buffer for the batch phases (grown on demand, kept for the process). Every phase takes
// it once, for all it needs, and ends with a stream wait before the next phase takes it again.
void* batch_pinned(size_t bytes) {
    static thread_local void* buf = nullptr;
    static thread_local size_t cap = 0;
    if (bytes > cap) {
        if (buf) (void)hipHostFree(buf);
        buf = nullptr;
        cap = 0;
        const size_t want = std::max<size_t>(bytes, 1 << 20);
        check(hipHostMalloc(&buf, want, hipHostMallocDefault), "hipHostMalloc");
        cap = want;
    }
    return buf;
}

// Ints of a batched job's level store: room for about six full levels (32,768 candidates of up to 10 items);
// a job that needs more ends its rounds as an overflow and takes the per-job loop.
constexpr int64_t kBatchOutInts = 2ll << 20;
// Jobs of a round graph's width come in multiples of this (the slots past the batch are marked done), so a
// few captured graphs serve every batch size.
constexpr int kBatchWidthStep = 16;
constexpr int kBatchInFlight = 2;  // rounds enqueued ahead of the host's read of the active count
constexpr int kBatchRoundsMax = bt::kMaxLevels + 2;

void enqueue_round(bt::Job* jobs, int nb, bt::Queue* q, hipStream_t s) {
    check(etm::detail::launch_batch_plan(jobs, nb, q, s), "batch plan");
    check(etm::detail::launch_batch_hash(jobs, nb, s), "batch hash");
    check(etm::detail::launch_batch_classify(jobs, nb, s), "batch classify");
    check(etm::detail::launch_batch_count(jobs, nb, q, s), "batch count");
    check(etm::detail::launch_batch_compact(jobs, nb, s), "batch compaction");
    check(etm::detail::launch_batch_groups(jobs, nb, q, false, s), "batch groups");
}

// Whether rounds replay a captured graph (ETM_BATCH_GRAPH=0: plain launches, for comparison).
bool batch_graphs() {
    static const bool on = [] {
        const char* v = std::getenv("ETM_BATCH_GRAPH");
        return !(v && v[0] == '0');
    }();
    return on;
}

struct RoundGraph {
    int dev;
    const void* jobs;
    const void* q;
    int nb;
    hipGraphExec_t exec;
};

// The executable graph of one round over (jobs, nb, q) on the current device: captured on a private stream of
// this thread the first time, then reused (the arena gives the records and the queue the same addresses call
// after call). At most eight are kept per thread.
hipGraphExec_t round_graph(bt::Job* jobs, int nb, bt::Queue* q) {
    static thread_local std::vector<RoundGraph> cache;
    static thread_local hipStream_t cap_stream = nullptr;
    static thread_local int cap_dev = -1;
    int dev = 0;
    check(hipGetDevice(&dev), "hipGetDevice");
    for (const RoundGraph& g : cache)
        if (g.dev == dev && g.jobs == jobs && g.q == q && g.nb == nb) return g.exec;
    if (cap_dev != dev) {
        check(hipStreamCreateWithFlags(&cap_stream, hipStreamNonBlocking), "capture stream");
        cap_dev = dev;
    }
    check(hipStreamBeginCapture(cap_stream, hipStreamCaptureModeThreadLocal), "begin capture");
    try {
        enqueue_round(jobs, nb, q, cap_stream);
    } catch (...) {
        hipGraph_t broken = nullptr;
        (void)hipStreamEndCapture(cap_stream, &broken);
        if (broken) (void)hipGraphDestroy(broken);
        throw;
    }
    hipGraph_t graph = nullptr;
    check(hipStreamEndCapture(cap_stream, &graph), "end capture");
    hipGraphExec_t exec = nullptr;
    const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    check(e, "graph instantiate");
    if (cache.size() >= 8) {
        (void)hipGraphExecDestroy(cache.front().exec);
        cache.erase(cache.begin());
    }
    cache.push_back(RoundGraph{dev, jobs, q, nb, exec});
    return exec;
}

// Device wall-clock ticks per millisecond (hipDeviceAttributeWallClockRate is in kHz).
double wall_ticks_per_ms() {
    static thread_local int dev_seen = -1;
    static thread_local double rate = 1e5;
    int dev = 0;
    check(hipGetDevice(&dev), "hipGetDevice");
    if (dev != dev_seen) {
        int khz = 0;
        check(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev), "wall clock rate");
        rate = khz > 0 ? (double)khz : 1e5;
        dev_seen = dev;
    }
    return rate;
}

// What a batch keeps per job on the host.
struct BatchHost {
    int64_t lo = 0, span = 0, hist_off = 0, n_words = 0, pitch = 0, pc_off = 0, mc = 1, max_len = 0;
    int32_t n_cols = 0;
    uint64_t* bv = nullptr;
    std::vector<int32_t> k1;
};

// Level 1 of a batched job into its result, as the per-job loop writes it (no free-set mode here).
void batch_level_one(etm_levels& out, const BatchHost& h, Clock::time_point t_call) {
    out.k1_rows = h.k1;
    out.k1_counts.resize(h.k1.size());
    for (size_t t = 0; t < h.k1.size(); ++t) out.k1_counts[t] = (int32_t)out.col_counts[(size_t)h.k1[t]];
    out.k.push_back(1);
    out.n.push_back((int64_t)out.k1_rows.size());
    out.rows.push_back(out.k1_rows.data());
    out.counts.push_back(out.k1_counts.data());
    etm_level_stat st1{};
    st1.k = 1;
    st1.n_candidates = h.n_cols;
    st1.n_counted = h.n_cols;
    st1.n_frequent = (int64_t)h.k1.size();
    st1.bytes_alg = (int64_t)h.n_cols * h.n_words * 8;
    st1.bytes_live = st1.bytes_alg;
    st1.wall_ms = (float)ms_since(t_call);
    out.stats.push_back(st1);
}

// Rounds over the batched jobs ids[0..nj): on return each job's levels K >= 2 are in its result, or the job is
// listed in `overflow`.
void batch_rounds(Arena& a, bt::Job* d_jobs, bt::Queue* d_q, const std::vector<int>& ids,
                  std::vector<BatchHost>& hj, std::vector<std::unique_ptr<etm_levels>>& res,
                  std::vector<int>& overflow, hipStream_t s) {
    const int nj = (int)ids.size();
    const int nb = std::min(bt::kMaxJobs, (nj + kBatchWidthStep - 1) / kBatchWidthStep * kBatchWidthStep);
    const int64_t cap = bt::kCapCands;
    int64_t l1_ints = 0;
    for (int i : ids) l1_ints += 2 * (int64_t)hj[(size_t)i].k1.size();
    const size_t rec_bytes = (size_t)nb * sizeof(bt::Job);
    const size_t l1_bytes = ((size_t)l1_ints * 4 + 255) & ~(size_t)255;
    char* hp = (char*)batch_pinned(rec_bytes + l1_bytes + 8 * kBatchRoundsMax);
    bt::Job* hr = (bt::Job*)hp;
    int32_t* h_l1 = (int32_t*)(hp + rec_bytes);
    int64_t* h_act = (int64_t*)(hp + rec_bytes + l1_bytes);
    std::memset(hr, 0, rec_bytes);
    int32_t* d_l1 = a.get<int32_t>(l1_ints);
    int64_t at = 0;
    for (int t = 0; t < nj; ++t) {
        const BatchHost& h = hj[(size_t)ids[(size_t)t]];
        bt::Job& J = hr[t];
        J.bv = h.bv;
        J.pitch = h.pitch;
        J.n_words = h.n_words;
        J.min_count = h.mc;
        J.max_len = (int32_t)std::min<int64_t>(h.max_len, bt::kMaxLevels + 1);
        J.small_max = 8;
        J.table = a.get<int32_t>(bt::kTableSize);
        J.status = a.get<uint8_t>(cap);
        J.counts = a.get<int32_t>(cap);
        J.gid = a.get<int32_t>(cap);
        for (int q = 0; q < 2; ++q) {
            J.gs[q] = a.get<int64_t>(cap + 1);
            J.cp[q] = a.get<int64_t>(cap + 1);
            J.ct[q] = a.get<int64_t>(cap + 1);
            J.list[q] = a.get<int32_t>(cap);
        }
        J.out_cap = kBatchOutInts;
        J.out = a.get<int32_t>(kBatchOutInts);
        const int64_t n1 = (int64_t)h.k1.size();
        for (int64_t c = 0; c < n1; ++c) {
            h_l1[at + c] = h.k1[(size_t)c];
            h_l1[at + n1 + c] = (int32_t)res[(size_t)ids[(size_t)t]]->col_counts[(size_t)h.k1[(size_t)c]];
        }
        J.prev_rows = d_l1 + at;
        J.prev_counts = d_l1 + at + n1;
        J.prev_n = n1;
        at += 2 * n1;
        J.k = 2;
        J.state = bt::kActive;
    }
    for (int t = nj; t < nb; ++t) hr[t].state = bt::kDone;
    check(hipMemcpyAsync(d_jobs, hr, rec_bytes, hipMemcpyHostToDevice, s), "upload job records");
    if (l1_ints) check(hipMemcpyAsync(d_l1, h_l1, (size_t)l1_ints * 4, hipMemcpyHostToDevice, s), "upload level 1");
    check(etm::detail::launch_batch_groups(d_jobs, nb, d_q, true, s), "batch groups");
    hipGraphExec_t exec = batch_graphs() ? round_graph(d_jobs, nb, d_q) : nullptr;
    int launched = 0, seen = 0;
    bool more = true;
    while (seen < launched || (more && launched < kBatchRoundsMax)) {
        if (more && launched < kBatchRoundsMax && launched - seen < kBatchInFlight) {
            if (exec) check(hipGraphLaunch(exec, s), "round graph");
            else enqueue_round(d_jobs, nb, d_q, s);
            check(hipMemcpyAsync(h_act + launched, &d_q->active_after, 8, hipMemcpyDeviceToHost, s), "readback");
            check(hipEventRecord(pooled_event((size_t)(launched % (2 * kBatchInFlight))), s), "event record");
            ++launched;
            continue;
        }
        const hipEvent_t ev = pooled_event((size_t)(seen % (2 * kBatchInFlight)));
        for (;;) {
            const hipError_t e = hipEventQuery(ev);
            if (e == hipSuccess) break;
            if (e != hipErrorNotReady) check(e, "round event");
            __builtin_ia32_pause();
        }
        if (h_act[seen] == 0) more = false;
        ++seen;
    }
    // The records back, then every job's level store in one gathered copy.
    bt::Job* back = (bt::Job*)batch_pinned(rec_bytes);
    check(hipMemcpyAsync(back, d_jobs, rec_bytes, hipMemcpyDeviceToHost, s), "readback job records");
    wait_stream(s);
    std::vector<bt::Job> rec(back, back + nj);
    int64_t total = 0;
    std::vector<int64_t> base((size_t)nj, 0);
    for (int t = 0; t < nj; ++t) {
        if (rec[(size_t)t].state != bt::kDone) continue;
        base[(size_t)t] = total;
        total += rec[(size_t)t].out_used;
    }
    int32_t* h_out = nullptr;
    if (total > 0) {
        int32_t* dev_stage = a.get<int32_t>(total);
        Segs segs{};
        int ns = 0;
        int64_t longest = 0;
        auto flush = [&]() {
            if (ns == 0) return;
            const int64_t bx = std::min<int64_t>(std::max<int64_t>((longest + 255) / 256, 1), 1024);
            k_gather<<<dim3((unsigned)bx, (unsigned)ns), 256, 0, s>>>(segs, dev_stage);
            check(hipGetLastError(), "gather");
            segs = Segs{};
            ns = 0;
            longest = 0;
        };
        for (int t = 0; t < nj; ++t) {
            const bt::Job& J = rec[(size_t)t];
            if (J.state != bt::kDone || J.out_used == 0) continue;
            segs.s[ns++] = Seg{J.out, J.out_used, base[(size_t)t]};
            longest = std::max(longest, J.out_used);
            if (ns == kGatherSegs) flush();
        }
        flush();
        h_out = (int32_t*)batch_pinned((size_t)total * 4);
        check(hipMemcpyAsync(h_out, dev_stage, (size_t)total * 4, hipMemcpyDeviceToHost, s), "copy out");
        wait_stream(s);
    }
    const double tpm = wall_ticks_per_ms();
    for (int t = 0; t < nj; ++t) {
        const bt::Job& J = rec[(size_t)t];
        const int i = ids[(size_t)t];
        if (J.state != bt::kDone) {
            overflow.push_back(i);
            continue;
        }
        etm_levels& out = *res[(size_t)i];
        out.owned.assign(h_out + base[(size_t)t], h_out + base[(size_t)t] + J.out_used);
        int64_t t_prev = J.t_init;
        for (int l = 0; l < J.n_levels; ++l) {
            const bt::Level& L = J.lv[l];
            const int32_t k = l + 2;
            out.k.push_back(k);
            out.n.push_back(L.n);
            out.rows.push_back(out.owned.data() + L.rows_off);
            out.counts.push_back(out.owned.data() + L.counts_off);
            etm_level_stat st{};
            st.k = k;
            st.kernel = ETM_KERNEL_VALU;
            st.n_candidates = L.cands;
            st.n_counted = L.counted;
            st.n_skipped = L.skipped;
            st.n_frequent = L.n;
            st.bytes_alg = L.bytes_alg;
            st.bytes_live = k == 2 ? L.bytes_alg : 0;
            st.wall_ms = (float)((double)(L.t_end - t_prev) / tpm);
            t_prev = L.t_end;
            out.stats.push_back(st);
        }
    }
}

void mine_batch(const etm_batch_job* in, int n, int64_t span_limit, const etm_mine_opts& o,
                std::vector<std::unique_ptr<etm_levels>>& res, hipStream_t s, Clock::time_point t_call) {
    Arena a(s);
    // The records and the queue first: their addresses then repeat from call to call on this thread, and so does
    // the round graph.
    bt::Job* d_jobs = a.get<bt::Job>(bt::kMaxJobs);
    bt::Queue* d_q = a.get<bt::Queue>(1);
    std::vector<BatchHost> hj((size_t)n);
    auto has_data = [&](int i) { return in[i].nnz > 0 && in[i].n_rows > 0; };
    // The input layer, one launch per phase over every job (etm::batch::InJob), one host sync per phase.
    std::vector<bt::InJob> ij((size_t)n);
    bt::InJob* d_in = a.get<bt::InJob>(n);
    int64_t* d_pre = a.get<int64_t>(4 * ((int64_t)n + 1));
    int64_t* d_mm = a.get<int64_t>(2 * (int64_t)n);
    const size_t table_bytes = (size_t)n * sizeof(bt::InJob);
    const size_t pre_bytes = ((size_t)n + 1) * 8;
    // Exclusive prefix of phase p's workgroups over the jobs into pre; returns the total.
    auto prefix = [&](int p, int64_t* pre) {
        int64_t t = 0;
        for (int i = 0; i < n; ++i) {
            pre[i] = t;
            t += has_data(i) ? etm::detail::batch_input_blocks(p, ij[(size_t)i]) : 0;
        }
        pre[n] = t;
        return t;
    };
    for (int i = 0; i < n; ++i) {
        bt::InJob& J = ij[(size_t)i];
        J.indptr = in[i].d_indptr;
        J.ids = in[i].d_ids;
        J.n_rows = in[i].n_rows;
        J.nnz = has_data(i) ? in[i].nnz : 0;
        J.id_bytes = in[i].id_bytes;
        J.mm = d_mm + 2 * i;
    }
    // Id ranges.
    int64_t total_span = 0;
    {
        char* hp = (char*)batch_pinned(table_bytes + pre_bytes + 16 * (size_t)n);
        int64_t* h_pre = (int64_t*)(hp + table_bytes);
        int64_t* h_mm = h_pre + n + 1;
        std::memcpy(hp, ij.data(), table_bytes);
        const int64_t blocks = prefix(0, h_pre);
        check(hipMemcpyAsync(d_in, hp, table_bytes, hipMemcpyHostToDevice, s), "upload input records");
        check(hipMemcpyAsync(d_pre, h_pre, pre_bytes, hipMemcpyHostToDevice, s), "upload input records");
        check(etm::detail::launch_batch_input(0, d_in, n, d_pre, blocks, 0, s), "batch id ranges");
        check(hipMemcpyAsync(h_mm, d_mm, (size_t)n * 16, hipMemcpyDeviceToHost, s), "readback");
        wait_stream(s);
        for (int i = 0; i < n; ++i) {
            if (!has_data(i)) continue;
            BatchHost& b = hj[(size_t)i];
            b.lo = h_mm[2 * i];
            b.span = h_mm[2 * i + 1] - h_mm[2 * i] + 1;
            if (b.span > span_limit)
                throw Error(ETM_EINVAL, "job " + std::to_string(i) + ": item ids span " + std::to_string(b.span) +
                                            " values: code them to dense ids first");
            b.hist_off = total_span;
            total_span += b.span;
        }
    }
    // Histograms down, LUTs up, bitvectors and K=1 counts. The pinned block: records, three prefixes, histograms,
    // LUTs, K=1 counts (at most one per id).
    int32_t* d_hist = a.get<int32_t>(total_span);
    int32_t* d_lut = a.get<int32_t>(total_span);
    char* hp = (char*)batch_pinned(table_bytes + 3 * pre_bytes + 16 * (size_t)std::max<int64_t>(total_span, 1));
    int64_t* h_pre1 = (int64_t*)(hp + table_bytes);
    int64_t* h_pre2 = h_pre1 + n + 1;
    int64_t* h_pre3 = h_pre2 + n + 1;
    int32_t* h_hist = (int32_t*)(h_pre3 + n + 1);
    int32_t* h_lut = h_hist + total_span;
    int64_t* h_pc = (int64_t*)(h_lut + total_span);
    for (int i = 0; i < n; ++i) {
        if (!has_data(i)) continue;
        ij[(size_t)i].lo = hj[(size_t)i].lo;
        ij[(size_t)i].span = hj[(size_t)i].span;
        ij[(size_t)i].hist = d_hist + hj[(size_t)i].hist_off;
    }
    if (total_span) {
        std::memcpy(hp, ij.data(), table_bytes);
        const int64_t blocks = prefix(1, h_pre1);
        check(etm::detail::launch_fill_i32(d_hist, total_span, 0, s), "zero histograms");
        check(hipMemcpyAsync(d_in, hp, table_bytes, hipMemcpyHostToDevice, s), "upload input records");
        check(hipMemcpyAsync(d_pre + (n + 1), h_pre1, pre_bytes, hipMemcpyHostToDevice, s), "upload input records");
        check(etm::detail::launch_batch_input(1, d_in, n, d_pre + (n + 1), blocks, 0, s), "batch histograms");
        check(hipMemcpyAsync(h_hist, d_hist, (size_t)total_span * 4, hipMemcpyDeviceToHost, s), "readback");
        wait_stream(s);
    }
    int64_t total_cols = 0;
    for (int i = 0; i < n; ++i) {
        if (!has_data(i)) continue;
        BatchHost& b = hj[(size_t)i];
        b.mc = std::max<int64_t>(1, in[i].min_count);
        etm_levels& out = *res[(size_t)i];
        for (int64_t v = 0; v < b.span; ++v) {
            h_lut[b.hist_off + v] = -1;
            if (h_hist[b.hist_off + v] >= b.mc) {
                h_lut[b.hist_off + v] = (int32_t)out.col_to_item.size();
                out.col_to_item.push_back(b.lo + v);
            }
        }
        b.n_cols = (int32_t)out.col_to_item.size();
        b.pc_off = total_cols;
        total_cols += b.n_cols;
    }
    int64_t* d_pc = a.get<int64_t>(total_cols);
    int64_t max_slab = 1;
    for (int i = 0; i < n; ++i) {
        BatchHost& b = hj[(size_t)i];
        bt::InJob& J = ij[(size_t)i];
        if (!has_data(i) || b.n_cols == 0) {
            J.n_cols = 0;
            continue;
        }
        b.n_words = (in[i].n_rows + 63) / 64;
        b.pitch = std::max<int64_t>(64, (b.n_words + 63) / 64 * 64);
        b.bv = a.get<uint64_t>((int64_t)b.n_cols * b.pitch);
        J.n_cols = b.n_cols;
        J.n_words = b.n_words;
        J.pitch = b.pitch;
        J.lut = d_lut + b.hist_off;
        J.bv = b.bv;
        J.pc = d_pc + b.pc_off;
        etm::detail::batch_input_shape(J);
        max_slab = std::max(max_slab, J.slab_cols);
    }
    if (total_cols) {
        std::memcpy(hp, ij.data(), table_bytes);
        const int64_t blocks2 = prefix(2, h_pre2);
        const int64_t blocks3 = prefix(3, h_pre3);
        check(hipMemcpyAsync(d_lut, h_lut, (size_t)total_span * 4, hipMemcpyHostToDevice, s), "upload lut");
        check(hipMemcpyAsync(d_in, hp, table_bytes, hipMemcpyHostToDevice, s), "upload input records");
        check(hipMemcpyAsync(d_pre + 2 * (n + 1), h_pre2, 2 * pre_bytes, hipMemcpyHostToDevice, s),
              "upload input records");
        check(etm::detail::launch_batch_input(2, d_in, n, d_pre + 2 * (n + 1), blocks2, max_slab, s),
              "batch bitvectors");
        check(etm::detail::launch_batch_input(3, d_in, n, d_pre + 3 * (n + 1), blocks3, 0, s), "batch K=1 counts");
        check(hipMemcpyAsync(h_pc, d_pc, (size_t)total_cols * 8, hipMemcpyDeviceToHost, s), "readback");
        wait_stream(s);
    }
    // K=1, and which jobs the rounds take.
    std::vector<int> batched, single;
    for (int i = 0; i < n; ++i) {
        BatchHost& b = hj[(size_t)i];
        if (!has_data(i) || b.n_cols == 0) continue;
        etm_levels& out = *res[(size_t)i];
        out.col_counts.assign(h_pc + b.pc_off, h_pc + b.pc_off + b.n_cols);
        for (int32_t c = 0; c < b.n_cols; ++c)
            if (out.col_counts[(size_t)c] >= b.mc) b.k1.push_back(c);
        b.max_len = in[i].max_length > 0 ? std::min<int64_t>(in[i].max_length, b.n_cols) : b.n_cols;
        const int64_t n1 = (int64_t)b.k1.size();
        if (n1 >= 2 && n1 <= bt::kMaxCols && b.max_len >= 2 && in[i].n_rows <= 0x7FFFFFFF) batched.push_back(i);
        else single.push_back(i);
    }
    std::vector<int> overflow;
    for (size_t b0 = 0; b0 < batched.size(); b0 += bt::kMaxJobs) {
        const size_t b1 = std::min(batched.size(), b0 + bt::kMaxJobs);
        std::vector<int> ids(batched.begin() + (std::ptrdiff_t)b0, batched.begin() + (std::ptrdiff_t)b1);
        for (int i : ids) batch_level_one(*res[(size_t)i], hj[(size_t)i], t_call);
        batch_rounds(a, d_jobs, d_q, ids, hj, res, overflow, s);
    }
    // The per-job loop for the rest, on the bitvectors built above.
    for (int i : overflow) {
        etm_levels& out = *res[(size_t)i];
        out.k.clear();
        out.n.clear();
        out.rows.clear();
        out.counts.clear();
        out.stats.clear();
        out.owned.clear();
        single.push_back(i);
    }
    for (int i : single) {
        const BatchHost& b = hj[(size_t)i];
        etm_mine_opts oi = o;
        oi.min_count = b.mc;
        oi.max_length = in[i].max_length;
        const RowsSrc src{in[i].d_indptr, in[i].d_ids, in[i].id_bytes, in[i].nnz, b.lo, b.span, d_lut + b.hist_off};
        etm_levels& out = *res[(size_t)i];
        mine_levels(b.bv, b.pitch, b.n_words, in[i].n_rows, out.col_counts.data(), b.n_cols, oi, out, s, t_call, &src,
                    false);
    }
    // Nothing of this call stays queued behind the return (the LUT upload reads the thread's pinned buffer).
    wait_stream(s);
}

template <class F>
int guarded(F&& f) {
    try {
        f();
        return ETM_OK;
    } catch (const Error& e) {
        etm::set_error("%s", e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        etm::set_error("host allocation failed");
        return ETM_ENOMEM;
    } catch (const std::exception& e) {
        etm::set_error("%s", e.what());
        return ETM_EHIP;
    }
}

}  // namespace

extern "C" int etm_mine_bitvecs(const uint64_t* d_bv, int64_t pitch, int64_t n_words, int64_t n_rows,
                                const int64_t* h_col_counts, int32_t n_cols, const etm_mine_opts* opts,
                                etm_levels** out, void* stream) {
    ETM_ARG(d_bv && opts && out && (n_cols == 0 || h_col_counts), "etm_mine_bitvecs: null pointer");
    ETM_ARG(pitch >= n_words && n_words == (n_rows + 63) / 64, "etm_mine_bitvecs: bad pitch / n_words");
    ETM_ARG(!opts->emit_levels || opts->level_cb, "etm_mine_bitvecs: emit_levels needs a level callback");
    const Clock::time_point t_call = Clock::now();
    *out = nullptr;
    auto res = std::make_unique<etm_levels>();
    res->col_counts.assign(h_col_counts, h_col_counts + n_cols);
    const int rc = guarded([&] {
        mine_levels(d_bv, pitch, n_words, n_rows, h_col_counts, n_cols, *opts, *res, ETM_STREAM(stream), t_call);
    });
    if (rc == ETM_OK) *out = res.release();
    return rc;
}

extern "C" int etm_mine_csr(const int64_t* d_indptr, const void* d_ids, int id_bytes, int64_t n_rows, int64_t nnz,
                            int64_t span_limit, const etm_mine_opts* opts, etm_levels** out, void* stream) {
    ETM_ARG(d_indptr && opts && out && (nnz == 0 || d_ids), "etm_mine_csr: null pointer");
    ETM_ARG(id_bytes == 4 || id_bytes == 8, "etm_mine_csr: id_bytes must be 4 or 8");
    ETM_ARG(n_rows >= 0 && nnz >= 0, "etm_mine_csr: negative size");
    ETM_ARG(!opts->emit_levels || opts->level_cb, "etm_mine_csr: emit_levels needs a level callback");
    const Clock::time_point t_call = Clock::now();
    *out = nullptr;
    auto res = std::make_unique<etm_levels>();
    hipStream_t s = ETM_STREAM(stream);
    const int rc = guarded([&] {
        if (nnz == 0 || n_rows == 0) return;
        const int64_t mc = std::max<int64_t>(1, opts->min_count);
        Arena a(s);
        Pinned host;
        int64_t* mm = a.get<int64_t>(2);
        check_rc(etm_ids_minmax(d_ids, id_bytes, nnz, mm, s), "etm_ids_minmax");
        check(hipMemcpyAsync(host.p, mm, 16, hipMemcpyDeviceToHost, s), "readback");
        wait_stream(s);
        const int64_t lo = host.p[0], span = host.p[1] - host.p[0] + 1;
        if (span > span_limit)
            throw Error(ETM_EINVAL, "item ids span " + std::to_string(span) + " values: code them to dense ids first");
        int32_t* hist = a.get<int32_t>(span);
        check_rc(etm_ids_hist(d_ids, id_bytes, nnz, lo, span, hist, s), "etm_ids_hist");
        // Histogram down and LUT up through the thread's pinned staging buffer (its two halves).
        int32_t* h_hist = staging(2 * span);
        int32_t* lut = h_hist + span;
        check(hipMemcpyAsync(h_hist, hist, span * 4, hipMemcpyDeviceToHost, s), "readback");
        wait_stream(s);
        for (int64_t v = 0; v < span; ++v) {
            lut[v] = -1;
            if (h_hist[v] >= mc) {
                lut[v] = (int32_t)res->col_to_item.size();
                res->col_to_item.push_back(lo + v);
            }
        }
        const int32_t n_cols = (int32_t)res->col_to_item.size();
        if (n_cols == 0) return;
        int32_t* d_lut = a.get<int32_t>(span);
        check(hipMemcpyAsync(d_lut, lut, span * 4, hipMemcpyHostToDevice, s), "upload lut");
        const int64_t n_words = (n_rows + 63) / 64;
        const int64_t pitch = std::max<int64_t>(64, (n_words + 63) / 64 * 64);
        // Row order (ETM_ROW_ORDER_LEX): rows sorted by the 64 most frequent columns they hold, most frequent
        // first, so that rows sharing frequent columns gather into few 64-word tiles and the group kernels skip
        // the tiles where a prefix is empty. The automatic choice keeps the rows: on deep_sparse_20m the skip
        // saves 34.2 -> 30.4 ms of K>=3 counting but the sort and the permuted build cost as much (62.5 ms
        // ordered, 63.2 ms kept), and on deep_k_1m the tiles stay occupied (3.16 -> 3.58 ms ordered).
        const bool ordered = opts->row_order == ETM_ROW_ORDER_LEX;
        int32_t* order = nullptr;
        if (ordered) {
            std::vector<int32_t> by_count((size_t)n_cols);
            for (int32_t c = 0; c < n_cols; ++c) by_count[(size_t)c] = c;
            std::stable_sort(by_count.begin(), by_count.end(), [&](int32_t x, int32_t y) {
                return h_hist[res->col_to_item[(size_t)x] - lo] > h_hist[res->col_to_item[(size_t)y] - lo];
            });
            const int32_t n_ranked = std::min<int32_t>(64, n_cols);
            int32_t* rank = lut;  // the LUT is on its way up: reuse its half of the staging buffer after the copy
            wait_stream(s);
            for (int32_t c = 0; c < n_cols; ++c) rank[c] = -1;
            for (int32_t t = 0; t < n_ranked; ++t) rank[by_count[(size_t)t]] = t;
            int32_t* d_rank = a.get<int32_t>(n_cols);
            check(hipMemcpyAsync(d_rank, rank, (size_t)n_cols * 4, hipMemcpyHostToDevice, s), "upload ranks");
            order = a.get<int32_t>(n_rows);
            const size_t ws_n = etm_row_order_ws(n_rows);
            void* ws = a.get<int64_t>((int64_t)(ws_n + 7) / 8);
            check_rc(etm_row_order(d_indptr, d_ids, id_bytes, n_rows, lo, span, d_lut, d_rank, n_ranked, order, ws,
                                   ws_n, s),
                     "etm_row_order");
        }
        uint64_t* bv = a.get<uint64_t>((int64_t)n_cols * pitch);
        // The K=1 counts come with the build (per-tile partials), not from a second pass over the bitvectors.
        int64_t* pc = a.get<int64_t>(n_cols);
        uint16_t* part = a.get<uint16_t>(etm::detail::bitvec_part_len(pitch, n_cols));
        check_rc(etm::detail::bitvec_build(d_indptr, d_ids, id_bytes, n_rows, lo, span, d_lut, n_cols, pitch, order,
                                           bv, part, pc, s),
                 "etm_bitvec_build");
        res->col_counts.resize(n_cols);
        check(hipMemcpyAsync(res->col_counts.data(), pc, n_cols * 8, hipMemcpyDeviceToHost, s), "readback");
        wait_stream(s);
        const RowsSrc src{d_indptr, d_ids, id_bytes, nnz, lo, span, d_lut};
        mine_levels(bv, pitch, n_words, n_rows, res->col_counts.data(), n_cols, *opts, *res, s, t_call, &src,
                    ordered);
    });
    if (rc == ETM_OK) *out = res.release();
    return rc;
}

extern "C" int etm_levels_count(const etm_levels* lv, int32_t* h_n_levels) {
    ETM_ARG(lv && h_n_levels, "etm_levels_count: null pointer");
    *h_n_levels = (int32_t)lv->k.size();
    return ETM_OK;
}

extern "C" int etm_levels_shape(const etm_levels* lv, int32_t level, int32_t* h_k, int64_t* h_n) {
    ETM_ARG(lv && h_k && h_n, "etm_levels_shape: null pointer");
    ETM_ARG(level >= 0 && level < (int32_t)lv->k.size(), "etm_levels_shape: level %d out of range", level);
    *h_k = lv->k[level];
    *h_n = lv->n[level];
    return ETM_OK;
}

extern "C" int etm_levels_copy(const etm_levels* lv, int32_t level, int32_t* h_rows, int64_t* h_counts) {
    ETM_ARG(lv, "etm_levels_copy: null result");
    ETM_ARG(level >= 0 && level < (int32_t)lv->k.size(), "etm_levels_copy: level %d out of range", level);
    ETM_ARG(level == 0 || !lv->rows_released, "etm_levels_copy: the levels were released by etm_levels_frame_view");
    const int64_t n = lv->n[level];
    if (h_rows && n) std::memcpy(h_rows, lv->rows[level], (size_t)(n * lv->k[level]) * 4);
    if (h_counts)
        for (int64_t i = 0; i < n; ++i) h_counts[i] = lv->counts[level][i];
    return ETM_OK;
}

extern "C" int etm_levels_columns(const etm_levels* lv, int32_t* h_n_cols, int64_t* h_col_to_item,
                                  int64_t* h_col_counts) {
    ETM_ARG(lv && h_n_cols, "etm_levels_columns: null pointer");
    *h_n_cols = (int32_t)lv->col_counts.size();
    if (h_col_to_item && !lv->col_to_item.empty())
        std::memcpy(h_col_to_item, lv->col_to_item.data(), lv->col_to_item.size() * 8);
    if (h_col_counts) std::memcpy(h_col_counts, lv->col_counts.data(), lv->col_counts.size() * 8);
    return ETM_OK;
}

extern "C" int etm_levels_stats(const etm_levels* lv, int32_t level, etm_level_stat* h_out) {
    ETM_ARG(lv && h_out, "etm_levels_stats: null pointer");
    ETM_ARG(level >= 0 && level < (int32_t)lv->stats.size(), "etm_levels_stats: level %d out of range", level);
    *h_out = lv->stats[level];
    return ETM_OK;
}

namespace {

// The flat export behind etm_levels_flat and etm_levels_frame: sizes into h_totals, then (when h_values and
// h_offsets are given) the item ids, offsets and, where given, counts and supports (count / n_rows in double).
int flat_export(const etm_levels* lv, int64_t* h_totals, int64_t* h_values, int64_t* h_offsets, int64_t* h_counts,
                double* h_support, int64_t n_rows) {
    int64_t n_sets = 0, n_vals = 0;
    for (size_t l = 0; l < lv->k.size(); ++l) {
        n_sets += lv->n[l];
        n_vals += lv->n[l] * lv->k[l];
    }
    h_totals[0] = n_sets;
    h_totals[1] = n_vals;
    if (!h_values || !h_offsets) return ETM_OK;
    const double rows_d = (double)n_rows;
    const int64_t* c2i = lv->col_to_item.empty() ? nullptr : lv->col_to_item.data();
    // Each level's rows [lo, hi) land at fixed places (itemsets before the level, items before them), so ranges of
    // rows are written by separate threads when the result is large.
    struct Range {
        size_t l;
        int64_t lo, hi, row0, at0;
    };
    std::vector<Range> ranges;
    const int64_t kRowsPerTask = 1 << 18;
    int64_t row = 0, at = 0;
    for (size_t l = 0; l < lv->k.size(); ++l) {
        const int64_t k = lv->k[l], n = lv->n[l];
        for (int64_t lo = 0; lo < n; lo += kRowsPerTask) {
            const int64_t hi = std::min(n, lo + kRowsPerTask);
            ranges.push_back(Range{l, lo, hi, row + lo, at + lo * k});
        }
        row += n;
        at += n * k;
    }
    h_offsets[0] = 0;
    auto fill = [&](const Range& g) {
        const int64_t k = lv->k[g.l];
        const int32_t* r = lv->rows[g.l];
        const int32_t* c = lv->counts[g.l];
        int64_t* v = h_values + g.at0;
        for (int64_t i = g.lo * k; i < g.hi * k; ++i) *v++ = c2i ? c2i[r[i]] : r[i];
        for (int64_t i = g.lo; i < g.hi; ++i) h_offsets[g.row0 + (i - g.lo) + 1] = g.at0 + (i - g.lo + 1) * k;
        if (h_counts)
            for (int64_t i = g.lo; i < g.hi; ++i) h_counts[g.row0 + (i - g.lo)] = c[i];
        if (h_support)  // IEEE division, as numpy's counts / n_rows
            for (int64_t i = g.lo; i < g.hi; ++i) h_support[g.row0 + (i - g.lo)] = (double)c[i] / rows_d;
    };
    // Threads pay off from a few million item ids (starting one costs tens of microseconds; one thread writes
    // about a billion ids a second).
    const int64_t kThreadedMinValues = 1 << 22;
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (ranges.size() <= 1 || hw == 1 || n_vals < kThreadedMinValues) {
        for (const Range& g : ranges) fill(g);
        return ETM_OK;
    }
    std::atomic<size_t> next{0};
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < std::min<size_t>(hw, ranges.size()); ++t)
        pool.emplace_back([&] {
            for (size_t i; (i = next.fetch_add(1)) < ranges.size();) fill(ranges[i]);
        });
    for (auto& th : pool) th.join();
    return ETM_OK;
}

}  // namespace

extern "C" int etm_levels_flat(const etm_levels* lv, int64_t* h_totals, int64_t* h_values, int64_t* h_offsets,
                               int64_t* h_counts) {
    ETM_ARG(lv && h_totals, "etm_levels_flat: null pointer");
    ETM_ARG(!lv->rows_released, "etm_levels_flat: the levels were released by etm_levels_frame_view");
    if (!h_values || !h_offsets || !h_counts) h_values = h_offsets = h_counts = nullptr;
    return flat_export(lv, h_totals, h_values, h_offsets, h_counts, nullptr, 1);
}

extern "C" int etm_levels_frame(const etm_levels* lv, int64_t* h_totals, int64_t* h_values, int64_t* h_offsets,
                                int64_t* h_counts, double* h_support, int64_t n_rows) {
    ETM_ARG(lv && h_totals, "etm_levels_frame: null pointer");
    ETM_ARG(!lv->rows_released, "etm_levels_frame: the levels were released by etm_levels_frame_view");
    ETM_ARG((h_values == nullptr) == (h_offsets == nullptr), "etm_levels_frame: values and offsets go together");
    ETM_ARG(h_support == nullptr || n_rows > 0, "etm_levels_frame: supports need n_rows > 0");
    return flat_export(lv, h_totals, h_values, h_offsets, h_counts, h_support, n_rows);
}

extern "C" int etm_levels_frame_view(etm_levels* lv, int64_t* h_totals, void** h_ptrs, int32_t release_rows) {
    ETM_ARG(lv && h_totals && h_ptrs, "etm_levels_frame_view: null pointer");
    if (lv->frame == nullptr) {
        h_totals[0] = h_totals[1] = -1;
        return ETM_OK;
    }
    const etm_frame& f = *lv->frame;
    h_totals[0] = f.n_sets;
    h_totals[1] = f.n_vals;
    h_ptrs[0] = f.values.data();
    h_ptrs[1] = f.offsets.data();
    h_ptrs[2] = f.counts.data();
    h_ptrs[3] = f.support.data();
    if (release_rows && !lv->rows_released) {
        pool_give(lv->block, lv->block_cap);
        lv->block = nullptr;
        lv->block_cap = 0;
        for (auto& b : lv->level_blocks) pool_give(b.first, b.second);
        lv->level_blocks.clear();
        lv->rows_released = true;
    }
    return ETM_OK;
}

etm_levels::~etm_levels() {
    pool_give(block, block_cap);
    for (auto& b : level_blocks) pool_give(b.first, b.second);
    delete frame;
}

extern "C" int etm_levels_free(etm_levels* lv) {
    delete lv;
    return ETM_OK;
}

extern "C" int etm_mine_batch(const etm_batch_job* jobs, int32_t n_jobs, int64_t span_limit, const etm_mine_opts* opts,
                              etm_levels** outs, void* stream) {
    ETM_ARG(opts && outs && (n_jobs == 0 || jobs), "etm_mine_batch: null pointer");
    ETM_ARG(n_jobs >= 0, "etm_mine_batch: negative job count");
    ETM_ARG(!opts->prune_non_free && !opts->infer_counts && !opts->emit_levels && opts->level_cb == nullptr,
            "etm_mine_batch: the free-set and inference modes and the level callback are mined per job "
            "(etm_mine_csr)");
    for (int32_t i = 0; i < n_jobs; ++i) {
        const etm_batch_job& j = jobs[i];
        ETM_ARG(j.d_indptr && (j.nnz == 0 || j.d_ids), "etm_mine_batch: job %d: null pointer", i);
        ETM_ARG(j.id_bytes == 4 || j.id_bytes == 8, "etm_mine_batch: job %d: id_bytes must be 4 or 8", i);
        ETM_ARG(j.n_rows >= 0 && j.nnz >= 0 && j.max_length >= 0, "etm_mine_batch: job %d: negative size", i);
        outs[i] = nullptr;
    }
    const Clock::time_point t_call = Clock::now();
    std::vector<std::unique_ptr<etm_levels>> res;
    res.reserve((size_t)n_jobs);
    for (int32_t i = 0; i < n_jobs; ++i) res.push_back(std::make_unique<etm_levels>());
    const int rc = guarded([&] {
        if (n_jobs > 0) mine_batch(jobs, n_jobs, span_limit, *opts, res, ETM_STREAM(stream), t_call);
    });
    if (rc == ETM_OK)
        for (int32_t i = 0; i < n_jobs; ++i) outs[i] = res[(size_t)i].release();
    return rc;
}
