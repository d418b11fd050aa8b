// dress_capi.hip -- the C ABI of libdress_hip.so: libdress's core API
// (libdress/include/dress/dress.h:80-107) and the dress_fit_hip entry point
// (counterpart of dress_fit_cuda, libdress/src/cuda/dress_cuda.cu:445-707).
//
// Device state of a graph lives in a side table keyed by the dress_graph_t
// pointer, so the struct layout the bindings read stays untouched.
#include "dress_hip_internal.h"
#include "dress_capi_internal.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <atomic>
#include <thread>
#include <chrono>
#include <sys/mman.h>
#include <mutex>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "dress/hip/dress_hip.h"

namespace dh {

// ---------------------------------------------------------------- errors
static thread_local std::string t_err;
static thread_local dress_hip_fit_stats_t t_stats;

void set_error(const std::string &msg)
{
    if (t_err.empty()) t_err = msg;      // keep the first (root) cause
    if (getenv("DRESS_HIP_QUIET") == nullptr) fprintf(stderr, "libdress_hip: %s\n", msg.c_str());
}
void clear_error() { t_err.clear(); }
bool has_error() { return !t_err.empty(); }

// ---------------------------------------------------------------- allocation
namespace {
struct AllocCtx {
    bool active = false;
    bool serving = false;              // a server worker's context (ctx_begin)
    hipStream_t stream = nullptr;
    std::unordered_map<void *, size_t> size_of;          // every block this context handed out
    std::unordered_map<size_t, std::vector<void *>> free_blocks;
    int64_t bytes = 0;
};
thread_local AllocCtx t_ctx;

// size class: at most 25 % above the request, 256 B granules
size_t size_class(size_t bytes)
{
    size_t b = bytes < 256 ? 256 : bytes;
    size_t top = (size_t)1 << (63 - __builtin_clzll((unsigned long long)b));
    size_t step = std::max<size_t>(256, top / 4);
    return (b + step - 1) / step * step;
}
}  // namespace

bool dmalloc(void **p, size_t bytes, const char *what)
{
    *p = nullptr;
    size_t want = bytes ? bytes : 16;
    if (t_ctx.active) {
        want = size_class(want);
        auto it = t_ctx.free_blocks.find(want);
        if (it != t_ctx.free_blocks.end() && !it->second.empty()) {
            *p = it->second.back();
            it->second.pop_back();
            return true;
        }
    }
    hipError_t e = hipMalloc(p, want);
    if (e != hipSuccess) {
        *p = nullptr;
        hipGetLastError();
        set_error(std::string("device allocation of ") + std::to_string(bytes) + " bytes for " + what +
                  " failed: " + hipGetErrorString(e));
        return false;
    }
    if (t_ctx.active) {
        t_ctx.size_of[*p] = want;
        t_ctx.bytes += (int64_t)want;
    }
    return true;
}

void dfree(void *p)
{
    if (!p) return;
    if (t_ctx.active) {
        auto it = t_ctx.size_of.find(p);
        if (it != t_ctx.size_of.end()) {
            t_ctx.free_blocks[it->second].push_back(p);
            return;
        }
    }
    hipFree(p);
}

// Host result arrays: large ones are 2 MiB aligned and advised onto
// transparent huge pages, so the copy that fills them takes a few hundred
// page faults instead of one per 4 KiB; free() releases either kind.
void *host_alloc(size_t bytes)
{
    const size_t huge = (size_t)2 << 20;
    if (bytes < 4 * huge) return malloc(bytes ? bytes : 16);
    size_t rounded = (bytes + huge - 1) / huge * huge;
    void *p = aligned_alloc(huge, rounded);
    if (p) madvise(p, rounded, MADV_HUGEPAGE);
    return p;
}

// Pinned host blocks for the batch entry points' result arrays (they have
// their own free functions).  A device-to-host copy into pinned memory runs
// at the link's rate; into pageable memory it is staged and several times
// slower.  Freed blocks are kept by size class for the next call, up to
// 8 GiB; small requests, and threads inside an allocation context, stay
// with pageable memory.
namespace {
struct PinnedPool {
    std::mutex mu;
    std::unordered_map<void *, size_t> size_of;                 // pinned blocks handed out
    std::unordered_map<size_t, std::vector<void *>> free_blocks;
    std::unordered_set<void *> touched;                          // every page written once by the host
    size_t cached = 0;
};
PinnedPool &pinned_pool()
{
    static PinnedPool *p = new PinnedPool();   // lives to process end
    return *p;
}
constexpr size_t PINNED_MIN = (size_t)1 << 20;
constexpr size_t PINNED_CACHE = (size_t)8 << 30;
}  // namespace

int copy_threads();

// The host's first write to each page of a new pinned block faults (84 MB:
// 31 ms against 2.3 ms for the same write later).  A caller that fills a
// block from host threads asks for it touched: each page is written once,
// by several threads for large blocks, the first time the block is handed
// out that way, so that the faults do not land inside its later calls.
void touch_pages(void *q, size_t bytes)
{
    constexpr size_t PAGE = 4096;
    const size_t pages = (bytes + PAGE - 1) / PAGE;
    const int T = bytes >= ((size_t)32 << 20) ? copy_threads() : 1;
    auto touch = [q, pages, T](int w) {
        for (size_t i = pages * (size_t)w / (size_t)T; i < pages * (size_t)(w + 1) / (size_t)T; i++)
            static_cast<volatile char *>(q)[i * PAGE] = 0;
    };
    std::vector<std::thread> team;
    for (int w = 1; w < T; w++) {
        try {
            team.emplace_back(touch, w);
        } catch (...) {
            touch(w);
        }
    }
    touch(0);
    for (auto &t : team) t.join();
}

void *pinned_alloc(size_t bytes, bool host_writes)
{
    // a thread with an allocation context (a server worker) stays with
    // pageable memory: pinning new blocks waits for the whole device
    if (bytes < PINNED_MIN || ctx_serving()) return host_alloc(bytes);
    const size_t want = size_class(bytes);
    PinnedPool &P = pinned_pool();
    void *q = nullptr;
    bool touched = false;
    {
        std::lock_guard<std::mutex> lk(P.mu);
        auto it = P.free_blocks.find(want);
        if (it != P.free_blocks.end() && !it->second.empty()) {
            q = it->second.back();
            it->second.pop_back();
            P.cached -= want;
            touched = P.touched.count(q) != 0;
        }
    }
    if (!q) {
        if (hipHostMalloc(&q, want, hipHostMallocDefault) != hipSuccess || !q) {
            hipGetLastError();
            return host_alloc(bytes);
        }
        std::lock_guard<std::mutex> lk(P.mu);
        P.size_of[q] = want;
    }
    if (host_writes && !touched) {
        touch_pages(q, want);
        std::lock_guard<std::mutex> lk(P.mu);
        P.touched.insert(q);
    }
    return q;
}

void pinned_free(void *p)
{
    if (!p) return;
    PinnedPool &P = pinned_pool();
    std::unique_lock<std::mutex> lk(P.mu);
    auto it = P.size_of.find(p);
    if (it == P.size_of.end()) {
        lk.unlock();
        free(p);
        return;
    }
    const size_t sz = it->second;
    if (P.cached + sz <= PINNED_CACHE) {
        P.free_blocks[sz].push_back(p);
        P.cached += sz;
        return;
    }
    P.size_of.erase(it);
    P.touched.erase(p);
    lk.unlock();
    hipHostFree(p);
}

// threads of the staged copies: 8 (first-touch faults and copies into the
// caller's fresh pageable memory are the limit: 790 MB took 20.5 ms with 8
// threads and 16.7 ms with 16 once warm, against 14 ms for the link, but 16
// threads' staging costs more to pin on a first call);
// DRESS_HIP_COPY_THREADS overrides
int copy_threads()
{
    static const int t = [] {
        const char *e = getenv("DRESS_HIP_COPY_THREADS");
        int v = e ? atoi(e) : 8;
        return v < 1 ? 1 : (v > 32 ? 32 : v);
    }();
    return t;
}

// Small pinned blocks for the state snapshots of running fits and passes,
// taken per call and returned: the pool holds at most as many as calls ran
// at once, kept for the process.
namespace {
struct SnapPool {
    std::mutex mu;
    std::vector<void *> free_blocks;
};
SnapPool &snap_pool()
{
    static SnapPool *p = new SnapPool();      // lives to process end
    return *p;
}
}  // namespace

void *snapshot_take()
{
    SnapPool &P = snap_pool();
    {
        std::lock_guard<std::mutex> lk(P.mu);
        if (!P.free_blocks.empty()) {
            void *q = P.free_blocks.back();
            P.free_blocks.pop_back();
            return q;
        }
    }
    void *q = nullptr;
    if (hipHostMalloc(&q, SNAPSHOT_BYTES, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        set_error("pinned allocation of a state snapshot block failed");
        return nullptr;
    }
    return q;
}

void snapshot_give(void *q)
{
    if (!q) return;
    SnapPool &P = snap_pool();
    std::lock_guard<std::mutex> lk(P.mu);
    P.free_blocks.push_back(q);
}

// Streams and events of the staged copies' host threads, kept for the
// process: creating a stream took 8 ms here, and every large copy needs one
// per thread.  Lane w belongs to copy thread w and is created by it on first
// use.  One copy at a time holds the lanes (for one device); a copy that
// finds them held creates and destroys its own.
namespace {
struct CopyLane {
    hipStream_t st = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
};
struct CopyLanes {
    std::mutex mu;
    int dev = -1;
    CopyLane lane[32];
};
CopyLanes &copy_lanes()
{
    static CopyLanes *p = new CopyLanes();   // lives to process end
    return *p;
}
bool make_lane(CopyLane &c)
{
    if (hipStreamCreateWithFlags(&c.st, hipStreamNonBlocking) == hipSuccess &&
        hipEventCreateWithFlags(&c.ev[0], hipEventDisableTiming) == hipSuccess &&
        hipEventCreateWithFlags(&c.ev[1], hipEventDisableTiming) == hipSuccess)
        return true;
    (void)hipGetLastError();
    if (c.st) (void)hipStreamDestroy(c.st);
    for (hipEvent_t &x : c.ev)
        if (x) (void)hipEventDestroy(x);
    c = CopyLane();
    return false;
}
void drop_lane(CopyLane &c)
{
    if (c.st) (void)hipStreamDestroy(c.st);
    for (hipEvent_t &x : c.ev)
        if (x) (void)hipEventDestroy(x);
    c = CopyLane();
}
// the lanes of one staged copy of T threads on device dev: the pooled ones
// when free, else each thread's own
struct LaneSet {
    std::unique_lock<std::mutex> lk;
    bool pooled = false;
    LaneSet(int dev, int T) : lk(copy_lanes().mu, std::try_to_lock)
    {
        CopyLanes &P = copy_lanes();
        pooled = lk.owns_lock() && T <= 32 && (P.dev < 0 || P.dev == dev);
        if (pooled) P.dev = dev;
    }
    // lane of thread w (created on first use); *own when it is the thread's
    // own, to be dropped after the copy
    bool get(int w, CopyLane &out, bool *own)
    {
        *own = false;
        if (pooled) {
            CopyLane &c = copy_lanes().lane[w];
            if (c.st || make_lane(c)) {
                out = c;
                return true;
            }
        }
        *own = true;
        return make_lane(out);
    }
};
}  // namespace

// Device-to-host copy of a large buffer (64 MiB and up) into pageable
// memory (results the caller frees with free()).  T host threads
// (copy_threads(), one per 32 MiB at most) each take a contiguous slice and
// move it in 2 MiB chunks through two pinned staging chunks of their own
// (one pooled block for all) on a stream of their own: the copy of chunk
// i+1 runs while chunk i is copied into place, and the first-touch faults
// of the destination are taken by T threads at once.  (300 MB took 8.4 ms
// with 2 MiB chunks against 10.5 ms with 8 MiB, whose 128 MiB of staging
// also cost 9 ms more to pin on first use.)  src must be complete (its
// stream idle).  Threads inside an allocation context copy directly.
bool d2h_parallel(void *dst, const void *src, size_t bytes)
{
    constexpr size_t CH = (size_t)2 << 20;
    constexpr size_t MIN_BYTES = (size_t)64 << 20, PER_THREAD = (size_t)32 << 20;
    if (bytes == 0) return true;
    if (bytes < MIN_BYTES || ctx_serving())
        return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess;
    const int T = (int)std::max<size_t>(1, std::min<size_t>({(size_t)copy_threads(), bytes / PER_THREAD,
                                                              (size_t)std::max(1u, std::thread::hardware_concurrency())}));
    int dev = 0;
    hipGetDevice(&dev);
    // one pinned block holds every thread's two staging chunks
    char *staging = (char *)pinned_alloc((size_t)T * 2 * CH);
    if (!staging) return false;
    std::atomic<bool> ok{true};
    LaneSet lanes(dev, T);
    auto slice = [&](int w) {
        hipSetDevice(dev);
        const size_t b = bytes * (size_t)w / (size_t)T, e = bytes * (size_t)(w + 1) / (size_t)T;
        char *stage[2] = {staging + (size_t)w * 2 * CH, staging + (size_t)w * 2 * CH + CH};
        CopyLane lane;
        bool own = false;
        bool good = lanes.get(w, lane, &own);
        hipStream_t st = lane.st;
        hipEvent_t *ev = lane.ev;
        auto issue = [&](size_t at, int j) {
            const size_t n = std::min(CH, e - at);
            return hipMemcpyAsync(stage[j], (const char *)src + at, n, hipMemcpyDeviceToHost, st) == hipSuccess &&
                   hipEventRecord(ev[j], st) == hipSuccess;
        };
        if (good && b < e) good = issue(b, 0);
        int j = 0;
        for (size_t at = b; good && at < e; at += CH, j ^= 1) {
            if (at + CH < e) good = issue(at + CH, j ^ 1);
            good = good && hipEventSynchronize(ev[j]) == hipSuccess;
            if (good) memcpy((char *)dst + at, stage[j], std::min(CH, e - at));
        }
        if (st && hipStreamSynchronize(st) != hipSuccess) good = false;
        if (own) drop_lane(lane);
        if (!good) ok = false;
    };
    std::vector<std::thread> pool;
    for (int w = 1; w < T; w++) pool.emplace_back(slice, w);
    slice(0);
    for (auto &th : pool) th.join();
    pinned_free(staging);
    return ok.load();
}

// The flat histogram (value bits, count; 2 * runs int64) into pageable dst
// from the device's values and its counts narrowed to `width` (1, 2 or 4)
// bytes, as d2h_parallel moves a buffer: T host threads each take a slice of
// the runs and move it in chunks of values and counts through two pinned
// staging pairs of their own on a stream of their own, interleaving each
// chunk into place while the next one crosses.  The link then carries
// 8 + width bytes a run instead of 16.  Chunks of 2^20 runs (9 MB at 1 byte
// a count; 18 MB of pooled staging a thread): the 49M runs of a delta^2
// histogram took 9.2 ms so, against 10.1 ms with 2^19 and 15.2 ms with
// 2^17.  vals and counts must be complete.
bool d2h_flat_narrow(int64_t *dst, const double *vals, const void *counts, int width, int64_t runs)
{
    constexpr int64_t CH = (int64_t)1 << 20;                  // runs a chunk
    if (runs <= 0) return true;
    if (width != 1 && width != 2 && width != 4) return false;
    const size_t bytes = (size_t)runs * 16;
    const int T = (int)std::max<size_t>(1, std::min<size_t>({(size_t)copy_threads(), bytes / ((size_t)32 << 20),
                                                              (size_t)std::max(1u, std::thread::hardware_concurrency())}));
    int dev = 0;
    hipGetDevice(&dev);
    const size_t pair = (size_t)CH * (8 + (size_t)width);       // one chunk's values, then its counts
    // (read by the host threads: pages touched)
    char *staging = (char *)pinned_alloc((size_t)T * 2 * pair, true);
    if (!staging) return false;
    std::atomic<bool> ok{true};
    LaneSet lanes(dev, T);
    auto slice = [&](int w) {
        hipSetDevice(dev);
        const int64_t b = runs * w / T, e = runs * (w + 1) / T;
        char *stage[2] = {staging + (size_t)w * 2 * pair, staging + (size_t)w * 2 * pair + pair};
        CopyLane lane;
        bool own = false;
        bool good = lanes.get(w, lane, &own);
        hipStream_t st = lane.st;
        hipEvent_t *ev = lane.ev;
        auto issue = [&](int64_t at, int j) {
            const int64_t n = std::min(CH, e - at);
            return hipMemcpyAsync(stage[j], vals + at, (size_t)n * 8, hipMemcpyDeviceToHost, st) == hipSuccess &&
                   hipMemcpyAsync(stage[j] + (size_t)CH * 8, static_cast<const char *>(counts) + (size_t)at * width,
                                  (size_t)n * width, hipMemcpyDeviceToHost, st) == hipSuccess &&
                   hipEventRecord(ev[j], st) == hipSuccess;
        };
        if (good && b < e) good = issue(b, 0);
        int j = 0;
        for (int64_t at = b; good && at < e; at += CH, j ^= 1) {
            if (at + CH < e) good = issue(at + CH, j ^ 1);
            good = good && hipEventSynchronize(ev[j]) == hipSuccess;
            if (!good) break;
            const int64_t n = std::min(CH, e - at);
            const int64_t *v = reinterpret_cast<const int64_t *>(stage[j]);
            const char *c = stage[j] + (size_t)CH * 8;
            // streaming stores: the destination is not read here, and
            // ordinary stores would first read every line of it
            int64_t *o = dst + 2 * at;
            auto put = [o](int64_t i, int64_t value, int64_t count) {
                __builtin_nontemporal_store(value, o + 2 * i);
                __builtin_nontemporal_store(count, o + 2 * i + 1);
            };
            if (width == 1)
                for (int64_t i = 0; i < n; i++) put(i, v[i], reinterpret_cast<const uint8_t *>(c)[i]);
            else if (width == 2)
                for (int64_t i = 0; i < n; i++) put(i, v[i], reinterpret_cast<const uint16_t *>(c)[i]);
            else
                for (int64_t i = 0; i < n; i++) put(i, v[i], reinterpret_cast<const uint32_t *>(c)[i]);
        }
        __builtin_ia32_sfence();                         // the stores visible before the join
        if (st && hipStreamSynchronize(st) != hipSuccess) good = false;
        if (own) drop_lane(lane);
        if (!good) ok = false;
    };
    std::vector<std::thread> pool;
    for (int w = 1; w < T; w++) pool.emplace_back(slice, w);
    slice(0);
    for (auto &th : pool) th.join();
    pinned_free(staging);
    return ok.load();
}

void fill_nan_parallel(double *p, size_t n) { fill_parallel(p, n, std::nan("")); }

// Every vertex id of an edge list lies in [0, N); otherwise the error is set
// and the call that builds from it returns before any device or host write
// indexed by an id.  Long lists are scanned by several host threads.
bool check_vertex_ids(int N, int64_t E, const int *U, const int *V, const char *who)
{
    if (E <= 0) return true;
    if (!U || !V) {
        set_error(std::string(who) + ": NULL edge list");
        return false;
    }
    const uint32_t n = (uint32_t)std::max(N, 0);
    const int64_t T = E < ((int64_t)1 << 22) ? 1
                                            : std::max<int64_t>(1, std::min<int64_t>((int64_t)copy_threads(),
                                                                   std::max(1u, std::thread::hardware_concurrency())));
    std::vector<int64_t> first_bad((size_t)T, -1);
    auto scan = [&](int64_t w) {
        const int64_t b = E * w / T, e = E * (w + 1) / T;
        unsigned bad = 0;
        for (int64_t i = b; i < e; i++) bad |= ((uint32_t)U[i] >= n) | ((uint32_t)V[i] >= n);
        if (!bad) return;
        for (int64_t i = b; i < e; i++)
            if ((uint32_t)U[i] >= n || (uint32_t)V[i] >= n) { first_bad[(size_t)w] = i; return; }
    };
    std::vector<std::thread> pool;
    for (int64_t w = 1; w < T; w++) {
        try {
            pool.emplace_back(scan, w);
        } catch (...) {
            scan(w);
        }
    }
    scan(0);
    for (auto &th : pool) th.join();
    for (int64_t i : first_bad)
        if (i >= 0) {
            set_error(std::string(who) + ": edge " + std::to_string(i) + " (" + std::to_string(U[i]) + ", " +
                      std::to_string(V[i]) + ") has a vertex id outside [0, " + std::to_string(N) + ")");
            return false;
        }
    return true;
}


void fill_parallel(double *p, size_t n, double value)
{
    const size_t T = n * sizeof(double) < ((size_t)64 << 20)
                         ? 1
                         : std::max<size_t>(1, std::min<size_t>((size_t)copy_threads(),
                                                                std::max(1u, std::thread::hardware_concurrency())));
    auto fill = [&](size_t w) {
        const size_t b = n * w / T, e = n * (w + 1) / T;
        for (size_t i = b; i < e; i++) p[i] = value;
    };
    std::vector<std::thread> pool;
    std::vector<size_t> left;                 // slices no thread could take
    for (size_t w = 1; w < T; w++) {
        try {
            pool.emplace_back(fill, w);
        } catch (...) {
            left.push_back(w);
        }
    }
    fill(0);
    for (size_t w : left) fill(w);
    for (auto &th : pool) th.join();
}

// Host-to-device copy of a large pageable buffer, the mirror of
// d2h_parallel: T host threads each move a contiguous slice through two
// pinned 8 MiB staging blocks on a stream of their own.  dst is complete
// when the call returns.
bool h2d_parallel(void *dst, const void *src, size_t bytes)
{
    constexpr size_t CH = (size_t)8 << 20;
    if (bytes == 0) return true;
    if (bytes < 8 * CH || ctx_serving())
        return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
    const int T = (int)std::max<size_t>(1, std::min<size_t>({(size_t)copy_threads(), bytes / (4 * CH),
                                                              (size_t)std::max(1u, std::thread::hardware_concurrency())}));
    int dev = 0;
    hipGetDevice(&dev);
    // one pinned block holds every thread's two staging chunks
    char *staging = (char *)pinned_alloc((size_t)T * 2 * CH);
    if (!staging) return false;
    std::atomic<bool> ok{true};
    LaneSet lanes(dev, T);
    auto slice = [&](int w) {
        hipSetDevice(dev);
        const size_t b = bytes * (size_t)w / (size_t)T, e = bytes * (size_t)(w + 1) / (size_t)T;
        char *stage[2] = {staging + (size_t)w * 2 * CH, staging + (size_t)w * 2 * CH + CH};
        CopyLane lane;
        bool own = false;
        bool good = lanes.get(w, lane, &own);
        hipStream_t st = lane.st;
        hipEvent_t *ev = lane.ev;
        int j = 0;
        bool used[2] = {false, false};
        for (size_t at = b; good && at < e; at += CH, j ^= 1) {
            const size_t n = std::min(CH, e - at);
            // the staging block's previous copy must be done before refilling it
            if (used[j]) good = hipEventSynchronize(ev[j]) == hipSuccess;
            if (!good) break;
            memcpy(stage[j], (const char *)src + at, n);
            good = hipMemcpyAsync((char *)dst + at, stage[j], n, hipMemcpyHostToDevice, st) == hipSuccess &&
                   hipEventRecord(ev[j], st) == hipSuccess;
            used[j] = true;
        }
        if (st && hipStreamSynchronize(st) != hipSuccess) good = false;
        if (own) drop_lane(lane);
        if (!good) ok = false;
    };
    std::vector<std::thread> pool;
    for (int w = 1; w < T; w++) pool.emplace_back(slice, w);
    slice(0);
    for (auto &th : pool) th.join();
    pinned_free(staging);
    return ok.load();
}

void ctx_begin(hipStream_t stream)
{
    ctx_end();
    t_ctx.active = true;
    t_ctx.serving = true;
    t_ctx.stream = stream;
}

bool ctx_serving() { return t_ctx.active && t_ctx.serving; }

// A thread's block cache for the batch entry points, kept between calls:
// between two calls it holds only free blocks whose last use finished (the
// leaving call synchronises its stream), so the next call may reuse them on
// any stream.  It keeps at most 8 GiB of free blocks (a batch call of
// ~10^7 subgraph values works in ~2 GiB); the thread's exit returns them.
namespace {
struct BatchCache {
    AllocCtx ctx;
    hipStream_t stream = nullptr;
    hipStream_t side[2] = {nullptr, nullptr};      // work that overlaps the main stream's
    ~BatchCache()
    {
        for (auto &kv : ctx.free_blocks)
            for (void *q : kv.second) hipFree(q);
        if (stream) hipStreamDestroy(stream);
        for (hipStream_t x : side)
            if (x) hipStreamDestroy(x);
    }
};
thread_local BatchCache t_batch;
}  // namespace

hipStream_t batch_side_stream(int i)
{
    if (!t_ctx.active || t_ctx.serving || i < 0 || i > 1) return nullptr;
    if (!t_batch.side[i] && hipStreamCreateWithFlags(&t_batch.side[i], hipStreamNonBlocking) != hipSuccess) {
        t_batch.side[i] = nullptr;
        hipGetLastError();
    }
    return t_batch.side[i];
}

hipStream_t batch_ctx_enter()
{
    if (t_ctx.active) return nullptr;       // inside a server context already
    if (!t_batch.stream && hipStreamCreateWithFlags(&t_batch.stream, hipStreamNonBlocking) != hipSuccess) {
        t_batch.stream = nullptr;
        hipGetLastError();
        return nullptr;
    }
    std::swap(t_ctx, t_batch.ctx);
    t_ctx.active = true;
    t_ctx.serving = false;
    t_ctx.stream = t_batch.stream;
    return t_batch.stream;
}

void batch_ctx_leave()
{
    if (!t_ctx.active || t_ctx.serving) return;
    hipStreamSynchronize(t_ctx.stream);
    ctx_trim(0, (size_t)8 << 30);
    t_ctx.active = false;
    std::swap(t_ctx, t_batch.ctx);
}

void ctx_end()
{
    if (!t_ctx.active) return;
    if (t_ctx.stream) hipStreamSynchronize(t_ctx.stream);
    for (auto &kv : t_ctx.free_blocks)
        for (void *q : kv.second) {
            hipFree(q);
            t_ctx.size_of.erase(q);
        }
    // blocks still in use are freed by their owner's later dfree()
    t_ctx.free_blocks.clear();
    t_ctx.size_of.clear();
    t_ctx.bytes = 0;
    t_ctx.active = false;
    t_ctx.stream = nullptr;
}

// returns the context's cached free blocks larger than max_block to the
// device once the cached free bytes exceed keep_bytes, so memory another
// thread was admitted for is not held here (hipFree waits for the device,
// so it is kept off the common path)
void ctx_trim(size_t max_block, size_t keep_bytes)
{
    if (!t_ctx.active) return;
    size_t cached = 0;
    for (auto &kv : t_ctx.free_blocks) cached += kv.first * kv.second.size();
    if (cached <= keep_bytes) return;
    bool synced = false;
    for (auto &kv : t_ctx.free_blocks) {
        if (kv.first <= max_block || kv.second.empty()) continue;
        if (!synced && t_ctx.stream) {
            hipStreamSynchronize(t_ctx.stream);
            synced = true;
        }
        for (void *q : kv.second) {
            hipFree(q);
            t_ctx.size_of.erase(q);
            t_ctx.bytes -= (int64_t)kv.first;
        }
        kv.second.clear();
    }
}

hipStream_t ctx_stream() { return t_ctx.active ? t_ctx.stream : nullptr; }
int64_t ctx_cached_bytes() { return t_ctx.active ? t_ctx.bytes : 0; }

int current_device()
{
    int d = 0;
    hipGetDevice(&d);
    return d;
}

int num_cus()
{
    static int cus[64] = {0};
    int d = current_device();
    if (d < 0 || d >= 64) d = 0;
    if (!cus[d]) {
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, d) == hipSuccess) cus[d] = p.multiProcessorCount;
        if (cus[d] <= 0) cus[d] = 256;
    }
    return cus[d];
}

int num_xcds()
{
    static int xcc[64] = {0};
    int d = current_device();
    if (d < 0 || d >= 64) d = 0;
    if (!xcc[d]) {
        int x = 0;
        if (hipDeviceGetAttribute(&x, hipDeviceAttributeNumberOfXccs, d) != hipSuccess || x < 1) {
            hipGetLastError();
            x = 1;
        }
        xcc[d] = x;
    }
    return xcc[d];
}

// ---------------------------------------------------------------- side table
static std::mutex g_mu;
static std::unordered_map<const dress_graph_t *, DevGraph *> g_table;
static int g_precision = -1;

static thread_local int t_precision = -1;

void set_thread_precision(int precision) { t_precision = precision; }

// share of the device (percent of its resident workgroups) the calling
// thread's multi-subgraph passes may fill
static thread_local int t_grid_share = 100;
void set_thread_grid_share(int percent) { t_grid_share = percent < 1 ? 1 : (percent > 100 ? 100 : percent); }
static thread_local const std::atomic<int> *t_pressure_for_share = nullptr;
// widest multi-subgraph pass the calling thread may run
static thread_local int t_max_slots = 256;
void set_thread_max_slots(int slots) { t_max_slots = slots < 16 ? 16 : slots; }
int thread_max_slots() { return t_max_slots; }
// the share applies while requests of higher priority are waiting or
// running (see set_thread_pressure); otherwise the whole device
int thread_grid_share()
{
    if (t_grid_share < 100 && t_pressure_for_share && t_pressure_for_share->load(std::memory_order_relaxed) == 0)
        return 100;
    return t_grid_share;
}

// requests of higher priority waiting or running (read by the calling
// thread's long loops between passes)
static thread_local const std::atomic<int> *t_pressure = nullptr;
void set_thread_pressure(const std::atomic<int> *p)
{
    t_pressure = p;
    t_pressure_for_share = p;
}
void yield_to_pressure(double max_seconds)
{
    if (!t_pressure) return;
    const auto until = std::chrono::steady_clock::now() + std::chrono::microseconds((int64_t)(max_seconds * 1e6));
    while (t_pressure->load(std::memory_order_relaxed) > 0 && std::chrono::steady_clock::now() < until)
        std::this_thread::sleep_for(std::chrono::microseconds(200));
}

// most subgraph values one chunk of a delta/nabla call holds on the device
// (the values and their histogram assembly): the calling thread's limit when
// set (server workers size it from their reservation), else
// DRESS_HIP_LIFT_CHUNK_VALUES, else what half of the free device memory holds
static thread_local int64_t t_lift_values = 0;
void set_thread_lifted_values(int64_t max_values) { t_lift_values = max_values > 0 ? max_values : 0; }
int64_t lifted_chunk_values()
{
    if (t_lift_values > 0) return t_lift_values;
    if (const char *e = getenv("DRESS_HIP_LIFT_CHUNK_VALUES")) {
        long long v = atoll(e);
        if (v > 0) return (int64_t)v;
    }
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
        (void)hipGetLastError();
        fr = (size_t)8 << 30;
    }
    return std::max<int64_t>((int64_t)1 << 20, (int64_t)(fr / 2 / LIFT_BYTES_PER_VALUE));
}

static thread_local const std::atomic<bool> *t_cancel = nullptr;

void set_thread_cancel(const std::atomic<bool> *flag) { t_cancel = flag; }
bool thread_cancelled() { return t_cancel && t_cancel->load(std::memory_order_relaxed); }

int default_precision()
{
    if (t_precision >= 0) return t_precision;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_precision < 0) {
        const char *e = getenv("DRESS_HIP_PRECISION");
        g_precision = (e && (!strcmp(e, "mixed") || !strcmp(e, "fp32_fp64"))) ? PREC_FP32_FP64 : PREC_FP64;
    }
    return g_precision;
}

DevGraph *lookup(const dress_graph_t *g)
{
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_table.find(g);
    return it == g_table.end() ? nullptr : it->second;
}

static void table_put(const dress_graph_t *g, DevGraph *G)
{
    std::lock_guard<std::mutex> lk(g_mu);
    g_table[g] = G;
}

static DevGraph *table_take(const dress_graph_t *g)
{
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_table.find(g);
    if (it == g_table.end()) return nullptr;
    DevGraph *G = it->second;
    g_table.erase(it);
    return G;
}

// Device graph for g: the one dress_init_graph built, or one built now from
// g's edge list (graphs constructed by another libdress build).
DevGraph *device_graph(p_dress_graph_t g)
{
    DevGraph *G = lookup(g);
    if (G) return G;
    G = new DevGraph();
    if (!build_graph(G, g->N, g->E, g->U, g->V, g->W, g->NW, (int)g->variant)) {
        free_graph(G);
        return nullptr;
    }
    release_reference_csr(G);
    table_put(g, G);
    return G;
}

void record_stats(DevGraph *G, const FitResult &r)
{
    t_stats.iterations_executed = r.executed;
    t_stats.fp32_iterations = r.fp32_iterations;
    t_stats.fp64_reruns = r.reruns;
    t_stats.device_ms = r.ms;
    t_stats.intersections = G->T;
    t_stats.half_edges = G->S;
}

}  // namespace dh

using namespace dh;

namespace {

__global__ void k_bin_terms(const int64_t *cnt, int E, unsigned long long *out)
{
    __shared__ unsigned long long acc[Bins::NB];
    if (threadIdx.x < Bins::NB) acc[threadIdx.x] = 0;
    __syncthreads();
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < E; e += gridDim.x * blockDim.x) {
        int64_t c = cnt[e];
        atomicAdd(&acc[term_bin(c + 1)], (unsigned long long)c);
    }
    __syncthreads();
    if (threadIdx.x < Bins::NB) atomicAdd(&out[threadIdx.x], acc[threadIdx.x]);
}

int fit_impl(p_dress_graph_t g, int max_iterations, double epsilon, int precision, int *iterations,
             double *delta)
{
    clear_error();
    if (!g) { set_error("dress_fit_hip: NULL graph"); return -1; }
    if (max_iterations <= 0) {
        // dress.c:607: no iteration runs; *delta is left untouched
        if (iterations) *iterations = max_iterations;
        return 0;
    }
    DevGraph *G = device_graph(g);
    if (!G) return -1;
    FitResult r;
    if (!refresh_self_terms(G, g->NW) || !fit_upload_values(G, g->edge_dress) ||
        !fit_run(G, max_iterations, epsilon, precision, &r) ||
        !fit_download(G, g->edge_dress, g->vertex_dress))
        return -1;
    G->fit_state_valid = true;
    record_stats(G, r);
    if (iterations) *iterations = r.iterations;
    if (delta) *delta = r.delta;
    return 0;
}

// plain streaming copy: the calibration workload for HBM byte counters and
// the practical bandwidth ceiling.  Each workgroup copies its own contiguous
// 64 KiB (4,096 16-byte words), four non-temporal loads in flight per thread
// before their stores: 5.4-5.7 TB/s for 1 GiB on MI355X, against 4.3-5.3
// for grid-stride forms with 2 to 32 workgroups per CU.
typedef double dvec2 __attribute__((ext_vector_type(2)));
constexpr int64_t COPY_WORDS_PER_BLOCK = 4096;
__global__ void __launch_bounds__(256) k_stream_copy(const double2 *__restrict__ a_, double2 *__restrict__ b_,
                                                     int64_t n)
{
    const dvec2 *__restrict__ a = reinterpret_cast<const dvec2 *>(a_);
    dvec2 *__restrict__ b = reinterpret_cast<dvec2 *>(b_);
    const int64_t s = (int64_t)blockIdx.x * COPY_WORDS_PER_BLOCK;
    const int64_t e = s + COPY_WORDS_PER_BLOCK < n ? s + COPY_WORDS_PER_BLOCK : n;
    for (int64_t i = s + threadIdx.x; i < e; i += 4 * 256) {
        dvec2 x[4];
#pragma unroll
        for (int j = 0; j < 4; j++)
            if (i + j * 256 < e) x[j] = __builtin_nontemporal_load(a + i + j * 256);
#pragma unroll
        for (int j = 0; j < 4; j++)
            if (i + j * 256 < e) __builtin_nontemporal_store(x[j], b + i + j * 256);
    }
}

// random 8-byte gathers: out[i] = src[idx[i]] -- the access pattern of the
// edge kernels' common-neighbour reads, for calibrating the byte counters
// on scattered requests
__global__ void __launch_bounds__(256) k_gather_probe(const double *__restrict__ src, const uint32_t *__restrict__ idx,
                                                      double *__restrict__ out, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = src[__builtin_nontemporal_load(idx + i)];
}

__global__ void k_fill_random_idx(uint32_t *idx, int64_t n, uint32_t range, uint64_t seed)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        uint64_t z = seed + (uint64_t)i * 0x9E3779B97F4A7C15ull;       // splitmix64
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        idx[i] = (uint32_t)(z % range);
    }
}

// LDS read bandwidth: every thread reads 16-byte words of a 32 KiB buffer,
// consecutive lanes at consecutive words (no bank conflicts), rotating the
// start each round so no read can be hoisted
constexpr int LDS_PROBE_TPB = 256;
constexpr int LDS_PROBE_WORDS = 2048;        // 32 KiB of double2
__global__ void __launch_bounds__(LDS_PROBE_TPB) k_lds_probe(int rounds, double *out)
{
    __shared__ double2 buf[LDS_PROBE_WORDS];
    for (int i = threadIdx.x; i < LDS_PROBE_WORDS; i += LDS_PROBE_TPB)
        buf[i] = make_double2((double)i, (double)(i ^ 7));
    __syncthreads();
    double sx = 0.0, sy = 0.0;
    for (int r = 0; r < rounds; r++) {
        const int base = (threadIdx.x + r * 64) & (LDS_PROBE_WORDS - 1);
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const double2 v = buf[(base + j * LDS_PROBE_TPB) & (LDS_PROBE_WORDS - 1)];
            sx += v.x;
            sy += v.y;
        }
    }
    if (sx + sy == -1.0) out[blockIdx.x] = sx;      // never true: keeps the reads
}

}  // namespace

extern "C" {

const char *dress_hip_last_error(void) { return t_err.c_str(); }

int dress_hip_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        hipGetLastError();
        return 0;
    }
    return n;
}

void dress_hip_set_precision(dress_hip_precision_t p)
{
    default_precision();
    std::lock_guard<std::mutex> lk(g_mu);
    g_precision = (p == DRESS_HIP_FP32_FP64) ? PREC_FP32_FP64 : PREC_FP64;
}

dress_hip_precision_t dress_hip_get_precision(void)
{
    return default_precision() == PREC_FP32_FP64 ? DRESS_HIP_FP32_FP64 : DRESS_HIP_FP64;
}

void dress_hip_last_fit_stats(dress_hip_fit_stats_t *out)
{
    if (out) *out = t_stats;
}

// dress.h:80-83 / dress.c:378-427.  The CSR and (when requested) the
// intercept arrays are built on the device and mirrored into the struct.
p_dress_graph_t dress_init_graph(int N, int E, int *U, int *V, double *W, double *NW,
                                 dress_variant_t variant, int precompute_intercepts)
{
    clear_error();
    p_dress_graph_t g = (p_dress_graph_t)calloc(1, sizeof(dress_graph_t));
    if (!g) { set_error("host allocation of dress_graph_t failed"); return nullptr; }
    g->variant = variant;
    g->N = N; g->E = E; g->U = U; g->V = V; g->W = W;
    g->precompute_intercepts = precompute_intercepts;
    // the host arrays of dress_graph_t (free() releases every one): large
    // ones on huge pages, filled by several host threads
    if (NW) g->NW = NW;
    else {
        g->NW = (double *)host_alloc(((size_t)N + 1) * sizeof(double));
        if (g->NW) fill_parallel(g->NW, (size_t)N, 1.0);
    }
    g->vertex_dress = (double *)calloc((size_t)N + 1, sizeof(double));
    g->edge_dress = (double *)host_alloc(((size_t)E + 1) * sizeof(double));
    g->edge_dress_next = (double *)host_alloc(((size_t)E + 1) * sizeof(double));
    g->adj_offset = (int *)host_alloc(((size_t)N + 1) * sizeof(int));
    g->edge_weight = (double *)host_alloc(((size_t)E + 1) * sizeof(double));
    DevGraph *G = new DevGraph();
    bool ok = g->NW && g->vertex_dress && g->edge_dress && g->edge_dress_next && g->adj_offset &&
              g->edge_weight;
    if (!ok) set_error("host allocation failed in dress_init_graph");
    ok = ok && build_graph(G, N, E, U, V, W, g->NW, (int)variant);
    if (ok) {
        g->adj_target = (int *)host_alloc(((size_t)G->S + 1) * sizeof(int));
        g->adj_edge_idx = (int *)host_alloc(((size_t)G->S + 1) * sizeof(int));
        ok = g->adj_target && g->adj_edge_idx;
        if (!ok) set_error("host allocation failed in dress_init_graph");
    }
    ok = ok && download_csr(G, g->adj_offset, g->adj_target, g->adj_edge_idx, g->edge_weight);
    if (ok) {
        g->max_degree = G->max_degree;
        fill_parallel(g->edge_dress, (size_t)E, 1.0);
        fill_parallel(g->edge_dress_next, (size_t)E, 1.0);
        if (precompute_intercepts)
            ok = download_index(G, &g->intercept_offset, &g->intercept_edge_ux, &g->intercept_edge_vx);
    }
    if (ok) release_reference_csr(G);
    if (!ok) {
        free_graph(G);
        free(g->adj_offset); free(g->adj_target); free(g->adj_edge_idx); free(g->edge_weight);
        free(g->edge_dress); free(g->edge_dress_next); free(g->vertex_dress);
        if (!NW) free(g->NW);
        free(g);
        return nullptr;
    }
    table_put(g, G);
    return g;
}

// dress.h:93 / dress.c:615-643, plus the device state
void dress_free_graph(p_dress_graph_t g)
{
    if (!g) return;
    DevGraph *G = table_take(g);
    if (G) free_graph(G);
    free(g->adj_offset); free(g->adj_target); free(g->adj_edge_idx);
    free(g->W); free(g->edge_weight); free(g->edge_dress); free(g->edge_dress_next);
    free(g->vertex_dress); free(g->NW);
    if (g->precompute_intercepts) {
        free(g->intercept_offset); free(g->intercept_edge_ux); free(g->intercept_edge_vx);
    }
    free(g->U); free(g->V);
    free(g);
}

void dress_fit_hip(p_dress_graph_t g, int max_iterations, double epsilon, int *iterations,
                   double *delta)
{
    if (fit_impl(g, max_iterations, epsilon, default_precision(), iterations, delta) != 0) {
        if (iterations) *iterations = -1;
        if (delta) *delta = NAN;
    }
}

// In libdress_hip.so the plain name runs on the GPU as well.
void dress_fit(p_dress_graph_t g, int max_iterations, double epsilon, int *iterations, double *delta)
{
    dress_fit_hip(g, max_iterations, epsilon, iterations, delta);
}

int dress_fit_hip_ex(p_dress_graph_t g, int max_iterations, double epsilon,
                     dress_hip_precision_t precision, int *iterations, double *delta)
{
    int r = fit_impl(g, max_iterations, epsilon,
                     precision == DRESS_HIP_FP32_FP64 ? PREC_FP32_FP64 : PREC_FP64, iterations, delta);
    if (r != 0) {
        if (iterations) *iterations = -1;
        if (delta) *delta = NAN;
    }
    return r;
}

int dress_hip_graph_info(p_dress_graph_t g, int64_t *info, int n)
{
    clear_error();
    DevGraph *G = g ? device_graph(g) : nullptr;
    if (!G || !build_bins(G)) return -1;
    int64_t v[32] = {0};
    v[0] = G->N; v[1] = G->E; v[2] = G->S; v[3] = G->T; v[4] = G->max_degree;
    v[5] = G->packed ? 1 : 0;
    for (int b = 0; b < Bins::NB; b++) {
        v[6 + b] = G->ebin[b + 1] - G->ebin[b];
        v[11 + b] = G->vbins.count[b];
    }
    unsigned long long *bt = nullptr;
    if (!dalloc(&bt, Bins::NB, "bin terms")) return -1;
    unsigned long long hb[Bins::NB] = {0, 0, 0, 0, 0};
    bool ok = hipMemsetAsync(bt, 0, sizeof(hb), G->stream) == hipSuccess;
    if (ok && G->E > 0)
        hipLaunchKernelGGL(k_bin_terms, dim3(1024), dim3(256), 0, G->stream, G->cnt_in, G->E, bt);
    ok = ok && hipMemcpyAsync(hb, bt, sizeof(hb), hipMemcpyDeviceToHost, G->stream) == hipSuccess &&
         hipStreamSynchronize(G->stream) == hipSuccess;
    dfree(bt);
    if (!ok) { set_error("dress_hip_graph_info failed"); return -1; }
    for (int b = 0; b < Bins::NB; b++) v[16 + b] = (int64_t)hb[b];
    for (int i = 0; i < n && i < 21; i++) info[i] = v[i];
    return 0;
}

int dress_hip_resident_reset(p_dress_graph_t g)
{
    clear_error();
    if (!g) { set_error("NULL graph"); return -1; }
    DevGraph *G = device_graph(g);
    if (!G) return -1;
    if (!refresh_self_terms(G, g->NW) || !fit_upload_values(G, g->edge_dress) || !build_bins(G)) return -1;
    if (hipStreamSynchronize(G->stream) != hipSuccess) { set_error("stream sync failed"); return -1; }
    return 0;
}

int dress_hip_fit_resident(p_dress_graph_t g, int max_iterations, double epsilon,
                           dress_hip_precision_t precision, int *iterations, double *delta)
{
    clear_error();
    DevGraph *G = g ? lookup(g) : nullptr;
    if (!G || !G->d[0]) { set_error("dress_hip_fit_resident: call dress_hip_resident_reset first"); return -1; }
    FitResult r;
    if (!fit_run(G, max_iterations, epsilon,
                 precision == DRESS_HIP_FP32_FP64 ? PREC_FP32_FP64 : PREC_FP64, &r))
        return -1;
    G->fit_state_valid = true;
    record_stats(G, r);
    if (iterations) *iterations = r.iterations;
    if (delta) *delta = r.delta;
    return 0;
}

int dress_hip_resident_fill(p_dress_graph_t g, double value)
{
    clear_error();
    DevGraph *G = g ? device_graph(g) : nullptr;
    if (!G) return -1;
    // only the values change: vertex weights and self terms stay as the
    // last reset or fit loaded them
    G->fit_state_valid = false;
    return fit_fill_values(G, value, true) ? 0 : -1;
}

int dress_hip_time_phases(p_dress_graph_t g, int reps, double *ms)
{
    clear_error();
    DevGraph *G = g ? lookup(g) : nullptr;
    if (!G || !G->d[0]) { set_error("dress_hip_time_phases: call dress_hip_resident_reset first"); return -1; }
    if (!ms || reps < 1) { set_error("dress_hip_time_phases: bad arguments"); return -1; }
    G->fit_state_valid = false;              // the timed phases rewrite norms and values
    return fit_time_phases(G, reps, ms) ? 0 : -1;
}

int dress_hip_copy_probe(int64_t n_bytes, int reps, double *ms)
{
    clear_error();
    if (n_bytes < 16 || (n_bytes & 15) || reps < 1 || !ms) {
        set_error("dress_hip_copy_probe: n_bytes must be a positive multiple of 16, reps >= 1");
        return -1;
    }
    void *a = nullptr, *b = nullptr;
    if (!dmalloc(&a, (size_t)n_bytes, "copy probe source") || !dmalloc(&b, (size_t)n_bytes, "copy probe target")) {
        dfree(a);
        return -1;
    }
    hipMemset(a, 0, (size_t)n_bytes);
    const int64_t n = n_bytes / 16;
    const int grid = (int)((n + COPY_WORDS_PER_BLOCK - 1) / COPY_WORDS_PER_BLOCK);
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, 0, (const double2 *)a, (double2 *)b, n);
    hipEventRecord(e0, 0);
    for (int r = 0; r < reps; r++)
        hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, 0, (const double2 *)a, (double2 *)b, n);
    hipEventRecord(e1, 0);
    hipError_t err = hipEventSynchronize(e1);
    float t = 0.f;
    hipEventElapsedTime(&t, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    dfree(a);
    dfree(b);
    if (err != hipSuccess) {
        set_error(std::string("dress_hip_copy_probe: ") + hipGetErrorString(err));
        return -1;
    }
    *ms = t / reps;
    return 0;
}

int dress_hip_gather_probe(int64_t src_bytes, int64_t n_gathers, int reps, double *ms)
{
    clear_error();
    if (src_bytes < 8 || src_bytes / 8 > (int64_t)UINT32_MAX || n_gathers < 1 || reps < 1 || !ms) {
        set_error("dress_hip_gather_probe: bad arguments");
        return -1;
    }
    double *src = nullptr, *out = nullptr;
    uint32_t *idx = nullptr;
    if (!dalloc(&src, (size_t)(src_bytes / 8), "gather probe source") ||
        !dalloc(&idx, (size_t)n_gathers, "gather probe indices") || !dalloc(&out, (size_t)n_gathers, "gather probe output")) {
        dfree(src); dfree(idx);
        return -1;
    }
    (void)hipMemset(src, 0, (size_t)src_bytes);
    const int grid = num_cus() * 16;
    hipLaunchKernelGGL(k_fill_random_idx, dim3(grid), dim3(256), 0, 0, idx, n_gathers, (uint32_t)(src_bytes / 8),
                       (uint64_t)12345);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    hipLaunchKernelGGL(k_gather_probe, dim3(grid), dim3(256), 0, 0, (const double *)src, (const uint32_t *)idx, out,
                       n_gathers);
    (void)hipEventRecord(e0, 0);
    for (int r = 0; r < reps; r++)
        hipLaunchKernelGGL(k_gather_probe, dim3(grid), dim3(256), 0, 0, (const double *)src, (const uint32_t *)idx,
                           out, n_gathers);
    (void)hipEventRecord(e1, 0);
    hipError_t err = hipEventSynchronize(e1);
    float t = 0.f;
    (void)hipEventElapsedTime(&t, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    dfree(src); dfree(idx); dfree(out);
    if (err != hipSuccess) {
        set_error(std::string("dress_hip_gather_probe: ") + hipGetErrorString(err));
        return -1;
    }
    *ms = t / reps;
    return 0;
}

int dress_hip_lds_probe(int rounds, double *ms, double *bytes)
{
    clear_error();
    if (rounds < 1 || !ms || !bytes) {
        set_error("dress_hip_lds_probe: bad arguments");
        return -1;
    }
    double *out = nullptr;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_lds_probe, LDS_PROBE_TPB, 0) != hipSuccess ||
        per_cu < 1) {
        (void)hipGetLastError();
        per_cu = 4;
    }
    const int grid = num_cus() * per_cu;
    if (!dalloc(&out, (size_t)grid, "lds probe output")) return -1;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    hipLaunchKernelGGL(k_lds_probe, dim3(grid), dim3(LDS_PROBE_TPB), 0, 0, rounds, out);
    (void)hipEventRecord(e0, 0);
    hipLaunchKernelGGL(k_lds_probe, dim3(grid), dim3(LDS_PROBE_TPB), 0, 0, rounds, out);
    (void)hipEventRecord(e1, 0);
    hipError_t err = hipEventSynchronize(e1);
    float t = 0.f;
    (void)hipEventElapsedTime(&t, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    dfree(out);
    if (err != hipSuccess) {
        set_error(std::string("dress_hip_lds_probe: ") + hipGetErrorString(err));
        return -1;
    }
    *ms = t;
    *bytes = (double)grid * LDS_PROBE_TPB * (double)rounds * 8.0 * 16.0;
    return 0;
}

int dress_hip_resident_download(p_dress_graph_t g)
{
    clear_error();
    DevGraph *G = g ? lookup(g) : nullptr;
    if (!G || !G->d[0]) { set_error("dress_hip_resident_download: no resident state"); return -1; }
    return fit_download(G, g->edge_dress, g->vertex_dress) ? 0 : -1;
}

int64_t dress_delta_binom(int n, int k)
{
    // delta_dress_impl.c:194-204
    if (k < 0 || k > n) return 0;
    if (k == 0 || k == n) return 1;
    if (k > n - k) k = n - k;
    int64_t r = 1;
    for (int i = 0; i < k; i++) r = r * (n - i) / (i + 1);
    return r;
}

int64_t dress_nabla_perm_count(int n, int k)
{
    // nabla_dress_impl.c:111-118
    if (k < 0 || k > n) return 0;
    int64_t r = 1;
    for (int i = 0; i < k; i++) r *= (n - i);
    return r;
}

}  // extern "C"
