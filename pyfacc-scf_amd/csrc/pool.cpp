// Large device scratch blocks kept across calls.
//
// hipMalloc and hipFree of a block of hundreds of MB cost about 0.3 ms each
// on the MI355X, which every SCF call paid for its per-class J'_cd buffers.
// An owner takes blocks with pool_acquire and returns them with pool_release
// only once no work that uses them is pending (the engine releases from its
// destructor, after synchronising its stream), so a block handed out again
// is idle.  Blocks stay allocated until the process ends (at most POOL_KEEP
// of them, per process; beyond that a released block is freed).
#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

#include "pf_common.h"

namespace pf {

namespace {

constexpr size_t POOL_ALIGN = size_t(2) << 20;
constexpr size_t POOL_KEEP = 12;

std::mutex &pool_mutex() {
    static std::mutex mu;
    return mu;
}
std::vector<PooledBlock> &pool_free() {
    static auto *v = new std::vector<PooledBlock>();  // never destroyed: blocks outlive static teardown
    return *v;
}

}  // namespace

PooledBlock pool_acquire(size_t bytes) {
    bytes = std::max(POOL_ALIGN, (bytes + POOL_ALIGN - 1) / POOL_ALIGN * POOL_ALIGN);
    int dev = 0;
    PF_HIP(hipGetDevice(&dev));
    {
        std::lock_guard<std::mutex> lock(pool_mutex());
        auto &fr = pool_free();
        int best = -1;
        for (int i = 0; i < (int)fr.size(); ++i)  // smallest block that fits, at most twice the request
            if (fr[i].dev == dev && fr[i].bytes >= bytes && fr[i].bytes <= 2 * bytes &&
                (best < 0 || fr[i].bytes < fr[best].bytes))
                best = i;
        if (best >= 0) {
            const PooledBlock b = fr[best];
            fr.erase(fr.begin() + best);
            return b;
        }
    }
    PooledBlock b;
    PF_HIP(hipMalloc(&b.p, bytes));
    b.bytes = bytes;
    b.dev = dev;
    return b;
}

void pool_release(PooledBlock &b) {
    if (!b.p) return;
    bool keep;
    {
        std::lock_guard<std::mutex> lock(pool_mutex());
        auto &fr = pool_free();
        keep = fr.size() < POOL_KEEP;
        if (keep) fr.push_back(b);
    }
    if (!keep) {
        int cur = 0;
        if (hipGetDevice(&cur) == hipSuccess && cur != b.dev) (void)hipSetDevice(b.dev);
        (void)hipFree(b.p);
        if (cur != b.dev) (void)hipSetDevice(cur);
    }
    b = PooledBlock{};
}

// ---- per-thread cache behind DevBuf
//
// Every call of the library issues its device work on one stream and
// synchronises that stream before it returns.  A buffer released during a
// call can therefore be reused by a later allocation of the same thread:
// within the call the new owner's work is ordered after the old owner's on
// the same stream, and across calls the old stream is idle.  Buffers are
// kept per thread and by exact size class, so a repeated SCF reuses every
// buffer of the previous one.  At most CACHE_MAX bytes stay cached per
// thread (beyond: hipFree); blocks over CACHE_BLOCK_MAX (the batched
// integral store) are never cached.
//
// When a thread ends, its blocks move to a process-wide list of idle blocks
// (every call of that thread has drained its stream, so nothing pending uses
// them), capped at ORPHAN_MAX bytes, the rest freed; any thread's allocation
// that misses its own cache takes from that list first.  Threads that come
// and go therefore reuse the same memory instead of leaking their caches.
// pf_release_device_cache() frees the calling thread's cache, the idle list
// and the idle J'_cd pool blocks.
namespace {

constexpr size_t CACHE_MAX = size_t(8) << 30;
constexpr size_t CACHE_BLOCK_MAX = size_t(2) << 30;
constexpr size_t ORPHAN_MAX = size_t(8) << 30;

size_t size_class(size_t bytes) {
    if (bytes <= 4096) return 4096;
    if (bytes <= (size_t(1) << 20)) {  // powers of two up to 1 MiB
        size_t c = 4096;
        while (c < bytes) c <<= 1;
        return c;
    }
    const size_t mib = size_t(1) << 20;  // then whole MiB
    return (bytes + mib - 1) / mib * mib;
}

using BlockMap = std::map<std::pair<int, size_t>, std::vector<void *>>;  // (device, class) -> blocks

struct Orphans {
    std::mutex mu;
    BlockMap free;
    size_t bytes = 0;
};
Orphans &orphans() {
    static auto *o = new Orphans();  // never destroyed: blocks outlive static teardown
    return *o;
}

void free_blocks(BlockMap &m) {
    int cur = -1;
    (void)hipGetDevice(&cur);
    for (auto &kv : m)
        for (void *p : kv.second) {
            (void)hipSetDevice(kv.first.first);
            (void)hipFree(p);
        }
    if (cur >= 0) (void)hipSetDevice(cur);
    m.clear();
}

struct ThreadCache {
    BlockMap free;
    size_t bytes = 0;
    ~ThreadCache() {  // thread exit: hand the idle blocks to the process-wide list
        BlockMap spill;
        {
            auto &o = orphans();
            std::lock_guard<std::mutex> lock(o.mu);
            for (auto &kv : free)
                for (void *p : kv.second) {
                    if (o.bytes + kv.first.second <= ORPHAN_MAX) {
                        o.free[kv.first].push_back(p);
                        o.bytes += kv.first.second;
                    } else {
                        spill[kv.first].push_back(p);
                    }
                }
        }
        free_blocks(spill);
    }
};
ThreadCache &thread_cache() {
    static thread_local ThreadCache c;
    return c;
}

}  // namespace

void *dev_alloc(size_t bytes) {
    if (bytes == 0) return nullptr;
    const size_t cls = size_class(bytes);
    int dev = 0;
    PF_HIP(hipGetDevice(&dev));
    if (cls <= CACHE_BLOCK_MAX) {
        auto &tc = thread_cache();
        auto it = tc.free.find({dev, cls});
        if (it != tc.free.end() && !it->second.empty()) {
            void *p = it->second.back();
            it->second.pop_back();
            tc.bytes -= cls;
            return p;
        }
        auto &o = orphans();
        std::lock_guard<std::mutex> lock(o.mu);
        auto jt = o.free.find({dev, cls});
        if (jt != o.free.end() && !jt->second.empty()) {
            void *p = jt->second.back();
            jt->second.pop_back();
            o.bytes -= cls;
            return p;
        }
    }
    void *p = nullptr;
    PF_HIP(hipMalloc(&p, cls));
    return p;
}

void dev_free(void *p, size_t bytes) {
    if (!p) return;
    const size_t cls = size_class(bytes);
    auto &tc = thread_cache();
    int dev = 0;
    if (cls <= CACHE_BLOCK_MAX && tc.bytes + cls <= CACHE_MAX && hipGetDevice(&dev) == hipSuccess) {
        tc.free[{dev, cls}].push_back(p);
        tc.bytes += cls;
        return;
    }
    (void)hipFree(p);
}

void release_device_cache() {
    auto &tc = thread_cache();
    free_blocks(tc.free);
    tc.bytes = 0;
    {
        auto &o = orphans();
        std::lock_guard<std::mutex> lock(o.mu);
        free_blocks(o.free);
        o.bytes = 0;
    }
    std::vector<PooledBlock> idle;
    {
        std::lock_guard<std::mutex> lock(pool_mutex());
        idle.swap(pool_free());
    }
    int cur = -1;
    (void)hipGetDevice(&cur);
    for (auto &b : idle) {
        (void)hipSetDevice(b.dev);
        (void)hipFree(b.p);
    }
    if (cur >= 0) (void)hipSetDevice(cur);
}

size_t cached_device_bytes() {
    size_t n = thread_cache().bytes;
    auto &o = orphans();
    std::lock_guard<std::mutex> lock(o.mu);
    return n + o.bytes;
}

// ---- per-thread pool of HIP events
//
// An SCF records a few hundred timing events; creating and destroying them one
// by one cost milliseconds of host time per call.  Events go back to the
// calling thread's pool once read (the work they bracket is complete) and are
// recorded again by later calls; like the block cache, the pool is never torn
// down.
namespace {
std::vector<hipEvent_t> &event_pool() {
    static thread_local std::vector<hipEvent_t> *v = new std::vector<hipEvent_t>();
    return *v;
}
}  // namespace

hipEvent_t event_acquire() {
    auto &pool = event_pool();
    if (!pool.empty()) {
        hipEvent_t e = pool.back();
        pool.pop_back();
        return e;
    }
    hipEvent_t e;
    PF_HIP(hipEventCreate(&e));
    return e;
}

void event_release(hipEvent_t e) {
    if (e) event_pool().push_back(e);
}

}  // namespace pf
