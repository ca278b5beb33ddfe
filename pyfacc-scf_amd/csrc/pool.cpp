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

}  // namespace pf
