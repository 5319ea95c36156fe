// Device memory cache behind DBuf: freed blocks are kept per device and handed out again, so a
// request loop (band after band, request after request) stops paying hipMalloc / hipFree -- fresh
// device memory has to be mapped and cleared by the driver, and hipFree waits for the device.
//
// Stream safety: a block remembers the stream of its last owner and an event recorded there when it
// was freed; a new owner on another stream waits on that event (hipStreamWaitEvent) before its first
// use, an owner on the same stream needs nothing (the stream runs in order).
#include <cstdlib>
#include <map>
#include <mutex>

#include "arh_internal.h"

namespace arh {

thread_local hipStream_t tl_stream = nullptr;

namespace {

struct Block {
    void *p;
    size_t bytes;
    hipStream_t s;
    hipEvent_t ev;
};

struct DevPool {
    std::mutex mu;
    std::multimap<size_t, Block> free;
    size_t cached = 0;
};

constexpr int kMaxDevices = 64;
constexpr size_t kSmall = size_t(1) << 20;  // below: power-of-two classes; above: 2 MiB granules
constexpr size_t kGranule = size_t(2) << 20;

DevPool &pool(int dev) {
    // never destroyed: blocks may be returned by static destructors at process exit
    static DevPool *pools = new DevPool[kMaxDevices];
    return pools[dev];
}

bool enabled() {
    static const bool on = [] {
        const char *e = std::getenv("ARH_POOL");
        return !(e && e[0] == '0');
    }();
    return on;
}

size_t round_size(size_t n) {
    if (n <= 4096) return 4096;
    if (n < kSmall) {
        size_t c = 4096;
        while (c < n) c <<= 1;
        return c;
    }
    return (n + kGranule - 1) / kGranule * kGranule;
}

void release(Block &b) {
    if (b.ev) {
        (void)hipEventSynchronize(b.ev);
        (void)hipEventDestroy(b.ev);
    }
    (void)hipFree(b.p);
}

}  // namespace

void *dev_alloc(size_t &n, hipStream_t s) {
    int dev = 0;
    ARH_HIP(hipGetDevice(&dev));
    ARH_CHECK(dev >= 0 && dev < kMaxDevices, ARH_EINVAL, "device index out of range");
    n = round_size(n);
    if (enabled()) {
        DevPool &P = pool(dev);
        Block b{};
        bool hit = false;
        {
            std::lock_guard<std::mutex> lk(P.mu);
            // best fit among blocks of at most 1.25x the request (exact class for small blocks)
            auto it = P.free.lower_bound(n);
            const size_t cap = n < kSmall ? n : n + n / 4;
            if (it != P.free.end() && it->first <= cap) {
                b = it->second;
                P.free.erase(it);
                P.cached -= b.bytes;
                hit = true;
            }
        }
        if (hit) {
            if (b.ev) {
                if (b.s != s) ARH_HIP(hipStreamWaitEvent(s, b.ev, 0));
                (void)hipEventDestroy(b.ev);
            }
            n = b.bytes;
            return b.p;
        }
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        dev_trim(dev);
        e = hipMalloc(&p, n);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            throw Error(ARH_ENOMEM, "hipMalloc(" + std::to_string(n) + " B): " + hipGetErrorString(e));
        }
    }
    return p;
}

void *dev_alloc(size_t &n, hipStream_t s, int *device) {
    ARH_HIP(hipGetDevice(device));
    return dev_alloc(n, s);
}

void dev_free(void *p, size_t n, hipStream_t s, int dev) {
    if (!p) return;
    int cur = 0;
    (void)hipGetDevice(&cur);
    if (cur != dev) (void)hipSetDevice(dev);
    struct Restore {
        int cur, dev;
        ~Restore() {
            if (cur != dev) (void)hipSetDevice(cur);
        }
    } restore{cur, dev};
    if (!enabled() || dev < 0 || dev >= kMaxDevices) {
        (void)hipFree(p);
        return;
    }
    Block b{p, n, s, nullptr};
    if (hipEventCreateWithFlags(&b.ev, hipEventDisableTiming) != hipSuccess || hipEventRecord(b.ev, s) != hipSuccess) {
        // no event: wait for the stream instead, then the block is free for anyone
        if (b.ev) (void)hipEventDestroy(b.ev);
        b.ev = nullptr;
        (void)hipStreamSynchronize(s);
    }
    DevPool &P = pool(dev);
    std::lock_guard<std::mutex> lk(P.mu);
    P.free.emplace(n, b);
    P.cached += n;
}

size_t dev_trim(int dev) {
    if (dev < 0 || dev >= kMaxDevices) return 0;
    DevPool &P = pool(dev);
    std::multimap<size_t, Block> blocks;
    {
        std::lock_guard<std::mutex> lk(P.mu);
        blocks.swap(P.free);
        P.cached = 0;
    }
    size_t freed = 0;
    for (auto &kv : blocks) {
        freed += kv.second.bytes;
        release(kv.second);
    }
    return freed;
}

size_t dev_cached(int dev) {
    if (dev < 0 || dev >= kMaxDevices) return 0;
    DevPool &P = pool(dev);
    std::lock_guard<std::mutex> lk(P.mu);
    return P.cached;
}

}  // namespace arh
