// Process-wide caching allocators for device and pinned host memory.
//
// An INSCY run allocates a few hundred buffers (tree lists, per-subspace
// arrays, batch scratch, staging).  hipFree waits for the whole device, so
// freeing them at the end of every run would stall every other stream, and
// hipMalloc/hipHostMalloc of large blocks is slow.  Blocks therefore go back to
// a free list instead (per device for device memory) and are reused by later
// runs, on any thread.  A block is handed out again only after the run that
// released it has synchronised its stream, so reuse needs no further waiting.
// inscy_release_memory() returns the cached blocks to the runtime.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <map>
#include <mutex>
#include <unordered_map>
#include <utility>
#include <vector>

namespace inscy {

class BlockCache {
public:
    explicit BlockCache(bool host) : host_(host) {}

    hipError_t get(void **p, size_t bytes) {
        const size_t sz = round(std::max<size_t>(bytes, 1));
        const int dev = device();
        {
            std::lock_guard<std::mutex> g(mu_);
            auto &fl = free_[dev];
            auto it = fl.lower_bound(sz);
            // best fit, but never more than twice the request (plus 4 MiB)
            if (it != fl.end() && it->first <= 2 * sz + ((size_t) 4 << 20)) {
                *p = it->second;
                fl.erase(it);
                return hipSuccess;
            }
        }
        hipError_t e = raw_alloc(p, sz);
        if (e != hipSuccess) {
            (void) hipGetLastError();
            release(dev);
            e = raw_alloc(p, sz);
        }
        if (e == hipSuccess) {
            std::lock_guard<std::mutex> g(mu_);
            size_[*p] = {sz, dev};
        }
        return e;
    }

    void put(void *p) {
        if (!p) return;
        std::lock_guard<std::mutex> g(mu_);
        auto it = size_.find(p);
        if (it == size_.end()) return;
        free_[it->second.second].insert({it->second.first, p});
    }

    // bytes of the current device's cached blocks that no run holds
    size_t idle_bytes() {
        const int dev = device();
        std::lock_guard<std::mutex> g(mu_);
        size_t sum = 0;
        auto it = free_.find(dev);
        if (it != free_.end())
            for (const auto &b : it->second) sum += b.first;
        return sum;
    }

    // frees the cached (unused) blocks of one device, or of all (dev < 0)
    void release(int dev) {
        std::lock_guard<std::mutex> g(mu_);
        for (auto &kv : free_) {
            if (dev >= 0 && kv.first != dev) continue;
            for (auto &b : kv.second) {
                if (host_)
                    (void) hipHostFree(b.second);
                else
                    (void) hipFree(b.second);
                size_.erase(b.second);
            }
            kv.second.clear();
        }
    }

private:
    static size_t round(size_t b) {
        const size_t q = b <= ((size_t) 64 << 10) ? 256 : b <= ((size_t) 16 << 20) ? ((size_t) 64 << 10)
                                                                                   : ((size_t) 2 << 20);
        return (b + q - 1) / q * q;
    }
    int device() const {
        if (host_) return -1;
        int d = 0;
        (void) hipGetDevice(&d);
        return d;
    }
    hipError_t raw_alloc(void **p, size_t sz) { return host_ ? hipHostMalloc(p, sz) : hipMalloc(p, sz); }

    bool host_;
    std::mutex mu_;
    std::unordered_map<int, std::multimap<size_t, void *>> free_;
    std::unordered_map<void *, std::pair<size_t, int>> size_;
};

// The caches live until the process ends (their blocks are not freed at exit:
// the runtime may already be gone by then).
inline BlockCache &device_cache() {
    static BlockCache *c = new BlockCache(false);
    return *c;
}
inline BlockCache &host_cache() {
    static BlockCache *c = new BlockCache(true);
    return *c;
}

// Timing events of the profiling switch, reused across runs: creating and
// destroying two events per timed launch costs more than the launches.
class EventPool {
public:
    hipError_t get(hipEvent_t *e) {
        const int dev = device();
        {
            std::lock_guard<std::mutex> g(mu_);
            auto &fl = free_[dev];
            if (!fl.empty()) {
                *e = fl.back();
                fl.pop_back();
                return hipSuccess;
            }
        }
        return hipEventCreate(e);
    }
    void put(hipEvent_t e) {
        if (!e) return;
        std::lock_guard<std::mutex> g(mu_);
        free_[device()].push_back(e);
    }

private:
    static int device() {
        int d = 0;
        (void) hipGetDevice(&d);
        return d;
    }
    std::mutex mu_;
    std::unordered_map<int, std::vector<hipEvent_t>> free_;
};
inline EventPool &event_pool() {
    static EventPool *p = new EventPool;
    return *p;
}

// Device memory a run can still take: what the runtime reports free plus the
// cached blocks no run holds (an earlier run's buffers stay cached and would
// otherwise shrink every later run's budgets).
inline size_t device_available_bytes() {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return 0;
    return free_b + device_cache().idle_bytes();
}

template <class T>
hipError_t dev_alloc(T **p, size_t bytes) {
    return device_cache().get(reinterpret_cast<void **>(p), bytes);
}
inline void dev_free(void *p) { device_cache().put(p); }
template <class T>
hipError_t host_alloc(T **p, size_t bytes) {
    return host_cache().get(reinterpret_cast<void **>(p), bytes);
}
inline void host_free(void *p) { host_cache().put(p); }

}  // namespace inscy
