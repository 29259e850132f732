// hipgrid_common.h -- host-side helpers shared by the engine's translation
// units: the per-thread error message behind hg_last_error, HIP error
// mapping, environment knobs, typed device allocation from a cache, pinned
// staging of host<->device copies.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <list>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <utility>
#include <vector>

#include "hipgrid.h"

namespace hgc {

inline thread_local std::string g_err;

inline int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

inline int env_int(const char* name, int dflt) {
    const char* v = std::getenv(name);
    if (!v || !*v) return dflt;
    const int x = std::atoi(v);
    return x > 0 ? x : dflt;
}

// Device memory cache.  Graphs and batches are created and destroyed per
// request in a serving loop, several lanes at once, and hipMalloc / hipFree
// are slow and serialise the lanes.  Freed blocks are kept per (device,
// size) for the next allocation of that size, up to HG_POOL_MB of cached
// memory per process (default 32768; the blocks freed longest ago go first):
// blocks of up to kPoolMaxBlock bytes are rounded up to a power of two,
// larger ones to 2 MB (requests of one shape reuse them).  When the device
// is full the cache is emptied and the allocation retried.  The owner frees
// only after its stream has drained, so a cached block is idle when it is
// handed out again; its contents are whatever the last owner left.
struct DevPool {
    using Key = std::pair<int, size_t>;  // (device, block bytes)
    std::mutex m;
    std::unordered_map<void*, Key> live;
    std::list<std::pair<Key, void*>> idle;  // least recently freed first
    std::map<Key, std::vector<std::list<std::pair<Key, void*>>::iterator>> by_key;
    size_t cached = 0, cap = 0;
};
constexpr size_t kPoolMaxBlock = size_t(512) << 20;

inline DevPool& dev_pool() {
    static DevPool* pool = [] {
        DevPool* d = new DevPool();
        const char* v = std::getenv("HG_POOL_MB");
        const long mb = v && *v ? std::atol(v) : 32768;
        d->cap = mb > 0 ? size_t(mb) << 20 : 0;
        return d;
    }();
    return *pool;
}

// Frees every cached block (when the runtime reports the device full).
inline void pool_trim() {
    DevPool& d = dev_pool();
    std::vector<void*> all;
    {
        std::lock_guard<std::mutex> lk(d.m);
        for (auto& kv : d.idle) all.push_back(kv.second);
        d.idle.clear();
        d.by_key.clear();
        d.cached = 0;
    }
    for (void* q : all) (void)hipFree(q);
}

inline hipError_t malloc_or_trim(void** p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes);
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        pool_trim();
        e = hipMalloc(p, bytes);
    }
    return e;
}

// Test hook (hg_debug_fail_alloc): while positive, counts device allocations
// down; the one that takes it to zero fails as if the device were full.
inline std::atomic<int64_t>& fail_countdown() {
    static std::atomic<int64_t> c{0};
    return c;
}

inline hipError_t pool_malloc(void** p, size_t bytes) {
    *p = nullptr;
    if (fail_countdown().load(std::memory_order_relaxed) > 0 && fail_countdown().fetch_sub(1) == 1)
        return hipErrorOutOfMemory;
    DevPool& d = dev_pool();
    if (d.cap == 0) return malloc_or_trim(p, bytes);  // cache off: exact sizes, freed at once
    size_t c = 256;
    if (bytes > kPoolMaxBlock) {
        const size_t g = size_t(2) << 20;
        c = (bytes + g - 1) / g * g;
    } else {
        while (c < bytes) c <<= 1;
    }
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    const DevPool::Key key{dev, c};
    {
        std::lock_guard<std::mutex> lk(d.m);
        auto it = d.by_key.find(key);
        if (it != d.by_key.end() && !it->second.empty()) {
            auto li = it->second.back();
            it->second.pop_back();
            *p = li->second;
            d.idle.erase(li);
            d.cached -= c;
            d.live[*p] = key;
            return hipSuccess;
        }
    }
    e = malloc_or_trim(p, c);
    if (e == hipSuccess) {
        std::lock_guard<std::mutex> lk(d.m);
        d.live[*p] = key;
    }
    return e;
}

// Returns a block to the cache (or to the runtime); p must be idle.  The
// blocks freed longest ago make room for it when the cache is full.
inline void dfree(void* p) {
    if (!p) return;
    DevPool& d = dev_pool();
    std::vector<void*> evict;
    bool kept = false;
    {
        std::lock_guard<std::mutex> lk(d.m);
        auto it = d.live.find(p);
        if (it != d.live.end()) {
            const DevPool::Key key = it->second;
            d.live.erase(it);
            if (key.second <= d.cap) {
                while (d.cached + key.second > d.cap && !d.idle.empty()) {
                    auto old = d.idle.begin();
                    auto& v = d.by_key[old->first];
                    v.erase(std::find(v.begin(), v.end(), old));
                    d.cached -= old->first.second;
                    evict.push_back(old->second);
                    d.idle.erase(old);
                }
                d.idle.emplace_back(key, p);
                d.by_key[key].push_back(std::prev(d.idle.end()));
                d.cached += key.second;
                kept = true;
            }
        }
    }
    for (void* q : evict) (void)hipFree(q);
    if (!kept) (void)hipFree(p);
}

template <typename T>
hipError_t dalloc(T** p, size_t count) {
    if (count == 0) count = 1;
    void* q = nullptr;
    const hipError_t e = pool_malloc(&q, count * sizeof(T));
    *p = static_cast<T*>(q);
    return e;
}

// Bytes the cache hands out for a request of `bytes` (blocks of up to
// kPoolMaxBlock rounded up to a power of two, at least 256; larger ones to
// 2 MB granules; exact when the cache is off): what an allocation really
// holds, which the byte accounting and the server's budget estimates count.
inline uint64_t pool_block_bytes(uint64_t bytes) {
    if (bytes == 0) bytes = 1;
    if (dev_pool().cap == 0) return bytes;
    if (bytes > kPoolMaxBlock) {
        const uint64_t g = uint64_t(2) << 20;
        return (bytes + g - 1) / g * g;
    }
    uint64_t c = 256;
    while (c < bytes) c <<= 1;
    return c;
}

// Binned apply of dense removal rounds (k_bin_*): used by graphs of at least
// HG_BIN_MIN directed entries (default 2^26) and fewer than 2^32; HG_BIN=2
// turns it off.  Its two record buffers hold bin_rcap(D) 4-byte records each.
inline bool bin_enabled(uint64_t D) {
    const int64_t lo = env_int("HG_BIN_MIN", 1 << 26);
    return env_int("HG_BIN", 1) == 1 && D >= uint64_t(lo) && D < (1ull << 32);
}
inline uint64_t bin_rcap(uint64_t D) { return std::min<uint64_t>(3ull << 29, 2 * D + (1ull << 20)); }
// Device bytes of the binned apply (record buffers and per-bin tables) for a
// graph of D entries and L layers.
inline uint64_t bin_bytes(uint64_t D, uint64_t L) {
    if (!bin_enabled(D)) return 0;
    return pool_block_bytes(8 * bin_rcap(D)) + 2 * pool_block_bytes(4 * 1024 * L) + pool_block_bytes(8 * 1024 * L) +
           pool_block_bytes(4 * 1025) + pool_block_bytes(2 * 4097) + pool_block_bytes(4 * 1025) +
           pool_block_bytes(2ull * 1024 * 512) + pool_block_bytes(8ull * 1024 * 512 * L);
}

// Core-index bounds per entry (adjh): for each lmd, one byte per entry holding
// min(h, 255), h = the lmd-th largest of the target's layer degrees counting
// only neighbours of z >= lmd (no vertex can have a core index above it for
// that lmd).  Kept for graphs of at most 16
// layers whose L * D bytes stay within 64 GiB; HG_HBOUND=0 turns them off.
inline bool hb_enabled(uint64_t D, uint64_t L) {
    return env_int("HG_HBOUND", 1) == 1 && L >= 1 && L <= 16 && D > 0 && L * D <= (1ull << 36);
}
inline uint64_t hb_bytes(uint64_t D, uint64_t L) { return hb_enabled(D, L) ? pool_block_bytes(L * D) : 0; }

// Runs fn(begin, end) over [0, count) in contiguous ranges of at least
// min_per_thread items on up to 16 host threads (the caller's included).
template <typename F>
void par_ranges(size_t count, size_t min_per_thread, F&& fn) {
    size_t T = std::min<size_t>(std::max(1u, std::thread::hardware_concurrency()), 16);
    T = std::min<size_t>(T, std::max<size_t>(1, count / std::max<size_t>(min_per_thread, 1)));
    if (T <= 1) {
        fn(size_t(0), count);
        return;
    }
    const size_t step = (count + T - 1) / T;
    std::vector<std::thread> th;
    th.reserve(T - 1);
    for (size_t t = 1; t < T; ++t) {
        const size_t b = t * step, e = std::min(count, b + step);
        if (b >= e) break;
        try {
            th.emplace_back([&fn, b, e] { fn(b, e); });
        } catch (...) {  // no thread available: run the range here
            fn(b, e);
        }
    }
    fn(size_t(0), std::min(count, step));
    for (auto& x : th) x.join();
}

inline void par_memcpy(void* dst, const void* src, size_t bytes) {
    par_ranges(bytes, size_t(2) << 20, [&](size_t b, size_t e) {
        std::memcpy(static_cast<char*>(dst) + b, static_cast<const char*>(src) + b, e - b);
    });
}

// Every copy between caller (pageable) memory and the device goes through
// pinned buffers the library owns: the host side is a plain memcpy into or
// out of a pinned slot, the device side a DMA on the handle's stream.  A
// caller buffer is never handed to the runtime, so its lifetime ends when
// the call returns, and uploads from several threads at once share no
// runtime staging.  Slot sets are pooled per device for the life of the
// process (kSlots x kSlotBytes of pinned memory each, one per thread that
// is copying at that moment).
struct SlotSet {
    static constexpr int kSlots = 4;
    static constexpr size_t kSlotBytes = size_t(16) << 20;
    int device = 0;
    char* buf[kSlots] = {};
    hipEvent_t ev[kSlots] = {};
};

inline std::mutex& slot_pool_mutex() {
    static std::mutex m;
    return m;
}
inline std::vector<SlotSet*>& slot_pool() {
    static std::vector<SlotSet*> p;
    return p;
}

class Staging {
public:
    explicit Staging(hipStream_t stream) : stream_(stream) {}
    Staging(const Staging&) = delete;
    Staging& operator=(const Staging&) = delete;
    ~Staging() {
        if (!set_) return;
        for (int s = 0; s < SlotSet::kSlots; ++s)
            if (busy_[s]) (void)hipEventSynchronize(set_->ev[s]);
        std::lock_guard<std::mutex> lk(slot_pool_mutex());
        slot_pool().push_back(set_);
    }

    // Queues dst[0, bytes) = src[0, bytes) on the stream; src may be reused
    // as soon as this returns.
    hipError_t h2d(void* dst, const void* src, size_t bytes) {
        hipError_t e = acquire();
        for (size_t off = 0; off < bytes && e == hipSuccess; off += SlotSet::kSlotBytes) {
            const size_t len = std::min(SlotSet::kSlotBytes, bytes - off);
            const int s = take();
            if ((e = settle(s)) != hipSuccess) break;
            par_memcpy(set_->buf[s], static_cast<const char*>(src) + off, len);
            e = hipMemcpyAsync(static_cast<char*>(dst) + off, set_->buf[s], len, hipMemcpyHostToDevice, stream_);
            if (e == hipSuccess) e = hipEventRecord(set_->ev[s], stream_);
            busy_[s] = e == hipSuccess;
        }
        return e;
    }

    // dst[0, bytes) = src[0, bytes) after the work queued before it on the
    // stream; complete when this returns.
    hipError_t d2h(void* dst, const void* src, size_t bytes) {
        hipError_t e = acquire();
        for (size_t off = 0; off < bytes && e == hipSuccess; off += SlotSet::kSlotBytes) {
            const size_t len = std::min(SlotSet::kSlotBytes, bytes - off);
            const int s = take();
            if ((e = settle(s)) != hipSuccess) break;
            e = hipMemcpyAsync(set_->buf[s], static_cast<const char*>(src) + off, len, hipMemcpyDeviceToHost, stream_);
            if (e == hipSuccess) e = hipEventRecord(set_->ev[s], stream_);
            busy_[s] = e == hipSuccess;
            out_[s] = static_cast<char*>(dst) + off;
            out_len_[s] = len;
        }
        for (int s = 0; s < SlotSet::kSlots; ++s) {
            const hipError_t es = settle(s);
            if (e == hipSuccess) e = es;
        }
        return e;
    }

private:
    hipError_t acquire() {
        if (set_) return hipSuccess;
        int dev = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e != hipSuccess) return e;
        {
            std::lock_guard<std::mutex> lk(slot_pool_mutex());
            auto& p = slot_pool();
            for (size_t i = 0; i < p.size(); ++i)
                if (p[i]->device == dev) {
                    set_ = p[i];
                    p.erase(p.begin() + std::ptrdiff_t(i));
                    return hipSuccess;
                }
        }
        SlotSet* s = new SlotSet();
        s->device = dev;
        for (int i = 0; i < SlotSet::kSlots && e == hipSuccess; ++i) {
            e = hipHostMalloc(reinterpret_cast<void**>(&s->buf[i]), SlotSet::kSlotBytes, hipHostMallocDefault);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&s->ev[i], hipEventDisableTiming);
        }
        if (e != hipSuccess) {
            for (int i = 0; i < SlotSet::kSlots; ++i) {
                if (s->buf[i]) (void)hipHostFree(s->buf[i]);
                if (s->ev[i]) (void)hipEventDestroy(s->ev[i]);
            }
            delete s;
            return e;
        }
        set_ = s;
        return hipSuccess;
    }
    int take() {
        const int s = next_;
        next_ = (next_ + 1) % SlotSet::kSlots;
        return s;
    }
    // Waits for the slot's last copy and delivers a finished download.
    hipError_t settle(int s) {
        hipError_t e = hipSuccess;
        if (busy_[s]) e = hipEventSynchronize(set_->ev[s]);
        busy_[s] = false;
        if (out_[s] && e == hipSuccess) par_memcpy(out_[s], set_->buf[s], out_len_[s]);
        out_[s] = nullptr;
        return e;
    }

    hipStream_t stream_;
    SlotSet* set_ = nullptr;
    int next_ = 0;
    bool busy_[SlotSet::kSlots] = {};
    char* out_[SlotSet::kSlots] = {};
    size_t out_len_[SlotSet::kSlots] = {};
};

}  // namespace hgc

#define HIP_TRY(expr)                                                                                  \
    do {                                                                                               \
        hipError_t _e = (expr);                                                                        \
        if (_e != hipSuccess) {                                                                        \
            return hgc::fail(_e == hipErrorOutOfMemory ? HG_ERR_NOMEM : HG_ERR_HIP,                    \
                             std::string(#expr) + ": " + hipGetErrorString(_e));                      \
        }                                                                                              \
    } while (0)
