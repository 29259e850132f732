// hipgrid_common.h -- host-side helpers shared by the engine's translation
// units: the per-thread error message behind hg_last_error, HIP error
// mapping, environment knobs, typed device allocation, pinned staging of
// host<->device copies.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
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

template <typename T>
hipError_t dalloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    return hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T));
}

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
