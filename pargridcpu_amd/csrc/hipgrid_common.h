// hipgrid_common.h -- host-side helpers shared by the engine's translation
// units: the per-thread error message behind hg_last_error, HIP error
// mapping, environment knobs, typed device allocation.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdlib>
#include <string>

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

}  // namespace hgc

#define HIP_TRY(expr)                                                                                  \
    do {                                                                                               \
        hipError_t _e = (expr);                                                                        \
        if (_e != hipSuccess) {                                                                        \
            return hgc::fail(_e == hipErrorOutOfMemory ? HG_ERR_NOMEM : HG_ERR_HIP,                    \
                             std::string(#expr) + ": " + hipGetErrorString(_e));                      \
        }                                                                                              \
    } while (0)
