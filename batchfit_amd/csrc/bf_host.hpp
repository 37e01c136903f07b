// bf_host.hpp — host-side pieces shared by the runtime translation units:
// the process-wide device/stream singleton and error helpers.
#pragma once

#include "bf_launch.hpp"

#include "../../include/batchfit_hip.h"

#include <cstdio>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>

namespace bfhip {

constexpr int kParams[4] = {3, 3, 4, 8};

inline int fail(char *err, size_t errlen, int code, const std::string &message) {
    if (err && errlen) {
        std::snprintf(err, errlen, "%s", message.c_str());
    }
    return code;
}

#define BF_TRY(expr, what)                                                                     \
    do {                                                                                       \
        const hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                                  \
            return fail(err, errlen, BF_RUNTIME_ERROR,                                         \
                        std::string(what) + " failed: " + hipGetErrorString(e_));             \
    } while (0)

struct Runtime {
    int status = BF_UNAVAILABLE;
    std::string why = "no HIP device";
    hipStream_t stream = nullptr;
    hipStream_t side[2] = {nullptr, nullptr};  // the host-array entry's chunk streams
    // Pinned staging slots of the host-array entry (hipHostMalloc, kept for
    // the process: pinning memory costs more than a chunk's copy).
    char *staging[2] = {nullptr, nullptr};
    size_t staging_bytes = 0;
    char *staging_out[2] = {nullptr, nullptr};  // a chunk's results, device → host
    size_t staging_out_bytes = 0;
    std::mutex host_calls;  // serialises the synchronous host-array entry (bindings.cpp:13)

    Runtime() {
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
            why = "no HIP device is visible";
            return;
        }
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, 0) != hipSuccess) {
            why = "hipGetDeviceProperties failed";
            return;
        }
        if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
            why = std::string("device 0 is ") + prop.gcnArchName + ", this library targets gfx950";
            return;
        }
        if (hipSetDevice(0) != hipSuccess ||
            hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) {
            why = "could not create a HIP stream";
            return;
        }
        hipMemPool_t pool;
        if (hipDeviceGetDefaultMemPool(&pool, 0) == hipSuccess) {
            uint64_t keep = std::numeric_limits<uint64_t>::max();
            (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
        }
        status = BF_OK;
        why.clear();
    }
    static Runtime &get() {
        static Runtime runtime;
        return runtime;
    }
};

inline int pow2_ceil(uint32_t v) {
    int r = 1;
    while (uint32_t(r) < v)
        r <<= 1;
    return r;
}

}  // namespace bfhip
