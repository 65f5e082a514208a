// Internal helpers shared by the HIP translation units of libsubpx_hip.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "subpx_hip.h"

namespace spx {

// Thread-local text of the last error; the C-ABI returns a code and keeps the text here.
void set_error(const std::string& msg);
void clear_error();

inline hipStream_t as_stream(spx_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

#define SPX_REQUIRE(cond, msg)              \
    do {                                    \
        if (!(cond)) {                      \
            ::spx::set_error(msg);          \
            return SPX_E_ARG;               \
        }                                   \
    } while (0)

#define SPX_HIP(call)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) {                                                            \
            ::spx::set_error(std::string(#call) + ": " + hipGetErrorString(e_));           \
            return SPX_E_HIP;                                                              \
        }                                                                                  \
    } while (0)

// Checks the launch status of the kernel just enqueued.
#define SPX_LAUNCHED()                                                                     \
    do {                                                                                   \
        hipError_t e_ = hipGetLastError();                                                 \
        if (e_ != hipSuccess) {                                                            \
            ::spx::set_error(std::string("kernel launch: ") + hipGetErrorString(e_));      \
            return SPX_E_HIP;                                                              \
        }                                                                                  \
    } while (0)

constexpr int kWave = 64;

// Scratch buffer owned by the library for one (device, stream) pair, grown on demand and freed
// by spx_release. Work on other streams never sees it, so calls on different streams do not
// overwrite each other's scratch. A call that grows it waits for that stream's queued work
// first; the memory stays valid for work ordered on the stream until the next call on it.
int workspace(size_t bytes, void** out, hipStream_t stream);

// The fit's own counters (kFitCounters x u64: the work counters, then the features per bucket
// at 8 + b) for one (device, stream) pair: a small buffer that no other entry point writes,
// kept until spx_release.
constexpr int kFitCounters = 32;
int fit_stats_buffer(unsigned long long** out, hipStream_t stream);

}  // namespace spx
