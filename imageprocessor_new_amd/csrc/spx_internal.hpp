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

// Per-device scratch buffer owned by the library, grown on demand and freed by spx_release.
// Returned memory is only valid for work ordered on the same stream before the next call
// that grows it; callers synchronise the stream before growing.
int workspace(size_t bytes, void** out);

}  // namespace spx
