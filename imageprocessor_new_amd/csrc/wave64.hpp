// Wave64 reductions shared by the per-feature kernels (one wave per feature).
//
// Each reduction returns the first lane's total through readfirstlane, so every branch
// taken on a reduced value is a uniform (scalar) branch of the whole wave.
#pragma once

#include <hip/hip_runtime.h>

namespace spx {
namespace wave {

__device__ __forceinline__ double uniform_d(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffll));
    const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ double sum_d(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return uniform_d(v);
}

__device__ __forceinline__ int sum_i(int v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return __builtin_amdgcn_readfirstlane(v);
}

__device__ __forceinline__ double max_d(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return uniform_d(v);
}

}  // namespace wave
}  // namespace spx
