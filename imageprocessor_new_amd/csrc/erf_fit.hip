// Batched erf-disk / erf-ellipse Levenberg-Marquardt fits: HIP kernels and launch.
// The per-wave fit logic is in erf_fit_device.hpp; this file supplies the CDNA4 wave
// primitives it is written against (DPP/bpermute shuffles, hardware transcendentals).
#include "spx_internal.hpp"

#define SPX_DEV_INLINE __device__ __forceinline__
#define SPX_DEV __device__

namespace spx {
namespace fitdev {

SPX_DEV_INLINE int spx_lane() { return threadIdx.x & 63; }
SPX_DEV_INLINE float spx_shfl_xor(float v, int off) { return __shfl_xor(v, off, 64); }
// readfirstlane: the compiler then knows the value (and every branch on it) is uniform.
SPX_DEV_INLINE float spx_uniform(float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
SPX_DEV_INLINE int spx_uniform_i(int v) { return __builtin_amdgcn_readfirstlane(v); }
// Every lane takes part in the atomic (lane 0 adds 1, the others 0; the atomic optimizer
// merges them into one), and the first lane's return is the wave's slot.
SPX_DEV_INLINE int spx_queue_grab(int32_t* q) {
    const int r = atomicAdd(q, spx_lane() == 0 ? 1 : 0);
    return __builtin_amdgcn_readfirstlane(r);
}
SPX_DEV_INLINE unsigned long long spx_atomic_add(unsigned long long* p, unsigned long long v) {
    return atomicAdd(p, v);
}
SPX_DEV_INLINE float hw_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
SPX_DEV_INLINE float hw_rsq(float x) { return __builtin_amdgcn_rsqf(x); }
SPX_DEV_INLINE float hw_exp2(float x) { return __builtin_amdgcn_exp2f(x); }
// v_rcp_f64 refined by two Newton steps to full double precision.
SPX_DEV_INLINE double hw_rcp64(double d) {
    double r = __builtin_amdgcn_rcp(d);
    r = r * fma(-d, r, 2.0);
    r = r * fma(-d, r, 2.0);
    return r;
}

}  // namespace fitdev
}  // namespace spx

#include "erf_fit_device.hpp"

namespace spx {
namespace {

using namespace fitdev;

template <int NPL>
__global__ __launch_bounds__(256) void erf_fit_kernel(FitArgs a) {
    erf_fit_wave<NPL>(a);
}

__global__ void zero_i32(int32_t* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Assigns every feature to its patch-size bucket. Features with an invalid half side or
// model get status SPX_FIT_TOO_FEW and are not fitted.
__global__ void bucket_kernel(const int16_t* half, const uint8_t* model, int K, int32_t* counts,
                              int32_t* order, float* params, float* rms, int16_t* iters,
                              uint8_t* status) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= K) return;
    const int hs = half[f];
    const int md = model[f];
    if (hs < 2 || hs > kMaxHalf || (md != SPX_MODEL_ERF_DISK6 && md != SPX_MODEL_ERF_ELLIPSE7)) {
        for (int i = 0; i < 7; ++i) params[(int64_t)f * 7 + i] = nanf("");
        rms[f] = nanf("");
        iters[f] = 0;
        status[f] = SPX_FIT_TOO_FEW;
        return;
    }
    const int b = bucket_of(hs);
    const int pos = atomicAdd(&counts[b], 1);
    order[(int64_t)b * K + pos] = f;
}

thread_local unsigned long long* g_last_stats = nullptr;

}  // namespace

int fit_stats_copy(uint64_t* out, hipStream_t stream) {
    if (g_last_stats == nullptr) {
        for (int i = 0; i < 4; ++i) out[i] = 0;
        return SPX_OK;
    }
    SPX_HIP(hipMemcpyAsync(out, g_last_stats, 4 * sizeof(uint64_t), hipMemcpyDeviceToHost, stream));
    SPX_HIP(hipStreamSynchronize(stream));
    return SPX_OK;
}

int fit_batch_launch(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, const float* seeds,
                     const int16_t* half, const uint8_t* model, const float* init,
                     const SpxFitOptions& opt, int32_t K, float* params, float* rms, int16_t* iters,
                     uint8_t* status, hipStream_t stream) {
    if (K == 0) return SPX_OK;
    // Workspace: stats u64[4] | counts[4] | queues[4] | order[4][K]
    void* ws = nullptr;
    const size_t bytes = 32 + sizeof(int32_t) * (2 * kBuckets + (size_t)kBuckets * K);
    int rc = workspace(bytes, &ws);
    if (rc != SPX_OK) return rc;
    unsigned long long* stats = static_cast<unsigned long long*>(ws);
    int32_t* counts = reinterpret_cast<int32_t*>(stats + 4);
    int32_t* queues = counts + kBuckets;
    int32_t* order = queues + kBuckets;
    g_last_stats = stats;
    zero_i32<<<1, 64, 0, stream>>>(reinterpret_cast<int32_t*>(stats), 8 + 2 * kBuckets);
    SPX_LAUNCHED();
    bucket_kernel<<<(K + 255) / 256, 256, 0, stream>>>(half, model, K, counts, order, params, rms,
                                                        iters, status);
    SPX_LAUNCHED();

    int dev = 0;
    SPX_HIP(hipGetDevice(&dev));
    int cus = 0;
    SPX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int waves_per_block = opt.tile_waves > 0 ? (opt.tile_waves < 4 ? opt.tile_waves : 4) : 4;
    const int threads = waves_per_block * fitdev::kWave;
    FitArgs a{img, H, W, pitch, seeds, half, model, init, opt, nullptr, nullptr, nullptr,
              stats, params, rms, iters, status};
    for (int b = 0; b < kBuckets; ++b) {
        a.order = order + (int64_t)b * K;
        a.count = counts + b;
        a.queue = queues + b;
        // Persistent grid: enough resident waves to fill the chip, never more than features.
        const int64_t want = ((int64_t)K + waves_per_block - 1) / waves_per_block;
        const int64_t cap = (int64_t)cus * (b >= 3 ? 4 : 8);
        const int blocks = (int)(want < cap ? want : cap);
        switch (b) {
            case 0: erf_fit_kernel<4><<<blocks, threads, 0, stream>>>(a); break;
            case 1: erf_fit_kernel<8><<<blocks, threads, 0, stream>>>(a); break;
            case 2: erf_fit_kernel<16><<<blocks, threads, 0, stream>>>(a); break;
            default: erf_fit_kernel<32><<<blocks, threads, 0, stream>>>(a); break;
        }
        SPX_LAUNCHED();
    }
    return SPX_OK;
}

}  // namespace spx
