// Batched erf-disk / erf-ellipse Levenberg-Marquardt fits: HIP kernels and launch.
// The per-wave fit logic is in erf_fit_device.hpp; this file supplies the CDNA4 wave
// primitives it is written against (DPP/bpermute shuffles, hardware transcendentals).
#include <map>
#include <mutex>
#include <tuple>

#include "spx_internal.hpp"

#define SPX_DEV_INLINE __device__ __forceinline__
#define SPX_DEV __device__

namespace spx {
namespace fitdev {

SPX_DEV_INLINE int spx_lane() { return threadIdx.x & 63; }

// readfirstlane: the compiler then knows the value (and every branch on it) is uniform.
SPX_DEV_INLINE float spx_uniform(float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
SPX_DEV_INLINE int spx_uniform_i(int v) { return __builtin_amdgcn_readfirstlane(v); }
SPX_DEV_INLINE double spx_uniform_d(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffll));
    const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Wave sum by DPP: butterfly within quads and rows (quad_perm, row_half_mirror,
// row_mirror), then row_bcast:15 / row_bcast:31 carry the row totals up to lane 63,
// whose value is read into a scalar register. No LDS traffic.
template <int CTRL, int ROW_MASK>
SPX_DEV_INLINE float dpp_f(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xF, false));
}
SPX_DEV_INLINE float spx_wave_sum(float v) {
    v += dpp_f<0xB1, 0xF>(v);   // quad_perm [1,0,3,2]
    v += dpp_f<0x4E, 0xF>(v);   // quad_perm [2,3,0,1]
    v += dpp_f<0x141, 0xF>(v);  // row_half_mirror
    v += dpp_f<0x140, 0xF>(v);  // row_mirror
    v += dpp_f<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
    v += dpp_f<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}
// Wave max: the same pattern; lanes outside the row mask see the identity (-inf) through
// the max with their own value, so only the masked rows change.
SPX_DEV_INLINE float spx_wave_max(float v) {
    v = fmaxf(v, dpp_f<0xB1, 0xF>(v));
    v = fmaxf(v, dpp_f<0x4E, 0xF>(v));
    v = fmaxf(v, dpp_f<0x141, 0xF>(v));
    v = fmaxf(v, dpp_f<0x140, 0xF>(v));
    const int r16 = __builtin_amdgcn_readlane(__float_as_int(v), 0);
    const int r32 = __builtin_amdgcn_readlane(__float_as_int(v), 16);
    const int r48 = __builtin_amdgcn_readlane(__float_as_int(v), 32);
    const int r64 = __builtin_amdgcn_readlane(__float_as_int(v), 48);
    return fmaxf(fmaxf(__int_as_float(r16), __int_as_float(r32)), fmaxf(__int_as_float(r48), __int_as_float(r64)));
}

// Butterfly steps of the many-value wave sum (wave_sum_many in erf_fit_device.hpp). Each
// pair step adds two per-lane values over lanes that differ in one index bit and keeps the
// first value's partial sum in lanes with that bit clear, the second's in lanes with it set.
// Bit 5 and bit 4: v_permlane32_swap / v_permlane16_swap exchange half-waves / rows of two
// registers, so the pair costs one swap and one add. Bits 3..0: row_mirror,
// row_half_mirror and quad_perm DPP partners (two selects and one DPP add).
SPX_DEV_INLINE float spx_pair_sum32(float a, float b) {
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
SPX_DEV_INLINE float spx_pair_sum16(float a, float b) {
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
template <int CTRL, int BIT>
SPX_DEV_INLINE float spx_pair_dpp(float a, float b) {
    const bool hi = (spx_lane() >> BIT) & 1;
    const float send = hi ? a : b;
    const float keep = hi ? b : a;
    return keep + dpp_f<CTRL, 0xF>(send);
}
template <int CTRL>
SPX_DEV_INLINE float spx_self_dpp(float a) {
    return a + dpp_f<CTRL, 0xF>(a);
}
SPX_DEV_INLINE float spx_readlane(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// Every lane takes part in the atomic (lane 0 adds 1, the others 0; the atomic optimizer
// merges them into one), and the first lane's return is the wave's slot.
SPX_DEV_INLINE int spx_queue_grab(int32_t* q, int n) {
    const int r = atomicAdd(q, spx_lane() == 0 ? n : 0);
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

// LDS patches take NPL x 64 words per wave from dynamic LDS sized at launch for the
// workgroup's waves, so small workgroups do not reserve LDS for absent waves.
template <int NPL, int P>
__global__ __launch_bounds__(256) void erf_fit_kernel(FitArgs a) {
    constexpr bool kLds = Patch<NPL>::kLds;
    extern __shared__ uint32_t s_words[];
    erf_fit_wave<NPL, P>(a, kLds ? s_words + (threadIdx.x >> 6) * NPL * spx::fitdev::kWave : nullptr);
}

// Bucket id: patch-size class x model (disk 6 / ellipse 7 parameters).
constexpr int kKernBuckets = kBuckets * 2;

__global__ void zero_i32(int32_t* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Assigns every feature to its patch-size bucket. Features with an invalid half side or
// model get status SPX_FIT_TOO_FEW and are not fitted. Slots are reserved per workgroup
// (LDS counters, then one global atomic per bucket and workgroup), so the bucket heads
// see K / 256 atomics instead of K.
__global__ __launch_bounds__(256) void bucket_kernel(const int16_t* half, const uint8_t* model, int K,
                                                     int32_t* counts, int32_t* order, float* params,
                                                     double* centers, float* rms, int16_t* iters,
                                                     uint8_t* status, int32_t mask) {
    __shared__ int32_t s_cnt[kKernBuckets];
    __shared__ int32_t s_base[kKernBuckets];
    if (threadIdx.x < kKernBuckets) s_cnt[threadIdx.x] = 0;
    __syncthreads();
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    int b = -1;
    int local = 0;
    if (f < K) {
        const int hs = half[f];
        const int md = model[f];
        const bool known = !(hs < 2 || hs > kMaxHalf || (md != SPX_MODEL_ERF_DISK6 && md != SPX_MODEL_ERF_ELLIPSE7));
        const int bk = known ? 2 * bucket_of(hs) + (md == SPX_MODEL_ERF_ELLIPSE7 ? 1 : 0) : 0;
        if (!known || (mask != 0 && !((mask >> bk) & 1))) {
            for (int i = 0; i < 7; ++i) params[(int64_t)f * 7 + i] = __builtin_nanf("");
            if (centers != nullptr) centers[2 * (int64_t)f] = centers[2 * (int64_t)f + 1] = __builtin_nan("");
            rms[f] = __builtin_nanf("");
            iters[f] = 0;
            status[f] = SPX_FIT_TOO_FEW;
        } else {
            b = bk;
            local = atomicAdd(&s_cnt[b], 1);
        }
    }
    __syncthreads();
    if (threadIdx.x < kKernBuckets) {
        const int c = s_cnt[threadIdx.x];
        s_base[threadIdx.x] = c ? atomicAdd(&counts[threadIdx.x], c) : 0;
    }
    __syncthreads();
    if (b >= 0) order[(int64_t)b * K + s_base[b] + local] = f;
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

namespace {

using KernelFn = void (*)(FitArgs);

template <int B>
constexpr KernelFn kernel_of() {
    constexpr int npl = 4 << (B / 2);
    constexpr int P = (B & 1) ? 7 : 6;
    return &erf_fit_kernel<npl, P>;
}

constexpr KernelFn kKernels[kKernBuckets] = {kernel_of<0>(), kernel_of<1>(), kernel_of<2>(), kernel_of<3>(),
                                             kernel_of<4>(), kernel_of<5>(), kernel_of<6>(), kernel_of<7>(),
                                             kernel_of<8>(), kernel_of<9>()};

// Resident workgroups per CU of each bucket kernel at a given workgroup size (cached per
// device): the persistent grid is sized to what the chip holds at once.
// Dynamic LDS of bucket b's kernel at `threads` per workgroup (the LDS slab of large patches).
size_t lds_bytes(int b, int threads) {
    const int npl = 4 << (b / 2);
    return npl > fitdev::kUnrollMax ? (size_t)(threads / fitdev::kWave) * npl * fitdev::kWave * sizeof(uint32_t) : 0;
}

int resident_blocks(int dev, int b, int threads) {
    static std::mutex mu;
    static std::map<std::tuple<int, int, int>, int> cache;
    std::lock_guard<std::mutex> lock(mu);
    auto key = std::make_tuple(dev, b, threads);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, reinterpret_cast<const void*>(kKernels[b]),
                                                     threads, lds_bytes(b, threads)) != hipSuccess ||
        blocks < 1) {
        (void)hipGetLastError();
        blocks = 1;
    }
    cache[key] = blocks;
    return blocks;
}

}  // namespace

int fit_batch_launch(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, const float* seeds,
                     const int16_t* half, const uint8_t* model, const float* init,
                     const SpxFitOptions& opt, int32_t K, const int32_t* nbr, int32_t nb, float* params,
                     double* centers, float* rms, int16_t* iters, uint8_t* status, hipStream_t stream) {
    if (K == 0) return SPX_OK;
    // Workspace: stats u64[4] | counts[10] | queues[10] | order[10][K]
    void* ws = nullptr;
    const size_t bytes = 32 + sizeof(int32_t) * (2 * kKernBuckets + (size_t)kKernBuckets * K);
    int rc = workspace(bytes, &ws);
    if (rc != SPX_OK) return rc;
    unsigned long long* stats = static_cast<unsigned long long*>(ws);
    int32_t* counts = reinterpret_cast<int32_t*>(stats + 4);
    int32_t* queues = counts + kKernBuckets;
    int32_t* order = queues + kKernBuckets;
    g_last_stats = stats;
    zero_i32<<<1, 64, 0, stream>>>(reinterpret_cast<int32_t*>(stats), 8 + 2 * kKernBuckets);
    SPX_LAUNCHED();
    bucket_kernel<<<(K + 255) / 256, 256, 0, stream>>>(half, model, K, counts, order, params, centers,
                                                        rms, iters, status, opt.bucket_mask);
    SPX_LAUNCHED();

    int dev = 0;
    SPX_HIP(hipGetDevice(&dev));
    int cus = 0;
    SPX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int waves_per_block = opt.tile_waves > 0 ? (opt.tile_waves < 4 ? opt.tile_waves : 4) : 4;
    const int threads = waves_per_block * fitdev::kWave;
    FitArgs a{img, H, W, pitch, seeds, half, model, init, opt, nullptr, nullptr, nullptr,
              stats, params, centers, rms, iters, status, nbr, nb, K};
    for (int b = 0; b < kKernBuckets; ++b) {
        if (opt.bucket_mask != 0 && !((opt.bucket_mask >> b) & 1)) continue;
        a.order = order + (int64_t)b * K;
        a.count = counts + b;
        a.queue = queues + b;
        // Persistent grid: the resident capacity of the chip, never more than features.
        // Waves of an empty bucket read its count and exit.
        // (persistent = 0: one wave per feature, the queue then hands each wave its chunk)
        const int64_t want = ((int64_t)K + waves_per_block - 1) / waves_per_block;
        const int64_t cap = opt.persistent ? (int64_t)cus * resident_blocks(dev, b, threads) : want;
        const int blocks = (int)(want < cap ? want : cap);
        hipLaunchKernelGGL(kKernels[b], dim3(blocks), dim3(threads), lds_bytes(b, threads), stream, a);
        SPX_LAUNCHED();
    }
    return SPX_OK;
}

}  // namespace spx
