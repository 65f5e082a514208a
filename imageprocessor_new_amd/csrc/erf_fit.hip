// Batched erf-disk / erf-ellipse Levenberg-Marquardt fits: HIP kernels and launch.
// The per-wave fit logic is in erf_fit_device.hpp; this file supplies the CDNA4 wave
// primitives it is written against (DPP/bpermute shuffles, hardware transcendentals).
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

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
SPX_DEV_INLINE int spx_shfl(int v, int src) { return __shfl(v, src, 64); }
SPX_DEV_INLINE int spx_readlane_i(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }
SPX_DEV_INLINE double spx_readlane_d(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
SPX_DEV_INLINE uint64_t spx_ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }
SPX_DEV_INLINE int spx_lanes_below(uint64_t m) {
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
// A wave's LDS operations complete in order; the fences keep the compiler from moving LDS
// accesses across this point, so words another lane wrote before it are read after it.
SPX_DEV_INLINE void spx_wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
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

// Every wave compacts its patches in NPL x 64 words of dynamic LDS sized at launch for the
// workgroup's waves, so small workgroups do not reserve LDS for absent waves.
// WPE: the waves per SIMD the register allocation must leave room for (2, 3 or 4). The LM
// state (FP64 solve, sums) takes most registers whatever the patch size; at two waves per SIMD
// a small patch's fit waits on its dependent FP64 chains, at four the compiler spills. Each
// bucket kernel is built for all three and the launch picks one (SpxFitOptions.bucket_waves).
template <int NPL, int P, int WPE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE))) void erf_fit_kernel(FitArgs a) {
    extern __shared__ uint32_t s_words[];
    erf_fit_wave<NPL, P>(a, s_words + (threadIdx.x >> 6) * NPL * spx::fitdev::kWave);
}

// Bucket id: patch-size class x model (disk 6 / ellipse 7 parameters).
constexpr int kKernBuckets = kBuckets * 2;

__global__ void zero_i32(int32_t* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Bucket of a feature with a valid half side and model from its pixel bound, or -1 when the
// mask leaves that bucket out.
__device__ inline int feature_bucket(int md, int bound, int32_t mask) {
    const int bk = 2 * bucket_of_pixels(bound) + (md == SPX_MODEL_ERF_ELLIPSE7 ? 1 : 0);
    return (mask != 0 && !((mask >> bk) & 1)) ? -1 : bk;
}

__device__ inline bool feature_valid(const int16_t* half, const uint8_t* model, int f) {
    const int hs = half[f];
    const int md = model[f];
    return hs >= 2 && hs <= kMaxHalf && (md == SPX_MODEL_ERF_DISK6 || md == SPX_MODEL_ERF_ELLIPSE7);
}

// Pass 1 of the bucketing, one wave per feature: a bound of the feature's fitted pixels (lane
// r bounds patch row r; kept in its iters entry until the fit writes it; -1 for a feature that
// is not fitted), the per-bucket feature counts (LDS counters, then one global atomic per
// bucket and workgroup), and the outputs of features that are not fitted: invalid half side
// or model, or a bucket the mask leaves out (status SPX_FIT_TOO_FEW and NaNs).
constexpr int kCountWaves = 4;
constexpr int kCountBlocks = 2048;

__global__ __launch_bounds__(64 * kCountWaves) void bucket_count_kernel(FitArgs a, int32_t* counts) {
    // counts[b] steers the launch; stats[8 + b] keeps a copy for the caller
    __shared__ int32_t s_cnt[kKernBuckets];
    if (threadIdx.x < kKernBuckets) s_cnt[threadIdx.x] = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63;
    // a grid of at most kCountBlocks workgroups strides over the features, so the per-bucket
    // totals take few global atomics on the same ten addresses
    for (int f = blockIdx.x * kCountWaves + (threadIdx.x >> 6); f < a.nfeat; f += gridDim.x * kCountWaves) {
        int b = -1;
        if (feature_valid(a.half, a.model, f)) {
            const PatchSpan p = patch_span(a, f);
            int cnt = p.side * p.side;
            if (!p.whole && !p.cells) {
                cnt = p.c1 >= p.c0 && p.r1 >= p.r0 ? (int)((p.c1 - p.c0 + 1) * (p.r1 - p.r0 + 1)) : 0;
            } else if (!p.whole) {
                cnt = patch_row_upper(lane_terms(a, f), p, lane);
                if (lane >= p.side) cnt = 0;
                for (int off = 32; off >= 1; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
            }
            b = feature_bucket(a.model[f], cnt, a.opt.bucket_mask);
            if (lane == 0) a.iters[f] = (int16_t)(b < 0 ? -1 : cnt);
        }
        if (lane == 0) {
            if (b < 0) {
                for (int i = 0; i < 7; ++i) a.params[(int64_t)f * 7 + i] = __builtin_nanf("");
                if (a.centers != nullptr) a.centers[2 * (int64_t)f] = a.centers[2 * (int64_t)f + 1] = __builtin_nan("");
                a.rms[f] = __builtin_nanf("");
                a.iters[f] = -1;
                a.status[f] = SPX_FIT_TOO_FEW;
            } else {
                atomicAdd(&s_cnt[b], 1);
            }
        }
    }
    __syncthreads();
    if (threadIdx.x < kKernBuckets && s_cnt[threadIdx.x] != 0) {
        atomicAdd(&counts[threadIdx.x], s_cnt[threadIdx.x]);
        atomicAdd(&a.stats[8 + threadIdx.x], (unsigned long long)s_cnt[threadIdx.x]);
    }
}

// Pass 2: every bucket's ids back to back in one K-entry array, bucket b after the ids of
// buckets 0 .. b-1. Each workgroup reserves its slots with one atomic per bucket. The iters
// entries of features that are not fitted go back to 0.
__global__ __launch_bounds__(256) void bucket_scatter_kernel(FitArgs a, const int32_t* counts, int32_t* cursors,
                                                             int32_t* order) {
    __shared__ int32_t s_cnt[kKernBuckets];
    __shared__ int32_t s_base[kKernBuckets];
    if (threadIdx.x < kKernBuckets) s_cnt[threadIdx.x] = 0;
    __syncthreads();
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    int b = -1, local = 0;
    if (f < a.nfeat) {
        const int bound = a.iters[f];
        if (bound < 0) {
            a.iters[f] = 0;
        } else {
            b = feature_bucket(a.model[f], bound, a.opt.bucket_mask);
            local = atomicAdd(&s_cnt[b], 1);
        }
    }
    __syncthreads();
    if (threadIdx.x < kKernBuckets) {
        int start = 0;
        for (int j = 0; j < (int)threadIdx.x; ++j) start += counts[j];
        const int c = s_cnt[threadIdx.x];
        s_base[threadIdx.x] = start + (c ? atomicAdd(&cursors[threadIdx.x], c) : 0);
    }
    __syncthreads();
    if (b >= 0) order[s_base[b] + local] = f;
}

// Counters of the last fit issued on this thread: where they live and the stream that
// writes them.
thread_local unsigned long long* g_last_stats = nullptr;
thread_local hipStream_t g_last_stats_stream = nullptr;

}  // namespace

int fit_stats_copy(uint64_t* out, int n, hipStream_t stream) {
    for (int i = 0; i < n; ++i) out[i] = 0;
    if (g_last_stats == nullptr) return SPX_OK;
    SPX_HIP(hipStreamSynchronize(g_last_stats_stream));
    SPX_HIP(hipMemcpyAsync(out, g_last_stats, (size_t)n * sizeof(uint64_t), hipMemcpyDeviceToHost, stream));
    SPX_HIP(hipStreamSynchronize(stream));
    return SPX_OK;
}

namespace {

using KernelFn = void (*)(FitArgs);

constexpr int kWaveVariants = 3;  // 2, 3, 4 waves per SIMD

template <int B, int V>
constexpr KernelFn kernel_of() {
    return &erf_fit_kernel<4 << (B / 2), (B & 1) ? 7 : 6, 2 + V>;
}

template <int B>
struct BucketKernels {
    static constexpr KernelFn fn[kWaveVariants] = {kernel_of<B, 0>(), kernel_of<B, 1>(), kernel_of<B, 2>()};
};

constexpr const KernelFn* kKernels[kKernBuckets] = {
    BucketKernels<0>::fn, BucketKernels<1>::fn, BucketKernels<2>::fn, BucketKernels<3>::fn, BucketKernels<4>::fn,
    BucketKernels<5>::fn, BucketKernels<6>::fn, BucketKernels<7>::fn, BucketKernels<8>::fn, BucketKernels<9>::fn};

// Default variant per bucket (index into 2, 3, 4 waves per SIMD) when the options leave it to
// the library: three waves for the LDS-resident disk buckets (a few dwords spilled outside the
// pixel loop cost less than the third wave gains), two elsewhere.
constexpr int kDefaultVariant[kKernBuckets] = {0, 0, 0, 0, 0, 0, 1, 0, 1, 0};

int wave_variant(const SpxFitOptions& opt, int b) {
    const int code = (opt.bucket_waves >> (2 * b)) & 3;
    return code == 0 ? kDefaultVariant[b] : code - 1;
}

// Resident workgroups per CU of each bucket kernel at a given workgroup size (cached per
// device): the persistent grid is sized to what the chip holds at once.
// Dynamic LDS of bucket b's kernel at `threads` per workgroup (the LDS slab of large patches).
size_t lds_bytes(int b, int threads) {
    const int npl = 4 << (b / 2);
    return (size_t)(threads / fitdev::kWave) * npl * fitdev::kWave * sizeof(uint32_t);
}

int resident_blocks(int dev, int b, int v, int threads) {
    static std::mutex mu;
    static std::map<std::tuple<int, int, int, int>, int> cache;
    std::lock_guard<std::mutex> lock(mu);
    auto key = std::make_tuple(dev, b, v, threads);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, reinterpret_cast<const void*>(kKernels[b][v]),
                                                     threads, lds_bytes(b, threads)) != hipSuccess ||
        blocks < 1) {
        (void)hipGetLastError();
        blocks = 1;
    }
    cache[key] = blocks;
    return blocks;
}

}  // namespace

// Layout of a fit's scratch: counters u64[kFitCounters] | counts[10] | cursors[10] |
// queues[10] | ids.
constexpr size_t kFitHead = 512;
constexpr int kFitHeadWords = (kFitCounters * 8 + 3 * kKernBuckets * 4) / 4;
static_assert(kFitHeadWords * 4 <= (int)kFitHead, "fit scratch head");

size_t fit_scratch_bytes(int32_t K) { return kFitHead + sizeof(int32_t) * (size_t)K; }

// The bucket kernels run on kForkStreams streams at once: the caller's and helpers that
// fork from it after the bucketing and join it before the call returns. A persistent kernel
// holds every CU it fits on until its queue drains, and the buckets of few or slowly
// converging features (ellipses, small patches) spend most of their time waiting on one
// feature's chain of dependent solves; side by side, those waits overlap the large buckets'
// work and tails. Each stream takes a group of buckets in turn.
constexpr int kForkStreams = 4;
constexpr int kForkOrder[kKernBuckets] = {6, 8, 4, 2, 7, 5, 9, 3, 1, 0};   // launch order
constexpr int kForkStream[kKernBuckets] = {3, 3, 1, 3, 1, 2, 0, 2, 0, 2};  // stream of bucket b

struct ForkSet {
    hipStream_t helper[kForkStreams] = {};  // [0] unused: the caller's stream
    hipEvent_t fork = nullptr;
    hipEvent_t join[kForkStreams] = {};
};

namespace {
std::mutex g_fork_mu;
std::map<std::pair<int, hipStream_t>, ForkSet> g_forks;

// The helper streams and events of (device, caller stream), created at first use (default
// priority, non-blocking) and kept until spx_release.
int fork_set(int dev, hipStream_t stream, ForkSet** out) {
    std::lock_guard<std::mutex> lock(g_fork_mu);
    ForkSet& f = g_forks[std::make_pair(dev, stream)];
    if (f.fork == nullptr) {
        SPX_HIP(hipEventCreateWithFlags(&f.fork, hipEventDisableTiming));
        for (int i = 1; i < kForkStreams; ++i) {
            SPX_HIP(hipStreamCreateWithFlags(&f.helper[i], hipStreamNonBlocking));
            SPX_HIP(hipEventCreateWithFlags(&f.join[i], hipEventDisableTiming));
        }
    }
    *out = &f;
    return SPX_OK;
}
}  // namespace

void fit_streams_release() {
    std::lock_guard<std::mutex> lock(g_fork_mu);
    for (auto& kv : g_forks) {
        if (hipSetDevice(kv.first.first) != hipSuccess) continue;
        for (int i = 1; i < kForkStreams; ++i) {
            if (kv.second.helper[i] != nullptr) {
                (void)hipStreamSynchronize(kv.second.helper[i]);
                (void)hipStreamDestroy(kv.second.helper[i]);
            }
            if (kv.second.join[i] != nullptr) (void)hipEventDestroy(kv.second.join[i]);
        }
        if (kv.second.fork != nullptr) (void)hipEventDestroy(kv.second.fork);
    }
    g_forks.clear();
}

// Enqueues the fit on `scratch` (fit_scratch_bytes(K) bytes of device memory) with its
// counters at `stats` (kFitCounters x u64, may lie inside scratch). With `forks`, the bucket
// kernels run on the caller's stream and the helpers (see kForkStreams); without, all in
// turn on the caller's stream. Results are the same either way.
int fit_batch_enqueue(void* scratch, unsigned long long* stats, const uint16_t* img, int64_t H, int64_t W,
                      int64_t pitch, const float* seeds, const int16_t* half, const uint8_t* model, const float* init,
                      const SpxFitOptions& opt, int32_t K, const int32_t* nbr, int32_t nb, float* params,
                      double* centers, float* rms, int16_t* iters, uint8_t* status, hipStream_t stream,
                      ForkSet* forks) {
    int32_t* head = static_cast<int32_t*>(scratch);
    int32_t* counts = head + 2 * kFitCounters;
    int32_t* cursors = counts + kKernBuckets;
    int32_t* queues = cursors + kKernBuckets;
    int32_t* order = reinterpret_cast<int32_t*>(static_cast<char*>(scratch) + kFitHead);
    zero_i32<<<(kFitHeadWords + 63) / 64, 64, 0, stream>>>(head, kFitHeadWords);
    SPX_LAUNCHED();
    zero_i32<<<(2 * kFitCounters + 63) / 64, 64, 0, stream>>>(reinterpret_cast<int32_t*>(stats), 2 * kFitCounters);
    SPX_LAUNCHED();
    g_last_stats = stats;
    g_last_stats_stream = stream;
    FitArgs a{img, H, W, pitch, seeds, half, model, init, opt, order, nullptr, nullptr,
              stats, params, centers, rms, iters, status, nbr, nb, K, counts, 0};
    const int blocks256 = (K + 255) / 256;
    const int count_blocks = (K + kCountWaves - 1) / kCountWaves < kCountBlocks ? (K + kCountWaves - 1) / kCountWaves
                                                                               : kCountBlocks;
    bucket_count_kernel<<<count_blocks, 64 * kCountWaves, 0, stream>>>(a, counts);
    SPX_LAUNCHED();
    bucket_scatter_kernel<<<blocks256, 256, 0, stream>>>(a, counts, cursors, order);
    SPX_LAUNCHED();

    int dev = 0;
    SPX_HIP(hipGetDevice(&dev));
    int cus = 0;
    SPX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int waves_per_block = opt.tile_waves > 0 ? (opt.tile_waves < 4 ? opt.tile_waves : 4) : 4;
    const int threads = waves_per_block * fitdev::kWave;
    hipStream_t streams[kForkStreams];
    for (int i = 0; i < kForkStreams; ++i) streams[i] = (forks != nullptr && i > 0) ? forks->helper[i] : stream;
    if (forks != nullptr) {
        SPX_HIP(hipEventRecord(forks->fork, stream));
        for (int i = 1; i < kForkStreams; ++i) SPX_HIP(hipStreamWaitEvent(forks->helper[i], forks->fork, 0));
    }
    for (int bi = 0; bi < kKernBuckets; ++bi) {
        const int b = kForkOrder[bi];
        if (opt.bucket_mask != 0 && !((opt.bucket_mask >> b) & 1)) continue;
        a.bucket = b;
        a.count = counts + b;
        a.queue = queues + b;
        // Persistent grid: the resident capacity of the chip, never more than features.
        // Waves of an empty bucket read its count and exit.
        // (persistent = 0: one wave per feature, the queue then hands each wave its chunk)
        const int v = wave_variant(opt, b);
        const int64_t want = ((int64_t)K + waves_per_block - 1) / waves_per_block;
        const int64_t cap = opt.persistent ? (int64_t)cus * resident_blocks(dev, b, v, threads) : want;
        const int blocks = (int)(want < cap ? want : cap);
        hipLaunchKernelGGL(kKernels[b][v], dim3(blocks), dim3(threads), lds_bytes(b, threads),
                           streams[kForkStream[b]], a);
        SPX_LAUNCHED();
    }
    if (forks != nullptr)
        for (int i = 1; i < kForkStreams; ++i) {
            SPX_HIP(hipEventRecord(forks->join[i], forks->helper[i]));
            SPX_HIP(hipStreamWaitEvent(stream, forks->join[i], 0));
        }
    return SPX_OK;
}

int fit_batch_launch(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, const float* seeds,
                     const int16_t* half, const uint8_t* model, const float* init,
                     const SpxFitOptions& opt, int32_t K, const int32_t* nbr, int32_t nb, float* params,
                     double* centers, float* rms, int16_t* iters, uint8_t* status, hipStream_t stream) {
    if (K == 0) return SPX_OK;
    void* ws = nullptr;
    int rc = workspace(fit_scratch_bytes(K), &ws, stream);
    if (rc != SPX_OK) return rc;
    unsigned long long* stats = nullptr;
    rc = fit_stats_buffer(&stats, stream);
    if (rc != SPX_OK) return rc;
    int dev = 0;
    SPX_HIP(hipGetDevice(&dev));
    ForkSet* forks = nullptr;
    rc = fork_set(dev, stream, &forks);
    if (rc != SPX_OK) return rc;
    return fit_batch_enqueue(ws, stats, img, H, W, pitch, seeds, half, model, init, opt, K, nbr, nb, params, centers,
                             rms, iters, status, stream, forks);
}

// ---- the graph-recorded fit --------------------------------------------------------------
// One batch recorded on a capture stream of the graph's own (the null stream cannot be
// captured), with helper streams of its own joined into the capture by the fork event. The
// recorded work uses no memset nodes: the scratch head and counters are cleared by zero_i32.

}  // namespace spx

struct SpxFitGraph {
    int dev = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void* scratch = nullptr;
    unsigned long long* stats = nullptr;
    hipStream_t capture = nullptr;
    spx::ForkSet forks;
};

namespace spx {

namespace {
void fit_graph_free(SpxFitGraph* g) {
    if (g == nullptr) return;
    (void)hipSetDevice(g->dev);
    if (g->exec != nullptr) (void)hipGraphExecDestroy(g->exec);
    if (g->graph != nullptr) (void)hipGraphDestroy(g->graph);
    for (int i = 1; i < kForkStreams; ++i) {
        if (g->forks.helper[i] != nullptr) {
            (void)hipStreamSynchronize(g->forks.helper[i]);
            (void)hipStreamDestroy(g->forks.helper[i]);
        }
        if (g->forks.join[i] != nullptr) (void)hipEventDestroy(g->forks.join[i]);
    }
    if (g->forks.fork != nullptr) (void)hipEventDestroy(g->forks.fork);
    if (g->capture != nullptr) {
        (void)hipStreamSynchronize(g->capture);
        (void)hipStreamDestroy(g->capture);
    }
    if (g->scratch != nullptr) (void)hipFree(g->scratch);
    if (g->stats != nullptr) (void)hipFree(g->stats);
    if (g_last_stats == g->stats) {
        g_last_stats = nullptr;
        g_last_stats_stream = nullptr;
    }
    delete g;
}

int fit_graph_record(SpxFitGraph* g, const uint16_t* img, int64_t H, int64_t W, int64_t pitch, const float* seeds,
                     const int16_t* half, const uint8_t* model, const float* init, const SpxFitOptions& opt,
                     int32_t K, const int32_t* nbr, int32_t nb, float* params, double* centers, float* rms,
                     int16_t* iters, uint8_t* status) {
    SPX_HIP(hipGetDevice(&g->dev));
    // the launch shapes are looked up before the capture (host calls only inside it)
    const int threads = (opt.tile_waves > 0 ? (opt.tile_waves < 4 ? opt.tile_waves : 4) : 4) * fitdev::kWave;
    for (int b = 0; b < kKernBuckets; ++b) (void)resident_blocks(g->dev, b, wave_variant(opt, b), threads);
    SPX_HIP(hipMalloc(&g->scratch, fit_scratch_bytes(K)));
    SPX_HIP(hipMalloc(reinterpret_cast<void**>(&g->stats), kFitCounters * sizeof(unsigned long long)));
    SPX_HIP(hipStreamCreateWithFlags(&g->capture, hipStreamNonBlocking));
    SPX_HIP(hipEventCreateWithFlags(&g->forks.fork, hipEventDisableTiming));
    for (int i = 1; i < kForkStreams; ++i) {
        SPX_HIP(hipStreamCreateWithFlags(&g->forks.helper[i], hipStreamNonBlocking));
        SPX_HIP(hipEventCreateWithFlags(&g->forks.join[i], hipEventDisableTiming));
    }
    unsigned long long* prev_stats = g_last_stats;
    hipStream_t prev_stream = g_last_stats_stream;
    SPX_HIP(hipStreamBeginCapture(g->capture, hipStreamCaptureModeThreadLocal));
    const int rc = fit_batch_enqueue(g->scratch, g->stats, img, H, W, pitch, seeds, half, model, init, opt, K, nbr,
                                     nb, params, centers, rms, iters, status, g->capture, &g->forks);
    // the capture ends whatever the recording returned
    const hipError_t end = hipStreamEndCapture(g->capture, &g->graph);
    g_last_stats = prev_stats;
    g_last_stats_stream = prev_stream;
    if (rc != SPX_OK) return rc;
    SPX_HIP(end);
    SPX_HIP(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
    return SPX_OK;
}
}  // namespace

int fit_graph_create(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, const float* seeds,
                     const int16_t* half, const uint8_t* model, const float* init, const SpxFitOptions& opt,
                     int32_t K, const int32_t* nbr, int32_t nb, float* params, double* centers, float* rms,
                     int16_t* iters, uint8_t* status, SpxFitGraph** out) {
    *out = nullptr;
    SpxFitGraph* g = new SpxFitGraph();
    const int rc = fit_graph_record(g, img, H, W, pitch, seeds, half, model, init, opt, K, nbr, nb, params, centers,
                                    rms, iters, status);
    if (rc != SPX_OK) {
        fit_graph_free(g);
        return rc;
    }
    *out = g;
    return SPX_OK;
}

int fit_graph_launch(SpxFitGraph* g, hipStream_t stream) {
    int dev = 0;
    SPX_HIP(hipGetDevice(&dev));
    if (dev != g->dev) {
        set_error("spx_fit_graph_launch: the graph was recorded on device " + std::to_string(g->dev) +
                  ", the current device is " + std::to_string(dev));
        return SPX_E_ARG;
    }
    SPX_HIP(hipGraphLaunch(g->exec, stream));
    g_last_stats = g->stats;
    g_last_stats_stream = stream;
    return SPX_OK;
}

int fit_graph_destroy(SpxFitGraph* g) {
    fit_graph_free(g);
    return SPX_OK;
}

// ---- the allocation-free fit -------------------------------------------------------------
// Scratch of spx_fit_batch_noalloc: a file-scope device array, one copy per GPU at its own
// address. Host code never passes the array's name to a kernel (on the host it names a
// stand-in, not device memory): it passes the address hipGetSymbolAddress returns for the
// current device, looked up once per device at spx_init or at the first call.
static __device__ __align__(256) unsigned char scratch[kFitHead + sizeof(int32_t) * SPX_FIT_NOALLOC_MAX_FEATURES];

namespace {
std::mutex g_scratch_mu;
std::vector<void*> g_scratch_ptr;  // by device ordinal
}  // namespace

int fit_scratch_init(int dev) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
    if (g_scratch_ptr[dev] != nullptr) return SPX_OK;
    void* p = nullptr;
    SPX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (p == nullptr) {
        set_error("fit scratch: hipGetSymbolAddress returned no address");
        return SPX_E_HIP;
    }
    g_scratch_ptr[dev] = p;
    // The launch shapes of every bucket kernel are looked up now, so that calls of the
    // allocation-free fit add no host-side cache entries either.
    int cus = 0;
    SPX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    for (int b = 0; b < kKernBuckets; ++b)
        for (int v = 0; v < kWaveVariants; ++v)
            for (int w = 1; w <= 4; ++w) (void)resident_blocks(dev, b, v, w * fitdev::kWave);
    return SPX_OK;
}

int fit_batch_noalloc_launch(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, const float* seeds,
                             const int16_t* half, const uint8_t* model, const float* init, const SpxFitOptions& opt,
                             int32_t K, const int32_t* nbr, int32_t nb, float* params, double* centers, float* rms,
                             int16_t* iters, uint8_t* status, hipStream_t stream) {
    if (K > SPX_FIT_NOALLOC_MAX_FEATURES) {
        set_error("spx_fit_batch_noalloc: " + std::to_string(K) + " features exceed the capacity of its fixed scratch ("
                  + std::to_string(SPX_FIT_NOALLOC_MAX_FEATURES) + ")");
        return SPX_E_ARG;
    }
    if (K == 0) return SPX_OK;
    int dev = 0;
    SPX_HIP(hipGetDevice(&dev));
    int rc = fit_scratch_init(dev);
    if (rc != SPX_OK) return rc;
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        base = g_scratch_ptr[dev];
    }
    // the counters sit at the head's start; counts, cursors and queues follow them
    unsigned long long* stats = static_cast<unsigned long long*>(base);
    // no helper streams here: creating them is work this entry point does not do
    return fit_batch_enqueue(base, stats, img, H, W, pitch, seeds, half, model, init, opt, K, nbr, nb, params,
                             centers, rms, iters, status, stream, nullptr);
}

}  // namespace spx
