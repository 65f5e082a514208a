// Host emulation of one wave64 running the fit kernel's device logic
// (imageprocessor_new_amd/csrc/erf_fit_device.hpp), for CPU tests only.
//
// Each lane is a std::thread; every cross-lane exchange (shuffle) passes two barriers, so
// lanes that take different control paths through a shuffle deadlock here. A watchdog
// turns that into exit code 3 with the shuffle count of every lane.
#include <atomic>
#include <barrier>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <algorithm>
#include <vector>

#include "subpx_hip.h"

#define SPX_DEV_INLINE inline
#define SPX_DEV

using std::fmaxf;
using std::isfinite;

namespace spx {
namespace fitdev {

struct WaveState {
    std::barrier<> bar{64};
    float fv[64];
    int iv[64];
    double dv[64];
    std::atomic<long> shuffles[64];
};
thread_local int t_lane = 0;
thread_local WaveState* t_wave = nullptr;

inline int spx_lane() { return t_lane; }
inline float spx_shfl_xor(float v, int off) {
    WaveState& w = *t_wave;
    w.shuffles[t_lane]++;
    w.fv[t_lane] = v;
    w.bar.arrive_and_wait();
    float r = w.fv[t_lane ^ off];
    w.bar.arrive_and_wait();
    return r;
}
inline int spx_shfl(int v, int src) {
    WaveState& w = *t_wave;
    w.shuffles[t_lane]++;
    w.iv[t_lane] = v;
    w.bar.arrive_and_wait();
    int r = w.iv[src];
    w.bar.arrive_and_wait();
    return r;
}
inline float spx_uniform(float v) {
    WaveState& w = *t_wave;
    w.shuffles[t_lane]++;
    w.fv[t_lane] = v;
    w.bar.arrive_and_wait();
    float r = w.fv[0];
    w.bar.arrive_and_wait();
    return r;
}
inline int spx_uniform_i(int v) { return spx_shfl(v, 0); }
inline double spx_uniform_d(double v) {
    WaveState& w = *t_wave;
    w.shuffles[t_lane]++;
    w.dv[t_lane] = v;
    w.bar.arrive_and_wait();
    double r = w.dv[0];
    w.bar.arrive_and_wait();
    return r;
}
// float read from an arbitrary lane (every lane must call it, like a hardware shuffle)
inline float spx_shfl_f(float v, int src) {
    WaveState& w = *t_wave;
    w.shuffles[t_lane]++;
    w.fv[t_lane] = v;
    w.bar.arrive_and_wait();
    float r = w.fv[src];
    w.bar.arrive_and_wait();
    return r;
}
// Butterfly steps of wave_sum_many: the first value's partial sum stays in lanes with the
// step's bit clear, the second's in lanes with it set (partner lanes as the hardware pairs them).
inline float emu_pair(float a, float b, int bit, int partner) {
    const bool hi = (t_lane >> bit) & 1;
    const float send = hi ? a : b;
    const float keep = hi ? b : a;
    return keep + spx_shfl_f(send, partner);
}
inline float spx_pair_sum32(float a, float b) { return emu_pair(a, b, 5, t_lane ^ 32); }
inline float spx_pair_sum16(float a, float b) { return emu_pair(a, b, 4, t_lane ^ 16); }
inline int dpp_partner(int ctrl) {
    const int l = t_lane;
    switch (ctrl) {
        case 0x140: return (l & ~15) | (15 - (l & 15));  // row_mirror
        case 0x141: return (l & ~7) | (7 - (l & 7));     // row_half_mirror
        case 0x4E: return l ^ 2;                         // quad_perm [2,3,0,1]
        default: return l ^ 1;                           // quad_perm [1,0,3,2]
    }
}
template <int CTRL, int BIT>
inline float spx_pair_dpp(float a, float b) { return emu_pair(a, b, BIT, dpp_partner(CTRL)); }
template <int CTRL>
inline float spx_self_dpp(float a) { return a + spx_shfl_f(a, dpp_partner(CTRL)); }
inline float spx_readlane(float v, int lane) { return spx_shfl_f(v, lane); }

inline float spx_wave_sum(float v) {
    for (int off = 32; off >= 1; off >>= 1) v += spx_shfl_xor(v, off);
    return spx_uniform(v);
}
inline float spx_wave_max(float v) {
    for (int off = 32; off >= 1; off >>= 1) v = std::max(v, spx_shfl_xor(v, off));
    return spx_uniform(v);
}
inline int spx_readlane_i(int v, int src) { return spx_shfl(v, src); }
inline double spx_readlane_d(double v, int src) {
    WaveState& w = *t_wave;
    w.shuffles[t_lane]++;
    w.dv[t_lane] = v;
    w.bar.arrive_and_wait();
    double r = w.dv[src];
    w.bar.arrive_and_wait();
    return r;
}
inline uint64_t spx_ballot(bool p) {
    WaveState& w = *t_wave;
    w.shuffles[t_lane]++;
    w.iv[t_lane] = p ? 1 : 0;
    w.bar.arrive_and_wait();
    uint64_t m = 0;
    for (int l = 0; l < 64; ++l) m |= (uint64_t)(w.iv[l] & 1) << l;
    w.bar.arrive_and_wait();
    return m;
}
inline int spx_lanes_below(uint64_t m) {
    return t_lane == 0 ? 0 : __builtin_popcountll(m & ((1ull << t_lane) - 1ull));
}
// every lane's slab writes before the barrier are seen by every lane after it
inline void spx_wave_lds_sync() {
    WaveState& w = *t_wave;
    w.shuffles[t_lane]++;
    w.bar.arrive_and_wait();
}
inline int spx_queue_grab(int32_t* q, int n) {
    const int mine = t_lane == 0 ? __atomic_fetch_add(q, n, __ATOMIC_SEQ_CST) : 0;
    return spx_shfl(mine, 0);
}
inline unsigned long long spx_atomic_add(unsigned long long* p, unsigned long long v) {
    return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST);
}
inline float hw_rcp(float x) { return 1.0f / x; }
inline float hw_rsq(float x) { return 1.0f / std::sqrt(x); }
inline float hw_exp2(float x) { return std::exp2(x); }
inline double hw_rcp64(double d) { return 1.0 / d; }

}  // namespace fitdev
}  // namespace spx

#include "erf_fit_device.hpp"

using namespace spx::fitdev;

template <int NPL, int P>
static void run_wave(const FitArgs& a, int timeout_s) {
    WaveState w;
    for (auto& s : w.shuffles) s = 0;
    std::vector<uint32_t> slab((size_t)NPL * 64);  // the wave's LDS slab
    std::atomic<int> done{0};
    std::vector<std::thread> lanes;
    for (int l = 0; l < 64; ++l) {
        lanes.emplace_back([&, l] {
            t_lane = l;
            t_wave = &w;
            erf_fit_wave<NPL, P>(a, slab.data());
            done++;
        });
    }
    auto t0 = std::chrono::steady_clock::now();
    while (done.load() < 64) {
        std::this_thread::sleep_for(std::chrono::milliseconds(5));
        if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(timeout_s)) {
            std::fprintf(stderr, "wave emulation stuck: %d of 64 lanes finished; shuffles per lane:", done.load());
            for (int l = 0; l < 64; ++l) std::fprintf(stderr, " %ld", w.shuffles[l].load());
            std::fprintf(stderr, "\n");
            std::fflush(stderr);
            std::_Exit(3);
        }
    }
    for (auto& t : lanes) t.join();
}

extern "C" int emu_fit_batch_cells(const uint16_t* img, int64_t H, int64_t W, int64_t pitch,
                                   const float* seeds, const int16_t* half, const uint8_t* model,
                                   const float* init, const SpxFitOptions* opt, int32_t K, const int32_t* nbr,
                                   int32_t nb, float* params, float* rms, int16_t* iters, uint8_t* status,
                                   int timeout_s) {
    constexpr int kKB = 2 * kBuckets;
    std::vector<std::vector<int32_t>> order(kKB);
    for (int f = 0; f < K; ++f) {
        const int hs = half[f];
        const int md = model[f];
        if (hs < 2 || hs > kMaxHalf || (md != SPX_MODEL_ERF_DISK6 && md != SPX_MODEL_ERF_ELLIPSE7)) {
            for (int i = 0; i < 7; ++i) params[(int64_t)f * 7 + i] = NAN;
            rms[f] = NAN;
            iters[f] = 0;
            status[f] = SPX_FIT_TOO_FEW;
            continue;
        }
        FitArgs pa{};
        pa.seeds = seeds;
        pa.half = half;
        pa.nbr = nbr;
        pa.nb = nb;
        pa.nfeat = K;
        pa.H = H;
        pa.W = W;
        const int bound = patch_pixel_bound(pa, f);
        order[2 * bucket_of_pixels(bound) + (md == SPX_MODEL_ERF_ELLIPSE7 ? 1 : 0)].push_back(f);
    }
    unsigned long long stats[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int b = 0; b < kKB; ++b) {
        int32_t count = (int32_t)order[b].size();
        int32_t queue = 0;
        if (count == 0) continue;
        FitArgs a{img, H, W, pitch, seeds, half, model, init, *opt, order[b].data(), &count, &queue,
                  stats, params, nullptr, rms, iters, status, nbr, nb, K};
        switch (b) {
            case 0: run_wave<4, 6>(a, timeout_s); break;
            case 1: run_wave<4, 7>(a, timeout_s); break;
            case 2: run_wave<8, 6>(a, timeout_s); break;
            case 3: run_wave<8, 7>(a, timeout_s); break;
            case 4: run_wave<16, 6>(a, timeout_s); break;
            case 5: run_wave<16, 7>(a, timeout_s); break;
            case 6: run_wave<32, 6>(a, timeout_s); break;
            case 7: run_wave<32, 7>(a, timeout_s); break;
            case 8: run_wave<64, 6>(a, timeout_s); break;
            default: run_wave<64, 7>(a, timeout_s); break;
        }
    }
    return (int)stats[5];  // features with more pixels than their bucket holds
}

// The bucketing's bound (out) and the exact count (exact) of every feature's fitted pixels
// (patch_pixel_bound / patch_pixel_count, plain per-thread code).
extern "C" void emu_pixel_bound(const float* seeds, const int16_t* half, const int32_t* nbr, int32_t nb, int32_t K,
                                int64_t H, int64_t W, int32_t* out, int32_t* exact) {
    FitArgs a{};
    a.seeds = seeds;
    a.half = half;
    a.nbr = nbr;
    a.nb = nb;
    a.nfeat = K;
    a.H = H;
    a.W = W;
    for (int f = 0; f < K; ++f) {
        out[f] = patch_pixel_bound(a, f);
        exact[f] = patch_pixel_count(a, f);
    }
}

extern "C" int emu_fit_batch(const uint16_t* img, int64_t H, int64_t W, int64_t pitch,
                             const float* seeds, const int16_t* half, const uint8_t* model,
                             const float* init, const SpxFitOptions* opt, int32_t K, float* params,
                             float* rms, int16_t* iters, uint8_t* status, int timeout_s) {
    return emu_fit_batch_cells(img, H, W, pitch, seeds, half, model, init, opt, K, nullptr, 0, params, rms, iters,
                               status, timeout_s);
}

// Cell decisions over feature f's whole patch (side x side, row-major) from the row bounds
// the gather uses (cell_row, packed as the kernel packs them): member and boundary flags.
extern "C" void emu_cell_rows(const float* seeds, const int16_t* half, const int32_t* nbr, int32_t nb, int32_t K,
                              int32_t f, uint8_t* in_out, uint8_t* edge_out) {
    using namespace spx::fitdev;
    FitArgs a{};
    a.seeds = seeds;
    a.half = half;
    a.nbr = nbr;
    a.nb = nb;
    a.nfeat = K;
    const int hs = half[f];
    const int side = 2 * hs + 1;
    const int64_t ox = (int64_t)rintf(seeds[2 * f]) - hs, oy = (int64_t)rintf(seeds[2 * f + 1]) - hs;
    for (int r = 0; r < side; ++r) {
        const uint32_t rb = pack_row(cell_row(MemTerms(a, f), ox, oy, side, r));
        for (int c = 0; c < side; ++c) {
            const int c1 = c + 1;
            const bool in = c1 >= (int)(rb & 0xffu) && c1 <= (int)((rb >> 8) & 0xffu);
            const bool interior = c1 >= (int)((rb >> 16) & 0xffu) && c1 <= (int)(rb >> 24);
            in_out[r * side + c] = in;
            edge_out[r * side + c] = in && !interior;
        }
    }
}

// The row walk's shortcut against the walk itself and a brute-force scan: for n random pairs
// of float seeds (fractional, quarter- and half-integer, integer, and 1e-15 off an integer, so
// that many roots fall on or next to a column), every row of a patch around the first seed,
// each neighbour rule's first (a > 0) or last (a < 0) kept and interior column, found (1) with
// the margin, (2) by the walk alone and (3) by scanning every column. Returns the number of
// decisions; *mismatch counts those where the three differ, *walked those the margin left to
// the walk.
extern "C" int64_t emu_row_walk_check(uint64_t seed, int32_t n, int64_t* mismatch, int64_t* walked) {
    using namespace spx::fitdev;
    uint64_t s = seed * 0x9E3779B97F4A7C15ull + 1;
    const auto next = [&]() {
        s ^= s << 13;
        s ^= s >> 7;
        s ^= s << 17;
        return s;
    };
    const auto unit = [&]() { return (double)(next() >> 11) * 0x1.0p-53 * 2.0 - 1.0; };
    const double fr[] = {0.0, 0.5, 0.25, 0.125, 1e-15, -1e-15, 1e-9, 0.1, 1.0 / 3.0};
    int64_t dec = 0, bad = 0, slow = 0;
    for (int it = 0; it < n; ++it) {
        const int mode = it % 4;
        const auto pick = [&](double base) {
            const double w = base + unit() * 20.0;
            if (mode == 0) return w;
            if (mode == 3) return std::round(w);
            return std::round(w) + fr[next() % 9] * (mode == 2 && unit() < 0.0 ? -1.0 : 1.0);
        };
        const double base = unit() * 8000.0 + 8000.0;
        const float sx = (float)pick(base), sy = (float)pick(base);
        const float jx = (float)pick(sx), jy = (float)pick(sy);
        if (sx == jx && sy == jy) continue;
        const double x0 = sx, y0 = sy;
        const CellTerm t = cell_term(x0, y0, x0 * x0 + y0 * y0, jx, jy, (it & 1) != 0);
        if (t.a == 0.0) continue;
        const int hs = 3 + (int)(next() % 20), side = 2 * hs + 1;
        const int64_t ox = (int64_t)rintf(sx) - hs, oy = (int64_t)rintf(sy) - hs;
        const double oxd = (double)ox;
        for (int r = 0; r < side; ++r) {
            const double by = t.b * (double)(oy + r);
            const auto keeps = [&](int c) { return cell_keeps(t, cell_f(t, oxd + (double)c, by)); };
            const auto interior = [&](int c) { return cell_interior(t, cell_f(t, oxd + (double)c, by)); };
            const double ek = -(by + t.c) * t.inv - oxd, ei = (t.m - by - t.c) * t.inv - oxd;
            const double mk = root_margin(t, by, oxd, side, ek), mi = root_margin(t, by, oxd, side, ei);
            int got[2][3];
            const bool up = t.a > 0.0;
            for (int q = 0; q < 2; ++q) {
                const double est = q ? ei : ek, mg = q ? mi : mk;
                const auto scan = [&](const auto& p) {
                    if (up) {
                        for (int c = 0; c < side; ++c)
                            if (p(c)) return c;
                        return side;
                    }
                    for (int c = side - 1; c >= 0; --c)
                        if (p(c)) return c;
                    return -1;
                };
                const auto run = [&](const auto& p) {
                    got[q][0] = up ? first_true(p, est, side, mg) : last_true(p, est, side, mg);
                    got[q][1] = up ? first_true(p, est, side, INFINITY) : last_true(p, est, side, INFINITY);
                    got[q][2] = scan(p);
                };
                if (q) run(interior);
                else run(keeps);
                const double near = std::fabs(est - std::round(est));
                slow += !(near > mg);
                bad += got[q][0] != got[q][1] || got[q][0] != got[q][2];
                ++dec;
            }
        }
    }
    *mismatch = bad;
    *walked = slow;
    return dec;
}
