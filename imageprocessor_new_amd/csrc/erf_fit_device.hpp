// Device-side logic of the batched erf-disk / erf-ellipse LM fit (one wave64 per feature).
//
// Included by erf_fit.hip (HIP build) after it defines the wave primitives:
//   SPX_DEV, SPX_DEV_INLINE, spx_lane(),
//   spx_wave_sum(float) / spx_wave_max(float): reductions over the wave, result uniform,
//   spx_uniform(float) / spx_uniform_i(int) / spx_uniform_d(double): the first lane's value,
//   spx_queue_grab(int32_t*, int n): one atomic add of n per wave, old value uniform,
//   spx_atomic_add(unsigned long long*, unsigned long long),
//   spx_pair_sum32 / spx_pair_sum16 / spx_pair_dpp<CTRL, BIT> / spx_self_dpp<CTRL> /
//   spx_readlane: the butterfly steps of wave_sum_many,
//   spx_shfl(int v, int src): lane src's v, spx_readlane_i / spx_readlane_d(v, lane): lane's v
//   as a uniform value,
//   spx_ballot(bool): the wave's 64-bit predicate mask (uniform), spx_lanes_below(mask): set
//   bits of the mask in lanes below this one, spx_wave_lds_sync(): LDS writes of every lane
//   visible to every lane of the wave,
//   hw_rcp / hw_rsq / hw_exp2 (float) and hw_rcp64 (double).
// Every value that steers control flow is wave-uniform by construction, so the compiler
// sees uniform (scalar) branches only. (A lane-dependent branch around a shuffle inside the
// persistent loop lets it give lanes different loop headers, and the wave then never
// reconverges.) The same text compiles against a host emulation of a wave (64 threads
// exchanging through barriers) in the tests, which checks on the CPU that every lane takes
// the same path.
//
// Model (patch-local coordinates, pixel centres at integers):
//   I(x, y) = B + A * E(u),  E(u) = erfc(u) / 2,  u = (rho - r) / (sqrt(2) * s)
//   rho = sqrt(dx^2 + dy^2) (disk, 6 parameters) or sqrt(dx^2 + (dy/q)^2) (ellipse, 7th q)
//   dx = x - x0, dy = y - y0.
// The reference has no erf-disk fit; the batched LM layout follows its wave-per-cell
// Gaussian LM kernel `lm_fit` (reference subpx/_gpu/deblend.py:55-152: butterfly totals
// :32-39/:107, LU solve :113-130, accept/reject :141-149) and the disk LM `_fit_disks`
// (reference subpx/_cpu/invariant_cd.py:165-246), redesigned for CDNA4:
//   * the patch is gathered once from HBM and compacted to the pixels the fit uses (inside
//     the image and the cell), one 32-bit word per pixel (column, row and grey level
//     packed), at most NPL words per lane, held in VGPRs (or LDS for large patches); the
//     feature's bucket (NPL) comes from that pixel count, found before the fit;
//   * residual and Jacobian are FP32 on the VALU, with hardware rsq / rcp / exp2;
//   * J^T J, J^T r and the cost are reduced over the wave with DPP (no LDS, no atomics),
//     so a result does not depend on scheduling;
//   * the damped normal equations are solved in FP64 (in-place packed LDL^T);
//   * the Jacobian pass at a trial point doubles as the trial cost evaluation, so an
//     accepted step costs one pixel pass;
//   * waves pull features from an atomic work queue per (patch size, model) bucket.
#pragma once

#include <stdint.h>

#include "subpx_hip.h"

namespace spx {
namespace fitdev {

constexpr int kWave = 64;

// Abramowitz & Stegun 7.1.26: erf(x) = 1 - t (a1 + t (a2 + t (a3 + t (a4 + t a5)))) e^{-x^2},
// t = 1 / (1 + p x), x >= 0, |error| <= 1.5e-7. So erfc(|u|) = t * poly(t) * e^{-u^2},
// and e^{-u^2} is shared with the Jacobian.
constexpr float kAsP = 0.3275911f;
constexpr float kAs1 = 0.254829592f;
constexpr float kAs2 = -0.284496736f;
constexpr float kAs3 = 1.421413741f;
constexpr float kAs4 = -1.453152027f;
constexpr float kAs5 = 1.061405429f;

constexpr float kLog2e = 1.4426950408889634f;
constexpr float kSqrt2 = 1.4142135623730951f;
constexpr float kInvSqrt2 = 0.7071067811865476f;
constexpr float kInvSqrt2Pi = 0.3989422804014327f;
constexpr float kAcceptRtol = 1.0e-6f;
constexpr int kGrab = 4;        // features claimed per work-queue atomic
constexpr int kUnrollMax = 16;  // pixel loops of up to this many slots per lane are unrolled

constexpr int kBuckets = 5;  // NPL = 4, 8, 16, 32, 64 pixels per lane
constexpr int kMaxHalf = 31;

// Sum layout for P parameters with the background B last (column of ones):
//   [ JtJ upper triangle over the P-1 non-B columns | sum J_a (a < P-1) | J^T r (P) | cost ]
// JtJ[B][B] is the pixel count, known from the gather.
template <int P>
struct Layout {
    static constexpr int kNB = P - 1;
    static constexpr int kTri = kNB * (kNB + 1) / 2;
    static constexpr int kColB = kTri;
    static constexpr int kRhs = kColB + kNB;
    static constexpr int kCost = kRhs + P;
    static constexpr int kN = kCost + 1;
};

// Uniform per-feature parameters in internal order: x0 y0 r s A [q] B.
template <int P>
struct Params {
    double p[P];
};

// Patch pixel word: bits 0-7 column, 8-15 row, 16-31 grey level.
SPX_DEV_INLINE uint32_t pack_px(int col, int row, uint32_t v) {
    return (uint32_t)col | ((uint32_t)row << 8) | (v << 16);
}

// Work of one patch: the fitted pixels compacted in patch order, pixel k in slot k / 64 of
// lane k % 64, so every slot below `nslots` - 1 is full; slots are weighted by `valid` (bit
// n = slot n), later slots are skipped. The words are compacted in the wave's LDS slab (word
// n of lane l at n * 64 + l, conflict free); patches of up to kUnrollMax slots per lane then
// copy them into VGPRs (fully unrolled loops), larger ones read them from the slab, so the
// LM state keeps its registers.
template <int NPL>
struct Patch {
    static constexpr bool kLds = NPL > kUnrollMax;
    uint32_t px[kLds ? 1 : NPL];
    uint32_t* lds;
    uint64_t valid;
    int nslots;
    float npix;
    SPX_DEV_INLINE uint32_t word(int n) const {
        if constexpr (kLds) return lds[n * kWave + spx_lane()];
        else return px[n];
    }
};

// One pixel pass at parameters `pp`: this lane's partial sums of J^T J, J^T r and the
// cost over its pixels (not yet reduced over the wave), with the Jacobian columns of x0, y0,
// r, sigma (and q) accumulated without their constant factors, which `sc` returns: the
// caller scales the reduced sums once per pass instead of every pixel once per pass.
// Per pixel: w = (rho - r) sqrt(log2 e) / (sqrt2 s), so that e^{-u^2} = 2^{-w^2} with
// u = (rho - r) / (sqrt2 s); erfc(|u|) / 2 = t p(t) 2^{-w^2} with the A&S coefficients halved
// and t = 1 / (1 + p |u|); empty lanes of a partly valid slot are zeroed by multiplying
// the residual, E and 2^{-w^2} by the slot's valid bit.
constexpr float kSqrtLog2e = 1.2011224087864498f;

template <int NPL, int P>
SPX_DEV_INLINE void pixel_pass(const Patch<NPL>& pt, int hs, const Params<P>& pp,
                               float (&acc)[Layout<P>::kN], float (&sc)[Layout<P>::kNB]) {
    using L = Layout<P>;
    constexpr bool kEll = (P == 7);
    // dx = col - hs - x0 = col - xo
    const float xo = (float)(pp.p[0] + (double)hs), yo = (float)(pp.p[1] + (double)hs);
    const float r = (float)pp.p[2];
    const float s = (float)pp.p[3], A = (float)pp.p[4], B = (float)pp.p[P - 1];
    const float q = kEll ? (float)pp.p[5] : 1.0f;
    const float inv_s = hw_rcp(s);
    const float kw = (kInvSqrt2 * kSqrtLog2e) * inv_s;
    const float rkw = r * kw;
    const float iq = kEll ? hw_rcp(q) : 1.0f;
    const float iq2 = iq * iq;
    // true columns: x0, y0, r: Ac * (2^{-w^2} rs dx, 2^{-w^2} rs dy q^-2, 2^{-w^2});
    // sigma: Ac sqrt2 / sqrt(log2 e) * 2^{-w^2} w; q: Ac * 2^{-w^2} rs dy^2 q^-3
    const float Ac = A * kInvSqrt2Pi * inv_s;
    sc[0] = Ac;
    sc[1] = Ac;
    sc[2] = Ac;
    sc[3] = Ac * (kSqrt2 / kSqrtLog2e);
    sc[4] = 1.0f;
    if constexpr (kEll) sc[5] = Ac;
    constexpr float kP = kAsP / kSqrtLog2e;
    constexpr float h1 = 0.5f * kAs1, h2 = 0.5f * kAs2, h3 = 0.5f * kAs3, h4 = 0.5f * kAs4, h5 = 0.5f * kAs5;

#pragma unroll
    for (int i = 0; i < L::kN; ++i) acc[i] = 0.0f;

    // Register patches: a uniform guard keeps the loop fully unrolled with px[] in VGPRs;
    // LDS patches run a rolled loop over their slots.
#pragma unroll(Patch<NPL>::kLds ? 4 : NPL)
    for (int n = 0; n < NPL; ++n) {
        if constexpr (Patch<NPL>::kLds) {
            if (n >= pt.nslots) break;
        }
        if (n < pt.nslots) {
            const uint32_t wd = pt.word(n);
            const float dx = (float)(wd & 0xffu) - xo;
            const float dy = (float)((wd >> 8) & 0xffu) - yo;
            const float val = (float)(wd >> 16);
            const float dyq = kEll ? dy * iq2 : dy;
            const float rho2 = fmaf(dy, dyq, fmaf(dx, dx, 1e-20f));
            const float rs = hw_rsq(rho2);
            const float rho = rho2 * rs;
            const float w = fmaf(rho, kw, -rkw);
            const float t = hw_rcp(fmaf(kP, fabsf(w), 1.0f));
            const float wm = (float)((uint32_t)(pt.valid >> n) & 1u);
            const float ez = hw_exp2(-(w * w)) * wm;
            const float hq = t * fmaf(t, fmaf(t, fmaf(t, fmaf(t, h5, h4), h3), h2), h1) * ez;
            const float E = (w >= 0.0f) ? hq : wm - hq;
            const float res = fmaf(val, wm, -fmaf(A, E, B * wm));
            const float Gr = ez * rs;
            float J[P - 1];
            J[0] = Gr * dx;
            J[1] = Gr * dyq;
            J[2] = ez;
            J[3] = ez * w;
            J[4] = E;
            if constexpr (kEll) J[5] = Gr * dy * dyq * iq;
            int idx = 0;
#pragma unroll
            for (int a = 0; a < L::kNB; ++a)
#pragma unroll
                for (int b = a; b < L::kNB; ++b) {
                    acc[idx] = fmaf(J[a], J[b], acc[idx]);
                    ++idx;
                }
#pragma unroll
            for (int a = 0; a < L::kNB; ++a) acc[L::kColB + a] += J[a];
#pragma unroll
            for (int a = 0; a < L::kNB; ++a) acc[L::kRhs + a] = fmaf(J[a], res, acc[L::kRhs + a]);
            acc[L::kRhs + P - 1] += res;
            acc[L::kCost] = fmaf(res, res, acc[L::kCost]);
        }
    }
}

// Wave totals of M per-lane values at once (a transposed butterfly). Six pair steps, one per
// lane-index bit from 5 down to 0, each add two registers over partner lanes and halve the
// register count; a register left without a partner is added to itself across the bit. After
// the sixth step every value's wave total sits in the lane whose index is the 6-bit reversal
// of the value's index, read out with readlane. For M = 26 this takes ~60 VALU operations plus
// 26 readlanes, against ~10 per value for separate reductions.
SPX_DEV_INLINE constexpr int bitrev6(int i) {
    return ((i & 1) << 5) | ((i & 2) << 3) | ((i & 4) << 1) | ((i & 8) >> 1) | ((i & 16) >> 3) | ((i & 32) >> 5);
}

template <int M>
SPX_DEV_INLINE void wave_sum_many(const float (&v)[M], float (&out)[M]) {
    static_assert(M >= 1 && M <= 64, "wave_sum_many: 1..64 values");
    constexpr int M1 = (M + 1) / 2, M2 = (M1 + 1) / 2, M3 = (M2 + 1) / 2;
    constexpr int M4 = (M3 + 1) / 2, M5 = (M4 + 1) / 2, M6 = (M5 + 1) / 2;
    float r1[M1], r2[M2], r3[M3], r4[M4], r5[M5], r6[M6];
#pragma unroll
    for (int j = 0; j < M1; ++j) r1[j] = (2 * j + 1 < M) ? spx_pair_sum32(v[2 * j], v[2 * j + 1]) : spx_pair_sum32(v[2 * j], v[2 * j]);
#pragma unroll
    for (int j = 0; j < M2; ++j) r2[j] = (2 * j + 1 < M1) ? spx_pair_sum16(r1[2 * j], r1[2 * j + 1]) : spx_pair_sum16(r1[2 * j], r1[2 * j]);
#pragma unroll
    for (int j = 0; j < M3; ++j)
        r3[j] = (2 * j + 1 < M2) ? spx_pair_dpp<0x140, 3>(r2[2 * j], r2[2 * j + 1]) : spx_self_dpp<0x140>(r2[2 * j]);
#pragma unroll
    for (int j = 0; j < M4; ++j)
        r4[j] = (2 * j + 1 < M3) ? spx_pair_dpp<0x141, 2>(r3[2 * j], r3[2 * j + 1]) : spx_self_dpp<0x141>(r3[2 * j]);
#pragma unroll
    for (int j = 0; j < M5; ++j)
        r5[j] = (2 * j + 1 < M4) ? spx_pair_dpp<0x4E, 1>(r4[2 * j], r4[2 * j + 1]) : spx_self_dpp<0x4E>(r4[2 * j]);
#pragma unroll
    for (int j = 0; j < M6; ++j)
        r6[j] = (2 * j + 1 < M5) ? spx_pair_dpp<0xB1, 0>(r5[2 * j], r5[2 * j + 1]) : spx_self_dpp<0xB1>(r5[2 * j]);
#pragma unroll
    for (int i = 0; i < M; ++i) out[i] = spx_readlane(r6[0], bitrev6(i));
}

// Wave totals of every partial sum except the cost (which the caller reduces first), with
// the constant column factors `sc` of the pass applied.
template <int P>
SPX_DEV_INLINE void reduce_sums(const float (&acc)[Layout<P>::kN], const float (&sc)[Layout<P>::kNB],
                                float (&S)[Layout<P>::kN]) {
    using L = Layout<P>;
    constexpr int M = L::kCost;
    float in[M], out[M];
#pragma unroll
    for (int i = 0; i < M; ++i) in[i] = acc[i];
    wave_sum_many<M>(in, out);
    int idx = 0;
#pragma unroll
    for (int a = 0; a < L::kNB; ++a)
#pragma unroll
        for (int b = a; b < L::kNB; ++b) {
            S[idx] = out[idx] * (sc[a] * sc[b]);
            ++idx;
        }
#pragma unroll
    for (int a = 0; a < L::kNB; ++a) S[L::kColB + a] = out[L::kColB + a] * sc[a];
#pragma unroll
    for (int a = 0; a < L::kNB; ++a) S[L::kRhs + a] = out[L::kRhs + a] * sc[a];
    S[L::kRhs + P - 1] = out[L::kRhs + P - 1];
}

// Solves (JtJ + lambda * diag) delta = J^T r in FP64 by an in-place LDL^T on the packed
// lower triangle. Returns false when a pivot is not positive.
template <int P>
SPX_DEV_INLINE bool damped_solve(const float (&S)[Layout<P>::kN], double npix, double lambda,
                                 double (&delta)[P]) {
    using L = Layout<P>;
    constexpr int kT = P * (P + 1) / 2;
    // packed lower triangle: element (i, j), j <= i, at i (i + 1) / 2 + j
    double T[kT];
    int idx = 0;
#pragma unroll
    for (int a = 0; a < L::kNB; ++a)
#pragma unroll
        for (int b = a; b < L::kNB; ++b) {
            T[b * (b + 1) / 2 + a] = (double)S[idx];
            ++idx;
        }
#pragma unroll
    for (int a = 0; a < L::kNB; ++a) T[(P - 1) * P / 2 + a] = (double)S[L::kColB + a];
    T[kT - 1] = npix;
    double dmax = 0.0;
#pragma unroll
    for (int a = 0; a < P; ++a) dmax = fmax(dmax, T[a * (a + 1) / 2 + a]);
    const double floor_d = 1e-12 * dmax;
#pragma unroll
    for (int a = 0; a < P; ++a) {
        const double d = T[a * (a + 1) / 2 + a];
        T[a * (a + 1) / 2 + a] = d + lambda * fmax(d, floor_d);
    }
    // LDL^T keeping both L (below the diagonal of T) and W = L D (the values before the
    // division by the pivot), so every elimination term is one FMA; the pivots' reciprocals
    // serve the back substitution too.
    double D[P], iD[P], Wm[kT];
#pragma unroll
    for (int j = 0; j < P; ++j) {
        const double mjj = T[j * (j + 1) / 2 + j];
        double d = mjj;
#pragma unroll
        for (int kk = 0; kk < j; ++kk) d = fma(-T[j * (j + 1) / 2 + kk], Wm[j * (j + 1) / 2 + kk], d);
        if (!(d > 1e-13 * mjj) || !(d > 0.0)) return false;
        D[j] = d;
        const double inv = hw_rcp64(d);
        iD[j] = inv;
#pragma unroll
        for (int i = j + 1; i < P; ++i) {
            double sacc = T[i * (i + 1) / 2 + j];
#pragma unroll
            for (int kk = 0; kk < j; ++kk) sacc = fma(-T[i * (i + 1) / 2 + kk], Wm[j * (j + 1) / 2 + kk], sacc);
            Wm[i * (i + 1) / 2 + j] = sacc;
            T[i * (i + 1) / 2 + j] = sacc * inv;
        }
    }
    double y[P];
#pragma unroll
    for (int i = 0; i < P; ++i) {
        double sacc = (double)S[L::kRhs + i];
#pragma unroll
        for (int kk = 0; kk < i; ++kk) sacc = fma(-T[i * (i + 1) / 2 + kk], y[kk], sacc);
        y[i] = sacc;
    }
#pragma unroll
    for (int i = P - 1; i >= 0; --i) {
        double sacc = y[i] * iD[i];
#pragma unroll
        for (int kk = i + 1; kk < P; ++kk) sacc = fma(-T[kk * (kk + 1) / 2 + i], delta[kk], sacc);
        delta[i] = sacc;
    }
    bool finite = true;
#pragma unroll
    for (int i = 0; i < P; ++i) finite = finite && isfinite(delta[i]);
    return finite;
}

struct FitArgs {
    const uint16_t* img;
    int64_t H, W, pitch;
    const float* seeds;
    const int16_t* half;
    const uint8_t* model;
    const float* init;
    SpxFitOptions opt;
    const int32_t* order;       // feature ids of this bucket
    const int32_t* count;       // number of ids in this bucket (device)
    int32_t* queue;             // work-queue head of this bucket (device, starts at 0)
    unsigned long long* stats;  // [pixel evaluations P=6, P=7, pixel passes, solves]
    float* params;
    double* centers;     // (K, 2) FP64 absolute centres, or null
    float* rms;
    int16_t* iters;
    uint8_t* status;
    const int32_t* nbr;  // (K, nb) neighbour feature ids, -1 = none; null = no cell mask
    int32_t nb;
    int32_t nfeat;       // K: ids outside [0, K) are ignored
    // When set, `order` is the start of one array holding every bucket's ids back to back and
    // this bucket's list starts after the ids of buckets 0 .. bucket - 1 (their counts here).
    const int32_t* bucket_counts;
    int32_t bucket;
};

// Voronoi cell of a feature against its listed neighbour seeds. A pixel q belongs to the
// cell when it is strictly nearer to the own seed s0 than to every neighbour s_j, ties going
// to the lower feature index (reference subpx/_cpu/cells.py:8-33). The test uses the linear
// form f_j(q) = |q - s_j|^2 - |q - s0|^2 = (a_j qx + b_j qy) + c_j with a_j = 2 (s0x - s_jx),
// b_j = 2 (s0y - s_jy), c_j = (s_jx^2 + s_jy^2) - (s0x^2 + s0y^2), in FP64 without contraction
// and in the oracle's operation order, so both take the same decisions: inside when every
// f_j > 0 (or = 0 towards a higher index). A pixel of the cell is on its boundary (a
// 4-neighbour lies outside) when some g_j = f_j(q) - max(|a_j|, |b_j|) is negative, or zero
// towards a lower index: the 4-neighbours shift f_j by +-a_j or +-b_j. Absent neighbours
// (id < 0 or >= K) never decide; only the first kMaxNb listed neighbours count.
constexpr int kMaxNb = 8;

struct CellTerm {
    double a, b, c, m;  // m = max(|a|, |b|)
    double inv;         // 1 / a (0 when a = 0): only to estimate where a rule changes
    bool lower;         // the neighbour's index is below the feature's
};

SPX_DEV_INLINE CellTerm cell_term(double x0, double y0, double n0, double xj, double yj, bool lower) {
#pragma clang fp contract(off)
    CellTerm t;
    t.a = 2.0 * (x0 - xj);
    t.b = 2.0 * (y0 - yj);
    t.c = (xj * xj + yj * yj) - n0;
    t.m = fmax(fabs(t.a), fabs(t.b));
    t.inv = t.a != 0.0 ? 1.0 / t.a : 0.0;
    t.lower = lower;
    return t;
}

// f_j at (qx, qy) given by = b_j qy, in the operation order above.
SPX_DEV_INLINE double cell_f(const CellTerm& t, double qx, double by) {
#pragma clang fp contract(off)
    return (t.a * qx + by) + t.c;
}
SPX_DEV_INLINE bool cell_keeps(const CellTerm& t, double fj) { return fj > 0.0 || (fj == 0.0 && !t.lower); }
SPX_DEV_INLINE bool cell_interior(const CellTerm& t, double fj) {
#pragma clang fp contract(off)
    const double g = fj - t.m;
    return !(g < 0.0 || (g == 0.0 && t.lower));
}

// The exact rules at one pixel, from the seeds in memory (edge: the boundary flag).
SPX_DEV_INLINE bool cell_exact(const FitArgs& a, int f, double qx, double qy, bool edge) {
#pragma clang fp contract(off)
    const double x0 = (double)a.seeds[2 * (int64_t)f], y0 = (double)a.seeds[2 * (int64_t)f + 1];
    const double n0 = x0 * x0 + y0 * y0;
    bool in = true, e = false;
    for (int j = 0; j < a.nb && j < kMaxNb; ++j) {
        const int id = a.nbr[(int64_t)f * a.nb + j];
        if (id < 0 || id >= a.nfeat) continue;
        const CellTerm t = cell_term(x0, y0, n0, (double)a.seeds[2 * (int64_t)id],
                                     (double)a.seeds[2 * (int64_t)id + 1], id < f);
        const double fj = cell_f(t, qx, t.b * qy);
        in = in && cell_keeps(t, fj);
        e = e || !cell_interior(t, fj);
    }
    return edge ? e : in;
}

// Along one row every rule is monotone in the column: f_j is a chain of FP64 operations each
// monotone in qx (non-decreasing for a_j > 0, non-increasing for a_j < 0, constant for 0),
// so the columns a neighbour keeps (or leaves interior) are a prefix or a suffix of the row.
// Its end is found from the real root of f_j (or g_j) and then settled by evaluating the
// exact rule at the columns next to it, moving at most side steps, so the result is exact
// whatever the root's rounding. The cell's part of the row is then the intersection of the
// neighbours' intervals: members in [lo, hi], interior members in [elo, ehi] (empty when lo >
// hi); the columns are patch-local, 0 .. side - 1, at qx = ox + column.
struct CellRow {
    int lo, hi, elo, ehi;
};

SPX_DEV_INLINE int imax(int a, int b) { return a > b ? a : b; }
SPX_DEV_INLINE int imin(int a, int b) { return a < b ? a : b; }

// How far (in columns) a column must lie from a rule's root estimate for the exact rule's
// decision there to be the root's side, so that the columns next to the estimate need not be
// evaluated. The rule's FP64 value at column x, ((a qx + by) + c) [- m] with a, b, qx, qy exact,
// is within 4 u (|a| (|ox| + side) + |by| + |c| + m) of its real value (u = 2^-53), and the
// estimate -(by + c) / a - ox (or (m - by - c) / a - ox) within 5 u (|by| + |c| + m) / |a| + u |est|
// of the real root; a column farther from the estimate than the sum over |a| has the real sign.
// The margin takes 64 u over those magnitudes (about 7 times the bound) plus 1e-12; a NaN or
// infinite margin (|a| near 0) never passes, and the exact walk decides.
SPX_DEV_INLINE double root_margin(const CellTerm& t, double by, double oxd, int side, double est) {
#pragma clang fp contract(off)
    constexpr double kU = 1.1102230246251565e-16;
    const double mag = fabs(t.a) * (fabs(oxd) + (double)side + 1.0) + fabs(by) + fabs(t.c) + t.m;
    return 64.0 * kU * (mag * fabs(t.inv) + fabs(est) + fabs(oxd) + (double)side) + 1e-12;
}

// first column in [0, side) where the monotone (false, then true) predicate holds, else side;
// the exact rule is evaluated only at a checked column within `margin` of the estimate
template <class Pred>
SPX_DEV_INLINE int first_true(const Pred& p, double est, int side, double margin) {
    int c = (est != est) ? 0 : (int)fmin(fmax(ceil(est), 0.0), (double)side);
    if ((c == 0 || est - (double)(c - 1) > margin) && (c == side || (double)c - est > margin)) return c;
    for (int k = 0; k <= side && c > 0 && p(c - 1); ++k) --c;
    for (int k = 0; k <= side && c < side && !p(c); ++k) ++c;
    return c;
}
// last column in [0, side) where the monotone (true, then false) predicate holds, else -1
template <class Pred>
SPX_DEV_INLINE int last_true(const Pred& p, double est, int side, double margin) {
    int c = (est != est) ? side - 1 : (int)fmin(fmax(floor(est), -1.0), (double)(side - 1));
    if ((c == side - 1 || (double)(c + 1) - est > margin) && (c == -1 || est - (double)c > margin)) return c;
    for (int k = 0; k <= side && c < side - 1 && p(c + 1); ++k) ++c;
    for (int k = 0; k <= side && c >= 0 && !p(c); ++k) --c;
    return c;
}

// The neighbour terms of feature f, term j read from memory (plain per-thread code).
struct MemTerms {
    const FitArgs& a;
    int f;
    double x0, y0, n0;
    SPX_DEV_INLINE MemTerms(const FitArgs& args, int feat) : a(args), f(feat) {
#pragma clang fp contract(off)
        x0 = (double)a.seeds[2 * (int64_t)f];
        y0 = (double)a.seeds[2 * (int64_t)f + 1];
        n0 = x0 * x0 + y0 * y0;
    }
    SPX_DEV_INLINE bool get(int j, CellTerm& t) const {
        if (j >= a.nb) return false;
        const int id = a.nbr[(int64_t)f * a.nb + j];
        if (id < 0 || id >= a.nfeat) return false;
        t = cell_term(x0, y0, n0, (double)a.seeds[2 * (int64_t)id], (double)a.seeds[2 * (int64_t)id + 1], id < f);
        return true;
    }
};

// The same terms spread over the wave: lane j < kMaxNb holds neighbour j's term, which every
// lane reads back with readlane (one pair of loads per lane instead of eight in turn).
struct LaneTerms {
    CellTerm mine;
    uint32_t act;  // bit j: neighbour j present
    SPX_DEV_INLINE bool get(int j, CellTerm& t) const {
        if (!((act >> j) & 1u)) return false;
        t.a = spx_readlane_d(mine.a, j);
        t.b = spx_readlane_d(mine.b, j);
        t.c = spx_readlane_d(mine.c, j);
        t.m = spx_readlane_d(mine.m, j);
        t.inv = spx_readlane_d(mine.inv, j);
        t.lower = spx_readlane_i(mine.lower ? 1 : 0, j) != 0;
        return true;
    }
};

SPX_DEV_INLINE LaneTerms lane_terms(const FitArgs& a, int f) {
#pragma clang fp contract(off)
    const int lane = spx_lane();
    const double x0 = (double)spx_uniform(a.seeds[2 * (int64_t)f]);
    const double y0 = (double)spx_uniform(a.seeds[2 * (int64_t)f + 1]);
    const double n0 = x0 * x0 + y0 * y0;
    int id = -1;
    if (lane < a.nb && lane < kMaxNb) id = a.nbr[(int64_t)f * a.nb + lane];
    const bool act = id >= 0 && id < a.nfeat;
    double xj = x0, yj = y0;
    if (act) {
        xj = (double)a.seeds[2 * (int64_t)id];
        yj = (double)a.seeds[2 * (int64_t)id + 1];
    }
    LaneTerms L;
    L.mine = cell_term(x0, y0, n0, xj, yj, act && id < f);
    L.act = (uint32_t)(spx_ballot(act) & 0xffull);
    return L;
}

template <class Terms>
SPX_DEV_INLINE CellRow cell_row(const Terms& terms, int64_t ox, int64_t oy, int side, int row) {
#pragma clang fp contract(off)
    CellRow r{0, side - 1, 0, side - 1};
    const double qy = (double)(oy + row);
    const double oxd = (double)ox;
#pragma unroll
    for (int j = 0; j < kMaxNb; ++j) {
        CellTerm t;
        if (!terms.get(j, t)) continue;
        const double by = t.b * qy;
        const auto keeps = [&](int c) { return cell_keeps(t, cell_f(t, oxd + (double)c, by)); };
        const auto interior = [&](int c) { return cell_interior(t, cell_f(t, oxd + (double)c, by)); };
        if (t.a != 0.0) {
            const double ek = -(by + t.c) * t.inv - oxd, ei = (t.m - by - t.c) * t.inv - oxd;
            const double mk = root_margin(t, by, oxd, side, ek), mi = root_margin(t, by, oxd, side, ei);
            if (t.a > 0.0) {
                r.lo = imax(r.lo, first_true(keeps, ek, side, mk));
                r.elo = imax(r.elo, first_true(interior, ei, side, mi));
            } else {
                r.hi = imin(r.hi, last_true(keeps, ek, side, mk));
                r.ehi = imin(r.ehi, last_true(interior, ei, side, mi));
            }
        } else {
            // a constant rule (a = 0, or not a number): the whole row or none of it
            if (!keeps(0)) {
                r.lo = side;
                r.hi = -1;
            }
            if (!interior(0)) {
                r.elo = side;
                r.ehi = -1;
            }
        }
    }
    return r;
}

// Row bounds packed for a shuffle: lo + 1, hi + 1, elo + 1, ehi + 1 in bytes 0..3.
SPX_DEV_INLINE uint32_t pack_row(const CellRow& r) {
    const auto b = [](int v) { return (uint32_t)(imin(imax(v, -1), 254) + 1); };
    return b(r.lo) | (b(r.hi) << 8) | (b(r.elo) << 16) | (b(r.ehi) << 24);
}

// The pixels the fit of feature f uses: those of its patch inside the image and, with a
// neighbour list, inside its cell (the exact rows above). An upper bound of their count
// chooses the feature's bucket before the fit. A seed too far out for an integer patch origin
// gets the whole patch as its count (also an upper bound: the fit keeps what its gather finds).
struct PatchSpan {
    int64_t ox, oy;          // patch origin
    int64_t c0, c1, r0, r1;  // columns and rows of the patch inside the image
    int side;
    bool cells;  // the feature has a listed neighbour
    bool whole;  // no integer origin: the count is side^2
};

SPX_DEV_INLINE PatchSpan patch_span(const FitArgs& a, int f) {
    PatchSpan p{};
    const int hs = a.half[f];
    p.side = 2 * hs + 1;
    const float sx = a.seeds[2 * (int64_t)f], sy = a.seeds[2 * (int64_t)f + 1];
    p.whole = !(fabsf(sx) < 1.0e15f) || !(fabsf(sy) < 1.0e15f);
    if (p.whole) return p;
    p.ox = (int64_t)rintf(sx) - hs;
    p.oy = (int64_t)rintf(sy) - hs;
    p.c0 = p.ox < 0 ? -p.ox : 0;
    p.r0 = p.oy < 0 ? -p.oy : 0;
    p.c1 = (a.W - 1 - p.ox) < (int64_t)(p.side - 1) ? (a.W - 1 - p.ox) : (int64_t)(p.side - 1);
    p.r1 = (a.H - 1 - p.oy) < (int64_t)(p.side - 1) ? (a.H - 1 - p.oy) : (int64_t)(p.side - 1);
    p.cells = false;
    if (a.nbr != nullptr)
        for (int j = 0; j < a.nb && j < kMaxNb; ++j) {
            const int id = a.nbr[(int64_t)f * a.nb + j];
            p.cells = p.cells || (id >= 0 && id < a.nfeat);
        }
    return p;
}

// Fitted pixels of patch row `row` (0 for a row outside the image). The cell's row is found
// for every row (the lane-spread terms need every lane of the wave).
template <class Terms>
SPX_DEV_INLINE int patch_row_count(const Terms& terms, const PatchSpan& p, int row) {
    int64_t lo = p.c0, hi = p.c1;
    if (p.cells) {
        const CellRow r = cell_row(terms, p.ox, p.oy, p.side, row < p.side ? row : p.side - 1);
        lo = r.lo > lo ? r.lo : lo;
        hi = r.hi < hi ? r.hi : hi;
    }
    if (row < p.r0 || row > p.r1) return 0;
    return hi >= lo ? (int)(hi - lo + 1) : 0;
}

// An upper bound of patch_row_count for the bucketing, from the real roots alone: each
// neighbour with |a| >= 1e-3 limits the row to its root's column widened by one column (the
// exact rule's end lies within a column of the root, whose rounding is far below a column);
// flatter rules are left out. About a tenth of the exact row's work.
template <class Terms>
SPX_DEV_INLINE int patch_row_upper(const Terms& terms, const PatchSpan& p, int row) {
#pragma clang fp contract(off)
    double lo = (double)p.c0, hi = (double)p.c1;
    if (p.cells) {
        const double qy = (double)(p.oy + (row < p.side ? row : p.side - 1));
        const double oxd = (double)p.ox;
#pragma unroll
        for (int j = 0; j < kMaxNb; ++j) {
            CellTerm t;
            if (!terms.get(j, t)) continue;
            if (!(fabs(t.a) >= 1e-3)) continue;
            const double root = -(t.b * qy + t.c) * t.inv - oxd;
            if (t.a > 0.0) lo = fmax(lo, floor(root) - 1.0);
            else hi = fmin(hi, ceil(root) + 1.0);
        }
    }
    if (row < p.r0 || row > p.r1 || !(hi >= lo)) return 0;
    return (int)(hi - lo + 1.0);
}

// The exact count by one thread (rows in turn).
SPX_DEV_INLINE int patch_pixel_count(const FitArgs& a, int f) {
    const PatchSpan p = patch_span(a, f);
    if (p.whole) return p.side * p.side;
    const MemTerms terms(a, f);
    int count = 0;
    for (int row = 0; row < p.side; ++row) count += patch_row_count(terms, p, row);
    return count;
}

// The bucketing's bound by one thread.
SPX_DEV_INLINE int patch_pixel_bound(const FitArgs& a, int f) {
    const PatchSpan p = patch_span(a, f);
    if (p.whole) return p.side * p.side;
    const MemTerms terms(a, f);
    int count = 0;
    for (int row = 0; row < p.side; ++row) count += patch_row_upper(terms, p, row);
    return count;
}

// Patch class of a pixel count: NPL = 4, 8, 16, 32, 64 words per lane.
SPX_DEV_INLINE int bucket_of_pixels(int npix) {
    const int npl = (npix + kWave - 1) / kWave;
    return npl <= 4 ? 0 : npl <= 8 ? 1 : npl <= 16 ? 2 : npl <= 32 ? 3 : 4;
}

// Work counters of one wave, flushed to FitArgs::stats once when the wave leaves its
// queue (one atomic per counter per wave instead of per feature).
struct WaveStats {
    unsigned long long evals;      // pixel passes x patch pixels
    unsigned long long passes;     // pixel passes
    unsigned long long steps;      // LM step attempts
    unsigned long long fallbacks;  // learned starts not converged within their budget
    unsigned long long overflow;   // features with more pixels than their bucket holds
};

// This lane's partial sums for the moment start, taken during the gather: grey-level sum and
// count of the ring pixels (the patch's outer ring and, with a cell, the cell's boundary) and
// the maximum over the fitted pixels.
struct StartSums {
    float bsum, bcnt, vmax;
};

// Residual judge of a budgeted fit (below): the fit is accepted when the Gauss-Newton model
// at its last point predicts that the minimum's cost lies within this fraction of the current
// cost, the FP32 rounding floor of the cost sums (kAcceptRtol), so its residual does not
// exceed the converged fit's by more than the sums can resolve.
constexpr double kJudgeRtol = 1.0e-6;

// Fit of one feature by the whole wave. The patch and the moment start's ring sums `ss` come
// from the caller's gather. `from_init`: start from the caller's parameters (a.init), else
// from the patch moments. `budget`: a learned-start pass (a.opt.learned_iters > 0) of at most
// learned_iters step attempts with the residual judge at the start point and after every
// accepted step: at the first point
// it accepts, its Gauss-Newton step is applied (as the plain rule applies its last
// sub-tolerance step) and the fit is converged; the step-size rule still ends a fit as well.
// A budgeted fit that is not converged then writes nothing and returns its attempt count: the
// caller fits the feature again from the moment start without a budget (the plain-start
// fallback), passing that count as `it_before`, so that iters_out counts the attempts of both.
// Returns -1 otherwise.
template <int NPL, int P>
SPX_DEV_INLINE int fit_feature(const FitArgs& a, int f, int hs, const Patch<NPL>& pt, const StartSums& ss,
                               float cx, float cy, bool from_init, bool budget, int it_before, WaveStats& ws) {
    using L = Layout<P>;
    const int lane = spx_lane();
    const float npix = pt.npix;
    Params<P> cur;
    uint8_t status = SPX_FIT_CONVERGED;
    int it = 0;
    int passes = 0;
    float cost_out = 0.0f;

    if (npix < (float)(P + 2)) {
        status = SPX_FIT_TOO_FEW;
        for (int i = 0; i < P; ++i) cur.p[i] = __builtin_nan("");
    } else {
        bool start_ok = true;
        if (from_init) {
            const float* in = a.init + (int64_t)f * 7;
            cur.p[0] = (double)spx_uniform(in[0]) - cx;
            cur.p[1] = (double)spx_uniform(in[1]) - cy;
            cur.p[2] = spx_uniform(in[2]);
            cur.p[3] = spx_uniform(in[3]);
            cur.p[4] = spx_uniform(in[4]);
            cur.p[P - 1] = spx_uniform(in[5]);
            if constexpr (P == 7) cur.p[5] = spx_uniform(in[6]);
        } else {
            // Moment start: background = mean of the patch's outer ring (with a cell mask:
            // also the cell's boundary pixels, those with a 4-neighbour outside the cell,
            // as the reference's _extract_voronoi_rois takes, subpx/_cpu/cells.py:72-80);
            // in-set = pixels above half the peak; centre = weighted centroid of the in-set;
            // radius from its area; amplitude = its mean; edge sigma = init_sigma; q = 1.
            // The ring's sum and count and the maximum were taken during the gather.
            const float bsum = spx_wave_sum(ss.bsum);
            const float bcnt = spx_wave_sum(ss.bcnt);
            const float B0 = bcnt > 0.0f ? bsum / bcnt : 0.0f;
            const float vmax = spx_wave_max(ss.vmax) - B0;
            const float half_peak = 0.5f * vmax;
            float s0 = 0.0f, sx = 0.0f, sy = 0.0f, nin = 0.0f;
#pragma unroll(Patch<NPL>::kLds ? 2 : NPL)
            for (int n = 0; n < NPL; ++n) {
                if constexpr (Patch<NPL>::kLds) {
                    if (n >= pt.nslots) break;
                }
                if (n < pt.nslots) {
                    const uint32_t w = pt.word(n);
                    const float wv = (float)(w >> 16) - B0;
                    if (((pt.valid >> n) & 1ull) && wv > half_peak) {
                        s0 += wv;
                        sx = fmaf(wv, (float)((int)(w & 0xffu) - hs), sx);
                        sy = fmaf(wv, (float)((int)((w >> 8) & 0xffu) - hs), sy);
                        nin += 1.0f;
                    }
                }
            }
            s0 = spx_wave_sum(s0);
            sx = spx_wave_sum(sx);
            sy = spx_wave_sum(sy);
            nin = spx_wave_sum(nin);
            if (!(vmax > 0.0f) || !(s0 > 0.0f) || nin < 1.0f) {
                start_ok = false;
            } else {
                cur.p[0] = (double)(sx / s0);
                cur.p[1] = (double)(sy / s0);
                cur.p[2] = (double)sqrtf(nin * 0.31830988618379067f);
                cur.p[3] = (double)a.opt.init_sigma;
                cur.p[4] = (double)(s0 / nin);
                cur.p[P - 1] = (double)B0;
                if constexpr (P == 7) cur.p[5] = 1.0;
            }
        }
        if (!start_ok) {
            status = SPX_FIT_BAD_START;
            for (int i = 0; i < P; ++i) cur.p[i] = __builtin_nan("");
        } else {
            float S[L::kN];
            {
                float acc[L::kN], sc[L::kNB];
                pixel_pass<NPL, P>(pt, hs, cur, acc, sc);
                S[L::kCost] = spx_wave_sum(acc[L::kCost]);
                reduce_sums<P>(acc, sc, S);
            }
            ++passes;
            float cost = S[L::kCost];
            double lambda = (double)a.opt.lambda0;
            const double lim = (double)hs + 0.5;
            status = SPX_FIT_MAX_ITER;
            if (!isfinite(cost)) status = SPX_FIT_DIVERGED;
            bool any_solve = false;
            const int max_iter = budget ? a.opt.learned_iters : a.opt.max_iter;
            bool fresh = budget;  // S belongs to a point the judge has not seen yet
            // One solve per turn: a judging turn (fresh sums of a budgeted fit) solves the
            // undamped system for the judge and counts no attempt; the others take an LM step.
            while (status == SPX_FIT_MAX_ITER && (it < max_iter || fresh)) {
                const bool judging = fresh;
                fresh = false;
                if (!judging) ++it;
                double delta[P];
                if (!damped_solve<P>(S, (double)npix, judging ? 0.0 : lambda, delta)) {
                    if (!judging) lambda *= (double)a.opt.lambda_up;
                    continue;
                }
                Params<P> tr;
#pragma unroll
                for (int i = 0; i < P; ++i) tr.p[i] = spx_uniform_d(cur.p[i] + delta[i]);
                bool inside = fabs(tr.p[0]) <= lim && fabs(tr.p[1]) <= lim && tr.p[2] > 0.25 &&
                              tr.p[2] < 2.0 * lim && tr.p[3] > 0.05 && tr.p[3] < lim;
                if constexpr (P == 7) inside = inside && tr.p[5] > 0.25 && tr.p[5] < 4.0;
                if (judging) {
                    // residual judge: the Gauss-Newton step predicts the cost decrease
                    // delta . J^T r; at most kJudgeRtol of the cost -> apply it and stop
                    double gain = 0.0;
#pragma unroll
                    for (int i = 0; i < P; ++i) gain = fma(delta[i], (double)S[L::kRhs + i], gain);
                    gain = spx_uniform_d(gain);
                    if (inside && gain >= 0.0 && gain <= kJudgeRtol * (double)cost) {
                        cur = tr;
                        status = SPX_FIT_CONVERGED;
                        break;
                    }
                    continue;
                }
                any_solve = true;
                double step = fmax(fmax(fabs(delta[0]), fabs(delta[1])),
                                   fmax(fabs(delta[2]), fabs(delta[3])));
                if constexpr (P == 7) step = fmax(step, fabs(delta[5]) * fabs(cur.p[2]));
                step = spx_uniform_d(step);
                if (!inside) {
                    lambda *= (double)a.opt.lambda_up;
                    continue;
                }
                // A step below the tolerance ends the fit and is applied without a trial
                // pass: taken or refused, it moves the result by less than step_tol, so the
                // pass would decide nothing (the rms stays that of the point before it).
                if (step < (double)a.opt.step_tol) {
                    cur = tr;
                    status = SPX_FIT_CONVERGED;
                    break;
                }
                float acc[L::kN], sc[L::kNB];
                pixel_pass<NPL, P>(pt, hs, tr, acc, sc);
                ++passes;
                const float cost_t = spx_wave_sum(acc[L::kCost]);
                // A trial within the FP32 rounding floor of the cost (1e-6 relative) is
                // taken: near the minimum the Gauss-Newton step is accurate while cost
                // differences drown in rounding. Only an accepted trial's sums are reduced.
                if (cost_t <= cost * (1.0f + kAcceptRtol)) {
                    cur = tr;
                    reduce_sums<P>(acc, sc, S);
                    S[L::kCost] = cost_t;
                    cost = cost_t;
                    lambda = fmax(lambda * (double)a.opt.lambda_down, 1e-12);
                    fresh = budget;
                } else {
                    lambda *= (double)a.opt.lambda_up;
                }
            }
            if (!any_solve && status == SPX_FIT_MAX_ITER) status = SPX_FIT_SINGULAR;
            if (!isfinite(cost)) status = SPX_FIT_DIVERGED;
            if (status <= SPX_FIT_MAX_ITER && (fabs(cur.p[0]) > (double)hs || fabs(cur.p[1]) > (double)hs))
                status = SPX_FIT_OUTSIDE;
            cost_out = cost;
        }
    }

    ws.evals += (unsigned long long)passes * (unsigned long long)npix;
    ws.passes += (unsigned long long)passes;
    ws.steps += (unsigned long long)it;
    if (budget && status != SPX_FIT_CONVERGED && status != SPX_FIT_TOO_FEW) {
        ++ws.fallbacks;
        return it;
    }
    it += it_before;

    // lanes 0..6 write one parameter each, lane 7 the fit summary
    float outv = 0.0f;
#pragma unroll
    for (int i = 0; i < 7; ++i) {
        float pv;
        if (i == 0) pv = (float)(cur.p[0] + cx);
        else if (i == 1) pv = (float)(cur.p[1] + cy);
        else if (i == 5) pv = (float)cur.p[P - 1];
        else if (i == 6) pv = (P == 7) ? (float)cur.p[5] : 1.0f;
        else pv = (float)cur.p[i];
        outv = (lane == i) ? pv : outv;
    }
    if (lane < 7) a.params[(int64_t)f * 7 + lane] = outv;
    // The FP64 centre: an FP32 coordinate near 16384 px rounds by up to 1e-3 px.
    if (a.centers != nullptr && (lane == 8 || lane == 9))
        a.centers[(int64_t)f * 2 + (lane - 8)] = lane == 8 ? cur.p[0] + (double)cx : cur.p[1] + (double)cy;
    if (lane == 7) {
        const float dof = fmaxf(npix - (float)P, 1.0f);
        a.rms[f] = (status == SPX_FIT_TOO_FEW || status == SPX_FIT_BAD_START) ? __builtin_nanf("")
                                                                              : sqrtf(cost_out / dof);
        a.iters[f] = (int16_t)it;
        a.status[f] = status;
    }
    return -1;
}

// Body of the persistent fit kernel for one wave: pull features of one (patch class,
// model) bucket from the work queue until it is empty. lds_words: the wave's NPL x 64-word
// LDS slab, where every patch is compacted.
template <int NPL, int P>
SPX_DEV void erf_fit_wave(const FitArgs& a, uint32_t* lds_words) {
    const int lane = spx_lane();
    const int count = spx_uniform_i(*a.count);
    if (count == 0) return;
    const int32_t* order = a.order;
    if (a.bucket_counts != nullptr) {
        int off = 0;
        for (int j = 0; j < a.bucket; ++j) off += a.bucket_counts[j];
        order += spx_uniform_i(off);
    }
    WaveStats ws{0ull, 0ull, 0ull, 0ull, 0ull};
    // Features are claimed kGrab at a time: one queue atomic per kGrab features.
    int base = 0, left = 0;
    // A feature whose budgeted fit needs the fallback is gathered again and fitted from the
    // moment start in the next turn of this loop (from the L2 lines it has just read).
    int redo = -1, redo_iters = 0;
    for (;;) {
        int f;
        const bool from_init = a.init != nullptr && redo < 0;
        const bool budget = a.opt.learned_iters > 0 && redo < 0;
        const int it_before = redo < 0 ? 0 : redo_iters;
        if (redo >= 0) {
            f = redo;
            redo = -1;
        } else {
            if (left == 0) {
                base = spx_queue_grab(a.queue, kGrab);
                left = kGrab;
            }
            const int slot = base + (kGrab - left);
            --left;
            if (slot >= count) break;
            f = spx_uniform_i(order[slot]);
        }
        const int hs = spx_uniform_i(a.half[f]);
        const int side = 2 * hs + 1;
        const int npx = side * side;
        const float cx = spx_uniform(rintf(a.seeds[2 * (int64_t)f]));
        const float cy = spx_uniform(rintf(a.seeds[2 * (int64_t)f + 1]));
        const int64_t ox = (int64_t)cx - hs;
        const int64_t oy = (int64_t)cy - hs;

        bool masked = false;
        if (a.nbr != nullptr)
            for (int j = 0; j < a.nb; ++j) {
                const int id = spx_uniform_i(a.nbr[(int64_t)f * a.nb + j]);
                masked = masked || (id >= 0 && id < a.nfeat);
            }
        // Gather and compaction: the wave walks the patch 64 pixels at a time, lane-strided
        // over the flattened patch, so each load instruction reads a few contiguous row
        // segments of the image. The loads of kBatch steps are issued back to back before
        // their pixels are tested (every lane loads from a clamped in-image address, so they
        // carry no branch). The pixels inside the image (and the cell) are appended to the LDS
        // slab in patch order: kept pixel k is word k, slot k / 64 of lane k % 64. Row and
        // column advance by the uniform quotient and remainder of 64 / side instead of a
        // division per step. With a cell, lane r first finds row r's member and interior
        // columns (cell_row); a pixel then reads its row's bounds by a shuffle. The moment
        // start's ring sums and maximum are taken on the way.
        constexpr int kCap = NPL * kWave;
        constexpr int kBatch = 8;
        Patch<NPL> pt;
        pt.lds = lds_words;
        StartSums ss{0.0f, 0.0f, -3.0e38f};
        int kept = 0;  // uniform
        {
            uint32_t rowb = 0u;
            if (masked) rowb = pack_row(cell_row(lane_terms(a, f), ox, oy, side, lane < side ? lane : side - 1));
            spx_wave_lds_sync();  // every lane is done with the previous feature's words
            const int raw = (npx + kWave - 1) / kWave;
            const int dq = kWave / side, dr = kWave - dq * side;
            int row = lane / side;
            int col = lane - row * side;
            const int64_t W1 = a.W - 1, H1 = a.H - 1;
            for (int n0 = 0; n0 < raw; n0 += kBatch) {
                uint32_t v[kBatch];
                int cs[kBatch], rs[kBatch];
                bool inside[kBatch];
#pragma unroll
                for (int j = 0; j < kBatch; ++j) {
                    const int pix = lane + kWave * (n0 + j);
                    const bool in_patch = pix < npx;
                    // lanes past the patch's end read its first row (lines the wave fetches anyway)
                    const int64_t gx = ox + col, gy = oy + (in_patch ? row : 0);
                    inside[j] = in_patch && gx >= 0 && gx <= W1 && gy >= 0 && gy <= H1;
                    const int64_t qx = gx < 0 ? 0 : (gx > W1 ? W1 : gx);
                    const int64_t qy = gy < 0 ? 0 : (gy > H1 ? H1 : gy);
                    v[j] = (uint32_t)a.img[qy * a.pitch + qx];
                    cs[j] = col;
                    rs[j] = row;
                    col += dr;
                    row += dq;
                    if (col >= side) {
                        col -= side;
                        ++row;
                    }
                }
#pragma unroll
                for (int j = 0; j < kBatch; ++j) {
                    if (n0 + j >= raw) break;
                    bool ok = inside[j];
                    bool ring = rs[j] == 0 || rs[j] == side - 1 || cs[j] == 0 || cs[j] == side - 1;
                    if (masked) {
                        // bytes: lo + 1, hi + 1, elo + 1, ehi + 1 of this pixel's row
                        const uint32_t rb = (uint32_t)spx_shfl(rowb, rs[j] < kWave ? rs[j] : kWave - 1);
                        const int c1 = cs[j] + 1;
                        ok = ok && c1 >= (int)(rb & 0xffu) && c1 <= (int)((rb >> 8) & 0xffu);
                        ring = ring || c1 < (int)((rb >> 16) & 0xffu) || c1 > (int)(rb >> 24);
                    }
                    if (!from_init && ok) {
                        const float val = (float)v[j];
                        ss.vmax = fmaxf(ss.vmax, val);
                        if (ring) {
                            ss.bsum += val;
                            ss.bcnt += 1.0f;
                        }
                    }
                    const uint64_t keep = spx_ballot(ok);
                    const int dst = kept + spx_lanes_below(keep);
                    if (ok && dst < kCap) lds_words[dst] = pack_px(cs[j], rs[j], v[j]);
                    kept += __builtin_popcountll(keep);
                }
            }
        }
        if (kept > kCap) {  // the bucket's bound was too small: never by construction
            ++ws.overflow;
            kept = kCap;
        }
        pt.nslots = (kept + kWave - 1) / kWave;
        // the last slot's empty words are zeroed (its valid bits exclude them as well)
        if (lane < pt.nslots * kWave - kept) lds_words[kept + lane] = 0u;
        spx_wave_lds_sync();
        {
            const int full = kept / kWave;
            const bool part = full < kWave && lane < kept - full * kWave;
            pt.valid = (full >= kWave ? ~0ull : ((1ull << full) - 1ull)) | (part ? (1ull << full) : 0ull);
        }
        if constexpr (!Patch<NPL>::kLds) {
#pragma unroll
            for (int n = 0; n < NPL; ++n) pt.px[n] = lds_words[n * kWave + lane];
        }
        pt.npix = (float)kept;
        const int again = fit_feature<NPL, P>(a, f, hs, pt, ss, cx, cy, from_init, budget, it_before, ws);
        if (again >= 0) {
            redo = f;
            redo_iters = again;
        }
    }
    if (lane == 8) spx_atomic_add(&a.stats[P == 7 ? 1 : 0], ws.evals);
    if (lane == 9) spx_atomic_add(&a.stats[2], ws.passes);
    if (lane == 10) spx_atomic_add(&a.stats[3], ws.steps);
    if (lane == 11 && ws.fallbacks != 0ull) spx_atomic_add(&a.stats[4], ws.fallbacks);
    if (lane == 12 && ws.overflow != 0ull) spx_atomic_add(&a.stats[5], ws.overflow);
}

}  // namespace fitdev
}  // namespace spx
