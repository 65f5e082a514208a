// Device-side logic of the batched erf-disk / erf-ellipse LM fit (one wave64 per feature).
//
// Included by erf_fit.hip (HIP build) after it defines the wave primitives:
//   SPX_DEV, SPX_DEV_INLINE, spx_lane(), spx_shfl_xor(float, int),
//   spx_uniform(float) / spx_uniform_i(int) (the first lane's value, made wave-uniform),
//   spx_queue_grab(int32_t*) (one atomic increment per wave, result uniform),
//   spx_atomic_add(unsigned long long*, unsigned long long),
//   hw_rcp / hw_rsq / hw_exp2 (float) and hw_rcp64 (double).
// Every value that steers control flow is wave-uniform by construction (reductions and
// the queue head pass through spx_uniform / readfirstlane), so the compiler sees uniform
// branches only: a lane-dependent branch around a shuffle inside the persistent loop
// lets it give lanes different loop headers, and the wave then never reconverges.
// The same text compiles against a host emulation of a wave (64 threads exchanging through
// barriers) in the tests, which checks on the CPU that every lane takes the same path.
//
// Model (patch-local coordinates, pixel centres at integers):
//   I(x, y) = B + A * E(u),  E(u) = erfc(u) / 2,  u = (rho - r) / (sqrt(2) * s)
//   rho = sqrt(dx^2 + dy^2) (disk, 6 parameters) or sqrt(dx^2 + (dy/q)^2) (ellipse, 7th q)
//   dx = x - x0, dy = y - y0.
// The reference has no erf-disk fit; the batched LM layout follows its wave-per-cell
// Gaussian LM kernel `lm_fit` (reference subpx/_gpu/deblend.py:55-152: butterfly totals
// :32-39/:107, LU solve :113-130, accept/reject :141-149) and the disk LM `_fit_disks`
// (reference subpx/_cpu/invariant_cd.py:165-246), redesigned for CDNA4:
//   * the patch lives in VGPRs (NPL pixels per lane), gathered once from HBM;
//   * residual and Jacobian are FP32 on the VALU, with hardware rsq / rcp / exp2;
//   * J^T J, J^T r and the cost are reduced by a wave butterfly, no atomics, so a result
//     does not depend on scheduling;
//   * the damped normal equations are solved in FP64 (LDL^T);
//   * the Jacobian pass at a trial point doubles as the trial cost evaluation, so an
//     accepted step costs one pixel pass;
//   * waves pull features from an atomic work queue per patch-size bucket (persistent).
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
constexpr float kFar = 1.0e4f;  // coordinate given to empty slots: E = 0, G = 0 exactly there
constexpr float kAcceptRtol = 1.0e-6f;

SPX_DEV_INLINE float wave_sum(float v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += spx_shfl_xor(v, off);
    return spx_uniform(v);
}

SPX_DEV_INLINE float wave_max(float v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fmaxf(v, spx_shfl_xor(v, off));
    return spx_uniform(v);
}

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

// One pixel pass at parameters `pp`: accumulates J^T J, J^T r and the cost over the
// patch held in registers, then reduces over the wave. Every lane returns the totals.
template <int NPL, int P>
SPX_DEV_INLINE void pixel_pass(const float (&v)[NPL], const float (&fx)[NPL],
                                           const float (&fy)[NPL], const float (&m)[NPL],
                                           const Params<P>& pp, float (&S)[Layout<P>::kN]) {
    using L = Layout<P>;
    constexpr bool kEll = (P == 7);
    const float x0 = (float)pp.p[0], y0 = (float)pp.p[1], r = (float)pp.p[2];
    const float s = (float)pp.p[3], A = (float)pp.p[4], B = (float)pp.p[P - 1];
    const float q = kEll ? (float)pp.p[5] : 1.0f;
    const float inv_s = 1.0f / s;
    const float k = kInvSqrt2 * inv_s;
    const float rk = r * k;
    const float Ac = A * kInvSqrt2Pi * inv_s;
    const float iq = 1.0f / q;
    const float iq2 = iq * iq;

#pragma unroll
    for (int i = 0; i < L::kN; ++i) S[i] = 0.0f;

#pragma unroll
    for (int n = 0; n < NPL; ++n) {
        const float dx = fx[n] - x0;
        const float dy = fy[n] - y0;
        const float dyq = kEll ? dy * iq2 : dy;
        float rho2 = fmaf(dy, dyq, dx * dx);
        rho2 = fmaxf(rho2, 1e-20f);
        const float rs = hw_rsq(rho2);
        const float rho = rho2 * rs;
        const float u = fmaf(rho, k, -rk);
        const float t = hw_rcp(fmaf(kAsP, fabsf(u), 1.0f));
        const float ez = hw_exp2(-(u * u) * kLog2e);
        const float poly = t * fmaf(t, fmaf(t, fmaf(t, fmaf(t, kAs5, kAs4), kAs3), kAs2), kAs1);
        const float hq = 0.5f * poly * ez;
        const float E = (u >= 0.0f) ? hq : 1.0f - hq;
        const float res = (v[n] - fmaf(A, E, B)) * m[n];
        const float G = Ac * ez;
        const float Gr = G * rs;
        float J[P - 1];
        J[0] = Gr * dx;
        J[1] = Gr * dyq;
        J[2] = G;
        J[3] = G * u * kSqrt2;
        J[4] = E;
        if constexpr (kEll) J[5] = Gr * dy * dyq * iq;
        int idx = 0;
#pragma unroll
        for (int a = 0; a < L::kNB; ++a)
#pragma unroll
            for (int b = a; b < L::kNB; ++b) {
                S[idx] = fmaf(J[a], J[b], S[idx]);
                ++idx;
            }
#pragma unroll
        for (int a = 0; a < L::kNB; ++a) S[L::kColB + a] += J[a];
#pragma unroll
        for (int a = 0; a < L::kNB; ++a) S[L::kRhs + a] = fmaf(J[a], res, S[L::kRhs + a]);
        S[L::kRhs + P - 1] += res;
        S[L::kCost] = fmaf(res, res, S[L::kCost]);
    }
#pragma unroll
    for (int i = 0; i < L::kN; ++i) S[i] = wave_sum(S[i]);
}

// Solves (JtJ + lambda * diag) delta = J^T r in FP64 by LDL^T. Returns false when a pivot
// is not positive.
template <int P>
SPX_DEV_INLINE bool damped_solve(const float (&S)[Layout<P>::kN], double npix,
                                             double lambda, double (&delta)[P]) {
    using L = Layout<P>;
    double M[P][P];
    double g[P];
    int idx = 0;
#pragma unroll
    for (int a = 0; a < L::kNB; ++a)
#pragma unroll
        for (int b = a; b < L::kNB; ++b) {
            M[a][b] = (double)S[idx];
            M[b][a] = (double)S[idx];
            ++idx;
        }
#pragma unroll
    for (int a = 0; a < L::kNB; ++a) {
        M[a][P - 1] = (double)S[L::kColB + a];
        M[P - 1][a] = (double)S[L::kColB + a];
    }
    M[P - 1][P - 1] = npix;
#pragma unroll
    for (int a = 0; a < P; ++a) g[a] = (double)S[L::kRhs + a];
    double dmax = 0.0;
#pragma unroll
    for (int a = 0; a < P; ++a) dmax = fmax(dmax, M[a][a]);
    const double floor_d = 1e-12 * dmax;
#pragma unroll
    for (int a = 0; a < P; ++a) M[a][a] += lambda * fmax(M[a][a], floor_d);

    double Lm[P][P];
    double D[P];
#pragma unroll
    for (int j = 0; j < P; ++j) {
        double d = M[j][j];
#pragma unroll
        for (int kk = 0; kk < j; ++kk) d -= Lm[j][kk] * Lm[j][kk] * D[kk];
        if (!(d > 1e-13 * M[j][j]) || !(d > 0.0)) return false;
        D[j] = d;
        const double inv = hw_rcp64(d);
#pragma unroll
        for (int i = j + 1; i < P; ++i) {
            double sacc = M[i][j];
#pragma unroll
            for (int kk = 0; kk < j; ++kk) sacc -= Lm[i][kk] * Lm[j][kk] * D[kk];
            Lm[i][j] = sacc * inv;
        }
    }
    double y[P];
#pragma unroll
    for (int i = 0; i < P; ++i) {
        double sacc = g[i];
#pragma unroll
        for (int kk = 0; kk < i; ++kk) sacc -= Lm[i][kk] * y[kk];
        y[i] = sacc;
    }
#pragma unroll
    for (int i = P - 1; i >= 0; --i) {
        double sacc = y[i] * hw_rcp64(D[i]);
#pragma unroll
        for (int kk = i + 1; kk < P; ++kk) sacc -= Lm[kk][i] * delta[kk];
        delta[i] = sacc;
    }
#pragma unroll
    for (int i = 0; i < P; ++i)
        if (!isfinite(delta[i])) return false;
    return true;
}

struct FitArgs {
    const uint16_t* img;
    int64_t H, W, pitch;
    const float* seeds;
    const int16_t* half;
    const uint8_t* model;
    const float* init;
    SpxFitOptions opt;
    const int32_t* order;  // feature ids of this bucket
    const int32_t* count;  // number of ids in this bucket (device)
    int32_t* queue;        // work-queue head of this bucket (device, starts at 0)
    unsigned long long* stats;  // [pixel evaluations P=6, P=7, pixel passes, solves]
    float* params;
    float* rms;
    int16_t* iters;
    uint8_t* status;
};

// Fit of one feature by the whole wave. Patch registers are filled by the caller.
template <int NPL, int P>
SPX_DEV void fit_feature(const FitArgs& a, int f, int hs, const float (&v)[NPL],
                            const float (&fx)[NPL], const float (&fy)[NPL],
                            const float (&m)[NPL], float npix, float cx, float cy) {
    using L = Layout<P>;
    const int lane = spx_lane();
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
        if (a.init != nullptr) {
            const float* in = a.init + (int64_t)f * 7;
            cur.p[0] = (double)spx_uniform(in[0]) - cx;
            cur.p[1] = (double)spx_uniform(in[1]) - cy;
            cur.p[2] = spx_uniform(in[2]);
            cur.p[3] = spx_uniform(in[3]);
            cur.p[4] = spx_uniform(in[4]);
            cur.p[P - 1] = spx_uniform(in[5]);
            if constexpr (P == 7) cur.p[5] = spx_uniform(in[6]);
        } else {
            // Moment start: background = mean of the patch's outer ring; in-set = pixels
            // above half the peak; centre = weighted centroid of the in-set; radius from
            // its area; amplitude = its mean; edge sigma = init_sigma; q = 1.
            float bsum = 0.0f, bcnt = 0.0f;
            const int side = 2 * hs + 1;
#pragma unroll
            for (int n = 0; n < NPL; ++n) {
                const int pix = lane + kWave * n;
                const int row = pix / side, col = pix - (pix / side) * side;
                const bool ring = (row == 0 || row == side - 1 || col == 0 || col == side - 1);
                if (ring && m[n] > 0.0f) {
                    bsum += v[n];
                    bcnt += 1.0f;
                }
            }
            bsum = wave_sum(bsum);
            bcnt = wave_sum(bcnt);
            const float B0 = bcnt > 0.0f ? bsum / bcnt : 0.0f;
            float vmax = -3.0e38f;
#pragma unroll
            for (int n = 0; n < NPL; ++n)
                if (m[n] > 0.0f) vmax = fmaxf(vmax, v[n] - B0);
            vmax = wave_max(vmax);
            const float half_peak = 0.5f * vmax;
            float s0 = 0.0f, sx = 0.0f, sy = 0.0f, nin = 0.0f;
#pragma unroll
            for (int n = 0; n < NPL; ++n) {
                const float w = v[n] - B0;
                if (m[n] > 0.0f && w > half_peak) {
                    s0 += w;
                    sx = fmaf(w, fx[n], sx);
                    sy = fmaf(w, fy[n], sy);
                    nin += 1.0f;
                }
            }
            s0 = wave_sum(s0);
            sx = wave_sum(sx);
            sy = wave_sum(sy);
            nin = wave_sum(nin);
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
            pixel_pass<NPL, P>(v, fx, fy, m, cur, S);
            ++passes;
            float cost = S[L::kCost];
            double lambda = (double)a.opt.lambda0;
            const double lim = (double)hs + 0.5;
            status = SPX_FIT_MAX_ITER;
            if (!isfinite(cost)) status = SPX_FIT_DIVERGED;
            bool any_solve = false;
            while (status == SPX_FIT_MAX_ITER && it < a.opt.max_iter) {
                ++it;
                double delta[P];
                if (!damped_solve<P>(S, (double)npix, lambda, delta)) {
                    lambda *= (double)a.opt.lambda_up;
                    continue;
                }
                any_solve = true;
                Params<P> tr;
#pragma unroll
                for (int i = 0; i < P; ++i) tr.p[i] = cur.p[i] + delta[i];
                double step = fmax(fmax(fabs(delta[0]), fabs(delta[1])),
                                   fmax(fabs(delta[2]), fabs(delta[3])));
                if constexpr (P == 7) step = fmax(step, fabs(delta[5]) * fabs(cur.p[2]));
                bool inside = fabs(tr.p[0]) <= lim && fabs(tr.p[1]) <= lim && tr.p[2] > 0.25 &&
                              tr.p[3] > 0.05 && tr.p[3] < (double)(2 * hs + 1);
                if constexpr (P == 7) inside = inside && tr.p[5] > 0.25 && tr.p[5] < 4.0;
                if (!inside) {
                    lambda *= (double)a.opt.lambda_up;
                    continue;
                }
                float St[L::kN];
                pixel_pass<NPL, P>(v, fx, fy, m, tr, St);
                ++passes;
                const float cost_t = St[L::kCost];
                // A trial within the FP32 rounding floor of the cost (1e-6 relative) is
                // taken: near the minimum the Gauss-Newton step is accurate while cost
                // differences drown in rounding.
                if (cost_t <= cost * (1.0f + kAcceptRtol)) {
                    cur = tr;
#pragma unroll
                    for (int i = 0; i < L::kN; ++i) S[i] = St[i];
                    cost = cost_t;
                    lambda = fmax(lambda * (double)a.opt.lambda_down, 1e-12);
                    if (step < (double)a.opt.step_tol) status = SPX_FIT_CONVERGED;
                } else {
                    lambda *= (double)a.opt.lambda_up;
                    if (step < (double)a.opt.step_tol) status = SPX_FIT_CONVERGED;
                }
            }
            if (!any_solve && status == SPX_FIT_MAX_ITER) status = SPX_FIT_SINGULAR;
            if (!isfinite(cost)) status = SPX_FIT_DIVERGED;
            if (status <= SPX_FIT_MAX_ITER && (fabs(cur.p[0]) > (double)hs || fabs(cur.p[1]) > (double)hs))
                status = SPX_FIT_OUTSIDE;
            cost_out = cost;
        }
    }

    if (lane == 0) {
        float* out = a.params + (int64_t)f * 7;
        out[0] = (float)(cur.p[0] + cx);
        out[1] = (float)(cur.p[1] + cy);
        out[2] = (float)cur.p[2];
        out[3] = (float)cur.p[3];
        out[4] = (float)cur.p[4];
        out[5] = (float)cur.p[P - 1];
        out[6] = (P == 7) ? (float)cur.p[5] : 1.0f;
        const float dof = fmaxf(npix - (float)P, 1.0f);
        a.rms[f] = (status == SPX_FIT_TOO_FEW || status == SPX_FIT_BAD_START) ? __builtin_nanf("")
                                                                              : sqrtf(cost_out / dof);
        a.iters[f] = (int16_t)it;
        a.status[f] = status;
        spx_atomic_add(&a.stats[P == 7 ? 1 : 0], (unsigned long long)passes * (unsigned long long)npix);
        spx_atomic_add(&a.stats[2], (unsigned long long)passes);
        spx_atomic_add(&a.stats[3], (unsigned long long)it);
    }
}

// Body of the persistent fit kernel for one wave: pull features of one patch-size bucket
// from the work queue until it is empty.
template <int NPL>
SPX_DEV void erf_fit_wave(const FitArgs& a) {
    const int lane = spx_lane();
    const int count = *a.count;
    for (;;) {
        const int slot = spx_queue_grab(a.queue);
        if (slot >= count) break;
        const int f = spx_uniform_i(a.order[slot]);
        const int hs = spx_uniform_i(a.half[f]);
        const int model = spx_uniform_i(a.model[f]);
        const int side = 2 * hs + 1;
        const int npx = side * side;
        const float cx = spx_uniform(rintf(a.seeds[2 * (int64_t)f]));
        const float cy = spx_uniform(rintf(a.seeds[2 * (int64_t)f + 1]));
        const int64_t ox = (int64_t)cx - hs;
        const int64_t oy = (int64_t)cy - hs;

        // Gather: lane-strided over the flattened patch, so each load instruction reads a
        // few contiguous row segments of the image.
        float v[NPL], fx[NPL], fy[NPL], m[NPL];
        float cnt = 0.0f;
#pragma unroll
        for (int n = 0; n < NPL; ++n) {
            const int pix = lane + kWave * n;
            const int row = pix / side;
            const int col = pix - row * side;
            const int64_t gx = ox + col;
            const int64_t gy = oy + row;
            const bool ok = pix < npx && gx >= 0 && gx < a.W && gy >= 0 && gy < a.H;
            v[n] = ok ? (float)a.img[gy * a.pitch + gx] : 0.0f;
            fx[n] = ok ? (float)(col - hs) : kFar;
            fy[n] = ok ? (float)(row - hs) : kFar;
            m[n] = ok ? 1.0f : 0.0f;
            cnt += m[n];
        }
        const float npix = wave_sum(cnt);
        if (model == SPX_MODEL_ERF_ELLIPSE7)
            fit_feature<NPL, 7>(a, f, hs, v, fx, fy, m, npix, cx, cy);
        else
            fit_feature<NPL, 6>(a, f, hs, v, fx, fy, m, npix, cx, cy);
    }
}


constexpr int kBuckets = 4;
constexpr int kMaxHalf = 22;

// Patch-size bucket of a half side: NPL = 4, 8, 16, 32 pixels per lane.
SPX_DEV_INLINE int bucket_of(int hs) {
    const int npx = (2 * hs + 1) * (2 * hs + 1);
    const int npl = (npx + kWave - 1) / kWave;
    return npl <= 4 ? 0 : npl <= 8 ? 1 : npl <= 16 ? 2 : 3;
}

}  // namespace fitdev
}  // namespace spx
