// Window log-quadratic centre refinement (the GPU backend's "logquad" method), one wave64
// per feature, FP64.
//
// Restates the reference's refine_centers_logquad_gpu_match_cpu (subpx/_gpu/centers.py:288-517)
// with _masked_median_lower (:145-153) and _plane_fit_batched (:264-285). Per feature:
//   outer window (2R+1)^2 around rint(centroid), R = win_rad + bg_win_pad, zero / label 0
//   outside the image; ROI = bbox +- pad clipped to the image;
//   background = lower median of ROI pixels outside the blob dilated by bg_dilate (clipped
//   to the window), then a tail re-median (g - bg0 <= bg_tail_frac * max over the fit window);
//   plane a x + b y + c by 3x3 normal equations in image coordinates over those pixels;
//   fit window (2 win_rad + 1)^2: blob & ROI dilated by fit_dilate, v = g - background > 0,
//   core = v >= core_frac * max v (if >= 6 pixels), weights v^p clipped at the order
//   statistic floor(0.99 (n - 1)); weighted least squares of log v on [xn^2, yn^2, xn yn,
//   xn, yn, 1] in normalised coordinates; negative-definite, cond and det gates; the
//   centre must lie within +-win_rad of the float centroid, else ok = 0 and the centroid.
// The reference runs this as ~100 batched CuPy launches with sorts; here the window lives
// in LDS, the two medians and the percentile are exact order statistics found by rank
// counting (every candidate counts the values below and at it), and the 3x3 / 6x6 systems
// are solved by Gaussian elimination with partial pivoting in every lane.
#include "spx_internal.hpp"
#include "wave64.hpp"

namespace spx {
namespace {

constexpr int kMaxR = 10;
constexpr int kMaxP = 2 * kMaxR + 1;
constexpr int kMaxN = kMaxP * kMaxP;
constexpr int kWaves = 4;

constexpr uint8_t kMatch = 1, kRoi = 2;        // outer-window flags
constexpr uint8_t kBg = 1, kBg2 = 2;           // background-set flags
constexpr uint8_t kFit = 1, kSel = 2;          // fit-window flags

struct Lqw {
    int R, P, N, win_rad, off, Pw, Nw;
    int bg_dilate, fit_dilate, min_bg_pts, min_plane_pts, plane_bg;
    double core_frac, weight_power, bg_tail_frac, max_cond, negdef_eps, eps;
};

// k-th smallest (0-based) of the values whose flag has `bit`; n = their count (> 0).
template <typename V>
__device__ double kth_flagged(const V* val, const uint8_t* flag, uint8_t bit, int N, int k, int lane) {
    double found = -__builtin_inf();
    for (int i = lane; i < N; i += 64) {
        if (!(flag[i] & bit)) continue;
        const V vi = val[i];
        int lt = 0, le = 0;
        for (int j = 0; j < N; ++j) {
            if (!(flag[j] & bit)) continue;
            const V vj = val[j];
            lt += vj < vi;
            le += vj <= vi;
        }
        if (lt <= k && k < le) found = (double)vi;
    }
    return wave::max_d(found);
}

// Gaussian elimination with partial pivoting on an n x n system (n <= 6), in place; returns
// the determinant (product of pivots with the row-swap sign). x receives the solution.
template <int n>
__device__ double ge_solve(double A[n][n], double b[n], double x[n]) {
    double det = 1.0;
#pragma unroll
    for (int c = 0; c < n; ++c) {
        int p = c;
        double pv = fabs(A[c][c]);
#pragma unroll
        for (int r = c + 1; r < n; ++r)
            if (fabs(A[r][c]) > pv) {
                pv = fabs(A[r][c]);
                p = r;
            }
        if (p != c) {
            det = -det;
#pragma unroll
            for (int j = 0; j < n; ++j) {
                const double t = A[c][j];
                A[c][j] = A[p][j];
                A[p][j] = t;
            }
            const double t = b[c];
            b[c] = b[p];
            b[p] = t;
        }
        det *= A[c][c];
#pragma unroll
        for (int r = c + 1; r < n; ++r) {
            const double f = A[r][c] / A[c][c];
#pragma unroll
            for (int j = c; j < n; ++j) A[r][j] -= f * A[c][j];
            b[r] -= f * b[c];
        }
    }
#pragma unroll
    for (int i = n - 1; i >= 0; --i) {
        double s = b[i];
#pragma unroll
        for (int j = i + 1; j < n; ++j) s -= A[i][j] * x[j];
        x[i] = s / A[i][i];
    }
    return det;
}

template <typename T>
__global__ __launch_bounds__(64 * kWaves) void logquad_window_kernel(
    const T* img, int64_t H, int64_t W, int64_t pitch, const int32_t* labels, const float* rows,
    const int32_t* lab_ids, int32_t K, int32_t pad, Lqw o, double* centers, uint8_t* ok_out) {
#pragma clang fp contract(off)
    __shared__ float s_g[kWaves][kMaxN];
    __shared__ uint8_t s_f[kWaves][kMaxN];
    __shared__ uint8_t s_b[kWaves][kMaxN];
    __shared__ double s_v[kWaves][kMaxN];
    __shared__ double s_w[kWaves][kMaxN];
    __shared__ uint8_t s_s[kWaves][kMaxN];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t f = (int64_t)blockIdx.x * kWaves + wv;
    const bool live = f < K;  // waves past the end work on the last feature and write nothing
    const int64_t fc = live ? f : (int64_t)K - 1;
    float* g = s_g[wv];
    uint8_t* fl = s_f[wv];
    uint8_t* fb = s_b[wv];
    double* vw = s_v[wv];
    double* ww = s_w[wv];
    uint8_t* fs = s_s[wv];

    const float* row = rows + 6 * fc;
    const int64_t x = (int64_t)(int32_t)row[0], y = (int64_t)(int32_t)row[1];
    const int64_t w = (int64_t)(int32_t)row[2], h = (int64_t)(int32_t)row[3];
    const double cxf = (double)row[4], cyf = (double)row[5];
    const int32_t lab = lab_ids[fc];
    const int64_t x0 = x - pad > 0 ? x - pad : 0, y0 = y - pad > 0 ? y - pad : 0;
    const int64_t x1 = x + w + pad < W ? x + w + pad : W, y1 = y + h + pad < H ? y + h + pad : H;
    const int64_t cxi = (int64_t)rint(cxf), cyi = (int64_t)rint(cyf);
    const int R = o.R, P = o.P, N = o.N;

    // 1. outer window: grey levels, label match, ROI membership
    for (int i = lane; i < N; i += 64) {
        const int iy = i / P, ix = i - (i / P) * P;
        const int64_t gx = cxi + ix - R, gy = cyi + iy - R;
        const bool inside = gx >= 0 && gx < W && gy >= 0 && gy < H;
        float gv = 0.0f;
        bool match = false;
        if (inside) {
            gv = (float)img[gy * pitch + gx];
            match = labels[gy * W + gx] == lab;
        }
        const bool roi = gx >= x0 && gx < x1 && gy >= y0 && gy < y1;
        g[i] = gv;
        fl[i] = (match ? kMatch : 0) | (roi ? kRoi : 0);
    }
    __syncthreads();

    // 2. background set: ROI pixels outside the dilated blob
    int nbg = 0;
    for (int i = lane; i < N; i += 64) {
        const int iy = i / P, ix = i - (i / P) * P;
        bool dil = false;
        const int d = o.bg_dilate;
        for (int ny = max(0, iy - d); ny <= min(P - 1, iy + d); ++ny)
            for (int nx = max(0, ix - d); nx <= min(P - 1, ix + d); ++nx) dil = dil || (fl[ny * P + nx] & kMatch);
        const bool bg = (fl[i] & kRoi) && !dil;
        fb[i] = bg ? kBg : 0;
        nbg += bg;
    }
    nbg = wave::sum_i(nbg);
    __syncthreads();
    double bg0 = nbg > 0 ? kth_flagged(g, fb, kBg, N, max((nbg - 1) / 2, 0), lane) : __builtin_inf();
    const bool ok_bg = isfinite(bg0) && nbg >= o.min_bg_pts;

    // 3. tail re-median over the background pixels that are not brighter than
    //    bg_tail_frac * (max over the fit window ROI of g - bg0)
    double vmax0 = -__builtin_inf();
    for (int iw = lane; iw < o.Nw; iw += 64) {
        const int i = (iw / o.Pw + o.off) * P + (iw - (iw / o.Pw) * o.Pw) + o.off;
        if (fl[i] & kRoi) vmax0 = fmax(vmax0, (double)g[i] - bg0);
    }
    vmax0 = wave::max_d(vmax0);
    uint8_t used = kBg;
    int nused = nbg;
    if (o.bg_tail_frac > 0.0) {
        const double thr = o.bg_tail_frac * vmax0;
        int n2 = 0;
        for (int i = lane; i < N; i += 64) {
            const bool b2 = (fb[i] & kBg) && ((double)g[i] - bg0 <= thr);
            if (b2) fb[i] |= kBg2;
            n2 += b2;
        }
        n2 = wave::sum_i(n2);
        __syncthreads();
        const double bg02 = n2 > 0 ? kth_flagged(g, fb, kBg2, N, max((n2 - 1) / 2, 0), lane) : __builtin_inf();
        if (isfinite(bg02) && n2 >= o.min_bg_pts) {
            bg0 = bg02;
            used = kBg2;
            nused = n2;
        }
    }

    // 4. plane background
    bool ok_plane = false;
    double pa = 0.0, pb = 0.0, pc = 0.0;
    if (o.plane_bg) {
        double S[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // 1, x, y, xx, yy, xy, v, xv, yv
        for (int i = lane; i < N; i += 64) {
            if (!(fb[i] & used)) continue;
            const int iy = i / P, ix = i - (i / P) * P;
            const double gx = (double)(cxi + ix - R), gy = (double)(cyi + iy - R), gv = (double)g[i];
            S[0] += 1.0;
            S[1] += gx;
            S[2] += gy;
            S[3] += gx * gx;
            S[4] += gy * gy;
            S[5] += gx * gy;
            S[6] += gv;
            S[7] += gx * gv;
            S[8] += gy * gv;
        }
#pragma unroll
        for (int q = 0; q < 9; ++q) S[q] = wave::sum_d(S[q]);
        double M[3][3] = {{S[3], S[5], S[1]}, {S[5], S[4], S[2]}, {S[1], S[2], S[0]}};
        double rhs[3] = {S[7], S[8], S[6]}, c3[3];
        const double det = ge_solve<3>(M, rhs, c3);
        const bool okp = isfinite(det) && fabs(det) > 1e-18 && S[0] >= 3.0;
        ok_plane = ok_bg && nused >= o.min_plane_pts && okp;
        if (okp) {
            pa = c3[0];
            pb = c3[1];
            pc = c3[2];
        }
    }

    // 5. fit window: dilated blob, residual above the background, fit / core selection
    double vmax = -__builtin_inf();
    for (int iw = lane; iw < o.Nw; iw += 64) {
        const int wy = iw / o.Pw, wx = iw - (iw / o.Pw) * o.Pw;
        const int i = (wy + o.off) * P + wx + o.off;
        bool inb = (fl[i] & kMatch) && (fl[i] & kRoi);
        if (o.fit_dilate > 0) {
            const int d = o.fit_dilate;
            inb = false;
            for (int ny = max(0, wy - d); ny <= min(o.Pw - 1, wy + d); ++ny)
                for (int nx = max(0, wx - d); nx <= min(o.Pw - 1, wx + d); ++nx) {
                    const uint8_t fn = fl[(ny + o.off) * P + nx + o.off];
                    inb = inb || ((fn & kMatch) && (fn & kRoi));
                }
            inb = inb && (fl[i] & kRoi);
        }
        const double gx = (double)(cxi + wx + o.off - R), gy = (double)(cyi + wy + o.off - R);
        const double v = ok_plane ? (double)g[i] - ((pa * gx + pb * gy) + pc) : (double)g[i] - bg0;
        const bool fit = inb && v > 0.0 && isfinite(v);
        vw[iw] = v;
        fs[iw] = fit ? kFit : 0;
        if (fit) vmax = fmax(vmax, v);
    }
    vmax = wave::max_d(vmax);
    const bool ok_fit = ok_bg && isfinite(vmax) && vmax > 0.0;
    const double core_thr = o.core_frac * vmax;
    int ncore = 0, nfit = 0;
    for (int iw = lane; iw < o.Nw; iw += 64) {
        const bool fit = fs[iw] & kFit;
        nfit += fit;
        ncore += fit && vw[iw] >= core_thr;
    }
    ncore = wave::sum_i(ncore);
    nfit = wave::sum_i(nfit);
    const bool use_core = ncore >= 6;
    const int nsel = use_core ? ncore : nfit;
    const bool ok_sel = ok_fit && nsel >= 6;
    for (int iw = lane; iw < o.Nw; iw += 64) {
        const bool fit = fs[iw] & kFit;
        const bool sel = use_core ? (fit && vw[iw] >= core_thr) : fit;
        const double vv = fmax(vw[iw], o.eps);
        const double wt = o.weight_power == 1.0 ? vv : pow(vv, o.weight_power);
        ww[iw] = sel ? wt : 0.0;
        if (sel) fs[iw] |= kSel;
    }
    __syncthreads();
    const double thr99 =
        nsel > 0 ? kth_flagged(ww, fs, kSel, o.Nw, max((int)floor(0.99 * ((double)nsel - 1.0)), 0), lane) : 0.0;

    // 6. normalised coordinates over the selection
    double m4[4] = {0, 0, 0, 0};  // x, y, xx, yy
    for (int iw = lane; iw < o.Nw; iw += 64) {
        if (!(fs[iw] & kSel)) continue;
        const int wy = iw / o.Pw, wx = iw - (iw / o.Pw) * o.Pw;
        const double gx = (double)(cxi + wx + o.off - R), gy = (double)(cyi + wy + o.off - R);
        m4[0] += gx;
        m4[1] += gy;
        m4[2] += gx * gx;
        m4[3] += gy * gy;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) m4[q] = wave::sum_d(m4[q]);
    const double n = fmax((double)nsel, 1.0);
    const double xm = m4[0] / n, ym = m4[1] / n;
    const double varx = fmax(m4[2] / n - xm * xm, 0.0), vary = fmax(m4[3] / n - ym * ym, 0.0);
    const double scl = fmax(fmax(sqrt(varx), sqrt(vary)), 1.0);

    // 7. weighted normal equations of log v on the quadratic monomials
    double A[21], bz[6];
#pragma unroll
    for (int q = 0; q < 21; ++q) A[q] = 0.0;
#pragma unroll
    for (int q = 0; q < 6; ++q) bz[q] = 0.0;
    for (int iw = lane; iw < o.Nw; iw += 64) {
        if (!(fs[iw] & kSel)) continue;
        const int wy = iw / o.Pw, wx = iw - (iw / o.Pw) * o.Pw;
        const double gx = (double)(cxi + wx + o.off - R), gy = (double)(cyi + wy + o.off - R);
        const double xn = (gx - xm) / scl, yn = (gy - ym) / scl;
        const double wt = fmin(ww[iw], thr99);
        const double z = log(fmax(vw[iw], o.eps));
        const double p[6] = {xn * xn, yn * yn, xn * yn, xn, yn, 1.0};
        int q = 0;
#pragma unroll
        for (int r = 0; r < 6; ++r) {
#pragma unroll
            for (int c = r; c < 6; ++c) A[q++] += (p[r] * wt) * p[c];
            bz[r] += p[r] * (wt * z);
        }
    }
#pragma unroll
    for (int q = 0; q < 21; ++q) A[q] = wave::sum_d(A[q]);
#pragma unroll
    for (int q = 0; q < 6; ++q) bz[q] = wave::sum_d(bz[q]);
    double coef[6] = {0, 0, 0, 0, 0, 0};
    if (ok_sel) {
        double M6[6][6];
        int q = 0;
#pragma unroll
        for (int r = 0; r < 6; ++r)
#pragma unroll
            for (int c = r; c < 6; ++c) {
                M6[r][c] = A[q];
                M6[c][r] = A[q];
                ++q;
            }
        ge_solve<6>(M6, bz, coef);
    }

    // 8. peak gates and the centre
    const double H00 = 2.0 * coef[0], H11 = 2.0 * coef[1], H01 = coef[2];
    const double tr = H00 + H11;
    const double disc = sqrt(fmax((H00 - H11) * (H00 - H11) + 4.0 * H01 * H01, 0.0));
    const double lam1 = 0.5 * (tr - disc), lam2 = 0.5 * (tr + disc);
    bool ok = ok_sel && lam1 < -o.negdef_eps && lam2 < -o.negdef_eps;
    ok = ok && fabs(lam2) / fmax(fabs(lam1), 1e-30) <= o.max_cond;
    const double det = H00 * H11 - H01 * H01;
    ok = ok && fabs(det) > 1e-20;
    const double cxg = xm + scl * ((-coef[3] * H11 + coef[4] * H01) / det);
    const double cyg = ym + scl * ((coef[3] * H01 - coef[4] * H00) / det);
    ok = ok && cxg >= cxf - o.win_rad && cxg <= cxf + o.win_rad && cyg >= cyf - o.win_rad && cyg <= cyf + o.win_rad;
    if (live && lane == 0) {
        centers[2 * f] = ok ? cxg : cxf;
        centers[2 * f + 1] = ok ? cyg : cyf;
        ok_out[f] = ok ? 1 : 0;
    }
}

}  // namespace

int logquad_window_launch(const void* img, int dtype, int64_t H, int64_t W, int64_t pitch, const int32_t* labels,
                          const float* rows, const int32_t* lab_ids, int32_t K, int32_t pad,
                          const SpxLogquadWindowOptions& opt, double* centers, uint8_t* ok, hipStream_t stream) {
    if (K == 0) return SPX_OK;
    SPX_REQUIRE(opt.win_rad >= 1 && opt.bg_win_pad >= 0 && opt.win_rad + opt.bg_win_pad <= kMaxR,
                "spx_refine_logquad_window: need win_rad >= 1, bg_win_pad >= 0, win_rad + bg_win_pad <= 10");
    SPX_REQUIRE(opt.bg_dilate >= 0 && opt.fit_dilate >= 0, "spx_refine_logquad_window: dilations must be >= 0");
    SPX_REQUIRE(dtype == SPX_DTYPE_U16 || dtype == SPX_DTYPE_F32, "spx_refine_logquad_window: dtype must be U16 or F32");
    Lqw o;
    o.R = opt.win_rad + opt.bg_win_pad;
    o.P = 2 * o.R + 1;
    o.N = o.P * o.P;
    o.win_rad = opt.win_rad;
    o.off = opt.bg_win_pad;
    o.Pw = 2 * opt.win_rad + 1;
    o.Nw = o.Pw * o.Pw;
    o.bg_dilate = opt.bg_dilate;
    o.fit_dilate = opt.fit_dilate;
    o.min_bg_pts = opt.min_bg_pts;
    o.min_plane_pts = opt.min_plane_pts;
    o.plane_bg = opt.plane_bg;
    o.core_frac = opt.core_frac;
    o.weight_power = opt.weight_power;
    o.bg_tail_frac = opt.bg_tail_frac;
    o.max_cond = opt.max_cond;
    o.negdef_eps = opt.negdef_eps;
    o.eps = opt.eps;
    const int blocks = (K + kWaves - 1) / kWaves;
    if (dtype == SPX_DTYPE_U16)
        logquad_window_kernel<uint16_t><<<blocks, 64 * kWaves, 0, stream>>>(
            static_cast<const uint16_t*>(img), H, W, pitch, labels, rows, lab_ids, K, pad, o, centers, ok);
    else
        logquad_window_kernel<float><<<blocks, 64 * kWaves, 0, stream>>>(
            static_cast<const float*>(img), H, W, pitch, labels, rows, lab_ids, K, pad, o, centers, ok);
    SPX_LAUNCHED();
    return SPX_OK;
}

}  // namespace spx
