// One-pass outline fit (fit_outline_fast) on gfx950: the device work of steps A to C.
//
//   seeds        separable Gaussian smoothing, local maxima, log-parabola refinement
//   geometry     nearest other hole through a cell list (the page's spacing and window)
//   stack        plain mean of the holes' windows on quarter-pixel bins, each hole against it
//   pair maps    every stacked hole's neighbours splatted as cubic B-splines into the map of
//                its sub-pixel class (followed by an FFT on the host side), cropped to the band
//   page fit     the outline's transform along its contour (trapezoid rule) and its shape
//                derivatives, the blur's transfer function and its derivatives, the stack's
//                model by class (two separable complex products), the variable-projection
//                cost, gradient and Gauss-Newton matrix per page, the grid of starts' costs,
//                and the Levenberg-Marquardt iterations themselves (bounded step, outline of
//                the trial point, accept and damping rules) with every page's state resident
//   templates    the outline polygon's exact transform x blur x pixel box on a band, written
//                into a half spectrum for the inverse FFT
//
// Reference: subpx/_gpu/outline_onepass.py:81-250 (stack_acc, hole_vs_stack, pair_splat,
// nearest), :463-501 (find_seeds), :663-698 (pair_map); subpx/_outline/onepass.py:44-126 (Spec,
// Band), :129-162 (outline_hat), :281-410 (otf, template_hat, synth, evaluate), :422-486
// (grid_start), :504-623 (fit). The reference sums with floating-point atomics, synthesises in
// single precision and runs its Levenberg-Marquardt loop on the host; here every sum that
// several threads share is a 64-bit fixed-point integer (order-independent, the same bits every
// run), the synthesis and the reductions are FP64 in a fixed order, and the iterations stay on
// the device (the host reads one flag per iteration).
//
// FP64 without FMA contraction, so that the smoothing reproduces the oracle's sums bit for bit.
#include <cmath>
#include <string>

#include "spx_internal.hpp"
#include "wave64.hpp"

#pragma clang fp contract(off)

namespace spx {
namespace {

constexpr int kBlock = 256;
constexpr double kTwoPi = 6.283185307179586;
constexpr double kPi = 3.141592653589793;
constexpr double kStackFix = 16777216.0;      // 2^24 units per grey level
constexpr double kPairFix = 1099511627776.0;  // 2^40 units per map count

__device__ inline double nan_d() { return __longlong_as_double(0x7FF8000000000000ll); }

__device__ inline double sinc_d(double x) {
    if (x == 0.0) return 1.0;
    const double px = kPi * x;
    return sin(px) / px;
}

__device__ inline void fix_add(unsigned long long* acc, double v, double scale) {
    atomicAdd(acc, (unsigned long long)llrint(v * scale));
}

__global__ void zero_u64(unsigned long long* p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

int zero_launch(unsigned long long* p, int64_t n, hipStream_t st) {
    if (n <= 0) return SPX_OK;
    hipLaunchKernelGGL(zero_u64, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, p, n);
    SPX_LAUNCHED();
    return SPX_OK;
}

// ------------------------------------------------------------------------------------- seeds
// One axis of the separable smoothing with clamped ('nearest') edges: out = w0 x + sum over
// k = r .. 1 of w_k (x[-k] + x[+k]), in that order. AXIS 0 runs down the columns.
template <int AXIS, typename T>
__global__ void smooth_axis(const T* __restrict__ src, int64_t H, int64_t W, const double* __restrict__ taps, int r,
                            double* __restrict__ dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= H * W) return;
    const int64_t y = i / W, x = i % W;
    const int64_t n = AXIS == 0 ? H : W;
    const int64_t p = AXIS == 0 ? y : x;
    auto at = [&](int64_t q) -> double {
        q = q < 0 ? 0 : (q >= n ? n - 1 : q);
        return (double)(AXIS == 0 ? src[q * W + x] : src[y * W + q]);
    };
    double acc = taps[0] * at(p);
    for (int k = r; k >= 1; --k) acc = acc + taps[k] * (at(p - k) + at(p + k));
    dst[i] = acc;
}

// Peak flags: not on the frame, above thr, equal to the maximum of its (2 nms + 1)^2 window
// (clamped edges).
__global__ void peak_flags(const double* __restrict__ sm, int64_t H, int64_t W, int nms, double thr,
                           uint8_t* __restrict__ flags) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= H * W) return;
    const int64_t y = i / W, x = i % W;
    uint8_t f = 0;
    const double v = sm[i];
    if (y > 0 && y < H - 1 && x > 0 && x < W - 1 && v > thr) {
        double m = v;
        for (int dy = -nms; dy <= nms; ++dy) {
            int64_t yy = y + dy;
            yy = yy < 0 ? 0 : (yy >= H ? H - 1 : yy);
            for (int dx = -nms; dx <= nms; ++dx) {
                int64_t xx = x + dx;
                xx = xx < 0 ? 0 : (xx >= W ? W - 1 : xx);
                m = fmax(m, sm[yy * W + xx]);
            }
        }
        f = (uint8_t)(v == m);
    }
    flags[i] = f;
}

// Peaks per row: one wave per row.
__global__ void row_counts(const uint8_t* __restrict__ flags, int64_t H, int64_t W, int32_t* __restrict__ counts) {
    const int64_t row = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= H) return;
    int c = 0;
    for (int64_t x = lane; x < W; x += 64) c += flags[row * W + x] != 0;
    c = wave::sum_i(c);
    if (lane == 0) counts[row] = c;
}

// The refined peaks in row-major order: one wave per row writes its peaks from the row's
// offset on, each refined by a parabola through log(max(v - base, 1e-6)) of the three samples
// along x and along y (the vertex offset clipped to +-0.5 px; 0 where the parabola is not
// concave).
__device__ inline double log_height(double v, double base) { return log(fmax(v - base, 1e-6)); }

__device__ inline double vertex(double l, double c0, double r) {
    const double den = l - 2.0 * c0 + r;
    const double o = den < 0.0 ? 0.5 * (l - r) / den : 0.0;
    return fmin(fmax(o, -0.5), 0.5);
}

__global__ void peak_list(const double* __restrict__ sm, const uint8_t* __restrict__ flags,
                          const int32_t* __restrict__ offsets, int64_t H, int64_t W, double base,
                          double* __restrict__ seeds) {
    const int64_t row = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= H) return;
    int base_pos = offsets[row];
    for (int64_t x0 = 0; x0 < W; x0 += 64) {
        const int64_t x = x0 + lane;
        const bool f = x < W && flags[row * W + x] != 0;
        const unsigned long long mask = __ballot(f);
        if (f) {
            const int pos = base_pos + __popcll(mask & ((1ull << lane) - 1ull));
            const int64_t i = row * W + x;
            const double c0 = log_height(sm[i], base);
            const double dx = vertex(log_height(sm[i - 1], base), c0, log_height(sm[i + 1], base));
            const double dy = vertex(log_height(sm[i - W], base), c0, log_height(sm[i + W], base));
            seeds[2 * pos] = (double)x + dx;
            seeds[2 * pos + 1] = (double)row + dy;
        }
        base_pos += __popcll(mask);
    }
}

// Pixels of image[::step, ::step] above thr.
__global__ void count_above(const float* __restrict__ img, int64_t H, int64_t W, int step, double thr,
                            unsigned long long* __restrict__ count) {
    const int64_t hs = (H + step - 1) / step, wsub = (W + step - 1) / step;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int c = 0;
    if (i < hs * wsub) c = (double)img[(i / wsub) * step * W + (i % wsub) * step] > thr;
    c = wave::sum_i(c);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(count, (unsigned long long)c);
}

// ------------------------------------------------------------------------------ geometry
// Nearest other hole among the 3 x 3 cells round the hole's cell; 1e15 when none is there.
__global__ void nearest_kernel(const double* __restrict__ c, int n, const int32_t* __restrict__ order,
                               const int32_t* __restrict__ start, int ncx, int ncy, double cs,
                               double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = c[2 * i], y = c[2 * i + 1];
    const int cx = (int)floor(x / cs), cy = (int)floor(y / cs);
    double best = 1e30;
    for (int yy = cy - 1; yy <= cy + 1; ++yy) {
        if (yy < 0 || yy >= ncy) continue;
        const int x0 = cx - 1 < 0 ? 0 : cx - 1, x1 = cx + 1 >= ncx ? ncx - 1 : cx + 1;
        if (x0 > x1) continue;
        for (int k = start[yy * ncx + x0]; k < start[yy * ncx + x1 + 1]; ++k) {
            const int j = order[k];
            if (j == i) continue;
            const double dx = c[2 * j] - x, dy = c[2 * j + 1] - y;
            best = fmin(best, dx * dx + dy * dy);
        }
    }
    out[i] = sqrt(best);
}

// ---------------------------------------------------------------------------------- stack
// One thread per (hole, window row) of the holes with use[i]: each pixel inside the frame whose
// bin rint((p - c) U) + Ms / 2 lies in the Ms x Ms stack adds its grey level and a count.
__global__ void stack_acc(const float* __restrict__ img, int64_t H, int64_t W, const double* __restrict__ c, int n,
                          const uint8_t* __restrict__ use, int R, int U, int Ms, unsigned long long* __restrict__ sums,
                          unsigned long long* __restrict__ counts) {
    const int side = 2 * R + 1;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)n * side) return;
    const int i = (int)(t / side);
    if (!use[i]) return;
    const int oy = (int)(t % side) - R;
    const double x = c[2 * i], y = c[2 * i + 1];
    const int64_t rx = (int64_t)rint(x), ry = (int64_t)rint(y);
    const int64_t py = ry + oy;
    if (py < 0 || py >= H) return;
    const int64_t by = (int64_t)rint(((double)py - y) * U) + Ms / 2;
    if (by < 0 || by >= Ms) return;
    for (int ox = -R; ox <= R; ++ox) {
        const int64_t px = rx + ox;
        if (px < 0 || px >= W) continue;
        const int64_t bx = (int64_t)rint(((double)px - x) * U) + Ms / 2;
        if (bx < 0 || bx >= Ms) continue;
        fix_add(sums + by * Ms + bx, (double)img[py * W + px], kStackFix);
        atomicAdd(counts + by * Ms + bx, 1ull);
    }
}

__global__ void stack_finish(const unsigned long long* __restrict__ sums, const unsigned long long* __restrict__ counts,
                             int64_t nb, double* __restrict__ stack, double* __restrict__ wsum) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const double cnt = (double)counts[b];
    const double s = (double)(long long)sums[b] / kStackFix;
    stack[b] = cnt > 0 ? s / cnt : nan_d();
    wsum[b] = cnt;
}

// The same sums with the bins of one workgroup's holes in LDS (64-bit fixed-point grey levels and
// 32-bit counts, LDS atomics), each workgroup's totals then stored as its own partial row
// (partial[g] = [sums (nb), counts (nb)] as u64): no global atomics, and the integer sums do not
// depend on the order of the threads. Holes [g per, (g + 1) per) of workgroup g.
constexpr int kStackLdsBytes = 64 * 1024;

__global__ __launch_bounds__(kBlock) void stack_acc_lds(const float* __restrict__ img, int64_t H, int64_t W,
                                                        const double* __restrict__ c, int n,
                                                        const uint8_t* __restrict__ use, int R, int U, int Ms, int per,
                                                        unsigned long long* __restrict__ partial) {
    extern __shared__ unsigned long long lds_stack[];
    const int nb = Ms * Ms;
    unsigned long long* ssum = lds_stack;
    unsigned int* scnt = reinterpret_cast<unsigned int*>(lds_stack + nb);
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
        ssum[i] = 0ull;
        scnt[i] = 0u;
    }
    __syncthreads();
    const int side = 2 * R + 1;
    const int h0 = blockIdx.x * per;
    const int h1 = h0 + per < n ? h0 + per : n;
    for (int t = threadIdx.x; t < (h1 - h0) * side; t += blockDim.x) {
        const int i = h0 + t / side;
        if (!use[i]) continue;
        const int oy = t % side - R;
        const double x = c[2 * i], y = c[2 * i + 1];
        const int64_t rx = (int64_t)rint(x), ry = (int64_t)rint(y);
        const int64_t py = ry + oy;
        if (py < 0 || py >= H) continue;
        const int64_t by = (int64_t)rint(((double)py - y) * U) + Ms / 2;
        if (by < 0 || by >= Ms) continue;
        for (int ox = -R; ox <= R; ++ox) {
            const int64_t px = rx + ox;
            if (px < 0 || px >= W) continue;
            const int64_t bx = (int64_t)rint(((double)px - x) * U) + Ms / 2;
            if (bx < 0 || bx >= Ms) continue;
            const int b = (int)(by * Ms + bx);
            atomicAdd(ssum + b, (unsigned long long)llrint((double)img[py * W + px] * kStackFix));
            atomicAdd(scnt + b, 1u);
        }
    }
    __syncthreads();
    unsigned long long* out = partial + (int64_t)blockIdx.x * 2 * nb;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
        out[i] = ssum[i];
        out[nb + i] = (unsigned long long)scnt[i];
    }
}

// stack and wsum from the G workgroups' partial rows (sums over g in index order).
__global__ void stack_reduce(const unsigned long long* __restrict__ partial, int G, int64_t nb,
                             double* __restrict__ stack, double* __restrict__ wsum) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    unsigned long long s = 0ull, cnt = 0ull;
    for (int g = 0; g < G; ++g) {
        s += partial[(int64_t)g * 2 * nb + b];
        cnt += partial[(int64_t)g * 2 * nb + nb + b];
    }
    const double cn = (double)cnt;
    stack[b] = cnt > 0 ? ((double)(long long)s / kStackFix) / cn : nan_d();
    wsum[b] = cn;
}

// Filled stack (NaN -> 0) and its central-difference gradient (one-sided at the edges), bins of
// 1 / U px.
__device__ inline double filled(const double* S, int64_t b) {
    const double v = S[b];
    return v == v ? v : 0.0;
}

__device__ inline double grad_axis(const double* S, int Ms, int by, int bx, int axis, double h) {
    const int p = axis == 0 ? by : bx;
    auto at = [&](int q) { return axis == 0 ? filled(S, (int64_t)q * Ms + bx) : filled(S, (int64_t)by * Ms + q); };
    if (p == 0) return (at(1) - at(0)) / h;
    if (p == Ms - 1) return (at(Ms - 1) - at(Ms - 2)) / h;
    return (at(p + 1) - at(p - 1)) / (2.0 * h);
}

// One wave per hole: brightness and level against the stack by 2 x 2 least squares over the
// core pixels (within core2 of the hole), leftover rms over the window, one Gauss-Newton step of
// the centre over the core pixels. out (n, 6) = a, level, rms, dx, dy, pixels.
__global__ void versus_kernel(const float* __restrict__ img, int64_t H, int64_t W, const double* __restrict__ c,
                              int n, const double* __restrict__ S, int R, int U, int Ms, double core2,
                              double* __restrict__ out) {
    const int i = (int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= n) return;
    const double x = c[2 * i], y = c[2 * i + 1];
    const int64_t rx = (int64_t)rint(x), ry = (int64_t)rint(y);
    const int side = 2 * R + 1;
    const int P = side * side;
    const double h = 1.0 / U;
    double cn = 0, s1 = 0, s2 = 0, sy = 0, y1 = 0;
    for (int q = lane; q < P; q += 64) {
        const int64_t py = ry + q / side - R, px = rx + q % side - R;
        if (py < 0 || py >= H || px < 0 || px >= W) continue;
        const double uy = (double)py - y, ux = (double)px - x;
        const int64_t by = (int64_t)rint(uy * U) + Ms / 2, bx = (int64_t)rint(ux * U) + Ms / 2;
        if (by < 0 || by >= Ms || bx < 0 || bx >= Ms) continue;
        if (ux * ux + uy * uy > core2) continue;
        const double s = S[by * Ms + bx];
        if (s != s) continue;
        const double v = (double)img[py * W + px];
        cn += 1.0;
        s1 += s;
        s2 += s * s;
        sy += s * v;
        y1 += v;
    }
    cn = wave::sum_d(cn);
    s1 = wave::sum_d(s1);
    s2 = wave::sum_d(s2);
    sy = wave::sum_d(sy);
    y1 = wave::sum_d(y1);
    const double det = cn * s2 - s1 * s1;
    double a = 1.0, lev = 0.0;
    if (det > 0) {
        a = (cn * sy - s1 * y1) / det;
        lev = (s2 * y1 - s1 * sy) / det;
    }
    double rss = 0, nf = 0, sxx = 0, sxy = 0, syy = 0, gx = 0, gy = 0;
    for (int q = lane; q < P; q += 64) {
        const int64_t py = ry + q / side - R, px = rx + q % side - R;
        if (py < 0 || py >= H || px < 0 || px >= W) continue;
        const double uy = (double)py - y, ux = (double)px - x;
        const int64_t by = (int64_t)rint(uy * U) + Ms / 2, bx = (int64_t)rint(ux * U) + Ms / 2;
        if (by < 0 || by >= Ms || bx < 0 || bx >= Ms) continue;
        const double s = S[by * Ms + bx];
        if (s != s) continue;
        const double res = (double)img[py * W + px] - a * s - lev;
        rss += res * res;
        nf += 1.0;
        if (ux * ux + uy * uy > core2) continue;
        const double jx = -a * grad_axis(S, Ms, (int)by, (int)bx, 1, h);
        const double jy = -a * grad_axis(S, Ms, (int)by, (int)bx, 0, h);
        sxx += jx * jx;
        sxy += jx * jy;
        syy += jy * jy;
        gx += jx * res;
        gy += jy * res;
    }
    rss = wave::sum_d(rss);
    nf = wave::sum_d(nf);
    sxx = wave::sum_d(sxx);
    sxy = wave::sum_d(sxy);
    syy = wave::sum_d(syy);
    gx = wave::sum_d(gx);
    gy = wave::sum_d(gy);
    if (lane == 0) {
        const double d2 = sxx * syy - sxy * sxy;
        double dx = 0, dy = 0;
        if (d2 > 0) {
            dx = (syy * gx - sxy * gy) / d2;
            dy = (sxx * gy - sxy * gx) / d2;
        }
        double* o = out + (int64_t)i * 6;
        o[0] = a;
        o[1] = lev;
        o[2] = sqrt(rss / (nf > 0 ? nf : 1.0));
        o[3] = dx;
        o[4] = dy;
        o[5] = nf;
    }
}

// ------------------------------------------------------------------------------- pair maps
// Weight of tap k (at -1, 0, 1, 2) of the cubic B-spline for the fraction t in [0, 1).
__device__ inline double bspline_tap(double t, int k) {
    const double u = 1.0 - t;
    if (k == 0) return u * u * u / 6.0;
    if (k == 1) return (3.0 * t * t * t - 6.0 * t * t + 4.0) / 6.0;
    if (k == 2) return (-3.0 * t * t * t + 3.0 * t * t + 3.0 * t + 1.0) / 6.0;
    return t * t * t / 6.0;
}

// Sub-pixel class and residual offset of hole i (outline_onepass.py:670-674): v = -c U,
// class (mod(rint(v_y), U), mod(rint(v_x), U)), s = (v - rint(v)) / U.
__device__ inline int pair_class(const double* c, int i, int U, double& sx, double& sy) {
    const double vx = -c[2 * i] * U, vy = -c[2 * i + 1] * U;
    const double nx = rint(vx), ny = rint(vy);
    sx = (vx - nx) / U;
    sy = (vy - ny) / U;
    const int qx = (int)(((long long)nx % U + U) % U), qy = (int)(((long long)ny % U + U) % U);
    return qy * U + qx;
}

// Stacked holes per class.
__global__ void pair_counts(const double* __restrict__ c, const uint8_t* __restrict__ use, int n, int U,
                            unsigned long long* __restrict__ counts) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !use[i]) return;
    double sx, sy;
    atomicAdd(counts + pair_class(c, i, U, sx, sy), 1ull);
}

// One wave per stacked hole i (g_i = 1 / the stacked holes of its class): every hole j of the
// cells within ceil(reach / cs) of i's cell with |c_j - c_i| <= reach per axis adds
// g_i a_j / amean at c_j - c_i - s_i to the map of i's class, as 4 x 4 cubic B-spline taps on
// Ug points per px, wrapped round the Ng x Ng map. The lanes take one (neighbour, tap) each,
// 4 neighbours per pass of the wave.
__global__ void pair_splat(const double* __restrict__ c, const double* __restrict__ a, double amean,
                           const uint8_t* __restrict__ use, const unsigned long long* __restrict__ counts, int n,
                           int U, const int32_t* __restrict__ order, const int32_t* __restrict__ start, int ncx,
                           int ncy, double cs, double reach, int Ug, int Ng, unsigned long long* __restrict__ hist) {
    const int i = (int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= n || !use[i]) return;
    double sx, sy;
    const int cls = pair_class(c, i, U, sx, sy);
    const double gi = 1.0 / fmax((double)counts[cls], 1e-300);
    const double x = c[2 * i], y = c[2 * i + 1];
    const int cx = (int)floor(x / cs), cy = (int)floor(y / cs);
    const int r = (int)ceil(reach / cs);
    unsigned long long* h = hist + (int64_t)cls * Ng * Ng;
    const int tq = (lane >> 2) & 3, tp = lane & 3;  // this lane's tap (row q, column p)
    for (int yy = cy - r; yy <= cy + r; ++yy) {
        if (yy < 0 || yy >= ncy) continue;
        const int x0 = cx - r < 0 ? 0 : cx - r, x1 = cx + r >= ncx ? ncx - 1 : cx + r;
        if (x0 > x1) continue;
        const int k0 = start[yy * ncx + x0], k1 = start[yy * ncx + x1 + 1];
        for (int k = k0 + (lane >> 4); k < k1; k += 4) {
            const int j = order[k];
            const double dx = c[2 * j] - x, dy = c[2 * j + 1] - y;
            if (fabs(dx) > reach || fabs(dy) > reach) continue;
            const double w = gi * (a[j] / amean);
            const double pxf = (dx - sx) * Ug + (double)(Ng / 2), pyf = (dy - sy) * Ug + (double)(Ng / 2);
            const double fx = floor(pxf), fy = floor(pyf);
            const int ry = (((int)fy - 1 + tq) % Ng + Ng) % Ng;
            const int rxx = (((int)fx - 1 + tp) % Ng + Ng) % Ng;
            fix_add(h + (int64_t)ry * Ng + rxx, w * bspline_tap(pyf - fy, tq) * bspline_tap(pxf - fx, tp), kPairFix);
        }
    }
}

__global__ void fix_to_double(const unsigned long long* __restrict__ acc, int64_t n, double scale,
                              double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (double)(long long)acc[i] / scale;
}

// Gamma_q(ky, kx) = spec_q[ky mod Ng][kx] / (centre x fix), ky in -kmax .. kmax, kx in 0 .. kmax;
// centre = exp(-2 pi i (fx + fy) (Ng / 2) / Ug), fix = sinc(fx / Ug)^4 sinc(fy / Ug)^4.
__global__ void pair_crop(const double* __restrict__ spec, int nclass, int Ng, int Ug, int L, int kmax,
                          double* __restrict__ gamma) {
    const int nky = 2 * kmax + 1, nkx = kmax + 1, nc = Ng / 2 + 1;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nclass * nky * nkx) return;
    const int q = (int)(t / (nky * nkx));
    const int iy = (int)(t / nkx % nky), kx = (int)(t % nkx);
    const int ky = iy - kmax;
    const int row = ((ky % Ng) + Ng) % Ng;
    const double fx = (double)kx / L, fy = (double)ky / L;
    const double sxv = sinc_d(fx / Ug), syv = sinc_d(fy / Ug);
    const double fix = (sxv * sxv * sxv * sxv) * (syv * syv * syv * syv);
    const double ph = -kTwoPi * (fx + fy) * ((double)(Ng / 2) / Ug);
    const double cr = cos(ph) * fix, ci = sin(ph) * fix;  // centre x fix
    const double* z = spec + (((int64_t)q * Ng + row) * nc + kx) * 2;
    const double den = cr * cr + ci * ci;
    gamma[t * 2] = (z[0] * cr + z[1] * ci) / den;
    gamma[t * 2 + 1] = (z[1] * cr - z[0] * ci) / den;
}

// --------------------------------------------------------------------------------- page fit
// One wave per (outline row, frequency), the lanes splitting the outline points:
// O^(f) = i dphi / (2 pi |f|^2) sum_m (fx yd_m - fy xd_m) exp(-2 pi i f . x_m) (the area at f = 0)
// and dO^/dp_k = dphi sum_m g_mk exp(-2 pi i f . x_m); the lanes' sums meet in a fixed-order wave
// reduction. geo = (4, rows, n) [x, y, xd, yd]; gk (rows, n, K); out (rows, F, 1 + K) complex.
constexpr int kMaxShape = 16;

// KK >= 0: the number of derivatives fixed at compile time (the accumulators stay in registers);
// KK < 0: K read at run time (up to kMaxShape).
template <int KK>
__global__ void outline_hat_kernel(const double* __restrict__ geo, const double* __restrict__ gk,
                                   const double* __restrict__ area, int rows, int n, int Kr, double dphi,
                                   const double* __restrict__ fxs, const double* __restrict__ fys, int F,
                                   double* __restrict__ out) {
    const int K = KK >= 0 ? KK : Kr;
    const int64_t t = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (t >= (int64_t)rows * F) return;
    const int b = (int)(t / F), f = (int)(t % F);
    const double fx = fxs[f], fy = fys[f];
    const double* xs = geo + (int64_t)b * n;
    const double* ys = geo + ((int64_t)rows + b) * n;
    const double* xd = geo + (2 * (int64_t)rows + b) * n;
    const double* yd = geo + (3 * (int64_t)rows + b) * n;
    const double* g = gk + (int64_t)b * n * K;
    double re0 = 0, im0 = 0;
    constexpr int KA = KK >= 0 ? (KK > 0 ? KK : 1) : kMaxShape;
    double re[KA], im[KA];
#pragma unroll
    for (int k = 0; k < KA; ++k) re[k] = im[k] = 0;
    for (int m = lane; m < n; m += 64) {
        const double ph = -kTwoPi * (fx * xs[m] + fy * ys[m]);
        double sn, cs;
        sincos(ph, &sn, &cs);
        const double w = fx * yd[m] - fy * xd[m];
        re0 += cs * w;
        im0 += sn * w;
        for (int k = 0; k < K; ++k) {
            const double gv = g[(int64_t)m * K + k];
            re[k] += cs * gv;
            im[k] += sn * gv;
        }
    }
    re0 = wave::sum_d(re0);
    im0 = wave::sum_d(im0);
    for (int k = 0; k < K; ++k) {
        re[k] = wave::sum_d(re[k]);
        im[k] = wave::sum_d(im[k]);
    }
    double* o = out + t * (1 + K) * 2;
    if (lane == 0) {
        const double f2 = fx * fx + fy * fy;
        if (f2 == 0.0) {
            o[0] = area[b];
            o[1] = 0.0;
        } else {
            const double q = dphi / (kTwoPi * f2);
            o[0] = -im0 * q;
            o[1] = re0 * q;
        }
    }
    for (int k = 0; k < K; ++k)
        if (lane == k) {
            o[2 * (1 + k)] = re[k] * dphi;
            o[2 * (1 + k) + 1] = im[k] * dphi;
        }
}

// Blur transfer function at (fx, fy) and (grads) its derivatives along the blur parameters bp.
// kind 0: two Gaussians (5 parameters), 1: Gaussian x circular pupil (3), 2: one Gaussian (2).
constexpr int kMaxBlur = 5;

__device__ inline int otf_eval(int kind, const double* bp, double fx, double fy, bool grads, double* o) {
    const double c = 2.0 * kPi * kPi;
    const double fx2 = fx * fx, fy2 = fy * fy;
    if (kind == 0) {
        const double s1x = exp(bp[0]), s1y = exp(bp[1]), s2x = exp(bp[2]), s2y = exp(bp[3]);
        const double w2 = 1.0 / (1.0 + exp(-bp[4]));
        const double g1 = exp(-c * (s1x * s1x * fx2 + s1y * s1y * fy2));
        const double g2 = exp(-c * (s2x * s2x * fx2 + s2y * s2y * fy2));
        o[0] = (1.0 - w2) * g1 + w2 * g2;
        if (grads) {
            o[1] = (1.0 - w2) * g1 * (-2.0 * c * s1x * s1x * fx2);
            o[2] = (1.0 - w2) * g1 * (-2.0 * c * s1y * s1y * fy2);
            o[3] = w2 * g2 * (-2.0 * c * s2x * s2x * fx2);
            o[4] = w2 * g2 * (-2.0 * c * s2y * s2y * fy2);
            o[5] = (g2 - g1) * w2 * (1.0 - w2);
        }
        return 5;
    }
    const double sx = exp(bp[0]), sy = exp(bp[1]);
    const double g = exp(-c * (sx * sx * fx2 + sy * sy * fy2));
    if (kind == 1) {
        const double s = fmin(fmax(sqrt(fx2 + fy2) / exp(bp[2]), 0.0), 1.0);
        const double root = sqrt(1.0 - s * s);
        const double lens = (2.0 / kPi) * (acos(s) - s * root);
        o[0] = g * lens;
        if (grads) {
            o[1] = g * lens * (-2.0 * c * sx * sx * fx2);
            o[2] = g * lens * (-2.0 * c * sy * sy * fy2);
            o[3] = g * (4.0 / kPi) * s * root;
        }
        return 3;
    }
    o[0] = g;
    if (grads) {
        o[1] = g * (-2.0 * c * sx * sx * fx2);
        o[2] = g * (-2.0 * c * sy * sy * fy2);
    }
    return 2;
}

// One thread per (page, frequency): the template's band coefficient and, with tm (P x nf), its
// derivatives along the stage's free directions (ohat rows of 1 + K entries; without tm only
// the first is read): coeff (B, 1 + nf, F) complex,
//   c0 = O0 B0 pix,  c_{1+i} = pix sum_k tm[k][i] full_k,  full = [O_{1..K} B0, O0 B_{1..nb}].
__global__ void coeff_kernel(const double* __restrict__ ohat, int K, const double* __restrict__ blur, int kind, int nb,
                             const double* __restrict__ tm, int nf, const double* __restrict__ fxs,
                             const double* __restrict__ fys, const double* __restrict__ pix, int B, int F,
                             double* __restrict__ coeff) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)B * F) return;
    const int b = (int)(t / F), f = (int)(t % F);
    const bool grads = tm != nullptr;
    double bo[1 + kMaxBlur];
    otf_eval(kind, blur + (int64_t)b * nb, fxs[f], fys[f], grads, bo);
    const double* o = ohat + t * (1 + K) * 2;
    const double p = pix[f];
    const int C = 1 + (grads ? nf : 0);
    double* out = coeff + (int64_t)b * C * F * 2;
    out[(int64_t)f * 2] = o[0] * bo[0] * p;
    out[(int64_t)f * 2 + 1] = o[1] * bo[0] * p;
    if (!grads) return;
    const int P = K + nb;
    for (int i = 0; i < nf; ++i) {
        double re = 0, im = 0;
        for (int k = 0; k < P; ++k) {
            const double w = tm[k * nf + i];
            if (w == 0.0) continue;
            double fr, fi;
            if (k < K) {
                fr = o[2 * (1 + k)] * bo[0];
                fi = o[2 * (1 + k) + 1] * bo[0];
            } else {
                // bo[1 + k - K] by selects: a register array indexed at run time would live in scratch
                const int q = 1 + k - K;
                double bq = bo[1];
#pragma unroll
                for (int r = 2; r <= kMaxBlur; ++r) bq = q == r ? bo[r] : bq;
                fr = o[0] * bq;
                fi = o[1] * bq;
            }
            re += fr * w;
            im += fi * w;
        }
        out[((int64_t)(1 + i) * F + f) * 2] = re * p;
        out[((int64_t)(1 + i) * F + f) * 2 + 1] = im * p;
    }
}

// The grid of starts' coefficients: coeff (B, ns nw, F) = O_{b, i} x OTF_j x pix, i < ns outlines
// of page b (ohat rows b ns + i, no derivatives), j < nw blur rows.
__global__ void grid_coeff_kernel(const double* __restrict__ ohat, int ns, const double* __restrict__ blur, int nw,
                                  int kind, int nb, const double* __restrict__ fxs, const double* __restrict__ fys,
                                  const double* __restrict__ pix, int B, int F, double* __restrict__ coeff) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)B * F) return;
    const int b = (int)(t / F), f = (int)(t % F);
    const double p = pix[f];
    double bo[1 + kMaxBlur];
    for (int j = 0; j < nw; ++j) {
        otf_eval(kind, blur + (int64_t)j * nb, fxs[f], fys[f], false, bo);
        for (int i = 0; i < ns; ++i) {
            const double* o = ohat + (((int64_t)b * ns + i) * F + f) * 2;
            double* out = coeff + (((int64_t)b * ns * nw + (int64_t)i * nw + j) * F + f) * 2;
            out[0] = o[0] * bo[0] * p;
            out[1] = o[1] * bo[0] * p;
        }
    }
}

// Stack model, first product: left[b][c][slot][p][kx] = sum_ky Ey[qy][p][ky] G[ky][kx] with
// G = coeff[b][c][f] dbl[f] gamma[b][cls][ky][kx] where f = fmap[ky][kx] >= 0, else 0.
// slot runs over the nclass classes (cls_sel null) or is the page's one class cls_sel[b].
struct SynthArgs {
    const double* coeff;
    const int32_t* fmap;
    const double* dbl;
    const double* gamma;
    const int32_t* cls_sel;
    const double* Ey;
    const double* ExT;
    double* left;
    double* M;
    int B, C, F, U, npts, nky, nkx, nslot;
};

__global__ void synth_left(SynthArgs A) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = (int64_t)A.B * A.C * A.nslot * A.npts * A.nkx;
    if (t >= total) return;
    const int kx = (int)(t % A.nkx);
    const int p = (int)(t / A.nkx % A.npts);
    const int slot = (int)(t / ((int64_t)A.nkx * A.npts) % A.nslot);
    const int cc = (int)(t / ((int64_t)A.nkx * A.npts * A.nslot) % A.C);
    const int b = (int)(t / ((int64_t)A.nkx * A.npts * A.nslot * A.C));
    const int cls = A.cls_sel ? A.cls_sel[b] : slot;
    const int qy = cls / A.U;
    const double* ey = A.Ey + ((int64_t)qy * A.npts + p) * A.nky * 2;
    const double* cf = A.coeff + ((int64_t)b * A.C + cc) * A.F * 2;
    const double* gm = A.gamma + ((int64_t)b * A.U * A.U + cls) * A.nky * A.nkx * 2;
    double re = 0, im = 0;
    for (int iy = 0; iy < A.nky; ++iy) {
        const int f = A.fmap[iy * A.nkx + kx];
        if (f < 0) continue;
        const double d = A.dbl[f];
        const double ar = cf[2 * f] * d, ai = cf[2 * f + 1] * d;
        const double gr = gm[((int64_t)iy * A.nkx + kx) * 2], gi = gm[((int64_t)iy * A.nkx + kx) * 2 + 1];
        const double zr = ar * gr - ai * gi, zi = ar * gi + ai * gr;
        const double er = ey[2 * iy], ei = ey[2 * iy + 1];
        re += er * zr - ei * zi;
        im += er * zi + ei * zr;
    }
    A.left[t * 2] = re;
    A.left[t * 2 + 1] = im;
}

// Second product: M[b][c][slot][p1][p2] = Re sum_kx left[b][c][slot][p1][kx] ExT[qx][kx][p2].
__global__ void synth_right(SynthArgs A) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = (int64_t)A.B * A.C * A.nslot * A.npts * A.npts;
    if (t >= total) return;
    const int p2 = (int)(t % A.npts);
    const int p1 = (int)(t / A.npts % A.npts);
    const int64_t row = t / ((int64_t)A.npts * A.npts);  // (b, c, slot)
    const int slot = (int)(row % A.nslot);
    const int b = (int)(row / ((int64_t)A.nslot * A.C));
    const int cls = A.cls_sel ? A.cls_sel[b] : slot;
    const int qx = cls % A.U;
    const double* l = A.left + (row * A.npts + p1) * A.nkx * 2;
    const double* e = A.ExT + (int64_t)qx * A.nkx * A.npts * 2;
    double acc = 0;
    for (int kx = 0; kx < A.nkx; ++kx) {
        const double lr = l[2 * kx], li = l[2 * kx + 1];
        const double er = e[((int64_t)kx * A.npts + p2) * 2], ei = e[((int64_t)kx * A.npts + p2) * 2 + 1];
        acc += lr * er - li * ei;
    }
    A.M[t] = acc;
}

// Block-wide sums of M doubles, every thread gets the totals (fixed order: lanes by xor
// swaps, then waves in index order).
template <int M>
__device__ inline void block_sum(double (&v)[M], double* red) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int m = 0; m < M; ++m)
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) v[m] += __shfl_xor(v[m], off, 64);
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int m = 0; m < M; ++m) red[wv * M + m] = v[m];
    __syncthreads();
    const int nw = blockDim.x >> 6;
#pragma unroll
    for (int m = 0; m < M; ++m) {
        double t = red[m];
        for (int w = 1; w < nw; ++w) t += red[w * M + m];
        v[m] = t;
    }
    __syncthreads();
}

// Variable projection, per page (one block): ws, the weighted means of model and data, scale,
// constant, cost, and for the derivatives D_k = M_k scale sqrt(w) the projections
// alpha_k = D_k . u1, beta_k = (D_k - alpha_k u1) . u2 with u1 = sqrt(w) / sqrt(ws),
// u2 = sqrt(w) dm / sqrt(max(var, 1e-300)). stats (B, 8 + 2 kMaxFree):
// [ws, mw, yw, var, scale, const, cost, -, alpha..., beta...].
constexpr int kMaxFree = 16;
constexpr int kStats = 8 + 2 * kMaxFree;
constexpr int kProjBlock = 1024;  // one page's bins per block: the passes are latency bound
constexpr int kProjScalarBlock = 512;  // 35 FP64 moment accumulators per thread: room for 256 VGPRs

__global__ __launch_bounds__(kProjScalarBlock) void project_scalars(const double* __restrict__ Mall, const double* __restrict__ y,
                                                          const double* __restrict__ w, int C, int nbins,
                                                          double* __restrict__ stats) {
    constexpr int kMom = 3 + 2 * kMaxFree;
    __shared__ double red[(kProjScalarBlock / 64) * kMom];
    const int b = blockIdx.x;
    const double* M = Mall + (int64_t)b * C * nbins;
    const double* yb = y + (int64_t)b * nbins;
    const double* wb = w + (int64_t)b * nbins;
    const int nf = C - 1;
    // first moments: ws, sum w m, sum w y and per derivative sum w M_k, sum w M_k m
    double mo[kMom];
#pragma unroll
    for (int k = 0; k < kMom; ++k) mo[k] = 0.0;
    for (int i = threadIdx.x; i < nbins; i += blockDim.x) {
        const double wi = wb[i], mi = M[i];
        mo[0] += wi;
        mo[1] += wi * mi;
        mo[2] += wi * yb[i];
#pragma unroll
        for (int k = 0; k < kMaxFree; ++k)
            if (k < nf) {
                const double wmk = wi * M[(int64_t)(1 + k) * nbins + i];
                mo[3 + k] += wmk;
                mo[3 + kMaxFree + k] += wmk * mi;
            }
    }
    block_sum<kMom>(mo, red);
    const double ws = mo[0], mw = mo[1] / ws, yw = mo[2] / ws;
    double s2[2] = {0, 0};
    for (int i = threadIdx.x; i < nbins; i += blockDim.x) {
        const double dm = M[i] - mw, dy = yb[i] - yw;
        s2[0] += wb[i] * dm * dm;
        s2[1] += wb[i] * dm * dy;
    }
    block_sum<2>(s2, red);
    const double var = s2[0];
    const double scale = s2[1] / (var > 0 ? var : 1.0);
    const double cst = yw - scale * mw;
    double* st = stats + (int64_t)b * kStats;
    double cost[1] = {0.0};
    if (nf == 0) {  // without derivatives there is no rows kernel to sum the cost
        for (int i = threadIdx.x; i < nbins; i += blockDim.x) {
            const double res = scale * (M[i] - mw) - (yb[i] - yw);
            cost[0] += wb[i] * res * res;
        }
        block_sum<1>(cost, red);
    }
    if (threadIdx.x == 0) {
        // alpha_k = D_k . u1 = scale sum w M_k / sqrt(ws); beta_k = (D_k - alpha_k u1) . u2 with
        // D_k . u2 = scale sum w M_k (m - mw) / rvar and u1 . u2 = sum w (m - mw) / (sqrt(ws) rvar)
        const double rws = sqrt(ws), rvar = sqrt(fmax(var, 1e-300));
        const double u12 = (mo[1] - mw * ws) / (rws * rvar);
#pragma unroll
        for (int k = 0; k < kMaxFree; ++k) {
            const double al = k < nf ? scale * mo[3 + k] / rws : 0.0;
            const double du2 = k < nf ? scale * (mo[3 + kMaxFree + k] - mw * mo[3 + k]) / rvar : 0.0;
            st[8 + k] = al;
            st[8 + kMaxFree + k] = du2 - al * u12;
        }
        st[0] = ws;
        st[1] = mw;
        st[2] = yw;
        st[3] = var;
        st[4] = scale;
        st[5] = cst;
        st[6] = cost[0];
        st[7] = 0.0;
    }
}

// One block per (page, k): row k of A = D D^T and g_k = D_k . r (r = sqrt(w) res), D projected.
// out (B, 3 + nf + nf^2) = [cost, scale, const, g, A].
__global__ __launch_bounds__(kProjBlock) void project_rows(const double* __restrict__ Mall, const double* __restrict__ y,
                                                       const double* __restrict__ w, int C, int nbins,
                                                       const double* __restrict__ stats, double* __restrict__ out) {
    __shared__ double red[(kProjBlock / 64) * (kMaxFree + 1)];
    const int nf = C - 1;
    const int b = blockIdx.x / nf, k = blockIdx.x % nf;
    const double* M = Mall + (int64_t)b * C * nbins;
    const double* yb = y + (int64_t)b * nbins;
    const double* wb = w + (int64_t)b * nbins;
    const double* st = stats + (int64_t)b * kStats;
    const double ws = st[0], mw = st[1], yw = st[2], var = st[3], scale = st[4];
    const double rws = sqrt(ws), rvar = sqrt(fmax(var, 1e-300));
    double acc[kMaxFree + 1];
#pragma unroll
    for (int l = 0; l <= kMaxFree; ++l) acc[l] = 0;
    for (int i = threadIdx.x; i < nbins; i += blockDim.x) {
        const double sw = sqrt(wb[i]);
        const double dm = M[i] - mw;
        const double u1 = sw / rws, u2 = sw * dm / rvar;
        const double r = sw * (scale * dm - (yb[i] - yw));
        const double dk = M[(int64_t)(1 + k) * nbins + i] * (scale * sw) - st[8 + k] * u1 - st[8 + kMaxFree + k] * u2;
#pragma unroll
        for (int l = 0; l < kMaxFree; ++l) {
            if (l < nf) {
                const double dl =
                    M[(int64_t)(1 + l) * nbins + i] * (scale * sw) - st[8 + l] * u1 - st[8 + kMaxFree + l] * u2;
                acc[l] += dk * dl;
            }
        }
        acc[kMaxFree] += dk * r;
    }
    block_sum<kMaxFree + 1>(acc, red);
    double cost[1] = {0.0};
    if (k == 0) {
        for (int i = threadIdx.x; i < nbins; i += blockDim.x) {
            const double res = scale * (M[i] - mw) - (yb[i] - yw);
            cost[0] += wb[i] * res * res;
        }
        block_sum<1>(cost, red);
    }
    if (threadIdx.x == 0) {
        double* o = out + (int64_t)b * (3 + nf + nf * nf);
        if (k == 0) {
            o[0] = cost[0];
            o[1] = st[4];
            o[2] = st[5];
        }
        o[3 + k] = acc[kMaxFree];
#pragma unroll
        for (int l = 0; l < kMaxFree; ++l)
            if (l < nf) o[3 + nf + k * nf + l] = acc[l];
    }
}

// Costs only (no derivatives): one block per page, out (B, 3) = [cost, scale, const].
__global__ void project_costs(const double* __restrict__ stats, int B, double* __restrict__ out) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const double* st = stats + (int64_t)b * kStats;
    out[b * 3] = st[6];
    out[b * 3 + 1] = st[4];
    out[b * 3 + 2] = st[5];
}

// The grid of starts' table: one wave per (page, model c) over the bins of the page's class:
// cost of the best scale and constant. M (B, C, nb); y, w (B, ncls, nb); table (B, C).
__global__ void grid_cost_kernel(const double* __restrict__ M, const double* __restrict__ y,
                                 const double* __restrict__ w, const int32_t* __restrict__ cls_sel, int B, int C,
                                 int ncls, int nb, double* __restrict__ table) {
    const int64_t wid = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (wid >= (int64_t)B * C) return;
    const int b = (int)(wid / C);
    const double* m = M + wid * nb;
    const double* yb = y + ((int64_t)b * ncls + cls_sel[b]) * nb;
    const double* wb = w + ((int64_t)b * ncls + cls_sel[b]) * nb;
    double ws = 0, sm = 0, sy = 0;
    for (int i = lane; i < nb; i += 64) {
        ws += wb[i];
        sm += wb[i] * m[i];
        sy += wb[i] * yb[i];
    }
    ws = wave::sum_d(ws);
    const double mw = wave::sum_d(sm) / ws, yw = wave::sum_d(sy) / ws;
    double var = 0, cov = 0;
    for (int i = lane; i < nb; i += 64) {
        const double dm = m[i] - mw, dy = yb[i] - yw;
        var += wb[i] * dm * dm;
        cov += wb[i] * dm * dy;
    }
    var = wave::sum_d(var);
    cov = wave::sum_d(cov);
    const double scale = cov / (var > 0 ? var : 1.0);
    double cost = 0;
    for (int i = lane; i < nb; i += 64) {
        const double res = scale * (m[i] - mw) - (yb[i] - yw);
        cost += wb[i] * res * res;
    }
    cost = wave::sum_d(cost);
    if (lane == 0) table[wid] = cost;
}

// ---------------------------------------------------------------------------------- templates
// One thread per (row, band sample (iy, kx)): T^(f) = O^_polygon(f) x OTF x sinc(fx) sinc(fy)
// [x sinc(fx box) sinc(fy box)] x exp(-2 pi i (fx + fy) (N / 2) / Uu), written into the half
// spectrum full (rows, N, N / 2 + 1) at row (ky mod N), column kx; ky = iy for iy <= kmax, else
// iy - (2 kmax + 1). xy (rows, 2, n) polygon vertices; blur (rows, nb).
__global__ void band_kernel(const double* __restrict__ xy, int rows, int n, const double* __restrict__ blur, int kind,
                            int nb, int L, int Uu, int kmax, double box, double* __restrict__ full) {
    const int nky = 2 * kmax + 1, nkx = kmax + 1;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)rows * nky * nkx) return;
    const int r = (int)(t / ((int64_t)nky * nkx));
    const int iy = (int)(t / nkx % nky), kx = (int)(t % nkx);
    const int ky = iy <= kmax ? iy : iy - nky;
    const int N = L * Uu;
    const double fx = (double)kx / L, fy = (double)ky / L;
    const double* xs = xy + (int64_t)r * 2 * n;
    const double* ys = xs + n;
    double re = 0, im = 0, area = 0;
    for (int m = 0; m < n; ++m) {
        const int m2 = m + 1 < n ? m + 1 : 0;
        const double x0 = xs[m], y0 = ys[m], x1 = xs[m2], y1 = ys[m2];
        const double ex = x1 - x0, ey = y1 - y0;
        const double mx = 0.5 * (x0 + x1), my = 0.5 * (y0 + y1);
        area += x0 * y1 - x1 * y0;
        const double cr = fx * ey - fy * ex;
        const double w = cr * sinc_d(fx * ex + fy * ey);
        double sn, cs;
        sincos(-kTwoPi * (fx * mx + fy * my), &sn, &cs);
        re += w * cs;
        im += w * sn;
    }
    const double f2 = fx * fx + fy * fy;
    double orr, oi;
    if (f2 == 0.0) {
        orr = 0.5 * area;
        oi = 0.0;
    } else {
        orr = -im / (kTwoPi * f2);
        oi = re / (kTwoPi * f2);
    }
    double bo[1 + kMaxBlur];
    otf_eval(kind, blur + (int64_t)r * nb, fx, fy, false, bo);
    double amp = bo[0] * sinc_d(fx) * sinc_d(fy);
    if (box > 0) amp = amp * sinc_d(fx * box) * sinc_d(fy * box);
    double sn, cs;
    sincos(-kTwoPi * (fx + fy) * ((double)(N / 2) / Uu), &sn, &cs);
    const double vr = orr * amp, vi = oi * amp;
    const int row = ky >= 0 ? ky : N + ky;
    double* o = full + (((int64_t)r * N + row) * (N / 2 + 1) + kx) * 2;
    o[0] = vr * cs - vi * sn;
    o[1] = vr * sn + vi * cs;
}

// ------------------------------------------------------------------ device Levenberg-Marquardt
// The page fit's iterations without a round trip per step (onepass.py:504-623): per page the
// damped Gauss-Newton step with the bounds as an active set, the trial point's theta and blur
// row, the outline geometry of every trial theta, and after the trial's evaluation the prior
// terms, the gain ratio, the accept / reject rule, the stopping rules and the damping update.
// State and stage constants live on the device; the host reads one flag per iteration.
struct LmState {
    int B, P, nf, nshape, nb, npr;
    const double* fixed;  // (B, P): theta entries no free parameter drives
    const double* tm;     // (P, nf): free parameter -> theta entries (ties included)
    const double* lo;     // (nf)
    const double* hi;     // (nf)
    const double* PL;     // (npr, P) prior rows
    const double* pm;     // (npr) prior means
    const double* pw;     // (npr) prior weights 1 / sd^2
    const double* pj;     // (npr, nf) = PL tm
    const double* pa;     // (nf, nf) = pj^T diag(pw) pj
    const double* ev;     // (B, 3 + nf + nf^2) the trial's evaluation [cost, scale, const, g, A]
    double* q;            // (B, nf) accepted point
    double* qn;           // (B, nf) trial point
    double* theta;        // (B, P) trial theta
    double* blur;         // (B, nb) trial theta's blur part
    double* cost;         // (B)
    double* scale;        // (B)
    double* cst;          // (B)
    double* g;            // (B, nf)
    double* A;            // (B, nf, nf)
    double* lam;          // (B)
    double* nu;           // (B)
    int32_t* done;        // (B)
    int32_t* alldone;     // (1)
    double ftol, xtol;
};

// One 64-thread block per page. first = 1: the stage's start (qn = q); else, for a page not
// done, the bounded step: parameters on a bound whose downhill direction points outside are
// frozen, the damped system of the others (diagonal floor 1e-10 of the largest) is solved by LU
// with partial pivoting (the first largest pivot; a zero pivot gives no step), a parameter the
// step carries past a bound is put on it and the others are solved again (at most nf rounds).
// The system sits in LDS; the lanes build it and eliminate one row each, thread 0 picks the
// pivots and back-substitutes, so every element sees the same operations in the same order as
// a serial elimination. Then theta = fixed + tm qn and its blur part.
__global__ __launch_bounds__(64) void lm_trial_kernel(LmState S, int first) {
    __shared__ double M[kMaxFree * kMaxFree];
    __shared__ double rhs[kMaxFree], held[kMaxFree], step[kMaxFree];
    __shared__ int frozen[kMaxFree], idx[kMaxFree], over[kMaxFree];
    __shared__ int m_sh, piv_sh, flag_sh, any_sh;
    __shared__ double dmax_sh;
    const int b = blockIdx.x, t = threadIdx.x;
    if (b >= S.B) return;
    const int nf = S.nf;
    const double* q = S.q + (int64_t)b * nf;
    double* qn = S.qn + (int64_t)b * nf;
    if (first || S.done[b]) {
        if (t < nf) qn[t] = q[t];
    } else {
        const double* g = S.g + (int64_t)b * nf;
        const double* A = S.A + (int64_t)b * nf * nf;
        const double lam = S.lam[b];
        if (t < nf) {
            frozen[t] = (q[t] <= S.lo[t] && g[t] > 0) || (q[t] >= S.hi[t] && g[t] < 0);
            held[t] = 0.0;
            step[t] = 0.0;
        }
        __syncthreads();
        for (int round = 0; round < nf; ++round) {
            if (t == 0) {
                int m = 0;
                double dmax = -1e300;
                for (int i = 0; i < nf; ++i)
                    if (!frozen[i]) {
                        idx[m++] = i;
                        dmax = fmax(dmax, A[i * nf + i]);
                    }
                m_sh = m;
                dmax_sh = dmax;
            }
            __syncthreads();
            const int m = m_sh;
            if (m == 0) {
                if (t < nf) step[t] = held[t];
                break;
            }
            for (int e = t; e < m * m; e += blockDim.x) {
                const int a = e / m, c = e % m;
                const double v = A[idx[a] * nf + idx[c]];
                M[a * kMaxFree + c] = a == c ? v + lam * fmax(v, 1e-10 * dmax_sh) : v;
            }
            if (t < m) {
                double r = g[idx[t]];
                for (int j = 0; j < nf; ++j)
                    if (frozen[j]) r += A[idx[t] * nf + j] * held[j];
                rhs[t] = -r;
            }
            __syncthreads();
            bool ok = true;
            for (int c = 0; c < m; ++c) {
                if (t == 0) {
                    int piv = c;
                    double best = fabs(M[c * kMaxFree + c]);
                    for (int r = c + 1; r < m; ++r) {
                        const double v = fabs(M[r * kMaxFree + c]);
                        if (v > best) {
                            best = v;
                            piv = r;
                        }
                    }
                    piv_sh = piv;
                    flag_sh = best == 0.0;
                }
                __syncthreads();
                if (flag_sh) {
                    ok = false;
                    break;
                }
                const int piv = piv_sh;
                if (piv != c) {
                    if (t < m) {
                        const double tmp = M[c * kMaxFree + t];
                        M[c * kMaxFree + t] = M[piv * kMaxFree + t];
                        M[piv * kMaxFree + t] = tmp;
                    }
                    if (t == 0) {
                        const double tmp = rhs[c];
                        rhs[c] = rhs[piv];
                        rhs[piv] = tmp;
                    }
                }
                __syncthreads();
                if (t > c && t < m) {
                    const double f = M[t * kMaxFree + c] / M[c * kMaxFree + c];
                    M[t * kMaxFree + c] = f;
                    for (int k = c + 1; k < m; ++k) M[t * kMaxFree + k] -= f * M[c * kMaxFree + k];
                    rhs[t] -= f * rhs[c];
                }
                __syncthreads();
            }
            if (ok && t == 0)
                for (int r = m - 1; r >= 0; --r) {
                    double v = rhs[r];
                    for (int k = r + 1; k < m; ++k) v -= M[r * kMaxFree + k] * rhs[k];
                    rhs[r] = v / M[r * kMaxFree + r];
                }
            __syncthreads();
            if (t < nf) step[t] = held[t];
            __syncthreads();
            if (t < m) {
                const double v = ok ? rhs[t] : 0.0;
                step[idx[t]] = isfinite(v) ? v : 0.0;
            }
            __syncthreads();
            if (t < nf) {
                const bool free_t = !frozen[t];
                const bool up = free_t && (q[t] + step[t] > S.hi[t]);
                const bool dn = free_t && (q[t] + step[t] < S.lo[t]);
                over[t] = up ? 1 : (dn ? -1 : 0);
            }
            __syncthreads();
            if (t == 0) {
                int any = 0;
                for (int i = 0; i < nf; ++i) any |= over[i] != 0;
                any_sh = any;
            }
            __syncthreads();
            if (!any_sh) break;
            if (t < nf) {
                held[t] = over[t] > 0 ? S.hi[t] - q[t] : (over[t] < 0 ? S.lo[t] - q[t] : held[t]);
                frozen[t] = frozen[t] || over[t] != 0;
            }
            __syncthreads();
        }
        __syncthreads();
        if (t < nf) qn[t] = fmin(fmax(q[t] + step[t], S.lo[t]), S.hi[t]);
    }
    __syncthreads();
    if (b == 0 && t == 0 && !first) S.alldone[0] = 1;
    if (t < S.P) {
        double v = 0.0;
        for (int i = 0; i < nf; ++i) v += qn[i] * S.tm[t * nf + i];
        const double th = S.fixed[(int64_t)b * S.P + t] + v;
        S.theta[(int64_t)b * S.P + t] = th;
        if (t >= S.nshape) S.blur[(int64_t)b * S.nb + (t - S.nshape)] = th;
    }
}

// The outline at n equal angles of every page's trial theta (one block per page, threads over
// the points): geo (4, B, n) = [x, y, dx/dphi, dy/dphi], gk (B, n, 4 + 2 nk) = r dr/d(shape
// parameter), area (B) = sum r^2 dphi / 2. Superellipse of half axes e^th0, e^th1, power
// 1 + e^th2, turned by th3, times 1 + sum c_k cos(k phi) + s_k sin(k phi) over the harmonics ks.
__global__ void lm_geometry_kernel(const double* __restrict__ theta, int B, int P, int n, int nk, int k0, int k1, int k2,
                                   int k3, double* __restrict__ geo, double* __restrict__ gk, double* __restrict__ area) {
    __shared__ double red[kBlock / 64];
    const int b = blockIdx.x;
    const double* th = theta + (int64_t)b * P;
    const int K = 4 + 2 * nk;
    const int ks[4] = {k0, k1, k2, k3};
    const double a = exp(th[0]), bb = exp(th[1]);
    const double p = 1.0 + exp(th[2]);
    const double dphi = kTwoPi / n;
    double part = 0.0;
    for (int m = threadIdx.x; m < n; m += blockDim.x) {
        const double phi = kTwoPi * m / n;
        const double t = phi - th[3];
        const double ct = cos(t), st = sin(t);
        const double ua = fabs(ct / a), ub = fabs(st / bb);
        const double Aa = pow(ua, p), Bq = pow(ub, p);
        const double S = Aa + Bq;
        const double base = pow(S, -1.0 / p);
        const double sct = (ct > 0) - (ct < 0), sst = (st > 0) - (st < 0);
        const double dbase = -(1.0 / p) * pow(base, 1.0 + p) *
                             (p * pow(ua, p - 1.0) * sct * (-st / a) + p * pow(ub, p - 1.0) * sst * (ct / bb));
        double mod = 1.0, dmod = 0.0;
        for (int i = 0; i < nk; ++i) {
            const double ck = th[4 + 2 * i], sk = th[5 + 2 * i];
            const double kp = ks[i] * phi;
            mod += ck * cos(kp) + sk * sin(kp);
            dmod += -ck * ks[i] * sin(kp) + sk * ks[i] * cos(kp);
        }
        const double r = base * mod;
        const double dr = dbase * mod + base * dmod;
        const double la = ua > 0 ? Aa * log(ua) : 0.0;
        const double lb = ub > 0 ? Bq * log(ub) : 0.0;
        double* gm = gk + ((int64_t)b * n + m) * K;
        gm[0] = base * (Aa / S) * mod * r;
        gm[1] = base * (Bq / S) * mod * r;
        gm[2] = (p - 1.0) * base * (log(S) / (p * p) - (la + lb) / (p * S)) * mod * r;
        gm[3] = -dbase * mod * r;
        for (int i = 0; i < nk; ++i) {
            const double kp = ks[i] * phi;
            gm[4 + 2 * i] = base * cos(kp) * r;
            gm[5 + 2 * i] = base * sin(kp) * r;
        }
        const double cp = cos(phi), sp = sin(phi);
        geo[(int64_t)b * n + m] = r * cp;
        geo[((int64_t)B + b) * n + m] = r * sp;
        geo[(2 * (int64_t)B + b) * n + m] = dr * cp - r * sp;
        geo[(3 * (int64_t)B + b) * n + m] = dr * sp + r * cp;
        part += r * r;
    }
    part = wave::sum_d(part);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = part;
    __syncthreads();
    if (threadIdx.x == 0) {
        double tot = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) tot += red[w];
        area[b] = 0.5 * tot * dphi;
    }
}

// The priors at the trial theta: dev = PL theta - pm (npr <= 4 rows).
constexpr int kMaxPrior = 4;

__device__ inline double prior_dev(const LmState& S, int b, double (&dv)[kMaxPrior]) {
    const double* th = S.theta + (int64_t)b * S.P;
    double c = 0.0;
#pragma unroll
    for (int r = 0; r < kMaxPrior; ++r) {
        dv[r] = 0.0;
        if (r < S.npr) {
            double v = -S.pm[r];
            for (int p = 0; p < S.P; ++p) v += S.PL[r * S.P + p] * th[p];
            dv[r] = v;
            c += S.pw[r] * v * v;
        }
    }
    return c;
}

// g_i of the trial with the priors: ev's g_i + sum_r dev_r pw_r pj[r][i].
__device__ inline double prior_g(const LmState& S, const double* ev, const double (&dv)[kMaxPrior], int i) {
    double v = ev[3 + i];
#pragma unroll
    for (int r = 0; r < kMaxPrior; ++r)
        if (r < S.npr) v += dv[r] * S.pw[r] * S.pj[r * S.nf + i];
    return v;
}

// The stage's start: the state from the evaluation of q (one block, threads over pages).
__global__ void lm_init_kernel(LmState S) {
    for (int b = threadIdx.x; b < S.B; b += blockDim.x) {
        const int nf = S.nf;
        const double* ev = S.ev + (int64_t)b * (3 + nf + nf * nf);
        double dv[kMaxPrior];
        S.cost[b] = ev[0] + prior_dev(S, b, dv);
        for (int i = 0; i < nf; ++i) S.g[(int64_t)b * nf + i] = prior_g(S, ev, dv, i);
        for (int i = 0; i < nf * nf; ++i) S.A[(int64_t)b * nf * nf + i] = ev[3 + nf + i] + S.pa[i];
        S.scale[b] = ev[1];
        S.cst[b] = ev[2];
        S.lam[b] = 1e-3;
        S.nu[b] = 2.0;
        S.done[b] = 0;
    }
    if (threadIdx.x == 0) S.alldone[0] = 0;
}

// After the trial's evaluation, one 64-lane block per page: the gain ratio against the quadratic
// model at the accepted point, accept when the cost falls, the stopping rules (a small fall or
// step, a rejected step at damping > 1e7, a predicted fall below ftol cost), Nielsen's damping
// update. The lanes take the rows of g . dq and dq^T A dq; lm_trial set *alldone = 1 and an
// open page clears it.
__global__ __launch_bounds__(64) void lm_update_kernel(LmState S) {
    __shared__ int better_sh;
    const int b = blockIdx.x, t = threadIdx.x;
    if (b >= S.B) return;
    const int nf = S.nf;
    const double* ev = S.ev + (int64_t)b * (3 + nf + nf * nf);
    double* q = S.q + (int64_t)b * nf;
    const double* qn = S.qn + (int64_t)b * nf;
    double* g = S.g + (int64_t)b * nf;
    double* A = S.A + (int64_t)b * nf * nf;
    double dv[kMaxPrior];
    const double cost2 = ev[0] + prior_dev(S, b, dv);
    double gd = 0.0, quad = 0.0, qmax = 0.0, dmax = 0.0;
    if (t < nf) {
        const double di = qn[t] - q[t];
        gd = g[t] * di;
        qmax = fabs(q[t]);
        dmax = fabs(di);
        double row = 0.0;
        for (int j = 0; j < nf; ++j) row += A[t * nf + j] * (qn[j] - q[j]);
        quad = di * row;
    }
    gd = wave::sum_d(gd);
    quad = wave::sum_d(quad);
    qmax = wave::max_d(qmax);
    dmax = wave::max_d(dmax);
    const double cost = S.cost[b];
    const double pred = -(2.0 * gd + quad);
    const double rho = (cost - cost2) / fmax(pred, 1e-300);
    const bool was_done = S.done[b] != 0;
    const bool better = (cost2 < cost) && !was_done;
    const bool small = ((cost - cost2) <= S.ftol * cost) || (dmax <= S.xtol * (S.xtol + qmax));
    const double lam = S.lam[b];
    if (t == 0) better_sh = better;
    __syncthreads();  // every lane has read q, g and A before any is overwritten
    if (better) {
        if (t < nf) {
            q[t] = qn[t];
            g[t] = prior_g(S, ev, dv, t);
        }
        for (int i = t; i < nf * nf; i += blockDim.x) A[i] = ev[3 + nf + i] + S.pa[i];
    }
    if (t == 0) {
        if (better) {
            S.scale[b] = ev[1];
            S.cst[b] = ev[2];
            S.cost[b] = cost2;
        }
        const bool done = was_done || (better && small) || (!better && lam > 1e7) || (pred <= S.ftol * cost);
        S.done[b] = done ? 1 : 0;
        const double tt = 2.0 * rho - 1.0;
        const double grow = fmax(1.0 / 3.0, 1.0 - tt * tt * tt);
        S.lam[b] = better ? lam * grow : lam * S.nu[b];
        S.nu[b] = better ? 2.0 : S.nu[b] * 2.0;
        if (!done) S.alldone[0] = 0;
    }
}

inline unsigned blocks(int64_t n, int per = kBlock) { return (unsigned)((n + per - 1) / per); }

}  // namespace
}  // namespace spx

using namespace spx;

namespace {
LmState to_state(const SpxLmStage& t) {
    LmState S{};
    S.B = t.B;
    S.P = t.P;
    S.nf = t.nf;
    S.nshape = t.nshape;
    S.nb = t.nb;
    S.npr = t.npr;
    S.fixed = t.fixed;
    S.tm = t.tm;
    S.lo = t.lo;
    S.hi = t.hi;
    S.PL = t.PL;
    S.pm = t.pm;
    S.pw = t.pw;
    S.pj = t.pj;
    S.pa = t.pa;
    S.ev = t.ev;
    S.q = t.q;
    S.qn = t.qn;
    S.theta = t.theta;
    S.blur = t.blur;
    S.cost = t.cost;
    S.scale = t.scale;
    S.cst = t.cst;
    S.g = t.g;
    S.A = t.A;
    S.lam = t.lam;
    S.nu = t.nu;
    S.done = t.done;
    S.alldone = t.alldone;
    S.ftol = t.ftol;
    S.xtol = t.xtol;
    return S;
}
}  // namespace

extern "C" {

int spx_onepass_smooth(const float* image, int64_t H, int64_t W, const double* taps, int32_t radius, double* tmp,
                       double* out, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(image && taps && tmp && out, "spx_onepass_smooth: null pointer");
    SPX_REQUIRE(H > 0 && W > 0 && radius >= 0, "spx_onepass_smooth: bad shape or radius");
    hipStream_t st = as_stream(stream);
    hipLaunchKernelGGL((smooth_axis<0, float>), dim3(blocks(H * W)), dim3(kBlock), 0, st, image, H, W, taps, radius,
                       tmp);
    SPX_LAUNCHED();
    hipLaunchKernelGGL((smooth_axis<1, double>), dim3(blocks(H * W)), dim3(kBlock), 0, st, tmp, H, W, taps, radius,
                       out);
    SPX_LAUNCHED();
    return SPX_OK;
}

int spx_onepass_peaks(const double* sm, int64_t H, int64_t W, int32_t nms, double thr, uint8_t* flags,
                      int32_t* row_counts_out, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(sm && flags && row_counts_out, "spx_onepass_peaks: null pointer");
    SPX_REQUIRE(H > 0 && W > 0 && nms >= 0, "spx_onepass_peaks: bad shape");
    hipStream_t st = as_stream(stream);
    hipLaunchKernelGGL(peak_flags, dim3(blocks(H * W)), dim3(kBlock), 0, st, sm, H, W, (int)nms, thr, flags);
    SPX_LAUNCHED();
    hipLaunchKernelGGL(row_counts, dim3(blocks(H * 64)), dim3(kBlock), 0, st, flags, H, W, row_counts_out);
    SPX_LAUNCHED();
    return SPX_OK;
}

int spx_onepass_peak_list(const double* sm, const uint8_t* flags, const int32_t* row_offsets, int64_t H, int64_t W,
                          double base, double* seeds, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(sm && flags && row_offsets && seeds, "spx_onepass_peak_list: null pointer");
    SPX_REQUIRE(H > 0 && W > 0, "spx_onepass_peak_list: bad shape");
    hipLaunchKernelGGL(peak_list, dim3(blocks(H * 64)), dim3(kBlock), 0, as_stream(stream), sm, flags, row_offsets, H,
                       W, base, seeds);
    SPX_LAUNCHED();
    return SPX_OK;
}

int spx_onepass_count_above(const float* image, int64_t H, int64_t W, int32_t step, double thr,
                            unsigned long long* count, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(image && count, "spx_onepass_count_above: null pointer");
    SPX_REQUIRE(H > 0 && W > 0 && step > 0, "spx_onepass_count_above: bad shape or step");
    hipStream_t st = as_stream(stream);
    int rc = zero_launch(count, 1, st);
    if (rc != SPX_OK) return rc;
    const int64_t m = ((H + step - 1) / step) * ((W + step - 1) / step);
    hipLaunchKernelGGL(count_above, dim3(blocks(m)), dim3(kBlock), 0, st, image, H, W, (int)step, thr, count);
    SPX_LAUNCHED();
    return SPX_OK;
}

int spx_onepass_nearest(const double* c, int32_t n, const int32_t* order, const int32_t* cell_start, int32_t ncx,
                        int32_t ncy, double cs, double* out, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(c && order && cell_start && out, "spx_onepass_nearest: null pointer");
    SPX_REQUIRE(n > 0 && ncx > 0 && ncy > 0 && cs > 0, "spx_onepass_nearest: bad sizes");
    hipLaunchKernelGGL(nearest_kernel, dim3(blocks(n)), dim3(kBlock), 0, as_stream(stream), c, (int)n, order,
                       cell_start, (int)ncx, (int)ncy, cs, out);
    SPX_LAUNCHED();
    return SPX_OK;
}

int spx_onepass_stack(const float* image, int64_t H, int64_t W, const double* c, int32_t n, const uint8_t* use,
                      int32_t R, int32_t U, int32_t Ms, double* stack, double* wsum, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(image && c && use && stack && wsum, "spx_onepass_stack: null pointer");
    SPX_REQUIRE(H > 0 && W > 0 && n > 0 && R >= 0 && U > 0 && Ms > 0, "spx_onepass_stack: bad sizes");
    hipStream_t st = as_stream(stream);
    const int64_t nb = (int64_t)Ms * Ms;
    void* ws = nullptr;
    const size_t lds = (size_t)nb * (sizeof(unsigned long long) + sizeof(unsigned int));
    if (lds <= (size_t)kStackLdsBytes) {
        // per-workgroup bins in LDS: about 8 (hole, row) items per thread
        const int64_t items = (int64_t)n * (2 * R + 1);
        int G = (int)((items + 8 * kBlock - 1) / (8 * kBlock));
        G = G < 1 ? 1 : (G > 512 ? 512 : G);
        const int per = (int)((n + G - 1) / G);
        G = (int)((n + per - 1) / per);
        int rc = spx::workspace((size_t)G * 2 * nb * sizeof(unsigned long long), &ws, st);
        if (rc != SPX_OK) return rc;
        auto* partial = static_cast<unsigned long long*>(ws);
        hipLaunchKernelGGL(stack_acc_lds, dim3(G), dim3(kBlock), lds, st, image, H, W, c, (int)n, use, (int)R, (int)U,
                           (int)Ms, per, partial);
        SPX_LAUNCHED();
        hipLaunchKernelGGL(stack_reduce, dim3(blocks(nb)), dim3(kBlock), 0, st, partial, G, nb, stack, wsum);
        SPX_LAUNCHED();
        return SPX_OK;
    }
    int rc = spx::workspace((size_t)nb * 2 * sizeof(unsigned long long), &ws, st);
    if (rc != SPX_OK) return rc;
    auto* sums = static_cast<unsigned long long*>(ws);
    rc = zero_launch(sums, 2 * nb, st);
    if (rc != SPX_OK) return rc;
    hipLaunchKernelGGL(stack_acc, dim3(blocks((int64_t)n * (2 * R + 1))), dim3(kBlock), 0, st, image, H, W, c, (int)n,
                       use, (int)R, (int)U, (int)Ms, sums, sums + nb);
    SPX_LAUNCHED();
    hipLaunchKernelGGL(stack_finish, dim3(blocks(nb)), dim3(kBlock), 0, st, sums, sums + nb, nb, stack, wsum);
    SPX_LAUNCHED();
    return SPX_OK;
}

int spx_onepass_versus(const float* image, int64_t H, int64_t W, const double* c, int32_t n, const double* stack,
                       int32_t R, int32_t U, int32_t Ms, double core2, double* out, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(image && c && stack && out, "spx_onepass_versus: null pointer");
    SPX_REQUIRE(H > 0 && W > 0 && n > 0 && R >= 0 && U > 0 && Ms > 1, "spx_onepass_versus: bad sizes");
    hipLaunchKernelGGL(versus_kernel, dim3(blocks((int64_t)n * 64)), dim3(kBlock), 0, as_stream(stream), image, H, W, c,
                       (int)n, stack, (int)R, (int)U, (int)Ms, core2, out);
    SPX_LAUNCHED();
    return SPX_OK;
}

int spx_onepass_pair_hist(const double* c, const double* a, double amean, const uint8_t* use, int32_t n, int32_t U,
                          const int32_t* order, const int32_t* cell_start, int32_t ncx, int32_t ncy, double cs,
                          double reach, int32_t Ug, int32_t Ng, double* hist, double* class_counts,
                          spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(c && a && use && order && cell_start && hist, "spx_onepass_pair_hist: null pointer");
    SPX_REQUIRE(n > 0 && U > 0 && U <= 16 && ncx > 0 && ncy > 0 && cs > 0 && reach >= 0 && Ug > 0 && Ng > 0 &&
                    amean > 0,
                "spx_onepass_pair_hist: bad sizes");
    hipStream_t st = as_stream(stream);
    const int nclass = U * U;
    const int64_t total = (int64_t)nclass * Ng * Ng;
    void* ws = nullptr;
    int rc = spx::workspace((size_t)(total + nclass) * sizeof(unsigned long long), &ws, st);
    if (rc != SPX_OK) return rc;
    auto* acc = static_cast<unsigned long long*>(ws);
    unsigned long long* counts = acc + total;
    rc = zero_launch(acc, total + nclass, st);
    if (rc != SPX_OK) return rc;
    hipLaunchKernelGGL(pair_counts, dim3(blocks(n)), dim3(kBlock), 0, st, c, use, (int)n, (int)U, counts);
    SPX_LAUNCHED();
    hipLaunchKernelGGL(pair_splat, dim3(blocks((int64_t)n * 64)), dim3(kBlock), 0, st, c, a, amean, use, counts, (int)n,
                       (int)U, order, cell_start, (int)ncx, (int)ncy, cs, reach, (int)Ug, (int)Ng, acc);
    SPX_LAUNCHED();
    hipLaunchKernelGGL(fix_to_double, dim3(blocks(total)), dim3(kBlock), 0, st, acc, total, kPairFix, hist);
    SPX_LAUNCHED();
    if (class_counts != nullptr) {
        hipLaunchKernelGGL(fix_to_double, dim3(1), dim3(kBlock), 0, st, counts, (int64_t)nclass, 1.0, class_counts);
        SPX_LAUNCHED();
    }
    return SPX_OK;
}

int spx_onepass_pair_crop(const double* spec, int32_t nclass, int32_t Ng, int32_t Ug, int32_t L, int32_t kmax,
                          double* gamma, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(spec && gamma, "spx_onepass_pair_crop: null pointer");
    SPX_REQUIRE(nclass > 0 && Ng > 0 && Ug > 0 && L > 0 && kmax >= 0 && 2 * kmax + 1 <= Ng && kmax + 1 <= Ng / 2 + 1,
                "spx_onepass_pair_crop: bad sizes");
    const int64_t total = (int64_t)nclass * (2 * kmax + 1) * (kmax + 1);
    hipLaunchKernelGGL(pair_crop, dim3(blocks(total)), dim3(kBlock), 0, as_stream(stream), spec, (int)nclass, (int)Ng,
                       (int)Ug, (int)L, (int)kmax, gamma);
    SPX_LAUNCHED();
    return SPX_OK;
}

int spx_onepass_outline_hat(const double* geo, const double* gk, const double* area, int32_t rows, int32_t n,
                            int32_t K, double dphi, const double* fx, const double* fy, int32_t F, double* out,
                            spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(geo && area && fx && fy && out && (K == 0 || gk), "spx_onepass_outline_hat: null pointer");
    SPX_REQUIRE(rows > 0 && n > 0 && F > 0 && K >= 0 && K <= kMaxShape, "spx_onepass_outline_hat: bad sizes");
    const dim3 grid(blocks((int64_t)rows * F * 64));
    hipStream_t st = as_stream(stream);
    if (K == 0)
        hipLaunchKernelGGL(outline_hat_kernel<0>, grid, dim3(kBlock), 0, st, geo, gk, area, (int)rows, (int)n, 0, dphi,
                           fx, fy, (int)F, out);
    else if (K == 8)
        hipLaunchKernelGGL(outline_hat_kernel<8>, grid, dim3(kBlock), 0, st, geo, gk, area, (int)rows, (int)n, 8, dphi,
                           fx, fy, (int)F, out);
    else
        hipLaunchKernelGGL(outline_hat_kernel<-1>, grid, dim3(kBlock), 0, st, geo, gk, area, (int)rows, (int)n, (int)K,
                           dphi, fx, fy, (int)F, out);
    SPX_LAUNCHED();
    return SPX_OK;
}

static bool blur_ok(int32_t kind, int32_t nb) {
    return (kind == 0 && nb == 5) || (kind == 1 && nb == 3) || (kind == 2 && nb == 2);
}

int spx_onepass_coeff(const double* ohat, int32_t K, const double* blur, int32_t kind, int32_t nb, const double* tm,
                      int32_t nf, const double* fx, const double* fy, const double* pix, int32_t B, int32_t F,
                      double* coeff, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(ohat && blur && fx && fy && pix && coeff, "spx_onepass_coeff: null pointer");
    SPX_REQUIRE(blur_ok(kind, nb), "spx_onepass_coeff: unknown blur kind or parameter count");
    SPX_REQUIRE(B > 0 && F > 0 && K >= 0 && K <= kMaxShape && (tm == nullptr || (nf > 0 && nf <= kMaxFree)),
                "spx_onepass_coeff: bad sizes");
    hipLaunchKernelGGL(coeff_kernel, dim3(blocks((int64_t)B * F)), dim3(kBlock), 0, as_stream(stream), ohat, (int)K,
                       blur, (int)kind, (int)nb, tm, (int)nf, fx, fy, pix, (int)B, (int)F, coeff);
    SPX_LAUNCHED();
    return SPX_OK;
}

int spx_onepass_grid_coeff(const double* ohat, int32_t ns, const double* blur, int32_t nw, int32_t kind, int32_t nb,
                           const double* fx, const double* fy, const double* pix, int32_t B, int32_t F, double* coeff,
                           spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(ohat && blur && fx && fy && pix && coeff, "spx_onepass_grid_coeff: null pointer");
    SPX_REQUIRE(blur_ok(kind, nb), "spx_onepass_grid_coeff: unknown blur kind or parameter count");
    SPX_REQUIRE(B > 0 && F > 0 && ns > 0 && nw > 0, "spx_onepass_grid_coeff: bad sizes");
    hipLaunchKernelGGL(grid_coeff_kernel, dim3(blocks((int64_t)B * F)), dim3(kBlock), 0, as_stream(stream), ohat,
                       (int)ns, blur, (int)nw, (int)kind, (int)nb, fx, fy, pix, (int)B, (int)F, coeff);
    SPX_LAUNCHED();
    return SPX_OK;
}

int spx_onepass_synth(const double* coeff, int32_t B, int32_t C, int32_t F, const int32_t* fmap, const double* dbl,
                      const double* gamma, int32_t U, const int32_t* cls_sel, const double* Ey, const double* ExT,
                      int32_t npts, int32_t nky, int32_t nkx, double* M, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(coeff && fmap && dbl && gamma && Ey && ExT && M, "spx_onepass_synth: null pointer");
    SPX_REQUIRE(B > 0 && C > 0 && F > 0 && U > 0 && npts > 0 && nky > 0 && nkx > 0, "spx_onepass_synth: bad sizes");
    hipStream_t st = as_stream(stream);
    SynthArgs A{};
    A.coeff = coeff;
    A.fmap = fmap;
    A.dbl = dbl;
    A.gamma = gamma;
    A.cls_sel = cls_sel;
    A.Ey = Ey;
    A.ExT = ExT;
    A.M = M;
    A.B = B;
    A.C = C;
    A.F = F;
    A.U = U;
    A.npts = npts;
    A.nky = nky;
    A.nkx = nkx;
    A.nslot = cls_sel ? 1 : U * U;
    const int64_t nleft = (int64_t)B * C * A.nslot * npts * nkx;
    void* ws = nullptr;
    int rc = spx::workspace((size_t)nleft * 2 * sizeof(double), &ws, st);
    if (rc != SPX_OK) return rc;
    A.left = static_cast<double*>(ws);
    hipLaunchKernelGGL(synth_left, dim3(blocks(nleft)), dim3(kBlock), 0, st, A);
    SPX_LAUNCHED();
    hipLaunchKernelGGL(synth_right, dim3(blocks((int64_t)B * C * A.nslot * npts * npts)), dim3(kBlock), 0, st, A);
    SPX_LAUNCHED();
    return SPX_OK;
}

int spx_onepass_project(const double* M, const double* y, const double* w, int32_t B, int32_t C, int32_t nbins,
                        double* out, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(M && y && w && out, "spx_onepass_project: null pointer");
    SPX_REQUIRE(B > 0 && C > 0 && C - 1 <= kMaxFree && nbins > 0, "spx_onepass_project: bad sizes");
    hipStream_t st = as_stream(stream);
    void* ws = nullptr;
    int rc = spx::workspace((size_t)B * kStats * sizeof(double), &ws, st);
    if (rc != SPX_OK) return rc;
    auto* stats = static_cast<double*>(ws);
    hipLaunchKernelGGL(project_scalars, dim3(B), dim3(kProjScalarBlock), 0, st, M, y, w, (int)C, (int)nbins, stats);
    SPX_LAUNCHED();
    if (C > 1) {
        hipLaunchKernelGGL(project_rows, dim3(B * (C - 1)), dim3(kProjBlock), 0, st, M, y, w, (int)C, (int)nbins,
                           stats, out);
    } else {
        hipLaunchKernelGGL(project_costs, dim3(blocks(B)), dim3(kBlock), 0, st, stats, (int)B, out);
    }
    SPX_LAUNCHED();
    return SPX_OK;
}

int spx_onepass_grid_costs(const double* M, const double* y, const double* w, const int32_t* cls_sel, int32_t B,
                           int32_t C, int32_t ncls, int32_t nb, double* table, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(M && y && w && cls_sel && table, "spx_onepass_grid_costs: null pointer");
    SPX_REQUIRE(B > 0 && C > 0 && ncls > 0 && nb > 0, "spx_onepass_grid_costs: bad sizes");
    hipLaunchKernelGGL(grid_cost_kernel, dim3(blocks((int64_t)B * C * 64)), dim3(kBlock), 0, as_stream(stream), M, y, w,
                       cls_sel, (int)B, (int)C, (int)ncls, (int)nb, table);
    SPX_LAUNCHED();
    return SPX_OK;
}

int spx_onepass_band(const double* xy, int32_t rows, int32_t n, const double* blur, int32_t kind, int32_t nb,
                     int32_t L, int32_t Uu, int32_t kmax, double box_px, double* full, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(xy && blur && full, "spx_onepass_band: null pointer");
    SPX_REQUIRE(blur_ok(kind, nb), "spx_onepass_band: unknown blur kind or parameter count");
    SPX_REQUIRE(rows > 0 && n > 2 && L > 0 && Uu > 0 && kmax >= 0 && 2 * kmax + 1 <= L * Uu,
                "spx_onepass_band: bad sizes");
    const int64_t total = (int64_t)rows * (2 * kmax + 1) * (kmax + 1);
    hipLaunchKernelGGL(band_kernel, dim3(blocks(total)), dim3(kBlock), 0, as_stream(stream), xy, (int)rows, (int)n, blur,
                       (int)kind, (int)nb, (int)L, (int)Uu, (int)kmax, box_px, full);
    SPX_LAUNCHED();
    return SPX_OK;
}

int spx_onepass_lm_trial(const SpxLmStage* st, int32_t first, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(st != nullptr, "spx_onepass_lm_trial: null stage");
    SPX_REQUIRE(st->B > 0 && st->nf > 0 && st->nf <= kMaxFree && st->P >= st->nf, "spx_onepass_lm_trial: bad sizes");
    SPX_REQUIRE(st->npr >= 0 && st->npr <= 4 && st->P <= 64, "spx_onepass_lm_trial: at most 4 prior rows, 64 parameters");
    LmState S = to_state(*st);
    hipLaunchKernelGGL(lm_trial_kernel, dim3(S.B), dim3(64), 0, as_stream(stream), S, (int)first);
    SPX_LAUNCHED();
    return SPX_OK;
}

int spx_onepass_lm_geometry(const double* theta, int32_t B, int32_t P, int32_t n, int32_t nk, const int32_t* ks,
                            double* geo, double* gk, double* area, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(theta && ks && geo && gk && area, "spx_onepass_lm_geometry: null pointer");
    SPX_REQUIRE(B > 0 && n > 0 && nk >= 0 && nk <= 4 && P >= 4 + 2 * nk, "spx_onepass_lm_geometry: bad sizes");
    int k[4] = {0, 0, 0, 0};
    for (int i = 0; i < nk; ++i) k[i] = ks[i];
    hipLaunchKernelGGL(lm_geometry_kernel, dim3(B), dim3(kBlock), 0, as_stream(stream), theta, (int)B, (int)P, (int)n,
                       (int)nk, k[0], k[1], k[2], k[3], geo, gk, area);
    SPX_LAUNCHED();
    return SPX_OK;
}

int spx_onepass_lm_init(const SpxLmStage* st, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(st != nullptr, "spx_onepass_lm_init: null stage");
    SPX_REQUIRE(st->B > 0 && st->nf > 0 && st->nf <= kMaxFree && st->npr >= 0 && st->npr <= 4,
                "spx_onepass_lm_init: bad sizes");
    hipLaunchKernelGGL(lm_init_kernel, dim3(1), dim3(kBlock), 0, as_stream(stream), to_state(*st));
    SPX_LAUNCHED();
    return SPX_OK;
}

int spx_onepass_lm_update(const SpxLmStage* st, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(st != nullptr, "spx_onepass_lm_update: null stage");
    SPX_REQUIRE(st->B > 0 && st->nf > 0 && st->nf <= kMaxFree && st->npr >= 0 && st->npr <= 4,
                "spx_onepass_lm_update: bad sizes");
    hipLaunchKernelGGL(lm_update_kernel, dim3(st->B), dim3(64), 0, as_stream(stream), to_state(*st));
    SPX_LAUNCHED();
    return SPX_OK;
}

}  // extern "C"
