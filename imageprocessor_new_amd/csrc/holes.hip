// Outline hole stage on gfx950: every hole's centre, brightness, level and edge offset by one
// Gauss-Newton step per sweep against the complete page model, one colour class at a time.
//
//   page model  P(p) = sum_i a_i [T0 + d_i T1 + d_i^2 T2](p - c_i) + far light   (FP64, H x W)
//   hole i      data(p) = P(p) + level_i on its (2R+1)^2 window
//
// Follows the reference's GPU hole pass (subpx/_gpu/outline_onepass.py:252-397, 399-427):
// the same normal equations, weights, step bounds and render rules. The layout and the work
// split differ:
//   - templates are one interleaved FP32 grid of 8 floats per fine point,
//     [T0, T1, T2, Tx, Ty, 0, 0, 0], so a tap is one 16-byte load (+ one dword for Ty);
//   - `hole_step_kernel` gives each hole a group of G lanes (a wave64). The lanes stride over
//     the window, keep FP32 partial sums of J^T W J (as the reference's
//     per-row sums) and FP64 sums of J^T W r, reduce them in FP64 over the group with
//     cross-lane swaps in a fixed order (deterministic, no atomics), then solve the 5 x 5
//     (4 x 4 without the edge offset) system by FP64 Cholesky and apply the bounded step:
//     the reference's hole_normal + hole_update in one launch;
//   - `hole_render_kernel` is one thread per (hole, render offset inside the light's reach):
//     old light out, new light in, by FP64 atomics into the page model (one atomic when the
//     rounded centre did not move), as the reference.
#include "spx_internal.hpp"

namespace spx {
namespace {

// Keys cubic (a = -0.5) weights of the taps at -1, 0, 1, 2 for a fraction t in [0, 1).
__device__ __forceinline__ void keys4(float t, float* w) {
    const float x0 = 1.0f + t, x1 = t, x2 = 1.0f - t, x3 = 2.0f - t;
    w[0] = ((-0.5f * x0 + 2.5f) * x0 - 4.0f) * x0 + 2.0f;
    w[1] = (1.5f * x1 - 2.5f) * x1 * x1 + 1.0f;
    w[2] = (1.5f * x2 - 2.5f) * x2 * x2 + 1.0f;
    w[3] = ((-0.5f * x3 + 2.5f) * x3 - 4.0f) * x3 + 2.0f;
}
__device__ __forceinline__ void keys4d(double t, double* w) {
    const double x0 = 1.0 + t, x1 = t, x2 = 1.0 - t, x3 = 2.0 - t;
    w[0] = ((-0.5 * x0 + 2.5) * x0 - 4.0) * x0 + 2.0;
    w[1] = (1.5 * x1 - 2.5) * x1 * x1 + 1.0;
    w[2] = (1.5 * x2 - 2.5) * x2 * x2 + 1.0;
    w[3] = ((-0.5 * x3 + 2.5) * x3 - 4.0) * x3 + 2.0;
}

// Cubic sample of the template grid at fine coordinates (sx, sy); taps off the grid read 0.
// NG = 3: (T0, T1, T2); NG = 5: (T0, T1, T2, Tx, Ty). Branch free: an off-grid tap reads a
// clamped in-grid point with weight 0, so all 16 taps' loads can be issued back to back.
template <int NG>
__device__ __forceinline__ void sample_tmpl(const float* __restrict__ tmpl, int N, double sx, double sy, float* v) {
    const double fx = floor(sx), fy = floor(sy);
    const int ix = (int)fx, iy = (int)fy;
    float wx[4], wy[4];
    keys4((float)(sx - fx), wx);
    keys4((float)(sy - fy), wy);
    int xs[4], ys[4];
#pragma unroll
    for (int l = 0; l < 4; ++l) {
        const int xx = ix + l - 1, yy = iy + l - 1;
        if ((unsigned)xx >= (unsigned)N) wx[l] = 0.0f;
        if ((unsigned)yy >= (unsigned)N) wy[l] = 0.0f;
        xs[l] = min(max(xx, 0), N - 1);
        ys[l] = min(max(yy, 0), N - 1);
    }
#pragma unroll
    for (int g = 0; g < NG; ++g) v[g] = 0.0f;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        float row[NG];
#pragma unroll
        for (int g = 0; g < NG; ++g) row[g] = 0.0f;
#pragma unroll
        for (int l = 0; l < 4; ++l) {
            const float* p = tmpl + ((int64_t)ys[k] * N + xs[l]) * 8;
            const float4 q = *reinterpret_cast<const float4*>(p);
            row[0] += wx[l] * q.x;
            row[1] += wx[l] * q.y;
            row[2] += wx[l] * q.z;
            if constexpr (NG == 5) {
                row[3] += wx[l] * q.w;
                row[4] += wx[l] * p[4];
            }
        }
#pragma unroll
        for (int g = 0; g < NG; ++g) v[g] += wy[k] * row[g];
    }
}

template <int G>
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
    for (int off = G / 2; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

struct HoleArgs {
    const float* image;
    int64_t H, W, pitch;
    const double* page;
    const float* tmpl;
    int N, U, R;
    const int32_t* idx;
    int n;
    double* c;
    double* a;
    double* level;
    double* delta;
    double* c_old;
    double nv0, nv1, rscale;
    int size;
    double max_move, max_delta, a_med, relax;
    double* stat;
};

template <int G, int K>
__global__ __launch_bounds__(256) void hole_step_kernel(HoleArgs A) {
    const int slot = (int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G);
    const int gl = threadIdx.x & (G - 1);
    const bool live = slot < A.n;
    const int h = live ? A.idx[slot] : 0;
    const double x = live ? A.c[2 * (int64_t)h] : 0.0, y = live ? A.c[2 * (int64_t)h + 1] : 0.0;
    const double aa = live ? A.a[h] : 0.0, lev = live ? A.level[h] : 0.0, de = live ? A.delta[h] : 0.0;
    const int64_t rx = (int64_t)rint(x), ry = (int64_t)rint(y);
    const int side = 2 * A.R + 1;
    const int npx = live ? side * side : 0;
    const float fa = (float)aa, fde = (float)de;
    const double half_n = (double)(A.N / 2);

    float m[15];
#pragma unroll
    for (int q = 0; q < 15; ++q) m[q] = 0.0f;
    double rhs[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    double rss = 0.0, cnt = 0.0;
    for (int p = gl; p < npx; p += G) {
        const int oy = p / side - A.R, ox = p - (p / side) * side - A.R;
        const int64_t px = rx + ox, py = ry + oy;
        if (px < 0 || px >= A.W || py < 0 || py >= A.H) continue;
        float v[5];
        sample_tmpl<5>(A.tmpl, A.N, ((double)px - x) * A.U + half_n, ((double)py - y) * A.U + half_n, v);
        const double model = A.page[py * A.W + px] + lev;
        const double resid = (double)A.image[py * A.pitch + px] - model;
        double wgt = 1.0 / (A.nv0 + A.nv1 * fmax(model, 0.0));
        rss += resid * resid;
        cnt += 1.0;
        if (A.rscale > 0.0) {
            const double q = resid / A.rscale;
            wgt = wgt / (1.0 + q * q);
        }
        float J[5];
        J[0] = -fa * v[3];
        J[1] = -fa * v[4];
        J[2] = v[0] + fde * (v[1] + fde * v[2]);
        J[3] = 1.0f;
        J[4] = (K == 5) ? fa * (v[1] + 2.0f * fde * v[2]) : 0.0f;
        const float fw = (float)wgt;
        const double wr = wgt * resid;
        int q = 0;
#pragma unroll
        for (int r = 0; r < 5; ++r) {
            const float jw = J[r] * fw;
#pragma unroll
            for (int s = r; s < 5; ++s) m[q++] += jw * J[s];
            rhs[r] += (double)J[r] * wr;
        }
    }
    // group totals (every lane of the group ends with them)
    double S[22];
#pragma unroll
    for (int q = 0; q < 15; ++q) S[q] = group_sum<G>((double)m[q]);
#pragma unroll
    for (int r = 0; r < 5; ++r) S[15 + r] = group_sum<G>(rhs[r]);
    S[20] = group_sum<G>(rss);
    S[21] = group_sum<G>(cnt);
    if (!live) return;

    // K x K system (K = 5 with the edge offset, 4 without) with the 1e-9 trace ridge, FP64 Cholesky
    constexpr int k = K;
    double L[5][5];
    {
        int q = 0;
#pragma unroll
        for (int r = 0; r < 5; ++r)
#pragma unroll
            for (int s = r; s < 5; ++s) {
                L[s][r] = S[q];
                ++q;
            }
    }
    double tr = 0.0;
    for (int r = 0; r < k; ++r) tr += L[r][r];
    const double ridge = 1e-9 * tr / k;
    bool ok = true;
#pragma unroll
    for (int j = 0; j < 5; ++j) {
        if (j >= k) break;
        double d = L[j][j] + ridge;
#pragma unroll
        for (int t = 0; t < j; ++t) d -= L[j][t] * L[j][t];
        if (!(d > 0.0)) ok = false;
        const double lj = sqrt(fmax(d, 1e-300));
        L[j][j] = lj;
#pragma unroll
        for (int i = j + 1; i < 5; ++i) {
            if (i >= k) break;
            double e = L[i][j];
#pragma unroll
            for (int t = 0; t < j; ++t) e -= L[i][t] * L[j][t];
            L[i][j] = e / lj;
        }
    }
    double step[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (ok) {
        double z[5];
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            if (i >= k) break;
            double e = S[15 + i];
#pragma unroll
            for (int t = 0; t < i; ++t) e -= L[i][t] * z[t];
            z[i] = e / L[i][i];
        }
#pragma unroll
        for (int i = 4; i >= 0; --i) {
            if (i >= k) continue;
            double e = z[i];
#pragma unroll
            for (int t = i + 1; t < 5; ++t)
                if (t < k) e -= L[t][i] * step[t];
            step[i] = e / L[i][i];
        }
#pragma unroll
        for (int r = 0; r < 5; ++r)
            if (!(step[r] == step[r]) || fabs(step[r]) > 1e30) step[r] = 0.0;
    }
    step[2] *= A.relax;
    step[3] *= A.relax;
    step[4] *= A.relax;
    if (gl != 0) return;
    double* st = A.stat + 4 * (int64_t)slot;
    st[2] = sqrt(S[20] / (S[21] > 0.0 ? S[21] : 1.0));
    // (A^-1)_44 = 1 / L_44^2 for the last index of a Cholesky factor
    st[3] = (ok && K == 5) ? sqrt(fmax(1.0 / (L[4][4] * L[4][4]), 0.0)) : __builtin_nan("");
    A.c_old[2 * (int64_t)slot] = x;
    A.c_old[2 * (int64_t)slot + 1] = y;
    const double mx = fmin(fmax(step[0], -A.max_move), A.max_move);
    const double my = fmin(fmax(step[1], -A.max_move), A.max_move);
    A.c[2 * (int64_t)h] = x + mx;
    A.c[2 * (int64_t)h + 1] = y + my;
    st[0] = fmax(fabs(mx), fabs(my));
    const double an = fmin(fmax(aa + step[2], 0.5 * aa), 2.0 * aa);
    A.a[h] = fmin(fmax(an, 1e-3), 5.0 * A.a_med);
    A.level[h] = lev + step[3];
    st[1] = 0.0;
    if (K == 5) {
        const double dn = fmin(fmax(de + step[4], -A.max_delta), A.max_delta);
        st[1] = fabs(dn - de);
        A.delta[h] = dn;
    }
}

struct RenderArgs {
    double* page;
    unsigned long long* acc;  // (H, W) fixed-point light changes, zero before and after a render
    int64_t H, W;
    float* cache;
    const float* tmpl;
    int N, U, Rr;
    double rl2;
    const int32_t* idx;
    int n;
    const double* c_old;
    const double* c;
    const double* a;
    const double* delta;
};

// Offsets of the render window inside the light's reach, ox^2 + oy^2 <= rl2, row by row: row
// oy holds ox in [-w, w] with w the largest such |ox|; w = -1 (no offset) when oy^2 > rl2.
__host__ __device__ inline int reach_half_width(int oy, int Rr, double rl2) {
    if ((double)(oy * oy) > rl2) return -1;
    int w = 0;
    while (w < Rr && (double)((w + 1) * (w + 1) + oy * oy) <= rl2) ++w;
    return w;
}

// Order-independent sums: a value v enters a 64-bit integer accumulator as llrint(v * scale)
// (two's complement), and integer addition does not depend on the order of the atomics, so a
// sum is the same bit for bit in every run. The accumulator is folded into its FP64 / FP32
// destination afterwards by one addition per element.
constexpr double kFixPage = 4294967296.0;  // 2^32 per count: page light below 2^31 counts
constexpr double kFixGrid = 4294967296.0;  // far-light splats
constexpr double kFixStack = 16777216.0;   // 2^24: stack sums below 2^39

__device__ __forceinline__ void fix_add(unsigned long long* acc, double v, double scale) {
    atomicAdd(acc, (unsigned long long)(long long)llrint(v * scale));
}

// dst[i] += acc[i] / scale, acc[i] = 0, for n contiguous elements.
template <typename T>
__global__ void fix_fold_kernel(unsigned long long* acc, T* dst, int64_t n, double scale) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long v = (long long)acc[i];
    if (v != 0) {
        dst[i] = (T)((double)dst[i] + (double)v / scale);
        acc[i] = 0ull;
    }
}

__global__ void zero_u64_kernel(unsigned long long* p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// The render thread's offset (ox, oy) of hole i inside the reach (row table in LDS).
__device__ __forceinline__ void reach_offset(const int* row_start, const int* row_w, int side, int Rr, int q, int& ox,
                                             int& oy) {
    int lo = 0, hi = side;  // row r with row_start[r] <= q < row_start[r + 1]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (row_start[mid] <= q)
            lo = mid;
        else
            hi = mid;
    }
    oy = lo - Rr;
    ox = q - row_start[lo] - row_w[lo];
}

// One thread per (hole, offset inside the reach); the block builds the row table in LDS.
__global__ __launch_bounds__(256) void hole_render_kernel(RenderArgs A, int Pn) {
    __shared__ int row_start[512];
    __shared__ int row_w[512];
    const int side = 2 * A.Rr + 1;
    // the rows' half widths in parallel, then their running sum by one thread
    for (int r = threadIdx.x; r < side; r += blockDim.x) row_w[r] = reach_half_width(r - A.Rr, A.Rr, A.rl2);
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int r = 0; r < side; ++r) {
            row_start[r] = acc;
            acc += row_w[r] >= 0 ? 2 * row_w[r] + 1 : 0;  // a row beyond the reach holds no offset
        }
        row_start[side] = acc;
    }
    __syncthreads();
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (unsigned)A.n * (unsigned)Pn) return;
    const int i = (int)(t / (unsigned)Pn);
    const int q = (int)(t - (unsigned)i * (unsigned)Pn);
    int ox, oy;
    reach_offset(row_start, row_w, side, A.Rr, q, ox, oy);
    const int64_t slot = (int64_t)i * side * side + (int64_t)(oy + A.Rr) * side + (ox + A.Rr);
    const int h = A.idx[i];
    const float old = A.cache[slot];
    const int64_t pxo = (int64_t)rint(A.c_old[2 * (int64_t)i]) + ox;
    const int64_t pyo = (int64_t)rint(A.c_old[2 * (int64_t)i + 1]) + oy;
    const double x = A.c[2 * (int64_t)h], y = A.c[2 * (int64_t)h + 1];
    const int64_t px = (int64_t)rint(x) + ox, py = (int64_t)rint(y) + oy;
    float light = 0.0f;
    const bool inside = px >= 0 && px < A.W && py >= 0 && py < A.H;
    if (inside) {
        float v[3];
        const double half_n = (double)(A.N / 2);
        sample_tmpl<3>(A.tmpl, A.N, ((double)px - x) * A.U + half_n, ((double)py - y) * A.U + half_n, v);
        const double de = A.delta[h];
        light = (float)(A.a[h] * ((double)v[0] + de * ((double)v[1] + de * (double)v[2])));
    }
    // old light out, new light in, as fixed-point integers: the cached value leaves exactly
    // the integer it entered as
    if (pxo == px && pyo == py) {
        // same pixel (the rounded centre did not move): one atomic for old out + new in
        if (light != old)
            atomicAdd(A.acc + py * A.W + px, (unsigned long long)(llrint((double)light * kFixPage) -
                                                                  llrint((double)old * kFixPage)));
    } else {
        if (old != 0.0f) fix_add(A.acc + pyo * A.W + pxo, -(double)old, kFixPage);
        if (inside) fix_add(A.acc + py * A.W + px, (double)light, kFixPage);
    }
    A.cache[slot] = light;
}

// After hole_render_kernel: every pixel the render touched takes its accumulated change into
// the page once (whichever thread exchanges it first) and leaves the accumulator at zero.
__global__ __launch_bounds__(256) void hole_fold_kernel(RenderArgs A, int Pn) {
    __shared__ int row_start[512];
    __shared__ int row_w[512];
    const int side = 2 * A.Rr + 1;
    for (int r = threadIdx.x; r < side; r += blockDim.x) row_w[r] = reach_half_width(r - A.Rr, A.Rr, A.rl2);
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int r = 0; r < side; ++r) {
            row_start[r] = acc;
            acc += row_w[r] >= 0 ? 2 * row_w[r] + 1 : 0;
        }
        row_start[side] = acc;
    }
    __syncthreads();
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (unsigned)A.n * (unsigned)Pn) return;
    const int i = (int)(t / (unsigned)Pn);
    const int q = (int)(t - (unsigned)i * (unsigned)Pn);
    int ox, oy;
    reach_offset(row_start, row_w, side, A.Rr, q, ox, oy);
    const int h = A.idx[i];
    const int64_t pxo = (int64_t)rint(A.c_old[2 * (int64_t)i]) + ox;
    const int64_t pyo = (int64_t)rint(A.c_old[2 * (int64_t)i + 1]) + oy;
    const int64_t px = (int64_t)rint(A.c[2 * (int64_t)h]) + ox, py = (int64_t)rint(A.c[2 * (int64_t)h + 1]) + oy;
    if (pxo >= 0 && pxo < A.W && pyo >= 0 && pyo < A.H) {
        const long long v = (long long)atomicExch(A.acc + pyo * A.W + pxo, 0ull);
        if (v != 0) A.page[pyo * A.W + pxo] += (double)v / kFixPage;
    }
    if (px >= 0 && px < A.W && py >= 0 && py < A.H) {
        const long long v = (long long)atomicExch(A.acc + py * A.W + px, 0ull);
        if (v != 0) A.page[py * A.W + px] += (double)v / kFixPage;
    }
}

// Keys splat of every hole's far-light amplitude a (1 + 4 delta / cd0) on the padded pixel grid.
__global__ void far_splat_kernel(const double* c, const double* a, const double* delta, double cd0, int n, int mx,
                                 int my, int hp, int wp, unsigned long long* acc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = c[2 * (int64_t)i] + mx, y = c[2 * (int64_t)i + 1] + my;
    const double fx = floor(x), fy = floor(y);
    // all 4 x 4 taps inside the grid (compared in FP64, so a non-finite centre is dropped too)
    if (!(fx >= 1.0 && fx + 2.0 < (double)wp && fy >= 1.0 && fy + 2.0 < (double)hp)) return;
    const int ix = (int)fx, iy = (int)fy;
    double wx[4], wy[4];
    keys4d(x - fx, wx);
    keys4d(y - fy, wy);
    const double am = a[i] * (1.0 + 4.0 * delta[i] / cd0);
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int p = 0; p < 4; ++p)
            fix_add(acc + (int64_t)(iy + q - 1) * wp + (ix + p - 1), (double)(float)(am * wy[q] * wx[p]), kFixGrid);
}

// Analytic near templates of an erf-edged disk 1/2 erfc((rho - r - d) / (sqrt2 s)) and its far kernel.
__global__ void disk_templates_kernel(double radius, double sigma, int N, int U, double r0, double ramp, int B,
                                      float* tmpl, float* far) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const double kPi = 3.14159265358979323846;
    const double s2 = 1.4142135623730951 * sigma;
    const int64_t nn = (int64_t)N * N;
    if (t < nn) {
        const int iy = (int)(t / N), ix = (int)(t - (int64_t)iy * N);
        const double X = (double)(ix - N / 2) / U, Y = (double)(iy - N / 2) / U;
        const double rho = sqrt(X * X + Y * Y);
        const double u = (rho - radius) / s2;
        const double t0 = 0.5 * erfc(u);
        const double g = exp(-u * u) / (2.5066282746310002 * sigma);
        const double t2 = g * u / s2;
        const double sr = fmin(fmax((rho - r0) / ramp, 0.0), 1.0);
        const double w = 0.5 * (1.0 + cos(kPi * sr));
        const double dw = (rho > r0 && rho < r0 + ramp) ? -0.5 * kPi * sin(kPi * sr) / ramp : 0.0;
        const double drho = -g * w + t0 * dw;
        const double ux = rho > 0.0 ? X / rho : 0.0, uy = rho > 0.0 ? Y / rho : 0.0;
        float* o = tmpl + t * 8;
        o[0] = (float)(t0 * w);
        o[1] = (float)(g * w);
        o[2] = (float)(t2 * w);
        o[3] = (float)(drho * ux);
        o[4] = (float)(drho * uy);
        o[5] = 0.0f;
        o[6] = 0.0f;
        o[7] = 0.0f;
    } else if (far != nullptr && t < nn + (int64_t)B * B) {
        const int64_t q = t - nn;
        const int iy = (int)(q / B), ix = (int)(q - (int64_t)iy * B);
        const double X = (double)(ix - B / 2), Y = (double)(iy - B / 2);
        const double rho = sqrt(X * X + Y * Y);
        const double t0 = 0.5 * erfc((rho - radius) / s2);
        const double wf = 1.0 - 0.5 * (1.0 + cos(kPi * fmin(fmax((rho - r0) / ramp, 0.0), 1.0)));
        far[q] = (float)(t0 * wf);
    }
}

// Page outline fit (subpx/_outline/page.py, model.py): the outline's pixel coverage on the
// template's fine grid, per-window gathers for the stacks, noise line and hole starts, and the
// far light's splat of a, a d, a d^2.

// Coverage of fine pixel (iy, ix): the mean over its aa x aa sub-samples of a one-sample
// linear ramp across the edge, clamp(0.5 + (edge[idx] - rho) * scale, 0, 1), with the
// sub-samples' radius rho and angle-table index idx precomputed (model.py:149-159).
__global__ void outline_coverage_kernel(const double* __restrict__ edge, const double* __restrict__ rho,
                                        const int32_t* __restrict__ ang, int N, int aa, double scale,
                                        double* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)N * N) return;
    // blockIdx.y: outline of the batch (its table of 8192 radii, its output grid)
    edge += (int64_t)blockIdx.y * 8192;
    out += (int64_t)blockIdx.y * N * N;
    const int iy = (int)(t / N), ix = (int)(t - (int64_t)iy * N);
    const int64_t M = (int64_t)N * aa;
    double acc = 0.0;
    for (int q = 0; q < aa; ++q) {
        const int64_t row = ((int64_t)iy * aa + q) * M + (int64_t)ix * aa;
        for (int p = 0; p < aa; ++p) {
            const double v = 0.5 + (edge[ang[row + p]] - rho[row + p]) * scale;
            acc += fmin(fmax(v, 0.0), 1.0);
        }
    }
    out[t] = acc / (double)(aa * aa);
}

struct WindowArgs {
    const float* image;
    int64_t H, W, pitch;
    const double* page;  // null: 0
    const float* tmpl;   // null: no template values
    int N, U, R;
    const int32_t* idx;
    int n;
    const double* c;
    const double* a;
    const double* level;  // null: 0
    const double* delta;
    double radius_fine;  // > 0: bins farther than this (fine units) from the template centre get -1
    int32_t* bin;        // outputs, (n, (2R+1)^2) each; any may be null
    uint8_t* valid;
    double* data;
    double* model;
    double* own;
    double* v0;
};

// One thread per (hole, window pixel): the pixel's grey level, the page model plus the hole's
// level, the hole's own near light a (T0 + d T1 + d^2 T2), T0 alone, and the stack bin
// rint(sx), rint(sy) of the pixel's fine offset from the hole (page.py:206-216, 294-299, 488-505,
// 732-766). Pixels off the frame have valid 0, data and model at the clamped pixel, bin -1.
__global__ __launch_bounds__(256) void outline_window_kernel(WindowArgs A) {
    const int side = 2 * A.R + 1;
    const int64_t P = (int64_t)side * side;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)A.n * P) return;
    const int i = (int)(t / P);
    const int q = (int)(t - (int64_t)i * P);
    const int h = A.idx[i];
    const double x = A.c[2 * (int64_t)h], y = A.c[2 * (int64_t)h + 1];
    const int64_t px = (int64_t)rint(x) + (q % side) - A.R, py = (int64_t)rint(y) + (q / side) - A.R;
    const bool ok = px >= 0 && px < A.W && py >= 0 && py < A.H;
    const int64_t cx = px < 0 ? 0 : (px >= A.W ? A.W - 1 : px), cy = py < 0 ? 0 : (py >= A.H ? A.H - 1 : py);
    if (A.valid) A.valid[t] = ok ? 1 : 0;
    if (A.data) A.data[t] = (double)A.image[cy * A.pitch + cx];
    if (A.model) A.model[t] = (A.page ? A.page[cy * A.W + cx] : 0.0) + (A.level ? A.level[h] : 0.0);
    const double half_n = (double)(A.N / 2);
    const double sx = ((double)px - x) * A.U + half_n, sy = ((double)py - y) * A.U + half_n;
    if (A.tmpl && (A.own || A.v0)) {
        float v[3];
        sample_tmpl<3>(A.tmpl, A.N, sx, sy, v);
        const double de = A.delta[h];
        if (A.v0) A.v0[t] = (double)v[0];
        if (A.own) A.own[t] = A.a[h] * ((double)v[0] + de * ((double)v[1] + de * (double)v[2]));
    }
    if (A.bin) {
        const double bx = rint(sx), by = rint(sy);
        bool in = ok && bx >= 0.0 && bx < (double)A.N && by >= 0.0 && by < (double)A.N;
        if (in && A.radius_fine > 0.0) in = hypot(sx - half_n, sy - half_n) <= A.radius_fine;
        A.bin[t] = in ? (int32_t)by * A.N + (int32_t)bx : -1;
    }
}

// Stacked average hole on the fine grid (page.py:488-505, 732-766), one thread per (hole, window
// pixel) of the holes idx with good[h] (null: all): the pixel's bin rint(sx), rint(sy) (inside
// the grid, and within radius_fine of the centre when radius_fine > 0) receives
//   cleaned = 0: the grey level, weight 1 (the plain stack);
//   cleaned = 1: (data - page - level + own) / a with weight a^2 / (nv0 + nv1 max(page + level, 0)),
//                own = a (T0 + d T1 + d^2 T2) (the neighbour-cleaned stack),
// as FP64 atomic sums into sums[bin] and wsum[bin].
struct StackArgs {
    const float* image;
    int64_t H, W, pitch;
    const double* page;
    const float* tmpl;
    int N, U, R;
    const int32_t* idx;
    int n;
    const uint8_t* good;
    const double* c;
    const double* a;
    const double* level;
    const double* delta;
    double nv0, nv1, radius_fine;
    int cleaned;
    unsigned long long* sums;  // fixed-point accumulators (N x N each), folded after the launch
    unsigned long long* wsum;
};

__global__ __launch_bounds__(256) void outline_stack_kernel(StackArgs A) {
    const int side = 2 * A.R + 1;
    const int64_t P = (int64_t)side * side;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)A.n * P) return;
    const int i = (int)(t / P);
    const int q = (int)(t - (int64_t)i * P);
    const int h = A.idx[i];
    if (A.good && !A.good[h]) return;
    const double x = A.c[2 * (int64_t)h], y = A.c[2 * (int64_t)h + 1];
    const int64_t px = (int64_t)rint(x) + (q % side) - A.R, py = (int64_t)rint(y) + (q / side) - A.R;
    if (px < 0 || px >= A.W || py < 0 || py >= A.H) return;
    const double half_n = (double)(A.N / 2);
    const double sx = ((double)px - x) * A.U + half_n, sy = ((double)py - y) * A.U + half_n;
    const double bx = rint(sx), by = rint(sy);
    if (!(bx >= 0.0 && bx < (double)A.N && by >= 0.0 && by < (double)A.N)) return;
    if (A.radius_fine > 0.0 && !(hypot(sx - half_n, sy - half_n) <= A.radius_fine)) return;
    const int64_t b = (int64_t)by * A.N + (int64_t)bx;
    const double data = (double)A.image[py * A.pitch + px];
    if (!A.cleaned) {
        fix_add(A.sums + b, data, kFixStack);
        fix_add(A.wsum + b, 1.0, kFixStack);
        return;
    }
    float v[3];
    sample_tmpl<3>(A.tmpl, A.N, sx, sy, v);
    const double ah = A.a[h], de = A.delta[h];
    const double own = ah * ((double)v[0] + de * ((double)v[1] + de * (double)v[2]));
    const double model = A.page[py * A.W + px] + A.level[h];
    const double wgt = ah * ah / (A.nv0 + A.nv1 * fmax(model, 0.0));
    fix_add(A.sums + b, (data - model + own) / ah * wgt, kFixStack);
    fix_add(A.wsum + b, wgt, kFixStack);
}

// Keys splat of a, a d and a d^2 of every hole on three padded pixel grids (FP64, grid k at
// grid + k hp wp); a hole whose 4 x 4 taps leave the grid is dropped (page.py:371-397).
__global__ void outline_far_splat_kernel(const double* c, const double* a, const double* delta, int n, int mx, int my,
                                         int hp, int wp, unsigned long long* acc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = c[2 * (int64_t)i] + mx, y = c[2 * (int64_t)i + 1] + my;
    const double fx = floor(x), fy = floor(y);
    if (!(fx >= 1.0 && fx + 2.0 < (double)wp && fy >= 1.0 && fy + 2.0 < (double)hp)) return;
    const int ix = (int)fx, iy = (int)fy;
    double wx[4], wy[4];
    keys4d(x - fx, wx);
    keys4d(y - fy, wy);
    const double m[3] = {a[i], a[i] * delta[i], a[i] * delta[i] * delta[i]};
    const int64_t plane = (int64_t)hp * wp;
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            const double w = wy[q] * wx[p];
            const int64_t o = (int64_t)(iy + q - 1) * wp + (ix + p - 1);
#pragma unroll
            for (int k = 0; k < 3; ++k) fix_add(acc + k * plane + o, m[k] * w, kFixGrid);
        }
}

// k-th smallest of x[i * stride], i < n (no NaNs), by an 8-round radix select over the 64-bit
// keys of the doubles (sign-flipped so that they order as unsigned integers). One block per k.
__device__ __forceinline__ uint64_t order_key(double v) {
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double from_key(uint64_t k) {
    const uint64_t b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    return __longlong_as_double((long long)b);
}

struct KthList {
    int64_t k[8];
    int64_t col[8];
};

__global__ __launch_bounds__(1024) void select_kth_kernel(const double* x, int64_t n, int64_t stride, KthList ks,
                                                          double* out) {
    __shared__ unsigned int hist[256];
    __shared__ uint64_t s_prefix;
    __shared__ int64_t s_k;
    const double* xc = x + ks.col[blockIdx.x];
    const int lane = threadIdx.x & 63;
    if (threadIdx.x == 0) {
        s_prefix = 0;
        s_k = ks.k[blockIdx.x];
    }
    uint64_t mask = 0;
    for (int shift = 56; shift >= 0; shift -= 8) {
        for (int b = threadIdx.x; b < 256; b += blockDim.x) hist[b] = 0;
        __syncthreads();
        const uint64_t prefix = s_prefix;
        // eight values per thread per step, their loads issued together (clamped indices keep
        // them unconditional), then binned
        constexpr int kU = 8;
        for (int64_t base = 0; base < n; base += (int64_t)blockDim.x * kU) {
            double vals[kU];
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const int64_t i = base + (int64_t)u * blockDim.x + threadIdx.x;
                vals[u] = xc[(i < n ? i : n - 1) * stride];
            }
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const int64_t i = base + (int64_t)u * blockDim.x + threadIdx.x;
                bool part = false;
                unsigned int d = 0;
                if (i < n) {
                    const uint64_t key = order_key(vals[u]);
                    if ((key & mask) == prefix) {
                        part = true;
                        d = (unsigned int)((key >> shift) & 255u);
                    }
                }
                // one LDS atomic per distinct digit of the wave (the high digits are mostly equal)
                uint64_t pending = __ballot(part);
                while (pending) {
                    const int leader = __ffsll((unsigned long long)pending) - 1;
                    const unsigned int dl = __shfl(d, leader, 64);
                    const uint64_t same = __ballot(part && d == dl);
                    if (lane == leader) atomicAdd(&hist[dl], (unsigned int)__popcll(same));
                    if (part && d == dl) part = false;
                    pending &= ~same;
                }
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int64_t k = s_k, below = 0;
            int d = 0;
            for (; d < 255; ++d) {
                if (below + (int64_t)hist[d] > k) break;
                below += hist[d];
            }
            s_k = k - below;
            s_prefix = prefix | ((uint64_t)d << shift);
        }
        mask |= (uint64_t)255 << shift;
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = from_key(s_prefix);
}

}  // namespace

// n zeroed 64-bit accumulators in the stream's workspace.
int zeroed_accumulator(int64_t n, unsigned long long** out, hipStream_t stream) {
    void* ws = nullptr;
    int rc = workspace(sizeof(unsigned long long) * (size_t)n, &ws, stream);
    if (rc != SPX_OK) return rc;
    *out = static_cast<unsigned long long*>(ws);
    zero_u64_kernel<<<(unsigned)((n + 255) / 256), 256, 0, stream>>>(*out, n);
    SPX_LAUNCHED();
    return SPX_OK;
}

int outline_coverage_launch(const double* edge, const double* rho, const int32_t* ang, int N, int aa, double scale,
                            int nbatch, double* out, hipStream_t stream) {
    const int64_t nn = (int64_t)N * N;
    if (nbatch == 0) return SPX_OK;
    dim3 grid((unsigned)((nn + 255) / 256), (unsigned)nbatch);
    outline_coverage_kernel<<<grid, 256, 0, stream>>>(edge, rho, ang, N, aa, scale, out);
    SPX_LAUNCHED();
    return SPX_OK;
}

int outline_window_launch(const float* image, int64_t H, int64_t W, int64_t pitch, const double* page,
                          const float* tmpl, int N, int U, int R, const int32_t* idx, int n, const double* c,
                          const double* a, const double* level, const double* delta, double radius_fine,
                          int32_t* bin, uint8_t* valid, double* data, double* model, double* own, double* v0,
                          hipStream_t stream) {
    if (n == 0) return SPX_OK;
    const int64_t total = (int64_t)n * (2 * R + 1) * (2 * R + 1);
    if ((total + 255) / 256 >= ((int64_t)1 << 32)) {
        set_error("outline_window: too many (hole, pixel) pairs for one launch");
        return SPX_E_ARG;
    }
    WindowArgs A{image, H, W, pitch, page, tmpl, N, U, R, idx, n, c, a, level, delta, radius_fine,
                 bin, valid, data, model, own, v0};
    outline_window_kernel<<<(unsigned)((total + 255) / 256), 256, 0, stream>>>(A);
    SPX_LAUNCHED();
    return SPX_OK;
}

int outline_stack_launch(const float* image, int64_t H, int64_t W, int64_t pitch, const double* page,
                         const float* tmpl, int N, int U, int R, const int32_t* idx, int n, const uint8_t* good,
                         const double* c, const double* a, const double* level, const double* delta, double nv0,
                         double nv1, double radius_fine, int cleaned, double* sums, double* wsum, hipStream_t stream) {
    if (n == 0) return SPX_OK;
    const int64_t total = (int64_t)n * (2 * R + 1) * (2 * R + 1);
    if ((total + 255) / 256 >= ((int64_t)1 << 32)) {
        set_error("outline_stack: too many (hole, pixel) pairs for one launch");
        return SPX_E_ARG;
    }
    const int64_t nn = (int64_t)N * N;
    unsigned long long* acc = nullptr;
    int rc = zeroed_accumulator(2 * nn, &acc, stream);
    if (rc != SPX_OK) return rc;
    StackArgs A{image, H, W, pitch, page, tmpl, N, U, R, idx, n, good, c, a, level, delta, nv0, nv1, radius_fine,
                cleaned, acc, acc + nn};
    outline_stack_kernel<<<(unsigned)((total + 255) / 256), 256, 0, stream>>>(A);
    SPX_LAUNCHED();
    fix_fold_kernel<double><<<(unsigned)((nn + 255) / 256), 256, 0, stream>>>(acc, sums, nn, kFixStack);
    SPX_LAUNCHED();
    fix_fold_kernel<double><<<(unsigned)((nn + 255) / 256), 256, 0, stream>>>(acc + nn, wsum, nn, kFixStack);
    SPX_LAUNCHED();
    return SPX_OK;
}

int outline_far_splat_launch(const double* c, const double* a, const double* delta, int n, int mx, int my, int hp,
                             int wp, double* grid, hipStream_t stream) {
    if (n == 0) return SPX_OK;
    const int64_t cells = 3 * (int64_t)hp * wp;
    unsigned long long* acc = nullptr;
    int rc = zeroed_accumulator(cells, &acc, stream);
    if (rc != SPX_OK) return rc;
    outline_far_splat_kernel<<<(n + 255) / 256, 256, 0, stream>>>(c, a, delta, n, mx, my, hp, wp, acc);
    SPX_LAUNCHED();
    fix_fold_kernel<double><<<(unsigned)((cells + 255) / 256), 256, 0, stream>>>(acc, grid, cells, kFixGrid);
    SPX_LAUNCHED();
    return SPX_OK;
}

int select_kth_launch(const double* x, int64_t n, int64_t stride, const int64_t* cols, const int64_t* ks, int nk,
                      double* out, hipStream_t stream) {
    if (nk == 0) return SPX_OK;
    KthList kl{};
    for (int i = 0; i < nk; ++i) {
        kl.k[i] = ks[i];
        kl.col[i] = cols ? cols[i] : 0;
    }
    select_kth_kernel<<<nk, 1024, 0, stream>>>(x, n, stride, kl, out);
    SPX_LAUNCHED();
    return SPX_OK;
}

int hole_step_launch(const float* image, int64_t H, int64_t W, int64_t pitch, const double* page, const float* tmpl,
                     int N, int U, int R, const int32_t* idx, int n, double* c, double* a, double* level,
                     double* delta, double* c_old, double nv0, double nv1, double rscale, int size, double max_move,
                     double max_delta, double a_med, double relax, double* stat, hipStream_t stream) {
    if (n == 0) return SPX_OK;
    HoleArgs A{image, H, W, pitch, page, tmpl, N, U, R, idx, n, c, a, level, delta, c_old, nv0, nv1, rscale, size,
               max_move, max_delta, a_med, relax, stat};
    // one wave per hole: at 15 x 15 windows 64 lanes per hole ran a sweep 20 % faster than 16
    constexpr int G = 64;
    const int64_t threads = (int64_t)n * G;
    const int blocks = (int)((threads + 255) / 256);
    if (size)
        hole_step_kernel<G, 5><<<blocks, 256, 0, stream>>>(A);
    else
        hole_step_kernel<G, 4><<<blocks, 256, 0, stream>>>(A);
    SPX_LAUNCHED();
    return SPX_OK;
}

int hole_render_launch(double* page, unsigned long long* acc, int64_t H, int64_t W, float* cache, const float* tmpl,
                       int N, int U, int Rr, double rl2, const int32_t* idx, int n, const double* c_old, const double* c,
                       const double* a, const double* delta, hipStream_t stream) {
    if (n == 0) return SPX_OK;
    RenderArgs A{page, acc, H, W, cache, tmpl, N, U, Rr, rl2, idx, n, c_old, c, a, delta};
    int Pn = 0;
    for (int oy = -Rr; oy <= Rr; ++oy) {
        const int w = reach_half_width(oy, Rr, rl2);
        if (w >= 0) Pn += 2 * w + 1;
    }
    if (Pn == 0) return SPX_OK;
    const int64_t total = (int64_t)n * Pn;
    if (total >= ((int64_t)1 << 32)) {
        set_error("hole_render: too many (hole, offset) pairs for one launch");
        return SPX_E_ARG;
    }
    hole_render_kernel<<<(unsigned)((total + 255) / 256), 256, 0, stream>>>(A, Pn);
    SPX_LAUNCHED();
    hole_fold_kernel<<<(unsigned)((total + 255) / 256), 256, 0, stream>>>(A, Pn);
    SPX_LAUNCHED();
    return SPX_OK;
}

int far_splat_launch(const double* c, const double* a, const double* delta, double cd0, int n, int mx, int my, int hp,
                     int wp, float* grid, hipStream_t stream) {
    if (n == 0) return SPX_OK;
    const int64_t cells = (int64_t)hp * wp;
    unsigned long long* acc = nullptr;
    int rc = zeroed_accumulator(cells, &acc, stream);
    if (rc != SPX_OK) return rc;
    far_splat_kernel<<<(n + 255) / 256, 256, 0, stream>>>(c, a, delta, cd0, n, mx, my, hp, wp, acc);
    SPX_LAUNCHED();
    fix_fold_kernel<float><<<(unsigned)((cells + 255) / 256), 256, 0, stream>>>(acc, grid, cells, kFixGrid);
    SPX_LAUNCHED();
    return SPX_OK;
}

int disk_templates_launch(double radius, double sigma, int N, int U, double r0, double ramp, int B, float* tmpl,
                          float* far, hipStream_t stream) {
    const int64_t total = (int64_t)N * N + (far ? (int64_t)B * B : 0);
    disk_templates_kernel<<<(unsigned)((total + 255) / 256), 256, 0, stream>>>(radius, sigma, N, U, r0, ramp, B, tmpl,
                                                                                far);
    SPX_LAUNCHED();
    return SPX_OK;
}

}  // namespace spx
