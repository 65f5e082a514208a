// Deblending behind detect_centers(background="model"): a smooth background map plus one 2-D
// Gaussian per hole, alternated until the hole centres settle, then every hole re-refined on
// its cell with the background and all other holes removed.
//
// Reference: subpx/_cpu/deblend.py:48-522 (model, fits, renders, tile medians, iterations) and
// subpx/_gpu/deblend.py:155-178, 364-404 (fixed-point render, tile medians), subpx/centers.py:
// 225-316 (cells and the final refinement), subpx/_gpu/centers.py:33-111 (batch refinements).
// The reference runs the fits as NumPy / CuPy array operations over (cells, pixels, 6) stacks;
// here one wave carries one cell's Levenberg-Marquardt fit with FP64 wave reductions, renders
// add fixed-point integers (order-independent sums), tile medians select their middle values by
// radix selection on the 64-bit keys, and one workgroup refines one cleaned cell.
//
// FP64 throughout, without FMA contraction, in the operation order of the NumPy restatement
// that serves as the oracle.
#include <cmath>
#include <string>

#include "spx_internal.hpp"
#include "wave64.hpp"

namespace spx {
namespace {

constexpr int kBlock = 256;
constexpr double kFix = 16777216.0;  // render: 2^24 fixed-point units per count
constexpr double kSigmaMin = 0.5;

__device__ inline double nan_d() { return __longlong_as_double(0x7FF8000000000000ll); }

// block-wide reduction of M doubles (sum or max), every thread gets the results
template <int M, bool kMax>
__device__ inline void block_reduce(double (&v)[M], double* red) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int m = 0; m < M; ++m)
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            const double o = __shfl_xor(v[m], off, 64);
            v[m] = kMax ? fmax(v[m], o) : v[m] + o;
        }
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int m = 0; m < M; ++m) red[wv * M + m] = v[m];
    __syncthreads();
    const int nw = blockDim.x >> 6;
#pragma unroll
    for (int m = 0; m < M; ++m) {
        double t = red[m];
        for (int w = 1; w < nw; ++w) t = kMax ? fmax(t, red[w * M + m]) : t + red[w * M + m];
        v[m] = t;
    }
    __syncthreads();
}

// k-th smallest (0-based) of n doubles at vals (global or LDS), by rank counting: the value v
// with #{< v} <= k < #{<= v}. Every thread returns it.
__device__ inline double kth_by_rank(const double* vals, int n, int k, double* slot) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double v = vals[i];
        int less = 0, le = 0;
        for (int q = 0; q < n; ++q) {
            const double u = vals[q];
            less += u < v;
            le += u <= v;
        }
        if (less <= k && k < le) *slot = v;  // equal values write the same number
    }
    __syncthreads();
    const double r = *slot;
    __syncthreads();
    return r;
}

// NumPy's linear percentile q (in [0, 1]) of n values.
__device__ inline double percentile_linear(const double* vals, int n, double q, double* slot) {
#pragma clang fp contract(off)
    const double vi = q * (double)(n - 1);
    const int lo = (int)floor(vi);
    const int hi = lo + 1 < n ? lo + 1 : n - 1;
    const double t = vi - (double)lo;
    const double a = kth_by_rank(vals, n, lo, slot);
    const double b = kth_by_rank(vals, n, hi, slot);
    const double d = b - a;
    return t >= 0.5 ? b - d * (1.0 - t) : a + d * t;
}

// ---- the Gaussian ------------------------------------------------------------------------
// Parameters [A, x0, y0, u1, u2, u3]: model A exp(-(a^2 + b^2) / 2), a = e^u1 dx + u2 dy,
// b = e^u3 dy; J = d model / d p.
__device__ inline double gauss_model(const double* p, double dx, double dy) {
#pragma clang fp contract(off)
    const double e1 = exp(p[3]), u2 = p[4], e3 = exp(p[5]);
    const double a = e1 * dx + u2 * dy;
    const double b = e3 * dy;
    return p[0] * exp(-0.5 * (a * a + b * b));
}

__device__ inline double gauss_jac(const double* p, double dx, double dy, double (&J)[6]) {
#pragma clang fp contract(off)
    const double e1 = exp(p[3]), u2 = p[4], e3 = exp(p[5]);
    const double a = e1 * dx + u2 * dy;
    const double b = e3 * dy;
    const double shape = exp(-0.5 * (a * a + b * b));
    const double model = p[0] * shape;
    const double half = -0.5 * model;
    J[0] = shape;
    J[1] = half * (-2.0 * a * e1);
    J[2] = half * (-2.0 * a * u2 - 2.0 * b * e3);
    J[3] = half * (2.0 * a * e1 * dx);
    J[4] = half * (2.0 * a * dy);
    J[5] = half * (2.0 * b * e3 * dy);
    return model;
}

// ---- first fit's data and start ---------------------------------------------------------------
// One workgroup per cell: the image window at the cell's origin (0 outside the image), the cell
// border (cell pixels whose 3 x 3 neighbourhood leaves the cell or the ROI), its 10th-percentile
// grey level as the background, data = window - background on the cell (0 elsewhere), and the
// start: amplitude = the cell's maximum, centre = the seed, width from the second moment of the
// pixels above 0.2 of the maximum (reference subpx/_cpu/deblend.py:152-171, 467-476).
__global__ __launch_bounds__(kBlock) void deblend_init_kernel(const double* image, int64_t H, int64_t W,
                                                              const uint8_t* masks, const int32_t* origins,
                                                              const double* seeds, int N, int Hm, int Wm,
                                                              double* data, double* params) {
#pragma clang fp contract(off)
    __shared__ double red[4 * 4];
    __shared__ double slot;
    __shared__ int s_n;
    const int j = blockIdx.x;
    if (j >= N) return;
    const int HW = Hm * Wm;
    const uint8_t* m = masks + (int64_t)j * HW;
    double* d = data + (int64_t)j * HW;
    const int64_t ox = origins[2 * j], oy = origins[2 * j + 1];
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    // the border's grey levels, listed in the cell's own data slots (overwritten below)
    for (int k = threadIdx.x; k < HW; k += kBlock) {
        if (!m[k]) continue;
        const int y = k / Wm, x = k - y * Wm;
        bool inner = y > 0 && x > 0 && y < Hm - 1 && x < Wm - 1;
        for (int dy = -1; dy <= 1 && inner; ++dy)
            for (int dx = -1; dx <= 1 && inner; ++dx) inner = m[k + dy * Wm + dx] != 0;
        if (!inner) {
            const int64_t gy = oy + y, gx = ox + x;
            d[atomicAdd(&s_n, 1)] = (gy >= 0 && gy < H && gx >= 0 && gx < W) ? image[gy * W + gx] : 0.0;
        }
    }
    __syncthreads();
    const int nb = s_n;
    const double bg0 = nb > 0 ? percentile_linear(d, nb, 0.1, &slot) : 0.0;
    __syncthreads();
    double mx[1] = {-__builtin_inf()};
    for (int k = threadIdx.x; k < HW; k += kBlock) {
        double v = 0.0;
        if (m[k]) {
            const int y = k / Wm, x = k - y * Wm;
            const int64_t gy = oy + y, gx = ox + x;
            const double raw = (gy >= 0 && gy < H && gx >= 0 && gx < W) ? image[gy * W + gx] : 0.0;
            v = raw - bg0;
            mx[0] = fmax(mx[0], v);
        }
        d[k] = v;
    }
    block_reduce<1, true>(mx, red);
    const double peak = mx[0];
    const double sx = seeds[2 * j] - (double)ox, sy = seeds[2 * j + 1] - (double)oy;
    double acc[2] = {0.0, 0.0};
    for (int k = threadIdx.x; k < HW; k += kBlock) {
        if (!m[k] || !(d[k] > 0.2 * peak)) continue;
        const int y = k / Wm, x = k - y * Wm;
        const double dx = (double)x - sx, dy = (double)y - sy;
        acc[0] += d[k];
        acc[1] += d[k] * (dx * dx + dy * dy);
    }
    block_reduce<2, false>(acc, red);
    if (threadIdx.x == 0) {
        const double tot = fmax(acc[0], 1e-300);
        const double var = (acc[1] / tot) / 2.0;
        const double sig = sqrt(fmax(var, 0.25));
        double* p = params + (int64_t)j * 6;
        p[0] = fmax(peak, 1e-12);
        p[1] = sx;
        p[2] = sy;
        p[3] = -log(sig);
        p[4] = 0.0;
        p[5] = -log(sig);
    }
}

// ---- the Gaussian fits ------------------------------------------------------------------
// 6x6 solve by LU with partial pivoting (first largest pivot); false when a pivot is zero.
__device__ inline bool solve6(double (&A)[6][6], double (&b)[6], double (&x)[6]) {
#pragma clang fp contract(off)
    for (int c = 0; c < 6; ++c) {
        int piv = c;
        double best = fabs(A[c][c]);
        for (int r = c + 1; r < 6; ++r)
            if (fabs(A[r][c]) > best) {
                best = fabs(A[r][c]);
                piv = r;
            }
        if (!(best > 0.0)) return false;
        if (piv != c) {
            for (int k = 0; k < 6; ++k) {
                const double t = A[c][k];
                A[c][k] = A[piv][k];
                A[piv][k] = t;
            }
            const double t = b[c];
            b[c] = b[piv];
            b[piv] = t;
        }
        for (int r = c + 1; r < 6; ++r) {
            const double l = A[r][c] / A[c][c];
            for (int k = c; k < 6; ++k) A[r][k] = A[r][k] - l * A[c][k];
            b[r] = b[r] - l * b[c];
        }
    }
    for (int r = 5; r >= 0; --r) {
        double s = b[r];
        for (int k = r + 1; k < 6; ++k) s = s - A[r][k] * x[k];
        x[r] = s / A[r][r];
    }
    return true;
}

__device__ inline double wave_cost(const double* d, const uint8_t* m, int Hm, int Wm, const double* p) {
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63;
    double c = 0.0;
    for (int k = lane; k < Hm * Wm; k += 64) {
        if (!m[k]) continue;
        const int y = k / Wm, x = k - y * Wm;
        const double r = d[k] - gauss_model(p, (double)x - p[1], (double)y - p[2]);
        c += r * r;
    }
    return wave::sum_d(c);
}

// One wave per cell: LM on the cell's data (reference subpx/_cpu/deblend.py:174-254): damping
// lam diag + 1e-12, amplitude clipped to [0, 3 max |data|], log-widths to [-log cap, -log 0.5]
// (cap = the cell's equivalent radius, at least 1 px), a trial taken when its cost does not
// rise (lam / 3, else lam x 4); the fit ends after a relative decrease <= 1e-12, a step below
// 1e-7, lam > 1e12 or max_iter steps. Every reduced value is uniform (readfirstlane), so the
// wave's control flow is uniform.
__global__ __launch_bounds__(kBlock) void gauss_fit_kernel(const double* data, const uint8_t* masks, int N, int Hm,
                                                           int Wm, const double* init, int max_iter, double* out) {
#pragma clang fp contract(off)
    const int j = (int)(blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6));
    if (j >= N) return;
    const int lane = threadIdx.x & 63;
    const int HW = Hm * Wm;
    const double* d = data + (int64_t)j * HW;
    const uint8_t* m = masks + (int64_t)j * HW;
    double p[6];
    for (int i = 0; i < 6; ++i) p[i] = wave::uniform_d(init[(int64_t)j * 6 + i]);
    double cnt = 0.0, amax = 0.0;
    for (int k = lane; k < HW; k += 64)
        if (m[k]) {
            cnt += 1.0;
            amax = fmax(amax, fabs(d[k]));
        }
    cnt = wave::sum_d(cnt);
    amax = wave::max_d(amax);
    const double cap = fmax(1.0, sqrt(cnt / 3.141592653589793));
    const double amp_cap = 3.0 * fmax(amax, 1e-12);
    const double low = -log(cap), high = -log(kSigmaMin);
    double cost = wave_cost(d, m, Hm, Wm, p);
    double lam = 1e-3;
    if (isfinite(cost)) {
        for (int it = 0; it < max_iter; ++it) {
            double acc[27];
            for (int i = 0; i < 27; ++i) acc[i] = 0.0;
            for (int k = lane; k < HW; k += 64) {
                if (!m[k]) continue;
                const int y = k / Wm, x = k - y * Wm;
                double J[6];
                const double model = gauss_jac(p, (double)x - p[1], (double)y - p[2], J);
                const double r = d[k] - model;
                int q = 0;
                for (int a = 0; a < 6; ++a)
                    for (int b = a; b < 6; ++b) acc[q++] += J[a] * J[b];
                for (int a = 0; a < 6; ++a) acc[21 + a] += J[a] * r;
            }
            double A[6][6], g[6], step[6];
            int q = 0;
            for (int a = 0; a < 6; ++a)
                for (int b = a; b < 6; ++b) {
                    const double s = wave::sum_d(acc[q++]);
                    A[a][b] = s;
                    A[b][a] = s;
                }
            for (int a = 0; a < 6; ++a) g[a] = wave::sum_d(acc[21 + a]);
            for (int a = 0; a < 6; ++a) A[a][a] = A[a][a] + (lam * A[a][a] + 1e-12);
            if (!solve6(A, g, step))
                for (int a = 0; a < 6; ++a) step[a] = 0.0;  // a singular damped matrix: no move
            double trial[6];
            for (int a = 0; a < 6; ++a) trial[a] = p[a] + step[a];
            trial[0] = fmin(fmax(trial[0], 0.0), amp_cap);
            trial[3] = fmin(fmax(trial[3], low), high);
            trial[5] = fmin(fmax(trial[5], low), high);
            const double nc = wave_cost(d, m, Hm, Wm, trial);
            const bool better = isfinite(nc) && nc <= cost;
            double smax = 0.0;
            for (int a = 0; a < 6; ++a) smax = fmax(smax, fabs(step[a]));
            bool done = smax < 1e-7;
            if (better) {
                done = done || (cost - nc <= 1e-12 * fmax(cost, 1e-300));
                for (int a = 0; a < 6; ++a) p[a] = trial[a];
                cost = nc;
                lam = lam / 3.0;
            } else {
                lam = lam * 4.0;
            }
            if (done || lam > 1e12) break;
        }
    }
    if (lane < 6) out[(int64_t)j * 6 + lane] = p[lane];
}

// ---- renders and cleaned cells ----------------------------------------------------------
// One workgroup per hole: its Gaussian at global pixels within `radius` of its centre in x and
// y (reference subpx/_cpu/deblend.py:74-107), added as 2^-24 fixed-point integers, so the sum
// does not depend on the order of the atomics.
__global__ __launch_bounds__(kBlock) void gauss_render_kernel(const double* pg, const double* radius, int N,
                                                              int64_t H, int64_t W, unsigned long long* acc) {
#pragma clang fp contract(off)
    const int j = blockIdx.x;
    if (j >= N) return;
    const double* p = pg + (int64_t)j * 6;
    const double rad = radius[j];
    bool ok = isfinite(rad);
    for (int i = 0; i < 6; ++i) ok = ok && isfinite(p[i]);
    if (!ok) return;
    const int r = (int)rad;
    const int64_t fx = (int64_t)floor(p[1]), fy = (int64_t)floor(p[2]);
    const int side = 2 * r + 1;
    for (int k = threadIdx.x; k < side * side; k += kBlock) {
        const int oy = k / side, ox = k - oy * side;
        const int64_t X = fx - r + ox, Y = fy - r + oy;
        if (X < 0 || X >= W || Y < 0 || Y >= H) continue;
        const double dx = (double)X - p[1], dy = (double)Y - p[2];
        if (!(fabs(dx) <= rad && fabs(dy) <= rad)) continue;
        const double v = gauss_model(p, dx, dy);
        const long long q = llrint(v * kFix);
        if (q != 0) atomicAdd(&acc[Y * W + X], (unsigned long long)q);
    }
}

__global__ void fixed_to_double(const unsigned long long* acc, int64_t n, double* out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (double)(long long)acc[i] / kFix;
}

__global__ void zero_u64(unsigned long long* p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

// Cell pixels minus the background map and every other hole's model: window(image) -
// window(background) - (window(model) - own truncated Gaussian); 0 off the cell.
__global__ void deblend_clean_kernel(const double* image, const double* background, const double* model,
                                     const double* pg, const double* radius, const uint8_t* masks,
                                     const int32_t* origins, int N, int Hm, int Wm, int64_t H, int64_t W,
                                     double* out) {
#pragma clang fp contract(off)
    const int64_t total = (int64_t)N * Hm * Wm;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        if (!masks[i]) {
            out[i] = 0.0;
            continue;
        }
        const int j = (int)(i / ((int64_t)Hm * Wm));
        const int k = (int)(i - (int64_t)j * Hm * Wm);
        const int y = k / Wm, x = k - y * Wm;
        const int64_t gx = origins[2 * j] + x, gy = origins[2 * j + 1] + y;
        const bool in = gx >= 0 && gx < W && gy >= 0 && gy < H;
        const double raw = in ? image[gy * W + gx] : 0.0;
        const double bgv = in ? background[gy * W + gx] : 0.0;
        const double mdl = in ? model[gy * W + gx] : 0.0;
        const double* p = pg + (int64_t)j * 6;
        const double dx = (double)gx - p[1], dy = (double)gy - p[2];
        const double rad = radius[j];
        const double own = (fabs(dx) <= rad && fabs(dy) <= rad) ? gauss_model(p, dx, dy) : 0.0;
        out[i] = (raw - bgv) - (mdl - own);
    }
}

// ---- background map ---------------------------------------------------------------------
__device__ inline unsigned long long order_key(double v) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__device__ inline double key_value(unsigned long long k) {
    const unsigned long long u = (k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
    return __longlong_as_double((long long)u);
}

// k-th smallest key of n keys at `keys` by radix selection, 8 bits at a time from the top.
__device__ inline unsigned long long radix_select(const unsigned long long* keys, int n, int k, int* hist,
                                                  unsigned long long* s_prefix, int* s_k) {
    if (threadIdx.x == 0) {
        *s_prefix = 0ull;
        *s_k = k;
    }
    unsigned long long mask = 0ull;
    for (int shift = 56; shift >= 0; shift -= 8) {
        for (int b = threadIdx.x; b < 256; b += blockDim.x) hist[b] = 0;
        __syncthreads();
        const unsigned long long prefix = *s_prefix;
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            const unsigned long long key = keys[i];
            if ((key & mask) == prefix) atomicAdd(&hist[(key >> shift) & 255ull], 1);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int kk = *s_k, b = 0;
            while (b < 255 && kk >= hist[b]) {
                kk -= hist[b];
                ++b;
            }
            *s_k = kk;
            *s_prefix = prefix | ((unsigned long long)b << shift);
        }
        mask |= 255ull << shift;
        __syncthreads();
    }
    return *s_prefix;
}

// One workgroup per tile: the residual image - model over the tile's pixels that are usable
// (model < model_below and dist >= dist_atleast; pass +inf / -inf for every pixel), their count
// and their median (the mean of the middle pair for an even count; NaN without pixels).
__global__ __launch_bounds__(kBlock) void tile_median_kernel(const double* image, const double* model,
                                                             const double* dist, int64_t H, int64_t W, int tile,
                                                             int nx, double model_below, double dist_atleast,
                                                             unsigned long long* scratch, double* medians,
                                                             long long* counts) {
#pragma clang fp contract(off)
    __shared__ int hist[256];
    __shared__ int s_n;
    __shared__ unsigned long long s_prefix;
    __shared__ int s_k;
    const int t = blockIdx.x;
    const int ty = t / nx, tx = t - ty * nx;
    unsigned long long* keys = scratch + (int64_t)t * tile * tile;
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    const int64_t y0 = (int64_t)ty * tile, x0 = (int64_t)tx * tile;
    for (int k = threadIdx.x; k < tile * tile; k += kBlock) {
        const int yy = k / tile, xx = k - yy * tile;
        const int64_t gy = y0 + yy, gx = x0 + xx;
        if (gy >= H || gx >= W) continue;
        const int64_t i = gy * W + gx;
        if (!(model[i] < model_below) || !(dist[i] >= dist_atleast)) continue;
        keys[atomicAdd(&s_n, 1)] = order_key(image[i] - model[i]);
    }
    __syncthreads();
    const int c = s_n;
    double med = nan_d();
    if (c > 0) {
        const double lo = key_value(radix_select(keys, c, (c - 1) / 2, hist, &s_prefix, &s_k));
        const double hi = key_value(radix_select(keys, c, c / 2, hist, &s_prefix, &s_k));
        med = (lo + hi) / 2.0;
    }
    if (threadIdx.x == 0) {
        medians[t] = med;
        counts[t] = c;
    }
}

// NumPy's interp at x of the points (xp[i], fp[i]), i < n (xp increasing).
__device__ inline double np_interp(double x, const double* xp, const double* fp, int n) {
#pragma clang fp contract(off)
    if (n == 1 || x <= xp[0]) return x < xp[0] ? fp[0] : (x == xp[0] ? fp[0] : fp[n - 1]);
    if (x >= xp[n - 1]) return fp[n - 1];
    int lo = 0, hi = n - 1;  // xp[lo] <= x < xp[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (xp[mid] <= x) lo = mid;
        else hi = mid;
    }
    if (xp[lo] == x) return fp[lo];
    const double slope = (fp[lo + 1] - fp[lo]) / (xp[lo + 1] - xp[lo]);
    double r = slope * (x - xp[lo]) + fp[lo];
    if (r != r) {
        r = slope * (x - xp[lo + 1]) + fp[lo + 1];
        if (r != r && fp[lo] == fp[lo + 1]) r = fp[lo];
    }
    return r;
}

// The background map from the tile medians (ny, nx): each tile row interpolated along x
// between the tile centres xc (np.interp, clamped), then linearly between the two tile rows
// around the pixel's row by the tile-row centres yc (reference subpx/_cpu/deblend.py:294-320).
__global__ void tile_map_kernel(const double* med, const double* yc, const double* xc, int ny, int nx, int64_t H,
                                int64_t W, double* out) {
#pragma clang fp contract(off)
    const int64_t total = H * W;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < total; q += (int64_t)gridDim.x * blockDim.x) {
        const int64_t y = q / W, x = q - y * W;
        const double fxv = (double)x, fyv = (double)y;
        if (ny == 1) {
            out[q] = np_interp(fxv, xc, med, nx);
            continue;
        }
        int lo = 0, hi = ny;  // first tile row whose centre is >= y, clipped to [1, ny - 1]
        while (lo < hi) {
            const int mid = (lo + hi) / 2;
            if (yc[mid] < fyv) lo = mid + 1;
            else hi = mid;
        }
        const int j1 = lo < 1 ? 1 : (lo > ny - 1 ? ny - 1 : lo);
        double t = (fyv - yc[j1 - 1]) / (yc[j1] - yc[j1 - 1]);
        t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
        const double r0 = np_interp(fxv, xc, med + (int64_t)(j1 - 1) * nx, nx);
        const double r1 = np_interp(fxv, xc, med + (int64_t)j1 * nx, nx);
        out[q] = (1.0 - t) * r0 + t * r1;
    }
}

// ---- refinement of cleaned cells -----------------------------------------------------------
// One workgroup per cell (reference subpx/_gpu/centers.py:33-111): background = median of the
// ROI's 2-px frame (top and bottom rows, left and right columns; the whole ROI when a side is 4
// px or less); method 0: the weighted centroid of max(v - bg, 0) on the mask (the mask's
// centroid when that weight is empty); 1: the log-quadratic fit of log(v - bg) on the mask's
// positive pixels (6x6 normal equations + 1e-9 ridge, negative definite, >= 6 pixels). Centres
// outside the ROI are NaN.
// With `amplitude` (the cells' fitted Gaussian amplitudes), the mask is first narrowed to the
// cell's bright core, the pixels above frac x peak with peak = the amplitude when finite and
// positive, else the cell's maximum (reference subpx/centers.py:255-260).
__global__ __launch_bounds__(kBlock) void cell_refine_kernel(const double* rois, const uint8_t* cell_masks, int N,
                                                             int Hm, int Wm, int method, const double* amplitude,
                                                             double frac, double* scratch, uint8_t* support,
                                                             double* centers) {
#pragma clang fp contract(off)
    __shared__ double red[4 * 27];
    __shared__ double slot;
    __shared__ int s_n;
    const int j = blockIdx.x;
    if (j >= N) return;
    const int HW = Hm * Wm;
    const double* r = rois + (int64_t)j * HW;
    const uint8_t* m = cell_masks + (int64_t)j * HW;
    if (amplitude != nullptr) {
        double pk[1] = {-__builtin_inf()};
        for (int k = threadIdx.x; k < HW; k += kBlock)
            if (m[k]) pk[0] = fmax(pk[0], r[k]);
        block_reduce<1, true>(pk, red);
        const double A = amplitude[j];
        const double peak = (isfinite(A) && A > 0.0) ? A : pk[0];
        uint8_t* sm = support + (int64_t)j * HW;
        for (int k = threadIdx.x; k < HW; k += kBlock) sm[k] = (m[k] && r[k] > frac * peak) ? 1 : 0;
        __syncthreads();
        m = sm;
    }
    double* ring = scratch + (int64_t)j * 2 * HW;
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    const int b = 2;
    int n;
    if (Hm <= 2 * b || Wm <= 2 * b) {
        for (int k = threadIdx.x; k < HW; k += kBlock) ring[k] = r[k];
        n = HW;
    } else {
        // rows 0..1 and Hm-2..Hm-1 (every column), columns 0..1 and Wm-2..Wm-1 (every row)
        const int nrow = 2 * b * Wm, ncol = 2 * b * Hm;
        for (int k = threadIdx.x; k < nrow + ncol; k += kBlock) {
            double v;
            if (k < nrow) {
                const int rr = k / Wm, cc = k - rr * Wm;
                v = r[(rr < b ? rr : Hm - 2 * b + rr) * Wm + cc];
            } else {
                const int kk = k - nrow;
                const int cc = kk / Hm, rr = kk - cc * Hm;
                v = r[rr * Wm + (cc < b ? cc : Wm - 2 * b + cc)];
            }
            ring[k] = v;
        }
        n = nrow + ncol;
    }
    __syncthreads();
    const double lo = kth_by_rank(ring, n, (n - 1) / 2, &slot);
    const double hi = kth_by_rank(ring, n, n / 2, &slot);
    const double bg = (lo + hi) / 2.0;
    double cx = nan_d(), cy = nan_d();
    if (method == 0) {
        double a[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        for (int k = threadIdx.x; k < HW; k += kBlock) {
            if (!m[k]) continue;
            const int y = k / Wm, x = k - y * Wm;
            const double w = fmax(r[k] - bg, 0.0);
            a[0] += w;
            a[1] += w * (double)x;
            a[2] += w * (double)y;
            a[3] += (double)x;
            a[4] += (double)y;
        }
        double cntv[1] = {0.0};
        for (int k = threadIdx.x; k < HW; k += kBlock) cntv[0] += m[k] ? 1.0 : 0.0;
        block_reduce<5, false>(a, red);
        block_reduce<1, false>(cntv, red);
        double x = a[1] / fmax(a[0], 1e-12), y = a[2] / fmax(a[0], 1e-12);
        if (a[0] <= 0.0 || !isfinite(x) || !isfinite(y)) {
            x = a[3] / fmax(cntv[0], 1e-12);
            y = a[4] / fmax(cntv[0], 1e-12);
        }
        if (x >= 0.0 && y >= 0.0 && x < (double)Wm && y < (double)Hm) {
            cx = x;
            cy = y;
        }
    } else {
        double a[28];
        for (int i = 0; i < 28; ++i) a[i] = 0.0;
        for (int k = threadIdx.x; k < HW; k += kBlock) {
            const double v = r[k] - bg;
            if (!m[k] || !(v > 0.0)) continue;
            const int y = k / Wm, x = k - y * Wm;
            const double X = (double)x, Y = (double)y;
            const double f[6] = {X * X, Y * Y, X * Y, X, Y, 1.0};
            const double z = log(fmax(v, 1e-12));
            int q = 0;
            for (int u = 0; u < 6; ++u)
                for (int w = u; w < 6; ++w) a[q++] += f[u] * f[w];
            for (int u = 0; u < 6; ++u) a[21 + u] += f[u] * z;
            a[27] += 1.0;
        }
        // 28 sums in groups that fit the reduction table
        block_reduce<27, false>(*reinterpret_cast<double(*)[27]>(a), red);
        double tail[1] = {a[27]};
        block_reduce<1, false>(tail, red);
        if (threadIdx.x == 0) {
            double G[6][6], rhs[6], c[6];
            int q = 0;
            for (int u = 0; u < 6; ++u)
                for (int w = u; w < 6; ++w) {
                    G[u][w] = a[q];
                    G[w][u] = a[q];
                    ++q;
                }
            for (int u = 0; u < 6; ++u) {
                G[u][u] = G[u][u] + 1e-9;
                rhs[u] = a[21 + u];
            }
            if (solve6(G, rhs, c)) {
                const double A2 = c[0], B2 = c[1], C2 = c[2], D2 = c[3], E2 = c[4];
                const double det = 4.0 * A2 * B2 - C2 * C2;
                if (fabs(det) > 1e-12) {
                    const double x = (C2 * E2 - 2.0 * B2 * D2) / det;
                    const double y = (C2 * D2 - 2.0 * A2 * E2) / det;
                    const bool negdef = A2 < -1e-12 && B2 < -1e-12 && det > 1e-12;
                    if (tail[0] >= 6.0 && negdef && x >= 0.0 && y >= 0.0 && x < (double)Wm && y < (double)Hm &&
                        isfinite(x) && isfinite(y)) {
                        cx = x;
                        cy = y;
                    }
                }
            }
        }
    }
    if (threadIdx.x == 0) {
        centers[2 * j] = cx;
        centers[2 * j + 1] = cy;
    }
}

// The Voronoi estimators' input after deblending (reference subpx/centers.py:241-252): the
// cleaned cell restricted to the pixels within reach + pad of the seed, as an FP32 stack and
// its mask.
__global__ void deblend_window_kernel(const double* cleaned, const uint8_t* masks, const double* seeds_local,
                                      const double* reach, int N, int Hm, int Wm, int pad, float* rois,
                                      uint8_t* out_masks) {
#pragma clang fp contract(off)
    const int64_t total = (int64_t)N * Hm * Wm;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(i / ((int64_t)Hm * Wm));
        const int k = (int)(i - (int64_t)j * Hm * Wm);
        const int y = k / Wm, x = k - y * Wm;
        const double dx = (double)x - seeds_local[2 * j], dy = (double)y - seeds_local[2 * j + 1];
        const bool near = hypot(dx, dy) <= reach[j] + (double)pad;
        const bool w = masks[i] && near;
        rois[i] = w ? (float)cleaned[i] : 0.0f;
        out_masks[i] = w ? 1 : 0;
    }
}

int grid_of(int64_t n) {
    const int64_t g = (n + 255) / 256;
    return (int)(g < 8192 ? (g > 0 ? g : 1) : 8192);
}

}  // namespace

int deblend_init_launch(const double* image, int64_t H, int64_t W, const uint8_t* masks, const int32_t* origins,
                        const double* seeds, int32_t N, int32_t Hm, int32_t Wm, double* data, double* params,
                        hipStream_t stream) {
    if (N == 0) return SPX_OK;
    deblend_init_kernel<<<N, kBlock, 0, stream>>>(image, H, W, masks, origins, seeds, N, Hm, Wm, data, params);
    SPX_LAUNCHED();
    return SPX_OK;
}

int gauss_fit_launch(const double* data, const uint8_t* masks, int32_t N, int32_t Hm, int32_t Wm, const double* init,
                     int32_t max_iter, double* out, hipStream_t stream) {
    if (N == 0) return SPX_OK;
    gauss_fit_kernel<<<(N + 3) / 4, kBlock, 0, stream>>>(data, masks, N, Hm, Wm, init, max_iter, out);
    SPX_LAUNCHED();
    return SPX_OK;
}

int gauss_render_launch(const double* pg, const double* radius, int32_t N, int64_t H, int64_t W, double* model,
                        hipStream_t stream) {
    void* ws = nullptr;
    int rc = workspace(sizeof(unsigned long long) * (size_t)(H * W), &ws, stream);
    if (rc != SPX_OK) return rc;
    unsigned long long* acc = static_cast<unsigned long long*>(ws);
    zero_u64<<<grid_of(H * W), 256, 0, stream>>>(acc, H * W);
    SPX_LAUNCHED();
    if (N > 0) {
        gauss_render_kernel<<<N, kBlock, 0, stream>>>(pg, radius, N, H, W, acc);
        SPX_LAUNCHED();
    }
    fixed_to_double<<<grid_of(H * W), 256, 0, stream>>>(acc, H * W, model);
    SPX_LAUNCHED();
    return SPX_OK;
}

int deblend_clean_launch(const double* image, const double* background, const double* model, const double* pg,
                         const double* radius, const uint8_t* masks, const int32_t* origins, int32_t N, int32_t Hm,
                         int32_t Wm, int64_t H, int64_t W, double* out, hipStream_t stream) {
    if (N == 0) return SPX_OK;
    deblend_clean_kernel<<<grid_of((int64_t)N * Hm * Wm), 256, 0, stream>>>(image, background, model, pg, radius,
                                                                          masks, origins, N, Hm, Wm, H, W, out);
    SPX_LAUNCHED();
    return SPX_OK;
}

int tile_medians_launch(const double* image, const double* model, const double* dist, int64_t H, int64_t W,
                        int32_t tile, double model_below, double dist_atleast, double* medians, int64_t* counts,
                        hipStream_t stream) {
    const int ny = (int)((H + tile - 1) / tile), nx = (int)((W + tile - 1) / tile);
    void* ws = nullptr;
    int rc = workspace(sizeof(unsigned long long) * (size_t)ny * nx * tile * tile, &ws, stream);
    if (rc != SPX_OK) return rc;
    tile_median_kernel<<<ny * nx, kBlock, 0, stream>>>(image, model, dist, H, W, tile, nx, model_below, dist_atleast,
                                                       static_cast<unsigned long long*>(ws), medians,
                                                       reinterpret_cast<long long*>(counts));
    SPX_LAUNCHED();
    return SPX_OK;
}

int tile_map_launch(const double* medians, const double* yc, const double* xc, int32_t ny, int32_t nx, int64_t H,
                    int64_t W, double* out, hipStream_t stream) {
    tile_map_kernel<<<grid_of(H * W), 256, 0, stream>>>(medians, yc, xc, ny, nx, H, W, out);
    SPX_LAUNCHED();
    return SPX_OK;
}

int cell_refine_launch(const double* rois, const uint8_t* masks, int32_t N, int32_t Hm, int32_t Wm, int32_t method,
                       const double* amplitude, double frac, double* centers, hipStream_t stream) {
    if (N == 0) return SPX_OK;
    void* ws = nullptr;
    const size_t ring = sizeof(double) * 2 * (size_t)N * Hm * Wm;
    int rc = workspace(ring + (amplitude != nullptr ? (size_t)N * Hm * Wm : 0), &ws, stream);
    if (rc != SPX_OK) return rc;
    uint8_t* support = amplitude != nullptr ? static_cast<uint8_t*>(ws) + ring : nullptr;
    cell_refine_kernel<<<N, kBlock, 0, stream>>>(rois, masks, N, Hm, Wm, method, amplitude, frac,
                                                 static_cast<double*>(ws), support, centers);
    SPX_LAUNCHED();
    return SPX_OK;
}

int deblend_window_launch(const double* cleaned, const uint8_t* masks, const double* seeds_local,
                          const double* reach, int32_t N, int32_t Hm, int32_t Wm, int32_t pad, float* rois,
                          uint8_t* out_masks, hipStream_t stream) {
    if (N == 0) return SPX_OK;
    deblend_window_kernel<<<grid_of((int64_t)N * Hm * Wm), 256, 0, stream>>>(cleaned, masks, seeds_local, reach, N,
                                                                           Hm, Wm, pad, rois, out_masks);
    SPX_LAUNCHED();
    return SPX_OK;
}

}  // namespace spx
