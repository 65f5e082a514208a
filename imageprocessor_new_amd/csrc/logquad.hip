// Per-label log-quadratic centre refinement, one wave64 per feature.
//
// Restates the reference CPU path (subpx/_cpu/centers.py:67-107, ROI rule :376-391,
// border median :31-42) for the GPU:
//   ROI = bbox +- pad clipped to the image; mask = labels == label inside the ROI;
//   bg = median of the ROI's 2-pixel ring (top 2 rows, bottom 2 rows, left 2 columns,
//        right 2 columns, corners counted twice), or of the whole ROI when a side <= 4;
//   v = g - bg on the mask; NaN when fewer than 6 mask pixels or max(v) <= 0;
//   least squares of z = log(max(v, 1e-9)) on [x^2, y^2, xy, x, y, 1];
//   H = [[2a, c], [c, 2b]]; NaN when cond(H) > 1e8; centre = H^-1 (-d, -e);
//   NaN unless the centre lies in [0, w) x [0, h) of the ROI.
// FP64 throughout. The median is an exact order statistic of integer grey levels found by
// bisection over the 16-bit value range (no sort buffer, any ROI size). The least squares
// uses normal equations in coordinates centred on the ROI and Jacobi-scaled, solved by
// LDL^T; the centre and cond(H) do not depend on the shift.
#include "spx_internal.hpp"
#include "wave64.hpp"

namespace spx {
namespace {

struct Roi {
    int64_t x0, y0;
    int rw, rh;
};

// j-th element of the ring multiset (or of the whole ROI when `whole`).
__device__ __forceinline__ void ring_pos(const Roi& r, bool whole, int j, int& ry, int& rx) {
    if (whole) {
        ry = j / r.rw;
        rx = j - ry * r.rw;
        return;
    }
    const int top = 2 * r.rw;
    if (j < top) {
        ry = j / r.rw;
        rx = j - ry * r.rw;
    } else if (j < 2 * top) {
        const int jj = j - top;
        ry = r.rh - 2 + jj / r.rw;
        rx = jj - (jj / r.rw) * r.rw;
    } else if (j < 2 * top + 2 * r.rh) {
        const int jj = j - 2 * top;
        ry = jj >> 1;
        rx = jj & 1;
    } else {
        const int jj = j - 2 * top - 2 * r.rh;
        ry = jj >> 1;
        rx = r.rw - 2 + (jj & 1);
    }
}

// k-th smallest (0-based) grey level of the ring multiset.
__device__ int ring_kth(const uint16_t* img, int64_t pitch, const Roi& r, bool whole, int n, int k) {
    const int lane = threadIdx.x & 63;
    int lo = 0, hi = 65535;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        int c = 0;
        for (int j = lane; j < n; j += 64) {
            int ry, rx;
            ring_pos(r, whole, j, ry, rx);
            c += (int)img[(r.y0 + ry) * pitch + r.x0 + rx] <= mid;
        }
        c = wave::sum_i(c);
        if (c > k)
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo;
}

// np.median of the ring (mean of the middle pair for an even count).
__device__ double border_median(const uint16_t* img, int64_t pitch, const Roi& r) {
    const bool whole = (r.rh <= 4 || r.rw <= 4);
    const int n = whole ? r.rw * r.rh : 4 * r.rw + 4 * r.rh;
    if (n & 1) return (double)ring_kth(img, pitch, r, whole, n, n >> 1);
    const int a = ring_kth(img, pitch, r, whole, n, (n >> 1) - 1);
    const int b = ring_kth(img, pitch, r, whole, n, n >> 1);
    return 0.5 * ((double)a + (double)b);
}

// ROI = bbox +- pad clipped to the image (reference subpx/_cpu/centers.py:381-384); false when empty.
__device__ __forceinline__ bool make_roi(const int32_t* row, int32_t pad, int64_t H, int64_t W, Roi& r,
                                         int32_t& lab) {
    const int64_t x = __builtin_amdgcn_readfirstlane(row[0]);
    const int64_t y = __builtin_amdgcn_readfirstlane(row[1]);
    const int64_t w = __builtin_amdgcn_readfirstlane(row[2]);
    const int64_t h = __builtin_amdgcn_readfirstlane(row[3]);
    lab = __builtin_amdgcn_readfirstlane(row[4]);
    r.x0 = x - pad > 0 ? x - pad : 0;
    r.y0 = y - pad > 0 ? y - pad : 0;
    const int64_t x1 = (x + w + pad) < W ? (x + w + pad) : W;
    const int64_t y1 = (y + h + pad) < H ? (y + h + pad) : H;
    r.rw = (int)(x1 - r.x0);
    r.rh = (int)(y1 - r.y0);
    return x1 > r.x0 && y1 > r.y0;
}

// Weighted centroid of the GPU detect path's "weighted" / "none" methods (reference
// subpx/_gpu/centers.py:43-66 via _refine_batch_gpu :931-955, background from the CPU
// border median :33-46): w = max(g - bg, 0) on the mask; centroid sum(w x) / max(sum w, 1e-12);
// the mask mean when sum w <= 0; NaN unless inside [0, rw) x [0, rh) of the ROI.
__global__ __launch_bounds__(256) void weighted_kernel(const uint16_t* img, int64_t H, int64_t W, int64_t pitch,
                                                       const int32_t* labels, const int32_t* rows, int32_t K,
                                                       int32_t pad, double* centers, uint8_t* ok_out) {
    const int lane = threadIdx.x & 63;
    const int64_t f = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (f >= K) return;
    Roi r;
    int32_t lab;
    bool good = make_roi(rows + 5 * f, pad, H, W, r, lab);
    double cx = __builtin_nan(""), cy = __builtin_nan("");
    if (good) {
        const double bg = border_median(img, pitch, r);
        double s[6] = {0, 0, 0, 0, 0, 0};  // sum w, sum w x, sum w y, count, sum x, sum y
        const int npx = r.rw * r.rh;
        for (int j = lane; j < npx; j += 64) {
            const int ry = j / r.rw, rx = j - (j / r.rw) * r.rw;
            const int64_t gy = r.y0 + ry, gx = r.x0 + rx;
            if (labels[gy * W + gx] != lab) continue;
            const double wt = fmax((double)img[gy * pitch + gx] - bg, 0.0);
            s[0] += wt;
            s[1] += wt * (double)rx;
            s[2] += wt * (double)ry;
            s[3] += 1.0;
            s[4] += (double)rx;
            s[5] += (double)ry;
        }
#pragma unroll
        for (int q = 0; q < 6; ++q) s[q] = wave::sum_d(s[q]);
        cx = s[1] / fmax(s[0], 1e-12);
        cy = s[2] / fmax(s[0], 1e-12);
        if (!(s[0] > 0.0) || !isfinite(cx) || !isfinite(cy)) {
            cx = s[4] / fmax(s[3], 1e-12);
            cy = s[5] / fmax(s[3], 1e-12);
        }
        good = cx >= 0.0 && cy >= 0.0 && cx < (double)r.rw && cy < (double)r.rh;
        cx += (double)r.x0;
        cy += (double)r.y0;
    }
    if (lane == 0) {
        centers[2 * f] = good ? cx : __builtin_nan("");
        centers[2 * f + 1] = good ? cy : __builtin_nan("");
        ok_out[f] = good ? 1 : 0;
    }
}

__global__ __launch_bounds__(256) void logquad_kernel(const uint16_t* img, int64_t H, int64_t W,
                                                      int64_t pitch, const int32_t* labels,
                                                      const int32_t* rows, int32_t K, int32_t pad,
                                                      double* centers, uint8_t* ok_out) {
    const int lane = threadIdx.x & 63;
    const int64_t f = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (f >= K) return;
    Roi r;
    int32_t lab;
    bool good = make_roi(rows + 5 * f, pad, H, W, r, lab);
    double cx = __builtin_nan(""), cy = __builtin_nan("");
    if (good) {
        const double bg = border_median(img, pitch, r);
        // mask statistics and normal equations in centred coordinates
        const double xc = 0.5 * (double)(r.rw - 1);
        const double yc = 0.5 * (double)(r.rh - 1);
        int cnt = 0;
        double vmax = -1.0e300;
        const int npx = r.rw * r.rh;
        for (int j = lane; j < npx; j += 64) {
            const int ry = j / r.rw, rx = j - (j / r.rw) * r.rw;
            const int64_t gy = r.y0 + ry, gx = r.x0 + rx;
            if (labels[gy * W + gx] == lab) {
                ++cnt;
                vmax = fmax(vmax, (double)img[gy * pitch + gx] - bg);
            }
        }
        cnt = wave::sum_i(cnt);
        vmax = wave::max_d(vmax);
        if (cnt < 6 || !(vmax > 0.0)) good = false;
        if (good) {
            double A[21];
            double bz[6];
#pragma unroll
            for (int i = 0; i < 21; ++i) A[i] = 0.0;
#pragma unroll
            for (int i = 0; i < 6; ++i) bz[i] = 0.0;
            for (int j = lane; j < npx; j += 64) {
                const int ry = j / r.rw, rx = j - (j / r.rw) * r.rw;
                const int64_t gy = r.y0 + ry, gx = r.x0 + rx;
                if (labels[gy * W + gx] != lab) continue;
                const double v = (double)img[gy * pitch + gx] - bg;
                const double z = log(fmax(v, 1e-9));
                const double u = (double)rx - xc, t = (double)ry - yc;
                const double c[6] = {u * u, t * t, u * t, u, t, 1.0};
                int idx = 0;
#pragma unroll
                for (int p = 0; p < 6; ++p) {
#pragma unroll
                    for (int q = p; q < 6; ++q) A[idx++] += c[p] * c[q];
                    bz[p] += c[p] * z;
                }
            }
#pragma unroll
            for (int i = 0; i < 21; ++i) A[i] = wave::sum_d(A[i]);
#pragma unroll
            for (int i = 0; i < 6; ++i) bz[i] = wave::sum_d(bz[i]);
            // Jacobi scaling + LDL^T
            double M[6][6], s[6], g6[6];
            int idx = 0;
#pragma unroll
            for (int p = 0; p < 6; ++p)
#pragma unroll
                for (int q = p; q < 6; ++q) {
                    M[p][q] = A[idx];
                    M[q][p] = A[idx];
                    ++idx;
                }
#pragma unroll
            for (int p = 0; p < 6; ++p) s[p] = M[p][p] > 0.0 ? 1.0 / sqrt(M[p][p]) : 0.0;
#pragma unroll
            for (int p = 0; p < 6; ++p) {
                g6[p] = bz[p] * s[p];
#pragma unroll
                for (int q = 0; q < 6; ++q) M[p][q] *= s[p] * s[q];
            }
            double L[6][6], D[6], yv[6], coef[6];
            bool spd = true;
#pragma unroll
            for (int j = 0; j < 6; ++j) {
                double d = M[j][j];
#pragma unroll
                for (int k = 0; k < j; ++k) d -= L[j][k] * L[j][k] * D[k];
                if (!(d > 1e-14)) spd = false;
                D[j] = d;
#pragma unroll
                for (int i = j + 1; i < 6; ++i) {
                    double acc = M[i][j];
#pragma unroll
                    for (int k = 0; k < j; ++k) acc -= L[i][k] * L[j][k] * D[k];
                    L[i][j] = acc / d;
                }
            }
            if (!spd) good = false;
            if (good) {
#pragma unroll
                for (int i = 0; i < 6; ++i) {
                    double acc = g6[i];
#pragma unroll
                    for (int k = 0; k < i; ++k) acc -= L[i][k] * yv[k];
                    yv[i] = acc;
                }
#pragma unroll
                for (int i = 5; i >= 0; --i) {
                    double acc = yv[i] / D[i];
#pragma unroll
                    for (int k = i + 1; k < 6; ++k) acc -= L[k][i] * coef[k];
                    coef[i] = acc;
                }
#pragma unroll
                for (int i = 0; i < 6; ++i) coef[i] *= s[i];
                const double a = coef[0], b = coef[1], c = coef[2], d = coef[3], e = coef[4];
                const double h11 = 2.0 * a, h22 = 2.0 * b, h12 = c;
                const double mean = 0.5 * (h11 + h22);
                const double rad = sqrt(0.25 * (h11 - h22) * (h11 - h22) + h12 * h12);
                const double l1 = fabs(mean + rad), l2 = fabs(mean - rad);
                const double lmax = fmax(l1, l2), lmin = fmin(l1, l2);
                const double cond = lmin > 0.0 ? lmax / lmin : __builtin_inf();
                if (!(cond <= 1e8)) {
                    good = false;
                } else {
                    const double det = h11 * h22 - h12 * h12;
                    const double ux = (-d * h22 + h12 * e) / det;
                    const double uy = (-e * h11 + h12 * d) / det;
                    const double px = ux + xc, py = uy + yc;
                    if (px >= 0.0 && px < (double)r.rw && py >= 0.0 && py < (double)r.rh) {
                        cx = (double)r.x0 + px;
                        cy = (double)r.y0 + py;
                    } else {
                        good = false;
                    }
                }
            }
        }
    }
    if (lane == 0) {
        centers[2 * f] = good ? cx : __builtin_nan("");
        centers[2 * f + 1] = good ? cy : __builtin_nan("");
        ok_out[f] = good ? 1 : 0;
    }
}

}  // namespace

int logquad_launch(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, const int32_t* labels,
                   const int32_t* rows, int32_t K, int32_t pad, double* centers, uint8_t* ok,
                   hipStream_t stream) {
    if (K == 0) return SPX_OK;
    const int waves_per_block = 4;
    const int blocks = (K + waves_per_block - 1) / waves_per_block;
    logquad_kernel<<<blocks, 64 * waves_per_block, 0, stream>>>(img, H, W, pitch, labels, rows, K, pad,
                                                                centers, ok);
    SPX_LAUNCHED();
    return SPX_OK;
}

int weighted_launch(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, const int32_t* labels,
                    const int32_t* rows, int32_t K, int32_t pad, double* centers, uint8_t* ok, hipStream_t stream) {
    if (K == 0) return SPX_OK;
    const int waves_per_block = 4;
    const int blocks = (K + waves_per_block - 1) / waves_per_block;
    weighted_kernel<<<blocks, 64 * waves_per_block, 0, stream>>>(img, H, W, pitch, labels, rows, K, pad, centers,
                                                                 ok);
    SPX_LAUNCHED();
    return SPX_OK;
}

}  // namespace spx
