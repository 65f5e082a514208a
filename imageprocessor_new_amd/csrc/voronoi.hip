// Voronoi-cell refinements of the GPU detect path ("radial_symmetry", "isophote_curvature"):
// nearest-seed labels over the image, the cell ROI of every component, and the two
// gradient-based centre estimators on 4x bicubic-upsampled ROIs.
//
// Reference: subpx/_gpu/centers.py:1545-1630 (dispatch), subpx/_gpu/voronoi.py:24-145
// (labels), subpx/_cpu/cells.py:34-84 (ROIs), subpx/_gpu/upsample.py:131-231 (resampling),
// subpx/_gpu/centers.py:978-1469 (estimators). The reference runs each estimator as ~30 CuPy
// array operations over a stack of upsampled ROIs in HBM; here one workgroup carries one ROI
// through every step with the upsampled ROI and its masks in LDS (global scratch per
// workgroup when they do not fit), and the sums are FP64 block reductions.
//
// Floating point follows the reference's split (FP32 resampling, gradients, slopes and votes;
// FP64 sums and solves) in its operation order, without FMA contraction, so results track the
// NumPy restatement used as the oracle to FP32 rounding.
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <string>

#include "spx_internal.hpp"

namespace spx {
namespace {

constexpr int kThreads = 256;
constexpr int kMaxGaussTaps = 65;

// ---- nearest-seed labels --------------------------------------------------------------------
// Seeds are binned into square cells of side `cell` (about one seed each). A pixel scans square
// rings of cells around its own cell; every seed outside the scanned square lies beyond one of
// the square's sides that the grid edge does not clip, so once the best squared distance is
// below the square of the nearest such side's distance, no unscanned seed can win (or tie).
struct SeedGrid {
    int gw, gh, cell;
};

__global__ void vgrid_zero(int32_t* p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

__device__ inline int vgrid_bin(const double* seeds, int i, SeedGrid g) {
    const double fx = floor(seeds[2 * i]), fy = floor(seeds[2 * i + 1]);
    long long bx = (long long)fx, by = (long long)fy;
    bx = bx < 0 ? 0 : bx / g.cell;
    by = by < 0 ? 0 : by / g.cell;
    if (bx > g.gw - 1) bx = g.gw - 1;
    if (by > g.gh - 1) by = g.gh - 1;
    return (int)(by * g.gw + bx);
}

__global__ void vgrid_count(const double* seeds, int n, SeedGrid g, int32_t* cnt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) atomicAdd(&cnt[vgrid_bin(seeds, i, g)], 1);
}

__global__ void vgrid_scatter(const double* seeds, int n, SeedGrid g, int32_t* cursor, int32_t* ids) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) ids[atomicAdd(&cursor[vgrid_bin(seeds, i, g)], 1)] = i;
}

__device__ inline void vscan_cell(const double* seeds, const int32_t* start, const int32_t* ids, int c, double px,
                                  double py, double& best, int& label) {
#pragma clang fp contract(off)
    for (int k = start[c]; k < start[c + 1]; ++k) {
        const int i = ids[k];
        const double dx = px - seeds[2 * i], dy = py - seeds[2 * i + 1];
        const double d2 = dx * dx + dy * dy;
        if (label < 0 || d2 < best || (d2 == best && i < label)) {
            best = d2;
            label = i;
        }
    }
}

// One lane per pixel; `areas` counts each seed's pixels, one atomic per run of equal labels
// among a wave's 64 consecutive pixels.
__global__ __launch_bounds__(kThreads) void voronoi_label_kernel(const double* seeds, const int32_t* start,
                                                                 const int32_t* ids, SeedGrid g, int64_t H,
                                                                 int64_t W, int32_t* labels, int32_t* areas,
                                                                 double* dist) {
#pragma clang fp contract(off)
    const int64_t HW = H * W;
    const int64_t idx = blockIdx.x * (int64_t)kThreads + threadIdx.x;
    int label = -1;
    if (idx < HW) {
        const int64_t row = idx / W, col = idx - row * W;
        const double px = (double)col, py = (double)row;
        const int cx = (int)(col / g.cell < g.gw - 1 ? col / g.cell : g.gw - 1);
        const int cy = (int)(row / g.cell < g.gh - 1 ? row / g.cell : g.gh - 1);
        double best = 0.0;
        for (int R = 0;; ++R) {
            const int ylo = cy - R, yhi = cy + R, xlo = cx - R, xhi = cx + R;
            for (int gy = ylo < 0 ? 0 : ylo; gy <= yhi && gy < g.gh; ++gy) {
                if (gy == ylo || gy == yhi) {
                    for (int gx = xlo < 0 ? 0 : xlo; gx <= xhi && gx < g.gw; ++gx)
                        vscan_cell(seeds, start, ids, gy * g.gw + gx, px, py, best, label);
                } else {
                    if (xlo >= 0) vscan_cell(seeds, start, ids, gy * g.gw + xlo, px, py, best, label);
                    if (xhi < g.gw) vscan_cell(seeds, start, ids, gy * g.gw + xhi, px, py, best, label);
                }
            }
            double bound = 1e300;
            bool covered = true;
            if (xlo > 0) { bound = fmin(bound, px - (double)xlo * g.cell); covered = false; }
            if (ylo > 0) { bound = fmin(bound, py - (double)ylo * g.cell); covered = false; }
            if (xhi < g.gw - 1) { bound = fmin(bound, (double)(xhi + 1) * g.cell - px); covered = false; }
            if (yhi < g.gh - 1) { bound = fmin(bound, (double)(yhi + 1) * g.cell - py); covered = false; }
            if (covered || (label >= 0 && best < bound * bound)) break;
        }
        if (labels != nullptr) labels[idx] = label;
        if (dist != nullptr) dist[idx] = sqrt(best);
    }
    if (areas != nullptr) {
        const int lane = threadIdx.x & 63;
        const int prev = __shfl_up(label, 1, 64);
        const bool head = label >= 0 && (lane == 0 || prev != label);
        // lanes where a run starts or that hold no label: a run ends before the next of them
        const unsigned long long stops = __ballot(head || label < 0);
        if (head) {
            const unsigned long long above = lane == 63 ? 0ull : (stops >> (lane + 1));
            const int len = above ? __builtin_ctzll(above) + 1 : 64 - lane;
            atomicAdd(&areas[label], len);
        }
    }
}

// ---- cell ROIs ----------------------------------------------------------------------------
// One workgroup per component row [x, y, w, h, cx, cy]: the ROI bbox +- pad clipped to the
// image, the cell = pixels labelled with the row's index, its border (cell pixels on the ROI
// frame or with a 4-neighbour outside the cell), the 10th percentile of the border's grey
// levels (NumPy's linear rule) as the fill value; the ROI is written into the (Hm, Wm) stack
// with the fill outside the cell and in the padding, the cell as the mask.
__global__ __launch_bounds__(kThreads) void voronoi_roi_kernel(const uint16_t* img, int64_t H, int64_t W,
                                                               int64_t pitch, const int32_t* labels, const double* rows,
                                                               int N, int pad, int Hm, int Wm, float* rois,
                                                               uint8_t* masks, int32_t* origins) {
#pragma clang fp contract(off)
    extern __shared__ unsigned char s_roi[];
    __shared__ int s_n;
    __shared__ double s_bg;
    uint8_t* cell = s_roi;                                           // Hm * Wm
    float* vals = reinterpret_cast<float*>(s_roi + ((Hm * Wm + 15) & ~15));  // border values
    const int j = blockIdx.x;
    if (j >= N) return;
    const double* r = rows + (int64_t)j * 6;
    const int64_t rx = (int64_t)(int32_t)r[0], ry = (int64_t)(int32_t)r[1];
    const int64_t x0 = rx - pad > 0 ? rx - pad : 0, y0 = ry - pad > 0 ? ry - pad : 0;
    int64_t x1 = (int64_t)(int32_t)(r[0] + r[2]) + pad, y1 = (int64_t)(int32_t)(r[1] + r[3]) + pad;
    x1 = x1 < W ? x1 : W;
    y1 = y1 < H ? y1 : H;
    // the stack's (Hm, Wm) holds every ROI when the caller sizes it from the same rows; a
    // smaller stack truncates the ROI rather than writing past the LDS tables
    const int h = (int)(y1 - y0) < Hm ? (int)(y1 - y0) : Hm, w = (int)(x1 - x0) < Wm ? (int)(x1 - x0) : Wm;
    if (threadIdx.x == 0) {
        s_n = 0;
        origins[2 * j] = (int32_t)x0;
        origins[2 * j + 1] = (int32_t)y0;
    }
    for (int k = threadIdx.x; k < h * w; k += kThreads) {
        const int yy = k / w, xx = k - yy * w;
        cell[k] = labels[(y0 + yy) * W + x0 + xx] == j ? 1 : 0;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < h * w; k += kThreads) {
        if (!cell[k]) continue;
        const int yy = k / w, xx = k - yy * w;
        const bool frame = yy == 0 || xx == 0 || yy == h - 1 || xx == w - 1;
        const bool inner = !frame && cell[k - w] && cell[k + w] && cell[k - 1] && cell[k + 1];
        if (!inner) vals[atomicAdd(&s_n, 1)] = (float)img[(y0 + yy) * pitch + x0 + xx];
    }
    __syncthreads();
    const int n = s_n;
    if (threadIdx.x == 0) s_bg = 0.0;
    __syncthreads();
    if (n > 0) {
        const double vi = 0.1 * (double)(n - 1);
        const int lo = (int)floor(vi);
        const int hi = lo + 1 < n ? lo + 1 : n - 1;
        const double t = vi - (double)lo;
        // order statistics lo and hi: the value v with #{< v} <= k < #{<= v}
        for (int i = threadIdx.x; i < n; i += kThreads) {
            const float v = vals[i];
            int less = 0, le = 0;
            for (int q = 0; q < n; ++q) {
                less += vals[q] < v;
                le += vals[q] <= v;
            }
            // lanes holding the same value agree; the write of either order statistic is
            // idempotent, and both are combined after the barrier
            if (less <= lo && lo < le) vals[n + 0] = v;  // slot past the list (sized for it)
            if (less <= hi && hi < le) vals[n + 1] = v;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            const double a = (double)vals[n], b = (double)vals[n + 1];
            const double d = b - a;
            s_bg = t >= 0.5 ? b - d * (1.0 - t) : a + d * t;
        }
    }
    __syncthreads();
    const float bg = (float)s_bg;
    float* out = rois + (int64_t)j * Hm * Wm;
    uint8_t* om = masks + (int64_t)j * Hm * Wm;
    for (int k = threadIdx.x; k < Hm * Wm; k += kThreads) {
        const int yy = k / Wm, xx = k - yy * Wm;
        bool in = false;
        float v = bg;
        if (yy < h && xx < w) {
            in = cell[yy * w + xx] != 0;
            if (in) v = (float)img[(y0 + yy) * pitch + x0 + xx];
        }
        out[k] = v;
        om[k] = in ? 1 : 0;
    }
}

// ---- shared steps of the estimators ---------------------------------------------------------
struct Gauss {
    int radius;  // taps = 2 radius + 1; 0 = no pre-smoothing
    double w[kMaxGaussTaps];
};

struct RoiShape {
    int Hm, Wm, f, Uh, Uw;  // ROI stack side, factor, upsampled side
    int margin;             // erosion half width actually used
};

// block-wide sums of M doubles; every thread gets the totals
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
#pragma unroll
    for (int m = 0; m < M; ++m) v[m] = ((red[m] + red[M + m]) + red[2 * M + m]) + red[3 * M + m];
}

__device__ inline void keys_weights(float f, float (&w)[4]) {
#pragma clang fp contract(off)
    float t = 1.0f + f;
    w[0] = ((-0.5f * t + 2.5f) * t - 4.0f) * t + 2.0f;
    t = f;
    w[1] = ((1.5f * t - 2.5f) * t) * t + 1.0f;
    t = 1.0f - f;
    w[2] = ((1.5f * t - 2.5f) * t) * t + 1.0f;
    t = 2.0f - f;
    w[3] = ((-0.5f * t + 2.5f) * t - 4.0f) * t + 2.0f;
}

// ROI j of the stack upsampled into up (Keys' cubic, edge-aligned) and its mask (nearest) into m.
__device__ inline void upsample_roi(const float* roi, const uint8_t* mask, RoiShape s, float* up, uint8_t* m) {
#pragma clang fp contract(off)
    const float scy = s.Uh > 1 ? ((float)s.Hm - 1.0f) / ((float)s.Uh - 1.0f) : 0.0f;
    const float scx = s.Uw > 1 ? ((float)s.Wm - 1.0f) / ((float)s.Uw - 1.0f) : 0.0f;
    for (int k = threadIdx.x; k < s.Uh * s.Uw; k += kThreads) {
        const int oy = k / s.Uw, ox = k - oy * s.Uw;
        float val;
        int ny, nx;
        if (s.f > 1) {
            const float sy = (float)oy * scy, sx = (float)ox * scx;
            const int iy = (int)floorf(sy), ix = (int)floorf(sx);
            float wy[4], wx[4];
            keys_weights(sy - (float)iy, wy);
            keys_weights(sx - (float)ix, wx);
            val = 0.0f;
            for (int dy = 0; dy < 4; ++dy) {
                int yy = iy - 1 + dy;
                yy = yy < 0 ? 0 : (yy >= s.Hm ? s.Hm - 1 : yy);
                for (int dx = 0; dx < 4; ++dx) {
                    int xx = ix - 1 + dx;
                    xx = xx < 0 ? 0 : (xx >= s.Wm ? s.Wm - 1 : xx);
                    val = val + (wy[dy] * wx[dx]) * roi[yy * s.Wm + xx];
                }
            }
            ny = (int)((float)oy * scy + 0.5f);
            nx = (int)((float)ox * scx + 0.5f);
            ny = ny < s.Hm ? ny : s.Hm - 1;
            nx = nx < s.Wm ? nx : s.Wm - 1;
        } else {
            val = roi[k];
            ny = oy;
            nx = ox;
        }
        up[k] = val;
        m[k] = mask[ny * s.Wm + nx];
    }
}

// Erosion of m by a (2 margin + 1)^2 square (outside the array counts as empty), t as temp.
__device__ inline void erode(uint8_t* m, uint8_t* t, RoiShape s) {
    const int r = s.margin;
    if (r <= 0) return;
    __syncthreads();
    for (int k = threadIdx.x; k < s.Uh * s.Uw; k += kThreads) {
        const int oy = k / s.Uw, ox = k - oy * s.Uw;
        uint8_t v = (ox - r >= 0 && ox + r < s.Uw) ? 1 : 0;
        for (int d = -r; d <= r && v; ++d) v = m[oy * s.Uw + ox + d];
        t[k] = v;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < s.Uh * s.Uw; k += kThreads) {
        const int oy = k / s.Uw, ox = k - oy * s.Uw;
        uint8_t v = (oy - r >= 0 && oy + r < s.Uh) ? 1 : 0;
        for (int d = -r; d <= r && v; ++d) v = t[(oy + d) * s.Uw + ox];
        m[k] = v;
    }
}

__device__ inline void to_roi(double xc, double yc, RoiShape s, double* centers, double* quality, int j,
                              double q) {
#pragma clang fp contract(off)
    double x = xc + ((double)s.Uw - 1.0) / 2.0;
    double y = yc + ((double)s.Uh - 1.0) / 2.0;
    if (s.f > 1) {
        x = x * ((double)s.Wm - 1.0) / fmax((double)s.Uw - 1.0, 1.0);
        y = y * ((double)s.Hm - 1.0) / fmax((double)s.Uh - 1.0, 1.0);
    }
    if (x < 0.0 || x >= (double)s.Wm || y < 0.0 || y >= (double)s.Hm) {
        x = y = __builtin_nan("");
        q = __builtin_inf();
    }
    centers[2 * j] = x;
    centers[2 * j + 1] = y;
    quality[j] = q;
}

// ---- radial symmetry ----------------------------------------------------------------------
// Diagonal gradient at midpoint q of the upsampled ROI (between rows qy, qy + 1 and columns
// qx, qx + 1).
__device__ inline void raw_grad(const float* up, int Uw, int qy, int qx, float& du, float& dv) {
#pragma clang fp contract(off)
    const float a = up[qy * Uw + qx], b = up[qy * Uw + qx + 1], c = up[(qy + 1) * Uw + qx],
                d = up[(qy + 1) * Uw + qx + 1];
    du = b - c;
    dv = a - d;
}

// The 3x3 edge-padded mean of both gradients at midpoint (py, px) (none when the grid is
// smaller than 3 in either direction).
__device__ inline void smooth_grad(const float* up, int Uw, int Hp, int Wp, int py, int px, float& du, float& dv) {
#pragma clang fp contract(off)
    if (Hp < 3 || Wp < 3) {
        raw_grad(up, Uw, py, px, du, dv);
        return;
    }
    float su = 0.0f, sv = 0.0f;
    for (int a = -1; a <= 1; ++a) {
        int qy = py + a;
        qy = qy < 0 ? 0 : (qy >= Hp ? Hp - 1 : qy);
        for (int b = -1; b <= 1; ++b) {
            int qx = px + b;
            qx = qx < 0 ? 0 : (qx >= Wp ? Wp - 1 : qx);
            float u, v;
            raw_grad(up, Uw, qy, qx, u, v);
            su = su + u;
            sv = sv + v;
        }
    }
    du = su / 9.0f;
    dv = sv / 9.0f;
}

struct RadialPoint {
    bool valid;
    float slope, b, g2, xm, ym;
};

__device__ inline RadialPoint radial_point(const float* up, const uint8_t* m, RoiShape s, int Hp, int Wp, int py,
                                           int px) {
#pragma clang fp contract(off)
    RadialPoint p;
    float du, dv;
    smooth_grad(up, s.Uw, Hp, Wp, py, px, du, dv);
    p.valid = m[py * s.Uw + px] && m[py * s.Uw + px + 1] && m[(py + 1) * s.Uw + px] && m[(py + 1) * s.Uw + px + 1];
    const float denom = du - dv;
    const float num = -(dv + du);
    const bool near = fabsf(denom) < 1e-6f;
    p.slope = near ? (num >= 0.0f ? 1e9f : -1e9f) : num / denom;
    p.xm = ((float)px + 0.5f) - ((float)s.Uw - 1.0f) / 2.0f;
    p.ym = ((float)py + 0.5f) - ((float)s.Uh - 1.0f) / 2.0f;
    p.b = p.ym - p.slope * p.xm;
    p.g2 = du * du + dv * dv;
    return p;
}

__global__ __launch_bounds__(kThreads) void radial_symmetry_kernel(const float* rois, const uint8_t* masks, int N,
                                                                   RoiShape s, unsigned char* scratch,
                                                                   size_t stride, double* centers, double* quality) {
#pragma clang fp contract(off)
    extern __shared__ unsigned char s_dyn[];
    __shared__ double red[4 * 5];
    unsigned char* base = scratch != nullptr ? scratch + (size_t)blockIdx.x * stride : s_dyn;
    const int UU = s.Uh * s.Uw;
    float* up = reinterpret_cast<float*>(base);
    uint8_t* m = base + (size_t)UU * 4;
    uint8_t* t = m + UU;
    const int Hp = s.Uh - 1, Wp = s.Uw - 1;
    for (int j = blockIdx.x; j < N; j += gridDim.x) {
        if (Hp < 1 || Wp < 1) {
            if (threadIdx.x == 0) to_roi(__builtin_nan(""), __builtin_nan(""), s, centers, quality, j, __builtin_inf());
            continue;
        }
        __syncthreads();  // the previous ROI's readers are done
        upsample_roi(rois + (int64_t)j * s.Hm * s.Wm, masks + (int64_t)j * s.Hm * s.Wm, s, up, m);
        erode(m, t, s);
        __syncthreads();
        // pass 1: the gradient centroid (weights |grad|^2 on valid midpoints)
        double a1[3] = {0.0, 0.0, 0.0};
        for (int k = threadIdx.x; k < Hp * Wp; k += kThreads) {
            const int py = k / Wp, px = k - py * Wp;
            const RadialPoint p = radial_point(up, m, s, Hp, Wp, py, px);
            const float wb = p.valid ? p.g2 : 0.0f;
            a1[0] += (double)wb;
            a1[1] += (double)(wb * p.xm);
            a1[2] += (double)(wb * p.ym);
        }
        block_sum<3>(a1, red);
        const double wsum = fmax(a1[0], 1e-30);
        const float gcx = (float)(a1[1] / wsum), gcy = (float)(a1[2] / wsum);
        // pass 2: the weighted least squares of the gradient lines' closest point
        double a2[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        for (int k = threadIdx.x; k < Hp * Wp; k += kThreads) {
            const int py = k / Wp, px = k - py * Wp;
            const RadialPoint p = radial_point(up, m, s, Hp, Wp, py, px);
            const float ex = p.xm - gcx, ey = p.ym - gcy;
            const float dist = sqrtf(fmaxf(ex * ex + ey * ey, 1e-30f));
            const double w = p.valid ? (double)(p.g2 / dist) : 0.0;
            const double sl = (double)p.slope, bb = (double)p.b;
            const double wm = w / (sl * sl + 1.0);
            a2[0] += wm;
            a2[1] += sl * sl * wm;
            a2[2] += sl * wm;
            a2[3] += sl * bb * wm;
            a2[4] += bb * wm;
        }
        block_sum<5>(a2, red);
        const double sw = a2[0], smm = a2[1], sm = a2[2], smb = a2[3], sb = a2[4];
        const double det = smm * sw - sm * sm;
        const bool bad = fabs(det) < 1e-30;
        const double xc = bad ? __builtin_nan("") : (-smb * sw + sm * sb) / det;
        const double yc = bad ? __builtin_nan("") : (smm * sb - smb * sm) / det;
        // pass 3: the weighted mean squared distance of the lines from that point
        double a3[1] = {0.0};
        for (int k = threadIdx.x; k < Hp * Wp; k += kThreads) {
            const int py = k / Wp, px = k - py * Wp;
            const RadialPoint p = radial_point(up, m, s, Hp, Wp, py, px);
            const float ex = p.xm - gcx, ey = p.ym - gcy;
            const float dist = sqrtf(fmaxf(ex * ex + ey * ey, 1e-30f));
            const double w = p.valid ? (double)(p.g2 / dist) : 0.0;
            const double sl = (double)p.slope, bb = (double)p.b;
            const double e = sl * xc - yc + bb;
            a3[0] += w * (e * e / (sl * sl + 1.0));
        }
        block_sum<1>(a3, red);
        if (threadIdx.x == 0) to_roi(xc, yc, s, centers, quality, j, a3[0] / fmax(sw, 1e-30));
    }
}

// ---- isophote curvature -------------------------------------------------------------------
__device__ inline int reflect_index(int i, int n) {
    const int period = 2 * n;
    int mm = i % period;
    if (mm < 0) mm += period;
    return mm >= n ? period - 1 - mm : mm;
}

// Separable Gaussian (taps g.w, FP64 sums, FP32 after each pass, mode "reflect"): along the
// rows' direction (y) from up into t, then along x from t back into up.
__device__ inline void gauss_yx(float* up, float* t, RoiShape s, const Gauss& g) {
#pragma clang fp contract(off)
    const int r = g.radius;
    __syncthreads();
    for (int k = threadIdx.x; k < s.Uh * s.Uw; k += kThreads) {
        const int oy = k / s.Uw, ox = k - oy * s.Uw;
        double acc = 0.0;
        for (int q = 0; q <= 2 * r; ++q) acc += g.w[q] * (double)up[reflect_index(oy + q - r, s.Uh) * s.Uw + ox];
        t[k] = (float)acc;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < s.Uh * s.Uw; k += kThreads) {
        const int oy = k / s.Uw, ox = k - oy * s.Uw;
        double acc = 0.0;
        for (int q = 0; q <= 2 * r; ++q) acc += g.w[q] * (double)t[oy * s.Uw + reflect_index(ox + q - r, s.Uw)];
        up[k] = (float)acc;
    }
}

struct Vote {
    float w, vx, vy;
};

__device__ inline Vote isophote_vote(const float* F, const uint8_t* m, RoiShape s, int iy, int ix) {
#pragma clang fp contract(off)
    const int U = s.Uw;
    const int y = iy + 1, x = ix + 1;
    const float c = F[y * U + x];
    const float l = F[y * U + x - 1], r = F[y * U + x + 1], u = F[(y - 1) * U + x], d = F[(y + 1) * U + x];
    const float Ix = (r - l) / 2.0f;
    const float Iy = (d - u) / 2.0f;
    const float Ixx = (r - 2.0f * c) + l;
    const float Iyy = (d - 2.0f * c) + u;
    const float Ixy = (((F[(y + 1) * U + x + 1] - F[(y + 1) * U + x - 1]) - F[(y - 1) * U + x + 1]) +
                       F[(y - 1) * U + x - 1]) / 4.0f;
    const float eps = 1e-6f;
    const float g2 = Ix * Ix + Iy * Iy;
    const float den = (Iy * Iy * Ixx - 2.0f * Ix * Ixy * Iy) + Ix * Ix * Iyy;
    bool valid = m[y * U + x] && fabsf(den) > eps && g2 > eps;
    const float sden = fabsf(den) > eps ? den : 1.0f;
    const float Dx = (-Ix * g2) / sden;
    const float Dy = (-Iy * g2) / sden;
    const float curv = sqrtf((Ixx * Ixx + 2.0f * Ixy * Ixy) + Iyy * Iyy);
    const float xp = ((float)ix + 1.0f) - ((float)s.Uw - 1.0f) / 2.0f;
    const float yp = ((float)iy + 1.0f) - ((float)s.Uh - 1.0f) / 2.0f;
    const float half = (float)(s.Uw > s.Uh ? s.Uw : s.Uh) / 2.0f;
    valid = valid && (Dx * Dx + Dy * Dy) < half * half;
    Vote v;
    v.w = valid ? curv : 0.0f;
    v.vx = xp + Dx;
    v.vy = yp + Dy;
    return v;
}

__global__ __launch_bounds__(kThreads) void isophote_kernel(const float* rois, const uint8_t* masks, int N,
                                                            RoiShape s, Gauss g, unsigned char* scratch,
                                                            size_t stride, double* centers, double* quality) {
#pragma clang fp contract(off)
    extern __shared__ unsigned char s_dyn[];
    __shared__ double red[4 * 3];
    unsigned char* base = scratch != nullptr ? scratch + (size_t)blockIdx.x * stride : s_dyn;
    const int UU = s.Uh * s.Uw;
    float* up = reinterpret_cast<float*>(base);
    float* tf = up + UU;
    uint8_t* m = reinterpret_cast<uint8_t*>(tf + (g.radius > 0 ? UU : 0));
    uint8_t* t = m + UU;
    const int Hd = s.Uh - 2, Wd = s.Uw - 2;
    for (int j = blockIdx.x; j < N; j += gridDim.x) {
        if (Hd < 1 || Wd < 1) {
            if (threadIdx.x == 0) to_roi(__builtin_nan(""), __builtin_nan(""), s, centers, quality, j, __builtin_inf());
            continue;
        }
        __syncthreads();
        upsample_roi(rois + (int64_t)j * s.Hm * s.Wm, masks + (int64_t)j * s.Hm * s.Wm, s, up, m);
        if (g.radius > 0) gauss_yx(up, tf, s, g);
        erode(m, t, s);
        __syncthreads();
        double a1[3] = {0.0, 0.0, 0.0};
        for (int k = threadIdx.x; k < Hd * Wd; k += kThreads) {
            const int iy = k / Wd, ix = k - iy * Wd;
            const Vote v = isophote_vote(up, m, s, iy, ix);
            const double w = (double)v.w;
            a1[0] += w;
            a1[1] += w * (double)v.vx;
            a1[2] += w * (double)v.vy;
        }
        block_sum<3>(a1, red);
        const double ws = fmax(a1[0], 1e-30);
        const double xc = a1[1] / ws, yc = a1[2] / ws;
        double a2[1] = {0.0};
        for (int k = threadIdx.x; k < Hd * Wd; k += kThreads) {
            const int iy = k / Wd, ix = k - iy * Wd;
            const Vote v = isophote_vote(up, m, s, iy, ix);
            const double ex = (double)v.vx - xc, ey = (double)v.vy - yc;
            a2[0] += (double)v.w * (ex * ex + ey * ey);
        }
        block_sum<1>(a2, red);
        if (threadIdx.x == 0) to_roi(xc, yc, s, centers, quality, j, sqrt(a2[0] / ws));
    }
}

constexpr size_t kLdsBudget = 60 * 1024;

}  // namespace

int voronoi_labels_launch(const double* seeds, int32_t n, int64_t H, int64_t W, int32_t* labels, int32_t* areas,
                          double* dist, hipStream_t stream) {
    const int64_t HW = H * W;
    if (areas != nullptr && n > 0) {
        vgrid_zero<<<(n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024, 256, 0, stream>>>(areas, n);
        SPX_LAUNCHED();
    }
    if (n == 0) {
        if (labels != nullptr) SPX_HIP(hipMemsetAsync(labels, 0xff, sizeof(int32_t) * (size_t)HW, stream));
        if (dist != nullptr) {
            set_error("spx_voronoi_labels: distances need at least one seed");
            return SPX_E_ARG;
        }
        return SPX_OK;
    }
    SeedGrid g;
    g.cell = (int)fmax(1.0, floor(sqrt((double)HW / (double)n)));
    g.gw = (int)((W + g.cell - 1) / g.cell);
    g.gh = (int)((H + g.cell - 1) / g.cell);
    const int64_t ncell = (int64_t)g.gw * g.gh;
    if (ncell >= ((int64_t)1 << 31) - 1) {
        set_error("spx_voronoi_labels: seed grid too large");
        return SPX_E_UNSUPPORTED;
    }
    size_t scan_bytes = 0;
    SPX_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (int32_t*)nullptr, (int32_t*)nullptr,
                                             (int)(ncell + 1), stream));
    // workspace: cnt[ncell + 1] | start[ncell + 1] | cursor[ncell + 1] | ids[n] | scan temp
    const size_t head = sizeof(int32_t) * (3 * (size_t)(ncell + 1) + (size_t)n);
    const size_t head_al = (head + 255) & ~(size_t)255;
    void* ws = nullptr;
    int rc = workspace(head_al + scan_bytes, &ws, stream);
    if (rc != SPX_OK) return rc;
    int32_t* cnt = static_cast<int32_t*>(ws);
    int32_t* start = cnt + (ncell + 1);
    int32_t* cursor = start + (ncell + 1);
    int32_t* ids = cursor + (ncell + 1);
    const int zb = (int)((ncell + 1 + 255) / 256 < 4096 ? (ncell + 1 + 255) / 256 : 4096);
    vgrid_zero<<<zb, 256, 0, stream>>>(cnt, ncell + 1);
    SPX_LAUNCHED();
    vgrid_count<<<(n + 255) / 256, 256, 0, stream>>>(seeds, n, g, cnt);
    SPX_LAUNCHED();
    // start[c] = seeds in cells before c; start[ncell] = n (the count of the extra cell is 0)
    SPX_HIP(hipcub::DeviceScan::ExclusiveSum(static_cast<char*>(ws) + head_al, scan_bytes, cnt, start,
                                             (int)(ncell + 1), stream));
    SPX_HIP(hipMemcpyAsync(cursor, start, sizeof(int32_t) * (size_t)(ncell + 1), hipMemcpyDeviceToDevice, stream));
    vgrid_scatter<<<(n + 255) / 256, 256, 0, stream>>>(seeds, n, g, cursor, ids);
    SPX_LAUNCHED();
    const int64_t blocks = (HW + kThreads - 1) / kThreads;
    voronoi_label_kernel<<<dim3((unsigned)blocks), kThreads, 0, stream>>>(seeds, start, ids, g, H, W, labels, areas,
                                                                          dist);
    SPX_LAUNCHED();
    return SPX_OK;
}

int voronoi_rois_launch(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, const int32_t* labels,
                        const double* rows, int32_t N, int32_t pad, int32_t Hm, int32_t Wm, float* rois,
                        uint8_t* masks, int32_t* origins, hipStream_t stream) {
    if (N == 0) return SPX_OK;
    const size_t lds = (((size_t)Hm * Wm + 15) & ~(size_t)15) + sizeof(float) * ((size_t)Hm * Wm + 2);
    if (lds > kLdsBudget) {
        set_error("spx_voronoi_rois: a " + std::to_string(Hm) + " x " + std::to_string(Wm) +
                  " ROI stack exceeds the kernel's LDS (at most about 12,000 ROI pixels)");
        return SPX_E_UNSUPPORTED;
    }
    voronoi_roi_kernel<<<N, kThreads, lds, stream>>>(img, H, W, pitch, labels, rows, N, pad, Hm, Wm, rois, masks,
                                                      origins);
    SPX_LAUNCHED();
    return SPX_OK;
}

namespace {
int estimator_launch(bool iso, const float* rois, const uint8_t* masks, int32_t N, int32_t Hm, int32_t Wm,
                     int32_t factor, int32_t margin, double sigma, double* centers, double* quality,
                     hipStream_t stream) {
    if (N == 0) return SPX_OK;
    RoiShape s;
    s.Hm = Hm;
    s.Wm = Wm;
    s.f = factor;
    s.Uh = Hm * factor;
    s.Uw = Wm * factor;
    const int min_up = s.Uh < s.Uw ? s.Uh : s.Uw;
    const int cap = (min_up - (iso ? 5 : 11)) / 2;
    s.margin = margin < (cap > 0 ? cap : 0) ? margin : (cap > 0 ? cap : 0);
    Gauss g{};
    g.radius = 0;
    if (iso && sigma > 0.0) {
        g.radius = (int)(4.0 * sigma + 0.5);
        if (2 * g.radius + 1 > kMaxGaussTaps) {
            set_error("spx_isophote_curvature: pre_smooth_sigma too large");
            return SPX_E_ARG;
        }
        double tot = 0.0;
        for (int q = -g.radius; q <= g.radius; ++q) {
            g.w[q + g.radius] = exp(-0.5 / (sigma * sigma) * (double)q * (double)q);
            tot += g.w[q + g.radius];
        }
        for (int q = 0; q <= 2 * g.radius; ++q) g.w[q] /= tot;
    }
    const size_t UU = (size_t)s.Uh * s.Uw;
    const size_t bytes = ((UU * 4 * (iso && g.radius > 0 ? 2 : 1) + UU * 2) + 255) & ~(size_t)255;
    unsigned char* scratch = nullptr;
    int grid = N;
    size_t lds = bytes;
    if (bytes > kLdsBudget) {
        int dev = 0, cus = 0;
        SPX_HIP(hipGetDevice(&dev));
        SPX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        grid = N < 4 * cus ? N : 4 * cus;
        void* ws = nullptr;
        int rc = workspace(bytes * (size_t)grid, &ws, stream);
        if (rc != SPX_OK) return rc;
        scratch = static_cast<unsigned char*>(ws);
        lds = 0;
    }
    if (iso)
        isophote_kernel<<<grid, kThreads, lds, stream>>>(rois, masks, N, s, g, scratch, bytes, centers, quality);
    else
        radial_symmetry_kernel<<<grid, kThreads, lds, stream>>>(rois, masks, N, s, scratch, bytes, centers, quality);
    SPX_LAUNCHED();
    return SPX_OK;
}
}  // namespace

int radial_symmetry_launch(const float* rois, const uint8_t* masks, int32_t N, int32_t Hm, int32_t Wm,
                           int32_t factor, int32_t margin, double* centers, double* quality, hipStream_t stream) {
    return estimator_launch(false, rois, masks, N, Hm, Wm, factor, margin, 0.0, centers, quality, stream);
}

int isophote_curvature_launch(const float* rois, const uint8_t* masks, int32_t N, int32_t Hm, int32_t Wm,
                              int32_t factor, int32_t margin, double sigma, double* centers, double* quality,
                              hipStream_t stream) {
    return estimator_launch(true, rois, masks, N, Hm, Wm, factor, margin, sigma, centers, quality, stream);
}

}  // namespace spx
