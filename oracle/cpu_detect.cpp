// CPU restatement of the reference's CPU detect path in C++ with OpenMP: Otsu threshold,
// 8-connected components with their stats, area filter and the per-label log-quadratic
// refinement. TEST INFRASTRUCTURE ONLY: the all-cores CPU baseline timed beside the GPU; it
// restates oracle/segment.py (detect_centers_logquad), which restates reference
// subpx/_cpu/centers.py:281-408 (Otsu :258, components subpx/_cpu/components.py:18-34,
// per-label ROI :376-391, _bg_from_border :31-42, refine_logquadratic_cpu :67-107).
//
// Differences of method, not of result: the least-squares fit solves the normal equations
// of the six monomials on coordinates centred in the ROI (the reference calls an SVD solver,
// numpy.linalg.lstsq); both give the same minimiser of a well-posed fit to rounding. The
// components are found by union-find over row strips; the output is ordered by each
// component's first pixel in raster order (the oracle's order differs; tests compare sets).
#include <stdint.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

int num_threads(int want) {
#ifdef _OPENMP
    return want > 0 ? want : omp_get_max_threads();
#else
    (void)want;
    return 1;
#endif
}

// OpenCV's Otsu recurrence over the full 65,536-bin histogram (oracle/segment.py
// otsu_from_hist): sequential FP64, first strict maximum, bins with a class weight within
// FLT_EPSILON of 0 or 1 skipped.
int otsu_from_hist(const std::vector<int64_t>& h) {
    int64_t N = 0;
    for (int64_t v : h) N += v;
    const double scale = 1.0 / (double)N;
    double mu = 0.0;
    for (size_t i = 0; i < h.size(); ++i) mu += (double)i * (double)h[i];
    mu *= scale;
    double mu1 = 0.0, q1 = 0.0, max_sigma = 0.0;
    int max_val = 0;
    const double eps = (double)FLT_EPSILON;
    for (size_t i = 0; i < h.size(); ++i) {
        const double p_i = (double)h[i] * scale;
        mu1 *= q1;
        q1 += p_i;
        const double q2 = 1.0 - q1;
        if (std::min(q1, q2) < eps || std::max(q1, q2) > 1.0 - eps) continue;
        mu1 = (mu1 + (double)i * p_i) / q1;
        const double mu2 = (mu - q1 * mu1) / q2;
        const double sigma = q1 * q2 * (mu1 - mu2) * (mu1 - mu2);
        if (sigma > max_sigma) {
            max_sigma = sigma;
            max_val = (int)i;
        }
    }
    return max_val;
}

int find(std::vector<int32_t>& parent, int32_t a) {
    while (parent[a] != a) {
        parent[a] = parent[parent[a]];
        a = parent[a];
    }
    return a;
}

void unite(std::vector<int32_t>& parent, int32_t a, int32_t b) {
    a = find(parent, a);
    b = find(parent, b);
    if (a == b) return;
    // the smaller index becomes the root, so every root is its component's first pixel
    if (a < b) parent[b] = a;
    else parent[a] = b;
}

double median(std::vector<double>& v) {
    const size_t n = v.size();
    std::sort(v.begin(), v.end());
    return (n % 2) ? v[n / 2] : 0.5 * (v[n / 2 - 1] + v[n / 2]);
}

// Solves the symmetric positive definite 6 x 6 system M x = b by Cholesky; false when a
// pivot is not positive.
bool solve6(double M[6][6], double b[6], double x[6]) {
    double L[6][6] = {};
    for (int j = 0; j < 6; ++j) {
        double d = M[j][j];
        for (int k = 0; k < j; ++k) d -= L[j][k] * L[j][k];
        if (!(d > 0.0)) return false;
        L[j][j] = std::sqrt(d);
        for (int i = j + 1; i < 6; ++i) {
            double s = M[i][j];
            for (int k = 0; k < j; ++k) s -= L[i][k] * L[j][k];
            L[i][j] = s / L[j][j];
        }
    }
    double y[6];
    for (int i = 0; i < 6; ++i) {
        double s = b[i];
        for (int k = 0; k < i; ++k) s -= L[i][k] * y[k];
        y[i] = s / L[i][i];
    }
    for (int i = 5; i >= 0; --i) {
        double s = y[i];
        for (int k = i + 1; k < 6; ++k) s -= L[k][i] * x[k];
        x[i] = s / L[i][i];
    }
    for (int i = 0; i < 6; ++i)
        if (!std::isfinite(x[i])) return false;
    return true;
}

// refine_logquadratic_cpu on the ROI [x0, x1) x [y0, y1) with the mask labels == lab:
// ROI-local centre, or false.
bool refine_logquad(const uint16_t* img, int64_t pitch, const int32_t* labels, int64_t W, int64_t x0, int64_t y0,
                    int64_t x1, int64_t y1, int32_t lab, double* cx, double* cy) {
    const int64_t w = x1 - x0, h = y1 - y0;
    int64_t n = 0;
    for (int64_t r = y0; r < y1; ++r)
        for (int64_t c = x0; c < x1; ++c) n += labels[r * W + c] == lab;
    if (n < 6) return false;
    // background: median of the 2-px ring (rows 0-1 and h-2..h-1 whole, columns 0-1 and
    // w-2..w-1 of every row, corners counted twice), or of the whole ROI when it is too small
    std::vector<double> ring;
    if (h <= 4 || w <= 4) {
        for (int64_t r = y0; r < y1; ++r)
            for (int64_t c = x0; c < x1; ++c) ring.push_back((double)img[r * pitch + c]);
    } else {
        for (int64_t r : {y0, y0 + 1, y1 - 2, y1 - 1})
            for (int64_t c = x0; c < x1; ++c) ring.push_back((double)img[r * pitch + c]);
        for (int64_t r = y0; r < y1; ++r)
            for (int64_t c : {x0, x0 + 1, x1 - 2, x1 - 1}) ring.push_back((double)img[r * pitch + c]);
    }
    const double bg = median(ring);
    double vmax = -1e300;
    for (int64_t r = y0; r < y1; ++r)
        for (int64_t c = x0; c < x1; ++c)
            if (labels[r * W + c] == lab) vmax = std::max(vmax, (double)img[r * pitch + c] - bg);
    if (!(vmax > 0.0)) return false;
    // normal equations of z = a x^2 + b y^2 + c x y + d x + e y + f on coordinates centred
    // at the ROI's middle (u = x - xm, v = y - ym)
    const double xm = 0.5 * (double)(w - 1), ym = 0.5 * (double)(h - 1);
    double M[6][6] = {}, rhs[6] = {};
    for (int64_t r = y0; r < y1; ++r)
        for (int64_t c = x0; c < x1; ++c) {
            if (labels[r * W + c] != lab) continue;
            const double v = (double)img[r * pitch + c] - bg;
            const double z = std::log(std::max(v, 1e-9));
            const double u = (double)(c - x0) - xm, t = (double)(r - y0) - ym;
            const double m[6] = {u * u, t * t, u * t, u, t, 1.0};
            for (int i = 0; i < 6; ++i) {
                rhs[i] += m[i] * z;
                for (int j = 0; j <= i; ++j) M[i][j] += m[i] * m[j];
            }
        }
    for (int i = 0; i < 6; ++i)
        for (int j = i + 1; j < 6; ++j) M[i][j] = M[j][i];
    double co[6];
    if (!solve6(M, rhs, co)) return false;
    const double a = co[0], b = co[1], c = co[2], d = co[3], e = co[4];
    // Hessian [[2a, c], [c, 2b]]: 2-norm condition number from its eigenvalues
    const double tr = 2.0 * a + 2.0 * b, det = 4.0 * a * b - c * c;
    const double disc = std::sqrt(std::max(0.25 * tr * tr - det, 0.0));
    const double l1 = std::fabs(0.5 * tr + disc), l2 = std::fabs(0.5 * tr - disc);
    const double lmax = std::max(l1, l2), lmin = std::min(l1, l2);
    if (!(lmin > 0.0) || lmax / lmin > 1e8) return false;
    // grad = 0: [2a c; c 2b] [u v]^T = -[d e]^T
    const double u = (-d * 2.0 * b + e * c) / det, v = (-e * 2.0 * a + d * c) / det;
    const double px = u + xm, py = v + ym;
    if (!std::isfinite(px) || !std::isfinite(py) || !(px >= 0.0 && px < (double)w && py >= 0.0 && py < (double)h))
        return false;
    *cx = px;
    *cy = py;
    return true;
}

}  // namespace

extern "C" {

// Otsu threshold of the image's full-resolution histogram (threads: per-thread histograms).
int32_t cpu_otsu_u16(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, int32_t threads) {
    const int T = num_threads(threads);
    std::vector<std::vector<int64_t>> hs(T, std::vector<int64_t>(65536, 0));
#pragma omp parallel for num_threads(T) schedule(static)
    for (int64_t r = 0; r < H; ++r) {
#ifdef _OPENMP
        std::vector<int64_t>& h = hs[omp_get_thread_num()];
#else
        std::vector<int64_t>& h = hs[0];
#endif
        for (int64_t c = 0; c < W; ++c) ++h[img[r * pitch + c]];
    }
    for (int t = 1; t < T; ++t)
        for (int i = 0; i < 65536; ++i) hs[0][i] += hs[t][i];
    return otsu_from_hist(hs[0]);
}

// The CPU detect path: Otsu (t < 0) or the given threshold, foreground = g > t, 8-connected
// components, area in [area_min, area_max], per-label log-quadratic on the bounding box grown
// by pad. Writes at most `cap` centres (x, y) to out and returns how many components were
// refined successfully (-1: cap too small); *threshold_out receives the threshold.
int64_t cpu_detect_logquad(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, int32_t t_in, int64_t area_min,
                           int64_t area_max, int32_t pad, int32_t threads, double* out, int64_t cap,
                           int32_t* threshold_out) {
    const int T = num_threads(threads);
    const int t = t_in >= 0 ? t_in : cpu_otsu_u16(img, H, W, pitch, T);
    if (threshold_out) *threshold_out = t;
    const int64_t N = H * W;
    std::vector<int32_t> parent(N);
    // pass 1: union-find inside row strips (each strip touches only its own pixels)
    const int64_t strips = std::max<int64_t>(1, std::min<int64_t>(T * 4, H));
#pragma omp parallel for num_threads(T) schedule(dynamic, 1)
    for (int64_t s = 0; s < strips; ++s) {
        const int64_t r0 = H * s / strips, r1 = H * (s + 1) / strips;
        for (int64_t r = r0; r < r1; ++r)
            for (int64_t c = 0; c < W; ++c) {
                const int64_t p = r * W + c;
                if (img[r * pitch + c] <= t) {
                    parent[p] = -1;
                    continue;
                }
                parent[p] = (int32_t)p;
                if (c > 0 && parent[p - 1] >= 0) unite(parent, (int32_t)p, (int32_t)(p - 1));
                if (r > r0) {
                    const int64_t q = p - W;
                    if (c > 0 && parent[q - 1] >= 0) unite(parent, (int32_t)p, (int32_t)(q - 1));
                    if (parent[q] >= 0) unite(parent, (int32_t)p, (int32_t)q);
                    if (c + 1 < W && parent[q + 1] >= 0) unite(parent, (int32_t)p, (int32_t)(q + 1));
                }
            }
    }
    // pass 2: merge across strip boundaries
    for (int64_t s = 1; s < strips; ++s) {
        const int64_t r = H * s / strips;
        if (r <= 0 || r >= H) continue;
        for (int64_t c = 0; c < W; ++c) {
            const int64_t p = r * W + c;
            if (parent[p] < 0) continue;
            const int64_t q = p - W;
            if (c > 0 && parent[q - 1] >= 0) unite(parent, (int32_t)p, (int32_t)(q - 1));
            if (parent[q] >= 0) unite(parent, (int32_t)p, (int32_t)q);
            if (c + 1 < W && parent[q + 1] >= 0) unite(parent, (int32_t)p, (int32_t)(q + 1));
        }
    }
    // pass 3: every pixel's root (parallel: roots no longer change)
#pragma omp parallel for num_threads(T) schedule(static)
    for (int64_t p = 0; p < N; ++p) {
        if (parent[p] < 0) continue;
        int32_t a = parent[p];
        while (parent[a] != a) a = parent[a];
        parent[p] = a;
    }
    // labels by root in raster order of the roots (each root is its component's first pixel)
    std::vector<int32_t> label_of(N, 0);
    int32_t nlab = 0;
    for (int64_t p = 0; p < N; ++p)
        if (parent[p] == p) label_of[p] = ++nlab;
    std::vector<int32_t> labels(N, 0);
#pragma omp parallel for num_threads(T) schedule(static)
    for (int64_t p = 0; p < N; ++p) labels[p] = parent[p] < 0 ? 0 : label_of[parent[p]];
    // stats per label: per-thread partial boxes and areas, merged
    struct Box {
        int64_t x0, y0, x1, y1, area;
    };
    std::vector<std::vector<Box>> part(T, std::vector<Box>(nlab + 1, Box{W, H, -1, -1, 0}));
#pragma omp parallel for num_threads(T) schedule(static)
    for (int64_t r = 0; r < H; ++r) {
#ifdef _OPENMP
        std::vector<Box>& bx = part[omp_get_thread_num()];
#else
        std::vector<Box>& bx = part[0];
#endif
        for (int64_t c = 0; c < W; ++c) {
            const int32_t l = labels[r * W + c];
            if (l == 0) continue;
            Box& b = bx[l];
            b.x0 = std::min(b.x0, c);
            b.y0 = std::min(b.y0, r);
            b.x1 = std::max(b.x1, c);
            b.y1 = std::max(b.y1, r);
            ++b.area;
        }
    }
    std::vector<Box> box(nlab + 1, Box{W, H, -1, -1, 0});
    for (int tt = 0; tt < T; ++tt)
        for (int32_t l = 1; l <= nlab; ++l) {
            const Box& b = part[tt][l];
            box[l].x0 = std::min(box[l].x0, b.x0);
            box[l].y0 = std::min(box[l].y0, b.y0);
            box[l].x1 = std::max(box[l].x1, b.x1);
            box[l].y1 = std::max(box[l].y1, b.y1);
            box[l].area += b.area;
        }
    std::vector<int32_t> keep;
    for (int32_t l = 1; l <= nlab; ++l)
        if (box[l].area >= area_min && box[l].area <= area_max) keep.push_back(l);
    const int64_t K = (int64_t)keep.size();
    std::vector<double> cen(2 * K, 0.0);
    std::vector<uint8_t> ok(K, 0);
#pragma omp parallel for num_threads(T) schedule(dynamic, 16)
    for (int64_t k = 0; k < K; ++k) {
        const Box& b = box[keep[k]];
        const int64_t x0 = std::max<int64_t>(0, b.x0 - pad), y0 = std::max<int64_t>(0, b.y0 - pad);
        const int64_t x1 = std::min<int64_t>(W, b.x1 + 1 + pad), y1 = std::min<int64_t>(H, b.y1 + 1 + pad);
        double cx = 0.0, cy = 0.0;
        if (refine_logquad(img, pitch, labels.data(), W, x0, y0, x1, y1, keep[k], &cx, &cy)) {
            cen[2 * k] = (double)x0 + cx;
            cen[2 * k + 1] = (double)y0 + cy;
            ok[k] = 1;
        }
    }
    int64_t n = 0;
    for (int64_t k = 0; k < K; ++k) n += ok[k];
    if (n > cap) return -1;
    int64_t j = 0;
    for (int64_t k = 0; k < K; ++k)
        if (ok[k]) {
            out[2 * j] = cen[2 * k];
            out[2 * j + 1] = cen[2 * k + 1];
            ++j;
        }
    return n;
}

}  // extern "C"
