// FP64 CPU restatement of the batched erf-disk / erf-ellipse LM fit, on every host core (OpenMP).
//
// TEST INFRASTRUCTURE ONLY: the CPU baseline that bench.py times beside the GPU (and that the
// tests check against oracle/erfdisk.py). Nothing in the product path loads it.
//
// The algorithm is oracle/erfdisk.py's, statement for statement: patch of (2h+1)^2 pixels at
// rint(seed), pixels outside the image or (with neighbour lists) outside the feature's Voronoi
// cell left out; moment start (ring mean, half-peak in-set centroid, r from its area, A its mean,
// s = init_sigma, q = 1); Marquardt damping lambda * max(M_aa, 1e-12 max M) on the FP64 normal
// equations; LDL^T solve with the oracle's pivot rule; steps refused outside the box; a step
// below step_tol ends the fit (applied without a trial); a trial is taken when its cost is at
// most cost (1 + 1e-6). The model uses std::erfc in FP64 throughout.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include <omp.h>

namespace {

constexpr double kSqrt2 = 1.4142135623730951;
constexpr double kInvSqrt2Pi = 0.3989422804014327;
enum { CONVERGED = 0, MAX_ITER = 1, SINGULAR = 2, DIVERGED = 3, TOO_FEW = 4, OUTSIDE = 5, BAD_START = 6 };

struct Opts {
    int max_iter;
    double step_tol, lambda0, lambda_up, lambda_down, init_sigma, accept_rtol;
};

// The oracle's cell forms of feature f: f_j(q) = (a qx + b qy) + c, m = max(|a|, |b|).
struct Cell {
    double a, b, c, m;
    bool lower;
};

void cell_forms(const float* seeds, int64_t K, int64_t f, const int32_t* row, int nb, std::vector<Cell>& out) {
    out.clear();
    const double x0 = seeds[2 * f], y0 = seeds[2 * f + 1];
    const double n0 = x0 * x0 + y0 * y0;
    for (int j = 0; j < nb; ++j) {
        const int64_t id = row[j];
        if (id < 0 || id >= K) continue;
        const double xj = seeds[2 * id], yj = seeds[2 * id + 1];
        const double a = 2.0 * (x0 - xj), b = 2.0 * (y0 - yj);
        out.push_back({a, b, (xj * xj + yj * yj) - n0, std::max(std::fabs(a), std::fabs(b)), id < f});
    }
}

bool in_cell(const std::vector<Cell>& cs, double px, double py) {
    for (const Cell& c : cs) {
        const double fj = (c.a * px + c.b * py) + c.c;
        if (!(fj > 0.0 || (fj == 0.0 && !c.lower))) return false;
    }
    return true;
}

bool on_edge(const std::vector<Cell>& cs, double px, double py) {
    for (const Cell& c : cs) {
        const double gj = ((c.a * px + c.b * py) + c.c) - c.m;
        if (gj < 0.0 || (gj == 0.0 && c.lower)) return true;
    }
    return false;
}

// Normal equations (P x P), J^T r and cost at p over the listed pixels.
void evaluate(const double* p, int P, bool ell, const std::vector<double>& fx, const std::vector<double>& fy,
              const std::vector<double>& v, double* M, double* g, double& cost) {
    const double x0 = p[0], y0 = p[1], r = p[2], s = p[3], A = p[4];
    const double q = ell ? p[5] : 1.0, B = p[P - 1];
    std::fill(M, M + P * P, 0.0);
    std::fill(g, g + P, 0.0);
    cost = 0.0;
    double J[7];
    for (size_t k = 0; k < v.size(); ++k) {
        const double dx = fx[k] - x0, dy = fy[k] - y0;
        const double dyq = dy / q;
        const double rho = std::sqrt(dx * dx + dyq * dyq);
        const double u = (rho - r) / (kSqrt2 * s);
        const double E = 0.5 * std::erfc(u);
        const double G = A * kInvSqrt2Pi / s * std::exp(-u * u);
        const bool pos = rho > 0.0;
        J[0] = pos ? G * dx / rho : 0.0;
        J[1] = pos ? G * dy / (q * q) / rho : 0.0;
        J[2] = G;
        J[3] = G * u * kSqrt2;
        J[4] = E;
        if (ell) J[5] = pos ? G * dy * dy / (q * q * q) / rho : 0.0;
        J[P - 1] = 1.0;
        const double res = v[k] - (B + A * E);
        for (int a = 0; a < P; ++a) {
            g[a] += J[a] * res;
            for (int b = 0; b <= a; ++b) M[a * P + b] += J[a] * J[b];
        }
        cost += res * res;
    }
    for (int a = 0; a < P; ++a)
        for (int b = a + 1; b < P; ++b) M[a * P + b] = M[b * P + a];
}

bool ldl_solve(const double* M, const double* g, int P, double* x) {
    double L[49] = {0.0}, D[7], y[7];
    for (int j = 0; j < P; ++j) {
        double d = M[j * P + j];
        for (int k = 0; k < j; ++k) d -= L[j * P + k] * L[j * P + k] * D[k];
        if (!(d > 1e-13 * M[j * P + j]) || !(d > 0.0)) return false;
        D[j] = d;
        for (int i = j + 1; i < P; ++i) {
            double sacc = M[i * P + j];
            for (int k = 0; k < j; ++k) sacc -= L[i * P + k] * L[j * P + k] * D[k];
            L[i * P + j] = sacc / d;
        }
    }
    for (int i = 0; i < P; ++i) {
        double sacc = g[i];
        for (int k = 0; k < i; ++k) sacc -= L[i * P + k] * y[k];
        y[i] = sacc;
    }
    for (int i = P - 1; i >= 0; --i) {
        double sacc = y[i] / D[i];
        for (int k = i + 1; k < P; ++k) sacc -= L[k * P + i] * x[k];
        x[i] = sacc;
    }
    for (int i = 0; i < P; ++i)
        if (!std::isfinite(x[i])) return false;
    return true;
}

void fit_one(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, const float* seeds, int64_t K, int64_t f, int h,
             bool ell, const int32_t* nbr, int nb, const Opts& o, double* out, double* rms_out, int32_t* it_out,
             uint8_t* st_out) {
    const int P = ell ? 7 : 6;
    for (int i = 0; i < 7; ++i) out[i] = NAN;
    *rms_out = NAN;
    *it_out = 0;
    if (h < 2 || h > 31) {
        *st_out = TOO_FEW;
        return;
    }
    const double cx = std::nearbyint((double)seeds[2 * f]), cy = std::nearbyint((double)seeds[2 * f + 1]);
    const int side = 2 * h + 1;
    std::vector<Cell> cs;
    if (nbr != nullptr) cell_forms(seeds, K, f, nbr + f * nb, nb, cs);
    std::vector<double> fx, fy, v;
    fx.reserve(side * side);
    fy.reserve(side * side);
    v.reserve(side * side);
    // moment start over the valid pixels
    double ring_sum = 0.0, vmax = -1e300;
    int64_t ring_n = 0;
    std::vector<uint8_t> edge;
    edge.reserve(side * side);
    for (int row = 0; row < side; ++row)
        for (int col = 0; col < side; ++col) {
            const int64_t gx = (int64_t)cx - h + col, gy = (int64_t)cy - h + row;
            if (gx < 0 || gx >= W || gy < 0 || gy >= H) continue;
            bool ring = row == 0 || row == side - 1 || col == 0 || col == side - 1;
            if (!cs.empty()) {
                if (!in_cell(cs, (double)gx, (double)gy)) continue;
                ring = ring || on_edge(cs, (double)gx, (double)gy);
            }
            const double val = img[gy * pitch + gx];
            fx.push_back(col - h);
            fy.push_back(row - h);
            v.push_back(val);
            edge.push_back(ring);
            if (ring) {
                ring_sum += val;
                ++ring_n;
            }
        }
    const int64_t npix = (int64_t)v.size();
    if (npix < P + 2) {
        *st_out = TOO_FEW;
        return;
    }
    const double B0 = ring_n ? ring_sum / (double)ring_n : 0.0;
    for (double val : v) vmax = std::max(vmax, val - B0);
    double s0 = 0.0, sx = 0.0, sy = 0.0;
    int64_t nin = 0;
    for (int64_t k = 0; k < npix; ++k) {
        const double w = v[k] - B0;
        if (w > 0.5 * vmax) {
            s0 += w;
            sx += w * fx[k];
            sy += w * fy[k];
            ++nin;
        }
    }
    if (!(vmax > 0.0) || !(s0 > 0.0) || nin < 1) {
        *st_out = BAD_START;
        return;
    }
    double p[7];
    p[0] = sx / s0;
    p[1] = sy / s0;
    p[2] = std::sqrt((double)nin / M_PI);
    p[3] = o.init_sigma;
    p[4] = s0 / (double)nin;
    if (ell) p[5] = 1.0;
    p[P - 1] = B0;
    double M[49], g[7], cost;
    evaluate(p, P, ell, fx, fy, v, M, g, cost);
    double lam = o.lambda0;
    const double lim = h + 0.5;
    int status = std::isfinite(cost) ? MAX_ITER : DIVERGED;
    int it = 0;
    bool any_solve = false;
    while (status == MAX_ITER && it < o.max_iter) {
        ++it;
        double Md[49], delta[7];
        std::copy(M, M + P * P, Md);
        double dmax = 0.0;
        for (int a = 0; a < P; ++a) dmax = std::max(dmax, M[a * P + a]);
        for (int a = 0; a < P; ++a) Md[a * P + a] += lam * std::max(M[a * P + a], 1e-12 * dmax);
        if (!ldl_solve(Md, g, P, delta)) {
            lam *= o.lambda_up;
            continue;
        }
        any_solve = true;
        double tr[7];
        for (int i = 0; i < P; ++i) tr[i] = p[i] + delta[i];
        double step = std::max(std::max(std::fabs(delta[0]), std::fabs(delta[1])),
                               std::max(std::fabs(delta[2]), std::fabs(delta[3])));
        if (ell) step = std::max(step, std::fabs(delta[5]) * std::fabs(p[2]));
        bool inside = std::fabs(tr[0]) <= lim && std::fabs(tr[1]) <= lim && tr[2] > 0.25 && tr[2] < 2.0 * lim &&
                      tr[3] > 0.05 && tr[3] < lim;
        if (ell) inside = inside && tr[5] > 0.25 && tr[5] < 4.0;
        if (!inside) {
            lam *= o.lambda_up;
            continue;
        }
        if (step < o.step_tol) {
            std::copy(tr, tr + P, p);
            status = CONVERGED;
            break;
        }
        double Mt[49], gt[7], cost_t;
        evaluate(tr, P, ell, fx, fy, v, Mt, gt, cost_t);
        if (cost_t <= cost * (1.0 + o.accept_rtol)) {
            std::copy(tr, tr + P, p);
            std::copy(Mt, Mt + P * P, M);
            std::copy(gt, gt + P, g);
            cost = cost_t;
            lam = std::max(lam * o.lambda_down, 1e-12);
        } else {
            lam *= o.lambda_up;
        }
    }
    if (!any_solve && status == MAX_ITER) status = SINGULAR;
    if (!std::isfinite(cost)) status = DIVERGED;
    if (status <= MAX_ITER && (std::fabs(p[0]) > h || std::fabs(p[1]) > h)) status = OUTSIDE;
    out[0] = p[0] + cx;
    out[1] = p[1] + cy;
    out[2] = p[2];
    out[3] = p[3];
    out[4] = p[4];
    out[5] = p[P - 1];
    out[6] = ell ? p[5] : 1.0;
    *rms_out = std::sqrt(cost / (double)std::max<int64_t>(npix - P, 1));
    *it_out = it;
    *st_out = (uint8_t)status;
}

}  // namespace

// Fits features idx[0 .. n) (all K when idx is NULL) on `threads` OpenMP threads (0: all the
// process may use). params (K, 7) [x0, y0, r, s, A, B, q], rms, iters, status are written for
// the fitted features only. nbr: (K, nb) neighbour ids or NULL. Returns the threads used.
extern "C" int cpu_fit_batch(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, const float* seeds,
                             const int16_t* half, const uint8_t* model, const int32_t* nbr, int32_t nb, int64_t K,
                             const int64_t* idx, int64_t n, int32_t max_iter, double step_tol, int32_t threads,
                             double* params, double* rms, int32_t* iters, uint8_t* status) {
    const Opts o{max_iter, step_tol, 1e-3, 10.0, 0.1, 1.0, 1e-6};
    const int nt = threads > 0 ? threads : omp_get_max_threads();
    const int64_t m = idx ? n : K;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt)
    for (int64_t t = 0; t < m; ++t) {
        const int64_t f = idx ? idx[t] : t;
        fit_one(img, H, W, pitch, seeds, K, f, half[f], model[f] == 1, nbr, nb, o, params + 7 * f, rms + f,
                iters + f, status + f);
    }
    return nt;
}
