// Edge-erf centre refinement: gradient-moment seeds + 1-D erf-edge Gauss-Newton fits of
// left / right / top / bottom edge profiles, FP64, one thread per feature (or per profile).
//
// Restates the reference's `_refine_edge_erf_core` (subpx/_gpu/centers.py:782-879) with
// `_smooth_binom5_gpu` (:518-523), `_edge_from_profile_gradmoment_gpu` (:526-544),
// `_edge_polarity_from_profile_gpu` (:648-662) and `_edge_from_profile_erf_gpu`
// (:665-779). The reference runs each step as batched CuPy expressions (about 60 launches
// per GN iteration); here one thread carries a feature through both passes and all four
// edges, with the 5x5 normal equations solved by Gaussian elimination with partial
// pivoting and the final cond(J^T J) gate from Jacobi eigenvalues (the 2-norm condition of
// a symmetric positive matrix). Profiles and seeds are FP64 here (the reference forms
// them in FP32).
#include "spx_internal.hpp"

namespace spx {
namespace {

constexpr int kLMax = 31;  // edge_rad <= 15
constexpr int kBandMax = 15;

// Profiles are formed with the same IEEE operations as the FP64 restatement (no FMA
// contraction, band mean as a division), so exact ties of |diff| in the argmax break the
// same way on both sides.
__device__ inline void smooth5(double* x, int L) {
#pragma clang fp contract(off)
    double y[kLMax];
    for (int j = 0; j < L; ++j) {
        const double xm2 = x[max(j - 2, 0)], xm1 = x[max(j - 1, 0)];
        const double xp1 = x[min(j + 1, L - 1)], xp2 = x[min(j + 2, L - 1)];
        y[j] = (xm2 + 4.0 * xm1 + 6.0 * x[j] + 4.0 * xp1 + xp2) * (1.0 / 16.0);
    }
    for (int j = 0; j < L; ++j) x[j] = y[j];
}

__device__ inline bool gradmoment(const double* prof, int L, double base, int loc, double power, double& x_edge) {
#pragma clang fp contract(off)
    const int Lm1 = L - 1;
    int ip = 0;
    double best = -1.0;
    for (int j = 0; j < Lm1; ++j) {
        const double d = fabs(prof[j + 1] - prof[j]);
        if (d > best) {
            best = d;
            ip = j;
        }
    }
    const int ip0 = min(max(ip, loc), max(loc, Lm1 - 1 - loc));
    double sw = 0.0, swx = 0.0;
    for (int t = -loc; t <= loc; ++t) {
        const int jj = min(max(ip0 + t, 0), Lm1 - 1);
        const double d = fabs(prof[jj + 1] - prof[jj]);
        const double w = (power == 8.0) ? (d * d) * (d * d) * ((d * d) * (d * d)) : pow(d, power);
        sw += w;
        swx += w * (base + (double)jj + 0.5);
    }
    sw += 1e-20;
    x_edge = swx / sw;
    return sw > 1e-12;
}

__device__ inline double polarity(const double* prof, int L, double base, double seed, int loc) {
    const int Lm1 = L - 1;
    int ip = 0;
    double best = 1e300;
    for (int j = 0; j < Lm1; ++j) {
        const double dd = fabs(base + (double)j + 0.5 - seed);
        if (dd < best) {
            best = dd;
            ip = j;
        }
    }
    ip = min(max(ip, 0), max(0, Lm1 - 1));
    double sgn = 0.0;
    for (int t = -loc; t <= loc; ++t) {
        const int jj = min(max(ip + t, 0), max(0, Lm1 - 1));
        sgn += prof[jj + 1] - prof[jj];
    }
    return sgn >= 0.0 ? 1.0 : -1.0;
}

// Gaussian elimination with partial pivoting on a 5x5 system; false when singular.
__device__ inline bool solve5(double A[5][5], double b[5], double x[5]) {
    double M[5][6];
    for (int i = 0; i < 5; ++i) {
        for (int j = 0; j < 5; ++j) M[i][j] = A[i][j];
        M[i][5] = b[i];
    }
    for (int c = 0; c < 5; ++c) {
        int p = c;
        double pv = fabs(M[c][c]);
        for (int r = c + 1; r < 5; ++r)
            if (fabs(M[r][c]) > pv) {
                pv = fabs(M[r][c]);
                p = r;
            }
        if (!(pv > 0.0)) return false;
        if (p != c)
            for (int j = 0; j < 6; ++j) {
                const double t = M[c][j];
                M[c][j] = M[p][j];
                M[p][j] = t;
            }
        for (int r = c + 1; r < 5; ++r) {
            const double f = M[r][c] / M[c][c];
            for (int j = c; j < 6; ++j) M[r][j] -= f * M[c][j];
        }
    }
    for (int i = 4; i >= 0; --i) {
        double s = M[i][5];
        for (int j = i + 1; j < 5; ++j) s -= M[i][j] * x[j];
        x[i] = s / M[i][i];
    }
    return true;
}

// 2-norm condition number of a symmetric matrix: ratio of extreme |eigenvalues| after
// cyclic Jacobi sweeps.
__device__ inline double cond_sym5(const double A[5][5]) {
    double M[5][5];
    for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 5; ++j) M[i][j] = A[i][j];
    for (int sweep = 0; sweep < 12; ++sweep) {
        double off = 0.0;
        for (int i = 0; i < 5; ++i)
            for (int j = i + 1; j < 5; ++j) off += M[i][j] * M[i][j];
        if (!(off > 1e-300)) break;
        for (int p = 0; p < 5; ++p)
            for (int q = p + 1; q < 5; ++q) {
                if (M[p][q] == 0.0) continue;
                const double theta = (M[q][q] - M[p][p]) / (2.0 * M[p][q]);
                const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < 5; ++k) {
                    const double mkp = M[k][p], mkq = M[k][q];
                    M[k][p] = c * mkp - s * mkq;
                    M[k][q] = s * mkp + c * mkq;
                }
                for (int k = 0; k < 5; ++k) {
                    const double mpk = M[p][k], mqk = M[q][k];
                    M[p][k] = c * mpk - s * mqk;
                    M[q][k] = s * mpk + c * mqk;
                }
            }
    }
    double lo = 1e300, hi = 0.0;
    for (int i = 0; i < 5; ++i) {
        const double a = fabs(M[i][i]);
        lo = fmin(lo, a);
        hi = fmax(hi, a);
        if (!isfinite(M[i][i])) return __builtin_nan("");
    }
    return lo > 0.0 ? hi / lo : __builtin_inf();
}

struct ErfOpts {
    int fit_rad, gn_iters, loc_rad;
    double init_sigma, sigma_min, sigma_max, max_shift, damp, cond_max;
};

__device__ bool erf_profile(const double* prof, int L, double base, double seed, const ErfOpts& o, double& x_out) {
    x_out = __builtin_nan("");
    if (L < 5) {
        x_out = seed;
        return false;
    }
    const double pol = polarity(prof, L, base, seed, o.loc_rad);
    double ymin = prof[0], ymax = prof[0];
    for (int j = 1; j < L; ++j) {
        ymin = fmin(ymin, prof[j]);
        ymax = fmax(ymax, prof[j]);
    }
    const double amp_floor = fmax((ymax - ymin) * 1e-6, 1e-9);
    double a = ymin, b = 0.0, c = fmax(ymax - ymin, amp_floor), x0 = seed;
    double lsig = log(fmax(o.init_sigma, 1e-6));
    const double frf = (double)max(2, o.fit_rad);
    const double msf = fmax(0.25, o.max_shift);
    const double lo_ls = log(o.sigma_min), hi_ls = log(o.sigma_max);
    bool ok = true;
    double JTJ[5][5];
    for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 5; ++j) JTJ[i][j] = 0.0;
    const double s2 = 1.4142135623730951, s2pi = 2.5066282746310002, spi = 1.7724538509055159;
    for (int it = 0; it < o.gn_iters; ++it) {
        const double sig = fmin(fmax(exp(lsig), o.sigma_min), o.sigma_max);
        double JTr[5] = {0, 0, 0, 0, 0};
        for (int i = 0; i < 5; ++i)
            for (int j = 0; j < 5; ++j) JTJ[i][j] = 0.0;
        for (int k = 0; k < L; ++k) {
            const double xs = base + (double)k;
            const double xr = xs - seed;
            const double z = (xs - x0) / (s2 * sig);
            const double s = 0.5 * (1.0 + pol * erf(z));
            const double model = a + b * xr + c * s;
            double w = (fabs(xs - x0) <= frf) ? 1.0 : 0.0;
            if (L >= 7) w += 0.15 * ((fabs(xs - x0) <= frf + 1.0) ? 1.0 : 0.0);
            const double sw = sqrt(w);
            const double r = (prof[k] - model) * sw;
            const double ez = exp(-(z * z));
            const double ds_dx0 = -pol * ez / (s2pi * sig);
            const double ds_dlsig = -pol * z * ez / spi;
            const double J[5] = {sw, sw * xr, sw * s, sw * (c * ds_dx0), sw * (c * ds_dlsig)};
            for (int i = 0; i < 5; ++i) {
                JTr[i] += J[i] * r;
                for (int j = 0; j < 5; ++j) JTJ[i][j] += J[i] * J[j];
            }
        }
        for (int i = 0; i < 5; ++i) JTJ[i][i] += o.damp * fmax(JTJ[i][i], 1e-9);
        double delta[5];
        bool good = solve5(JTJ, JTr, delta);
        for (int i = 0; i < 5 && good; ++i) good = isfinite(delta[i]);
        if (!good) {
            for (int i = 0; i < 5; ++i) delta[i] = 0.0;
            ok = false;
        }
        delta[3] = fmin(fmax(delta[3], -0.75), 0.75);
        delta[4] = fmin(fmax(delta[4], -0.35), 0.35);
        a += delta[0];
        b += delta[1];
        c = fmax(c + delta[2], amp_floor);
        x0 = fmin(fmax(x0 + delta[3], seed - msf), seed + msf);
        lsig = fmin(fmax(lsig + delta[4], lo_ls), hi_ls);
    }
    const double cond = cond_sym5(JTJ);
    ok = ok && isfinite(cond) && cond < o.cond_max;
    const double sig = fmin(fmax(exp(lsig), o.sigma_min), o.sigma_max);
    double ssq = 0.0, nfit = 0.0;
    for (int k = 0; k < L; ++k) {
        const double xs = base + (double)k;
        const double z = (xs - x0) / (s2 * sig);
        const double model = a + b * (xs - seed) + c * 0.5 * (1.0 + pol * erf(z));
        const double w = (fabs(xs - x0) <= frf) ? 1.0 : 0.0;
        ssq += w * (prof[k] - model) * (prof[k] - model);
        nfit += w;
    }
    const double rms = sqrt(ssq / fmax(nfit, 1.0));
    const double dyn = fmax(ymax - ymin, 1e-9);
    ok = ok && isfinite(x0) && isfinite(sig) && isfinite(rms);
    ok = ok && fabs(x0 - seed) <= msf + 1e-6;
    ok = ok && sig >= o.sigma_min && sig <= o.sigma_max;
    ok = ok && c > amp_floor;
    ok = ok && rms <= 0.35 * dyn + 1e-6;
    x_out = ok ? x0 : __builtin_nan("");
    return ok;
}

__global__ void edge_profile_kernel(const double* prof, int K, int L, const double* base, const double* seed,
                                    ErfOpts o, double* x_out, uint8_t* ok_out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    double p[kLMax];
    for (int j = 0; j < L; ++j) p[j] = prof[(int64_t)k * L + j];
    double x;
    const bool ok = erf_profile(p, L, base[k], seed[k], o, x);
    x_out[k] = x;
    ok_out[k] = ok ? 1 : 0;
}

__device__ inline int64_t toint(double v) {
    const double r = rint(v);
    return isfinite(r) && fabs(r) < 9.0e18 ? (int64_t)r : 0;
}

// kErf: gradient-moment seeds refined by erf fits (_refine_edge_erf_core); otherwise the
// seeds are the edges (_refine_edge_moment_core, :547-604, no xR > xL / yB > yT test).
template <typename T, bool kErf>
__global__ void edge_erf_kernel(const T* img, int64_t H, int64_t W, int64_t pitch, const float* rows, int K,
                                SpxEdgeErfOptions opt, ErfOpts o, double* centers, uint8_t* ok_out) {
#pragma clang fp contract(off)
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const float* row = rows + 6 * (int64_t)k;
    const double w = (double)row[2], h = (double)row[3];
    double xc = (double)row[4], yc = (double)row[5];
    const int br = opt.band_rad, er = opt.edge_rad;
    const int L = 2 * er + 1;
    const double nband = (double)(2 * br + 1);
    bool ok_all = true;
    for (int pass = 0; pass < opt.iters; ++pass) {
        const int64_t e0[4] = {toint(xc - 0.5 * (w - 1.0)), toint(xc + 0.5 * (w - 1.0)),
                               toint(yc - 0.5 * (h - 1.0)), toint(yc + 0.5 * (h - 1.0))};
        const int64_t ybase = toint(yc), xbase = toint(xc);
        double edge[4];
        bool oks = true;
        for (int side = 0; side < 4; ++side) {
            double prof[kLMax];
            for (int j = 0; j < L; ++j) {
                double s = 0.0;
                for (int t = -br; t <= br; ++t) {
                    int64_t x, y;
                    if (side < 2) {
                        x = min(max(e0[side] + j - er, (int64_t)0), W - 1);
                        y = min(max(ybase + t, (int64_t)0), H - 1);
                    } else {
                        x = min(max(xbase + t, (int64_t)0), W - 1);
                        y = min(max(e0[side] + j - er, (int64_t)0), H - 1);
                    }
                    s += (double)img[y * pitch + x];
                }
                prof[j] = s / nband;
            }
            for (int sp = 0; sp < opt.smooth_passes; ++sp) smooth5(prof, L);
            const double base = (double)e0[side] - (double)er;
            double seed;
            const bool ok0 = gradmoment(prof, L, base, opt.loc_rad, (double)opt.grad_power, seed);
            if (kErf) {
                double e;
                const bool ok1 = erf_profile(prof, L, base, seed, o, e);
                edge[side] = e;
                oks = oks && ok0 && ok1;
            } else {
                edge[side] = seed;
                oks = oks && ok0;
            }
        }
        if (kErf) oks = oks && (edge[1] > edge[0]) && (edge[3] > edge[2]);
        ok_all = ok_all && oks;
        xc = 0.5 * (edge[0] + edge[1]);
        yc = 0.5 * (edge[2] + edge[3]);
    }
    centers[2 * (int64_t)k] = xc;
    centers[2 * (int64_t)k + 1] = yc;
    ok_out[k] = ok_all ? 1 : 0;
}

ErfOpts erf_opts(const SpxEdgeErfOptions& opt) {
    return ErfOpts{opt.erf_fit_rad, opt.erf_gn_iters, opt.loc_rad - 1 > 1 ? opt.loc_rad - 1 : 1,
                   (double)opt.erf_init_sigma, (double)opt.erf_sigma_min, (double)opt.erf_sigma_max,
                   (double)opt.erf_max_shift, (double)opt.erf_damp, (double)opt.erf_cond_max};
}

}  // namespace

template <bool kErf>
int edge_launch(const char* name, const void* img, int dtype, int64_t H, int64_t W, int64_t pitch, const float* rows,
                int K, const SpxEdgeErfOptions& opt, double* centers, uint8_t* ok, hipStream_t stream) {
    if (K == 0) return SPX_OK;
    if (!(opt.edge_rad >= 2 && opt.edge_rad <= (kLMax - 1) / 2)) {
        set_error(std::string(name) + ": edge_rad must lie in [2, 15]");
        return SPX_E_ARG;
    }
    if (!(opt.band_rad >= 0 && opt.band_rad <= kBandMax)) {
        set_error(std::string(name) + ": band_rad must lie in [0, 15]");
        return SPX_E_ARG;
    }
    if (dtype != SPX_DTYPE_U16 && dtype != SPX_DTYPE_F32) {
        set_error(std::string(name) + ": dtype must be U16 or F32");
        return SPX_E_ARG;
    }
    if (dtype == SPX_DTYPE_U16)
        edge_erf_kernel<uint16_t, kErf><<<(K + 127) / 128, 128, 0, stream>>>(
            static_cast<const uint16_t*>(img), H, W, pitch, rows, K, opt, erf_opts(opt), centers, ok);
    else
        edge_erf_kernel<float, kErf><<<(K + 127) / 128, 128, 0, stream>>>(
            static_cast<const float*>(img), H, W, pitch, rows, K, opt, erf_opts(opt), centers, ok);
    SPX_LAUNCHED();
    return SPX_OK;
}

int edge_erf_launch(const void* img, int dtype, int64_t H, int64_t W, int64_t pitch, const float* rows, int K,
                    const SpxEdgeErfOptions& opt, double* centers, uint8_t* ok, hipStream_t stream) {
    return edge_launch<true>("spx_edge_erf", img, dtype, H, W, pitch, rows, K, opt, centers, ok, stream);
}

int edge_moment_launch(const void* img, int dtype, int64_t H, int64_t W, int64_t pitch, const float* rows, int K,
                       const SpxEdgeErfOptions& opt, double* centers, uint8_t* ok, hipStream_t stream) {
    return edge_launch<false>("spx_edge_moment", img, dtype, H, W, pitch, rows, K, opt, centers, ok, stream);
}

int edge_profile_launch(const double* prof, int K, int L, const double* base, const double* seed,
                        const SpxEdgeErfOptions& opt, double* x, uint8_t* ok, hipStream_t stream) {
    if (K == 0) return SPX_OK;
    SPX_REQUIRE(L >= 2 && L <= kLMax, "spx_edge_profile_erf: profile length must lie in [2, 31]");
    ErfOpts o = erf_opts(opt);
    o.loc_rad = opt.loc_rad;  // the per-profile entry takes the polarity window as given
    edge_profile_kernel<<<(K + 127) / 128, 128, 0, stream>>>(prof, K, L, base, seed, o, x, ok);
    SPX_LAUNCHED();
    return SPX_OK;
}

}  // namespace spx
