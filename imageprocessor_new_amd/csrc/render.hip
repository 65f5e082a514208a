// Synthetic training tiles for the learned start, rendered on the GPU.
//
// Each tile (side T, placed in a mosaic image of `cols` tiles per row) holds one hole at
// its centre plus the ring of nearest neighbours of its array (6 for a staggered array,
// 8 for a square one) at distance `pitch`, each with its own draws. Holes are the erf
// disks / ellipses of the fits, point-sampled at pixel centres (same model as
// imageprocessor_new_amd/synthetic.py):
//   I = B + sum_h A erfc((rho_h - r_h) / (sqrt(2) s_h)) / 2,
// with photon noise Poisson(I) plus N(0, read_sigma^2), rounded half-even and clipped to
// uint16 (the noise model of reference subpx/synthetic_arrays.py:327-350).
// Draws come from a counter-based generator (splitmix64 of seed, tile, hole, field), so
// a tile is the same whatever else is rendered with it.
#include <hipcub/hipcub.hpp>

#include "spx_internal.hpp"

namespace spx {
namespace {

__host__ __device__ inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// uniform in (0, 1): 53 random bits
__device__ inline double uni(uint64_t seed, uint64_t a, uint64_t b, uint64_t c) {
    uint64_t h = splitmix64(seed ^ splitmix64(a ^ splitmix64(b ^ splitmix64(c))));
    return ((double)(h >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

__device__ inline double normal(uint64_t seed, uint64_t a, uint64_t b, uint64_t c) {
    const double u1 = uni(seed, a, b, 2 * c);
    const double u2 = uni(seed, a, b, 2 * c + 1);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

// Poisson(lam) from the counter-based stream (seed, a, b, 16 + k), k = 0, 1, ...
// lam < 12: inversion by sequential search of the CDF with one uniform.
// lam >= 12: Hoermann's transformed rejection with squeeze, PTRS (W. Hoermann, "The
// transformed rejection method for generating Poisson random variables", Insurance:
// Mathematics and Economics 12 (1993) 39-45); its acceptance rate is above 0.9, and the loop
// gives up after 64 candidates (never reached in practice) with the rounded mean.
__device__ inline double poisson(uint64_t seed, uint64_t a, uint64_t b, double lam) {
    if (!(lam > 0.0)) return 0.0;
    if (lam < 12.0) {
        const double u = uni(seed, a, b, 16);
        double p = exp(-lam), F = p, k = 0.0;
        while (u > F && k < 400.0) {
            k += 1.0;
            p *= lam / k;
            F += p;
        }
        return k;
    }
    const double slam = sqrt(lam), loglam = log(lam);
    const double bb = 0.931 + 2.53 * slam;
    const double aa = -0.059 + 0.02483 * bb;
    const double inv_alpha = 1.1239 + 1.1328 / (bb - 3.4);
    const double vr = 0.9277 - 3.6224 / (bb - 2.0);
    for (int it = 0; it < 64; ++it) {
        const double U = uni(seed, a, b, 16 + 2 * it) - 0.5;
        const double V = uni(seed, a, b, 17 + 2 * it);
        const double us = 0.5 - fabs(U);
        const double k = floor((2.0 * aa / us + bb) * U + lam + 0.43);
        if (us >= 0.07 && V <= vr) return k;
        if (k < 0.0 || (us < 0.013 && V > us)) continue;
        if (log(V) + log(inv_alpha) - log(aa / (us * us) + bb) <= -lam + k * loglam - lgamma(k + 1.0)) return k;
    }
    return rint(lam);
}

constexpr int kMaxHoles = 9;

// Per tile and hole: [x, y, r, s, q, A] in tile coordinates; hole 0 is the centre hole.
__global__ void tile_holes_kernel(uint64_t seed, int n, int T, SpxTileSpec spec, float* holes, float* truth) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int nh = spec.layout == 0 ? 9 : 7;
    const double c = 0.5 * (double)(T - 1);
    for (int h = 0; h < kMaxHoles; ++h) {
        double x = c, y = c;
        if (h > 0 && h < nh) {
            if (spec.layout == 0) {
                // square array: 8 neighbours
                const int dx[8] = {1, -1, 0, 0, 1, 1, -1, -1};
                const int dy[8] = {0, 0, 1, -1, 1, -1, 1, -1};
                x += dx[h - 1] * spec.pitch;
                y += dy[h - 1] * spec.pitch;
            } else {
                const double ang = 1.0471975511965976 * (h - 1);
                x += cos(ang) * spec.pitch;
                y += sin(ang) * spec.pitch;
            }
        }
        x += spec.jitter * normal(seed, t, h, 0);
        y += spec.jitter * normal(seed, t, h, 1);
        const double r = spec.r_lo + (spec.r_hi - spec.r_lo) * uni(seed, t, h, 10);
        const double s = spec.s_lo + (spec.s_hi - spec.s_lo) * uni(seed, t, h, 11);
        const bool ell = uni(seed, t, h, 12) < spec.ellipse_frac;
        const double q = ell ? 0.8 + 0.2 * uni(seed, t, h, 13) : 1.0;
        float* o = holes + ((int64_t)t * kMaxHoles + h) * 6;
        const bool present = h < nh;
        o[0] = (float)x;
        o[1] = (float)y;
        o[2] = (float)r;
        o[3] = (float)s;
        o[4] = (float)q;
        o[5] = present ? spec.amplitude : 0.0f;
        if (h == 0) {
            const int64_t ox = (int64_t)(t % spec.cols) * T, oy = (int64_t)(t / spec.cols) * T;
            float* tr = truth + (int64_t)t * 7;
            tr[0] = (float)(x + ox);
            tr[1] = (float)(y + oy);
            tr[2] = (float)r;
            tr[3] = (float)s;
            tr[4] = spec.amplitude;
            tr[5] = spec.background;
            tr[6] = (float)q;
        }
    }
}

__global__ void tile_pixels_kernel(uint64_t seed, int n, int T, SpxTileSpec spec, const float* holes,
                                   uint16_t* img, int64_t pitch) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t per = (int64_t)T * T;
    if (i >= (int64_t)n * per) return;
    const int t = (int)(i / per);
    const int p = (int)(i - (int64_t)t * per);
    const int py = p / T, px = p - (p / T) * T;
    double I = spec.background;
    const float* hp = holes + (int64_t)t * kMaxHoles * 6;
    for (int h = 0; h < kMaxHoles; ++h) {
        const float* o = hp + h * 6;
        if (o[5] == 0.0f) continue;
        const double dx = px - o[0], dy = (py - o[1]) / o[4];
        const double rho = sqrt(dx * dx + dy * dy);
        I += (double)o[5] * 0.5 * erfc((rho - o[2]) / (1.4142135623730951 * o[3]));
    }
    const uint64_t key = (uint64_t)p;
    double v = poisson(seed ^ 0x5bd1e995ull, t, key, I);
    if (spec.read_sigma > 0.0f) v += spec.read_sigma * normal(seed ^ 0x1b873593ull, t, key, 0);
    v = rint(v);
    v = fmin(fmax(v, 0.0), 65535.0);
    const int64_t gy = (int64_t)(t / spec.cols) * T + py, gx = (int64_t)(t % spec.cols) * T + px;
    img[gy * pitch + gx] = (uint16_t)v;
}

// Whole pages: a hole list (x, y, r, s, q, A per hole, image coordinates) binned into a grid
// of kPageCell-px cells by a counting sort; each pixel adds the holes of its 3 x 3 cells
// whose support (r + 8 s, where erfc/2 < 1e-15) reaches it, then the tile renderer's noise
// (Poisson(I) + N(0, read_sigma^2)).
constexpr int kPageCell = 64;

__global__ void page_count_kernel(const float* holes, int n, int gw, int gh, int32_t* cnt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int cx = min(max((int)floorf(holes[6 * i] / kPageCell), 0), gw - 1);
    const int cy = min(max((int)floorf(holes[6 * i + 1] / kPageCell), 0), gh - 1);
    atomicAdd(&cnt[cy * gw + cx], 1);
}

__global__ void page_scatter_kernel(const float* holes, int n, int gw, int gh, int32_t* cursor, int32_t* ids) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int cx = min(max((int)floorf(holes[6 * i] / kPageCell), 0), gw - 1);
    const int cy = min(max((int)floorf(holes[6 * i + 1] / kPageCell), 0), gh - 1);
    ids[atomicAdd(&cursor[cy * gw + cx], 1)] = i;
}

__global__ void page_pixels_kernel(uint64_t seed, const float* holes, const int32_t* start, const int32_t* cnt,
                                   const int32_t* ids, int gw, int gh, int64_t H, int64_t W, double background,
                                   double read_sigma, uint16_t* img, int64_t pitch) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= H * W) return;
    const int64_t py = i / W, px = i - (i / W) * W;
    const int cx = (int)(px / kPageCell), cy = (int)(py / kPageCell);
    double I = background;
    for (int dy = -1; dy <= 1; ++dy) {
        const int yy = cy + dy;
        if (yy < 0 || yy >= gh) continue;
        for (int dx = -1; dx <= 1; ++dx) {
            const int xx = cx + dx;
            if (xx < 0 || xx >= gw) continue;
            const int c = yy * gw + xx;
            const int s0 = start[c], s1 = s0 + cnt[c];
            for (int k = s0; k < s1; ++k) {
                const float* o = holes + 6 * (int64_t)ids[k];
                const double ex = (double)px - o[0], ey = ((double)py - o[1]) / o[4];
                const double rho = sqrt(ex * ex + ey * ey);
                if (rho > (double)o[2] + 8.0 * (double)o[3]) continue;
                I += (double)o[5] * 0.5 * erfc((rho - o[2]) / (1.4142135623730951 * o[3]));
            }
        }
    }
    double v = poisson(seed ^ 0x5bd1e995ull, (uint64_t)py, (uint64_t)px, I);
    if (read_sigma > 0.0) v += read_sigma * normal(seed ^ 0x1b873593ull, (uint64_t)py, (uint64_t)px, 0);
    v = rint(v);
    v = fmin(fmax(v, 0.0), 65535.0);
    img[py * pitch + px] = (uint16_t)v;
}

}  // namespace

int render_page_launch(uint64_t seed, const float* holes, int n, int64_t H, int64_t W, float background,
                       float read_sigma, uint16_t* img, int64_t pitch, hipStream_t stream) {
    const int gw = (int)((W + kPageCell - 1) / kPageCell), gh = (int)((H + kPageCell - 1) / kPageCell);
    const int64_t ncell = (int64_t)gw * gh;
    size_t scan_bytes = 0;
    SPX_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (int32_t*)nullptr, (int32_t*)nullptr, (int)ncell,
                                             stream));
    const size_t head = sizeof(int32_t) * (3 * (size_t)ncell + (size_t)(n > 0 ? n : 1));
    const size_t head_al = (head + 255) & ~(size_t)255;
    void* ws = nullptr;
    int rc = workspace(head_al + scan_bytes, &ws, stream);
    if (rc != SPX_OK) return rc;
    int32_t* cnt = static_cast<int32_t*>(ws);
    int32_t* start = cnt + ncell;
    int32_t* cursor = start + ncell;
    int32_t* ids = cursor + ncell;
    SPX_HIP(hipMemsetAsync(cnt, 0, sizeof(int32_t) * ncell, stream));
    if (n > 0) {
        page_count_kernel<<<(n + 255) / 256, 256, 0, stream>>>(holes, n, gw, gh, cnt);
        SPX_LAUNCHED();
    }
    SPX_HIP(hipcub::DeviceScan::ExclusiveSum(static_cast<char*>(ws) + head_al, scan_bytes, cnt, start, (int)ncell,
                                             stream));
    SPX_HIP(hipMemcpyAsync(cursor, start, sizeof(int32_t) * ncell, hipMemcpyDeviceToDevice, stream));
    if (n > 0) {
        page_scatter_kernel<<<(n + 255) / 256, 256, 0, stream>>>(holes, n, gw, gh, cursor, ids);
        SPX_LAUNCHED();
    }
    const int64_t total = H * W;
    page_pixels_kernel<<<(unsigned)((total + 255) / 256), 256, 0, stream>>>(
        seed, holes, start, cnt, ids, gw, gh, H, W, (double)background, (double)read_sigma, img, pitch);
    SPX_LAUNCHED();
    return SPX_OK;
}

int render_tiles_launch(uint64_t seed, int n, int T, const SpxTileSpec& spec, uint16_t* img, int64_t pitch,
                        float* truth, hipStream_t stream) {
    if (n == 0) return SPX_OK;
    void* ws = nullptr;
    int rc = workspace(sizeof(float) * (size_t)n * kMaxHoles * 6, &ws, stream);
    if (rc != SPX_OK) return rc;
    float* holes = static_cast<float*>(ws);
    tile_holes_kernel<<<(n + 255) / 256, 256, 0, stream>>>(seed, n, T, spec, holes, truth);
    SPX_LAUNCHED();
    const int64_t total = (int64_t)n * T * T;
    tile_pixels_kernel<<<(unsigned)((total + 255) / 256), 256, 0, stream>>>(seed, n, T, spec, holes, img, pitch);
    SPX_LAUNCHED();
    return SPX_OK;
}

}  // namespace spx
