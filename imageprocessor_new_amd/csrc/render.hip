// Synthetic training tiles for the learned start, rendered on the GPU.
//
// Each tile (side T, placed in a mosaic image of `cols` tiles per row) holds one hole at
// its centre plus the ring of nearest neighbours of its array (6 for a staggered array,
// 8 for a square one) at distance `pitch`, each with its own draws. Holes are the erf
// disks / ellipses of the fits, point-sampled at pixel centres (same model as
// imageprocessor_new_amd/synthetic.py):
//   I = B + sum_h A erfc((rho_h - r_h) / (sqrt(2) s_h)) / 2,
// with photon noise approximated by its normal limit N(I, I) (I >= 500 everywhere) plus
// N(0, read_sigma^2), rounded half-even and clipped to uint16.
// Draws come from a counter-based generator (splitmix64 of seed, tile, hole, field), so
// a tile is the same whatever else is rendered with it.
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
    double v = I + sqrt(fmax(I, 0.0)) * normal(seed ^ 0x5bd1e995ull, t, key, 0);
    if (spec.read_sigma > 0.0f) v += spec.read_sigma * normal(seed ^ 0x1b873593ull, t, key, 0);
    v = rint(v);
    v = fmin(fmax(v, 0.0), 65535.0);
    const int64_t gy = (int64_t)(t / spec.cols) * T + py, gx = (int64_t)(t % spec.cols) * T + px;
    img[gy * pitch + gx] = (uint16_t)v;
}

}  // namespace

int render_tiles_launch(uint64_t seed, int n, int T, const SpxTileSpec& spec, uint16_t* img, int64_t pitch,
                        float* truth, hipStream_t stream) {
    if (n == 0) return SPX_OK;
    void* ws = nullptr;
    int rc = workspace(sizeof(float) * (size_t)n * kMaxHoles * 6, &ws);
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
