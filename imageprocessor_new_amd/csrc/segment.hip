// Integer front end on the GPU: 65,536-bin histogram + Otsu threshold, threshold +
// connected components with stats in OpenCV's layout.
//
// Reference calls replaced: cv2.threshold(g, 0, 255, BINARY|OTSU) (reference
// subpx/_cpu/centers.py:258, global variant subpx/_gpu/centers.py:1789-1796) and
// cv2.connectedComponentsWithStats (reference subpx/_cpu/components.py:18-34, GPU twin
// subpx/_gpu/components.py:64-134 which uses cupyx label + an atomic stats kernel).
//
// Layout: HBM-bound byte work. The histogram is privatised per workgroup in LDS as packed
// 16-bit counters (65,536 bins = 128 KiB) over at most 65,535 pixels per workgroup, then
// flushed with one global atomic per non-empty bin. Components use block-based union-find:
// for 8-connectivity every 2x2 pixel block is one node (its foreground pixels are
// mutually adjacent), which quarters the nodes and makes the smallest root the component's
// first block in block-raster order, i.e. OpenCV's numbering; 4-connectivity uses pixels.
// Labels are numbered by an exclusive scan over root flags.
#include <hipcub/hipcub.hpp>

#include "spx_internal.hpp"

namespace spx {
namespace {

constexpr int kHistBins = 65536;
constexpr int kHistThreads = 1024;
constexpr int kHistPixelsPerThread = 63;  // 1024 * 63 < 65536: a 16-bit counter cannot wrap

__global__ void fill_u32(uint32_t* p, int64_t n, uint32_t v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

// hist[v] += count of img[y*step][x*step] == v over the strided view (Hs x Ws).
__global__ __launch_bounds__(kHistThreads) void hist_u16_kernel(const uint16_t* img, int64_t Hs,
                                                                int64_t Ws, int64_t pitch,
                                                                int step, uint32_t* hist) {
    __shared__ uint32_t lds[kHistBins / 2];
    for (int i = threadIdx.x; i < kHistBins / 2; i += blockDim.x) lds[i] = 0;
    __syncthreads();
    const int64_t n = Hs * Ws;
    const int64_t per_block = (int64_t)kHistThreads * kHistPixelsPerThread;
    const int64_t base = (int64_t)blockIdx.x * per_block;
    for (int k = 0; k < kHistPixelsPerThread; ++k) {
        const int64_t idx = base + (int64_t)k * kHistThreads + threadIdx.x;
        if (idx < n) {
            const int64_t y = idx / Ws;
            const int64_t x = idx - y * Ws;
            const uint32_t v = img[y * step * pitch + x * step];
            atomicAdd(&lds[v >> 1], 1u << ((v & 1u) * 16u));
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kHistBins / 2; i += blockDim.x) {
        const uint32_t w = lds[i];
        if (w & 0xFFFFu) atomicAdd(&hist[2 * i], w & 0xFFFFu);
        if (w >> 16) atomicAdd(&hist[2 * i + 1], w >> 16);
    }
}

// Otsu over the histogram, one workgroup of 1024 threads, 64 bins each.
// With exact integer class counts cnt(i) = sum_{j<=i} h_j and moments, OpenCV's
// recurrence (q1 += p_i; mu1 = (mu1 q1_prev + i p_i) / q1; skip while min(q1, q2) <
// FLT_EPSILON) evaluates sigma(i) = q1 q2 (mu1 - mu2)^2 with mu1 = m1 / cnt and
// mu2 = (S - m1) / (N - cnt), where m1 sums j h_j from the first bin that is not skipped
// (bins skipped at the start never enter mu1). The first maximum wins.
__global__ __launch_bounds__(1024) void otsu_kernel(const uint32_t* hist, int32_t* out) {
    __shared__ unsigned long long s_cnt[1024];
    __shared__ unsigned long long s_mom[1024];
    __shared__ double s_best[1024];
    __shared__ int s_arg[1024];
    __shared__ int s_first;
    const int t = threadIdx.x;
    const int b0 = t * 64;
    unsigned long long c = 0, m = 0;
    for (int i = 0; i < 64; ++i) {
        const unsigned long long h = hist[b0 + i];
        c += h;
        m += h * (unsigned long long)(b0 + i);
    }
    s_cnt[t] = c;
    s_mom[t] = m;
    if (t == 0) s_first = kHistBins;
    __syncthreads();
    // inclusive scan over the 1024 chunk totals (Hillis-Steele)
    for (int off = 1; off < 1024; off <<= 1) {
        unsigned long long a = 0, bm = 0;
        if (t >= off) {
            a = s_cnt[t - off];
            bm = s_mom[t - off];
        }
        __syncthreads();
        s_cnt[t] += a;
        s_mom[t] += bm;
        __syncthreads();
    }
    const unsigned long long N = s_cnt[1023];
    const unsigned long long S = s_mom[1023];
    const unsigned long long cnt_before = t ? s_cnt[t - 1] : 0ull;
    const double scale = 1.0 / (double)N;
    const double eps = (double)__FLT_EPSILON__;
    // first bin that is not skipped
    {
        unsigned long long cc = cnt_before;
        for (int i = 0; i < 64; ++i) {
            cc += hist[b0 + i];
            const double q1 = (double)cc * scale;
            const double q2 = 1.0 - q1;
            if (!(fmin(q1, q2) < eps || fmax(q1, q2) > 1.0 - eps)) {
                atomicMin(&s_first, b0 + i);
                break;
            }
        }
    }
    __syncthreads();
    const int i0 = s_first;
    // moment of the skipped prefix (bins < i0) is excluded from m1
    unsigned long long mom_skip = 0;
    if (i0 < kHistBins) {
        // all threads compute the same prefix moment below i0 from the chunk scan
        const int tc = i0 / 64;
        mom_skip = tc ? s_mom[tc - 1] : 0ull;
        for (int i = tc * 64; i < i0; ++i) mom_skip += (unsigned long long)hist[i] * (unsigned long long)i;
    }
    double best = 0.0;
    int arg = 0x7fffffff;
    {
        unsigned long long cc = cnt_before;
        unsigned long long mm = t ? s_mom[t - 1] : 0ull;
        for (int i = 0; i < 64; ++i) {
            const int bin = b0 + i;
            const unsigned long long h = hist[bin];
            cc += h;
            mm += h * (unsigned long long)bin;
            if (bin < i0) continue;
            const double q1 = (double)cc * scale;
            const double q2 = 1.0 - q1;
            if (fmin(q1, q2) < eps || fmax(q1, q2) > 1.0 - eps) continue;
            const double m1 = (double)(mm - mom_skip);
            const double mu1 = m1 / (double)cc;
            const double mu2 = ((double)S - m1) / (double)(N - cc);
            const double d = mu1 - mu2;
            const double sigma = q1 * q2 * d * d;
            if (sigma > best) {
                best = sigma;
                arg = bin;
            }
        }
    }
    s_best[t] = best;
    s_arg[t] = arg;
    __syncthreads();
    for (int off = 512; off > 0; off >>= 1) {
        if (t < off) {
            const double ob = s_best[t + off];
            const int oa = s_arg[t + off];
            if (ob > s_best[t] || (ob == s_best[t] && oa < s_arg[t])) {
                s_best[t] = ob;
                s_arg[t] = oa;
            }
        }
        __syncthreads();
    }
    if (t == 0) out[0] = (s_best[0] > 0.0 && s_arg[0] != 0x7fffffff) ? s_arg[0] : 0;
}

// ---------------------------------------------------------------------------------------
// Connected components. BS = 2: 2x2 pixel blocks (8-connectivity); BS = 1: pixels (4).

struct CCGeom {
    int64_t H, W, pitch;
    int64_t NH, NW;  // node grid
};

__device__ __forceinline__ bool is_fg(const uint16_t* img, const CCGeom& g, int64_t y, int64_t x,
                                      int32_t thresh, int invert) {
    if (y < 0 || y >= g.H || x < 0 || x >= g.W) return false;
    const int32_t v = img[y * g.pitch + x];
    return invert ? (v <= thresh) : (v > thresh);
}

// Foreground bits of a node: bit (2*dy + dx) for BS = 2, bit 0 for BS = 1.
template <int BS>
__device__ __forceinline__ uint32_t node_bits(const uint16_t* img, const CCGeom& g, int64_t ny,
                                              int64_t nx, int32_t thresh, int invert) {
    if (BS == 1) return is_fg(img, g, ny, nx, thresh, invert) ? 1u : 0u;
    const int64_t y = 2 * ny, x = 2 * nx;
    return (is_fg(img, g, y, x, thresh, invert) ? 1u : 0u) |
           (is_fg(img, g, y, x + 1, thresh, invert) ? 2u : 0u) |
           (is_fg(img, g, y + 1, x, thresh, invert) ? 4u : 0u) |
           (is_fg(img, g, y + 1, x + 1, thresh, invert) ? 8u : 0u);
}

__device__ __forceinline__ int32_t uf_find(const int32_t* parent, int32_t a) {
    int32_t p = parent[a];
    while (p != a) {
        a = p;
        p = parent[a];
    }
    return a;
}

// Lock-free union: link the larger root under the smaller one with atomicMin; if the
// larger root was re-linked meanwhile, continue from what it points to.
__device__ __forceinline__ void uf_union(int32_t* parent, int32_t a, int32_t b) {
    for (;;) {
        a = uf_find(parent, a);
        b = uf_find(parent, b);
        if (a == b) return;
        if (a > b) {
            const int32_t t = a;
            a = b;
            b = t;
        }
        const int32_t old = atomicMin(&parent[b], a);
        if (old == b) return;
        b = old;
    }
}

// Union-find inside one 32 x 32-node tile in LDS (same lock-free rule as the global one);
// 256 threads, each holding the nodes of rows ly, ly + 8, ly + 16 and ly + 24 of its column, so
// that a wave reads 128 contiguous bytes of every pixel row.
constexpr int kUfTile = 32;
constexpr int kUfThreads = 256;
constexpr int kUfPer = kUfTile * kUfTile / kUfThreads;

__device__ __forceinline__ int uf_find_lds(const int32_t* lp, int a) {
    int p = lp[a];
    while (p != a) {
        a = p;
        p = lp[a];
    }
    return a;
}
__device__ __forceinline__ void uf_union_lds(int32_t* lp, int a, int b) {
    for (;;) {
        a = uf_find_lds(lp, a);
        b = uf_find_lds(lp, b);
        if (a == b) return;
        if (a > b) {
            const int t = a;
            a = b;
            b = t;
        }
        const int old = atomicMin(&lp[b], a);
        if (old == b) return;
        b = old;
    }
}

// Which of a node's four backward neighbours (left, up, up-left, up-right) are adjacent to it:
// bit 0 left, bit 1 up, bit 2 up-left, bit 3 up-right, from the nodes' foreground bits.
template <int BS>
__device__ __forceinline__ uint32_t cc_links(uint32_t b, uint32_t l, uint32_t u, uint32_t ul, uint32_t ur) {
    if (BS == 1) return (l ? 1u : 0u) | (u ? 2u : 0u);
    uint32_t m = 0;
    // left block: its right column (bits 1, 3) touches our left column (bits 0, 2)
    if ((b & 5u) && (l & 10u)) m |= 1u;
    // upper block: its bottom row (bits 4, 8) touches our top row (bits 1, 2)
    if ((b & 3u) && (u & 12u)) m |= 2u;
    // upper-left: its bottom-right pixel (bit 8) touches our top-left (bit 1)
    if ((b & 1u) && (ul & 8u)) m |= 4u;
    // upper-right: its bottom-left pixel (bit 4) touches our top-right (bit 2)
    if ((b & 2u) && (ur & 4u)) m |= 8u;
    return m;
}

// Foreground bits of every node, then the union-find of the edges inside each 32 x 32-node
// tile in LDS: parent[n] = the node index of n's tile-local root (the smallest node index of
// its component within the tile; tile-local order is raster order), -1 for background. Only
// edges between tiles are left for cc_merge_kernel, so the global atomics run on tile borders.
template <int BS>
__global__ __launch_bounds__(kUfThreads) void cc_local_kernel(const uint16_t* img, CCGeom g, int32_t thresh,
                                                              int invert, int32_t* parent, uint8_t* bits,
                                                              int64_t tiles_x) {
    __shared__ int32_t lp[kUfTile * kUfTile];
    __shared__ uint8_t lb[kUfTile * kUfTile];
    const int64_t ty = blockIdx.x / tiles_x, tx = blockIdx.x - (blockIdx.x / tiles_x) * tiles_x;
    uint32_t bk[kUfPer];
#pragma unroll
    for (int k = 0; k < kUfPer; ++k) {
        const int i = threadIdx.x + k * kUfThreads, ly = i / kUfTile, lx = i % kUfTile;
        const int64_t ny = ty * kUfTile + ly, nx = tx * kUfTile + lx;
        const bool in = ny < g.NH && nx < g.NW;
        const uint32_t b = in ? node_bits<BS>(img, g, ny, nx, thresh, invert) : 0u;
        if (in) bits[ny * g.NW + nx] = (uint8_t)b;
        bk[k] = b;
        lb[i] = (uint8_t)b;
        lp[i] = b ? i : -1;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kUfPer; ++k) {
        const uint32_t b = bk[k];
        if (!b) continue;
        const int i = threadIdx.x + k * kUfThreads, ly = i / kUfTile, lx = i % kUfTile;
        const uint32_t l = lx > 0 ? lb[i - 1] : 0u;
        const uint32_t u = ly > 0 ? lb[i - kUfTile] : 0u;
        const uint32_t ul = (lx > 0 && ly > 0) ? lb[i - kUfTile - 1] : 0u;
        const uint32_t ur = (lx + 1 < kUfTile && ly > 0) ? lb[i - kUfTile + 1] : 0u;
        const uint32_t m = cc_links<BS>(b, l, u, ul, ur);
        if (m & 1u) uf_union_lds(lp, i, i - 1);
        if (m & 2u) uf_union_lds(lp, i, i - kUfTile);
        if (m & 4u) uf_union_lds(lp, i, i - kUfTile - 1);
        if (m & 8u) uf_union_lds(lp, i, i - kUfTile + 1);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kUfPer; ++k) {
        const int i = threadIdx.x + k * kUfThreads, ly = i / kUfTile, lx = i % kUfTile;
        const int64_t ny = ty * kUfTile + ly, nx = tx * kUfTile + lx;
        if (ny >= g.NH || nx >= g.NW) continue;
        int32_t out = -1;
        if (bk[k]) {
            const int r = uf_find_lds(lp, i);
            out = (int32_t)((ty * kUfTile + r / kUfTile) * g.NW + tx * kUfTile + r % kUfTile);
        }
        parent[ny * g.NW + nx] = out;
    }
}

// The edges between tiles, by the global lock-free union-find: one block of kUfMergeThreads
// threads per tile, thread t taking one of the tile's 3 kUfTile - 2 border nodes (left column,
// top row without its first node, right column without its first node); interior nodes have
// no such edge.
constexpr int kUfMergeThreads = (3 * kUfTile - 2 + 63) / 64 * 64;
template <int BS>
__global__ __launch_bounds__(kUfMergeThreads) void cc_merge_kernel(CCGeom g, int32_t* parent, const uint8_t* bits,
                                                      int64_t tiles_x) {
    const int t = threadIdx.x;
    if (t >= 3 * kUfTile - 2) return;
    int lx, ly;
    if (t < kUfTile) {
        ly = t;
        lx = 0;
    } else if (t < 2 * kUfTile - 1) {
        ly = 0;
        lx = t - kUfTile + 1;
    } else {
        ly = t - 2 * kUfTile + 2;
        lx = kUfTile - 1;
    }
    const int64_t ty = blockIdx.x / tiles_x, tx = blockIdx.x - (blockIdx.x / tiles_x) * tiles_x;
    const int64_t ny = ty * kUfTile + ly, nx = tx * kUfTile + lx;
    if (ny >= g.NH || nx >= g.NW) return;
    const int64_t n = ny * g.NW + nx;
    const uint32_t b = bits[n];
    if (!b) return;
    const int64_t u = n - g.NW;
    const uint32_t l = nx > 0 ? bits[n - 1] : 0u;
    const uint32_t up = ny > 0 ? bits[u] : 0u;
    const uint32_t ul = (nx > 0 && ny > 0) ? bits[u - 1] : 0u;
    const uint32_t ur = (nx + 1 < g.NW && ny > 0) ? bits[u + 1] : 0u;
    uint32_t m = cc_links<BS>(b, l, up, ul, ur);
    // keep the edges that leave the tile
    if (lx != 0) m &= ~1u;
    if (ly != 0) m &= ~2u;
    if (lx != 0 && ly != 0) m &= ~4u;
    if (lx != kUfTile - 1 && ly != 0) m &= ~8u;
    if (m & 1u) uf_union(parent, (int32_t)n, (int32_t)(n - 1));
    if (m & 2u) uf_union(parent, (int32_t)n, (int32_t)u);
    if (m & 4u) uf_union(parent, (int32_t)n, (int32_t)(u - 1));
    if (m & 8u) uf_union(parent, (int32_t)n, (int32_t)(u + 1));
}

__global__ void cc_compress_kernel(int64_t nn, int32_t* parent, int32_t* root_flag) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nn) return;
    const int32_t p = parent[n];
    if (p < 0) {
        root_flag[n] = 0;
        return;
    }
    const int32_t r = uf_find(parent, (int32_t)n);
    parent[n] = r;
    root_flag[n] = (r == (int32_t)n) ? 1 : 0;
}

struct CCAcc {
    int32_t xmin, ymin, xmax, ymax;
    uint32_t area, pad;
    unsigned long long sx, sy;
};

__global__ void cc_acc_init_kernel(CCAcc* acc, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    acc[i] = CCAcc{0x7fffffff, 0x7fffffff, -1, -1, 0u, 0u, 0ull, 0ull};
}

// Wave reductions for the per-label aggregation (every lane takes part; lanes outside the
// current label contribute the identity). Within each row of 16 lanes by four DPP steps
// (quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror, row_mirror: every lane ends with
// its row's total), then the four row totals read into scalars: no LDS traffic. The result is
// uniform.
template <int CTRL>
__device__ __forceinline__ int32_t dpp_i32(int32_t v) {
    return __builtin_amdgcn_update_dpp(v, v, CTRL, 0xF, 0xF, false);
}
__device__ __forceinline__ int32_t wave_min_i32(int32_t v) {
    v = min(v, dpp_i32<0xB1>(v));
    v = min(v, dpp_i32<0x4E>(v));
    v = min(v, dpp_i32<0x141>(v));
    v = min(v, dpp_i32<0x140>(v));
    return min(min(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
               min(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}
__device__ __forceinline__ int32_t wave_max_i32(int32_t v) {
    v = max(v, dpp_i32<0xB1>(v));
    v = max(v, dpp_i32<0x4E>(v));
    v = max(v, dpp_i32<0x141>(v));
    v = max(v, dpp_i32<0x140>(v));
    return max(max(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
               max(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}
// Sum of per-lane values whose wave total fits 32 bits (pixel counts and coordinate sums of
// one wave's 8 x 8 nodes: below 2^32 for images narrower and shorter than 2^24 px).
__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
    int32_t w = (int32_t)v;
    w += dpp_i32<0xB1>(w);
    w += dpp_i32<0x4E>(w);
    w += dpp_i32<0x141>(w);
    w += dpp_i32<0x140>(w);
    return (uint32_t)__builtin_amdgcn_readlane(w, 0) + (uint32_t)__builtin_amdgcn_readlane(w, 16) +
           (uint32_t)__builtin_amdgcn_readlane(w, 32) + (uint32_t)__builtin_amdgcn_readlane(w, 48);
}

// Background pixels (label 0) are reduced per block into bg_part[block] (no atomics);
// cc_bg_reduce_kernel folds those into row 0.
constexpr int kLabTile = 16;  // nodes per block side

template <int BS>
__global__ __launch_bounds__(256) void cc_label_kernel(CCGeom g, const int32_t* parent,
                                                       const uint8_t* bits, const int32_t* ids,
                                                       int32_t* labels, CCAcc* acc, int32_t cap,
                                                       CCAcc* bg_part, int64_t tiles_x) {
    __shared__ CCAcc s_bg[4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t bx = (int64_t)blockIdx.x % tiles_x, by = (int64_t)blockIdx.x / tiles_x;
    const int64_t nx = bx * kLabTile + (wv & 1) * 8 + (lane & 7);
    const int64_t ny = by * kLabTile + (wv >> 1) * 8 + (lane >> 3);
    int32_t bxmin = 0x7fffffff, bymin = 0x7fffffff, bxmax = -1, bymax = -1;
    uint32_t bcnt = 0, bsx = 0, bsy = 0;
    int32_t lab = 0;
    int32_t xmin = 0x7fffffff, ymin = 0x7fffffff, xmax = -1, ymax = -1;
    uint32_t cnt = 0, sx = 0, sy = 0;
    if (nx < g.NW && ny < g.NH) {
        const int64_t n = ny * g.NW + nx;
        const uint32_t b = bits[n];
        if (b) lab = ids[parent[n]] + 1;
        for (int dy = 0; dy < BS; ++dy) {
            for (int dx = 0; dx < BS; ++dx) {
                const int64_t y = ny * BS + dy, x = nx * BS + dx;
                if (y >= g.H || x >= g.W) continue;
                const bool fg = (b >> (BS == 2 ? (2 * dy + dx) : 0)) & 1u;
                if (labels != nullptr) labels[y * g.W + x] = fg ? lab : 0;
                if (fg) {
                    xmin = min(xmin, (int32_t)x);
                    ymin = min(ymin, (int32_t)y);
                    xmax = max(xmax, (int32_t)x);
                    ymax = max(ymax, (int32_t)y);
                    ++cnt;
                    sx += (uint32_t)x;
                    sy += (uint32_t)y;
                } else {
                    bxmin = min(bxmin, (int32_t)x);
                    bymin = min(bymin, (int32_t)y);
                    bxmax = max(bxmax, (int32_t)x);
                    bymax = max(bymax, (int32_t)y);
                    ++bcnt;
                    bsx += (uint32_t)x;
                    bsy += (uint32_t)y;
                }
            }
        }
    }
    // foreground: one aggregation round per distinct label of the wave
    bool pending = cnt > 0 && lab < cap;
    for (;;) {
        const unsigned long long m = __ballot(pending);
        if (m == 0ull) break;
        const int leader = __ffsll((long long)m) - 1;
        const int32_t L = __builtin_amdgcn_readlane(lab, leader);
        const bool mine = pending && lab == L;
        const int32_t rxmin = wave_min_i32(mine ? xmin : 0x7fffffff);
        const int32_t rymin = wave_min_i32(mine ? ymin : 0x7fffffff);
        const int32_t rxmax = wave_max_i32(mine ? xmax : -1);
        const int32_t rymax = wave_max_i32(mine ? ymax : -1);
        const uint32_t rcnt = wave_sum_u32(mine ? cnt : 0u);
        const uint32_t rsx = wave_sum_u32(mine ? sx : 0u);
        const uint32_t rsy = wave_sum_u32(mine ? sy : 0u);
        if (lane == leader) {
            CCAcc* a = acc + L;
            atomicMin(&a->xmin, rxmin);
            atomicMin(&a->ymin, rymin);
            atomicMax(&a->xmax, rxmax);
            atomicMax(&a->ymax, rymax);
            atomicAdd(&a->area, rcnt);
            atomicAdd(&a->sx, (unsigned long long)rsx);
            atomicAdd(&a->sy, (unsigned long long)rsy);
        }
        pending = pending && !mine;
    }
    // background: per-wave reduction, then the block's four waves into bg_part[block]
    const int32_t wxmin = wave_min_i32(bxmin), wymin = wave_min_i32(bymin);
    const int32_t wxmax = wave_max_i32(bxmax), wymax = wave_max_i32(bymax);
    const uint32_t wcnt = wave_sum_u32(bcnt), wsx = wave_sum_u32(bsx), wsy = wave_sum_u32(bsy);
    if (lane == 0) s_bg[wv] = CCAcc{wxmin, wymin, wxmax, wymax, wcnt, 0u, wsx, wsy};
    __syncthreads();
    if (threadIdx.x == 0) {
        CCAcc r = s_bg[0];
        for (int k = 1; k < 4; ++k) {
            r.xmin = min(r.xmin, s_bg[k].xmin);
            r.ymin = min(r.ymin, s_bg[k].ymin);
            r.xmax = max(r.xmax, s_bg[k].xmax);
            r.ymax = max(r.ymax, s_bg[k].ymax);
            r.area += s_bg[k].area;
            r.sx += s_bg[k].sx;
            r.sy += s_bg[k].sy;
        }
        bg_part[blockIdx.x] = r;
    }
}

// Row 0 (background) from the per-block partials: one block, grid-stride over the parts.
__global__ __launch_bounds__(1024) void cc_bg_reduce_kernel(const CCAcc* part, int64_t n, CCAcc* acc0) {
    __shared__ int32_t s_i[4][1024];
    __shared__ unsigned long long s_u[3][1024];
    const int t = threadIdx.x;
    int32_t xmin = 0x7fffffff, ymin = 0x7fffffff, xmax = -1, ymax = -1;
    unsigned long long area = 0, sx = 0, sy = 0;
    for (int64_t i = t; i < n; i += 1024) {
        const CCAcc p = part[i];
        xmin = min(xmin, p.xmin);
        ymin = min(ymin, p.ymin);
        xmax = max(xmax, p.xmax);
        ymax = max(ymax, p.ymax);
        area += p.area;
        sx += p.sx;
        sy += p.sy;
    }
    s_i[0][t] = xmin;
    s_i[1][t] = ymin;
    s_i[2][t] = xmax;
    s_i[3][t] = ymax;
    s_u[0][t] = area;
    s_u[1][t] = sx;
    s_u[2][t] = sy;
    __syncthreads();
    for (int off = 512; off > 0; off >>= 1) {
        if (t < off) {
            s_i[0][t] = min(s_i[0][t], s_i[0][t + off]);
            s_i[1][t] = min(s_i[1][t], s_i[1][t + off]);
            s_i[2][t] = max(s_i[2][t], s_i[2][t + off]);
            s_i[3][t] = max(s_i[3][t], s_i[3][t + off]);
            s_u[0][t] += s_u[0][t + off];
            s_u[1][t] += s_u[1][t + off];
            s_u[2][t] += s_u[2][t + off];
        }
        __syncthreads();
    }
    if (t == 0) *acc0 = CCAcc{s_i[0][0], s_i[1][0], s_i[2][0], s_i[3][0], (uint32_t)s_u[0][0], 0u, s_u[1][0], s_u[2][0]};
}

__global__ void cc_finalize_kernel(const CCAcc* acc, int32_t rows, int32_t* stats, double* cents) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows) return;
    const CCAcc a = acc[i];
    if (a.area == 0) {
        stats[5 * i + 0] = 0;
        stats[5 * i + 1] = 0;
        stats[5 * i + 2] = 0;
        stats[5 * i + 3] = 0;
        stats[5 * i + 4] = 0;
        cents[2 * i] = __builtin_nan("");
        cents[2 * i + 1] = __builtin_nan("");
        return;
    }
    stats[5 * i + 0] = a.xmin;
    stats[5 * i + 1] = a.ymin;
    stats[5 * i + 2] = a.xmax - a.xmin + 1;
    stats[5 * i + 3] = a.ymax - a.ymin + 1;
    stats[5 * i + 4] = (int32_t)a.area;
    cents[2 * i] = (double)a.sx / (double)a.area;
    cents[2 * i + 1] = (double)a.sy / (double)a.area;
}

inline int grid_for(int64_t n, int threads) { return (int)((n + threads - 1) / threads); }

template <int BS>
int threshold_cc_impl(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, int32_t thresh,
                      int32_t invert, int32_t* labels, int32_t* stats, double* cents, int32_t cap,
                      int32_t* n_out, hipStream_t stream) {
    CCGeom g{H, W, pitch, (H + BS - 1) / BS, (W + BS - 1) / BS};
    const int64_t nn = g.NH * g.NW;
    SPX_REQUIRE(nn < (int64_t)0x7fffffff, "spx_threshold_cc: image too large for int32 node ids");
    // workspace: parent i32[nn] | flag i32[nn] | ids i32[nn] | acc[cap] | bits u8[nn] | scan temp
    size_t scan_bytes = 0;
    SPX_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (int32_t*)nullptr, (int32_t*)nullptr,
                                             (int)nn, stream));
    auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t o_parent = 0;
    const size_t o_flag = o_parent + align(4 * nn);
    const size_t o_ids = o_flag + align(4 * nn);
    const size_t o_acc = o_ids + align(4 * nn);
    const size_t n_acc = (size_t)(cap > 0 ? cap : 1);
    const size_t o_bits = o_acc + align(sizeof(CCAcc) * n_acc);
    const int64_t tiles_x = (g.NW + kLabTile - 1) / kLabTile;
    const int64_t tiles = tiles_x * ((g.NH + kLabTile - 1) / kLabTile);
    const size_t o_bg = o_bits + align(nn);
    const size_t o_scan = o_bg + align(sizeof(CCAcc) * tiles);
    const size_t total = o_scan + align(scan_bytes) + 256;
    void* ws = nullptr;
    int rc = workspace(total, &ws, stream);
    if (rc != SPX_OK) return rc;
    char* base = static_cast<char*>(ws);
    int32_t* parent = reinterpret_cast<int32_t*>(base + o_parent);
    int32_t* flag = reinterpret_cast<int32_t*>(base + o_flag);
    int32_t* ids = reinterpret_cast<int32_t*>(base + o_ids);
    CCAcc* acc = reinterpret_cast<CCAcc*>(base + o_acc);
    uint8_t* bits = reinterpret_cast<uint8_t*>(base + o_bits);
    CCAcc* bg_part = reinterpret_cast<CCAcc*>(base + o_bg);
    void* scan_tmp = base + o_scan;

    const int T = 256;
    const int64_t uf_tiles_x = (g.NW + kUfTile - 1) / kUfTile;
    const int64_t uf_tiles = uf_tiles_x * ((g.NH + kUfTile - 1) / kUfTile);
    cc_local_kernel<BS><<<(unsigned)uf_tiles, kUfThreads, 0, stream>>>(img, g, thresh, invert, parent, bits,
                                                                       uf_tiles_x);
    SPX_LAUNCHED();
    cc_merge_kernel<BS><<<(unsigned)uf_tiles, kUfMergeThreads, 0, stream>>>(g, parent, bits, uf_tiles_x);
    SPX_LAUNCHED();
    cc_compress_kernel<<<grid_for(nn, T), T, 0, stream>>>(nn, parent, flag);
    SPX_LAUNCHED();
    SPX_HIP(hipcub::DeviceScan::ExclusiveSum(scan_tmp, scan_bytes, flag, ids, (int)nn, stream));
    cc_acc_init_kernel<<<grid_for((int64_t)n_acc, T), T, 0, stream>>>(acc, (int64_t)n_acc);
    SPX_LAUNCHED();
    SPX_REQUIRE(tiles < (int64_t)0x7fffffff, "spx_threshold_cc: image too large");
    cc_label_kernel<BS><<<(unsigned)tiles, T, 0, stream>>>(g, parent, bits, ids, labels, acc, cap, bg_part, tiles_x);
    SPX_LAUNCHED();
    if (cap > 0) {
        cc_bg_reduce_kernel<<<1, 1024, 0, stream>>>(bg_part, tiles, acc);
        SPX_LAUNCHED();
    }
    // number of components = ids[nn-1] + flag[nn-1]
    int32_t tail[2] = {0, 0};
    SPX_HIP(hipMemcpyAsync(&tail[0], ids + nn - 1, 4, hipMemcpyDeviceToHost, stream));
    SPX_HIP(hipMemcpyAsync(&tail[1], flag + nn - 1, 4, hipMemcpyDeviceToHost, stream));
    SPX_HIP(hipStreamSynchronize(stream));
    const int32_t num = tail[0] + tail[1] + 1;
    *n_out = num;
    const int32_t rows = num < cap ? num : cap;
    if (rows > 0) {
        cc_finalize_kernel<<<grid_for(rows, T), T, 0, stream>>>(acc, rows, stats, cents);
        SPX_LAUNCHED();
    }
    if (num > cap) {
        set_error("spx_threshold_cc: " + std::to_string(num) + " labels exceed capacity " +
                  std::to_string(cap));
        return SPX_E_CAPACITY;
    }
    return SPX_OK;
}

}  // namespace

int otsu_launch(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, int32_t step,
                int32_t* thresh_out, hipStream_t stream) {
    const int64_t Hs = (H + step - 1) / step;
    const int64_t Ws = (W + step - 1) / step;
    void* ws = nullptr;
    int rc = workspace(sizeof(uint32_t) * kHistBins + 256, &ws, stream);
    if (rc != SPX_OK) return rc;
    uint32_t* hist = static_cast<uint32_t*>(ws);
    int32_t* dev_out = reinterpret_cast<int32_t*>(hist + kHistBins);
    fill_u32<<<64, 256, 0, stream>>>(hist, kHistBins, 0u);
    SPX_LAUNCHED();
    const int64_t n = Hs * Ws;
    const int64_t per_block = (int64_t)kHistThreads * kHistPixelsPerThread;
    const int64_t blocks = (n + per_block - 1) / per_block;
    hist_u16_kernel<<<(unsigned)blocks, kHistThreads, 0, stream>>>(img, Hs, Ws, pitch, step, hist);
    SPX_LAUNCHED();
    otsu_kernel<<<1, 1024, 0, stream>>>(hist, dev_out);
    SPX_LAUNCHED();
    SPX_HIP(hipMemcpyAsync(thresh_out, dev_out, sizeof(int32_t), hipMemcpyDeviceToHost, stream));
    SPX_HIP(hipStreamSynchronize(stream));
    return SPX_OK;
}

namespace {

// Triangle threshold of a 256-bin histogram (OpenCV's THRESH_TRIANGLE, 8-bit images only):
// bounds = first / last non-empty bins widened by one, peak = first maximum; the longer
// side of the peak is taken (the histogram mirrored when the right side is longer); the
// threshold is the bin of largest distance to the line from (left bound, 0) to the peak,
// minus one, mirrored back. out = -1 when a grey level above 255 is present.
__global__ __launch_bounds__(1024) void triangle_kernel(const uint32_t* hist, int32_t* out) {
    __shared__ int s_wide;
    const int t = threadIdx.x;
    if (t == 0) s_wide = 0;
    __syncthreads();
    int wide = 0;
    for (int i = 256 + t; i < kHistBins; i += 1024) wide |= hist[i] != 0;
    if (wide) atomicOr(&s_wide, 1);
    __syncthreads();
    if (t != 0) return;
    if (s_wide) {
        out[0] = -1;
        return;
    }
    constexpr int N = 256;
    int lo = 0, hi = 0, peak = 0;
    uint32_t hmax = 0;
    for (int i = 0; i < N; ++i)
        if (hist[i] > 0) {
            lo = i;
            break;
        }
    if (lo > 0) --lo;
    for (int i = N - 1; i > 0; --i)
        if (hist[i] > 0) {
            hi = i;
            break;
        }
    if (hi < N - 1) ++hi;
    for (int i = 0; i < N; ++i)
        if (hist[i] > hmax) {
            hmax = hist[i];
            peak = i;
        }
    const bool flip = peak - lo < hi - peak;
    if (flip) {
        lo = N - 1 - hi;
        peak = N - 1 - peak;
    }
    // distance to the line, up to a positive factor and a constant: hmax * i + (lo - peak) * h(i)
    const double a = (double)hmax, b = (double)(lo - peak);
    double best = 0.0;
    int th = lo;
    for (int i = lo + 1; i <= peak; ++i) {
        const double hv = (double)hist[flip ? N - 1 - i : i];
        const double d = a * (double)i + b * hv;
        if (d > best) {
            best = d;
            th = i;
        }
    }
    --th;
    out[0] = flip ? N - 1 - th : th;
}

// mask = (invert ? v <= t : v > t) as 0 / 1.
__global__ void threshold_mask_kernel(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, int32_t thresh,
                                      int32_t invert, uint16_t* out) {
    const int64_t n = H * W;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t y = i / W, x = i - (i / W) * W;
        const int32_t v = img[y * pitch + x];
        out[i] = (invert ? v <= thresh : v > thresh) ? 1 : 0;
    }
}

// One erosion (op 0) or dilation (op 1) of a 0 / non-zero mask by the 3x3 cross (OpenCV's
// 3x3 MORPH_ELLIPSE). Pixels outside the image do not take part (OpenCV's default border:
// +inf for erosion, -inf for dilation).
__global__ void morph_cross_kernel(const uint16_t* in, int64_t H, int64_t W, int64_t pitch, int32_t op,
                                   uint16_t* out) {
    const int64_t n = H * W;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t y = i / W, x = i - (i / W) * W;
        const bool c = in[y * pitch + x] != 0;
        const bool l = x > 0 ? in[y * pitch + x - 1] != 0 : !op;
        const bool r = x + 1 < W ? in[y * pitch + x + 1] != 0 : !op;
        const bool u = y > 0 ? in[(y - 1) * pitch + x] != 0 : !op;
        const bool d = y + 1 < H ? in[(y + 1) * pitch + x] != 0 : !op;
        const bool v = op ? (c || l || r || u || d) : (c && l && r && u && d);
        out[i] = v ? 1 : 0;
    }
}

}  // namespace

int triangle_launch(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, int32_t step, int32_t* thresh_out,
                    hipStream_t stream) {
    const int64_t Hs = (H + step - 1) / step;
    const int64_t Ws = (W + step - 1) / step;
    void* ws = nullptr;
    int rc = workspace(sizeof(uint32_t) * kHistBins + 256, &ws, stream);
    if (rc != SPX_OK) return rc;
    uint32_t* hist = static_cast<uint32_t*>(ws);
    int32_t* dev_out = reinterpret_cast<int32_t*>(hist + kHistBins);
    fill_u32<<<64, 256, 0, stream>>>(hist, kHistBins, 0u);
    SPX_LAUNCHED();
    const int64_t n = Hs * Ws;
    const int64_t per_block = (int64_t)kHistThreads * kHistPixelsPerThread;
    const int64_t blocks = (n + per_block - 1) / per_block;
    hist_u16_kernel<<<(unsigned)blocks, kHistThreads, 0, stream>>>(img, Hs, Ws, pitch, step, hist);
    SPX_LAUNCHED();
    triangle_kernel<<<1, 1024, 0, stream>>>(hist, dev_out);
    SPX_LAUNCHED();
    SPX_HIP(hipMemcpyAsync(thresh_out, dev_out, sizeof(int32_t), hipMemcpyDeviceToHost, stream));
    SPX_HIP(hipStreamSynchronize(stream));
    if (*thresh_out < 0) {
        set_error("spx_triangle_u8: grey levels above 255; the triangle method is defined for 8-bit images");
        return SPX_E_ARG;
    }
    return SPX_OK;
}

int threshold_mask_launch(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, int32_t thresh, int32_t invert,
                          uint16_t* out, hipStream_t stream) {
    const int64_t n = H * W;
    const unsigned blocks = (unsigned)((n + 255) / 256 < 65536 ? (n + 255) / 256 : 65536);
    threshold_mask_kernel<<<blocks, 256, 0, stream>>>(img, H, W, pitch, thresh, invert, out);
    SPX_LAUNCHED();
    return SPX_OK;
}

int morph_cross_launch(const uint16_t* in, int64_t H, int64_t W, int64_t pitch, int32_t op, uint16_t* out,
                       hipStream_t stream) {
    const int64_t n = H * W;
    const unsigned blocks = (unsigned)((n + 255) / 256 < 65536 ? (n + 255) / 256 : 65536);
    morph_cross_kernel<<<blocks, 256, 0, stream>>>(in, H, W, pitch, op, out);
    SPX_LAUNCHED();
    return SPX_OK;
}

int threshold_cc_launch(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, int32_t thresh,
                        int32_t invert, int32_t connectivity, int32_t* labels, int32_t* stats,
                        double* cents, int32_t cap, int32_t* n_out, hipStream_t stream) {
    if (connectivity == 8)
        return threshold_cc_impl<2>(img, H, W, pitch, thresh, invert, labels, stats, cents, cap, n_out, stream);
    return threshold_cc_impl<1>(img, H, W, pitch, thresh, invert, labels, stats, cents, cap, n_out, stream);
}

}  // namespace spx
