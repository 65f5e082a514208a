// Nearest-seed lists on the device: the nb nearest other seeds of every seed within a radius,
// by a uniform-grid hash (cell side >= radius, so a 3 x 3 block of cells covers the disk; the
// side grows past the radius when the radius is small against the image, to bound the grid).
//
// Counting sort of the seeds into cells (atomic counts, exclusive scan, atomic scatter),
// then one thread per seed scans its 3 x 3 cells and keeps the nb nearest in a sorted
// register list. Candidates are ordered by (squared distance in FP64, index), a total
// order, so the result does not depend on the scatter order.
#include <hipcub/hipcub.hpp>

#include "spx_internal.hpp"

namespace spx {
namespace {

constexpr int kMaxNb = 8;

struct Grid {
    int gw, gh;
    float inv;
};

__device__ __forceinline__ int cell_of(const Grid& g, float x, float y) {
    int cx = (int)floorf(x * g.inv), cy = (int)floorf(y * g.inv);
    cx = cx < 0 ? 0 : (cx >= g.gw ? g.gw - 1 : cx);
    cy = cy < 0 ? 0 : (cy >= g.gh ? g.gh - 1 : cy);
    return cy * g.gw + cx;
}

__global__ void zero_cells(int32_t* p, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

// Seeds with a non-finite coordinate take no part: not binned, no neighbours of their own.
__device__ __forceinline__ bool seed_ok(const float* seeds, int i) {
    return isfinite(seeds[2 * i]) && isfinite(seeds[2 * i + 1]);
}

__global__ void count_cells(const float* seeds, int K, Grid g, int32_t* cnt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < K && seed_ok(seeds, i)) atomicAdd(&cnt[cell_of(g, seeds[2 * i], seeds[2 * i + 1])], 1);
}

__global__ void scatter_cells(const float* seeds, int K, Grid g, int32_t* cursor, int32_t* ids) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < K && seed_ok(seeds, i)) ids[atomicAdd(&cursor[cell_of(g, seeds[2 * i], seeds[2 * i + 1])], 1)] = i;
}

__global__ void query_cells(const float* seeds, int K, Grid g, const int32_t* start, const int32_t* cnt,
                            const int32_t* ids, double r2, int nb, int32_t* out) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= K) return;
    if (!seed_ok(seeds, i)) {
        for (int q = 0; q < nb; ++q) out[(int64_t)i * nb + q] = -1;
        return;
    }
    const double xi = (double)seeds[2 * i], yi = (double)seeds[2 * i + 1];
    const int c = cell_of(g, seeds[2 * i], seeds[2 * i + 1]);
    const int cx = c % g.gw, cy = c / g.gw;
    double bd[kMaxNb];
    int bi[kMaxNb];
    int n = 0;
    for (int dy = -1; dy <= 1; ++dy) {
        const int yy = cy + dy;
        if (yy < 0 || yy >= g.gh) continue;
        for (int dx = -1; dx <= 1; ++dx) {
            const int xx = cx + dx;
            if (xx < 0 || xx >= g.gw) continue;
            const int cc = yy * g.gw + xx;
            const int s = start[cc], e = s + cnt[cc];
            for (int k = s; k < e; ++k) {
                const int j = ids[k];
                if (j == i) continue;
                const double ex = (double)seeds[2 * j] - xi, ey = (double)seeds[2 * j + 1] - yi;
                const double d = ex * ex + ey * ey;
                if (!(d <= r2)) continue;
                // insertion into the sorted list (d, j), keeping the nb smallest
                int pos = n;
                while (pos > 0 && (bd[pos - 1] > d || (bd[pos - 1] == d && bi[pos - 1] > j))) --pos;
                if (pos >= nb) continue;
                const int last = n < nb ? n : nb - 1;
                for (int q = last; q > pos; --q) {
                    bd[q] = bd[q - 1];
                    bi[q] = bi[q - 1];
                }
                bd[pos] = d;
                bi[pos] = j;
                if (n < nb) ++n;
            }
        }
    }
    for (int q = 0; q < nb; ++q) out[(int64_t)i * nb + q] = q < n ? bi[q] : -1;
}

}  // namespace

int seed_neighbours_launch(const float* seeds, int32_t K, int64_t H, int64_t W, float radius, int32_t nb,
                           int32_t* out, hipStream_t stream) {
    if (K == 0) return SPX_OK;
    Grid g;
    const double min_cell = sqrt((double)H * (double)W / (double)(1 << 24));
    const float cell = (float)fmax((double)radius, min_cell);
    g.inv = 1.0f / cell;
    g.gw = (int)(W / cell) + 1;
    g.gh = (int)(H / cell) + 1;
    const int64_t ncell = (int64_t)g.gw * g.gh;
    size_t scan_bytes = 0;
    SPX_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (int32_t*)nullptr, (int32_t*)nullptr, (int)ncell,
                                             stream));
    // workspace: cnt[ncell] | start[ncell] | cursor[ncell] | ids[K] | scan temp
    const size_t head = sizeof(int32_t) * (3 * (size_t)ncell + (size_t)K);
    const size_t head_al = (head + 255) & ~(size_t)255;
    void* ws = nullptr;
    int rc = workspace(head_al + scan_bytes, &ws, stream);
    if (rc != SPX_OK) return rc;
    int32_t* cnt = static_cast<int32_t*>(ws);
    int32_t* start = cnt + ncell;
    int32_t* cursor = start + ncell;
    int32_t* ids = cursor + ncell;
    void* temp = static_cast<char*>(ws) + head_al;
    const int blocks_k = (K + 255) / 256;
    zero_cells<<<(int)((ncell + 255) / 256 < 4096 ? (ncell + 255) / 256 : 4096), 256, 0, stream>>>(cnt, (int)ncell);
    SPX_LAUNCHED();
    count_cells<<<blocks_k, 256, 0, stream>>>(seeds, K, g, cnt);
    SPX_LAUNCHED();
    SPX_HIP(hipcub::DeviceScan::ExclusiveSum(temp, scan_bytes, cnt, start, (int)ncell, stream));
    SPX_HIP(hipMemcpyAsync(cursor, start, sizeof(int32_t) * ncell, hipMemcpyDeviceToDevice, stream));
    scatter_cells<<<blocks_k, 256, 0, stream>>>(seeds, K, g, cursor, ids);
    SPX_LAUNCHED();
    query_cells<<<blocks_k, 256, 0, stream>>>(seeds, K, g, start, cnt, ids, (double)radius * (double)radius, nb, out);
    SPX_LAUNCHED();
    return SPX_OK;
}

}  // namespace spx
