// Learned initial guess: a 4-layer CNN on 32x32 BF16 patches + 2-layer MLP head, run and
// trained with hand-written BF16 MFMA kernels (FP32 accumulation) on gfx950.
//
// No reference counterpart (the learned start is part of the task definition). Layers
// (NHWC activations, BF16; weights as K x N matrices, K = 9 Cin for 3x3 convolutions):
//   conv 1->16  s2 p1 : 32x32 -> 16x16     conv 16->32 s2 p1 : 16x16 -> 8x8
//   conv 32->64 s2 p1 : 8x8   -> 4x4       conv 64->64 s2 p1 : 4x4   -> 2x2
//   fc 256 -> 128 (ReLU), fc 128 -> 7.
// Every layer is an implicit GEMM (im2col into a BF16 column buffer, then C = A B on the
// matrix cores with v_mfma_f32_32x32x8bf16_1k), with the bias and ReLU fused into the GEMM
// epilogue. The backward pass uses the same GEMM with transposed operand strides
// (dW = col^T dZ with split-K partial sums reduced in a fixed order, dcol = dZ W^T) and a
// gather form of col2im, so training is deterministic. Adam keeps FP32 master weights
// and moments; a BF16 copy feeds the GEMMs.
#include <algorithm>

#include "spx_internal.hpp"

namespace spx {
namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// FP32 -> BF16, round to nearest even: gfx950's v_cvt_pk_bf16_f32.
__device__ __forceinline__ uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
typedef float f32x2v __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2v __attribute__((ext_vector_type(2)));
// Two FP32 values -> two BF16 in one 32-bit word (lo = a), one v_cvt_pk_bf16_f32.
__device__ __forceinline__ uint32_t f2bf2(float a, float b) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2v{a, b}), bf16x2v));
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// ---------------------------------------------------------------------------------------
// GEMM: C[z](m, n) = sum_k A(m, k) B(k, n) over the k-slice of split z, A(m, k) at
// A[m*sam + k*sak], B(k, n) at B[k*sbk + n*sbn] (BF16). Epilogue (splits == 1 only):
// + bias[n], ReLU, FP32 store to C (row-major M x N) and/or BF16 store to Cb.
// Block 256 threads = 4 waves, 64 x 64 tile, each wave a 32 x 32 quadrant; K step 32.
constexpr int kBM = 64, kBN = 64, kBK = 32;

struct GemmArgs {
    const uint16_t* A;
    int64_t sam, sak;
    const uint16_t* B;
    int64_t sbk, sbn;
    float* C;
    uint16_t* Cb;
    const float* bias;
    int relu;
    int M, N, K;
    int k_per_split;
};

__global__ __launch_bounds__(256) void gemm_bf16_kernel(GemmArgs g) {
    __shared__ uint16_t As[kBM][kBK + 4];
    __shared__ uint16_t Bs[kBN][kBK + 4];  // transposed: Bs[n][k]
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int m0 = blockIdx.y * kBM, n0 = blockIdx.x * kBN;
    const int kbeg = blockIdx.z * g.k_per_split;
    const int kend = min(g.K, kbeg + g.k_per_split);
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    f32x16 acc;
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.0f;

    for (int k0 = kbeg; k0 < kend; k0 += kBK) {
        // stage A (64 x 32) and B^T (64 x 32): 8 elements per thread each
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int idx = tid + 256 * e;
            const int r = idx / kBK, c = idx % kBK;
            const int gm = m0 + r, gk = k0 + c;
            As[r][c] = (gm < g.M && gk < kend) ? g.A[(int64_t)gm * g.sam + (int64_t)gk * g.sak] : (uint16_t)0;
            const int gn = n0 + r;
            Bs[r][c] = (gn < g.N && gk < kend) ? g.B[(int64_t)gk * g.sbk + (int64_t)gn * g.sbn] : (uint16_t)0;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kBK; kk += 8) {
            bf16x4 a, b;
            const int kq = kk + 4 * (lane >> 5);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                a[j] = (short)As[wm + (lane & 31)][kq + j];
                b[j] = (short)Bs[wn + (lane & 31)][kq + j];
            }
            acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    // D(i, j): j = lane % 32, i = 8 (r / 4) + 4 (lane / 32) + r % 4
    const int64_t zoff = (int64_t)blockIdx.z * g.M * g.N;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int i = wm + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const int j = wn + (lane & 31);
        const int gm = m0 + i, gn = n0 + j;
        if (gm < g.M && gn < g.N) {
            float v = acc[r];
            if (g.bias) v += g.bias[gn];
            if (g.relu) v = fmaxf(v, 0.0f);
            if (g.C) g.C[zoff + (int64_t)gm * g.N + gn] = v;
            if (g.Cb) g.Cb[(int64_t)gm * g.N + gn] = f2bf(v);
        }
    }
}

// out[i] = sum_z part[z * n + i] in split order (deterministic split-K reduction).
__global__ void reduce_splits_kernel(const float* part, int splits, int64_t n, float* out, int accumulate) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
    for (int z = 0; z < splits; ++z) s += part[(int64_t)z * n + i];
    out[i] = accumulate ? out[i] + s : s;
}

// im2col for a 3x3 stride-2 pad-1 convolution on NHWC BF16:
// col[(b Ho + oy) Wo + ox][(ky 3 + kx) Cin + c] = act[b][2oy + ky - 1][2ox + kx - 1][c].
__global__ void im2col_s2_kernel(const uint16_t* act, int Bn, int H, int W, int C, uint16_t* col) {
    const int Ho = H / 2, Wo = W / 2;
    const int64_t total = (int64_t)Bn * Ho * Wo * 9 * C;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        int64_t t = i / C;
        const int kk = (int)(t % 9);
        t /= 9;
        const int ox = (int)(t % Wo);
        t /= Wo;
        const int oy = (int)(t % Ho);
        const int b = (int)(t / Ho);
        const int y = 2 * oy + kk / 3 - 1, x = 2 * ox + kk % 3 - 1;
        col[i] = (y >= 0 && y < H && x >= 0 && x < W) ? act[(((int64_t)b * H + y) * W + x) * C + c] : (uint16_t)0;
    }
}

// Gather form of col2im for the same convolution (FP32):
// dact[b][y][x][c] = sum over (ky, kx) with y = 2oy + ky - 1, x = 2ox + kx - 1 of dcol.
__global__ void col2im_s2_kernel(const float* dcol, int Bn, int H, int W, int C, float* dact) {
    const int Ho = H / 2, Wo = W / 2;
    const int64_t total = (int64_t)Bn * H * W * C;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        int64_t t = i / C;
        const int x = (int)(t % W);
        t /= W;
        const int y = (int)(t % H);
        const int b = (int)(t / H);
        float s = 0.0f;
#pragma unroll
        for (int ky = 0; ky < 3; ++ky) {
            const int yy = y + 1 - ky;
            if (yy < 0 || (yy & 1) || (yy >> 1) >= Ho) continue;
#pragma unroll
            for (int kx = 0; kx < 3; ++kx) {
                const int xx = x + 1 - kx;
                if (xx < 0 || (xx & 1) || (xx >> 1) >= Wo) continue;
                s += dcol[((((int64_t)b * Ho + (yy >> 1)) * Wo + (xx >> 1)) * 9 + ky * 3 + kx) * C + c];
            }
        }
        dact[i] = s;
    }
}

// dZ = dA * (a > 0) (ReLU backward on the post-activation), BF16 copy for the GEMMs, and
// optionally the FP32 dZ itself.
__global__ void relu_bwd_kernel(const float* dA, const uint16_t* a, int64_t n, int relu, uint16_t* dZb, float* dZ) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        float v = dA[i];
        if (relu && !(bf2f(a[i]) > 0.0f)) v = 0.0f;
        dZb[i] = f2bf(v);
        if (dZ) dZ[i] = v;
    }
}

// db[n] = sum_m dZ[m][n] (one block per column, fixed-order tree).
__global__ __launch_bounds__(256) void colsum_kernel(const float* dZ, int M, int N, float* db) {
    __shared__ float s[256];
    const int n = blockIdx.x;
    float acc = 0.0f;
    for (int m = threadIdx.x; m < M; m += 256) acc += dZ[(int64_t)m * N + n];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[n] = s[0];
}

// MSE loss over the head outputs: dY = 2 w (y - t) / (B * 7), per-row loss terms.
__global__ void mse_kernel(const float* y, const float* t, const float* w, int n, float scale, float* dy,
                           float* terms) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float wi = w ? w[i] : 1.0f;
    const float d = y[i] - t[i];
    dy[i] = 2.0f * wi * d * scale;
    terms[i] = wi * d * d * scale;
}

__global__ void adam_kernel(float* theta, uint16_t* theta_bf, const float* grad, float* m, float* v, int64_t n,
                            float lr, float b1, float b2, float eps, float bc1, float bc2) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float g = grad[i];
        const float mi = b1 * m[i] + (1.0f - b1) * g;
        const float vi = b2 * v[i] + (1.0f - b2) * g * g;
        m[i] = mi;
        v[i] = vi;
        const float th = theta[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
        theta[i] = th;
        theta_bf[i] = f2bf(th);
    }
}

__global__ void to_bf16_kernel(const float* x, int64_t n, uint16_t* y) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (int64_t)gridDim.x * blockDim.x) y[i] = f2bf(x[i]);
}

// Patch preparation: the (2h+1)^2 fit patch around rint(seed), resampled bilinearly onto a
// 32 x 32 grid covering it, normalised as (v - B0) / A0 with B0 = mean of the patch's
// outer ring and A0 = max - B0 (1 when flat). norm[k] = {cx, cy, side, B0, A0}.
// One wave per patch, the patch staged in LDS.
constexpr int kPrepWaves = 4;
constexpr int kMaxSide = 63;

// Position of input pixel (i, j) of patch slot `slot` in the fused kernels' LDS input: the
// four stride-2 phases of the 32 x 32 plane apart (see the fused inference below).
__device__ __forceinline__ int input_phase_idx(int slot, int i, int j) {
    return (slot * 4 + (((i & 1) << 1) | (j & 1))) * 256 + (i >> 1) * 16 + (j >> 1);
}

// Patch preparation by NTHR threads (one or two waves): the patch's pixels go into the u16
// staging area P (side^2 values) and the threads keep their maximum. A patch whose 16-byte
// aligned row span lies inside the image is read in 8-pixel chunks (below); otherwise thread t
// gathers pixels t, t + NTHR, ... of the flattened patch. A patch inside the image is read
// without clamps; the loads of kB steps are issued back to
// back before their values are used (a border patch clamps every address into the image, so
// its loads carry no branch either). Row and column advance by the quotient and remainder of
// NTHR / side, the image offset by the matching number of rows and columns.
template <int NTHR>
__device__ __forceinline__ void prep_gather(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, float cx,
                                            float cy, int hs, int t, uint16_t* P, float& vmax) {
    const int side = 2 * hs + 1;
    const int npx = side * side;
    const int64_t ox = (int64_t)cx - hs, oy = (int64_t)cy - hs;
    constexpr int kB = 8;
    const int dq = NTHR / side, dr = NTHR - dq * side;
    int row = t / side, col = t - row * side;
    uint32_t m = 0u;
    // 16-byte chunks of 8 pixels, aligned in the image: [x0, x1) covers the patch's columns
    const int64_t x0 = ox & ~(int64_t)7, x1 = (ox + side + 7) & ~(int64_t)7;
    if (ox >= 0 && oy >= 0 && x1 <= W && oy + side <= H && (pitch & 7) == 0 &&
        (reinterpret_cast<uintptr_t>(img) & 15u) == 0u) {
        // Rows inside the image, every chunk inside its row: thread t loads chunks t, t + NTHR,
        // ... (row c / cpr, chunk c % cpr) with one 16-byte load each and puts the pixels that
        // fall in the patch into P (a load of eight pixels replaces eight 2-byte loads)
        const int cpr = (int)((x1 - x0) >> 3), n = cpr * side;
        const uint16_t* base = img + oy * pitch + x0;
        constexpr int kC = 4;
        for (int c0 = 0; c0 < n; c0 += NTHR * kC) {
            uint4 v[kC];
            int rr[kC], cc[kC];
#pragma unroll
            for (int j = 0; j < kC; ++j) {
                const int c = c0 + NTHR * j + t;
                const int cs = c < n ? c : 0;  // past the end: the first chunk again
                rr[j] = cs / cpr;
                cc[j] = (int)(x0 - ox) + 8 * (cs - rr[j] * cpr);  // patch column of the chunk's first pixel
                v[j] = *reinterpret_cast<const uint4*>(base + (int64_t)rr[j] * pitch + 8 * (cs - rr[j] * cpr));
            }
#pragma unroll
            for (int j = 0; j < kC; ++j) {
                if (c0 + NTHR * j + t >= n) continue;
                const uint32_t w[4] = {v[j].x, v[j].y, v[j].z, v[j].w};
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int col = cc[j] + e;
                    const uint32_t px = (w[e >> 1] >> (16 * (e & 1))) & 0xffffu;
                    if (col >= 0 && col < side) {
                        P[rr[j] * side + col] = (uint16_t)px;
                        m = max(m, px);
                    }
                }
            }
        }
    } else if (ox >= 0 && oy >= 0 && ox + side <= W && oy + side <= H) {
        const uint16_t* first = img + oy * pitch + ox;
        const uint16_t* q = first + (int64_t)row * pitch + col;
        const int64_t step = (int64_t)dq * pitch + dr, wrap = pitch - side;
        for (int p0 = 0; p0 < npx; p0 += NTHR * kB) {
            uint16_t raw[kB];
#pragma unroll
            for (int j = 0; j < kB; ++j) {
                // lanes past the patch's end read its first pixel
                raw[j] = *(p0 + NTHR * j + t < npx ? q : first);
                q += step;
                col += dr;
                if (col >= side) {
                    col -= side;
                    q += wrap;
                }
            }
#pragma unroll
            for (int j = 0; j < kB; ++j) {
                const int p = p0 + NTHR * j + t;
                if (p < npx) P[p] = raw[j];
                m = max(m, (uint32_t)raw[j]);
            }
        }
    } else {
        for (int p0 = 0; p0 < npx; p0 += NTHR * kB) {
            uint16_t raw[kB];
#pragma unroll
            for (int j = 0; j < kB; ++j) {
                const int64_t gx = min(max(ox + col, (int64_t)0), W - 1);
                const int64_t gy = min(max(oy + min(row, side - 1), (int64_t)0), H - 1);
                raw[j] = img[gy * pitch + gx];
                col += dr;
                row += dq;
                if (col >= side) {
                    col -= side;
                    ++row;
                }
            }
#pragma unroll
            for (int j = 0; j < kB; ++j) {
                const int p = p0 + NTHR * j + t;
                if (p < npx) {
                    P[p] = raw[j];
                    m = max(m, (uint32_t)raw[j]);
                }
            }
        }
    }
    vmax = fmaxf(vmax, (float)m);
}

// Partial ring sum of the staged patch (its outer rows and columns, 4 side - 4 pixels) over
// threads t of NTHR. Ring values are integers of at most 65535 and at most 248 of them, so
// every partial sum is an integer below 2^24: exact in FP32 whichever threads add it up.
template <int NTHR>
__device__ __forceinline__ float prep_ring(const uint16_t* P, int side, int t) {
    float rsum = 0.0f;
    const int n = 4 * side - 4;
    for (int r = t; r < n; r += NTHR) {
        // top row, bottom row, then the left and right columns between them
        int y, x;
        if (r < side) {
            y = 0;
            x = r;
        } else if (r < 2 * side) {
            y = side - 1;
            x = r - side;
        } else {
            const int e = r - 2 * side;
            y = 1 + (e >> 1);
            x = (e & 1) ? side - 1 : 0;
        }
        rsum += (float)P[y * side + x];
    }
    return rsum;
}

// Wave totals of the partial ring sums and maxima (every lane).
__device__ __forceinline__ void prep_wave_reduce(float& rsum, float& vmax) {
    for (int off = 32; off >= 1; off >>= 1) {
        rsum += __shfl_xor(rsum, off, 64);
        vmax = fmaxf(vmax, __shfl_xor(vmax, off, 64));
    }
}

// B0 = ring mean, A0 = max - B0 (1 when flat).
__device__ __forceinline__ void prep_levels(float rsum, int side, float vmax, float& b0, float& a0) {
    b0 = rsum / (float)(4 * side - 4);
    a0 = (vmax - b0) > 1.0f ? (vmax - b0) : 1.0f;
}

// Bilinear resampling of the staged patch onto the 32 x 32 grid, written as BF16
// (v - B0) * (1 / A0) to xdst: row-major (slot < 0; global memory) or at
// input_phase_idx(slot, ., .) (slot >= 0; the fused kernels' LDS). Thread t takes column
// j = t % 32 (its x weights fixed) and rows t / 32 + k NTHR / 32. Grid point (j, i) lies at
// patch coordinate (j + 0.5) side / 32 - 0.5, clamped into the patch; its left / upper
// neighbour is at most side - 2, so the right / lower one is always the next pixel.
template <int NTHR>
__device__ __forceinline__ void prep_resample(const uint16_t* P, int side, float B0, float A0, int t,
                                              uint16_t* xdst, int slot) {
    static_assert(NTHR % 32 == 0, "a thread keeps one grid column");
    constexpr int kStep = NTHR / 32;
    const float scale = (float)side / 32.0f;
    const float rA = 1.0f / A0;
    const int j = t & 31, i0 = t >> 5;
    const float px = fminf(fmaxf(((float)j + 0.5f) * scale - 0.5f, 0.0f), (float)(side - 1));
    const int x0 = min((int)px, side - 2);
    const float fx = px - (float)x0;
    const int o0 = slot < 0 ? i0 * 32 + j : input_phase_idx(slot, i0, j);
    constexpr int kDo = 32 * kStep;  // index step per row step, row-major
    constexpr int kDp = 8 * kStep;   // the same in the phase-split layout (the row's phase is fixed)
    const int dstep = slot < 0 ? kDo : kDp;
#pragma unroll 4
    for (int k = 0; k < 1024 / NTHR; ++k) {
        const int i = i0 + kStep * k;
        const float py = fminf(fmaxf(((float)i + 0.5f) * scale - 0.5f, 0.0f), (float)(side - 1));
        const int y0 = min((int)py, side - 2);
        const float fy = py - (float)y0;
        const uint16_t* r0 = P + y0 * side + x0;
        const uint16_t* r1 = r0 + side;
        const float v = (1.0f - fy) * ((1.0f - fx) * (float)r0[0] + fx * (float)r0[1]) +
                        fy * ((1.0f - fx) * (float)r1[0] + fx * (float)r1[1]);
        xdst[o0 + dstep * k] = f2bf((v - B0) * rA);
    }
}

// One wave prepares one patch (gather, ring mean and maximum, resampling); returns {B0, A0}
// through b0 / a0 (every lane).
__device__ __forceinline__ void prep_patch(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, float cx,
                                           float cy, int hs, int lane, uint16_t* P, uint16_t* xdst, int slot,
                                           float& b0, float& a0) {
    const int side = 2 * hs + 1;
    float vmax = -3.0e38f;
    prep_gather<64>(img, H, W, pitch, cx, cy, hs, lane, P, vmax);
    __builtin_amdgcn_wave_barrier();
    float rsum = prep_ring<64>(P, side, lane);
    prep_wave_reduce(rsum, vmax);
    prep_levels(rsum, side, vmax, b0, a0);
    prep_resample<64>(P, side, b0, a0, lane, xdst, slot);
}

__global__ __launch_bounds__(256) void prepare_kernel(const uint16_t* img, int64_t H, int64_t W, int64_t pitch,
                                                      const float* seeds, const int16_t* half, int K,
                                                      uint16_t* x_out, float* norm) {
    __shared__ uint16_t patch[kPrepWaves][kMaxSide * kMaxSide];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int f = blockIdx.x * kPrepWaves + wave;
    if (f >= K) return;
    const int hs = __builtin_amdgcn_readfirstlane(max(1, min((int)half[f], (kMaxSide - 1) / 2)));
    const float cx = rintf(seeds[2 * (int64_t)f]), cy = rintf(seeds[2 * (int64_t)f + 1]);
    float B0, A0;
    prep_patch(img, H, W, pitch, cx, cy, hs, lane, patch[wave], x_out + (int64_t)f * 1024, -1, B0, A0);
    if (lane == 0) {
        norm[5 * (int64_t)f + 0] = cx;
        norm[5 * (int64_t)f + 1] = cy;
        norm[5 * (int64_t)f + 2] = (float)(2 * hs + 1);
        norm[5 * (int64_t)f + 3] = B0;
        norm[5 * (int64_t)f + 4] = A0;
    }
}

// Network outputs t -> fit parameters [x0, y0, r, s, A, B, q] and back:
//   x0 = cx + side t0 / 8, y0 = cy + side t1 / 8, r = side t2, s = side exp(t3) / 16,
//   A = A0 t4, B = B0 + A0 t5, q = 1 + t6 / 8.
__device__ __forceinline__ void decode_one(const float* o, const float* n, float* p) {
    const float side = n[2];
    p[0] = n[0] + side * o[0] * 0.125f;
    p[1] = n[1] + side * o[1] * 0.125f;
    p[2] = side * fmaxf(o[2], 0.01f);
    p[3] = side * expf(fminf(fmaxf(o[3], -6.0f), 3.0f)) * 0.0625f;
    p[4] = n[4] * o[4];
    p[5] = n[3] + n[4] * o[5];
    p[6] = 1.0f + o[6] * 0.125f;
}

__global__ void decode_kernel(const float* t, const float* norm, int K, float* params) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const float* n = norm + 5 * (int64_t)k;
    const float* o = t + 7 * (int64_t)k;
    float* p = params + 7 * (int64_t)k;
    const float side = n[2];
    p[0] = n[0] + side * o[0] * 0.125f;
    p[1] = n[1] + side * o[1] * 0.125f;
    p[2] = side * fmaxf(o[2], 0.01f);
    p[3] = side * expf(fminf(fmaxf(o[3], -6.0f), 3.0f)) * 0.0625f;
    p[4] = n[4] * o[4];
    p[5] = n[3] + n[4] * o[5];
    p[6] = 1.0f + o[6] * 0.125f;
}

__global__ void encode_kernel(const float* params, const float* norm, int K, float* t) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const float* n = norm + 5 * (int64_t)k;
    const float* p = params + 7 * (int64_t)k;
    float* o = t + 7 * (int64_t)k;
    const float side = n[2];
    o[0] = 8.0f * (p[0] - n[0]) / side;
    o[1] = 8.0f * (p[1] - n[1]) / side;
    o[2] = p[2] / side;
    o[3] = logf(fmaxf(16.0f * p[3] / side, 1e-6f));
    o[4] = p[4] / n[4];
    o[5] = (p[5] - n[3]) / n[4];
    o[6] = 8.0f * (p[6] - 1.0f);
}

inline int grid1(int64_t n, int threads = 256, int cap = 65536) {
    int64_t g = (n + threads - 1) / threads;
    return (int)(g < cap ? (g > 0 ? g : 1) : cap);
}

// ---------------------------------------------------------------------------------------
// Network description
struct Layer {
    int kd, n;            // GEMM K and N
    int H, W, Cin;        // input (conv) or 0
    bool conv;
    bool relu;
};
constexpr int kLayers = 6;
constexpr Layer kNet[kLayers] = {
    {9 * 1, 16, 32, 32, 1, true, true},   {9 * 16, 32, 16, 16, 16, true, true},
    {9 * 32, 64, 8, 8, 32, true, true},   {9 * 64, 64, 4, 4, 64, true, true},
    {256, 128, 0, 0, 0, false, true},     {128, 7, 0, 0, 0, false, false},
};

__host__ __device__ constexpr int64_t param_offset_dev(int l) {  // start of layer l's weights (K x N)
    int64_t o = 0;
    for (int i = 0; i < l; ++i) o += (int64_t)kNet[i].kd * kNet[i].n + kNet[i].n;
    return o;
}
int64_t param_offset(int l) { return param_offset_dev(l); }  // bias follows the K x N weights
int64_t param_total() { return param_offset(kLayers); }

int64_t rows_of(int l, int Bn) {  // GEMM M of layer l
    if (!kNet[l].conv) return Bn;
    return (int64_t)Bn * (kNet[l].H / 2) * (kNet[l].W / 2);
}

int launch_gemm(const uint16_t* A, int64_t sam, int64_t sak, const uint16_t* B, int64_t sbk, int64_t sbn,
                float* C, uint16_t* Cb, const float* bias, int relu, int M, int N, int K, int splits,
                hipStream_t stream) {
    GemmArgs g{A, sam, sak, B, sbk, sbn, C, Cb, bias, relu, M, N, K, 0};
    if (splits < 1) splits = 1;
    int kps = (K + splits - 1) / splits;
    kps = ((kps + kBK - 1) / kBK) * kBK;
    splits = (K + kps - 1) / kps;
    g.k_per_split = kps;
    dim3 grid((N + kBN - 1) / kBN, (M + kBM - 1) / kBM, splits);
    hipLaunchKernelGGL(gemm_bf16_kernel, grid, dim3(256), 0, stream, g);
    SPX_LAUNCHED();
    return SPX_OK;
}

// Workspace layout for a batch of Bn patches: per layer the column buffer (BF16), the
// post-activation (BF16) and, for training, FP32 scratch for gradients.
struct CnnBuffers {
    uint16_t* col[kLayers];
    uint16_t* act[kLayers];   // output of layer l (BF16)
    float* out;               // head output (FP32, Bn x 7)
    float* g_act;             // FP32 gradient w.r.t. a layer output (max size)
    float* g_col;             // FP32 gradient w.r.t. a column buffer (max size)
    uint16_t* g_zb;           // BF16 dZ
    float* g_z;               // FP32 dZ
    float* g_part;            // split-K partial sums of dW
    float* grad;              // parameter gradients (FP32)
    float* loss_terms;
    float* dy;
};

int64_t max_col(int Bn) {
    int64_t m = 0;
    for (int l = 0; l < kLayers; ++l) m = std::max(m, rows_of(l, Bn) * kNet[l].kd);
    return m;
}
int64_t max_act(int Bn) {
    int64_t m = (int64_t)Bn * 1024;
    for (int l = 0; l < kLayers; ++l) m = std::max(m, rows_of(l, Bn) * kNet[l].n);
    return m;
}
constexpr int kSplits = 16;

int cnn_buffers(int Bn, bool train, CnnBuffers& b, hipStream_t stream) {
    auto al = [](int64_t x) { return (size_t)((x + 255) & ~(int64_t)255); };
    size_t total = 0;
    size_t off_col[kLayers], off_act[kLayers];
    for (int l = 0; l < kLayers; ++l) {
        off_col[l] = total;
        total += al(2 * rows_of(l, Bn) * kNet[l].kd);
        off_act[l] = total;
        total += al(2 * rows_of(l, Bn) * kNet[l].n);
    }
    const size_t off_out = total;
    total += al(4 * (int64_t)Bn * 7);
    size_t off_gact = 0, off_gcol = 0, off_gzb = 0, off_gz = 0, off_part = 0, off_grad = 0, off_loss = 0, off_dy = 0;
    if (train) {
        off_gact = total;
        total += al(4 * max_act(Bn));
        off_gcol = total;
        total += al(4 * max_col(Bn));
        off_gzb = total;
        total += al(2 * max_act(Bn));
        off_gz = total;
        total += al(4 * max_act(Bn));
        off_part = total;
        int64_t mp = 0;
        for (int l = 0; l < kLayers; ++l) mp = std::max(mp, (int64_t)kNet[l].kd * kNet[l].n);
        total += al(4 * mp * kSplits);
        off_grad = total;
        total += al(4 * param_total());
        off_loss = total;
        total += al(4 * (int64_t)Bn * 7);
        off_dy = total;
        total += al(4 * (int64_t)Bn * 7);
    }
    void* ws = nullptr;
    int rc = workspace(total, &ws, stream);
    if (rc != SPX_OK) return rc;
    char* base = static_cast<char*>(ws);
    for (int l = 0; l < kLayers; ++l) {
        b.col[l] = reinterpret_cast<uint16_t*>(base + off_col[l]);
        b.act[l] = reinterpret_cast<uint16_t*>(base + off_act[l]);
    }
    b.out = reinterpret_cast<float*>(base + off_out);
    if (train) {
        b.g_act = reinterpret_cast<float*>(base + off_gact);
        b.g_col = reinterpret_cast<float*>(base + off_gcol);
        b.g_zb = reinterpret_cast<uint16_t*>(base + off_gzb);
        b.g_z = reinterpret_cast<float*>(base + off_gz);
        b.g_part = reinterpret_cast<float*>(base + off_part);
        b.grad = reinterpret_cast<float*>(base + off_grad);
        b.loss_terms = reinterpret_cast<float*>(base + off_loss);
        b.dy = reinterpret_cast<float*>(base + off_dy);
    }
    return SPX_OK;
}

// Forward over Bn patches x (Bn x 1024 BF16) with BF16 weights wbf and FP32 biases.
int forward(const uint16_t* x, int Bn, const uint16_t* wbf, const float* theta, CnnBuffers& b,
            hipStream_t stream) {
    const uint16_t* in = x;
    for (int l = 0; l < kLayers; ++l) {
        const Layer& L = kNet[l];
        const int64_t M = rows_of(l, Bn);
        const uint16_t* A;
        if (L.conv) {
            const int64_t n = M * L.kd;
            im2col_s2_kernel<<<grid1(n), 256, 0, stream>>>(in, Bn, L.H, L.W, L.Cin, b.col[l]);
            SPX_LAUNCHED();
            A = b.col[l];
        } else {
            A = in;  // previous activation, row-major Bn x kd
        }
        const int64_t wo = param_offset(l);
        const bool last = (l == kLayers - 1);
        int rc = launch_gemm(A, L.kd, 1, wbf + wo, L.n, 1, last ? b.out : nullptr, last ? nullptr : b.act[l],
                             theta + wo + (int64_t)L.kd * L.n, L.relu ? 1 : 0, (int)M, L.n, L.kd, 1, stream);
        if (rc != SPX_OK) return rc;
        in = b.act[l];
    }
    return SPX_OK;
}

// ---------------------------------------------------------------------------------------
// Fused inference: one workgroup of 8 waves takes kNP = 4 patches through all six layers,
// activations staying in LDS (49 KB, three workgroups per CU). conv1 (1 -> 16, 74 kflop per
// patch) runs on v_mfma_f32_16x16x16_bf16, two waves per patch. The other conv layers and fc1
// are implicit GEMMs on gfx950's v_mfma_f32_16x16x32_bf16 (rows = (patch, output pixel),
// columns = output channels, k = (ky, kx, c)); a wave owns GC 16-column blocks and G row
// blocks, so each weight fragment feeds G MFMAs and each activation fragment GC. The weights
// are a transposed, K-padded copy W^T in fragment order: the 16 x 32 block of a (column block,
// k chunk) is 512 contiguous BF16 values, lane r + 16 c holding row r, k 8 c .. 8 c + 7, so a
// wave reads a fragment with one coalesced 1 KB load (with rows of W^T apart, each load
// touched sixteen 128-byte lines for 1 KB and the layers waited on L2: 1.66 -> 1.00 ms for
// 199,440 patches). Bias + ReLU + BF16 rounding in the epilogue, as the GEMM path of training.
//
// LDS layout of a layer input of side S and C channels (C >= 8): the four stride-2 phases of
// the plane apart (phase (y & 1, x & 1), sub-pixel (y >> 1, x >> 1)), so the 16 rows of an A
// fragment (16 consecutive output pixels) read consecutive sub-pixels for every tap, and the
// 8-channel chunks of a sub-pixel sp XOR-permuted by sp, so those 16 reads of 16 bytes do not
// pile onto the same banks:
//   idx(pt, y, x, c) = ((pt * 4 + ph) * (S / 2)^2 + sp) * C + (((c >> 3) ^ (sp & (C / 8 - 1))) << 3 | (c & 7)).
// The 32 x 32 input (C = 1) uses the phase split alone; the fc1 input (the 2 x 2 x 64 NHWC
// flatten, 256 values per patch) permutes its 8-value chunks by the patch index.
typedef __bf16 bf16x8v __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
constexpr int kNP = 4;          // patches per workgroup
constexpr int kIW = 8;          // waves per workgroup (two per patch in conv1)
static_assert(kIW % kNP == 0, "conv1 splits each patch's rows evenly over the waves");
constexpr int kBufX = kNP * 2048;  // BF16 values: input, L2 out, L4 out
constexpr int kZeroElems = 128;    // the zeroed block that reads outside a plane (or beyond K) go to
constexpr int kBufY = kNP * 4096;  // BF16 values: raw patch staging, L1 out, L3 out, fc1 out
static_assert(kNP * 1024 <= kBufX && kNP * 8 * 8 * 32 <= kBufX, "bufX");
static_assert(kNP * 16 * 16 * 16 <= kBufY, "bufY");

struct InferWeights {
    const uint16_t* wt;      // W^T per layer 0..5 in MFMA fragment order (16 x 32 blocks of 512 values), K
                             // padded to a multiple of 32 with zeros; the head's 7 columns padded to 16 zero rows
    int64_t off[6];          // element offset of layer l in wt
    int kp[6];               // padded K of layer l
    const float* theta;      // FP32 parameters (biases; K x N layout)
    const uint16_t* wbf;     // BF16 K x N weights (head)
};

// Columns of layer l in its transposed copy (the head's 7 padded to one 16-column block).
__host__ __device__ constexpr int padded_cols(int l) { return kNet[l].n < 16 ? 16 : kNet[l].n; }

__global__ void transpose_weights_kernel(const uint16_t* wbf, InferWeights iw, uint16_t* wt) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t base = 0;
#pragma unroll
    for (int l = 0; l < 6; ++l) {
        const int K = kNet[l].kd, N = kNet[l].n, Kp = iw.kp[l];
        const int64_t cnt = (int64_t)padded_cols(l) * Kp;
        if (t >= base && t < base + cnt) {
            // fragment order: the 16 x 32 block (nb, kc) of W^T is 512 contiguous values, lane
            // r + 16 c holding W^T[16 nb + r][32 kc + 8 c .. + 7], so a wave reads a block
            // with one fully coalesced 1 KB load
            const int64_t q = t - base;
            const int j = (int)(q & 7), ln = (int)((q >> 3) & 63);
            const int64_t blk = q >> 9;
            const int KC = Kp / 32;
            const int n = (int)(blk / KC) * 16 + (ln & 15), k = (int)(blk % KC) * 32 + 8 * (ln >> 4) + j;
            wt[iw.off[l] + q] = (k < K && n < N) ? wbf[param_offset_dev(l) + (int64_t)k * N + n] : (uint16_t)0;
        }
        base += cnt;
    }
}

// Index of element (pt, y, x, c) of a phase-split layer input of side S, C channels.
template <int S, int C>
__device__ __forceinline__ int act_idx(int pt, int y, int x, int c) {
    constexpr int HS = S / 2;
    const int sp = (y >> 1) * HS + (x >> 1);
    const int ph = ((y & 1) << 1) | (x & 1);
    if constexpr (C == 1) {
        static_assert(S == 32 || C != 1, "the C = 1 layout is the 32 x 32 input's");
        return (pt * 4 + ph) * HS * HS + sp;  // = input_phase_idx(pt, y, x)
    } else {
        constexpr int CM = C / 8 - 1;
        return ((pt * 4 + ph) * HS * HS + sp) * C + ((((c >> 3) ^ (sp & CM)) << 3) | (c & 7));
    }
}
// Index of value k of patch pt in the fc1 input (2 x 2 x 64 NHWC flatten).
__device__ __forceinline__ int fc_idx(int pt, int k) { return pt * 256 + ((((k >> 3) ^ (pt & 31)) << 3) | (k & 7)); }

__device__ __forceinline__ bf16x8v load8(const uint16_t* p) { return *reinterpret_cast<const bf16x8v*>(p); }

// One 3x3 stride-2 pad-1 conv layer of the kNP patches, `in` (side HIN, CIN channels) to `out`
// (the next layer's input layout; NEXT_FC: the fc1 input). Wave w takes column blocks
// GC (w % (NB / GC)) .. + GC - 1 and G consecutive row blocks. The k loop is unrolled, so a chunk's tap, phase and
// sub-pixel shift are constants (for CIN >= 32): tap (ky, kx) reads phase ((ky + 1) & 1,
// (kx + 1) & 1) at the output pixel's sub-pixel moved up / left by one for ky = 0 / kx = 0,
// outside the plane only there and only for the first output row / column. A fragment outside
// the plane (or beyond K) is read from `zb`, a zeroed block in LDS, so every load is
// unconditional: the G loads of a chunk issue back to back ahead of its G x GC MFMAs.
template <int CIN, int COUT, int HIN, int G, int GC, bool NEXT_FC>
__device__ __forceinline__ void conv_layer(const uint16_t* __restrict__ in, uint16_t* __restrict__ out,
                                           const uint16_t* __restrict__ wt, const float* __restrict__ bias, int wave,
                                           int lane, const uint16_t* zb) {
    constexpr int HOUT = HIN / 2, PIX = HOUT * HOUT, M = kNP * PIX, MB = M / 16, NB = COUT / 16;
    constexpr int KIN = 9 * CIN, KC = (KIN + 31) / 32;
    constexpr int HS = HIN / 2, PLANE = HS * HS * CIN, CM = CIN / 8 - 1;
    constexpr int NG = NB / GC;  // column-block groups
    static_assert(NB % GC == 0 && MB % G == 0 && (MB / G) * NG <= kIW,
                  "a wave takes G row blocks of GC column blocks");
    if (wave >= (MB / G) * NG) return;  // wave-uniform: more waves than tiles
    static_assert(CIN % 8 == 0 && HS == HOUT, "8-channel chunks; the output pixel indexes the sub-plane");
    const int li = lane & 15, g4 = lane >> 4;
    const int nb0 = (wave % NG) * GC, rb0 = (wave / NG) * G;
    const uint16_t* wcol = wt + ((int64_t)nb0 * KC * 64 + lane) * 8;  // fragment-ordered W^T
    int base[G], sp0[G];
    bool top[G], left[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const int m = (rb0 + g) * 16 + li;
        const int pt = m / PIX, pix = m - (m / PIX) * PIX;
        base[g] = pt * 4 * PLANE;
        sp0[g] = pix;  // oy * HS + ox
        top[g] = pix < HOUT;
        left[g] = (pix % HOUT) == 0;
    }
    f32x4 acc[G][GC];
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
        for (int c = 0; c < GC; ++c) acc[g][c] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kc = 0; kc < KC; ++kc) {
        bf16x8v b[GC];
#pragma unroll
        for (int c = 0; c < GC; ++c) b[c] = load8(wcol + (c * KC + kc) * 512);
        const int k0 = kc * 32 + 8 * g4;
        const int tap = k0 / CIN, chunk = (k0 % CIN) >> 3;
        const int ky = tap / 3, kx = tap - 3 * (tap / 3);
        const int off = ((((ky + 1) & 1) << 1) | ((kx + 1) & 1)) * PLANE;
        const int dsp = (ky == 0 ? -HS : 0) + (kx == 0 ? -1 : 0);
        const bool live = k0 < KIN;
        bf16x8v a[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int sp = sp0[g] + dsp;
            const bool ok = live && !(ky == 0 && top[g]) && !(kx == 0 && left[g]);
            a[g] = load8(ok ? in + base[g] + off + sp * CIN + ((chunk ^ (sp & CM)) << 3) : zb);
        }
        // the transposed product W^T X^T: the weight fragment is the A operand, the
        // activation fragment the B operand (the same registers as X W's B and A)
#pragma unroll
        for (int g = 0; g < G; ++g)
#pragma unroll
            for (int c = 0; c < GC; ++c)
                acc[g][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b[c], a[g], acc[g][c], 0, 0, 0);
    }
    // D^T(i, j): output channel nb * 16 + 4 (lane / 16) + r of output pixel (rb0 + g) * 16 +
    // lane % 16, so a lane holds four consecutive channels of one pixel: one 8-byte store
#pragma unroll
    for (int c = 0; c < GC; ++c) {
        const int n0 = (nb0 + c) * 16 + 4 * g4;
        float bn[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) bn[r] = bias[n0 + r];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int m = (rb0 + g) * 16 + li;
            const int p = m / PIX, pix = m - (m / PIX) * PIX;
            const int y = pix / HOUT, x = pix - (pix / HOUT) * HOUT;
            const uint32_t lo = f2bf2(fmaxf(acc[g][c][0] + bn[0], 0.0f), fmaxf(acc[g][c][1] + bn[1], 0.0f));
            const uint32_t hi = f2bf2(fmaxf(acc[g][c][2] + bn[2], 0.0f), fmaxf(acc[g][c][3] + bn[3], 0.0f));
            const int idx = NEXT_FC ? fc_idx(p, pix * COUT + n0) : act_idx<HOUT, COUT>(p, y, x, n0);
            *reinterpret_cast<uint2*>(out + idx) = make_uint2(lo, hi);
        }
    }
}

// conv1 (1 -> 16, 3x3 stride 2 pad 1, 32 x 32 -> 16 x 16) of patch `wave % kNP`, output rows split
// over kIW / kNP waves, on the matrix cores: one v_mfma_f32_16x16x16_bf16 product W^T X^T per output row j, with the 16
// channels as rows, the row's 16 output pixels as columns and the 9 taps (padded to 16) as K.
// Lane l holds taps 4 (l / 16) .. + 3 of output pixel (j, l % 16) as the B operand and the
// same taps' weights of channel l % 16 as the A operand (fixed for the patch); taps 9..15 are
// zero. The result gives lane l channels 4 (l / 16) .. + 3 of pixel (j, l % 16): bias, ReLU,
// BF16 and one 8-byte store. Products of BF16 values are exact; the FP32 sums are the MFMA's.
// w1: the 9 x 16 weights as FP32 (BF16 values) and the 16 biases, in LDS. Taps outside the
// plane read the zeroed block zb; the taps of four output rows are loaded before their four
// products.
__device__ __forceinline__ void conv1_mfma(const uint16_t* __restrict__ in, uint16_t* __restrict__ out,
                                           const float* __restrict__ w1, int wave, int lane, const uint16_t* zb) {
    const int li = lane & 15, g4 = lane >> 4;
    constexpr int kRowsPerWave = 16 * kNP / kIW;
    const int pt = wave % kNP, jb = (wave / kNP) * kRowsPerWave;  // patch and first output row
    bf16x4 wa;
    short tap_base[4];
    bool tap_live[4], tap_top[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int t = 4 * g4 + i;  // tap index, live when < 9
        const int ky = t / 3, kx = t - 3 * (t / 3);
        tap_live[i] = t < 9 && !(kx == 0 && li == 0);
        tap_top[i] = ky == 0;
        // input pixel (2 j + ky - 1, 2 li + kx - 1): phase ((ky + 1) & 1, (kx + 1) & 1), sub-pixel
        // (j - (ky == 0), li - (kx == 0)); the row offset 16 j is added per output row
        const int ph = (((ky + 1) & 1) << 1) | ((kx + 1) & 1);
        tap_base[i] = (short)((pt * 4 + ph) * 256 + (ky == 0 ? -16 : 0) + li - (kx == 0 ? 1 : 0));
        wa[i] = t < 9 ? (short)f2bf(w1[t * 16 + li]) : (short)0;
    }
    float bn[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) bn[r] = w1[144 + 4 * g4 + r];
    // tap i of output row jb + r is read at tp[i] + 16 r: a dead tap reads the zeroed rows zb
    // (kZeroElems values: every row offset 16 r stays inside it), the taps above the plane (ky = 0 for output row 0) read zb too
    static_assert(16 * kRowsPerWave <= kZeroElems, "the zeroed block covers every row offset");
    const uint16_t* tp[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) tp[i] = tap_live[i] ? in + tap_base[i] + 16 * jb : zb;
    constexpr int kRows = 4;
#pragma unroll
    for (int j0 = 0; j0 < kRowsPerWave; j0 += kRows) {
        bf16x4 xb[kRows];
#pragma unroll
        for (int r = 0; r < kRows; ++r) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const uint16_t* a = tp[i];
                if (j0 + r == 0) a = (jb == 0 && tap_top[i]) ? zb : a;
                xb[r][i] = (short)a[16 * (j0 + r)];
            }
        }
        f32x4 d[kRows];
#pragma unroll
        for (int r = 0; r < kRows; ++r)
            d[r] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wa, xb[r], f32x4{0.0f, 0.0f, 0.0f, 0.0f}, 0, 0, 0);
#pragma unroll
        for (int r = 0; r < kRows; ++r) {
            const uint32_t lo = f2bf2(fmaxf(d[r][0] + bn[0], 0.0f), fmaxf(d[r][1] + bn[1], 0.0f));
            const uint32_t hi = f2bf2(fmaxf(d[r][2] + bn[2], 0.0f), fmaxf(d[r][3] + bn[3], 0.0f));
            *reinterpret_cast<uint2*>(out + act_idx<16, 16>(pt, jb + j0 + r, li, 4 * g4)) = make_uint2(lo, hi);
        }
    }
}

// fc1 (256 -> 128, ReLU) of the kNP patches: rows 0 .. kNP-1 of one 16-row block; the 8
// column blocks go round the waves, the 8 k chunks alternating between two accumulators.
__device__ __forceinline__ void fc1_layer(const uint16_t* __restrict__ in, uint16_t* __restrict__ out,
                                          const uint16_t* __restrict__ wt, const float* __restrict__ bias, int wave,
                                          int lane, const uint16_t* zb) {
    const int li = lane & 15, g4 = lane >> 4;
    for (int nb = wave; nb < 8; nb += kIW) {
        const uint16_t* wcol = wt + ((int64_t)nb * 8 * 64 + lane) * 8;  // fragment-ordered W^T, 8 k blocks
        f32x4 acc0 = f32x4{0.0f, 0.0f, 0.0f, 0.0f}, acc1 = acc0;
#pragma unroll
        for (int kc = 0; kc < 8; kc += 2) {
            const bf16x8v b0 = load8(wcol + kc * 512), b1 = load8(wcol + kc * 512 + 512);
            const bool row = li < kNP;  // rows beyond the patches read the zeroed block
            const bf16x8v a0 = load8(row ? in + fc_idx(li, kc * 32 + 8 * g4) : zb);
            const bf16x8v a1 = load8(row ? in + fc_idx(li, kc * 32 + 32 + 8 * g4) : zb);
            acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b0, acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1, acc1, 0, 0, 0);
        }
        const int n = nb * 16 + li;
        const float bn = bias[n];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int p = 4 * g4 + r;
            if (p < kNP) out[p * 128 + n] = f2bf(fmaxf(acc0[r] + acc1[r] + bn, 0.0f));
        }
    }
}

// conv1's BF16 weights (as FP32) and biases into LDS, and the zeroed block that fragments
// outside a plane read; both read after the first barrier.
__device__ __forceinline__ void stage_conv1(const InferWeights& iw, float* w1, uint16_t* zb) {
    const int tid = threadIdx.x;
    if (tid < 144) w1[tid] = bf2f(iw.wbf[param_offset_dev(0) + tid]);
    else if (tid < 160) w1[tid] = iw.theta[param_offset_dev(0) + 144 + (tid - 144)];
    else if (tid < 160 + kZeroElems) zb[tid - 160] = 0;
}

// Layers 1-6 on the kNP patches in bufX (input layout), then the outputs of patches p0 ..
// p0 + kNP - 1 (< K): decoded with nrm[patch][5] when `decode`, raw otherwise.
__device__ __forceinline__ void infer_net(uint16_t* bufX, uint16_t* bufY, float* head, const float* w1,
                                          const uint16_t* zb, const InferWeights& iw, int K, int p0, bool decode,
                                          const float* nrm, float* out) {
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const float* th = iw.theta;
    conv1_mfma(bufX, bufY, w1, wave, lane, zb);
    __syncthreads();
    conv_layer<16, 32, 16, 2, 2, false>(bufY, bufX, iw.wt + iw.off[1], th + param_offset_dev(1) + 144 * 32, wave, lane,
                                     zb);
    __syncthreads();
    conv_layer<32, 64, 8, 2, 1, false>(bufX, bufY, iw.wt + iw.off[2], th + param_offset_dev(2) + 288 * 64, wave, lane,
                                    zb);
    __syncthreads();
    conv_layer<64, 64, 4, 1, 1, true>(bufY, bufX, iw.wt + iw.off[3], th + param_offset_dev(3) + 576 * 64, wave, lane,
                                   zb);
    __syncthreads();
    fc1_layer(bufX, bufY, iw.wt + iw.off[4], th + param_offset_dev(4) + 256 * 128, wave, lane, zb);
    __syncthreads();
    // head 128 -> 7: one wave, rows = patches, columns = outputs (7 of a 16-column block), 4 k
    // chunks; FP32 outputs, no ReLU
    if (wave == 0) {
        const int li = lane & 15, g4 = lane >> 4;
        const uint16_t* wcol = iw.wt + iw.off[5] + lane * 8;  // fragment-ordered W^T, 4 k blocks
        f32x4 acc = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kc = 0; kc < 4; ++kc) {
            const bf16x8v b = load8(wcol + kc * 512);
            const bf16x8v a = load8(li < kNP ? bufY + li * 128 + kc * 32 + 8 * g4 : zb);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
        }
        const int64_t wo = param_offset_dev(5);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int pt = 4 * g4 + r;
            if (pt < kNP && li < 7) head[pt * 8 + li] = acc[r] + th[wo + 128 * 7 + li];
        }
    }
    __syncthreads();
    if (tid < kNP && p0 + tid < K) {
        const int64_t k = p0 + tid;
        if (decode) {
            decode_one(head + tid * 8, nrm + 5 * tid, out + 7 * k);
        } else {
#pragma unroll
            for (int j = 0; j < 7; ++j) out[7 * k + j] = head[tid * 8 + j];
        }
    }
}

__global__ __launch_bounds__(64 * kIW) __attribute__((amdgpu_waves_per_eu(6))) void cnn_infer_kernel(const uint16_t* __restrict__ x, int K, InferWeights iw,
                                                        const float* __restrict__ norm, float* __restrict__ out) {
    __shared__ __attribute__((aligned(16))) uint16_t bufX[kBufX];
    __shared__ __attribute__((aligned(16))) uint16_t bufY[kBufY];
    __shared__ float head[kNP * 8];
    __shared__ float w1[160];
    __shared__ __attribute__((aligned(16))) uint16_t zb[kZeroElems];
    const int tid = threadIdx.x;
    const int p0 = blockIdx.x * kNP;
    stage_conv1(iw, w1, zb);
    // input patches (row-major 32 x 32 BF16) into the phase-split layout; missing patches read 0.
    // With x 16-byte aligned a thread moves 8 values of a row in one load: the even columns go
    // to one phase plane and the odd ones to the other, 4 consecutive values each.
    static_assert(kNP * 1024 == 8 * 64 * kIW, "one 8-value piece per thread");
    if ((reinterpret_cast<uintptr_t>(x) & 15u) == 0u) {
        const int q = 8 * tid, pt = q >> 10, e = q & 1023, i = e >> 5, j = e & 31;
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (p0 + pt < K) v = *reinterpret_cast<const uint4*>(x + (int64_t)(p0 + pt) * 1024 + e);
        const uint2 even = make_uint2(__builtin_amdgcn_perm(v.y, v.x, 0x05040100u),
                                      __builtin_amdgcn_perm(v.w, v.z, 0x05040100u));
        const uint2 odd = make_uint2(__builtin_amdgcn_perm(v.y, v.x, 0x07060302u),
                                     __builtin_amdgcn_perm(v.w, v.z, 0x07060302u));
        *reinterpret_cast<uint2*>(bufX + act_idx<32, 1>(pt, i, j, 0)) = even;
        *reinterpret_cast<uint2*>(bufX + act_idx<32, 1>(pt, i, j + 1, 0)) = odd;
    } else {
        for (int q = tid; q < kNP * 1024; q += 64 * kIW) {
            const int pt = q >> 10, e = q & 1023;
            const uint16_t v = p0 + pt < K ? x[(int64_t)(p0 + pt) * 1024 + e] : (uint16_t)0;
            bufX[act_idx<32, 1>(pt, e >> 5, e & 31, 0)] = v;
        }
    }
    __syncthreads();
    infer_net(bufX, bufY, head, w1, zb, iw, K, p0, norm != nullptr, norm ? norm + 5 * (int64_t)p0 : nullptr, out);
}

// Proposal in one launch: wave w prepares patch p0 + w (raw pixels staged in bufY, free until
// the first layer writes it), writing the resampled input straight into bufX's layout, then
// the network and the decode.
static_assert(kNP * kMaxSide * kMaxSide <= kBufY, "patch staging must fit in bufY");

__global__ __launch_bounds__(64 * kIW) __attribute__((amdgpu_waves_per_eu(6))) void cnn_propose_kernel(const uint16_t* __restrict__ img, int64_t H, int64_t W,
                                                          int64_t pitch, const float* __restrict__ seeds,
                                                          const int16_t* __restrict__ half, int K, InferWeights iw,
                                                          float* __restrict__ out) {
    __shared__ __attribute__((aligned(16))) uint16_t bufX[kBufX];
    __shared__ __attribute__((aligned(16))) uint16_t bufY[kBufY];
    __shared__ float head[kNP * 8];
    __shared__ float s_norm[kNP * 5];
    __shared__ float s_part[kIW * 2];
    __shared__ float w1[160];
    __shared__ __attribute__((aligned(16))) uint16_t zb[kZeroElems];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int p0 = blockIdx.x * kNP;
    stage_conv1(iw, w1, zb);
    // waves pt and pt + kNP prepare patch pt together: the gather, a barrier, then the
    // resampling (the staged pixels in bufY, free until conv1 writes it)
    static_assert(kIW == 2 * kNP, "two waves per patch in the preparation");
    const int pt = wave % kNP, h = wave / kNP, t = lane + 64 * h;
    const int f = p0 + pt;
    int hs = 1;
    float cx = 0.0f, cy = 0.0f;
    uint16_t* P = bufY + pt * kMaxSide * kMaxSide;
    float vmax = -3.0e38f;
    if (f < K) {
        hs = __builtin_amdgcn_readfirstlane(max(1, min((int)half[f], (kMaxSide - 1) / 2)));
        cx = rintf(seeds[2 * (int64_t)f]);
        cy = rintf(seeds[2 * (int64_t)f + 1]);
        prep_gather<128>(img, H, W, pitch, cx, cy, hs, t, P, vmax);
    }
    __syncthreads();
    if (f < K) {
        float rsum = prep_ring<128>(P, 2 * hs + 1, t);
        prep_wave_reduce(rsum, vmax);
        if (lane == 0) {
            s_part[2 * wave + 0] = rsum;
            s_part[2 * wave + 1] = vmax;
        }
    }
    __syncthreads();
    if (f < K) {
        const int o = 2 * pt, q = 2 * (pt + kNP);
        float B0, A0;
        prep_levels(s_part[o] + s_part[q], 2 * hs + 1, fmaxf(s_part[o + 1], s_part[q + 1]), B0, A0);
        prep_resample<128>(P, 2 * hs + 1, B0, A0, t, bufX, pt);
        if (t == 0) {
            s_norm[5 * pt + 0] = cx;
            s_norm[5 * pt + 1] = cy;
            s_norm[5 * pt + 2] = (float)(2 * hs + 1);
            s_norm[5 * pt + 3] = B0;
            s_norm[5 * pt + 4] = A0;
        }
    } else {
        for (int o = t; o < 1024; o += 128) bufX[act_idx<32, 1>(pt, o >> 5, o & 31, 0)] = 0;
    }
    __syncthreads();
    infer_net(bufX, bufY, head, w1, zb, iw, K, p0, true, s_norm, out);
}

}  // namespace

int64_t cnn_param_count() { return param_total(); }

int gemm_bf16_launch(const uint16_t* A, int64_t sam, int64_t sak, const uint16_t* B, int64_t sbk, int64_t sbn,
                     float* C, int M, int N, int K, hipStream_t stream) {
    return launch_gemm(A, sam, sak, B, sbk, sbn, C, nullptr, nullptr, 0, M, N, K, 1, stream);
}

int cnn_prepare_launch(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, const float* seeds,
                       const int16_t* half, int K, uint16_t* x_out, float* norm, hipStream_t stream) {
    if (K == 0) return SPX_OK;
    prepare_kernel<<<(K + kPrepWaves - 1) / kPrepWaves, 64 * kPrepWaves, 0, stream>>>(img, H, W, pitch, seeds,
                                                                                     half, K, x_out, norm);
    SPX_LAUNCHED();
    return SPX_OK;
}

int cnn_encode_launch(const float* params, const float* norm, int K, float* t, hipStream_t stream) {
    if (K == 0) return SPX_OK;
    encode_kernel<<<grid1(K), 256, 0, stream>>>(params, norm, K, t);
    SPX_LAUNCHED();
    return SPX_OK;
}

// Inference: out = decoded parameters (K x 7) when norm != NULL, raw network outputs
// otherwise. The fused kernel reads a transposed, K-padded copy of the BF16 weights made in
// the library workspace first.
namespace {
// The transposed, K-padded BF16 weight copy of the fused kernels, in the library workspace.
int infer_weights(const float* theta, const uint16_t* theta_bf, InferWeights& iw, hipStream_t stream) {
    iw = InferWeights{};
    int64_t total = 0;
    for (int l = 0; l < 6; ++l) {
        iw.kp[l] = (kNet[l].kd + 31) / 32 * 32;
        iw.off[l] = total;
        total += (int64_t)padded_cols(l) * iw.kp[l];
    }
    void* ws = nullptr;
    int rc = workspace(sizeof(uint16_t) * (size_t)total, &ws, stream);
    if (rc != SPX_OK) return rc;
    uint16_t* wt = static_cast<uint16_t*>(ws);
    iw.wt = wt;
    iw.theta = theta;
    iw.wbf = theta_bf;
    transpose_weights_kernel<<<grid1(total), 256, 0, stream>>>(theta_bf, iw, wt);
    SPX_LAUNCHED();
    return SPX_OK;
}
}  // namespace

int cnn_forward_launch(const uint16_t* x, int K, const float* theta, const uint16_t* theta_bf, const float* norm,
                       float* out, hipStream_t stream) {
    if (K == 0) return SPX_OK;
    InferWeights iw;
    int rc = infer_weights(theta, theta_bf, iw, stream);
    if (rc != SPX_OK) return rc;
    cnn_infer_kernel<<<(K + kNP - 1) / kNP, 64 * kIW, 0, stream>>>(x, K, iw, norm, out);
    SPX_LAUNCHED();
    return SPX_OK;
}

// Patch preparation + inference + decode in one launch: params (K x 7).
int cnn_propose_launch(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, const float* seeds,
                       const int16_t* half, int K, const float* theta, const uint16_t* theta_bf, float* params,
                       hipStream_t stream) {
    if (K == 0) return SPX_OK;
    InferWeights iw;
    int rc = infer_weights(theta, theta_bf, iw, stream);
    if (rc != SPX_OK) return rc;
    cnn_propose_kernel<<<(K + kNP - 1) / kNP, 64 * kIW, 0, stream>>>(img, H, W, pitch, seeds, half, K, iw, params);
    SPX_LAUNCHED();
    return SPX_OK;
}

// One Adam step on a batch: forward, MSE against targets t (Bn x 7, optional weights w),
// backward through every layer, Adam update of theta (FP32) and theta_bf (BF16). With
// adam_m == NULL only the gradient is formed (copied to grad_out when given).
int cnn_train_step_launch(const uint16_t* x, const float* t, const float* w, int Bn, float* theta,
                          uint16_t* theta_bf, float* adam_m, float* adam_v, int step, float lr, float* loss_out,
                          float* grad_out, hipStream_t stream) {
    CnnBuffers b;
    int rc = cnn_buffers(Bn, true, b, stream);
    if (rc != SPX_OK) return rc;
    rc = forward(x, Bn, theta_bf, theta, b, stream);
    if (rc != SPX_OK) return rc;
    const int n_out = Bn * 7;
    mse_kernel<<<grid1(n_out), 256, 0, stream>>>(b.out, t, w, n_out, 1.0f / (float)n_out, b.dy, b.loss_terms);
    SPX_LAUNCHED();
    // backward
    const float* dA = b.dy;  // gradient w.r.t. the output of layer l (FP32, M x N)
    for (int l = kLayers - 1; l >= 0; --l) {
        const Layer& L = kNet[l];
        const int64_t M = rows_of(l, Bn);
        const int64_t n = M * L.n;
        // dZ (relu mask against the stored post-activation; the head has no ReLU)
        relu_bwd_kernel<<<grid1(n), 256, 0, stream>>>(dA, l == kLayers - 1 ? nullptr : b.act[l], n,
                                                      L.relu ? 1 : 0, b.g_zb, b.g_z);
        SPX_LAUNCHED();
        const int64_t wo = param_offset(l);
        // dW = A^T dZ (K-dim = M), split-K partials then a fixed-order reduction
        const uint16_t* A = L.conv ? b.col[l] : (l == 0 ? x : b.act[l - 1]);
        rc = launch_gemm(A, 1, L.kd, b.g_zb, L.n, 1, b.g_part, nullptr, nullptr, 0, L.kd, L.n, (int)M, kSplits,
                         stream);
        if (rc != SPX_OK) return rc;
        {
            int kps = (int)((M + kSplits - 1) / kSplits);
            kps = ((kps + kBK - 1) / kBK) * kBK;
            const int splits = (int)((M + kps - 1) / kps);
            reduce_splits_kernel<<<grid1((int64_t)L.kd * L.n), 256, 0, stream>>>(b.g_part, splits,
                                                                                (int64_t)L.kd * L.n, b.grad + wo, 0);
            SPX_LAUNCHED();
        }
        colsum_kernel<<<L.n, 256, 0, stream>>>(b.g_z, (int)M, L.n, b.grad + wo + (int64_t)L.kd * L.n);
        SPX_LAUNCHED();
        if (l == 0) break;
        // dA of the layer input: dcol = dZ W^T (M x kd)
        rc = launch_gemm(b.g_zb, L.n, 1, theta_bf + wo, 1, L.n, b.g_col, nullptr, nullptr, 0, (int)M, L.kd, L.n, 1,
                         stream);
        if (rc != SPX_OK) return rc;
        if (L.conv) {
            const int64_t nin = (int64_t)Bn * L.H * L.W * L.Cin;
            col2im_s2_kernel<<<grid1(nin), 256, 0, stream>>>(b.g_col, Bn, L.H, L.W, L.Cin, b.g_act);
            SPX_LAUNCHED();
            dA = b.g_act;
        } else {
            // fc input gradient is dcol itself (Bn x kd), in NHWC order of the previous layer
            SPX_HIP(hipMemcpyAsync(b.g_act, b.g_col, sizeof(float) * M * L.kd, hipMemcpyDeviceToDevice, stream));
            dA = b.g_act;
        }
    }
    if (grad_out) {
        SPX_HIP(hipMemcpyAsync(grad_out, b.grad, sizeof(float) * param_total(), hipMemcpyDeviceToDevice, stream));
    }
    if (adam_m) {
        const float b1 = 0.9f, b2 = 0.999f;
        const float bc1 = 1.0f - powf(b1, (float)step), bc2 = 1.0f - powf(b2, (float)step);
        adam_kernel<<<grid1(param_total()), 256, 0, stream>>>(theta, theta_bf, b.grad, adam_m, adam_v, param_total(),
                                                              lr, b1, b2, 1e-8f, bc1, bc2);
        SPX_LAUNCHED();
    }
    if (loss_out) {
        // loss = sum of the terms (small: Bn x 7), summed on the device by a colsum over 1 column
        colsum_kernel<<<1, 256, 0, stream>>>(b.loss_terms, n_out, 1, b.dy);
        SPX_LAUNCHED();
        SPX_HIP(hipMemcpyAsync(loss_out, b.dy, sizeof(float), hipMemcpyDeviceToHost, stream));
        SPX_HIP(hipStreamSynchronize(stream));
    }
    return SPX_OK;
}

int cnn_to_bf16_launch(const float* theta, int64_t n, uint16_t* theta_bf, hipStream_t stream) {
    to_bf16_kernel<<<grid1(n), 256, 0, stream>>>(theta, n, theta_bf);
    SPX_LAUNCHED();
    return SPX_OK;
}

}  // namespace spx
