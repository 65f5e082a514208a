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

__device__ __forceinline__ uint16_t f2bf(float f) {
    // round to nearest even
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
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

// One wave prepares one patch: gather into the u16 staging area P (side^2 values), ring
// mean and maximum, bilinear resampling to 32 x 32 written as BF16 to xdst (global memory or
// LDS). Returns {B0, A0} through b0 / a0 (every lane).
__device__ __forceinline__ void prep_patch(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, float cx,
                                           float cy, int hs, int lane, uint16_t* P, uint16_t* xdst, float& b0,
                                           float& a0) {
    const int side = 2 * hs + 1;
    const int64_t ox = (int64_t)cx - hs, oy = (int64_t)cy - hs;
    float rsum = 0.0f, rcnt = 0.0f, vmax = -3.0e38f;
    for (int p = lane; p < side * side; p += 64) {
        const int row = p / side, col = p - (p / side) * side;
        const int64_t gx = min(max(ox + col, (int64_t)0), W - 1);
        const int64_t gy = min(max(oy + row, (int64_t)0), H - 1);
        const uint16_t raw = img[gy * pitch + gx];
        const float v = (float)raw;
        P[p] = raw;
        if (row == 0 || row == side - 1 || col == 0 || col == side - 1) {
            rsum += v;
            rcnt += 1.0f;
        }
        vmax = fmaxf(vmax, v);
    }
    for (int off = 32; off >= 1; off >>= 1) {
        rsum += __shfl_xor(rsum, off, 64);
        rcnt += __shfl_xor(rcnt, off, 64);
        vmax = fmaxf(vmax, __shfl_xor(vmax, off, 64));
    }
    const float B0 = rsum / fmaxf(rcnt, 1.0f);
    const float A0 = (vmax - B0) > 1.0f ? (vmax - B0) : 1.0f;
    __builtin_amdgcn_wave_barrier();
    const float scale = (float)side / 32.0f;
    for (int o = lane; o < 1024; o += 64) {
        const int i = o >> 5, j = o & 31;
        // grid point (j, i) at patch coordinate (j + 0.5) * side / 32 - 0.5
        float px = ((float)j + 0.5f) * scale - 0.5f, py = ((float)i + 0.5f) * scale - 0.5f;
        px = fminf(fmaxf(px, 0.0f), (float)(side - 1));
        py = fminf(fmaxf(py, 0.0f), (float)(side - 1));
        const int x0 = min((int)px, side - 2 >= 0 ? side - 2 : 0), y0 = min((int)py, side - 2 >= 0 ? side - 2 : 0);
        const float fx = px - (float)x0, fy = py - (float)y0;
        const int x1 = min(x0 + 1, side - 1), y1 = min(y0 + 1, side - 1);
        const float v = (1.0f - fy) * ((1.0f - fx) * (float)P[y0 * side + x0] + fx * (float)P[y0 * side + x1]) +
                        fy * ((1.0f - fx) * (float)P[y1 * side + x0] + fx * (float)P[y1 * side + x1]);
        xdst[o] = f2bf((v - B0) / A0);
    }
    b0 = B0;
    a0 = A0;
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
    prep_patch(img, H, W, pitch, cx, cy, hs, lane, patch[wave], x_out + (int64_t)f * 1024, B0, A0);
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
// Fused inference: one workgroup (4 waves) takes kIP patches through all six layers, the
// activations staying in LDS (NHWC, BF16), the weights read as MFMA B fragments from a
// transposed, K-padded BF16 copy (Wt[n][k], k padded to a multiple of 16, zeros beyond K).
// Each layer is an implicit GEMM on v_mfma_f32_16x16x16bf16_1k tiles: rows = (patch, output
// pixel), columns = output channels, k = (ky, kx, c). A wave owns one column block and a group
// of up to kGM row blocks, so one B fragment feeds kGM MFMAs. Bias + ReLU + BF16 rounding in
// the epilogue, as the GEMM path. The 128 -> 7 head runs on the VALU in FP32.
typedef short bf16x4s __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
constexpr int kIP = 4;   // patches per workgroup
constexpr int kGM = 8;   // row blocks per wave item (B fragment reuse)

struct InferWeights {
    const uint16_t* wt;      // transposed, K-padded BF16 weights, layers 0..4
    int64_t off[5];          // element offset of layer l in wt
    int kp[5];               // padded K of layer l
    const float* theta;      // FP32 parameters (biases; K x N layout)
    const uint16_t* wbf;     // BF16 K x N weights (head)
};

__global__ void transpose_weights_kernel(const uint16_t* wbf, InferWeights iw, uint16_t* wt) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t base = 0;
#pragma unroll
    for (int l = 0; l < 5; ++l) {
        const int K = kNet[l].kd, N = kNet[l].n, Kp = iw.kp[l];
        const int64_t cnt = (int64_t)N * Kp;
        if (t >= base && t < base + cnt) {
            const int64_t q = t - base;
            const int n = (int)(q / Kp), k = (int)(q - (int64_t)n * Kp);
            wt[iw.off[l] + q] = k < K ? wbf[param_offset_dev(l) + (int64_t)k * N + n] : (uint16_t)0;
        }
        base += cnt;
    }
}

// One conv layer (3x3, stride 2, pad 1) or the first fc layer (as a 1x1 "conv" on the NHWC
// flatten: Hout = 1, Cin = 256) from LDS `in` to LDS `out`.
template <int CIN, int COUT, int HIN, bool FC>
__device__ __forceinline__ void infer_layer(const uint16_t* __restrict__ in, uint16_t* __restrict__ out,
                                            const uint16_t* __restrict__ wt, int kp, const float* __restrict__ bias,
                                            int wave, int lane) {
    constexpr int HOUT = FC ? 1 : HIN / 2;
    constexpr int PIX = HOUT * HOUT;
    constexpr int M = kIP * PIX;
    constexpr int MB = (M + 15) / 16;
    constexpr int NB = COUT / 16;
    constexpr int MG = (MB + kGM - 1) / kGM;
    constexpr int KIN = FC ? CIN : 9 * CIN;
    const int kc_n = kp / 16;
    const int li = lane & 15, lq = lane >> 4;
    for (int item = wave; item < NB * MG; item += 4) {
        const int nb = item % NB, mg = item / NB;
        const int mb0 = mg * kGM;
        const int nmb = MB - mb0 < kGM ? MB - mb0 : kGM;
        f32x4 acc[kGM];
#pragma unroll
        for (int g = 0; g < kGM; ++g) acc[g] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        const uint16_t* wcol = wt + (int64_t)(nb * 16 + li) * kp;
        if constexpr (!FC && CIN >= 16) {
            // tap-major: the row geometry once per item, the tap offset and bounds per tap
            constexpr int CPT = CIN / 16;  // 16-wide k chunks per tap
            int base[kGM], y0[kGM], x0[kGM];
#pragma unroll
            for (int g = 0; g < kGM; ++g) {
                const int row = (mb0 + g) * 16 + li;
                const int pt = row / PIX, pix = row - pt * PIX;
                const int oy = pix / HOUT, ox = pix - (pix / HOUT) * HOUT;
                const bool live = g < nmb && row < M;
                base[g] = pt * HIN * HIN * CIN;
                y0[g] = live ? 2 * oy - 1 : -1000;  // an absent row fails every bounds test
                x0[g] = 2 * ox - 1;
            }
#pragma unroll
            for (int tap = 0; tap < 9; ++tap) {
                const int ky = tap / 3, kx = tap % 3;
#pragma unroll
                for (int cc = 0; cc < CPT; ++cc) {
                    const int c0 = cc * 16 + 4 * lq;
                    const bf16x4s b = *reinterpret_cast<const bf16x4s*>(wcol + (tap * CPT + cc) * 16 + 4 * lq);
#pragma unroll
                    for (int g = 0; g < kGM; ++g) {
                        if (g >= nmb) break;
                        const int iy = y0[g] + ky, ix = x0[g] + kx;
                        bf16x4s a = bf16x4s{0, 0, 0, 0};
                        if ((unsigned)iy < (unsigned)HIN && (unsigned)ix < (unsigned)HIN)
                            a = *reinterpret_cast<const bf16x4s*>(in + base[g] + (iy * HIN + ix) * CIN + c0);
                        acc[g] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[g], 0, 0, 0);
                    }
                }
            }
        } else
        for (int kc = 0; kc < kc_n; ++kc) {
            const int k0 = kc * 16 + 4 * lq;
            const bf16x4s b = *reinterpret_cast<const bf16x4s*>(wcol + k0);
#pragma unroll
            for (int g = 0; g < kGM; ++g) {
                if (g >= nmb) break;
                const int row = (mb0 + g) * 16 + li;
                bf16x4s a = bf16x4s{0, 0, 0, 0};
                if (row < M) {
                    const int pt = row / PIX, pix = row - pt * PIX;
                    if constexpr (FC) {
                        a = *reinterpret_cast<const bf16x4s*>(in + pt * CIN + k0);
                    } else if constexpr (CIN >= 4) {
                        const int oy = pix / HOUT, ox = pix - oy * HOUT;
                        const int tap = k0 / CIN, c0 = k0 - tap * CIN;
                        const int iy = 2 * oy + tap / 3 - 1, ix = 2 * ox + tap % 3 - 1;
                        if (k0 < KIN && iy >= 0 && iy < HIN && ix >= 0 && ix < HIN)
                            a = *reinterpret_cast<const bf16x4s*>(in + ((pt * HIN + iy) * HIN + ix) * CIN + c0);
                    } else {
                        const int oy = pix / HOUT, ox = pix - oy * HOUT;
#pragma unroll
                        for (int t = 0; t < 4; ++t) {
                            const int k = k0 + t;  // CIN == 1: k = tap
                            const int iy = 2 * oy + k / 3 - 1, ix = 2 * ox + k % 3 - 1;
                            if (k < KIN && iy >= 0 && iy < HIN && ix >= 0 && ix < HIN)
                                a[t] = (short)in[(pt * HIN + iy) * HIN + ix];
                        }
                    }
                }
                acc[g] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[g], 0, 0, 0);
            }
        }
        // D(i, j): i = 4 (lane / 16) + r, j = lane % 16
        const int n = nb * 16 + li;
        const float bn = bias[n];
#pragma unroll
        for (int g = 0; g < kGM; ++g) {
            if (g >= nmb) break;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = (mb0 + g) * 16 + 4 * lq + r;
                if (row < M) out[row * COUT + n] = f2bf(fmaxf(acc[g][r] + bn, 0.0f));
            }
        }
    }
}

// Layers 1-5 on the patches in bufB (NHWC BF16, 32 x 32 x 1 each), the head, then the
// outputs of patches p0 .. p0 + kIP - 1 (< K): decoded with nrm[patch][5] when `decode`, raw
// otherwise.
__device__ __forceinline__ void infer_tail(uint16_t* bufA, uint16_t* bufB, float* head, const InferWeights& iw,
                                           int K, int p0, bool decode, const float* nrm, float* out) {
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const float* th = iw.theta;
    infer_layer<1, 16, 32, false>(bufB, bufA, iw.wt + iw.off[0], iw.kp[0], th + param_offset_dev(0) + 9 * 16, wave,
                                  lane);
    __syncthreads();
    infer_layer<16, 32, 16, false>(bufA, bufB, iw.wt + iw.off[1], iw.kp[1], th + param_offset_dev(1) + 144 * 32,
                                   wave, lane);
    __syncthreads();
    infer_layer<32, 64, 8, false>(bufB, bufA, iw.wt + iw.off[2], iw.kp[2], th + param_offset_dev(2) + 288 * 64,
                                  wave, lane);
    __syncthreads();
    infer_layer<64, 64, 4, false>(bufA, bufB, iw.wt + iw.off[3], iw.kp[3], th + param_offset_dev(3) + 576 * 64,
                                  wave, lane);
    __syncthreads();
    infer_layer<256, 128, 1, true>(bufB, bufA, iw.wt + iw.off[4], iw.kp[4], th + param_offset_dev(4) + 256 * 128,
                                   wave, lane);
    __syncthreads();
    // head 128 -> 7 (FP32 on the VALU): thread (patch, output)
    if (tid < kIP * 7) {
        const int pt = tid / 7, j = tid - pt * 7;
        const int64_t wo = param_offset_dev(5);
        float acc = 0.0f;
        for (int k = 0; k < 128; ++k) acc = fmaf(bf2f(bufA[pt * 128 + k]), bf2f(iw.wbf[wo + k * 7 + j]), acc);
        head[pt * 8 + j] = acc + th[wo + 128 * 7 + j];
    }
    __syncthreads();
    if (tid < kIP && p0 + tid < K) {
        const int64_t k = p0 + tid;
        if (decode) {
            decode_one(head + tid * 8, nrm + 5 * tid, out + 7 * k);
        } else {
#pragma unroll
            for (int j = 0; j < 7; ++j) out[7 * k + j] = head[tid * 8 + j];
        }
    }
}

__global__ __launch_bounds__(256) void cnn_infer_kernel(const uint16_t* __restrict__ x, int K, InferWeights iw,
                                                        const float* __restrict__ norm, float* __restrict__ out) {
    __shared__ __attribute__((aligned(16))) uint16_t bufA[kIP * 16 * 16 * 16];  // L1 out, L3 out, fc1 out
    __shared__ __attribute__((aligned(16))) uint16_t bufB[kIP * 8 * 8 * 32];    // input, L2 out, L4 out
    __shared__ float head[kIP * 8];
    const int tid = threadIdx.x;
    const int p0 = blockIdx.x * kIP;
    // input patches (2 KB each) -> bufB, 16 bytes per thread per step; missing patches read 0
    for (int q = tid; q < kIP * 128; q += 256) {
        const int pt = q >> 7, e = (q & 127) * 8;
        uint4 v = make_uint4(0, 0, 0, 0);
        if (p0 + pt < K) v = *reinterpret_cast<const uint4*>(x + (int64_t)(p0 + pt) * 1024 + e);
        *reinterpret_cast<uint4*>(bufB + pt * 1024 + e) = v;
    }
    __syncthreads();
    infer_tail(bufA, bufB, head, iw, K, p0, norm != nullptr, norm ? norm + 5 * (int64_t)p0 : nullptr, out);
}

// Proposal in one launch: each wave prepares its patch straight into LDS (the u16 staging
// area aliases bufA, free until the first layer writes it), then the network and the decode.
static_assert(kIP * kMaxSide * kMaxSide <= kIP * 16 * 16 * 16, "patch staging must fit in bufA");
static_assert(kIP == 256 / 64, "one wave per patch");

__global__ __launch_bounds__(256) void cnn_propose_kernel(const uint16_t* __restrict__ img, int64_t H, int64_t W,
                                                          int64_t pitch, const float* __restrict__ seeds,
                                                          const int16_t* __restrict__ half, int K, InferWeights iw,
                                                          float* __restrict__ out) {
    __shared__ __attribute__((aligned(16))) uint16_t bufA[kIP * 16 * 16 * 16];
    __shared__ __attribute__((aligned(16))) uint16_t bufB[kIP * 8 * 8 * 32];
    __shared__ float head[kIP * 8];
    __shared__ float s_norm[kIP * 5];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int p0 = blockIdx.x * kIP;
    const int f = p0 + wave;
    if (f < K) {
        const int hs = __builtin_amdgcn_readfirstlane(max(1, min((int)half[f], (kMaxSide - 1) / 2)));
        const float cx = rintf(seeds[2 * (int64_t)f]), cy = rintf(seeds[2 * (int64_t)f + 1]);
        float B0, A0;
        prep_patch(img, H, W, pitch, cx, cy, hs, lane, bufA + wave * kMaxSide * kMaxSide, bufB + wave * 1024, B0, A0);
        if (lane == 0) {
            s_norm[5 * wave + 0] = cx;
            s_norm[5 * wave + 1] = cy;
            s_norm[5 * wave + 2] = (float)(2 * hs + 1);
            s_norm[5 * wave + 3] = B0;
            s_norm[5 * wave + 4] = A0;
        }
    } else {
        for (int o = lane; o < 1024; o += 64) bufB[wave * 1024 + o] = 0;
    }
    __syncthreads();
    infer_tail(bufA, bufB, head, iw, K, p0, true, s_norm, out);
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
    for (int l = 0; l < 5; ++l) {
        iw.kp[l] = (kNet[l].kd + 15) / 16 * 16;
        iw.off[l] = total;
        total += (int64_t)kNet[l].n * iw.kp[l];
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
    cnn_infer_kernel<<<(K + kIP - 1) / kIP, 256, 0, stream>>>(x, K, iw, norm, out);
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
    cnn_propose_kernel<<<(K + kIP - 1) / kIP, 256, 0, stream>>>(img, H, W, pitch, seeds, half, K, iw, params);
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
