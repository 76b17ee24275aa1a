// Field MLP 32 -> 64 -> 64 -> n_out (n_out <= 16) with SELU hidden activations, on MFMA.
//
// Reference head: Linear -> SELU -> Linear -> SELU -> Linear (models/ingp_color.py:141-151),
// weights row-major [out, in] as torch.nn.Linear. Each wave owns tiles of 16 query points; every
// layer is a chain of 16x16 MFMA tiles with FP32 accumulation:
//   precision 0: v_mfma_f32_16x16x4_f32 (FP32 operands),
//   precision 1: v_mfma_f32_16x16x16_bf16 (BF16 operands, rounded to nearest even).
// Operand fragment layouts (64 lanes, q = lane / 16, c = lane % 16):
//   A[i][k]: i = c, k = 4*step + q (f32) or k = 16*step + 4*q + e, e in 0..3 (bf16)
//   B[k][j]: j = c, k as for A
//   D[i][j]: j = c, i = 4*q + r for accumulator register r in 0..3
// Weights and per-wave activation tiles live in LDS (rows padded by one float against bank
// conflicts). The backward recomputes the forward per tile (as the reference's checkpointed
// field), forms dX, and accumulates weight gradients gW = dH^T A as MFMA tiles with the 16
// points of a tile as the contraction dimension, in registers across all tiles of the wave;
// each wave adds its totals to global memory with one float atomic per element at the end.
#include "common.h"

namespace rmesh {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef short s4 __attribute__((ext_vector_type(4)));

constexpr int kIn = 32, kHid = 64, kOutMax = 16;
constexpr int kWaves = 2, kThreads = 64 * kWaves;
constexpr int SW1 = kIn + 1, SW2 = kHid + 1, SW3 = kHid + 1;   // LDS row strides (floats)
constexpr int SX = kIn + 1, SA = kHid + 1, SY = kOutMax + 1;

constexpr float kSeluAlpha = 1.6732632423543772f;
constexpr float kSeluScale = 1.0507009873554805f;

// Orders this wave's LDS writes before its following LDS reads by other lanes.
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ inline float selu(float x) { return kSeluScale * (x > 0.f ? x : kSeluAlpha * expm1f(x)); }
__device__ inline float dselu(float x) { return kSeluScale * (x > 0.f ? 1.f : kSeluAlpha * __expf(x)); }

__device__ inline short to_bf16(float f) {  // round to nearest even
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

// Fragment loader: element (row, col) of an LDS matrix with row stride s, scaled access pattern
// chosen by the caller through a functor get(k, c) returning operand[k-index][c].
template <int P>
struct MM;

template <>
struct MM<0> {  // FP32 operands, K step 4
    static constexpr int KS = 4;
    template <typename GA, typename GB>
    __device__ static f4 run(int K, GA ga, GB gb, f4 c) {
        const int lane = threadIdx.x & 63, q = lane >> 4, j = lane & 15;
        for (int s = 0; s < K / 4; ++s) {
            int k = 4 * s + q;
            c = __builtin_amdgcn_mfma_f32_16x16x4f32(ga(j, k), gb(k, j), c, 0, 0, 0);
        }
        return c;
    }
};

template <>
struct MM<1> {  // BF16 operands, K step 16
    static constexpr int KS = 16;
    template <typename GA, typename GB>
    __device__ static f4 run(int K, GA ga, GB gb, f4 c) {
        const int lane = threadIdx.x & 63, q = lane >> 4, j = lane & 15;
        for (int s = 0; s < K / 16; ++s) {
            s4 a, b;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                int k = 16 * s + 4 * q + e;
                a[e] = to_bf16(ga(j, k));
                b[e] = to_bf16(gb(k, j));
            }
            c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
        }
        return c;
    }
};

struct Smem {
    float w1[kHid * SW1];
    float w2[kHid * SW2];
    float w3[kOutMax * SW3];
    float b1[kHid], b2[kHid], b3[kOutMax];
};

struct WaveBuf {
    float x[16 * SX];
    float a1[16 * SA];
    float a2[16 * SA];     // second hidden activations; reused for dH1 in the backward
    float dh2[16 * SA];
    float dy[16 * SY];
};

__device__ void load_weights(Smem& S, const float* w1, const float* b1, const float* w2, const float* b2,
                             const float* w3, const float* b3, int n_out) {
    for (int i = threadIdx.x; i < kHid * kIn; i += blockDim.x) S.w1[(i / kIn) * SW1 + i % kIn] = w1[i];
    for (int i = threadIdx.x; i < kHid * kHid; i += blockDim.x) S.w2[(i / kHid) * SW2 + i % kHid] = w2[i];
    for (int i = threadIdx.x; i < kOutMax * kHid; i += blockDim.x) {
        int o = i / kHid;
        S.w3[o * SW3 + i % kHid] = o < n_out ? w3[i] : 0.f;
    }
    for (int i = threadIdx.x; i < kHid; i += blockDim.x) {
        S.b1[i] = b1[i];
        S.b2[i] = b2[i];
    }
    for (int i = threadIdx.x; i < kOutMax; i += blockDim.x) S.b3[i] = i < n_out ? b3[i] : 0.f;
}

// Stage the tile's inputs x[16, 32] (rows beyond n are zero).
__device__ inline void load_x(WaveBuf& B, const float* __restrict__ x, int64_t row0, int64_t n) {
    const int lane = threadIdx.x & 63;
    for (int i = lane; i < 16 * kIn; i += 64) {
        int r = i / kIn, k = i % kIn;
        int64_t g = row0 + r;
        B.x[r * SX + k] = g < n ? x[g * kIn + k] : 0.f;
    }
}

// Forward of one 16-point tile. Leaves pre-activations h1, h2 (D layout) in registers and the
// activations in B.a1 / B.a2; returns the output tile (D layout).
template <int P>
__device__ inline f4 tile_forward(const Smem& S, WaveBuf& B, f4 h1[4], f4 h2[4]) {
    const int lane = threadIdx.x & 63, q = lane >> 4, j = lane & 15;
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        f4 c = {0.f, 0.f, 0.f, 0.f};
        c = MM<P>::run(kIn, [&](int i, int k) { return B.x[i * SX + k]; },
                       [&](int k, int jj) { return S.w1[(16 * nt + jj) * SW1 + k]; }, c);
        float bias = S.b1[16 * nt + j];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            c[r] += bias;
            B.a1[(4 * q + r) * SA + 16 * nt + j] = selu(c[r]);
        }
        h1[nt] = c;
    }
    wave_sync();
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        f4 c = {0.f, 0.f, 0.f, 0.f};
        c = MM<P>::run(kHid, [&](int i, int k) { return B.a1[i * SA + k]; },
                       [&](int k, int jj) { return S.w2[(16 * nt + jj) * SW2 + k]; }, c);
        float bias = S.b2[16 * nt + j];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            c[r] += bias;
            B.a2[(4 * q + r) * SA + 16 * nt + j] = selu(c[r]);
        }
        h2[nt] = c;
    }
    wave_sync();
    f4 y = {0.f, 0.f, 0.f, 0.f};
    y = MM<P>::run(kHid, [&](int i, int k) { return B.a2[i * SA + k]; },
                   [&](int k, int jj) { return S.w3[jj * SW3 + k]; }, y);
    float bias = S.b3[j];
#pragma unroll
    for (int r = 0; r < 4; ++r) y[r] += bias;
    return y;
}

template <int P>
__global__ __launch_bounds__(kThreads) void mlp_fwd_kernel(const float* __restrict__ x, int64_t n,
                                                           const float* w1, const float* b1, const float* w2,
                                                           const float* b2, const float* w3, const float* b3,
                                                           int n_out, float* __restrict__ y) {
    __shared__ Smem S;
    __shared__ WaveBuf WB[kWaves];
    load_weights(S, w1, b1, w2, b2, w3, b3, n_out);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, q = lane >> 4, j = lane & 15;
    WaveBuf& B = WB[wave];
    const int64_t ntiles = (n + 15) / 16;
    for (int64_t t = (int64_t)blockIdx.x * kWaves + wave; t < ntiles; t += (int64_t)gridDim.x * kWaves) {
        load_x(B, x, t * 16, n);
        wave_sync();
        f4 h1[4], h2[4];
        f4 out = tile_forward<P>(S, B, h1, h2);
        if (j < n_out) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                int64_t row = t * 16 + 4 * q + r;
                if (row < n) y[row * n_out + j] = out[r];
            }
        }
        wave_sync();
    }
}

template <int P>
__global__ __launch_bounds__(kThreads) void mlp_bwd_kernel(
    const float* __restrict__ x, int64_t n, const float* w1, const float* b1, const float* w2, const float* b2,
    const float* w3, const float* b3, int n_out, const float* __restrict__ gy, float* __restrict__ gx,
    float* gw1, float* gb1, float* gw2, float* gb2, float* gw3, float* gb3) {
    __shared__ Smem S;
    __shared__ WaveBuf WB[kWaves];
    load_weights(S, w1, b1, w2, b2, w3, b3, n_out);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, q = lane >> 4, j = lane & 15;
    WaveBuf& B = WB[wave];
    // weight-gradient accumulators (D layout): gW3 [16 x 64], gW2 [64 x 64], gW1 [64 x 32]
    f4 g3[4], g2[4][4], g1[4][2];
    float gb1p[4] = {0, 0, 0, 0}, gb2p[4] = {0, 0, 0, 0}, gb3p = 0.f;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        g3[a] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int b = 0; b < 4; ++b) g2[a][b] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int b = 0; b < 2; ++b) g1[a][b] = f4{0.f, 0.f, 0.f, 0.f};
    }
    const int64_t ntiles = (n + 15) / 16;
    for (int64_t t = (int64_t)blockIdx.x * kWaves + wave; t < ntiles; t += (int64_t)gridDim.x * kWaves) {
        load_x(B, x, t * 16, n);
        for (int i = lane; i < 16 * kOutMax; i += 64) {
            int r = i / kOutMax, o = i % kOutMax;
            int64_t g = t * 16 + r;
            B.dy[r * SY + o] = (g < n && o < n_out) ? gy[g * n_out + o] : 0.f;
        }
        wave_sync();
        f4 h1[4], h2[4];
        tile_forward<P>(S, B, h1, h2);
        // dH2 = (dY W3) * selu'(H2)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            f4 c = {0.f, 0.f, 0.f, 0.f};
            c = MM<P>::run(kOutMax, [&](int i, int k) { return B.dy[i * SY + k]; },
                           [&](int k, int jj) { return S.w3[k * SW3 + 16 * nt + jj]; }, c);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                float d = c[r] * dselu(h2[nt][r]);
                B.dh2[(4 * q + r) * SA + 16 * nt + j] = d;
                gb2p[nt] += d;
            }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) gb3p += B.dy[(4 * q + r) * SY + j];
        wave_sync();
        // gW3 += dY^T A2 (contract over the tile's points)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
            g3[nt] = MM<P>::run(16, [&](int i, int k) { return B.dy[k * SY + i]; },
                                [&](int k, int jj) { return B.a2[k * SA + 16 * nt + jj]; }, g3[nt]);
        wave_sync();
        // dH1 = (dH2 W2) * selu'(H1), stored over A2
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            f4 c = {0.f, 0.f, 0.f, 0.f};
            c = MM<P>::run(kHid, [&](int i, int k) { return B.dh2[i * SA + k]; },
                           [&](int k, int jj) { return S.w2[k * SW2 + 16 * nt + jj]; }, c);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                float d = c[r] * dselu(h1[nt][r]);
                B.a2[(4 * q + r) * SA + 16 * nt + j] = d;
                gb1p[nt] += d;
            }
        }
        // gW2 += dH2^T A1
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int nt = 0; nt < 4; ++nt)
                g2[mt][nt] = MM<P>::run(16, [&](int i, int k) { return B.dh2[k * SA + 16 * mt + i]; },
                                        [&](int k, int jj) { return B.a1[k * SA + 16 * nt + jj]; }, g2[mt][nt]);
        wave_sync();
        // dX = dH1 W1
        if (gx) {
#pragma unroll
            for (int nt = 0; nt < 2; ++nt) {
                f4 c = {0.f, 0.f, 0.f, 0.f};
                c = MM<P>::run(kHid, [&](int i, int k) { return B.a2[i * SA + k]; },
                               [&](int k, int jj) { return S.w1[k * SW1 + 16 * nt + jj]; }, c);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    int64_t row = t * 16 + 4 * q + r;
                    if (row < n) gx[row * kIn + 16 * nt + j] = c[r];
                }
            }
        }
        // gW1 += dH1^T X
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int nt = 0; nt < 2; ++nt)
                g1[mt][nt] = MM<P>::run(16, [&](int i, int k) { return B.a2[k * SA + 16 * mt + i]; },
                                        [&](int k, int jj) { return B.x[k * SX + 16 * nt + jj]; }, g1[mt][nt]);
        wave_sync();
    }
    // flush: D layout element (4q + r, j) of tile (mt, nt)
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            int o = 4 * q + r;
            if (o < n_out && g3[nt][r] != 0.f) unsafeAtomicAdd(gw3 + o * kHid + 16 * nt + j, g3[nt][r]);
        }
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                unsafeAtomicAdd(gw2 + (16 * mt + 4 * q + r) * kHid + 16 * nt + j, g2[mt][nt][r]);
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int nt = 0; nt < 2; ++nt)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                unsafeAtomicAdd(gw1 + (16 * mt + 4 * q + r) * kIn + 16 * nt + j, g1[mt][nt][r]);
    // biases: lanes j, j+16, j+32, j+48 hold partial sums of column j
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        float v1 = gb1p[nt], v2 = gb2p[nt];
        v1 += __shfl_xor(v1, 16, 64); v1 += __shfl_xor(v1, 32, 64);
        v2 += __shfl_xor(v2, 16, 64); v2 += __shfl_xor(v2, 32, 64);
        if (q == 0) {
            unsafeAtomicAdd(gb1 + 16 * nt + j, v1);
            unsafeAtomicAdd(gb2 + 16 * nt + j, v2);
        }
    }
    gb3p += __shfl_xor(gb3p, 16, 64);
    gb3p += __shfl_xor(gb3p, 32, 64);
    if (q == 0 && j < n_out) unsafeAtomicAdd(gb3 + j, gb3p);
}

inline unsigned grid_for(int64_t n, int cap) {
    int64_t tiles = (n + 15) / 16;
    int64_t blocks = (tiles + kWaves - 1) / kWaves;
    return (unsigned)(blocks < cap ? blocks : cap);
}

}  // namespace
}  // namespace rmesh

using namespace rmesh;

extern "C" int rm_mlp_fwd(const float* x_dev, int64_t n, const float* w1, const float* b1, const float* w2,
                          const float* b2, const float* w3, const float* b3, int32_t n_out, int32_t precision,
                          float* y_dev, rm_stream_t stream) {
    RM_CHECK_ARG(n >= 0, "rm_mlp_fwd: negative n");
    RM_CHECK_ARG(n_out >= 1 && n_out <= kOutMax, "rm_mlp_fwd: n_out must be in [1, 16]");
    RM_CHECK_ARG(precision == 0 || precision == 1, "rm_mlp_fwd: precision must be 0 (fp32) or 1 (bf16)");
    if (n == 0) return RM_OK;
    RM_CHECK_ARG(x_dev && w1 && b1 && w2 && b2 && w3 && b3 && y_dev, "rm_mlp_fwd: null buffer");
    dim3 grid(grid_for(n, 4096)), block(kThreads);
    if (precision == 0)
        hipLaunchKernelGGL(mlp_fwd_kernel<0>, grid, block, 0, as_stream(stream), x_dev, n, w1, b1, w2, b2, w3, b3,
                           n_out, y_dev);
    else
        hipLaunchKernelGGL(mlp_fwd_kernel<1>, grid, block, 0, as_stream(stream), x_dev, n, w1, b1, w2, b2, w3, b3,
                           n_out, y_dev);
    RM_CHECK_LAUNCH("mlp_fwd_kernel");
    return RM_OK;
}

extern "C" int rm_mlp_bwd(const float* x_dev, int64_t n, const float* w1, const float* b1, const float* w2,
                          const float* b2, const float* w3, const float* b3, int32_t n_out, int32_t precision,
                          const float* grad_y_dev, float* grad_x_dev, float* gw1, float* gb1, float* gw2,
                          float* gb2, float* gw3, float* gb3, rm_stream_t stream) {
    RM_CHECK_ARG(n >= 0, "rm_mlp_bwd: negative n");
    RM_CHECK_ARG(n_out >= 1 && n_out <= kOutMax, "rm_mlp_bwd: n_out must be in [1, 16]");
    RM_CHECK_ARG(precision == 0 || precision == 1, "rm_mlp_bwd: precision must be 0 (fp32) or 1 (bf16)");
    if (n == 0) return RM_OK;
    RM_CHECK_ARG(x_dev && w1 && b1 && w2 && b2 && w3 && b3 && grad_y_dev && gw1 && gb1 && gw2 && gb2 && gw3 && gb3,
                 "rm_mlp_bwd: null buffer");
    dim3 grid(grid_for(n, 1024)), block(kThreads);
    if (precision == 0)
        hipLaunchKernelGGL(mlp_bwd_kernel<0>, grid, block, 0, as_stream(stream), x_dev, n, w1, b1, w2, b2, w3, b3,
                           n_out, grad_y_dev, grad_x_dev, gw1, gb1, gw2, gb2, gw3, gb3);
    else
        hipLaunchKernelGGL(mlp_bwd_kernel<1>, grid, block, 0, as_stream(stream), x_dev, n, w1, b1, w2, b2, w3, b3,
                           n_out, grad_y_dev, grad_x_dev, gw1, gb1, gw2, gb2, gw3, gb3);
    RM_CHECK_LAUNCH("mlp_bwd_kernel");
    return RM_OK;
}
