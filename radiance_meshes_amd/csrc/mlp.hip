// Field MLP heads n_in -> 64 -> 64 -> n_out (n_in 32 or 64, n_out <= 48) with SELU hidden
// activations, on MFMA.
//
// Reference head: Linear -> SELU -> Linear -> SELU -> Linear (models/ingp_color.py:141-151),
// weights row-major [out, in] as torch.nn.Linear. Every product is a chain of 16x16 MFMA tiles
// with FP32 accumulation:
//   precision 0: v_mfma_f32_16x16x4_f32 (FP32 operands),
//   precision 1: v_mfma_f32_16x16x16_bf16 (BF16 operands, rounded to nearest even).
// Operand fragments (64 lanes, q = lane / 16, c = lane % 16):
//   A[i][k]: i = c, k = 4*step + q (f32) or k = 16*step + 4*q + e, e in 0..3 (bf16)
//   B[k][j]: j = c, k as for A
//   D[i][j]: j = c, i = 4*q + r for accumulator register r in 0..3
//
// Layers are computed transposed, out^T[unit][point] = W[unit][k] x act^T[k][point]: the weights
// are the A operand (read from LDS), the 16 points of a tile are the columns. A layer's result then
// sits in the lanes as D[unit 16*ot + 4q + r][point c], and feeding it to the next layer only needs
// the contraction order of k chosen to match: at step s the lane supplies unit
// 16*(s/4) + 4q + (s%4) (f32) or 16*s + 4q + e (bf16), which is exactly its own register r (or e)
// of block s/4 (or s). Activations never leave registers; the input rows come from global memory
// in that layout (two float4 per lane) and dX leaves the same way.
// The backward recomputes the forward per tile (as the reference's checkpointed field) keeping
// the hidden pre-activations (selu' is taken from them), chains dY -> dH2 -> dH1 -> dX through
// the transposed weights, and forms the weight gradients
// gW = dH act^T as MFMA tiles contracted over the tile's points; those need the point index on
// the lanes' k, so dH and act blocks are transposed through a 16x17 LDS block per 16 units.
// Weight-gradient tiles are split among the workgroup's waves (each owns one hidden block of 16
// units and contracts it over the four tiles the waves work on together), accumulate in registers
// across the workgroup's steps and are added to global memory with one float atomic per element.
#include <atomic>

#include "common.h"

namespace rmesh {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef short s4 __attribute__((ext_vector_type(4)));

// Shapes: KI input k-tiles of 16 (n_in = 16 KI: 32 for the training-step field, 64 for the
// reference model's 8 x 8 encoding), hidden width 64, KO output blocks of 16 (n_out <= 16 KO).
constexpr int kHid = 64;
constexpr int kMaxDevices = 64;
constexpr int kWaves = 4, kThreads = 64 * kWaves;
constexpr int SW2 = kHid + 1, SW3 = kHid + 1;                  // LDS row strides (floats)
constexpr int kTS = 17, kTBlock = 16 * kTS;                    // transpose block (16 x 17 floats)
constexpr int kTBlocks = 8;                                    // transpose blocks per wave
template <int KI, int KO>
struct Shape {
    static constexpr int IN = 16 * KI, OUT = 16 * KO, SW1 = IN + 1;
    static constexpr int NGRAD = OUT * kHid + kHid * kHid + kHid * IN + kHid + kHid + OUT;
    static constexpr int TB = kWaves * kTBlocks * kTBlock > NGRAD ? kWaves * kTBlocks * kTBlock : NGRAD;
};

constexpr float kSeluAlpha = 1.6732632423543772f;
constexpr float kSeluScale = 1.0507009873554805f;

// Orders this wave's LDS writes before its following LDS reads by other lanes.
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// e^x - 1 for x <= 0: above -0.25, x + x^2 q(x) with q a degree-3 polynomial fitted to
// (expm1(x)/x - 1)/x at Chebyshev nodes of [-0.25, 0] (within 1e-7 relative in float32, so small
// negative inputs keep their relative precision); below it the hardware exponential minus one
// (the result is at least 0.22 in magnitude there, so the exponential's ~1e-7 error stays ~3e-7
// relative). Six instructions beside the exponential, where expm1f takes 24; this matters here
// (32 SELUs per lane per 16-query tile next to the MFMAs).
// expm1_from takes the exponential e = e^x already formed.
__device__ inline float expm1_from(float x, float e) {
    const float q = 0.49999952f + x * (0.16665113f + x * (0.041491002f + x * 0.0075143012f));
    const float p = x + (x * x) * q;
    return x >= -0.25f ? p : e - 1.0f;
}
__device__ inline float expm1_nonpos(float x) { return expm1_from(x, __expf(x)); }
__device__ inline float selu(float x) { return kSeluScale * (x > 0.f ? x : kSeluAlpha * expm1_nonpos(x)); }
// selu'(h) from the pre-activation, scale alpha e^h for h <= 0 as torch's SELU backward takes it.
// (Taking it from a = selu(h) as a + scale alpha would cancel: for h << 0 the float a keeps e^h
// only to the absolute precision of scale alpha, 1e-3 relative at h = -10.)
// selu(h) and selu'(h) of a block from one exponential per element: selu' is scale alpha e^h for
// h <= 0, as torch's SELU backward takes it from the pre-activation. (From a = selu(h) as
// a + scale alpha it would cancel: for h << 0 the float a holds e^h only to the absolute precision
// of scale alpha, 1e-3 relative at h = -10.)
__device__ inline void selu_pair(f4 h, f4& a, f4& d) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const float x = h[r];
        const float e = __expf(x);
        const float em1 = expm1_from(x, e);
        a[r] = x > 0.f ? kSeluScale * x : kSeluScale * (kSeluAlpha * em1);
        d[r] = x > 0.f ? kSeluScale : (kSeluScale * kSeluAlpha) * e;
    }
}

__device__ inline short to_bf16(float f) {  // round to nearest even
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

__device__ inline f4 mfma_step(float a, float b, f4 d) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, d, 0, 0, 0); }
__device__ inline f4 mfma_step(const float (&a)[4], const float (&b)[4], f4 d) {
    s4 av, bv;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        av[e] = to_bf16(a[e]);
        bv[e] = to_bf16(b[e]);
    }
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, d, 0, 0, 0);
}

// One output block of a transposed layer: D[unit 16*ot + 4q + r][point c] = sum_k A(16*ot + c, k)
// act[k][point c], with act in the chain layout (lane (q, c) holds unit 16*kt + 4q + r of point c
// in act[kt][r]) and wa(i, k) the weight of output row i (relative to the block) and input k.
template <int P, int KT, typename WA>
__device__ inline f4 chain(const f4 (&act)[KT], WA wa) {
    const int lane = threadIdx.x & 63, q = lane >> 4, c = lane & 15;
    f4 d = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kt = 0; kt < KT; ++kt) {
        if constexpr (P == 0) {
#pragma unroll
            for (int r = 0; r < 4; ++r) d = mfma_step(wa(c, 16 * kt + 4 * q + r), act[kt][r], d);
        } else {
            float a[4], b[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                a[e] = wa(c, 16 * kt + 4 * q + e);
                b[e] = act[kt][e];
            }
            d = mfma_step(a, b, d);
        }
    }
    return d;
}

// D[unit 16*ot + 4q + r][unit' 16*kt + c] += sum over the tile's 16 points of a[.][p] b[.][p],
// with a and b as point fragments: lane (q, c) holds unit c of its block at points 4q + s.
template <int P>
__device__ inline f4 outer(const float (&a)[4], const float (&b)[4], f4 d) {
    if constexpr (P == 0) {
#pragma unroll
        for (int s = 0; s < 4; ++s) d = mfma_step(a[s], b[s], d);
    } else {
        d = mfma_step(a, b, d);
    }
    return d;
}

// Chain layout -> point fragment of one 16-unit block through LDS block tb.
__device__ inline void put_block(float* tb, f4 v) {
    const int lane = threadIdx.x & 63, q = lane >> 4, c = lane & 15;
#pragma unroll
    for (int r = 0; r < 4; ++r) tb[(4 * q + r) * kTS + c] = v[r];
}
__device__ inline void get_block(const float* tb, float (&out)[4]) {
    const int lane = threadIdx.x & 63, q = lane >> 4, c = lane & 15;
#pragma unroll
    for (int s = 0; s < 4; ++s) out[s] = tb[c * kTS + 4 * q + s];
}

template <int KI, int KO>
struct Smem {
    using Sh = Shape<KI, KO>;
    float w1[kHid * Sh::SW1];
    float w2[kHid * SW2];
    float w3[Sh::OUT * SW3];
    float b1[kHid], b2[kHid], b3[Sh::OUT];
};

template <int KI, int KO>
__device__ void load_weights(Smem<KI, KO>& S, const float* w1, const float* b1, const float* w2, const float* b2,
                             const float* w3, const float* b3, int n_out) {
    using Sh = Shape<KI, KO>;
    for (int i = threadIdx.x; i < kHid * Sh::IN; i += blockDim.x) S.w1[(i / Sh::IN) * Sh::SW1 + i % Sh::IN] = w1[i];
    for (int i = threadIdx.x; i < kHid * kHid; i += blockDim.x) S.w2[(i / kHid) * SW2 + i % kHid] = w2[i];
    for (int i = threadIdx.x; i < Sh::OUT * kHid; i += blockDim.x) {
        int o = i / kHid;
        S.w3[o * SW3 + i % kHid] = o < n_out ? w3[i] : 0.f;
    }
    for (int i = threadIdx.x; i < kHid; i += blockDim.x) {
        S.b1[i] = b1[i];
        S.b2[i] = b2[i];
    }
    for (int i = threadIdx.x; i < Sh::OUT; i += blockDim.x) S.b3[i] = i < n_out ? b3[i] : 0.f;
}

// Input rows of the tile in the chain layout: xin[kt][r] = x[point c][16 kt + 4q + r].
template <int KI>
__device__ inline void load_x_chain(const float* __restrict__ x, int64_t p, int64_t n, f4 (&xin)[KI]) {
    const int lane = threadIdx.x & 63, q = lane >> 4;
#pragma unroll
    for (int kt = 0; kt < KI; ++kt) {
        if (p < n) {
            const float4 v = *reinterpret_cast<const float4*>(x + p * (16 * KI) + 16 * kt + 4 * q);
            xin[kt] = f4{v.x, v.y, v.z, v.w};
        } else {
            xin[kt] = f4{0.f, 0.f, 0.f, 0.f};
        }
    }
}

// Forward of one tile in the chain layout and, with OUT, the output blocks. h1, h2 receive the
// hidden activations selu(H), or with PRE the pre-activations H (the backward takes selu' from H
// and recomputes selu(H) where it needs the activation).
template <int P, int KI, int KO, bool OUT, bool PRE = false>
__device__ inline void tile_forward(const Smem<KI, KO>& S, const f4 (&xin)[KI], f4 (&h1)[4], f4 (&h2)[4],
                                    f4 (&y)[KO]) {
    constexpr int SW1 = Shape<KI, KO>::SW1;
    const int lane = threadIdx.x & 63, q = lane >> 4;
    f4 a1[4];
#pragma unroll
    for (int ot = 0; ot < 4; ++ot) {
        f4 d = chain<P, KI>(xin, [&](int i, int k) { return S.w1[(16 * ot + i) * SW1 + k]; });
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const float h = d[r] + S.b1[16 * ot + 4 * q + r];
            a1[ot][r] = selu(h);
            h1[ot][r] = PRE ? h : a1[ot][r];
        }
    }
    f4 a2[4];
#pragma unroll
    for (int ot = 0; ot < 4; ++ot) {
        f4 d = chain<P, 4>(a1, [&](int i, int k) { return S.w2[(16 * ot + i) * SW2 + k]; });
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const float h = d[r] + S.b2[16 * ot + 4 * q + r];
            a2[ot][r] = PRE ? 0.f : selu(h);
            h2[ot][r] = PRE ? h : a2[ot][r];
        }
    }
    if constexpr (!OUT) return;
#pragma unroll
    for (int ob = 0; ob < KO; ++ob) {
        y[ob] = chain<P, 4>(a2, [&](int i, int k) { return S.w3[(16 * ob + i) * SW3 + k]; });
#pragma unroll
        for (int r = 0; r < 4; ++r) y[ob][r] += S.b3[16 * ob + 4 * q + r];
    }
}

template <int P, int KI, int KO>
__global__ __launch_bounds__(kThreads) void mlp_fwd_kernel(const float* __restrict__ x, int64_t n,
                                                           const float* w1, const float* b1, const float* w2,
                                                           const float* b2, const float* w3, const float* b3,
                                                           int n_out, float* __restrict__ y) {
    __shared__ Smem<KI, KO> S;
    load_weights<KI, KO>(S, w1, b1, w2, b2, w3, b3, n_out);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, q = lane >> 4, c = lane & 15;
    const int64_t ntiles = (n + 15) / 16;
    for (int64_t t = (int64_t)blockIdx.x * kWaves + wave; t < ntiles; t += (int64_t)gridDim.x * kWaves) {
        asm volatile("" ::: "memory");  // weights are re-read from LDS per tile, not held in registers
        const int64_t p = t * 16 + c;
        f4 xin[KI], a1[4], a2[4], out[KO];
        load_x_chain<KI>(x, p, n, xin);
        tile_forward<P, KI, KO, true>(S, xin, a1, a2, out);
        if (p < n) {
#pragma unroll
            for (int ob = 0; ob < KO; ++ob)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int o = 16 * ob + 4 * q + r;
                    if (o < n_out) y[p * n_out + o] = out[ob][r];
                }
        }
    }
}

// KI * KO > 4 (the up-to-48-output heads) runs one wave per SIMD, the others two.
// The workgroup's four waves take four consecutive 16-point tiles per step and split the
// weight-gradient tiles among themselves: wave w holds gW3's column block w, gW2's row block w and
// gW1's row block w, each contracted over all 64 points of the step from the transposed fragments
// that every wave leaves in LDS. A wave so keeps 4 (KO + 4 + KI) accumulator registers (28 for the
// 32-input, 16-output head) instead of four times as many, and the kernel needs no scratch. Five
// workgroup barriers per step order the shared blocks: A2 in blocks 0-3 (read for gW3), then dH2 in
// 4-7 and A1 in 0-3 (read for gW2), then dH1 in 4-7 (read for gW1).
template <int P, int KI, int KO>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(KI * KO > 4 ? 1 : 2))) void mlp_bwd_kernel(
    const float* __restrict__ x, int64_t n, const float* w1, const float* b1, const float* w2, const float* b2,
    const float* w3, const float* b3, int n_out, const float* __restrict__ gy, float* __restrict__ gx,
    float* gw1, float* gb1, float* gw2, float* gb2, float* gw3, float* gb3) {
    using Sh = Shape<KI, KO>;
    constexpr int SW1 = Sh::SW1, IN = Sh::IN, OUT = Sh::OUT;
    __shared__ Smem<KI, KO> S;
    __shared__ float TB[Sh::TB];
    static_assert(Sh::TB >= Sh::NGRAD && Sh::TB >= kWaves * kTBlocks * kTBlock, "reduction buffer too small");
    static_assert(kWaves == 4, "one hidden block of 16 units per wave");
    load_weights<KI, KO>(S, w1, b1, w2, b2, w3, b3, n_out);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, q = lane >> 4, c = lane & 15;
    float* tb = TB + wave * kTBlocks * kTBlock;
    // block b of wave w's transpose area
    auto blk = [&](int w, int b) -> const float* { return TB + (w * kTBlocks + b) * kTBlock; };
    // weight-gradient tiles (D layout): g3[ob] = gW3[16 ob.., 16 wave..], g2[kt] = gW2[16 wave.., 16 kt..],
    // g1[kt] = gW1[16 wave.., 16 kt..]; bias partials of lane (q, c) over its points 4q + s:
    // gb1s, gb2s for unit 16 wave + c, gb3s[ob] for output 16 ob + c (every wave forms them, wave 0 adds)
    f4 g3[KO], g2[4], g1[KI];
    float gb1s = 0.f, gb2s = 0.f, gb3s[KO];
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) g2[kt] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kt = 0; kt < KI; ++kt) g1[kt] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ob = 0; ob < KO; ++ob) {
        g3[ob] = f4{0.f, 0.f, 0.f, 0.f};
        gb3s[ob] = 0.f;
    }
    const int64_t ntiles = (n + 15) / 16;
    const int64_t nsteps = (ntiles + kWaves - 1) / kWaves;
    // the loop bound is uniform over the workgroup (every wave reaches every barrier); tiles past
    // the end contribute zeros
    for (int64_t step = blockIdx.x; step < nsteps; step += gridDim.x) {
        asm volatile("" ::: "memory");  // weights are re-read from LDS per tile, not held in registers
        const int64_t t0 = step * kWaves;
        const int64_t p = (t0 + wave) * 16 + c;
        f4 xin[KI], h1[4], h2[4], unused[KO];
        load_x_chain<KI>(x, p, n, xin);
        tile_forward<P, KI, KO, false, true>(S, xin, h1, h2, unused);
        // dY of this wave's tile in the chain layout (unit 16 ob + 4q + r of point c), and as point
        // fragments of all four tiles (output 16 ob + c, points 4q + s of tile t0 + w)
        f4 dyc[KO];
        float dyp[kWaves][KO][4];
#pragma unroll
        for (int ob = 0; ob < KO; ++ob) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int o = 16 * ob + 4 * q + r;
                dyc[ob][r] = (p < n && o < n_out) ? gy[p * n_out + o] : 0.f;
            }
#pragma unroll
            for (int w = 0; w < kWaves; ++w)
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const int64_t pp = (t0 + w) * 16 + 4 * q + s;
                    const int oc = 16 * ob + c;
                    dyp[w][ob][s] = (pp < n && oc < n_out) ? gy[pp * n_out + oc] : 0.f;
                }
        }
        // A2 = selu(H2) to blocks 0-3
        f4 d2[4];
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) {
            f4 a;
            selu_pair(h2[kt], a, d2[kt]);
            put_block(tb + kt * kTBlock, a);
        }
        __syncthreads();
        // gW3[:, 16 wave..] += dY A2^T over the 64 points
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            float a2p[4];
            get_block(blk(w, wave), a2p);
#pragma unroll
            for (int ob = 0; ob < KO; ++ob) {
                gb3s[ob] += (dyp[w][ob][0] + dyp[w][ob][1]) + (dyp[w][ob][2] + dyp[w][ob][3]);
                g3[ob] = outer<P>(dyp[w][ob], a2p, g3[ob]);
            }
        }
        // dH2 = (W3^T dY) * selu'(H2) of this wave's tile
        f4 dh2[4];
#pragma unroll
        for (int ot = 0; ot < 4; ++ot) {
            f4 d = chain<P, KO>(dyc, [&](int i, int k) { return S.w3[k * SW3 + 16 * ot + i]; });
#pragma unroll
            for (int r = 0; r < 4; ++r) dh2[ot][r] = d[r] * d2[ot][r];
        }
        f4 a1[4], d1[4];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            selu_pair(h1[b], a1[b], d1[b]);
            put_block(tb + (4 + b) * kTBlock, dh2[b]);
        }
        __syncthreads();  // every wave is done with the A2 blocks
#pragma unroll
        for (int b = 0; b < 4; ++b) put_block(tb + b * kTBlock, a1[b]);
        __syncthreads();
        // gW2[16 wave.., :] += dH2 A1^T over the 64 points
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            float dh2p[4];
            get_block(blk(w, 4 + wave), dh2p);
            gb2s += (dh2p[0] + dh2p[1]) + (dh2p[2] + dh2p[3]);
#pragma unroll
            for (int kt = 0; kt < 4; ++kt) {
                float a1p[4];
                get_block(blk(w, kt), a1p);
                g2[kt] = outer<P>(dh2p, a1p, g2[kt]);
            }
        }
        // dH1 = (W2^T dH2) * selu'(H1)
        f4 dh1[4];
#pragma unroll
        for (int ot = 0; ot < 4; ++ot) {
            f4 d = chain<P, 4>(dh2, [&](int i, int k) { return S.w2[k * SW2 + 16 * ot + i]; });
#pragma unroll
            for (int r = 0; r < 4; ++r) dh1[ot][r] = d[r] * d1[ot][r];
        }
        // dX = W1^T dH1
        if (gx) {
#pragma unroll
            for (int ot = 0; ot < KI; ++ot) {
                f4 d = chain<P, 4>(dh1, [&](int i, int k) { return S.w1[k * SW1 + 16 * ot + i]; });
                if (p < n) *reinterpret_cast<float4*>(gx + p * IN + 16 * ot + 4 * q) = make_float4(d[0], d[1], d[2], d[3]);
            }
        }
        __syncthreads();  // every wave is done with the dH2 and A1 blocks
#pragma unroll
        for (int b = 0; b < 4; ++b) put_block(tb + (4 + b) * kTBlock, dh1[b]);
        __syncthreads();
        // gW1[16 wave.., :] += dH1 X^T over the 64 points
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            float dh1p[4];
            get_block(blk(w, 4 + wave), dh1p);
            gb1s += (dh1p[0] + dh1p[1]) + (dh1p[2] + dh1p[3]);
#pragma unroll
            for (int kt = 0; kt < KI; ++kt) {
                float xp[4];
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const int64_t pp = (t0 + w) * 16 + 4 * q + s;
                    xp[s] = pp < n ? x[pp * IN + 16 * kt + c] : 0.f;
                }
                g1[kt] = outer<P>(dh1p, xp, g1[kt]);
            }
        }
    }
    // each weight-gradient element has one owning wave; the bias partials are summed over the lanes
    // in LDS; then one global atomic per element
    __syncthreads();
    float* G3 = TB;                       // [OUT][64]
    float* G2 = G3 + OUT * kHid;          // [64][64]
    float* G1 = G2 + kHid * kHid;         // [64][IN]
    float* B1 = G1 + kHid * IN;
    float* B2 = B1 + kHid;
    float* B3 = B2 + kHid;
    for (int i = threadIdx.x; i < Sh::NGRAD; i += blockDim.x) TB[i] = 0.f;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int o = 4 * q + r;
#pragma unroll
        for (int ob = 0; ob < KO; ++ob) G3[(16 * ob + o) * kHid + 16 * wave + c] = g3[ob][r];
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) G2[(16 * wave + o) * kHid + 16 * kt + c] = g2[kt][r];
#pragma unroll
        for (int kt = 0; kt < KI; ++kt) G1[(16 * wave + o) * IN + 16 * kt + c] = g1[kt][r];
    }
    atomicAdd(&B1[16 * wave + c], gb1s);
    atomicAdd(&B2[16 * wave + c], gb2s);
    if (wave == 0)
#pragma unroll
        for (int ob = 0; ob < KO; ++ob) atomicAdd(&B3[16 * ob + c], gb3s[ob]);
    __syncthreads();
    for (int i = threadIdx.x; i < Sh::NGRAD; i += blockDim.x) {
        const float v = TB[i];
        if (v == 0.f) continue;
        if (i < OUT * kHid) {
            if (i / kHid < n_out) unsafeAtomicAdd(gw3 + i, v);
        } else if (i < OUT * kHid + kHid * kHid) {
            unsafeAtomicAdd(gw2 + (i - OUT * kHid), v);
        } else if (i < OUT * kHid + kHid * kHid + kHid * IN) {
            unsafeAtomicAdd(gw1 + (i - OUT * kHid - kHid * kHid), v);
        } else {
            const int b = i - (OUT * kHid + kHid * kHid + kHid * IN);
            if (b < kHid) unsafeAtomicAdd(gb1 + b, v);
            else if (b < 2 * kHid) unsafeAtomicAdd(gb2 + (b - kHid), v);
            else if (b - 2 * kHid < n_out) unsafeAtomicAdd(gb3 + (b - 2 * kHid), v);
        }
    }
}

// Workgroups of `fn` that are resident on the current device at once (CUs x blocks per CU), cached
// per kernel and device; `fallback` if the runtime cannot say. The kernels loop over the tiles in
// grid strides, so a grid larger than this would run its last workgroups as a second, partly
// occupied pass over the same amount of work per workgroup.
inline int resident_blocks(const void* fn, std::atomic<int>* cache, int fallback) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return fallback;
    int v = cache[dev].load(std::memory_order_relaxed);
    if (v > 0) return v;
    int cus = 0, per_cu = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kThreads, 0) != hipSuccess || cus <= 0 ||
        per_cu <= 0) {
        (void)hipGetLastError();
        return fallback;
    }
    v = cus * per_cu;
    cache[dev].store(v, std::memory_order_relaxed);
    return v;
}

inline unsigned grid_for(int64_t n, int cap) {
    int64_t tiles = (n + 15) / 16;
    int64_t blocks = (tiles + kWaves - 1) / kWaves;
    return (unsigned)(blocks < cap ? blocks : cap);
}

// resident-workgroup caches: [shape][fwd/bwd][precision], shapes (KI, KO) = (2, 1), (2, 3), (4, 1), (4, 3)
std::atomic<int> g_resident[4][2][2][kMaxDevices];

template <int KI, int KO>
constexpr int shape_slot() { return (KI == 4 ? 2 : 0) + (KO == 3 ? 1 : 0); }

template <int P, int KI, int KO>
void launch_fwd(const float* x, int64_t n, const float* w1, const float* b1, const float* w2, const float* b2,
                const float* w3, const float* b3, int n_out, float* y, hipStream_t st) {
    auto* cache = g_resident[shape_slot<KI, KO>()][0][P];
    dim3 grid(grid_for(n, resident_blocks(reinterpret_cast<const void*>(&mlp_fwd_kernel<P, KI, KO>), cache, 1024)));
    hipLaunchKernelGGL((mlp_fwd_kernel<P, KI, KO>), grid, dim3(kThreads), 0, st, x, n, w1, b1, w2, b2, w3, b3, n_out,
                       y);
}

template <int P, int KI, int KO>
void launch_bwd(const float* x, int64_t n, const float* w1, const float* b1, const float* w2, const float* b2,
                const float* w3, const float* b3, int n_out, const float* gy, float* gx, float* gw1, float* gb1,
                float* gw2, float* gb2, float* gw3, float* gb3, hipStream_t st) {
    auto* cache = g_resident[shape_slot<KI, KO>()][1][P];
    dim3 grid(grid_for(n, resident_blocks(reinterpret_cast<const void*>(&mlp_bwd_kernel<P, KI, KO>), cache, 512)));
    hipLaunchKernelGGL((mlp_bwd_kernel<P, KI, KO>), grid, dim3(kThreads), 0, st, x, n, w1, b1, w2, b2, w3, b3, n_out,
                       gy, gx, gw1, gb1, gw2, gb2, gw3, gb3);
}

template <int KI, int KO>
void dispatch_fwd(int precision, const float* x, int64_t n, const float* w1, const float* b1, const float* w2,
                  const float* b2, const float* w3, const float* b3, int n_out, float* y, hipStream_t st) {
    if (precision == 0) launch_fwd<0, KI, KO>(x, n, w1, b1, w2, b2, w3, b3, n_out, y, st);
    else launch_fwd<1, KI, KO>(x, n, w1, b1, w2, b2, w3, b3, n_out, y, st);
}

template <int KI, int KO>
void dispatch_bwd(int precision, const float* x, int64_t n, const float* w1, const float* b1, const float* w2,
                  const float* b2, const float* w3, const float* b3, int n_out, const float* gy, float* gx,
                  float* gw1, float* gb1, float* gw2, float* gb2, float* gw3, float* gb3, hipStream_t st) {
    if (precision == 0) launch_bwd<0, KI, KO>(x, n, w1, b1, w2, b2, w3, b3, n_out, gy, gx, gw1, gb1, gw2, gb2, gw3, gb3, st);
    else launch_bwd<1, KI, KO>(x, n, w1, b1, w2, b2, w3, b3, n_out, gy, gx, gw1, gb1, gw2, gb2, gw3, gb3, st);
}

// Argument checks of the head entry points; the message names the entry point.
#define HEAD_CHECK(cond, msg)                                 \
    do {                                                      \
        if (!(cond)) {                                        \
            ::rmesh::set_error(std::string(who) + ": " + msg); \
            return RM_ERR_INVALID_ARGUMENT;                   \
        }                                                     \
    } while (0)

int head_fwd(const char* who, const float* x, int64_t n, int n_in, const float* w1, const float* b1, const float* w2,
             const float* b2, const float* w3, const float* b3, int n_out, int precision, float* y, hipStream_t st) {
    HEAD_CHECK(n >= 0, "negative n");
    HEAD_CHECK(n_in == 32 || n_in == 64, "n_in must be 32 or 64");
    HEAD_CHECK(n_out >= 1 && n_out <= 48, "n_out must be in [1, 48]");
    HEAD_CHECK(precision == 0 || precision == 1, "precision must be 0 (fp32) or 1 (bf16)");
    if (n == 0) return RM_OK;
    HEAD_CHECK(x && w1 && b1 && w2 && b2 && w3 && b3 && y, "null buffer");
    HEAD_CHECK(((uintptr_t)x & 15) == 0, "x must be 16-byte aligned");
    const bool wide = n_out > 16;
    if (n_in == 32) {
        if (wide) dispatch_fwd<2, 3>(precision, x, n, w1, b1, w2, b2, w3, b3, n_out, y, st);
        else dispatch_fwd<2, 1>(precision, x, n, w1, b1, w2, b2, w3, b3, n_out, y, st);
    } else {
        if (wide) dispatch_fwd<4, 3>(precision, x, n, w1, b1, w2, b2, w3, b3, n_out, y, st);
        else dispatch_fwd<4, 1>(precision, x, n, w1, b1, w2, b2, w3, b3, n_out, y, st);
    }
    RM_CHECK_LAUNCH(who);
    return RM_OK;
}

int head_bwd(const char* who, const float* x, int64_t n, int n_in, const float* w1, const float* b1, const float* w2,
             const float* b2, const float* w3, const float* b3, int n_out, int precision, const float* gy, float* gx,
             float* gw1, float* gb1, float* gw2, float* gb2, float* gw3, float* gb3, hipStream_t st) {
    HEAD_CHECK(n >= 0, "negative n");
    HEAD_CHECK(n_in == 32 || n_in == 64, "n_in must be 32 or 64");
    HEAD_CHECK(n_out >= 1 && n_out <= 48, "n_out must be in [1, 48]");
    HEAD_CHECK(precision == 0 || precision == 1, "precision must be 0 (fp32) or 1 (bf16)");
    if (n == 0) return RM_OK;
    HEAD_CHECK(x && w1 && b1 && w2 && b2 && w3 && b3 && gy && gw1 && gb1 && gw2 && gb2 && gw3 && gb3,
                 "null buffer");
    HEAD_CHECK(((uintptr_t)x & 15) == 0 && ((uintptr_t)gx & 15) == 0, "x and grad_x must be 16-byte aligned");
    const bool wide = n_out > 16;
    if (n_in == 32) {
        if (wide) dispatch_bwd<2, 3>(precision, x, n, w1, b1, w2, b2, w3, b3, n_out, gy, gx, gw1, gb1, gw2, gb2, gw3, gb3, st);
        else dispatch_bwd<2, 1>(precision, x, n, w1, b1, w2, b2, w3, b3, n_out, gy, gx, gw1, gb1, gw2, gb2, gw3, gb3, st);
    } else {
        if (wide) dispatch_bwd<4, 3>(precision, x, n, w1, b1, w2, b2, w3, b3, n_out, gy, gx, gw1, gb1, gw2, gb2, gw3, gb3, st);
        else dispatch_bwd<4, 1>(precision, x, n, w1, b1, w2, b2, w3, b3, n_out, gy, gx, gw1, gb1, gw2, gb2, gw3, gb3, st);
    }
    RM_CHECK_LAUNCH(who);
    return RM_OK;
}

#undef HEAD_CHECK

}  // namespace
}  // namespace rmesh

using namespace rmesh;

extern "C" int rm_mlp_fwd(const float* x_dev, int64_t n, const float* w1, const float* b1, const float* w2,
                          const float* b2, const float* w3, const float* b3, int32_t n_out, int32_t precision,
                          float* y_dev, rm_stream_t stream) {
    RM_CHECK_ARG(n_out >= 1 && n_out <= 16, "rm_mlp_fwd: n_out must be in [1, 16]");
    return head_fwd("rm_mlp_fwd", x_dev, n, 32, w1, b1, w2, b2, w3, b3, n_out, precision, y_dev, as_stream(stream));
}

extern "C" int rm_mlp_bwd(const float* x_dev, int64_t n, const float* w1, const float* b1, const float* w2,
                          const float* b2, const float* w3, const float* b3, int32_t n_out, int32_t precision,
                          const float* grad_y_dev, float* grad_x_dev, float* gw1, float* gb1, float* gw2,
                          float* gb2, float* gw3, float* gb3, rm_stream_t stream) {
    RM_CHECK_ARG(n_out >= 1 && n_out <= 16, "rm_mlp_bwd: n_out must be in [1, 16]");
    return head_bwd("rm_mlp_bwd", x_dev, n, 32, w1, b1, w2, b2, w3, b3, n_out, precision, grad_y_dev, grad_x_dev,
                    gw1, gb1, gw2, gb2, gw3, gb3, as_stream(stream));
}

extern "C" int rm_field_head_fwd(const float* x_dev, int64_t n, int32_t n_in, const float* w1, const float* b1,
                                 const float* w2, const float* b2, const float* w3, const float* b3, int32_t n_out,
                                 int32_t precision, float* y_dev, rm_stream_t stream) {
    return head_fwd("rm_field_head_fwd", x_dev, n, n_in, w1, b1, w2, b2, w3, b3, n_out, precision, y_dev,
                    as_stream(stream));
}

extern "C" int rm_field_head_bwd(const float* x_dev, int64_t n, int32_t n_in, const float* w1, const float* b1,
                                 const float* w2, const float* b2, const float* w3, const float* b3, int32_t n_out,
                                 int32_t precision, const float* grad_y_dev, float* grad_x_dev, float* gw1, float* gb1,
                                 float* gw2, float* gb2, float* gw3, float* gb3, rm_stream_t stream) {
    return head_bwd("rm_field_head_bwd", x_dev, n, n_in, w1, b1, w2, b2, w3, b3, n_out, precision, grad_y_dev,
                    grad_x_dev, gw1, gb1, gw2, gb2, gw3, gb3, as_stream(stream));
}
