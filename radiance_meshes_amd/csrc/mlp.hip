// Field MLP 32 -> 64 -> 64 -> n_out (n_out <= 16) with SELU hidden activations, on MFMA.
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
// the hidden activations (selu' is a function of them), chains dY -> dH2 -> dH1 -> dX through
// the transposed weights, and forms the weight gradients
// gW = dH act^T as MFMA tiles contracted over the tile's points; those need the point index on
// the lanes' k, so dH and act blocks are transposed through a 16x17 LDS block per 16 units.
// Weight-gradient tiles accumulate in registers across all tiles of a wave, are summed across the
// workgroup in LDS and added to global memory with one float atomic per element.
#include <atomic>

#include "common.h"

namespace rmesh {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef short s4 __attribute__((ext_vector_type(4)));

constexpr int kIn = 32, kHid = 64, kOutMax = 16;
constexpr int kMaxDevices = 64;
constexpr int kWaves = 4, kThreads = 64 * kWaves;
constexpr int SW1 = kIn + 1, SW2 = kHid + 1, SW3 = kHid + 1;   // LDS row strides (floats)
constexpr int kTS = 17, kTBlock = 16 * kTS;                    // transpose block (16 x 17 floats)
constexpr int kTBlocks = 8;                                    // transpose blocks per wave
constexpr int kNGrad = kOutMax * kHid + kHid * kHid + kHid * kIn + kHid + kHid + kOutMax;

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
__device__ inline float expm1_nonpos(float x) {
    const float q = 0.49999952f + x * (0.16665113f + x * (0.041491002f + x * 0.0075143012f));
    const float p = x + (x * x) * q;
    return x >= -0.25f ? p : __expf(x) - 1.0f;
}
__device__ inline float selu(float x) { return kSeluScale * (x > 0.f ? x : kSeluAlpha * expm1_nonpos(x)); }
// selu'(h) from a = selu(h): a has the sign of h, and for h <= 0 selu'(h) = scale alpha e^h = a + scale alpha
__device__ inline float dselu_of_act(float a) { return a > 0.f ? kSeluScale : a + kSeluScale * kSeluAlpha; }

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

struct Smem {
    float w1[kHid * SW1];
    float w2[kHid * SW2];
    float w3[kOutMax * SW3];
    float b1[kHid], b2[kHid], b3[kOutMax];
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

// Input rows of the tile in the chain layout: xin[kt][r] = x[point c][16 kt + 4q + r].
__device__ inline void load_x_chain(const float* __restrict__ x, int64_t p, int64_t n, f4 (&xin)[2]) {
    const int lane = threadIdx.x & 63, q = lane >> 4;
#pragma unroll
    for (int kt = 0; kt < 2; ++kt) {
        if (p < n) {
            const float4 v = *reinterpret_cast<const float4*>(x + p * kIn + 16 * kt + 4 * q);
            xin[kt] = f4{v.x, v.y, v.z, v.w};
        } else {
            xin[kt] = f4{0.f, 0.f, 0.f, 0.f};
        }
    }
}

// Forward of one tile: hidden activations a1, a2 (chain layout) and the output block. The
// backward takes selu' from the activations (dselu_of_act), so no pre-activation is kept.
template <int P>
__device__ inline f4 tile_forward(const Smem& S, const f4 (&xin)[2], f4 (&a1)[4], f4 (&a2)[4]) {
    const int lane = threadIdx.x & 63, q = lane >> 4;
#pragma unroll
    for (int ot = 0; ot < 4; ++ot) {
        f4 d = chain<P, 2>(xin, [&](int i, int k) { return S.w1[(16 * ot + i) * SW1 + k]; });
#pragma unroll
        for (int r = 0; r < 4; ++r) a1[ot][r] = selu(d[r] + S.b1[16 * ot + 4 * q + r]);
    }
#pragma unroll
    for (int ot = 0; ot < 4; ++ot) {
        f4 d = chain<P, 4>(a1, [&](int i, int k) { return S.w2[(16 * ot + i) * SW2 + k]; });
#pragma unroll
        for (int r = 0; r < 4; ++r) a2[ot][r] = selu(d[r] + S.b2[16 * ot + 4 * q + r]);
    }
    f4 y = chain<P, 4>(a2, [&](int i, int k) { return S.w3[i * SW3 + k]; });
#pragma unroll
    for (int r = 0; r < 4; ++r) y[r] += S.b3[4 * q + r];
    return y;
}

template <int P>
__global__ __launch_bounds__(kThreads) void mlp_fwd_kernel(const float* __restrict__ x, int64_t n,
                                                           const float* w1, const float* b1, const float* w2,
                                                           const float* b2, const float* w3, const float* b3,
                                                           int n_out, float* __restrict__ y) {
    __shared__ Smem S;
    load_weights(S, w1, b1, w2, b2, w3, b3, n_out);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, q = lane >> 4, c = lane & 15;
    const int64_t ntiles = (n + 15) / 16;
    for (int64_t t = (int64_t)blockIdx.x * kWaves + wave; t < ntiles; t += (int64_t)gridDim.x * kWaves) {
        asm volatile("" ::: "memory");  // weights are re-read from LDS per tile, not held in registers
        const int64_t p = t * 16 + c;
        f4 xin[2], a1[4], a2[4];
        load_x_chain(x, p, n, xin);
        const f4 out = tile_forward<P>(S, xin, a1, a2);
        if (p < n) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int o = 4 * q + r;
                if (o < n_out) y[p * n_out + o] = out[r];
            }
        }
    }
}

template <int P>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(2))) void mlp_bwd_kernel(
    const float* __restrict__ x, int64_t n, const float* w1, const float* b1, const float* w2, const float* b2,
    const float* w3, const float* b3, int n_out, const float* __restrict__ gy, float* __restrict__ gx,
    float* gw1, float* gb1, float* gw2, float* gb2, float* gw3, float* gb3) {
    __shared__ Smem S;
    __shared__ float TB[kWaves * kTBlocks * kTBlock];
    static_assert(kWaves * kTBlocks * kTBlock >= kNGrad, "gradient sums reuse the transpose blocks");
    load_weights(S, w1, b1, w2, b2, w3, b3, n_out);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, q = lane >> 4, c = lane & 15;
    float* tb = TB + wave * kTBlocks * kTBlock;
    // weight-gradient tiles (D layout): gW3 [16 x 64], gW2 [64 x 64], gW1 [64 x 32]; bias partials
    // per lane for the units it holds in the chain layout
    f4 g3[4], g2[4][4], g1[4][2];
    // bias partials from the point fragments: lane (q, c) sums unit 16*ot + c over its points 4q + s
    float gb1s[4], gb2s[4], gb3s = 0.f;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        g3[a] = f4{0.f, 0.f, 0.f, 0.f};
        gb1s[a] = 0.f;
        gb2s[a] = 0.f;
#pragma unroll
        for (int b = 0; b < 4; ++b) g2[a][b] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int b = 0; b < 2; ++b) g1[a][b] = f4{0.f, 0.f, 0.f, 0.f};
    }
    const int64_t ntiles = (n + 15) / 16;
    for (int64_t t = (int64_t)blockIdx.x * kWaves + wave; t < ntiles; t += (int64_t)gridDim.x * kWaves) {
        asm volatile("" ::: "memory");  // weights are re-read from LDS per tile, not held in registers
        const int64_t p = t * 16 + c;
        f4 xin[2], a1[4], a2[4];
        load_x_chain(x, p, n, xin);
        tile_forward<P>(S, xin, a1, a2);
        // dY in the chain layout (unit 4q + r of point c) and as a point fragment (unit c, points 4q + s)
        f4 dyc[1];
        float dyp[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int o = 4 * q + r;
            dyc[0][r] = (p < n && o < n_out) ? gy[p * n_out + o] : 0.f;
            const int64_t pp = t * 16 + 4 * q + r;
            dyp[r] = (pp < n && c < n_out) ? gy[pp * n_out + c] : 0.f;
        }
        gb3s += (dyp[0] + dyp[1]) + (dyp[2] + dyp[3]);
        // gW3 += dY A2^T
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) put_block(tb + kt * kTBlock, a2[kt]);
        wave_sync();
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) {
            float a2p[4];
            get_block(tb + kt * kTBlock, a2p);
            g3[kt] = outer<P>(dyp, a2p, g3[kt]);
        }
        // dH2 = (W3^T dY) * selu'(H2)
        f4 dh2[4];
#pragma unroll
        for (int ot = 0; ot < 4; ++ot) {
            f4 d = chain<P, 1>(dyc, [&](int i, int k) { return S.w3[k * SW3 + 16 * ot + i]; });
#pragma unroll
            for (int r = 0; r < 4; ++r) dh2[ot][r] = d[r] * dselu_of_act(a2[ot][r]);
        }
        wave_sync();
        // gW2 += dH2 A1^T
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            put_block(tb + b * kTBlock, dh2[b]);
            put_block(tb + (4 + b) * kTBlock, a1[b]);
        }
        wave_sync();
        {
            // the dH2 fragments are read one block at a time, to keep the live registers down
            float a1p[4][4];
#pragma unroll
            for (int b = 0; b < 4; ++b) get_block(tb + (4 + b) * kTBlock, a1p[b]);
#pragma unroll
            for (int ot = 0; ot < 4; ++ot) {
                float dh2p[4];
                get_block(tb + ot * kTBlock, dh2p);
                gb2s[ot] += (dh2p[0] + dh2p[1]) + (dh2p[2] + dh2p[3]);
#pragma unroll
                for (int kt = 0; kt < 4; ++kt) g2[ot][kt] = outer<P>(dh2p, a1p[kt], g2[ot][kt]);
            }
        }
        // dH1 = (W2^T dH2) * selu'(H1)
        f4 dh1[4];
#pragma unroll
        for (int ot = 0; ot < 4; ++ot) {
            f4 d = chain<P, 4>(dh2, [&](int i, int k) { return S.w2[k * SW2 + 16 * ot + i]; });
#pragma unroll
            for (int r = 0; r < 4; ++r) dh1[ot][r] = d[r] * dselu_of_act(a1[ot][r]);
        }
        // dX = W1^T dH1
        if (gx) {
#pragma unroll
            for (int ot = 0; ot < 2; ++ot) {
                f4 d = chain<P, 4>(dh1, [&](int i, int k) { return S.w1[k * SW1 + 16 * ot + i]; });
                if (p < n) *reinterpret_cast<float4*>(gx + p * kIn + 16 * ot + 4 * q) = make_float4(d[0], d[1], d[2], d[3]);
            }
        }
        wave_sync();
        // gW1 += dH1 X^T
#pragma unroll
        for (int b = 0; b < 4; ++b) put_block(tb + b * kTBlock, dh1[b]);
        wave_sync();
        {
            float xp[2][4];
#pragma unroll
            for (int kt = 0; kt < 2; ++kt)
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const int64_t pp = t * 16 + 4 * q + s;
                    xp[kt][s] = pp < n ? x[pp * kIn + 16 * kt + c] : 0.f;
                }
#pragma unroll
            for (int ot = 0; ot < 4; ++ot) {
                float dh1p[4];
                get_block(tb + ot * kTBlock, dh1p);
                gb1s[ot] += (dh1p[0] + dh1p[1]) + (dh1p[2] + dh1p[3]);
#pragma unroll
                for (int kt = 0; kt < 2; ++kt) g1[ot][kt] = outer<P>(dh1p, xp[kt], g1[ot][kt]);
            }
        }
        wave_sync();
    }
    // sum the waves' gradients in LDS (over the transpose blocks), then one global atomic per element
    __syncthreads();
    float* G3 = TB;                       // [16][64]
    float* G2 = G3 + kOutMax * kHid;      // [64][64]
    float* G1 = G2 + kHid * kHid;         // [64][32]
    float* B1 = G1 + kHid * kIn;
    float* B2 = B1 + kHid;
    float* B3 = B2 + kHid;
    for (int i = threadIdx.x; i < kNGrad; i += blockDim.x) TB[i] = 0.f;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int o = 4 * q + r;
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) atomicAdd(&G3[o * kHid + 16 * kt + c], g3[kt][r]);
#pragma unroll
        for (int ot = 0; ot < 4; ++ot) {
#pragma unroll
            for (int kt = 0; kt < 4; ++kt) atomicAdd(&G2[(16 * ot + o) * kHid + 16 * kt + c], g2[ot][kt][r]);
#pragma unroll
            for (int kt = 0; kt < 2; ++kt) atomicAdd(&G1[(16 * ot + o) * kIn + 16 * kt + c], g1[ot][kt][r]);
        }
    }
#pragma unroll
    for (int ot = 0; ot < 4; ++ot) {
        atomicAdd(&B1[16 * ot + c], gb1s[ot]);
        atomicAdd(&B2[16 * ot + c], gb2s[ot]);
    }
    atomicAdd(&B3[c], gb3s);
    __syncthreads();
    for (int i = threadIdx.x; i < kNGrad; i += blockDim.x) {
        const float v = TB[i];
        if (v == 0.f) continue;
        if (i < kOutMax * kHid) {
            if (i / kHid < n_out) unsafeAtomicAdd(gw3 + i, v);
        } else if (i < kOutMax * kHid + kHid * kHid) {
            unsafeAtomicAdd(gw2 + (i - kOutMax * kHid), v);
        } else if (i < kOutMax * kHid + kHid * kHid + kHid * kIn) {
            unsafeAtomicAdd(gw1 + (i - kOutMax * kHid - kHid * kHid), v);
        } else {
            const int b = i - (kOutMax * kHid + kHid * kHid + kHid * kIn);
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

std::atomic<int> g_resident[4][kMaxDevices];   // mlp_fwd<0>, mlp_fwd<1>, mlp_bwd<0>, mlp_bwd<1>

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
    RM_CHECK_ARG(((uintptr_t)x_dev & 15) == 0, "rm_mlp_fwd: x must be 16-byte aligned");
    const void* fn = precision == 0 ? reinterpret_cast<const void*>(&mlp_fwd_kernel<0>)
                                    : reinterpret_cast<const void*>(&mlp_fwd_kernel<1>);
    dim3 grid(grid_for(n, resident_blocks(fn, g_resident[precision], 1024))), block(kThreads);
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
    RM_CHECK_ARG(((uintptr_t)x_dev & 15) == 0 && ((uintptr_t)grad_x_dev & 15) == 0,
                 "rm_mlp_bwd: x and grad_x must be 16-byte aligned");
    const void* fn = precision == 0 ? reinterpret_cast<const void*>(&mlp_bwd_kernel<0>)
                                    : reinterpret_cast<const void*>(&mlp_bwd_kernel<1>);
    dim3 grid(grid_for(n, resident_blocks(fn, g_resident[2 + precision], 512))), block(kThreads);
    if (precision == 0)
        hipLaunchKernelGGL(mlp_bwd_kernel<0>, grid, block, 0, as_stream(stream), x_dev, n, w1, b1, w2, b2, w3, b3,
                           n_out, grad_y_dev, grad_x_dev, gw1, gb1, gw2, gb2, gw3, gb3);
    else
        hipLaunchKernelGGL(mlp_bwd_kernel<1>, grid, block, 0, as_stream(stream), x_dev, n, w1, b1, w2, b2, w3, b3,
                           n_out, grad_y_dev, grad_x_dev, gw1, gb1, gw2, gb2, gw3, gb3);
    RM_CHECK_LAUNCH("mlp_bwd_kernel");
    return RM_OK;
}
