// Hash-grid field: multi-level grid encode (forward) and its scatter-add adjoint (backward).
//
// Semantics: HashEmbedderOptimized (utils/hashgrid.py:5-124) on the [0,1]^3 box with
// conditional hashing: level l has resolution res[l]; a level with res^3 <= 2^log2_size is
// dense (index = x%r + (y%r) r + (z%r) r^2), otherwise hashed as
// (x * 1) ^ (y * 2654435761) ^ (z * 805459861) & (2^log2_size - 1) (the low 32 bits of the
// reference's int64 arithmetic). pos = x * res + 0.5 rounded as two float ops, as torch does;
// trilinear interpolation in the reference's x -> y -> z order. Output [N, L * n_feat],
// level-major. One thread per (point, level): consecutive threads write consecutive features.
#include "common.h"

namespace rmesh {
namespace {

constexpr int kMaxLevels = 32;

struct GridArgs {
    int res[kMaxLevels];
    int64_t offs[kMaxLevels];
    int dense[kMaxLevels];
    int L, F;
    uint32_t mask;
};

__device__ inline uint32_t corner_index(const GridArgs& g, int l, int cx, int cy, int cz) {
    if (g.dense[l]) {
        int r = g.res[l];
        return (uint32_t)((cx % r) + (cy % r) * r + (cz % r) * r * r);
    }
    return ((uint32_t)cx * 1u ^ (uint32_t)cy * 2654435761u ^ (uint32_t)cz * 805459861u) & g.mask;
}

// Cell corner and fractions of a query point already clamped to [0, 1].
__device__ inline void cell_of(const float v[3], const GridArgs& g, int l, int& bx, int& by, int& bz, float& fx,
                               float& fy, float& fz) {
#pragma clang fp contract(off)
    // x * r and + 0.5 round separately, as the reference's float32 torch ops do
    const float r = (float)g.res[l];
    float p[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float xr = v[k] * r;
        p[k] = xr + 0.5f;
    }
    float f0 = floorf(p[0]), f1 = floorf(p[1]), f2 = floorf(p[2]);
    bx = (int)f0; by = (int)f1; bz = (int)f2;
    fx = p[0] - f0; fy = p[1] - f1; fz = p[2] - f2;
}

__device__ inline void load_point(const float* __restrict__ x, int64_t n, float v[3]) {
#pragma unroll
    for (int k = 0; k < 3; ++k) v[k] = fminf(fmaxf(x[3 * n + k], 0.0f), 1.0f);
}

__device__ inline void cell(const float* __restrict__ x, int64_t n, const GridArgs& g, int l, int& bx, int& by,
                            int& bz, float& fx, float& fy, float& fz) {
    float v[3];
    load_point(x, n, v);
    cell_of(v, g, l, bx, by, bz, fx, fy, fz);
}

template <int F>
__global__ __launch_bounds__(256) void hashgrid_fwd_kernel(const float* __restrict__ x, int64_t N,
                                                           const float* __restrict__ table, GridArgs g,
                                                           float* __restrict__ out) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N * g.L) return;
    int64_t n = t / g.L;
    int l = (int)(t - n * g.L);
    int bx, by, bz;
    float fx, fy, fz;
    cell(x, n, g, l, bx, by, bz, fx, fy, fz);
    const float* tb = table + g.offs[l] * F;
    float e[8][F];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        uint32_t h = corner_index(g, l, bx + ((c >> 2) & 1), by + ((c >> 1) & 1), bz + (c & 1));
#pragma unroll
        for (int f = 0; f < F; ++f) e[c][f] = tb[(int64_t)h * F + f];
    }
    const float wx0 = 1.0f - fx, wy0 = 1.0f - fy, wz0 = 1.0f - fz;
#pragma unroll
    for (int f = 0; f < F; ++f) {
        float c0 = e[0][f] * wx0 + e[4][f] * fx;
        float c1 = e[1][f] * wx0 + e[5][f] * fx;
        float c2 = e[2][f] * wx0 + e[6][f] * fx;
        float c3 = e[3][f] * wx0 + e[7][f] * fx;
        c0 = c0 * wy0 + c2 * fy;
        c1 = c1 * wy0 + c3 * fy;
        out[(n * g.L + l) * F + f] = c0 * wz0 + c1 * fz;
    }
}

// Backward, level-major: blockIdx.y is the level and consecutive lanes take consecutive query
// points, so spatially ordered queries give adjacent lanes the same table slots at coarse levels.
// Per corner, a wave whose lanes hit at most 32 distinct slot runs sums each run of equal slot
// indices with a segmented scan and issues one atomic per run (its last lane); otherwise every
// lane adds its own contribution directly.
template <int F>
__global__ __launch_bounds__(256) void hashgrid_bwd_kernel(const float* __restrict__ x, int64_t N,
                                                           const float* __restrict__ gout, GridArgs g,
                                                           float* __restrict__ gtable) {
    const int l = blockIdx.y;
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const bool valid = n < N;
    float go[F];
#pragma unroll
    for (int f = 0; f < F; ++f) go[f] = valid ? gout[(n * g.L + l) * F + f] : 0.f;
    int bx = 0, by = 0, bz = 0;
    float fx = 0.f, fy = 0.f, fz = 0.f;
    if (valid) cell(x, n, g, l, bx, by, bz, fx, fy, fz);
    float* tb = gtable + g.offs[l] * F;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const int ox = (c >> 2) & 1, oy = (c >> 1) & 1, oz = c & 1;
        const float w = (ox ? fx : 1.0f - fx) * (oy ? fy : 1.0f - fy) * (oz ? fz : 1.0f - fz);
        const uint32_t h = valid ? corner_index(g, l, bx + ox, by + oy, bz + oz) : 0xFFFFFFFFu;
        float v[F];
#pragma unroll
        for (int f = 0; f < F; ++f) v[f] = w * go[f];
        const uint32_t hprev = __shfl_up(h, 1, 64);
        const bool head = lane == 0 || h != hprev;
        const uint64_t heads = __ballot(head);
        if (__popcll(heads) > 32) {
            if (valid)
#pragma unroll
                for (int f = 0; f < F; ++f)
                    if (v[f] != 0.f) unsafeAtomicAdd(tb + (int64_t)h * F + f, v[f]);
            continue;
        }
        // segmented inclusive scan over runs of equal h
        bool seg = head;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            float u[F];
#pragma unroll
            for (int f = 0; f < F; ++f) u[f] = __shfl_up(v[f], o, 64);
            const bool su = __shfl_up((int)seg, o, 64);
            if (lane >= o && !seg) {
#pragma unroll
                for (int f = 0; f < F; ++f) v[f] += u[f];
            }
            if (lane >= o) seg = seg || su;
        }
        const bool next_head = (lane == 63) || ((heads >> (lane + 1)) & 1ull);
        if (next_head && h != 0xFFFFFFFFu)
#pragma unroll
            for (int f = 0; f < F; ++f)
                if (v[f] != 0.f) unsafeAtomicAdd(tb + (int64_t)h * F + f, v[f]);
    }
}

// Backward, chunk-major with workgroup aggregation (n_feat 1 or 2): a workgroup owns 256
// consecutive query points, one per lane (a Morton-ordered mesh makes them a compact region), and
// walks every level. Per level its 8 x 256 corner updates are summed in an LDS open-addressing
// table keyed by table slot and then flushed with one global atomic per distinct (slot, feature).
// Each grad_out row is read once (a level-major pass re-reads every point's row once per level),
// and the global atomics drop to the distinct slots a chunk touches (about a third of the
// per-lane count on a 6.7M-point Morton-ordered query set).
constexpr int kAggPoints = 256;
constexpr int kAggLog2Slots = 12;
constexpr int kAggSlots = 1 << kAggLog2Slots;  // > 8 corners x 256 points: an insert always finds room
constexpr uint32_t kEmptySlot = 0xFFFFFFFFu;   // table slots are < 2^31

template <int F>
__global__ __launch_bounds__(kAggPoints) void hashgrid_bwd_agg_kernel(const float* __restrict__ x, int64_t N,
                                                                       const float* __restrict__ gout, GridArgs g,
                                                                       float* __restrict__ gtable) {
    static_assert(F == 1 || F == 2, "LDS table sized for n_feat <= 2");
    __shared__ uint32_t s_key[kAggSlots];
    __shared__ float s_val[kAggSlots * F];
    const int tid = threadIdx.x;
    // dispatch slots go round-robin over the 8 XCDs: give each XCD a contiguous band of chunks so
    // the coarse-level slots that neighbouring chunks share stay in one L2
    const int nb = (int)gridDim.x, per = nb >> 3, bi = (int)blockIdx.x;
    const int b = bi < (per << 3) ? (bi & 7) * per + (bi >> 3) : bi;
    const int64_t n = (int64_t)b * kAggPoints + tid;
    const bool valid = n < N;
    float xv[3] = {0.f, 0.f, 0.f};
    if (valid) load_point(x, n, xv);
    for (int p = tid; p < kAggSlots; p += kAggPoints) s_key[p] = kEmptySlot;
    for (int p = tid; p < kAggSlots * F; p += kAggPoints) s_val[p] = 0.f;
    __syncthreads();
    for (int l = 0; l < g.L; ++l) {
        if (valid) {
            float go[F];
#pragma unroll
            for (int f = 0; f < F; ++f) go[f] = gout[(n * g.L + l) * F + f];
            int bx, by, bz;
            float fx, fy, fz;
            cell_of(xv, g, l, bx, by, bz, fx, fy, fz);
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const int ox = (c >> 2) & 1, oy = (c >> 1) & 1, oz = c & 1;
                const float w = (ox ? fx : 1.0f - fx) * (oy ? fy : 1.0f - fy) * (oz ? fz : 1.0f - fz);
                float v[F];
                bool nz = false;
#pragma unroll
                for (int f = 0; f < F; ++f) {
                    v[f] = w * go[f];
                    nz = nz || v[f] != 0.f;
                }
                if (!nz) continue;
                const uint32_t h = corner_index(g, l, bx + ox, by + oy, bz + oz);
                uint32_t p = (h * 2654435761u) >> (32 - kAggLog2Slots);
                bool placed = false;
                for (int probe = 0; probe < kAggSlots && !placed; ++probe) {
                    const uint32_t k = atomicCAS(&s_key[p], kEmptySlot, h);
                    if (k == kEmptySlot || k == h) {
#pragma unroll
                        for (int f = 0; f < F; ++f) atomicAdd(&s_val[p * F + f], v[f]);
                        placed = true;
                    }
                    p = (p + 1) & (kAggSlots - 1);
                }
                if (!placed)  // unreachable while the table holds more slots than a level can insert
#pragma unroll
                    for (int f = 0; f < F; ++f) unsafeAtomicAdd(gtable + (g.offs[l] + (int64_t)h) * F + f, v[f]);
            }
        }
        __syncthreads();
        float* tb = gtable + g.offs[l] * F;
        for (int p = tid; p < kAggSlots; p += kAggPoints) {
            const uint32_t k = s_key[p];
            if (k == kEmptySlot) continue;
#pragma unroll
            for (int f = 0; f < F; ++f) {
                const float v = s_val[p * F + f];
                if (v != 0.f) unsafeAtomicAdd(tb + (int64_t)k * F + f, v);
                s_val[p * F + f] = 0.f;
            }
            s_key[p] = kEmptySlot;
        }
        __syncthreads();
    }
}

int make_grid_args(const int32_t* res, const int64_t* offs, int32_t L, int32_t F, int32_t log2_size, GridArgs& g,
                   const char* who) {
    std::string w(who);
    if (L <= 0 || L > kMaxLevels) { set_error(w + ": n_levels must be in [1, 32]"); return RM_ERR_INVALID_ARGUMENT; }
    if (F != 1 && F != 2 && F != 4 && F != 8) { set_error(w + ": n_feat must be 1, 2, 4 or 8"); return RM_ERR_INVALID_ARGUMENT; }
    if (log2_size <= 0 || log2_size > 31) { set_error(w + ": log2_size out of range"); return RM_ERR_INVALID_ARGUMENT; }
    if (!res || !offs) { set_error(w + ": null level table"); return RM_ERR_INVALID_ARGUMENT; }
    g.L = L;
    g.F = F;
    g.mask = (uint32_t)((1ull << log2_size) - 1ull);
    for (int l = 0; l < L; ++l) {
        if (res[l] <= 0) { set_error(w + ": resolution must be positive"); return RM_ERR_INVALID_ARGUMENT; }
        g.res[l] = res[l];
        g.offs[l] = offs[l];
        double cells = (double)res[l] * res[l] * res[l];
        g.dense[l] = cells <= (double)(1ull << log2_size);
        int64_t need = g.dense[l] ? (int64_t)res[l] * res[l] * res[l] : (1ll << log2_size);
        if (offs[l + 1] - offs[l] < need) { set_error(w + ": level table smaller than the level needs"); return RM_ERR_INVALID_ARGUMENT; }
    }
    return RM_OK;
}

}  // namespace
}  // namespace rmesh

using namespace rmesh;

#define RM_FEAT_DISPATCH(F_, KERNEL, ...)                                                     \
    switch (F_) {                                                                             \
        case 1: hipLaunchKernelGGL(KERNEL<1>, __VA_ARGS__); break;                            \
        case 2: hipLaunchKernelGGL(KERNEL<2>, __VA_ARGS__); break;                            \
        case 4: hipLaunchKernelGGL(KERNEL<4>, __VA_ARGS__); break;                            \
        default: hipLaunchKernelGGL(KERNEL<8>, __VA_ARGS__); break;                           \
    }

extern "C" int rm_hashgrid_fwd(const float* x_dev, int64_t n, const float* table_dev, const int32_t* res,
                               const int64_t* offs, int32_t n_levels, int32_t n_feat, int32_t log2_size,
                               float* out_dev, rm_stream_t stream) {
    RM_CHECK_ARG(n >= 0, "rm_hashgrid_fwd: negative n");
    GridArgs g;
    int e = make_grid_args(res, offs, n_levels, n_feat, log2_size, g, "rm_hashgrid_fwd");
    if (e) return e;
    if (n == 0) return RM_OK;
    RM_CHECK_ARG(x_dev && table_dev && out_dev, "rm_hashgrid_fwd: null buffer");
    int64_t total = n * n_levels;
    dim3 grid((unsigned)ceil_div(total, 256)), block(256);
    RM_FEAT_DISPATCH(n_feat, hashgrid_fwd_kernel, grid, block, 0, as_stream(stream), x_dev, n, table_dev, g, out_dev);
    RM_CHECK_LAUNCH("hashgrid_fwd_kernel");
    return RM_OK;
}

extern "C" int rm_hashgrid_bwd(const float* x_dev, int64_t n, const float* grad_out_dev, const int32_t* res,
                               const int64_t* offs, int32_t n_levels, int32_t n_feat, int32_t log2_size,
                               float* grad_table_dev, rm_stream_t stream) {
    RM_CHECK_ARG(n >= 0, "rm_hashgrid_bwd: negative n");
    GridArgs g;
    int e = make_grid_args(res, offs, n_levels, n_feat, log2_size, g, "rm_hashgrid_bwd");
    if (e) return e;
    if (n == 0) return RM_OK;
    RM_CHECK_ARG(x_dev && grad_out_dev && grad_table_dev, "rm_hashgrid_bwd: null buffer");
    if (n_feat <= 2) {
        dim3 grid((unsigned)ceil_div(n, kAggPoints)), block(kAggPoints);
        if (n_feat == 1)
            hipLaunchKernelGGL(hashgrid_bwd_agg_kernel<1>, grid, block, 0, as_stream(stream), x_dev, n, grad_out_dev, g,
                               grad_table_dev);
        else
            hipLaunchKernelGGL(hashgrid_bwd_agg_kernel<2>, grid, block, 0, as_stream(stream), x_dev, n, grad_out_dev, g,
                               grad_table_dev);
        RM_CHECK_LAUNCH("hashgrid_bwd_agg_kernel");
        return RM_OK;
    }
    dim3 grid((unsigned)ceil_div(n, 256), (unsigned)n_levels), block(256);
    RM_FEAT_DISPATCH(n_feat, hashgrid_bwd_kernel, grid, block, 0, as_stream(stream), x_dev, n, grad_out_dev, g,
                     grad_table_dev);
    RM_CHECK_LAUNCH("hashgrid_bwd_kernel");
    return RM_OK;
}

// Field MLP entry points: filled in by mlp.hip.
