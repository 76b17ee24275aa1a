// Hash-grid field: multi-level grid encode (forward) and its scatter-add adjoint (backward).
//
// Semantics: HashEmbedderOptimized (utils/hashgrid.py:5-124) on the [0,1]^3 box with
// conditional hashing: level l has resolution res[l]; a level with res^3 <= 2^log2_size is
// dense (index = x%r + (y%r) r + (z%r) r^2), otherwise hashed as
// (x * 1) ^ (y * 2654435761) ^ (z * 805459861) & (2^log2_size - 1) (the low 32 bits of the
// reference's int64 arithmetic). pos = x * res + 0.5 rounded as two float ops, as torch does;
// trilinear interpolation in the reference's x -> y -> z order. Output [N, L * n_feat],
// level-major. One thread per (point, level): consecutive threads write consecutive features.
#include <mutex>

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

// c mod r for the corner coordinates a clamped query produces: 0 <= c <= r + 1 (the cell index is
// at most r, plus the corner offset), so at most two subtractions (two only when r = 1) replace
// the integer division a % by a runtime divisor costs.
__device__ inline int wrap_corner(int c, int r) {
    c = c >= r ? c - r : c;
    return c >= r ? c - r : c;
}

__device__ inline uint32_t corner_index(const GridArgs& g, int l, int cx, int cy, int cz) {
    if (g.dense[l]) {
        int r = g.res[l];
        return (uint32_t)(wrap_corner(cx, r) + wrap_corner(cy, r) * r + wrap_corner(cz, r) * r * r);
    }
    return ((uint32_t)cx * 1u ^ (uint32_t)cy * 2654435761u ^ (uint32_t)cz * 805459861u) & g.mask;
}

// corner_index of a hashed level (no dense branch), for loops over levels known to be hashed.
__device__ inline uint32_t hashed_index(const GridArgs& g, int cx, int cy, int cz) {
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

// Encoding of one clamped point at level l: the 8 corner rows gathered, trilinear in x -> y -> z.
template <int F>
__device__ inline void encode_level(const float v[3], const GridArgs& g, int l, const float* __restrict__ table,
                                    float (&res)[F]) {
    int bx, by, bz;
    float fx, fy, fz;
    cell_of(v, g, l, bx, by, bz, fx, fy, fz);
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
        res[f] = c0 * wz0 + c1 * fz;
    }
}

// One thread per (point, level), point-major (any L * F).
template <int F>
__global__ __launch_bounds__(256) void hashgrid_fwd_kernel(const float* __restrict__ x, int64_t N,
                                                           const float* __restrict__ table, GridArgs g,
                                                           float* __restrict__ out) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N * g.L) return;
    int64_t n = t / g.L;
    int l = (int)(t - n * g.L);
    float v[3];
    load_point(x, n, v);
    float r[F];
    encode_level<F>(v, g, l, table, r);
#pragma unroll
    for (int f = 0; f < F; ++f) out[(n * g.L + l) * F + f] = r[f];
}

// Level-major forward for L * F <= kLmMaxRow: a workgroup takes 64 consecutive points, wave w
// levels w, w + 4, ..., one point per lane, so each gather instruction serves 64 neighbouring
// points at one level (spatially ordered points share corner rows at the coarser levels, and
// the level's parameters are wave-uniform). The rows are assembled in LDS and written out as one
// contiguous block.
constexpr int kLmPoints = 64, kLmWaves = 4, kLmMaxRow = 64;

template <int F>
__global__ __launch_bounds__(kLmPoints * kLmWaves) void hashgrid_fwd_lm_kernel(const float* __restrict__ x,
                                                                               int64_t N,
                                                                               const float* __restrict__ table,
                                                                               GridArgs g, float* __restrict__ out) {
    __shared__ float rows[kLmPoints * (kLmMaxRow + 1)];
    const int LF = g.L * F, stride = LF + 1;   // odd stride: a wave's 64 row writes hit distinct banks
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t n0 = (int64_t)blockIdx.x * kLmPoints;
    const int64_t n = n0 + lane;
    if (n < N) {
        float v[3];
        load_point(x, n, v);
        for (int l = wave; l < g.L; l += kLmWaves) {
            float r[F];
            encode_level<F>(v, g, l, table, r);
#pragma unroll
            for (int f = 0; f < F; ++f) rows[lane * stride + l * F + f] = r[f];
        }
    }
    __syncthreads();
    const int64_t left = N - n0;
    const int total = (int)(left < kLmPoints ? left : kLmPoints) * LF;
    float* __restrict__ o = out + n0 * LF;
    for (int i = threadIdx.x; i < total; i += kLmPoints * kLmWaves) {
        const int pt = i / LF;
        o[i] = rows[pt * stride + (i - pt * LF)];
    }
}

// Backward, level-major: blockIdx.y is the level and consecutive lanes take consecutive query
// points, so spatially ordered queries give adjacent lanes the same table slots at coarse levels.
// Per corner, a wave whose lanes hit at most 32 distinct slot runs sums each run of equal slot
// indices with a segmented scan and issues one atomic per run (its last lane); otherwise every
// lane adds its own contribution directly.
// A float atomic costs one memory transaction per cache line it touches, whatever the number of
// lanes that hit the line. With n_feat >= 4 each update is a row of F contiguous floats, so the
// wave's updates of one corner go through LDS and out transposed: F consecutive lanes add the F
// features of one update, one transaction per update instead of F.
template <int F>
__device__ inline void flush_rows(float* __restrict__ tb, uint32_t h, const float (&v)[F], bool writes, float* s_t,
                                  uint32_t* s_h, int lane) {
    constexpr int S = F + 1;
#pragma unroll
    for (int f = 0; f < F; ++f) s_t[lane * S + f] = v[f];
    s_h[lane] = writes ? h : 0xFFFFFFFFu;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    constexpr int PER = 64 / F;   // updates per instruction
#pragma unroll
    for (int j = 0; j < F; ++j) {
        const int e = j * PER + lane / F, f = lane % F;
        const uint32_t he = s_h[e];
        const float val = s_t[e * S + f];
        if (he != 0xFFFFFFFFu && val != 0.f) unsafeAtomicAdd(tb + (int64_t)he * F + f, val);
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

template <int F>
__global__ __launch_bounds__(256) void hashgrid_bwd_kernel(const float* __restrict__ x, int64_t N,
                                                           const float* __restrict__ gout, GridArgs g,
                                                           float* __restrict__ gtable) {
    constexpr bool kRows = F >= 4;
    __shared__ float s_t[kRows ? 4 * 64 * (F + 1) : 1];
    __shared__ uint32_t s_h[kRows ? 4 * 64 : 1];
    float* wt = s_t + (kRows ? (threadIdx.x >> 6) * 64 * (F + 1) : 0);
    uint32_t* wh = s_h + (kRows ? (threadIdx.x >> 6) * 64 : 0);
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
            if constexpr (kRows) {
                flush_rows<F>(tb, h, v, valid, wt, wh, lane);
            } else if (valid) {
#pragma unroll
                for (int f = 0; f < F; ++f)
                    if (v[f] != 0.f) unsafeAtomicAdd(tb + (int64_t)h * F + f, v[f]);
            }
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
        if constexpr (kRows) {
            flush_rows<F>(tb, h, v, next_head && h != 0xFFFFFFFFu, wt, wh, lane);
        } else if (next_head && h != 0xFFFFFFFFu) {
#pragma unroll
            for (int f = 0; f < F; ++f)
                if (v[f] != 0.f) unsafeAtomicAdd(tb + (int64_t)h * F + f, v[f]);
        }
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
constexpr int kAggLog2Slots = 11;
constexpr int kAggSlots = 1 << kAggLog2Slots;  // = 8 corners x 256 points: an insert always finds room
constexpr uint32_t kEmptySlot = 0xFFFFFFFFu;   // table slots are < 2^31

// DPP lane moves (VALU, no LDS): lanes without a source, and rows outside ROWS, read 0.
template <int CTRL, int ROWS = 0xF>
__device__ inline int dpp_i(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, ROWS, 0xF, false);
}
template <int CTRL, int ROWS = 0xF>
__device__ inline float dpp_f(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWS, 0xF, false));
}

// One step of a segmented inclusive scan over the wave: the lane combines with the partial
// (value, segment-started) pair that CTRL brings in, unless its own segment already started.
// row_shr 1, 2, 4, 8 scan each row of 16 lanes; row_bcast:15 (rows 1, 3) and row_bcast:31
// (rows 2, 3) then carry the rows' totals across the wave.
template <int F, int CTRL, int ROWS>
__device__ inline void seg_scan_step(float (&v)[F], int& seg) {
    float u[F];
#pragma unroll
    for (int f = 0; f < F; ++f) u[f] = dpp_f<CTRL, ROWS>(v[f]);
    const int su = dpp_i<CTRL, ROWS>(seg);
#pragma unroll
    for (int f = 0; f < F; ++f) v[f] = seg ? v[f] : v[f] + u[f];
    seg |= su;
}

template <int F>
__global__ __launch_bounds__(kAggPoints) void hashgrid_bwd_agg_kernel(const float* __restrict__ x, int64_t N,
                                                                       const float* __restrict__ gout, GridArgs g,
                                                                       int l_end, float* __restrict__ gtable) {
    static_assert(F == 1 || F == 2, "LDS table sized for n_feat <= 2");
    __shared__ uint32_t s_key[kAggSlots];
    __shared__ double s_val[kAggSlots * F];   // FP64: LDS FP64 adds run ~4x the rate of FP32 ones
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
    for (int p = tid; p < kAggSlots * F; p += kAggPoints) s_val[p] = 0.0;
    __syncthreads();
    const int lane = tid & 63;
    for (int l = 0; l < l_end; ++l) {
        float go[F];
        int bx = 0, by = 0, bz = 0;
        float fx = 0.f, fy = 0.f, fz = 0.f;
#pragma unroll
        for (int f = 0; f < F; ++f) go[f] = valid ? gout[(n * g.L + l) * F + f] : 0.f;
        if (valid) cell_of(xv, g, l, bx, by, bz, fx, fy, fz);
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const int ox = (c >> 2) & 1, oy = (c >> 1) & 1, oz = c & 1;
            const float w = (ox ? fx : 1.0f - fx) * (oy ? fy : 1.0f - fy) * (oz ? fz : 1.0f - fz);
            const uint32_t h = valid ? corner_index(g, l, bx + ox, by + oy, bz + oz) : kEmptySlot;
            float v[F];
#pragma unroll
            for (int f = 0; f < F; ++f) v[f] = w * go[f];
            // consecutive lanes are neighbouring points: at coarse levels runs of lanes share a
            // slot, so a wave with at most 32 runs first sums each run (segmented scan) and only
            // the run's last lane inserts
            const uint32_t hprev = (uint32_t)dpp_i<0x138>((int)h);   // wave_shr:1
            const bool head = lane == 0 || h != hprev;
            const uint64_t heads = __ballot(head);
            bool tail = true;
            if (__popcll(heads) <= 32) {
                int seg = head ? 1 : 0;
                seg_scan_step<F, 0x111, 0xF>(v, seg);   // row_shr:1
                seg_scan_step<F, 0x112, 0xF>(v, seg);   // row_shr:2
                seg_scan_step<F, 0x114, 0xF>(v, seg);   // row_shr:4
                seg_scan_step<F, 0x118, 0xF>(v, seg);   // row_shr:8
                seg_scan_step<F, 0x142, 0xA>(v, seg);   // row_bcast:15 into rows 1 and 3
                seg_scan_step<F, 0x143, 0xC>(v, seg);   // row_bcast:31 into rows 2 and 3
                tail = lane == 63 || ((heads >> (lane + 1)) & 1ull);
            }
            bool nz = false;
#pragma unroll
            for (int f = 0; f < F; ++f) nz = nz || v[f] != 0.f;
            if (!valid || !tail || !nz) continue;
            uint32_t p = (h * 2654435761u) >> (32 - kAggLog2Slots);
            bool placed = false;
            for (int probe = 0; probe < kAggSlots && !placed; ++probe) {
                const uint32_t k = atomicCAS(&s_key[p], kEmptySlot, h);
                if (k == kEmptySlot || k == h) {
#pragma unroll
                    for (int f = 0; f < F; ++f) atomicAdd(&s_val[p * F + f], (double)v[f]);
                    placed = true;
                }
                p = (p + 1) & (kAggSlots - 1);
            }
            if (!placed)  // unreachable while the table holds more slots than a level can insert
#pragma unroll
                for (int f = 0; f < F; ++f) unsafeAtomicAdd(gtable + (g.offs[l] + (int64_t)h) * F + f, v[f]);
        }
        __syncthreads();
        // consecutive lanes take the F features of one slot, so they share one atomic transaction;
        // the lanes of a slot are in one wave and read its key in one instruction, so the first of
        // them clears the key afterwards for the next level
        float* tb = gtable + g.offs[l] * F;
        for (int p = tid; p < kAggSlots * F; p += kAggPoints) {
            const int slot = p / F;
            const uint32_t k = s_key[slot];
            if (k == kEmptySlot) continue;
            const float v = (float)s_val[p];
            if (v != 0.f) unsafeAtomicAdd(tb + (int64_t)k * F + (p - slot * F), v);
            s_val[p] = 0.0;
            if (p == slot * F) s_key[slot] = kEmptySlot;
        }
        __syncthreads();
    }
}

// Backward of the fine hashed levels by slot ownership (n_feat 2). Past a few hundred cells per
// axis nearly every corner update of a 256-point chunk hits its own table slot, so the
// aggregation above issues one global atomic per update. Instead the updates of levels
// [lp, L) are sorted by bucket (4096 consecutive slots of one level) with a counting sort, and one
// workgroup per bucket sums its updates in LDS and adds them to the table with plain stores:
//  1. part_count: per 512-point chunk, the number of updates per bucket (LDS histogram),
//  2. part_column / part_bucket_scan: exclusive offsets per (chunk, bucket), bucket-major,
//  3. part_scatter: per chunk and level, the updates sorted by bucket in LDS, then each bucket's
//     run written to its place as (slot within the bucket, value pair) with coalesced stores,
//  4. part_accumulate: per bucket, the sum of its updates added to the table.
// Staging a chunk-level in LDS keeps every run's cache lines whole: writing the records straight
// from the lanes left each of the ~900 runs a workgroup feeds partly written over the workgroup's
// lifetime, more open lines than the L2 holds, and the partial evictions wrote ~1.45x the bytes.
constexpr int kPartPoints = 512;               // query points per chunk (one per lane)
constexpr int kPartLog2Slots = 12;
constexpr int kPartSlots = 1 << kPartLog2Slots;  // table slots per bucket (64 KB of FP64 LDS sums)
constexpr int kPartMaxBuckets = 8192;
constexpr int kPartMaxLevelBuckets = 512;      // buckets of one level: one per lane
constexpr int kPartAccThreads = 512;
constexpr int kPartMaxLevels = 8;              // levels in the partitioned suffix (their grad rows preloaded)

__global__ __launch_bounds__(kPartPoints) void part_count_kernel(const float* __restrict__ x, int64_t N, GridArgs g,
                                                                 int lp, int nb_level, uint32_t* __restrict__ hist) {
    __shared__ uint32_t s_hist[kPartMaxBuckets];
    const int nb = (g.L - lp) * nb_level;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) s_hist[b] = 0;
    __syncthreads();
    const int64_t n = (int64_t)blockIdx.x * kPartPoints + threadIdx.x;
    if (n < N) {
        float xv[3];
        load_point(x, n, xv);
        for (int l = lp; l < g.L; ++l) {
            int bx, by, bz;
            float fx, fy, fz;
            cell_of(xv, g, l, bx, by, bz, fx, fy, fz);
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const uint32_t h = hashed_index(g, bx + ((c >> 2) & 1), by + ((c >> 1) & 1), bz + (c & 1));
                atomicAdd(&s_hist[(l - lp) * nb_level + (int)(h >> kPartLog2Slots)], 1u);
            }
        }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nb; b += blockDim.x) hist[(int64_t)b * gridDim.x + blockIdx.x] = s_hist[b];
}

// hist[bucket][chunk] -> exclusive offset of each chunk's updates within the bucket (one workgroup
// per bucket); bucket_total gets each bucket's update count.
__global__ __launch_bounds__(256) void part_column_kernel(uint32_t* __restrict__ hist, int64_t n_chunks,
                                                          uint32_t* __restrict__ bucket_total) {
    __shared__ uint32_t s_wave[4];
    __shared__ uint32_t s_carry;
    uint32_t* h = hist + (int64_t)blockIdx.x * n_chunks;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
    for (int64_t c0 = 0; c0 < n_chunks; c0 += 256) {
        const int64_t k = c0 + threadIdx.x;
        const uint32_t v = k < n_chunks ? h[k] : 0u;
        uint32_t incl = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t u = __shfl_up(incl, o, 64);
            if (lane >= o) incl += u;
        }
        if (lane == 63) s_wave[w] = incl;
        __syncthreads();
        uint32_t pre = s_carry;
        for (int i = 0; i < w; ++i) pre += s_wave[i];
        if (k < n_chunks) h[k] = pre + incl - v;
        __syncthreads();
        if (threadIdx.x == 255) s_carry = pre + incl;
        __syncthreads();
    }
    if (threadIdx.x == 0) bucket_total[blockIdx.x] = s_carry;
}

__global__ __launch_bounds__(256) void part_bucket_scan_kernel(const uint32_t* __restrict__ bucket_total, int nb,
                                                               uint32_t* __restrict__ bucket_start) {
    __shared__ uint32_t s_part[256];
    // each lane owns a contiguous run of ceil(nb / 256) buckets
    const int per = (nb + 255) / 256;
    const int b0 = threadIdx.x * per;
    uint32_t sum = 0;
    for (int b = b0; b < b0 + per && b < nb; ++b) sum += bucket_total[b];
    s_part[threadIdx.x] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int t = 0; t < 256; ++t) {
            const uint32_t v = s_part[t];
            s_part[t] = run;
            run += v;
        }
        bucket_start[nb] = run;
    }
    __syncthreads();
    uint32_t run = s_part[threadIdx.x];
    for (int b = b0; b < b0 + per && b < nb; ++b) {
        bucket_start[b] = run;
        run += bucket_total[b];
    }
}

__global__ __launch_bounds__(kPartPoints) void part_scatter_kernel(const float* __restrict__ x, int64_t N,
                                                                   const float* __restrict__ gout, GridArgs g, int lp,
                                                                   int nb_level, int64_t n_chunks,
                                                                   const uint32_t* __restrict__ hist,
                                                                   const uint32_t* __restrict__ bucket_start,
                                                                   uint16_t* __restrict__ rslot,
                                                                   float2* __restrict__ rval) {
    constexpr int kRec = kPartPoints * 8;
    __shared__ uint16_t s_slot[kRec];
    __shared__ uint16_t s_bk[kRec];       // bucket of each staged record
    __shared__ float2 s_val[kRec];
    __shared__ uint32_t s_off[kPartMaxLevelBuckets + 1];   // counts, then exclusive offsets
    __shared__ uint32_t s_dst[kPartMaxLevelBuckets];       // each bucket's first record in rslot/rval
    __shared__ uint32_t s_wave[kPartPoints / 64];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int64_t n = (int64_t)blockIdx.x * kPartPoints + t;
    const bool valid = n < N;
    float xv[3] = {0.f, 0.f, 0.f};
    if (valid) load_point(x, n, xv);
    // the query's gradient rows of all partitioned levels at once: they are contiguous, and read
    // level by level they were fetched from HBM once per level (the lines do not survive a level)
    float2 gor[kPartMaxLevels];
#pragma unroll
    for (int i = 0; i < kPartMaxLevels; ++i)
        gor[i] = (valid && lp + i < g.L) ? *reinterpret_cast<const float2*>(gout + (n * g.L + lp + i) * 2)
                                         : make_float2(0.f, 0.f);
#pragma unroll
    for (int i = 0; i < kPartMaxLevels; ++i) {
        const int l = lp + i;
        if (l >= g.L) break;
        for (int b = t; b <= nb_level; b += blockDim.x) s_off[b] = 0;
        __syncthreads();
        int bx = 0, by = 0, bz = 0;
        float fx = 0.f, fy = 0.f, fz = 0.f;
        if (valid) cell_of(xv, g, l, bx, by, bz, fx, fy, fz);
        uint32_t h[8], pos[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            h[c] = hashed_index(g, bx + ((c >> 2) & 1), by + ((c >> 1) & 1), bz + (c & 1));
            pos[c] = valid ? atomicAdd(&s_off[h[c] >> kPartLog2Slots], 1u) : 0u;
        }
        __syncthreads();
        // exclusive scan of the counts (nb_level <= blockDim.x: one bucket per lane)
        const int lb = l - lp;
        const uint32_t cnt = t < nb_level ? s_off[t] : 0u;
        uint32_t incl = cnt;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t u = __shfl_up(incl, o, 64);
            if (lane >= o) incl += u;
        }
        if (lane == 63) s_wave[wave] = incl;
        __syncthreads();
        uint32_t pre = 0;
        for (int i = 0; i < wave; ++i) pre += s_wave[i];
        if (t < nb_level) {
            s_off[t] = pre + incl - cnt;
            const int64_t bk = (int64_t)lb * nb_level + t;
            s_dst[t] = bucket_start[bk] + hist[bk * n_chunks + blockIdx.x];
        }
        if (t == nb_level - 1) s_off[nb_level] = pre + incl;
        __syncthreads();
        if (valid) {
            const float2 go = gor[i];
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const int ox = (c >> 2) & 1, oy = (c >> 1) & 1, oz = c & 1;
                const float w = (ox ? fx : 1.0f - fx) * (oy ? fy : 1.0f - fy) * (oz ? fz : 1.0f - fz);
                const uint32_t bk = h[c] >> kPartLog2Slots;
                const uint32_t j = s_off[bk] + pos[c];
                s_slot[j] = (uint16_t)(h[c] & (kPartSlots - 1));
                s_bk[j] = (uint16_t)bk;
                s_val[j] = make_float2(w * go.x, w * go.y);
            }
        }
        __syncthreads();
        // write-out: consecutive staged records of a bucket go to consecutive records of its run
        const uint32_t total = s_off[nb_level];
        for (uint32_t j = t; j < total; j += blockDim.x) {
            const int bk = s_bk[j];
            const uint32_t dst = s_dst[bk] + (j - s_off[bk]);
            rslot[dst] = s_slot[j];
            rval[dst] = s_val[j];
        }
        __syncthreads();
    }
}

// The sums are kept in FP64: LDS float adds (ds_add_f32) run at about a quarter of the rate of
// FP64 ones (ds_add_f64) on this GPU, and the wider sum is rounded once when added to the table.
__global__ __launch_bounds__(kPartAccThreads) void part_accumulate_kernel(const uint32_t* __restrict__ bucket_start,
                                                                          const uint16_t* __restrict__ rslot,
                                                                          const float2* __restrict__ rval, GridArgs g,
                                                                          int lp, int nb_level,
                                                                          float* __restrict__ gtable) {
    // the two features in separate arrays: a wave's random slots then spread over the banks
    __shared__ double s_ax[kPartSlots], s_ay[kPartSlots];
    for (int s = threadIdx.x; s < kPartSlots; s += blockDim.x) {
        s_ax[s] = 0.0;
        s_ay[s] = 0.0;
    }
    __syncthreads();
    const int b = blockIdx.x;
    const uint32_t lo = bucket_start[b], hi = bucket_start[b + 1];
    // records [lo, a0) and [a1, hi) one per lane; [a0, a1) eight per lane with 16-byte loads (lane
    // t of a round reads records 8t..8t+7 of it, so a wave reads one contiguous stretch)
    uint32_t a0 = (lo + 7u) & ~7u;
    a0 = a0 < hi ? a0 : hi;
    const uint32_t a1 = a0 + ((hi - a0) & ~7u);
    for (uint32_t p = lo + threadIdx.x; p < a0; p += blockDim.x) {
        const int s = rslot[p];
        const float2 v = rval[p];
        atomicAdd(&s_ax[s], (double)v.x);
        atomicAdd(&s_ay[s], (double)v.y);
    }
    for (uint32_t p = a0 + 8u * threadIdx.x; p < a1; p += 8u * blockDim.x) {
        const uint4 sl = *reinterpret_cast<const uint4*>(rslot + p);
        float4 v[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) v[k] = *reinterpret_cast<const float4*>(rval + p + 2 * k);
        const uint32_t w4[4] = {sl.x, sl.y, sl.z, sl.w};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int s0 = (int)(w4[k] & 0xFFFFu), s1 = (int)(w4[k] >> 16);
            atomicAdd(&s_ax[s0], (double)v[k].x);
            atomicAdd(&s_ay[s0], (double)v[k].y);
            atomicAdd(&s_ax[s1], (double)v[k].z);
            atomicAdd(&s_ay[s1], (double)v[k].w);
        }
    }
    for (uint32_t p = a1 + threadIdx.x; p < hi; p += blockDim.x) {
        const int s = rslot[p];
        const float2 v = rval[p];
        atomicAdd(&s_ax[s], (double)v.x);
        atomicAdd(&s_ay[s], (double)v.y);
    }
    __syncthreads();
    const int l = lp + b / nb_level;
    float* tb = gtable + (g.offs[l] + (int64_t)(b % nb_level) * kPartSlots) * 2;
    for (int s = threadIdx.x; s < kPartSlots; s += blockDim.x) {   // this workgroup owns these slots
        const double vx = s_ax[s], vy = s_ay[s];
        if (vx != 0.0 || vy != 0.0) {
            float2 t = reinterpret_cast<float2*>(tb)[s];
            t.x = (float)((double)t.x + vx);
            t.y = (float)((double)t.y + vy);
            reinterpret_cast<float2*>(tb)[s] = t;
        }
    }
}

// First level of the partitioned suffix, or L when the partitioned backward does not apply:
// n_feat 2, hashed levels with res >= 512 (a suffix, as resolutions grow with the level), whole
// buckets per level, and few enough buckets for the LDS histograms.
int partition_first_level(const GridArgs& g, int log2_size) {
    if (g.F != 2 || log2_size < kPartLog2Slots) return g.L;
    int lp = g.L;
    while (lp > 0 && !g.dense[lp - 1] && g.res[lp - 1] >= 512 && g.L - lp < kPartMaxLevels) --lp;
    const int64_t nb = (int64_t)(g.L - lp) << (log2_size - kPartLog2Slots);
    if (lp == g.L || nb > kPartMaxBuckets || (1 << (log2_size - kPartLog2Slots)) > kPartMaxLevelBuckets) return g.L;
    return lp;
}

struct PartLayout {
    size_t hist, total, start, rslot, rval, bytes;
    int64_t n_chunks, n_records;
    int nb_level, nb;
};

PartLayout part_layout(const GridArgs& g, int log2_size, int lp, int64_t n) {
    PartLayout P;
    auto a256 = [](size_t v) { return (v + 255) & ~size_t(255); };
    P.nb_level = 1 << (log2_size - kPartLog2Slots);
    P.nb = (g.L - lp) * P.nb_level;
    P.n_chunks = (n + kPartPoints - 1) / kPartPoints;
    P.n_records = n * 8 * (int64_t)(g.L - lp);
    size_t o = 0;
    P.hist = o; o += a256(sizeof(uint32_t) * (size_t)P.n_chunks * P.nb);
    P.total = o; o += a256(sizeof(uint32_t) * (size_t)P.nb);
    P.start = o; o += a256(sizeof(uint32_t) * (size_t)(P.nb + 1));
    P.rslot = o; o += a256(sizeof(uint16_t) * (size_t)P.n_records);
    P.rval = o; o += a256(sizeof(float2) * (size_t)P.n_records);
    P.bytes = o;
    return P;
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

#define RM_FEAT_DISPATCH(F_, KERNEL, ...)                                                     \
    switch (F_) {                                                                             \
        case 1: hipLaunchKernelGGL(KERNEL<1>, __VA_ARGS__); break;                            \
        case 2: hipLaunchKernelGGL(KERNEL<2>, __VA_ARGS__); break;                            \
        case 4: hipLaunchKernelGGL(KERNEL<4>, __VA_ARGS__); break;                            \
        default: hipLaunchKernelGGL(KERNEL<8>, __VA_ARGS__); break;                           \
    }

// Backward of levels [0, l_end): workgroup aggregation for n_feat <= 2, else level-major.
int hashgrid_bwd_levels(const float* x, int64_t n, const float* gout, const GridArgs& g, int l_end, float* gtable,
                        hipStream_t st) {
    if (g.F <= 2) {
        dim3 grid((unsigned)ceil_div(n, kAggPoints)), block(kAggPoints);
        if (g.F == 1)
            hipLaunchKernelGGL(hashgrid_bwd_agg_kernel<1>, grid, block, 0, st, x, n, gout, g, l_end, gtable);
        else
            hipLaunchKernelGGL(hashgrid_bwd_agg_kernel<2>, grid, block, 0, st, x, n, gout, g, l_end, gtable);
        RM_CHECK_LAUNCH("hashgrid_bwd_agg_kernel");
        return RM_OK;
    }
    dim3 grid((unsigned)ceil_div(n, 256), (unsigned)l_end), block(256);  // level = blockIdx.y < l_end
    RM_FEAT_DISPATCH(g.F, hashgrid_bwd_kernel, grid, block, 0, st, x, n, gout, g, gtable);
    RM_CHECK_LAUNCH("hashgrid_bwd_kernel");
    return RM_OK;
}

// A second stream per device for the coarse-level aggregation, which then runs beside the
// fine-level ownership kernels: the aggregation is bound by vector and LDS work, the ownership
// kernels by HBM, and they write disjoint levels of the table. Created once per device (default
// priority, non-blocking), with the fork and join events; the mutex keeps one caller's
// record/wait sequence whole.
constexpr int kMaxSideDevices = 64;
struct SideStream {
    hipStream_t stream = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
};
std::mutex g_side_mu;
SideStream g_side[kMaxSideDevices];

// The device's side stream, or nullptr (run everything on the caller's stream) when it cannot be
// made.
SideStream* side_stream() {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxSideDevices) return nullptr;
    SideStream& s = g_side[dev];
    if (s.stream) return &s;
    hipStream_t st = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return nullptr;
    if (hipEventCreateWithFlags(&a, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b, hipEventDisableTiming) != hipSuccess) {
        if (a) (void)hipEventDestroy(a);
        (void)hipStreamDestroy(st);
        return nullptr;
    }
    s.fork = a;
    s.join = b;
    s.stream = st;
    return &s;
}

}  // namespace
}  // namespace rmesh

using namespace rmesh;

extern "C" int rm_hashgrid_fwd(const float* x_dev, int64_t n, const float* table_dev, const int32_t* res,
                               const int64_t* offs, int32_t n_levels, int32_t n_feat, int32_t log2_size,
                               float* out_dev, rm_stream_t stream) {
    RM_CHECK_ARG(n >= 0, "rm_hashgrid_fwd: negative n");
    GridArgs g;
    int e = make_grid_args(res, offs, n_levels, n_feat, log2_size, g, "rm_hashgrid_fwd");
    if (e) return e;
    if (n == 0) return RM_OK;
    RM_CHECK_ARG(x_dev && table_dev && out_dev, "rm_hashgrid_fwd: null buffer");
    if (n_levels * n_feat <= kLmMaxRow) {
        dim3 grid((unsigned)ceil_div(n, kLmPoints)), block(kLmPoints * kLmWaves);
        RM_FEAT_DISPATCH(n_feat, hashgrid_fwd_lm_kernel, grid, block, 0, as_stream(stream), x_dev, n, table_dev, g,
                         out_dev);
        RM_CHECK_LAUNCH("hashgrid_fwd_lm_kernel");
        return RM_OK;
    }
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
    return hashgrid_bwd_levels(x_dev, n, grad_out_dev, g, n_levels, grad_table_dev, as_stream(stream));
}

extern "C" size_t rm_hashgrid_bwd_workspace_size(int64_t n, const int32_t* res, const int64_t* offs, int32_t n_levels,
                                                 int32_t n_feat, int32_t log2_size) {
    GridArgs g;
    if (n <= 0 || make_grid_args(res, offs, n_levels, n_feat, log2_size, g, "rm_hashgrid_bwd_workspace_size"))
        return 0;
    const int lp = partition_first_level(g, log2_size);
    if (lp == g.L) return 0;
    const PartLayout P = part_layout(g, log2_size, lp, n);
    return P.n_records < (1ll << 31) ? P.bytes : 0;
}

extern "C" int rm_hashgrid_bwd_ws(const float* x_dev, int64_t n, const float* grad_out_dev, const int32_t* res,
                                  const int64_t* offs, int32_t n_levels, int32_t n_feat, int32_t log2_size,
                                  float* grad_table_dev, void* workspace_dev, size_t workspace_bytes,
                                  rm_stream_t stream) {
    RM_CHECK_ARG(n >= 0, "rm_hashgrid_bwd_ws: negative n");
    GridArgs g;
    int e = make_grid_args(res, offs, n_levels, n_feat, log2_size, g, "rm_hashgrid_bwd_ws");
    if (e) return e;
    if (n == 0) return RM_OK;
    RM_CHECK_ARG(x_dev && grad_out_dev && grad_table_dev, "rm_hashgrid_bwd_ws: null buffer");
    const size_t need = rm_hashgrid_bwd_workspace_size(n, res, offs, n_levels, n_feat, log2_size);
    hipStream_t st = as_stream(stream);
    if (need == 0) return hashgrid_bwd_levels(x_dev, n, grad_out_dev, g, n_levels, grad_table_dev, st);
    if (!workspace_dev || workspace_bytes < need) {
        set_error("rm_hashgrid_bwd_ws: workspace too small");
        return RM_ERR_WORKSPACE_TOO_SMALL;
    }
    RM_CHECK_ARG(((uintptr_t)workspace_dev & 255) == 0 && ((uintptr_t)grad_out_dev & 7) == 0,
                 "rm_hashgrid_bwd_ws: misaligned workspace or grad_out");
    const int lp = partition_first_level(g, log2_size);
    const PartLayout P = part_layout(g, log2_size, lp, n);
    char* w = static_cast<char*>(workspace_dev);
    uint32_t* hist = reinterpret_cast<uint32_t*>(w + P.hist);
    uint32_t* total = reinterpret_cast<uint32_t*>(w + P.total);
    uint32_t* start = reinterpret_cast<uint32_t*>(w + P.start);
    uint16_t* rslot = reinterpret_cast<uint16_t*>(w + P.rslot);
    float2* rval = reinterpret_cast<float2*>(w + P.rval);
    // coarse levels: workgroup aggregation, on the side stream between a fork and a join on st
    // (the lock holds from the fork's record to the join's wait)
    std::unique_lock<std::mutex> side_lock(g_side_mu);
    SideStream* side = lp > 0 ? side_stream() : nullptr;
    if (side && (hipEventRecord(side->fork, st) != hipSuccess ||
                 hipStreamWaitEvent(side->stream, side->fork, 0) != hipSuccess)) {
        (void)hipGetLastError();
        side = nullptr;
    }
    if (!side) side_lock.unlock();
    if (lp > 0) {
        int r = hashgrid_bwd_levels(x_dev, n, grad_out_dev, g, lp, grad_table_dev, side ? side->stream : st);
        if (r == RM_OK && side && hipEventRecord(side->join, side->stream) != hipSuccess) {
            set_error("rm_hashgrid_bwd_ws: recording the aggregation stream's join failed");
            r = RM_ERR_HIP;
        }
        if (r) return r;
    }
    hipLaunchKernelGGL(part_count_kernel, dim3((unsigned)P.n_chunks), dim3(kPartPoints), 0, st, x_dev, n, g, lp,
                       P.nb_level, hist);
    hipLaunchKernelGGL(part_column_kernel, dim3((unsigned)P.nb), dim3(256), 0, st, hist, P.n_chunks, total);
    hipLaunchKernelGGL(part_bucket_scan_kernel, dim3(1), dim3(256), 0, st, (const uint32_t*)total, P.nb, start);
    hipLaunchKernelGGL(part_scatter_kernel, dim3((unsigned)P.n_chunks), dim3(kPartPoints), 0, st, x_dev, n,
                       grad_out_dev, g, lp, P.nb_level, P.n_chunks, (const uint32_t*)hist, (const uint32_t*)start,
                       rslot, rval);
    hipLaunchKernelGGL(part_accumulate_kernel, dim3((unsigned)P.nb), dim3(kPartAccThreads), 0, st, (const uint32_t*)start,
                       (const uint16_t*)rslot, (const float2*)rval, g, lp, P.nb_level, grad_table_dev);
    if (side && hipStreamWaitEvent(st, side->join, 0) != hipSuccess) {
        set_error("rm_hashgrid_bwd_ws: waiting for the aggregation stream failed");
        return RM_ERR_HIP;
    }
    RM_CHECK_LAUNCH("rm_hashgrid_bwd_ws");
    return RM_OK;
}

// Field MLP entry points: filled in by mlp.hip.
