// Scans, stable LSD radix sort, key generation, tile ranges and the fused binning path.
//
// Radix sort: 8-bit (or narrower) digits, 2048 keys per 256-thread workgroup. Each pass is
//   (1) per-workgroup digit histogram in LDS,
//   (2) device-wide exclusive scan of the digit-major [digit][workgroup] count table,
//   (3) stable scatter: wave-level digit matching with ballots gives each key its rank among
//       equal digits in original order, keys are first sorted by digit in LDS, then written
//       out so that consecutive lanes store consecutive addresses within a digit run.
// The result equals a stable sort (equal keys keep input order), the contract of
// cub::DeviceRadixSort::SortPairs used by sort_by_keys.cu:36-48.
#include "common.h"

namespace rmesh {
namespace {

constexpr int kThreads = 256;
constexpr int kScanItems = 4;                       // per thread
constexpr int kScanTile = kThreads * kScanItems;    // 1024
constexpr int kRsItems = 8;
constexpr int kRsTile = kThreads * kRsItems;        // 2048

__device__ inline int lane_id() { return threadIdx.x & 63; }

// Element count of a launch sized for a capacity: the capacity itself, or a count the device
// wrote earlier in the stream (graph-capturable launches never read sizes back to the host).
__device__ inline int64_t eff_n(int64_t cap, const int64_t* n_dev) {
    if (!n_dev) return cap;
    int64_t v = *n_dev;
    return v < cap ? v : cap;
}

template <typename T>
__device__ inline T wave_incl_scan(T v) {
    int lane = lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        T u = __shfl_up(v, o, 64);
        if (lane >= o) v += u;
    }
    return v;
}

// Exclusive scan over a 256-thread block; lds must hold 4 elements.
template <typename T>
__device__ inline T block_excl_scan(T v, T& total, T* lds) {
    int lane = lane_id(), w = threadIdx.x >> 6;
    T inc = wave_incl_scan(v);
    if (lane == 63) lds[w] = inc;
    __syncthreads();
    T pre = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        T s = lds[i];
        if (i < w) pre += s;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return pre + inc - v;
}

// ---------------------------------------------------------------- device-wide scan ------

struct LoadU32 {
    const uint32_t* p;
    __device__ uint32_t operator()(int64_t i) const { return p[i]; }
};
struct LoadI32 {
    const int32_t* p;
    __device__ uint32_t operator()(int64_t i) const { return (uint32_t)p[i]; }
};
struct LoadGather {  // tiles_touched[order[i]]
    const int32_t* tiles;
    const uint32_t* order;
    __device__ uint32_t operator()(int64_t i) const { return (uint32_t)tiles[order[i]]; }
};

template <typename Load>
__global__ __launch_bounds__(kThreads) void scan_reduce_kernel(Load ld, int64_t cap, const int64_t* n_dev,
                                                               uint32_t* partial) {
    __shared__ uint32_t lds[4];
    const int64_t n = eff_n(cap, n_dev);
    int64_t base = (int64_t)blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (base + k < n) s += ld(base + k);
    uint32_t tot;
    block_excl_scan(s, tot, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// Single workgroup: exclusive scan of partial[0..m) in place; writes the grand total.
__global__ __launch_bounds__(kThreads) void scan_partials_kernel(uint32_t* partial, int64_t m,
                                                                 int64_t* total_out) {
    __shared__ uint32_t lds[4];
    uint32_t carry = 0;
    for (int64_t c = 0; c < m; c += kScanTile) {
        int64_t base = c + threadIdx.x * kScanItems;
        uint32_t v[kScanItems];
        uint32_t s = 0;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            v[k] = (base + k < m) ? partial[base + k] : 0u;
            s += v[k];
        }
        uint32_t tot;
        uint32_t pre = block_excl_scan(s, tot, lds) + carry;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            if (base + k < m) partial[base + k] = pre;
            pre += v[k];
        }
        carry += tot;
    }
    if (threadIdx.x == 0 && total_out) *total_out = (int64_t)carry;
}

template <typename Load, typename Out>
__global__ __launch_bounds__(kThreads) void scan_down_kernel(Load ld, int64_t cap, const int64_t* n_dev,
                                                             const uint32_t* partial, Out* out, int inclusive) {
    __shared__ uint32_t lds[4];
    const int64_t n = eff_n(cap, n_dev);
    int64_t base = (int64_t)blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = (base + k < n) ? ld(base + k) : 0u;
        s += v[k];
    }
    uint32_t tot;
    uint32_t pre = block_excl_scan(s, tot, lds) + partial[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < n) out[base + k] = (Out)(inclusive ? pre + v[k] : pre);
        pre += v[k];
    }
}

inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

inline int64_t scan_blocks(int64_t n) { return ceil_div(n, kScanTile); }
inline size_t scan_ws_bytes(int64_t n) { return align256(sizeof(uint32_t) * (size_t)(scan_blocks(n) + 1)); }

// n is the element count, or the capacity when n_dev points at the device-side count.
template <typename Load, typename Out>
hipError_t device_scan(Load ld, int64_t n, Out* out, bool inclusive, uint32_t* partial,
                       int64_t* total_out, hipStream_t st, const int64_t* n_dev = nullptr) {
    if (n <= 0) return hipSuccess;
    int64_t nb = scan_blocks(n);
    hipLaunchKernelGGL(scan_reduce_kernel<Load>, dim3((unsigned)nb), dim3(kThreads), 0, st, ld, n, n_dev, partial);
    hipLaunchKernelGGL(scan_partials_kernel, dim3(1), dim3(kThreads), 0, st, partial, nb, total_out);
    hipLaunchKernelGGL((scan_down_kernel<Load, Out>), dim3((unsigned)nb), dim3(kThreads), 0, st, ld, n, n_dev,
                       (const uint32_t*)partial, out, inclusive ? 1 : 0);
    return hipGetLastError();
}

// ---------------------------------------------------------------- radix sort ------------

// Per-tile digit counts, digit-major over the tiles in use: counts[d * nbu + tile] with nbu the
// tiles that hold keys (all nb of a launch without a device count; with one, a launch sized for a
// capacity writes and scans only the counts of the tiles in use, and len_out receives their number
// times the digit count, the length of the scan).
template <typename K>
__global__ __launch_bounds__(kThreads) void rs_hist_kernel(const K* __restrict__ keys, int64_t cap,
                                                           const int64_t* n_dev, int shift, int bits,
                                                           uint32_t* counts, int64_t* len_out) {
    __shared__ uint32_t h[256];
    const int64_t n = eff_n(cap, n_dev);
    const int64_t nbu = (n + kRsTile - 1) / kRsTile;
    if (len_out && blockIdx.x == 0 && threadIdx.x == 0) *len_out = ((int64_t)1 << bits) * nbu;
    if ((int64_t)blockIdx.x >= nbu) return;
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t mask = (1u << bits) - 1u;
    int64_t base = (int64_t)blockIdx.x * kRsTile + threadIdx.x;
#pragma unroll
    for (int r = 0; r < kRsItems; ++r) {
        int64_t i = base + r * kThreads;
        if (i < n) atomicAdd(&h[(uint32_t)(keys[i] >> shift) & mask], 1u);
    }
    __syncthreads();
    if (threadIdx.x < (1 << bits)) counts[(int64_t)threadIdx.x * nbu + blockIdx.x] = h[threadIdx.x];
}

template <typename K>
__global__ __launch_bounds__(kThreads) void rs_scatter_kernel(
    const K* __restrict__ kin, const uint32_t* __restrict__ vin, K* __restrict__ kout,
    uint32_t* __restrict__ vout, int64_t cap, const int64_t* n_dev, int shift, int bits,
    const uint32_t* __restrict__ offsets) {
    const int64_t n = eff_n(cap, n_dev);
    const int64_t nbu = (n + kRsTile - 1) / kRsTile;   // tiles in use (the offsets' row length)
    __shared__ uint32_t s_cnt[256];
    __shared__ uint32_t s_wcnt[4][256];
    __shared__ uint32_t s_loc[256];
    __shared__ uint32_t s_glob[256];
    __shared__ uint32_t s_tmp[4];
    __shared__ K s_keys[kRsTile];
    __shared__ uint32_t s_vals[kRsTile];

    const int tid = threadIdx.x, lane = lane_id(), w = tid >> 6;
    const uint32_t mask = (1u << bits) - 1u;
    const int nbins = 1 << bits;
    const int64_t base = (int64_t)blockIdx.x * kRsTile;
    if (base >= n) return;   // a launch sized for a capacity: nothing of this tile is in use
    s_cnt[tid] = 0;
    s_glob[tid] = (tid < nbins) ? offsets[(int64_t)tid * nbu + blockIdx.x] : 0u;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));

    K key[kRsItems];
    uint32_t val[kRsItems], rank[kRsItems], dig[kRsItems];
#pragma unroll
    for (int r = 0; r < kRsItems; ++r) {
        int64_t i = base + r * kThreads + tid;
        bool valid = i < n;
        key[r] = valid ? kin[i] : K(0);
        val[r] = valid ? vin[i] : 0u;
        dig[r] = valid ? ((uint32_t)(key[r] >> shift) & mask) : 0xFFFFFFFFu;
#pragma unroll
        for (int q = 0; q < 4; ++q) s_wcnt[q][tid] = 0;
        __syncthreads();
        uint64_t peers = __ballot(valid);
        for (int b = 0; b < bits; ++b) {
            bool bit = valid && ((dig[r] >> b) & 1u);
            uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        uint32_t wr = (uint32_t)__popcll(peers & lt);
        if (valid && wr == 0) s_wcnt[w][dig[r]] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pre = s_cnt[dig[r]];
            for (int q = 0; q < w; ++q) pre += s_wcnt[q][dig[r]];
            rank[r] = pre + wr;
        } else {
            rank[r] = 0;
        }
        __syncthreads();
        s_cnt[tid] += s_wcnt[0][tid] + s_wcnt[1][tid] + s_wcnt[2][tid] + s_wcnt[3][tid];
    }
    __syncthreads();
    uint32_t tot;
    s_loc[tid] = block_excl_scan(s_cnt[tid], tot, s_tmp);
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kRsItems; ++r) {
        if (dig[r] != 0xFFFFFFFFu) {
            uint32_t pos = s_loc[dig[r]] + rank[r];
            s_keys[pos] = key[r];
            s_vals[pos] = val[r];
        }
    }
    __syncthreads();
    int64_t cnt = n - base < kRsTile ? n - base : kRsTile;
    if (cnt <= 0) return;
#pragma unroll
    for (int r = 0; r < kRsItems; ++r) {
        int j = r * kThreads + tid;
        if (j < cnt) {
            K k = s_keys[j];
            uint32_t d = (uint32_t)(k >> shift) & mask;
            uint32_t o = s_glob[d] + ((uint32_t)j - s_loc[d]);
            kout[o] = k;
            vout[o] = s_vals[j];
        }
    }
}

inline int64_t rs_blocks(int64_t n) { return ceil_div(n, kRsTile); }
// the digit counts, then the scan length of a sort with a device count (one int64)
inline size_t rs_counts_bytes(int64_t n) {
    return align256(sizeof(uint32_t) * 256 * (size_t)rs_blocks(n) + sizeof(int64_t));
}

// Stable LSD sort of (key, value) over key bits [begin_bit, end_bit). kin/vin are not modified;
// the result lands in kout/vout. ktmp/vtmp are a ping-pong buffer of n entries; all three
// buffer pairs must be distinct.
template <typename K>
hipError_t radix_sort(const K* kin, const uint32_t* vin, K* kout, uint32_t* vout, K* ktmp, uint32_t* vtmp,
                      int64_t n, int begin_bit, int end_bit, uint32_t* counts, uint32_t* partial,
                      hipStream_t st, const int64_t* n_dev = nullptr) {
    if (n <= 0) return hipSuccess;
    int nbits = end_bit - begin_bit;
    if (nbits <= 0 && !n_dev) {
        hipMemcpyAsync(kout, kin, sizeof(K) * n, hipMemcpyDeviceToDevice, st);
        hipMemcpyAsync(vout, vin, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, st);
        return hipGetLastError();
    }
    int passes = (nbits + 7) / 8;
    int64_t nb = rs_blocks(n);
    int64_t* len_dev = n_dev ? reinterpret_cast<int64_t*>(counts + 256 * nb) : nullptr;
    const K* ks = kin;
    const uint32_t* vs = vin;
    int shift = begin_bit;
    for (int p = 0; p < passes; ++p) {
        int bits = (nbits - (shift - begin_bit) + (passes - p) - 1) / (passes - p);  // split evenly
        bool to_out = ((passes - 1 - p) % 2) == 0;
        K* kd = to_out ? kout : ktmp;
        uint32_t* vd = to_out ? vout : vtmp;
        hipLaunchKernelGGL(rs_hist_kernel<K>, dim3((unsigned)nb), dim3(kThreads), 0, st, ks, n, n_dev, shift,
                           bits, counts, len_dev);
        hipError_t e = device_scan(LoadU32{counts}, (int64_t)(1 << bits) * nb, counts, false, partial,
                                   (int64_t*)nullptr, st, (const int64_t*)len_dev);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(rs_scatter_kernel<K>, dim3((unsigned)nb), dim3(kThreads), 0, st, ks, vs, kd, vd, n,
                           n_dev, shift, bits, (const uint32_t*)counts);
        ks = kd;
        vs = vd;
        shift += bits;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------- keys / ranges ----------

__device__ inline uint32_t sortable_key(float v) {  // tile_shader.slang:18-47
    uint32_t b = __float_as_uint(v);
    return (b & 0x80000000u) ? ~b : (b ^ 0x80000000u);
}

__global__ __launch_bounds__(kThreads) void generate_keys_kernel(
    const float* __restrict__ vs, const int4* __restrict__ rect, const int32_t* __restrict__ offs,
    int64_t T, int32_t gw, int64_t* __restrict__ keys, int32_t* __restrict__ vals) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    int64_t o = (t == 0) ? 0 : offs[t - 1];
    int4 r = rect[t];
    uint64_t fk = sortable_key(vs[3 * t + 2]);
    for (int32_t y = r.y; y < r.w; ++y)
        for (int32_t x = r.x; x < r.z; ++x) {
            uint64_t key = (uint64_t)(uint32_t)(y * gw + x);
            keys[o] = (int64_t)((key << 32) | fk);
            vals[o] = (int32_t)t;
            ++o;
        }
}

__global__ void zero_i64_kernel(int64_t* p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

__global__ void zero_ranges_kernel(int2* ranges, int64_t G) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < G) ranges[i] = make_int2(0, 0);
}

// Grid-stride over the sorted entries (a launch sized for a capacity may use fewer workgroups).
template <typename TileOf>
__global__ __launch_bounds__(kThreads) void ranges_kernel(TileOf tile_of, int64_t cap, const int64_t* n_dev,
                                                          int32_t* ranges) {
    const int64_t n = eff_n(cap, n_dev);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        uint32_t cur = tile_of(i);
        if (i == 0) ranges[2 * (int64_t)cur] = 0;
        else {
            uint32_t prev = tile_of(i - 1);
            if (cur != prev) {
                ranges[2 * (int64_t)prev + 1] = (int32_t)i;
                ranges[2 * (int64_t)cur] = (int32_t)i;
            }
        }
        if (i == n - 1) ranges[2 * (int64_t)cur + 1] = (int32_t)n;
    }
}

struct TileOfKey {
    const int64_t* k;
    __device__ uint32_t operator()(int64_t i) const { return (uint32_t)((uint64_t)k[i] >> 32); }
};
struct TileOfU32 {
    const uint32_t* t;
    __device__ uint32_t operator()(int64_t i) const { return t[i]; }
};

// ---------------------------------------------------------------- fused binning ----------

// Per 1024-tet block: number of visible tets and their pair count.
__global__ __launch_bounds__(kThreads) void bin_count_kernel(const int32_t* __restrict__ tiles, int64_t T,
                                                             uint32_t* blk_vis, uint32_t* blk_pairs) {
    __shared__ uint32_t lds[4];
    int64_t base = (int64_t)blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    uint32_t nv = 0, np = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (base + k < T) {
            int32_t c = tiles[base + k];
            nv += c > 0;
            np += c > 0 ? (uint32_t)c : 0u;
        }
    uint32_t tv, tp;
    block_excl_scan(nv, tv, lds);
    block_excl_scan(np, tp, lds);
    if (threadIdx.x == 0) {
        blk_vis[blockIdx.x] = tv;
        blk_pairs[blockIdx.x] = tp;
    }
}

__device__ inline uint32_t blocks_of_rect(int4 r);

// Order-preserving compaction of the visible tets into (sortable depth key, tet id); with bcnt,
// also each visible tet's number of overlapped 8x8-tile blocks (indexed by tet).
__global__ __launch_bounds__(kThreads) void compact_kernel(const int32_t* __restrict__ tiles,
                                                           const float* __restrict__ vs, int dstride, int64_t T,
                                                           const uint32_t* __restrict__ blk_vis_off,
                                                           uint32_t* __restrict__ vkey, uint32_t* __restrict__ vidx,
                                                           const int4* __restrict__ rect = nullptr,
                                                           uint32_t* __restrict__ bcnt = nullptr) {
    __shared__ uint32_t lds[4];
    int64_t base = (int64_t)blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    bool vis[kScanItems];
    uint32_t nv = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        vis[k] = (base + k < T) && tiles[base + k] > 0;
        nv += vis[k];
    }
    uint32_t tot;
    uint32_t pos = block_excl_scan(nv, tot, lds) + blk_vis_off[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (vis[k]) {
            vkey[pos] = sortable_key(vs[dstride * (base + k) + (dstride - 1)]);
            vidx[pos] = (uint32_t)(base + k);
            if (bcnt) bcnt[base + k] = blocks_of_rect(rect[base + k]);
            ++pos;
        }
}

// Expands depth-ordered visible tets into (tile id, tet id) pairs, row-major over each rect.
// Work is balanced by pairs: a workgroup takes chunks of K pairs [c*K, (c+1)*K) in grid strides;
// its first lane finds the tets covering a chunk by binary search over the exclusive pair offsets,
// their offsets and rects are staged in LDS (at most K tets, each owning at least one pair), and
// lanes take the chunk's pairs round-robin, locating their tet by binary search in LDS. A
// screen-filling tet (tens of thousands of pairs) is thus spread over many chunks.
constexpr int kPairsPerBlock = 1024;
// Workgroups of the grid-stride kernels whose launches are sized by a pair capacity (pair
// expansion, block ranges, segment kernels): enough to fill the chip many times over, few enough
// that a capacity far above the pairs in use launches no flood of idle workgroups.
constexpr int64_t kMaxStrideGrid = 16384;

__device__ inline int64_t last_le(const uint32_t* __restrict__ a, int64_t n, uint32_t v) {
    int64_t lo = 0, hi = n - 1;  // a[0] == 0 <= v
    while (lo < hi) {
        int64_t mid = (lo + hi + 1) >> 1;
        if (a[mid] <= v) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Id maps of the pair expansion: for a tet, the (first id, ids per row) of its view's grid, so
// that cell (x, y) of the tet's rect gets id first + y * row + x; looked up once per tet.
struct TileMapSingle {  // one view's grid of gw cells per row
    static constexpr bool kPerTet = false;
    int32_t gw;
    __device__ uint2 operator()(uint32_t) const { return make_uint2(0u, (uint32_t)gw); }
};

struct RectTiles {  // the tet's rect of tiles
    __device__ int4 operator()(int4 r) const { return r; }
};

template <typename TileMap, typename RectMap = RectTiles>
__global__ __launch_bounds__(kThreads) void emit_pairs_kernel(
    const uint32_t* __restrict__ order, const uint32_t* __restrict__ poff, const int4* __restrict__ rect,
    int64_t nvis_cap, const int64_t* nvis_dev, int64_t P_cap, const int64_t* P_dev, TileMap tile_of,
    uint32_t* __restrict__ ptile, uint32_t* __restrict__ ptet, RectMap rect_of = RectMap{}) {
    __shared__ uint32_t s_off[kPairsPerBlock];
    __shared__ int4 s_rect[kPairsPerBlock];
    __shared__ uint32_t s_tet[kPairsPerBlock];
    __shared__ uint2 s_map[TileMap::kPerTet ? kPairsPerBlock : 1];
    __shared__ int64_t s_lo;
    __shared__ int s_n;
    const int64_t nvis = eff_n(nvis_cap, nvis_dev);
    const int64_t P = eff_n(P_cap, P_dev);
    if (nvis <= 0) return;
    // chunks of kPairsPerBlock pairs in grid strides (a launch sized for a capacity stays small)
    for (int64_t p0 = (int64_t)blockIdx.x * kPairsPerBlock; p0 < P; p0 += (int64_t)gridDim.x * kPairsPerBlock) {
        const int64_t p1 = (p0 + kPairsPerBlock < P) ? p0 + kPairsPerBlock : P;
        if (threadIdx.x == 0) {
            int64_t lo = last_le(poff, nvis, (uint32_t)p0);
            int64_t hi = last_le(poff, nvis, (uint32_t)(p1 - 1));
            s_lo = lo;
            s_n = (int)(hi - lo + 1);
        }
        __syncthreads();
        const int64_t lo = s_lo;
        const int n = s_n;
        for (int j = threadIdx.x; j < n; j += kThreads) {
            uint32_t t = order[lo + j];
            s_off[j] = poff[lo + j];
            s_tet[j] = t;
            s_rect[j] = rect_of(rect[t]);
            if constexpr (TileMap::kPerTet) s_map[j] = tile_of(t);
        }
        __syncthreads();
        for (int64_t p = p0 + threadIdx.x; p < p1; p += kThreads) {
            const uint32_t pp = (uint32_t)p;
            int a = 0, b = n - 1;  // last j with s_off[j] <= p
            while (a < b) {
                int mid = (a + b + 1) >> 1;
                if (s_off[mid] <= pp) a = mid;
                else b = mid - 1;
            }
            const int4 r = s_rect[a];
            const uint32_t m = pp - s_off[a];
            const uint32_t wdt = (uint32_t)(r.z - r.x);
            const uint32_t y = (uint32_t)r.y + m / wdt, x = (uint32_t)r.x + m % wdt;
            const uint2 mp = TileMap::kPerTet ? s_map[a] : tile_of(0u);
            ptile[p] = mp.x + y * mp.y + x;
            ptet[p] = s_tet[a];
        }
        __syncthreads();   // the chunk's LDS records are rewritten by the next one
    }
}

// ---- tile lists without a sort of the pairs: blocks of 8x8 tiles --------------------------
//
// The depth-ordered visible tets are first expanded into (block, tet) pairs over the blocks of
// 8x8 tiles their rects overlap (about 1.8 per tet at 4x4-pixel tiles, a quarter of the tile
// pairs) and stably sorted by block (two radix passes over the block id). One workgroup per block
// then walks the block's tets, still in depth order: a first walk counts the pairs of each of its
// 64 tiles, a scan over all tiles gives each tile's first position, and a second walk writes every
// tet to its tiles' lists directly. Within a chunk of 256 entries an entry's place in a tile's
// list is the number of earlier entries that cover the tile: each lane packs its coverage of the
// 64 tiles as 16 registers of four 8-bit counters and a wave-wide inclusive scan of those
// registers sums all 64 counters at once (a field never exceeds 255). The lists equal those of a
// stable sort of the (tile, tet) pairs by tile.
constexpr int kBinBlock = kBinBlockTiles;            // tiles per block side
constexpr int kBinLocal = kBinBlock * kBinBlock;     // tiles per block
constexpr int kBinRegs = kBinLocal / 4;              // packed 8-bit counters per lane

__device__ inline int4 block_rect_of(int4 r) {       // rect of tiles (half-open) -> rect of blocks
    return make_int4(r.x / kBinBlock, r.y / kBinBlock, (r.z - 1) / kBinBlock + 1, (r.w - 1) / kBinBlock + 1);
}
struct RectBlocks {
    __device__ int4 operator()(int4 r) const { return block_rect_of(r); }
};
__device__ inline uint32_t blocks_of_rect(int4 r) {
    const int4 b = block_rect_of(r);
    return (uint32_t)((b.z - b.x) * (b.w - b.y));
}
struct LoadBlockCount {  // blocks overlapped by tet order[i] (bcnt is indexed by tet)
    const uint32_t* bcnt;
    const uint32_t* order;
    __device__ uint32_t operator()(int64_t i) const { return bcnt[order[i]]; }
};

// The tet's tiles within block (bx, by), as local tile coordinates [lx0, lx1) x [ly0, ly1).
__device__ inline int4 local_rect(int4 r, int bx, int by) {
    const int ox = bx * kBinBlock, oy = by * kBinBlock;
    return make_int4(max(r.x - ox, 0), max(r.y - oy, 0), min(r.z - ox, kBinBlock), min(r.w - oy, kBinBlock));
}

// Tile grid of the view that a block or a tile belongs to: tile grid gw x gh, blocks per row bgw,
// and the view's first global tile and block (0 for a single view; a batch numbers the tiles and
// blocks of its views one view after the other).
struct ViewGrid {
    int32_t gw, gh, bgw;
    int64_t tile_base, block_base;
};
struct GridSingle {
    int32_t gw, gh, bgw;
    __device__ ViewGrid of_block(int64_t) const { return {gw, gh, bgw, 0, 0}; }
    __device__ ViewGrid of_tile(int64_t) const { return {gw, gh, bgw, 0, 0}; }
    __device__ int64_t n_tiles() const { return (int64_t)gw * gh; }
};
struct GridBatch {   // views of a batch plan, looked up by global block or tile id
    const void* plan;
    __device__ ViewGrid at(int v) const {
        const ViewRec& r = plan_views(plan)[v];
        return {r.cam.gw, r.cam.gh, bin_blocks_per_row(r.cam.gw), r.tile_offset, r.bin_offset};
    }
    template <class Off>
    __device__ int last_view(Off off, int64_t id) const {   // last view whose first id is <= id
        const ViewRec* v = plan_views(plan);
        int lo = 0, hi = static_cast<const PlanHeader*>(plan)->n_views - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (off(v[mid]) <= id) lo = mid;
            else hi = mid - 1;
        }
        return lo;
    }
    __device__ ViewGrid of_block(int64_t b) const {
        return at(last_view([](const ViewRec& r) { return r.bin_offset; }, b));
    }
    __device__ ViewGrid of_tile(int64_t t) const {
        return at(last_view([](const ViewRec& r) { return r.tile_offset; }, t));
    }
    __device__ int64_t n_tiles() const { return static_cast<const PlanHeader*>(plan)->total_tiles; }
};
struct BlockMapBatch {  // blocks of the tet's view in a batch (first global block, blocks per row),
    static constexpr bool kPerTet = true;   // written per tet by the batched projection
    const uint2* tet_blocks;
    __device__ uint2 operator()(uint32_t tet) const { return tet_blocks[tet]; }
};

// A block's list is cut into segments of kSegEntries; seg_first[b] (an exclusive scan over the
// blocks) numbers them, so that every segment is one workgroup of the count and emit kernels.
constexpr int kSegEntries = kThreads;

struct LoadSegCount {  // segments of block i
    const int2* branges;
    __device__ uint32_t operator()(int64_t i) const {
        const int2 r = branges[i];
        return (uint32_t)((r.y - r.x + kSegEntries - 1) / kSegEntries);
    }
};

// seg_desc[g] = (first entry, end entry, block, block's first segment) of segment g; one thread
// per block writes its segments' descriptors.
__global__ __launch_bounds__(kThreads) void seg_desc_kernel(const int2* __restrict__ branges,
                                                            const uint32_t* __restrict__ seg_first, int64_t nb,
                                                            int4* __restrict__ seg_desc) {
    const int64_t b = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (b >= nb) return;
    const int2 r = branges[b];
    const uint32_t f = seg_first[b], n = (uint32_t)((r.y - r.x + kSegEntries - 1) / kSegEntries);
    for (uint32_t j = 0; j < n; ++j) {
        const int e0 = r.x + (int)j * kSegEntries;
        seg_desc[f + j] = make_int4(e0, min(e0 + kSegEntries, r.y), (int)b, (int)f);
    }
}

// Workgroups of the segment kernels: one per segment up to an estimate of the segment count from
// the pair count (the blocks hold about a quarter of the tile pairs at 8x8 tiles per block), each
// looping over further segments in grid strides when there are more.
inline int64_t seg_grid(int64_t n_pairs, int64_t nb) {
    return std::min<int64_t>(ceil_div(n_pairs, 4 * kSegEntries) + nb, kMaxStrideGrid);
}

// Column-major per-segment counts, 0 past the last segment (those entries are never written).
struct LoadSegColumn {
    const uint32_t* cnt;
    int64_t stride;
    const int64_t* nseg_dev;
    __device__ uint32_t operator()(int64_t i) const { return (i % stride) < *nseg_dev ? cnt[i] : 0u; }
};

// Pairs of each of the block's 64 tiles within segment g, column-major: cnt[t * stride + g].
template <class Grid>
__global__ __launch_bounds__(kThreads) void seg_count_kernel(const uint32_t* __restrict__ blist,
                                                             const int4* __restrict__ seg_desc,
                                                             const int64_t* __restrict__ nseg_dev,
                                                             const int4* __restrict__ rect, Grid grid,
                                                             int64_t stride, uint32_t* __restrict__ cnt) {
    __shared__ uint32_t s_cnt[kBinLocal];
    const int64_t nseg = *nseg_dev;
    const int tid = threadIdx.x;
    for (int64_t g = blockIdx.x; g < nseg; g += gridDim.x) {
        const int4 d = seg_desc[g];
        const ViewGrid vg = grid.of_block(d.z);
        const int lb = (int)(d.z - vg.block_base);
        if (tid < kBinLocal) s_cnt[tid] = 0;
        __syncthreads();
        const int64_t i = (int64_t)d.x + tid;
        if (i < d.y) {
            const int4 l = local_rect(rect[blist[i]], lb % vg.bgw, lb / vg.bgw);
            for (int y = l.y; y < l.w; ++y)
                for (int x = l.x; x < l.z; ++x) atomicAdd(&s_cnt[y * kBinBlock + x], 1u);
        }
        __syncthreads();
        if (tid < kBinLocal) cnt[(int64_t)tid * stride + g] = s_cnt[tid];
    }
}

// Pairs per tile (from the column prefix sums at the block's first and one-past-last segment);
// 0 for the tiles of a launch sized for a capacity past the views' tiles.
template <class Grid>
__global__ __launch_bounds__(kThreads) void tile_total_kernel(const uint32_t* __restrict__ cpre,
                                                              const int2* __restrict__ branges,
                                                              const uint32_t* __restrict__ seg_first,
                                                              Grid grid, int64_t G, int64_t stride,
                                                              uint32_t* __restrict__ tcount) {
    const int64_t tile = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (tile >= G) return;
    if (tile >= grid.n_tiles()) {
        tcount[tile] = 0;
        return;
    }
    const ViewGrid vg = grid.of_tile(tile);
    const int64_t lt = tile - vg.tile_base;
    const int tx = (int)(lt % vg.gw), ty = (int)(lt / vg.gw);
    const int64_t b = vg.block_base + (ty / kBinBlock) * vg.bgw + tx / kBinBlock;
    const int t = (ty % kBinBlock) * kBinBlock + tx % kBinBlock;
    const int2 r = branges[b];
    const uint32_t f = seg_first[b], e = f + (uint32_t)((r.y - r.x + kSegEntries - 1) / kSegEntries);
    tcount[tile] = cpre[t * stride + e] - cpre[t * stride + f];
}

// ranges[t] = [tstart, tstart + tcount) for tiles with pairs, (0, 0) otherwise.
__global__ __launch_bounds__(kThreads) void tile_ranges_kernel(const uint32_t* __restrict__ tcount,
                                                               const uint32_t* __restrict__ tstart, int64_t G,
                                                               int32_t* __restrict__ ranges) {
    const int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (t >= G) return;
    const uint32_t c = tcount[t], st = tstart[t];
    ranges[2 * t] = c ? (int32_t)st : 0;
    ranges[2 * t + 1] = c ? (int32_t)(st + c) : 0;
}

// Inclusive sum over the wave of packed counters (DPP row shifts, then row broadcasts).
template <int CTRL, int ROWMASK>
__device__ inline uint32_t dpp_u(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROWMASK, 0xF, false);
}
__device__ inline uint32_t wave_scan_u32(uint32_t x) {
    x += dpp_u<0x111, 0xF>(x);   // row_shr:1
    x += dpp_u<0x112, 0xF>(x);   // row_shr:2
    x += dpp_u<0x114, 0xF>(x);   // row_shr:4
    x += dpp_u<0x118, 0xF>(x);   // row_shr:8
    x += dpp_u<0x142, 0xA>(x);   // row_bcast:15 into rows 1 and 3
    x += dpp_u<0x143, 0xC>(x);   // row_bcast:31 into rows 2 and 3
    return x;
}

// Bits 0..3 of b spread to bytes 0..3.
__device__ inline uint32_t spread4(uint32_t b) {
    return (b & 1u) | ((b & 2u) << 7) | ((b & 4u) << 14) | ((b & 8u) << 21);
}

// Writes each tet of segment g into the lists of the block's tiles it covers. A tile's first free
// position: its start, plus the pairs of the block's earlier segments (column prefix sums), plus
// the earlier entries of this segment that cover it (packed wave scans).
template <class Grid>
__global__ __launch_bounds__(kThreads) void seg_emit_kernel(const uint32_t* __restrict__ blist,
                                                            const int4* __restrict__ seg_desc,
                                                            const int64_t* __restrict__ nseg_dev,
                                                            const int4* __restrict__ rect, Grid grid, int64_t stride,
                                                            const uint32_t* __restrict__ cpre,
                                                            const uint32_t* __restrict__ tstart, int64_t n_pairs,
                                                            int32_t* __restrict__ out) {
    __shared__ uint32_t s_run[kBinLocal];                 // first free position of each tile's list
    __shared__ uint32_t s_wtot[kThreads / 64][kBinRegs];  // per wave: the segment's packed counts
    __shared__ uint32_t s_excl[kThreads][kBinRegs + 1];   // per lane: earlier entries covering each tile
    const int64_t nseg = *nseg_dev;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    for (int64_t g = blockIdx.x; g < nseg; g += gridDim.x) {
        const int4 d = seg_desc[g];
        const ViewGrid vg = grid.of_block(d.z);
        const int lb = (int)(d.z - vg.block_base);
        const int bx = lb % vg.bgw, by = lb / vg.bgw;
        const int64_t i = (int64_t)d.x + tid;
        int4 l = make_int4(0, 0, 0, 0);
        uint32_t tet = 0;
        if (i < d.y) {
            tet = blist[i];
            l = local_rect(rect[tet], bx, by);
        }
        if (tid < kBinLocal) {
            const int tx = bx * kBinBlock + (tid % kBinBlock), ty = by * kBinBlock + (tid / kBinBlock);
            uint32_t st = 0;
            if (tx < vg.gw && ty < vg.gh)
                st = tstart[vg.tile_base + (int64_t)ty * vg.gw + tx] + cpre[tid * stride + g] - cpre[tid * stride + d.w];
            s_run[tid] = st;
        }
        const uint32_t xb = ((1u << l.z) - 1u) & ~((1u << l.x) - 1u);   // covered columns
        const uint32_t lo = spread4(xb & 15u), hi = spread4(xb >> 4);
        uint32_t m[kBinRegs], sc[kBinRegs];
#pragma unroll
        for (int k = 0; k < kBinRegs; ++k) {
            const int y = k >> 1;
            m[k] = (y >= l.y && y < l.w) ? ((k & 1) ? hi : lo) : 0u;
            sc[k] = wave_scan_u32(m[k]);
        }
        if (lane == 63) {
#pragma unroll
            for (int k = 0; k < kBinRegs; ++k) s_wtot[w][k] = sc[k];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kBinRegs; ++k) {
            uint32_t pre = 0;
            for (int q = 0; q < w; ++q) pre += s_wtot[q][k];   // at most 192 per field
            s_excl[tid][k] = sc[k] - m[k] + pre;                 // at most 255 per field
        }
        if (i < d.y) {
            for (int y = l.y; y < l.w; ++y)
                for (int x = l.x; x < l.z; ++x) {
                    const int t = y * kBinBlock + x;
                    const uint32_t rank = (s_excl[tid][t >> 2] >> (8 * (t & 3))) & 255u;
                    const int64_t pos = (int64_t)s_run[t] + rank;
                    if (pos < n_pairs) out[pos] = (int32_t)tet;   // always, when the counts add up to n_pairs
                }
        }
        __syncthreads();   // s_run, s_wtot and s_excl are rewritten for the next segment
    }
}

// ---- tile order: tiles with longer lists first ----------------------------------------------
//
// The compositing kernels take their tiles in this order, so that the longest walks start first
// and the short ones fill in behind them (a tile's walk runs on one wave; one late-starting
// tile with a long list would otherwise keep the chip partly idle at the end). Keys 0xFFFF minus
// the tile's count (clamped to 0xFFFF) in 16 bits, values the tile ids, sorted stably: tiles of
// equal count stay in tile order.
template <typename Count>
__global__ __launch_bounds__(kThreads) void tile_order_keys_kernel(Count count_of, int64_t G, uint32_t* __restrict__ key,
                                                                   uint32_t* __restrict__ val) {
    const int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (t >= G) return;
    const uint32_t c = count_of(t);
    key[t] = 0xFFFFu - (c < 0xFFFFu ? c : 0xFFFFu);
    val[t] = (uint32_t)t;
}
struct CountOfRange {   // list length of tile t
    const int2* ranges;
    __device__ uint32_t operator()(int64_t t) const {
        const int2 r = ranges[t];
        return r.y > r.x ? (uint32_t)(r.y - r.x) : 0u;
    }
};
struct CountOfArray {
    const uint32_t* c;
    __device__ uint32_t operator()(int64_t t) const { return c[t]; }
};

struct OrderLayout {
    size_t key, val, key2, key3, val3, counts, partial, total;
};
inline OrderLayout order_layout(int64_t G) {
    OrderLayout L;
    size_t o = 0;
    const int64_t g = G > 0 ? G : 1;
    L.key = o; o += align256(4 * (size_t)g);
    L.val = o; o += align256(4 * (size_t)g);
    L.key2 = o; o += align256(4 * (size_t)g);
    L.key3 = o; o += align256(4 * (size_t)g);
    L.val3 = o; o += align256(4 * (size_t)g);
    L.counts = o; o += rs_counts_bytes(g);
    L.partial = o; o += scan_ws_bytes(256 * rs_blocks(g) > g ? 256 * rs_blocks(g) : g);
    L.total = o;
    return L;
}

// order[0..G) = the tiles by decreasing count of count_of, over a workspace of order_layout(G).
template <typename Count>
hipError_t tile_order(Count count_of, int64_t G, int32_t* order, char* w, hipStream_t st) {
    if (G <= 0) return hipSuccess;
    const OrderLayout L = order_layout(G);
    auto W = [&](size_t off) { return reinterpret_cast<uint32_t*>(w + off); };
    hipLaunchKernelGGL(tile_order_keys_kernel<Count>, dim3((unsigned)ceil_div(G, kThreads)), dim3(kThreads), 0, st,
                       count_of, G, W(L.key), W(L.val));
    return radix_sort<uint32_t>(W(L.key), W(L.val), W(L.key2), reinterpret_cast<uint32_t*>(order), W(L.key3),
                                W(L.val3), G, 0, 16, W(L.counts), W(L.partial), st);
}

// Fused-path workspace layout (offsets in bytes). The count phase only uses the first region.
struct BinLayout {
    size_t blk_vis, blk_pairs, partial, counts2, vkey, vidx, vkey2, vidx2, vkey3, vidx3, poff, ptile, ptet,
        ptile2, ptile3, ptet3, rcounts, ptet2, branges, tcount, tstart, seg_first, seg_block, seg_cnt, seg_pre,
        bcnt, total;
};

// Segments of the block lists: at most ceil(P / kSegEntries) plus one partial segment per block
// (at most one block per tile); each of the 64 per-tile columns has one extra entry past that.
inline int64_t seg_capacity(int64_t P, int64_t G) { return ceil_div(P, kSegEntries) + G; }

// G (tiles) sizes the block binning's per-tile and per-block buffers; 0 leaves them out.
inline BinLayout bin_layout(int64_t T, int64_t P, int64_t G = 0) {
    BinLayout L;
    size_t o = 0;
    int64_t nbT = scan_blocks(T > 0 ? T : 1);
    int64_t big = T > P ? T : P;
    int64_t nb_rs = rs_blocks(big > 0 ? big : 1);
    int64_t scan_n = (256 * nb_rs) > (T + 1) ? 256 * nb_rs : T + 1;
    scan_n = scan_n > G + 1 ? scan_n : G + 1;
    const int64_t seg_stride = G > 0 ? seg_capacity(P, G) + 1 : 0;
    scan_n = scan_n > kBinLocal * seg_stride ? scan_n : kBinLocal * seg_stride;
    L.blk_vis = o; o += align256(4 * (size_t)nbT);
    L.blk_pairs = o; o += align256(4 * (size_t)nbT);
    L.partial = o; o += scan_ws_bytes(scan_n);
    L.counts2 = o; o += align256(16);
    L.vkey = o; o += align256(4 * (size_t)T);
    L.vidx = o; o += align256(4 * (size_t)T);
    L.vkey2 = o; o += align256(4 * (size_t)T);
    L.vidx2 = o; o += align256(4 * (size_t)T);
    L.vkey3 = o; o += align256(4 * (size_t)T);
    L.vidx3 = o; o += align256(4 * (size_t)T);
    L.poff = o; o += align256(4 * (size_t)T);
    L.rcounts = o; o += rs_counts_bytes(big > 0 ? big : 1);
    L.ptile = o; o += align256(4 * (size_t)P);
    L.ptet = o; o += align256(4 * (size_t)P);
    L.ptile2 = o; o += align256(4 * (size_t)P);
    L.ptile3 = o; o += align256(4 * (size_t)P);
    L.ptet3 = o; o += align256(4 * (size_t)P);
    L.ptet2 = o; o += G > 0 ? align256(4 * (size_t)P) : 0;
    L.branges = o; o += align256(8 * (size_t)G);       // blocks: at most one per tile
    L.tcount = o; o += align256(4 * (size_t)G);
    L.tstart = o; o += align256(4 * (size_t)G);
    L.seg_first = o; o += align256(4 * (size_t)G);
    L.seg_block = o; o += G > 0 ? align256(16 * (size_t)seg_stride) : 0;   // segment descriptors (int4)
    L.bcnt = o; o += G > 0 ? align256(4 * (size_t)T) : 0;
    L.seg_cnt = o; o += align256(4 * (size_t)(kBinLocal * seg_stride));
    L.seg_pre = o; o += align256(4 * (size_t)(kBinLocal * seg_stride));
    L.total = o;
    return L;
}

inline int bit_length(int64_t v) {
    int b = 0;
    while (v > 0) { ++b; v >>= 1; }
    return b;
}

// Tile lists from the depth-ordered visible tets `order` (n_vis of them: nvis_cap, or the count at
// nvis_dev), through blocks of 8x8 tiles: (block, tet) pair offsets in depth order (their count
// stays on the device), the pairs expanded and stably sorted by block, the block lists cut into
// segments, per segment and tile the pairs, prefix sums of those per tile over the segments,
// pairs and first position of each tile, then every tet written into its tiles' lists. bmap gives
// a tet's global block id from its block coordinates, grid the view of a block or tile; pair_cap
// bounds the tile pairs (the exact count for a single view), G the tiles and NB the blocks.
template <class BlockMap, class Grid>
hipError_t block_bin_lists(const uint32_t* order, int64_t nvis_cap, const int64_t* nvis_dev, const int4* rect,
                           const BinLayout& L, char* w, int64_t pair_cap, int64_t G, int64_t NB, BlockMap bmap,
                           Grid grid, int32_t* sorted_tet, int32_t* ranges, hipStream_t st) {
    auto W = [&](size_t off) { return reinterpret_cast<uint32_t*>(w + off); };
    uint32_t* partial = W(L.partial);
    uint32_t* rcounts = W(L.rcounts);
    int64_t* nbp = reinterpret_cast<int64_t*>(w + L.counts2);   // [0] block pairs, [1] segments
    device_scan(LoadBlockCount{(const uint32_t*)W(L.bcnt), order}, nvis_cap, W(L.poff), false, partial, nbp, st,
                nvis_dev);
    hipLaunchKernelGGL((emit_pairs_kernel<BlockMap, RectBlocks>),
                       dim3((unsigned)std::min<int64_t>(ceil_div(pair_cap, kPairsPerBlock), kMaxStrideGrid)),
                       dim3(kThreads), 0, st, order, (const uint32_t*)W(L.poff), rect, nvis_cap, nvis_dev, pair_cap,
                       (const int64_t*)nbp, bmap, W(L.ptile), W(L.ptet), RectBlocks{});
    const int bbits = bit_length(NB - 1) > 0 ? bit_length(NB - 1) : 1;
    hipError_t e = radix_sort<uint32_t>(W(L.ptile), W(L.ptet), W(L.ptile2), W(L.ptet2), W(L.ptile3), W(L.ptet3),
                                        pair_cap, 0, bbits, rcounts, partial, st, nbp);
    if (e != hipSuccess) return e;
    int2* branges = reinterpret_cast<int2*>(w + L.branges);
    hipLaunchKernelGGL(zero_ranges_kernel, dim3((unsigned)ceil_div(NB, kThreads)), dim3(kThreads), 0, st, branges, NB);
    hipLaunchKernelGGL(ranges_kernel<TileOfU32>, dim3((unsigned)std::min<int64_t>(ceil_div(pair_cap, kThreads), 8192)),
                       dim3(kThreads), 0, st, TileOfU32{W(L.ptile2)}, pair_cap, (const int64_t*)nbp,
                       (int32_t*)branges);
    const uint32_t* blist = W(L.ptet2);
    uint32_t* seg_first = W(L.seg_first);
    int64_t* nsegp = nbp + 1;
    device_scan(LoadSegCount{(const int2*)branges}, NB, seg_first, false, partial, nsegp, st);
    const int64_t segcap = seg_capacity(pair_cap, G), stride = segcap + 1;
    uint32_t* cnt = W(L.seg_cnt);
    uint32_t* cpre = W(L.seg_pre);
    int4* seg_desc = reinterpret_cast<int4*>(w + L.seg_block);
    const unsigned sgrid = (unsigned)std::min(seg_grid(pair_cap, NB), segcap);
    hipLaunchKernelGGL(seg_desc_kernel, dim3((unsigned)ceil_div(NB, kThreads)), dim3(kThreads), 0, st,
                       (const int2*)branges, (const uint32_t*)seg_first, NB, seg_desc);
    hipLaunchKernelGGL((seg_count_kernel<Grid>), dim3(sgrid), dim3(kThreads), 0, st, blist, (const int4*)seg_desc,
                       (const int64_t*)nsegp, rect, grid, stride, cnt);
    device_scan(LoadSegColumn{cnt, stride, (const int64_t*)nsegp}, kBinLocal * stride, cpre, false, partial,
                (int64_t*)nullptr, st);
    uint32_t* tcount = W(L.tcount);
    uint32_t* tstart = W(L.tstart);
    hipLaunchKernelGGL((tile_total_kernel<Grid>), dim3((unsigned)ceil_div(G, kThreads)), dim3(kThreads), 0, st,
                       (const uint32_t*)cpre, (const int2*)branges, (const uint32_t*)seg_first, grid, G, stride,
                       tcount);
    device_scan(LoadU32{tcount}, G, tstart, false, partial, (int64_t*)nullptr, st);
    hipLaunchKernelGGL(tile_ranges_kernel, dim3((unsigned)ceil_div(G, kThreads)), dim3(kThreads), 0, st,
                       (const uint32_t*)tcount, (const uint32_t*)tstart, G, ranges);
    hipLaunchKernelGGL((seg_emit_kernel<Grid>), dim3(sgrid), dim3(kThreads), 0, st, blist, (const int4*)seg_desc,
                       (const int64_t*)nsegp, rect, grid, stride, (const uint32_t*)cpre, (const uint32_t*)tstart,
                       pair_cap, sorted_tet);
    return hipGetLastError();
}

}  // namespace

size_t batch_bin_workspace(const PlanHeader& h, int64_t pair_cap) {
    return align256(bin_layout(h.total_tets, pair_cap, h.total_tiles).total) + order_layout(h.total_tiles).total;
}

// Binning of a whole batch with every size read from device memory: grids are sized by the
// batch's tet, tile and block capacities and the pair capacity; counts_dev receives {n_visible,
// n_pairs} (n_pairs may exceed pair_cap: the lists are then incomplete, ranges past the capacity
// are dropped by the compositing, and the caller re-renders with a larger capacity). The views'
// tiles and blocks are numbered one view after the other, so the lists equal those of a stable
// sort of the (global tile, tet) pairs by tile.
hipError_t batch_bin(const void* plan_dev, const PlanHeader& h, int64_t pair_cap, const int32_t* tiles,
                     const int4* rect, const float* depth, const uint2* tet_blocks, int64_t* counts_dev,
                     int32_t* sorted_tet, int32_t* ranges, int32_t* tile_order_out, void* ws, hipStream_t st) {
    const int64_t T = h.total_tets, G = h.total_tiles, NB = h.total_bin_blocks;
    BinLayout L = bin_layout(T, pair_cap, G);
    char* w = static_cast<char*>(ws);
    auto W = [&](size_t off) { return reinterpret_cast<uint32_t*>(w + off); };
    const int64_t nb = scan_blocks(T);
    hipLaunchKernelGGL(bin_count_kernel, dim3((unsigned)nb), dim3(kThreads), 0, st, tiles, T, W(L.blk_vis),
                       W(L.blk_pairs));
    hipLaunchKernelGGL(scan_partials_kernel, dim3(1), dim3(kThreads), 0, st, W(L.blk_vis), nb, counts_dev);
    hipLaunchKernelGGL(scan_partials_kernel, dim3(1), dim3(kThreads), 0, st, W(L.blk_pairs), nb, counts_dev + 1);
    hipLaunchKernelGGL(compact_kernel, dim3((unsigned)nb), dim3(kThreads), 0, st, tiles, depth, 1, T,
                       (const uint32_t*)W(L.blk_vis), W(L.vkey), W(L.vidx), rect, W(L.bcnt));
    hipError_t e = radix_sort<uint32_t>(W(L.vkey), W(L.vidx), W(L.vkey2), W(L.vidx2), W(L.vkey3), W(L.vidx3), T, 0,
                                        32, W(L.rcounts), W(L.partial), st, counts_dev);
    if (e != hipSuccess) return e;
    e = block_bin_lists(W(L.vidx2), T, counts_dev, rect, L, w, pair_cap, G, NB, BlockMapBatch{tet_blocks},
                        GridBatch{plan_dev}, sorted_tet, ranges, st);
    if (e != hipSuccess || !tile_order_out) return e;
    // the global tiles by decreasing pair count, for the compositing (its capacity tiles count 0)
    return tile_order(CountOfArray{W(L.tcount)}, G, tile_order_out, w + align256(L.total), st);
}

}  // namespace rmesh

using namespace rmesh;

#define WS(T_, off) reinterpret_cast<T_*>(static_cast<char*>(workspace_dev) + (off))

extern "C" size_t rm_scan_workspace_size(int64_t n) { return scan_ws_bytes(n > 0 ? n : 1); }

extern "C" int rm_inclusive_scan_i32(const int32_t* in_dev, int32_t* out_dev, int64_t n, void* workspace_dev,
                                     size_t workspace_bytes, rm_stream_t stream) {
    RM_CHECK_ARG(n >= 0 && n < (1ll << 31), "rm_inclusive_scan_i32: n out of range");
    if (n == 0) return RM_OK;
    RM_CHECK_ARG(in_dev && out_dev && workspace_dev, "rm_inclusive_scan_i32: null buffer");
    if (workspace_bytes < rm_scan_workspace_size(n)) {
        set_error("rm_inclusive_scan_i32: workspace too small");
        return RM_ERR_WORKSPACE_TOO_SMALL;
    }
    device_scan(LoadI32{in_dev}, n, out_dev, true, (uint32_t*)workspace_dev, (int64_t*)nullptr,
                as_stream(stream));
    RM_CHECK_LAUNCH("rm_inclusive_scan_i32");
    return RM_OK;
}

extern "C" int rm_generate_keys(const float* vs_dev, const int32_t* rect_dev, const int32_t* offsets_dev,
                                int64_t n_tets, int32_t grid_height, int32_t grid_width, int64_t* keys_dev,
                                int32_t* values_dev, rm_stream_t stream) {
    RM_CHECK_ARG(n_tets >= 0 && grid_height > 0 && grid_width > 0, "rm_generate_keys: bad size");
    if (n_tets == 0) return RM_OK;
    RM_CHECK_ARG(vs_dev && rect_dev && offsets_dev && keys_dev && values_dev, "rm_generate_keys: null buffer");
    RM_CHECK_ARG(((uintptr_t)rect_dev & 15) == 0, "rm_generate_keys: rect must be 16-byte aligned");
    hipLaunchKernelGGL(generate_keys_kernel, dim3((unsigned)ceil_div(n_tets, kThreads)), dim3(kThreads), 0,
                       as_stream(stream), vs_dev, (const int4*)rect_dev, offsets_dev, n_tets, grid_width,
                       keys_dev, values_dev);
    RM_CHECK_LAUNCH("generate_keys_kernel");
    return RM_OK;
}

extern "C" size_t rm_sort_pairs_workspace_size(int64_t n) {
    if (n < 1) n = 1;
    return align256(8 * (size_t)n) + align256(4 * (size_t)n) + rs_counts_bytes(n) +
           scan_ws_bytes(256 * rs_blocks(n));
}

extern "C" int rm_sort_pairs(const int64_t* keys_in_dev, const int32_t* values_in_dev, int64_t* keys_out_dev,
                             int32_t* values_out_dev, int64_t n, int32_t end_bit, void* workspace_dev,
                             size_t workspace_bytes, rm_stream_t stream) {
    RM_CHECK_ARG(n >= 0 && n < (1ll << 31), "rm_sort_pairs: n out of range");
    RM_CHECK_ARG(end_bit >= 0 && end_bit <= 64, "rm_sort_pairs: end_bit must be in [0, 64]");
    if (n == 0) return RM_OK;
    RM_CHECK_ARG(keys_in_dev && values_in_dev && keys_out_dev && values_out_dev && workspace_dev,
                 "rm_sort_pairs: null buffer");
    if (workspace_bytes < rm_sort_pairs_workspace_size(n)) {
        set_error("rm_sort_pairs: workspace too small");
        return RM_ERR_WORKSPACE_TOO_SMALL;
    }
    size_t o = 0;
    uint64_t* ktmp = WS(uint64_t, o); o += align256(8 * (size_t)n);
    uint32_t* vtmp = WS(uint32_t, o); o += align256(4 * (size_t)n);
    uint32_t* counts = WS(uint32_t, o); o += rs_counts_bytes(n);
    uint32_t* partial = WS(uint32_t, o);
    radix_sort<uint64_t>((const uint64_t*)keys_in_dev, (const uint32_t*)values_in_dev, (uint64_t*)keys_out_dev,
                         (uint32_t*)values_out_dev, ktmp, vtmp, n, 0, end_bit, counts, partial,
                         as_stream(stream));
    RM_CHECK_LAUNCH("rm_sort_pairs");
    return RM_OK;
}

extern "C" int rm_tile_ranges(const int64_t* sorted_keys_dev, int64_t n, int32_t* ranges_dev, int64_t n_tiles,
                              rm_stream_t stream) {
    RM_CHECK_ARG(n >= 0 && n_tiles > 0, "rm_tile_ranges: bad size");
    RM_CHECK_ARG(ranges_dev != nullptr, "rm_tile_ranges: null ranges");
    hipStream_t st = as_stream(stream);
    hipLaunchKernelGGL(zero_ranges_kernel, dim3((unsigned)ceil_div(n_tiles, kThreads)), dim3(kThreads), 0, st,
                       (int2*)ranges_dev, n_tiles);
    if (n > 0) {
        RM_CHECK_ARG(sorted_keys_dev != nullptr, "rm_tile_ranges: null keys");
        hipLaunchKernelGGL(ranges_kernel<TileOfKey>, dim3((unsigned)ceil_div(n, kThreads)), dim3(kThreads), 0,
                           st, TileOfKey{sorted_keys_dev}, n, (const int64_t*)nullptr, ranges_dev);
    }
    RM_CHECK_LAUNCH("rm_tile_ranges");
    return RM_OK;
}

extern "C" size_t rm_bin_workspace_size(int64_t n_tets, int64_t n_pairs, int64_t n_tiles) {
    return bin_layout(n_tets, n_pairs, n_tiles > 0 ? n_tiles : 1).total;
}

extern "C" int rm_bin_count(const int32_t* tiles_touched_dev, int64_t n_tets, int64_t* counts_dev,
                            void* workspace_dev, size_t workspace_bytes, rm_stream_t stream) {
    RM_CHECK_ARG(n_tets >= 0 && n_tets < (1ll << 31), "rm_bin_count: n_tets out of range");
    RM_CHECK_ARG(counts_dev && workspace_dev, "rm_bin_count: null buffer");
    BinLayout L = bin_layout(n_tets, 0);
    if (workspace_bytes < L.total) {
        set_error("rm_bin_count: workspace too small");
        return RM_ERR_WORKSPACE_TOO_SMALL;
    }
    hipStream_t st = as_stream(stream);
    if (n_tets == 0) {
        hipLaunchKernelGGL(zero_i64_kernel, dim3(1), dim3(64), 0, st, counts_dev, 2);
        RM_CHECK_LAUNCH("rm_bin_count");
        return RM_OK;
    }
    RM_CHECK_ARG(tiles_touched_dev != nullptr, "rm_bin_count: null tiles");
    int64_t nb = scan_blocks(n_tets);
    hipLaunchKernelGGL(bin_count_kernel, dim3((unsigned)nb), dim3(kThreads), 0, st, tiles_touched_dev, n_tets,
                       WS(uint32_t, L.blk_vis), WS(uint32_t, L.blk_pairs));
    hipLaunchKernelGGL(scan_partials_kernel, dim3(1), dim3(kThreads), 0, st, WS(uint32_t, L.blk_vis), nb,
                       counts_dev);
    hipLaunchKernelGGL(scan_partials_kernel, dim3(1), dim3(kThreads), 0, st, WS(uint32_t, L.blk_pairs), nb,
                       counts_dev + 1);
    RM_CHECK_LAUNCH("rm_bin_count");
    return RM_OK;
}

extern "C" int rm_bin_sort(const int32_t* tiles_touched_dev, const int32_t* rect_dev, const float* vs_dev,
                           int64_t n_tets, int64_t n_visible, int64_t n_pairs, int32_t grid_height,
                           int32_t grid_width, int32_t* sorted_tet_dev, int32_t* ranges_dev, void* workspace_dev,
                           size_t workspace_bytes, rm_stream_t stream) {
    RM_CHECK_ARG(n_tets >= 0 && n_visible >= 0 && n_visible <= n_tets, "rm_bin_sort: bad n_visible");
    RM_CHECK_ARG(n_pairs >= 0 && n_pairs < (1ll << 31), "rm_bin_sort: n_pairs out of range");
    RM_CHECK_ARG(grid_height > 0 && grid_width > 0, "rm_bin_sort: bad grid");
    RM_CHECK_ARG(ranges_dev && workspace_dev, "rm_bin_sort: null buffer");
    int64_t G = (int64_t)grid_height * grid_width;
    BinLayout L = bin_layout(n_tets, n_pairs, G);
    if (workspace_bytes < L.total) {
        set_error("rm_bin_sort: workspace too small");
        return RM_ERR_WORKSPACE_TOO_SMALL;
    }
    hipStream_t st = as_stream(stream);
    if (n_pairs == 0 || n_visible == 0) {
        hipLaunchKernelGGL(zero_ranges_kernel, dim3((unsigned)ceil_div(G, kThreads)), dim3(kThreads), 0, st,
                           (int2*)ranges_dev, G);
        RM_CHECK_LAUNCH("rm_bin_sort");
        return RM_OK;
    }
    RM_CHECK_ARG(tiles_touched_dev && rect_dev && vs_dev && sorted_tet_dev, "rm_bin_sort: null buffer");
    RM_CHECK_ARG(((uintptr_t)rect_dev & 15) == 0, "rm_bin_sort: rect must be 16-byte aligned");
    int64_t nb = scan_blocks(n_tets);
    uint32_t* partial = WS(uint32_t, L.partial);
    uint32_t* rcounts = WS(uint32_t, L.rcounts);
    // 1. compact visible tets (blk_vis holds exclusive offsets after rm_bin_count)
    hipLaunchKernelGGL(compact_kernel, dim3((unsigned)nb), dim3(kThreads), 0, st, tiles_touched_dev, vs_dev, 3,
                       n_tets, (const uint32_t*)WS(uint32_t, L.blk_vis), WS(uint32_t, L.vkey),
                       WS(uint32_t, L.vidx), (const int4*)rect_dev, WS(uint32_t, L.bcnt));
    // 2. stable sort of the visible tets by depth key (32 bits)
    radix_sort<uint32_t>(WS(uint32_t, L.vkey), WS(uint32_t, L.vidx), WS(uint32_t, L.vkey2), WS(uint32_t, L.vidx2),
                         WS(uint32_t, L.vkey3), WS(uint32_t, L.vidx3), n_visible, 0, 32, rcounts, partial, st);
    // 3.-5. (block, tet) pairs in depth order, sorted by block; tile lists through block segments
    const int32_t bgw = bin_blocks_per_row(grid_width), bgh = bin_blocks_per_row(grid_height);
    block_bin_lists(WS(uint32_t, L.vidx2), n_visible, (const int64_t*)nullptr, (const int4*)rect_dev, L,
                    static_cast<char*>(workspace_dev), n_pairs, G, (int64_t)bgw * bgh, TileMapSingle{bgw},
                    GridSingle{grid_width, grid_height, bgw}, sorted_tet_dev, ranges_dev, st);
    RM_CHECK_LAUNCH("rm_bin_sort");
    return RM_OK;
}

extern "C" size_t rm_tile_order_workspace_size(int64_t n_tiles) { return order_layout(n_tiles).total; }

extern "C" int rm_tile_order(const int32_t* ranges_dev, int64_t n_tiles, int32_t* order_dev, void* workspace_dev,
                             size_t workspace_bytes, rm_stream_t stream) {
    RM_CHECK_ARG(n_tiles >= 0 && n_tiles < (1ll << 31), "rm_tile_order: n_tiles out of range");
    if (n_tiles == 0) return RM_OK;
    RM_CHECK_ARG(ranges_dev && order_dev && workspace_dev, "rm_tile_order: null buffer");
    RM_CHECK_ARG(((uintptr_t)ranges_dev & 7) == 0 && ((uintptr_t)workspace_dev & 255) == 0,
                 "rm_tile_order: misaligned ranges or workspace");
    if (workspace_bytes < order_layout(n_tiles).total) {
        set_error("rm_tile_order: workspace too small");
        return RM_ERR_WORKSPACE_TOO_SMALL;
    }
    hipError_t e = tile_order(CountOfRange{reinterpret_cast<const int2*>(ranges_dev)}, n_tiles, order_dev,
                              static_cast<char*>(workspace_dev), as_stream(stream));
    if (e != hipSuccess) {
        set_error(std::string("rm_tile_order: ") + hipGetErrorString(e));
        return RM_ERR_HIP;
    }
    RM_CHECK_LAUNCH("rm_tile_order");
    return RM_OK;
}
