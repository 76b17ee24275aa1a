// Delaunay tetrahedralisation of a point set on the GPU, by Voronoi cell clipping (gfx950, wave64).
//
// Replaces the triangulation step of models/ingp_color.py:523-544 (gdel3d `Del(V).compute` of the
// vertices, or scipy.spatial.Delaunay on the high-precision path) together with its
// post-processing: tets with the point at infinity are not produced, and every tet is positively
// oriented, det(v1-v0, v2-v0, v3-v0) > 0 as utils/topo_utils.py:9-24 measures it.
//
// Method. The Delaunay tets are the vertices of the Voronoi diagram. One wave computes the Voronoi
// cell of one point p: it starts from p's cell among four far points (the corners of a tetrahedron
// 1e8 times the size of the point set, indices n..n+3, treated as ordinary points) and clips it
// with the bisector plane of each neighbour q, visiting the cells of a uniform grid in shells of growing Chebyshev radius until no
// unvisited point can reach the cell (q cuts a cell vertex v only if |q-p| < 2|v-p|). The cell is
// held in dual form in LDS: a table of planes and a list of triangles, one per cell vertex, naming
// the three planes that meet there, ordered counter-clockwise seen from outside. A vertex whose
// three planes are bisectors with points a, b, c is the circumcentre of the Delaunay tet
// (p, a, b, c); p writes that tet when it is the smallest of the four indices and none is a far
// point (tets with a far point are the reference's tets at infinity, which it drops).
//
// Consistency. Whether q cuts vertex (p, a, b, c) is the insphere predicate of five points. Where
// the cached vertex position cannot decide it safely, it is evaluated the same way in every cell
// that asks: points sorted by index and translated to the first, FP64 with an error bound, else
// exactly in 512-bit integers, and exact ties are broken by symbolic perturbation of the lifted
// coordinate (lift_k -= eps^k for point index k). That is a real (infinitesimally weighted)
// configuration, so co-spherical sets (lattices, cube corners) and duplicate points (the lowest
// index keeps its cell) come out as one valid triangulation, the same seen from every cell.
#include "common.h"

#include <float.h>

namespace rmesh {
namespace {

constexpr int kMaxT = 256;                  // cell vertices (dual triangles) per cell
constexpr int kMaxP = 128;                  // planes per cell, the 4 far points' included
constexpr int kTPerLane = kMaxT / kWave;    // 4
constexpr double kQuickMargin = 1e-7;       // relative margin of the geometric cut test
constexpr double kCullMargin = 1e-10;       // relative margin of the conservative "cannot cut" tests
constexpr double kLost = 1e-3;              // vertex error estimate beyond which its position is not used
constexpr double kReach = 1.0 + 1e-6;       // safety factor on the 2|v-p| reach bound

// info_dev slots (int64)
enum { kInfoTets = 0, kInfoRealVerts, kInfoOverflow, kInfoDup, kInfoExact, kInfoRange, kInfoRetried, kInfoMaxT,
       kInfoCount = 8 };

struct GridParams {
    double lo[3], h[3], inv_h[3];
    float far[4][3];  // the four far points
    int g;            // cells per axis
    int pad;
};

// ------------------------------------------------------------------------------------------------
// Exact arithmetic: 512-bit two's complement integers (rare path: loops are kept rolled).
constexpr int kLimbs = 8;
struct Big {
    uint64_t w[kLimbs];
};

__host__ __device__ inline Big big_from(int64_t v) {
    Big b;
    const uint64_t s = v < 0 ? ~0ull : 0ull;
    b.w[0] = (uint64_t)v;
#pragma unroll 1
    for (int i = 1; i < kLimbs; ++i) b.w[i] = s;
    return b;
}
__host__ __device__ inline bool big_isneg(const Big& a) { return (int64_t)a.w[kLimbs - 1] < 0; }
__host__ __device__ inline Big big_add(const Big& a, const Big& b) {
    Big r;
    uint64_t c = 0;
#pragma unroll 1
    for (int i = 0; i < kLimbs; ++i) {
        const uint64_t t = a.w[i] + b.w[i];
        const uint64_t c1 = t < a.w[i];
        r.w[i] = t + c;
        c = c1 + (r.w[i] < t);
    }
    return r;
}
__host__ __device__ inline Big big_neg(const Big& a) {
    Big r;
    uint64_t c = 1;
#pragma unroll 1
    for (int i = 0; i < kLimbs; ++i) {
        const uint64_t t = ~a.w[i];
        r.w[i] = t + c;
        c = r.w[i] < t;
    }
    return r;
}
__host__ __device__ inline Big big_sub(const Big& a, const Big& b) { return big_add(a, big_neg(b)); }
// Non-negative a shifted left by sh < 64 * kLimbs bits.
__host__ __device__ inline Big big_shl(const Big& a, int sh) {
    Big r;
    const int q = sh >> 6, m = sh & 63;
#pragma unroll 1
    for (int i = kLimbs - 1; i >= 0; --i) {
        const int j = i - q;
        const uint64_t hi = j >= 0 ? a.w[j] : 0ull;
        const uint64_t lo = (j - 1 >= 0 && m) ? a.w[j - 1] >> (64 - m) : 0ull;
        r.w[i] = (m ? hi << m : hi) | lo;
    }
    return r;
}
// Product truncated to 512 bits: exact whenever the true product fits, as every product here does.
__host__ __device__ inline Big big_mul(const Big& a, const Big& b) {
    const bool na = big_isneg(a), nb = big_isneg(b);
    const Big x = na ? big_neg(a) : a;
    const Big y = nb ? big_neg(b) : b;
    Big r;
#pragma unroll 1
    for (int i = 0; i < kLimbs; ++i) r.w[i] = 0;
#pragma unroll 1
    for (int i = 0; i < kLimbs; ++i) {
        if (x.w[i] == 0) continue;
        uint64_t carry = 0;
#pragma unroll 1
        for (int j = 0; i + j < kLimbs; ++j) {
            const unsigned __int128 pr = (unsigned __int128)x.w[i] * y.w[j];
            const uint64_t lo = (uint64_t)pr, hi = (uint64_t)(pr >> 64);
            const uint64_t t = r.w[i + j] + lo;
            uint64_t c1 = t < lo;
            const uint64_t t2 = t + carry;
            c1 += t2 < t;
            r.w[i + j] = t2;
            carry = hi + c1;
        }
    }
    return na != nb ? big_neg(r) : r;
}
__host__ __device__ inline int big_sign(const Big& a) {
    if (big_isneg(a)) return -1;
    uint64_t o = 0;
#pragma unroll 1
    for (int i = 0; i < kLimbs; ++i) o |= a.w[i];
    return o ? 1 : 0;
}
__host__ __device__ inline Big big_det3(const Big* a, const Big* b, const Big* c) {
    const Big m0 = big_sub(big_mul(b[1], c[2]), big_mul(b[2], c[1]));
    const Big m1 = big_sub(big_mul(b[0], c[2]), big_mul(b[2], c[0]));
    const Big m2 = big_sub(big_mul(b[0], c[1]), big_mul(b[1], c[0]));
    return big_add(big_sub(big_mul(a[0], m0), big_mul(a[1], m1)), big_mul(a[2], m2));
}

struct Five {
    float x[5][3];   // points sorted by index
};

// Exact sign of the lifted determinant of five points (rows sorted by point index), with ties
// broken by lift_k -= eps^k. The coordinates are scaled to integers (every float is an integer
// multiple of the smallest ulp among them); magnitudes below 2^W with W <= 99 keep every
// intermediate (at most 5 W + 12 bits) inside 512 bits. Returns 2 beyond that range.
__host__ __device__ __attribute__((noinline)) int insphere_exact(const Five f) {
    int emin = 1 << 20, emax = -(1 << 20);
    for (int i = 0; i < 5; ++i)
        for (int c = 0; c < 3; ++c) {
            const uint32_t bits = __builtin_bit_cast(uint32_t, f.x[i][c]);
            if ((bits & 0x7fffffffu) == 0) continue;
            const int e = (bits >> 23) & 255;
            const int lsb = e == 0 ? -149 : e - 150;
            const int top = e == 0 ? -126 : e - 126;
            emin = lsb < emin ? lsb : emin;
            emax = top > emax ? top : emax;
        }
    if (emax < emin) return 0;   // all five points at the origin
    if (emax - emin > 99) return 2;
    Big X[5][3];
#pragma unroll 1
    for (int i = 0; i < 5; ++i)
#pragma unroll 1
        for (int c = 0; c < 3; ++c) {
            const uint32_t bits = __builtin_bit_cast(uint32_t, f.x[i][c]);
            Big v = big_from(0);
            if (bits & 0x7fffffffu) {
                const int e = (bits >> 23) & 255;
                const int64_t m = e == 0 ? (int64_t)(bits & 0x7fffffu) : (int64_t)((bits & 0x7fffffu) | 0x800000u);
                const int lsb = e == 0 ? -149 : e - 150;
                v = big_shl(big_from(m), lsb - emin);
                if (bits >> 31) v = big_neg(v);
            }
            X[i][c] = v;
        }
    Big A[4][3];
#pragma unroll 1
    for (int r = 0; r < 4; ++r)
#pragma unroll 1
        for (int c = 0; c < 3; ++c) A[r][c] = big_sub(X[r + 1][c], X[0][c]);
    // det[a_r, |a_r|^2] expanded along the lift column (signs -, +, -, +), one row at a time
    Big D = big_from(0);
    int sM[4];
#pragma unroll 1
    for (int r = 0; r < 4; ++r) {
        int rows[3], k = 0;
        for (int j = 0; j < 4; ++j)
            if (j != r) rows[k++] = j;
        const Big M = big_det3(A[rows[0]], A[rows[1]], A[rows[2]]);
        sM[r] = big_sign(M);
        const Big L = big_add(big_add(big_mul(A[r][0], A[r][0]), big_mul(A[r][1], A[r][1])), big_mul(A[r][2], A[r][2]));
        const Big t = big_mul(L, M);
        D = (r & 1) ? big_add(D, t) : big_sub(D, t);
    }
    const int s = big_sign(D);
    if (s != 0) return s;
    // D_eps = D - sum_k eps^k C_k: the first non-zero cofactor, lowest point index first, decides.
    // C_0 = det3(x2-x1, x3-x1, x4-x1); C_k = (-1)^k M_{k-1} for k >= 1.
    Big B[3][3];
#pragma unroll 1
    for (int c = 0; c < 3; ++c) {
        B[0][c] = big_sub(A[1][c], A[0][c]);
        B[1][c] = big_sub(A[2][c], A[0][c]);
        B[2][c] = big_sub(A[3][c], A[0][c]);
    }
    const int c0 = big_sign(big_det3(B[0], B[1], B[2]));
    if (c0 != 0) return -c0;
    for (int k = 1; k < 5; ++k) {
        const int ck = sM[k - 1];
        if (ck != 0) return (k & 1) ? ck : -ck;
    }
    return 0;   // all five coplanar
}

__host__ __device__ inline double det3d(const double* a, const double* b, const double* c) {
    return a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) +
           a[2] * (b[0] * c[1] - b[1] * c[0]);
}
__host__ __device__ inline double perm3d(const double* a, const double* b, const double* c) {
    return fabs(a[0]) * (fabs(b[1] * c[2]) + fabs(b[2] * c[1])) + fabs(a[1]) * (fabs(b[0] * c[2]) + fabs(b[2] * c[0])) +
           fabs(a[2]) * (fabs(b[0] * c[1]) + fabs(b[1] * c[0]));
}

// Sign of the perturbed lifted determinant of five points given in increasing index order.
// *exact counts evaluations that needed the integer path; *range those outside its range.
__host__ __device__ inline int insphere_sorted(const Five& f, int* exact, int* range) {
    double a[4][3], L[4];
    for (int r = 0; r < 4; ++r) {
        for (int c = 0; c < 3; ++c) a[r][c] = (double)f.x[r + 1][c] - (double)f.x[0][c];
        L[r] = a[r][0] * a[r][0] + a[r][1] * a[r][1] + a[r][2] * a[r][2];
    }
    const double M0 = det3d(a[1], a[2], a[3]), M1 = det3d(a[0], a[2], a[3]);
    const double M2 = det3d(a[0], a[1], a[3]), M3 = det3d(a[0], a[1], a[2]);
    const double D = (L[1] * M1 - L[0] * M0) + (L[3] * M3 - L[2] * M2);
    const double P = L[0] * perm3d(a[1], a[2], a[3]) + L[1] * perm3d(a[0], a[2], a[3]) +
                     L[2] * perm3d(a[0], a[1], a[3]) + L[3] * perm3d(a[0], a[1], a[2]);
    if (fabs(D) > 1e-13 * P) return D > 0 ? 1 : -1;
    ++*exact;
    const int s = insphere_exact(f);
    if (s == 2) {
        ++*range;
        return D > 0 ? 1 : (D < 0 ? -1 : 0);
    }
    return s;
}

// True when q lies strictly inside the circumsphere of the positively oriented tet (p, a, b, c),
// under the perturbation: the lifted determinant D(p, a, b, c, q) is negative.
__host__ __device__ __attribute__((noinline)) bool insphere_inside(const int* id_in, const float (*x_in)[3], int* exact, int* range) {
    int id[5];
    int o[5] = {0, 1, 2, 3, 4};
    for (int i = 0; i < 5; ++i) id[i] = id_in[i];
    bool odd = false;
    // insertion sort of five indices (tracking the permutation parity)
    for (int i = 1; i < 5; ++i)
        for (int j = i; j > 0 && id[j - 1] > id[j]; --j) {
            const int t = id[j]; id[j] = id[j - 1]; id[j - 1] = t;
            const int u = o[j]; o[j] = o[j - 1]; o[j - 1] = u;
            odd = !odd;
        }
    // a point that is one of the tet's own vertices is never strictly inside its sphere
    for (int i = 1; i < 5; ++i)
        if (id[i] == id[i - 1]) return false;
    Five f;
    for (int i = 0; i < 5; ++i)
        for (int c = 0; c < 3; ++c) f.x[i][c] = x_in[o[i]][c];
    const int s = insphere_sorted(f, exact, range);
    return (odd ? -s : s) < 0;
}

// ------------------------------------------------------------------------------------------------
// Grid construction.
__global__ __launch_bounds__(256) void dt_stats_kernel(const float* __restrict__ pts, int64_t n, double* partial) {
    double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX}, s[3] = {0, 0, 0},
           s2[3] = {0, 0, 0};
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        for (int c = 0; c < 3; ++c) {
            const double v = pts[3 * i + c];
            mn[c] = fmin(mn[c], v);
            mx[c] = fmax(mx[c], v);
            s[c] += v;
            s2[c] += v * v;
        }
    }
    __shared__ double red[12][256];
    for (int c = 0; c < 3; ++c) {
        red[c][threadIdx.x] = mn[c];
        red[3 + c][threadIdx.x] = mx[c];
        red[6 + c][threadIdx.x] = s[c];
        red[9 + c][threadIdx.x] = s2[c];
    }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 12; ++k) {
                const double u = red[k][threadIdx.x + w], &v = red[k][threadIdx.x];
                red[k][threadIdx.x] = k < 3 ? fmin(v, u) : (k < 6 ? fmax(v, u) : v + u);
            }
        __syncthreads();
    }
    if (threadIdx.x < 12) partial[12 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void dt_grid_kernel(const double* partial, int nb, int64_t n, int g, GridParams* gp,
                               int64_t* info) {
    if (threadIdx.x != 0) return;
    double mn[3], mx[3], s[3], s2[3];
    for (int c = 0; c < 3; ++c) {
        mn[c] = DBL_MAX; mx[c] = -DBL_MAX; s[c] = 0; s2[c] = 0;
    }
    for (int b = 0; b < nb; ++b)
        for (int c = 0; c < 3; ++c) {
            mn[c] = fmin(mn[c], partial[12 * b + c]);
            mx[c] = fmax(mx[c], partial[12 * b + 3 + c]);
            s[c] += partial[12 * b + 6 + c];
            s2[c] += partial[12 * b + 9 + c];
        }
    GridParams p;
    double ext = 0, mag = 0;
    for (int c = 0; c < 3; ++c) {
        ext = fmax(ext, mx[c] - mn[c]);
        mag = fmax(mag, fmax(fabs(mn[c]), fabs(mx[c])));
    }
    for (int c = 0; c < 3; ++c) {
        // robust extent: outliers beyond mean +- 3 sigma are clamped into the boundary cells
        const double mean = s[c] / (double)n;
        const double sd = sqrt(fmax(s2[c] / (double)n - mean * mean, 0.0));
        double lo = fmax(mn[c], mean - 3 * sd), hi = fmin(mx[c], mean + 3 * sd);
        if (!(hi - lo > 1e-30 + 1e-12 * mag)) {
            lo = mn[c];
            hi = lo + fmax(ext, fmax(1e-30, 1e-12 * mag));
        }
        p.lo[c] = lo;
        p.h[c] = (hi - lo) / g;
        p.inv_h[c] = 1.0 / p.h[c];
    }
    // far points: a regular tetrahedron around the bounding box, 1e8 times its size
    const double R = 1e8 * fmax(ext, fmax(1e-30, 1e-12 * mag));
    const double u[4][3] = {{1, 1, 1}, {1, -1, -1}, {-1, 1, -1}, {-1, -1, 1}};
    for (int k = 0; k < 4; ++k)
        for (int c = 0; c < 3; ++c) p.far[k][c] = (float)(0.5 * (mn[c] + mx[c]) + R * u[k][c]);
    p.g = g;
    p.pad = 0;
    *gp = p;
    for (int k = 0; k < kInfoCount; ++k) info[k] = 0;
}

// Extreme points of the set in 26 directions (the 3x3x3 stencil without its centre): seeds for
// cells of points outside the grid's range, whose neighbours are the set's outer points.
constexpr int kExt = 26;

__device__ inline void ext_dir(int k, double u[3]) {
    const int m = k < 13 ? k : k + 1;   // skip the centre of the stencil
    u[0] = m % 3 - 1;
    u[1] = (m / 3) % 3 - 1;
    u[2] = m / 9 - 1;
}

__global__ __launch_bounds__(256) void dt_extreme_kernel(const float* __restrict__ pts, int64_t n,
                                                         unsigned long long* keys) {
    const int64_t i = blockIdx.x * 256ll + threadIdx.x;
    const bool ok = i < n;
    double x[3] = {0, 0, 0};
    if (ok)
        for (int c = 0; c < 3; ++c) x[c] = pts[3 * i + c];
    for (int k = 0; k < kExt; ++k) {
        double u[3];
        ext_dir(k, u);
        unsigned long long key = 0ull;
        if (ok) {
            // largest projection first, lowest index among equal (float-rounded) projections
            const uint32_t b = __float_as_uint((float)(x[0] * u[0] + x[1] * u[1] + x[2] * u[2]));
            const uint32_t ord = (b >> 31) ? ~b : (b | 0x80000000u);
            key = ((unsigned long long)ord << 32) | (0xffffffffull - (unsigned long long)i);
        }
        for (int o = 32; o > 0; o >>= 1) {
            const unsigned long long y = __shfl_xor(key, o, 64);
            key = y > key ? y : key;
        }
        if ((threadIdx.x & 63) == 0 && key) atomicMax(&keys[k], key);
    }
}

__global__ void dt_extreme_ids_kernel(const unsigned long long* keys, int64_t n, int32_t* ids) {
    const int k = threadIdx.x;
    if (k < kExt) {
        const int64_t i = 0xffffffffll - (int64_t)(keys[k] & 0xffffffffull);
        ids[k] = (i >= 0 && i < n) ? (int32_t)i : 0;
    }
}

__device__ inline int cell_coord(double v, const GridParams& g, int c) {
    const double t = floor((v - g.lo[c]) * g.inv_h[c]);
    return t < 0 ? 0 : (t > g.g - 1 ? g.g - 1 : (int)t);
}

__global__ __launch_bounds__(256) void dt_keys_kernel(const float* __restrict__ pts, int64_t n,
                                                      const GridParams* __restrict__ gp, int64_t* keys,
                                                      int32_t* vals) {
    const int64_t i = blockIdx.x * 256ll + threadIdx.x;
    if (i >= n) return;
    const GridParams g = *gp;
    const int cx = cell_coord(pts[3 * i], g, 0), cy = cell_coord(pts[3 * i + 1], g, 1),
              cz = cell_coord(pts[3 * i + 2], g, 2);
    // cell id above 20 bits of an index hash: within a cell the points come in a scrambled order.
    // A cell's points are clipped 64 at a time; in index order an input sorted along a space-
    // filling curve would hand over 64 points of one corner of a dense cell at once, each cutting
    // a sliver off a still-large cell, and the cell would outgrow its capacity.
    const uint64_t h = ((uint32_t)i * 2654435761u) >> 12;   // 20 bits
    keys[i] = (int64_t)(((uint64_t)(((int64_t)cz * g.g + cy) * g.g + cx) << 20) | h);
    vals[i] = (int32_t)i;
}

__global__ __launch_bounds__(256) void dt_zero_ranges_kernel(int2* r, int64_t m) {
    const int64_t i = blockIdx.x * 256ll + threadIdx.x;
    if (i < m) r[i] = make_int2(0, 0);
}

__global__ __launch_bounds__(256) void dt_ranges_kernel(const int64_t* __restrict__ keys, int64_t n, int2* r) {
    const int64_t i = blockIdx.x * 256ll + threadIdx.x;
    if (i >= n) return;
    const int64_t k = keys[i] >> 20;   // the cell id (the low bits order points within a cell)
    if (i == 0 || (keys[i - 1] >> 20) != k) r[k].x = (int)i;
    if (i == n - 1 || (keys[i + 1] >> 20) != k) r[k].y = (int)(i + 1);
}

// ------------------------------------------------------------------------------------------------
// Cell kernels: one wave (one workgroup) per point.
//
// The cell lives in LDS: plane table pl (point position and index), triangles tri (three 8-bit
// plane slots and a flag for ill-conditioned vertices) with their vertex positions relative to p.
// The main pass holds up to 128 vertices and 128 planes (9.9 KB of LDS: 3 waves per SIMD; a
// typical cell has 26 vertices, 1 in 1000 more than 60); cells that outgrow it are queued and
// recomputed by a retry pass holding up to 960 vertices and 1024 planes (a workgroup's LDS stays
// within 64 KB).
template <int MAXT, int MAXP>
struct __align__(16) CellSmem {
    static_assert(MAXP <= 1024, "plane slots are at most 10-bit");
    float4 pl[MAXP];           // xyz: neighbour position; w: its index (int bits; far points n..n+3)
    uint32_t tri[MAXT];        // plane slots a | b << SB | c << 2 SB (SB = 8, or 10 beyond 256 planes)
    double4 vt[MAXT];          // vertex position relative to p (xyz), relative error bound of it (w)
    uint32_t newtri[MAXT];     // triangles created by one clip
    uint32_t ctri[MAXT];       // triangles removed by one clip
    uint8_t clip[MAXT];
    double cd2[kWave];         // candidates of one round, unsorted
    int cid[kWave];
    double sd2[kWave];         // the same, by increasing (distance, index)
    int sid[kWave];
    float4 spt[kWave];
    int bpre[kWave];           // a batch of grid cells: first flat index of each lane's cell
    int bst[kWave];            // and its first position in the sorted point list
};

// Whole-wave reductions: DPP within each 16-lane row (quad_perm [1,0,3,2], [2,3,0,1], row_ror 4,
// row_ror 8), then the four rows' results by readlane. Every lane must be active.
template <int CTRL>
__device__ inline int dpp_i32(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xf, 0xf, false);
}
template <int CTRL>
__device__ inline double dpp_f64(double v) {
    return __hiloint2double(dpp_i32<CTRL>(__double2hiint(v)), dpp_i32<CTRL>(__double2loint(v)));
}
__device__ inline double lane_f64(double v, int l) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                            __builtin_amdgcn_readlane(__double2loint(v), l));
}
__device__ inline double wave_max(double v) {   // of non-negative values
    v = fmax(v, dpp_f64<0xB1>(v));
    v = fmax(v, dpp_f64<0x4E>(v));
    v = fmax(v, dpp_f64<0x124>(v));
    v = fmax(v, dpp_f64<0x128>(v));
    return fmax(fmax(lane_f64(v, 0), lane_f64(v, 16)), fmax(lane_f64(v, 32), lane_f64(v, 48)));
}
__device__ inline uint32_t wave_or(uint32_t v) {
    int x = (int)v;
    x |= dpp_i32<0xB1>(x);
    x |= dpp_i32<0x4E>(x);
    x |= dpp_i32<0x124>(x);
    x |= dpp_i32<0x128>(x);
    return (uint32_t)(__builtin_amdgcn_readlane(x, 0) | __builtin_amdgcn_readlane(x, 16) |
                      __builtin_amdgcn_readlane(x, 32) | __builtin_amdgcn_readlane(x, 48));
}
__device__ inline int lanes_below(uint64_t m, int lane) {
    return __popcll(m & ((1ull << lane) - 1ull));
}

// Bisector of p and the point in plane slot `slot`, as n . x <= r with x relative to p.
__device__ inline void plane_of(const float4* pl, int slot, double px, double py, double pz, double n[3],
                                double& r) {
    const float4 P = pl[slot];
    n[0] = (double)P.x - px;
    n[1] = (double)P.y - py;
    n[2] = (double)P.z - pz;
    r = 0.5 * (n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
}

// Vertex where three planes meet (Cramer's rule). err: bound on its relative position error,
// 1e-15 / kappa with kappa = |det(n1,n2,n3)| / (|n1||n2||n3|) (the conditioning of the system
// with unit normals, ten times the FP64 rounding estimate); 1 when the planes are degenerate.
__device__ inline void vertex_of(const double a[3], double ra, const double b[3], double rb, const double c[3],
                                 double rc, double& x, double& y, double& z, float& err) {
    const double bc[3] = {b[1] * c[2] - b[2] * c[1], b[2] * c[0] - b[0] * c[2], b[0] * c[1] - b[1] * c[0]};
    const double ca[3] = {c[1] * a[2] - c[2] * a[1], c[2] * a[0] - c[0] * a[2], c[0] * a[1] - c[1] * a[0]};
    const double ab[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    const double det = a[0] * bc[0] + a[1] * bc[1] + a[2] * bc[2];
    const double k2 = det * det / ((a[0] * a[0] + a[1] * a[1] + a[2] * a[2]) * (b[0] * b[0] + b[1] * b[1] + b[2] * b[2]) *
                                   (c[0] * c[0] + c[1] * c[1] + c[2] * c[2]));   // kappa^2
    err = k2 > 1e-30 ? fminf(1e-15f / sqrtf((float)k2), 1.0f) : 1.0f;
    if (!(fabs(det) > 0.0)) {
        x = y = 0.0;
        z = 1e150;
        err = 1.0f;
        return;
    }
    const double inv = 1.0 / det;
    x = (ra * bc[0] + rb * ca[0] + rc * ab[0]) * inv;
    y = (ra * bc[1] + rb * ca[1] + rc * ab[1]) * inv;
    z = (ra * bc[2] + rb * ca[2] + rc * ab[2]) * inv;
    if (!(fabs(x) < 1e150 && fabs(y) < 1e150 && fabs(z) < 1e150)) {
        x = y = 0.0;
        z = 1e150;
        err = 1.0f;
    }
}

// Shell cell idx (0 <= idx < 24 r^2 + 2) at Chebyshev radius r >= 1.
__device__ inline void shell_cell(int r, int idx, int& dx, int& dy, int& dz) {
    const int a = 2 * r + 1, b = 2 * r - 1;
    if (idx < 2 * a * a) {
        dz = idx < a * a ? -r : r;
        const int k = idx % (a * a);
        dx = k % a - r;
        dy = k / a - r;
        return;
    }
    idx -= 2 * a * a;
    if (idx < 2 * a * b) {
        dy = idx < a * b ? -r : r;
        const int k = idx % (a * b);
        dx = k % a - r;
        dz = k / a - (r - 1);
        return;
    }
    idx -= 2 * a * b;
    dx = idx < b * b ? -r : r;
    const int k = idx % (b * b);
    dy = k % b - (r - 1);
    dz = k / b - (r - 1);
}

struct CellArgs {
    const float* pts;
    int64_t n;
    const int32_t* cell_pts;
    const int2* cell_rng;
    const GridParams* gp;
    int4* app;
    int64_t app_cap;
    unsigned long long* app_count;
    int32_t* pt_count;
    int64_t* pt_off;
    int64_t* info;
    const int32_t* ext;                // kExt extreme points
    const int32_t* order;              // main pass: block b computes point order[b]
    int32_t* retry;                    // cells that outgrew the main pass (null in the retry pass)
    unsigned long long* retry_count;
};

template <int MAXT, int MAXP>
__device__ void cell_run(CellSmem<MAXT, MAXP>& S, const int p, const CellArgs& A) {
    constexpr int TPL = MAXT / kWave;
    constexpr int UW = MAXP / 32;
    constexpr int SB = MAXP > 256 ? 10 : 8;   // bits per plane slot in a triangle word
    constexpr uint32_t SM = (1u << SB) - 1u;
    const int lane = threadIdx.x;
    const GridParams& gpr = *A.gp;
    const int gg = gpr.g;
    const int64_t n = A.n;
    const float* pts = A.pts;
    const float pfx = pts[3 * (int64_t)p], pfy = pts[3 * (int64_t)p + 1], pfz = pts[3 * (int64_t)p + 2];
    const double px = pfx, py = pfy, pz = pfz;
    const double pc[3] = {px, py, pz};
    const int cx = cell_coord(px, gpr, 0), cy = cell_coord(py, gpr, 1), cz = cell_coord(pz, gpr, 2);
    const int cc[3] = {cx, cy, cz};

    // starting cell: p's Voronoi cell among the four far points (plane slots 0..3), a tetrahedron
    // with one vertex per triple of far points, each triple ordered counter-clockwise seen from
    // outside (det of the three plane normals > 0)
    if (lane < 4)
        S.pl[lane] = make_float4(gpr.far[lane][0], gpr.far[lane][1], gpr.far[lane][2], __int_as_float((int)n + lane));
    __syncthreads();
    if (lane < 4) {
        uint32_t s3[3];
        int k = 0;
        for (int j = 0; j < 4; ++j)
            if (j != lane) s3[k++] = (uint32_t)j;
        double na[3], nb[3], nc[3], ra, rb, rc;
        plane_of(S.pl, s3[0], px, py, pz, na, ra);
        plane_of(S.pl, s3[1], px, py, pz, nb, rb);
        plane_of(S.pl, s3[2], px, py, pz, nc, rc);
        if (det3d(na, nb, nc) < 0) {
            const uint32_t t = s3[0]; s3[0] = s3[1]; s3[1] = t;
            for (int c = 0; c < 3; ++c) { const double u = na[c]; na[c] = nb[c]; nb[c] = u; }
            const double u = ra; ra = rb; rb = u;
        }
        double x, y, z;
        float er;
        vertex_of(na, ra, nb, rb, nc, rc, x, y, z, er);
        S.tri[lane] = s3[0] | s3[1] << SB | s3[2] << (2 * SB);
        S.vt[lane] = make_double4(x, y, z, er);
    }
    int nt = 4;
    uint32_t used[UW];
#pragma unroll
    for (int w = 0; w < UW; ++w) used[w] = w == 0 ? 0xfu : 0u;
    bool dead = false, fail = false;
    int exact = 0, range = 0;
    __syncthreads();
    // squared reach of the cell: the largest |v - p|^2 over its vertices (widened by their error
    // bounds); a point farther than twice the reach cannot cut it. Recomputed once per round of
    // candidates, so between two recomputations it only overestimates (the cell only shrinks).
    auto cell_reach2 = [&]() -> double {
        double r2 = 0.0;
#pragma unroll
        for (int j = 0; j < TPL; ++j) {
            const int t = lane + kWave * j;
            if (t < nt) {
                const double4 V = S.vt[t];
                const double rr = V.x * V.x + V.y * V.y + V.z * V.z;
                const double er = V.w;
                r2 = fmax(r2, er < kLost ? rr * (1.0 + 16.0 * er) : 4.0 * rr);
            }
        }
        return wave_max(r2);
    };
    double maxr2 = cell_reach2();

    // a box (relative to p) can hold a point that cuts the cell only if it meets some ball
    // B(v, |v - p|) over the cell's vertices v
    auto box_reach = [&](const double* bl, const double* bh) -> bool {
        bool reach = false;
        for (int t = 0; t < nt && !reach; ++t) {
            const double4 V = S.vt[t];
            double vx = V.x, vy = V.y, vz = V.z;
            double rr = vx * vx + vy * vy + vz * vz;
            const double er = V.w;
            if (er < kLost) {
                rr *= 1.0 + 16.0 * er;
            } else {   // position unreliable: the ball B(p, 2|v - p|)
                vx = vy = vz = 0.0;
                rr *= 4.0;
            }
            const double ex = fmax(fmax(bl[0] - vx, vx - bh[0]), 0.0);
            const double ey = fmax(fmax(bl[1] - vy, vy - bh[1]), 0.0);
            const double ez = fmax(fmax(bl[2] - vz, vz - bh[2]), 0.0);
            reach = ex * ex + ey * ey + ez * ez <= rr * (1.0 + 1e-9);
        }
        return reach;
    };
    // box of grid cells [i0, i1] per axis, relative to p; boundary cells are open on their outer
    // side (they hold the clamped outliers)
    auto cell_box = [&](const int* i0, const int* i1, double* bl, double* bh) {
        for (int c = 0; c < 3; ++c) {
            bl[c] = i0[c] == 0 ? -DBL_MAX : gpr.lo[c] + i0[c] * gpr.h[c] - pc[c];
            bh[c] = i1[c] == gg - 1 ? DBL_MAX : gpr.lo[c] + (i1[c] + 1) * gpr.h[c] - pc[c];
        }
    };
    // clip with every point of the lanes' point ranges [s, e) (one grid cell per lane)
    auto process_batch = [&](const int32_t* list, int s, int e) {
        // the batch's points as one flat list spread over all lanes
        const int cntl = e - s;
        int incl = cntl;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            const int y = __shfl_up(incl, o, 64);
            if (lane >= o) incl += y;
        }
        const int total = __shfl(incl, kWave - 1, 64);
        if (total == 0) return;
        __syncthreads();   // the previous batch's readers are done
        S.bpre[lane] = incl - cntl;
        S.bst[lane] = s;
        __syncthreads();
        for (int c0 = 0; c0 < total; c0 += kWave) {
            maxr2 = cell_reach2();
            const int kf = c0 + lane;
            const bool have = kf < total;
            int q = -1;
            float qx = 0.f, qy = 0.f, qz = 0.f;
            double d2 = 0.0;
            bool valid = false, dup = false;
            if (have) {
                // the last lane whose cell starts at or before kf holds it
                int lo = 0, hi = kWave - 1;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (S.bpre[mid] <= kf) lo = mid;
                    else hi = mid - 1;
                }
                q = list[S.bst[lo] + (kf - S.bpre[lo])];
                qx = pts[3 * (int64_t)q];
                qy = pts[3 * (int64_t)q + 1];
                qz = pts[3 * (int64_t)q + 2];
                const double dx = (double)qx - px, dy = (double)qy - py, dz = (double)qz - pz;
                d2 = dx * dx + dy * dy + dz * dz;
                if (q != p) {
                    if (d2 == 0.0) {
                        dup = q < p;   // a lower-index copy of p owns this position
                    } else if (d2 < 4.0 * maxr2 * kReach) {
                        // q can cut the cell only if some vertex is closer to q than to p
                        for (int t = 0; t < nt && !valid; ++t) {
                            const double4 V = S.vt[t];
                            const double vx = V.x, vy = V.y, vz = V.z;
                            const double sdist = vx * dx + vy * dy + vz * dz - 0.5 * d2;
                            const double mag = sqrt((vx * vx + vy * vy + vz * vz) * d2) + d2;
                            const double er = V.w;
                            if (er < kLost) valid = sdist > -(kCullMargin + 8.0 * er) * mag;
                            else valid = d2 < 16.0 * (vx * vx + vy * vy + vz * vz) * kReach;
                        }
                    }
                }
            }
            if (__ballot(dup)) {
                dead = true;
                break;
            }
            const int nv = __popcll(__ballot(valid));
            if (nv == 0) continue;
            S.cd2[lane] = valid ? d2 : DBL_MAX;
            S.cid[lane] = valid ? q : 0x7fffffff;
            __syncthreads();
            if (valid) {
                int rank = 0;
                for (int j = 0; j < kWave; ++j) {
                    const double dj = S.cd2[j];
                    const int ij = S.cid[j];
                    // ties (the same point offered by two lanes, as the extreme-point seeds can)
                    // broken by lane, so the ranks are a permutation and every slot gets written
                    rank += (dj < d2) || (dj == d2 && (ij < q || (ij == q && j < lane)));
                }
                S.sd2[rank] = d2;
                S.sid[rank] = q;
                S.spt[rank] = make_float4(qx, qy, qz, 0.f);
            }
            __syncthreads();
            for (int k = 0; k < nv && !fail; ++k) {
                // ---- clip the cell with the bisector of p and candidate k
                const int cq = S.sid[k];
                const float4 Q = S.spt[k];
                const double ddx = (double)Q.x - px, ddy = (double)Q.y - py, ddz = (double)Q.z - pz;
                const double cd2 = ddx * ddx + ddy * ddy + ddz * ddz;
                if (!(cd2 < 4.0 * maxr2 * kReach)) continue;
                // 1. which vertices does the plane cut (q strictly inside their circumsphere)
                int mine = 0;
#pragma unroll
                for (int j = 0; j < TPL; ++j) {
                    const int t = lane + kWave * j;
                    if (t < nt) {
                        const uint32_t tw = S.tri[t];
                        const double4 V = S.vt[t];
                        const double vx = V.x, vy = V.y, vz = V.z;
                        const double sdist = vx * ddx + vy * ddy + vz * ddz - 0.5 * cd2;
                        const double mag = sqrt((vx * vx + vy * vy + vz * vz) * cd2) + cd2;
                        const double er = V.w;
                        bool c;
                        if (er < kLost && fabs(sdist) > (kQuickMargin + 8.0 * er) * mag) {
                            c = sdist > 0.0;
                        } else {
                            const float4 Pa = S.pl[tw & SM], Pb = S.pl[(tw >> SB) & SM], Pc = S.pl[(tw >> (2 * SB)) & SM];
                            const int ids[5] = {p, __float_as_int(Pa.w), __float_as_int(Pb.w), __float_as_int(Pc.w), cq};
                            const float xs[5][3] = {{pfx, pfy, pfz}, {Pa.x, Pa.y, Pa.z}, {Pb.x, Pb.y, Pb.z},
                                                    {Pc.x, Pc.y, Pc.z}, {Q.x, Q.y, Q.z}};
                            c = insphere_inside(ids, xs, &exact, &range);
                        }
                        S.clip[t] = c ? 1 : 0;
                        mine += c;
                    }
                }
                if (!__ballot(mine > 0)) continue;
                // a free plane slot for q
                int slot = -1;
#pragma unroll
                for (int w = UW - 1; w >= 0; --w)
                    if (used[w] != 0xffffffffu) slot = 32 * w + __builtin_ctz(~used[w]);
                if (slot < 0) {
                    fail = true;
                    break;
                }
                // 2. list the removed triangles
                int nclip = 0;
#pragma unroll
                for (int j = 0; j < TPL; ++j) {
                    const int t = lane + kWave * j;
                    const bool c = t < nt && S.clip[t];
                    const uint64_t cm = __ballot(c);
                    if (c) S.ctri[nclip + lanes_below(cm, lane)] = S.tri[t];
                    nclip += __popcll(cm);
                }
                if (lane == 0) S.pl[slot] = make_float4(Q.x, Q.y, Q.z, __int_as_float(cq));
                __syncthreads();
                // 3. boundary of the removed region: edge (i -> j) of a removed triangle whose
                // neighbour (the triangle holding j -> i) is not removed gives the new triangle
                // (i, j, slot); the neighbour is removed iff some removed triangle holds j -> i
                int n_new = 0;
                for (int i0 = 0; i0 < nclip; i0 += kWave) {
                    const int i = i0 + lane;
                    uint32_t a = 0, b = 0, c = 0;
                    bool e0 = false, e1 = false, e2 = false;
                    if (i < nclip) {
                        const uint32_t tw = S.ctri[i];
                        a = tw & SM;
                        b = (tw >> SB) & SM;
                        c = (tw >> (2 * SB)) & SM;
                        bool f0 = false, f1 = false, f2 = false;
                        for (int u = 0; u < nclip; ++u) {
                            const uint32_t uw = S.ctri[u];
                            const uint32_t u0 = uw & SM, u1 = (uw >> SB) & SM, u2 = (uw >> (2 * SB)) & SM;
                            f0 |= (u0 == b && u1 == a) || (u1 == b && u2 == a) || (u2 == b && u0 == a);
                            f1 |= (u0 == c && u1 == b) || (u1 == c && u2 == b) || (u2 == c && u0 == b);
                            f2 |= (u0 == a && u1 == c) || (u1 == a && u2 == c) || (u2 == a && u0 == c);
                        }
                        e0 = !f0;
                        e1 = !f1;
                        e2 = !f2;
                    }
                    // positions in the new-triangle list from ballots, edge by edge
                    const uint32_t sl = (uint32_t)slot << (2 * SB);
                    const uint64_t m0 = __ballot(e0), m1 = __ballot(e1), m2 = __ballot(e2);
                    int pos = n_new + lanes_below(m0, lane);
                    if (e0 && pos < MAXT) S.newtri[pos] = a | b << SB | sl;
                    n_new += __popcll(m0);
                    pos = n_new + lanes_below(m1, lane);
                    if (e1 && pos < MAXT) S.newtri[pos] = b | c << SB | sl;
                    n_new += __popcll(m1);
                    pos = n_new + lanes_below(m2, lane);
                    if (e2 && pos < MAXT) S.newtri[pos] = c | a << SB | sl;
                    n_new += __popcll(m2);
                }
                __syncthreads();
                const int kept = nt - nclip;
                if (kept + n_new > MAXT) {
                    fail = true;
                    break;
                }
                // 4. compaction, one chunk of 64 at a time (writes of a chunk land below it)
                int base2 = 0;
                for (int c0 = 0; c0 < nt; c0 += kWave) {
                    const int t = c0 + lane;
                    const bool keep = t < nt && !S.clip[t];
                    uint32_t tw = 0;
                    double4 V = make_double4(0.0, 0.0, 0.0, 0.0);
                    if (keep) {
                        tw = S.tri[t];
                        V = S.vt[t];
                    }
                    const uint64_t km = __ballot(keep);
                    __syncthreads();
                    if (keep) {
                        const int o = base2 + lanes_below(km, lane);
                        S.tri[o] = tw;
                        S.vt[o] = V;
                    }
                    base2 += __popcll(km);
                }
                // 5. the new triangles after the survivors
                for (int t = lane; t < n_new; t += kWave) {
                    const uint32_t tw = S.newtri[t];
                    double na[3], nb[3], nc[3], ra, rb, rc;
                    plane_of(S.pl, tw & SM, px, py, pz, na, ra);
                    plane_of(S.pl, (tw >> SB) & SM, px, py, pz, nb, rb);
                    plane_of(S.pl, (tw >> (2 * SB)) & SM, px, py, pz, nc, rc);
                    double x, y, z;
                    float er;
                    vertex_of(na, ra, nb, rb, nc, rc, x, y, z, er);
                    S.tri[kept + t] = tw;
                    S.vt[kept + t] = make_double4(x, y, z, er);
                }
                nt = kept + n_new;
                __syncthreads();
                // 6. planes still in use, and the reach of the cell
                uint32_t uu[UW];
#pragma unroll
                for (int w = 0; w < UW; ++w) uu[w] = 0u;
#pragma unroll
                for (int j = 0; j < TPL; ++j) {
                    const int t = lane + kWave * j;
                    if (t < nt) {
                        const uint32_t tw = S.tri[t];
                        const uint32_t sa = tw & SM, sb = (tw >> SB) & SM, sc = (tw >> (2 * SB)) & SM;
#pragma unroll
                        for (int w = 0; w < UW; ++w) {
                            uu[w] |= (sa >> 5) == (uint32_t)w ? 1u << (sa & 31) : 0u;
                            uu[w] |= (sb >> 5) == (uint32_t)w ? 1u << (sb & 31) : 0u;
                            uu[w] |= (sc >> 5) == (uint32_t)w ? 1u << (sc & 31) : 0u;
                        }
                    }
                }
#pragma unroll
                for (int w = 0; w < UW; ++w) used[w] = wave_or(uu[w]);
            }
            __syncthreads();
            if (fail) break;
        }
    };

    // a point outside the grid's range (clamped into a boundary cell) starts from the set's
    // extreme points, which bound its cell far sooner than the shells around its clamped cell
    {
        bool outside = false;
        for (int c = 0; c < 3; ++c) outside |= pc[c] < gpr.lo[c] || pc[c] > gpr.lo[c] + gg * gpr.h[c];
        if (outside) process_batch(A.ext, lane < kExt ? lane : 0, lane < kExt ? lane + 1 : 0);
    }

    // fine shells of grid cells around p's cell, then shells of 4x4x4-cell blocks
    constexpr int kFineRings = 3;
    bool done = false;
    for (int r = 0; r <= kFineRings && !dead && !fail; ++r) {
        maxr2 = cell_reach2();
        if (r > 0) {
            // lower bound on |q - p| for every point in a cell at Chebyshev distance >= r
            double L = DBL_MAX;
            for (int c = 0; c < 3; ++c) {
                if (cc[c] + r <= gg - 1) L = fmin(L, gpr.lo[c] + (cc[c] + r) * gpr.h[c] - pc[c]);
                if (cc[c] - r >= 0) L = fmin(L, pc[c] - (gpr.lo[c] + (cc[c] - r + 1) * gpr.h[c]));
            }
            if (L == DBL_MAX) {   // every cell visited
                done = true;
                break;
            }
            L *= 1.0 - 1e-12;
            if (L > 0 && L * L > 4.0 * maxr2 * kReach) {
                done = true;
                break;
            }
        }
        const int count = r == 0 ? 1 : 24 * r * r + 2;
        bool ring_reached = false;
        for (int base = 0; base < count && !dead && !fail; base += kWave) {
            int s = 0, e = 0;
            const int idx = base + lane;
            if (idx < count) {
                int dx = 0, dy = 0, dz = 0;
                if (r > 0) shell_cell(r, idx, dx, dy, dz);
                const int ci[3] = {cx + dx, cy + dy, cz + dz};
                if (ci[0] >= 0 && ci[1] >= 0 && ci[2] >= 0 && ci[0] < gg && ci[1] < gg && ci[2] < gg) {
                    double bl[3], bh[3];
                    cell_box(ci, ci, bl, bh);
                    const bool reach = box_reach(bl, bh);
                    if (reach) {
                        const int2 rg = A.cell_rng[((int64_t)ci[2] * gg + ci[1]) * gg + ci[0]];
                        s = rg.x;
                        e = rg.y;
                    }
                    ring_reached |= reach;
                }
            }
            process_batch(A.cell_pts, s, e);
        }
        // every ball touches p, so a ball reaching ring r + 1 crosses ring r: a ring no ball
        // reaches ends the search
        if (r > 0 && !__ballot(ring_reached)) {
            done = true;
            break;
        }
    }
    if (!done && !dead && !fail) {
        const int gb = (gg + 3) >> 2;
        const int bc[3] = {cx >> 2, cy >> 2, cz >> 2};
        for (int rb = 0; !dead && !fail; ++rb) {
            maxr2 = cell_reach2();
            if (rb > 0) {
                double L = DBL_MAX;
                for (int c = 0; c < 3; ++c) {
                    const int hiCell = (bc[c] + rb) * 4, loCell = (bc[c] - rb) * 4 + 3;
                    if (bc[c] + rb <= gb - 1) L = fmin(L, gpr.lo[c] + hiCell * gpr.h[c] - pc[c]);
                    if (bc[c] - rb >= 0) L = fmin(L, pc[c] - (gpr.lo[c] + (loCell + 1) * gpr.h[c]));
                }
                if (L == DBL_MAX) break;   // every block visited
                L *= 1.0 - 1e-12;
                if (L > 0 && L * L > 4.0 * maxr2 * kReach) break;
            }
            const int count = rb == 0 ? 1 : 24 * rb * rb + 2;
            bool ring_reached = false;
            for (int base = 0; base < count && !dead && !fail; base += kWave) {
                // one block per lane: does it meet a ball?
                const int idx = base + lane;
                int b3[3] = {0, 0, 0};
                bool reach = false;
                if (idx < count) {
                    int dx = 0, dy = 0, dz = 0;
                    if (rb > 0) shell_cell(rb, idx, dx, dy, dz);
                    b3[0] = bc[0] + dx;
                    b3[1] = bc[1] + dy;
                    b3[2] = bc[2] + dz;
                    if (b3[0] >= 0 && b3[1] >= 0 && b3[2] >= 0 && b3[0] < gb && b3[1] < gb && b3[2] < gb) {
                        int i0[3], i1[3];
                        for (int c = 0; c < 3; ++c) {
                            i0[c] = b3[c] * 4;
                            i1[c] = min(b3[c] * 4 + 3, gg - 1);
                        }
                        double bl[3], bh[3];
                        cell_box(i0, i1, bl, bh);
                        reach = box_reach(bl, bh);
                    }
                }
                ring_reached |= reach;
                uint64_t bm = __ballot(reach);
                while (bm && !dead && !fail) {
                    const int j = __builtin_ctzll(bm);
                    bm &= bm - 1;
                    const int bx = __shfl(b3[0], j, 64), by = __shfl(b3[1], j, 64), bz = __shfl(b3[2], j, 64);
                    // the block's 64 cells, one per lane; cells of the fine shells are done
                    const int ci[3] = {bx * 4 + (lane & 3), by * 4 + ((lane >> 2) & 3), bz * 4 + (lane >> 4)};
                    int s = 0, e = 0;
                    if (ci[0] < gg && ci[1] < gg && ci[2] < gg &&
                        max(max(abs(ci[0] - cx), abs(ci[1] - cy)), abs(ci[2] - cz)) > kFineRings) {
                        double bl[3], bh[3];
                        cell_box(ci, ci, bl, bh);
                        if (box_reach(bl, bh)) {
                            const int2 rg = A.cell_rng[((int64_t)ci[2] * gg + ci[1]) * gg + ci[0]];
                            s = rg.x;
                            e = rg.y;
                        }
                    }
                    process_batch(A.cell_pts, s, e);
                }
            }
            if (rb > 0 && !__ballot(ring_reached)) break;
        }
    }

    // the tets this point owns: vertices of three real planes with indices all above p
    int nreal = 0, nemit = 0;
    if (!dead && !fail) {
#pragma unroll
        for (int j = 0; j < TPL; ++j) {
            const int t = lane + kWave * j;
            bool real = false, em = false;
            if (t < nt) {
                const uint32_t tw = S.tri[t];
                const int a = __float_as_int(S.pl[tw & SM].w), b = __float_as_int(S.pl[(tw >> SB) & SM].w),
                          c = __float_as_int(S.pl[(tw >> (2 * SB)) & SM].w);
                real = a < n && b < n && c < n;   // no far point
                em = real && p < a && p < b && p < c;
            }
            nreal += __popcll(__ballot(real));
            nemit += __popcll(__ballot(em));
        }
    }
    const bool requeue = fail && A.retry != nullptr;
    unsigned long long off = 0;
    if (lane == 0) {
        if (nemit > 0) off = atomicAdd(A.app_count, (unsigned long long)nemit);
        A.pt_count[p] = nemit;
        A.pt_off[p] = (int64_t)off;
        if (nreal) atomicAdd((unsigned long long*)&A.info[kInfoRealVerts], (unsigned long long)nreal);
        if (dead) atomicAdd((unsigned long long*)&A.info[kInfoDup], 1ull);
        if (requeue) {
            atomicAdd((unsigned long long*)&A.info[kInfoRetried], 1ull);
            const unsigned long long slot = atomicAdd(A.retry_count, 1ull);
            A.retry[slot] = p;   // at most one entry per point: the list holds n
        } else if (fail) {
            atomicAdd((unsigned long long*)&A.info[kInfoOverflow], 1ull);
        }
        atomicMax((unsigned long long*)&A.info[kInfoMaxT], (unsigned long long)nt);
    }
    off = __shfl(off, 0, 64);
    if (!dead && !fail && nemit > 0) {
        int before = 0;
#pragma unroll
        for (int j = 0; j < TPL; ++j) {
            const int t = lane + kWave * j;
            bool em = false;
            int4 tet = make_int4(0, 0, 0, 0);
            if (t < nt) {
                const uint32_t tw = S.tri[t];
                const int a = __float_as_int(S.pl[tw & SM].w), b = __float_as_int(S.pl[(tw >> SB) & SM].w),
                          c = __float_as_int(S.pl[(tw >> (2 * SB)) & SM].w);
                em = a < n && b < n && c < n && p < a && p < b && p < c;
                // rotate (a, b, c) cyclically to put the smallest first: orientation is kept
                if (a < b && a < c) tet = make_int4(p, a, b, c);
                else if (b < c) tet = make_int4(p, b, c, a);
                else tet = make_int4(p, c, a, b);
            }
            const uint64_t m = __ballot(em);
            if (em) {
                const unsigned long long o = off + before + lanes_below(m, lane);
                if (o < (unsigned long long)A.app_cap) A.app[o] = tet;
            }
            before += __popcll(m);
        }
    }
    if (exact) atomicAdd((unsigned long long*)&A.info[kInfoExact], (unsigned long long)exact);
    if (range) atomicAdd((unsigned long long*)&A.info[kInfoRange], (unsigned long long)range);
    __syncthreads();   // the LDS cell may be reused by the next point of this workgroup
}

__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void dt_cell_kernel(CellArgs A) {
    __shared__ CellSmem<128, 128> S;
    cell_run(S, A.order[blockIdx.x], A);
}

// Cells that outgrew the main pass, with 3.75 times the vertices and eight times the planes; a
// fixed grid walks the queue.
static_assert(sizeof(CellSmem<128, 128>) <= 65536 && sizeof(CellSmem<960, 1024>) <= 65536,
              "a cell's LDS must fit one workgroup's 64 KB");
__global__ __launch_bounds__(64) void dt_cell_retry_kernel(CellArgs A, const int32_t* __restrict__ queue,
                                                           const unsigned long long* __restrict__ queue_len) {
    __shared__ CellSmem<960, 1024> S;
    const unsigned long long len = *queue_len;
    for (unsigned long long i = blockIdx.x; i < len; i += gridDim.x) cell_run(S, queue[i], A);
}

__global__ __launch_bounds__(256) void dt_gather_kernel(const int4* __restrict__ app, int64_t app_cap,
                                                        const int32_t* __restrict__ pt_count,
                                                        const int32_t* __restrict__ pt_incl,
                                                        const int64_t* __restrict__ pt_off, int64_t n,
                                                        int4* __restrict__ out, int64_t cap, int64_t* info) {
    const int64_t p = blockIdx.x * 256ll + threadIdx.x;
    if (p >= n) return;
    const int c = pt_count[p];
    const int64_t dst = (int64_t)pt_incl[p] - c, src = pt_off[p];
    for (int j = 0; j < c; ++j)
        if (dst + j < cap && src + j < app_cap) out[dst + j] = app[src + j];
    if (p == n - 1) info[kInfoTets] = pt_incl[p];
}

inline size_t al(size_t x) { return (x + 255) & ~size_t(255); }

inline int grid_dim(int64_t n) {
    int g = (int)llround(cbrt((double)n / 2.0));
    return g < 1 ? 1 : (g > 512 ? 512 : g);
}

struct DtLayout {
    size_t keys, keys2, vals, vals2, sortws, ranges, partial, gp, app, cnt, incl, off, scanws, appcount, retry, ext, order,
        total;
    int g, nb;
};

inline DtLayout dt_layout(int64_t n, int64_t cap) {
    DtLayout L;
    L.g = grid_dim(n);
    L.nb = (int)std::min<int64_t>(1024, (n + 255) / 256 > 0 ? (n + 255) / 256 : 1);
    const int64_t cells = (int64_t)L.g * L.g * L.g;
    size_t o = 0;
    L.keys = o; o += al(8 * (size_t)n);
    L.keys2 = o; o += al(8 * (size_t)n);
    L.vals = o; o += al(4 * (size_t)n);
    L.vals2 = o; o += al(4 * (size_t)n);
    L.sortws = o; o += al(rm_sort_pairs_workspace_size(n));
    L.ranges = o; o += al(8 * (size_t)cells);
    L.partial = o; o += al(12 * 8 * (size_t)L.nb);
    L.gp = o; o += al(sizeof(GridParams));
    L.app = o; o += al(16 * (size_t)(cap > 0 ? cap : 1));
    L.cnt = o; o += al(4 * (size_t)n);
    L.incl = o; o += al(4 * (size_t)n);
    L.off = o; o += al(8 * (size_t)n);
    L.scanws = o; o += al(rm_scan_workspace_size(n));
    L.appcount = o; o += al(16);   // append counter, retry-queue length
    L.retry = o; o += al(4 * (size_t)n);
    L.ext = o; o += al(8 * kExt + 4 * kExt);   // keys, then ids
    L.order = o; o += al(4 * (size_t)n);
    L.total = o;
    return L;
}

// Work estimate per point for the dispatch order: the number of points in its 3x3x3 block of grid
// cells (sparse neighbourhoods make large, slow cells; points outside the grid's range get 0).
// Sorting by it starts the slow cells first, so no long cell runs alone at the end of the grid.
__global__ __launch_bounds__(256) void dt_weight_kernel(const float* __restrict__ pts, int64_t n,
                                                        const GridParams* __restrict__ gp,
                                                        const int2* __restrict__ cell_rng, int64_t* keys,
                                                        int32_t* vals) {
    const int64_t i = blockIdx.x * 256ll + threadIdx.x;
    if (i >= n) return;
    const GridParams g = *gp;
    const double x[3] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
    bool outside = false;
    int c[3];
    for (int k = 0; k < 3; ++k) {
        outside |= x[k] < g.lo[k] || x[k] > g.lo[k] + g.g * g.h[k];
        c[k] = cell_coord(x[k], g, k);
    }
    int64_t w = 0;
    if (!outside)
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const int X = c[0] + dx, Y = c[1] + dy, Z = c[2] + dz;
                    if (X < 0 || Y < 0 || Z < 0 || X >= g.g || Y >= g.g || Z >= g.g) continue;
                    const int2 r = cell_rng[((int64_t)Z * g.g + Y) * g.g + X];
                    w += r.y - r.x;
                }
    // equal estimates in a hashed order: inputs sorted along a space-filling curve would otherwise
    // start the cells of one region together
    const uint32_t h = ((uint32_t)i * 2654435761u) >> 20;   // 12 bits
    keys[i] = ((w < (1 << 20) - 1 ? w : (1 << 20) - 1) << 12) | h;
    vals[i] = (int32_t)i;
}

__global__ void dt_zero_u64_kernel(unsigned long long* p, unsigned long long* ext_keys) {
    p[0] = 0ull;
    p[1] = 0ull;
    for (int k = 0; k < kExt; ++k) ext_keys[k] = 0ull;
}

__global__ void dt_zero_info_kernel(int64_t* info) {
    if (threadIdx.x < kInfoCount) info[threadIdx.x] = 0;
}

}  // namespace
}  // namespace rmesh

using namespace rmesh;

extern "C" size_t rm_delaunay_workspace_size(int64_t n_points, int64_t tet_capacity) {
    if (n_points < 1) n_points = 1;
    return dt_layout(n_points, tet_capacity).total;
}

extern "C" int rm_delaunay(const float* points_dev, int64_t n_points, int32_t* tets_dev, int64_t tet_capacity,
                           int64_t* info_dev, void* workspace_dev, size_t workspace_bytes, rm_stream_t stream) {
    RM_CHECK_ARG(n_points >= 0 && n_points < (1ll << 30), "rm_delaunay: n_points out of range");
    RM_CHECK_ARG(tet_capacity >= 0 && tet_capacity < (1ll << 31), "rm_delaunay: tet_capacity out of range");
    RM_CHECK_ARG(info_dev != nullptr, "rm_delaunay: null info");
    hipStream_t st = as_stream(stream);
    const int64_t n = n_points;
    if (n == 0) {
        // nothing to triangulate: clear the counters so the caller reads zero tets
        hipLaunchKernelGGL(dt_zero_info_kernel, dim3(1), dim3(kInfoCount), 0, st, info_dev);
        RM_CHECK_LAUNCH("rm_delaunay");
        return RM_OK;
    }
    RM_CHECK_ARG(points_dev && workspace_dev, "rm_delaunay: null buffer");
    RM_CHECK_ARG(tet_capacity == 0 || tets_dev != nullptr, "rm_delaunay: null tets");
    const DtLayout L = dt_layout(n, tet_capacity);
    if (workspace_bytes < L.total) {
        set_error("rm_delaunay: workspace too small");
        return RM_ERR_WORKSPACE_TOO_SMALL;
    }
    char* ws = (char*)workspace_dev;
    int64_t* keys = (int64_t*)(ws + L.keys);
    int64_t* keys2 = (int64_t*)(ws + L.keys2);
    int32_t* vals = (int32_t*)(ws + L.vals);
    int32_t* vals2 = (int32_t*)(ws + L.vals2);
    int2* ranges = (int2*)(ws + L.ranges);
    double* partial = (double*)(ws + L.partial);
    GridParams* gp = (GridParams*)(ws + L.gp);
    int4* app = (int4*)(ws + L.app);
    int32_t* cnt = (int32_t*)(ws + L.cnt);
    int32_t* incl = (int32_t*)(ws + L.incl);
    int64_t* off = (int64_t*)(ws + L.off);
    unsigned long long* appcount = (unsigned long long*)(ws + L.appcount);
    unsigned long long* retrycount = appcount + 1;
    int32_t* retry = (int32_t*)(ws + L.retry);
    unsigned long long* ext_keys = (unsigned long long*)(ws + L.ext);
    int32_t* ext_ids = (int32_t*)(ws + L.ext + 8 * kExt);
    const int64_t cells = (int64_t)L.g * L.g * L.g;
    const int64_t app_cap = tet_capacity > 0 ? tet_capacity : 1;

    hipLaunchKernelGGL(dt_stats_kernel, dim3(L.nb), dim3(256), 0, st, points_dev, n, partial);
    hipLaunchKernelGGL(dt_grid_kernel, dim3(1), dim3(1), 0, st, (const double*)partial, L.nb, n, L.g, gp, info_dev);
    hipLaunchKernelGGL(dt_keys_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, points_dev, n,
                       (const GridParams*)gp, keys, vals);
    RM_CHECK_LAUNCH("rm_delaunay grid");
    int end_bit = 1;
    while ((1ll << end_bit) < cells) ++end_bit;
    end_bit += 20;
    int rc = rm_sort_pairs(keys, vals, keys2, vals2, n, end_bit, ws + L.sortws, rm_sort_pairs_workspace_size(n),
                           stream);
    if (rc != RM_OK) return rc;
    hipLaunchKernelGGL(dt_zero_ranges_kernel, dim3((unsigned)ceil_div(cells, 256)), dim3(256), 0, st, ranges, cells);
    hipLaunchKernelGGL(dt_ranges_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, (const int64_t*)keys2,
                       n, ranges);
    hipLaunchKernelGGL(dt_zero_u64_kernel, dim3(1), dim3(1), 0, st, appcount, ext_keys);
    hipLaunchKernelGGL(dt_extreme_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, points_dev, n, ext_keys);
    hipLaunchKernelGGL(dt_extreme_ids_kernel, dim3(1), dim3(64), 0, st, (const unsigned long long*)ext_keys, n,
                       ext_ids);
    CellArgs ca;
    ca.pts = points_dev;
    ca.n = n;
    ca.cell_pts = vals2;
    ca.cell_rng = ranges;
    ca.gp = gp;
    ca.app = app;
    ca.app_cap = app_cap;
    ca.app_count = appcount;
    ca.pt_count = cnt;
    ca.pt_off = off;
    ca.info = info_dev;
    ca.ext = ext_ids;
    // dispatch order: most expensive cells first (the sorted-by-cell keys and the input key/value
    // buffers are free once the cell ranges exist)
    int32_t* order = (int32_t*)(ws + L.order);
    hipLaunchKernelGGL(dt_weight_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, points_dev, n,
                       (const GridParams*)gp, (const int2*)ranges, keys, vals);
    rc = rm_sort_pairs(keys, vals, keys2, order, n, 32, ws + L.sortws, rm_sort_pairs_workspace_size(n), stream);
    if (rc != RM_OK) return rc;
    ca.order = order;
    ca.retry = retry;
    ca.retry_count = retrycount;
    hipLaunchKernelGGL(dt_cell_kernel, dim3((unsigned)n), dim3(kWave), 0, st, ca);
    CellArgs cr = ca;
    cr.retry = nullptr;
    cr.retry_count = nullptr;
    hipLaunchKernelGGL(dt_cell_retry_kernel, dim3(1024), dim3(kWave), 0, st, cr, (const int32_t*)retry,
                       (const unsigned long long*)retrycount);
    RM_CHECK_LAUNCH("rm_delaunay cells");
    rc = rm_inclusive_scan_i32(cnt, incl, n, ws + L.scanws, rm_scan_workspace_size(n), stream);
    if (rc != RM_OK) return rc;
    hipLaunchKernelGGL(dt_gather_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, (const int4*)app,
                       app_cap, (const int32_t*)cnt, (const int32_t*)incl, (const int64_t*)off, n, (int4*)tets_dev,
                       tet_capacity, info_dev);
    RM_CHECK_LAUNCH("rm_delaunay gather");
    return RM_OK;
}

// Host evaluation of the cell kernel's insphere predicate (same code), for tests without a GPU:
// 1 when x[4] lies strictly inside the circumsphere of the positively oriented tet x[0..3] under
// the index perturbation, else 0. ids[5], xyz[15].
extern "C" int rm_delaunay_insphere_host(const int32_t* ids, const float* xyz) {
    float x[5][3];
    for (int i = 0; i < 5; ++i)
        for (int c = 0; c < 3; ++c) x[i][c] = xyz[3 * i + c];
    int e = 0, r = 0;
    return insphere_inside(ids, x, &e, &r) ? 1 : 0;
}
