// Per-tile front-to-back compositing of tetrahedra and its backward.
//
// Forward semantics: splat_tiled / alpha_blend (alphablend_shader_interp.slang:194-271,371-431),
// ray_tetrahedron_intersect_fused (intersect.slang:393-624), integrate_tetra_segment
// (interp_version.slang:21-98), get_ray (camera.slang:298-346). Per pixel, the tile's tets are
// visited in sorted order; a hit with sigma*dt > 0 is blended with the exact integral of a
// linear colour under constant density, n_contrib counts every visited tet up to and including
// the one after which log-transmittance drops below -5.5.
//
// MI355X layout: at 8x8 and 16x16 tiles a workgroup of tile^2 lanes holds one tile, a lane per
// pixel. At 4x4 tiles one wave takes a tile: the forward gives each pixel a quad of lanes that
// walk four entries per step (composite_fwd_quads_kernel); the backward and the error votes put
// the lanes on list entries, a lane evaluating its tet for each of the tile's 16 pixels, with each
// pixel's walk state from prefix sums across the lanes. Each chunk of a tile's list is staged as
// per-view plane records computed once per (tile, tet): the four outward face planes n_i with
// offsets e_i = n_i . (cam - v_a), and
// the colour folded to c(s) = base + g.(cam - v0) + s * g.d, where s is the distance from the
// camera centre along the unit ray. A plane crossing is then s_i = -e_i / (n_i . d) and the
// reference's ray parameter is t = s - min_t (its origin is cam + min_t * d).
//
// Backward: the reference's structure (bwd_alpha_blend, alphablend_shader_interp.slang:273-369):
// each pixel walks its list back to front from the stored final state, undoing log-transmittance
// and the distortion moments, skipping entries at or beyond its n_contrib. Per-tet gradients
// are accumulated as adjoints of the plane records (per plane: sum of u = ds/den and of
// u (x - v0), x the crossing point), summed across the wavefront (or 16-lane tile group) with cross-lane shuffles,
// and converted once per (wave, tet) to vertex and feature gradients, which lanes of the
// wave then add with one float atomic each. At 4x4 tiles with per-tet rows the lanes hold tets
// instead of pixels (composite_bwd_tet_lanes_kernel): each lane sums its tet's adjoints over the
// tile's 16 pixels and the walk state comes from prefix sums across the lanes.
#include <mutex>
#include <string>

#include "common.h"

namespace rmesh {
namespace {

template <int TS>
struct Cfg {
    static constexpr int GROUP = TS * TS;                 // pixels per tile
    static constexpr int WG = GROUP < 64 ? 64 : GROUP;    // lanes per workgroup
    static constexpr int NG = WG / GROUP;                 // tiles per workgroup
    static constexpr int GX = NG >= 2 ? 2 : 1;            // tiles per workgroup along x
    static constexpr int GY = NG / GX;
    static constexpr int RED = GROUP < 64 ? GROUP : 64;   // lanes summed per reduction
};

template <int F>
struct NG_ {
    static constexpr int NGRD = F == 7 ? 3 : 9;           // colour-gradient floats per tet
};

template <typename T>
struct RayT {
    T dx, dy, dz;
};

// Plane records: float4 (n, e) in FP32, or double4 in FP64 for the accurate intersection path.
template <bool P64>
struct Prec;
template <>
struct Prec<false> {
    using R = float;
    using PL = float4;
};
template <>
struct Prec<true> {
    using R = double;
    using PL = double4;
};

__device__ inline float3 f3(float x, float y, float z) { return make_float3(x, y, z); }
__device__ inline float3 operator-(float3 a, float3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ inline float3 operator+(float3 a, float3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ inline float3 operator*(float s, float3 a) { return f3(s * a.x, s * a.y, s * a.z); }
__device__ inline float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline float3 cross3(float3 a, float3 b) {
    return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// camera.slang:241-296 undistort (Newton, 10 iterations) and 298-346 get_ray, evaluated in T.
template <typename T>
__device__ inline RayT<T> make_ray(const CamArgs& c, int ix, int iy, float jx, float jy) {
    // no implicit FMA contraction: every kernel that casts this pixel's ray gets the same bits
#pragma clang fp contract(off)
    const T px = (T)ix + (T)jx, py = (T)iy + (T)jy;
    T x = (px - (T)c.K[2]) / (T)c.K[0];
    T y = (py - (T)c.K[5]) / (T)c.K[4];
    if ((c.type == 0 || c.type == 1) && c.distorted) {
        const T k1 = c.dist[0], k2 = c.dist[1], p1 = c.dist[2], p2 = c.dist[3];
        const T xd = x, yd = y;
        for (int i = 0; i < 10; ++i) {
            T r2 = x * x + y * y;
            T d = (T)1 + r2 * (k1 + r2 * k2);
            T fx = d * x + (T)2 * p1 * x * y + p2 * (r2 + (T)2 * x * x) - xd;
            T fy = d * y + (T)2 * p2 * x * y + p1 * (r2 + (T)2 * y * y) - yd;
            T d_r = k1 + r2 * ((T)2 * k2);
            T d_x = (T)2 * x * d_r, d_y = (T)2 * y * d_r;
            T fx_x = d + d_x * x + (T)2 * p1 * y + (T)6 * p2 * x;
            T fx_y = d_y * x + (T)2 * p1 * x + (T)2 * p2 * y;
            T fy_x = d_x * y + (T)2 * p2 * y + (T)2 * p1 * x;
            T fy_y = d + d_y * y + (T)2 * p2 * x + (T)6 * p1 * y;
            T den = fy_x * fx_y - fx_x * fy_y;
            if (fabs(den) < (T)1e-9) break;
            x += (fx * fy_y - fy * fx_y) / den;
            y += (fy * fx_x - fx * fy_x) / den;
        }
    }
    T dcx, dcy, dcz;
    if (c.type == 1) {
        T th = sqrt(x * x + y * y);
        th = th < (T)3.14159265 ? th : (T)3.14159265;
        T s = th < (T)1e-6 ? (T)1 : sin(th) / th;
        dcx = x * s; dcy = y * s; dcz = cos(th);
    } else {
        T inv = (T)1 / sqrt(x * x + y * y + (T)1);
        dcx = x * inv; dcy = y * inv; dcz = inv;
    }
    const float* R = c.ray_rot;
    T dx = (T)R[0] * dcx + (T)R[1] * dcy + (T)R[2] * dcz;
    T dy = (T)R[3] * dcx + (T)R[4] * dcy + (T)R[5] * dcz;
    T dz = (T)R[6] * dcx + (T)R[7] * dcy + (T)R[8] * dcz;
    T inv = (T)1 / sqrt(dx * dx + dy * dy + dz * dz);
    return {dx * inv, dy * inv, dz * inv};
}

__device__ inline float3 load_v(const float* __restrict__ v, int i) {
    const float* p = v + 3 * (int64_t)i;
    return f3(p[0], p[1], p[2]);
}

template <typename T>
struct HitT {
    bool hit;
    float se, sx;      // clamped entry / exit distance from the camera along the ray
    float dt;          // sx - se, formed in the plane precision before rounding
    float se_raw;      // entering plane distance before clamping at min_t
    float den_e, den_x;
    int ie, ix;
    T se_p, sx_p;      // entering / exiting plane distances in the plane precision, unclamped
};

__device__ inline float rcp_(float x) { return __builtin_amdgcn_rcpf(x); }
// v_rcp_f64 (max relative error 2^-24.4 measured on MI355X over 1M inputs spanning 2^+-43) refined
// by one Newton step (2^-48.7): a crossing distance s = -e / den then carries a relative error far
// below the FP32 rounding of the segment end points it feeds.
__device__ inline double rcp_(double x) {
    const double r = __builtin_amdgcn_rcp(x);
    return fma(fma(-x, r, 1.0), r, r);
}

// ray_tetrahedron_intersect_fused on plane records (n_i, e_i) for a ray from cam along d.
template <typename PL, typename T>
__device__ inline HitT<T> intersect(const PL pl[4], RayT<T> r, float mtf) {
    // explicit FMAs and no implicit contraction: the forward, the backward and the error votes
    // reach bit-identical hit decisions for the same ray and plane records
#pragma clang fp contract(off)
    HitT<T> h;
    const T mt = (T)mtf;
    T se = (T)-1e30, sx = (T)1e30, de = 1, dx = 1;
    int ie = -1, ix = -1;
    bool miss = false;
    // branch-free: a plane nearly parallel to the ray (|den| < 1e-10) only decides a miss; the
    // others are entering (den < 0) or exiting candidates (s of a parallel plane is not used)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const T den = fma(pl[i].z, r.dz, fma(pl[i].y, r.dy, pl[i].x * r.dx));
        const bool par = fabs(den) < (T)1e-10;
        miss = miss || (par && fma(mt, den, pl[i].w) > (T)0);
        const T s = -pl[i].w * rcp_(den);
        const bool ent = !par && den < (T)0 && s > se;
        const bool ext = !par && !(den < (T)0) && s < sx;
        se = ent ? s : se; de = ent ? den : de; ie = ent ? i : ie;
        sx = ext ? s : sx; dx = ext ? den : dx; ix = ext ? i : ix;
    }
    h.hit = !miss && !(se > sx) && (sx > mt);
    T sec = se > mt ? se : mt;
    T sxc = sx > sec ? sx : sec;
    h.se_raw = (float)se;
    h.se = (float)sec;
    h.sx = (float)sxc;
    h.dt = (float)(sxc - sec);
    h.den_e = (float)de;
    h.den_x = (float)dx;
    h.ie = ie;
    h.ix = ix;
    h.se_p = se;
    h.sx_p = sx;
    return h;
}

// e^-x - 1 for x >= 0, to a few ulp relative: below 0.25 the Taylor series of expm1 to degree 7
// (truncation under 2e-9 relative), above it the hardware exponential minus one (the result is
// at least 0.22 in magnitude there, so the exponential's ~1e-7 error stays ~3e-7 relative). About
// half the instructions of expm1f, on the per-(pixel, tet) path of compositing.
__device__ inline float expm1_neg(float x) {
    const float u = -x;
    const float p = u * (1.0f + u * (0.5f + u * (1.0f / 6.0f + u * (1.0f / 24.0f + u * (1.0f / 120.0f +
                    u * (1.0f / 720.0f + u * (1.0f / 5040.0f)))))));
    const float e = __expf(u) - 1.0f;
    return x <= 0.25f ? p : e;
}


// Weights of the exit and entry colours in the segment integral, w0 = phi(x) - e^-x and
// w1 = 1 - phi(x) with phi(x) = (1 - e^-x) / x (interp_version.slang:21-69); their sum is
// 1 - e^-x. Both are O(x): formed as phi - alpha and 1 - phi they lose about eps / x of relative
// precision (1% at x = 1e-5, a thin sliver of a crossing), so below x = 0.05 they come from their
// own series, w1 = sum (-1)^(k+1) x^k / (k+1)! and w0 = sum (-1)^(k+1) k x^k / (k+1)!, to x^5
// (truncation under 1e-9 relative).
__device__ inline void seg_weights(float x, float em1, float alpha, float& w0, float& w1) {
    if (x < 0.05f) {
        w1 = x * (0.5f - x * (1.f / 6.f - x * (1.f / 24.f - x * (1.f / 120.f - x * (1.f / 720.f)))));
        w0 = x * (0.5f - x * (1.f / 3.f - x * (1.f / 8.f - x * (1.f / 30.f - x * (1.f / 144.f)))));
    } else {
        const float ph = -em1 * rcp_(x);
        w0 = ph - alpha;
        w1 = 1.0f - ph;
    }
}

// seg_weights and phi'(x) for the backward, on one branch: series below 0.05 (phi' truncated after
// x^4, under 1e-9 relative), closed forms above through one hardware reciprocal (1 ulp).
__device__ inline void seg_weights_dphi(float x, float em1, float alpha, float& w0, float& w1, float& dphi) {
    if (x < 0.05f) {
        w1 = x * (0.5f - x * (1.f / 6.f - x * (1.f / 24.f - x * (1.f / 120.f - x * (1.f / 720.f)))));
        w0 = x * (0.5f - x * (1.f / 3.f - x * (1.f / 8.f - x * (1.f / 30.f - x * (1.f / 144.f)))));
        dphi = -0.5f + x * (1.f / 3.f + x * (-1.f / 8.f + x * (1.f / 30.f - x * (1.f / 144.f))));
    } else {
        const float r = rcp_(x);
        const float ph = -em1 * r;
        w0 = ph - alpha;
        w1 = 1.0f - ph;
        dphi = (x * alpha + em1) * (r * r);
    }
}

// DPP lane permute of a 64-bit bit mask within a row of 16 lanes.
template <int CTRL>
__device__ inline uint64_t bits_dpp(uint64_t v) {
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)v, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(v >> 32), CTRL, 0xF, 0xF, false);
    return (uint64_t)(uint32_t)lo | ((uint64_t)(uint32_t)hi << 32);
}

// XCD-aware workgroup order: consecutive dispatch slots go round-robin over the 8 XCDs, so
// slot b is remapped to make each XCD own a contiguous band of tiles (shared tets stay in its L2).
__device__ inline int remap_block(int b, int nb) {
    int per = nb >> 3;
    if (b < (per << 3)) return (b & 7) * per + (b >> 3);
    return b;
}

// A tile's [start, end) range; ranges outside [0, n_sorted] are treated as empty.
__device__ inline int2 load_range(const int2* __restrict__ ranges, bool tile_ok, int tile, int64_t n_sorted) {
    if (!tile_ok) return make_int2(0, 0);
    int2 rg = ranges[tile];
    if (rg.x < 0 || rg.y < rg.x || (int64_t)rg.y > n_sorted) return make_int2(0, 0);
    return rg;
}

// Tile group and pixel of this lane for workgroup b of a view (b already XCD-remapped).
template <int TS>
__device__ inline void locate(const CamArgs& cam, int b, int& g, int& r, int& tx, int& ty, int& px, int& py,
                              bool& tile_ok, bool& inside) {
    using C = Cfg<TS>;
    int sgw = (cam.gw + C::GX - 1) / C::GX;
    int sx = b % sgw, sy = b / sgw;
    g = threadIdx.x / C::GROUP;
    r = threadIdx.x % C::GROUP;
    tx = sx * C::GX + (g % C::GX);
    ty = sy * C::GY + (g / C::GX);
    tile_ok = tx < cam.gw && ty < cam.gh;
    px = tx * TS + (r % TS);
    py = ty * TS + (r / TS);
    inside = tile_ok && px < cam.W && py < cam.H;
}

// LDS staging of one chunk: per slot, 4 planes, (sigma, folded base rgb), colour gradients; the
// backward adds the tet id and, when it converts gradients per slot (VTX), the vertices and the
// face sign bits.
template <int WG, int F, bool BWD, bool P64, bool VTX = BWD>
struct Stage {
    typename Prec<P64>::PL pl[4][WG];
    float4 fa[WG];                     // sigma, base' r, g, b
    float4 fg[(NG_<F>::NGRD + 3) / 4][WG];
    float4 vtx[VTX ? 3 : 1][VTX ? WG : 1];  // v0..v3 packed (backward epilogue)
    int meta[BWD ? (VTX ? 2 : 1) : 1][BWD ? WG : 1];  // tet id, face sign bits
    int kidx[BWD ? WG : 1];            // position of the staged entry in its tile's list (backward)
    typename Prec<P64>::R cv[BWD ? 3 : 1][BWD ? WG : 1];  // cam - v0 in the plane precision (backward)
};

template <typename T>
struct V3 {
    T x, y, z;
};
template <typename T>
__device__ inline V3<T> v3sub(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename T>
__device__ inline T v3dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename T>
__device__ inline V3<T> v3cross(V3<T> a, V3<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// Outward plane record of face (a, b, c) opposite o, anchored at a, evaluated in T:
// n = +-cross(b - a, c - a) flipped away from o (intersect.slang:455-479), e = n . (cam - a).
template <typename T, typename PL>
__device__ inline int plane_rec(float3 af, float3 bf, float3 cf, float3 of, float3 camf, PL& out) {
#pragma clang fp contract(off)
    V3<T> a = {(T)af.x, (T)af.y, (T)af.z}, b = {(T)bf.x, (T)bf.y, (T)bf.z};
    V3<T> c = {(T)cf.x, (T)cf.y, (T)cf.z}, o = {(T)of.x, (T)of.y, (T)of.z};
    V3<T> cam = {(T)camf.x, (T)camf.y, (T)camf.z};
    V3<T> n = v3cross(v3sub(b, a), v3sub(c, a));
    int flip = v3dot(n, v3sub(o, a)) > (T)0;
    if (flip) n = {-n.x, -n.y, -n.z};
    out.x = n.x; out.y = n.y; out.z = n.z;
    out.w = v3dot(n, v3sub(cam, a));
    return flip;
}

// The global data of one list entry: its four vertices and its feature row.
template <int F>
struct TetData {
    float3 v[4];
    float f[F];
};

template <int F>
__device__ inline void load_tet_data(TetData<F>& d, int4 q, int tet, const float* __restrict__ verts,
                                     const float* __restrict__ feats) {
    d.v[0] = load_v(verts, q.x);
    d.v[1] = load_v(verts, q.y);
    d.v[2] = load_v(verts, q.z);
    d.v[3] = load_v(verts, q.w);
    const float* f = feats + (int64_t)F * tet;
#pragma unroll
    for (int k = 0; k < F; ++k) d.f[k] = f[k];
}

// The per-view records of one tet: its four outward face planes (n_i, e_i) in the plane
// precision, (sigma, folded base rgb) and the colour gradients; returns the face sign bits.
template <int F, bool P64>
__device__ inline int tet_record(const TetData<F>& d, const CamArgs& cam, typename Prec<P64>::PL (&pl)[4],
                                 float4& fa, float (&g)[NG_<F>::NGRD]) {
    // the roundings of the base colour are fixed (no contraction chosen by the compiler), so that
    // every kernel staging a tet, whatever its surrounding code, stages the same bits
#pragma clang fp contract(off)
    using T = typename Prec<P64>::R;
    constexpr int NGRD = NG_<F>::NGRD;
    const float3 v0 = d.v[0], v1 = d.v[1], v2 = d.v[2], v3 = d.v[3];
    float3 cp = f3(cam.cam_pos[0], cam.cam_pos[1], cam.cam_pos[2]);
    int signs = 0;
    signs |= plane_rec<T>(v0, v1, v2, v3, cp, pl[0]) << 0;
    signs |= plane_rec<T>(v0, v1, v3, v2, cp, pl[1]) << 1;
    signs |= plane_rec<T>(v0, v2, v3, v1, cp, pl[2]) << 2;
    signs |= plane_rec<T>(v1, v2, v3, v0, cp, pl[3]) << 3;
    const float* f = d.f;
#pragma unroll
    for (int k = 0; k < NGRD; ++k) g[k] = f[4 + k];
    float3 cv = cp - v0;
    // g . (cam - v0) as two fused multiply-adds on the first product
    auto gdot = [&](float a, float b, float c) { return fmaf(c, cv.z, fmaf(b, cv.y, a * cv.x)); };
    float b0, b1, b2;
    if (F == 7) {
        float g0 = gdot(g[0], g[1], g[2]);
        b0 = f[1] + g0; b1 = f[2] + g0; b2 = f[3] + g0;
    } else {
        b0 = f[1] + gdot(g[0], g[1], g[2]);
        b1 = f[2] + gdot(g[3 % NGRD], g[4 % NGRD], g[5 % NGRD]);
        b2 = f[3] + gdot(g[6 % NGRD], g[7 % NGRD], g[8 % NGRD]);
    }
    fa = make_float4(fmaxf(f[0], 0.f), b0, b1, b2);
    return signs;
}

// Writes the slot's LDS records (face planes, base colour, gradients, ...) of tet from its data.
template <int WG, int F, bool BWD, bool P64, bool VTX>
__device__ inline void stage_data(Stage<WG, F, BWD, P64, VTX>& S, int slot, int tet, const CamArgs& cam,
                                  const TetData<F>& d) {
    using T = typename Prec<P64>::R;
    constexpr int NGRD = NG_<F>::NGRD;
    const float3 v0 = d.v[0], v1 = d.v[1], v2 = d.v[2], v3 = d.v[3];
    float3 cp = f3(cam.cam_pos[0], cam.cam_pos[1], cam.cam_pos[2]);
    typename Prec<P64>::PL pl[4];
    float4 fa;
    float g[NGRD];
    const int signs = tet_record<F, P64>(d, cam, pl, fa, g);
#pragma unroll
    for (int i = 0; i < 4; ++i) S.pl[i][slot] = pl[i];
    S.fa[slot] = fa;
#pragma unroll
    for (int k = 0; k < (NGRD + 3) / 4; ++k)
        S.fg[k][slot] = make_float4(g[4 * k], 4 * k + 1 < NGRD ? g[(4 * k + 1) % NGRD] : 0.f,
                                    4 * k + 2 < NGRD ? g[(4 * k + 2) % NGRD] : 0.f,
                                    4 * k + 3 < NGRD ? g[(4 * k + 3) % NGRD] : 0.f);
    if constexpr (BWD) {
        S.meta[0][slot] = tet;
        S.cv[0][slot] = (T)cp.x - (T)v0.x;
        S.cv[1][slot] = (T)cp.y - (T)v0.y;
        S.cv[2][slot] = (T)cp.z - (T)v0.z;
    }
    if constexpr (VTX) {
        S.vtx[0][slot] = make_float4(v0.x, v0.y, v0.z, v1.x);
        S.vtx[1][slot] = make_float4(v1.y, v1.z, v2.x, v2.y);
        S.vtx[2][slot] = make_float4(v2.z, v3.x, v3.y, v3.z);
        S.meta[1][slot] = signs;
    }
}

template <int WG, int F, bool BWD, bool P64, bool VTX>
__device__ inline void stage_tet(Stage<WG, F, BWD, P64, VTX>& S, int slot, int tet, const CamArgs& cam,
                                 const int4* __restrict__ indices, const float* __restrict__ verts,
                                 const float* __restrict__ feats) {
    TetData<F> d;
    load_tet_data<F>(d, indices[tet], tet, verts, feats);
    stage_data(S, slot, tet, cam, d);
}

// Load pipeline over a tile's chunks of list entries, walked front to back. Each lane owns at most
// one entry per chunk; entry(c) gives the lane's tet in chunk c, or -1 when it has none. While
// chunk c is walked, the entries of chunk c + 3, the index rows of chunk c + 2 and the vertices and
// features of chunk c + 1 are in flight, so the three dependent fetches (list entry, index row,
// vertex and feature rows) of a chunk overlap the walks of the chunks before it (18 + F registers).
// The index, vertex and feature loads are issued by every lane, a lane with no entry reading the
// rows of tet 0 (so the pipeline is only started when the workgroup has entries, hence tets): with
// loads under a divergent branch the number of loads issued between two points depends on the
// path, and the compiler's wait counting then falls back to waiting for every load in flight,
// which put a wait for the prefetched chunk at the top of the walk.
template <int F>
struct ChunkPipe {
    int t0, t1, t2;   // this lane's entries in the next chunk to stage and the two after it
    int4 q1;          // index row of t1
    TetData<F> d0;    // vertices and features of t0

    template <class E>
    __device__ inline void start(E entry, const int4* __restrict__ indices, const float* __restrict__ verts,
                                 const float* __restrict__ feats) {
        t0 = entry(0);
        t1 = entry(1);
        t2 = entry(2);
        q1 = indices[t1 >= 0 ? t1 : 0];
        const int u0 = t0 >= 0 ? t0 : 0;
        load_tet_data<F>(d0, indices[u0], u0, verts, feats);
    }

    // stages this lane's entry of the current chunk (a lane with no entry stages nothing)
    template <class St>
    __device__ inline void stage(St& S, int slot, const CamArgs& cam) {
        if (t0 >= 0) stage_data(S, slot, t0, cam, d0);
    }

    // issues the loads of the chunks after chunk c, the one just staged
    template <class E>
    __device__ inline void advance(E entry, int c, const int4* __restrict__ indices,
                                   const float* __restrict__ verts, const float* __restrict__ feats) {
        const int tn = entry(c + 3);
        const int4 qn = indices[t2 >= 0 ? t2 : 0];
        load_tet_data<F>(d0, q1, t1 >= 0 ? t1 : 0, verts, feats);
        t0 = t1;
        t1 = t2;
        t2 = tn;
        q1 = qn;
    }
};

template <int WG, int F, bool BWD, bool P64, bool VTX>
__device__ inline void read_grads(const Stage<WG, F, BWD, P64, VTX>& S, int slot, float* g) {
    constexpr int NGRD = NG_<F>::NGRD;
#pragma unroll
    for (int k = 0; k < (NGRD + 3) / 4; ++k) {
        float4 v = S.fg[k][slot];
        if (4 * k + 0 < NGRD) g[4 * k + 0] = v.x;
        if (4 * k + 1 < NGRD) g[(4 * k + 1) % NGRD] = v.y;
        if (4 * k + 2 < NGRD) g[(4 * k + 2) % NGRD] = v.z;
        if (4 * k + 3 < NGRD) g[(4 * k + 3) % NGRD] = v.w;
    }
}

// ------------------------------------------------------------------------------- forward --

// Blend of one hit into the pixel state (integrate_tetra_segment and the distortion moments);
// returns true when the pixel terminates (log-transmittance below -5.5) after this tet.
// log T is summed in FP64 and stored rounded once: the backward rebuilds each tet's log T from the
// stored final value and the x of the tets behind it, so that value carries no accumulated error.
template <int WG, int F, typename St, typename H>
__device__ __forceinline__ bool blend_step(const St& S, int slot, const H& h, float rdx, float rdy, float rdz,
                                           float mt, float& cr, float& cg, float& cb, double& logT, float& X,
                                           float& Y, float& Z, float& W, float& V, bool& blended) {
    constexpr int NGRD = NG_<F>::NGRD;
    const float4 fa = S.fa[slot];
    const float dt = h.dt;
    const float x = fa.x * dt;
    blended = h.hit && x > 0.f;
    if (!blended) return false;
    float gr[NGRD];
    read_grads(S, slot, gr);
    const float em1 = expm1_neg(x);
    const float alpha = 1.0f + em1;
    float w0, w1;
    seg_weights(x, em1, alpha, w0, w1);
    const float T = __expf((float)logT);
    const float gd0 = gr[0] * rdx + gr[1] * rdy + gr[2] * rdz;
    float gd1 = gd0, gd2 = gd0;
    if (F == 13) {
        gd1 = gr[3 % NGRD] * rdx + gr[4 % NGRD] * rdy + gr[5 % NGRD] * rdz;
        gd2 = gr[6 % NGRD] * rdx + gr[7 % NGRD] * rdy + gr[8 % NGRD] * rdz;
    }
    // w0 * c(s_x) + w1 * c(s_e) with c(s) = base' + s * gd
    const float wsum = w0 + w1, wsl = w0 * h.sx + w1 * h.se;
    cr += T * (wsum * fa.y + wsl * gd0);
    cg += T * (wsum * fa.z + wsl * gd1);
    cb += T * (wsum * fa.w + wsl * gd2);
    logT -= (double)x;
    const float te = h.se - mt, tx_ = h.sx - mt;
    const float m1 = fa.x * (te + tx_) * 0.5f;
    X += 2.0f * m1 * Z;
    Y += 2.0f * fa.x * W;
    Z += fa.x;
    W += m1;
    V += fa.x * fa.x * dt * (1.0f / 3.0f);
    return logT < -5.5;
}

// Forward of the tiles of workgroup b of one view. Tile ids are offset by tile_base, the sorted
// list holds tet ids offset by tet_base, outputs start at pixel pix_base; jitter may be null
// (pixel centres).
template <int TS, int F, bool P64>
__device__ __forceinline__ void fwd_tiles(Stage<Cfg<TS>::WG, F, false, P64>& S, int& s_chunks, const CamArgs& cam,
                                          int b, const int* __restrict__ sorted, int64_t n_sorted,
                                          const int2* __restrict__ ranges, int64_t tile_base, int64_t tet_base,
                                          const int4* __restrict__ indices, const float* __restrict__ verts,
                                          const float* __restrict__ feats, const float2* __restrict__ jitter,
                                          float4* __restrict__ out_img, float* __restrict__ out_dist,
                                          int* __restrict__ out_nc, int64_t pix_base,
                                          uint8_t* __restrict__ flags = nullptr) {
    using C = Cfg<TS>;
    constexpr int NGRD = NG_<F>::NGRD;
    using T = typename Prec<P64>::R;
    using PL = typename Prec<P64>::PL;

    int g, r, tx, ty, px, py;
    bool tile_ok, inside;
    locate<TS>(cam, b, g, r, tx, ty, px, py, tile_ok, inside);
    int2 rg = load_range(ranges, tile_ok, (int)(tile_base + ty * cam.gw + tx), n_sorted);
    int n = rg.y - rg.x;
    if (threadIdx.x == 0) s_chunks = 0;
    __syncthreads();
    if (r == 0) atomicMax(&s_chunks, (n + C::GROUP - 1) / C::GROUP);
    const int64_t pix = pix_base + (int64_t)py * cam.W + px;
    RayT<T> ray = {0, 0, 1};
    if (inside) {
        float2 j = jitter ? jitter[pix - pix_base] : make_float2(0.5f, 0.5f);
        ray = make_ray<T>(cam, px, py, j.x, j.y);
    }
    const float rdx = (float)ray.dx, rdy = (float)ray.dy, rdz = (float)ray.dz;
    const float mt = cam.min_t;
    float cr = 0.f, cg = 0.f, cb = 0.f;
    double logT = 0.0;
    float X = 0.f, Y = 0.f, Z = 0.f, W = 0.f, V = 0.f;
    int nc = 0;
    bool alive = inside;
    __syncthreads();
    const int nchunks = s_chunks;
    auto entry = [&](int cc) {
        const int kk = cc * C::GROUP + r;
        return kk < n ? (int)(sorted[rg.x + kk] - tet_base) : -1;
    };
    ChunkPipe<F> pipe;
    if (nchunks > 0) pipe.start(entry, indices, verts, feats);
    for (int c = 0; c < nchunks; ++c) {
        if (!__syncthreads_or(alive)) break;
        uint64_t blend_bits = 0;
        pipe.stage(S, threadIdx.x, cam);
        __syncthreads();
        pipe.advance(entry, c, indices, verts, feats);
        if (alive) {
            int cnt = n - c * C::GROUP;
            cnt = cnt < C::GROUP ? cnt : C::GROUP;
            const int base = g * C::GROUP;
            // this lane's blends of the chunk's entries, one bit per entry (see the flags below)
            auto record = [&](int jj, bool bl) { blend_bits |= (uint64_t)bl << jj; };
            // two slots per step: both intersections (independent of the pixel state) first,
            // then the two blends in list order
            for (int jj = 0; jj < cnt; jj += 2) {
                const bool two = jj + 1 < cnt;
                const int s0 = base + jj, s1 = two ? s0 + 1 : s0;
                const PL p0[4] = {S.pl[0][s0], S.pl[1][s0], S.pl[2][s0], S.pl[3][s0]};
                const PL p1[4] = {S.pl[0][s1], S.pl[1][s1], S.pl[2][s1], S.pl[3][s1]};
                const auto h0 = intersect(p0, ray, mt);
                const auto h1 = intersect(p1, ray, mt);
                ++nc;
                bool bl;
                const bool t0 = blend_step<C::WG, F>(S, s0, h0, rdx, rdy, rdz, mt, cr, cg, cb, logT, X, Y, Z, W, V, bl);
                record(jj, bl);
                if (t0) {
                    alive = false;
                    break;
                }
                if (!two) break;
                ++nc;
                const bool t1 = blend_step<C::WG, F>(S, s1, h1, rdx, rdy, rdz, mt, cr, cg, cb, logT, X, Y, Z, W, V, bl);
                record(jj + 1, bl);
                if (t1) {
                    alive = false;
                    break;
                }
            }
        }
        if constexpr (C::GROUP <= 64) {
            if (flags) {
                // entry r of the chunk: 1 when some pixel of the tile blended it; written by lane r of
                // the tile group for every entry of the chunk (entries past every pixel's last one are
                // never read back)
                uint64_t any = blend_bits;
                any |= bits_dpp<0xB1>(any);      // quad_perm [1,0,3,2]
                any |= bits_dpp<0x4E>(any);      // quad_perm [2,3,0,1]
                any |= bits_dpp<0x124>(any);     // row_ror:4
                any |= bits_dpp<0x128>(any);     // row_ror:8
                if (C::GROUP == 64) {
                    any |= __shfl_xor(any, 16, 64);
                    any |= __shfl_xor(any, 32, 64);
                }
                const int cntc = n - c * C::GROUP;
                if (r < cntc) flags[rg.x + c * C::GROUP + r] = (uint8_t)((any >> r) & 1);
            }
        }
    }
    if (inside) {
        out_img[pix] = make_float4(cr, cg, cb, (float)logT);
        float* d = out_dist + 5 * pix;
        d[0] = X; d[1] = Y; d[2] = Z; d[3] = W; d[4] = V;
        out_nc[pix] = nc;
    }
}


template <int TS, int F, bool P64>
__global__ __launch_bounds__(Cfg<TS>::WG) void composite_fwd_kernel(
    CamArgs cam, const int* __restrict__ sorted, int64_t n_sorted, const int2* __restrict__ ranges,
    const int4* __restrict__ indices, const float* __restrict__ verts, const float* __restrict__ feats,
    const float2* __restrict__ jitter, float4* __restrict__ out_img, float* __restrict__ out_dist,
    int* __restrict__ out_nc, uint8_t* __restrict__ flags) {
    __shared__ Stage<Cfg<TS>::WG, F, false, P64> S;
    __shared__ int s_chunks;
    fwd_tiles<TS, F, P64>(S, s_chunks, cam, remap_block(blockIdx.x, gridDim.x), sorted, n_sorted, ranges, 0, 0,
                          indices, verts, feats, jitter, out_img, out_dist, out_nc, 0, flags);
}

// Batched forward: the workgroups of all views of a plan in one launch (pixel-centre rays).
template <int TS, int F, bool P64>
__global__ __launch_bounds__(Cfg<TS>::WG) void batch_composite_fwd_kernel(
    const void* __restrict__ plan, const int* __restrict__ sorted, int64_t n_sorted, const int2* __restrict__ ranges,
    float4* __restrict__ out_img, float* __restrict__ out_dist, int* __restrict__ out_nc) {
    __shared__ Stage<Cfg<TS>::WG, F, false, P64> S;
    __shared__ int s_chunks;
    const PlanHeader* h = static_cast<const PlanHeader*>(plan);
    const ViewRec* views = plan_views(plan);
    const int bg = remap_block(blockIdx.x, gridDim.x);
    if (bg >= h->total_blocks) return;  // launch sized for a capacity
    int lo = 0, hi = h->n_views - 1;  // last view with block_offset <= bg
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (views[mid].block_offset <= bg) lo = mid;
        else hi = mid - 1;
    }
    const ViewRec& v = views[lo];
    fwd_tiles<TS, F, P64>(S, s_chunks, v.cam, (int)(bg - v.block_offset), sorted, n_sorted, ranges, v.tile_offset,
                          v.tet_offset, uniform_global_ptr(reinterpret_cast<const int4*>(v.indices)),
                          uniform_global_ptr(v.verts), uniform_global_ptr(v.feats), nullptr, out_img, out_dist,
                          out_nc, v.pixel_offset);
}

// ------------------------------------------------------------------------------ backward --

// DPP lane permutes within a row of 16 lanes (full-rate VALU, no LDS traffic).
template <int CTRL>
__device__ inline float dpp(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// Sum of v over the reduction group, returned in every lane of the group. Within a row of 16:
// quad_perm xor 1 and xor 2 give each lane its quad's sum, row_ror 4 and 8 add the other quads.
// A 64-lane group then adds the other rows with two cross-row shuffles.
template <int N>
__device__ inline void group_sum(float (&v)[N], int width) {
#pragma unroll
    for (int i = 0; i < N; ++i) {
        float x = v[i];
        x += dpp<0xB1>(x);   // quad_perm [1,0,3,2]
        x += dpp<0x4E>(x);   // quad_perm [2,3,0,1]
        x += dpp<0x124>(x);  // row_ror:4
        x += dpp<0x128>(x);  // row_ror:8
        if (width > 16) {
            x += __shfl_xor(x, 16, 64);
            x += __shfl_xor(x, 32, 64);
        }
        v[i] = x;
    }
}

// Per-tet gradients from the reduced per-slot sums acc = [dsig, dbase[3], Sq[NCH], Rg[3*NCH],
// u[4], q[12]]: Sq / Rg are the sums of the colour adjoints at the entry and exit points and of
// those adjoints times the points' offsets from v0; per plane u is the sum of ds/den and q the
// sum of u (x - v0) over its crossing points x. Outputs 0..11 are the four vertices' gradients,
// 12.. the feature gradients. Colour gradients enter through the base colour at v0; each face's
// plane record n = +-cross(b - a, c - a), e = n . (cam - a) passes its adjoint to a, b, c.
// The face normals are recomputed from the vertices in the plane precision (the same cross
// product plane_rec formed, flipped by the recorded sign), so no plane record is read here.
// The map from acc to the outputs is linear with coefficients fixed by the tet and the camera,
// so sums of acc over tiles may be converted once per tet.
template <int F, bool P64, int NV, int NOUT>
__device__ inline void tet_grads(const float3 (&V4)[4], int signs, const float (&gr)[NG_<F>::NGRD],
                                 const float (&acc)[NV], float3 cp, float (&outv)[NOUT]) {
    using T = typename Prec<P64>::R;
    constexpr int NGRD = NG_<F>::NGRD;
    constexpr int NCH = F == 7 ? 1 : 3;
    float3 gv[4] = {f3(0, 0, 0), f3(0, 0, 0), f3(0, 0, 0), f3(0, 0, 0)};
    outv[12] = acc[0];
    outv[13] = acc[1]; outv[14] = acc[2]; outv[15] = acc[3];
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch) {
        float Sq = acc[4 + ch];
        float3 grow = f3(gr[3 * ch], gr[(3 * ch + 1) % NGRD], gr[(3 * ch + 2) % NGRD]);
        gv[0] = gv[0] - Sq * grow;
        outv[16 + 3 * ch] = acc[4 + NCH + 3 * ch];
        outv[17 + 3 * ch] = acc[5 + NCH + 3 * ch];
        outv[18 + 3 * ch] = acc[6 + NCH + 3 * ch];
    }
    const int PU = 4 + NCH + 3 * NCH;
    const int fa_[4] = {0, 0, 0, 1}, fb_[4] = {1, 1, 2, 2}, fc_[4] = {2, 3, 3, 3};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const float u = acc[PU + i];
        const float3 rr = f3(acc[PU + 4 + 3 * i], acc[PU + 5 + 3 * i], acc[PU + 6 + 3 * i]);
        const int a = fa_[i], b = fb_[i], cc = fc_[i];
        const float sgn = ((signs >> i) & 1) ? -1.f : 1.f;
        const V3<T> va = {(T)V4[a].x, (T)V4[a].y, (T)V4[a].z}, vb = {(T)V4[b].x, (T)V4[b].y, (T)V4[b].z};
        const V3<T> vc = {(T)V4[cc].x, (T)V4[cc].y, (T)V4[cc].z};
        const V3<T> n = v3cross(v3sub(vb, va), v3sub(vc, va));
        const float3 nn = sgn * f3((float)n.x, (float)n.y, (float)n.z);
        // d/dn of -(e / den) summed over pixels: -sum u (x - v_a) = -(sum u (x - v0) + u (v0 - v_a))
        const float3 gn = a == 0 ? f3(-rr.x, -rr.y, -rr.z) : (-u) * (V4[0] - V4[a]) - rr;
        gv[a] = gv[a] + u * nn;
        const float3 gcr = sgn * gn;
        const float3 e1 = V4[b] - V4[a], e2 = V4[cc] - V4[a];
        const float3 ge1 = cross3(e2, gcr), ge2 = cross3(gcr, e1);
        gv[b] = gv[b] + ge1;
        gv[cc] = gv[cc] + ge2;
        gv[a] = gv[a] - (ge1 + ge2);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        outv[3 * v] = gv[v].x;
        outv[3 * v + 1] = gv[v].y;
        outv[3 * v + 2] = gv[v].z;
    }
}

template <int WG, int F, bool P64, int NV, int NOUT>
__device__ inline void bwd_outputs(const Stage<WG, F, true, P64>& S, int slot, const float (&acc)[NV], float3 cp,
                                   float (&outv)[NOUT]) {
    const float4 q0 = S.vtx[0][slot], q1 = S.vtx[1][slot], q2 = S.vtx[2][slot];
    const float3 V4[4] = {f3(q0.x, q0.y, q0.z), f3(q0.w, q1.x, q1.y), f3(q1.z, q1.w, q2.x), f3(q2.y, q2.z, q2.w)};
    float gr[NG_<F>::NGRD];
    read_grads(S, slot, gr);
    tet_grads<F, P64, NV, NOUT>(V4, S.meta[1][slot], gr, acc, cp, outv);
}

// Per-tet sums of the plane-record adjoints, one 128-byte row per tet (at most 32 floats: the
// 16-float halves are whole 64-byte segments, the unit in which float atomics are serviced).
// Row position pos holds sum j = row_sum_index(pos): the order in which group16_transpose_sum
// leaves the sums in the lanes, so that each half-row is added by one instruction.
constexpr int kRowFloats = 32;

__host__ __device__ constexpr int row_sum_index(int pos) {
    return 16 * ((pos >> 3) & 1) + 8 * ((pos >> 2) & 1) + 4 * ((pos >> 4) & 1) + (pos & 3);
}

// Sum over a 16-lane group (one DPP row) of 32 values per lane, transposed: afterwards lane r
// holds the group sums of values row_sum_index(r) (o0) and row_sum_index(16 + r) (o1). Two
// halving steps exchange half of the values with a partner lane, r ^ 8 (row_ror:8) and then
// r ^ 7 (row_half_mirror), and add only the half the lane keeps; the quad's 4 lanes are then
// summed with two quad permutes: 24 exchange adds with 48 selects, 16 quad adds and 6 final
// selects in place of the 4 x 32 adds of summing every value in every lane.
__device__ inline void group16_transpose_sum(const float (&a)[32], int r, float& o0, float& o1) {
    const bool h8 = (r & 8) != 0;
    float k[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const float keep = h8 ? a[16 + i] : a[i];
        const float send = h8 ? a[i] : a[16 + i];
        k[i] = keep + dpp<0x128>(send);   // row_ror:8
    }
    const bool h4 = (r & 4) != 0;
    float l[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const float keep = h4 ? k[8 + i] : k[i];
        const float send = h4 ? k[i] : k[8 + i];
        l[i] = keep + dpp<0x141>(send);   // row_half_mirror
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        l[i] += dpp<0xB1>(l[i]);          // quad_perm [1,0,3,2]
        l[i] += dpp<0x4E>(l[i]);          // quad_perm [2,3,0,1]
    }
    const int p = r & 3;
    o0 = p == 0 ? l[0] : p == 1 ? l[1] : p == 2 ? l[2] : l[3];
    o1 = p == 0 ? l[4] : p == 1 ? l[5] : p == 2 ? l[6] : l[7];
}

// ROWS: instead of converting each (tile, tet) sum to vertex and feature gradients, add the sums
// themselves to the tet's row of `rows` (kRowFloats floats per tet); composite_tet_grads_kernel
// converts the rows once per tet. Otherwise gradients go straight to gverts / gfeats.
// FLAGS (tile groups within one wave): only list entries whose forward blend flag is set are
// staged and walked; a chunk's flagged entries are compacted in list order.
template <int TS, int F, bool P64, bool ROWS, bool FLAGS = false>
__global__ __launch_bounds__(Cfg<TS>::WG) void composite_bwd_kernel(
    CamArgs cam, const int* __restrict__ sorted, int64_t n_sorted, const int2* __restrict__ ranges,
    const int4* __restrict__ indices, const float* __restrict__ verts, const float* __restrict__ feats,
    const float2* __restrict__ jitter, const float4* __restrict__ img, const float* __restrict__ dist,
    const int* __restrict__ ncontrib, const float4* __restrict__ gimg, const float* __restrict__ gdist,
    float* __restrict__ gverts, float* __restrict__ gfeats, float* __restrict__ rows,
    const uint8_t* __restrict__ flags) {
    using C = Cfg<TS>;
    static_assert(!FLAGS || C::GROUP <= 64, "blend flags need the tile group within one wave");
    constexpr int NGRD = NG_<F>::NGRD;
    constexpr int NCH = F == 7 ? 1 : 3;        // independent colour-gradient rows
    // reduced per-tet values: dsig, dbase[3], Sq[NCH], Rg[3*NCH], u[4], q[12]
    constexpr int NV = 1 + 3 + NCH + 3 * NCH + 4 + 12;
    constexpr int NOUT = 12 + F;                // vertex + feature gradients per tet
    using T = typename Prec<P64>::R;
    using PL = typename Prec<P64>::PL;
    __shared__ Stage<C::WG, F, true, P64, !ROWS> S;
    __shared__ int s_nc[C::NG];
    static_assert(NV <= kRowFloats, "row too small");
    // deferred per-lane epilogue of the direct path: the reduction group is the tile group, the
    // workgroup one wave, and the slots' gradients fit over the plane records (FP64 records;
    // FP32 ones are too small)
    constexpr bool kDefer = !ROWS && C::RED == C::GROUP && C::WG == 64 &&
                            sizeof(float) * C::WG * (NOUT + 1) <= sizeof(S.pl);

    int g, r, tx, ty, px, py;
    bool tile_ok, inside;
    locate<TS>(cam, remap_block(blockIdx.x, gridDim.x), g, r, tx, ty, px, py, tile_ok, inside);
    int2 rg = load_range(ranges, tile_ok, ty * cam.gw + tx, n_sorted);
    const int64_t pix = (int64_t)py * cam.W + px;
    if (threadIdx.x < C::NG) s_nc[threadIdx.x] = 0;
    __syncthreads();
    int nc = 0;
    RayT<T> ray = {0, 0, 1};
    double A = 0.0;   // log T, stepped back in FP64 from the stored final value
    float Z = 0.f, W = 0.f;
    float dCr = 0.f, dCg = 0.f, dCb = 0.f, dA = 0.f, dX = 0.f, dY = 0.f, dZ = 0.f, dW = 0.f, dV = 0.f;
    if (inside) {
        nc = ncontrib[pix];
        float2 j = jitter[pix];
        ray = make_ray<T>(cam, px, py, j.x, j.y);
        float4 im = img[pix];
        A = im.w;
        Z = dist[5 * pix + 2];
        W = dist[5 * pix + 3];
        float4 gi = gimg[pix];
        dCr = gi.x; dCg = gi.y; dCb = gi.z; dA = gi.w;
        const float* gd = gdist + 5 * pix;
        dX = gd[0]; dY = gd[1]; dZ = gd[2]; dW = gd[3]; dV = gd[4];
        atomicMax(&s_nc[g], nc);
    }
    __syncthreads();
    const float rdx = (float)ray.dx, rdy = (float)ray.dy, rdz = (float)ray.dz;
    const int nmax_g = min(s_nc[g], rg.y - rg.x);   // list entries this tile must revisit
    int nchunks_wg = 0;
#pragma unroll
    for (int q = 0; q < C::NG; ++q) {
        int v = (s_nc[q] + C::GROUP - 1) / C::GROUP;
        nchunks_wg = v > nchunks_wg ? v : nchunks_wg;
    }
    const int nchunks_g = (nmax_g + C::GROUP - 1) / C::GROUP;
    const float mt = cam.min_t;
    const float3 cp = f3(cam.cam_pos[0], cam.cam_pos[1], cam.cam_pos[2]);
    const int lane = threadIdx.x & 63;
    const int red_base = lane & ~(C::RED - 1);
    const uint64_t red_mask = (C::RED == 64) ? ~0ull : (((1ull << C::RED) - 1ull) << red_base);
    float kept[NV];   // reduced sums of the slot with in-group index r (deferred epilogue)
    bool kept_on = false;

    for (int c = nchunks_wg - 1; c >= 0; --c) {
        __syncthreads();
        int k = c * C::GROUP + r;
        const bool valid = c < nchunks_g && k < nmax_g;
        const int base = g * C::GROUP;
        int cnt;
        if constexpr (FLAGS) {
            // compact the chunk to the entries some pixel of the tile blended in the forward
            const bool f = valid && flags[rg.x + k] != 0;
            const int gb = lane & ~(C::GROUP - 1);
            const uint64_t gm = (C::GROUP == 64 ? ~0ull : ((1ull << C::GROUP) - 1ull)) << gb;
            const uint64_t m = __ballot(f) & gm;
            if (f) {
                const int ci = __popcll(m & ((1ull << lane) - 1ull));
                stage_tet(S, base + ci, sorted[rg.x + k], cam, indices, verts, feats);
                S.kidx[base + ci] = k;
            }
            cnt = __popcll(m);
        } else {
            if (valid) stage_tet(S, threadIdx.x, sorted[rg.x + k], cam, indices, verts, feats);
            // tile groups past their last chunk run no slots (cnt = 0) but still reach the barriers below
            cnt = c < nchunks_g ? nmax_g - c * C::GROUP : 0;
            cnt = cnt < C::GROUP ? cnt : C::GROUP;
        }
        __syncthreads();
        // slot body for slot jj with its intersection h (computed ahead, see below)
        auto slot_body = [&](int jj, const HitT<typename Prec<P64>::R>& h) {
            using TP = typename Prec<P64>::R;
            const int slot = base + jj;
            const int kk = FLAGS ? S.kidx[slot] : c * C::GROUP + jj;
            float acc[NV];
#pragma unroll
            for (int i = 0; i < NV; ++i) acc[i] = 0.f;
            bool active = false;
            if (inside && kk < nc) {
                float4 fa = S.fa[slot];
                const float sig = fa.x;
                float dt = h.dt;
                float x = sig * dt;
                if (h.hit && x > 0.f) {
                    active = true;
                    float gr[NGRD];
                    read_grads(S, slot, gr);
                    float em1 = expm1_neg(x);
                    float alpha = 1.0f + em1;
                    float w0, w1, dphi;
                    seg_weights_dphi(x, em1, alpha, w0, w1, dphi);
                    float gd[3];
                    gd[0] = gr[0] * rdx + gr[1] * rdy + gr[2] * rdz;
                    gd[1] = gd[0];
                    gd[2] = gd[0];
                    if (F == 13) {
                        gd[1] = gr[3 % NGRD] * rdx + gr[4 % NGRD] * rdy + gr[5 % NGRD] * rdz;
                        gd[2] = gr[6 % NGRD] * rdx + gr[7 % NGRD] * rdy + gr[8 % NGRD] * rdz;
                    }
                    float bse[3] = {fa.y, fa.z, fa.w};
                    float ce[3], cx[3], rgb[3];
#pragma unroll
                    for (int ch = 0; ch < 3; ++ch) {
                        ce[ch] = bse[ch] + h.se * gd[ch];
                        cx[ch] = bse[ch] + h.sx * gd[ch];
                        rgb[ch] = w0 * cx[ch] + w1 * ce[ch];
                    }
                    // undo this tet's blend
                    const double Ap = A + (double)x;
                    const float T = __expf((float)Ap);
                    const float te = h.se - mt, tx_ = h.sx - mt;
                    const float m1 = sig * (te + tx_) * 0.5f;
                    const float zp = Z - sig, wp = W - m1;
                    const float dC[3] = {dCr, dCg, dCb};
                    float drgb[3];
                    float dT = 0.f;
#pragma unroll
                    for (int ch = 0; ch < 3; ++ch) {
                        drgb[ch] = dC[ch] * T;
                        dT += dC[ch] * rgb[ch];
                    }
                    float dxx = -dA;
                    const float dm1 = dX * 2.0f * zp + dW;
                    const float dm0 = dY * 2.0f * wp + dZ;
                    const float dZn = dZ + dX * 2.0f * m1;
                    const float dWn = dW + dY * 2.0f * sig;
                    float dsig = dm0 + dm1 * (te + tx_) * 0.5f + dV * 2.0f * sig * dt * (1.0f / 3.0f);
                    float ddt = dV * sig * sig * (1.0f / 3.0f);
                    float dse = dm1 * sig * 0.5f, dsx = dm1 * sig * 0.5f;
                    // entry / exit points relative to v0, formed in the plane precision from
                    // cam - v0 and s d (each ~ the camera distance, far larger than the tet), so
                    // the sums below carry no cancellation between the two
                    const TP c0 = S.cv[0][slot], c1 = S.cv[1][slot], c2 = S.cv[2][slot];
                    const TP sec = h.se_p > (TP)mt ? h.se_p : (TP)mt;
                    const float qex = (float)(c0 + sec * ray.dx), qey = (float)(c1 + sec * ray.dy),
                                qez = (float)(c2 + sec * ray.dz);
                    const float qxx = (float)(c0 + h.sx_p * ray.dx), qxy = (float)(c1 + h.sx_p * ray.dy),
                                qxz = (float)(c2 + h.sx_p * ray.dz);
                    float dw0 = 0.f, dw1 = 0.f;
                    float dqe[3], dqx[3];
#pragma unroll
                    for (int ch = 0; ch < 3; ++ch) {
                        acc[1 + ch] = drgb[ch] * (w0 + w1);
                        dqx[ch] = drgb[ch] * w0;
                        dqe[ch] = drgb[ch] * w1;
                        dw0 += drgb[ch] * cx[ch];
                        dw1 += drgb[ch] * ce[ch];
                    }
                    dxx += alpha * dw0 + (dw0 - dw1) * dphi;
                    dsig += dxx * dt;
                    ddt += dxx * sig;
                    dsx += ddt;
                    dse -= ddt;
                    acc[0] = dsig;
                    if (F == 7) {
                        float qe = dqe[0] + dqe[1] + dqe[2], qx = dqx[0] + dqx[1] + dqx[2];
                        acc[4] = qe + qx;
                        acc[5] = qe * qex + qx * qxx;
                        acc[6] = qe * qey + qx * qxy;
                        acc[7] = qe * qez + qx * qxz;
                        dse += qe * gd[0];
                        dsx += qx * gd[0];
                    } else {
#pragma unroll
                        for (int ch = 0; ch < 3; ++ch) {
                            acc[4 + ch] = dqe[ch] + dqx[ch];
                            acc[7 + 3 * ch] = dqe[ch] * qex + dqx[ch] * qxx;
                            acc[8 + 3 * ch] = dqe[ch] * qey + dqx[ch] * qxy;
                            acc[9 + 3 * ch] = dqe[ch] * qez + dqx[ch] * qxz;
                            dse += dqe[ch] * gd[ch];
                            dsx += dqx[ch] * gd[ch];
                        }
                    }
                    // plane adjoints: s_i = -e_i / den_i; per plane the sums of u = ds / den and of
                    // u (x - v0) over its crossing points x (the entering plane only counts when
                    // its crossing lies beyond min_t, where the clamped entry is the crossing)
                    const int PU = 4 + NCH + 3 * NCH;
                    float ue = (h.se_raw > mt) ? dse * rcp_(h.den_e) : 0.f;
                    float ux = dsx * rcp_(h.den_x);
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const bool en = h.ie == i, ex = h.ix == i;
                        acc[PU + i] = (en ? ue : 0.f) + (ex ? ux : 0.f);
                        acc[PU + 4 + 3 * i] = (en ? ue * qex : 0.f) + (ex ? ux * qxx : 0.f);
                        acc[PU + 5 + 3 * i] = (en ? ue * qey : 0.f) + (ex ? ux * qxy : 0.f);
                        acc[PU + 6 + 3 * i] = (en ? ue * qez : 0.f) + (ex ? ux * qxz : 0.f);
                    }
                    // step the pixel state back over this tet
                    A = Ap; Z = zp; W = wp;
                    dA += dT * T;
                    dZ = dZn; dW = dWn;
                }
            }
            uint64_t any = __ballot(active) & red_mask;
            if (any == 0) return;
            if constexpr (ROWS && C::RED == 16) {
                // transposed group sum: lane r ends with row positions r and 16 + r
                float a32[kRowFloats];
#pragma unroll
                for (int i = 0; i < kRowFloats; ++i) a32[i] = i < NV ? acc[i] : 0.f;
                float o0, o1;
                group16_transpose_sum(a32, r, o0, o1);
                float* row = rows + (int64_t)S.meta[0][slot] * kRowFloats;
                if (o0 != 0.f) unsafeAtomicAdd(row + r, o0);
                if (o1 != 0.f) unsafeAtomicAdd(row + 16 + r, o1);
                return;
            }
            group_sum<NV>(acc, C::RED);
            if constexpr (kDefer) {
                // the lane whose in-group index is this slot keeps the sums; the chunk's epilogues
                // then run once per lane instead of once per slot in every lane of the group
                if (r == jj) {
#pragma unroll
                    for (int i = 0; i < NV; ++i) kept[i] = acc[i];
                    kept_on = true;
                }
                return;
            }
            if constexpr (ROWS) {
                // lane l of the reduction group adds row position l (the sum row_sum_index(l))
                const int tet = S.meta[0][slot];
#pragma unroll
                for (int l0 = 0; l0 < kRowFloats; l0 += C::RED) {
                    const int l = lane - red_base + l0;
                    if (l < kRowFloats) {
                        const int j = row_sum_index(l);
                        float val = 0.f;
#pragma unroll
                        for (int i = 0; i < NV; ++i) val = (j == i) ? acc[i] : val;
                        if (val != 0.f) unsafeAtomicAdd(rows + (int64_t)tet * kRowFloats + l, val);
                    }
                }
            } else {
                float outv[NOUT];
                bwd_outputs<C::WG, F, P64, NV, NOUT>(S, slot, acc, cp, outv);
                const int tet = S.meta[0][slot];
                // lane l of the reduction group adds outputs l, l + RED, ... (NOUT may exceed RED)
#pragma unroll
                for (int l0 = 0; l0 < NOUT; l0 += C::RED) {
                    const int l = lane - red_base + l0;
                    if (l < NOUT) {
                        float val = 0.f;
#pragma unroll
                        for (int i = 0; i < NOUT; ++i) val = (l == i) ? outv[i] : val;
                        if (val != 0.f) {
                            if (l < 12) {
                                const int4 qi = indices[tet];
                                const int vsel = l / 3;
                                const int vid = vsel == 0 ? qi.x : vsel == 1 ? qi.y : vsel == 2 ? qi.z : qi.w;
                                unsafeAtomicAdd(gverts + 3 * (int64_t)vid + (l % 3), val);
                            } else {
                                unsafeAtomicAdd(gfeats + (int64_t)F * tet + (l - 12), val);
                            }
                        }
                    }
                }
            }
        };
        // two slots per step: both intersections (independent of the pixel state) first, then the
        // two slots back to front
        for (int jj = cnt - 1; jj >= 0; jj -= 2) {
            const int s0 = base + jj, s1 = jj > 0 ? s0 - 1 : s0;
            const PL p0[4] = {S.pl[0][s0], S.pl[1][s0], S.pl[2][s0], S.pl[3][s0]};
            const PL p1[4] = {S.pl[0][s1], S.pl[1][s1], S.pl[2][s1], S.pl[3][s1]};
            const auto h0 = intersect(p0, ray, mt);
            const auto h1 = intersect(p1, ray, mt);
            slot_body(jj, h0);
            if (jj > 0) slot_body(jj - 1, h1);
        }
        if constexpr (kDefer && !ROWS) {
            // each lane turns its kept slot's sums into the tet's gradients; the gradients then go
            // through LDS (over the plane/colour records, dead by then) so that the atomics are
            // issued tet by tet with consecutive lanes on consecutive outputs, as in the path above
            // the gradients overlay the plane records only, which bwd_outputs does not read
            float* mine = reinterpret_cast<float*>(&S.pl[0][0]) + (base + r) * (NOUT + 1);
            if (kept_on) {
                float outv[NOUT];
                bwd_outputs<C::WG, F, P64, NV, NOUT>(S, base + r, kept, cp, outv);
#pragma unroll
                for (int i = 0; i < NOUT; ++i) mine[i] = outv[i];
            }
            mine[NOUT] = kept_on ? 1.f : 0.f;
            kept_on = false;
            __syncthreads();
#pragma unroll 1
            for (int jj = 0; jj < cnt; ++jj) {
                const float* out = reinterpret_cast<const float*>(&S.pl[0][0]) + (base + jj) * (NOUT + 1);
                if (out[NOUT] == 0.f) continue;          // uniform per tile group
                const int tet = S.meta[0][base + jj];
#pragma unroll
                for (int l0 = 0; l0 < NOUT; l0 += C::GROUP) {
                    const int l = r + l0;
                    if (l < NOUT) {
                        const float val = out[l];
                        if (val != 0.f) {
                            if (l < 12) {
                                const int4 qi = indices[tet];
                                const int vsel = l / 3;
                                const int vid = vsel == 0 ? qi.x : vsel == 1 ? qi.y : vsel == 2 ? qi.z : qi.w;
                                unsafeAtomicAdd(gverts + 3 * (int64_t)vid + (l % 3), val);
                            } else {
                                unsafeAtomicAdd(gfeats + (int64_t)F * tet + (l - 12), val);
                            }
                        }
                    }
                }
            }
        }
    }
}

// ------------------------------------------------------------- backward, one lane per tet --
//
// The 4x4-tile backward with the roles of lanes and pixels exchanged: one wave per tile, a chunk
// holds up to 64 of the tile's list entries (with FLAGS only the entries some pixel blended in the
// forward, gathered densely across the list), lane j the j-th entry counted from the back of what
// is left, its plane records, colour and cam - v0 in registers. The wave then walks the tile's 16
// pixels; every lane evaluates its tet for the pixel and adds the (pixel, tet) adjoints to its own
// 32 sums, so a (tile, tet) pair's sums are complete in one lane with no cross-lane reduction.
//
// The back-to-front walk of a pixel (bwd_alpha_blend, alphablend_shader_interp.slang:273-369)
// carries log T, the distortion moments Z and W and the adjoints dA, dZ, dW from tet to tet; each
// changes by a term of the tet it steps over. At lane j the walk has stepped over the entries of
// the later chunks and lanes 0..j-1 of this one, so its state is the state at the back of the
// chunk plus exclusive prefix sums over the lanes: one wave scan of (x, sigma, m1) and, as dA's
// terms dT e^A need the transmittance, a second one of dT e^A. The state at the back of the chunk
// lives in lane p of six registers for pixel p.

// Inclusive prefix sum over the 64 lanes of a wave: DPP row shifts sum within each row of 16 lanes,
// row broadcasts then carry the totals of rows 0 (into 1 and 3) and 0-1 (into 2 and 3).
template <int CTRL, int ROWMASK>
__device__ inline float dpp_rm(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWMASK, 0xF, false));
}
__device__ inline float wave_scan_incl(float x) {
    x += dpp_rm<0x111, 0xF>(x);   // row_shr:1
    x += dpp_rm<0x112, 0xF>(x);   // row_shr:2
    x += dpp_rm<0x114, 0xF>(x);   // row_shr:4
    x += dpp_rm<0x118, 0xF>(x);   // row_shr:8
    x += dpp_rm<0x142, 0xA>(x);   // row_bcast:15 into rows 1 and 3
    x += dpp_rm<0x143, 0xC>(x);   // row_bcast:31 into rows 2 and 3
    return x;
}

// Exclusive prefix sum over the wave: the values shifted one lane up (wave_shr:1; lane 0 reads 0),
// then the inclusive scan, so that lane j's sum never contains its own term (subtracting it from
// the inclusive sum would leave a rounding error of the size of that term).
__device__ inline float wave_scan_excl(float x) { return wave_scan_incl(dpp_rm<0x138, 0xF>(x)); }

// m ? a : b for a lane mask m of all ones or all zeros, as bit operations (kept branch-free)
__device__ inline float bitsel(int m, float a, float b) {
    return __int_as_float((__float_as_int(a) & m) | (__float_as_int(b) & ~m));
}
__device__ inline float lane_f(float v, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l)); }
__device__ inline double lane_f(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(uint32_t)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), l);
    return __longlong_as_double((long long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}

template <int F, bool P64, bool FLAGS>
__global__ __launch_bounds__(64, 3) void composite_bwd_tet_lanes_kernel(
    CamArgs cam, const int* __restrict__ sorted, int64_t n_sorted, const int2* __restrict__ ranges,
    const int4* __restrict__ indices, const float* __restrict__ verts, const float* __restrict__ feats,
    const float2* __restrict__ jitter, const float4* __restrict__ img, const float* __restrict__ dist,
    const int* __restrict__ ncontrib, const float4* __restrict__ gimg, const float* __restrict__ gdist,
    float* __restrict__ rows, const uint8_t* __restrict__ flags, const int* __restrict__ tile_order) {
    constexpr int NGRD = NG_<F>::NGRD;
    constexpr int NCH = F == 7 ? 1 : 3;
    constexpr int NV = 1 + 3 + NCH + 3 * NCH + 4 + 12;   // dsig, dbase[3], Sq[NCH], Rg[3*NCH], u[4], q[12]
    constexpr int PU = 4 + NCH + 3 * NCH;
    static_assert(NV <= kRowFloats, "row too small");
    using T = typename Prec<P64>::R;
    using PL = typename Prec<P64>::PL;
    __shared__ int s_pos[128];                      // gathered list positions, a ring
    // each lane's plane records while the pixels are walked, then the chunk's per-lane sums read
    // back row-wise (the records live in LDS rather than registers: two more waves per CU)
    __shared__ union {
        PL pl[4][64];
        float sum[64][kRowFloats + 1];
    } s_u;
    __shared__ int s_tet[64];
    auto& s_sum = s_u.sum;

    const int lane = threadIdx.x;
    const int tile = tile_order ? tile_order[blockIdx.x] : remap_block(blockIdx.x, gridDim.x);
    const int tx = tile % cam.gw, ty = tile / cam.gw;
    const int2 rg = load_range(ranges, true, tile, n_sorted);
    const int n = rg.y - rg.x;

    // pixel lane p < 16: the pixel's constants and its walk state at the back of the next chunk
    const int px = tx * 4 + (lane & 3), py = ty * 4 + ((lane >> 2) & 3);
    const bool inside = lane < 16 && px < cam.W && py < cam.H;
    const int64_t pix = (int64_t)py * cam.W + px;
    int pnc = 0;
    RayT<T> pray = {0, 0, 1};
    float sA = 0.f, sZ = 0.f, sW = 0.f, sdA = 0.f, sdZ = 0.f, sdW = 0.f;
    float pdCr = 0.f, pdCg = 0.f, pdCb = 0.f, pdX = 0.f, pdY = 0.f, pdV = 0.f;
    if (inside) {
        pnc = ncontrib[pix];
        const float2 j = jitter[pix];
        pray = make_ray<T>(cam, px, py, j.x, j.y);
        sA = img[pix].w;
        sZ = dist[5 * pix + 2];
        sW = dist[5 * pix + 3];
        const float4 gi = gimg[pix];
        pdCr = gi.x; pdCg = gi.y; pdCb = gi.z; sdA = gi.w;
        const float* gd = gdist + 5 * pix;
        pdX = gd[0]; pdY = gd[1]; sdZ = gd[2]; sdW = gd[3]; pdV = gd[4];
    }
    // entries at or beyond every pixel's n_contrib are never revisited
    int nmax = pnc;
    nmax = max(nmax, __shfl_xor(nmax, 1, 64));
    nmax = max(nmax, __shfl_xor(nmax, 2, 64));
    nmax = max(nmax, __shfl_xor(nmax, 4, 64));
    nmax = max(nmax, __shfl_xor(nmax, 8, 64));
    nmax = min(__builtin_amdgcn_readfirstlane(nmax), n);
    const float mt = cam.min_t;

    int hi = nmax;         // list positions [0, hi) not yet gathered
    int head = 0, npend = 0;
    for (;;) {
        // gather up to 64 entries, back to front, into the ring
        while (npend < 64 && hi > 0) {
            const int k = hi - 1 - lane;
            const bool f = k >= 0 && (!FLAGS || flags[rg.x + k] != 0);
            const uint64_t m = __ballot(f);
            if (f) s_pos[(head + npend + __popcll(m & ((1ull << lane) - 1ull))) & 127] = k;
            npend += __popcll(m);
            hi -= 64;
        }
        if (npend == 0) break;
        const int cnt = npend < 64 ? npend : 64;
        __syncthreads();
        const bool valid = lane < cnt;
        const int kpos = valid ? s_pos[(head + lane) & 127] : 0x7FFFFFFF;
        head += cnt;
        npend -= cnt;
        const int kmin = __builtin_amdgcn_readlane(kpos, cnt - 1);   // the chunk's first list position

        // this lane's tet: plane records (to LDS), colour and cam - v0 in registers
        PL pl[4];
        float4 fa = make_float4(0.f, 0.f, 0.f, 0.f);
        float gr[NGRD];
        T cv0 = 0, cv1 = 0, cv2 = 0;
        int tet = -1;
        if (valid) {
            tet = sorted[rg.x + kpos];
            TetData<F> d;
            load_tet_data<F>(d, indices[tet], tet, verts, feats);
            tet_record<F, P64>(d, cam, pl, fa, gr);
            cv0 = (T)cam.cam_pos[0] - (T)d.v[0].x;
            cv1 = (T)cam.cam_pos[1] - (T)d.v[0].y;
            cv2 = (T)cam.cam_pos[2] - (T)d.v[0].z;
        } else {
#pragma unroll
            for (int i = 0; i < 4; ++i) {   // planes no ray crosses ahead of the camera: a miss
                pl[i].x = 0; pl[i].y = 0; pl[i].z = 0; pl[i].w = 1;
            }
#pragma unroll
            for (int k = 0; k < NGRD; ++k) gr[k] = 0.f;
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) s_u.pl[i][lane] = pl[i];
        const float sig = fa.x;
        float acc[NV];
#pragma unroll
        for (int i = 0; i < NV; ++i) acc[i] = 0.f;
        bool touched = false;

        for (int p = 0; p < 16; ++p) {
            const int ncp = __builtin_amdgcn_readlane(pnc, p);
            if (ncp <= kmin) continue;                         // no entry of the chunk is revisited
            const RayT<T> ray = {lane_f(pray.dx, p), lane_f(pray.dy, p), lane_f(pray.dz, p)};
            const float rdx = (float)ray.dx, rdy = (float)ray.dy, rdz = (float)ray.dz;
            const PL plp[4] = {s_u.pl[0][lane], s_u.pl[1][lane], s_u.pl[2][lane], s_u.pl[3][lane]};
            const auto h = intersect(plp, ray, mt);
            const float dt = h.dt;
            const float x = sig * dt;
            const bool act = valid && kpos < ncp && h.hit && x > 0.f;
            const float te = h.se - mt, tx_ = h.sx - mt;
            const float m1 = sig * (te + tx_) * 0.5f;
            const float xa = act ? x : 0.f, sa = act ? sig : 0.f, ma = act ? m1 : 0.f;
            const float ex = wave_scan_excl(xa), es = wave_scan_excl(sa), em = wave_scan_excl(ma);
            const float A = lane_f(sA, p) + ex;
            const float Z = lane_f(sZ, p) - es;
            const float W = lane_f(sW, p) - em;
            const float dXp = lane_f(pdX, p), dYp = lane_f(pdY, p), dVp = lane_f(pdV, p);
            const float dZ = lane_f(sdZ, p) + dXp * 2.0f * em;
            const float dW = lane_f(sdW, p) + dYp * 2.0f * es;
            const float dC[3] = {lane_f(pdCr, p), lane_f(pdCg, p), lane_f(pdCb, p)};
            float em1 = 0.f, alpha = 0.f, w0 = 0.f, w1 = 0.f, dphi = 0.f, Tr = 0.f, dT = 0.f;
            float gd[3] = {0.f, 0.f, 0.f}, ce[3] = {0.f, 0.f, 0.f}, cx[3] = {0.f, 0.f, 0.f};
            if (act) {
                em1 = expm1_neg(x);
                alpha = 1.0f + em1;
                seg_weights_dphi(x, em1, alpha, w0, w1, dphi);
                gd[0] = gr[0] * rdx + gr[1] * rdy + gr[2] * rdz;
                gd[1] = gd[0];
                gd[2] = gd[0];
                if (F == 13) {
                    gd[1] = gr[3 % NGRD] * rdx + gr[4 % NGRD] * rdy + gr[5 % NGRD] * rdz;
                    gd[2] = gr[6 % NGRD] * rdx + gr[7 % NGRD] * rdy + gr[8 % NGRD] * rdz;
                }
                const float bse[3] = {fa.y, fa.z, fa.w};
#pragma unroll
                for (int ch = 0; ch < 3; ++ch) {
                    ce[ch] = bse[ch] + h.se * gd[ch];
                    cx[ch] = bse[ch] + h.sx * gd[ch];
                    dT += dC[ch] * (w0 * cx[ch] + w1 * ce[ch]);
                }
                Tr = __expf(A + x);    // transmittance in front of this tet
            }
            const float dtt = act ? dT * Tr : 0.f;
            const float edt = wave_scan_excl(dtt);
            const float dA = lane_f(sdA, p) + edt;
            if (act) {
                touched = true;
                const float zp = Z - sig, wp = W - m1;
                float drgb[3];
#pragma unroll
                for (int ch = 0; ch < 3; ++ch) drgb[ch] = dC[ch] * Tr;
                float dxx = -dA;
                const float dm1 = dXp * 2.0f * zp + dW;
                const float dm0 = dYp * 2.0f * wp + dZ;
                float dsig = dm0 + dm1 * (te + tx_) * 0.5f + dVp * 2.0f * sig * dt * (1.0f / 3.0f);
                float ddt = dVp * sig * sig * (1.0f / 3.0f);
                float dse = dm1 * sig * 0.5f, dsx = dm1 * sig * 0.5f;
                // entry / exit points relative to v0, formed in the plane precision from cam - v0
                // and s d, so the sums below carry no cancellation against the camera distance
                const T sec = h.se_p > (T)mt ? h.se_p : (T)mt;
                const float qex = (float)(cv0 + sec * ray.dx), qey = (float)(cv1 + sec * ray.dy),
                            qez = (float)(cv2 + sec * ray.dz);
                const float qxx = (float)(cv0 + h.sx_p * ray.dx), qxy = (float)(cv1 + h.sx_p * ray.dy),
                            qxz = (float)(cv2 + h.sx_p * ray.dz);
                float dw0 = 0.f, dw1 = 0.f;
                float dqe[3], dqx[3];
#pragma unroll
                for (int ch = 0; ch < 3; ++ch) {
                    acc[1 + ch] += drgb[ch] * (w0 + w1);
                    dqx[ch] = drgb[ch] * w0;
                    dqe[ch] = drgb[ch] * w1;
                    dw0 += drgb[ch] * cx[ch];
                    dw1 += drgb[ch] * ce[ch];
                }
                dxx += alpha * dw0 + (dw0 - dw1) * dphi;
                dsig += dxx * dt;
                ddt += dxx * sig;
                dsx += ddt;
                dse -= ddt;
                acc[0] += dsig;
                if (F == 7) {
                    const float qe = dqe[0] + dqe[1] + dqe[2], qx = dqx[0] + dqx[1] + dqx[2];
                    acc[4] += qe + qx;
                    acc[5] += qe * qex + qx * qxx;
                    acc[6] += qe * qey + qx * qxy;
                    acc[7] += qe * qez + qx * qxz;
                    dse += qe * gd[0];
                    dsx += qx * gd[0];
                } else {
#pragma unroll
                    for (int ch = 0; ch < 3; ++ch) {
                        acc[4 + ch] += dqe[ch] + dqx[ch];
                        acc[7 + 3 * ch] += dqe[ch] * qex + dqx[ch] * qxx;
                        acc[8 + 3 * ch] += dqe[ch] * qey + dqx[ch] * qxy;
                        acc[9 + 3 * ch] += dqe[ch] * qez + dqx[ch] * qxz;
                        dse += dqe[ch] * gd[ch];
                        dsx += dqx[ch] * gd[ch];
                    }
                }
                // plane adjoints: s_i = -e_i / den_i; per plane the sums of u = ds / den and of
                // u (x - v0) over its crossing points x (the entering plane only counts when its
                // crossing lies beyond min_t, where the clamped entry is the crossing)
                const float ue = (h.se_raw > mt) ? dse * rcp_(h.den_e) : 0.f;
                const float ux = dsx * rcp_(h.den_x);
                // the entering plane (den < 0) and the exiting one (den >= 0) are never the same
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int me = -(int)(h.ie == i), mx = -(int)(h.ix == i);
                    const float u = __int_as_float((__float_as_int(ue) & me) | (__float_as_int(ux) & mx));
                    acc[PU + i] += u;
                    acc[PU + 4 + 3 * i] += u * bitsel(me, qex, qxx);
                    acc[PU + 5 + 3 * i] += u * bitsel(me, qey, qxy);
                    acc[PU + 6 + 3 * i] += u * bitsel(me, qez, qxz);
                }
            }
            // pixel p's state at the back of the next chunk: add the whole chunk's terms
            const float tX = lane_f(ex + xa, 63), tS = lane_f(es + sa, 63), tM = lane_f(em + ma, 63),
                        tD = lane_f(edt + dtt, 63);
            if (lane == p) {
                sA += tX;
                sZ -= tS;
                sW -= tM;
                sdZ += dXp * 2.0f * tM;
                sdW += dYp * 2.0f * tS;
                sdA += tD;
            }
        }

        // the chunk's sums leave through LDS, so that each atomic instruction adds two whole rows
        // (consecutive lanes on consecutive floats of a 128-byte row)
        const uint64_t tm = __ballot(touched);
        __syncthreads();   // every lane's plane records are read before the sums overwrite them
        if (tm != 0) {
#pragma unroll
            for (int i = 0; i < NV; ++i) s_sum[lane][i] = acc[i];
            s_tet[lane] = tet;
            __syncthreads();
            const int half = lane >> 5, pos = lane & 31;
            const int j = row_sum_index(pos);
            for (int s0 = 0; s0 < cnt; s0 += 2) {
                if (((tm >> s0) & 3ull) == 0) continue;
                const int s = s0 + half;
                if (s < cnt && ((tm >> s) & 1ull) && j < NV) {
                    const float v = s_sum[s][j];
                    if (v != 0.f) unsafeAtomicAdd(rows + (int64_t)s_tet[s] * kRowFloats + pos, v);
                }
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------- forward, pixel quads --
//
// The 4x4-tile forward with four lanes per pixel: lane 4p + k walks pixel p over entry 4g + k of
// each group g of four consecutive list entries, so a pixel's walk advances four entries per step.
// Chunks of the tile's next 64 entries are staged in LDS, one entry per lane, through the load
// pipeline of the 8x8 and 16x16 forward. Within a step the four entries' intersections and blend
// terms do not depend on the pixel's state; the state in front of each entry (log T, Z, W) comes
// from exclusive prefix sums over the quad of lanes, the walk ends at the first blended entry
// behind which log T falls below -5.5 (as the sequential walk does), and sums over the quad of
// the entries up to it update the pixel's state, which all four lanes hold.
template <int CTRL>
__device__ inline float dpp_q(float v) {   // a quad permutation (all lanes of a quad are active together)
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
// N sums over the quad, every lane receiving them (butterflies, the N chains interleaved).
template <int N>
__device__ inline void quad_sum_n(float (&v)[N]) {
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] += dpp_q<0xB1>(v[i]);   // quad_perm [1,0,3,2]
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] += dpp_q<0x4E>(v[i]);   // quad_perm [2,3,0,1]
}
// N exclusive prefix sums over the quad: the values shifted one lane up (quad_perm [0,0,1,2],
// lane 0 of the quad takes 0), then the inclusive sum in two steps (lanes 1.. add lane k-1, lanes
// 2.. add lane k-2, through quad_perm [0,0,0,1]). The lanes that take nothing are masked with bit
// operations, never with a select: the compiler may turn a select into a branch and move the DPP
// read into it, and a DPP read from a lane outside exec returns the old value, not the lane's.
template <int N>
__device__ inline void quad_scan_excl_n(float (&v)[N], int k) {
    const int m1 = -(int)(k >= 1), m2 = -(int)(k >= 2);
    auto keep = [](float x, int m) { return __int_as_float(__float_as_int(x) & m); };
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = keep(dpp_q<0x90>(v[i]), m1);
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] += keep(dpp_q<0x90>(v[i]), m1);
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] += keep(dpp_q<0x40>(v[i]), m2);
}

template <int F, bool P64, bool FLAGS>
__device__ __forceinline__ void fwd_quads_tile(
    Stage<64, F, false, P64>& S, const CamArgs& cam, int tile, const int* __restrict__ sorted, int64_t n_sorted,
    const int2* __restrict__ ranges, int64_t tile_base, int64_t tet_base, const int4* __restrict__ indices,
    const float* __restrict__ verts, const float* __restrict__ feats, const float2* __restrict__ jitter,
    float4* __restrict__ out_img, float* __restrict__ out_dist, int* __restrict__ out_nc, int64_t pix_base,
    uint8_t* __restrict__ flags) {
    constexpr int NGRD = NG_<F>::NGRD;
    using T = typename Prec<P64>::R;
    using PL = typename Prec<P64>::PL;
    const int lane = threadIdx.x, k = lane & 3, p = lane >> 2;
    const int tx = tile % cam.gw, ty = tile / cam.gw;
    const int2 rg = load_range(ranges, true, (int)(tile_base + tile), n_sorted);
    const int n = rg.y - rg.x;
    const int px = tx * 4 + (p & 3), py = ty * 4 + (p >> 2);
    const bool inside = px < cam.W && py < cam.H;
    const int64_t pix = pix_base + (int64_t)py * cam.W + px;
    RayT<T> ray = {0, 0, 1};
    if (inside) {
        const float2 j = jitter ? jitter[pix - pix_base] : make_float2(0.5f, 0.5f);
        ray = make_ray<T>(cam, px, py, j.x, j.y);
    }
    const float rdx = (float)ray.dx, rdy = (float)ray.dy, rdz = (float)ray.dz;
    const float mt = cam.min_t;
    float cr = 0.f, cg = 0.f, cb = 0.f, X = 0.f, Y = 0.f, Z = 0.f, W = 0.f, V = 0.f;
    double L = 0.0;
    int nc = 0;
    bool alive = inside;
    const int nchunks = (n + 63) / 64;
    auto entry = [&](int cc) {
        const int kk = cc * 64 + lane;
        return kk < n ? (int)(sorted[rg.x + kk] - tet_base) : -1;
    };
    ChunkPipe<F> pipe;
    if (nchunks > 0) pipe.start(entry, indices, verts, feats);
    for (int c = 0; c < nchunks; ++c) {
        if (!__syncthreads_or(alive)) break;
        pipe.stage(S, lane, cam);
        __syncthreads();
        pipe.advance(entry, c, indices, verts, feats);
        const int cnt = n - c * 64 < 64 ? n - c * 64 : 64;
        uint32_t act_bits = 0;   // bit g: this lane's entry of group g was blended
        for (int g = 0; 4 * g < cnt; ++g) {
            if (__ballot(alive) == 0) break;
            if (alive) {
                const int e = 4 * g + k;
                const bool valid = e < cnt;
                const PL plp[4] = {S.pl[0][e], S.pl[1][e], S.pl[2][e], S.pl[3][e]};
                const auto h = intersect(plp, ray, mt);
                const float4 fa = S.fa[e];
                const float sig = fa.x;
                const float x = sig * h.dt;
                const bool bl = valid && h.hit && x > 0.f;
                const float te = h.se - mt, tx_ = h.sx - mt;
                const float m1 = sig * (te + tx_) * 0.5f;
                const float xa = bl ? x : 0.f;
                float sc[3] = {xa, bl ? sig : 0.f, bl ? m1 : 0.f};
                quad_scan_excl_n<3>(sc, k);
                const float ex = sc[0], es = sc[1], em = sc[2];
                const uint64_t stop = __ballot(bl && L - (double)(ex + xa) < -5.5);
                const uint32_t qb = (uint32_t)(stop >> (lane & ~3)) & 0xFu;
                const int first = qb ? (int)__builtin_ctz(qb) : 4;
                const bool act = bl && k <= first;
                act_bits |= (uint32_t)act << g;
                float rd[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
                if (act) {
                    float gr[NGRD];
                    read_grads(S, e, gr);
                    const float em1 = expm1_neg(x);
                    const float alpha = 1.0f + em1;
                    float w0, w1;
                    seg_weights(x, em1, alpha, w0, w1);
                    const float Tr = __expf((float)(L - (double)ex));
                    const float gd0 = gr[0] * rdx + gr[1] * rdy + gr[2] * rdz;
                    float gd1 = gd0, gd2 = gd0;
                    if (F == 13) {
                        gd1 = gr[3 % NGRD] * rdx + gr[4 % NGRD] * rdy + gr[5 % NGRD] * rdz;
                        gd2 = gr[6 % NGRD] * rdx + gr[7 % NGRD] * rdy + gr[8 % NGRD] * rdz;
                    }
                    const float wsum = w0 + w1, wsl = w0 * h.sx + w1 * h.se;
                    rd[0] = Tr * (wsum * fa.y + wsl * gd0);
                    rd[1] = Tr * (wsum * fa.z + wsl * gd1);
                    rd[2] = Tr * (wsum * fa.w + wsl * gd2);
                    rd[3] = 2.0f * m1 * (Z + es);
                    rd[4] = 2.0f * sig * (W + em);
                    rd[5] = sig * sig * h.dt * (1.0f / 3.0f);
                    rd[6] = x;
                    rd[7] = sig;
                    rd[8] = m1;
                }
                quad_sum_n<9>(rd);
                cr += rd[0]; cg += rd[1]; cb += rd[2];
                X += rd[3]; Y += rd[4]; V += rd[5];
                L -= (double)rd[6];
                Z += rd[7]; W += rd[8];
                const int left = cnt - 4 * g;
                nc += first < 4 ? first + 1 : (left < 4 ? left : 4);
                if (first < 4) alive = false;
            }
        }
        if constexpr (FLAGS) {
            // entry e of the chunk: 1 when some pixel blended it; lane e holds slot e % 4's bits
            uint32_t any = act_bits;
            any |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)any, 0x124, 0xF, 0xF, false);   // row_ror:4
            any |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)any, 0x128, 0xF, 0xF, false);   // row_ror:8
            any |= (uint32_t)__shfl_xor((int)any, 16, 64);
            any |= (uint32_t)__shfl_xor((int)any, 32, 64);
            if (lane < cnt) flags[rg.x + c * 64 + lane] = (uint8_t)((any >> (lane >> 2)) & 1u);
        }
    }
    if (inside && k == 0) {
        out_img[pix] = make_float4(cr, cg, cb, (float)L);
        float* d = out_dist + 5 * pix;
        d[0] = X; d[1] = Y; d[2] = Z; d[3] = W; d[4] = V;
        out_nc[pix] = nc;
    }
}

template <int F, bool P64, bool FLAGS>
__global__ __launch_bounds__(64) void composite_fwd_quads_kernel(
    CamArgs cam, const int* __restrict__ sorted, int64_t n_sorted, const int2* __restrict__ ranges,
    const int4* __restrict__ indices, const float* __restrict__ verts, const float* __restrict__ feats,
    const float2* __restrict__ jitter, float4* __restrict__ out_img, float* __restrict__ out_dist,
    int* __restrict__ out_nc, uint8_t* __restrict__ flags, const int* __restrict__ tile_order) {
    __shared__ Stage<64, F, false, P64> S;
    const int tile = tile_order ? tile_order[blockIdx.x] : remap_block(blockIdx.x, gridDim.x);
    fwd_quads_tile<F, P64, FLAGS>(S, cam, tile, sorted, n_sorted, ranges, 0, 0, indices, verts, feats, jitter, out_img,
                                  out_dist, out_nc, 0, flags);
}

// Batched 4x4-tile forward: one workgroup per tile of every view of a plan (pixel-centre rays).
template <int F, bool P64>
__global__ __launch_bounds__(64) void batch_composite_fwd_quads_kernel(
    const void* __restrict__ plan, const int* __restrict__ sorted, int64_t n_sorted, const int2* __restrict__ ranges,
    const int* __restrict__ tile_order, float4* __restrict__ out_img, float* __restrict__ out_dist,
    int* __restrict__ out_nc) {
    __shared__ Stage<64, F, false, P64> S;
    const PlanHeader* h = static_cast<const PlanHeader*>(plan);
    const ViewRec* views = plan_views(plan);
    // one workgroup per tile at 4x4 tiles, so a view's first workgroup is its first tile
    const int bg = tile_order ? tile_order[blockIdx.x] : remap_block(blockIdx.x, gridDim.x);
    if (bg >= h->total_blocks) return;  // launch sized for a capacity
    int lo = 0, hi = h->n_views - 1;  // last view with block_offset <= bg
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (views[mid].block_offset <= bg) lo = mid;
        else hi = mid - 1;
    }
    const ViewRec& v = views[lo];
    fwd_quads_tile<F, P64, false>(S, v.cam, (int)(bg - v.block_offset), sorted, n_sorted, ranges, v.tile_offset,
                                  v.tet_offset, uniform_global_ptr(reinterpret_cast<const int4*>(v.indices)),
                                  uniform_global_ptr(v.verts), uniform_global_ptr(v.feats), nullptr, out_img, out_dist,
                                  out_nc, v.pixel_offset, nullptr);
}

template <int TS, int F, bool P64>
int launch_fwd(const CamArgs& a, const int* sorted, int64_t n_sorted, const int* ranges, const int* indices, const float* verts,
               const float* feats, const float* jitter, float* img, float* dist, int* nc, hipStream_t st,
               uint8_t* flags = nullptr, const int* tile_order = nullptr) {
    using C = Cfg<TS>;
    if constexpr (TS == 4) {
        if (flags)
            hipLaunchKernelGGL((composite_fwd_quads_kernel<F, P64, true>), dim3(a.gw * a.gh), dim3(64), 0, st, a,
                               sorted, n_sorted, (const int2*)ranges, (const int4*)indices, verts, feats,
                               (const float2*)jitter, (float4*)img, dist, nc, flags, tile_order);
        else
            hipLaunchKernelGGL((composite_fwd_quads_kernel<F, P64, false>), dim3(a.gw * a.gh), dim3(64), 0, st, a,
                               sorted, n_sorted, (const int2*)ranges, (const int4*)indices, verts, feats,
                               (const float2*)jitter, (float4*)img, dist, nc, flags, tile_order);
    } else {
        int sgw = (a.gw + C::GX - 1) / C::GX, sgh = (a.gh + C::GY - 1) / C::GY;
        hipLaunchKernelGGL((composite_fwd_kernel<TS, F, P64>), dim3(sgw * sgh), dim3(C::WG), 0, st, a, sorted,
                           n_sorted, (const int2*)ranges, (const int4*)indices, verts, feats, (const float2*)jitter,
                           (float4*)img, dist, nc, flags);
    }
    return 0;
}

// Zeroes n4 float4s (a kernel rather than a memset, so that the sequence may be graph-captured).
__global__ __launch_bounds__(256) void zero_f4_kernel(float4* __restrict__ p, int64_t n4) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n4) p[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// Vertex and feature gradients of each tet from its row of plane-record adjoint sums (the
// layout of acc in composite_bwd_kernel), one thread per tet of a 256-tet workgroup; tets whose
// row is all zero add nothing. Feature gradients are added in place (one writer per row). Vertex
// gradients are first summed per workgroup in an LDS table keyed by vertex id (FP64 sums): the
// tets of a Morton-ordered mesh that share a workgroup share most of their vertices (about 120
// distinct vertices per 1024 corners), so one float atomic per distinct (vertex, axis) follows.
constexpr int kTetGradTets = 256;
constexpr int kTetGradLog2Slots = 10;
constexpr int kTetGradSlots = 1 << kTetGradLog2Slots;   // = 4 corners x 256 tets: an insert always finds room
constexpr uint32_t kNoVertex = 0xFFFFFFFFu;

template <int F, bool P64>
__global__ __launch_bounds__(kTetGradTets) void composite_tet_grads_kernel(
    const float4* __restrict__ rows, int64_t n_tets, const int4* __restrict__ indices, const float* __restrict__ verts,
    const float* __restrict__ feats, float3 cp, float* __restrict__ gverts, float* __restrict__ gfeats) {
    using T = typename Prec<P64>::R;
    using PL = typename Prec<P64>::PL;
    constexpr int NGRD = NG_<F>::NGRD;
    constexpr int NCH = F == 7 ? 1 : 3;
    constexpr int NV = 1 + 3 + NCH + 3 * NCH + 4 + 12;
    constexpr int NOUT = 12 + F;
    static_assert(NV % 4 == 0 && NV <= kRowFloats, "row layout");
    static_assert(kTetGradSlots >= 4 * kTetGradTets, "vertex table too small");
    __shared__ uint32_t s_vid[kTetGradSlots];
    __shared__ double s_g[3][kTetGradSlots];
    for (int p = threadIdx.x; p < kTetGradSlots; p += kTetGradTets) {
        s_vid[p] = kNoVertex;
        s_g[0][p] = 0.0; s_g[1][p] = 0.0; s_g[2][p] = 0.0;
    }
    __syncthreads();
    const int64_t t = (int64_t)blockIdx.x * kTetGradTets + threadIdx.x;
    bool any = false;
    float acc[NV];
    if (t < n_tets) {
        float row[kRowFloats];
#pragma unroll
        for (int k = 0; k < kRowFloats / 4; ++k) {
            const float4 v = rows[t * (kRowFloats / 4) + k];
            row[4 * k] = v.x; row[4 * k + 1] = v.y; row[4 * k + 2] = v.z; row[4 * k + 3] = v.w;
            any = any || v.x != 0.f || v.y != 0.f || v.z != 0.f || v.w != 0.f;
        }
#pragma unroll
        for (int pos = 0; pos < kRowFloats; ++pos)
            if (row_sum_index(pos) < NV) acc[row_sum_index(pos)] = row[pos];
    }
    if (any) {
        const int4 q = indices[t];
        const float3 V4[4] = {load_v(verts, q.x), load_v(verts, q.y), load_v(verts, q.z), load_v(verts, q.w)};
        PL unused;
        int signs = 0;
        signs |= plane_rec<T>(V4[0], V4[1], V4[2], V4[3], cp, unused) << 0;
        signs |= plane_rec<T>(V4[0], V4[1], V4[3], V4[2], cp, unused) << 1;
        signs |= plane_rec<T>(V4[0], V4[2], V4[3], V4[1], cp, unused) << 2;
        signs |= plane_rec<T>(V4[1], V4[2], V4[3], V4[0], cp, unused) << 3;
        float* gf = gfeats + (int64_t)F * t;
        const float* f = feats + (int64_t)F * t;
        float gr[NGRD];
#pragma unroll
        for (int k = 0; k < NGRD; ++k) gr[k] = f[4 + k];
        float outv[NOUT];
        tet_grads<F, P64, NV, NOUT>(V4, signs, gr, acc, cp, outv);
#pragma unroll
        for (int k = 0; k < F; ++k) gf[k] += outv[12 + k];
        const uint32_t vid[4] = {(uint32_t)q.x, (uint32_t)q.y, (uint32_t)q.z, (uint32_t)q.w};
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            uint32_t p = (vid[v] * 2654435761u) >> (32 - kTetGradLog2Slots);
            for (int probe = 0; probe < kTetGradSlots; ++probe) {
                const uint32_t k = atomicCAS(&s_vid[p], kNoVertex, vid[v]);
                if (k == kNoVertex || k == vid[v]) {
#pragma unroll
                    for (int a = 0; a < 3; ++a) atomicAdd(&s_g[a][p], (double)outv[3 * v + a]);
                    break;
                }
                p = (p + 1) & (kTetGradSlots - 1);
            }
        }
    }
    __syncthreads();
    for (int p = threadIdx.x; p < kTetGradSlots; p += kTetGradTets) {
        const uint32_t k = s_vid[p];
        if (k == kNoVertex) continue;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const float v = (float)s_g[a][p];
            if (v != 0.f) unsafeAtomicAdd(gverts + 3 * (int64_t)k + a, v);
        }
    }
}

template <int TS, int F, bool P64, bool ROWS>
int launch_bwd(const CamArgs& a, const int* sorted, int64_t n_sorted, const int* ranges, const int* indices, const float* verts,
               const float* feats, const float* jitter, const float* img, const float* dist, const int* nc,
               const float* gimg, const float* gdist, float* gverts, float* gfeats, float* rows, int64_t n_tets,
               hipStream_t st, const uint8_t* flags = nullptr, const int* tile_order = nullptr) {
    using C = Cfg<TS>;
    int sgw = (a.gw + C::GX - 1) / C::GX, sgh = (a.gh + C::GY - 1) / C::GY;
    if (ROWS && n_tets > 0) {
        const int64_t n4 = n_tets * (kRowFloats / 4);
        hipLaunchKernelGGL(zero_f4_kernel, dim3((unsigned)ceil_div(n4, 256)), dim3(256), 0, st, (float4*)rows, n4);
    }
    if constexpr (ROWS && TS == 4) {
        if (flags)
            hipLaunchKernelGGL((composite_bwd_tet_lanes_kernel<F, P64, true>), dim3(a.gw * a.gh), dim3(64), 0, st, a,
                               sorted, n_sorted, (const int2*)ranges, (const int4*)indices, verts, feats,
                               (const float2*)jitter, (const float4*)img, dist, nc, (const float4*)gimg, gdist, rows,
                               flags, tile_order);
        else
            hipLaunchKernelGGL((composite_bwd_tet_lanes_kernel<F, P64, false>), dim3(a.gw * a.gh), dim3(64), 0, st, a,
                               sorted, n_sorted, (const int2*)ranges, (const int4*)indices, verts, feats,
                               (const float2*)jitter, (const float4*)img, dist, nc, (const float4*)gimg, gdist, rows,
                               (const uint8_t*)nullptr, tile_order);
    } else if constexpr (ROWS && C::GROUP <= 64) {
        if (flags) {
            hipLaunchKernelGGL((composite_bwd_kernel<TS, F, P64, ROWS, true>), dim3(sgw * sgh), dim3(C::WG), 0, st,
                               a, sorted, n_sorted, (const int2*)ranges, (const int4*)indices, verts, feats,
                               (const float2*)jitter, (const float4*)img, dist, nc, (const float4*)gimg, gdist,
                               gverts, gfeats, rows, flags);
        } else {
            hipLaunchKernelGGL((composite_bwd_kernel<TS, F, P64, ROWS>), dim3(sgw * sgh), dim3(C::WG), 0, st, a,
                               sorted, n_sorted, (const int2*)ranges, (const int4*)indices, verts, feats,
                               (const float2*)jitter, (const float4*)img, dist, nc, (const float4*)gimg, gdist,
                               gverts, gfeats, rows, (const uint8_t*)nullptr);
        }
    } else {
        hipLaunchKernelGGL((composite_bwd_kernel<TS, F, P64, ROWS>), dim3(sgw * sgh), dim3(C::WG), 0, st, a, sorted,
                           n_sorted, (const int2*)ranges, (const int4*)indices, verts, feats, (const float2*)jitter,
                           (const float4*)img, dist, nc, (const float4*)gimg, gdist, gverts, gfeats, rows,
                           (const uint8_t*)nullptr);
    }
    if (ROWS && n_tets > 0)
        hipLaunchKernelGGL((composite_tet_grads_kernel<F, P64>), dim3((unsigned)ceil_div(n_tets, kTetGradTets)),
                           dim3(kTetGradTets), 0,
                           st, (const float4*)rows, n_tets, (const int4*)indices, verts, feats,
                           make_float3(a.cam_pos[0], a.cam_pos[1], a.cam_pos[2]), gverts, gfeats);
    return 0;
}

// ------------------------------------------------------------------------- error votes --

// Integer sum / max over the reduction group (the DPP row pattern of group_sum).
template <bool MAX>
__device__ inline int group_red_i(int x, int width) {
    auto op = [](int a, int b) { return MAX ? (a > b ? a : b) : a + b; };
    x = op(x, __builtin_amdgcn_update_dpp(0, x, 0xB1, 0xF, 0xF, false));
    x = op(x, __builtin_amdgcn_update_dpp(0, x, 0x4E, 0xF, 0xF, false));
    x = op(x, __builtin_amdgcn_update_dpp(0, x, 0x124, 0xF, 0xF, false));
    x = op(x, __builtin_amdgcn_update_dpp(0, x, 0x128, 0xF, 0xF, false));
    if (width > 16) {
        x = op(x, __shfl_xor(x, 16, 64));
        x = op(x, __shfl_xor(x, 32, 64));
    }
    return x;
}

// Per-tet error votes of one view: kern_calc_tet_err / calc_tet_err
// (alphablend_shader_interp.slang:433-599), launched by render_err (render_err.py:138-150).
// Each pixel ray (pixel centre) walks its tile's list front to back starting from
// log-transmittance 1 -- that kernel's initial state, unlike splat_tiled's 0 -- and for every
// blended tet (hit with sigma*dt > 0) votes w = exp(logT) * sigma * dt:
//   tet_err[t] += (w, w*e, w*e^2, 0, 0, w*s, w*p_in, w*p_out, w, w*(p_in + p_out)/2)
//   tet_count[t] = (count + 1, max(count_max, int(65535 w)))
// with e / s the pixel's L1 / SSIM error and p_in / p_out the segment's end points; the walk
// stops after a blend that takes log-transmittance below -5.5. Every lane of a reduction group
// (one 4x4 tile per 16-lane DPP row, or a 64-lane wave of a larger tile) visits the same slot at
// the same step -- terminated lanes vote zero -- so the group's votes are summed with DPP row
// ops and lane k of the group adds channel k: one 64-byte atomic segment per (tile, tet) and
// group instead of 15 atomics per (pixel, tet).
template <int TS, int F, bool P64>
__global__ __launch_bounds__(Cfg<TS>::WG) void tet_err_kernel(
    CamArgs cam, const int* __restrict__ sorted, int64_t n_sorted, const int2* __restrict__ ranges,
    const int4* __restrict__ indices, const float* __restrict__ verts, const float* __restrict__ feats,
    const float* __restrict__ pixel_err, const float* __restrict__ ssim_err, float* __restrict__ tet_err,
    int* __restrict__ tet_count) {
    using C = Cfg<TS>;
    using T = typename Prec<P64>::R;
    using PL = typename Prec<P64>::PL;
    __shared__ Stage<C::WG, F, true, P64, false> S;
    __shared__ int s_chunks;
    int g, r, tx, ty, px, py;
    bool tile_ok, inside;
    locate<TS>(cam, remap_block(blockIdx.x, gridDim.x), g, r, tx, ty, px, py, tile_ok, inside);
    const int2 rg = load_range(ranges, tile_ok, ty * cam.gw + tx, n_sorted);
    const int n = rg.y - rg.x;
    if (threadIdx.x == 0) s_chunks = 0;
    __syncthreads();
    if (r == 0) atomicMax(&s_chunks, (n + C::GROUP - 1) / C::GROUP);
    const int64_t pix = (int64_t)py * cam.W + px;
    RayT<T> ray = {0, 0, 1};
    float perr = 0.f, serr = 0.f;
    if (inside) {
        ray = make_ray<T>(cam, px, py, 0.5f, 0.5f);
        perr = pixel_err[pix];
        serr = ssim_err[pix];
    }
    const float rdx = (float)ray.dx, rdy = (float)ray.dy, rdz = (float)ray.dz;
    const float cx = cam.cam_pos[0], cy = cam.cam_pos[1], cz = cam.cam_pos[2];
    const int lane = threadIdx.x & 63;
    const int gl = lane & (C::RED - 1);          // lane within the reduction group
    const int gshift = C::RED == 16 ? (lane & 48) : 0;
    float logT = 1.0f;
    bool alive = inside;
    __syncthreads();
    const int nchunks = s_chunks;
    for (int c = 0; c < nchunks; ++c) {
        if (!__syncthreads_or(alive)) break;
        const int k = c * C::GROUP + r;
        if (k < n) stage_tet(S, threadIdx.x, sorted[rg.x + k], cam, indices, verts, feats);
        __syncthreads();
        int cnt = n - c * C::GROUP;
        cnt = cnt < C::GROUP ? cnt : C::GROUP;
        const int base = g * C::GROUP;
        // the trip count is uniform over each reduction group (one tile), so every lane of a
        // group reaches the ballot and the DPP sums of each step
        for (int j = 0; j < cnt; ++j) {
            const int s = base + j;
            float w = 0.f, se = 0.f, sx = 0.f;
            int hit = 0;
            if (alive) {
                const PL p[4] = {S.pl[0][s], S.pl[1][s], S.pl[2][s], S.pl[3][s]};
                const auto h = intersect(p, ray, cam.min_t);
                const float x = S.fa[s].x * h.dt;
                if (h.hit && x > 0.f) {
                    w = __expf(logT) * x;   // hardware exponential: ~4e-7 relative for logT in [-5.5, 1]
                    se = h.se;
                    sx = h.sx;
                    hit = 1;
                    logT -= x;
                    alive = !(logT < -5.5f);
                }
            }
            const uint64_t votes = (__ballot(hit) >> gshift) & (C::RED == 16 ? 0xFFFFull : ~0ull);
            if (votes == 0) continue;
            float v[10] = {w, w * perr, w * perr * perr, w * serr,
                           w * (cx + se * rdx), w * (cy + se * rdy), w * (cz + se * rdz),
                           w * (cx + sx * rdx), w * (cy + sx * rdy), w * (cz + sx * rdz)};
            group_sum(v, C::RED);
            const int n_hit = __popcll(votes);
            const int wmax = group_red_i<true>(hit ? (int)(65535.f * w) : 0, C::RED);
            if (gl < 16) {
                const int tet = S.meta[0][s];
                float val;
                switch (gl) {
                    case 0: case 12: val = v[0]; break;
                    case 1: val = v[1]; break;
                    case 2: val = v[2]; break;
                    case 3: val = (float)n_hit; break;   // the vote count, moved out by the finish kernel
                    case 5: val = v[3]; break;
                    case 6: val = v[4]; break;
                    case 7: val = v[5]; break;
                    case 8: val = v[6]; break;
                    case 9: val = v[7]; break;
                    case 10: val = v[8]; break;
                    case 11: val = v[9]; break;
                    default: val = 0.f; break;   // channel 4 carries no vote
                }
                // channels 12-15 (w again, w * segment midpoint) follow from 0 and 6-11: written
                // once per tet by tet_err_finish_kernel instead of four more atomics per vote row.
                // The count rides in channel 3 of the same row (one atomic transaction for the
                // row's lanes instead of a second one for the count); the finish kernel moves it.
                if (gl != 4 && gl < 12) unsafeAtomicAdd(tet_err + 16 * (int64_t)tet + gl, val);
                if (gl == 1) atomicMax(tet_count + 2 * (int64_t)tet + 1, wmax);
            }
        }
    }
}

// Channels 12-15 of the vote rows from the accumulated 0 and 6-11: 12 = sum w, 13-15 = sum of
// w * (entry + exit) / 2, the reference's per-vote midpoint terms summed. Channel 3 held this
// call's vote count (integer sums, exact below 2^24): it is added to count[t][0] and cleared.
__global__ __launch_bounds__(256) void tet_err_finish_kernel(float4* __restrict__ err, int* __restrict__ count,
                                                             int64_t n_tets) {
    const int64_t t = blockIdx.x * 256ll + threadIdx.x;
    if (t >= n_tets) return;
    const float4 r0 = err[4 * t], r1 = err[4 * t + 1], r2 = err[4 * t + 2];
    if (r0.w != 0.f) {
        count[2 * t] += (int)r0.w;
        err[4 * t] = make_float4(r0.x, r0.y, r0.z, 0.f);
    }
    // r1 = channels 4-7, r2 = 8-11: entry (6, 7, 8) and exit (9, 10, 11)
    err[4 * t + 3] = make_float4(r0.x, 0.5f * (r1.z + r2.y), 0.5f * (r1.w + r2.z), 0.5f * (r2.x + r2.w));
}

// The 4x4-tile votes with one lane per tet: one wave per tile walks the list front to back in
// chunks of 64 entries, lane j holding entry j of the chunk (its plane records and sigma in
// registers) and summing its tet's votes over the tile's 16 pixels. A pixel's log T in front of
// lane j is its log T at the chunk start minus the exclusive prefix sum of the blended x over lanes
// 0..j-1; the walk's stop (after the blend that takes log T below -5.5) makes every lane from the
// first one whose log T in front is below -5.5 vote nothing, the pixel ending there.
template <int F, bool P64>
__global__ __launch_bounds__(64) void tet_err_tet_lanes_kernel(
    CamArgs cam, const int* __restrict__ sorted, int64_t n_sorted, const int2* __restrict__ ranges,
    const int4* __restrict__ indices, const float* __restrict__ verts, const float* __restrict__ feats,
    const float* __restrict__ pixel_err, const float* __restrict__ ssim_err, float* __restrict__ tet_err,
    int* __restrict__ tet_count, const int* __restrict__ tile_order) {
    using T = typename Prec<P64>::R;
    using PL = typename Prec<P64>::PL;
    constexpr int NVOTE = 12;                   // row channels 0-11 (3: the vote count, 4: none)
    __shared__ float s_row[64][NVOTE + 1];
    __shared__ int s_tet[64];
    const int lane = threadIdx.x;
    const int tile = tile_order ? tile_order[blockIdx.x] : remap_block(blockIdx.x, gridDim.x);
    const int tx = tile % cam.gw, ty = tile / cam.gw;
    const int2 rg = load_range(ranges, true, tile, n_sorted);
    const int n = rg.y - rg.x;
    // pixel lane p < 16: its ray, errors and walk state
    const int px = tx * 4 + (lane & 3), py = ty * 4 + ((lane >> 2) & 3);
    const bool inside = lane < 16 && px < cam.W && py < cam.H;
    RayT<T> pray = {0, 0, 1};
    float perr = 0.f, serr = 0.f, plogT = 1.0f;
    if (inside) {
        const int64_t pix = (int64_t)py * cam.W + px;
        pray = make_ray<T>(cam, px, py, 0.5f, 0.5f);
        perr = pixel_err[pix];
        serr = ssim_err[pix];
    }
    uint32_t alive = (uint32_t)__ballot(inside) & 0xFFFFu;   // pixels still walking
    const float mt = cam.min_t;
    const float3 cp = f3(cam.cam_pos[0], cam.cam_pos[1], cam.cam_pos[2]);
    for (int c0 = 0; c0 < n && alive != 0; c0 += 64) {
        const int k = c0 + lane;
        const bool valid = k < n;
        PL pl[4];
        float sig = 0.f;
        int tet = -1;
        if (valid) {
            tet = sorted[rg.x + k];
            const int4 q = indices[tet];
            const float3 v0 = load_v(verts, q.x), v1 = load_v(verts, q.y), v2 = load_v(verts, q.z),
                         v3 = load_v(verts, q.w);
            plane_rec<T>(v0, v1, v2, v3, cp, pl[0]);
            plane_rec<T>(v0, v1, v3, v2, cp, pl[1]);
            plane_rec<T>(v0, v2, v3, v1, cp, pl[2]);
            plane_rec<T>(v1, v2, v3, v0, cp, pl[3]);
            sig = fmaxf(feats[(int64_t)F * tet], 0.f);
        } else {
#pragma unroll
            for (int i = 0; i < 4; ++i) {   // planes no ray crosses ahead of the camera: a miss
                pl[i].x = 0; pl[i].y = 0; pl[i].z = 0; pl[i].w = 1;
            }
        }
        float v[10];
#pragma unroll
        for (int i = 0; i < 10; ++i) v[i] = 0.f;
        int votes = 0, wmax = 0;
        for (int p = 0; p < 16; ++p) {
            if (!((alive >> p) & 1u)) continue;
            const RayT<T> ray = {lane_f(pray.dx, p), lane_f(pray.dy, p), lane_f(pray.dz, p)};
            const auto h = intersect(pl, ray, mt);
            const float x = sig * h.dt;
            const bool hit = valid && h.hit && x > 0.f;
            const float xa = hit ? x : 0.f;
            const float ex = wave_scan_excl(xa);
            const float lt = lane_f(plogT, p) - ex;           // log T in front of this entry
            const uint64_t stop = __ballot(valid && lt < -5.5f);
            const int first = stop ? __builtin_ctzll(stop) : 64;
            if (hit && lane < first) {
                const float w = __expf(lt) * x;   // hardware exponential: ~4e-7 relative for log T in [-5.5, 1]
                const float e = lane_f(perr, p), s2 = lane_f(serr, p);
                const float rdx = (float)ray.dx, rdy = (float)ray.dy, rdz = (float)ray.dz;
                v[0] += w;
                v[1] += w * e;
                v[2] += w * e * e;
                v[3] += w * s2;
                v[4] += w * (cp.x + h.se * rdx);
                v[5] += w * (cp.y + h.se * rdy);
                v[6] += w * (cp.z + h.se * rdz);
                v[7] += w * (cp.x + h.sx * rdx);
                v[8] += w * (cp.y + h.sx * rdy);
                v[9] += w * (cp.z + h.sx * rdz);
                ++votes;
                wmax = max(wmax, (int)(65535.f * w));
            }
            if (first < 64) {
                alive &= ~(1u << p);                          // the pixel stopped inside this chunk
            } else {
                const float nl = lane_f(plogT, p) - lane_f(ex + xa, 63);
                if (lane == p) plogT = nl;
                if (nl < -5.5f) alive &= ~(1u << p);
            }
        }
        // rows leave through LDS: each atomic instruction adds four 64-byte rows, 16 lanes each
        const uint64_t vm = __ballot(votes > 0);
        if (vm != 0) {
            s_row[lane][0] = v[0]; s_row[lane][1] = v[1]; s_row[lane][2] = v[2];
            s_row[lane][3] = (float)votes;    // the vote count, moved out by the finish kernel
            s_row[lane][4] = 0.f;             // channel 4 carries no vote
            s_row[lane][5] = v[3];
#pragma unroll
            for (int i = 0; i < 6; ++i) s_row[lane][6 + i] = v[4 + i];
            s_tet[lane] = tet;
            if (votes > 0) atomicMax(tet_count + 2 * (int64_t)tet + 1, wmax);
            __syncthreads();
            const int ch = lane & 15;
            for (int s0 = 0; s0 < 64; s0 += 4) {
                if (((vm >> s0) & 0xFull) == 0) continue;
                const int s = s0 + (lane >> 4);
                if (((vm >> s) & 1ull) && ch < NVOTE && ch != 4) {
                    const float val = s_row[s][ch];
                    if (val != 0.f) unsafeAtomicAdd(tet_err + 16 * (int64_t)s_tet[s] + ch, val);
                }
            }
            __syncthreads();
        }
    }
}

template <int TS, int F, bool P64>
int launch_tet_err(const CamArgs& a, const int* sorted, int64_t n_sorted, const int* ranges, const int* indices,
                   const float* verts, const float* feats, const float* perr, const float* serr, float* err,
                   int* count, hipStream_t st, const int* tile_order) {
    using C = Cfg<TS>;
    if constexpr (TS == 4) {
        hipLaunchKernelGGL((tet_err_tet_lanes_kernel<F, P64>), dim3(a.gw * a.gh), dim3(64), 0, st, a, sorted,
                           n_sorted, (const int2*)ranges, (const int4*)indices, verts, feats, perr, serr, err, count,
                           tile_order);
        return 0;
    }
    int sgw = (a.gw + C::GX - 1) / C::GX, sgh = (a.gh + C::GY - 1) / C::GY;
    hipLaunchKernelGGL((tet_err_kernel<TS, F, P64>), dim3(sgw * sgh), dim3(C::WG), 0, st, a, sorted, n_sorted,
                       (const int2*)ranges, (const int4*)indices, verts, feats, perr, serr, err, count);
    return 0;
}

int check_common(const rm_camera* cam, int32_t nf, const void* sorted, const void* ranges, const void* indices,
                 const void* verts, const void* feats, const void* jitter, const char* who) {
    std::string w(who);
    if (!cam) { set_error(w + ": camera is null"); return RM_ERR_INVALID_ARGUMENT; }
    if (nf != 7 && nf != 13) { set_error(w + ": n_features must be 7 or 13"); return RM_ERR_INVALID_ARGUMENT; }
    if (cam->tile_width != cam->tile_height ||
        (cam->tile_width != 4 && cam->tile_width != 8 && cam->tile_width != 16)) {
        set_error(w + ": tile size must be square 4, 8 or 16");
        return RM_ERR_UNSUPPORTED_TILE;
    }
    if (cam->image_height <= 0 || cam->image_width <= 0 ||
        cam->grid_height != (cam->image_height + cam->tile_height - 1) / cam->tile_height ||
        cam->grid_width != (cam->image_width + cam->tile_width - 1) / cam->tile_width) {
        set_error(w + ": grid does not match image and tile size");
        return RM_ERR_INVALID_ARGUMENT;
    }
    if (!sorted || !ranges || !indices || !verts || !feats || !jitter) {
        set_error(w + ": null buffer");
        return RM_ERR_INVALID_ARGUMENT;
    }
    if (((uintptr_t)indices & 15) || ((uintptr_t)ranges & 7) || ((uintptr_t)jitter & 7)) {
        set_error(w + ": misaligned indices/ranges/jitter");
        return RM_ERR_INVALID_ARGUMENT;
    }
    return RM_OK;
}

// Ray/plane intersection precision: FP64 plane records (default) or FP32.
int g_fp64_planes = 1;

template <int TS, int F, bool P64>
void launch_batch(const void* plan, const PlanHeader& h, const int32_t* sorted, int64_t cap, const int32_t* ranges,
                  const int32_t* order, float* img, float* dist, int32_t* nc, hipStream_t st) {
    if constexpr (TS == 4)
        hipLaunchKernelGGL((batch_composite_fwd_quads_kernel<F, P64>), dim3((unsigned)h.total_blocks), dim3(64),
                           0, st, plan, sorted, cap, (const int2*)ranges, order, (float4*)img, dist, nc);
    else
        hipLaunchKernelGGL((batch_composite_fwd_kernel<TS, F, P64>), dim3((unsigned)h.total_blocks),
                           dim3(Cfg<TS>::WG), 0, st, plan, sorted, cap, (const int2*)ranges, (float4*)img, dist, nc);
}

}  // namespace

// one workgroup per tile at every tile size (at 4x4 tiles one wave, lanes on list entries)
int batch_workgroups_per_view(int, int gh, int gw) { return gw * gh; }

hipError_t batch_composite(const void* plan_dev, const PlanHeader& h, const int32_t* sorted_tet, int64_t pair_cap,
                           const int32_t* ranges, const int32_t* tile_order, float* img, float* dist, int32_t* nc,
                           hipStream_t st) {
    if (h.total_blocks <= 0) return hipSuccess;
#define RM_BATCH(TS, F) (g_fp64_planes ? launch_batch<TS, F, true>(plan_dev, h, sorted_tet, pair_cap, ranges, tile_order, \
                                                                      img, dist, nc, st) \
                                       : launch_batch<TS, F, false>(plan_dev, h, sorted_tet, pair_cap, ranges, tile_order, \
                                                                    img, dist, nc, st))
    if (h.n_features == 7) {
        if (h.tile == 4) RM_BATCH(4, 7); else if (h.tile == 8) RM_BATCH(8, 7); else RM_BATCH(16, 7);
    } else {
        if (h.tile == 4) RM_BATCH(4, 13); else if (h.tile == 8) RM_BATCH(8, 13); else RM_BATCH(16, 13);
    }
#undef RM_BATCH
    return hipGetLastError();
}

}  // namespace rmesh

using namespace rmesh;

extern "C" int rm_set_intersect_precision(int fp64) {
    g_fp64_planes = fp64 ? 1 : 0;
    return RM_OK;
}

namespace {
int composite_fwd_entry(const rm_camera* cam, const int32_t* sorted_tet_dev, int64_t n_sorted,
                        const int32_t* ranges_dev, const int32_t* indices_dev, const float* vertices_dev,
                        const float* features_dev, int32_t n_features, const float* jitter_dev, float* out_img_dev,
                        float* out_dist_dev, int32_t* out_ncontrib_dev, uint8_t* flags_dev, rm_stream_t stream,
                        const char* who, const int32_t* tile_order_dev = nullptr) {
    const std::string w(who);
    RM_CHECK_ARG(n_sorted >= 0, w + ": negative n_sorted");
    if (n_sorted == 0 && sorted_tet_dev == nullptr) sorted_tet_dev = ranges_dev;  // empty list
    int e = check_common(cam, n_features, sorted_tet_dev, ranges_dev, indices_dev, vertices_dev, features_dev,
                         jitter_dev, who);
    if (e) return e;
    RM_CHECK_ARG(out_img_dev && out_dist_dev && out_ncontrib_dev, w + ": null output");
    RM_CHECK_ARG(((uintptr_t)out_img_dev & 15) == 0, w + ": out_img must be 16-byte aligned");
    CamArgs a = make_cam_args(*cam);
    hipStream_t st = as_stream(stream);
    const int ts = cam->tile_width;
#define RM_FWD(TS, F, P) launch_fwd<TS, F, P>(a, sorted_tet_dev, n_sorted, ranges_dev, indices_dev, vertices_dev, \
                                        features_dev, jitter_dev, out_img_dev, out_dist_dev, out_ncontrib_dev, st, \
                                        flags_dev, tile_order_dev)
#define RM_FWD_P(TS, F) (g_fp64_planes ? RM_FWD(TS, F, true) : RM_FWD(TS, F, false))
    if (n_features == 7) {
        if (ts == 4) RM_FWD_P(4, 7); else if (ts == 8) RM_FWD_P(8, 7); else RM_FWD_P(16, 7);
    } else {
        if (ts == 4) RM_FWD_P(4, 13); else if (ts == 8) RM_FWD_P(8, 13); else RM_FWD_P(16, 13);
    }
#undef RM_FWD_P
#undef RM_FWD
    RM_CHECK_LAUNCH("composite_fwd_kernel");
    return RM_OK;
}
}  // namespace

extern "C" int rm_composite_fwd(const rm_camera* cam, const int32_t* sorted_tet_dev, int64_t n_sorted,
                                const int32_t* ranges_dev,
                                const int32_t* indices_dev, const float* vertices_dev, const float* features_dev,
                                int32_t n_features, const float* jitter_dev, float* out_img_dev,
                                float* out_dist_dev, int32_t* out_ncontrib_dev, rm_stream_t stream) {
    return composite_fwd_entry(cam, sorted_tet_dev, n_sorted, ranges_dev, indices_dev, vertices_dev, features_dev,
                               n_features, jitter_dev, out_img_dev, out_dist_dev, out_ncontrib_dev, nullptr, stream,
                               "rm_composite_fwd");
}

extern "C" int rm_composite_fwd_flags(const rm_camera* cam, const int32_t* sorted_tet_dev, int64_t n_sorted,
                                      const int32_t* ranges_dev, const int32_t* indices_dev,
                                      const float* vertices_dev, const float* features_dev, int32_t n_features,
                                      const float* jitter_dev, float* out_img_dev, float* out_dist_dev,
                                      int32_t* out_ncontrib_dev, uint8_t* pair_flags_dev,
                                      const int32_t* tile_order_dev, rm_stream_t stream) {
    return composite_fwd_entry(cam, sorted_tet_dev, n_sorted, ranges_dev, indices_dev, vertices_dev, features_dev,
                               n_features, jitter_dev, out_img_dev, out_dist_dev, out_ncontrib_dev, pair_flags_dev,
                               stream, "rm_composite_fwd_flags", tile_order_dev);
}

extern "C" int rm_composite_bwd(const rm_camera* cam, const int32_t* sorted_tet_dev, int64_t n_sorted,
                                const int32_t* ranges_dev,
                                const int32_t* indices_dev, const float* vertices_dev, const float* features_dev,
                                int32_t n_features, const float* jitter_dev, const float* img_dev,
                                const float* dist_dev, const int32_t* ncontrib_dev, const float* grad_img_dev,
                                const float* grad_dist_dev, float* grad_vertices_dev, float* grad_features_dev,
                                rm_stream_t stream) {
    RM_CHECK_ARG(n_sorted >= 0, "rm_composite_bwd: negative n_sorted");
    if (n_sorted == 0 && sorted_tet_dev == nullptr) sorted_tet_dev = ranges_dev;  // empty list
    int e = check_common(cam, n_features, sorted_tet_dev, ranges_dev, indices_dev, vertices_dev, features_dev,
                         jitter_dev, "rm_composite_bwd");
    if (e) return e;
    RM_CHECK_ARG(img_dev && dist_dev && ncontrib_dev && grad_img_dev && grad_dist_dev && grad_vertices_dev &&
                     grad_features_dev,
                 "rm_composite_bwd: null buffer");
    RM_CHECK_ARG(((uintptr_t)img_dev & 15) == 0 && ((uintptr_t)grad_img_dev & 15) == 0,
                 "rm_composite_bwd: img and grad_img must be 16-byte aligned");
    CamArgs a = make_cam_args(*cam);
    hipStream_t st = as_stream(stream);
    const int ts = cam->tile_width;
#define RM_BWD(TS, F, P) launch_bwd<TS, F, P, false>(a, sorted_tet_dev, n_sorted, ranges_dev, indices_dev, \
                                        vertices_dev, features_dev, jitter_dev, img_dev, dist_dev, ncontrib_dev, \
                                        grad_img_dev, grad_dist_dev, grad_vertices_dev, grad_features_dev, nullptr, \
                                        0, st)
#define RM_BWD_P(TS, F) (g_fp64_planes ? RM_BWD(TS, F, true) : RM_BWD(TS, F, false))
    if (n_features == 7) {
        if (ts == 4) RM_BWD_P(4, 7); else if (ts == 8) RM_BWD_P(8, 7); else RM_BWD_P(16, 7);
    } else {
        if (ts == 4) RM_BWD_P(4, 13); else if (ts == 8) RM_BWD_P(8, 13); else RM_BWD_P(16, 13);
    }
#undef RM_BWD_P
#undef RM_BWD
    RM_CHECK_LAUNCH("composite_bwd_kernel");
    return RM_OK;
}

extern "C" size_t rm_composite_bwd_workspace_size(int64_t n_tets) {
    return n_tets > 0 ? (size_t)n_tets * kRowFloats * sizeof(float) : 0;
}

namespace {
int composite_bwd_rows_entry(const rm_camera* cam, const int32_t* sorted_tet_dev, int64_t n_sorted,
                             const int32_t* ranges_dev, const int32_t* indices_dev, int64_t n_tets,
                             const float* vertices_dev, const float* features_dev, int32_t n_features,
                             const float* jitter_dev, const float* img_dev, const float* dist_dev,
                             const int32_t* ncontrib_dev, const float* grad_img_dev, const float* grad_dist_dev,
                             float* grad_vertices_dev, float* grad_features_dev, const uint8_t* flags_dev,
                             void* workspace_dev, size_t workspace_bytes, rm_stream_t stream, const char* who,
                             const int32_t* tile_order_dev = nullptr) {
    const std::string w(who);
    RM_CHECK_ARG(n_sorted >= 0, w + ": negative n_sorted");
    RM_CHECK_ARG(n_tets >= 0, w + ": negative n_tets");
    if (n_sorted == 0 && sorted_tet_dev == nullptr) sorted_tet_dev = ranges_dev;  // empty list
    int e = check_common(cam, n_features, sorted_tet_dev, ranges_dev, indices_dev, vertices_dev, features_dev,
                         jitter_dev, who);
    if (e) return e;
    RM_CHECK_ARG(img_dev && dist_dev && ncontrib_dev && grad_img_dev && grad_dist_dev && grad_vertices_dev &&
                     grad_features_dev,
                 w + ": null buffer");
    RM_CHECK_ARG(((uintptr_t)img_dev & 15) == 0 && ((uintptr_t)grad_img_dev & 15) == 0,
                 w + ": img and grad_img must be 16-byte aligned");
    RM_CHECK_ARG(workspace_bytes >= rm_composite_bwd_workspace_size(n_tets),
                 w + ": workspace smaller than rm_composite_bwd_workspace_size(n_tets)");
    RM_CHECK_ARG(n_tets == 0 || (workspace_dev && ((uintptr_t)workspace_dev & 127) == 0),
                 w + ": workspace must be 128-byte aligned");
    CamArgs a = make_cam_args(*cam);
    hipStream_t st = as_stream(stream);
    const int ts = cam->tile_width;
    float* rows = static_cast<float*>(workspace_dev);
#define RM_BWD(TS, F, P) launch_bwd<TS, F, P, true>(a, sorted_tet_dev, n_sorted, ranges_dev, indices_dev, \
                                        vertices_dev, features_dev, jitter_dev, img_dev, dist_dev, ncontrib_dev, \
                                        grad_img_dev, grad_dist_dev, grad_vertices_dev, grad_features_dev, rows, \
                                        n_tets, st, flags_dev, tile_order_dev)
#define RM_BWD_P(TS, F) (g_fp64_planes ? RM_BWD(TS, F, true) : RM_BWD(TS, F, false))
    if (n_features == 7) {
        if (ts == 4) RM_BWD_P(4, 7); else if (ts == 8) RM_BWD_P(8, 7); else RM_BWD_P(16, 7);
    } else {
        if (ts == 4) RM_BWD_P(4, 13); else if (ts == 8) RM_BWD_P(8, 13); else RM_BWD_P(16, 13);
    }
#undef RM_BWD_P
#undef RM_BWD
    RM_CHECK_LAUNCH("composite_bwd_kernel");
    return RM_OK;
}
}  // namespace

extern "C" int rm_composite_bwd_ws(const rm_camera* cam, const int32_t* sorted_tet_dev, int64_t n_sorted,
                                   const int32_t* ranges_dev, const int32_t* indices_dev, int64_t n_tets,
                                   const float* vertices_dev, const float* features_dev, int32_t n_features,
                                   const float* jitter_dev, const float* img_dev, const float* dist_dev,
                                   const int32_t* ncontrib_dev, const float* grad_img_dev,
                                   const float* grad_dist_dev, float* grad_vertices_dev, float* grad_features_dev,
                                   void* workspace_dev, size_t workspace_bytes, rm_stream_t stream) {
    return composite_bwd_rows_entry(cam, sorted_tet_dev, n_sorted, ranges_dev, indices_dev, n_tets, vertices_dev,
                                    features_dev, n_features, jitter_dev, img_dev, dist_dev, ncontrib_dev,
                                    grad_img_dev, grad_dist_dev, grad_vertices_dev, grad_features_dev, nullptr,
                                    workspace_dev, workspace_bytes, stream, "rm_composite_bwd_ws");
}

extern "C" int rm_composite_bwd_flags(const rm_camera* cam, const int32_t* sorted_tet_dev, int64_t n_sorted,
                                      const int32_t* ranges_dev, const int32_t* indices_dev, int64_t n_tets,
                                      const float* vertices_dev, const float* features_dev, int32_t n_features,
                                      const float* jitter_dev, const float* img_dev, const float* dist_dev,
                                      const int32_t* ncontrib_dev, const uint8_t* pair_flags_dev,
                                      const int32_t* tile_order_dev, const float* grad_img_dev,
                                      const float* grad_dist_dev, float* grad_vertices_dev, float* grad_features_dev,
                                      void* workspace_dev, size_t workspace_bytes, rm_stream_t stream) {
    RM_CHECK_ARG(pair_flags_dev != nullptr || n_sorted == 0, "rm_composite_bwd_flags: null pair_flags");
    return composite_bwd_rows_entry(cam, sorted_tet_dev, n_sorted, ranges_dev, indices_dev, n_tets, vertices_dev,
                                    features_dev, n_features, jitter_dev, img_dev, dist_dev, ncontrib_dev,
                                    grad_img_dev, grad_dist_dev, grad_vertices_dev, grad_features_dev,
                                    pair_flags_dev, workspace_dev, workspace_bytes, stream,
                                    "rm_composite_bwd_flags", tile_order_dev);
}

// ---- no-allocation variant of rm_composite_bwd_ws -------------------------------------------
// The per-tet rows live in a file-scope device array instead of a caller workspace, for callers
// that cannot allocate (code under graph capture, serving loops). Each GPU holds its own copy of
// the array from module load until the process ends; its device address is looked up once per
// device with hipGetSymbolAddress and only that address is handed to kernels (the array's name in
// host code is a host-side stand-in, not a device pointer).
constexpr int64_t kNoAllocMaxTets = int64_t(1) << 19;   // 524,288 tets: 64 MiB of rows per GPU
static __device__ __align__(256) float scratch[kNoAllocMaxTets * kRowFloats];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch_ptr[kMaxDevices] = {};

int scratch_of_current_device(float** out, const char* who) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        set_error(std::string(who) + ": no current HIP device (" + hipGetErrorString(e) + ")");
        return RM_ERR_HIP;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr) {
            set_error(std::string(who) + ": hipGetSymbolAddress(scratch) failed: " + hipGetErrorString(e));
            return RM_ERR_HIP;
        }
        g_scratch_ptr[dev] = static_cast<float*>(p);
    }
    *out = g_scratch_ptr[dev];
    return RM_OK;
}
}  // namespace

extern "C" int64_t rm_composite_bwd_noalloc_max_tets(void) { return kNoAllocMaxTets; }

extern "C" int rm_composite_bwd_noalloc(const rm_camera* cam, const int32_t* sorted_tet_dev, int64_t n_sorted,
                                        const int32_t* ranges_dev, const int32_t* indices_dev, int64_t n_tets,
                                        const float* vertices_dev, const float* features_dev, int32_t n_features,
                                        const float* jitter_dev, const float* img_dev, const float* dist_dev,
                                        const int32_t* ncontrib_dev, const float* grad_img_dev,
                                        const float* grad_dist_dev, float* grad_vertices_dev,
                                        float* grad_features_dev, rm_stream_t stream) {
    RM_CHECK_ARG(n_tets >= 0, "rm_composite_bwd_noalloc: negative n_tets");
    if (n_tets > kNoAllocMaxTets) {
        set_error("rm_composite_bwd_noalloc: n_tets " + std::to_string(n_tets) + " exceeds the static scratch (" +
                  std::to_string(kNoAllocMaxTets) + " tets); use rm_composite_bwd_ws");
        return RM_ERR_CAPACITY;
    }
    float* rows = nullptr;
    if (n_tets > 0) {
        int e = scratch_of_current_device(&rows, "rm_composite_bwd_noalloc");
        if (e) return e;
    }
    // same kernels as rm_composite_bwd_ws; the rows are cleared by a fill kernel before use
    return rm_composite_bwd_ws(cam, sorted_tet_dev, n_sorted, ranges_dev, indices_dev, n_tets, vertices_dev,
                               features_dev, n_features, jitter_dev, img_dev, dist_dev, ncontrib_dev, grad_img_dev,
                               grad_dist_dev, grad_vertices_dev, grad_features_dev, rows,
                               rm_composite_bwd_workspace_size(n_tets), stream);
}

extern "C" int rm_tet_err(const rm_camera* cam, const int32_t* sorted_tet_dev, int64_t n_sorted,
                          const int32_t* ranges_dev, const int32_t* indices_dev, const float* vertices_dev,
                          const float* features_dev, int32_t n_features, const float* pixel_err_dev,
                          const float* ssim_err_dev, float* tet_err_dev, int32_t* tet_count_dev,
                          int64_t n_tets, const int32_t* tile_order_dev, rm_stream_t stream) {
    RM_CHECK_ARG(n_sorted >= 0, "rm_tet_err: negative n_sorted");
    RM_CHECK_ARG(n_tets >= 0, "rm_tet_err: negative n_tets");
    if (n_sorted == 0 && sorted_tet_dev == nullptr) sorted_tet_dev = ranges_dev;  // empty list
    // the votes use pixel-centre rays: any non-null, aligned pointer stands in for the jitter check
    int e = check_common(cam, n_features, sorted_tet_dev, ranges_dev, indices_dev, vertices_dev, features_dev,
                         ranges_dev, "rm_tet_err");
    if (e) return e;
    RM_CHECK_ARG(pixel_err_dev && ssim_err_dev && tet_err_dev && tet_count_dev, "rm_tet_err: null buffer");
    CamArgs a = make_cam_args(*cam);
    hipStream_t st = as_stream(stream);
    const int ts = cam->tile_width;
#define RM_ERRV(TS, F, P) launch_tet_err<TS, F, P>(a, sorted_tet_dev, n_sorted, ranges_dev, indices_dev, vertices_dev, \
                                        features_dev, pixel_err_dev, ssim_err_dev, tet_err_dev, tet_count_dev, st, \
                                        tile_order_dev)
#define RM_ERRV_P(TS, F) (g_fp64_planes ? RM_ERRV(TS, F, true) : RM_ERRV(TS, F, false))
    if (n_features == 7) {
        if (ts == 4) RM_ERRV_P(4, 7); else if (ts == 8) RM_ERRV_P(8, 7); else RM_ERRV_P(16, 7);
    } else {
        if (ts == 4) RM_ERRV_P(4, 13); else if (ts == 8) RM_ERRV_P(8, 13); else RM_ERRV_P(16, 13);
    }
#undef RM_ERRV_P
#undef RM_ERRV
    if (n_tets > 0)
        hipLaunchKernelGGL(tet_err_finish_kernel, dim3((unsigned)ceil_div(n_tets, 256)), dim3(256), 0, st,
                           (float4*)tet_err_dev, tet_count_dev, n_tets);
    RM_CHECK_LAUNCH("tet_err_kernel");
    return RM_OK;
}
