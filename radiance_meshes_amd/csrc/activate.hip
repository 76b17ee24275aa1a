// Per-tet feature activation of the reference's field model, fused into one pass per direction.
//
// Semantics: activate_output (utils/model_util.py:93-105) with offset_normalize (:81-91),
// compute_vertex_colors_from_field (:62-78), eval_sh (utils/eval_sh_py.py:137-195) and safe_div
// (utils/safe_math.py:143-171). Per tet t with vertices v0..v3, circumcenter c, raw colour rgb,
// raw gradient g, SH coefficients sh[k][ch] (k < K, first (deg+1)^2 - 1 used), camera centre o:
//   m      = min_ch rgb[ch]                       (gradient to the first minimal channel)
//   r      = |v0 - c|,  ng = safe_div(g m, r)     (output features 4..6)
//   b[ch]  = rgb[ch] + ng . (v0 - c)              (v0 and c taken as constants here, as the reference
//                                                  detaches them in compute_vertex_colors_from_field)
//   dir    = (mean(v) - o) / max(|mean(v) - o|, 1e-8)
//   raw    = b[ch] + sum_k Y_k(dir) sh[k][ch]     (eval_sh of RGB2SH(b): SH_C0 * sh0 + 0.5 == b)
//   col    = softplus(raw, beta 10, threshold 20) (output features 1..3)
// features = [density, col, ng]. The backward adds the vertex gradients (through r and through
// dir) with float atomics and writes the others.
#include "common.h"

namespace rmesh {
namespace {

constexpr int kMaxSh = 15;   // degree 3

struct Sh3 {
    static constexpr float c1 = 0.48860251190291992f;    // sqrt(3 / 4pi)
    static constexpr float c2a = 1.0925484305920792f;    // sqrt(15 / pi) / 2
    static constexpr float c2b = 0.31539156525252005f;   // sqrt(5 / pi) / 4
    static constexpr float c2c = 0.54627421529603959f;   // sqrt(15 / pi) / 4
    static constexpr float c3a = 0.59004358992664352f;   // sqrt(35 / 2pi) / 4
    static constexpr float c3b = 2.8906114426405538f;    // sqrt(105 / pi) / 2
    static constexpr float c3c = 0.45704579946446572f;   // sqrt(21 / 2pi) / 4
    static constexpr float c3d = 0.37317633259011540f;   // sqrt(7 / pi) / 4
    static constexpr float c3e = 1.4453057213202769f;    // sqrt(105 / pi) / 4
};

// Real SH basis (degrees 1..3, m = -l..l; sqrt(2) Im / Re of the Condon-Shortley harmonics) and,
// when G is set, its gradient with respect to (x, y, z) taken as free variables.
template <bool G>
__device__ inline void sh_basis(float x, float y, float z, int n, float (&Y)[kMaxSh], float (&dY)[kMaxSh][3]) {
    using S = Sh3;
    const float xx = x * x, yy = y * y, zz = z * z;
    auto set = [&](int k, float v, float gx, float gy, float gz) {
        if (k < n) {
            Y[k] = v;
            if (G) { dY[k][0] = gx; dY[k][1] = gy; dY[k][2] = gz; }
        }
    };
    set(0, -S::c1 * y, 0.f, -S::c1, 0.f);
    set(1, S::c1 * z, 0.f, 0.f, S::c1);
    set(2, -S::c1 * x, -S::c1, 0.f, 0.f);
    set(3, S::c2a * x * y, S::c2a * y, S::c2a * x, 0.f);
    set(4, -S::c2a * y * z, 0.f, -S::c2a * z, -S::c2a * y);
    set(5, S::c2b * (2.f * zz - xx - yy), -2.f * S::c2b * x, -2.f * S::c2b * y, 4.f * S::c2b * z);
    set(6, -S::c2a * x * z, -S::c2a * z, 0.f, -S::c2a * x);
    set(7, S::c2c * (xx - yy), 2.f * S::c2c * x, -2.f * S::c2c * y, 0.f);
    set(8, -S::c3a * y * (3.f * xx - yy), -6.f * S::c3a * x * y, -3.f * S::c3a * (xx - yy), 0.f);
    set(9, S::c3b * x * y * z, S::c3b * y * z, S::c3b * x * z, S::c3b * x * y);
    set(10, -S::c3c * y * (4.f * zz - xx - yy), 2.f * S::c3c * x * y, -S::c3c * (4.f * zz - xx - 3.f * yy),
        -8.f * S::c3c * y * z);
    set(11, S::c3d * z * (2.f * zz - 3.f * xx - 3.f * yy), -6.f * S::c3d * x * z, -6.f * S::c3d * y * z,
        S::c3d * (6.f * zz - 3.f * xx - 3.f * yy));
    set(12, -S::c3c * x * (4.f * zz - xx - yy), -S::c3c * (4.f * zz - 3.f * xx - yy), 2.f * S::c3c * x * y,
        -8.f * S::c3c * x * z);
    set(13, S::c3e * z * (xx - yy), 2.f * S::c3e * x * z, -2.f * S::c3e * y * z, S::c3e * (xx - yy));
    set(14, -S::c3a * x * (xx - 3.f * yy), -3.f * S::c3a * (xx - yy), 6.f * S::c3a * x * y, 0.f);
}

constexpr float kTiny = 1.17549435e-38f;   // torch.finfo(float32).tiny
constexpr float kFMax = 3.40282347e+38f;

__device__ inline float remove_zero(float d) { return fabsf(d) < kTiny ? kTiny : d; }
__device__ inline float clampf(float v) { return fminf(fmaxf(v, -kFMax), kFMax); }
__device__ inline float softplus10(float x) { return 10.f * x > 20.f ? x : log1pf(expf(10.f * x)) / 10.f; }
__device__ inline float dsoftplus10(float x) { return 10.f * x > 20.f ? 1.f : 1.f / (1.f + expf(-10.f * x)); }

struct TetIn {
    float3 v[4], c;
    float rgb[3], g[3], dens;
};

__device__ inline TetIn load_tet(int64_t t, const float* dens, const float* rgb, const float* grd, const int4* idx,
                                 const float* cc, const float* verts) {
    TetIn r;
    const int4 q = idx[t];
    const int vid[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float* p = verts + 3 * (int64_t)vid[k];
        r.v[k] = make_float3(p[0], p[1], p[2]);
    }
    r.c = make_float3(cc[3 * t], cc[3 * t + 1], cc[3 * t + 2]);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        r.rgb[k] = rgb[3 * t + k];
        r.g[k] = grd[3 * t + k];
    }
    r.dens = dens[t];
    return r;
}

// Shared forward quantities of one tet.
struct TetFwd {
    float m, r, ng[3], off[3], b[3], md[3], dlen, dir[3], raw[3];
    int amin;
};

__device__ inline void tet_forward(const TetIn& I, const float* __restrict__ sh, int64_t t, int K, int nsh,
                                   const float* __restrict__ o, TetFwd& F, float (&Y)[kMaxSh]) {
    F.amin = 0;
    F.m = I.rgb[0];
    if (I.rgb[1] < F.m) { F.m = I.rgb[1]; F.amin = 1; }
    if (I.rgb[2] < F.m) { F.m = I.rgb[2]; F.amin = 2; }
    F.off[0] = I.v[0].x - I.c.x; F.off[1] = I.v[0].y - I.c.y; F.off[2] = I.v[0].z - I.c.z;
    F.r = sqrtf(F.off[0] * F.off[0] + F.off[1] * F.off[1] + F.off[2] * F.off[2]);
    const bool zero = fabsf(F.r) < kTiny;
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        F.ng[k] = zero ? 0.f : clampf((I.g[k] * F.m) / remove_zero(F.r));
        s += F.ng[k] * F.off[k];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) F.b[k] = I.rgb[k] + s;
    F.md[0] = 0.25f * (I.v[0].x + I.v[1].x + I.v[2].x + I.v[3].x) - o[0];
    F.md[1] = 0.25f * (I.v[0].y + I.v[1].y + I.v[2].y + I.v[3].y) - o[1];
    F.md[2] = 0.25f * (I.v[0].z + I.v[1].z + I.v[2].z + I.v[3].z) - o[2];
    F.dlen = fmaxf(sqrtf(F.md[0] * F.md[0] + F.md[1] * F.md[1] + F.md[2] * F.md[2]), 1e-8f);
#pragma unroll
    for (int k = 0; k < 3; ++k) F.dir[k] = F.md[k] / F.dlen;
    float dY[kMaxSh][3];
    sh_basis<false>(F.dir[0], F.dir[1], F.dir[2], nsh, Y, dY);
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
        float v = F.b[ch];
        for (int k = 0; k < nsh; ++k) v += Y[k] * sh[((int64_t)t * K + k) * 3 + ch];
        F.raw[ch] = v;
    }
}

__global__ __launch_bounds__(256) void activate_fwd_kernel(const float* __restrict__ o, const float* __restrict__ dens,
                                                           const float* __restrict__ rgb, const float* __restrict__ grd,
                                                           const float* __restrict__ sh, int K, int nsh,
                                                           const int4* __restrict__ idx, const float* __restrict__ cc,
                                                           const float* __restrict__ verts, int64_t T,
                                                           float* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    const TetIn I = load_tet(t, dens, rgb, grd, idx, cc, verts);
    TetFwd F;
    float Y[kMaxSh];
    tet_forward(I, sh, t, K, nsh, o, F, Y);
    float* f = out + 7 * t;
    f[0] = I.dens;
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) f[1 + ch] = softplus10(F.raw[ch]);
#pragma unroll
    for (int k = 0; k < 3; ++k) f[4 + k] = F.ng[k];
}

__global__ __launch_bounds__(256) void activate_bwd_kernel(
    const float* __restrict__ o, const float* __restrict__ dens, const float* __restrict__ rgb,
    const float* __restrict__ grd, const float* __restrict__ sh, int K, int nsh, const int4* __restrict__ idx,
    const float* __restrict__ cc, const float* __restrict__ verts, int64_t T, const float* __restrict__ gout,
    float* __restrict__ gdens, float* __restrict__ grgb, float* __restrict__ ggrd, float* __restrict__ gsh,
    float* __restrict__ gverts, float* __restrict__ gcc) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    const TetIn I = load_tet(t, dens, rgb, grd, idx, cc, verts);
    TetFwd F;
    float Y[kMaxSh], dY[kMaxSh][3];
    tet_forward(I, sh, t, K, nsh, o, F, Y);
    sh_basis<true>(F.dir[0], F.dir[1], F.dir[2], nsh, Y, dY);
    const float* G = gout + 7 * t;
    if (gdens) gdens[t] = G[0];
    float graw[3];
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) graw[ch] = G[1 + ch] * dsoftplus10(F.raw[ch]);
    // SH coefficients and the direction
    float gdir[3] = {0.f, 0.f, 0.f};
    for (int k = 0; k < nsh; ++k) {
        float a = 0.f;
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
            const int64_t e = ((int64_t)t * K + k) * 3 + ch;
            if (gsh) gsh[e] = graw[ch] * Y[k];
            a += graw[ch] * sh[e];
        }
#pragma unroll
        for (int j = 0; j < 3; ++j) gdir[j] += a * dY[k][j];
    }
    if (gsh)
        for (int k = nsh; k < K; ++k)
#pragma unroll
            for (int ch = 0; ch < 3; ++ch) gsh[((int64_t)t * K + k) * 3 + ch] = 0.f;
    // dir = md / max(|md|, eps): d dir = (I - dir dir^T) d md / |md| above eps, d md / eps below
    float gmd[3];
    const bool normed = F.dlen > 1e-8f;
    const float dd = gdir[0] * F.dir[0] + gdir[1] * F.dir[1] + gdir[2] * F.dir[2];
#pragma unroll
    for (int j = 0; j < 3; ++j) gmd[j] = (normed ? gdir[j] - dd * F.dir[j] : gdir[j]) / F.dlen;
    // b[ch] = rgb[ch] + ng . off  (off constant)
    const float gs = graw[0] + graw[1] + graw[2];
    float grgb_l[3] = {graw[0], graw[1], graw[2]};
    float gng[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) gng[k] = gs * F.off[k] + G[4 + k];
    // ng = safe_div(g m, r): d(gm) = clamp(gng / r'), d r = sum clamp(-gng ng / r')
    const float rz = remove_zero(F.r);
    float gr = 0.f, gm = 0.f;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float gn = clampf(gng[k] / rz);
        gr += clampf(-gng[k] * F.ng[k] / rz);
        if (ggrd) ggrd[3 * t + k] = gn * F.m;
        gm += gn * I.g[k];
    }
    grgb_l[F.amin] += gm;
    if (grgb)
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) grgb[3 * t + ch] = grgb_l[ch];
    // r = |v0 - c|
    float grv[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) grv[k] = F.r > 0.f ? gr * F.off[k] / F.r : 0.f;
    if (gcc)
#pragma unroll
        for (int k = 0; k < 3; ++k) gcc[3 * t + k] = -grv[k];
    if (gverts) {
        const int4 q = idx[t];
        const int vid[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const float a = 0.25f * gmd[k] + (v == 0 ? grv[k] : 0.f);
                if (a != 0.f) unsafeAtomicAdd(gverts + 3 * (int64_t)vid[v] + k, a);
            }
    }
}

int check_activate(const float* o, const float* dens, const float* rgb, const float* grd, const float* sh, int K,
                   int deg, const int32_t* idx, const float* cc, const float* verts, int64_t T, const char* who) {
    std::string w(who);
    if (T < 0) { set_error(w + ": negative n_tets"); return RM_ERR_INVALID_ARGUMENT; }
    if (deg < 0 || deg > 3) { set_error(w + ": SH degree must be in [0, 3]"); return RM_ERR_INVALID_ARGUMENT; }
    if (K < (deg + 1) * (deg + 1) - 1 || K > kMaxSh) {
        set_error(w + ": sh must hold (deg+1)^2 - 1 <= K <= 15 coefficients per channel");
        return RM_ERR_INVALID_ARGUMENT;
    }
    if (T > 0 && (!o || !dens || !rgb || !grd || !idx || !cc || !verts || (K > 0 && !sh))) {
        set_error(w + ": null buffer");
        return RM_ERR_INVALID_ARGUMENT;
    }
    if (((uintptr_t)idx & 15) != 0) { set_error(w + ": indices must be 16-byte aligned"); return RM_ERR_INVALID_ARGUMENT; }
    return RM_OK;
}

}  // namespace
}  // namespace rmesh

using namespace rmesh;

extern "C" int rm_activate_fwd(const float* camera_center_dev, const float* density_dev, const float* rgb_dev,
                               const float* grd_dev, const float* sh_dev, int32_t sh_coeffs, int32_t sh_degree,
                               const int32_t* indices_dev, const float* circumcenters_dev, const float* vertices_dev,
                               int64_t n_tets, float* features_dev, rm_stream_t stream) {
    int e = check_activate(camera_center_dev, density_dev, rgb_dev, grd_dev, sh_dev, sh_coeffs, sh_degree, indices_dev,
                           circumcenters_dev, vertices_dev, n_tets, "rm_activate_fwd");
    if (e) return e;
    if (n_tets == 0) return RM_OK;
    RM_CHECK_ARG(features_dev != nullptr, "rm_activate_fwd: null features");
    const int nsh = (sh_degree + 1) * (sh_degree + 1) - 1;
    hipLaunchKernelGGL(activate_fwd_kernel, dim3((unsigned)ceil_div(n_tets, 256)), dim3(256), 0, as_stream(stream),
                       camera_center_dev, density_dev, rgb_dev, grd_dev, sh_dev, (int)sh_coeffs, nsh,
                       (const int4*)indices_dev, circumcenters_dev, vertices_dev, n_tets, features_dev);
    RM_CHECK_LAUNCH("activate_fwd_kernel");
    return RM_OK;
}

extern "C" int rm_activate_bwd(const float* camera_center_dev, const float* density_dev, const float* rgb_dev,
                               const float* grd_dev, const float* sh_dev, int32_t sh_coeffs, int32_t sh_degree,
                               const int32_t* indices_dev, const float* circumcenters_dev, const float* vertices_dev,
                               int64_t n_tets, const float* grad_features_dev, float* grad_density_dev,
                               float* grad_rgb_dev, float* grad_grd_dev, float* grad_sh_dev, float* grad_vertices_dev,
                               float* grad_circumcenters_dev, rm_stream_t stream) {
    int e = check_activate(camera_center_dev, density_dev, rgb_dev, grd_dev, sh_dev, sh_coeffs, sh_degree, indices_dev,
                           circumcenters_dev, vertices_dev, n_tets, "rm_activate_bwd");
    if (e) return e;
    if (n_tets == 0) return RM_OK;
    RM_CHECK_ARG(grad_features_dev != nullptr, "rm_activate_bwd: null grad_features");
    const int nsh = (sh_degree + 1) * (sh_degree + 1) - 1;
    hipLaunchKernelGGL(activate_bwd_kernel, dim3((unsigned)ceil_div(n_tets, 256)), dim3(256), 0, as_stream(stream),
                       camera_center_dev, density_dev, rgb_dev, grd_dev, sh_dev, (int)sh_coeffs, nsh,
                       (const int4*)indices_dev, circumcenters_dev, vertices_dev, n_tets, grad_features_dev,
                       grad_density_dev, grad_rgb_dev, grad_grd_dev, grad_sh_dev, grad_vertices_dev,
                       grad_circumcenters_dev);
    RM_CHECK_LAUNCH("activate_bwd_kernel");
    return RM_OK;
}
