// The glue of the reference model's field (iNGPDW, models/ingp_color.py:100-205) around the hash
// grid and the MLP heads:
//   field_query   per tet: query anchor (centroid, or circumcenter), circumradius in FP64, the
//                 contraction of the normalised anchor and of the radius (contract_mean_std,
//                 utils/contraction.py:6-20), the grid coordinate (cv/2 + 1)/2
//                 (compute_batch_features, models/ingp_color.py:314-357);
//   downweight    the erf down-weighting of the encoding (_encode, models/ingp_color.py:174-183);
//                 the same per-element scale is the adjoint, so one kernel serves both directions;
//   outputs       the heads' raw outputs -> density = safe_exp(sigma + offset), rgb + 0.5,
//                 grd / sqrt(|grd|^2 + 1), sh rounded to FP16 (forward, models/ingp_color.py:
//                 186-205), and the adjoint.
// Every float op rounds once, in the reference's order (contraction off), FP32 where the
// reference computes in float32 and FP64 where its tensors are double (the circumradius, and so
// the contracted radius and the down-weighting).
#include <hip/hip_fp16.h>

#include "common.h"

namespace rmesh {
namespace {

constexpr float kTinyF = 1.17549435e-38f;          // float32 tiny (smallest normal)
constexpr float kMaxF = 3.40282347e+38f;           // float32 max
constexpr double kTiny = 1.1754943508222875e-38;
constexpr double kMax = 3.4028234663852886e+38;
// nextafter(log(float32 max) rounded to float32, 0): the upper clamp of safe_exp
constexpr float kLogMaxF = 88.72283172607422f;
constexpr double kFourOverPi = 1.2732395447351628;  // 4 / pi in double

// safe_div (utils/safe_math.py:151-174) in precision T: 0 where |d| < tiny, else n/d clamped to
// the float32 range.
template <typename T>
__device__ inline T safe_div(T n, T d) {
    const T tiny = (T)kTinyF, big = (T)kMaxF;
    if (fabs(d) < tiny) return (T)0;
    T r = n / d;
    r = r < -big ? -big : (r > big ? big : r);
    return r;
}

struct QueryArgs {
    float center[3];
    float scene_scaling;
    int anchor;   // 0: centroid (ablate_circumsphere), 1: circumcenter
};

__global__ __launch_bounds__(256) void field_query_kernel(const int32_t* __restrict__ indices,
                                                          const float* __restrict__ vertices, int64_t n,
                                                          QueryArgs q, float* __restrict__ x_out,
                                                          double* __restrict__ cr_out, float* __restrict__ anchor_out) {
#pragma clang fp contract(off)
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    float v[4][3];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int64_t vi = indices[4 * t + k];
#pragma unroll
        for (int d = 0; d < 3; ++d) v[k][d] = vertices[3 * vi + d];
    }
    // circumcenter and radius of the tet in FP64 (calculate_circumcenters_torch,
    // utils/topo_utils.py:120-164, on tets.double())
    double a[3], b[3], c[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        a[d] = (double)v[1][d] - (double)v[0][d];
        b[d] = (double)v[2][d] - (double)v[0][d];
        c[d] = (double)v[3][d] - (double)v[0][d];
    }
    const double aa = (a[0] * a[0] + a[1] * a[1]) + a[2] * a[2];
    const double bb = (b[0] * b[0] + b[1] * b[1]) + b[2] * b[2];
    const double cc = (c[0] * c[0] + c[1] * c[1]) + c[2] * c[2];
    const double xbc[3] = {b[1] * c[2] - b[2] * c[1], b[2] * c[0] - b[0] * c[2], b[0] * c[1] - b[1] * c[0]};
    const double xca[3] = {c[1] * a[2] - c[2] * a[1], c[2] * a[0] - c[0] * a[2], c[0] * a[1] - c[1] * a[0]};
    const double xab[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    const double den = 2.0 * ((a[0] * xbc[0] + a[1] * xbc[1]) + a[2] * xbc[2]);
    double rel[3], dr[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        rel[d] = safe_div<double>((aa * xbc[d] + bb * xca[d]) + cc * xab[d], den);
        dr[d] = a[d] - rel[d];
    }
    const double radius = sqrt((dr[0] * dr[0] + dr[1] * dr[1]) + dr[2] * dr[2]);
    const double std0 = radius / (double)q.scene_scaling;

    float xo[3];
    double cr;
    if (q.anchor == 0) {
        // centroid in float32 (tets.mean(dim=1)), then the float32 contraction
        float cen[3], z[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            cen[d] = (((v[0][d] + v[1][d]) + v[2][d]) + v[3][d]) / 4.0f;
            anchor_out[3 * t + d] = cen[d];
            z[d] = (cen[d] - q.center[d]) / q.scene_scaling;
        }
        float msq = (z[0] * z[0] + z[1] * z[1]) + z[2] * z[2];
        msq = fmaxf(msq, 1.1920928955078125e-07f);
        const float mag = sqrtf(msq);
        const bool inside = msq <= 1.0f;
        if (!inside) {
            const float s = safe_div<float>(2.0f * sqrtf(msq) - 1.0f, msq);
#pragma unroll
            for (int d = 0; d < 3; ++d) z[d] = s * z[d];
        }
        float top = 2.0f * mag - 1.0f;
        top = fminf(fmaxf(top, 1e-6f), 1e6f);
        const float r13 = powf(top, 1.0f / 3.0f) / mag;
        const float det13 = r13 * r13;
        cr = inside ? std0 : (double)det13 * std0;
#pragma unroll
        for (int d = 0; d < 3; ++d) xo[d] = (z[d] / 2.0f + 1.0f) / 2.0f;
    } else {
        // circumcenter in FP64 and the FP64 contraction; the grid coordinate is rounded to float32
        double z[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double cen = (double)v[0][d] + rel[d];
            anchor_out[3 * t + d] = (float)cen;
            z[d] = (cen - (double)q.center[d]) / (double)q.scene_scaling;
        }
        double msq = (z[0] * z[0] + z[1] * z[1]) + z[2] * z[2];
        msq = fmax(msq, 1.1920928955078125e-07);
        const double mag = sqrt(msq);
        const bool inside = msq <= 1.0;
        if (!inside) {
            const double s = safe_div<double>(2.0 * sqrt(msq) - 1.0, msq);
#pragma unroll
            for (int d = 0; d < 3; ++d) z[d] = s * z[d];
        }
        double top = 2.0 * mag - 1.0;
        top = fmin(fmax(top, 1e-6), 1e6);
        const double r13 = pow(top, 1.0 / 3.0) / mag;
        const double det13 = r13 * r13;
        cr = inside ? std0 : det13 * std0;
#pragma unroll
        for (int d = 0; d < 3; ++d) xo[d] = (float)((z[d] / 2.0 + 1.0) / 2.0);
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) x_out[3 * t + d] = xo[d];
    cr_out[t] = cr;
}

// Scale of encoding column group k (0 <= k < L) at contracted radius cr, in FP64:
// approx_erf(safe_div(1, safe_sqrt(per_level_scale * 4 * k * cr * scale_multi))) with
// approx_erf(x) = sign(x) sqrt(1 - exp(-(4/pi) x^2)) (models/ingp_color.py:84-86). k = 0 gives
// safe_sqrt's floor (tiny), a huge ratio and a scale of exactly 1.
__device__ inline double erf_scale(int k, double cr, double scale_multi, double pls4) {
#pragma clang fp contract(off)
    const double crs = cr * scale_multi;
    double d = (pls4 * (double)k) * crs;
    d = fmin(fmax(d, kTiny), kMax);
    const double r = safe_div<double>(1.0, sqrt(d));
    const double e = sqrt(1.0 - exp(-(kFourOverPi * (r * r))));
    return r > 0.0 ? e : (r < 0.0 ? -e : 0.0);
}

// One thread per four consecutive elements of the [n, L*F] encoding (float4 in and out, so a wave
// reads and writes one contiguous kilobyte); column i belongs to group i % L (the reference views
// the level-major row as [F][L] and scales the last axis). Each thread evaluates the scales of its
// four columns itself (four FP64 exponentials per 16 bytes in and out).
__global__ __launch_bounds__(256) void downweight4_kernel(const float4* __restrict__ in, const double* __restrict__ cr,
                                                          int64_t n4, int C, int L, double scale_multi, double pls4,
                                                          float4* __restrict__ out) {
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n4) return;
    const int64_t e = 4 * i;
    const int64_t p = e / C;
    const int col = (int)(e - p * C);
    const double c = cr[p];
    const float4 v = in[i];
    float4 o;
    o.x = (float)((double)v.x * erf_scale(col % L, c, scale_multi, pls4));
    o.y = (float)((double)v.y * erf_scale((col + 1) % L, c, scale_multi, pls4));
    o.z = (float)((double)v.z * erf_scale((col + 2) % L, c, scale_multi, pls4));
    o.w = (float)((double)v.w * erf_scale((col + 3) % L, c, scale_multi, pls4));
    out[i] = o;
}

// Rows whose width is not a multiple of 4: one thread per element.
__global__ __launch_bounds__(256) void downweight_kernel(const float* __restrict__ in, const double* __restrict__ cr,
                                                         int64_t n_el, int C, int L, double scale_multi, double pls4,
                                                         float* __restrict__ out) {
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_el) return;
    const int64_t p = i / C;
    const int col = (int)(i - p * C);
    out[i] = (float)((double)in[i] * erf_scale(col % L, cr[p], scale_multi, pls4));
}

// One thread per element of the [n, n_sh] SH block (coalesced); threads below n also do that
// query's density, colour and gradient.
__global__ __launch_bounds__(256) void outputs_fwd_kernel(const float* __restrict__ sigma, const float* __restrict__ rgb,
                                                          const float* __restrict__ grd, const float* __restrict__ sh,
                                                          int64_t n, int n_sh, float density_offset,
                                                          float* __restrict__ density, float* __restrict__ rgb_out,
                                                          float* __restrict__ grd_out, __half* __restrict__ sh_out) {
#pragma clang fp contract(off)
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n * n_sh) sh_out[p] = __float2half_rn(sh[p]);
    if (p >= n) return;
    float s = sigma[p] + density_offset;
    s = fminf(fmaxf(s, -kMaxF), kLogMaxF);
    density[p] = expf(s);
    float g[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        rgb_out[3 * p + d] = rgb[3 * p + d] + 0.5f;
        g[d] = grd[3 * p + d];
    }
    const float den = sqrtf(((g[0] * g[0] + g[1] * g[1]) + g[2] * g[2]) + 1.0f);
#pragma unroll
    for (int d = 0; d < 3; ++d) grd_out[3 * p + d] = g[d] / den;
}

__global__ __launch_bounds__(256) void outputs_bwd_kernel(const float* __restrict__ density,
                                                          const float* __restrict__ grd, const float* __restrict__ g_density,
                                                          const float* __restrict__ g_rgb, const float* __restrict__ g_grd,
                                                          const __half* __restrict__ g_sh, int64_t n, int n_sh,
                                                          float* __restrict__ g_sigma, float* __restrict__ g_rgb_raw,
                                                          float* __restrict__ g_grd_raw, float* __restrict__ g_sh_raw) {
#pragma clang fp contract(off)
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    // the FP16 cast passes its (FP16) gradient back as float32
    if (p < n * n_sh) g_sh_raw[p] = g_sh ? __half2float(g_sh[p]) : 0.0f;
    if (p >= n) return;
    // safe_exp's backward: y * grad (utils/safe_math.py:231-233), also where the input was clamped
    g_sigma[p] = g_density ? density[p] * g_density[p] : 0.0f;
    float g[3], gy[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        g_rgb_raw[3 * p + d] = g_rgb ? g_rgb[3 * p + d] : 0.0f;
        g[d] = grd[3 * p + d];
        gy[d] = g_grd ? g_grd[3 * p + d] : 0.0f;
    }
    // y = g / s, s = sqrt(|g|^2 + 1): dL/dg = gy / s - g (g . gy) / s^3
    const float s = sqrtf(((g[0] * g[0] + g[1] * g[1]) + g[2] * g[2]) + 1.0f);
    const float dot = (g[0] * gy[0] + g[1] * gy[1]) + g[2] * gy[2];
    const float k = dot / ((s * s) * s);
#pragma unroll
    for (int d = 0; d < 3; ++d) g_grd_raw[3 * p + d] = gy[d] / s - g[d] * k;
}

inline unsigned blocks_for(int64_t n) { return (unsigned)ceil_div(n, 256); }

}  // namespace
}  // namespace rmesh

using namespace rmesh;

extern "C" int rm_field_query(const int32_t* indices_dev, const float* vertices_dev, int64_t n_tets,
                              const float* center, float scene_scaling, int32_t anchor, float* x_dev,
                              double* cr_dev, float* anchor_dev, rm_stream_t stream) {
    RM_CHECK_ARG(n_tets >= 0, "rm_field_query: negative n_tets");
    RM_CHECK_ARG(anchor == 0 || anchor == 1, "rm_field_query: anchor must be 0 (centroid) or 1 (circumcenter)");
    RM_CHECK_ARG(center != nullptr, "rm_field_query: null center");
    RM_CHECK_ARG(scene_scaling > 0.0f, "rm_field_query: scene_scaling must be positive");
    if (n_tets == 0) return RM_OK;
    RM_CHECK_ARG(indices_dev && vertices_dev && x_dev && cr_dev && anchor_dev, "rm_field_query: null buffer");
    QueryArgs q;
    for (int d = 0; d < 3; ++d) q.center[d] = center[d];
    q.scene_scaling = scene_scaling;
    q.anchor = anchor;
    hipLaunchKernelGGL(field_query_kernel, dim3(blocks_for(n_tets)), dim3(256), 0, as_stream(stream), indices_dev,
                       vertices_dev, n_tets, q, x_dev, cr_dev, anchor_dev);
    RM_CHECK_LAUNCH("field_query_kernel");
    return RM_OK;
}

extern "C" int rm_field_downweight(const float* in_dev, const double* cr_dev, int64_t n, int32_t n_levels,
                                   int32_t n_feat, double scale_multi, double per_level_scale, float* out_dev,
                                   rm_stream_t stream) {
    RM_CHECK_ARG(n >= 0, "rm_field_downweight: negative n");
    RM_CHECK_ARG(n_levels >= 1 && n_levels <= 64 && n_feat >= 1 && n_feat <= 64,
                 "rm_field_downweight: n_levels and n_feat must be in [1, 64]");
    if (n == 0) return RM_OK;
    RM_CHECK_ARG(in_dev && cr_dev && out_dev, "rm_field_downweight: null buffer");
    const int C = n_levels * n_feat;
    if (C % 4 == 0 && ((uintptr_t)in_dev & 15) == 0 && ((uintptr_t)out_dev & 15) == 0) {
        const int64_t n4 = n * C / 4;
        hipLaunchKernelGGL(downweight4_kernel, dim3(blocks_for(n4)), dim3(256), 0, as_stream(stream),
                           reinterpret_cast<const float4*>(in_dev), cr_dev, n4, C, n_levels, scale_multi,
                           per_level_scale * 4.0, reinterpret_cast<float4*>(out_dev));
    } else {
        hipLaunchKernelGGL(downweight_kernel, dim3(blocks_for(n * C)), dim3(256), 0, as_stream(stream), in_dev,
                           cr_dev, n * C, C, n_levels, scale_multi, per_level_scale * 4.0, out_dev);
    }
    RM_CHECK_LAUNCH("downweight_kernel");
    return RM_OK;
}

extern "C" int rm_field_outputs_fwd(const float* sigma_dev, const float* rgb_dev, const float* grd_dev,
                                    const float* sh_dev, int64_t n, int32_t n_sh, float density_offset,
                                    float* density_dev, float* rgb_out_dev, float* grd_out_dev, uint16_t* sh_out_dev,
                                    rm_stream_t stream) {
    RM_CHECK_ARG(n >= 0 && n_sh >= 0, "rm_field_outputs_fwd: negative size");
    if (n == 0) return RM_OK;
    RM_CHECK_ARG(sigma_dev && rgb_dev && grd_dev && density_dev && rgb_out_dev && grd_out_dev,
                 "rm_field_outputs_fwd: null buffer");
    RM_CHECK_ARG(n_sh == 0 || (sh_dev && sh_out_dev), "rm_field_outputs_fwd: null sh buffer");
    hipLaunchKernelGGL(outputs_fwd_kernel, dim3(blocks_for(n_sh > 1 ? n * n_sh : n)), dim3(256), 0, as_stream(stream), sigma_dev, rgb_dev,
                       grd_dev, sh_dev, n, n_sh, density_offset, density_dev, rgb_out_dev, grd_out_dev,
                       reinterpret_cast<__half*>(sh_out_dev));
    RM_CHECK_LAUNCH("outputs_fwd_kernel");
    return RM_OK;
}

extern "C" int rm_field_outputs_bwd(const float* density_dev, const float* grd_raw_dev, const float* g_density_dev,
                                    const float* g_rgb_dev, const float* g_grd_dev, const uint16_t* g_sh_dev,
                                    int64_t n, int32_t n_sh, float* g_sigma_dev, float* g_rgb_raw_dev,
                                    float* g_grd_raw_dev, float* g_sh_raw_dev, rm_stream_t stream) {
    RM_CHECK_ARG(n >= 0 && n_sh >= 0, "rm_field_outputs_bwd: negative size");
    if (n == 0) return RM_OK;
    RM_CHECK_ARG(density_dev && grd_raw_dev && g_sigma_dev && g_rgb_raw_dev && g_grd_raw_dev,
                 "rm_field_outputs_bwd: null buffer");
    RM_CHECK_ARG(n_sh == 0 || g_sh_raw_dev, "rm_field_outputs_bwd: null sh gradient buffer");
    hipLaunchKernelGGL(outputs_bwd_kernel, dim3(blocks_for(n_sh > 1 ? n * n_sh : n)), dim3(256), 0, as_stream(stream), density_dev,
                       grd_raw_dev, g_density_dev, g_rgb_dev, g_grd_dev, reinterpret_cast<const __half*>(g_sh_dev), n,
                       n_sh, g_sigma_dev, g_rgb_raw_dev, g_grd_raw_dev, g_sh_raw_dev);
    RM_CHECK_LAUNCH("outputs_bwd_kernel");
    return RM_OK;
}
