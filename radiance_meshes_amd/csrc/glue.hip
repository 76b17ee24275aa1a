// Per-tet field glue around the hash grid and the MLP, fused into three passes over the visible
// tets (the reference does these steps as separate PyTorch ops):
//  - query points: the tet centroid (models/ingp_color.py:376, vertices[indices].mean(dim=1))
//    contracted into the unit box (utils/contraction.py:6-20 contract_mean_std position part,
//    |z| <= 1 kept, else (2|z| - 1)/|z|^2 z, then (z/2 + 1)/2), one FP32 point per tet;
//  - head activation written straight into the tet's feature row (models/ingp_color.py:200-203:
//    density = exp(raw) with its argument clamped at log(1e20), base colour = raw + 0.5, colour
//    gradient rows unchanged), rows of tets that are not listed are left as they are;
//  - the adjoint of that activation, gathered back from the feature-row gradients.
#include "common.h"

namespace rmesh {
namespace {

constexpr float kLogMaxDensity = 46.0517018598809f;  // log(1e20) rounded to float

__global__ __launch_bounds__(256) void field_points_kernel(const int4* __restrict__ indices,
                                                           const float* __restrict__ verts,
                                                           const int32_t* __restrict__ tet_ids, int64_t n,
                                                           float scale, float* __restrict__ out) {
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int4 q = indices[tet_ids[i]];
    const int vid[4] = {q.x, q.y, q.z, q.w};
    float c[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int k = 0; k < 3; ++k) c[k] = c[k] + verts[3 * (int64_t)vid[v] + k];
    float z[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) z[k] = (c[k] * 0.25f) / scale;
    const float zz0 = z[0] * z[0], zz1 = z[1] * z[1], zz2 = z[2] * z[2];
    const float mag_sq = fmaxf((zz0 + zz1) + zz2, 1.1920928955078125e-07f);
    const float f = mag_sq <= 1.0f ? 1.0f : (2.0f * sqrtf(mag_sq) - 1.0f) / mag_sq;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float zc = mag_sq <= 1.0f ? z[k] : f * z[k];
        out[3 * i + k] = (zc / 2.0f + 1.0f) / 2.0f;
    }
}

// Four consecutive elements of the [n, F] head output per thread (one 16-byte load or store of
// the dense side); F is a template parameter so that splitting an element index is a multiply.
template <int F>
__global__ __launch_bounds__(256) void field_scatter_fwd_kernel(const float* __restrict__ raw,
                                                                const int32_t* __restrict__ tet_ids, int64_t n,
                                                                float* __restrict__ feats) {
    const int64_t nf = n * F;
    const int64_t e0 = 4 * ((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (e0 >= nf) return;
    float v[4];
    if (e0 + 3 < nf) {
        const float4 q = *reinterpret_cast<const float4*>(raw + e0);
        v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) v[k] = e0 + k < nf ? raw[e0 + k] : 0.f;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int64_t e = e0 + k;
        if (e >= nf) break;
        const int64_t i = e / F;
        const int f = (int)(e - i * F);
        const float o = f == 0 ? expf(fminf(v[k], kLogMaxDensity)) : (f < 4 ? v[k] + 0.5f : v[k]);
        feats[(int64_t)tet_ids[i] * F + f] = o;
    }
}

template <int F>
__global__ __launch_bounds__(256) void field_scatter_bwd_kernel(const float* __restrict__ raw,
                                                                const int32_t* __restrict__ tet_ids, int64_t n,
                                                                const float* __restrict__ gfeats,
                                                                float* __restrict__ graw) {
    const int64_t nf = n * F;
    const int64_t e0 = 4 * ((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (e0 >= nf) return;
    float o[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int64_t e = e0 + k;
        o[k] = 0.f;
        if (e < nf) {
            const int64_t i = e / F;
            const int f = (int)(e - i * F);
            const float g = gfeats[(int64_t)tet_ids[i] * F + f];
            if (f == 0) {
                const float vr = raw[e];
                o[k] = vr <= kLogMaxDensity ? g * expf(vr) : 0.f;
            } else {
                o[k] = g;
            }
        }
    }
    if (e0 + 3 < nf) {
        *reinterpret_cast<float4*>(graw + e0) = make_float4(o[0], o[1], o[2], o[3]);
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (e0 + k < nf) graw[e0 + k] = o[k];
    }
}

}  // namespace
}  // namespace rmesh

using namespace rmesh;

extern "C" int rm_field_points(const int32_t* indices_dev, const float* vertices_dev, const int32_t* tet_ids_dev,
                               int64_t n, float scale, float* points_dev, rm_stream_t stream) {
    RM_CHECK_ARG(n >= 0, "rm_field_points: negative n");
    RM_CHECK_ARG(scale > 0.f, "rm_field_points: scale must be positive");
    if (n == 0) return RM_OK;
    RM_CHECK_ARG(indices_dev && vertices_dev && tet_ids_dev && points_dev, "rm_field_points: null buffer");
    RM_CHECK_ARG(((uintptr_t)indices_dev & 15) == 0, "rm_field_points: indices must be 16-byte aligned");
    hipLaunchKernelGGL(field_points_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, as_stream(stream),
                       (const int4*)indices_dev, vertices_dev, tet_ids_dev, n, scale, points_dev);
    RM_CHECK_LAUNCH("field_points_kernel");
    return RM_OK;
}

extern "C" int rm_field_scatter_fwd(const float* raw_dev, const int32_t* tet_ids_dev, int64_t n, int32_t n_features,
                                    float* features_dev, rm_stream_t stream) {
    RM_CHECK_ARG(n >= 0, "rm_field_scatter_fwd: negative n");
    RM_CHECK_ARG(n_features == 7 || n_features == 13, "rm_field_scatter_fwd: n_features must be 7 or 13");
    if (n == 0) return RM_OK;
    RM_CHECK_ARG(raw_dev && tet_ids_dev && features_dev, "rm_field_scatter_fwd: null buffer");
    RM_CHECK_ARG(((uintptr_t)raw_dev & 15) == 0, "rm_field_scatter_fwd: raw must be 16-byte aligned");
    if (n_features == 7)
        hipLaunchKernelGGL(field_scatter_fwd_kernel<7>, dim3((unsigned)ceil_div(n * 7, 1024)), dim3(256), 0,
                           as_stream(stream), raw_dev, tet_ids_dev, n, features_dev);
    else
        hipLaunchKernelGGL(field_scatter_fwd_kernel<13>, dim3((unsigned)ceil_div(n * 13, 1024)), dim3(256), 0,
                           as_stream(stream), raw_dev, tet_ids_dev, n, features_dev);
    RM_CHECK_LAUNCH("field_scatter_fwd_kernel");
    return RM_OK;
}

extern "C" int rm_field_scatter_bwd(const float* raw_dev, const int32_t* tet_ids_dev, int64_t n, int32_t n_features,
                                    const float* grad_features_dev, float* grad_raw_dev, rm_stream_t stream) {
    RM_CHECK_ARG(n >= 0, "rm_field_scatter_bwd: negative n");
    RM_CHECK_ARG(n_features == 7 || n_features == 13, "rm_field_scatter_bwd: n_features must be 7 or 13");
    if (n == 0) return RM_OK;
    RM_CHECK_ARG(raw_dev && tet_ids_dev && grad_features_dev && grad_raw_dev, "rm_field_scatter_bwd: null buffer");
    RM_CHECK_ARG(((uintptr_t)grad_raw_dev & 15) == 0, "rm_field_scatter_bwd: grad_raw must be 16-byte aligned");
    if (n_features == 7)
        hipLaunchKernelGGL(field_scatter_bwd_kernel<7>, dim3((unsigned)ceil_div(n * 7, 1024)), dim3(256), 0,
                           as_stream(stream), raw_dev, tet_ids_dev, n, grad_features_dev, grad_raw_dev);
    else
        hipLaunchKernelGGL(field_scatter_bwd_kernel<13>, dim3((unsigned)ceil_div(n * 13, 1024)), dim3(256), 0,
                           as_stream(stream), raw_dev, tet_ids_dev, n, grad_features_dev, grad_raw_dev);
    RM_CHECK_LAUNCH("field_scatter_bwd_kernel");
    return RM_OK;
}
