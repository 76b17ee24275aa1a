// Per-tet projection and tile binning (forward) and its FP64 adjoint (backward).
//
// Forward semantics follow delaunay_rasterization/internal/slang/vertex_shader.slang:20-89:
// cull planar tets (intersect.slang:76-89), transform to view space (camera.slang:104-108),
// conservative screen extent with near-plane clipping and edge sampling (camera.slang:416-508),
// cull if the extent lies behind the camera, FP64 circumsphere power as the sort depth
// (stable_power.slang:18-85), tile rectangle, cull if empty.
//
// Every float operation rounds once in source order (this file is compiled with
// -ffp-contract=off and correctly rounded division/sqrt), so tile counts, rectangles and depth
// keys are bit-identical to a strict IEEE single-precision evaluation of the same expressions.
#include "common.h"

#pragma clang fp contract(off)

namespace rmesh {
namespace {

struct F3 {
    float x, y, z;
};
__device__ inline F3 sub(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline F3 cross(F3 a, F3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

constexpr float kTiny = 1.0754944e-20f;  // safe-math.slang:2

__device__ inline float safe_clip(float v, float lo, float hi) { return fmaxf(fminf(v, hi), lo); }
__device__ inline float safe_div(float a, float b) {
    if (fabsf(b) < kTiny) return safe_clip(a / kTiny, -1e20f, 1e20f);
    return safe_clip(a / b, -1e20f, 1e20f);
}

__device__ inline bool is_planar(const F3 v[4]) {
    F3 n = cross(sub(v[1], v[0]), sub(v[2], v[0]));
    float len = sqrtf(dot(n, n));
    F3 nn = {n.x / len, n.y / len, n.z / len};
    float d = dot(nn, sub(v[3], v[0]));
    return fabsf(d) < 1e-3f;
}

__device__ inline F3 to_view(const CamArgs& c, F3 p) {
    float h[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const float* m = c.w2v + 4 * i;
        h[i] = m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3] * 1.0f;
    }
    return {safe_div(h[0], h[3]), safe_div(h[1], h[3]), safe_div(h[2], h[3])};
}

__device__ inline void distort(const CamArgs& c, float& x, float& y) {
    const float* k = c.dist;
    if ((k[0] * k[0] + k[1] * k[1] + k[2] * k[2] + k[3] * k[3]) < 1e-9f) return;
    float r2 = x * x + y * y;
    float radial = r2 * (k[0] + r2 * k[1]);
    float tx = 2.0f * k[2] * x * y + k[3] * (r2 + 2.0f * x * x);
    float ty = 2.0f * k[3] * x * y + k[2] * (r2 + 2.0f * y * y);
    float nx = x + x * radial + tx;
    float ny = y + y * radial + ty;
    x = nx;
    y = ny;
}

__device__ inline F3 to_image(const CamArgs& c, F3 p) {
    float x, y;
    if (c.type == 1) {
        float len = sqrtf(dot(p, p));
        F3 d = {p.x / len, p.y / len, p.z / len};
        float z = fminf(fmaxf(d.z, 0.0f), 1.0f);
        float theta = acosf(z);
        float s = sinf(theta);
        float ratio = (s < 1e-6f) ? 1.0f : theta / s;
        x = d.x * ratio;
        y = d.y * ratio;
    } else {
        x = p.x / p.z;
        y = p.y / p.z;
    }
    distort(c, x, y);
    return {x * c.K[0] + c.K[2], y * c.K[4] + c.K[5], p.z};
}

__device__ inline F3 lerp(F3 a, F3 b, float t) {
    float u = 1.0f - t;
    return {a.x * u + b.x * t, a.y * u + b.y * t, a.z * u + b.z * t};
}

__device__ inline void edge_bounds(const CamArgs& c, int n, F3 a, F3 b, F3& mn, F3& mx) {
    for (int i = 0; i < n; ++i) {
        float t = (float)i / (float)(n - 1);
        F3 px = to_image(c, lerp(a, b, t));
        mn = {fminf(mn.x, px.x), fminf(mn.y, px.y), fminf(mn.z, px.z)};
        mx = {fmaxf(mx.x, px.x), fmaxf(mx.y, px.y), fmaxf(mx.z, px.z)};
    }
}

__device__ inline void find_extent(const CamArgs& c, const F3 v[4], F3& mn, F3& mx) {
    const int faces[4][3] = {{0, 1, 2}, {0, 2, 3}, {0, 3, 1}, {1, 3, 2}};
    const float eps = 1e-4f;
    const float* k = c.dist;
    int nsamp = 2;
    if (c.type == 1) nsamp = 4;
    else if ((k[0] * k[0] + k[1] * k[1] + k[2] * k[2] + k[3] * k[3]) > 1e-9f) nsamp = 3;
    mn = {1e20f, 1e20f, 1e20f};
    mx = {-1e20f, -1e20f, -1e20f};
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        F3 fin[4];
        int cnt = 0;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            F3 a = v[faces[f][j]];
            F3 b = v[faces[f][(j + 1) % 3]];
            bool ain = a.z > eps, bin = b.z > eps;
            if (ain) {
                bool found = false;
                for (int q = 0; q < cnt; ++q)
                    if (fin[q].x == a.x && fin[q].y == a.y && fin[q].z == a.z) found = true;
                if (!found) fin[cnt++] = a;
            }
            if (ain != bin) {
                float t = (eps - a.z) / (b.z - a.z);
                fin[cnt++] = lerp(a, b, t);
            }
        }
        if (cnt > 1)
            for (int j = 0; j < cnt; ++j) edge_bounds(c, nsamp, fin[j], fin[(j + 1) % cnt], mn, mx);
    }
}

struct D3 {
    double x, y, z;
};
__device__ inline D3 dsub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline D3 dadd(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline D3 dscale(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ inline double ddot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline D3 dcross(D3 a, D3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ inline double dlen(D3 a) { return sqrt(ddot(a, a)); }
__device__ inline D3 todouble(F3 a) { return {(double)a.x, (double)a.y, (double)a.z}; }

// stable_power.slang:18-85
__device__ double power_of_circumsphere(D3 A, D3 B, D3 C, D3 D, D3 P, D3& center) {
    D3 a = dsub(B, A), b = dsub(C, A), c = dsub(D, A);
    double aa = ddot(a, a), bb = ddot(b, b), cc = ddot(c, c);
    D3 bc = dcross(b, c), ca = dcross(c, a), ab = dcross(a, b);
    double den = 2.0 * ddot(a, bc);
    if (fabs(den) < 1e-10) {
        D3 pts[4] = {A, B, C, D};
        double md = 0.0;
        for (int i = 0; i < 4; ++i)
            for (int j = i + 1; j < 4; ++j) md = fmax(md, dlen(dsub(pts[j], pts[i])));
        if (md < 1e-7) {
            center = {0.25 * (A.x + B.x + C.x + D.x), 0.25 * (A.y + B.y + C.y + D.y),
                      0.25 * (A.z + B.z + C.z + D.z)};
            double r = 1e300;
            for (int k = 0; k < 4; ++k) r = fmin(r, dlen(dsub(pts[k], center)));
            D3 diff = dsub(center, P);
            return ddot(diff, diff) - r * r;
        }
        return -1e20;
    }
    D3 rel = {(aa * bc.x + bb * ca.x + cc * ab.x) / den, (aa * bc.y + bb * ca.y + cc * ab.y) / den,
              (aa * bc.z + bb * ca.z + cc * ab.z) / den};
    center = dadd(A, rel);
    double radius = dlen(rel);
    D3 diff = dsub(center, P);
    return ddot(diff, diff) - radius * radius;
}

__device__ inline F3 load_vert(const float* __restrict__ v, int i) {
    const float* p = v + 3 * (int64_t)i;
    return {p[0], p[1], p[2]};
}

// One tet: tile count, rect, clamped depth and circumcenter (zeros when culled).
__device__ inline void project_one(const CamArgs& cam, int4 q, const float* __restrict__ verts, int32_t& n,
                                   int4& r, float& depth_out, F3& ctr_out) {
    F3 v[4] = {load_vert(verts, q.x), load_vert(verts, q.y), load_vert(verts, q.z), load_vert(verts, q.w)};
    n = 0;
    r = {0, 0, 0, 0};
    depth_out = 0.f;
    ctr_out = {0.f, 0.f, 0.f};
    if (is_planar(v)) return;
    F3 vc[4] = {to_view(cam, v[0]), to_view(cam, v[1]), to_view(cam, v[2]), to_view(cam, v[3])};
    F3 mn, mx;
    find_extent(cam, vc, mn, mx);
    if (mx.z <= 0.0f) return;
    D3 P = {(double)cam.cam_pos[0], (double)cam.cam_pos[1], (double)cam.cam_pos[2]};
    D3 ctr = {0, 0, 0};
    float depth = (float)power_of_circumsphere(todouble(v[0]), todouble(v[1]), todouble(v[2]), todouble(v[3]), P, ctr);
    if (depth < -1e19f) return;
    float gw = (float)cam.gw, gh = (float)cam.gh;
    int4 rr;
    rr.x = (int32_t)floorf(fmaxf(0.0f, fminf(gw, mn.x / (float)cam.tw)));
    rr.y = (int32_t)floorf(fmaxf(0.0f, fminf(gh, mn.y / (float)cam.th)));
    rr.z = (int32_t)ceilf(fmaxf(0.0f, fminf(gw, mx.x / (float)cam.tw)));
    rr.w = (int32_t)ceilf(fmaxf(0.0f, fminf(gh, mx.y / (float)cam.th)));
    int32_t nt = (rr.z - rr.x) * (rr.w - rr.y);
    if (nt == 0) return;
    n = nt;
    r = rr;
    depth_out = fminf(fmaxf(depth, -1e20f), 1e20f);
    ctr_out = {(float)ctr.x, (float)ctr.y, (float)ctr.z};
}

__global__ __launch_bounds__(256) void project_fwd_kernel(
    CamArgs cam, const int4* __restrict__ indices, const float* __restrict__ verts, int64_t T,
    int32_t* __restrict__ tiles, int4* __restrict__ rect, float* __restrict__ vs,
    float* __restrict__ cc) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    int32_t n;
    int4 r;
    float depth;
    F3 ctr;
    project_one(cam, indices[t], verts, n, r, depth, ctr);
    tiles[t] = n;
    rect[t] = r;
    vs[3 * t + 0] = 0.f;
    vs[3 * t + 1] = 0.f;
    vs[3 * t + 2] = depth;
    cc[3 * t + 0] = ctr.x;
    cc[3 * t + 1] = ctr.y;
    cc[3 * t + 2] = ctr.z;
}

// Batched projection: lane g projects global tet g of the batch with its view's camera. Views
// start at multiples of the workgroup size, so a workgroup's tets belong to one view, found once
// per workgroup: the camera is read with scalar loads rather than once per lane.
static_assert(kPlanTetAlign == 256, "one projection workgroup per aligned block of a view's tets");
__global__ __launch_bounds__(256) void batch_project_kernel(const void* __restrict__ plan, int64_t total,
                                                            int32_t* __restrict__ tiles, int4* __restrict__ rect,
                                                            float* __restrict__ depth, uint2* __restrict__ tet_blocks) {
    const int64_t g0 = (int64_t)blockIdx.x * blockDim.x;
    const int64_t g = g0 + threadIdx.x;
    if (g >= total) return;
    const PlanHeader* h = static_cast<const PlanHeader*>(plan);
    const ViewRec* views = plan_views(plan);
    int v = 0, hi = h->n_views - 1;  // last view with tet_offset <= g0 (uniform)
    while (v < hi) {
        const int mid = (v + hi + 1) >> 1;
        if (views[mid].tet_offset <= g0) v = mid;
        else hi = mid - 1;
    }
    const ViewRec& r = views[v];
    const int64_t t = g - r.tet_offset;
    int32_t n = 0;
    int4 rc = make_int4(0, 0, 0, 0);
    float d = 0.f;
    // slots past the view's tets (alignment gaps, or a launch sized for a capacity) stay empty
    if (g < h->total_tets && t < r.n_tets) {
        F3 ctr;
        project_one(r.cam, uniform_global_ptr(reinterpret_cast<const int4*>(r.indices))[t],
                    uniform_global_ptr(r.verts), n, rc, d, ctr);
    }
    tiles[g] = n;
    rect[g] = rc;
    depth[g] = d;
    // the view's binning blocks: first global block and blocks per row of its tile grid
    tet_blocks[g] = make_uint2((uint32_t)r.bin_offset, (uint32_t)bin_blocks_per_row(r.cam.gw));
}

// Adjoint of (clamped float(power), float3(center)) w.r.t. the four vertices, in FP64.
// Tets the forward culled produced constant zeros and receive no gradient.
__global__ __launch_bounds__(256) void project_bwd_kernel(
    CamArgs cam, const int4* __restrict__ indices, const float* __restrict__ verts, int64_t T,
    const int32_t* __restrict__ tiles, const float* __restrict__ gvs,
    const float* __restrict__ gcc, float* __restrict__ gverts) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    if (tiles[t] <= 0) return;
    int4 q = indices[t];
    D3 V[4] = {todouble(load_vert(verts, q.x)), todouble(load_vert(verts, q.y)),
               todouble(load_vert(verts, q.z)), todouble(load_vert(verts, q.w))};
    D3 P = {(double)cam.cam_pos[0], (double)cam.cam_pos[1], (double)cam.cam_pos[2]};
    double g_pow = gvs ? (double)gvs[3 * t + 2] : 0.0;
    D3 g_ctr = gcc ? D3{(double)gcc[3 * t], (double)gcc[3 * t + 1], (double)gcc[3 * t + 2]}
                   : D3{0, 0, 0};
    D3 A = V[0];
    D3 a = dsub(V[1], A), b = dsub(V[2], A), c = dsub(V[3], A);
    double aa = ddot(a, a), bb = ddot(b, b), cc2 = ddot(c, c);
    D3 bc = dcross(b, c), ca = dcross(c, a), ab = dcross(a, b);
    double den = 2.0 * ddot(a, bc);
    D3 gV[4] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    if (fabs(den) < 1e-10) {
        // small-tet branch: center = mean, radius = min distance to a corner
        D3 ctr = dscale(dadd(dadd(V[0], V[1]), dadd(V[2], V[3])), 0.25);
        int kmin = 0;
        double r = 1e300;
        for (int k = 0; k < 4; ++k) {
            double l = dlen(dsub(V[k], ctr));
            if (l < r) { r = l; kmin = k; }
        }
        D3 diff = dsub(ctr, P);
        D3 gct = dadd(g_ctr, dscale(diff, 2.0 * g_pow));
        double gr = -2.0 * r * g_pow;
        D3 u = r > 0 ? dscale(dsub(V[kmin], ctr), 1.0 / r) : D3{0, 0, 0};
        gV[kmin] = dadd(gV[kmin], dscale(u, gr));
        gct = dsub(gct, dscale(u, gr));
        for (int k = 0; k < 4; ++k) gV[k] = dadd(gV[k], dscale(gct, 0.25));
    } else {
        D3 num = {aa * bc.x + bb * ca.x + cc2 * ab.x, aa * bc.y + bb * ca.y + cc2 * ab.y,
                  aa * bc.z + bb * ca.z + cc2 * ab.z};
        D3 rel = dscale(num, 1.0 / den);
        D3 ctr = dadd(A, rel);
        double radius = dlen(rel);
        D3 diff = dsub(ctr, P);
        D3 gct = dadd(g_ctr, dscale(diff, 2.0 * g_pow));
        double grr = -2.0 * radius * g_pow;
        D3 grel = dadd(gct, radius > 0 ? dscale(rel, grr / radius) : D3{0, 0, 0});
        D3 gnum = dscale(grel, 1.0 / den);
        double gden = -ddot(grel, rel) / den;
        double gaa = ddot(gnum, bc), gbb = ddot(gnum, ca), gcc = ddot(gnum, ab);
        D3 gbc = dadd(dscale(gnum, aa), dscale(a, 2.0 * gden));
        D3 gca = dscale(gnum, bb), gab = dscale(gnum, cc2);
        D3 ga = dadd(dscale(a, 2.0 * gaa), dscale(bc, 2.0 * gden));
        D3 gb = dscale(b, 2.0 * gbb);
        D3 gc = dscale(c, 2.0 * gcc);
        gb = dadd(gb, dcross(c, gbc));
        gc = dadd(gc, dcross(gbc, b));
        gc = dadd(gc, dcross(a, gca));
        ga = dadd(ga, dcross(gca, c));
        ga = dadd(ga, dcross(b, gab));
        gb = dadd(gb, dcross(gab, a));
        gV[0] = dsub(gct, dadd(dadd(ga, gb), gc));
        gV[1] = ga;
        gV[2] = gb;
        gV[3] = gc;
    }
    int vi[4] = {q.x, q.y, q.z, q.w};
    for (int k = 0; k < 4; ++k) {
        float* g = gverts + 3 * (int64_t)vi[k];
        unsafeAtomicAdd(g + 0, (float)gV[k].x);
        unsafeAtomicAdd(g + 1, (float)gV[k].y);
        unsafeAtomicAdd(g + 2, (float)gV[k].z);
    }
}

}  // namespace

hipError_t batch_project(const void* plan_dev, const PlanHeader& h, int32_t* tiles, int4* rect, float* depth,
                         uint2* tet_blocks, hipStream_t st) {
    if (h.total_tets <= 0) return hipSuccess;
    hipLaunchKernelGGL(batch_project_kernel, dim3((unsigned)ceil_div(h.total_tets, 256)), dim3(256), 0, st, plan_dev,
                       h.total_tets, tiles, rect, depth, tet_blocks);
    return hipGetLastError();
}

}  // namespace rmesh

using namespace rmesh;

extern "C" int rm_project_fwd(const rm_camera* cam, const int32_t* indices_dev,
                              const float* vertices_dev, int64_t n_tets, int64_t n_verts,
                              int32_t* tiles_touched_dev, int32_t* rect_dev, float* vs_dev,
                              float* circumcenter_dev, rm_stream_t stream) {
    RM_CHECK_ARG(cam != nullptr, "rm_project_fwd: camera is null");
    RM_CHECK_ARG(n_tets >= 0 && n_verts >= 0, "rm_project_fwd: negative size");
    RM_CHECK_ARG(cam->tile_width > 0 && cam->tile_height > 0, "rm_project_fwd: bad tile size");
    if (n_tets == 0) return RM_OK;
    RM_CHECK_ARG(indices_dev && vertices_dev && tiles_touched_dev && rect_dev && vs_dev &&
                     circumcenter_dev,
                 "rm_project_fwd: null buffer");
    RM_CHECK_ARG(((uintptr_t)indices_dev & 15) == 0 && ((uintptr_t)rect_dev & 15) == 0,
                 "rm_project_fwd: indices and rect must be 16-byte aligned");
    CamArgs a = make_cam_args(*cam);
    hipLaunchKernelGGL(project_fwd_kernel, dim3((unsigned)ceil_div(n_tets, 256)), dim3(256), 0,
                       as_stream(stream), a, (const int4*)indices_dev, vertices_dev, n_tets,
                       tiles_touched_dev, (int4*)rect_dev, vs_dev, circumcenter_dev);
    RM_CHECK_LAUNCH("project_fwd_kernel");
    return RM_OK;
}

extern "C" int rm_project_bwd(const rm_camera* cam, const int32_t* indices_dev,
                              const float* vertices_dev, int64_t n_tets, int64_t n_verts,
                              const int32_t* tiles_touched_dev, const float* grad_vs_dev,
                              const float* grad_circumcenter_dev, float* grad_vertices_dev,
                              rm_stream_t stream) {
    RM_CHECK_ARG(cam != nullptr, "rm_project_bwd: camera is null");
    RM_CHECK_ARG(n_tets >= 0 && n_verts >= 0, "rm_project_bwd: negative size");
    if (n_tets == 0) return RM_OK;
    RM_CHECK_ARG(indices_dev && vertices_dev && tiles_touched_dev && grad_vertices_dev,
                 "rm_project_bwd: null buffer");
    RM_CHECK_ARG(((uintptr_t)indices_dev & 15) == 0, "rm_project_bwd: indices must be 16-byte aligned");
    CamArgs a = make_cam_args(*cam);
    hipLaunchKernelGGL(project_bwd_kernel, dim3((unsigned)ceil_div(n_tets, 256)), dim3(256), 0,
                       as_stream(stream), a, (const int4*)indices_dev, vertices_dev, n_tets,
                       tiles_touched_dev, grad_vs_dev, grad_circumcenter_dev, grad_vertices_dev);
    RM_CHECK_LAUNCH("project_bwd_kernel");
    return RM_OK;
}
