/*
 * rmesh_oracle.c -- CPU oracle for the tetrahedral tile rasterizer.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker: tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it; the product (radiance_meshes_amd) never does.
 *
 * Contents, each restating the reference's algorithm (paths relative to the reference repo):
 *   orc_project        FP32 restatement of vertex_shader (delaunay_rasterization/internal/slang/
 *                      vertex_shader.slang:20-89, camera.slang:104-108,357-508, intersect.slang:76-89)
 *                      with the circumsphere power in FP64 (stable_power.slang:18-85). Built with
 *                      -ffp-contract=off: every float op rounds once, in source order.
 *   orc_bin            keys (tile_shader.slang:18-97), stable sort (sort_by_keys.cu:36-48 sorts
 *                      stably by key), tile ranges (tile_shader.slang:99-125).
 *   orc_composite_fwd  FP64 exact per-ray segment compositor: get_ray (camera.slang:298-346),
 *                      ray_tetrahedron_intersect_fused (intersect.slang:393-624),
 *                      integrate_tetra_segment (interp_version.slang:21-98), update_alpha_out
 *                      and the termination rule (alphablend_shader_interp.slang:97-126,153-183,
 *                      194-271), evaluated in double on the same per-tile lists.
 *   (the *_tiles variants restrict either pass to a list of tiles)
 *   orc_composite_bwd  FP64 adjoint of orc_composite_fwd with the reference's backward structure
 *                      (bwd_alpha_blend, alphablend_shader_interp.slang:273-369): back-to-front
 *                      walk from the stored final state, offsets >= n_contrib skipped.
 *   orc_quadrature     FP64 midpoint quadrature over all tets per sample, the method of
 *                      jaxutil/tetra_quad.py:74-99,142-224 with the colour read as
 *                      base + (gx,gy,gz).(p - v0) (tetra_quad.py:192-195 reads indices 4..6 of a
 *                      6-wide array; JAX clamps the gather, so it uses (gy,gz,gz)).
 *   orc_project_bwd    FP64 adjoint of the circumsphere power / circumcenter map.
 *   orc_tet_err        FP64 per-tet error votes of calc_tet_err (alphablend_shader_interp.slang:
 *                      433-599, render_err.py:138-150).
 *
 * Pinned by tests/test_oracle_*.py against closed forms, brute force and finite differences.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct orc_camera {
    float w2v[16];
    float K[9];
    float cam_pos[3];
    float dist[4];
    float min_t, fovy, fovx;
    int32_t H, W, gh, gw, th, tw, camera_type, pad_;
} orc_camera;

/* ------------------------------------------------------------------ FP32 projection ---- */

static const float TINY_VAL = 1.0754944e-20f; /* safe-math.slang:2 */

static float clipf(float v, float lo, float hi) { return fmaxf(lo, fminf(hi, v)); }
static float safe_clipf(float v, float lo, float hi) { return fmaxf(fminf(v, hi), lo); }
static float safe_divf(float a, float b) {
    if (fabsf(b) < TINY_VAL) return safe_clipf(a / TINY_VAL, -1e20f, 1e20f);
    return safe_clipf(a / b, -1e20f, 1e20f);
}
static float dot3f(const float* a, const float* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static void sub3f(const float* a, const float* b, float* r) { r[0] = a[0] - b[0]; r[1] = a[1] - b[1]; r[2] = a[2] - b[2]; }
static void cross3f(const float* a, const float* b, float* r) {
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}

/* intersect.slang:76-89; normalize(x) taken as x / sqrt(dot(x, x)). */
static int is_planar_f(const float v[4][3]) {
    float e1[3], e2[3], n[3], d3[3];
    sub3f(v[1], v[0], e1);
    sub3f(v[2], v[0], e2);
    cross3f(e1, e2, n);
    float len = sqrtf(dot3f(n, n));
    float nn[3] = {n[0] / len, n[1] / len, n[2] / len};
    sub3f(v[3], v[0], d3);
    float dist = dot3f(nn, d3);
    return fabsf(dist) < 1e-3f;
}

/* camera.slang:104-108 */
static void point2cam(const orc_camera* c, const float* p, float* r) {
    float h[4];
    for (int i = 0; i < 4; ++i) {
        const float* m = c->w2v + 4 * i;
        h[i] = m[0] * p[0] + m[1] * p[1] + m[2] * p[2] + m[3] * 1.0f;
    }
    r[0] = safe_divf(h[0], h[3]);
    r[1] = safe_divf(h[1], h[3]);
    r[2] = safe_divf(h[2], h[3]);
}

static int cam_distorted(const orc_camera* c) {
    const float* k = c->dist;
    return (k[0] * k[0] + k[1] * k[1] + k[2] * k[2] + k[3] * k[3]) > 1e-9f;
}

/* camera.slang:357-377 */
static void distort_f(const orc_camera* c, float* p) {
    const float* k = c->dist;
    if ((k[0] * k[0] + k[1] * k[1] + k[2] * k[2] + k[3] * k[3]) < 1e-9f) return;
    float x = p[0], y = p[1];
    float r2 = x * x + y * y;
    float radial = r2 * (k[0] + r2 * k[1]);
    float tx = 2.0f * k[2] * x * y + k[3] * (r2 + 2.0f * x * x);
    float ty = 2.0f * k[3] * x * y + k[2] * (r2 + 2.0f * y * y);
    p[0] = x + x * radial + tx;
    p[1] = y + y * radial + ty;
}

/* camera.slang:382-412 */
static void cam2image_f(const orc_camera* c, const float* pc, float* r) {
    float p[2];
    if (c->camera_type == 1) {
        float len = sqrtf(dot3f(pc, pc));
        float d[3] = {pc[0] / len, pc[1] / len, pc[2] / len};
        float z = fminf(fmaxf(d[2], 0.0f), 1.0f);
        float theta = acosf(z);
        float s = sinf(theta);
        float ratio = (s < 1e-6f) ? 1.0f : theta / s;
        p[0] = d[0] * ratio;
        p[1] = d[1] * ratio;
    } else {
        p[0] = pc[0] / pc[2];
        p[1] = pc[1] / pc[2];
    }
    distort_f(c, p);
    r[0] = p[0] * c->K[0] + c->K[2];
    r[1] = p[1] * c->K[4] + c->K[5];
    r[2] = pc[2];
}

/* lerp(a, b, t) taken as a * (1 - t) + b * t (exact at both endpoints). */
static void lerp3f(const float* a, const float* b, float t, float* r) {
    float u = 1.0f - t;
    for (int i = 0; i < 3; ++i) r[i] = a[i] * u + b[i] * t;
}

/* camera.slang:416-439 */
static void bounds_edge(const orc_camera* c, const float* a, const float* b, float* mn, float* mx) {
    int n = 2;
    if (c->camera_type == 1) n = 4;
    else if (cam_distorted(c)) n = 3;
    for (int i = 0; i < n; ++i) {
        float t = (float)i / (float)(n - 1);
        float v[3], px[3];
        lerp3f(a, b, t, v);
        cam2image_f(c, v, px);
        for (int k = 0; k < 3; ++k) {
            mn[k] = fminf(mn[k], px[k]);
            mx[k] = fmaxf(mx[k], px[k]);
        }
    }
}

/* camera.slang:443-508 */
static void find_extent_f(const orc_camera* c, const float v[4][3], float* mn, float* mx) {
    static const int faces[4][3] = {{0, 1, 2}, {0, 2, 3}, {0, 3, 1}, {1, 3, 2}};
    const float eps = 1e-4f;
    for (int k = 0; k < 3; ++k) { mn[k] = 1e20f; mx[k] = -1e20f; }
    for (int f = 0; f < 4; ++f) {
        float fin[4][3];
        int cnt = 0;
        for (int j = 0; j < 3; ++j) {
            const float* a = v[faces[f][j]];
            const float* b = v[faces[f][(j + 1) % 3]];
            int ain = a[2] > eps, bin = b[2] > eps;
            if (ain) {
                int found = 0;
                for (int q = 0; q < cnt; ++q)
                    if (fin[q][0] == a[0] && fin[q][1] == a[1] && fin[q][2] == a[2]) found = 1;
                if (!found) { memcpy(fin[cnt], a, sizeof(float) * 3); cnt++; }
            }
            if (ain != bin) {
                float t = (eps - a[2]) / (b[2] - a[2]);
                lerp3f(a, b, t, fin[cnt]);
                cnt++;
            }
        }
        if (cnt > 1)
            for (int j = 0; j < cnt; ++j) bounds_edge(c, fin[j], fin[(j + 1) % cnt], mn, mx);
    }
}

static double dot3d(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static void cross3d(const double* a, const double* b, double* r) {
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}
static double len3d(const double* a) { return sqrt(dot3d(a, a)); }

/* stable_power.slang:18-85 (FP64). Returns the power; writes the circumcenter. */
double orc_power_of_circumsphere(const double* A, const double* B, const double* C, const double* D,
                                 const double* P, double* center) {
    double a[3], b[3], c[3];
    for (int i = 0; i < 3; ++i) { a[i] = B[i] - A[i]; b[i] = C[i] - A[i]; c[i] = D[i] - A[i]; }
    double aa = dot3d(a, a), bb = dot3d(b, b), cc = dot3d(c, c);
    double bc[3], ca[3], ab[3];
    cross3d(b, c, bc);
    cross3d(c, a, ca);
    cross3d(a, b, ab);
    double den = 2.0 * dot3d(a, bc);
    double radius;
    if (fabs(den) < 1e-10) {
        const double* pts[4] = {A, B, C, D};
        double md = 0.0;
        for (int i = 0; i < 4; ++i)
            for (int j = i + 1; j < 4; ++j) {
                double d[3] = {pts[j][0] - pts[i][0], pts[j][1] - pts[i][1], pts[j][2] - pts[i][2]};
                double l = len3d(d);
                if (l > md) md = l;
            }
        if (md < 1e-7) {
            for (int i = 0; i < 3; ++i) center[i] = 0.25 * (A[i] + B[i] + C[i] + D[i]);
            double r = 1e300;
            for (int k = 0; k < 4; ++k) {
                double d[3] = {pts[k][0] - center[0], pts[k][1] - center[1], pts[k][2] - center[2]};
                double l = len3d(d);
                if (l < r) r = l;
            }
            double diff[3] = {center[0] - P[0], center[1] - P[1], center[2] - P[2]};
            return dot3d(diff, diff) - r * r;
        }
        return -1e20;
    }
    double rel[3];
    for (int i = 0; i < 3; ++i) rel[i] = (aa * bc[i] + bb * ca[i] + cc * ab[i]) / den;
    for (int i = 0; i < 3; ++i) center[i] = A[i] + rel[i];
    radius = len3d(rel);
    double diff[3] = {center[0] - P[0], center[1] - P[1], center[2] - P[2]};
    return dot3d(diff, diff) - radius * radius;
}

void orc_project(const orc_camera* c, const int32_t* idx, const float* verts, int64_t T,
                 int32_t* tiles, int32_t* rect, float* vs, float* cc) {
#pragma omp parallel for schedule(static, 4096)
    for (int64_t t = 0; t < T; ++t) {
        tiles[t] = 0;
        for (int k = 0; k < 4; ++k) rect[4 * t + k] = 0;
        for (int k = 0; k < 3; ++k) { vs[3 * t + k] = 0.f; cc[3 * t + k] = 0.f; }
        float v[4][3], vc[4][3];
        for (int i = 0; i < 4; ++i) memcpy(v[i], verts + 3 * (int64_t)idx[4 * t + i], sizeof(float) * 3);
        if (is_planar_f(v)) continue;
        for (int i = 0; i < 4; ++i) point2cam(c, v[i], vc[i]);
        float mn[3], mx[3];
        find_extent_f(c, vc, mn, mx);
        if (mx[2] <= 0.0f) continue;
        double A[3], B[3], C[3], D[3], P[3], ctr[3];
        for (int i = 0; i < 3; ++i) {
            A[i] = v[0][i]; B[i] = v[1][i]; C[i] = v[2][i]; D[i] = v[3][i]; P[i] = c->cam_pos[i];
        }
        float depth = (float)orc_power_of_circumsphere(A, B, C, D, P, ctr);
        if (depth < -1e19f) continue;
        int32_t r0 = (int32_t)floorf(clipf(mn[0] / (float)c->tw, 0.0f, (float)c->gw));
        int32_t r1 = (int32_t)floorf(clipf(mn[1] / (float)c->th, 0.0f, (float)c->gh));
        int32_t r2 = (int32_t)ceilf(clipf(mx[0] / (float)c->tw, 0.0f, (float)c->gw));
        int32_t r3 = (int32_t)ceilf(clipf(mx[1] / (float)c->th, 0.0f, (float)c->gh));
        int32_t n = (r2 - r0) * (r3 - r1);
        if (n == 0) continue;
        tiles[t] = n;
        rect[4 * t + 0] = r0; rect[4 * t + 1] = r1; rect[4 * t + 2] = r2; rect[4 * t + 3] = r3;
        vs[3 * t + 2] = fminf(fmaxf(depth, -1e20f), 1e20f);
        cc[3 * t + 0] = (float)ctr[0]; cc[3 * t + 1] = (float)ctr[1]; cc[3 * t + 2] = (float)ctr[2];
    }
}

/* ------------------------------------------------------------------ keys / sort / ranges -- */

uint32_t orc_sortable(float value) { /* tile_shader.slang:18-47 */
    uint32_t bits;
    memcpy(&bits, &value, 4);
    if (bits & 0x80000000u) bits = ~bits;
    else bits ^= 0x80000000u;
    return bits;
}

typedef struct { uint32_t depth; int32_t tet; } orc_entry;

static int cmp_entry(const void* a, const void* b) {
    const orc_entry* x = (const orc_entry*)a;
    const orc_entry* y = (const orc_entry*)b;
    if (x->depth != y->depth) return x->depth < y->depth ? -1 : 1;
    return (x->tet > y->tet) - (x->tet < y->tet);
}

/* Generates keys in the reference's buffer order (tet order, row-major tiles per tet) and sorts
 * them stably by the 64-bit key (tile << 32) | sortable(depth), writing sorted keys/values and
 * [G,2] ranges (empty tiles = [0,0]). A stable sort by that key is the same as grouping the
 * entries by tile in buffer order and sorting each tile's entries by depth key, ties kept in
 * buffer order, i.e. by increasing tet index (a tet adds at most one entry per tile); the tiles
 * are sorted in parallel. Returns P. keys_out/vals_out must hold sum(tiles) entries. */
int64_t orc_bin(const int32_t* tiles, const int32_t* rect, const float* vs, int64_t T, int32_t gh,
                int32_t gw, int64_t* keys_out, int32_t* vals_out, int32_t* ranges) {
    int64_t G = (int64_t)gh * gw;
    int64_t* off = (int64_t*)calloc((size_t)G + 1, sizeof(int64_t));
    for (int64_t t = 0; t < T; ++t) {
        if (tiles[t] <= 0) continue;
        for (int32_t y = rect[4 * t + 1]; y < rect[4 * t + 3]; ++y)
            for (int32_t x = rect[4 * t + 0]; x < rect[4 * t + 2]; ++x) off[(int64_t)y * gw + x + 1]++;
    }
    for (int64_t g = 0; g < G; ++g) off[g + 1] += off[g];
    int64_t P = off[G];
    orc_entry* e = (orc_entry*)malloc(sizeof(orc_entry) * (size_t)(P > 0 ? P : 1));
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)(G > 0 ? G : 1));
    for (int64_t g = 0; g < G; ++g) fill[g] = off[g];
    for (int64_t t = 0; t < T; ++t) {
        if (tiles[t] <= 0) continue;
        uint32_t fk = orc_sortable(vs[3 * t + 2]);
        for (int32_t y = rect[4 * t + 1]; y < rect[4 * t + 3]; ++y)
            for (int32_t x = rect[4 * t + 0]; x < rect[4 * t + 2]; ++x) {
                int64_t g = (int64_t)y * gw + x;
                e[fill[g]].depth = fk;
                e[fill[g]].tet = (int32_t)t;
                fill[g]++;
            }
    }
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t g = 0; g < G; ++g) {
        int64_t n = off[g + 1] - off[g];
        if (n > 1) qsort(e + off[g], (size_t)n, sizeof(orc_entry), cmp_entry);
        for (int64_t i = off[g]; i < off[g + 1]; ++i) {
            keys_out[i] = (int64_t)(((uint64_t)(uint32_t)g << 32) | e[i].depth);
            vals_out[i] = e[i].tet;
        }
        ranges[2 * g] = n > 0 ? (int32_t)off[g] : 0;
        ranges[2 * g + 1] = n > 0 ? (int32_t)off[g + 1] : 0;
    }
    free(e);
    free(fill);
    free(off);
    return P;
}

/* ------------------------------------------------------------------ FP64 compositor ------ */

typedef struct { double o[3], d[3]; } ray_d;

/* camera.slang:241-296 in double */
static void undistort_d(const orc_camera* c, double* p) {
    const float* kf = c->dist;
    if ((kf[0] * kf[0] + kf[1] * kf[1] + kf[2] * kf[2] + kf[3] * kf[3]) < 1e-9f) return;
    double k1 = kf[0], k2 = kf[1], p1 = kf[2], p2 = kf[3];
    double px = p[0], py = p[1], x = px, y = py;
    for (int i = 0; i < 10; ++i) {
        double r2 = x * x + y * y;
        double d = 1.0 + r2 * (k1 + r2 * k2);
        double fx = d * x + 2.0 * p1 * x * y + p2 * (r2 + 2.0 * x * x) - px;
        double fy = d * y + 2.0 * p2 * x * y + p1 * (r2 + 2.0 * y * y) - py;
        double d_r = k1 + r2 * (2.0 * k2);
        double d_x = 2.0 * x * d_r, d_y = 2.0 * y * d_r;
        double fx_x = d + d_x * x + 2.0 * p1 * y + 6.0 * p2 * x;
        double fx_y = d_y * x + 2.0 * p1 * x + 2.0 * p2 * y;
        double fy_x = d_x * y + 2.0 * p2 * y + 2.0 * p1 * x;
        double fy_y = d + d_y * y + 2.0 * p2 * x + 6.0 * p1 * y;
        double den = fy_x * fx_y - fx_x * fy_y;
        if (fabs(den) < 1e-9) break;
        x += (fx * fy_y - fy * fx_y) / den;
        y += (fy * fx_x - fx * fy_x) / den;
    }
    p[0] = x; p[1] = y;
}

/* camera.slang:298-346 in double */
static ray_d get_ray_d(const orc_camera* c, double px, double py) {
    double p[2] = {(px - c->K[2]) / c->K[0], (py - c->K[5]) / c->K[4]};
    if (c->camera_type == 0 || c->camera_type == 1) undistort_d(c, p);
    double dc[3];
    if (c->camera_type == 1) {
        double th = sqrt(p[0] * p[0] + p[1] * p[1]);
        th = th < 3.14159265 ? th : 3.14159265;
        double s = th < 1e-6 ? 1.0 : sin(th) / th;
        dc[0] = p[0] * s; dc[1] = p[1] * s; dc[2] = cos(th);
    } else {
        double l = sqrt(p[0] * p[0] + p[1] * p[1] + 1.0);
        dc[0] = p[0] / l; dc[1] = p[1] / l; dc[2] = 1.0 / l;
    }
    ray_d r;
    double d[3];
    for (int i = 0; i < 3; ++i) /* V_rot = transpose(W2V[:3,:3]) */
        d[i] = (double)c->w2v[0 + i] * dc[0] + (double)c->w2v[4 + i] * dc[1] + (double)c->w2v[8 + i] * dc[2];
    double l = len3d(d);
    for (int i = 0; i < 3; ++i) {
        r.d[i] = d[i] / l;
        r.o[i] = (double)c->cam_pos[i] + r.d[i] * (double)c->min_t;
    }
    return r;
}

typedef struct {
    int hit;
    double t_enter_raw, t_exit;   /* before clamping */
    double enter, exit;
    int ie, ix;                   /* entering / exiting plane, -1 when none */
    double n[4][3], dist[4], denom[4];
} isect_d;

static const int kFaceA[4] = {0, 0, 0, 1};

/* intersect.slang:393-624 in double; plane i keeps the reference's vertex choice. */
static void planes_d(const double v[4][3], double n[4][3]) {
    double e[3], f[3], w[3];
    /* face 0: (v0, v1, v2), outward away from v3 */
    for (int i = 0; i < 3; ++i) { e[i] = v[1][i] - v[0][i]; f[i] = v[2][i] - v[0][i]; w[i] = v[3][i] - v[0][i]; }
    cross3d(e, f, n[0]);
    if (dot3d(n[0], w) > 0) for (int i = 0; i < 3; ++i) n[0][i] = -n[0][i];
    for (int i = 0; i < 3; ++i) { e[i] = v[1][i] - v[0][i]; f[i] = v[3][i] - v[0][i]; w[i] = v[2][i] - v[0][i]; }
    cross3d(e, f, n[1]);
    if (dot3d(n[1], w) > 0) for (int i = 0; i < 3; ++i) n[1][i] = -n[1][i];
    for (int i = 0; i < 3; ++i) { e[i] = v[2][i] - v[0][i]; f[i] = v[3][i] - v[0][i]; w[i] = v[1][i] - v[0][i]; }
    cross3d(e, f, n[2]);
    if (dot3d(n[2], w) > 0) for (int i = 0; i < 3; ++i) n[2][i] = -n[2][i];
    for (int i = 0; i < 3; ++i) { e[i] = v[2][i] - v[1][i]; f[i] = v[3][i] - v[1][i]; w[i] = v[0][i] - v[1][i]; }
    cross3d(e, f, n[3]);
    if (dot3d(n[3], w) > 0) for (int i = 0; i < 3; ++i) n[3][i] = -n[3][i];
}

static isect_d intersect_d(const double v[4][3], const ray_d* r) {
    isect_d s;
    memset(&s, 0, sizeof(s));
    planes_d(v, s.n);
    double te = -1e30, tx = 1e30;
    s.ie = -1; s.ix = -1;
    for (int i = 0; i < 4; ++i) {
        const double* va = v[kFaceA[i]];
        double vo[3] = {va[0] - r->o[0], va[1] - r->o[1], va[2] - r->o[2]};
        s.dist[i] = -dot3d(s.n[i], vo);
        s.denom[i] = dot3d(s.n[i], r->d);
        if (fabs(s.denom[i]) < 1e-10) {
            if (s.dist[i] > 0.0) { s.hit = 0; return s; }
        } else {
            double tp = -s.dist[i] / s.denom[i];
            if (s.denom[i] < 0.0) { if (tp > te) { te = tp; s.ie = i; } }
            else { if (tp < tx) { tx = tp; s.ix = i; } }
        }
    }
    if (te > tx || tx <= 0.0) { s.hit = 0; return s; }
    s.hit = 1;
    s.t_enter_raw = te;
    s.t_exit = tx;
    s.enter = te > 0.0 ? te : 0.0;
    s.exit = tx > s.enter ? tx : s.enter;
    return s;
}

/* interp_version.slang:21-28 */
static double phi_d(double x) {
    if (fabs(x) < 1e-6) return 1.0 - x / 2.0;
    return -expm1(-x) / x;
}
static double dphi_d(double x) {
    if (fabs(x) < 1e-6) return -0.5;
    if (x < 1e-2) return -0.5 + x / 3.0 - x * x / 8.0 + x * x * x / 30.0 - x * x * x * x / 144.0;
    return (x * exp(-x) + expm1(-x)) / (x * x);
}

static void load_tet_d(const int32_t* idx, const double* verts, int32_t tet, double v[4][3]) {
    for (int i = 0; i < 4; ++i)
        for (int k = 0; k < 3; ++k) v[i][k] = verts[3 * (int64_t)idx[4 * (int64_t)tet + i] + k];
}

/* colour gradient row for channel ch: F=7 shares (f4,f5,f6); F=13 uses f[4+3ch .. 6+3ch] */
static const double* grd_row(const double* f, int F, int ch) { return F == 7 ? f + 4 : f + 4 + 3 * ch; }

/* Forward. out_img [H,W,4], out_dist [H,W,5], ncontrib [H,W], margin [H,W] (min |logT+5.5|
 * after any blend before termination; marks pixels whose termination decision is within the
 * margin of the threshold). */
/* The pixels of the listed tiles (every pixel when tile_list is NULL) as (x, y) pairs in raster
 * order within each tile; tiles are visited in list order. Caller frees *out. */
static int64_t pixel_list(const orc_camera* c, const int32_t* tile_list, int64_t n_list, int32_t** out) {
    int64_t n = 0;
    if (!tile_list) {
        n = (int64_t)c->H * c->W;
        int32_t* p = (int32_t*)malloc(sizeof(int32_t) * 2 * (size_t)(n > 0 ? n : 1));
        int64_t q = 0;
        for (int y = 0; y < c->H; ++y)
            for (int x = 0; x < c->W; ++x) { p[2 * q] = x; p[2 * q + 1] = y; q++; }
        *out = p;
        return n;
    }
    int32_t* p = (int32_t*)malloc(sizeof(int32_t) * 2 * (size_t)(n_list * c->th * c->tw + 1));
    for (int64_t i = 0; i < n_list; ++i) {
        int ty = tile_list[i] / c->gw, tx = tile_list[i] % c->gw;
        for (int y = ty * c->th; y < (ty + 1) * c->th && y < c->H; ++y)
            for (int x = tx * c->tw; x < (tx + 1) * c->tw && x < c->W; ++x) { p[2 * n] = x; p[2 * n + 1] = y; n++; }
    }
    *out = p;
    return n;
}

/* The forward over the pixels of tile_list [n_list] only (all tiles when tile_list is NULL);
 * other pixels' outputs are left untouched. */
void orc_composite_fwd_tiles(const orc_camera* c, const int32_t* sorted, const int32_t* ranges,
                             const int32_t* idx, const double* verts, const double* feats, int F,
                             const float* jitter, const int32_t* tile_list, int64_t n_list,
                             double* out_img, double* out_dist, int32_t* ncontrib, double* margin) {
    int W = c->W;
    int32_t* pl;
    int64_t n_pix = pixel_list(c, tile_list, n_list, &pl);
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t q = 0; q < n_pix; ++q) {
        {
            int x = pl[2 * q], y = pl[2 * q + 1];
            int tile = (y / c->th) * c->gw + (x / c->tw);
            int s0 = ranges[2 * tile], s1 = ranges[2 * tile + 1];
            int64_t pix = (int64_t)y * W + x;
            ray_d r = get_ray_d(c, (double)x + (double)jitter[2 * pix], (double)y + (double)jitter[2 * pix + 1]);
            double C[3] = {0, 0, 0}, A = 0.0, X = 0, Y = 0, Z = 0, Wd = 0, Vd = 0, mg = 1e30;
            int n = 0;
            for (int k = s0; k < s1; ++k) {
                n++;
                int32_t tet = sorted[k];
                double v[4][3];
                load_tet_d(idx, verts, tet, v);
                const double* f = feats + (int64_t)F * tet;
                isect_d s = intersect_d(v, &r);
                if (!s.hit) continue;
                double sig = f[0] > 0.0 ? f[0] : 0.0;
                double dt = s.exit - s.enter;
                if (dt < 0) dt = 0;
                double xx = sig * dt;
                if (!(xx > 0.0)) continue;
                double alpha = exp(-xx), ph = phi_d(xx);
                double w0 = ph - alpha, w1 = 1.0 - ph;
                double pe[3], px[3];
                for (int i = 0; i < 3; ++i) {
                    pe[i] = s.enter * r.d[i] + r.o[i] - v[0][i];
                    px[i] = s.exit * r.d[i] + r.o[i] - v[0][i];
                }
                double T = exp(A);
                for (int ch = 0; ch < 3; ++ch) {
                    const double* g = grd_row(f, F, ch);
                    double ce = dot3d(g, pe) + f[1 + ch];
                    double cx = dot3d(g, px) + f[1 + ch];
                    C[ch] += (w0 * cx + w1 * ce) * T;
                }
                A -= xx;
                double m0 = sig, m1 = sig * (s.exit + s.enter) / 2.0;
                X += 2.0 * m1 * Z;
                Y += 2.0 * m0 * Wd;
                Z += m0;
                Wd += m1;
                Vd += sig * sig * dt / 3.0;
                double am = fabs(A + 5.5);
                if (am < mg) mg = am;
                if (A < -5.5) break;
            }
            out_img[4 * pix + 0] = C[0]; out_img[4 * pix + 1] = C[1]; out_img[4 * pix + 2] = C[2];
            out_img[4 * pix + 3] = A;
            out_dist[5 * pix + 0] = X; out_dist[5 * pix + 1] = Y; out_dist[5 * pix + 2] = Z;
            out_dist[5 * pix + 3] = Wd; out_dist[5 * pix + 4] = Vd;
            ncontrib[pix] = n;
            if (margin) margin[pix] = mg;
        }
    }
    free(pl);
}

void orc_composite_fwd(const orc_camera* c, const int32_t* sorted, const int32_t* ranges,
                       const int32_t* idx, const double* verts, const double* feats, int F,
                       const float* jitter, double* out_img, double* out_dist, int32_t* ncontrib,
                       double* margin) {
    orc_composite_fwd_tiles(c, sorted, ranges, idx, verts, feats, F, jitter, NULL, 0, out_img, out_dist,
                            ncontrib, margin);
}

static void atomic_addd(double* p, double v) {
#pragma omp atomic
    *p += v;
}

/* Backward of n = s * cross(e1, e2), e1 = vb - va, e2 = vc - va, with gradient gn on n. */
static void cross_bwd(const double* e1, const double* e2, const double* gn, double* ge1, double* ge2) {
    double t[3];
    cross3d(e2, gn, t);             /* d<gn, e1 x e2>/de1 = e2 x gn */
    for (int i = 0; i < 3; ++i) ge1[i] = t[i];
    cross3d(gn, e1, t);             /* d<gn, e1 x e2>/de2 = gn x e1 */
    for (int i = 0; i < 3; ++i) ge2[i] = t[i];
}

/* face i = (a, b, c) with n_i = sign * cross(b - a, c - a), as planes_d */
static const int kFace[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};

/* Backward. Gradients are accumulated (+=) into grad_verts [V,3], grad_feats [T,F]. */
/* The backward over the pixels of tile_list [n_list] only (all tiles when tile_list is NULL). */
void orc_composite_bwd_tiles(const orc_camera* c, const int32_t* sorted, const int32_t* ranges,
                             const int32_t* idx, const double* verts, const double* feats, int F,
                             const float* jitter, const int32_t* tile_list, int64_t n_list,
                             const double* img, const double* dist, const int32_t* ncontrib,
                             const double* gimg, const double* gdist, double* grad_verts, double* grad_feats) {
    int W = c->W;
    int32_t* pl;
    int64_t n_pix = pixel_list(c, tile_list, n_list, &pl);
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t q = 0; q < n_pix; ++q) {
        {
            int x = pl[2 * q], y = pl[2 * q + 1];
            int tile = (y / c->th) * c->gw + (x / c->tw);
            int s0 = ranges[2 * tile];
            int64_t pix = (int64_t)y * W + x;
            ray_d r = get_ray_d(c, (double)x + (double)jitter[2 * pix], (double)y + (double)jitter[2 * pix + 1]);
            double A = img[4 * pix + 3], Z = dist[5 * pix + 2], Wd = dist[5 * pix + 3];
            double dC[3] = {gimg[4 * pix], gimg[4 * pix + 1], gimg[4 * pix + 2]};
            double dA = gimg[4 * pix + 3];
            double dX = gdist[5 * pix], dY = gdist[5 * pix + 1], dZ = gdist[5 * pix + 2];
            double dW = gdist[5 * pix + 3], dV = gdist[5 * pix + 4];
            int n = ncontrib[pix];
            for (int k = n - 1; k >= 0; --k) {
                int32_t tet = sorted[s0 + k];
                double v[4][3];
                load_tet_d(idx, verts, tet, v);
                const double* f = feats + (int64_t)F * tet;
                isect_d s = intersect_d(v, &r);
                if (!s.hit) continue;
                double sig = f[0] > 0.0 ? f[0] : 0.0;
                double dt = s.exit - s.enter;
                if (dt < 0) dt = 0;
                double xx = sig * dt;
                if (!(xx > 0.0)) continue;
                double alpha = exp(-xx), ph = phi_d(xx);
                double w0 = ph - alpha, w1 = 1.0 - ph;
                double pe[3], px[3];
                for (int i = 0; i < 3; ++i) {
                    pe[i] = s.enter * r.d[i] + r.o[i] - v[0][i];
                    px[i] = s.exit * r.d[i] + r.o[i] - v[0][i];
                }
                /* undo */
                double Ap = A + xx, T = exp(Ap);
                double m0 = sig, m1 = sig * (s.exit + s.enter) / 2.0;
                double zp = Z - m0, wp = Wd - m1;
                /* pixel state adjoint */
                double drgb[3], rgb[3], ce[3], cx[3];
                for (int ch = 0; ch < 3; ++ch) {
                    const double* g = grd_row(f, F, ch);
                    ce[ch] = dot3d(g, pe) + f[1 + ch];
                    cx[ch] = dot3d(g, px) + f[1 + ch];
                    rgb[ch] = w0 * cx[ch] + w1 * ce[ch];
                    drgb[ch] = dC[ch] * T;
                }
                double dT = dC[0] * rgb[0] + dC[1] * rgb[1] + dC[2] * rgb[2];
                double dxx = -dA;
                double dm1 = dX * 2.0 * zp + dW, dm0 = dY * 2.0 * wp + dZ;
                double dZn = dZ + dX * 2.0 * m1, dWn = dW + dY * 2.0 * m0;
                double dsig = dm0 + dm1 * (s.exit + s.enter) / 2.0 + dV * 2.0 * sig * dt / 3.0;
                double ddt = dV * sig * sig / 3.0;
                double den_ = dm1 * sig / 2.0, dex = dm1 * sig / 2.0;
                /* colour integral adjoint */
                double dw0 = 0, dw1 = 0, dqe[3], dqx[3], dbase[3];
                for (int ch = 0; ch < 3; ++ch) {
                    dbase[ch] = drgb[ch] * (w0 + w1);
                    dqx[ch] = drgb[ch] * w0;
                    dqe[ch] = drgb[ch] * w1;
                    dw0 += drgb[ch] * cx[ch];
                    dw1 += drgb[ch] * ce[ch];
                }
                double dph = dw0 - dw1, dal = -dw0;
                dxx += -alpha * dal + dph * dphi_d(xx);
                dsig += dxx * dt;
                ddt += dxx * sig;
                dex += ddt;
                den_ -= ddt;
                double gv[4][3];
                memset(gv, 0, sizeof(gv));
                double dgrd[9];
                memset(dgrd, 0, sizeof(dgrd));
                for (int ch = 0; ch < 3; ++ch) {
                    const double* g = grd_row(f, F, ch);
                    double gd = dot3d(g, r.d);
                    double* dg = F == 7 ? dgrd : dgrd + 3 * ch;
                    for (int i = 0; i < 3; ++i) {
                        dg[i] += dqe[ch] * pe[i] + dqx[ch] * px[i];
                        gv[0][i] -= (dqe[ch] + dqx[ch]) * g[i];
                    }
                    den_ += dqe[ch] * gd;
                    dex += dqx[ch] * gd;
                }
                /* intersection adjoint: enter = max(t_enter, 0), exit = t_exit */
                double dtp[4] = {0, 0, 0, 0};
                if (s.ie >= 0 && s.t_enter_raw > 0.0) dtp[s.ie] += den_;
                if (s.ix >= 0) dtp[s.ix] += dex;
                for (int i = 0; i < 4; ++i) {
                    if (dtp[i] == 0.0) continue;
                    int a = kFace[i][0], b = kFace[i][1], cc = kFace[i][2];
                    double dn = s.denom[i], ds = s.dist[i];
                    double ddist = -dtp[i] / dn;
                    double dden = dtp[i] * ds / (dn * dn);
                    const double* va = v[kFaceA[i]];
                    double gn[3];
                    for (int q = 0; q < 3; ++q) {
                        gn[q] = -ddist * (va[q] - r.o[q]) + dden * r.d[q];
                        gv[kFaceA[i]][q] += -ddist * s.n[i][q];
                    }
                    /* n_i = sign * cross(v_b - v_a, v_c - v_a); recover sign from the stored n */
                    double e1[3], e2[3], raw[3];
                    for (int q = 0; q < 3; ++q) { e1[q] = v[b][q] - v[a][q]; e2[q] = v[cc][q] - v[a][q]; }
                    cross3d(e1, e2, raw);
                    double sg = (dot3d(raw, s.n[i]) < 0.0) ? -1.0 : 1.0;
                    for (int q = 0; q < 3; ++q) gn[q] *= sg;
                    double ge1[3], ge2[3];
                    cross_bwd(e1, e2, gn, ge1, ge2);
                    for (int q = 0; q < 3; ++q) {
                        gv[b][q] += ge1[q];
                        gv[cc][q] += ge2[q];
                        gv[a][q] -= ge1[q] + ge2[q];
                    }
                }
                for (int i = 0; i < 4; ++i)
                    for (int q = 0; q < 3; ++q)
                        atomic_addd(&grad_verts[3 * (int64_t)idx[4 * (int64_t)tet + i] + q], gv[i][q]);
                double* gf = grad_feats + (int64_t)F * tet;
                if (f[0] > 0.0) atomic_addd(&gf[0], dsig);
                for (int ch = 0; ch < 3; ++ch) atomic_addd(&gf[1 + ch], dbase[ch]);
                for (int q = 0; q < (F == 7 ? 3 : 9); ++q) atomic_addd(&gf[4 + q], dgrd[q]);
                /* step the state back */
                A = Ap; Z = zp; Wd = wp;
                dA = dA + dT * T;
                dZ = dZn; dW = dWn;
            }
        }
    }
    free(pl);
}

void orc_composite_bwd(const orc_camera* c, const int32_t* sorted, const int32_t* ranges,
                       const int32_t* idx, const double* verts, const double* feats, int F,
                       const float* jitter, const double* img, const double* dist,
                       const int32_t* ncontrib, const double* gimg, const double* gdist,
                       double* grad_verts, double* grad_feats) {
    orc_composite_bwd_tiles(c, sorted, ranges, idx, verts, feats, F, jitter, NULL, 0, img, dist, ncontrib,
                            gimg, gdist, grad_verts, grad_feats);
}

/* ------------------------------------------------------------------ FP64 error votes ----- */

/* Per-tet error votes of one view: kern_calc_tet_err / calc_tet_err
 * (alphablend_shader_interp.slang:433-599) as launched by render_err (render_err.py:138-150),
 * in double. Pixel-centre rays (:435-450, ray jitter 0.5), the same per-tile lists and cull set
 * as orc_composite_fwd, log-transmittance starting at 1 (curr_pixel_state = (0,0,0,1), :450),
 * w = exp(logT) * sigma * dt for each blended tet (hit and sigma*dt > 0, :487-491), votes
 * (:497-525) into err[T,16] and count[T,2] (count, max int(65535 w)), termination after a blend
 * with logT < -5.5 (:529). Pixels are visited in raster order, serially (deterministic sums).
 * margin[H*W] (optional) receives min |logT + 5.5| after any blend: the distance of the
 * pixel's termination decision from the threshold; grazes[T] (optional) counts, per tet, the
 * pixels whose ray hits it with a chord dt < 1e-12 before stopping (a vote needs sigma*dt > 0,
 * so such a hit votes or not depending on the precision of the crossing distances). */
void orc_tet_err(const orc_camera* c, const int32_t* sorted, const int32_t* ranges, const int32_t* idx,
                 const double* verts, const double* feats, int F, const double* pixel_err,
                 const double* ssim_err, double* err, int32_t* count, double* margin, int32_t* grazes) {
    int H = c->H, W = c->W;
    for (int y = 0; y < H; ++y) {
        for (int x = 0; x < W; ++x) {
            int tile = (y / c->th) * c->gw + (x / c->tw);
            int s0 = ranges[2 * tile], s1 = ranges[2 * tile + 1];
            int64_t pix = (int64_t)y * W + x;
            ray_d r = get_ray_d(c, (double)x + 0.5, (double)y + 0.5);
            double A = 1.0, mg = 1e30;
            double pe = pixel_err[pix], ps = ssim_err[pix];
            for (int k = s0; k < s1; ++k) {
                int32_t tet = sorted[k];
                double v[4][3];
                load_tet_d(idx, verts, tet, v);
                const double* f = feats + (int64_t)F * tet;
                isect_d s = intersect_d(v, &r);
                if (!s.hit) continue;
                double sig = f[0] > 0.0 ? f[0] : 0.0;
                double dt = s.exit - s.enter;
                if (dt < 0) dt = 0;
                if (dt < 1e-12 && grazes) grazes[tet] += 1;
                double xx = sig * dt;
                if (!(xx > 0.0)) continue;
                double w = exp(A) * xx;
                double* e = err + 16 * (int64_t)tet;
                e[0] += w;
                e[1] += w * pe;
                e[2] += w * pe * pe;
                e[5] += w * ps;
                for (int i = 0; i < 3; ++i) {
                    double p1 = s.enter * r.d[i] + r.o[i];
                    double p2 = s.exit * r.d[i] + r.o[i];
                    double pc = 0.5 * (s.enter + s.exit) * r.d[i] + r.o[i];
                    e[6 + i] += w * p1;
                    e[9 + i] += w * p2;
                    e[13 + i] += w * pc;
                }
                e[12] += w;
                count[2 * (int64_t)tet] += 1;
                double wi = 65535.0 * w;
                int32_t iw = wi >= 2147483647.0 ? 2147483647 : (int32_t)wi;
                if (iw > count[2 * (int64_t)tet + 1]) count[2 * (int64_t)tet + 1] = iw;
                A -= xx;
                double am = fabs(A + 5.5);
                if (am < mg) mg = am;
                if (A < -5.5) break;
            }
            if (margin) margin[pix] = mg;
        }
    }
}

/* ------------------------------------------------------------------ midpoint quadrature -- */

/* tetra_quad.py:101-133 same-side test (strict sign product) */
static int same_side_d(const double* v1, const double* v2, const double* v3, const double* v4, const double* p) {
    double e1[3], e2[3], n[3], a[3], b[3];
    for (int i = 0; i < 3; ++i) { e1[i] = v2[i] - v1[i]; e2[i] = v3[i] - v1[i]; a[i] = v4[i] - v1[i]; b[i] = p[i] - v1[i]; }
    cross3d(e1, e2, n);
    double s4 = dot3d(n, a), sp = dot3d(n, b);
    double g4 = (s4 > 0) - (s4 < 0), gp = (sp > 0) - (sp < 0);
    return g4 * gp > 0;
}

/* Midpoint quadrature of one ray over [t0, t1] with n_samples intervals; every tet is tested
 * at every sample (tetra_quad.py:142-224), weights (1 - e^{-sigma dt}) T (:58-72), no early
 * termination, background 0. out4 = (rgb, sum of weights). */
void orc_quadrature_ray(const double* o, const double* d, double t0, double t1, int n_samples,
                        const int32_t* idx, const double* verts, const double* feats, int F,
                        int64_t T, double* out4) {
    double C[3] = {0, 0, 0}, acc = 0.0, logT = 0.0;
    double dtq = (t1 - t0) / n_samples;
    for (int s = 0; s < n_samples; ++s) {
        double t = t0 + (s + 0.5) * dtq;
        double p[3] = {o[0] + t * d[0], o[1] + t * d[1], o[2] + t * d[2]};
        double dens = 0.0, col[3] = {0, 0, 0};
        for (int64_t k = 0; k < T; ++k) {
            double v[4][3];
            load_tet_d(idx, verts, (int32_t)k, v);
            if (!(same_side_d(v[0], v[1], v[2], v[3], p) && same_side_d(v[1], v[2], v[3], v[0], p) &&
                  same_side_d(v[2], v[3], v[0], v[1], p) && same_side_d(v[3], v[0], v[1], v[2], p)))
                continue;
            const double* f = feats + (int64_t)F * k;
            double sig = f[0] > 0.0 ? f[0] : 0.0;
            double q[3] = {p[0] - v[0][0], p[1] - v[0][1], p[2] - v[0][2]};
            dens += sig;
            for (int ch = 0; ch < 3; ++ch) col[ch] += sig * (f[1 + ch] + dot3d(grd_row(f, F, ch), q));
        }
        if (dens > 1e-6)
            for (int ch = 0; ch < 3; ++ch) col[ch] /= dens;
        else
            col[0] = col[1] = col[2] = 0.0;
        double w = -expm1(-dens * dtq) * exp(logT);
        for (int ch = 0; ch < 3; ++ch) C[ch] += w * col[ch];
        acc += w;
        logT -= dens * dtq;
    }
    out4[0] = C[0]; out4[1] = C[1]; out4[2] = C[2]; out4[3] = acc;
}

/* Quadrature for the pixels listed in pix[n_pix] (flat y*W+x) with the rasterizer's rays. */
void orc_quadrature(const orc_camera* c, const float* jitter, const int64_t* pix, int64_t n_pix,
                    double t1, int n_samples, const int32_t* idx, const double* verts,
                    const double* feats, int F, int64_t T, double* out4) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t i = 0; i < n_pix; ++i) {
        int64_t p = pix[i];
        int x = (int)(p % c->W), y = (int)(p / c->W);
        ray_d r = get_ray_d(c, (double)x + (double)jitter[2 * p], (double)y + (double)jitter[2 * p + 1]);
        orc_quadrature_ray(r.o, r.d, 0.0, t1, n_samples, idx, verts, feats, F, T, out4 + 4 * i);
    }
}

/* Ray (origin, direction) in double for a pixel, exported for tests. */
void orc_get_ray(const orc_camera* c, double px, double py, double* o, double* d) {
    ray_d r = get_ray_d(c, px, py);
    memcpy(o, r.o, sizeof(r.o));
    memcpy(d, r.d, sizeof(r.d));
}

/* Single ray/tet intersection in double, exported for tests: returns hit, writes enter/exit. */
int orc_intersect(const double* verts12, const double* o, const double* d, double* enter, double* exit_) {
    double v[4][3];
    memcpy(v, verts12, sizeof(v));
    ray_d r;
    memcpy(r.o, o, sizeof(r.o));
    memcpy(r.d, d, sizeof(r.d));
    isect_d s = intersect_d(v, &r);
    *enter = s.enter;
    *exit_ = s.exit;
    return s.hit;
}

/* ------------------------------------------------------------------ projection adjoint --- */

/* Adjoint of (power, center) = power_of_circumsphere(A,B,C,D,P) w.r.t. A..D given gradients
 * g_power and g_center[3]; adds into g[4][3]. Non-degenerate branch only (den != 0). */
void orc_power_bwd(const double* V12, const double* P, double g_power, const double* g_center, double* g12) {
    const double* A = V12;
    double a[3], b[3], c[3];
    for (int i = 0; i < 3; ++i) { a[i] = V12[3 + i] - A[i]; b[i] = V12[6 + i] - A[i]; c[i] = V12[9 + i] - A[i]; }
    double aa = dot3d(a, a), bb = dot3d(b, b), cc = dot3d(c, c);
    double bc[3], ca[3], ab[3];
    cross3d(b, c, bc); cross3d(c, a, ca); cross3d(a, b, ab);
    double den = 2.0 * dot3d(a, bc);
    if (fabs(den) < 1e-10) {
        /* stable_power.slang:41-58: a very small tet (largest edge below 1e-7) uses the centroid
         * and its smallest corner distance; otherwise the power is the constant -1e20. */
        const double* pts[4] = {V12, V12 + 3, V12 + 6, V12 + 9};
        double md = 0.0;
        for (int i = 0; i < 4; ++i)
            for (int j = i + 1; j < 4; ++j) {
                double d[3] = {pts[j][0] - pts[i][0], pts[j][1] - pts[i][1], pts[j][2] - pts[i][2]};
                double l = len3d(d);
                if (l > md) md = l;
            }
        if (md >= 1e-7) return;
        double ctr[3];
        for (int i = 0; i < 3; ++i) ctr[i] = 0.25 * (pts[0][i] + pts[1][i] + pts[2][i] + pts[3][i]);
        int km = 0;
        double r = 1e300;
        for (int k = 0; k < 4; ++k) {
            double d[3] = {pts[k][0] - ctr[0], pts[k][1] - ctr[1], pts[k][2] - ctr[2]};
            double l = len3d(d);
            if (l < r) { r = l; km = k; }
        }
        /* power = |ctr - P|^2 - r^2, r = |V_km - ctr| */
        double gct[3], u[3];
        for (int i = 0; i < 3; ++i) {
            gct[i] = g_center[i] + 2.0 * (ctr[i] - P[i]) * g_power;
            u[i] = r > 0 ? (pts[km][i] - ctr[i]) / r : 0.0;
        }
        double gr = -2.0 * r * g_power;
        for (int i = 0; i < 3; ++i) {
            g12[3 * km + i] += gr * u[i];
            double gc = gct[i] - gr * u[i];
            for (int k = 0; k < 4; ++k) g12[3 * k + i] += 0.25 * gc;
        }
        return;
    }
    double num[3], rel[3], ctr[3];
    for (int i = 0; i < 3; ++i) { num[i] = aa * bc[i] + bb * ca[i] + cc * ab[i]; rel[i] = num[i] / den; ctr[i] = A[i] + rel[i]; }
    double radius = len3d(rel);
    double diff[3] = {ctr[0] - P[0], ctr[1] - P[1], ctr[2] - P[2]};
    double gct[3], grel[3], gA[3];
    for (int i = 0; i < 3; ++i) gct[i] = g_center[i] + 2.0 * diff[i] * g_power;
    double grad_r = -2.0 * radius * g_power;
    for (int i = 0; i < 3; ++i) { gA[i] = gct[i]; grel[i] = gct[i] + (radius > 0 ? grad_r * rel[i] / radius : 0.0); }
    double gnum[3], gden = 0.0;
    for (int i = 0; i < 3; ++i) { gnum[i] = grel[i] / den; gden -= grel[i] * rel[i] / den; }
    double gaa = dot3d(gnum, bc), gbb = dot3d(gnum, ca), gcc = dot3d(gnum, ab);
    double gbc[3], gca[3], gab[3], ga[3], gb[3], gc[3], t[3];
    for (int i = 0; i < 3; ++i) {
        gbc[i] = aa * gnum[i] + 2.0 * a[i] * gden;
        gca[i] = bb * gnum[i];
        gab[i] = cc * gnum[i];
        ga[i] = 2.0 * a[i] * gaa + 2.0 * bc[i] * gden;
        gb[i] = 2.0 * b[i] * gbb;
        gc[i] = 2.0 * c[i] * gcc;
    }
    cross3d(c, gbc, t); for (int i = 0; i < 3; ++i) gb[i] += t[i];   /* bc = b x c */
    cross3d(gbc, b, t); for (int i = 0; i < 3; ++i) gc[i] += t[i];
    cross3d(a, gca, t); for (int i = 0; i < 3; ++i) gc[i] += t[i];   /* ca = c x a */
    cross3d(gca, c, t); for (int i = 0; i < 3; ++i) ga[i] += t[i];
    cross3d(b, gab, t); for (int i = 0; i < 3; ++i) ga[i] += t[i];   /* ab = a x b */
    cross3d(gab, a, t); for (int i = 0; i < 3; ++i) gb[i] += t[i];
    for (int i = 0; i < 3; ++i) {
        g12[i] += gA[i] - ga[i] - gb[i] - gc[i];
        g12[3 + i] += ga[i];
        g12[6 + i] += gb[i];
        g12[9 + i] += gc[i];
    }
}
