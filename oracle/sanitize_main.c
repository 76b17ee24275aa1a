/* Test infrastructure only: runs the oracle's rasterizer functions on one scene under
 * AddressSanitizer and UndefinedBehaviorSanitizer (built by `make -C oracle sanitize`).
 *
 * Reads a scene file (written by tests/test_oracle_sanitized.py), calls orc_project, orc_bin,
 * orc_composite_fwd, orc_composite_bwd, orc_tet_err and orc_quadrature on it, and writes every
 * output array to the output file in that order, so the test can compare them with the same
 * functions run from the regular (unsanitized) library. Any out-of-bounds access, use of
 * uninitialised heap through the sanitizer's checks, signed overflow or invalid shift aborts the
 * run with the sanitizer's report and a non-zero exit status.
 *
 * Scene file: int64 magic 0x52534e45, T, V, F, H, W, n_pix; the orc_camera struct as raw bytes;
 * idx int32[T*4]; verts float32[V*3]; feats float64[T*F]; jitter float32[H*W*2];
 * gimg float64[H*W*4]; gdist float64[H*W*5]; pixel_err float64[H*W]; ssim_err float64[H*W];
 * pix int64[n_pix] (pixels for the quadrature). */
#include <stdint.h>
#include <stdio.h>
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

void orc_project(const orc_camera* c, const int32_t* idx, const float* verts, int64_t T, int32_t* tiles,
                 int32_t* rect, float* vs, float* cc);
int64_t orc_bin(const int32_t* tiles, const int32_t* rect, const float* vs, int64_t T, int32_t gh, int32_t gw,
                int64_t* keys_out, int32_t* vals_out, int32_t* ranges);
void orc_composite_fwd(const orc_camera* c, const int32_t* sorted, const int32_t* ranges, const int32_t* idx,
                       const double* verts, const double* feats, int F, const float* jitter, double* out_img,
                       double* out_dist, int32_t* ncontrib, double* margin);
void orc_composite_bwd(const orc_camera* c, const int32_t* sorted, const int32_t* ranges, const int32_t* idx,
                       const double* verts, const double* feats, int F, const float* jitter, const double* img,
                       const double* dist, const int32_t* ncontrib, const double* gimg, const double* gdist,
                       double* grad_verts, double* grad_feats);
void orc_tet_err(const orc_camera* c, const int32_t* sorted, const int32_t* ranges, const int32_t* idx,
                 const double* verts, const double* feats, int F, const double* pixel_err,
                 const double* ssim_err, double* err, int32_t* count, double* margin, int32_t* grazes);
void orc_quadrature(const orc_camera* c, const float* jitter, const int64_t* pix, int64_t n_pix, double t1,
                    int n_samples, const int32_t* idx, const double* verts, const double* feats, int F,
                    int64_t T, double* out4);

static void* take(FILE* f, size_t n) {
    void* p = malloc(n ? n : 1);
    if (!p || (n && fread(p, 1, n, f) != n)) {
        fprintf(stderr, "sanitize_main: short read\n");
        exit(2);
    }
    return p;
}

static void put(FILE* f, const void* p, size_t n) {
    if (n && fwrite(p, 1, n, f) != n) {
        fprintf(stderr, "sanitize_main: short write\n");
        exit(2);
    }
}

int main(int argc, char** argv) {
    if (argc != 3) {
        fprintf(stderr, "usage: %s scene.bin out.bin\n", argv[0]);
        return 2;
    }
    FILE* in = fopen(argv[1], "rb");
    if (!in) return 2;
    int64_t* hdr = (int64_t*)take(in, 7 * sizeof(int64_t));
    if (hdr[0] != 0x52534e45) return 2;
    const int64_t T = hdr[1], V = hdr[2], F = hdr[3], H = hdr[4], W = hdr[5], NP = hdr[6];
    orc_camera* cam = (orc_camera*)take(in, sizeof(orc_camera));
    int32_t* idx = (int32_t*)take(in, sizeof(int32_t) * T * 4);
    float* vf = (float*)take(in, sizeof(float) * V * 3);
    double* feats = (double*)take(in, sizeof(double) * T * F);
    float* jit = (float*)take(in, sizeof(float) * H * W * 2);
    double* gimg = (double*)take(in, sizeof(double) * H * W * 4);
    double* gdist = (double*)take(in, sizeof(double) * H * W * 5);
    double* perr = (double*)take(in, sizeof(double) * H * W);
    double* serr = (double*)take(in, sizeof(double) * H * W);
    int64_t* pix = (int64_t*)take(in, sizeof(int64_t) * NP);
    fclose(in);

    const int64_t G = (int64_t)cam->gh * cam->gw;
    int32_t* tiles = (int32_t*)malloc(sizeof(int32_t) * T);
    int32_t* rect = (int32_t*)malloc(sizeof(int32_t) * T * 4);
    float* vs = (float*)malloc(sizeof(float) * T * 3);
    float* cc = (float*)malloc(sizeof(float) * T * 3);
    orc_project(cam, idx, vf, T, tiles, rect, vs, cc);
    int64_t P = 0;
    for (int64_t t = 0; t < T; ++t) P += tiles[t];
    int64_t* keys = (int64_t*)malloc(sizeof(int64_t) * (P ? P : 1));
    int32_t* vals = (int32_t*)malloc(sizeof(int32_t) * (P ? P : 1));
    int32_t* ranges = (int32_t*)malloc(sizeof(int32_t) * G * 2);
    const int64_t P2 = orc_bin(tiles, rect, vs, T, cam->gh, cam->gw, keys, vals, ranges);
    if (P2 != P) return 3;

    double* vd = (double*)malloc(sizeof(double) * V * 3);
    for (int64_t i = 0; i < V * 3; ++i) vd[i] = vf[i];
    double* img = (double*)calloc((size_t)(H * W * 4), sizeof(double));
    double* dist = (double*)calloc((size_t)(H * W * 5), sizeof(double));
    int32_t* nc = (int32_t*)calloc((size_t)(H * W), sizeof(int32_t));
    double* margin = (double*)calloc((size_t)(H * W), sizeof(double));
    orc_composite_fwd(cam, vals, ranges, idx, vd, feats, (int)F, jit, img, dist, nc, margin);
    double* gv = (double*)calloc((size_t)(V * 3), sizeof(double));
    double* gf = (double*)calloc((size_t)(T * F), sizeof(double));
    orc_composite_bwd(cam, vals, ranges, idx, vd, feats, (int)F, jit, img, dist, nc, gimg, gdist, gv, gf);
    double* err = (double*)calloc((size_t)(T * 16), sizeof(double));
    int32_t* count = (int32_t*)calloc((size_t)(T * 2), sizeof(int32_t));
    double* emargin = (double*)calloc((size_t)(H * W), sizeof(double));
    int32_t* grazes = (int32_t*)calloc((size_t)T, sizeof(int32_t));
    orc_tet_err(cam, vals, ranges, idx, vd, feats, (int)F, perr, serr, err, count, emargin, grazes);
    double* quad = (double*)calloc((size_t)(NP * 4 + 1), sizeof(double));
    orc_quadrature(cam, jit, pix, NP, 12.0, 32, idx, vd, feats, (int)F, T, quad);

    FILE* out = fopen(argv[2], "wb");
    if (!out) return 2;
    put(out, &P, sizeof(P));
    put(out, tiles, sizeof(int32_t) * T);
    put(out, rect, sizeof(int32_t) * T * 4);
    put(out, vs, sizeof(float) * T * 3);
    put(out, cc, sizeof(float) * T * 3);
    put(out, keys, sizeof(int64_t) * P);
    put(out, vals, sizeof(int32_t) * P);
    put(out, ranges, sizeof(int32_t) * G * 2);
    put(out, img, sizeof(double) * H * W * 4);
    put(out, dist, sizeof(double) * H * W * 5);
    put(out, nc, sizeof(int32_t) * H * W);
    put(out, margin, sizeof(double) * H * W);
    put(out, gv, sizeof(double) * V * 3);
    put(out, gf, sizeof(double) * T * F);
    put(out, err, sizeof(double) * T * 16);
    put(out, count, sizeof(int32_t) * T * 2);
    put(out, quad, sizeof(double) * NP * 4);
    fclose(out);

    free(hdr); free(cam); free(idx); free(vf); free(feats); free(jit); free(gimg); free(gdist); free(perr);
    free(serr); free(pix); free(tiles); free(rect); free(vs); free(cc); free(keys); free(vals); free(ranges);
    free(vd); free(img); free(dist); free(nc); free(margin); free(gv); free(gf); free(err); free(count);
    free(emargin); free(grazes); free(quad);
    return 0;
}
