/*
 * rmesh.h -- C ABI of the MI355X (gfx950) tetrahedral tile rasterizer and hash-grid field.
 *
 * Every pointer argument named *_dev is device memory owned by the caller; every entry point
 * enqueues work on the given HIP stream (0 = the null stream), returns 0 on success and a
 * non-zero RM_ERR_* code otherwise. rm_last_error() returns a thread-local message for the last
 * failure. No entry point allocates device memory: operations that need scratch take a
 * caller-provided workspace sized by the matching *_workspace_size() query (two-phase, as
 * CUB's temp-storage protocol). Forward passes are deterministic (no atomics); backward passes
 * accumulate with float atomics into caller-zeroed gradient buffers.
 *
 * Camera convention (reference: delaunay_rasterization/internal/slang/camera.slang:13-76 and
 * data/camera.py:72-91): w2v is the row-major 4x4 world->view matrix applied to column vectors,
 * K the row-major 3x3 intrinsics [[fx,0,cx],[0,fy,cy],[0,0,1]], dist = (k1,k2,p1,p2),
 * camera_type 0 = pinhole (optionally distorted), 1 = fisheye.
 */
#ifndef RMESH_H
#define RMESH_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* rm_stream_t; /* == hipStream_t */

enum {
    RM_OK = 0,
    RM_ERR_INVALID_ARGUMENT = 1,
    RM_ERR_UNSUPPORTED_TILE = 2,
    RM_ERR_WORKSPACE_TOO_SMALL = 3,
    RM_ERR_HIP = 4,
    RM_ERR_CAPACITY = 5
};

typedef struct rm_camera {
    float w2v[16];
    float K[9];
    float cam_pos[3];
    float dist[4];
    float min_t;
    float fovy;
    float fovx;
    int32_t image_height;
    int32_t image_width;
    int32_t grid_height;
    int32_t grid_width;
    int32_t tile_height;
    int32_t tile_width;
    int32_t camera_type;
    int32_t _pad;
} rm_camera;

/* ---- library info ---------------------------------------------------------------------- */
const char* rm_last_error(void);
int rm_version(void);

/* ---- per-tet projection and tile binning ------------------------------------------------
 * Replaces the Slang kernel vertex_shader launched by VertexShader.forward
 * (delaunay_rasterization/internal/tile_shader_slang.py:75-125,
 *  delaunay_rasterization/internal/slang/vertex_shader.slang:20-89).
 * Writes every output element (culled tets get zeros), so outputs need no zero-init.
 * vs_dev is [T,3] f32 of which only column 2 (the circumsphere power) is non-zero. */
int rm_project_fwd(const rm_camera* cam, const int32_t* indices_dev, const float* vertices_dev,
                   int64_t n_tets, int64_t n_verts,
                   int32_t* tiles_touched_dev, int32_t* rect_dev, float* vs_dev,
                   float* circumcenter_dev, rm_stream_t stream);

/* Replaces vertex_shader.bwd (tile_shader_slang.py:127-150): gradients of vs[:,2] and the
 * circumcenter w.r.t. the vertices (FP64 adjoint), atomically added into grad_vertices_dev. */
int rm_project_bwd(const rm_camera* cam, const int32_t* indices_dev, const float* vertices_dev,
                   int64_t n_tets, int64_t n_verts, const int32_t* tiles_touched_dev,
                   const float* grad_vs_dev, const float* grad_circumcenter_dev,
                   float* grad_vertices_dev, rm_stream_t stream);

/* ---- scan / keys / sort / ranges: the reference's own decomposition ---------------------
 * rm_inclusive_scan_i32 replaces torch.cumsum(tiles_touched) (tile_shader_slang.py:36). */
size_t rm_scan_workspace_size(int64_t n);
int rm_inclusive_scan_i32(const int32_t* in_dev, int32_t* out_dev, int64_t n,
                          void* workspace_dev, size_t workspace_bytes, rm_stream_t stream);

/* Replaces tile_shader.generate_keys (tile_shader.slang:49-97). */
int rm_generate_keys(const float* vs_dev, const int32_t* rect_dev, const int32_t* offsets_dev,
                     int64_t n_tets, int32_t grid_height, int32_t grid_width,
                     int64_t* keys_dev, int32_t* values_dev, rm_stream_t stream);

/* Replaces sort_by_keys_cub.sort_by_keys (sort_by_keys/sort_by_keys.cu:9-53): stable LSD
 * radix sort of (u64 key, i32 value) pairs over key bits [0, end_bit). */
size_t rm_sort_pairs_workspace_size(int64_t n);
int rm_sort_pairs(const int64_t* keys_in_dev, const int32_t* values_in_dev,
                  int64_t* keys_out_dev, int32_t* values_out_dev, int64_t n, int32_t end_bit,
                  void* workspace_dev, size_t workspace_bytes, rm_stream_t stream);

/* Replaces tile_shader.compute_tile_ranges (tile_shader.slang:99-125). ranges_dev [G,2]
 * is fully written: empty tiles get [0,0] as the reference's zero-initialised buffer. */
int rm_tile_ranges(const int64_t* sorted_keys_dev, int64_t n, int32_t* ranges_dev,
                   int64_t n_tiles, rm_stream_t stream);

/* ---- fused binning: the MI355X path behind vertex_and_tile_shader ------------------------
 * Replaces tile_shader_slang.py:36-68 (cumsum + generate_keys + CUB sort + ranges) by a
 * stable radix sort of the visible tets by depth key, a stable partition of the depth-ordered
 * tets into blocks of 8x8 tiles, and per block a direct write of every tet into the lists of
 * the tiles it covers. The result is bit-identical to the reference's stable sort of
 * (tile << 32 | sortable(depth)) keys (equal keys keep tet-index order).
 * Phase 1 (rm_bin_count) compacts visible tets, scans pair offsets and writes the totals
 * {n_visible, n_pairs} to counts_dev (int64[2]); the caller reads n_pairs to size phase 2.
 * rm_bin_workspace_size needs the grid's tile count n_tiles (grid_height * grid_width) for
 * phase 2; its first rm_bin_workspace_size(n_tets, 0, n_tiles) bytes hold phase 1's state. */
size_t rm_bin_workspace_size(int64_t n_tets, int64_t n_pairs, int64_t n_tiles);
int rm_bin_count(const int32_t* tiles_touched_dev, int64_t n_tets, int64_t* counts_dev,
                 void* workspace_dev, size_t workspace_bytes, rm_stream_t stream);
int rm_bin_sort(const int32_t* tiles_touched_dev, const int32_t* rect_dev, const float* vs_dev,
                int64_t n_tets, int64_t n_visible, int64_t n_pairs,
                int32_t grid_height, int32_t grid_width,
                int32_t* sorted_tet_dev, int32_t* ranges_dev,
                void* workspace_dev, size_t workspace_bytes, rm_stream_t stream);

/* ---- per-tile compositing ---------------------------------------------------------------
 * Replaces splat_tiled (alphablend_shader_interp.slang:371-431, launched by
 * AlphaBlendTiledRender.forward, alphablend_tiled_slang_interp.py:11-67).
 * n_features = 7 ([density, rgb, grad]: one colour gradient shared by rgb, the reference's
 * layout) or 13 ([density, rgb, 3x3 per-channel gradient rows]). Tiles: square 4, 8 or 16.
 * out_img [H,W,4] = (rgb, log-transmittance), out_dist [H,W,5], out_ncontrib [H,W] i32.
 * n_sorted is the length of sorted_tet_dev; tile ranges outside [0, n_sorted] count as empty. */
int rm_composite_fwd(const rm_camera* cam, const int32_t* sorted_tet_dev, int64_t n_sorted,
                     const int32_t* ranges_dev, const int32_t* indices_dev, const float* vertices_dev,
                     const float* features_dev, int32_t n_features, const float* jitter_dev,
                     float* out_img_dev, float* out_dist_dev, int32_t* out_ncontrib_dev,
                     rm_stream_t stream);

/* rm_composite_fwd that also records, for the backward, which list entries matter: for every entry
 * k (an index into sorted_tet_dev) that the pixels of its tile walked, pair_flags_dev[k] = 1 when some
 * pixel of the tile blended that tet (a hit with sigma * dt > 0) and 0 otherwise; entries no pixel
 * reached are left unwritten. pair_flags_dev holds n_sorted bytes, or is null (nothing recorded).
 * Square tiles 4 and 8 (a tile within one wavefront); with tile 16 nothing is written. With 4x4
 * tiles, tile_order_dev (null, or the gw*gh tile ids in any order, e.g. from rm_tile_order) is the
 * order in which the tiles are taken up. Same image outputs as rm_composite_fwd, whatever the order.
 * No counterpart in the reference (its backward revisits every entry). */
int rm_composite_fwd_flags(const rm_camera* cam, const int32_t* sorted_tet_dev, int64_t n_sorted,
                           const int32_t* ranges_dev, const int32_t* indices_dev, const float* vertices_dev,
                           const float* features_dev, int32_t n_features, const float* jitter_dev,
                           float* out_img_dev, float* out_dist_dev, int32_t* out_ncontrib_dev,
                           uint8_t* pair_flags_dev, const int32_t* tile_order_dev, rm_stream_t stream);

/* The tiles by decreasing list length (ranges [n_tiles, 2] as written by rm_bin_sort; lengths
 * above 65535 count as 65535; tiles of equal length in increasing id) into order_dev [n_tiles]:
 * a tile's walk runs on one wavefront, so taking the longest lists first keeps a late long walk
 * from running alone at the end of a launch. The workspace holds rm_tile_order_workspace_size
 * bytes (256-byte aligned). No counterpart in the reference. */
size_t rm_tile_order_workspace_size(int64_t n_tiles);
int rm_tile_order(const int32_t* ranges_dev, int64_t n_tiles, int32_t* order_dev, void* workspace_dev,
                  size_t workspace_bytes, rm_stream_t stream);

/* Precision of the per-pixel ray/plane intersection in rm_composite_fwd/bwd (process-wide):
 * 1 (default) = FP64 plane records and crossing distances, rounded to FP32 once the segment
 * [entry, exit] is known; 0 = FP32 throughout (faster; rays that graze a face then carry the FP32
 * conditioning of the crossing, up to ~1e-4 in colour on the test scenes). Blending, colour and
 * transmittance are FP32 in both modes. No counterpart in the reference (always FP32). */
int rm_set_intersect_precision(int fp64);

/* Replaces splat_tiled.bwd / bwd_alpha_blend (alphablend_shader_interp.slang:273-369,
 * alphablend_tiled_slang_interp.py:70-112). Walks each pixel's list back to front from the
 * stored final state, reduces per-tet gradients across the wavefront, then adds them with
 * one atomic per (wave, tet, parameter) into grad_vertices_dev [V,3] / grad_features_dev [T,F]. */
int rm_composite_bwd(const rm_camera* cam, const int32_t* sorted_tet_dev, int64_t n_sorted,
                     const int32_t* ranges_dev, const int32_t* indices_dev, const float* vertices_dev,
                     const float* features_dev, int32_t n_features, const float* jitter_dev,
                     const float* img_dev, const float* dist_dev, const int32_t* ncontrib_dev,
                     const float* grad_img_dev, const float* grad_dist_dev,
                     float* grad_vertices_dev, float* grad_features_dev, rm_stream_t stream);

/* The same backward as rm_composite_bwd (same reference interface) in two passes: each (tile, tet)
 * adds its reduced plane-record adjoint sums (24 or 32 floats) to the tet's 128-byte row of the
 * workspace, then one thread per tet converts its row to vertex and feature gradients. Float
 * atomics on this GPU are serviced per 64-byte segment, so a row costs 2 segments per (tile, tet)
 * where the direct path's scattered vertex and feature rows cost about 6. n_tets is the number
 * of rows of indices_dev / features_dev; the workspace (rm_composite_bwd_workspace_size(n_tets)
 * bytes, 128-byte aligned) needs no initial contents. Gradients are added into
 * grad_vertices_dev / grad_features_dev as by rm_composite_bwd. */
size_t rm_composite_bwd_workspace_size(int64_t n_tets);
int rm_composite_bwd_ws(const rm_camera* cam, const int32_t* sorted_tet_dev, int64_t n_sorted,
                        const int32_t* ranges_dev, const int32_t* indices_dev, int64_t n_tets,
                        const float* vertices_dev, const float* features_dev, int32_t n_features,
                        const float* jitter_dev, const float* img_dev, const float* dist_dev,
                        const int32_t* ncontrib_dev, const float* grad_img_dev, const float* grad_dist_dev,
                        float* grad_vertices_dev, float* grad_features_dev, void* workspace_dev,
                        size_t workspace_bytes, rm_stream_t stream);

/* rm_composite_bwd_ws (the same backward, alphablend_shader_interp.slang:273-369) that walks only the
 * list entries rm_composite_fwd_flags flagged for the same forward: an entry no pixel of its tile
 * blended contributes no gradient, so staging, intersecting and reducing it is skipped. Same
 * gradients as rm_composite_bwd_ws up to the order of float additions. With tile 16 the flags are
 * not read and every entry is walked. tile_order_dev as for rm_composite_fwd_flags (4x4 tiles). */
int rm_composite_bwd_flags(const rm_camera* cam, const int32_t* sorted_tet_dev, int64_t n_sorted,
                           const int32_t* ranges_dev, const int32_t* indices_dev, int64_t n_tets,
                           const float* vertices_dev, const float* features_dev, int32_t n_features,
                           const float* jitter_dev, const float* img_dev, const float* dist_dev,
                           const int32_t* ncontrib_dev, const uint8_t* pair_flags_dev,
                           const int32_t* tile_order_dev, const float* grad_img_dev, const float* grad_dist_dev,
                           float* grad_vertices_dev, float* grad_features_dev, void* workspace_dev,
                           size_t workspace_bytes, rm_stream_t stream);

/* No counterpart in the reference: rm_composite_bwd_ws without a workspace argument, for callers
 * that cannot allocate device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc). The per-tet rows live in a static device array of
 * rm_composite_bwd_noalloc_max_tets() rows (524,288 tets, 64 MiB), one copy per GPU held from
 * library load until the process ends; n_tets above that returns RM_ERR_CAPACITY. No allocation
 * inside. All calls on one GPU share its array: calls that overlap on one GPU (different streams
 * or host threads) race and must be serialised by the caller. Same gradients as rm_composite_bwd_ws. */
int64_t rm_composite_bwd_noalloc_max_tets(void);
int rm_composite_bwd_noalloc(const rm_camera* cam, const int32_t* sorted_tet_dev, int64_t n_sorted,
                             const int32_t* ranges_dev, const int32_t* indices_dev, int64_t n_tets,
                             const float* vertices_dev, const float* features_dev, int32_t n_features,
                             const float* jitter_dev, const float* img_dev, const float* dist_dev,
                             const int32_t* ncontrib_dev, const float* grad_img_dev, const float* grad_dist_dev,
                             float* grad_vertices_dev, float* grad_features_dev, rm_stream_t stream);

/* ---- per-tet feature activation of the reference's field model ------------------------
 * Replaces activate_output (utils/model_util.py:93-105) with offset_normalize (:81-91),
 * compute_vertex_colors_from_field (:62-78), eval_sh (utils/eval_sh_py.py:137-195) and safe_div
 * (utils/safe_math.py:143-171): features_dev [T,7] = [density, softplus_10(colour), normed_grd].
 * camera_center_dev: 3 floats in device memory; density [T], rgb [T,3], grd [T,3], sh [T,K,3]
 * (K >= (sh_degree+1)^2 - 1, K <= 15, only the first (sh_degree+1)^2 - 1 used; may be null when
 * K = 0), circumcenters [T,3], vertices [V,3]. The backward writes the per-tet gradients (any may
 * be null) and adds the vertex gradients into grad_vertices_dev (caller-zeroed, may be null). */
int rm_activate_fwd(const float* camera_center_dev, const float* density_dev, const float* rgb_dev,
                    const float* grd_dev, const float* sh_dev, int32_t sh_coeffs, int32_t sh_degree,
                    const int32_t* indices_dev, const float* circumcenters_dev, const float* vertices_dev,
                    int64_t n_tets, float* features_dev, rm_stream_t stream);
int rm_activate_bwd(const float* camera_center_dev, const float* density_dev, const float* rgb_dev,
                    const float* grd_dev, const float* sh_dev, int32_t sh_coeffs, int32_t sh_degree,
                    const int32_t* indices_dev, const float* circumcenters_dev, const float* vertices_dev,
                    int64_t n_tets, const float* grad_features_dev, float* grad_density_dev, float* grad_rgb_dev,
                    float* grad_grd_dev, float* grad_sh_dev, float* grad_vertices_dev,
                    float* grad_circumcenters_dev, rm_stream_t stream);

/* ---- per-tet error votes (densification statistics) ------------------------------------
 * Replaces calc_tet_err / kern_calc_tet_err (alphablend_shader_interp.slang:433-599), launched
 * by render_err (delaunay_rasterization/internal/render_err.py:138-150). Pixel-centre rays walk
 * the same binned lists as rm_composite_fwd, starting from log-transmittance 1 (that kernel's
 * initial state) and stopping after a blend below -5.5; every blended tet receives
 * w = exp(logT) * sigma * dt. Adds into tet_err_dev [n_tets,16] f32 (caller-zeroed):
 *   0: w, 1: w*pixel_err, 2: w*pixel_err^2, 3, 4: untouched (3 carries the call's vote counts
 *   while it runs and is 0 again on return), 5: w*ssim_err,
 *   6-8: w*entry point, 9-11: w*exit point,
 * then writes 12: channel 0 and 13-15: (channels 6-8 + channels 9-11) / 2 -- the sums of the
 * reference's per-vote w and w*segment midpoint, so they stay consistent when calls accumulate --
 * and adds into tet_count_dev [n_tets,2] i32 (caller-zeroed): 0: blended pixels, 1: max int(65535 w).
 * tet_err_dev must be 16-byte aligned.
 * pixel_err_dev / ssim_err_dev are [H,W] f32. The reference's gt, output_img and n_contributors
 * arguments are read by neither kernel and do not cross this boundary. tile_order_dev as for
 * rm_composite_fwd_flags (4x4 tiles; null takes the tiles in id order). */
int rm_tet_err(const rm_camera* cam, const int32_t* sorted_tet_dev, int64_t n_sorted,
               const int32_t* ranges_dev, const int32_t* indices_dev, const float* vertices_dev,
               const float* features_dev, int32_t n_features, const float* pixel_err_dev,
               const float* ssim_err_dev, float* tet_err_dev, int32_t* tet_count_dev, int64_t n_tets,
               const int32_t* tile_order_dev, rm_stream_t stream);

/* ---- batched forward rendering (serving) -----------------------------------------------
 * No counterpart in the reference (which renders one view per call, evaluate.py:50-63 /
 * tile_benchmark.py:59-94). Many views of device-resident scenes, with mixed resolutions, render
 * as one fixed launch sequence: rm_batch_plan packs a host-side plan (copy it to device memory
 * before rendering); rm_batch_render runs projection, binning and compositing of every view with
 * all sizes taken from the plan and pair_capacity, leaving {n_visible, n_pairs} in counts_dev
 * (n_pairs > pair_capacity means pairs were dropped: render again with a larger capacity). No
 * host synchronisation and no memset inside, so the call can be captured in a HIP graph.
 * Outputs are the views' images concatenated ([sum H*W, 4], [sum H*W, 5], [sum H*W]), pixel-centre
 * rays. All views share one tile size; all scenes one feature layout. The info passed to
 * rm_batch_render sets the launch geometry: the plan's totals, or capacities at least as large
 * (kernels read the plan's actual totals from device memory and leave the rest idle), so one
 * captured graph per resolution bucket replays for any plan that fits it. */
typedef struct rm_scene {
    const int32_t* indices_dev;
    const float* vertices_dev;
    const float* features_dev;
    int64_t n_tets;
    int64_t n_verts;
    int32_t n_features;
    int32_t _pad;
} rm_scene;

typedef struct rm_batch_info {
    int64_t n_views;
    int64_t total_tets;         /* tet slots: each view's tets start at a multiple of 256 */
    int64_t total_tiles;
    int64_t total_pixels;
    int64_t total_workgroups;
    int64_t plan_bytes;
    int32_t tile;
    int32_t n_features;
    int64_t total_bin_blocks;   /* blocks of 8x8 tiles over all views (the binning's unit) */
} rm_batch_info;

int rm_batch_plan(const rm_scene* scenes, int32_t n_scenes, const rm_camera* cams,
                  const int32_t* scene_of_view, int32_t n_views, void* plan_host, size_t plan_bytes,
                  rm_batch_info* info);
size_t rm_batch_workspace_size(const rm_batch_info* info, int64_t pair_capacity);
int rm_batch_render(const void* plan_dev, const rm_batch_info* info, int64_t pair_capacity,
                    float* out_img_dev, float* out_dist_dev, int32_t* out_ncontrib_dev,
                    int64_t* counts_dev, void* workspace_dev, size_t workspace_bytes,
                    rm_stream_t stream);

/* ---- hash-grid field ----------------------------------------------------------------------
 * Replaces HashEmbedderOptimized.forward (utils/hashgrid.py:74-124): multi-level grid over
 * [0,1]^3, per-level resolution res[l] and table offset offs[l] (entries, level-concatenated
 * table of n_feat floats per entry; offs has n_levels + 1 entries); a level is dense (modulo wrap) when res^3 <= 2^log2_size,
 * else hashed with primes (1, 2654435761, 805459861). Output [N, L*n_feat] level-major. */
int rm_hashgrid_fwd(const float* x_dev, int64_t n, const float* table_dev, const int32_t* res,
                    const int64_t* offs, int32_t n_levels, int32_t n_feat, int32_t log2_size,
                    float* out_dev, rm_stream_t stream);
int rm_hashgrid_bwd(const float* x_dev, int64_t n, const float* grad_out_dev, const int32_t* res,
                    const int64_t* offs, int32_t n_levels, int32_t n_feat, int32_t log2_size,
                    float* grad_table_dev, rm_stream_t stream);
/* The same backward (the adjoint of HashEmbedderOptimized.forward, utils/hashgrid.py:117-124)
 * with a workspace: for n_feat 2, the hashed levels with res >= 512 are summed by slot ownership
 * (updates sorted into 4096-slot buckets, one workgroup per bucket, no float atomics) instead of
 * by atomics. rm_hashgrid_bwd_workspace_size returns the bytes needed (0: no level qualifies,
 * and rm_hashgrid_bwd_ws then runs rm_hashgrid_bwd's path without a workspace). The coarser levels
 * are summed on a second stream of the library's own (one per device, made on first use, default
 * priority) between an event recorded on `stream` and a wait of `stream` on its completion, so all
 * of the call's work is ordered on `stream` as for any other entry point. */
size_t rm_hashgrid_bwd_workspace_size(int64_t n, const int32_t* res, const int64_t* offs, int32_t n_levels,
                                      int32_t n_feat, int32_t log2_size);
int rm_hashgrid_bwd_ws(const float* x_dev, int64_t n, const float* grad_out_dev, const int32_t* res,
                       const int64_t* offs, int32_t n_levels, int32_t n_feat, int32_t log2_size,
                       float* grad_table_dev, void* workspace_dev, size_t workspace_bytes, rm_stream_t stream);

/* Field MLP 32 -> 64 -> 64 -> n_out (n_out <= 16), SELU hidden activations, weights row-major
 * [out,in] as torch.nn.Linear (models/ingp_color.py:141-151). precision 0 = FP32 operands,
 * 1 = BF16 operands; both contract on MFMA with FP32 accumulation. The backward recomputes the
 * hidden activations from x (the reference checkpoints the field, models/ingp_color.py:351)
 * and adds weight and bias gradients into caller-zeroed gw1..gb3; grad_x_dev may be null. */
int rm_mlp_fwd(const float* x_dev, int64_t n, const float* w1, const float* b1, const float* w2,
               const float* b2, const float* w3, const float* b3, int32_t n_out, int32_t precision,
               float* y_dev, rm_stream_t stream);
int rm_mlp_bwd(const float* x_dev, int64_t n, const float* w1, const float* b1, const float* w2,
               const float* b2, const float* w3, const float* b3, int32_t n_out, int32_t precision,
               const float* grad_y_dev, float* grad_x_dev, float* gw1, float* gb1, float* gw2,
               float* gb2, float* gw3, float* gb3, rm_stream_t stream);

/* One head of the reference model's field, n_in -> 64 -> 64 -> n_out with SELU: mk_head
 * (models/ingp_color.py:141-151) for n_in = L * n_feat in {32, 64} and n_out in [1, 48] (the
 * reference's density, colour, gradient and SH heads are 1, 3, 3 and 45 wide). Same layout,
 * precision and gradient contract as rm_mlp_fwd/rm_mlp_bwd, which are this entry point at
 * n_in = 32. */
int rm_field_head_fwd(const float* x_dev, int64_t n, int32_t n_in, const float* w1, const float* b1,
                      const float* w2, const float* b2, const float* w3, const float* b3, int32_t n_out,
                      int32_t precision, float* y_dev, rm_stream_t stream);
int rm_field_head_bwd(const float* x_dev, int64_t n, int32_t n_in, const float* w1, const float* b1,
                      const float* w2, const float* b2, const float* w3, const float* b3, int32_t n_out,
                      int32_t precision, const float* grad_y_dev, float* grad_x_dev, float* gw1, float* gb1,
                      float* gw2, float* gb2, float* gw3, float* gb3, rm_stream_t stream);

/* ---- the reference model's field glue -------------------------------------------------------
 * rm_field_query replaces the query-point steps of Model.compute_batch_features
 * (models/ingp_color.py:314-349) for tets [0, n_tets) of indices_dev: the circumradius of each
 * tet in FP64 (calculate_circumcenters_torch, utils/topo_utils.py:120-164), the anchor
 * (anchor 0: the float32 centroid, as with ablate_circumsphere; 1: the FP64 circumcenter),
 * normalised by (center, scene_scaling) and contracted with its radius by contract_mean_std
 * (utils/contraction.py:6-20). Writes x_dev [n,3] = (cv/2 + 1)/2 (float32; the circumcenter
 * anchor's FP64 coordinate is rounded), cr_dev [n] (FP64, as the reference's double radius
 * makes it) and anchor_dev [n,3] (float32). center is a host pointer to 3 floats. */
int rm_field_query(const int32_t* indices_dev, const float* vertices_dev, int64_t n_tets, const float* center,
                   float scene_scaling, int32_t anchor, float* x_dev, double* cr_dev, float* anchor_dev,
                   rm_stream_t stream);
/* The erf down-weighting of iNGPDW._encode (models/ingp_color.py:174-183): out = in * s, per
 * element of the [n, n_levels * n_feat] encoding, s = approx_erf(1 / sqrt(per_level_scale * 4 *
 * k * cr * scale_multi)) for column i with k = i mod n_levels (the reference views a row as
 * [n_feat][n_levels]), computed and applied in FP64 and rounded to float32 once. s does not
 * depend on the encoding and cr is detached, so the same call maps the encoding's gradient. */
int rm_field_downweight(const float* in_dev, const double* cr_dev, int64_t n, int32_t n_levels, int32_t n_feat,
                        double scale_multi, double per_level_scale, float* out_dev, rm_stream_t stream);
/* The output activations of iNGPDW.forward (models/ingp_color.py:186-205) from the four heads'
 * raw outputs: density = safe_exp(sigma + density_offset) (clamped to log(float32 max)),
 * rgb + 0.5, grd / sqrt(|grd|^2 + 1) and sh rounded to FP16 (sh_out_dev holds IEEE half bits).
 * The backward takes the saved density and raw grd; each incoming gradient may be null (zero);
 * g_sh_dev holds FP16 bits, as the reference's half tensor receives its gradient. */
int rm_field_outputs_fwd(const float* sigma_dev, const float* rgb_dev, const float* grd_dev, const float* sh_dev,
                         int64_t n, int32_t n_sh, float density_offset, float* density_dev, float* rgb_out_dev,
                         float* grd_out_dev, uint16_t* sh_out_dev, rm_stream_t stream);
int rm_field_outputs_bwd(const float* density_dev, const float* grd_raw_dev, const float* g_density_dev,
                         const float* g_rgb_dev, const float* g_grd_dev, const uint16_t* g_sh_dev, int64_t n,
                         int32_t n_sh, float* g_sigma_dev, float* g_rgb_raw_dev, float* g_grd_raw_dev,
                         float* g_sh_raw_dev, rm_stream_t stream);

/* ---- field glue -----------------------------------------------------------------------------
 * Replaces the PyTorch steps around the field for a list of visible tets (tet_ids_dev, n ids):
 * rm_field_points writes the query point of each listed tet, its centroid
 * (models/ingp_color.py:376, vertices[indices].mean(dim=1)) divided by scale and contracted into
 * [0,1]^3 (utils/contraction.py:6-20 position part: |z| <= 1 kept, else (2|z| - 1)/|z|^2 z; then
 * (z/2 + 1)/2), [n, 3].
 * rm_field_scatter_fwd writes the activated head outputs raw [n, n_features] into the listed tets'
 * rows of features [T, n_features] (models/ingp_color.py:200-203: density exp(raw) with raw clamped
 * at log(1e20), base colour raw + 0.5, colour-gradient rows as they are); other rows are untouched.
 * rm_field_scatter_bwd writes grad_raw [n, n_features] from grad_features [T, n_features].
 * raw (forward) and grad_raw (backward) must be 16-byte aligned. */
int rm_field_points(const int32_t* indices_dev, const float* vertices_dev, const int32_t* tet_ids_dev, int64_t n,
                    float scale, float* points_dev, rm_stream_t stream);
int rm_field_scatter_fwd(const float* raw_dev, const int32_t* tet_ids_dev, int64_t n, int32_t n_features,
                         float* features_dev, rm_stream_t stream);
int rm_field_scatter_bwd(const float* raw_dev, const int32_t* tet_ids_dev, int64_t n, int32_t n_features,
                         const float* grad_features_dev, float* grad_raw_dev, rm_stream_t stream);

/* ---- Delaunay tetrahedralisation ------------------------------------------------------------
 * Replaces the triangulation step of models/ingp_color.py:523-544: gdel3d `Del(V).compute(verts)`
 * (or scipy.spatial.Delaunay(verts).simplices when high_precision), the filter dropping tets with
 * the point at infinity, and the swap of columns 0 and 1 of negatively oriented tets
 * (topo_utils.tet_volumes, utils/topo_utils.py:9-24). points_dev is [n_points, 3] f32; tets_dev
 * receives the Delaunay tets as [n, 4] i32, each with positive volume, grouped by their smallest
 * vertex index in increasing order, at most tet_capacity of them. info_dev [8] i64 receives:
 *   0: n (the full count, which may exceed tet_capacity: run again with a larger capacity),
 *   1: Voronoi vertices found over all cells (4 n when the cells agree),
 *   2: cells over the per-cell capacity (960 vertices, 1024 planes; non-zero = failed),
 *   3: duplicate points dropped (the copy with the lowest index keeps the cell),
 *   4: predicate evaluations that needed exact arithmetic,
 *   5: of those, inputs outside its range (coordinate exponents spanning more than 99 bits),
 *   6: cells recomputed by the larger second pass, 7: largest cell (in vertices) of the first pass.
 * Exact ties (co-spherical or coplanar points) are resolved by a symbolic perturbation, so the
 * result is one valid Delaunay tetrahedralisation also for degenerate input. The point at infinity
 * is modelled by four far points (a tetrahedron 1e8 times the point-set extent): a hull tet whose
 * circumsphere would reach them (circumradius above ~5e7 extents) is replaced by tets at infinity.
 * No host synchronisation inside. */
size_t rm_delaunay_workspace_size(int64_t n_points, int64_t tet_capacity);
int rm_delaunay(const float* points_dev, int64_t n_points, int32_t* tets_dev, int64_t tet_capacity,
                int64_t* info_dev, void* workspace_dev, size_t workspace_bytes, rm_stream_t stream);
/* Host evaluation of the insphere predicate rm_delaunay uses (for tests without a GPU): 1 when
 * point 4 of xyz [5,3] is strictly inside the circumsphere of the positively oriented tet of
 * points 0-3 under the perturbation indexed by ids [5], else 0. */
int rm_delaunay_insphere_host(const int32_t* ids, const float* xyz);

#ifdef __cplusplus
}
#endif

#endif /* RMESH_H */
