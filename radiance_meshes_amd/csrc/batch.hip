// Batched forward rendering of many views (mixed scenes and resolutions) as one fixed launch
// sequence: projection of every view's tets, one joint binning with per-view tile offsets, and one
// compositing launch over every view's tiles. All sizes inside come from the plan and a pair
// capacity, counts stay on the device, and buffers are cleared with kernels, so the sequence
// can be captured in a HIP graph and replayed per resolution bucket.
#include "common.h"

namespace rmesh {
namespace {

inline size_t a256(size_t x) { return (x + 255) & ~size_t(255); }

struct BatchLayout {
    size_t tiles, rect, depth, tet_blocks, sorted, ranges, order, bin, total;
};

BatchLayout batch_layout(const PlanHeader& h, int64_t cap) {
    BatchLayout L;
    size_t o = 0;
    int64_t T = h.total_tets > 0 ? h.total_tets : 1;
    L.tiles = o; o += a256(4 * (size_t)T);
    L.rect = o; o += a256(16 * (size_t)T);
    L.depth = o; o += a256(4 * (size_t)T);
    L.tet_blocks = o; o += a256(8 * (size_t)T);
    L.sorted = o; o += a256(4 * (size_t)(cap > 0 ? cap : 1));
    L.ranges = o; o += a256(8 * (size_t)(h.total_tiles > 0 ? h.total_tiles : 1));
    L.order = o; o += a256(4 * (size_t)(h.total_tiles > 0 ? h.total_tiles : 1));
    L.bin = o; o += a256(batch_bin_workspace(h, cap));
    L.total = o;
    return L;
}

PlanHeader header_of(const rm_batch_info* info) {
    PlanHeader h;
    h.n_views = (int32_t)info->n_views;
    h.tile = info->tile;
    h.n_features = info->n_features;
    h._pad = 0;
    h.total_tets = info->total_tets;
    h.total_tiles = info->total_tiles;
    h.total_pixels = info->total_pixels;
    h.total_blocks = info->total_workgroups;
    h.total_bin_blocks = info->total_bin_blocks;
    return h;
}

}  // namespace
}  // namespace rmesh

using namespace rmesh;

extern "C" int rm_batch_plan(const rm_scene* scenes, int32_t n_scenes, const rm_camera* cams,
                             const int32_t* scene_of_view, int32_t n_views, void* plan_host, size_t plan_bytes,
                             rm_batch_info* info) {
    RM_CHECK_ARG(info != nullptr, "rm_batch_plan: info is null");
    RM_CHECK_ARG(n_views >= 1 && n_scenes >= 1, "rm_batch_plan: need at least one view and one scene");
    RM_CHECK_ARG(scenes && cams && scene_of_view, "rm_batch_plan: null input");
    const size_t need = sizeof(PlanHeader) + sizeof(ViewRec) * (size_t)n_views;
    PlanHeader h = {};
    h.n_views = n_views;
    h.tile = cams[0].tile_width;
    h.n_features = scenes[scene_of_view[0] < n_scenes && scene_of_view[0] >= 0 ? scene_of_view[0] : 0].n_features;
    if (h.tile != 4 && h.tile != 8 && h.tile != 16) {
        set_error("rm_batch_plan: tile size must be 4, 8 or 16");
        return RM_ERR_UNSUPPORTED_TILE;
    }
    RM_CHECK_ARG(h.n_features == 7 || h.n_features == 13, "rm_batch_plan: n_features must be 7 or 13");
    ViewRec* views = plan_host ? reinterpret_cast<ViewRec*>(static_cast<char*>(plan_host) + sizeof(PlanHeader)) : nullptr;
    if (plan_host) RM_CHECK_ARG(plan_bytes >= need, "rm_batch_plan: plan buffer too small");
    for (int32_t v = 0; v < n_views; ++v) {
        const rm_camera& c = cams[v];
        const int32_t s = scene_of_view[v];
        RM_CHECK_ARG(s >= 0 && s < n_scenes, "rm_batch_plan: scene index out of range");
        const rm_scene& sc = scenes[s];
        RM_CHECK_ARG(c.tile_width == h.tile && c.tile_height == h.tile, "rm_batch_plan: all views need one tile size");
        RM_CHECK_ARG(sc.n_features == h.n_features, "rm_batch_plan: all scenes need one feature layout");
        RM_CHECK_ARG(c.image_height > 0 && c.image_width > 0, "rm_batch_plan: empty view");
        RM_CHECK_ARG(c.grid_height == (c.image_height + h.tile - 1) / h.tile &&
                         c.grid_width == (c.image_width + h.tile - 1) / h.tile,
                     "rm_batch_plan: grid does not match image and tile size");
        RM_CHECK_ARG(sc.n_tets >= 0 && sc.indices_dev && sc.vertices_dev && sc.features_dev,
                     "rm_batch_plan: incomplete scene");
        RM_CHECK_ARG(((uintptr_t)sc.indices_dev & 15) == 0, "rm_batch_plan: indices must be 16-byte aligned");
        const int64_t blocks = batch_workgroups_per_view(h.tile, c.grid_height, c.grid_width);
        const int64_t bin_blocks = bin_blocks_per_view(c.grid_height, c.grid_width);
        if (views) {
            ViewRec& r = views[v];
            r.cam = make_cam_args(c);
            r.indices = sc.indices_dev;
            r.verts = sc.vertices_dev;
            r.feats = sc.features_dev;
            r.tet_offset = kPlanTetAlign * ceil_div(h.total_tets, (int64_t)kPlanTetAlign);
            r.n_tets = sc.n_tets;
            r.tile_offset = h.total_tiles;
            r.pixel_offset = h.total_pixels;
            r.block_offset = h.total_blocks;
            r.n_blocks = blocks;
            r.bin_offset = h.total_bin_blocks;
        }
        h.total_tets = kPlanTetAlign * ceil_div(h.total_tets, (int64_t)kPlanTetAlign) + sc.n_tets;
        h.total_tiles += (int64_t)c.grid_height * c.grid_width;
        h.total_pixels += (int64_t)c.image_height * c.image_width;
        h.total_blocks += blocks;
        h.total_bin_blocks += bin_blocks;
    }
    RM_CHECK_ARG(h.total_tets < (1ll << 31) && h.total_tiles < (1ll << 31), "rm_batch_plan: batch too large");
    if (plan_host) *reinterpret_cast<PlanHeader*>(plan_host) = h;
    info->n_views = n_views;
    info->tile = h.tile;
    info->n_features = h.n_features;
    info->total_tets = h.total_tets;
    info->total_tiles = h.total_tiles;
    info->total_pixels = h.total_pixels;
    info->total_workgroups = h.total_blocks;
    info->total_bin_blocks = h.total_bin_blocks;
    info->plan_bytes = (int64_t)need;
    return RM_OK;
}

extern "C" size_t rm_batch_workspace_size(const rm_batch_info* info, int64_t pair_capacity) {
    if (!info) return 0;
    return batch_layout(header_of(info), pair_capacity).total;
}

// info gives the launch geometry: the plan's own totals, or larger capacities (a resolution
// bucket's slot count times per-view sizes) so one captured graph serves any plan that fits.
extern "C" int rm_batch_render(const void* plan_dev, const rm_batch_info* info, int64_t pair_capacity,
                               float* out_img_dev, float* out_dist_dev, int32_t* out_ncontrib_dev,
                               int64_t* counts_dev, void* workspace_dev, size_t workspace_bytes,
                               rm_stream_t stream) {
    RM_CHECK_ARG(plan_dev && info, "rm_batch_render: null plan");
    RM_CHECK_ARG(pair_capacity > 0 && pair_capacity < (1ll << 31), "rm_batch_render: pair capacity out of range");
    RM_CHECK_ARG(out_img_dev && out_dist_dev && out_ncontrib_dev && counts_dev && workspace_dev,
                 "rm_batch_render: null buffer");
    RM_CHECK_ARG(((uintptr_t)out_img_dev & 15) == 0 && ((uintptr_t)workspace_dev & 255) == 0,
                 "rm_batch_render: misaligned output or workspace");
    RM_CHECK_ARG(info->total_bin_blocks > 0 && info->total_bin_blocks <= info->total_tiles,
                 "rm_batch_render: total_bin_blocks must be set (rm_batch_plan's value or a capacity)");
    RM_CHECK_ARG(info->total_workgroups <= info->total_tiles,
                 "rm_batch_render: total_workgroups exceeds total_tiles (one workgroup per tile)");
    const PlanHeader h = header_of(info);
    const BatchLayout L = batch_layout(h, pair_capacity);
    if (workspace_bytes < L.total) {
        set_error("rm_batch_render: workspace too small");
        return RM_ERR_WORKSPACE_TOO_SMALL;
    }
    hipStream_t st = as_stream(stream);
    char* w = static_cast<char*>(workspace_dev);
    int32_t* tiles = reinterpret_cast<int32_t*>(w + L.tiles);
    int4* rect = reinterpret_cast<int4*>(w + L.rect);
    float* depth = reinterpret_cast<float*>(w + L.depth);
    uint2* tet_blocks = reinterpret_cast<uint2*>(w + L.tet_blocks);
    int32_t* sorted = reinterpret_cast<int32_t*>(w + L.sorted);
    int32_t* ranges = reinterpret_cast<int32_t*>(w + L.ranges);
    int32_t* order = reinterpret_cast<int32_t*>(w + L.order);   // tiles by decreasing pair count
    hipError_t e = batch_project(plan_dev, h, tiles, rect, depth, tet_blocks, st);
    if (e == hipSuccess)
        e = batch_bin(plan_dev, h, pair_capacity, tiles, rect, depth, tet_blocks, counts_dev, sorted, ranges, order,
                      w + L.bin, st);
    if (e == hipSuccess)
        e = batch_composite(plan_dev, h, sorted, pair_capacity, ranges, order, out_img_dev, out_dist_dev,
                            out_ncontrib_dev, st);
    if (e != hipSuccess) {
        set_error(std::string("rm_batch_render: ") + hipGetErrorString(e));
        return RM_ERR_HIP;
    }
    return RM_OK;
}
