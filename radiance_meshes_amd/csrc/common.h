// Shared device/host helpers for the rmesh HIP library (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "../../include/rmesh.h"

namespace rmesh {

constexpr int kWave = 64;

// Thread-local error message behind rm_last_error().
void set_error(const std::string& msg);

#define RM_CHECK_ARG(cond, msg)                         \
    do {                                                \
        if (!(cond)) {                                  \
            ::rmesh::set_error(msg);                    \
            return RM_ERR_INVALID_ARGUMENT;             \
        }                                               \
    } while (0)

#define RM_CHECK_LAUNCH(what)                                                         \
    do {                                                                              \
        hipError_t e_ = hipGetLastError();                                            \
        if (e_ != hipSuccess) {                                                       \
            ::rmesh::set_error(std::string(what) + ": " + hipGetErrorString(e_));     \
            return RM_ERR_HIP;                                                        \
        }                                                                             \
    } while (0)

inline hipStream_t as_stream(rm_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Camera as passed to kernels by value. ray_rot is transpose(w2v[:3,:3]) (the view->world
// rotation of camera.slang:200-207), precomputed on the host.
struct CamArgs {
    float w2v[16];
    float K[9];
    float cam_pos[3];
    float dist[4];
    float ray_rot[9];
    float min_t;
    int H, W, gh, gw, th, tw, type;
    int distorted;   // dot(dist, dist) >= 1e-9 (camera.slang:244,360)
};

inline CamArgs make_cam_args(const rm_camera& c) {
    CamArgs a;
    for (int i = 0; i < 16; ++i) a.w2v[i] = c.w2v[i];
    for (int i = 0; i < 9; ++i) a.K[i] = c.K[i];
    for (int i = 0; i < 3; ++i) a.cam_pos[i] = c.cam_pos[i];
    for (int i = 0; i < 4; ++i) a.dist[i] = c.dist[i];
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k) a.ray_rot[3 * r + k] = c.w2v[4 * k + r];
    a.min_t = c.min_t;
    a.H = c.image_height;
    a.W = c.image_width;
    a.gh = c.grid_height;
    a.gw = c.grid_width;
    a.th = c.tile_height;
    a.tw = c.tile_width;
    a.type = c.camera_type;
    const float* k = c.dist;
    a.distorted = !((k[0] * k[0] + k[1] * k[1] + k[2] * k[2] + k[3] * k[3]) < 1e-9f);
    return a;
}

// One view of a batched render (device-resident plan record, packed by rm_batch_plan).
// A view's first global tet in a batch plan is a multiple of this (the batched projection's
// workgroup size), so that no projection workgroup straddles two views; the slots between one
// view's last tet and the next view's first stay empty.
constexpr int kPlanTetAlign = 256;

struct ViewRec {
    CamArgs cam;
    const int32_t* indices;    // [n_tets, 4] of the view's scene
    const float* verts;        // [V, 3]
    const float* feats;        // [n_tets, F]
    int64_t tet_offset;        // first global tet id of this view in the batch (kPlanTetAlign-aligned)
    int64_t n_tets;
    int64_t tile_offset;       // first global tile id
    int64_t pixel_offset;      // first pixel of the view in the batch image buffers
    int64_t block_offset;      // first compositing workgroup of the view
    int64_t n_blocks;
    int64_t bin_offset;        // first global 8x8-tile binning block of the view
};

struct PlanHeader {
    int32_t n_views, tile, n_features, _pad;
    int64_t total_tets, total_tiles, total_pixels, total_blocks, total_bin_blocks;
};

// A wave-uniform pointer read from memory (a plan record) marked as a global-memory pointer, so
// that the compiler loads through it with global instructions, as through a kernel argument,
// rather than flat ones (which also count against the LDS wait counter: an LDS wait then waits for
// them too). The empty asm statement (no instruction) keeps the compiler from folding the two
// address-space casts away; its operand lives in scalar registers, hence wave-uniform pointers only.
template <typename T>
__device__ inline const T* uniform_global_ptr(const T* p) {
    using G = const __attribute__((address_space(1))) T*;
    G g = (G)p;
    asm volatile("" : "+s"(g));
    return (const T*)g;
}

__host__ __device__ inline const ViewRec* plan_views(const void* plan) {
    return reinterpret_cast<const ViewRec*>(static_cast<const char*>(plan) + sizeof(PlanHeader));
}

// Batched stages (defined in project.hip, sort.hip, composite.hip), all on one stream, sized by
// the plan and a pair capacity: no host synchronisation inside.
hipError_t batch_project(const void* plan_dev, const PlanHeader& h, int32_t* tiles, int4* rect, float* depth,
                         uint2* tet_blocks, hipStream_t st);
size_t batch_bin_workspace(const PlanHeader& h, int64_t pair_cap);
hipError_t batch_bin(const void* plan_dev, const PlanHeader& h, int64_t pair_cap, const int32_t* tiles,
                     const int4* rect, const float* depth, const uint2* tet_blocks, int64_t* counts_dev,
                     int32_t* sorted_tet, int32_t* ranges, int32_t* tile_order_out, void* ws, hipStream_t st);
int batch_workgroups_per_view(int tile, int gh, int gw);
// The binning's unit: blocks of kBinBlockTiles x kBinBlockTiles tiles over a view's tile grid.
constexpr int kBinBlockTiles = 8;
__host__ __device__ inline int32_t bin_blocks_per_row(int gw) { return (gw + kBinBlockTiles - 1) / kBinBlockTiles; }
inline int64_t bin_blocks_per_view(int gh, int gw) { return (int64_t)bin_blocks_per_row(gw) * bin_blocks_per_row(gh); }
hipError_t batch_composite(const void* plan_dev, const PlanHeader& h, const int32_t* sorted_tet,
                           int64_t pair_cap, const int32_t* ranges, const int32_t* tile_order, float* img,
                           float* dist, int32_t* nc, hipStream_t st);

}  // namespace rmesh
