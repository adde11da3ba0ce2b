// Frame state shared by the forward pipeline (render.hip) and the adjoint
// (adjoint.hip): camera and light arguments, the 8x8-tile sample layout, the
// per-frame workspace that holds every path record of a chunk of tiles, and
// the plan that walks a frame chunk by chunk.
#pragma once

#include "rt_internal.hpp"

namespace rt {

constexpr int kMaxLights = 8;

struct CamArgs {
    float px, py, pz;
    float xx, xy, xz, yx, yy, yz, zx, zy, zz;
    float x_m, y_m, x_inc, y_inc, t;
    int32_t x_res, y_res;
};

struct LightArgs {
    float4 pos_d[kMaxLights];  // position, falloff d
    float4 rgb[kMaxLights];    // rgb / 255
    int n;
};

// Samples are laid out in 8x8-pixel tiles (the reference's packet shape,
// rt/raytracer.cc:464-483), so the 64 lanes of a wavefront trace one tile:
// slot = tile * 64 + (y % 8) * 8 + x % 8, sample i of a chunk starting at
// slot0 -> slot slot0 + i / spp.
struct TileMap {
    int32_t x_res, y_res, tiles_x;
    __host__ __device__ __forceinline__ bool pixel_of_slot(int64_t slot, int *x, int *y) const {
        const int64_t tile = slot >> 6;
        const int local = static_cast<int>(slot & 63);
        *x = static_cast<int>(tile % tiles_x) * 8 + (local & 7);
        *y = static_cast<int>(tile / tiles_x) * 8 + (local >> 3);
        return *x < x_res && *y < y_res;
    }
    __host__ __device__ __forceinline__ int64_t slot_of_pixel(int x, int y) const {
        const int64_t tile = static_cast<int64_t>(y >> 3) * tiles_x + (x >> 3);
        return (tile << 6) | ((y & 7) << 3) | (x & 7);
    }
};

// Path records of one chunk: the rays, hits, queue indices, shading normals,
// shadow rays and their visibility for the primary and the bounce level. The
// compose kernel and the adjoint both read them.
struct Workspace {
    int64_t nsamples = 0;
    int nlights = 0;
    void *mem = nullptr;
    size_t bytes = 0;
    rt_ray *rays1, *shadow1, *bounce, *shadow2;
    rt_hit *hits1, *hits2;
    int32_t *qidx1, *qidx2, *hit_sample1, *hit_sample2;
    float4 *nshade1, *nshade2;
    uint8_t *occ1, *occ2;
    float *rgb;
    rt_hit *bounce_out;
    uint32_t *blocks1, *blocks2;  // per-block hit counts, then prefixes; total at [nblocks]
    uint32_t *skeys, *svals, *skeys_t, *svals_t, *shist, *ssums;  // bounce regrouping sort
};

// Carves the arrays for ns samples and nl lights out of base; returns the size.
size_t workspace_layout(Workspace &w, char *base, int64_t ns, int nl);
// Allocates (or reuses) a workspace for ns samples and nl lights.
int32_t workspace_reserve(rt_ctx *ctx, Workspace *w, int64_t ns, int nl);
void workspace_release(Workspace *w);

// One frame: camera, lights, options, the tile grid and the chunking.
struct FramePlan {
    rt_accel *a = nullptr;
    CamArgs cam{};
    LightArgs lights{};
    TileMap tm{};
    rt_render_params p{};
    int spp = 1;
    int nlights = 0;
    int64_t npix = 0;
    int64_t tiles = 0;        // 8x8 tiles covering the image
    int64_t chunk_tiles = 0;  // tiles per chunk
};

// Per-chunk statistics (filled when the chunk is timed).
struct ChunkStats {
    uint64_t hits1 = 0, hits2 = 0;
    float ms[7] = {};  // elapsed between consecutive frame events 0..7
};

int32_t plan_frame(rt_accel *a, const rt_camera &cam, const rt_light *lights, int nlights, const rt_render_params &p,
                   FramePlan *plan);
// Largest number of tiles per chunk whose workspace fits in budget bytes
// (0 if even one tile does not fit).
int64_t tiles_within(const FramePlan &plan, int64_t budget);
// Bytes of the workspace for a chunk of ntiles tiles.
int64_t chunk_bytes(const FramePlan &plan, int64_t ntiles);
// Traces and shades the samples of tiles [tile0, tile0 + ntiles) into w
// (sized for plan.chunk_tiles); timed records the context events 0..7.
int32_t trace_chunk(const FramePlan &plan, Workspace *w, int64_t tile0, int64_t ntiles, bool timed);
// After a timed trace_chunk: hit counts and stage times (synchronises).
int32_t collect_chunk(const FramePlan &plan, Workspace *w, int64_t ntiles, ChunkStats *stats);
// Pixel radiance of the chunk's pixels into rgb (device, pixel order).
int32_t compose_chunk(const FramePlan &plan, Workspace *w, int64_t tile0, int64_t ntiles, float *rgb);

}  // namespace rt
