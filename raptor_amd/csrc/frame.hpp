// Frame state shared by the forward pipeline (render.hip) and the adjoint
// (adjoint.hip): camera and light arguments, the 8x8-tile sample layout, the
// per-frame workspace that holds every path record of a chunk of tiles, and
// the plan that walks a frame chunk by chunk.
#pragma once

#include "rt_device.hpp"
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

// Primary ray of pixel (x, y), sample s (rt/camera.h:127-145,
// EXACT_NORMALISE): screen = ((X * x_t) + (Y * y_t)) + (Z * t) component by
// component; samples after the first of a multi-sample pixel are jittered
// inside it by the counter-based sampler.
__device__ __forceinline__ rtd::V3 primary_dir(const CamArgs &c, int x, int y, int64_t p, int s, int spp,
                                               uint64_t seed) {
    float xf = static_cast<float>(x), yf = static_cast<float>(y);
    if (spp > 1) {
        const uint64_t k = rtd::sample_key(seed, static_cast<uint32_t>(p), static_cast<uint32_t>(s), 255u);
        xf = xf + rtd::unit24(static_cast<uint32_t>(k >> 40));
        yf = yf + rtd::unit24(static_cast<uint32_t>((k >> 16) & 0xFFFFFFu));
    }
    const float x_t = (xf * c.x_inc) + c.x_m;
    const float y_t = (yf * c.y_inc) + c.y_m;
    const rtd::V3 scr = rtd::mk(((c.xx * x_t) + (c.yx * y_t)) + (c.zx * c.t), ((c.xy * x_t) + (c.yy * y_t)) + (c.zy * c.t),
                                ((c.xz * x_t) + (c.yz * y_t)) + (c.zz * c.t));
    return rtd::normalise_exact(scr);
}

// rt/ray.cc:15-20 + rt/triangle.h:464-498: move to the hit, re-project onto
// the triangle's plane, take the normal facing the ray.
__device__ __forceinline__ void surface_at(rtd::V3 o, rtd::V3 d, float t, rtd::V3 n, rtd::V3 c, rtd::V3 *dst,
                                           rtd::V3 *geo, rtd::V3 *shade) {
    rtd::V3 p = rtd::add(o, rtd::mul(d, t));
    const float denom = rtd::dot(n, d);
    const float num = rtd::dot(n, rtd::sub(c, p));
    const float dl = num / denom;
    p = rtd::add(p, rtd::mul(d, dl));
    *dst = p;
    if (denom > 0.0f) {
        *geo = rtd::neg(n);
        *shade = rtd::neg(n);
    } else {
        *geo = n;
        *shade = n;
    }
}

// Shadow ray to a light (ray::set_up(o, dst), rt/ray.h:44-54): direction and
// length from the offset origin to the light centre.
__device__ __forceinline__ void light_dir(rtd::V3 org, float4 light_pos, rtd::V3 *dn, float *len) {
    const rtd::V3 dv = rtd::sub(rtd::f4v(light_pos), org);
    *len = rtd::magnitude(dv);
    *dn = rtd::divs(dv, *len);
}

struct Rgb {
    float r, g, b;
};

// One light's Lambert term, rt/materials/phong_shader.cc:21-64 for kd = 255 *
// albedo and ks = ka = 0, with rt/light.h:44-59: adds to c unless the light is
// below the surface.
__device__ __forceinline__ void lambert_term(Rgb &c, rtd::V3 dn, float len, bool occluded, rtd::V3 nsh, float4 kd,
                                             float falloff, float4 rgb) {
    const float shade = rtd::dot(dn, nsh);
    if (shade < 0.0f) return;
    const float scr = shade * kd.x, scg = shade * kd.y, scb = shade * kd.z;
    const float magn = (occluded ? 0.0f : 1.0f) / 1.0f;
    const float dist = len * falloff;
    const float d_scale = (dist * dist) + 1.0f;
    c.r = c.r + ((rgb.x / d_scale) * (scr * magn));
    c.g = c.g + ((rgb.y / d_scale) * (scg * magn));
    c.b = c.b + ((rgb.z / d_scale) * (scb * magn));
}

// kd of a triangle from its albedo: phong_shader(rgb = 255 * albedo, kd = 1.0).
__device__ __forceinline__ float4 kd_of_albedo(const float *albedo, int32_t prim) {
    const float *a = albedo + 3 * static_cast<int64_t>(prim);
    return make_float4((a[0] * 255.0f) * 1.0f, (a[1] * 255.0f) * 1.0f, (a[2] * 255.0f) * 1.0f, 0.0f);
}

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
    // Frame-sized image and hit staging of a chunked rt_render (grow-only).
    void *frame_mem = nullptr;
    size_t frame_bytes = 0;
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
// Most tiles one chunk may hold: the hit compaction's single-block scan
// covers 1024 * kScanPer blocks of 256 samples.
int64_t scan_chunk_tiles(const FramePlan &plan);
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
