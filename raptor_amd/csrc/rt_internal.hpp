// Internal types of libraptor_rt: the context, scene and acceleration
// structure behind the opaque C handles, and the wide-node layout.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <atomic>
#include <string>
#include <vector>

#include "../../include/raptor_rt.h"

namespace rt {

// Child encoding inside a wide node: >= 0 is an interior node index;
// kLeafBit | (first << 4) | (count - 1) is a leaf of count triangles starting
// at position first of the BVH-ordered triangle array; kEmpty is an unused slot.
constexpr int32_t kEmpty = -1;
constexpr uint32_t kLeafBit = 0x80000000u;
constexpr int kMaxLeafSize = 16;

__host__ __device__ inline bool is_leaf(int32_t c) { return c != kEmpty && (static_cast<uint32_t>(c) & kLeafBit); }
__host__ __device__ inline int32_t leaf_code(int32_t first, int32_t count) {
    return static_cast<int32_t>(kLeafBit | (static_cast<uint32_t>(first) << 4) | static_cast<uint32_t>(count - 1));
}
__host__ __device__ inline int32_t leaf_first(int32_t c) { return static_cast<int32_t>((static_cast<uint32_t>(c) & 0x7fffffffu) >> 4); }
__host__ __device__ inline int32_t leaf_count(int32_t c) { return static_cast<int32_t>(static_cast<uint32_t>(c) & 0xfu) + 1; }

// Wide node of W children, structure of arrays so one lane fetches the whole
// node with 16-byte loads: lo_x[W], hi_x[W], lo_y[W], hi_y[W], lo_z[W], hi_z[W],
// child[W]. 112 bytes for W = 4 (padded to 128), 224 for W = 8 (padded to 256).
// An empty slot has child kEmpty and the inverted box lo = +inf, hi = -inf.
template <int W>
struct alignas(W == 4 ? 128 : 256) WideNode {
    float lo_x[W], hi_x[W], lo_y[W], hi_y[W], lo_z[W], hi_z[W];
    int32_t child[W];
    int32_t pad[W == 4 ? 4 : 8];
};
static_assert(sizeof(WideNode<4>) == 128, "BVH4 node is one 128-byte line");
static_assert(sizeof(WideNode<8>) == 256, "BVH8 node is two 128-byte lines");

// BVH-ordered triangle for the leaf test: a (xyz) + emplace id (w bits),
// e1 = c - a, e2 = b - a. 48 bytes.
struct alignas(16) LeafTri {
    float4 a_id;
    float4 e1;
    float4 e2;
};

}  // namespace rt

namespace rt {

struct GlareCache;  // glare.hip: filters and their spectra, per light level

// Scratch memory of BVH builds, kept on the context between builds: device
// chunks handed out by a bump pointer. A build that outgrows the chunks takes
// another; the next build merges them into one chunk of the high-water mark,
// so repeated builds of one scene allocate nothing.
struct BuildArena {
    std::vector<std::pair<void *, size_t>> chunks;
    size_t cur = 0, used = 0, taken = 0, high = 0;
    hipError_t begin();
    hipError_t take(void **p, size_t bytes);
    void release();
};

}  // namespace rt

struct rt_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // Second stream (default priority) for a launch that can run beside the
    // main stream's: the frame's shadow any-hit launch overlaps its bounce
    // closest-hit launch. fork / join order the two streams.
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    std::string last_error;
    bool counting = false;
    hipEvent_t ev[16] = {};
    // Device counters when counting is on: (node visits, triangle tests) for
    // slots 0..3, the stack-overflow flag at index 15.
    unsigned long long *d_counters = nullptr;
    unsigned long long *d_frame_rays = nullptr;  // rays of RT_ASYNC frames by stage (4)
    uint32_t *d_queue = nullptr;  // chunk counters of the persistent trace kernels (kQueueWords)
    void *d_static = nullptr;     // this device's static scratch (rt_render_static), not owned
    rt::GlareCache *glare = nullptr;  // glare filters (glare.hip)
    rt::BuildArena arena;         // build scratch (bvh_build.hip)
    int count_slot = 0;
    std::atomic<int> refs{1};  // the caller's handle + one per scene
};

namespace rt {
// Words of rt_ctx::d_queue: eight per-XCD chunk counters 128 bytes apart and
// the persistent trace kernel's exit counter.
constexpr int kQueueWords = 1024;
// Device material row: {ka rgb, shininess}, {kd rgb, transmittance},
// {ks rgb, refractive index}, {reflectance, kind, chroma x, chroma y},
// {diffuse, roughness, extinction, 0}, {tex ka, kd, ks, ns},
// {tex refl, d, mapper kd, mapper ks}, {clrh, clrf, layers 0, 1},
// {layers 2, 3, 4, 0} (texture slots 1 + map, layer lists first << 8 |
// count, as int bits).
constexpr int kMaterialFloat4s = 9;
// Largest Mandelbrot iteration limit and Perlin octave count a scene takes:
// a hit's shading loops over them.
constexpr int kMaxMandelbrotIterations = 65536;
constexpr int kMaxPerlinOctaves = 32;
struct DevTexMap;  // textures.hpp
}  // namespace rt

struct rt_scene {
    rt_ctx *ctx = nullptr;
    int64_t n = 0;
    float *d_abc = nullptr;    // n x 9, emplace order
    float4 *d_n = nullptr;     // shading normal (xyz), flags in w bits
    float4 *d_c = nullptr;     // vertex c (xyz) for the plane re-projection
    float4 *d_kd = nullptr;    // kd = (255 * albedo) * 1.0
    uint8_t *d_flags = nullptr;
    // Materials (rt_scene_set_materials): kMaterialFloat4s float4 per
    // material, the material of each triangle (-1: Lambert from the albedo);
    // both null when none are set.
    float4 *d_mat = nullptr;
    int32_t *d_tri_mat = nullptr;
    int32_t nmat = 0;
    // Vertex normals (rt_scene_set_vertex_normals): 9 floats per triangle,
    // FLT_MAX as the first marks a triangle without them; null when none.
    float *d_vn = nullptr;
    // Texture maps (rt_scene_set_textures): maps with their images' shapes,
    // texels, mapper_shader layer lists; texture coordinates
    // (rt_scene_set_texture_coords), 6 floats per triangle. Null when none.
    rt::DevTexMap *d_maps = nullptr;
    float *d_texels = nullptr;
    int32_t *d_layers = nullptr;
    int32_t nmaps = 0, nlayers = 0;
    // What the current materials reference: maps 1..maps_used, layers
    // [0, layers_used).
    int32_t maps_used = 0, layers_used = 0;
    float *d_vt = nullptr;
    std::atomic<int> refs{1};  // the caller's handle + one per accel
};

struct rt_accel {
    rt_scene *scene = nullptr;
    int width = 4;
    int leaf_size = 4;
    int sah_bins = 0;          // the build's binned-SAH refinement (0: LBVH)
    int lds_stack = 16;        // traversal stack entries kept in LDS (4, 8 or 16)
    int32_t root = 0;          // child code of the root (interior index or leaf)
    int64_t nodes = 0, leaves = 0;
    int depth = 0;
    int binary_depth = 0;
    void *d_nodes = nullptr;   // WideNode<W>[nodes]
    rt::LeafTri *d_tris = nullptr;  // BVH order
    float build_ms = 0.0f;
    float refit_ms = 0.0f;
    float sah_cost = 0.0f;
    float box_pad = 0.0f;
    float near_limit = 0.0f;
    float scene_lo[3] = {0, 0, 0}, scene_inv[3] = {0, 0, 0};  // scene box origin and 1/extent
    void *frame_ws = nullptr;  // frame workspace (render.hip)
    int64_t triangles = 0;
    std::vector<int64_t> level_start;  // wide nodes of tree level L: [level_start[L], level_start[L + 1])
    float *d_refit = nullptr;          // refit scratch: scene bounds, padding, partial bounds
    std::atomic<int> refs{1};  // the caller's handle + one per batch item
};

namespace rt {

// Handles are reference counted so destroy calls may come in any order: an
// object is freed once its own handle and every object built on it are gone.
void release_ctx(rt_ctx *ctx);
void release_scene(rt_scene *scene);
void release_accel(rt_accel *accel);

// Error helpers: record text on the context and return a code.
int32_t fail(rt_ctx *ctx, int32_t code, const std::string &msg);

#define RT_HIP_CHECK(ctx, expr)                                                                          \
    do {                                                                                                 \
        hipError_t _e = (expr);                                                                          \
        if (_e != hipSuccess) {                                                                          \
            return ::rt::fail((ctx), (_e == hipErrorOutOfMemory) ? RT_ERR_OOM : RT_ERR_HIP,              \
                              std::string(#expr) + ": " + hipGetErrorString(_e));                      \
        }                                                                                                \
    } while (0)

// Build entry (bvh_build.hip).
int32_t build_accel(rt_scene *scene, const rt_build_params &p, rt_accel **out);

// The reference builder's binary tree (aac_build.hip), on the host: node
// boxes, children (interior: left and right node id; leaf: the range [b, e)
// of indirect), leaf flags, the root id, the final triangle order, and the
// Morton codes (emplace order) with the order the sort left them in.
struct AacTree {
    int64_t n = 0, nodes = 0;
    int32_t root = -1;
    int32_t split_nodes = 0;  // nodes of the top-down split
    float build_ms = 0.0f;
    std::vector<uint32_t> codes, order;
    std::vector<int32_t> indirect;
    std::vector<float4> lo, hi;
    std::vector<int2> kid;
    std::vector<uint8_t> leaf;
    std::vector<float4> tri_lo, tri_hi;  // triangle boxes, emplace order
};
int32_t build_aac_tree(rt_scene *scene, float alpha, float delta, AacTree *out);
// A wide BVH over the reference builder's tree (rt_build_params.builder
// RT_BUILDER_AAC).
int32_t build_accel_aac(rt_scene *scene, const rt_build_params &p, rt_accel **out);
// LeafTri records of the triangles in the given order (bvh_build.hip).
hipError_t make_leaf_tris(const float *abc, const uint8_t *flags, const uint32_t *order, int64_t n, LeafTri *out,
                          hipStream_t s);

// Trace launchers (trace.hip). Rays, hits and flags are device pointers.
// stage tags the kernel symbol: 0 primary, 1 shadow, 2 bounce, 3 bounce
// shadow, 4 / 5 direct API calls (closest / any hit).
int32_t launch_closest(rt_accel *a, const rt_ray *d_rays, int64_t n, rt_hit *d_hits, hipStream_t s, int stage = 4);
int32_t launch_occluded(rt_accel *a, const rt_ray *d_rays, int64_t n, uint8_t *d_occ, hipStream_t s, int stage = 5);
// Variants that read the ray count from device memory (queues of unknown
// size): n = d_count[0] * mul, at most max_n.
int32_t launch_closest_q(rt_accel *a, const rt_ray *d_rays, const uint32_t *d_count, uint32_t mul, int64_t max_n,
                         rt_hit *d_hits, hipStream_t s, int stage, const uint32_t *order = nullptr);
int32_t launch_occluded_q(rt_accel *a, const rt_ray *d_rays, const uint32_t *d_count, uint32_t mul, int64_t max_n,
                          uint8_t *d_occ, hipStream_t s, int stage);

// Frame (render.hip).
void free_frame_workspace(rt_accel *a);
int32_t refit_accel(rt_accel *a);
// fixed_ws: a workspace of fixed_bytes to use instead of the accel's own
// (rt_render_static); the outputs must then be device pointers.
int32_t render_frame(rt_accel *a, const rt_camera &cam, const rt_light *lights, int nlights,
                     const rt_render_params &p, float *rgb, rt_hit *primary, rt_hit *bounce, rt_stats *stats,
                     void *fixed_ws = nullptr, size_t fixed_bytes = 0);
// The static scratch of rt_render_static (render.hip): kStaticScratchBytes
// per GPU; its device address on the current device.
constexpr size_t kStaticScratchBytes = size_t(64) << 20;
int32_t static_scratch_address(rt_ctx *ctx, void **ptr);
// Camera post-process (postprocess.hip).
int32_t image_tone_map(rt_ctx *ctx, float *rgb, int32_t x_res, int32_t y_res, rt_tone_params *tp, uint32_t flags,
                       float *ms);
int32_t image_gamma(rt_ctx *ctx, float *rgb, int64_t npix, float gamma, uint32_t flags, float *ms);
// Glare filter (glare.hip): perform_glare_filter on a tone-mapped device
// image at adapted luminance yi; the cached filter of a light level; release.
int32_t apply_glare(rt_ctx *ctx, float *img, int x_res, int y_res, float yi, float x_angle);
int32_t glare_filter(rt_ctx *ctx, int x_res, int y_res, float x_angle, int level, const float **gf);
void release_glare(rt_ctx *ctx);
int32_t image_clip(rt_ctx *ctx, const float *rgb, int32_t x_res, int32_t y_res, int32_t xs, int32_t ys, int32_t bgr,
                   uint8_t *out, uint32_t flags, float *ms);
int32_t write_png(const char *path, const uint8_t *rgb, int32_t x_res, int32_t y_res);
int32_t write_jpeg(rt_ctx *ctx, const char *path, const uint8_t *rgb, int32_t x_res, int32_t y_res, int32_t quality,
                   uint32_t flags, float *ms);
int32_t write_tga(const char *path, const uint8_t *bgr, int32_t x_res, int32_t y_res);
// Frame of a scene with phong materials (materials.hip).
int32_t render_frame_materials(rt_accel *a, const rt_camera &cam, const rt_light *lights, int nlights,
                               const rt_render_params &p, float *rgb, rt_hit *primary, rt_stats *stats);

}  // namespace rt
