// Batched inverse rendering: many scenes' frames, their adjoint and an Adam
// step on the per-triangle albedo, one persistent launch per step.
//
// The reference renders one scene forward per ray_tracer() call
// (rt/raytracer.cc:563-598) and has no adjoint. Here k_path is a path
// megakernel: a fixed grid of single-wave workgroups takes (scene, 8x8 tile)
// work items from one global counter (scenes of every size share the launch);
// each lane owns one pixel and traces all of its samples: primary ray, one
// shadow ray per light at every vertex, up to `bounces` cosine-weighted
// diffuse bounces (the same arithmetic as the wavefront frame in render.hip,
// so one bounce reproduces rt_render bit for bit). With the loss
// L = sum (I - I_target)^2 the pixel's gradient 2 (I - I_target) is known as
// soon as its samples are done, so the adjoint runs in the same lane from the
// path vertices it kept in private memory: no path records are written and
// nothing is re-traced. Per sample, with D_k = kd_k s_k the direct light at
// vertex k, T_k the throughput (product of kd / 255 over earlier vertices)
// and R_k = D_k + (kd_k / 255) R_{k+1},
//   dc / dkd_k = T_k (s_k + R_{k+1} / 255),
// accumulated into the albedo gradient with float atomics. k_adam then
// updates the albedo (clamped to [0, 1]) and clears the gradient; the reset,
// the path launch, the update and the step counter are captured in one
// hipGraph per step.
#include <algorithm>
#include <cmath>
#include <vector>

#include "frame.hpp"
#include "traverse.hpp"

struct rt_invrender {
    rt_ctx *ctx = nullptr;
    int width = 4;
    int lds_stack = 8;
    rt_invrender_params params{};
    std::vector<rt_invrender_item> items;
    std::vector<int64_t> param_offset;  // first albedo scalar of each scene in the Adam state
    int64_t nparams = 0;
    int nscenes = 0;
    uint32_t total_chunks = 0;
    int grid = 0;
    rt::LightArgs lights{};
    void *d_scenes = nullptr;           // PathScene[nscenes]
    uint32_t *d_counter = nullptr;      // [0] queue head, [1] stack overflow, [2..3] Adam step (int64)
    unsigned long long *d_rays = nullptr;
    double *d_loss = nullptr;           // per scene
    float *d_m = nullptr, *d_v = nullptr;  // Adam moments over all scenes' albedo
    void *d_adam = nullptr;             // AdamSeg[nscenes]
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

namespace rt {

namespace {

constexpr int kMaxSpp = 8;
constexpr int kMaxBounces = 3;
constexpr int kMaxVerts = kMaxSpp * (kMaxBounces + 1);

struct PathScene {
    const void *nodes;
    const LeafTri *tris;
    int32_t root;
    float near_limit;
    const float4 *tri_n, *tri_c;
    const float *albedo;
    float *d_albedo;
    const float *target;
    float *image;
    CamArgs cam;
    TileMap tm;
    uint32_t chunk_begin;
    uint32_t pad;
};

struct PathArgs {
    const PathScene *scenes;
    int scene_lo, scene_hi;   // scenes [lo, hi) of this launch
    uint32_t chunk_lo, chunk_hi;
    uint32_t *counter;
    uint32_t *overflow;
    unsigned long long *rays;
    double *loss;
    LightArgs lights;
    int spp;
    int bounces;
    uint64_t seed;
    int adjoint;
};

// One path vertex kept for the adjoint.
struct Vertex {
    float ox, oy, oz;      // shadow-ray origin
    float nx, ny, nz;      // shading normal
    int32_t prim;
    uint32_t vis;          // bit l: light l unoccluded
};

// Queue head = *first; with loss != nullptr also a new frame set: clears the
// per-scene losses and the overflow flag. (A fill kernel, not a memset, so
// that it can sit in a captured graph.)
__global__ void k_path_reset(uint32_t *counter, const uint32_t *first, double *loss, int nscenes) {
    const int t = threadIdx.x;
    if (t == 0) counter[0] = first[0];
    if (!loss) return;
    if (t == 0) counter[1] = 0u;
    for (int k = t; k < nscenes; k += blockDim.x) loss[k] = 0.0;
}

// At its natural 134 VGPRs the kernel fits 3 waves per SIMD; capping it at
// 128 (4 waves, a few spills) measured 21% faster on 512x512x2-spp frames of
// 1M-4M-triangle scenes, 96 (5 waves, 40 spills) 23%.
template <int W, int kLdsDepth>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4, 4))) k_path(PathArgs a) {
    const int lane = threadIdx.x;
    Vertex verts[kMaxVerts];
    unsigned long long my_rays = 0;
    bool overflow = false;
    const int levels = a.bounces + 1;
    while (true) {
        uint32_t g = 0;
        if (lane == 0) g = atomicAdd(a.counter, 1u);
        g = __builtin_amdgcn_readfirstlane(__shfl(g, 0, 64));
        if (g >= a.chunk_hi) break;
        int lo = a.scene_lo, hi = a.scene_hi - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (a.scenes[mid].chunk_begin <= g)
                lo = mid;
            else
                hi = mid - 1;
        }
        const PathScene &sc = a.scenes[lo];
        const WideNode<W> *nodes = reinterpret_cast<const WideNode<W> *>(sc.nodes);
        const int64_t slot = static_cast<int64_t>(g - sc.chunk_begin) * 64 + lane;
        int x = 0, y = 0;
        const bool valid = sc.tm.pixel_of_slot(slot, &x, &y);
        const int64_t p = static_cast<int64_t>(y) * sc.tm.x_res + x;
        Rgb acc{0.0f, 0.0f, 0.0f};
        int nverts[kMaxSpp];
        for (int s = 0; s < a.spp; ++s) {
            nverts[s] = 0;
            rtd::V3 o = rtd::mk(sc.cam.px, sc.cam.py, sc.cam.pz);
            rtd::V3 d = primary_dir(sc.cam, x, y, p, s, a.spp, a.seed);
            Rgb c{0.0f, 0.0f, 0.0f};
            Rgb T{1.0f, 1.0f, 1.0f};
            bool alive = valid;
            for (int k = 0; k < levels; ++k) {
                float4 hit = make_float4(rtd::kMaxDist, 0.0f, 0.0f, __int_as_float(-1));
                if (alive) {
                    uint32_t vv = 0, tt = 0;
                    trace_ray<W, true, false, kLdsDepth>(nodes, sc.tris, sc.root, sc.near_limit, o, d, rtd::kEpsilon,
                                                         rtd::kMaxDist, lane, hit, vv, tt, overflow);
                    ++my_rays;
                }
                const int32_t prim = __float_as_int(hit.w);
                alive = alive && prim >= 0;
                if (!alive) break;
                rtd::V3 dst, geo, shn;
                surface_at(o, d, hit.x, rtd::f4v(sc.tri_n[prim]), rtd::f4v(sc.tri_c[prim]), &dst, &geo, &shn);
                const rtd::V3 org = rtd::offset_start_point(dst, geo, 1);
                const float4 kd = kd_of_albedo(sc.albedo, prim);
                uint32_t vis = 0;
                Rgb dk{0.0f, 0.0f, 0.0f};
                for (int l = 0; l < a.lights.n; ++l) {
                    rtd::V3 dn;
                    float len;
                    light_dir(org, a.lights.pos_d[l], &dn, &len);
                    float4 sh = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(-1));
                    uint32_t vv = 0, tt = 0;
                    trace_ray<W, false, false, kLdsDepth>(nodes, sc.tris, sc.root, sc.near_limit, org, dn,
                                                          rtd::kEpsilon, len, lane, sh, vv, tt, overflow);
                    ++my_rays;
                    const bool occ = __float_as_int(sh.w) >= 0;
                    if (!occ) vis |= 1u << l;
                    lambert_term(dk, dn, len, occ, shn, kd, a.lights.pos_d[l].w, a.lights.rgb[l]);
                }
                dk.r = dk.r + 0.0f;  // + ka (zero)
                dk.g = dk.g + 0.0f;
                dk.b = dk.b + 0.0f;
                if (k == 0) {
                    c = dk;
                    T = Rgb{kd.x / 255.0f, kd.y / 255.0f, kd.z / 255.0f};
                } else {
                    c.r = c.r + (T.r * dk.r);
                    c.g = c.g + (T.g * dk.g);
                    c.b = c.b + (T.b * dk.b);
                    T = Rgb{T.r * (kd.x / 255.0f), T.g * (kd.y / 255.0f), T.b * (kd.z / 255.0f)};
                }
                if (a.adjoint) {
                    Vertex &v = verts[s * (kMaxBounces + 1) + k];
                    v.ox = org.x; v.oy = org.y; v.oz = org.z;
                    v.nx = shn.x; v.ny = shn.y; v.nz = shn.z;
                    v.prim = prim;
                    v.vis = vis;
                    nverts[s] = k + 1;
                }
                if (k + 1 < levels) {
                    o = org;
                    d = rtd::cosine_direction(rtd::sample_key(a.seed, static_cast<uint32_t>(p), static_cast<uint32_t>(s),
                                                              static_cast<uint32_t>(k + 1)),
                                              shn);
                }
            }
            acc.r = acc.r + c.r;
            acc.g = acc.g + c.g;
            acc.b = acc.b + c.b;
        }
        const float fs = static_cast<float>(a.spp);
        const Rgb I{acc.r / fs, acc.g / fs, acc.b / fs};
        if (valid && sc.image) {
            sc.image[3 * p + 0] = I.r;
            sc.image[3 * p + 1] = I.g;
            sc.image[3 * p + 2] = I.b;
        }
        if (!a.adjoint) continue;
        // L = sum (I - target)^2: dL/dI = 2 (I - target); I = (sum_s c_s) / spp.
        double l2 = 0.0;
        if (valid) {
            const float er = I.r - sc.target[3 * p + 0], eg = I.g - sc.target[3 * p + 1],
                        eb = I.b - sc.target[3 * p + 2];
            l2 = static_cast<double>(er) * er + static_cast<double>(eg) * eg + static_cast<double>(eb) * eb;
            const Rgb w{(2.0f * er) / fs, (2.0f * eg) / fs, (2.0f * eb) / fs};
            for (int s = 0; s < a.spp; ++s) {
                const int nv = nverts[s];
                // Throughput into each vertex.
                Rgb Tk[kMaxBounces + 1];
                Tk[0] = Rgb{1.0f, 1.0f, 1.0f};
                for (int k = 1; k < nv; ++k) {
                    const float4 kp = kd_of_albedo(sc.albedo, verts[s * (kMaxBounces + 1) + k - 1].prim);
                    Tk[k] = Rgb{Tk[k - 1].r * (kp.x / 255.0f), Tk[k - 1].g * (kp.y / 255.0f),
                                Tk[k - 1].b * (kp.z / 255.0f)};
                }
                Rgb R{0.0f, 0.0f, 0.0f};  // R_{k+1}
                for (int k = nv - 1; k >= 0; --k) {
                    const Vertex &v = verts[s * (kMaxBounces + 1) + k];
                    const rtd::V3 org = rtd::mk(v.ox, v.oy, v.oz);
                    const rtd::V3 shn = rtd::mk(v.nx, v.ny, v.nz);
                    const float4 kd = kd_of_albedo(sc.albedo, v.prim);
                    // s_k = sum_l V S I (the direct light with kd = 1).
                    Rgb sk{0.0f, 0.0f, 0.0f};
                    const float4 one = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
                    for (int l = 0; l < a.lights.n; ++l) {
                        rtd::V3 dn;
                        float len;
                        light_dir(org, a.lights.pos_d[l], &dn, &len);
                        lambert_term(sk, dn, len, ((v.vis >> l) & 1u) == 0, shn, one, a.lights.pos_d[l].w,
                                     a.lights.rgb[l]);
                    }
                    const Rgb gk{w.r * Tk[k].r * (sk.r + R.r / 255.0f), w.g * Tk[k].g * (sk.g + R.g / 255.0f),
                                 w.b * Tk[k].b * (sk.b + R.b / 255.0f)};
                    float *ga = sc.d_albedo + 3 * static_cast<int64_t>(v.prim);
                    // kd = 255 * albedo
                    unsafeAtomicAdd(ga + 0, 255.0f * gk.r);
                    unsafeAtomicAdd(ga + 1, 255.0f * gk.g);
                    unsafeAtomicAdd(ga + 2, 255.0f * gk.b);
                    R = Rgb{kd.x * sk.r + (kd.x / 255.0f) * R.r, kd.y * sk.g + (kd.y / 255.0f) * R.g,
                            kd.z * sk.b + (kd.z / 255.0f) * R.b};
                }
            }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) l2 += __shfl_xor(l2, off, 64);
        if (lane == 0) unsafeAtomicAdd(a.loss + lo, l2);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) my_rays += __shfl_xor(my_rays, off, 64);
    if (lane == 0) atomicAdd(a.rays, my_rays);
    if (overflow) atomicOr(a.overflow, 1u);
}

struct AdamSeg {
    float *albedo;
    float *grad;
    int64_t first;   // offset into the moment arrays
    int64_t count;   // scalars
};

// Adam (Kingma & Ba 2015) over every scene's albedo; the step number t lives
// on the device so that a captured graph can be replayed step after step.
__global__ void __launch_bounds__(256) k_adam(const AdamSeg *__restrict__ segs, int nsegs, int64_t total,
                                              float *__restrict__ m, float *__restrict__ v,
                                              const int64_t *__restrict__ step, float lr, float b1, float b2,
                                              float eps) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= total) return;
    int lo = 0, hi = nsegs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (segs[mid].first <= i)
            lo = mid;
        else
            hi = mid - 1;
    }
    const AdamSeg sg = segs[lo];
    const int64_t j = i - sg.first;
    const float g = sg.grad[j];
    const float t = static_cast<float>(step[0] + 1);
    const float mi = b1 * m[i] + (1.0f - b1) * g;
    const float vi = b2 * v[i] + (1.0f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / (1.0f - powf(b1, t));
    const float vh = vi / (1.0f - powf(b2, t));
    const float x = sg.albedo[j] - lr * mh / (sqrtf(vh) + eps);
    sg.albedo[j] = fminf(fmaxf(x, 0.0f), 1.0f);
    sg.grad[j] = 0.0f;
}

__global__ void k_step_inc(int64_t *step) {
    if (threadIdx.x == 0) step[0] += 1;
}

__global__ void __launch_bounds__(256) k_zero_f32(float *p, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

template <int W>
void launch_path_w(int depth, dim3 grid, const PathArgs &pa, hipStream_t s) {
    if (depth <= 4)
        hipLaunchKernelGGL((k_path<W, 4>), grid, dim3(64), 0, s, pa);
    else if (depth <= 8)
        hipLaunchKernelGGL((k_path<W, 8>), grid, dim3(64), 0, s, pa);
    else
        hipLaunchKernelGGL((k_path<W, 16>), grid, dim3(64), 0, s, pa);
}

int path_occupancy(int width, int depth) {
    int n = 0;
    hipError_t e;
    if (width == 4)
        e = depth <= 4 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_path<4, 4>, 64, 0)
            : depth <= 8 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_path<4, 8>, 64, 0)
                         : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_path<4, 16>, 64, 0);
    else
        e = depth <= 4 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_path<8, 4>, 64, 0)
            : depth <= 8 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_path<8, 8>, 64, 0)
                         : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_path<8, 16>, 64, 0);
    return e == hipSuccess ? n : 0;
}

PathArgs path_args(rt_invrender *r, int adjoint) {
    PathArgs pa{};
    pa.scenes = reinterpret_cast<const PathScene *>(r->d_scenes);
    pa.scene_lo = 0;
    pa.scene_hi = r->nscenes;
    pa.chunk_lo = 0;
    pa.chunk_hi = r->total_chunks;
    pa.counter = r->d_counter;
    pa.overflow = r->d_counter + 1;
    pa.rays = r->d_rays;
    pa.loss = r->d_loss;
    pa.lights = r->lights;
    pa.spp = r->params.spp;
    pa.bounces = r->params.bounces;
    pa.seed = r->params.seed;
    pa.adjoint = adjoint;
    return pa;
}

// The frame(s) of one step: persistent (one launch over all scenes) or plain
// (one launch per scene, one workgroup per tile).
void enqueue_frames(rt_invrender *r, int adjoint, bool plain, const std::vector<uint32_t> &chunk_first,
                    hipStream_t s) {
    PathArgs pa = path_args(r, adjoint);
    if (!plain) {
        hipLaunchKernelGGL(k_path_reset, dim3(1), dim3(256), 0, s, r->d_counter, r->d_counter + 4, r->d_loss,
                           r->nscenes);
        if (r->width == 4)
            launch_path_w<4>(r->lds_stack, dim3(r->grid), pa, s);
        else
            launch_path_w<8>(r->lds_stack, dim3(r->grid), pa, s);
        return;
    }
    for (int k = 0; k < r->nscenes; ++k) {
        hipLaunchKernelGGL(k_path_reset, dim3(1), dim3(256), 0, s, r->d_counter, r->d_counter + 5 + k,
                           k == 0 ? r->d_loss : nullptr, r->nscenes);
        pa.scene_lo = k;
        pa.scene_hi = k + 1;
        pa.chunk_lo = chunk_first[k];
        pa.chunk_hi = chunk_first[k + 1];
        const dim3 grid(chunk_first[k + 1] - chunk_first[k]);
        if (r->width == 4)
            launch_path_w<4>(r->lds_stack, grid, pa, s);
        else
            launch_path_w<8>(r->lds_stack, grid, pa, s);
    }
}

void enqueue_adam(rt_invrender *r, hipStream_t s) {
    const int64_t total = r->nparams;
    int64_t *step = reinterpret_cast<int64_t *>(r->d_counter + 2);
    hipLaunchKernelGGL(k_adam, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0, s,
                       reinterpret_cast<const AdamSeg *>(r->d_adam), r->nscenes, total, r->d_m, r->d_v,
                       static_cast<const int64_t *>(step), r->params.lr, r->params.beta1, r->params.beta2,
                       r->params.eps);
    hipLaunchKernelGGL(k_step_inc, dim3(1), dim3(64), 0, s, step);
}

void free_invrender(rt_invrender *r) {
    if (!r) return;
    if (r->exec) (void)hipGraphExecDestroy(r->exec);
    if (r->graph) (void)hipGraphDestroy(r->graph);
    for (void *p : {static_cast<void *>(r->d_scenes), static_cast<void *>(r->d_counter), static_cast<void *>(r->d_rays),
                    static_cast<void *>(r->d_loss), static_cast<void *>(r->d_m), static_cast<void *>(r->d_v),
                    r->d_adam})
        if (p) (void)hipFree(p);
    if (r->ev0) (void)hipEventDestroy(r->ev0);
    if (r->ev1) (void)hipEventDestroy(r->ev1);
    delete r;
}

std::vector<uint32_t> chunk_firsts(const rt_invrender *r) {
    std::vector<uint32_t> f(r->nscenes + 1, 0);
    uint32_t c = 0;
    for (int k = 0; k < r->nscenes; ++k) {
        f[k] = c;
        const rt_camera &cam = r->items[k].cam;
        c += static_cast<uint32_t>(((cam.x_res + 7) / 8) * ((cam.y_res + 7) / 8));
    }
    f[r->nscenes] = c;
    return f;
}

}  // namespace
}  // namespace rt

using rt::fail;

extern "C" {

int32_t rt_invrender_create(rt_ctx *ctx, const rt_invrender_item *items, int32_t nitems, const rt_light *lights,
                            int32_t nlights, const rt_invrender_params *params, rt_invrender **out) {
    if (!ctx || !items || !out || nitems <= 0 || !params) return RT_ERR_INVALID;
    *out = nullptr;
    RT_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    const rt_invrender_params &p = *params;
    if (p.spp < 1 || p.spp > rt::kMaxSpp) return fail(ctx, RT_ERR_INVALID, "rt_invrender_create: spp must be 1..8");
    if (p.bounces < 0 || p.bounces > rt::kMaxBounces)
        return fail(ctx, RT_ERR_INVALID, "rt_invrender_create: bounces must be 0..3");
    if (nlights < 0 || nlights > rt::kMaxLights || (nlights > 0 && !lights))
        return fail(ctx, RT_ERR_INVALID, "rt_invrender_create: nlights must be 0..8");
    if (!items[0].accel) return fail(ctx, RT_ERR_INVALID, "rt_invrender_create: null accel");
    const int width = items[0].accel->width;
    std::vector<rt::PathScene> host(nitems);
    uint64_t chunks = 0;
    int64_t nparams = 0;
    for (int k = 0; k < nitems; ++k) {
        const rt_invrender_item &it = items[k];
        const rt_accel *a = it.accel;
        if (!a || a->width != width) return fail(ctx, RT_ERR_INVALID, "rt_invrender_create: items must share a width");
        if (a->scene->ctx->device != ctx->device)
            return fail(ctx, RT_ERR_INVALID, "rt_invrender_create: items must live on the context's device");
        if (a->root == rt::kEmpty || a->root < 0)
            return fail(ctx, RT_ERR_UNSUPPORTED, "rt_invrender_create: scenes need at least two triangles");
        if (a->scene->d_mat || a->scene->d_vn)
            return fail(ctx, RT_ERR_UNSUPPORTED,
                        "rt_invrender_create: scenes with materials or vertex normals are not differentiable");
        if (!it.albedo || !it.d_albedo || !it.image)
            return fail(ctx, RT_ERR_INVALID, "rt_invrender_create: albedo, gradient and image buffers are required");
        if (it.cam.x_res <= 0 || it.cam.y_res <= 0)
            return fail(ctx, RT_ERR_INVALID, "rt_invrender_create: camera resolution must be positive");
        rt::PathScene &ps = host[k];
        ps.nodes = a->d_nodes;
        ps.tris = a->d_tris;
        ps.root = a->root;
        ps.near_limit = a->near_limit;
        ps.tri_n = a->scene->d_n;
        ps.tri_c = a->scene->d_c;
        ps.albedo = it.albedo;
        ps.d_albedo = it.d_albedo;
        ps.target = it.target;
        ps.image = it.image;
        rt::FramePlan plan;
        const rt_render_params rp{0, p.spp, p.seed, 0, 0};
        if (int32_t rc = rt::plan_frame(const_cast<rt_accel *>(a), it.cam, lights, nlights, rp, &plan)) return rc;
        ps.cam = plan.cam;
        ps.tm = plan.tm;
        ps.chunk_begin = static_cast<uint32_t>(chunks);
        chunks += static_cast<uint64_t>(plan.tiles);
        nparams += 3 * a->scene->n;
    }
    if (chunks >= (uint64_t(1) << 31)) return fail(ctx, RT_ERR_INVALID, "rt_invrender_create: too many tiles");
    rt_invrender *r = new rt_invrender();
    r->ctx = ctx;
    r->width = width;
    r->lds_stack = items[0].accel->lds_stack;
    r->params = p;
    r->items.assign(items, items + nitems);
    r->nscenes = nitems;
    r->total_chunks = static_cast<uint32_t>(chunks);
    r->nparams = nparams;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess || cus <= 0)
        cus = 256;
    int per_cu = p.waves_per_cu;
    if (per_cu <= 0) {
        // As many single-wave workgroups as the kernel's registers allow.
        per_cu = rt::path_occupancy(width, r->lds_stack);
        if (per_cu <= 0) per_cu = 8;
    }
    r->grid = cus * per_cu;
    r->lights.n = nlights;
    for (int l = 0; l < nlights; ++l) {
        r->lights.pos_d[l] = make_float4(lights[l].pos[0], lights[l].pos[1], lights[l].pos[2], lights[l].d);
        r->lights.rgb[l] =
            make_float4(lights[l].rgb[0] / 255.0f, lights[l].rgb[1] / 255.0f, lights[l].rgb[2] / 255.0f, 0.0f);
    }
    const std::vector<uint32_t> first = rt::chunk_firsts(r);
    std::vector<rt::AdamSeg> segs(nitems);
    int64_t off = 0;
    for (int k = 0; k < nitems; ++k) {
        segs[k] = rt::AdamSeg{items[k].albedo, items[k].d_albedo, off, 3 * items[k].accel->scene->n};
        off += segs[k].count;
    }
    // Counter words: [0] queue head, [1] overflow, [2..3] Adam step, [4] 0,
    // [5 + k] first work item of scene k (for per-scene launches).
    std::vector<uint32_t> ctr(5 + nitems, 0u);
    for (int k = 0; k < nitems; ++k) ctr[5 + k] = first[k];
    hipError_t e = hipMalloc(&r->d_scenes, nitems * sizeof(rt::PathScene));
    if (e == hipSuccess) e = hipMalloc(&r->d_counter, ctr.size() * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&r->d_rays, sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMalloc(&r->d_loss, nitems * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&r->d_m, std::max<int64_t>(1, nparams) * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&r->d_v, std::max<int64_t>(1, nparams) * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&r->d_adam, nitems * sizeof(rt::AdamSeg));
    if (e == hipSuccess) e = hipEventCreate(&r->ev0);
    if (e == hipSuccess) e = hipEventCreate(&r->ev1);
    hipStream_t s = ctx->stream;
    if (e == hipSuccess)
        e = hipMemcpyAsync(r->d_scenes, host.data(), nitems * sizeof(rt::PathScene), hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(r->d_counter, ctr.data(), ctr.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(r->d_adam, segs.data(), nitems * sizeof(rt::AdamSeg), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        const unsigned blocks = static_cast<unsigned>((std::max<int64_t>(1, nparams) + 255) / 256);
        hipLaunchKernelGGL(rt::k_zero_f32, dim3(blocks), dim3(256), 0, s, r->d_m, nparams);
        hipLaunchKernelGGL(rt::k_zero_f32, dim3(blocks), dim3(256), 0, s, r->d_v, nparams);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e == hipSuccess && p.use_graph) {
        // One Adam step: reset, persistent forward + adjoint, update, step + 1.
        e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            rt::enqueue_frames(r, 1, false, first, s);
            rt::enqueue_adam(r, s);
            e = hipStreamEndCapture(s, &r->graph);
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&r->exec, r->graph, nullptr, nullptr, 0);
    }
    if (e != hipSuccess) {
        rt::free_invrender(r);
        return fail(ctx, e == hipErrorOutOfMemory ? RT_ERR_OOM : RT_ERR_HIP,
                    std::string("rt_invrender_create: ") + hipGetErrorString(e));
    }
    ctx->refs.fetch_add(1);
    for (int k = 0; k < nitems; ++k) items[k].accel->refs.fetch_add(1);
    *out = r;
    return RT_OK;
}

int32_t rt_invrender_run(rt_invrender *r, int32_t mode, int32_t steps, int32_t plain, double *loss, uint64_t *rays,
                         float *ms) {
    if (!r || mode < 0 || mode > 2 || steps < 1) return RT_ERR_INVALID;
    rt_ctx *ctx = r->ctx;
    RT_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    if (mode > 0)
        for (const rt_invrender_item &it : r->items)
            if (!it.target) return fail(ctx, RT_ERR_INVALID, "rt_invrender_run: the adjoint needs target images");
    hipStream_t s = ctx->stream;
    const std::vector<uint32_t> first = rt::chunk_firsts(r);
    RT_HIP_CHECK(ctx, hipMemsetAsync(r->d_rays, 0, sizeof(unsigned long long), s));
    if (mode == 1) {
        // Gradients only: clear them first.
        for (const rt_invrender_item &it : r->items) {
            const int64_t nparam = 3 * it.accel->scene->n;
            hipLaunchKernelGGL(rt::k_zero_f32, dim3(static_cast<unsigned>((nparam + 255) / 256)), dim3(256), 0, s,
                               it.d_albedo, nparam);
        }
    }
    RT_HIP_CHECK(ctx, hipEventRecord(r->ev0, s));
    const int run_steps = mode == 2 ? steps : 1;
    for (int t = 0; t < run_steps; ++t) {
        if (mode == 2 && r->exec && !plain) {
            RT_HIP_CHECK(ctx, hipGraphLaunch(r->exec, s));
        } else {
            rt::enqueue_frames(r, mode > 0 ? 1 : 0, plain != 0, first, s);
            if (mode == 2) rt::enqueue_adam(r, s);
        }
    }
    RT_HIP_CHECK(ctx, hipGetLastError());
    RT_HIP_CHECK(ctx, hipEventRecord(r->ev1, s));
    RT_HIP_CHECK(ctx, hipEventSynchronize(r->ev1));
    if (ms) RT_HIP_CHECK(ctx, hipEventElapsedTime(ms, r->ev0, r->ev1));
    uint32_t flags[2];
    RT_HIP_CHECK(ctx, hipMemcpy(flags, r->d_counter, sizeof(flags), hipMemcpyDeviceToHost));
    if (flags[1]) return fail(ctx, RT_ERR_INVALID, "rt_invrender_run: traversal stack overflow");
    if (rays) RT_HIP_CHECK(ctx, hipMemcpy(rays, r->d_rays, sizeof(uint64_t), hipMemcpyDeviceToHost));
    if (loss) {
        RT_HIP_CHECK(ctx, hipMemcpy(loss, r->d_loss, r->nscenes * sizeof(double), hipMemcpyDeviceToHost));
    }
    return RT_OK;
}

void rt_invrender_destroy(rt_invrender *r) {
    if (!r) return;
    rt_ctx *ctx = r->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    std::vector<rt_invrender_item> items;
    items.swap(r->items);
    rt::free_invrender(r);
    for (const rt_invrender_item &it : items) rt::release_accel(it.accel);
    rt::release_ctx(ctx);
}

}  // extern "C"
