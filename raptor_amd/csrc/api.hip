// C-ABI entry points of libraptor_rt (declared in include/raptor_rt.h).
#include <algorithm>
#include <cstring>
#include <new>

#include "rt_device.hpp"
#include "rt_internal.hpp"
#include "textures.hpp"

namespace rt {

int32_t fail(rt_ctx *ctx, int32_t code, const std::string &msg) {
    if (ctx) ctx->last_error = msg;
    return code;
}

namespace {

// rt/triangle.h:171-175: n = normalise(cross(b - a, c - a)) (EXACT_NORMALISE),
// plus the shading inputs of each triangle.
__global__ void __launch_bounds__(256) k_scene_prep(const float *__restrict__ abc, const float *__restrict__ albedo,
                                                    const uint8_t *__restrict__ flags, int64_t n,
                                                    float4 *__restrict__ tri_n, float4 *__restrict__ tri_c,
                                                    float4 *__restrict__ kd) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float *t = abc + 9 * i;
    const rtd::V3 a = rtd::mk(t[0], t[1], t[2]);
    const rtd::V3 b = rtd::mk(t[3], t[4], t[5]);
    const rtd::V3 c = rtd::mk(t[6], t[7], t[8]);
    const rtd::V3 dir_b = rtd::sub(b, a);
    const rtd::V3 dir_c = rtd::sub(c, a);
    const rtd::V3 nn = rtd::normalise_exact(rtd::cross(dir_b, dir_c));
    const int32_t f = flags ? static_cast<int32_t>(flags[i]) : 0;
    tri_n[i] = make_float4(nn.x, nn.y, nn.z, __int_as_float(f));
    tri_c[i] = make_float4(c.x, c.y, c.z, 0.0f);
    // phong_shader(rgb = 255 * albedo, kd = 1.0): kd = rgb * 1.0f
    // (rt/materials/phong_shader.h:135-137).
    float4 k = make_float4(255.0f, 255.0f, 255.0f, 0.0f);
    if (albedo) {
        k.x = (albedo[3 * i + 0] * 255.0f) * 1.0f;
        k.y = (albedo[3 * i + 1] * 255.0f) * 1.0f;
        k.z = (albedo[3 * i + 2] * 255.0f) * 1.0f;
    }
    kd[i] = k;
}

// New vertex positions: normal and vertex c again, the flags kept (tri_n.w).
__global__ void __launch_bounds__(256) k_scene_move(const float *__restrict__ abc, int64_t n,
                                                    float4 *__restrict__ tri_n, float4 *__restrict__ tri_c) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float *t = abc + 9 * i;
    const rtd::V3 a = rtd::mk(t[0], t[1], t[2]);
    const rtd::V3 b = rtd::mk(t[3], t[4], t[5]);
    const rtd::V3 c = rtd::mk(t[6], t[7], t[8]);
    const rtd::V3 nn = rtd::normalise_exact(rtd::cross(rtd::sub(b, a), rtd::sub(c, a)));
    tri_n[i] = make_float4(nn.x, nn.y, nn.z, tri_n[i].w);
    tri_c[i] = make_float4(c.x, c.y, c.z, 0.0f);
}

// Surface rays: triangle k = floor(r0 * n); point by the square-root
// barycentric map (Osada et al. 2002); origin offset along the face normal as
// rt/ray.h:111-157 offsets shadow-ray origins; cosine-weighted direction.
__global__ void __launch_bounds__(256) k_surface_rays(const float *__restrict__ abc, const float4 *__restrict__ tri_n,
                                                      int64_t ntri, uint64_t seed, int64_t n, float4 *__restrict__ rays) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t k0 = rtd::sample_key(seed, static_cast<uint32_t>(i), static_cast<uint32_t>(i >> 32), 254u);
    const float r0 = rtd::unit24(static_cast<uint32_t>(k0 >> 40));
    const float r1 = rtd::unit24(static_cast<uint32_t>((k0 >> 16) & 0xFFFFFFu));
    const float r2 = rtd::unit24(static_cast<uint32_t>(k0 & 0xFFFFu) << 8);
    int64_t tri = static_cast<int64_t>(static_cast<double>(r0) * static_cast<double>(ntri));
    if (tri >= ntri) tri = ntri - 1;
    const float *t = abc + 9 * tri;
    const float s = sqrtf(r1);
    const float wa = 1.0f - s, wb = r2 * s, wc = 1.0f - (wa + wb);
    const rtd::V3 p = rtd::mk((wa * t[0] + wb * t[3]) + wc * t[6], (wa * t[1] + wb * t[4]) + wc * t[7],
                              (wa * t[2] + wb * t[5]) + wc * t[8]);
    const float4 nn = tri_n[tri];
    const rtd::V3 nv = rtd::mk(nn.x, nn.y, nn.z);
    const rtd::V3 o = rtd::offset_start_point(p, nv, 1);
    const rtd::V3 d = rtd::cosine_direction(rtd::sample_key(seed, static_cast<uint32_t>(i), static_cast<uint32_t>(i >> 32), 1u), nv);
    rays[2 * i] = make_float4(o.x, o.y, o.z, rtd::kEpsilon);
    rays[2 * i + 1] = make_float4(d.x, d.y, d.z, rtd::kMaxDist);
}

template <class T>
hipError_t upload(T **dst, const T *src, int64_t count, hipStream_t s) {
    hipError_t e = hipMalloc(dst, std::max<int64_t>(1, count) * sizeof(T));
    if (e != hipSuccess) return e;
    if (count > 0 && src) e = hipMemcpyAsync(*dst, src, count * sizeof(T), hipMemcpyHostToDevice, s);
    return e;
}

void free_scene(rt_scene *sc) {
    if (!sc) return;
    if (sc->d_abc) (void)hipFree(sc->d_abc);
    if (sc->d_n) (void)hipFree(sc->d_n);
    if (sc->d_c) (void)hipFree(sc->d_c);
    if (sc->d_kd) (void)hipFree(sc->d_kd);
    if (sc->d_flags) (void)hipFree(sc->d_flags);
    if (sc->d_mat) (void)hipFree(sc->d_mat);
    if (sc->d_tri_mat) (void)hipFree(sc->d_tri_mat);
    if (sc->d_vn) (void)hipFree(sc->d_vn);
    if (sc->d_maps) (void)hipFree(sc->d_maps);
    if (sc->d_texels) (void)hipFree(sc->d_texels);
    if (sc->d_layers) (void)hipFree(sc->d_layers);
    if (sc->d_vt) (void)hipFree(sc->d_vt);
    delete sc;
}

int32_t set_device(rt_ctx *ctx) {
    RT_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    return RT_OK;
}

// An RT_ASYNC frame returns before the GPU has run it, so everything it
// reports must stay on the device.
int32_t check_async(rt_ctx *ctx, const rt_accel *a, const rt_render_params &p, const rt_stats *stats) {
    if (!(p.flags & RT_ASYNC)) return RT_OK;
    if (!(p.flags & RT_DEVICE_PTRS))
        return fail(ctx, RT_ERR_INVALID, "rt_render: RT_ASYNC needs RT_DEVICE_PTRS (device outputs)");
    if (stats) return fail(ctx, RT_ERR_INVALID, "rt_render: RT_ASYNC frames report rays by rt_ctx_frame_rays, not stats");
    if (ctx->counting) return fail(ctx, RT_ERR_INVALID, "rt_render: RT_ASYNC with visit counting on");
    if (a->scene->d_mat || a->scene->d_vn)
        return fail(ctx, RT_ERR_UNSUPPORTED, "rt_render: RT_ASYNC frames without materials or vertex normals only");
    return RT_OK;
}

}  // namespace

void release_ctx(rt_ctx *ctx) {
    if (!ctx || ctx->refs.fetch_sub(1) != 1) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->side) (void)hipStreamSynchronize(ctx->side);
    for (auto &e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    rt::release_glare(ctx);
    if (ctx->fork) (void)hipEventDestroy(ctx->fork);
    if (ctx->join) (void)hipEventDestroy(ctx->join);
    if (ctx->d_counters) (void)hipFree(ctx->d_counters);
    if (ctx->d_frame_rays) (void)hipFree(ctx->d_frame_rays);
    if (ctx->d_queue) (void)hipFree(ctx->d_queue);
    ctx->arena.release();
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    if (ctx->side) (void)hipStreamDestroy(ctx->side);
    delete ctx;
}

void release_scene(rt_scene *scene) {
    if (!scene || scene->refs.fetch_sub(1) != 1) return;
    rt_ctx *ctx = scene->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    free_scene(scene);
    release_ctx(ctx);
}

void release_accel(rt_accel *accel) {
    if (!accel || accel->refs.fetch_sub(1) != 1) return;
    rt_scene *scene = accel->scene;
    (void)hipSetDevice(scene->ctx->device);
    (void)hipStreamSynchronize(scene->ctx->stream);
    free_frame_workspace(accel);
    if (accel->d_nodes) (void)hipFree(accel->d_nodes);
    if (accel->d_tris) (void)hipFree(accel->d_tris);
    if (accel->d_refit) (void)hipFree(accel->d_refit);
    delete accel;
    release_scene(scene);
}

}  // namespace rt

using rt::fail;

extern "C" {

int32_t rt_ctx_create(int32_t hip_device, rt_ctx **out) {
    if (!out) return RT_ERR_INVALID;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return RT_ERR_HIP;
    if (hip_device < 0 || hip_device >= count) return RT_ERR_INVALID;
    rt_ctx *ctx = new (std::nothrow) rt_ctx();
    if (!ctx) return RT_ERR_OOM;
    ctx->device = hip_device;
    if (hipSetDevice(hip_device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return RT_ERR_HIP;
    }
    if (hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->join, hipEventDisableTiming) != hipSuccess) {
        rt_ctx_destroy(ctx);
        return RT_ERR_HIP;
    }
    for (auto &e : ctx->ev) {
        if (hipEventCreate(&e) != hipSuccess) {
            rt_ctx_destroy(ctx);
            return RT_ERR_HIP;
        }
    }
    if (hipMalloc(&ctx->d_counters, 16 * sizeof(unsigned long long)) != hipSuccess ||
        hipMalloc(&ctx->d_frame_rays, 4 * sizeof(unsigned long long)) != hipSuccess ||
        hipMalloc(&ctx->d_queue, rt::kQueueWords * sizeof(uint32_t)) != hipSuccess) {
        rt_ctx_destroy(ctx);
        return RT_ERR_HIP;
    }
    (void)hipMemset(ctx->d_counters, 0, 16 * sizeof(unsigned long long));
    (void)hipMemset(ctx->d_frame_rays, 0, 4 * sizeof(unsigned long long));
    // The persistent trace kernels expect their chunk counters at zero and
    // leave them at zero.
    (void)hipMemset(ctx->d_queue, 0, rt::kQueueWords * sizeof(uint32_t));
    // The device address of this device's static scratch, looked up once
    // with the device current (each GPU holds its own copy of the array).
    if (rt::static_scratch_address(ctx, &ctx->d_static) != RT_OK) {
        rt_ctx_destroy(ctx);
        return RT_ERR_HIP;
    }
    *out = ctx;
    return RT_OK;
}

void rt_ctx_destroy(rt_ctx *ctx) { rt::release_ctx(ctx); }

const char *rt_last_error(const rt_ctx *ctx) { return ctx ? ctx->last_error.c_str() : "null context"; }

int32_t rt_scene_create(rt_ctx *ctx, const float *abc, const float *albedo, const uint8_t *flags, int64_t n,
                        rt_scene **out) {
    if (!ctx || !out) return RT_ERR_INVALID;
    *out = nullptr;
    if (n < 0 || (n > 0 && !abc)) return fail(ctx, RT_ERR_INVALID, "rt_scene_create: bad triangle array");
    if (int32_t rc = rt::set_device(ctx)) return rc;
    rt_scene *sc = new (std::nothrow) rt_scene();
    if (!sc) return fail(ctx, RT_ERR_OOM, "rt_scene_create: host allocation");
    sc->ctx = ctx;
    sc->n = n;
    hipStream_t s = ctx->stream;
    float *d_alb = nullptr;
    hipError_t e = rt::upload(&sc->d_abc, abc, 9 * n, s);
    if (e == hipSuccess && flags) e = rt::upload(&sc->d_flags, flags, n, s);
    if (e == hipSuccess && albedo) e = rt::upload(&d_alb, albedo, 3 * n, s);
    if (e == hipSuccess) e = hipMalloc(&sc->d_n, std::max<int64_t>(1, n) * sizeof(float4));
    if (e == hipSuccess) e = hipMalloc(&sc->d_c, std::max<int64_t>(1, n) * sizeof(float4));
    if (e == hipSuccess) e = hipMalloc(&sc->d_kd, std::max<int64_t>(1, n) * sizeof(float4));
    if (e == hipSuccess && n > 0) {
        hipLaunchKernelGGL(rt::k_scene_prep, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, s, sc->d_abc,
                           d_alb, sc->d_flags, n, sc->d_n, sc->d_c, sc->d_kd);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (d_alb) (void)hipFree(d_alb);
    if (e != hipSuccess) {
        rt::free_scene(sc);
        return fail(ctx, e == hipErrorOutOfMemory ? RT_ERR_OOM : RT_ERR_HIP,
                    std::string("rt_scene_create: ") + hipGetErrorString(e));
    }
    ctx->refs.fetch_add(1);
    *out = sc;
    return RT_OK;
}

void rt_scene_destroy(rt_scene *scene) { rt::release_scene(scene); }

int32_t rt_scene_normals(rt_scene *scene, float *n3) {
    if (!scene || !n3) return RT_ERR_INVALID;
    rt_ctx *ctx = scene->ctx;
    if (int32_t rc = rt::set_device(ctx)) return rc;
    std::vector<float4> tmp(std::max<int64_t>(1, scene->n));
    RT_HIP_CHECK(ctx, hipMemcpyAsync(tmp.data(), scene->d_n, scene->n * sizeof(float4), hipMemcpyDeviceToHost,
                                     ctx->stream));
    RT_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    for (int64_t i = 0; i < scene->n; ++i) {
        n3[3 * i + 0] = tmp[i].x;
        n3[3 * i + 1] = tmp[i].y;
        n3[3 * i + 2] = tmp[i].z;
    }
    return RT_OK;
}

int32_t rt_accel_build(rt_scene *scene, const rt_build_params *params, rt_accel **out) {
    if (!scene || !out) return RT_ERR_INVALID;
    *out = nullptr;
    rt_ctx *ctx = scene->ctx;
    if (int32_t rc = rt::set_device(ctx)) return rc;
    rt_build_params p{4, 2, 16, 8};
    if (params) p = *params;
    const int32_t rc = rt::build_accel(scene, p, out);
    if (rc == RT_OK) scene->refs.fetch_add(1);
    return rc;
}

void rt_accel_destroy(rt_accel *accel) { rt::release_accel(accel); }

int32_t rt_scene_update(rt_scene *scene, const float *abc, uint32_t flags) {
    if (!scene || (!abc && scene->n > 0)) return RT_ERR_INVALID;
    rt_ctx *ctx = scene->ctx;
    if (int32_t rc = rt::set_device(ctx)) return rc;
    if (scene->n == 0) return RT_OK;
    hipStream_t s = ctx->stream;
    const bool dev = (flags & RT_DEVICE_PTRS) != 0;
    RT_HIP_CHECK(ctx, hipMemcpyAsync(scene->d_abc, abc, 9 * scene->n * sizeof(float),
                                     dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(rt::k_scene_move, dim3(static_cast<unsigned>((scene->n + 255) / 256)), dim3(256), 0, s,
                       scene->d_abc, scene->n, scene->d_n, scene->d_c);
    RT_HIP_CHECK(ctx, hipGetLastError());
    // The source (host or device) is borrowed for the call only: it must stay
    // untouched until the copy has read it.
    RT_HIP_CHECK(ctx, hipStreamSynchronize(s));
    return RT_OK;
}

int32_t rt_image_tone_map(rt_ctx *ctx, float *rgb, int32_t x_res, int32_t y_res, rt_tone_params *params,
                          uint32_t flags, float *ms) {
    if (!ctx || !rgb || !params) return RT_ERR_INVALID;
    if (int32_t rc = rt::set_device(ctx)) return rc;
    if (x_res <= 0 || y_res <= 0) return fail(ctx, RT_ERR_INVALID, "rt_image_tone_map: resolution must be positive");
    return rt::image_tone_map(ctx, rgb, x_res, y_res, params, flags, ms);
}

int32_t rt_image_gamma(rt_ctx *ctx, float *rgb, int64_t npix, float gamma, uint32_t flags, float *ms) {
    if (!ctx || (npix > 0 && !rgb) || npix < 0) return RT_ERR_INVALID;
    if (int32_t rc = rt::set_device(ctx)) return rc;
    if (npix == 0) return RT_OK;
    if (!(gamma > 0.0f)) return fail(ctx, RT_ERR_INVALID, "rt_image_gamma: gamma must be positive");
    return rt::image_gamma(ctx, rgb, npix, gamma, flags, ms);
}

int32_t rt_image_clip(rt_ctx *ctx, const float *rgb, int32_t x_res, int32_t y_res, int32_t x_samples,
                      int32_t y_samples, int32_t bgr, uint8_t *out, uint32_t flags, float *ms) {
    if (!ctx || !rgb || !out) return RT_ERR_INVALID;
    if (int32_t rc = rt::set_device(ctx)) return rc;
    if (x_res <= 0 || y_res <= 0 || x_samples < 1 || y_samples < 1 || x_samples > x_res || y_samples > y_res)
        return fail(ctx, RT_ERR_INVALID, "rt_image_clip: bad resolution or sample counts");
    return rt::image_clip(ctx, rgb, x_res, y_res, x_samples, y_samples, bgr ? 1 : 0, out, flags, ms);
}

int32_t rt_write_png(const char *path, const uint8_t *rgb, int32_t x_res, int32_t y_res) {
    if (!path || !rgb || x_res <= 0 || y_res <= 0) return RT_ERR_INVALID;
    return rt::write_png(path, rgb, x_res, y_res);
}

int32_t rt_write_tga(const char *path, const uint8_t *bgr, int32_t x_res, int32_t y_res) {
    if (!path || !bgr || x_res <= 0 || y_res <= 0 || x_res > 65535 || y_res > 65535) return RT_ERR_INVALID;
    return rt::write_tga(path, bgr, x_res, y_res);
}

static_assert(sizeof(rt_material) == 160, "rt_material is 40 four-byte fields");

int32_t rt_scene_set_materials(rt_scene *scene, const rt_material *materials, int32_t nmaterials,
                               const int32_t *tri_material) {
    if (!scene) return RT_ERR_INVALID;
    rt_ctx *ctx = scene->ctx;
    if (int32_t rc = rt::set_device(ctx)) return rc;
    if (nmaterials < 0 || (nmaterials > 0 && (!materials || (scene->n > 0 && !tri_material))))
        return fail(ctx, RT_ERR_INVALID, "rt_scene_set_materials: bad material arrays");
    hipStream_t s = ctx->stream;
    auto drop = [&]() {
        (void)hipStreamSynchronize(s);
        if (scene->d_mat) (void)hipFree(scene->d_mat);
        if (scene->d_tri_mat) (void)hipFree(scene->d_tri_mat);
        scene->d_mat = nullptr;
        scene->d_tri_mat = nullptr;
        scene->nmat = 0;
        scene->maps_used = scene->layers_used = 0;
    };
    if (nmaterials == 0) {
        drop();
        return RT_OK;
    }
    // The ranges the shaders assert (rt/materials/phong_shader.h:38-51,
    // rt/materials/cook_torrance_cxy.h:20-29).
    std::vector<float4> table(rt::kMaterialFloat4s * static_cast<size_t>(nmaterials));
    auto in = [](float v, float lo, float hi) { return v >= lo && v <= hi; };
    for (int32_t m = 0; m < nmaterials; ++m) {
        const rt_material &q = materials[m];
        bool ok = true;
        std::string range;
        if (q.kind == RT_MATERIAL_PHONG || q.kind == RT_MATERIAL_COLOURED) {
            // coloured_mapper_shader asserts tran and rf up to 1
            // (rt/materials/coloured_mapper_shader.h:26-40).
            const float top = (q.kind == RT_MATERIAL_PHONG) ? 0.95f : 1.0f;
            ok = in(q.transmittance, 0.0f, top) && in(q.reflectance, 0.0f, top) &&
                 (q.transmittance == 0.0f || in(q.refractive_index, 1.0f, 5.0f));
            for (int k = 0; k < 3; ++k)
                for (const float v : {q.ka[k], q.kd[k], q.ks[k]}) ok = ok && in(v, 0.0f, 255.0f);
            range = (q.kind == RT_MATERIAL_PHONG) ? "colours 0..255, tran and rf 0..0.95, ri 1..5 when tran > 0"
                                                  : "colours 0..255, tran and rf 0..1, ri 1..5 when tran > 0, "
                                                    "texture slots 0 or 1 + a map of the scene";
            for (int k = 0; k < 6; ++k) {
                const bool none = q.tex[k] == 0;
                ok = ok && (none || (q.kind == RT_MATERIAL_COLOURED && q.tex[k] >= 1 && q.tex[k] <= scene->nmaps));
            }
        } else if (q.kind == RT_MATERIAL_MAPPER) {
            // rt/materials/mapper_shader.h:21-33.
            ok = q.mapper_kd >= 0.0f && in(q.mapper_ks, 0.0f, 1.0f) && in(q.transmittance, 0.0f, 1.0f) &&
                 in(q.reflectance, 0.0f, 1.0f) && (q.transmittance == 0.0f || in(q.refractive_index, 1.0f, 5.0f));
            for (int k = 0; k < 3; ++k) ok = ok && in(q.kd[k], 0.0f, 255.0f);
            for (int k = 0; k < 5; ++k)
                ok = ok && q.layer_first[k] >= 0 && q.layer_count[k] >= 0 && q.layer_count[k] < 256 &&
                     q.layer_first[k] < (1 << 23) && q.layer_first[k] + q.layer_count[k] <= scene->nlayers;
            range = "rgb 0..255, kd >= 0, ks, tran and rf 0..1, ri 1..5 when tran > 0, layer lists inside the "
                    "scene's layers (rt_scene_set_textures first)";
        } else if (q.kind == RT_MATERIAL_COOK_TORRANCE) {
            ok = in(q.chroma[0], 0.0f, 1.0f) && in(q.chroma[1], 0.0f, 1.0f) && in(q.diffuse, 0.0f, 1.0f) &&
                 in(q.reflectance, 0.0f, 1.0f) && in(q.transmittance, 0.0f, 1.0f) && in(q.roughness, 0.0f, 0.5f) &&
                 in(q.refractive_index, 0.0f, 5.0f) && in(q.extinction, 0.0f, 1.0f);
            range = "x, y, rd, rs, ts, ri_i 0..1, sr 0..0.5, ri_r 0..5";
        } else if (q.kind == RT_MATERIAL_MANDELBROT) {
            // mandelbrot_shader is opaque and spawns no secondary rays
            // (rt/materials/mandelbrot_shader.h:14-15); a hit iterates at most
            // twice the limit, which bounds the shading kernel's time.
            ok = q.iterations >= 1 && q.iterations <= rt::kMaxMandelbrotIterations && q.transmittance == 0.0f &&
                 q.reflectance == 0.0f;
            range = "iterations 1.." + std::to_string(rt::kMaxMandelbrotIterations) + ", transmittance and "
                    "reflectance 0";
        } else {
            ok = false;
            range = "unknown kind " + std::to_string(q.kind);
        }
        if (!ok)
            return fail(ctx, RT_ERR_INVALID,
                        "rt_scene_set_materials: material " + std::to_string(m) + " out of range (" + range + ")");
        float4 *row = &table[rt::kMaterialFloat4s * static_cast<size_t>(m)];
        row[0] = make_float4(q.ka[0], q.ka[1], q.ka[2], q.shininess);
        row[1] = make_float4(q.kd[0], q.kd[1], q.kd[2], q.transmittance);
        row[2] = make_float4(q.ks[0], q.ks[1], q.ks[2], q.refractive_index);
        row[3] = make_float4(q.reflectance, static_cast<float>(q.kind), q.chroma[0], q.chroma[1]);
        row[4] = make_float4(q.diffuse, q.roughness, q.extinction, 0.0f);
        auto bits = [](int32_t v) {
            float f;
            std::memcpy(&f, &v, sizeof(f));
            return f;
        };
        const bool mapper = q.kind == RT_MATERIAL_MAPPER;
        int32_t lists[5];
        for (int k = 0; k < 5; ++k) lists[k] = mapper ? ((q.layer_first[k] << 8) | q.layer_count[k]) : 0;
        const bool col = q.kind == RT_MATERIAL_COLOURED;
        row[5] = make_float4(bits(col ? q.tex[0] : 0), bits(col ? q.tex[1] : 0), bits(col ? q.tex[2] : 0),
                             bits(col ? q.tex[3] : 0));
        row[6] = make_float4(bits(col ? q.tex[4] : 0), bits(col ? q.tex[5] : 0), q.mapper_kd, q.mapper_ks);
        row[7] = make_float4(q.clrh, q.clrf, bits(lists[0]), bits(lists[1]));
        row[8] = make_float4(bits(lists[2]), bits(lists[3]), bits(lists[4]),
                             bits(q.kind == RT_MATERIAL_MANDELBROT ? q.iterations : 0));
    }
    // A triangle is transparent to shadow rays exactly when its material is
    // (material(tran > 0), rt/materials/material.h:50-66).
    std::vector<uint8_t> flags(std::max<int64_t>(1, scene->n), 0);
    if (scene->d_flags && scene->n > 0) {
        RT_HIP_CHECK(ctx, hipMemcpyAsync(flags.data(), scene->d_flags, scene->n, hipMemcpyDeviceToHost, s));
        RT_HIP_CHECK(ctx, hipStreamSynchronize(s));
    }
    for (int64_t i = 0; i < scene->n; ++i) {
        const int32_t m = tri_material[i];
        if (m < -1 || m >= nmaterials)
            return fail(ctx, RT_ERR_INVALID,
                        "rt_scene_set_materials: triangle " + std::to_string(i) + " has no material " +
                            std::to_string(m));
        const bool transparent = m >= 0 && materials[m].transmittance > 0.0f;
        if (transparent != ((flags[i] & 2) != 0))
            return fail(ctx, RT_ERR_INVALID,
                        "rt_scene_set_materials: triangle " + std::to_string(i) +
                            (transparent ? " has a transparent material but no transparent flag"
                                         : " is flagged transparent but its material is not"));
    }
    drop();
    float4 *d_mat = nullptr;
    int32_t *d_tri = nullptr;
    hipError_t e = rt::upload(&d_mat, table.data(), static_cast<int64_t>(table.size()), s);
    if (e == hipSuccess) e = rt::upload(&d_tri, tri_material, scene->n, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        if (d_mat) (void)hipFree(d_mat);
        if (d_tri) (void)hipFree(d_tri);
        return fail(ctx, e == hipErrorOutOfMemory ? RT_ERR_OOM : RT_ERR_HIP,
                    std::string("rt_scene_set_materials: ") + hipGetErrorString(e));
    }
    scene->d_mat = d_mat;
    scene->d_tri_mat = d_tri;
    scene->nmat = nmaterials;
    for (int32_t m = 0; m < nmaterials; ++m) {
        const rt_material &q = materials[m];
        if (q.kind == RT_MATERIAL_COLOURED)
            for (int k = 0; k < 6; ++k) scene->maps_used = std::max(scene->maps_used, q.tex[k]);
        if (q.kind == RT_MATERIAL_MAPPER)
            for (int k = 0; k < 5; ++k)
                if (q.layer_count[k] > 0)
                    scene->layers_used = std::max(scene->layers_used, q.layer_first[k] + q.layer_count[k]);
    }
    return RT_OK;
}

int32_t rt_scene_set_textures(rt_scene *scene, const rt_texture_map *maps, int32_t nmaps,
                              const rt_texture_image *images, int32_t nimages, const float *texels, int64_t ntexels,
                              const int32_t *layers, int32_t nlayers) {
    if (!scene) return RT_ERR_INVALID;
    rt_ctx *ctx = scene->ctx;
    if (int32_t rc = rt::set_device(ctx)) return rc;
    hipStream_t s = ctx->stream;
    auto bad = [&](const std::string &why) { return fail(ctx, RT_ERR_INVALID, "rt_scene_set_textures: " + why); };
    if (nmaps < 0 || nimages < 0 || ntexels < 0 || nlayers < 0 || (nmaps > 0 && !maps) ||
        (nimages > 0 && !images) || (ntexels > 0 && !texels) || (nlayers > 0 && !layers))
        return bad("bad arrays");
    std::vector<rt::DevTexMap> dev(static_cast<size_t>(nmaps));
    for (int32_t i = 0; i < nmaps; ++i) {
        const rt_texture_map &m = maps[i];
        rt::DevTexMap &d = dev[i];
        std::memset(&d, 0, sizeof(d));
        d.m = m;
        const std::string who = "map " + std::to_string(i) + ": ";
        if (m.kind < RT_MAP_PLANAR || m.kind > RT_MAP_GRADIENT) return bad(who + "unknown kind");
        if (m.falloff < RT_FALLOFF_CUBIC || m.falloff > RT_FALLOFF_NONE) return bad(who + "unknown falloff");
        if (m.kind == RT_MAP_CHECKER_BOARD) continue;
        if (m.kind == RT_MAP_PERLIN_3D) {
            if (m.octaves < 0 || m.octaves > rt::kMaxPerlinOctaves)
                return bad(who + "octaves must be 0.." + std::to_string(rt::kMaxPerlinOctaves));
            continue;
        }
        if (m.kind == RT_MAP_GRADIENT) {
            // gradient_mapper.cc asserts on every key source but the
            // incidence angle and on spline interpolation; upper_bound needs
            // sorted keys.
            if (m.grad_of != RT_GRAD_OF_INCIDENCE_ANGLE) return bad(who + "only the incidence-angle gradient exists");
            if (m.key_count < 1 || m.key_offset < 0 || m.key_offset + 6 * static_cast<int64_t>(m.key_count) > ntexels)
                return bad(who + "its keys lie outside the texels");
            const float *k = texels + m.key_offset;
            for (int32_t j = 0; j < m.key_count; ++j) {
                const int interp = static_cast<int>(k[6 * j + 5]);
                if (static_cast<float>(interp) != k[6 * j + 5] || (interp != RT_GRAD_LINE && interp != RT_GRAD_STEP))
                    return bad(who + "key " + std::to_string(j) + ": interpolation must be line or step");
                if (j > 0 && !(k[6 * (j - 1)] <= k[6 * j])) return bad(who + "keys must be non-decreasing");
            }
            continue;
        }
        if (m.image < 0 || m.image >= nimages) return bad(who + "no image " + std::to_string(m.image));
        const rt_texture_image &im = images[m.image];
        if (im.width <= 0 || im.height <= 0 || (im.cpp != 1 && im.cpp != 3 && im.cpp != 4) || im.offset < 0 ||
            im.offset + static_cast<int64_t>(im.width) * im.height * im.cpp > ntexels)
            return bad(who + "image " + std::to_string(m.image) + " has a bad shape or lies outside the texels");
        for (const int32_t wm : {m.u_wrap, m.v_wrap})
            if (wm < RT_WRAP_BLANK || wm > RT_WRAP_MIRROR) return bad(who + "unknown wrapping mode");
        // Tile and mirror divide by size - 1 (texture_mapper.h apply_wrapping_mode).
        if (((m.u_wrap == RT_WRAP_TILE || m.u_wrap == RT_WRAP_MIRROR) && im.width < 2) ||
            ((m.v_wrap == RT_WRAP_TILE || m.v_wrap == RT_WRAP_MIRROR) && im.height < 2))
            return bad(who + "tile and mirror wrapping need 2 texels along the axis");
        d.w = im.width;
        d.h = im.height;
        d.cpp = im.cpp;
        d.offset = im.offset;
    }
    for (int32_t i = 0; i < nlayers; ++i)
        if (layers[i] < 0 || layers[i] >= nmaps) return bad("layer " + std::to_string(i) + " names no map");
    if (nmaps < scene->maps_used || nlayers < scene->layers_used)
        return bad("the scene's materials use " + std::to_string(scene->maps_used) + " maps and " +
                   std::to_string(scene->layers_used) + " layers; change the materials first");
    RT_HIP_CHECK(ctx, hipStreamSynchronize(s));
    for (void *p : {static_cast<void *>(scene->d_maps), static_cast<void *>(scene->d_texels),
                    static_cast<void *>(scene->d_layers)})
        if (p) (void)hipFree(p);
    scene->d_maps = nullptr;
    scene->d_texels = nullptr;
    scene->d_layers = nullptr;
    scene->nmaps = scene->nlayers = 0;
    if (nmaps == 0) return RT_OK;
    hipError_t e = rt::upload(&scene->d_maps, dev.data(), nmaps, s);
    if (e == hipSuccess) e = rt::upload(&scene->d_texels, ntexels > 0 ? texels : nullptr, ntexels, s);
    if (e == hipSuccess && nlayers > 0) e = rt::upload(&scene->d_layers, layers, nlayers, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess)
        return fail(ctx, e == hipErrorOutOfMemory ? RT_ERR_OOM : RT_ERR_HIP,
                    std::string("rt_scene_set_textures: ") + hipGetErrorString(e));
    scene->nmaps = nmaps;
    scene->nlayers = nlayers;
    return RT_OK;
}

int32_t rt_scene_set_texture_coords(rt_scene *scene, const float *vt) {
    if (!scene) return RT_ERR_INVALID;
    rt_ctx *ctx = scene->ctx;
    if (int32_t rc = rt::set_device(ctx)) return rc;
    hipStream_t s = ctx->stream;
    if (scene->d_vt) {
        RT_HIP_CHECK(ctx, hipStreamSynchronize(s));
        (void)hipFree(scene->d_vt);
        scene->d_vt = nullptr;
    }
    if (!vt || scene->n == 0) return RT_OK;
    float *d = nullptr;
    hipError_t e = rt::upload(&d, vt, 6 * scene->n, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        if (d) (void)hipFree(d);
        return fail(ctx, e == hipErrorOutOfMemory ? RT_ERR_OOM : RT_ERR_HIP,
                    std::string("rt_scene_set_texture_coords: ") + hipGetErrorString(e));
    }
    scene->d_vt = d;
    return RT_OK;
}

int32_t rt_scene_set_vertex_normals(rt_scene *scene, const float *vn) {
    if (!scene) return RT_ERR_INVALID;
    rt_ctx *ctx = scene->ctx;
    if (int32_t rc = rt::set_device(ctx)) return rc;
    hipStream_t s = ctx->stream;
    if (scene->d_vn) {
        RT_HIP_CHECK(ctx, hipStreamSynchronize(s));
        (void)hipFree(scene->d_vn);
        scene->d_vn = nullptr;
    }
    if (!vn || scene->n == 0) return RT_OK;
    float *d = nullptr;
    hipError_t e = rt::upload(&d, vn, 9 * scene->n, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        if (d) (void)hipFree(d);
        return fail(ctx, e == hipErrorOutOfMemory ? RT_ERR_OOM : RT_ERR_HIP,
                    std::string("rt_scene_set_vertex_normals: ") + hipGetErrorString(e));
    }
    scene->d_vn = d;
    return RT_OK;
}

int32_t rt_accel_refit(rt_accel *accel) {
    if (!accel) return RT_ERR_INVALID;
    if (int32_t rc = rt::set_device(accel->scene->ctx)) return rc;
    return rt::refit_accel(accel);
}

int32_t rt_accel_info_get(const rt_accel *a, rt_accel_info *info) {
    if (!a || !info) return RT_ERR_INVALID;
    info->nodes = a->nodes;
    info->leaves = a->leaves;
    info->triangles = a->scene->n;
    info->width = a->width;
    info->depth = a->depth;
    const int64_t node_bytes = (a->width == 4) ? 128 : 256;
    info->bytes = a->nodes * node_bytes + a->scene->n * static_cast<int64_t>(sizeof(rt::LeafTri));
    info->build_ms = a->build_ms;
    info->sah_cost = a->sah_cost;
    info->refit_ms = a->refit_ms;
    return RT_OK;
}

static int32_t trace_common(rt_accel *accel, const rt_ray *rays, int64_t n, void *out, size_t out_elem,
                            uint32_t flags, bool closest) {
    if (!accel || (n > 0 && (!rays || !out)) || n < 0) return RT_ERR_INVALID;
    rt_ctx *ctx = accel->scene->ctx;
    if (int32_t rc = rt::set_device(ctx)) return rc;
    if (n == 0) return RT_OK;
    hipStream_t s = ctx->stream;
    const bool dev = (flags & RT_DEVICE_PTRS) != 0;
    const rt_ray *d_rays = rays;
    void *d_out = out;
    rt_ray *tmp_rays = nullptr;
    void *tmp_out = nullptr;
    if (!dev) {
        RT_HIP_CHECK(ctx, hipMalloc(&tmp_rays, n * sizeof(rt_ray)));
        hipError_t e = hipMalloc(&tmp_out, n * out_elem);
        if (e != hipSuccess) {
            (void)hipFree(tmp_rays);
            return fail(ctx, RT_ERR_OOM, "trace staging buffers");
        }
        (void)hipMemcpyAsync(tmp_rays, rays, n * sizeof(rt_ray), hipMemcpyHostToDevice, s);
        d_rays = tmp_rays;
        d_out = tmp_out;
    }
    int32_t rc = closest ? rt::launch_closest(accel, d_rays, n, reinterpret_cast<rt_hit *>(d_out), s)
                         : rt::launch_occluded(accel, d_rays, n, reinterpret_cast<uint8_t *>(d_out), s);
    if (rc == RT_OK && !dev) {
        hipError_t e = hipMemcpyAsync(out, d_out, n * out_elem, hipMemcpyDeviceToHost, s);
        if (e != hipSuccess) rc = fail(ctx, RT_ERR_HIP, hipGetErrorString(e));
    }
    hipError_t e = hipStreamSynchronize(s);
    if (rc == RT_OK && e != hipSuccess) rc = fail(ctx, RT_ERR_HIP, std::string("trace: ") + hipGetErrorString(e));
    if (tmp_rays) (void)hipFree(tmp_rays);
    if (tmp_out) (void)hipFree(tmp_out);
    return rc;
}

int32_t rt_trace_closest(rt_accel *accel, const rt_ray *rays, int64_t n, rt_hit *out, uint32_t flags) {
    return trace_common(accel, rays, n, out, sizeof(rt_hit), flags, true);
}

int32_t rt_trace_occluded(rt_accel *accel, const rt_ray *rays, int64_t n, uint8_t *occluded, uint32_t flags) {
    return trace_common(accel, rays, n, occluded, 1, flags, false);
}

int32_t rt_render(rt_accel *accel, const rt_camera *cam, const rt_light *lights, int32_t nlights,
                  const rt_render_params *params, float *rgb, rt_hit *primary, rt_hit *bounce, rt_stats *stats) {
    if (!accel || !cam || (nlights > 0 && !lights)) return RT_ERR_INVALID;
    rt_ctx *ctx = accel->scene->ctx;
    if (int32_t rc = rt::set_device(ctx)) return rc;
    rt_render_params p{0, 1, 0, 0, 0};
    if (params) p = *params;
    if (int32_t rc = rt::check_async(ctx, accel, p, stats)) return rc;
    if (ctx->counting)
        RT_HIP_CHECK(ctx, hipMemsetAsync(ctx->d_counters, 0, 16 * sizeof(unsigned long long), ctx->stream));
    int32_t rc = RT_OK;
    if (accel->scene->d_mat || accel->scene->d_vn) {
        rc = bounce ? fail(ctx, RT_ERR_UNSUPPORTED,
                           "rt_render: the diffuse bounce is not defined with materials or vertex normals")
                    : rt::render_frame_materials(accel, *cam, lights, nlights, p, rgb, primary, stats);
    } else {
        rc = rt::render_frame(accel, *cam, lights, nlights, p, rgb, primary, bounce, stats);
    }
    ctx->count_slot = 0;
    if (rc == RT_OK && stats) {
        for (auto &v : stats->visits) v = 0;
        if (ctx->counting) {
            unsigned long long v[16];
            RT_HIP_CHECK(ctx, hipMemcpy(v, ctx->d_counters, sizeof(v), hipMemcpyDeviceToHost));
            for (int k = 0; k < 8; ++k) stats->visits[k] = v[k];
            if (v[15]) return fail(ctx, RT_ERR_INVALID, "traversal stack overflow");
        }
    }
    return rc;
}

int32_t rt_render_static(rt_accel *accel, const rt_camera *cam, const rt_light *lights, int32_t nlights,
                         const rt_render_params *params, float *rgb, rt_hit *primary, rt_hit *bounce,
                         rt_stats *stats) {
    if (!accel || !cam || !rgb || (nlights > 0 && !lights)) return RT_ERR_INVALID;
    rt_ctx *ctx = accel->scene->ctx;
    if (int32_t rc = rt::set_device(ctx)) return rc;
    if (!ctx->d_static) return fail(ctx, RT_ERR_HIP, "rt_render_static: no static scratch on this device");
    if (accel->scene->d_mat || accel->scene->d_vn)
        return fail(ctx, RT_ERR_UNSUPPORTED,
                    "rt_render_static: frames with materials or vertex normals size their generations at run time; "
                    "use rt_render");
    rt_render_params p{0, 1, 0, 0, 0};
    if (params) p = *params;
    p.flags |= RT_DEVICE_PTRS;
    if (int32_t rc = rt::check_async(ctx, accel, p, stats)) return rc;
    if (ctx->counting)
        RT_HIP_CHECK(ctx, hipMemsetAsync(ctx->d_counters, 0, 16 * sizeof(unsigned long long), ctx->stream));
    int32_t rc = rt::render_frame(accel, *cam, lights, nlights, p, rgb, primary, bounce, stats, ctx->d_static,
                                  rt::kStaticScratchBytes);
    ctx->count_slot = 0;
    if (rc == RT_OK && stats) {
        for (auto &v : stats->visits) v = 0;
        if (ctx->counting) {
            unsigned long long v[16];
            RT_HIP_CHECK(ctx, hipMemcpy(v, ctx->d_counters, sizeof(v), hipMemcpyDeviceToHost));
            for (int k = 0; k < 8; ++k) stats->visits[k] = v[k];
            if (v[15]) return fail(ctx, RT_ERR_INVALID, "traversal stack overflow");
        }
    }
    return rc;
}

int32_t rt_gen_surface_rays(rt_scene *scene, uint64_t seed, int64_t n, rt_ray *rays) {
    if (!scene || n < 0 || (n > 0 && !rays)) return RT_ERR_INVALID;
    rt_ctx *ctx = scene->ctx;
    if (scene->n <= 0) return fail(ctx, RT_ERR_INVALID, "rt_gen_surface_rays: empty scene");
    if (int32_t rc = rt::set_device(ctx)) return rc;
    if (n == 0) return RT_OK;
    hipLaunchKernelGGL(rt::k_surface_rays, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, ctx->stream,
                       scene->d_abc, scene->d_n, scene->n, seed, n, reinterpret_cast<float4 *>(rays));
    RT_HIP_CHECK(ctx, hipGetLastError());
    RT_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return RT_OK;
}

int32_t rt_ctx_frame_rays(rt_ctx *ctx, uint64_t rays[4]) {
    if (!ctx || !rays) return RT_ERR_INVALID;
    if (int32_t rc = rt::set_device(ctx)) return rc;
    unsigned long long v[4];
    RT_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    RT_HIP_CHECK(ctx, hipMemcpy(v, ctx->d_frame_rays, sizeof(v), hipMemcpyDeviceToHost));
    RT_HIP_CHECK(ctx, hipMemset(ctx->d_frame_rays, 0, sizeof(v)));
    for (int k = 0; k < 4; ++k) rays[k] = v[k];
    return RT_OK;
}

int32_t rt_set_counting(rt_ctx *ctx, int32_t on) {
    if (!ctx) return RT_ERR_INVALID;
    ctx->counting = on != 0;
    return RT_OK;
}

}  // extern "C"
