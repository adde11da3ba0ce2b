// Adjoint of the frame's shading and transport with respect to per-triangle
// albedo and the point lights' positions and intensities.
//
// The reference renders only forward (rt/raytracer.cc:20-87); this is the
// vector-Jacobian product of that frame. Per sample the radiance is
//   c = sum_l V1l S1l kd1 I1l + (kd1 / 255) * sum_l V2l S2l kd2 I2l
// with V the shadow-ray visibility, S = max-gated dot(dir_l, n) (lights
// below the surface are skipped), I = rgb_l / 255 / ((len_l d_l)^2 + 1) and
// kd = 255 * albedo (rt/materials/phong_shader.cc:21-64, rt/light.h:44-59);
// pixel = sum_s c_s / spp. Hit points, normals and visibility are held fixed
// (traversal and hit ids are the forward's, exactly), so the derivatives are
//   dc/dkd1   = sum_l V1 S1 I1 + c2 / 255          dc/dkd2 = (kd1 / 255) sum_l V2 S2 I2
//   dc/drgb_l = V S kd / (255 D)                   D = (len d)^2 + 1
//   dc/dpos_l = V kd I [(n - S dir) / len - S (2 d^2 len / D) dir]
// (the bounce terms carry the extra factor kd1 / 255).
//
// Records: the frame workspace of the forward pass holds every path record
// (queues, hits, normals, shadow rays, visibility). When the whole frame's
// records fit the caller's budget the tape keeps them and the adjoint is one
// pass over them; otherwise the tape keeps only the frame's inputs and the
// adjoint re-traces the frame chunk by chunk (path replay), each chunk's
// records within the budget, and runs the same pass on each chunk.
//
// Albedo gradients accumulate with float atomics; light gradients are summed
// per block into doubles and then over blocks in a fixed order.
#include <algorithm>
#include <vector>

#include "frame.hpp"
#include "rt_device.hpp"

struct rt_tape {
    rt_accel *accel = nullptr;
    rt::FramePlan plan;
    bool replay = false;
    int64_t record_bytes = 0;  // the whole frame's records
    int64_t chunk_tiles = 0;
    rt::Workspace *ws = nullptr;
    double *d_light = nullptr;  // per-block light partials, then the accumulator
    int64_t light_blocks = 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
};

namespace rt {

namespace {

constexpr int kLightGrad = 6;  // position xyz, rgb
constexpr int kLightSlots = kMaxLights * kLightGrad;

__global__ void __launch_bounds__(256) k_set_kd(const float *__restrict__ albedo, int64_t n, float4 *__restrict__ kd) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // phong_shader(rgb = 255 * albedo, kd = 1.0): kd = rgb * 1.0f.
    kd[i] = make_float4((albedo[3 * i + 0] * 255.0f) * 1.0f, (albedo[3 * i + 1] * 255.0f) * 1.0f,
                        (albedo[3 * i + 2] * 255.0f) * 1.0f, 0.0f);
}

struct AdjArgs {
    TileMap tm;
    int64_t slot0, nslots;
    int spp;
    int bounce;
    const rt_hit *hits1;
    const int32_t *qidx1;
    const float4 *nshade1;
    const rt_ray *shadow1;
    const uint8_t *occ1;
    const int32_t *qidx2;
    const rt_hit *hits2;
    const float4 *nshade2;
    const rt_ray *shadow2;
    const uint8_t *occ2;
    const float4 *kd;
    LightArgs lights;
    const float *grad;     // dL/d pixel radiance, (y * x_res + x) * 3
    float *d_albedo;       // n x 3, accumulated
    double *light_part;    // per block: kLightSlots partial sums
};

struct Col3 {
    float r, g, b;
};

// One surface's direct light: adds w . dc/dkd to dk, the light gradients to
// acc, and returns the surface's direct radiance.
__device__ __forceinline__ Col3 direct_adjoint(const rt_ray *shadow, const uint8_t *occ, int64_t q, rtd::V3 n,
                                               float4 kd, const LightArgs &L, Col3 w, Col3 &dk,
                                               float (&acc)[kMaxLights][kLightGrad]) {
    Col3 c{0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int l = 0; l < kMaxLights; ++l) {
        if (l >= L.n) break;
        const int64_t k = q * L.n + l;
        if (occ[k]) continue;
        const float4 s1 = reinterpret_cast<const float4 *>(shadow)[2 * k + 1];
        const rtd::V3 dir = rtd::f4v(s1);
        const float shade = rtd::dot(dir, n);
        if (shade < 0.0f) continue;
        const float len = s1.w;
        const float d = L.pos_d[l].w;
        const float dist = len * d;
        const float D = (dist * dist) + 1.0f;
        const float4 rgb = L.rgb[l];
        const float ir = rgb.x / D, ig = rgb.y / D, ib = rgb.z / D;
        c.r += ir * (shade * kd.x);
        c.g += ig * (shade * kd.y);
        c.b += ib * (shade * kd.z);
        dk.r += w.r * (shade * ir);
        dk.g += w.g * (shade * ig);
        dk.b += w.b * (shade * ib);
        const float sD = shade / (255.0f * D);
        acc[l][3] += w.r * kd.x * sD;
        acc[l][4] += w.g * kd.y * sD;
        acc[l][5] += w.b * kd.z * sD;
        const float T = (w.r * kd.x) * ir + (w.g * kd.y) * ig + (w.b * kd.z) * ib;
        const float inv_len = 1.0f / len;
        const float fall = shade * (2.0f * d * d * len) / D;
        acc[l][0] += T * ((n.x - shade * dir.x) * inv_len - fall * dir.x);
        acc[l][1] += T * ((n.y - shade * dir.y) * inv_len - fall * dir.y);
        acc[l][2] += T * ((n.z - shade * dir.z) * inv_len - fall * dir.z);
    }
    return c;
}

__device__ __forceinline__ void add_albedo(float *d_albedo, int32_t prim, Col3 dk) {
    float *g = d_albedo + 3 * static_cast<int64_t>(prim);
    // kd = 255 * albedo
    unsafeAtomicAdd(g + 0, 255.0f * dk.r);
    unsafeAtomicAdd(g + 1, 255.0f * dk.g);
    unsafeAtomicAdd(g + 2, 255.0f * dk.b);
}

// One thread per pixel slot of the chunk (all spp samples of the pixel).
__global__ void __launch_bounds__(256) k_adjoint(AdjArgs a) {
    __shared__ float part[4][kLightSlots];
    float acc[kMaxLights][kLightGrad];
#pragma unroll
    for (int l = 0; l < kMaxLights; ++l)
#pragma unroll
        for (int k = 0; k < kLightGrad; ++k) acc[l][k] = 0.0f;
    const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    int x = 0, y = 0;
    if (j < a.nslots && a.tm.pixel_of_slot(a.slot0 + j, &x, &y)) {
        const int64_t p = static_cast<int64_t>(y) * a.tm.x_res + x;
        const float fs = static_cast<float>(a.spp);
        // pixel = (sum_s c_s) / spp
        const Col3 g{a.grad[3 * p + 0] / fs, a.grad[3 * p + 1] / fs, a.grad[3 * p + 2] / fs};
        for (int s = 0; s < a.spp; ++s) {
            const int64_t i = j * a.spp + s;
            const int32_t q = a.qidx1[i];
            if (q < 0) continue;
            const int32_t prim = __float_as_int(reinterpret_cast<const float4 *>(a.hits1)[i].w);
            const float4 kd1 = a.kd[prim];
            Col3 dk1{0.0f, 0.0f, 0.0f};
            if (a.bounce) {
                const int32_t q2 = a.qidx2[q];
                if (q2 >= 0) {
                    const int32_t prim2 = __float_as_int(reinterpret_cast<const float4 *>(a.hits2)[q].w);
                    const Col3 w2{g.r * (kd1.x / 255.0f), g.g * (kd1.y / 255.0f), g.b * (kd1.z / 255.0f)};
                    Col3 dk2{0.0f, 0.0f, 0.0f};
                    const Col3 c2 = direct_adjoint(a.shadow2, a.occ2, q2, rtd::f4v(a.nshade2[q2]), a.kd[prim2],
                                                   a.lights, w2, dk2, acc);
                    add_albedo(a.d_albedo, prim2, dk2);
                    dk1.r += g.r * (c2.r / 255.0f);
                    dk1.g += g.g * (c2.g / 255.0f);
                    dk1.b += g.b * (c2.b / 255.0f);
                }
            }
            (void)direct_adjoint(a.shadow1, a.occ1, q, rtd::f4v(a.nshade1[q]), kd1, a.lights, g, dk1, acc);
            add_albedo(a.d_albedo, prim, dk1);
        }
    }
    // Light gradients: wave sums by shuffles, then the block's four waves.
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int l = 0; l < kMaxLights; ++l) {
        if (l >= a.lights.n) break;
#pragma unroll
        for (int k = 0; k < kLightGrad; ++k) {
            float v = acc[l][k];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
            if (lane == 0) part[wave][l * kLightGrad + k] = v;
        }
    }
    __syncthreads();
    const int t = threadIdx.x;
    if (t < a.lights.n * kLightGrad) {
        const double v = (static_cast<double>(part[0][t]) + static_cast<double>(part[1][t])) +
                         (static_cast<double>(part[2][t]) + static_cast<double>(part[3][t]));
        a.light_part[static_cast<int64_t>(blockIdx.x) * kLightSlots + t] = v;
    }
}

// acc[t] += sum over blocks of part[b][t]: one 256-thread block per slot t,
// each thread a fixed stride of blocks, then a fixed-shape tree (the result
// does not depend on scheduling).
__global__ void __launch_bounds__(256) k_light_sum(const double *__restrict__ part, int64_t nblocks,
                                                   double *__restrict__ acc) {
    __shared__ double wsum[4];
    const int t = blockIdx.x;
    double s = 0.0;
    for (int64_t b = threadIdx.x; b < nblocks; b += 256) s += part[b * kLightSlots + t];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) acc[t] += (wsum[0] + wsum[1]) + (wsum[2] + wsum[3]);
}

__global__ void k_fill_f64(double *p, int n) {
    if (static_cast<int>(threadIdx.x) < n) p[threadIdx.x] = 0.0;
}

__global__ void k_light_out(const double *__restrict__ acc, int n, float *__restrict__ d_lights) {
    const int t = threadIdx.x;
    if (t < n) d_lights[t] += static_cast<float>(acc[t]);
}

void free_tape(rt_tape *t) {
    if (!t) return;
    if (t->ws) {
        // Hand the records' memory back to the accel for its next frame.
        if (!t->accel->frame_ws)
            t->accel->frame_ws = t->ws;
        else
            workspace_release(t->ws);
    }
    if (t->d_light) (void)hipFree(t->d_light);
    if (t->e0) (void)hipEventDestroy(t->e0);
    if (t->e1) (void)hipEventDestroy(t->e1);
    release_accel(t->accel);
    delete t;
}

int32_t adjoint_chunk(rt_tape *t, int64_t tile0, int64_t ntiles, const float *grad, float *d_albedo) {
    const FramePlan &f = t->plan;
    rt_ctx *ctx = f.a->scene->ctx;
    const Workspace *w = t->ws;
    AdjArgs aa{};
    aa.tm = f.tm;
    aa.slot0 = tile0 * 64;
    aa.nslots = ntiles * 64;
    aa.spp = f.spp;
    aa.bounce = f.p.bounce;
    aa.hits1 = w->hits1;
    aa.qidx1 = w->qidx1;
    aa.nshade1 = w->nshade1;
    aa.shadow1 = w->shadow1;
    aa.occ1 = w->occ1;
    aa.qidx2 = w->qidx2;
    aa.hits2 = w->hits2;
    aa.nshade2 = w->nshade2;
    aa.shadow2 = w->shadow2;
    aa.occ2 = w->occ2;
    aa.kd = f.a->scene->d_kd;
    aa.lights = f.lights;
    aa.grad = grad;
    aa.d_albedo = d_albedo;
    aa.light_part = t->d_light;
    const int64_t blocks = (aa.nslots + 255) / 256;
    if (blocks > t->light_blocks) return fail(ctx, RT_ERR_INVALID, "adjoint: chunk larger than the tape's buffers");
    hipLaunchKernelGGL(k_adjoint, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, ctx->stream, aa);
    if (f.nlights > 0)
        hipLaunchKernelGGL(k_light_sum, dim3(f.nlights * kLightGrad), dim3(256), 0, ctx->stream, t->d_light, blocks,
                           t->d_light + t->light_blocks * kLightSlots);
    RT_HIP_CHECK(ctx, hipGetLastError());
    return RT_OK;
}

}  // namespace
}  // namespace rt

using rt::fail;

extern "C" {

int32_t rt_scene_set_albedo(rt_scene *scene, const float *albedo, uint32_t flags) {
    if (!scene || (!albedo && scene->n > 0)) return RT_ERR_INVALID;
    rt_ctx *ctx = scene->ctx;
    RT_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    if (scene->n == 0) return RT_OK;
    hipStream_t s = ctx->stream;
    const float *src = albedo;
    float *tmp = nullptr;
    if (!(flags & RT_DEVICE_PTRS)) {
        RT_HIP_CHECK(ctx, hipMalloc(&tmp, 3 * scene->n * sizeof(float)));
        hipError_t e = hipMemcpyAsync(tmp, albedo, 3 * scene->n * sizeof(float), hipMemcpyHostToDevice, s);
        if (e != hipSuccess) {
            (void)hipFree(tmp);
            return fail(ctx, RT_ERR_HIP, std::string("rt_scene_set_albedo: ") + hipGetErrorString(e));
        }
        src = tmp;
    }
    hipLaunchKernelGGL(rt::k_set_kd, dim3(static_cast<unsigned>((scene->n + 255) / 256)), dim3(256), 0, s, src,
                       scene->n, scene->d_kd);
    hipError_t e = hipGetLastError();
    // The source is borrowed for the call only: wait until the kernel has
    // read it (the host copy's staging buffer is freed, a device source may
    // be reused by the caller as soon as the call returns).
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (tmp) (void)hipFree(tmp);
    if (e != hipSuccess) return fail(ctx, RT_ERR_HIP, std::string("rt_scene_set_albedo: ") + hipGetErrorString(e));
    return RT_OK;
}

int32_t rt_render_tape(rt_accel *accel, const rt_camera *cam, const rt_light *lights, int32_t nlights,
                       const rt_render_params *params, float *rgb, int64_t max_record_bytes, rt_tape **out,
                       rt_stats *stats) {
    if (!accel || !cam || !rgb || !out || (nlights > 0 && !lights)) return RT_ERR_INVALID;
    *out = nullptr;
    rt_ctx *ctx = accel->scene->ctx;
    RT_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    if (accel->root == rt::kEmpty) return fail(ctx, RT_ERR_UNSUPPORTED, "rt_render_tape: empty scene");
    if (accel->scene->d_mat || accel->scene->d_vn)
        return fail(ctx, RT_ERR_UNSUPPORTED,
                    "rt_render_tape: scenes with materials or vertex normals are not differentiable");
    rt_render_params p{0, 1, 0, 0, 0};
    if (params) p = *params;
    p.flags |= RT_DEVICE_PTRS;
    rt_tape *t = new rt_tape();
    accel->refs.fetch_add(1);
    t->accel = accel;
    int32_t rc = rt::plan_frame(accel, *cam, lights, nlights, p, &t->plan);
    if (rc) {
        rt::free_tape(t);
        return rc;
    }
    rt::FramePlan &f = t->plan;
    t->record_bytes = rt::chunk_bytes(f, f.tiles);
    // A frame beyond one hit-compaction scan cannot keep all its records in
    // one chunk: it replays, in chunks of at most one scan.
    const int64_t scan_tiles = rt::scan_chunk_tiles(f);
    t->replay = (max_record_bytes > 0 && t->record_bytes > max_record_bytes) || f.tiles > scan_tiles;
    int64_t ct = f.tiles;
    if (t->replay && max_record_bytes > 0) ct = rt::tiles_within(f, max_record_bytes);
    t->chunk_tiles = std::min(ct, scan_tiles);
    if (t->chunk_tiles <= 0) {
        rt::free_tape(t);
        return fail(ctx, RT_ERR_INVALID, "rt_render_tape: max_record_bytes is below one tile's records");
    }
    f.chunk_tiles = t->chunk_tiles;
    // Reuse the accel's frame workspace for the records when it has one.
    t->ws = accel->frame_ws ? reinterpret_cast<rt::Workspace *>(accel->frame_ws) : new rt::Workspace();
    accel->frame_ws = nullptr;
    t->light_blocks = (t->chunk_tiles * 64 + 255) / 256;
    hipError_t e = hipMalloc(&t->d_light, (t->light_blocks + 1) * rt::kLightSlots * sizeof(double));
    if (e == hipSuccess) e = hipEventCreate(&t->e0);
    if (e == hipSuccess) e = hipEventCreate(&t->e1);
    if (e != hipSuccess) {
        rt::free_tape(t);
        return fail(ctx, RT_ERR_HIP, std::string("rt_render_tape: ") + hipGetErrorString(e));
    }
    if ((rc = rt::workspace_reserve(ctx, t->ws, t->chunk_tiles * 64 * f.spp, std::max(1, nlights)))) {
        rt::free_tape(t);
        return rc;
    }
    hipStream_t s = ctx->stream;
    uint64_t rays[4] = {0, 0, 0, 0};
    float stage[5] = {0, 0, 0, 0, 0};
    const bool single = t->chunk_tiles >= f.tiles;
    auto add_stats = [&](int64_t nt, const rt::ChunkStats &cs) {
        rays[0] += static_cast<uint64_t>(nt) * 64 * f.spp;
        rays[1] += cs.hits1 * nlights;
        rays[2] += p.bounce ? cs.hits1 : 0;
        rays[3] += p.bounce ? cs.hits2 * nlights : 0;
        stage[1] += cs.ms[1];
        stage[2] += cs.ms[3];
        stage[3] += cs.ms[4];
        stage[4] += cs.ms[6];
    };
    if (hipEventRecord(t->e0, s) != hipSuccess) rc = fail(ctx, RT_ERR_HIP, "hipEventRecord");
    for (int64_t t0 = 0; rc == RT_OK && t0 < f.tiles; t0 += t->chunk_tiles) {
        const int64_t nt = std::min(t->chunk_tiles, f.tiles - t0);
        rc = rt::trace_chunk(f, t->ws, t0, nt, stats != nullptr);
        if (rc == RT_OK) rc = rt::compose_chunk(f, t->ws, t0, nt, rgb);
        // Several chunks share the workspace: read each chunk's counts before
        // the next overwrites them (this synchronises).
        if (rc == RT_OK && stats && !single) {
            rt::ChunkStats cs;
            rc = rt::collect_chunk(f, t->ws, nt, &cs);
            add_stats(nt, cs);
        }
    }
    if (rc == RT_OK && hipEventRecord(t->e1, s) != hipSuccess) rc = fail(ctx, RT_ERR_HIP, "hipEventRecord");
    if (rc == RT_OK && stats && single) {
        rt::ChunkStats cs;
        rc = rt::collect_chunk(f, t->ws, f.tiles, &cs);
        add_stats(f.tiles, cs);
    }
    if (rc == RT_OK && hipEventSynchronize(t->e1) != hipSuccess) rc = fail(ctx, RT_ERR_HIP, "forward failed");
    if (rc) {
        rt::free_tape(t);
        return rc;
    }
    if (stats) {
        for (int k = 0; k < 4; ++k) stats->rays[k] = rays[k];
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, t->e0, t->e1);
        stats->ms[0] = ms;
        for (int k = 1; k < 5; ++k) stats->ms[k] = stage[k];
        stats->ms[5] = ms - (stage[1] + stage[2] + stage[3] + stage[4]);
        for (auto &v : stats->visits) v = 0;
    }
    *out = t;
    return RT_OK;
}

int32_t rt_tape_info(const rt_tape *tape, int32_t *replay, int64_t *record_bytes, int64_t *chunk_tiles) {
    if (!tape) return RT_ERR_INVALID;
    if (replay) *replay = tape->replay ? 1 : 0;
    if (record_bytes) *record_bytes = tape->record_bytes;
    if (chunk_tiles) *chunk_tiles = tape->chunk_tiles;
    return RT_OK;
}

int32_t rt_render_adjoint(rt_tape *t, const float *d_rgb, float *d_albedo, float *d_lights, float *ms) {
    if (!t || !d_rgb || !d_albedo || (t->plan.nlights > 0 && !d_lights)) return RT_ERR_INVALID;
    const rt::FramePlan &f = t->plan;
    rt_ctx *ctx = f.a->scene->ctx;
    RT_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    double *acc = t->d_light + t->light_blocks * rt::kLightSlots;
    RT_HIP_CHECK(ctx, hipEventRecord(t->e0, s));
    hipLaunchKernelGGL(rt::k_fill_f64, dim3(1), dim3(64), 0, s, acc, rt::kLightSlots);
    int32_t rc = RT_OK;
    if (!t->replay) {
        rc = rt::adjoint_chunk(t, 0, f.tiles, d_rgb, d_albedo);
    } else {
        for (int64_t t0 = 0; rc == RT_OK && t0 < f.tiles; t0 += t->chunk_tiles) {
            const int64_t nt = std::min(t->chunk_tiles, f.tiles - t0);
            rc = rt::trace_chunk(f, t->ws, t0, nt, false);
            if (rc == RT_OK) rc = rt::adjoint_chunk(t, t0, nt, d_rgb, d_albedo);
        }
    }
    if (rc) return rc;
    if (f.nlights > 0)
        hipLaunchKernelGGL(rt::k_light_out, dim3(1), dim3(64), 0, s, acc, f.nlights * rt::kLightGrad, d_lights);
    RT_HIP_CHECK(ctx, hipGetLastError());
    RT_HIP_CHECK(ctx, hipEventRecord(t->e1, s));
    RT_HIP_CHECK(ctx, hipEventSynchronize(t->e1));
    if (ms) RT_HIP_CHECK(ctx, hipEventElapsedTime(ms, t->e0, t->e1));
    return RT_OK;
}

void rt_tape_destroy(rt_tape *tape) {
    if (!tape) return;
    (void)hipSetDevice(tape->accel->scene->ctx->device);
    (void)hipStreamSynchronize(tape->accel->scene->ctx->stream);
    rt::free_tape(tape);
}

}  // extern "C"
