// One frame of the tracer as a wavefront pipeline of kernels on the context
// stream. Replaces the per-pixel loop of ray_tracer() (rt/raytracer.cc:563-598)
// and ray_trace_engine::ray_trace (rt/raytracer.cc:20-87, 168-461):
//
//   k_raygen        primary rays, rt/camera.h:127-145 (EXACT_NORMALISE)
//   closest hit     primary rays (trace.hip)
//   k_shade_hit<0>  for each primary hit: move to the hit and re-project onto
//                   the plane (rt/ray.cc:15-20, rt/triangle.h:464-498), one
//                   shadow ray per light (rt/ray.cc:91-104, rt/ray.h:111-157,
//                   rt/light.h:96-100) and the diffuse bounce ray; hits are
//                   compacted into queues with one atomic per wavefront
//   any hit         shadow queue
//   closest hit     bounce queue
//   k_shade_hit<1>  shadow rays of the bounce hits (same arithmetic)
//   any hit         bounce-shadow queue
//   k_compose       Lambert shading, rt/materials/phong_shader.cc:21-64 with
//                   rt/light.h:44-59, bounce = (kd1 / 255) * direct(x2)
//
// Queue lengths stay on the device; trace kernels read them, so the frame has
// no host round trip until its statistics are read back.
#include <algorithm>

#include "frame.hpp"
#include "radix_sort.hpp"
#include "rt_device.hpp"

namespace rt {

namespace {

__global__ void __launch_bounds__(256) k_raygen(CamArgs c, TileMap tm, int spp, uint64_t seed, int64_t slot0,
                                                int64_t nsamples, rt_ray *__restrict__ rays) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= nsamples) return;
    const int64_t slot = slot0 + i / spp;
    const int s = static_cast<int>(i % spp);
    int x, y;
    float4 *r = reinterpret_cast<float4 *>(rays) + 2 * i;
    if (!tm.pixel_of_slot(slot, &x, &y)) {
        // Padding slot outside the image: tmax < tmin, so nothing is accepted.
        r[0] = make_float4(c.px, c.py, c.pz, rtd::kEpsilon);
        r[1] = make_float4(0.0f, 0.0f, 1.0f, -1.0f);
        return;
    }
    const int64_t p = static_cast<int64_t>(y) * c.x_res + x;
    const rtd::V3 d = primary_dir(c, x, y, p, s, spp, seed);
    r[0] = make_float4(c.px, c.py, c.pz, rtd::kEpsilon);
    r[1] = make_float4(d.x, d.y, d.z, rtd::kMaxDist);
}

struct ShadeArgs {
    TileMap tm;
    const rt_ray *rays;      // rays that were traced
    const rt_hit *hits;      // their hits
    const uint32_t *count;   // number of rays (nullptr: n_static)
    int64_t n_static;
    const float4 *tri_n, *tri_c;
    LightArgs lights;
    int bounce;
    int spp;
    uint64_t seed;
    int64_t slot0;              // first slot of the chunk
    const int32_t *sample_of;   // for bounce hits: the sample index of each bounce ray
    int32_t *qidx;              // per ray: compacted hit index or -1
    float4 *nshade;             // per compacted hit: shading normal
    int32_t *hit_sample;        // per compacted hit: sample index
    rt_ray *shadow_q;           // per compacted hit x light
    rt_ray *bounce_q;           // per compacted hit (stage 0 with bounce)
    const uint32_t *block_base; // exclusive prefix of the per-block hit counts
};

// Deterministic stream compaction of hits (no contended atomics: a single
// counter hit once per wavefront serialises across the eight XCDs):
//   k_count_hits    hits per 256-thread block
//   k_scan_blocks   exclusive prefix over the block counts, total in [nblocks]
//   k_shade_hit     each hit's slot = block base + rank inside the block
// The compacted order is the sample order, so queues are reproducible.
__global__ void __launch_bounds__(256) k_count_hits(const rt_hit *__restrict__ hits, const uint32_t *count,
                                                    int64_t n_static, uint32_t *__restrict__ block_counts) {
    __shared__ uint32_t wsum[4];
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t n = count ? static_cast<int64_t>(count[0]) : n_static;
    bool hit = false;
    if (i < n) hit = __float_as_int(reinterpret_cast<const float4 *>(hits)[i].w) >= 0;
    const uint64_t m = __ballot(hit);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = static_cast<uint32_t>(__popcll(m));
    __syncthreads();
    if (threadIdx.x == 0) block_counts[blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

// Rays of a frame chunk enqueued with RT_ASYNC, added to the context's sums
// (primary, shadow, bounce, bounce shadow) from the compaction totals: one
// shadow ray per hit per light, one bounce ray per primary hit. Launches on
// one stream run in order, so one thread updates the sums.
__global__ void k_add_frame_rays(const uint32_t *hits1, const uint32_t *hits2, uint64_t primary, int nlights,
                                 unsigned long long *sums) {
    if (threadIdx.x != 0) return;
    const unsigned long long h1 = hits1[0];
    sums[0] += primary;
    sums[1] += h1 * static_cast<unsigned long long>(nlights);
    if (hits2) {
        sums[2] += h1;
        sums[3] += static_cast<unsigned long long>(hits2[0]) * static_cast<unsigned long long>(nlights);
    }
}

// One block of 1024 threads scans up to 1024 * kScanPer block counts.
constexpr int kScanPer = 32;
__global__ void __launch_bounds__(1024) k_scan_blocks(uint32_t *__restrict__ v, int n) {
    __shared__ uint32_t wtot[16];
    const int t = threadIdx.x;
    uint32_t loc[kScanPer];
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        const int idx = t * kScanPer + k;
        loc[k] = (idx < n) ? v[idx] : 0u;
        sum += loc[k];
    }
    const int lane = t & 63, wave = t >> 6;
    uint32_t x = sum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wtot[wave] = x;
    __syncthreads();
    uint32_t wbase = 0, total = 0;
    for (int w = 0; w < 16; ++w) {
        if (w < wave) wbase += wtot[w];
        total += wtot[w];
    }
    uint32_t run = wbase + x - sum;
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        const int idx = t * kScanPer + k;
        if (idx < n) v[idx] = run;
        run += loc[k];
    }
    if (t == 0) v[n] = total;
}

template <int STAGE>
__global__ void __launch_bounds__(256) k_shade_hit(ShadeArgs a) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t n = a.count ? static_cast<int64_t>(a.count[0]) : a.n_static;
    const bool in = i < n;
    float4 h = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(-1));
    if (in) h = reinterpret_cast<const float4 *>(a.hits)[i];
    const int32_t prim = __float_as_int(h.w);
    const bool hit = in && prim >= 0;
    // Rank inside the block: ballot within the wave, wave totals through LDS.
    __shared__ uint32_t wcount[4];
    const uint64_t mask = __ballot(hit);
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    if (lane == 0) wcount[wave] = static_cast<uint32_t>(__popcll(mask));
    __syncthreads();
    uint32_t base = a.block_base[blockIdx.x];
    for (int w = 0; w < wave; ++w) base += wcount[w];
    if (!in) return;
    if (!hit) {
        a.qidx[i] = -1;
        return;
    }
    const uint32_t q = base + static_cast<uint32_t>(__popcll(mask & lt));
    a.qidx[i] = static_cast<int32_t>(q);
    const float4 r0 = reinterpret_cast<const float4 *>(a.rays)[2 * i];
    const float4 r1 = reinterpret_cast<const float4 *>(a.rays)[2 * i + 1];
    const float4 tn = a.tri_n[prim];
    const float4 tc = a.tri_c[prim];
    rtd::V3 dst, geo, shn;
    surface_at(rtd::f4v(r0), rtd::f4v(r1), h.x, rtd::f4v(tn), rtd::f4v(tc), &dst, &geo, &shn);
    a.nshade[q] = make_float4(shn.x, shn.y, shn.z, 0.0f);
    const int32_t sample = (STAGE == 0) ? static_cast<int32_t>(i) : a.sample_of[i];
    a.hit_sample[q] = sample;
    const rtd::V3 org = rtd::offset_start_point(dst, geo, 1);
    for (int l = 0; l < a.lights.n; ++l) {
        rtd::V3 dn;
        float len;
        light_dir(org, a.lights.pos_d[l], &dn, &len);
        float4 *sr = reinterpret_cast<float4 *>(a.shadow_q) + 2 * (static_cast<int64_t>(q) * a.lights.n + l);
        sr[0] = make_float4(org.x, org.y, org.z, rtd::kEpsilon);
        sr[1] = make_float4(dn.x, dn.y, dn.z, len);
    }
    if (STAGE == 0 && a.bounce) {
        const int64_t slot = a.slot0 + i / a.spp;
        const int s = static_cast<int>(i % a.spp);
        int x, y;
        a.tm.pixel_of_slot(slot, &x, &y);
        const int64_t p = static_cast<int64_t>(y) * a.tm.x_res + x;
        const rtd::V3 bd = rtd::cosine_direction(
            rtd::sample_key(a.seed, static_cast<uint32_t>(p), static_cast<uint32_t>(s), 1u), shn);
        float4 *br = reinterpret_cast<float4 *>(a.bounce_q) + 2 * static_cast<int64_t>(q);
        br[0] = make_float4(org.x, org.y, org.z, rtd::kEpsilon);
        br[1] = make_float4(bd.x, bd.y, bd.z, rtd::kMaxDist);
    }
}

// Regrouping of the incoherent bounce rays before they are traced: key =
// direction octant (3 bits) above a 27-bit Morton code of the origin in the
// scene box, so a wavefront gets rays that leave nearby points in similar
// directions. Entries past the bounce count sort last (key 0xffffffff).
__global__ void __launch_bounds__(256) k_bounce_keys(const rt_ray *__restrict__ rays, const uint32_t *count,
                                                     int64_t n_max, float4 scene_lo, float4 scene_inv,
                                                     uint32_t *__restrict__ keys, uint32_t *__restrict__ vals) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n_max) return;
    uint32_t key = 0xffffffffu;
    if (i < static_cast<int64_t>(count[0])) {
        const float4 r0 = reinterpret_cast<const float4 *>(rays)[2 * i];
        const float4 r1 = reinterpret_cast<const float4 *>(rays)[2 * i + 1];
        const uint32_t oct = (r1.x < 0.0f ? 1u : 0u) | (r1.y < 0.0f ? 2u : 0u) | (r1.z < 0.0f ? 4u : 0u);
        auto q = [](float v) {
            return static_cast<uint32_t>(fminf(fmaxf(v * 512.0f, 0.0f), 511.0f));
        };
        const uint32_t qx = q((r0.x - scene_lo.x) * scene_inv.x);
        const uint32_t qy = q((r0.y - scene_lo.y) * scene_inv.y);
        const uint32_t qz = q((r0.z - scene_lo.z) * scene_inv.z);
        uint32_t m = 0;
#pragma unroll
        for (int b = 0; b < 9; ++b)
            m |= (((qx >> b) & 1u) << (3 * b)) | (((qy >> b) & 1u) << (3 * b + 1)) | (((qz >> b) & 1u) << (3 * b + 2));
        key = (oct << 27) | m;
    }
    keys[i] = key;
    vals[i] = static_cast<uint32_t>(i);
}

struct ComposeArgs {
    TileMap tm;
    int64_t slot0, nslots;
    int spp;
    int bounce;
    const rt_hit *hits1;
    const int32_t *qidx1;        // per sample
    const float4 *nshade1;       // per primary hit
    const rt_ray *shadow1;       // per primary hit x light
    const uint8_t *occ1;
    const int32_t *qidx2;        // per primary hit (bounce ray): bounce hit index or -1
    const rt_hit *hits2;         // per primary hit (bounce ray)
    const float4 *nshade2;       // per bounce hit
    const rt_ray *shadow2;
    const uint8_t *occ2;
    const float4 *kd;
    LightArgs lights;
    float bg_r, bg_g, bg_b;  // the camera's background colour
    float *rgb;
};

using Col = Rgb;

// Direct light at one surface from its shadow rays and their visibility.
__device__ __forceinline__ Col direct(const rt_ray *shadow, const uint8_t *occ, int64_t q, rtd::V3 nsh, float4 kd,
                                      const LightArgs &L) {
    Col c{0.0f, 0.0f, 0.0f};
    for (int l = 0; l < L.n; ++l) {
        const int64_t k = q * L.n + l;
        const float4 s1 = reinterpret_cast<const float4 *>(shadow)[2 * k + 1];
        lambert_term(c, rtd::f4v(s1), s1.w, occ[k] != 0, nsh, kd, L.pos_d[l].w, L.rgb[l]);
    }
    c.r = c.r + 0.0f;  // + ka (zero)
    c.g = c.g + 0.0f;
    c.b = c.b + 0.0f;
    return c;
}

__global__ void __launch_bounds__(256) k_compose(ComposeArgs a) {
    const int64_t slot = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;  // in the chunk
    if (slot >= a.nslots) return;
    int x, y;
    if (!a.tm.pixel_of_slot(a.slot0 + slot, &x, &y)) return;
    const int64_t p = static_cast<int64_t>(y) * a.tm.x_res + x;
    Col acc{0.0f, 0.0f, 0.0f};
    for (int s = 0; s < a.spp; ++s) {
        const int64_t i = slot * a.spp + s;
        Col c{a.bg_r, a.bg_g, a.bg_b};  // a miss takes the background (camera::shade)
        const int32_t q = a.qidx1[i];
        if (q >= 0) {
            const int32_t prim = __float_as_int(reinterpret_cast<const float4 *>(a.hits1)[i].w);
            const float4 kd1 = a.kd[prim];
            c = direct(a.shadow1, a.occ1, q, rtd::f4v(a.nshade1[q]), kd1, a.lights);
            if (a.bounce) {
                const int32_t q2 = a.qidx2[q];
                if (q2 >= 0) {
                    const int32_t prim2 = __float_as_int(reinterpret_cast<const float4 *>(a.hits2)[q].w);
                    const Col c2 = direct(a.shadow2, a.occ2, q2, rtd::f4v(a.nshade2[q2]), a.kd[prim2], a.lights);
                    c.r = c.r + ((kd1.x / 255.0f) * c2.r);
                    c.g = c.g + ((kd1.y / 255.0f) * c2.g);
                    c.b = c.b + ((kd1.z / 255.0f) * c2.b);
                }
            }
        }
        acc.r = acc.r + c.r;
        acc.g = acc.g + c.g;
        acc.b = acc.b + c.b;
    }
    const float fs = static_cast<float>(a.spp);
    a.rgb[3 * p + 0] = acc.r / fs;
    a.rgb[3 * p + 1] = acc.g / fs;
    a.rgb[3 * p + 2] = acc.b / fs;
}

// Hits back in pixel order (p * spp + s) for the caller, for the samples of
// one chunk (slots [slot0, slot0 + nslots)). For bounce hits the queue is in
// compaction order, so go through the per-sample queue index.
__global__ void __launch_bounds__(256) k_hits_out(TileMap tm, int spp, int64_t slot0, int64_t nslots,
                                                  const int32_t *qidx1, const rt_hit *hits, rt_hit *out) {
    const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= nslots * spp) return;
    const int64_t slot = slot0 + j / spp;
    const int s = static_cast<int>(j % spp);
    int x, y;
    if (!tm.pixel_of_slot(slot, &x, &y)) return;  // tile padding
    int64_t k = j;
    if (qidx1) k = qidx1[j];
    float4 v = make_float4(rtd::kMaxDist, 0.0f, 0.0f, __int_as_float(-1));
    if (k >= 0) v = reinterpret_cast<const float4 *>(hits)[k];
    reinterpret_cast<float4 *>(out)[(static_cast<int64_t>(y) * tm.x_res + x) * spp + s] = v;
}

template <class T>
T *carve(char *&p, int64_t count) {
    T *r = reinterpret_cast<T *>(p);
    p += ((count * static_cast<int64_t>(sizeof(T)) + 255) / 256) * 256;
    return r;
}

}  // namespace

size_t workspace_layout(Workspace &w, char *base, int64_t ns, int nl) {
    char *p = base;
    w.rays1 = carve<rt_ray>(p, ns);
    w.hits1 = carve<rt_hit>(p, ns);
    w.qidx1 = carve<int32_t>(p, ns);
    w.nshade1 = carve<float4>(p, ns);
    w.hit_sample1 = carve<int32_t>(p, ns);
    w.shadow1 = carve<rt_ray>(p, ns * nl);
    w.occ1 = carve<uint8_t>(p, ns * nl);
    w.bounce = carve<rt_ray>(p, ns);
    w.hits2 = carve<rt_hit>(p, ns);
    w.qidx2 = carve<int32_t>(p, ns);
    w.nshade2 = carve<float4>(p, ns);
    w.hit_sample2 = carve<int32_t>(p, ns);
    w.shadow2 = carve<rt_ray>(p, ns * nl);
    w.occ2 = carve<uint8_t>(p, ns * nl);
    w.rgb = carve<float>(p, 3 * ns);
    const int64_t nb = (ns + 255) / 256 + 1;
    w.blocks1 = carve<uint32_t>(p, nb);
    w.blocks2 = carve<uint32_t>(p, nb);
    w.skeys = carve<uint32_t>(p, ns);
    w.svals = carve<uint32_t>(p, ns);
    w.skeys_t = carve<uint32_t>(p, ns);
    w.svals_t = carve<uint32_t>(p, ns);
    w.shist = carve<uint32_t>(p, 256 * ((ns + kSortTile - 1) / kSortTile));
    w.ssums = carve<uint32_t>(p, kScanTile);
    return static_cast<size_t>(p - base) + 256;
}

int32_t workspace_reserve(rt_ctx *ctx, Workspace *w, int64_t ns, int nl) {
    if (w->nsamples >= ns && w->nlights >= nl && w->mem) {
        workspace_layout(*w, reinterpret_cast<char *>(w->mem), w->nsamples, w->nlights);
        return RT_OK;
    }
    // Frames enqueued without waiting (RT_ASYNC) may still use the old buffer.
    if (w->mem) RT_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (w->mem) (void)hipFree(w->mem);
    w->mem = nullptr;
    w->nsamples = 0;
    Workspace probe;
    const size_t bytes = workspace_layout(probe, reinterpret_cast<char *>(0), ns, nl);
    hipError_t e = hipMalloc(&w->mem, bytes);
    if (e != hipSuccess) {
        w->mem = nullptr;
        return fail(ctx, e == hipErrorOutOfMemory ? RT_ERR_OOM : RT_ERR_HIP,
                    std::string("frame workspace: ") + hipGetErrorString(e));
    }
    w->bytes = bytes;
    w->nsamples = ns;
    w->nlights = nl;
    workspace_layout(*w, reinterpret_cast<char *>(w->mem), ns, nl);
    return RT_OK;
}

void workspace_release(Workspace *w) {
    if (!w) return;
    if (w->mem) (void)hipFree(w->mem);
    if (w->frame_mem) (void)hipFree(w->frame_mem);
    delete w;
}

// Workspace of rt_render_static: one array per GPU, allocated when the
// library's code object is loaded on that device and held until the process
// ends, so a frame needs no hipMalloc. In host code `scratch` is only a
// stand-in for the symbol; its device address comes from
// static_scratch_address, once per device, and only that address is handed
// to kernels.
static __device__ __align__(256) char scratch[kStaticScratchBytes];

int32_t static_scratch_address(rt_ctx *ctx, void **ptr) {
    *ptr = nullptr;
    void *p = nullptr;
    RT_HIP_CHECK(ctx, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(ctx, RT_ERR_HIP, "static scratch: no device address");
    *ptr = p;
    return RT_OK;
}

int64_t scan_chunk_tiles(const FramePlan &plan) {
    // The hit compaction scans at most 1024 * kScanPer blocks of 256 samples.
    return std::max<int64_t>(1, (int64_t(1024) * kScanPer * 256) / (int64_t(64) * plan.spp));
}

int32_t plan_frame(rt_accel *a, const rt_camera &cam, const rt_light *lights, int nlights, const rt_render_params &p,
                   FramePlan *plan) {
    rt_ctx *ctx = a->scene->ctx;
    if (nlights < 0 || nlights > kMaxLights) return fail(ctx, RT_ERR_INVALID, "nlights must be 0..8");
    if (nlights > 0 && !lights) return fail(ctx, RT_ERR_INVALID, "null light array");
    if (cam.x_res <= 0 || cam.y_res <= 0) return fail(ctx, RT_ERR_INVALID, "camera resolution must be positive");
    FramePlan &f = *plan;
    f.a = a;
    f.p = p;
    f.spp = std::max(1, p.spp);
    f.nlights = nlights;
    f.npix = static_cast<int64_t>(cam.x_res) * cam.y_res;
    f.tm.x_res = cam.x_res;
    f.tm.y_res = cam.y_res;
    f.tm.tiles_x = (cam.x_res + 7) / 8;
    f.tiles = static_cast<int64_t>(f.tm.tiles_x) * ((cam.y_res + 7) / 8);
    f.chunk_tiles = f.tiles;
    CamArgs &c = f.cam;
    c.px = cam.pos[0]; c.py = cam.pos[1]; c.pz = cam.pos[2];
    c.xx = cam.x_axis[0]; c.xy = cam.x_axis[1]; c.xz = cam.x_axis[2];
    c.yx = cam.y_axis[0]; c.yy = cam.y_axis[1]; c.yz = cam.y_axis[2];
    c.zx = cam.z_axis[0]; c.zy = cam.z_axis[1]; c.zz = cam.z_axis[2];
    // rt/camera.h:54-62
    c.x_m = -cam.w;
    c.y_m = -cam.h;
    c.x_inc = (cam.w * 2.0f) / static_cast<float>(cam.x_res);
    c.y_inc = (cam.h * 2.0f) / static_cast<float>(cam.y_res);
    c.t = cam.t;
    c.x_res = cam.x_res;
    c.y_res = cam.y_res;
    f.lights.n = nlights;
    for (int l = 0; l < nlights; ++l) {
        f.lights.pos_d[l] = make_float4(lights[l].pos[0], lights[l].pos[1], lights[l].pos[2], lights[l].d);
        f.lights.rgb[l] =
            make_float4(lights[l].rgb[0] / 255.0f, lights[l].rgb[1] / 255.0f, lights[l].rgb[2] / 255.0f, 0.0f);
    }
    return RT_OK;
}

int64_t chunk_bytes(const FramePlan &plan, int64_t ntiles) {
    Workspace probe;
    return static_cast<int64_t>(
        workspace_layout(probe, reinterpret_cast<char *>(0), ntiles * 64 * plan.spp, std::max(1, plan.nlights)));
}

int64_t tiles_within(const FramePlan &plan, int64_t budget) {
    int64_t lo = 0, hi = plan.tiles;  // largest k in [0, tiles] with chunk_bytes(k) <= budget
    if (chunk_bytes(plan, 1) > budget) return 0;
    lo = 1;
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo + 1) / 2;
        if (chunk_bytes(plan, mid) <= budget)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

static dim3 grid256(int64_t n) { return dim3(static_cast<unsigned>((n + 255) / 256)); }

int32_t trace_chunk(const FramePlan &f, Workspace *w, int64_t tile0, int64_t ntiles, bool timed) {
    rt_accel *a = f.a;
    rt_ctx *ctx = a->scene->ctx;
    hipStream_t s = ctx->stream;
    const int spp = f.spp;
    const int nlights = f.nlights;
    const int64_t ns = ntiles * 64 * spp;  // samples of the chunk, tile padding included
    if (ns >= (int64_t(1) << 31)) return fail(ctx, RT_ERR_INVALID, "too many samples per chunk");
    const int nblocks = static_cast<int>((ns + 255) / 256);
    if (nblocks > 1024 * kScanPer) return fail(ctx, RT_ERR_INVALID, "chunk too large for the hit compaction");
    hipEvent_t *ev = ctx->ev;
    auto rec = [&](int k) -> int32_t {
        if (timed) RT_HIP_CHECK(ctx, hipEventRecord(ev[k], s));
        return RT_OK;
    };
    int32_t rc;
    const int64_t slot0 = tile0 * 64;

    if ((rc = rec(0))) return rc;
    hipLaunchKernelGGL(k_raygen, grid256(ns), dim3(256), 0, s, f.cam, f.tm, spp, f.p.seed, slot0, ns, w->rays1);
    if ((rc = rec(1))) return rc;
    ctx->count_slot = 0;
    if ((rc = launch_closest(a, w->rays1, ns, w->hits1, s, 0))) return rc;
    if ((rc = rec(2))) return rc;
    hipLaunchKernelGGL(k_count_hits, dim3(nblocks), dim3(256), 0, s, w->hits1, nullptr, ns, w->blocks1);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, s, w->blocks1, nblocks);
    ShadeArgs sa{};
    sa.tm = f.tm;
    sa.rays = w->rays1;
    sa.hits = w->hits1;
    sa.count = nullptr;
    sa.n_static = ns;
    sa.tri_n = a->scene->d_n;
    sa.tri_c = a->scene->d_c;
    sa.lights = f.lights;
    sa.bounce = f.p.bounce;
    sa.spp = spp;
    sa.seed = f.p.seed;
    sa.slot0 = slot0;
    sa.sample_of = nullptr;
    sa.qidx = w->qidx1;
    sa.nshade = w->nshade1;
    sa.hit_sample = w->hit_sample1;
    sa.shadow_q = w->shadow1;
    sa.bounce_q = w->bounce;
    sa.block_base = w->blocks1;
    hipLaunchKernelGGL(k_shade_hit<0>, grid256(ns), dim3(256), 0, s, sa);
    RT_HIP_CHECK(ctx, hipGetLastError());
    if ((rc = rec(3))) return rc;
    // With a bounce, the shadow any-hit launch runs on the side stream beside
    // the bounce closest-hit launch: both read only the primary shading's
    // queues, and the bounce launch's last waves leave most of the chip idle
    // while the shadow rays fill it. The main stream waits for the side
    // stream before anything after this function (the join below is enqueued
    // on every return path, so no later call can touch the shadow buffers
    // while they are in use).
    struct Join {
        rt_ctx *ctx = nullptr;
        ~Join() {
            if (ctx) (void)hipStreamWaitEvent(ctx->stream, ctx->join, 0);
        }
    } join;
    const bool overlap = f.p.bounce && nlights > 0;
    hipStream_t ss = s;
    if (overlap) {
        RT_HIP_CHECK(ctx, hipEventRecord(ctx->fork, s));
        RT_HIP_CHECK(ctx, hipStreamWaitEvent(ctx->side, ctx->fork, 0));
        ss = ctx->side;
    }
    if (timed) RT_HIP_CHECK(ctx, hipEventRecord(ev[10], ss));
    ctx->count_slot = 1;
    if (nlights > 0 &&
        (rc = launch_occluded_q(a, w->shadow1, w->blocks1 + nblocks, nlights, ns * nlights, w->occ1, ss, 1))) {
        if (overlap) {
            (void)hipEventRecord(ctx->join, ss);
            join.ctx = ctx;
        }
        return rc;
    }
    if (timed) RT_HIP_CHECK(ctx, hipEventRecord(ev[11], ss));
    if (overlap) {
        RT_HIP_CHECK(ctx, hipEventRecord(ctx->join, ss));
        join.ctx = ctx;
    }
    if ((rc = rec(4))) return rc;
    if (f.p.bounce) {
        const uint32_t *order = nullptr;
        if (f.p.flags & RT_REGROUP_BOUNCE) {
            const float4 lo = make_float4(a->scene_lo[0], a->scene_lo[1], a->scene_lo[2], 0.0f);
            const float4 inv = make_float4(a->scene_inv[0], a->scene_inv[1], a->scene_inv[2], 0.0f);
            hipLaunchKernelGGL(k_bounce_keys, grid256(ns), dim3(256), 0, s, w->bounce, w->blocks1 + nblocks, ns, lo,
                               inv, w->skeys, w->svals);
            RT_HIP_CHECK(ctx, radix_sort_pairs(w->skeys, w->svals, w->skeys_t, w->svals_t, ns, 30, w->shist,
                                               w->ssums, s));
            order = w->svals;
        }
        ctx->count_slot = 2;
        if ((rc = launch_closest_q(a, w->bounce, w->blocks1 + nblocks, 1, ns, w->hits2, s, 2, order))) return rc;
        if ((rc = rec(5))) return rc;
        hipLaunchKernelGGL(k_count_hits, dim3(nblocks), dim3(256), 0, s, w->hits2, w->blocks1 + nblocks, ns,
                           w->blocks2);
        hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, s, w->blocks2, nblocks);
        ShadeArgs sb{};
        sb.tm = f.tm;
        sb.rays = w->bounce;
        sb.hits = w->hits2;
        sb.count = w->blocks1 + nblocks;
        sb.n_static = ns;
        sb.tri_n = a->scene->d_n;
        sb.tri_c = a->scene->d_c;
        sb.lights = f.lights;
        sb.bounce = 0;
        sb.spp = spp;
        sb.seed = f.p.seed;
        sb.slot0 = slot0;
        sb.sample_of = w->hit_sample1;
        sb.qidx = w->qidx2;
        sb.nshade = w->nshade2;
        sb.hit_sample = w->hit_sample2;
        sb.shadow_q = w->shadow2;
        sb.bounce_q = nullptr;
        sb.block_base = w->blocks2;
        hipLaunchKernelGGL(k_shade_hit<1>, grid256(ns), dim3(256), 0, s, sb);
        RT_HIP_CHECK(ctx, hipGetLastError());
        if ((rc = rec(6))) return rc;
        ctx->count_slot = 3;
        if (nlights > 0 &&
            (rc = launch_occluded_q(a, w->shadow2, w->blocks2 + nblocks, nlights, ns * nlights, w->occ2, s, 3)))
            return rc;
        if ((rc = rec(7))) return rc;
    } else {
        if ((rc = rec(5))) return rc;
        if ((rc = rec(6))) return rc;
        if ((rc = rec(7))) return rc;
    }
    return RT_OK;
}

int32_t collect_chunk(const FramePlan &f, Workspace *w, int64_t ntiles, ChunkStats *stats) {
    rt_ctx *ctx = f.a->scene->ctx;
    hipStream_t s = ctx->stream;
    const int nblocks = static_cast<int>((ntiles * 64 * f.spp + 255) / 256);
    uint32_t counts[2] = {0, 0};
    RT_HIP_CHECK(ctx, hipMemcpyAsync(counts, w->blocks1 + nblocks, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    if (f.p.bounce)
        RT_HIP_CHECK(ctx, hipMemcpyAsync(counts + 1, w->blocks2 + nblocks, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    RT_HIP_CHECK(ctx, hipStreamSynchronize(s));
    stats->hits1 = counts[0];
    stats->hits2 = counts[1];
    for (int k = 0; k < 7; ++k) RT_HIP_CHECK(ctx, hipEventElapsedTime(&stats->ms[k], ctx->ev[k], ctx->ev[k + 1]));
    // The shadow any-hit launch has its own events (it may run on the side
    // stream); [3] stays its time, [4] the bounce launch's alone.
    RT_HIP_CHECK(ctx, hipEventElapsedTime(&stats->ms[3], ctx->ev[10], ctx->ev[11]));
    return RT_OK;
}

int32_t compose_chunk(const FramePlan &f, Workspace *w, int64_t tile0, int64_t ntiles, float *rgb) {
    rt_ctx *ctx = f.a->scene->ctx;
    ComposeArgs ca{};
    ca.tm = f.tm;
    ca.slot0 = tile0 * 64;
    ca.nslots = ntiles * 64;
    ca.spp = f.spp;
    ca.bounce = f.p.bounce;
    ca.bg_r = f.p.background[0];
    ca.bg_g = f.p.background[1];
    ca.bg_b = f.p.background[2];
    ca.hits1 = w->hits1;
    ca.qidx1 = w->qidx1;
    ca.nshade1 = w->nshade1;
    ca.shadow1 = w->shadow1;
    ca.occ1 = w->occ1;
    ca.qidx2 = w->qidx2;
    ca.hits2 = w->hits2;
    ca.nshade2 = w->nshade2;
    ca.shadow2 = w->shadow2;
    ca.occ2 = w->occ2;
    ca.kd = f.a->scene->d_kd;
    ca.lights = f.lights;
    ca.rgb = rgb;
    hipLaunchKernelGGL(k_compose, grid256(ca.nslots), dim3(256), 0, ctx->stream, ca);
    RT_HIP_CHECK(ctx, hipGetLastError());
    return RT_OK;
}

Workspace *accel_workspace(rt_accel *a) {
    if (!a->frame_ws) a->frame_ws = new Workspace();
    return reinterpret_cast<Workspace *>(a->frame_ws);
}

int32_t render_frame(rt_accel *a, const rt_camera &cam, const rt_light *lights, int nlights,
                     const rt_render_params &p, float *rgb, rt_hit *primary, rt_hit *bounce_hits,
                     rt_stats *stats, void *fixed_ws, size_t fixed_bytes) {
    rt_ctx *ctx = a->scene->ctx;
    hipStream_t s = ctx->stream;
    FramePlan f;
    int32_t rc = plan_frame(a, cam, lights, nlights, p, &f);
    if (rc) return rc;
    if (bounce_hits && !p.bounce) return fail(ctx, RT_ERR_INVALID, "bounce hits requested without bounce");
    // A frame larger than one compaction scan (or than a fixed workspace)
    // runs as consecutive chunks of tiles through one chunk-sized workspace;
    // sample order and arithmetic do not depend on the chunking.
    int64_t chunk = std::min(f.tiles, scan_chunk_tiles(f));
    Workspace fixed;
    Workspace *w = &fixed;
    if (fixed_ws) {
        const int64_t fit = tiles_within(f, static_cast<int64_t>(fixed_bytes));
        if (fit <= 0)
            return fail(ctx, RT_ERR_INVALID,
                        "rt_render_static: one 8x8 tile at " + std::to_string(f.spp) + " samples per pixel and " +
                            std::to_string(nlights) + " lights needs " + std::to_string(chunk_bytes(f, 1)) +
                            " bytes of workspace; the static scratch holds " + std::to_string(fixed_bytes));
        chunk = std::min(chunk, fit);
        workspace_layout(fixed, reinterpret_cast<char *>(fixed_ws), chunk * 64 * f.spp, std::max(1, nlights));
    } else {
        w = accel_workspace(a);
        if ((rc = workspace_reserve(ctx, w, chunk * 64 * f.spp, std::max(1, nlights)))) return rc;
    }

    // Frame-sized outputs: the caller's device buffers, else staging kept on
    // the workspace (grow-only).
    const bool dev_out = (p.flags & RT_DEVICE_PTRS) != 0;
    const bool async = (p.flags & RT_ASYNC) != 0;  // checked by the entry points: device outputs, no stats
    const int64_t npix = f.npix;
    const int spp = f.spp;
    const int64_t nout = npix * spp;
    float *img = (dev_out && rgb) ? rgb : nullptr;
    rt_hit *hp = (dev_out && primary) ? primary : nullptr;
    rt_hit *hb = (dev_out && bounce_hits) ? bounce_hits : nullptr;
    if (fixed_ws && !(dev_out && img))
        return fail(ctx, RT_ERR_INVALID, "rt_render_static: rgb and the hit outputs must be device buffers");
    if (!fixed_ws) {
        const size_t img_bytes = img ? 0 : ((3 * npix * sizeof(float) + 255) / 256) * 256;
        const size_t hits_bytes = static_cast<size_t>(nout) * sizeof(rt_hit);
        const size_t need = img_bytes + ((primary && !hp) ? hits_bytes : 0) + ((bounce_hits && !hb) ? hits_bytes : 0);
        if (need > w->frame_bytes) {
            if (w->frame_mem) RT_HIP_CHECK(ctx, hipStreamSynchronize(s));
            if (w->frame_mem) (void)hipFree(w->frame_mem);
            w->frame_mem = nullptr;
            w->frame_bytes = 0;
            hipError_t e = hipMalloc(&w->frame_mem, need);
            if (e != hipSuccess)
                return fail(ctx, e == hipErrorOutOfMemory ? RT_ERR_OOM : RT_ERR_HIP,
                            std::string("frame outputs: ") + hipGetErrorString(e));
            w->frame_bytes = need;
        }
        char *q = reinterpret_cast<char *>(w->frame_mem);
        if (!img) {
            img = reinterpret_cast<float *>(q);
            q += img_bytes;
        }
        if (primary && !hp) {
            hp = reinterpret_cast<rt_hit *>(q);
            q += hits_bytes;
        }
        if (bounce_hits && !hb) hb = reinterpret_cast<rt_hit *>(q);
    }

    uint64_t hits1 = 0, hits2 = 0;
    float ms[6] = {0, 0, 0, 0, 0, 0};
    for (int64_t t0 = 0; t0 < f.tiles; t0 += chunk) {
        const int64_t nt = std::min(chunk, f.tiles - t0);
        if ((rc = trace_chunk(f, w, t0, nt, stats != nullptr))) return rc;
        if (async) {
            const int nb = static_cast<int>((nt * 64 * spp + 255) / 256);
            hipLaunchKernelGGL(k_add_frame_rays, dim3(1), dim3(64), 0, s, w->blocks1 + nb,
                               p.bounce ? w->blocks2 + nb : nullptr, t0 == 0 ? static_cast<uint64_t>(nout) : 0ull,
                               nlights, ctx->d_frame_rays);
        }
        if (stats) RT_HIP_CHECK(ctx, hipEventRecord(ctx->ev[8], s));
        if ((rc = compose_chunk(f, w, t0, nt, img))) return rc;
        if (stats) RT_HIP_CHECK(ctx, hipEventRecord(ctx->ev[9], s));
        if (hp)
            hipLaunchKernelGGL(k_hits_out, grid256(nt * 64 * spp), dim3(256), 0, s, f.tm, spp, t0 * 64, nt * 64,
                               nullptr, w->hits1, hp);
        if (hb)
            hipLaunchKernelGGL(k_hits_out, grid256(nt * 64 * spp), dim3(256), 0, s, f.tm, spp, t0 * 64, nt * 64,
                               w->qidx1, w->hits2, hb);
        RT_HIP_CHECK(ctx, hipGetLastError());
        if (stats) {
            ChunkStats cs;
            if ((rc = collect_chunk(f, w, nt, &cs))) return rc;  // synchronises
            float compose_ms = 0.0f, frame_ms = 0.0f;
            RT_HIP_CHECK(ctx, hipEventElapsedTime(&compose_ms, ctx->ev[8], ctx->ev[9]));
            RT_HIP_CHECK(ctx, hipEventElapsedTime(&frame_ms, ctx->ev[0], ctx->ev[9]));
            hits1 += cs.hits1;
            hits2 += cs.hits2;
            // Trace stages: primary, shadow, bounce, bounce shadow; the rest
            // is ray generation, shading and compose. The frame time runs
            // from ray generation to the end of compose: with a bounce the
            // shadow launch overlaps the bounce launch, so the stages can sum
            // to more than it.
            ms[1] += cs.ms[1];
            ms[2] += cs.ms[3];
            ms[3] += cs.ms[4];
            ms[4] += cs.ms[6];
            ms[5] += cs.ms[0] + cs.ms[2] + cs.ms[5] + compose_ms;
            ms[0] += frame_ms;
        }
    }

    // Outputs.
    if (rgb && !dev_out)
        RT_HIP_CHECK(ctx, hipMemcpyAsync(rgb, img, 3 * npix * sizeof(float), hipMemcpyDeviceToHost, s));
    if (primary && !dev_out)
        RT_HIP_CHECK(ctx, hipMemcpyAsync(primary, hp, nout * sizeof(rt_hit), hipMemcpyDeviceToHost, s));
    if (bounce_hits && !dev_out)
        RT_HIP_CHECK(ctx, hipMemcpyAsync(bounce_hits, hb, nout * sizeof(rt_hit), hipMemcpyDeviceToHost, s));
    if (!async) RT_HIP_CHECK(ctx, hipStreamSynchronize(s));
    if (stats) {
        stats->rays[0] = static_cast<uint64_t>(nout);
        stats->rays[1] = hits1 * nlights;
        stats->rays[2] = p.bounce ? hits1 : 0;
        stats->rays[3] = p.bounce ? hits2 * nlights : 0;
        for (int k = 0; k < 6; ++k) stats->ms[k] = ms[k];
    }
    return RT_OK;
}

void free_frame_workspace(rt_accel *a) {
    workspace_release(reinterpret_cast<Workspace *>(a->frame_ws));
    a->frame_ws = nullptr;
}

}  // namespace rt
