// Oracle — TEST INFRASTRUCTURE ONLY (see vec.hpp for the rules).
//
// One frame of the scalar (IEEE) tracer: rt/raytracer.cc:20-87 for a hit,
// rt/triangle.h:464-498 for the hit normal, rt/ray.cc:91-104 and
// rt/light.h:62-100 for the shadow ray, rt/materials/phong_shader.cc:21-64
// with rt/light.h:44-59 for the shading; primary rays per
// rt/camera.h:127-158. The diffuse bounce is this project's definition
// (sampling.hpp). Pixels are independent, so the frame is split across
// threads with a per-thread traversal stack (the reference stack at
// ssd/bvh.h:84 is a member and is not re-entrant).
#include <atomic>
#include <thread>
#include <vector>

#include "oracle.hpp"
#include "rt_oracle.h"
#include "sampling.hpp"

namespace oracle {

struct Colour {
    float r = 0.0f, g = 0.0f, b = 0.0f;
};

struct PointLight {
    V3 c;
    Colour rgb;  // already divided by 255 (rt/light.h:17-18)
    float d;
};

struct CameraState {
    V3 pos, X, Y, Z;
    float x_m, y_m, x_inc, y_inc, t;
    int x_res, y_res;
};

CameraState make_camera(const oracle_camera &c) {
    CameraState s;
    s.pos = V3(c.pos[0], c.pos[1], c.pos[2]);
    s.X = V3(c.x_axis[0], c.x_axis[1], c.x_axis[2]);
    s.Y = V3(c.y_axis[0], c.y_axis[1], c.y_axis[2]);
    s.Z = V3(c.z_axis[0], c.z_axis[1], c.z_axis[2]);
    // rt/camera.h:54-62 (x_a_res = y_a_res = 1)
    s.x_m = -c.w;
    s.y_m = -c.h;
    s.x_inc = (c.w * 2.0f) / static_cast<float>(c.x_res);
    s.y_inc = (c.h * 2.0f) / static_cast<float>(c.y_res);
    s.t = c.t;
    s.x_res = c.x_res;
    s.y_res = c.y_res;
    return s;
}

// rt/camera.h:127-145: screen = X*x_t + Y*y_t + Z*t, normalised exactly.
Ray primary_ray(const CameraState &c, float xf, float yf) {
    const float x_t = (xf * c.x_inc) + c.x_m;
    const float y_t = (yf * c.y_inc) + c.y_m;
    const V3 screen = ((c.X * x_t) + (c.Y * y_t)) + (c.Z * c.t);
    Ray r;
    r.ogn = c.pos;
    r.dir = normalise_exact(screen);
    return r;
}

namespace {

struct SurfaceHit {
    V3 dst, geo_norm, shade_norm;
};

// rt/ray.cc:15-20 then rt/triangle.h:464-498: move to the hit, re-project onto
// the triangle plane, pick the normal facing the ray.
SurfaceHit surface_at(const Triangle &tri, Ray *r, float d) {
    r->dst = r->ogn + (r->dir * d);
    r->length = d;
    const float denom = dot(tri.n, r->dir);
    const float num = dot(tri.n, (tri.c - r->dst));
    const float dl = num / denom;
    r->dst = r->dst + (r->dir * dl);
    r->length = r->length + dl;
    SurfaceHit s;
    s.dst = r->dst;
    if (denom > 0.0f) {
        s.shade_norm = -tri.n;
        s.geo_norm = -tri.n;
    } else {
        s.shade_norm = tri.n;
        s.geo_norm = tri.n;
    }
    return s;
}

// rt/ray.cc:91-104 + rt/light.h:96-100 + rt/ray.h:44-54: offset origin, then
// direction and length to the light centre.
Ray shadow_ray(const SurfaceHit &s, const V3 &light_c) {
    Ray sr;
    sr.ogn = offset_start_point(s.dst, s.geo_norm, 1);
    sr.dst = light_c;
    sr.dir = sr.dst - sr.ogn;
    sr.length = magnitude(sr.dir);
    sr.dir = sr.dir / sr.length;
    return sr;
}

// rt/materials/phong_shader.cc:21-64 for phong_shader(255*albedo, kd = 1):
// ks = ka = 0, so only the diffuse term and the zero ambient remain.
constexpr int kMaxLights = 8;

Colour shade_direct(const Colour &kd, const SurfaceHit &s, const Ray *illum, const std::vector<PointLight> &lights) {
    Colour c;
    for (size_t l = 0; l < lights.size(); ++l) {
        const Ray &ir = illum[l];
        const float shade = dot(ir.dir, s.shade_norm);
        if (shade < 0.0f) continue;
        // Specular: ray_dot_reflection > 0 adds pow(.)*ks with ks = 0, a no-op.
        const Colour sc{shade * kd.r, shade * kd.g, shade * kd.b};
        // rt/light.h:44-59 (no spot): rgb / ((len * d)^2 + 1).
        const float dist = ir.length * lights[l].d;
        const float d_scale = (dist * dist) + 1.0f;
        const Colour li{lights[l].rgb.r / d_scale, lights[l].rgb.g / d_scale, lights[l].rgb.b / d_scale};
        c.r += li.r * (sc.r * ir.magn);
        c.g += li.g * (sc.g * ir.magn);
        c.b += li.b * (sc.b * ir.magn);
    }
    // + ka (zero)
    c.r += 0.0f;
    c.g += 0.0f;
    c.b += 0.0f;
    return c;
}

struct FrameCtx {
    const Scene *s;
    const float *albedo;
    CameraState cam;
    std::vector<PointLight> lights;
    oracle_render_params p;
    oracle_render_out out;
    int64_t ns = 0;  // samples per frame (records are level-major)
};

Colour kd_of(const FrameCtx &f, int prim) {
    // phong_shader(rgb = 255 * albedo, kd = 1.0): kd = rgb * 1.0f.
    const float *a = f.albedo + 3 * static_cast<int64_t>(prim);
    return Colour{(a[0] * 255.0f) * 1.0f, (a[1] * 255.0f) * 1.0f, (a[2] * 255.0f) * 1.0f};
}

struct LocalCounters {
    uint64_t rays[4] = {0, 0, 0, 0};
    Counters trav;
};

// Direct light at a surface hit: one shadow ray per light, visibility as the
// fraction of unoccluded shadow rays (rt/raytracer.cc:44-64).
Colour direct_light(const FrameCtx &f, const SurfaceHit &sh, const Colour &kd, uint8_t *vis_out, int ray_kind,
                    LocalCounters *lc, uint8_t *vis_rec) {
    Ray illum[kMaxLights];
    for (size_t l = 0; l < f.lights.size(); ++l) {
        illum[l] = shadow_ray(sh, f.lights[l].c);
        const bool occluded = found_nearer_object(*f.s, illum[l], illum[l].length, &lc->trav);
        ++lc->rays[ray_kind];
        const float made_it = occluded ? 0.0f : 1.0f;
        illum[l].magn = made_it / 1.0f;
        if (vis_out) vis_out[l] = occluded ? 0 : 1;
        if (vis_rec) vis_rec[l] = occluded ? 0 : 1;
    }
    return shade_direct(kd, sh, illum, f.lights);
}

// Path record of one surface of one sample.
void record_surface(int32_t *prim_out, float *o_out, float *n_out, uint8_t *vis_out, int64_t k, int nl, int prim,
                    const SurfaceHit *sh) {
    if (prim_out) prim_out[k] = prim;
    const V3 o = sh ? offset_start_point(sh->dst, sh->geo_norm, 1) : V3(0.0f, 0.0f, 0.0f);
    const V3 n = sh ? sh->shade_norm : V3(0.0f, 0.0f, 0.0f);
    if (o_out) {
        o_out[3 * k + 0] = o.x;
        o_out[3 * k + 1] = o.y;
        o_out[3 * k + 2] = o.z;
    }
    if (n_out) {
        n_out[3 * k + 0] = n.x;
        n_out[3 * k + 1] = n.y;
        n_out[3 * k + 2] = n.z;
    }
    if (vis_out && !sh)
        for (int l = 0; l < nl; ++l) vis_out[nl * k + l] = 0;
}

void render_pixel(const FrameCtx &f, int px, int py, LocalCounters *lc) {
    const int W = f.cam.x_res;
    const int64_t p = static_cast<int64_t>(py) * W + px;
    const int nl = static_cast<int>(f.lights.size());
    const int spp = std::max(1, f.p.spp);
    const oracle_render_out &o = f.out;
    Colour acc;
    for (int s = 0; s < spp; ++s) {
        const int64_t k = p * spp + s;
        float xf = static_cast<float>(px), yf = static_cast<float>(py);
        if (spp > 1) {
            float jx, jy;
            pixel_jitter(f.p.seed, static_cast<uint32_t>(p), static_cast<uint32_t>(s), &jx, &jy);
            xf = xf + jx;
            yf = yf + jy;
        }
        Ray r = primary_ray(f.cam, xf, yf);
        Hit h;
        const int prim = find_nearest_object(*f.s, r, &h, &lc->trav);
        ++lc->rays[0];
        const bool first = (s == 0);
        if (first && f.out.prim) f.out.prim[p] = prim;
        if (first && f.out.tuv) {
            f.out.tuv[3 * p + 0] = h.d;
            f.out.tuv[3 * p + 1] = h.u;
            f.out.tuv[3 * p + 2] = h.v;
        }
        if (first && f.out.bounce_prim) f.out.bounce_prim[p] = -1;
        Colour c;  // background (0, 0, 0)
        const int levels = f.p.bounce + 1;
        for (int lv = 0; lv < o.rec_levels; ++lv)
            record_surface(o.rec_prim, o.rec_o, o.rec_n, o.rec_vis, static_cast<int64_t>(lv) * f.ns + k, nl, -1,
                           nullptr);
        if (h.d < kMaxDist) {
            const Triangle &tri = f.s->tris[prim];
            SurfaceHit sh = surface_at(tri, &r, h.d);
            const Colour kd1 = kd_of(f, prim);
            if (o.rec_levels > 0) record_surface(o.rec_prim, o.rec_o, o.rec_n, nullptr, k, nl, prim, &sh);
            c = direct_light(f, sh, kd1, (first && f.out.vis) ? f.out.vis + nl * p : nullptr, 1, lc,
                             o.rec_levels > 0 && o.rec_vis ? o.rec_vis + nl * k : nullptr);
            // Diffuse bounces: c += T (.) direct(x_k) with the throughput
            // T = prod over earlier vertices of kd / 255 (the Lambert
            // cosine / pdf ratio cancels); one bounce gives
            // c + (kd1 / 255) * direct(x2).
            Colour T{kd1.r / 255.0f, kd1.g / 255.0f, kd1.b / 255.0f};
            for (int bn = 1; bn < levels; ++bn) {
                Ray br;
                br.ogn = offset_start_point(sh.dst, sh.geo_norm, 1);
                br.dir = cosine_direction(bounce_key(f.p.seed, static_cast<uint32_t>(p), static_cast<uint32_t>(s),
                                                     static_cast<uint32_t>(bn)),
                                          sh.shade_norm);
                Hit bh;
                const int bprim = find_nearest_object(*f.s, br, &bh, &lc->trav);
                ++lc->rays[2];
                if (bn == 1 && first && f.out.bounce_prim) f.out.bounce_prim[p] = bprim;
                if (bn == 1 && first && f.out.bounce_tuv) {
                    f.out.bounce_tuv[3 * p + 0] = bh.d;
                    f.out.bounce_tuv[3 * p + 1] = bh.u;
                    f.out.bounce_tuv[3 * p + 2] = bh.v;
                }
                if (!(bh.d < kMaxDist)) break;
                const Triangle &btri = f.s->tris[bprim];
                sh = surface_at(btri, &br, bh.d);
                const Colour kdb = kd_of(f, bprim);
                const int64_t rk = static_cast<int64_t>(bn) * f.ns + k;
                if (bn < o.rec_levels) record_surface(o.rec_prim, o.rec_o, o.rec_n, nullptr, rk, nl, bprim, &sh);
                const Colour cb = direct_light(f, sh, kdb,
                                               (bn == 1 && first && f.out.bounce_vis) ? f.out.bounce_vis + nl * p
                                                                                      : nullptr,
                                               3, lc, (bn < o.rec_levels && o.rec_vis) ? o.rec_vis + nl * rk : nullptr);
                c.r = c.r + (T.r * cb.r);
                c.g = c.g + (T.g * cb.g);
                c.b = c.b + (T.b * cb.b);
                T = Colour{T.r * (kdb.r / 255.0f), T.g * (kdb.g / 255.0f), T.b * (kdb.b / 255.0f)};
            }
        }
        acc.r += c.r;
        acc.g += c.g;
        acc.b += c.b;
    }
    if (f.out.rgb) {
        const float inv = static_cast<float>(spp);
        f.out.rgb[3 * p + 0] = acc.r / inv;
        f.out.rgb[3 * p + 1] = acc.g / inv;
        f.out.rgb[3 * p + 2] = acc.b / inv;
    }
}

}  // namespace

int render_frame(const Scene &s, const float *albedo, const oracle_camera &cam, const oracle_light *lights,
                 int nlights, const oracle_render_params &p, const oracle_render_out &out) {
    FrameCtx f;
    f.s = &s;
    f.albedo = albedo;
    f.cam = make_camera(cam);
    for (int l = 0; l < nlights; ++l) {
        PointLight pl;
        pl.c = V3(lights[l].pos[0], lights[l].pos[1], lights[l].pos[2]);
        pl.rgb = Colour{lights[l].rgb[0] / 255.0f, lights[l].rgb[1] / 255.0f, lights[l].rgb[2] / 255.0f};
        pl.d = lights[l].d;
        f.lights.push_back(pl);
    }
    f.p = p;
    f.out = out;
    f.ns = static_cast<int64_t>(cam.x_res) * cam.y_res * std::max(1, p.spp);
    const int64_t npix = static_cast<int64_t>(f.cam.x_res) * f.cam.y_res;
    const int stride = std::max(1, p.pixel_stride);
    const int nthreads = std::max(1, p.nthreads);
    std::atomic<int64_t> next{0};
    std::vector<LocalCounters> counters(nthreads);
    auto worker = [&](int tid) {
        const int64_t chunk = 64;
        while (true) {
            const int64_t b = next.fetch_add(chunk);
            if (b * stride >= npix) break;
            for (int64_t k = b; k < b + chunk; ++k) {
                const int64_t pix = k * stride;
                if (pix >= npix) break;
                render_pixel(f, static_cast<int>(pix % f.cam.x_res), static_cast<int>(pix / f.cam.x_res),
                             &counters[tid]);
            }
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nthreads; ++t) pool.emplace_back(worker, t);
    worker(0);
    for (auto &th : pool) th.join();
    if (out.counters) {
        for (int k = 0; k < 6; ++k) out.counters[k] = 0;
        for (auto &c : counters) {
            for (int k = 0; k < 4; ++k) out.counters[k] += c.rays[k];
            out.counters[4] += c.trav.box;
            out.counters[5] += c.trav.tri;
        }
    }
    return 0;
}

}  // namespace oracle
