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
#include <cfloat>
#include <cmath>
#include <thread>
#include <vector>

#include "oracle.hpp"
#include "packet.hpp"
#include "rt_oracle.h"
#include "sampling.hpp"
#include "texture.hpp"

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
        Colour c{f.p.background[0], f.p.background[1], f.p.background[2]};  // a miss: camera::shade
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


// ---------------------------------------------------------------------------
// Phong materials with mirror and refraction rays: rt/raytracer.cc:20-87 (one
// hit: normal, shadow rays, shade, secondary rays), phong_shader::shade
// (rt/materials/phong_shader.cc:21-111) and ray::reflect / ray::refract
// without DIFFUSE_REFLECTIONS (rt/ray.cc:118-258): one reflected and one
// refracted ray while the ray's power stays above MIN_REFLECTIVE_POWER.

// common/common.h:30-32. A double literal: the float power is compared in
// double, so a power of exactly 0.1f (above 0.1) still recurses.
constexpr double kMinReflectivePower = 0.1;

constexpr int kMaterialFloats = 40;
constexpr int kCookTorrance = 1;
constexpr int kColoured = 2;
constexpr int kMapper = 3;
constexpr int kMandelbrot = 4;

struct Phong {
    Colour ka, kd, ks;
    float s = 0.0f, tran = 0.0f, ri = 1.0f, rf = 0.0f;
    int kind = 0;  // 0 phong_shader, 1 cook_torrance_cxy, 2 coloured_mapper_shader, 3 mapper_shader
    float cx = 0.33f, cy = 0.33f, rd = 0.5f, sr = 0.0f, rii = 0.0f;
    int tex[6] = {0, 0, 0, 0, 0, 0};  // 1 + map: t_ka, t_kd, t_ks, t_ns, t_refl, t_d
    float mkd = 0.0f, mks = 0.0f, clrh = 0.0f, clrf = 0.0f;
    int first[5] = {0, 0, 0, 0, 0}, count[5] = {0, 0, 0, 0, 0};  // ctex, dtex, stex, rtex, ttex
    int iters = 0;  // mandelbrot_shader's iteration limit
};

// XYZ -> RGB for the primaries and white point of rt/materials/material.h:17-45,
// evaluated in float in the order written there.
struct Cie {
    float m[3][3];
};

Cie xyz_to_rgb_matrix() {
    const float xr = 0.640f, yr = 0.330f, xg = 0.290f, yg = 0.600f, xb = 0.150f, yb = 0.060f;
    const float xw = 0.3333f, yw = 0.3333f;
    const float crd = ((1.0f / yw) * (xw * (yg - yb) - yw * (xg - xb) + xg * yb - xb * yg));
    const float cgd = ((1.0f / yw) * (xw * (yb - yr) - yw * (xb - xr) - xr * yb + xb * yr));
    const float cbd = ((1.0f / yw) * (xw * (yr - yg) - yw * (xr - xg) + xr * yg - xg * yr));
    Cie c;
    c.m[0][0] = (yg - yb - xb * yg + yb * xg) / crd;
    c.m[0][1] = (xb - xg - xb * yg + xg * yb) / crd;
    c.m[0][2] = (xg * yb - xb * yg) / crd;
    c.m[1][0] = (yb - yr - yb * xr + yr * xb) / cgd;
    c.m[1][1] = (xr - xb - xr * yb + xb * yr) / cgd;
    c.m[1][2] = (xb * yr - xr * yb) / cgd;
    c.m[2][0] = (yr - yg - yr * xg + yg * xr) / cbd;
    c.m[2][1] = (xg - xr - xg * yr + xr * yg) / cbd;
    c.m[2][2] = (xr * yg - xg * yr) / cbd;
    return c;
}

const Cie kCie = xyz_to_rgb_matrix();

// cxy_to_rgb (rt/materials/material.h:326-343): chroma (x, y) at luminance
// lum, clamped to [0, 1] and scaled to 0-255.
Colour cxy_to_rgb(float x, float y, float lum) {
    float cie[3];
    cie[0] = lum * x / y;
    cie[1] = lum;
    cie[2] = lum * (1.0f / y - 1.0f) - cie[0];
    float out[3];
    for (int k = 0; k < 3; ++k)
        out[k] = (kCie.m[k][0] * cie[0]) + (kCie.m[k][1] * cie[1]) + (kCie.m[k][2] * cie[2]);
    return Colour{std::max(0.0f, std::min(1.0f, out[0])) * 255.0f, std::max(0.0f, std::min(1.0f, out[1])) * 255.0f,
                  std::max(0.0f, std::min(1.0f, out[2])) * 255.0f};
}

// cook_torrance_cxy::shade for one light (rt/materials/cook_torrance_cxy.cc:33-76)
// with schlick_fresnell, geometric_attenuation_factor and
// gaussian_facet_distribution (rt/materials/material.h:146-215).
void cook_torrance_term(const Phong &m, const V3 &n, const V3 &dir, float nv, const V3 &ldir, float shade,
                        float magn, const Colour &li, Colour *c) {
    const V3 half((ldir.x + dir.x) * 0.5f, (ldir.y + dir.y) * 0.5f, (ldir.z + dir.z) * 0.5f);
    const float nh = dot(n, half);
    const float vh = dot(half, dir);
    const float nl = shade;
    const float f = (m.sr + (1.0f - m.sr) * std::pow(1.0f - nv, 5.0f));
    const float g0 = (2.0f * nh * nv) / vh;
    const float g1 = (2.0f * nh * nl) / vh;
    const float g = std::min(1.0f, std::max(0.0f, std::min(g0, g1)));
    const float a_div_m = nh / m.sr;
    const float ro = 5.0f * std::exp(-(a_div_m * a_div_m));
    const float s = m.rf * ((g * f * ro) / (nl * nv));
    const Colour rgb = cxy_to_rgb(m.cx, m.cy, (shade * (m.rd + s) * magn));
    c->r += rgb.r * li.r;
    c->g += rgb.g * li.g;
    c->b += rgb.b * li.b;
}

Phong material_of(const FrameCtx &f, int prim) {
    Phong m;
    const int32_t mi = f.p.tri_material ? f.p.tri_material[prim] : -1;
    if (f.p.materials && mi >= 0 && mi < f.p.nmaterials) {
        const float *q = f.p.materials + kMaterialFloats * static_cast<int64_t>(mi);
        m.ka = Colour{q[0], q[1], q[2]};
        m.kd = Colour{q[3], q[4], q[5]};
        m.ks = Colour{q[6], q[7], q[8]};
        m.s = q[9];
        m.tran = q[10];
        m.ri = q[11];
        m.rf = q[12];
        m.kind = static_cast<int>(q[13]);
        m.cx = q[14];
        m.cy = q[15];
        m.rd = q[16];
        m.sr = q[17];
        m.rii = q[18];
        for (int k = 0; k < 6; ++k) m.tex[k] = static_cast<int>(q[19 + k]);
        m.mkd = q[25];
        m.mks = q[26];
        m.clrh = q[27];
        m.clrf = q[28];
        for (int k = 0; k < 5; ++k) {
            m.first[k] = static_cast<int>(q[29 + k]);
            m.count[k] = static_cast<int>(q[34 + k]);
        }
        m.iters = static_cast<int>(q[39]);
    } else {
        m.kd = kd_of(f, prim);  // phong_shader(255 * albedo, 1): ks = 0 * rgb, ka = 0
    }
    return m;
}

// ray::reflect (rt/ray.cc:121-131): dir - n * 2 (dir . n).
V3 reflect_dir(const V3 &dir, const V3 &n) {
    const float rdn = 2.0f * dot(dir, n);
    return dir - n * rdn;
}

// ray::refract (rt/ray.cc:200-227): Snell through n (facing the ray), the
// index inverted for a ray entering the front face (hit_t::out_in); false on
// total internal reflection.
bool refract_dir(const V3 &dir, const V3 &n, float ri, bool leaving, V3 *out) {
    const float rdn = -dot(dir, n);
    if (!leaving) ri = 1.0f / ri;
    const float cos_t2 = 1.0f - ri * ri * (1.0f - rdn * rdn);
    if (cos_t2 < 0.0f) return false;
    const float offset = (ri * rdn - std::sqrt(cos_t2));
    *out = (dir * ri) + (n * offset);
    return true;
}

// mandelbrot_shader::shade's colour (rt/materials/mandelbrot_shader.cc:41-123)
// before the light term: the fractal coordinates of the hit point on the
// plane picked by the shading normal, the escape-time loop, one retry on
// shifted coordinates for a point that does not escape (black if it again
// does not), and the colour ramp.
Colour mandelbrot_colour(const V3 &dst, const V3 &n, int limit) {
    float x0;
    float y0;
    if ((n.x == 1.0f) || (n.x == -1.0f)) {
        x0 = std::fmod(std::fabs(dst.z), 2.4f) - 1.5f;
        y0 = std::fmod(std::fabs(dst.y) - 1.5f, 2.4f) - 1.0f;
    } else if ((n.z == 1.0f) || (n.z == -1.0f)) {
        x0 = std::fmod(std::fabs(dst.x), 2.4f) - 1.5f;
        y0 = std::fmod(std::fabs(dst.y) - 1.5f, 2.4f) - 1.0f;
    } else {
        x0 = std::fmod(std::fabs(dst.x) - 1.5f, 2.4f) - 1.0f;
        y0 = std::fmod(std::fabs(dst.z), 2.4f) - 1.5f;
    }
    float x = x0;
    float y = y0;
    bool doomed = false;
    for (;;) {
        int iteration = 0;
        while ((x * x + y * y <= (10.0f * 10.0f)) && (iteration < limit)) {
            const float xtemp = x * x - y * y + x0;
            y = 2.0f * x * y + y0;
            x = xtemp;
            ++iteration;
        }
        if (doomed && (iteration == limit)) return Colour{0.0f, 0.0f, 0.0f};
        if (iteration == limit) {
            doomed = true;
            if ((n.x == 1.0f) || (n.x == -1.0f)) {
                x0 = std::fmod(std::fabs(dst.z) - 1.0f, 2.4f) * 1.5f - 0.4f;
                y0 = std::fmod(std::fabs(dst.y) - 2.0f, 2.4f) * 1.5f - 0.7f;
            } else if ((n.z == 1.0f) || (n.z == -1.0f)) {
                x0 = std::fmod(std::fabs(dst.x) - 1.0f, 2.4f) * 1.5f - 0.4f;
                y0 = std::fmod(std::fabs(dst.y) - 2.0f, 2.4f) * 1.5f - 0.7f;
            } else {
                x0 = std::fmod(std::fabs(dst.x) - 2.0f, 2.4f) * 1.5f - 0.7f;
                y0 = std::fmod(std::fabs(dst.z) - 1.0f, 2.4f) * 1.5f - 0.4f;
            }
            x = x0;
            y = y0;
            continue;
        }
        const float colour = 767.0f - ((iteration * 767.0f) / limit);
        if (colour < 256.0f) return Colour{25.0f - colour, 255.0f, 255.0f};
        if (colour < 512.0f) return Colour{0.0f, 256.0f - (colour - 256.0f), 255.0f};
        return Colour{0.0f, 0.0f, 256.0f - (colour - 767.0f)};
    }
}

struct PhongCounters {
    uint64_t secondary = 0, secondary_shadow = 0;
};

Colour trace_phong(const FrameCtx &f, Ray &r, bool primary, LocalCounters *lc, PhongCounters *pc, int32_t *prim_out,
                   float *tuv_out, uint8_t *vis_out) {
    Hit h;
    const int prim = find_nearest_object(*f.s, r, &h, &lc->trav);
    if (primary) {
        ++lc->rays[0];
        if (prim_out) *prim_out = prim;
        if (tuv_out) {
            tuv_out[0] = h.d;
            tuv_out[1] = h.u;
            tuv_out[2] = h.v;
        }
    } else {
        ++pc->secondary;
    }
    if (!(h.d < kMaxDist)) return Colour{f.p.background[0], f.p.background[1], f.p.background[2]};
    Colour c;
    const Triangle &tri = f.s->tris[prim];
    const bool leaving = dot(tri.n, r.dir) > 0.0f;  // hit_t::in_out (rt/triangle.h:487-496)
    SurfaceHit sh = surface_at(tri, &r, h.d);
    if (f.p.vertex_normals && f.p.vertex_normals[9 * static_cast<int64_t>(prim)] != FLT_MAX) {
        // triangle::normal_at_point (rt/triangle.h:474-477): the vertex
        // normals weighted u (c), v (b), 1 - (u + v) (a), not renormalised.
        const float *q = f.p.vertex_normals + 9 * static_cast<int64_t>(prim);
        const V3 na(q[0], q[1], q[2]), nb(q[3], q[4], q[5]), nc(q[6], q[7], q[8]);
        const V3 sn = ((nc * h.u) + (nb * h.v)) + (na * (1.0f - (h.u + h.v)));
        sh.shade_norm = leaving ? -sn : sn;
    }
    const Phong m = material_of(f, prim);
    const V3 &n = sh.shade_norm;
    // The hit's texture coordinates (rt/triangle.h:72-79): u vt_c + v vt_b +
    // (1 - (u + v)) vt_a, when the triangle has them.
    const oracle_tex::Tables T{f.p.maps, f.p.images, f.p.texels, f.p.layers};
    bool has_vt = false;
    float vtu = 0.0f, vtv = 0.0f;
    if (f.p.texture_coords && f.p.texture_coords[6 * static_cast<int64_t>(prim)] != FLT_MAX) {
        const float *q = f.p.texture_coords + 6 * static_cast<int64_t>(prim);
        vtu = (h.u * q[4]) + (h.v * q[2]) + ((1.0f - (h.u + h.v)) * q[0]);
        vtv = (h.u * q[5]) + (h.v * q[3]) + ((1.0f - (h.u + h.v)) * q[1]);
        has_vt = true;
    }
    // mapper_shader's monochrome layer lists (rt/materials/mapper_shader.cc:41-66, 100-126).
    auto layered = [&](int list, float value) {
        float alpha = 1.0f;
        for (int k = 0; k < m.count[list]; ++k) {
            float z;
            alpha *= oracle_tex::texture_map_monochrome(T, f.p.layers[m.first[list] + k], &z, sh.dst, r.dir, sh.shade_norm, has_vt, vtu, vtv);
            value = z * (1.0f / 255.0f);
            if (alpha == 0.0f) break;
        }
        return value;
    };
    // generate_rays: the reflection and refraction coefficients
    // (coloured_mapper_shader.cc:8-38, mapper_shader.cc:41-82).
    float refl_c = m.rf, tran_c = m.tran;
    if (m.kind == kColoured) {
        if (m.tex[4]) {
            oracle_tex::Rgb3 tmp;
            oracle_tex::texture_map(T, m.tex[4] - 1, &tmp, sh.dst, r.dir, sh.shade_norm, has_vt, vtu, vtv);
            refl_c = tmp.r * (m.rf * (1.0f / 255.0f));
        }
        if (m.tex[5]) {
            oracle_tex::Rgb3 tmp;
            oracle_tex::texture_map(T, m.tex[5] - 1, &tmp, sh.dst, r.dir, sh.shade_norm, has_vt, vtu, vtv);
            tran_c = tmp.r * (m.tran * (1.0f / 255.0f));
        }
    } else if (m.kind == kMapper) {
        refl_c = layered(3, m.rf);
        tran_c = layered(4, m.tran);
    }
    // shade: the material's colours at the hit (coloured_mapper_shader.cc:41-80,
    // mapper_shader.cc:84-126).
    Colour ka = m.ka, kd = m.kd, ks = m.ks;
    float ns = m.s, mkd = m.mkd;
    if (m.kind == kColoured) {
        auto mapped = [&](int slot, Colour col) {
            if (!slot) return col;
            oracle_tex::Rgb3 t{col.r, col.g, col.b};
            oracle_tex::texture_map(T, slot - 1, &t, sh.dst, r.dir, sh.shade_norm, has_vt, vtu, vtv);
            return Colour{t.r * (col.r * (1.0f / 255.0f)), t.g * (col.g * (1.0f / 255.0f)),
                          t.b * (col.b * (1.0f / 255.0f))};
        };
        if (m.tex[3]) {
            oracle_tex::Rgb3 cur_ka;
            ns = oracle_tex::texture_map(T, m.tex[3] - 1, &cur_ka, sh.dst, r.dir, sh.shade_norm, has_vt, vtu, vtv);
            ns *= (m.s * (1.0f / 255.0f));
        }
        ka = mapped(m.tex[0], m.ka);
        kd = mapped(m.tex[1], m.kd);
        ks = mapped(m.tex[2], m.ks);
    } else if (m.kind == kMapper) {
        oracle_tex::Rgb3 rgb{m.kd.r, m.kd.g, m.kd.b};
        float alpha = 1.0f;
        for (int k = 0; k < m.count[0]; ++k) {
            alpha *= oracle_tex::texture_map(T, f.p.layers[m.first[0] + k], &rgb, sh.dst, r.dir, sh.shade_norm, has_vt, vtu, vtv);
            if (alpha == 0.0f) break;
        }
        kd = Colour{rgb.r, rgb.g, rgb.b};
        mkd = layered(1, m.mkd);
        ns = layered(2, m.s);
    }
    const float nv = dot(n, r.dir);  // cook_torrance_cxy: one dot product for every light
    float total_shade = 0.0f;        // mandelbrot_shader: sum of n.l x visibility
    for (size_t l = 0; l < f.lights.size(); ++l) {
        Ray illum = shadow_ray(sh, f.lights[l].c);
        const bool occluded = found_nearer_object(*f.s, illum, illum.length, &lc->trav);
        if (primary) {
            ++lc->rays[1];
        } else {
            ++pc->secondary_shadow;
        }
        illum.magn = (occluded ? 0.0f : 1.0f) / 1.0f;
        if (vis_out) vis_out[l] = occluded ? 0 : 1;
        // rt/light.h:44-59 (no spot): rgb / ((len * d)^2 + 1).
        const float dist = illum.length * f.lights[l].d;
        const float d_scale = (dist * dist) + 1.0f;
        const Colour li{f.lights[l].rgb.r / d_scale, f.lights[l].rgb.g / d_scale, f.lights[l].rgb.b / d_scale};
        if (m.kind == kMapper) {
            // mapper_shader.cc:128-170: the reflection is taken about n
            // scaled by 2 (n.l) kd, and there is no ambient term.
            const float diff_shade = dot(illum.dir, n) * mkd;
            if (diff_shade < 0.0f) continue;
            const float shade_x2 = 2.0f * diff_shade;
            const V3 ref_dir(illum.dir.x - n.x * shade_x2, illum.dir.y - n.y * shade_x2, illum.dir.z - n.z * shade_x2);
            const float ray_dot_reflection = dot(r.dir, ref_dir);
            float spec_shade = 0.0f;
            if (ray_dot_reflection > 0.0f) spec_shade = std::pow(ray_dot_reflection, ns) * m.mks;
            c.r += kd.r * diff_shade * li.r * illum.magn;
            c.g += kd.g * diff_shade * li.g * illum.magn;
            c.b += kd.b * diff_shade * li.b * illum.magn;
            c.r += ((m.clrh * kd.r) + ((1.0f - m.clrh) * li.r)) * spec_shade * illum.magn;
            c.g += ((m.clrh * kd.g) + ((1.0f - m.clrh) * li.g)) * spec_shade * illum.magn;
            c.b += ((m.clrh * kd.b) + ((1.0f - m.clrh) * li.b)) * spec_shade * illum.magn;
            continue;
        }
        const float shade = dot(illum.dir, n);
        if (shade < 0.0f) continue;
        if (m.kind == kMandelbrot) {  // mandelbrot_shader.cc:23-38: diffuse only, no light colour
            total_shade += shade * illum.magn;
            continue;
        }
        if (m.kind == kCookTorrance) {
            cook_torrance_term(m, n, r.dir, nv, illum.dir, shade, illum.magn, li, &c);
            continue;
        }
        // Reflection of the light ray about the normal.
        const float shade_x2 = 2.0f * shade;
        const V3 ref_dir(illum.dir.x - n.x * shade_x2, illum.dir.y - n.y * shade_x2, illum.dir.z - n.z * shade_x2);
        const float ray_dot_reflection = dot(r.dir, ref_dir);
        Colour sc{shade * kd.r, shade * kd.g, shade * kd.b};
        if (ray_dot_reflection > 0.0f) {
            const float sp = std::pow(ray_dot_reflection, ns);
            sc.r += sp * ks.r;
            sc.g += sp * ks.g;
            sc.b += sp * ks.b;
        }
        c.r += li.r * (sc.r * illum.magn);
        c.g += li.g * (sc.g * illum.magn);
        c.b += li.b * (sc.b * illum.magn);
    }
    if (m.kind == kMandelbrot) {
        c = mandelbrot_colour(sh.dst, n, m.iters);
        c.r *= total_shade;
        c.g *= total_shade;
        c.b *= total_shade;
    } else if (m.kind != kCookTorrance && m.kind != kMapper) {  // no ambient term in those two
        c.r += ka.r;
        c.g += ka.g;
        c.b += ka.b;
    }
    {
        // ray::reflect: dir - n * 2 (dir . n) from a point moved off the front.
        const float power = r.magn * refl_c;
        if (static_cast<double>(power) > kMinReflectivePower) {
            Ray rl;
            rl.ogn = offset_start_point(sh.dst, sh.geo_norm, 1);
            rl.dir = reflect_dir(r.dir, n);
            rl.magn = power;
            const Colour pix = trace_phong(f, rl, false, lc, pc, nullptr, nullptr, nullptr);
            // phong and coloured_mapper: rf (untextured); mapper_shader: its layered value.
            const float w = ((m.kind == kMapper) ? refl_c : m.rf) / 1.0f;
            c.r += pix.r * w;
            c.g += pix.g * w;
            c.b += pix.b * w;
        }
    }
    {
        const float power = r.magn * tran_c;
        V3 rdir;
        if (static_cast<double>(power) > kMinReflectivePower && refract_dir(r.dir, n, m.ri, leaving, &rdir)) {
            Ray rr;
            rr.ogn = offset_start_point(sh.dst, sh.geo_norm, -1);
            rr.dir = rdir;
            rr.magn = power;
            const Colour pix = trace_phong(f, rr, false, lc, pc, nullptr, nullptr, nullptr);
            if (m.kind == kMapper) {
                // mapper_shader.cc:179-183: c += clrf c + (1 - clrf) refracted trans.
                const Colour add{(m.clrf * c.r) + ((1.0f - m.clrf) * pix.r * tran_c),
                                 (m.clrf * c.g) + ((1.0f - m.clrf) * pix.g * tran_c),
                                 (m.clrf * c.b) + ((1.0f - m.clrf) * pix.b * tran_c)};
                c.r += add.r;
                c.g += add.g;
                c.b += add.b;
            } else {
                const float w = m.tran / 1.0f;
                c.r += pix.r * w;
                c.g += pix.g * w;
                c.b += pix.b * w;
            }
        }
    }
    return c;
}

void render_pixel_phong(const FrameCtx &f, int px, int py, LocalCounters *lc, PhongCounters *pc) {
    const int64_t p = static_cast<int64_t>(py) * f.cam.x_res + px;
    const int nl = static_cast<int>(f.lights.size());
    const int spp = std::max(1, f.p.spp);
    Colour acc;
    for (int s = 0; s < spp; ++s) {
        float xf = static_cast<float>(px), yf = static_cast<float>(py);
        if (spp > 1) {
            float jx, jy;
            pixel_jitter(f.p.seed, static_cast<uint32_t>(p), static_cast<uint32_t>(s), &jx, &jy);
            xf = xf + jx;
            yf = yf + jy;
        }
        Ray r = primary_ray(f.cam, xf, yf);
        const bool first = (s == 0);
        const Colour c = trace_phong(f, r, true, lc, pc, first ? f.out.prim + (f.out.prim ? p : 0) : nullptr,
                                     (first && f.out.tuv) ? f.out.tuv + 3 * p : nullptr,
                                     (first && f.out.vis) ? f.out.vis + nl * p : nullptr);
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

// ---------------------------------------------------------------------------
// The SSE packet frame (#ifdef SIMD_PACKET_TRACING, MAXIMUM_PACKET_SIZE 16):
// one 8x8 tile is 16 packets of 2x2 pixels, rt/raytracer.cc:464-483.

constexpr int kPackets = 16;
constexpr int kTileRays = kPackets * 4;

// rt/packet_ray.h:255-293 (packet -> tile offset) and :204-250 (ray ->
// pixel), built the same way: interleaved address bits.
struct PacketLuts {
    int xoff[kPackets], yoff[kPackets];
    int ray_to_pixel[kTileRays];
    PacketLuts() {
        for (int i = 0; i < kPackets; ++i) {
            unsigned x = 0, y = 0, addr = static_cast<unsigned>(i) << 1, mask = 0x2;
            for (int j = 0; j < 2; ++j) {
                x += addr & mask;
                addr >>= 1;
                y += addr & mask;
                mask <<= 1;
            }
            xoff[i] = static_cast<int>(x);
            yoff[i] = static_cast<int>(y);
        }
        for (int i = 0; i < kTileRays; ++i) {
            unsigned x = 0, y = 0, addr = static_cast<unsigned>(i), mask = 0x1;
            for (int j = 0; j < 8; ++j) {
                x += addr & mask;
                addr >>= 1;
                y += addr & mask;
                mask <<= 1;
            }
            ray_to_pixel[i] = static_cast<int>(y * 8 + x);
        }
    }
};
const PacketLuts &luts() {
    static const PacketLuts l;
    return l;
}

float lane(__m128 v, int k) {
    alignas(16) float f[4];
    _mm_store_ps(f, v);
    return f[k];
}

Ray lane_ray(const PacketRay &p, int k) {
    Ray r;
    r.ogn = V3(lane(p.ox, k), lane(p.oy, k), lane(p.oz, k));
    r.dir = V3(lane(p.dx, k), lane(p.dy, k), lane(p.dz, k));
    return r;
}

// packet_ray::pack (rt/packet_ray.h:71-92): four rays, their lengths.
PacketRay pack4(const Ray *const *rr, __m128 *t) {
    PacketRay p;
    p.ox = _mm_setr_ps(rr[0]->ogn.x, rr[1]->ogn.x, rr[2]->ogn.x, rr[3]->ogn.x);
    p.oy = _mm_setr_ps(rr[0]->ogn.y, rr[1]->ogn.y, rr[2]->ogn.y, rr[3]->ogn.y);
    p.oz = _mm_setr_ps(rr[0]->ogn.z, rr[1]->ogn.z, rr[2]->ogn.z, rr[3]->ogn.z);
    p.dx = _mm_setr_ps(rr[0]->dir.x, rr[1]->dir.x, rr[2]->dir.x, rr[3]->dir.x);
    p.dy = _mm_setr_ps(rr[0]->dir.y, rr[1]->dir.y, rr[2]->dir.y, rr[3]->dir.y);
    p.dz = _mm_setr_ps(rr[0]->dir.z, rr[1]->dir.z, rr[2]->dir.z, rr[3]->dir.z);
    if (t) *t = _mm_setr_ps(rr[0]->length, rr[1]->length, rr[2]->length, rr[3]->length);
    return p;
}

// Coherency test of rt/raytracer.cc:172-200 and :93-121: a packet whose
// lanes disagree in the sign of a direction component goes ray by ray
// (size 0); the packets are "coherent" when there are more than four and
// all share packet 0's signs.
bool coherency(const PacketRay *r, int s, int *size) {
    bool coherent = s > 4;
    const int px = _mm_movemask_ps(r[0].dx), py = _mm_movemask_ps(r[0].dy), pz = _mm_movemask_ps(r[0].dz);
    for (int i = 0; i < s; ++i) {
        const int x = _mm_movemask_ps(r[i].dx), y = _mm_movemask_ps(r[i].dy), z = _mm_movemask_ps(r[i].dz);
        if ((x != 0 && x != 0xf) || (y != 0 && y != 0xf) || (z != 0 && z != 0xf)) {
            coherent = false;
            size[i] = 0;
        } else if (x != px || y != py || z != pz) {
            coherent = false;
            size[i] = 1;
        } else {
            size[i] = 1;
        }
    }
    return coherent;
}

// rt/raytracer.cc:168-231: closest hits of s packets. The coherent case
// (frustum wrapper, ssd/bvh.cc:107-117 under #if 1) is a per-packet loop.
void engine_closest(const Scene &sc, const PacketRay *r, int s, int *ids, Hit *hits, Counters *c) {
    int size[kPackets];
    const bool coherent = coherency(r, s, size);
    for (int i = 0; i < s; ++i) {
        if (!coherent && size[i] == 0) {
            for (int k = 0; k < 4; ++k) {
                const Ray ray = lane_ray(r[i], k);
                Hit h;
                ids[4 * i + k] = find_nearest_object(sc, ray, &h, c);
                hits[4 * i + k] = h;
            }
        } else {
            PacketHit h;
            packet_nearest(sc, r[i], &h, c);
            alignas(16) float d[4], u[4], v[4];
            alignas(16) int id[4];
            _mm_store_ps(d, h.d);
            _mm_store_ps(u, h.u);
            _mm_store_ps(v, h.v);
            _mm_store_si128(reinterpret_cast<__m128i *>(id), h.id);
            for (int k = 0; k < 4; ++k) {
                ids[4 * i + k] = d[k] < kMaxDist ? id[k] : -1;
                hits[4 * i + k].d = d[k];
                hits[4 * i + k].u = u[k];
                hits[4 * i + k].v = v[k];
            }
        }
    }
}

// rt/raytracer.cc:91-165 shoot_shadow_packet: visibility of s packed shadow
// packets (sr: their rays, for the ray-by-ray case).
void engine_occluded(const Scene &sc, const PacketRay *r, const __m128 *t, const Ray *const *sr, int s, bool *occ,
                     Counters *c) {
    int size[kPackets];
    const bool coherent = coherency(r, s, size);
    for (int j = 0; j < s; ++j) {
        if (!coherent && size[j] == 0) {
            for (int k = 0; k < 4; ++k) occ[4 * j + k] = found_nearer_object(sc, *sr[4 * j + k], sr[4 * j + k]->length, c);
        } else {
            const int m = _mm_movemask_ps(packet_nearer(sc, r[j], t[j], c));
            for (int k = 0; k < 4; ++k) occ[4 * j + k] = (m >> k) & 1;
        }
    }
}

// Visibility of n shadow rays packed in order, 16 packets at a time; the
// remainder (n mod 4) goes ray by ray (rt/raytracer.cc:286-323).
void shadow_batch(const Scene &sc, const Ray *const *rays, int n, bool *occ, Counters *c) {
    PacketRay pk[kPackets];
    __m128 t[kPackets];
    int done = 0;
    while (n - done >= 4) {
        const int np = std::min(kPackets, (n - done) / 4);
        for (int j = 0; j < np; ++j) pk[j] = pack4(rays + done + 4 * j, &t[j]);
        engine_occluded(sc, pk, t, rays + done, np, occ + done, c);
        done += 4 * np;
    }
    for (; done < n; ++done) occ[done] = found_nearer_object(sc, *rays[done], rays[done]->length, c);
}

// Shading of a batch of hit rays: shadow rays per light packed over the
// batch, then Lambert (the same arithmetic as the scalar frame).
void shade_batch(const FrameCtx &f, const SurfaceHit *sh, const Colour *kd, int n, Colour *out, LocalCounters *lc,
                 int ray_kind) {
    const int nl = static_cast<int>(f.lights.size());
    std::vector<Ray> illum(static_cast<size_t>(n) * std::max(1, nl));
    std::vector<const Ray *> ptr(n);
    std::vector<uint8_t> occ_b(n);
    bool occ[kTileRays];
    for (int l = 0; l < nl; ++l) {
        for (int k = 0; k < n; ++k) {
            illum[static_cast<size_t>(k) * nl + l] = shadow_ray(sh[k], f.lights[l].c);
            ptr[k] = &illum[static_cast<size_t>(k) * nl + l];
        }
        for (int b = 0; b < n; b += kTileRays) {
            const int m = std::min(kTileRays, n - b);
            shadow_batch(*f.s, ptr.data() + b, m, occ, &lc->trav);
            for (int k = 0; k < m; ++k) illum[static_cast<size_t>(b + k) * nl + l].magn = (occ[k] ? 0.0f : 1.0f) / 1.0f;
        }
        lc->rays[ray_kind] += static_cast<uint64_t>(n);
    }
    for (int k = 0; k < n; ++k) out[k] = shade_direct(kd[k], sh[k], &illum[static_cast<size_t>(k) * nl], f.lights);
}

void render_tile_packet(const FrameCtx &f, int tx, int ty, LocalCounters *lc) {
    const PacketLuts &L = luts();
    const CameraState &cam = f.cam;
    PacketRay pk[kPackets];
    // rt/camera.h:185-229: directions normalised with inverse_sqrt.
    for (int i = 0; i < kPackets; ++i) {
        const float x = static_cast<float>(tx + L.xoff[i]), y = static_cast<float>(ty + L.yoff[i]);
        const __m128 vx = _mm_setr_ps(x, x + 1.0f, x, x + 1.0f);
        const __m128 vy = _mm_setr_ps(y, y, y + 1.0f, y + 1.0f);
        const __m128 vx_t = _mm_add_ps(_mm_mul_ps(vx, _mm_set1_ps(cam.x_inc)), _mm_set1_ps(cam.x_m));
        const __m128 vy_t = _mm_add_ps(_mm_mul_ps(vy, _mm_set1_ps(cam.y_inc)), _mm_set1_ps(cam.y_m));
        __m128 dx = _mm_mul_ps(_mm_set1_ps(cam.X.x), vx_t);
        __m128 dy = _mm_mul_ps(_mm_set1_ps(cam.X.y), vx_t);
        __m128 dz = _mm_mul_ps(_mm_set1_ps(cam.X.z), vx_t);
        dx = _mm_add_ps(dx, _mm_mul_ps(_mm_set1_ps(cam.Y.x), vy_t));
        dy = _mm_add_ps(dy, _mm_mul_ps(_mm_set1_ps(cam.Y.y), vy_t));
        dz = _mm_add_ps(dz, _mm_mul_ps(_mm_set1_ps(cam.Y.z), vy_t));
        dx = _mm_add_ps(dx, _mm_mul_ps(_mm_set1_ps(cam.Z.x), _mm_set1_ps(cam.t)));
        dy = _mm_add_ps(dy, _mm_mul_ps(_mm_set1_ps(cam.Z.y), _mm_set1_ps(cam.t)));
        dz = _mm_add_ps(dz, _mm_mul_ps(_mm_set1_ps(cam.Z.z), _mm_set1_ps(cam.t)));
        const __m128 len = packet_inverse_sqrt(
            _mm_add_ps(_mm_add_ps(_mm_mul_ps(dx, dx), _mm_mul_ps(dy, dy)), _mm_mul_ps(dz, dz)));
        pk[i].ox = _mm_set1_ps(cam.pos.x);
        pk[i].oy = _mm_set1_ps(cam.pos.y);
        pk[i].oz = _mm_set1_ps(cam.pos.z);
        pk[i].dx = _mm_mul_ps(dx, len);
        pk[i].dy = _mm_mul_ps(dy, len);
        pk[i].dz = _mm_mul_ps(dz, len);
    }
    int ids[kTileRays];
    Hit hits[kTileRays];
    engine_closest(*f.s, pk, kPackets, ids, hits, &lc->trav);
    lc->rays[0] += kTileRays;
    // Hit rays in ray order: surface, kd.
    int hit_ray[kTileRays];
    SurfaceHit sh[kTileRays];
    Colour kd[kTileRays];
    Ray rays[kTileRays];
    int nh = 0;
    for (int k = 0; k < kTileRays; ++k) {
        rays[k] = lane_ray(pk[k >> 2], k & 3);
        if (!(hits[k].d < kMaxDist)) continue;
        hit_ray[nh] = k;
        sh[nh] = surface_at(f.s->tris[ids[k]], &rays[k], hits[k].d);
        kd[nh] = kd_of(f, ids[k]);
        ++nh;
    }
    Colour direct[kTileRays];
    shade_batch(f, sh, kd, nh, direct, lc, 1);
    Colour col[kTileRays];
    int bounce_id[kTileRays];  // bounce hit of each ray, -1: none
    for (int k = 0; k < kTileRays; ++k) {
        col[k] = Colour{};
        bounce_id[k] = -1;
    }
    for (int j = 0; j < nh; ++j) col[hit_ray[j]] = direct[j];
    if (f.p.bounce) {
        // Bounce rays of the hit rays, packed in order and traced through
        // the same engine (as the reference packs secondary rays,
        // rt/raytracer.cc:353-394); the remainder goes ray by ray.
        Ray br[kTileRays];
        const Ray *bp[kTileRays];
        for (int j = 0; j < nh; ++j) {
            const int k = hit_ray[j];
            const int px = tx + (L.ray_to_pixel[k] & 7), py = ty + (L.ray_to_pixel[k] >> 3);
            const int64_t p = static_cast<int64_t>(py) * cam.x_res + px;
            br[j].ogn = offset_start_point(sh[j].dst, sh[j].geo_norm, 1);
            br[j].dir = cosine_direction(bounce_key(f.p.seed, static_cast<uint32_t>(p), 0u, 1u), sh[j].shade_norm);
            bp[j] = &br[j];
        }
        int bid[kTileRays];
        Hit bh[kTileRays];
        const int np = nh / 4;
        if (np > 0) {
            PacketRay bpk[kPackets];
            for (int i = 0; i < np; ++i) bpk[i] = pack4(bp + 4 * i, nullptr);
            engine_closest(*f.s, bpk, np, bid, bh, &lc->trav);
        }
        for (int j = 4 * np; j < nh; ++j) bid[j] = find_nearest_object(*f.s, br[j], &bh[j], &lc->trav);
        lc->rays[2] += static_cast<uint64_t>(nh);
        SurfaceHit bsh[kTileRays];
        Colour bkd[kTileRays];
        int bsrc[kTileRays];
        int nb = 0;
        for (int j = 0; j < nh; ++j) {
            if (!(bh[j].d < kMaxDist)) continue;
            bounce_id[hit_ray[j]] = bid[j];
            bsh[nb] = surface_at(f.s->tris[bid[j]], &br[j], bh[j].d);
            bkd[nb] = kd_of(f, bid[j]);
            bsrc[nb] = j;
            ++nb;
        }
        Colour c2[kTileRays];
        shade_batch(f, bsh, bkd, nb, c2, lc, 3);
        for (int b = 0; b < nb; ++b) {
            const int j = bsrc[b];
            Colour &c = col[hit_ray[j]];
            c.r = c.r + ((kd[j].r / 255.0f) * c2[b].r);
            c.g = c.g + ((kd[j].g / 255.0f) * c2[b].g);
            c.b = c.b + ((kd[j].b / 255.0f) * c2[b].b);
        }
    }
    for (int k = 0; k < kTileRays; ++k) {
        const int px = tx + (L.ray_to_pixel[k] & 7), py = ty + (L.ray_to_pixel[k] >> 3);
        if (px >= cam.x_res || py >= cam.y_res) continue;
        const int64_t p = static_cast<int64_t>(py) * cam.x_res + px;
        if (f.out.prim) f.out.prim[p] = ids[k];
        if (f.out.bounce_prim) f.out.bounce_prim[p] = f.p.bounce ? bounce_id[k] : -1;
        if (f.out.tuv) {
            f.out.tuv[3 * p + 0] = hits[k].d;
            f.out.tuv[3 * p + 1] = hits[k].u;
            f.out.tuv[3 * p + 2] = hits[k].v;
        }
        if (f.out.rgb) {
            f.out.rgb[3 * p + 0] = col[k].r / 1.0f;
            f.out.rgb[3 * p + 1] = col[k].g / 1.0f;
            f.out.rgb[3 * p + 2] = col[k].b / 1.0f;
        }
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
    std::vector<PhongCounters> pcounters(nthreads);
    const int tiles_x = (f.cam.x_res + 7) / 8;
    const int64_t tiles = static_cast<int64_t>(tiles_x) * ((f.cam.y_res + 7) / 8);
    auto packet_worker = [&](int tid) {
        while (true) {
            const int64_t t = next.fetch_add(1) * stride;  // the packet frame strides over tiles
            if (t >= tiles) break;
            render_tile_packet(f, static_cast<int>(t % tiles_x) * 8, static_cast<int>(t / tiles_x) * 8,
                               &counters[tid]);
        }
    };
    auto worker = [&](int tid) {
        if (p.packet) {
            packet_worker(tid);
            return;
        }
        const int64_t chunk = 64;
        while (true) {
            const int64_t b = next.fetch_add(chunk);
            if (b * stride >= npix) break;
            for (int64_t k = b; k < b + chunk; ++k) {
                const int64_t pix = k * stride;
                if (pix >= npix) break;
                if (p.materials || p.vertex_normals)
                    render_pixel_phong(f, static_cast<int>(pix % f.cam.x_res), static_cast<int>(pix / f.cam.x_res),
                                       &counters[tid], &pcounters[tid]);
                else
                    render_pixel(f, static_cast<int>(pix % f.cam.x_res), static_cast<int>(pix / f.cam.x_res),
                                 &counters[tid]);
            }
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nthreads; ++t) pool.emplace_back(worker, t);
    worker(0);
    for (auto &th : pool) th.join();
    if (out.sec_counters) {
        out.sec_counters[0] = out.sec_counters[1] = 0;
        for (auto &c : pcounters) {
            out.sec_counters[0] += c.secondary;
            out.sec_counters[1] += c.secondary_shadow;
        }
    }
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

extern "C" {

// One texture_map (mono = 0: rgb in/out) or texture_map_monochrome (mono =
// 1: rgb[0] out) of map `map` at hit point dst; returns the alpha.
float oracle_texture_sample(const oracle_texture_map *maps, const oracle_texture_image *images, const float *texels,
                            int32_t map, const float *dst, const float *dir, const float *n, int32_t has_vt,
                            float vtu, float vtv, int32_t mono, float *rgb) {
    const oracle_tex::Tables T{maps, images, texels, nullptr};
    const oracle::V3 d(dst[0], dst[1], dst[2]), rd(dir[0], dir[1], dir[2]), nn(n[0], n[1], n[2]);
    if (mono) return oracle_tex::texture_map_monochrome(T, map, rgb, d, rd, nn, has_vt != 0, vtu, vtv);
    oracle_tex::Rgb3 c{rgb[0], rgb[1], rgb[2]};
    const float a = oracle_tex::texture_map(T, map, &c, d, rd, nn, has_vt != 0, vtu, vtv);
    rgb[0] = c.r;
    rgb[1] = c.g;
    rgb[2] = c.b;
    return a;
}


float oracle_perlin_noise(float x, float y, float z, int32_t seed, int32_t lattice) {
    if (lattice)
        return oracle_tex::perlin_noise(static_cast<uint32_t>(static_cast<int32_t>(x)),
                                        static_cast<uint32_t>(static_cast<int32_t>(y)),
                                        static_cast<uint32_t>(static_cast<int32_t>(z)), static_cast<uint32_t>(seed));
    return oracle_tex::perlin_interpolated_noise(x, y, z, static_cast<uint32_t>(seed));
}

void oracle_mandelbrot_colour(const float *dst, const float *n, int32_t iterations, float *rgb) {
    const oracle::Colour c =
        oracle::mandelbrot_colour(oracle::V3(dst[0], dst[1], dst[2]), oracle::V3(n[0], n[1], n[2]), iterations);
    rgb[0] = c.r;
    rgb[1] = c.g;
    rgb[2] = c.b;
}

void oracle_reflect_dir(const float *d, const float *n, float *out) {
    const oracle::V3 r = oracle::reflect_dir(oracle::V3(d[0], d[1], d[2]), oracle::V3(n[0], n[1], n[2]));
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
}

int32_t oracle_refract_dir(const float *d, const float *n, float ri, int32_t leaving, float *out) {
    oracle::V3 r;
    if (!oracle::refract_dir(oracle::V3(d[0], d[1], d[2]), oracle::V3(n[0], n[1], n[2]), ri, leaving != 0, &r))
        return 0;
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
    return 1;
}

}  // extern "C"
