// Frames with materials: the reference's recursive shading of one hit
// (rt/raytracer.cc:20-87) with phong_shader::shade
// (rt/materials/phong_shader.cc:21-111) or cook_torrance_cxy::shade
// (rt/materials/cook_torrance_cxy.cc:12-90) and the mirror and refraction rays of
// ray::reflect / ray::refract without DIFFUSE_REFLECTIONS (rt/ray.cc:118-258),
// run as a wavefront of ray generations:
//
//   generation 0     the primary rays of every sample (8x8 tiles, as rt_render)
//   closest hit
//   k_mat_classify   hit flag and number of children of each ray (mirror ray
//                    while the ray power times rf stays above
//                    MIN_REFLECTIVE_POWER, refraction ray likewise unless
//                    totally reflected)
//   scans            compacted hit slots and child slots, in ray order
//   k_mat_emit       per hit: surface record, one shadow ray per light, the
//                    children into the next generation's queue
//   any hit          shadow rays
//   k_mat_local      local colour of every hit (phong: diffuse, specular,
//                    ambient; Cook-Torrance: diffuse and microfacet specular
//                    through the chroma-to-RGB conversion)
//
// until a generation has no children. Then k_mat_gather walks the
// generations backwards and folds each hit's children into it in the
// reference's order, c = (local + rf c_reflected) + tran c_refracted, so the
// colour of a ray is the one the recursion returns, rounding included.
// Children sit in the next generation in the order of their parents and
// every sum has a fixed order: frames are reproducible bit for bit.
#include <algorithm>
#include <cfloat>
#include <type_traits>
#include <vector>

#include "frame.hpp"
#include "radix_sort.hpp"
#include "rt_device.hpp"
#include "textures.hpp"

namespace rt {

namespace {

// common/common.h:30-32 is a double literal, so the float power is compared
// in double (a power of exactly 0.1f lies above it and still recurses).
constexpr double kMinReflectivePower = 0.1;

constexpr int kCookTorrance = 1;
constexpr int kColoured = 2;
constexpr int kMapper = 3;
constexpr int kMandelbrot = 4;

// A material row (kMaterialFloat4s float4): {ka rgb, shininess},
// {kd rgb, transmittance}, {ks rgb, refractive index},
// {reflectance, kind, chroma x, chroma y}, {diffuse, roughness, extinction, 0},
// {tex ka, kd, ks, ns}, {tex refl, d, mapper kd, mapper ks},
// {clrh, clrf, layers 0, layers 1}, {layers 2, 3, 4, iterations}; texture
// slots are 1 + map index (0 none), layer lists first << 8 | count and the
// Mandelbrot iteration limit as int bits.
// Cook-Torrance keeps rs, ts, ri_r where phong keeps rf, tran, ri.
struct Mat {
    float4 ka_s, kd_tran, ks_ri;
    float rf;
    int kind;
    float cx, cy, rd, sr;
    int tex[6];
    float mkd, mks, clrh, clrf;
    int layers[5];
    int iters;
};

// The triangle's material: its table entry, or phong_shader(255 * albedo, 1)
// (kd from the scene, ks = ka = 0, s = 0, no mirror or refraction).
__device__ __forceinline__ Mat mat_of(const float4 *mats, const int32_t *tri_mat, const float4 *kd, int32_t prim) {
    Mat m;
    const int32_t mi = tri_mat ? tri_mat[prim] : -1;
    if (mi < 0) {
        const float4 k = kd[prim];
        m.ka_s = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        m.kd_tran = make_float4(k.x, k.y, k.z, 0.0f);
        m.ks_ri = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
        m.rf = 0.0f;
        m.kind = 0;
        m.cx = m.cy = m.rd = m.sr = 0.0f;
        for (int k = 0; k < 6; ++k) m.tex[k] = 0;
        m.mkd = m.mks = m.clrh = m.clrf = 0.0f;
        for (int k = 0; k < 5; ++k) m.layers[k] = 0;
        m.iters = 0;
    } else {
        const float4 *q = mats + kMaterialFloat4s * static_cast<int64_t>(mi);
        m.ka_s = q[0];
        m.kd_tran = q[1];
        m.ks_ri = q[2];
        const float4 q3 = q[3], q4 = q[4];
        m.rf = q3.x;
        m.kind = static_cast<int>(q3.y);
        m.cx = q3.z;
        m.cy = q3.w;
        m.rd = q4.x;
        m.sr = q4.y;
        const float4 q5 = q[5], q6 = q[6], q7 = q[7], q8 = q[8];
        m.tex[0] = __float_as_int(q5.x);
        m.tex[1] = __float_as_int(q5.y);
        m.tex[2] = __float_as_int(q5.z);
        m.tex[3] = __float_as_int(q5.w);
        m.tex[4] = __float_as_int(q6.x);
        m.tex[5] = __float_as_int(q6.y);
        m.mkd = q6.z;
        m.mks = q6.w;
        m.clrh = q7.x;
        m.clrf = q7.y;
        m.layers[0] = __float_as_int(q7.z);
        m.layers[1] = __float_as_int(q7.w);
        m.layers[2] = __float_as_int(q8.x);
        m.layers[3] = __float_as_int(q8.y);
        m.layers[4] = __float_as_int(q8.z);
        m.iters = __float_as_int(q8.w);
    }
    return m;
}

using TexPoint = TexHit;

// A monochrome layer list of mapper_shader (rt/materials/mapper_shader.cc:
// 41-66, 89-126): each layer sets the value to its texel mean / 255; the
// list stops once the product of the opacities is 0.
__device__ __forceinline__ float mono_layers(const TexTables &T, int list, float value, const TexPoint &p) {
    const int first = list >> 8, count = list & 0xff;
    float alpha = 1.0f;
    for (int k = 0; k < count; ++k) {
        float z;
        alpha *= texture_mono(T, T.layers[first + k], &z, p);
        value = z * (1.0f / 255.0f);
        if (alpha == 0.0f) break;
    }
    return value;
}

// The mirror and refraction coefficients a hit's rays are spawned with:
// coloured_mapper_shader scales rf and tran by red / 255 of t_refl and t_d
// (rt/materials/coloured_mapper_shader.cc:8-38); mapper_shader takes its
// reflectance and transmittance layers (rt/materials/mapper_shader.cc:41-66).
__device__ __forceinline__ void child_coefficients(const Mat &m, const TexTables &T, const TexPoint &p, float *refl,
                                                   float *tran) {
    *refl = m.rf;
    *tran = m.kd_tran.w;
    if (m.kind == kColoured) {
        if (m.tex[4]) {
            TexColour t{0.0f, 0.0f, 0.0f};
            texture_colour(T, m.tex[4] - 1, &t, p);
            *refl = t.r * (m.rf * (1.0f / 255.0f));
        }
        if (m.tex[5]) {
            TexColour t{0.0f, 0.0f, 0.0f};
            texture_colour(T, m.tex[5] - 1, &t, p);
            *tran = t.r * (m.kd_tran.w * (1.0f / 255.0f));
        }
    } else if (m.kind == kMapper) {
        *refl = mono_layers(T, m.layers[3], m.rf, p);
        *tran = mono_layers(T, m.layers[4], m.kd_tran.w, p);
    }
}

// A coloured_mapper_shader colour: the map's colour times colour / 255
// (rt/materials/coloured_mapper_shader.cc:58-80).
__device__ __forceinline__ float4 mapped_colour(const TexTables &T, int slot, float4 colour, const TexPoint &p) {
    if (!slot) return colour;
    TexColour t{colour.x, colour.y, colour.z};
    texture_colour(T, slot - 1, &t, p);
    return make_float4(t.r * (colour.x * (1.0f / 255.0f)), t.g * (colour.y * (1.0f / 255.0f)),
                       t.b * (colour.z * (1.0f / 255.0f)), colour.w);
}

// XYZ -> RGB for the primaries and white point of
// rt/materials/material.h:17-45, in float and in the order written there.
struct CieMatrix {
    float m[3][3];
};

__device__ __forceinline__ CieMatrix cie_matrix() {
    const float xr = 0.640f, yr = 0.330f, xg = 0.290f, yg = 0.600f, xb = 0.150f, yb = 0.060f;
    const float xw = 0.3333f, yw = 0.3333f;
    const float crd = ((1.0f / yw) * (xw * (yg - yb) - yw * (xg - xb) + xg * yb - xb * yg));
    const float cgd = ((1.0f / yw) * (xw * (yb - yr) - yw * (xb - xr) - xr * yb + xb * yr));
    const float cbd = ((1.0f / yw) * (xw * (yr - yg) - yw * (xr - xg) + xr * yg - xg * yr));
    CieMatrix c;
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

// cxy_to_rgb (rt/materials/material.h:326-343): chroma (x, y) at luminance
// lum, clamped to [0, 1] and scaled to 0-255.
__device__ __forceinline__ Rgb cxy_to_rgb(const CieMatrix &cm, float x, float y, float lum) {
    const float c0 = lum * x / y;
    const float c1 = lum;
    const float c2 = lum * (1.0f / y - 1.0f) - c0;
    float o[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) o[k] = (cm.m[k][0] * c0) + (cm.m[k][1] * c1) + (cm.m[k][2] * c2);
    return Rgb{fmaxf(0.0f, fminf(1.0f, o[0])) * 255.0f, fmaxf(0.0f, fminf(1.0f, o[1])) * 255.0f,
               fmaxf(0.0f, fminf(1.0f, o[2])) * 255.0f};
}

// cook_torrance_cxy::shade for one light (rt/materials/cook_torrance_cxy.cc:33-76)
// with schlick_fresnell, geometric_attenuation_factor and
// gaussian_facet_distribution (rt/materials/material.h:146-215); adds
// rgb x light intensity to c.
__device__ __forceinline__ void cook_torrance_term(const Mat &m, const CieMatrix &cm, rtd::V3 n, rtd::V3 d, float nv,
                                                   rtd::V3 l, float shade, float magn, float4 li, Rgb &c) {
    const rtd::V3 half = rtd::mk((l.x + d.x) * 0.5f, (l.y + d.y) * 0.5f, (l.z + d.z) * 0.5f);
    const float nh = rtd::dot(n, half);
    const float vh = rtd::dot(half, d);
    const float nl = shade;
    const float f = (m.sr + (1.0f - m.sr) * powf(1.0f - nv, 5.0f));
    const float g0 = (2.0f * nh * nv) / vh;
    const float g1 = (2.0f * nh * nl) / vh;
    const float g = fminf(1.0f, fmaxf(0.0f, fminf(g0, g1)));
    const float a_div_m = nh / m.sr;
    const float ro = 5.0f * expf(-(a_div_m * a_div_m));
    const float s = m.rf * ((g * f * ro) / (nl * nv));
    const Rgb rgb = cxy_to_rgb(cm, m.cx, m.cy, (shade * (m.rd + s) * magn));
    c.r += rgb.r * li.x;
    c.g += rgb.g * li.y;
    c.b += rgb.b * li.z;
}

// ray::refract (rt/ray.cc:200-227): n faces the ray; the index is inverted
// for a ray entering the front face (hit_t::out_in); false on total internal
// reflection.
__device__ __forceinline__ bool refract_dir(rtd::V3 d, rtd::V3 n, float ri, bool leaving, rtd::V3 *out) {
    const float rdn = -rtd::dot(d, n);
    if (!leaving) ri = 1.0f / ri;
    const float cos_t2 = 1.0f - ri * ri * (1.0f - rdn * rdn);
    if (cos_t2 < 0.0f) return false;
    const float offset = (ri * rdn - sqrtf(cos_t2));
    *out = rtd::add(rtd::mul(d, ri), rtd::mul(n, offset));
    return true;
}

// ray::reflect (rt/ray.cc:121-131): dir - n * 2 (dir . n).
__device__ __forceinline__ rtd::V3 reflect_dir(rtd::V3 d, rtd::V3 n) {
    const float rdn = 2.0f * rtd::dot(d, n);
    return rtd::sub(d, rtd::mul(n, rdn));
}

// One generation: rays (2 float4 each), their power, their closest hits.
struct GenArgs {
    const float4 *rays;
    const float *power;
    const float4 *hits;  // t, u, v, prim
    int64_t n;
    const float4 *tri_n, *tri_c, *kd, *mats;
    const int32_t *tri_mat;
    const float *vn;  // vertex normals, 9 per triangle, or null
    TexTables tex;    // texture maps and coordinates (maps null: none)
};

// Padding slots past the image edge get tmax < tmin, so they never hit.
__global__ void __launch_bounds__(256) k_mat_raygen(CamArgs c, TileMap tm, int spp, uint64_t seed, int64_t ns,
                                                    float4 *__restrict__ rays, float *__restrict__ power) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= ns) return;
    const int64_t slot = i / spp;
    const int s = static_cast<int>(i % spp);
    power[i] = 1.0f;
    int x, y;
    if (!tm.pixel_of_slot(slot, &x, &y)) {
        rays[2 * i] = make_float4(c.px, c.py, c.pz, rtd::kEpsilon);
        rays[2 * i + 1] = make_float4(0.0f, 0.0f, 1.0f, -1.0f);
        return;
    }
    const int64_t p = static_cast<int64_t>(y) * c.x_res + x;
    const rtd::V3 d = primary_dir(c, x, y, p, s, spp, seed);
    rays[2 * i] = make_float4(c.px, c.py, c.pz, rtd::kEpsilon);
    rays[2 * i + 1] = make_float4(d.x, d.y, d.z, rtd::kMaxDist);
}

// The surface at ray i's hit and which children it spawns, with the
// coefficients (possibly textured) the children are spawned with.
struct HitSurface {
    rtd::V3 dst, geo, shn, rdir;
    Mat m;
    TexPoint tp;
    float refl_c, tran_c;
    bool refl, refr;
};

__device__ __forceinline__ HitSurface surface_of(const GenArgs &a, int64_t i) {
    HitSurface h;
    const float4 r0 = a.rays[2 * i], r1 = a.rays[2 * i + 1];
    const float4 hit = a.hits[i];
    const int32_t prim = __float_as_int(hit.w);
    const float4 tn = a.tri_n[prim];
    const rtd::V3 d = rtd::f4v(r1);
    const bool leaving = rtd::dot(rtd::f4v(tn), d) > 0.0f;  // hit_t::in_out (rt/triangle.h:487-496)
    surface_at(rtd::f4v(r0), d, hit.x, rtd::f4v(tn), rtd::f4v(a.tri_c[prim]), &h.dst, &h.geo, &h.shn);
    if (a.vn && a.vn[9 * static_cast<int64_t>(prim)] != FLT_MAX) {
        // rt/triangle.h:474-477: u weights vertex c, v vertex b, 1 - (u + v)
        // vertex a; not renormalised; turned like the geometric normal.
        const float *q = a.vn + 9 * static_cast<int64_t>(prim);
        const float u = hit.y, v = hit.z, w = 1.0f - (u + v);
        rtd::V3 s = rtd::mk(((u * q[6]) + (v * q[3])) + (w * q[0]), ((u * q[7]) + (v * q[4])) + (w * q[1]),
                            ((u * q[8]) + (v * q[5])) + (w * q[2]));
        h.shn = leaving ? rtd::neg(s) : s;
    }
    h.m = mat_of(a.mats, a.tri_mat, a.kd, prim);
    h.tp.dst = h.dst;
    h.tp.dir = d;
    h.tp.n = h.shn;
    h.tp.has_vt = hit_texture_coords(a.tex, prim, hit.y, hit.z, &h.tp.vtu, &h.tp.vtv);
    if (a.tex.maps) {
        child_coefficients(h.m, a.tex, h.tp, &h.refl_c, &h.tran_c);
    } else {
        h.refl_c = h.m.rf;
        h.tran_c = h.m.kd_tran.w;
    }
    // ray::reflect / ray::refract spawn a ray while power x coefficient
    // stays above MIN_REFLECTIVE_POWER (rt/ray.cc:121-128, 200-206).
    const float power = a.power[i];
    h.refl = h.refl_c > 0.0f && static_cast<double>(power * h.refl_c) > kMinReflectivePower;
    h.refr = h.tran_c > 0.0f && static_cast<double>(power * h.tran_c) > kMinReflectivePower &&
             refract_dir(d, h.shn, h.m.ks_ri.w, leaving, &h.rdir);
    return h;
}

// The escape-time iterations of z -> z^2 + (x0, y0) from z = (x0, y0) while
// |z|^2 <= 100, at most limit (rt/materials/mandelbrot_shader.cc:66-73).
__device__ __forceinline__ int mandel_escape(float x0, float y0, int limit) {
    float x = x0, y = y0;
    int iteration = 0;
    while (((x * x) + (y * y) <= (10.0f * 10.0f)) && (iteration < limit)) {
        const float xtemp = ((x * x) - (y * y)) + x0;
        y = ((2.0f * x) * y) + y0;
        x = xtemp;
        ++iteration;
    }
    return iteration;
}

// mandelbrot_shader::shade's colour (rt/materials/mandelbrot_shader.cc:
// 41-120): fractal coordinates from the hit point on the plane the shading
// normal picks (exact +-x: y-z, exact +-z: x-y, else x-z); a point that does
// not escape is tried once more on shifted, 1.5x scaled coordinates and is
// black if it again does not; the escape count picks a band of the ramp.
// The light term multiplies in later.
__device__ __forceinline__ float4 mandel_colour(rtd::V3 dst, rtd::V3 n, int limit) {
    const bool nx = (n.x == 1.0f) || (n.x == -1.0f);
    const bool nz = (n.z == 1.0f) || (n.z == -1.0f);
    float x0, y0;
    if (nx) {
        x0 = fmodf(fabsf(dst.z), 2.4f) - 1.5f;
        y0 = fmodf(fabsf(dst.y) - 1.5f, 2.4f) - 1.0f;
    } else if (nz) {
        x0 = fmodf(fabsf(dst.x), 2.4f) - 1.5f;
        y0 = fmodf(fabsf(dst.y) - 1.5f, 2.4f) - 1.0f;
    } else {
        x0 = fmodf(fabsf(dst.x) - 1.5f, 2.4f) - 1.0f;
        y0 = fmodf(fabsf(dst.z), 2.4f) - 1.5f;
    }
    int iteration = mandel_escape(x0, y0, limit);
    if (iteration == limit) {
        if (nx) {
            x0 = fmodf(fabsf(dst.z) - 1.0f, 2.4f) * 1.5f - 0.4f;
            y0 = fmodf(fabsf(dst.y) - 2.0f, 2.4f) * 1.5f - 0.7f;
        } else if (nz) {
            x0 = fmodf(fabsf(dst.x) - 1.0f, 2.4f) * 1.5f - 0.4f;
            y0 = fmodf(fabsf(dst.y) - 2.0f, 2.4f) * 1.5f - 0.7f;
        } else {
            x0 = fmodf(fabsf(dst.x) - 2.0f, 2.4f) * 1.5f - 0.7f;
            y0 = fmodf(fabsf(dst.z) - 1.0f, 2.4f) * 1.5f - 0.4f;
        }
        iteration = mandel_escape(x0, y0, limit);
        if (iteration == limit) return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    const float colour = 767.0f - ((static_cast<float>(iteration) * 767.0f) / static_cast<float>(limit));
    if (colour < 256.0f) return make_float4(25.0f - colour, 255.0f, 255.0f, 0.0f);
    if (colour < 512.0f) return make_float4(0.0f, 256.0f - (colour - 256.0f), 255.0f, 0.0f);
    return make_float4(0.0f, 0.0f, 256.0f - (colour - 767.0f), 0.0f);
}

// The shading record of a hit, its material resolved at the hit:
// r0 {ka rgb, shininess}, r1 {kd rgb (mapper: base colour), mapper kd},
// r2 {ks rgb (mapper: ks, clrh, 0), reflected-colour weight},
// r3 {refracted-colour weight, clrf, kind, 0}.
struct ShadeRec {
    float4 r0, r1, r2, r3;
};

// coloured_mapper_shader::shade's colours and shininess
// (rt/materials/coloured_mapper_shader.cc:41-80) and mapper_shader::shade's
// colour, kd and s layers (rt/materials/mapper_shader.cc:84-126); phong and
// Cook-Torrance take their fields. The children's colours come back weighted
// by rf and tran, except mapper_shader's, weighted by its layered values.
__device__ __forceinline__ ShadeRec shade_record(const HitSurface &h, const TexTables &T) {
    const Mat &m = h.m;
    ShadeRec r;
    const float kind = static_cast<float>(m.kind);
    if (m.kind == kMandelbrot) {
        r.r0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        r.r1 = mandel_colour(h.dst, h.shn, m.iters);
        r.r2 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        r.r3 = make_float4(0.0f, 0.0f, kind, 0.0f);
        return r;
    }
    if (m.kind == kMapper) {
        float4 rgb = make_float4(m.kd_tran.x, m.kd_tran.y, m.kd_tran.z, 0.0f);
        const int first = m.layers[0] >> 8, count = m.layers[0] & 0xff;
        float alpha = 1.0f;
        TexColour c{rgb.x, rgb.y, rgb.z};
        for (int k = 0; k < count; ++k) {
            alpha *= texture_colour(T, T.layers[first + k], &c, h.tp);
            if (alpha == 0.0f) break;
        }
        const float kd = mono_layers(T, m.layers[1], m.mkd, h.tp);
        const float sh = mono_layers(T, m.layers[2], m.ka_s.w, h.tp);
        r.r0 = make_float4(0.0f, 0.0f, 0.0f, sh);
        r.r1 = make_float4(c.r, c.g, c.b, kd);
        r.r2 = make_float4(m.mks, m.clrh, 0.0f, h.refl_c);
        r.r3 = make_float4(h.tran_c, m.clrf, kind, 0.0f);
        return r;
    }
    float4 ka = m.ka_s, kd = m.kd_tran, ks = m.ks_ri;
    if (m.kind == kColoured && T.maps) {
        if (m.tex[3]) {
            TexColour dummy{ka.x, ka.y, ka.z};
            const float alpha = texture_colour(T, m.tex[3] - 1, &dummy, h.tp);
            ka.w = alpha * (m.ka_s.w * (1.0f / 255.0f));
        }
        ka = mapped_colour(T, m.tex[0], ka, h.tp);
        kd = mapped_colour(T, m.tex[1], kd, h.tp);
        ks = mapped_colour(T, m.tex[2], ks, h.tp);
    }
    r.r0 = make_float4(ka.x, ka.y, ka.z, ka.w);
    r.r1 = make_float4(kd.x, kd.y, kd.z, 0.0f);
    r.r2 = make_float4(ks.x, ks.y, ks.z, m.rf);
    r.r3 = make_float4(m.kd_tran.w, 0.0f, kind, 0.0f);
    return r;
}

// flags[i] = 1 for a hit, kids[i] = its number of children; entry n of both
// is 0 so that the exclusive scans over n + 1 entries end with the totals.
__global__ void __launch_bounds__(256) k_mat_classify(GenArgs a, uint32_t *__restrict__ flags,
                                                      uint32_t *__restrict__ kids) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i > a.n) return;
    uint32_t f = 0, k = 0;
    if (i < a.n && __float_as_int(a.hits[i].w) >= 0) {
        f = 1;
        const HitSurface h = surface_of(a, i);
        k = static_cast<uint32_t>(h.refl) + static_cast<uint32_t>(h.refr);
    }
    flags[i] = f;
    kids[i] = k;
}

struct EmitArgs {
    GenArgs g;
    const uint32_t *hit_pos, *kid_pos;  // exclusive scans of the flags and the child counts
    LightArgs lights;
    float4 *rec_n;      // per hit: shading normal, triangle (int bits)
    float4 *rec_d;      // per hit: incoming direction
    int2 *rec_kids;     // per hit: first child slot, children (bit 0 mirror, bit 1 refraction)
    ShadeRec *rec_m;    // per hit: the resolved material
    float4 *shadow;     // per hit x light: 2 float4
    float4 *next_rays;  // children, 2 float4 each
    float *next_power;
};

__global__ void __launch_bounds__(256) k_mat_emit(EmitArgs e) {
    const GenArgs &a = e.g;
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= a.n || __float_as_int(a.hits[i].w) < 0) return;
    const HitSurface h = surface_of(a, i);
    const int64_t q = e.hit_pos[i];
    const float4 r1 = a.rays[2 * i + 1];
    const uint32_t c0 = e.kid_pos[i];
    e.rec_n[q] = make_float4(h.shn.x, h.shn.y, h.shn.z, a.hits[i].w);
    e.rec_d[q] = make_float4(r1.x, r1.y, r1.z, 0.0f);
    e.rec_kids[q] = make_int2(static_cast<int32_t>(c0),
                              static_cast<int32_t>(h.refl) | (static_cast<int32_t>(h.refr) << 1));
    e.rec_m[q] = shade_record(h, a.tex);
    // Shadow rays and the mirror ray start off the front of the surface.
    const rtd::V3 org = rtd::offset_start_point(h.dst, h.geo, 1);
    for (int l = 0; l < e.lights.n; ++l) {
        rtd::V3 dn;
        float len;
        light_dir(org, e.lights.pos_d[l], &dn, &len);
        float4 *sr = e.shadow + 2 * (q * e.lights.n + l);
        sr[0] = make_float4(org.x, org.y, org.z, rtd::kEpsilon);
        sr[1] = make_float4(dn.x, dn.y, dn.z, len);
    }
    const float power = a.power[i];
    int64_t c = c0;
    if (h.refl) {
        const rtd::V3 rd = reflect_dir(rtd::f4v(r1), h.shn);
        e.next_rays[2 * c] = make_float4(org.x, org.y, org.z, rtd::kEpsilon);
        e.next_rays[2 * c + 1] = make_float4(rd.x, rd.y, rd.z, rtd::kMaxDist);
        e.next_power[c] = power * h.refl_c;
        ++c;
    }
    if (h.refr) {
        // The refraction ray starts off the back of the surface.
        const rtd::V3 o2 = rtd::offset_start_point(h.dst, h.geo, -1);
        e.next_rays[2 * c] = make_float4(o2.x, o2.y, o2.z, rtd::kEpsilon);
        e.next_rays[2 * c + 1] = make_float4(h.rdir.x, h.rdir.y, h.rdir.z, rtd::kMaxDist);
        e.next_power[c] = power * h.tran_c;
    }
}

struct LocalArgs {
    const float4 *rec_n, *rec_d;
    const ShadeRec *rec_m;
    const float4 *shadow;
    const uint8_t *occ;
    int64_t nhits;
    LightArgs lights;
    const float4 *kd, *mats;
    const int32_t *tri_mat;
    float4 *colour;  // per hit: local colour
};

// The material's shade without the secondary rays, with the light intensity
// of rt/light.h:44-59: phong_shader (rt/materials/phong_shader.cc:21-64) and
// coloured_mapper_shader (rt/materials/coloured_mapper_shader.cc:82-126),
// diffuse and specular per light, then the ambient term; cook_torrance_cxy
// (rt/materials/cook_torrance_cxy.cc:33-76) per light; mapper_shader
// (rt/materials/mapper_shader.cc:128-170) per light, no ambient term.
__global__ void __launch_bounds__(256) k_mat_local(LocalArgs a) {
    const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= a.nhits) return;
    const float4 nn = a.rec_n[q];
    const rtd::V3 n = rtd::f4v(nn), d = rtd::f4v(a.rec_d[q]);
    const ShadeRec rm = a.rec_m[q];
    const int kind = static_cast<int>(rm.r3.z);
    const bool ct = kind == kCookTorrance;
    const bool mapper = kind == kMapper;
    const bool mandel = kind == kMandelbrot;
    float total_shade = 0.0f;  // mandelbrot_shader: sum of n.l x visibility
    Mat m;
    if (ct) m = mat_of(a.mats, a.tri_mat, a.kd, __float_as_int(nn.w));
    const CieMatrix cm = cie_matrix();
    const float nv = rtd::dot(n, d);  // cook_torrance_cxy: one dot product for every light
    Rgb c{0.0f, 0.0f, 0.0f};
    for (int l = 0; l < a.lights.n; ++l) {
        const int64_t k = q * a.lights.n + l;
        const float4 sd = a.shadow[2 * k + 1];
        const rtd::V3 dn = rtd::f4v(sd);
        const float magn = (a.occ[k] ? 0.0f : 1.0f) / 1.0f;
        const float dist = sd.w * a.lights.pos_d[l].w;
        const float d_scale = (dist * dist) + 1.0f;
        const float4 lrgb = a.lights.rgb[l];
        if (mapper) {
            const float diff_shade = rtd::dot(dn, n) * rm.r1.w;
            if (diff_shade < 0.0f) continue;
            const float shade_x2 = 2.0f * diff_shade;
            const rtd::V3 ref_dir = rtd::mk(dn.x - n.x * shade_x2, dn.y - n.y * shade_x2, dn.z - n.z * shade_x2);
            const float rdr = rtd::dot(d, ref_dir);
            float spec = 0.0f;
            if (rdr > 0.0f) spec = powf(rdr, rm.r0.w) * rm.r2.x;
            const float lr = lrgb.x / d_scale, lg = lrgb.y / d_scale, lb = lrgb.z / d_scale;
            c.r += ((rm.r1.x * diff_shade) * lr) * magn;
            c.g += ((rm.r1.y * diff_shade) * lg) * magn;
            c.b += ((rm.r1.z * diff_shade) * lb) * magn;
            const float clrh = rm.r2.y, keep = 1.0f - clrh;
            c.r += (((clrh * rm.r1.x) + (keep * lr)) * spec) * magn;
            c.g += (((clrh * rm.r1.y) + (keep * lg)) * spec) * magn;
            c.b += (((clrh * rm.r1.z) + (keep * lb)) * spec) * magn;
            continue;
        }
        const float shade = rtd::dot(dn, n);
        if (shade < 0.0f) continue;
        if (mandel) {
            total_shade += shade * magn;
            continue;
        }
        if (ct) {
            const float4 li = make_float4(lrgb.x / d_scale, lrgb.y / d_scale, lrgb.z / d_scale, 0.0f);
            cook_torrance_term(m, cm, n, d, nv, dn, shade, magn, li, c);
            continue;
        }
        // The light ray reflected about the normal.
        const float shade_x2 = 2.0f * shade;
        const rtd::V3 ref_dir = rtd::mk(dn.x - n.x * shade_x2, dn.y - n.y * shade_x2, dn.z - n.z * shade_x2);
        const float rdr = rtd::dot(d, ref_dir);
        float scr = shade * rm.r1.x, scg = shade * rm.r1.y, scb = shade * rm.r1.z;
        if (rdr > 0.0f) {
            const float sp = powf(rdr, rm.r0.w);
            scr += sp * rm.r2.x;
            scg += sp * rm.r2.y;
            scb += sp * rm.r2.z;
        }
        c.r = c.r + ((lrgb.x / d_scale) * (scr * magn));
        c.g = c.g + ((lrgb.y / d_scale) * (scg * magn));
        c.b = c.b + ((lrgb.z / d_scale) * (scb * magn));
    }
    if (mandel) {  // mandelbrot_shader.cc:126: the colour times the light term
        c = Rgb{rm.r1.x * total_shade, rm.r1.y * total_shade, rm.r1.z * total_shade};
    } else if (!ct && !mapper) {  // cook_torrance_cxy and mapper_shader have no ambient term
        c.r += rm.r0.x;
        c.g += rm.r0.y;
        c.b += rm.r0.z;
    }
    a.colour[q] = make_float4(c.r, c.g, c.b, 0.0f);
}

struct GatherArgs {
    float4 *colour;  // this generation's hits: local colour in, full colour out
    const int2 *rec_kids;
    const ShadeRec *rec_m;
    int64_t nhits;
    const uint32_t *child_hit_pos;  // next generation: exclusive scan of its hit flags
    const float4 *child_colour;     // next generation: full colour of its hits
    float4 bg;                      // the camera's background colour
};

// The colour a child ray returns: its hit's colour, or the background.
__device__ __forceinline__ float4 child_colour(const GatherArgs &a, int64_t j) {
    const uint32_t k = a.child_hit_pos[j];
    if (a.child_hit_pos[j + 1] == k) return a.bg;
    return a.child_colour[k];
}

// rt/materials/phong_shader.cc:66-111: c += reflected * (rf / 1), then
// c += refracted * (tran / 1); mapper_shader (rt/materials/mapper_shader.cc:
// 172-184): c += reflected * refl, then c += clrf c + (1 - clrf) refracted
// trans.
__global__ void __launch_bounds__(256) k_mat_gather(GatherArgs a) {
    const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= a.nhits) return;
    const int2 kids = a.rec_kids[q];
    if (kids.y == 0) return;
    const ShadeRec rm = a.rec_m[q];
    float4 c = a.colour[q];
    int64_t j = kids.x;
    if (kids.y & 1) {
        const float4 p = child_colour(a, j++);
        const float w = rm.r2.w / 1.0f;
        c.x += p.x * w;
        c.y += p.y * w;
        c.z += p.z * w;
    }
    if (kids.y & 2) {
        const float4 p = child_colour(a, j);
        const float w = rm.r3.x / 1.0f;
        if (static_cast<int>(rm.r3.z) == kMapper) {
            const float clrf = rm.r3.y, keep = 1.0f - clrf;
            const float tx = (clrf * c.x) + ((keep * p.x) * w);
            const float ty = (clrf * c.y) + ((keep * p.y) * w);
            const float tz = (clrf * c.z) + ((keep * p.z) * w);
            c.x += tx;
            c.y += ty;
            c.z += tz;
        } else {
            c.x += p.x * w;
            c.y += p.y * w;
            c.z += p.z * w;
        }
    }
    a.colour[q] = c;
}

// Pixel = (sum of its samples' colours) / spp, samples in order.
__global__ void __launch_bounds__(256) k_mat_resolve(TileMap tm, int spp, int64_t npix,
                                                     const uint32_t *__restrict__ hit_pos,
                                                     const float4 *__restrict__ colour, float4 bg,
                                                     float *__restrict__ rgb) {
    const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= npix) return;
    const int x = static_cast<int>(p % tm.x_res), y = static_cast<int>(p / tm.x_res);
    const int64_t i0 = tm.slot_of_pixel(x, y) * spp;
    float r = 0.0f, g = 0.0f, b = 0.0f;
    for (int s = 0; s < spp; ++s) {
        const uint32_t k = hit_pos[i0 + s];
        const float4 v = (hit_pos[i0 + s + 1] == k) ? bg : colour[k];
        r += v.x;
        g += v.y;
        b += v.z;
    }
    const float inv = static_cast<float>(spp);
    rgb[3 * p + 0] = r / inv;
    rgb[3 * p + 1] = g / inv;
    rgb[3 * p + 2] = b / inv;
}

// Primary hits of every sample in pixel order (pixel * spp + sample).
__global__ void __launch_bounds__(256) k_mat_hits_out(TileMap tm, int spp, int64_t npix,
                                                      const float4 *__restrict__ hits, float4 *__restrict__ out) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= npix * spp) return;
    const int64_t p = k / spp;
    const int s = static_cast<int>(k % spp);
    const int x = static_cast<int>(p % tm.x_res), y = static_cast<int>(p / tm.x_res);
    out[k] = hits[tm.slot_of_pixel(x, y) * spp + s];
}

hipError_t scan_u32(uint32_t *data, int64_t n, uint32_t *tile_sums, hipStream_t s) {
    const int64_t blocks = (n + kScanTile - 1) / kScanTile;
    if (n <= 0 || blocks > kScanTile) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_scan_tiles, dim3(static_cast<unsigned>(blocks)), dim3(kScanThreads), 0, s, data, n,
                       tile_sums);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanThreads), 0, s, tile_sums, static_cast<int>(blocks));
    hipLaunchKernelGGL(k_scan_add, dim3(static_cast<unsigned>(blocks)), dim3(kScanThreads), 0, s, data, n,
                       tile_sums);
    return hipGetLastError();
}

dim3 grid_of(int64_t n) { return dim3(static_cast<unsigned>(std::max<int64_t>(1, (n + 255) / 256))); }

// What the backward pass needs of one generation.
struct Gen {
    int64_t n = 0, nhits = 0;
    uint32_t *hit_pos = nullptr;
    float4 *colour = nullptr, *rec_n = nullptr;
    int2 *rec_kids = nullptr;
    ShadeRec *rec_m = nullptr;
};

// The power of a ray falls by a factor of at least 0.95 (the largest rf and
// transmittance) per generation, so a chain from a primary ray of power 1
// stays above 0.1 for at most 45 generations.
constexpr int kMaxGenerations = 64;

}  // namespace

int32_t render_frame_materials(rt_accel *a, const rt_camera &cam, const rt_light *lights, int nlights,
                               const rt_render_params &p, float *rgb, rt_hit *primary, rt_stats *stats) {
    rt_scene *sc = a->scene;
    rt_ctx *ctx = sc->ctx;
    hipStream_t s = ctx->stream;
    if (p.bounce) return fail(ctx, RT_ERR_UNSUPPORTED, "rt_render: the diffuse bounce is not defined with materials");
    FramePlan f;
    int32_t rc = plan_frame(a, cam, lights, nlights, p, &f);
    if (rc) return rc;
    const int nl = f.lights.n;
    const int64_t ns = f.tiles * 64 * f.spp;
    if (ns + 1 > int64_t(kScanTile) * kScanTile) return fail(ctx, RT_ERR_INVALID, "rt_render: frame too large");
    const float4 background = make_float4(p.background[0], p.background[1], p.background[2], 0.0f);
    const TexTables tex{sc->d_maps, sc->d_texels, sc->d_layers, sc->d_vt};
    BuildArena &ar = ctx->arena;
    RT_HIP_CHECK(ctx, ar.begin());
    auto take = [&](auto **ptr, int64_t count) -> int32_t {
        void *v = nullptr;
        const hipError_t e = ar.take(&v, static_cast<size_t>(std::max<int64_t>(1, count)) * sizeof(**ptr));
        if (e != hipSuccess) return fail(ctx, RT_ERR_OOM, std::string("rt_render: ") + hipGetErrorString(e));
        *ptr = reinterpret_cast<std::remove_reference_t<decltype(*ptr)>>(v);
        return RT_OK;
    };
    float4 *rays = nullptr;
    float *power = nullptr;
    uint32_t *sums = nullptr;
    if ((rc = take(&rays, 2 * ns)) || (rc = take(&power, ns)) || (rc = take(&sums, kScanTile))) return rc;
    const bool timed = stats != nullptr;
    if (timed) RT_HIP_CHECK(ctx, hipEventRecord(ctx->ev[0], s));
    hipLaunchKernelGGL(k_mat_raygen, grid_of(ns), dim3(256), 0, s, f.cam, f.tm, f.spp, p.seed, ns, rays, power);
    uint64_t ray_count[4] = {static_cast<uint64_t>(f.npix * f.spp), 0, 0, 0};
    float trace_ms[2] = {0.0f, 0.0f};  // closest hit and any hit, summed over generations
    std::vector<Gen> gens;
    float4 *primary_hits = nullptr;
    for (int64_t n = ns; n > 0;) {
        if (static_cast<int>(gens.size()) == kMaxGenerations)
            return fail(ctx, RT_ERR_INVALID, "rt_render: secondary rays did not die out");
        const bool first = gens.empty();
        Gen gen;
        gen.n = n;
        float4 *hits = nullptr;
        uint32_t *kids = nullptr;
        if ((rc = take(&hits, n)) || (rc = take(&gen.hit_pos, n + 1)) || (rc = take(&kids, n + 1))) return rc;
        if (first) primary_hits = hits;
        if (timed) RT_HIP_CHECK(ctx, hipEventRecord(ctx->ev[1], s));
        ctx->count_slot = first ? 0 : 2;  // node-visit counters: primary or secondary closest hit
        if ((rc = launch_closest(a, reinterpret_cast<const rt_ray *>(rays), n, reinterpret_cast<rt_hit *>(hits), s,
                                 first ? 0 : 2)))
            return rc;
        if (timed) RT_HIP_CHECK(ctx, hipEventRecord(ctx->ev[2], s));
        const GenArgs g{rays,      power,          hits,     n,        sc->d_n, sc->d_c,
                        sc->d_kd,  sc->d_mat,      sc->d_tri_mat, sc->d_vn, tex};
        hipLaunchKernelGGL(k_mat_classify, grid_of(n + 1), dim3(256), 0, s, g, gen.hit_pos, kids);
        RT_HIP_CHECK(ctx, scan_u32(gen.hit_pos, n + 1, sums, s));
        RT_HIP_CHECK(ctx, scan_u32(kids, n + 1, sums, s));
        uint32_t tot[2];
        RT_HIP_CHECK(ctx, hipMemcpyAsync(&tot[0], gen.hit_pos + n, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        RT_HIP_CHECK(ctx, hipMemcpyAsync(&tot[1], kids + n, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        RT_HIP_CHECK(ctx, hipStreamSynchronize(s));
        if (timed) {
            float ms = 0.0f;
            RT_HIP_CHECK(ctx, hipEventElapsedTime(&ms, ctx->ev[1], ctx->ev[2]));
            trace_ms[0] += ms;
        }
        const int64_t nh = tot[0], nk = tot[1];
        gen.nhits = nh;
        if (!first) ray_count[2] += static_cast<uint64_t>(n);
        ray_count[first ? 1 : 3] += static_cast<uint64_t>(nh * nl);
        float4 *next_rays = nullptr;
        float *next_power = nullptr;
        if (nh > 0) {
            float4 *rec_d = nullptr, *shadow = nullptr;
            uint8_t *occ = nullptr;
            if ((rc = take(&gen.rec_n, nh)) || (rc = take(&rec_d, nh)) || (rc = take(&gen.rec_kids, nh)) ||
                (rc = take(&gen.rec_m, nh)) ||
                (rc = take(&gen.colour, nh)) || (rc = take(&shadow, 2 * nh * nl)) || (rc = take(&occ, nh * nl)) ||
                (rc = take(&next_rays, 2 * nk)) || (rc = take(&next_power, nk)))
                return rc;
            const EmitArgs e{g,      gen.hit_pos, kids,      f.lights,  gen.rec_n, rec_d,
                             gen.rec_kids, gen.rec_m, shadow, next_rays, next_power};
            hipLaunchKernelGGL(k_mat_emit, grid_of(n), dim3(256), 0, s, e);
            if (nl > 0) {
                if (timed) RT_HIP_CHECK(ctx, hipEventRecord(ctx->ev[3], s));
                ctx->count_slot = first ? 1 : 3;
                if ((rc = launch_occluded(a, reinterpret_cast<const rt_ray *>(shadow), nh * nl, occ, s,
                                          first ? 1 : 3)))
                    return rc;
                if (timed) RT_HIP_CHECK(ctx, hipEventRecord(ctx->ev[4], s));
            }
            const LocalArgs la{gen.rec_n, rec_d,     gen.rec_m, shadow,        occ,       nh,
                               f.lights,  sc->d_kd, sc->d_mat, sc->d_tri_mat, gen.colour};
            hipLaunchKernelGGL(k_mat_local, grid_of(nh), dim3(256), 0, s, la);
            RT_HIP_CHECK(ctx, hipGetLastError());
            if (timed && nl > 0) {
                RT_HIP_CHECK(ctx, hipEventSynchronize(ctx->ev[4]));
                float ms = 0.0f;
                RT_HIP_CHECK(ctx, hipEventElapsedTime(&ms, ctx->ev[3], ctx->ev[4]));
                trace_ms[1] += ms;
            }
        }
        gens.push_back(gen);
        rays = next_rays;
        power = next_power;
        n = (nh > 0) ? nk : 0;
    }
    // Fold every generation's colours into its parents, deepest first.
    for (int k = static_cast<int>(gens.size()) - 2; k >= 0; --k) {
        const Gen &g = gens[k], &c = gens[k + 1];
        if (g.nhits == 0) continue;
        const GatherArgs ga{g.colour, g.rec_kids, g.rec_m, g.nhits, c.hit_pos, c.colour, background};
        hipLaunchKernelGGL(k_mat_gather, grid_of(g.nhits), dim3(256), 0, s, ga);
    }
    const bool dev_out = (p.flags & RT_DEVICE_PTRS) != 0;
    float *rgb_dev = nullptr;
    if (rgb && dev_out) {
        rgb_dev = rgb;
    } else if ((rc = take(&rgb_dev, 3 * f.npix))) {
        return rc;
    }
    hipLaunchKernelGGL(k_mat_resolve, grid_of(f.npix), dim3(256), 0, s, f.tm, f.spp, f.npix, gens[0].hit_pos,
                       gens[0].colour, background, rgb_dev);
    RT_HIP_CHECK(ctx, hipGetLastError());
    if (timed) RT_HIP_CHECK(ctx, hipEventRecord(ctx->ev[5], s));
    if (rgb && !dev_out)
        RT_HIP_CHECK(ctx, hipMemcpyAsync(rgb, rgb_dev, 3 * f.npix * sizeof(float), hipMemcpyDeviceToHost, s));
    if (primary) {
        float4 *hout = nullptr;
        if ((rc = take(&hout, f.npix * f.spp))) return rc;
        hipLaunchKernelGGL(k_mat_hits_out, grid_of(f.npix * f.spp), dim3(256), 0, s, f.tm, f.spp, f.npix,
                           primary_hits, hout);
        RT_HIP_CHECK(ctx, hipMemcpyAsync(primary, hout, f.npix * f.spp * sizeof(rt_hit),
                                         dev_out ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
    }
    RT_HIP_CHECK(ctx, hipGetLastError());
    RT_HIP_CHECK(ctx, hipStreamSynchronize(s));
    if (stats) {
        for (int k = 0; k < 4; ++k) stats->rays[k] = ray_count[k];
        float total = 0.0f;
        RT_HIP_CHECK(ctx, hipEventElapsedTime(&total, ctx->ev[0], ctx->ev[5]));
        for (auto &v : stats->ms) v = 0.0f;
        stats->ms[0] = total;
        stats->ms[1] = trace_ms[0];  // closest hit of every generation
        stats->ms[2] = trace_ms[1];  // any hit of every generation
        stats->ms[5] = total - trace_ms[0] - trace_ms[1];
    }
    return RT_OK;
}

}  // namespace rt
