// Texture maps on the device: what texture_mapper::texture_map and
// texture_map_monochrome (rt/materials/texture_mapper.h) return for one hit,
// for the image mappers (planar, cubic, cylindrical: rt/materials/
// planar_mapper.cc, cubic_mapper.cc, cylindrical_mapper.cc), the procedural
// ones (checker_board_mapper.cc, perlin_noise_3d_mapper.cc with
// common/perlin_noise_3d.cc, gradient_mapper.cc) and the falloff
// (mapper_falloff.cc). Every float expression keeps the reference's operand
// order.
#pragma once

#include <cfloat>
#include <climits>
#include <cstdint>

#include "../../include/raptor_rt.h"
#include "rt_device.hpp"

namespace rt {

// A map as uploaded: the caller's state plus its image's shape and texels.
struct DevTexMap {
    rt_texture_map m;
    int32_t w, h, cpp, pad;
    int64_t offset;  // first texel in the scene's texel array
};

struct TexTables {
    const DevTexMap *maps;
    const float *texels;
    const int32_t *layers;
    const float *vt;  // 6 floats per triangle, or null
};

struct TexColour {
    float r, g, b;
};

// Where a hit samples its maps: the (re-projected) hit point, the incident
// ray's direction, the shading normal and the interpolated texture
// coordinates, if the triangle has them.
struct TexHit {
    rtd::V3 dst, dir, n;
    bool has_vt;
    float vtu, vtv;
};

namespace tex {

constexpr float kPi = 3.1415926535897932384626433f;  // common/common.cc:10

// static_cast<int>(float) as x86's cvttss2si gives it: truncation, and
// INT_MIN for NaN and values outside int's range (a degenerate mapper,
// e.g. one of size 0 used without texture coordinates).
__device__ __forceinline__ int to_int(float x) {
    return (x >= -2147483648.0f && x < 2147483648.0f) ? static_cast<int>(x) : INT_MIN;
}

// apply_wrapping_mode (texture_mapper.h): false for a texel outside a
// blank-wrapped image.
__device__ __forceinline__ bool wrap(int *c, int s, int mode) {
    int v = *c;
    switch (mode) {
        case RT_WRAP_BLANK:
            if ((v < 0) || (v >= s)) return false;
            break;
        case RT_WRAP_CLAMP:
            if (v < 0) {
                v = 0;
            } else if (v >= s) {
                v = s - 1;
            }
            break;
        case RT_WRAP_TILE:
            if (v < 0) {
                v = v % (s - 1);
                v += s - 1;
            } else if (v >= s) {
                v = v % s;
            }
            break;
        default:  // RT_WRAP_MIRROR
            if (v < 0) {
                v = v % ((s - 1) << 1);
                v += s - 1;
                v = abs(v);
            } else if (v >= s) {
                v = v % ((s - 1) << 1);
                v -= s - 1;
                v = abs(v);
            }
            break;
    }
    *c = v;
    return true;
}

// Bilinear texel lookup (image_texture_mapper::texel_lookup): the upper
// neighbours clamped to the image, weights from the fractions of u_co and
// v_co, rgb of the first three channels or grey for one.
__device__ __forceinline__ TexColour texel_lookup(const DevTexMap &t, const float *texels, float u_co, float v_co,
                                                  int u0, int v0) {
    const int u1 = min(u0 + 1, t.w - 1);
    const int v1 = min(v0 + 1, t.h - 1);
    const float fu = u_co - floorf(u_co);
    const float fv = v_co - floorf(v_co);
    const float w0 = (1.0f - fu) * (1.0f - fv);
    const float w1 = fu * (1.0f - fv);
    const float w2 = fv * (1.0f - fu);
    const float w3 = fu * fv;
    const float *img = texels + t.offset;
    const int64_t a0 = static_cast<int64_t>(u0) + static_cast<int64_t>(v0) * t.w;
    const int64_t a1 = static_cast<int64_t>(u1) + static_cast<int64_t>(v0) * t.w;
    const int64_t a2 = static_cast<int64_t>(u0) + static_cast<int64_t>(v1) * t.w;
    const int64_t a3 = static_cast<int64_t>(u1) + static_cast<int64_t>(v1) * t.w;
    if (t.cpp != 1) {
        const float *c0 = img + a0 * t.cpp, *c1 = img + a1 * t.cpp, *c2 = img + a2 * t.cpp, *c3 = img + a3 * t.cpp;
        return TexColour{(((c0[0] * w0) + (c1[0] * w1)) + (c2[0] * w2)) + (c3[0] * w3),
                         (((c0[1] * w0) + (c1[1] * w1)) + (c2[1] * w2)) + (c3[1] * w3),
                         (((c0[2] * w0) + (c1[2] * w1)) + (c2[2] * w2)) + (c3[2] * w3)};
    }
    const float g = (((img[a0] * w0) + (img[a1] * w1)) + (img[a2] * w2)) + (img[a3] * w3);
    return TexColour{g, g, g};
}

__device__ __forceinline__ rtd::V3 v3(const float *p) { return rtd::mk(p[0], p[1], p[2]); }
__device__ __forceinline__ rtd::V3 cmul(rtd::V3 a, rtd::V3 b) { return rtd::mk(a.x * b.x, a.y * b.y, a.z * b.z); }

// The image mappers' texture_coordinates of hit point dst.
__device__ __forceinline__ void coordinates(const DevTexMap &t, rtd::V3 dst, float *u_co, float *v_co) {
    const rtd::V3 c = v3(t.m.c), s = v3(t.m.s), u = v3(t.m.u), v = v3(t.m.v);
    const float wf = static_cast<float>(static_cast<unsigned>(t.w)), hf = static_cast<float>(static_cast<unsigned>(t.h));
    if (t.m.kind == RT_MAP_CYLINDRICAL) {
        const rtd::V3 dist = rtd::sub(dst, c);
        const float h_off = rtd::dot(dist, v);
        const float v_len = rtd::dot(s, v);
        *v_co = (h_off + (0.5f * v_len)) * (hf / v_len);
        const rtd::V3 height_align = rtd::add(c, rtd::mul(v, h_off));
        const float cosang = rtd::dot(rtd::normalise_exact(rtd::sub(dst, height_align)), u);
        const float theta = acosf(fmaxf(-1.0f, fminf(1.0f, cosang)));
        const float phase = theta * (0.5f / kPi);
        *u_co = ((static_cast<float>(t.m.cycles) * phase) * wf) / rtd::dot(s, u);
        return;
    }
    const rtd::V3 dist = rtd::add(rtd::sub(dst, c), rtd::mul(rtd::add(cmul(s, u), cmul(s, v)), 0.5f));
    float dist_u = rtd::dot(dist, u);
    if (t.m.kind == RT_MAP_CUBIC) dist_u = dist_u + rtd::dot(dist, v3(t.m.n));
    const float dist_v = rtd::dot(dist, v);
    *u_co = dist_u * (wf / fabsf(rtd::dot(s, u)));
    *v_co = dist_v * (hf / fabsf(rtd::dot(s, v)));
}

// mapper_falloff::falloff of a colour at dst.
__device__ __forceinline__ void falloff(const DevTexMap &t, TexColour *c, rtd::V3 dst) {
    const rtd::V3 cnt = v3(t.m.falloff_centre), grad = v3(t.m.falloff_grad);
    switch (t.m.falloff) {
        case RT_FALLOFF_CUBIC: {
            const rtd::V3 fall = cmul(rtd::sub(dst, cnt), grad);
            const float all_fall = ((fabsf(fall.x) + fabsf(fall.y)) + fabsf(fall.z)) * 255.0f;
            c->r = fmaxf(0.0f, c->r - all_fall);
            c->g = fmaxf(0.0f, c->g - all_fall);
            c->b = fmaxf(0.0f, c->b - all_fall);
            break;
        }
        case RT_FALLOFF_SPHERICAL: {
            const float dist = rtd::magnitude(rtd::sub(dst, cnt));
            const rtd::V3 fall = rtd::mk(dist * grad.x, dist * grad.y, dist * grad.z);
            const float all_fall =
                (fmaxf(0.0f, 1.0f - fall.x) * fmaxf(0.0f, 1.0f - fall.y)) * fmaxf(0.0f, 1.0f - fall.z);
            c->r *= all_fall;
            c->g *= all_fall;
            c->b *= all_fall;
            break;
        }
        case RT_FALLOFF_LINEAR_X:
        case RT_FALLOFF_LINEAR_Y:
        case RT_FALLOFF_LINEAR_Z: {
            const int a = t.m.falloff - RT_FALLOFF_LINEAR_X;
            const float d = (a == 0 ? dst.x : (a == 1 ? dst.y : dst.z)) - (a == 0 ? cnt.x : (a == 1 ? cnt.y : cnt.z));
            const float fall = d * (a == 0 ? grad.x : (a == 1 ? grad.y : grad.z));
            const float k = fmaxf(0.0f, 1.0f - fall);
            c->r *= k;
            c->g *= k;
            c->b *= k;
            break;
        }
        default:
            break;
    }
}

// checker_board_mapper::run_procedure: the colour where the x cell is odd
// (the z cell is xored with itself); the map's opacity.
__device__ __forceinline__ float checker(const DevTexMap &t, TexColour *c, rtd::V3 dst) {
    const float gx = (dst.x - t.m.c[0]) / t.m.s[0];
    const float gz = (dst.z - t.m.c[2]) / t.m.s[2];
    if ((to_int(gx) ^ to_int(gz) ^ to_int(gz)) & 0x1) *c = TexColour{t.m.rgb[0], t.m.rgb[1], t.m.rgb[2]};
    return t.m.opacity;
}

// perlin_noise_3d::noise (common/perlin_noise_3d.cc:10-15): the integer
// hash in 32-bit wrapping arithmetic (the reference's int overflow wraps the
// same way on x86), the masked value over 2^30 in float.
__device__ __forceinline__ float lattice_noise(uint32_t x, uint32_t y, uint32_t z, uint32_t s) {
    uint32_t n = ((x + y * 57u) + z * 13u) + s;
    n = (n << 13) ^ n;
    const int32_t m = static_cast<int32_t>((n * (n * n * 15731u + 789221u) + 1376312589u) & 0x7fffffffu);
    return 1.0f - static_cast<float>(m) / 1073741824.0f;
}

// perlin_noise_3d::smooth_noise (perlin_noise_3d.cc:26-32): the six face
// neighbours summed left to right times 1/16, plus the centre times 1/4.
__device__ __forceinline__ float smooth_noise(uint32_t x, uint32_t y, uint32_t z, uint32_t s) {
    const float sides = (((((lattice_noise(x - 1u, y, z, s) + lattice_noise(x + 1u, y, z, s)) +
                            lattice_noise(x, y - 1u, z, s)) +
                           lattice_noise(x, y + 1u, z, s)) +
                          lattice_noise(x, y, z - 1u, s)) +
                         lattice_noise(x, y, z + 1u, s)) *
                        (1.0f / 16.0f);
    const float center = lattice_noise(x, y, z, s) * (1.0f / 4.0f);
    return sides + center;
}

// perlin_noise_3d::interpolate (perlin_noise_3d.cc:41-46): cosine blend.
__device__ __forceinline__ float cos_blend(float a, float b, float x) {
    const float ft = x * kPi;
    const float f = (1.0f - cosf(ft)) * 0.5f;
    return (a * (1.0f - f)) + (b * f);
}

// perlin_noise_3d::interpolated_noise (perlin_noise_3d.cc:55-80): the
// lattice cell from int(float(int(x)) + 0.5) (which moves negative
// coordinates one cell up), smoothed noise at its 8 corners, blended along
// x, then y, then z.
__device__ __forceinline__ float interpolated_noise(float x, float y, float z, uint32_t s) {
    const int wx = to_int(static_cast<float>(to_int(x)) + 0.5f);
    const float px = x - static_cast<float>(wx);
    const int wy = to_int(static_cast<float>(to_int(y)) + 0.5f);
    const float py = y - static_cast<float>(wy);
    const int wz = to_int(static_cast<float>(to_int(z)) + 0.5f);
    const float pz = z - static_cast<float>(wz);
    const uint32_t ux = static_cast<uint32_t>(wx), uy = static_cast<uint32_t>(wy), uz = static_cast<uint32_t>(wz);
    const float v1 = smooth_noise(ux, uy, uz, s);
    const float v2 = smooth_noise(ux + 1u, uy, uz, s);
    const float v3 = smooth_noise(ux, uy + 1u, uz, s);
    const float v4 = smooth_noise(ux + 1u, uy + 1u, uz, s);
    const float v5 = smooth_noise(ux, uy, uz + 1u, s);
    const float v6 = smooth_noise(ux + 1u, uy, uz + 1u, s);
    const float v7 = smooth_noise(ux, uy + 1u, uz + 1u, s);
    const float v8 = smooth_noise(ux + 1u, uy + 1u, uz + 1u, s);
    const float ix0 = cos_blend(v1, v2, px);
    const float ix1 = cos_blend(v3, v4, px);
    const float ix2 = cos_blend(v5, v6, px);
    const float ix3 = cos_blend(v7, v8, px);
    const float iy0 = cos_blend(ix0, ix1, py);
    const float iy1 = cos_blend(ix2, ix3, py);
    return cos_blend(iy0, iy1, pz);
}

// perlin_noise_3d_mapper::run_procedure (rt/materials/perlin_noise_3d_mapper.cc:12-40).
__device__ __forceinline__ float perlin(const DevTexMap &t, TexColour *c, rtd::V3 dst) {
    float total = 0.0f;
    float frequency = 1.0f / t.m.zoom;
    float amplitude = 1.0f;
    const uint32_t seed = static_cast<uint32_t>(t.m.seed);
    for (int i = 0; i < t.m.octaves; ++i) {
        total += interpolated_noise(dst.x * frequency, dst.y * frequency, dst.z * frequency, seed) * amplitude;
        frequency *= 2.0f;
        amplitude *= t.m.persistence;
    }
    total = total + 1.0f;
    *c = TexColour{(t.m.rgb[0] * 0.5f) * total, (t.m.rgb[1] * 0.5f) * total, (t.m.rgb[2] * 0.5f) * total};
    return (total > 1.0f) ? 0.0f : t.m.opacity;
}

// gradient_mapper::run_procedure (rt/materials/gradient_mapper.cc:12-105)
// for the incidence angle: the keys' upper bound, the end values outside
// them, linear (alpha the lower key's: the reference's alpha difference is
// zero) or step between; the mapper's own inversion.
__device__ __forceinline__ float gradient(const DevTexMap &t, const float *texels, TexColour *c, rtd::V3 dir,
                                          rtd::V3 n) {
    const float key = acosf(fabsf(rtd::dot(n, dir))) * (180.0f / kPi);
    const float *k = texels + t.m.key_offset;
    const int count = t.m.key_count;
    int up = 0;
    while (up < count && !(key < k[6 * up])) ++up;
    float alpha;
    if (up == 0) {
        *c = TexColour{k[1], k[2], k[3]};
        alpha = k[4];
    } else if (up == count) {
        const float *b = k + 6 * (count - 1);
        *c = TexColour{b[1], b[2], b[3]};
        alpha = b[4];
    } else {
        const float *lo = k + 6 * (up - 1), *hi = k + 6 * up;
        if (static_cast<int>(hi[5]) == RT_GRAD_LINE) {
            const float key_weight = (key - lo[0]) / (hi[0] - lo[0]);
            *c = TexColour{lo[1] + ((hi[1] - lo[1]) * key_weight), lo[2] + ((hi[2] - lo[2]) * key_weight),
                           lo[3] + ((hi[3] - lo[3]) * key_weight)};
            const float alp_0 = lo[4];
            const float alp_dist = alp_0 - lo[4];
            alpha = alp_0 + (alp_dist * key_weight);
        } else {  // RT_GRAD_STEP (spline is refused at upload)
            *c = TexColour{hi[1], hi[2], hi[3]};
            alpha = hi[4];
        }
    }
    if (t.m.inverse) *c = TexColour{255.0f - c->r, 255.0f - c->g, 255.0f - c->b};
    return alpha;
}

// A procedural map's run_procedure: checker board, Perlin or gradient.
__device__ __forceinline__ float procedure(const DevTexMap &t, const float *texels, TexColour *c, const TexHit &p,
                                           rtd::V3 dst) {
    if (t.m.kind == RT_MAP_CHECKER_BOARD) return checker(t, c, dst);
    if (t.m.kind == RT_MAP_PERLIN_3D) return perlin(t, c, dst);
    return gradient(t, texels, c, p.dir, p.n);
}

__device__ __forceinline__ bool is_procedural(int kind) {
    return kind == RT_MAP_CHECKER_BOARD || kind == RT_MAP_PERLIN_3D || kind == RT_MAP_GRADIENT;
}

// The image mappers' texture address of a hit: the interpolated texture
// coordinates scaled to the image when the triangle has them, else the
// mapper's projection of dst.
__device__ __forceinline__ void address(const DevTexMap &t, rtd::V3 dst, bool has_vt, float vtu, float vtv,
                                        float *u_co, float *v_co) {
    if (has_vt) {
        *u_co = vtu * static_cast<float>(static_cast<unsigned>(t.w));
        *v_co = vtv * static_cast<float>(static_cast<unsigned>(t.h));
    } else {
        coordinates(t, dst, u_co, v_co);
    }
}

}  // namespace tex

// texture_mapper::texture_map: sample into *c (a checker board leaves it
// where it does not paint), then falloff and inverse; returns the alpha.
__device__ __forceinline__ float texture_colour(const TexTables &T, int map, TexColour *c, const TexHit &p) {
    const DevTexMap &t = T.maps[map];
    const rtd::V3 dst = p.dst;
    float alpha = 1.0f;
    if (tex::is_procedural(t.m.kind)) {
        alpha = tex::procedure(t, T.texels, c, p, dst);
    } else {
        float u_co, v_co;
        tex::address(t, dst, p.has_vt, p.vtu, p.vtv, &u_co, &v_co);
        int u0 = tex::to_int(u_co) + t.m.u_off;
        int v0 = tex::to_int(v_co) + t.m.v_off;
        if (!tex::wrap(&u0, t.w, t.m.u_wrap) || !tex::wrap(&v0, t.h, t.m.v_wrap)) {
            *c = TexColour{0.0f, 0.0f, 0.0f};
        } else {
            *c = tex::texel_lookup(t, T.texels, u_co, v_co, u0, v0);
        }
    }
    tex::falloff(t, c, dst);
    if (t.m.inverse) *c = TexColour{255.0f - c->r, 255.0f - c->g, 255.0f - c->b};
    return alpha;
}

// texture_mapper::texture_map_monochrome at offset (0, 0): the texel's mean
// (its grey for one channel), 0 outside a blank-wrapped image; then falloff
// and inverse; returns the alpha.
__device__ __forceinline__ float texture_mono(const TexTables &T, int map, float *z, const TexHit &p) {
    const DevTexMap &t = T.maps[map];
    const rtd::V3 dst = p.dst;
    float alpha = 1.0f;
    float pz;
    if (tex::is_procedural(t.m.kind)) {
        // procedural_texture_mapper::sample_texture_monochrome at offset
        // (0, 0): the point dst + u 0 + v 0 (which turns a -0 coordinate
        // into +0), the colour's mean.
        const rtd::V3 u = rtd::mk(p.n.y + p.n.z, 0.0f, p.n.x);
        const rtd::V3 v = rtd::mk(0.0f, -p.n.x - p.n.z, -p.n.y);
        const rtd::V3 q = rtd::add(rtd::add(dst, rtd::mul(u, 0.0f)), rtd::mul(v, 0.0f));
        TexColour c{0.0f, 0.0f, 0.0f};
        alpha = tex::procedure(t, T.texels, &c, p, q);
        pz = ((c.r + c.g) + c.b) * (1.0f / 3.0f);
    } else {
        float u_co, v_co;
        tex::address(t, dst, p.has_vt, p.vtu, p.vtv, &u_co, &v_co);
        u_co += 0.0f;
        v_co += 0.0f;
        int u0 = tex::to_int(u_co) + t.m.u_off;
        int v0 = tex::to_int(v_co) + t.m.v_off;
        if (!tex::wrap(&u0, t.w, t.m.u_wrap) || !tex::wrap(&v0, t.h, t.m.v_wrap)) {
            pz = 0.0f;
        } else {
            const TexColour c = tex::texel_lookup(t, T.texels, u_co, v_co, u0, v0);
            pz = (t.cpp != 1) ? ((c.r + c.g) + c.b) * (1.0f / 3.0f) : c.r;
        }
    }
    if (t.m.falloff != RT_FALLOFF_NONE) {
        TexColour c{pz, pz, pz};
        tex::falloff(t, &c, dst);
        pz = c.r;
    }
    if (t.m.inverse) pz = 255.0f - pz;
    *z = pz;
    return alpha;
}

// The texture coordinates of a hit on triangle prim at barycentrics (u, v):
// u vt_c + v vt_b + (1 - (u + v)) vt_a (rt/triangle.h:72-79); false when
// the triangle has none.
__device__ __forceinline__ bool hit_texture_coords(const TexTables &T, int32_t prim, float u, float v, float *vtu,
                                                   float *vtv) {
    if (!T.vt) return false;
    const float *q = T.vt + 6 * static_cast<int64_t>(prim);
    if (q[0] == FLT_MAX) return false;
    const float w = 1.0f - (u + v);
    *vtu = ((u * q[4]) + (v * q[2])) + (w * q[0]);
    *vtv = ((u * q[5]) + (v * q[3])) + (w * q[1]);
    return *vtu != FLT_MAX;
}

}  // namespace rt
