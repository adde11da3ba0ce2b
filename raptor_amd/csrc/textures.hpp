// Texture maps on the device: what texture_mapper::texture_map and
// texture_map_monochrome (rt/materials/texture_mapper.h) return for one hit,
// for the image mappers (planar, cubic, cylindrical: rt/materials/
// planar_mapper.cc, cubic_mapper.cc, cylindrical_mapper.cc), the checker
// board (checker_board_mapper.cc) and the falloff (mapper_falloff.cc).
// Every float expression keeps the reference's operand order.
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
__device__ __forceinline__ float texture_colour(const TexTables &T, int map, TexColour *c, rtd::V3 dst, bool has_vt,
                                                float vtu, float vtv) {
    const DevTexMap &t = T.maps[map];
    float alpha = 1.0f;
    if (t.m.kind == RT_MAP_CHECKER_BOARD) {
        alpha = tex::checker(t, c, dst);
    } else {
        float u_co, v_co;
        tex::address(t, dst, has_vt, vtu, vtv, &u_co, &v_co);
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
__device__ __forceinline__ float texture_mono(const TexTables &T, int map, float *z, rtd::V3 dst, bool has_vt,
                                              float vtu, float vtv) {
    const DevTexMap &t = T.maps[map];
    float alpha = 1.0f;
    float pz;
    if (t.m.kind == RT_MAP_CHECKER_BOARD) {
        TexColour c{0.0f, 0.0f, 0.0f};
        alpha = tex::checker(t, &c, dst);
        pz = ((c.r + c.g) + c.b) * (1.0f / 3.0f);
    } else {
        float u_co, v_co;
        tex::address(t, dst, has_vt, vtu, vtv, &u_co, &v_co);
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
