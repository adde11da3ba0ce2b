// Oracle — TEST INFRASTRUCTURE ONLY (see vec.hpp for the rules).
//
// Texture maps: texture_mapper::texture_map / texture_map_monochrome
// (rt/materials/texture_mapper.h), image_texture_mapper's sample_texture,
// sample_texture_monochrome and texel_lookup, apply_wrapping_mode, the
// planar, cubic and cylindrical texture_coordinates
// (rt/materials/planar_mapper.cc:8-17, cubic_mapper.cc:8-21,
// cylindrical_mapper.cc:8-20), checker_board_mapper::run_procedure
// (rt/materials/checker_board_mapper.cc:8-21), procedural_texture_mapper's
// samples and mapper_falloff::falloff (rt/materials/mapper_falloff.cc:9-61).
// The reference's operand order is kept in every float expression.
#pragma once

#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdlib>

#include "rt_oracle.h"
#include "vec.hpp"

namespace oracle_tex {

using oracle::V3;
using oracle::dot;
using oracle::magnitude;
using oracle::normalise_exact;
using oracle::ref_max;

const float kPi = 3.1415926535897932384626433f;  // common/common.cc:10

struct Rgb3 {
    float r = 0.0f, g = 0.0f, b = 0.0f;
};

struct Tables {
    const oracle_texture_map *maps = nullptr;
    const oracle_texture_image *images = nullptr;
    const float *texels = nullptr;
    const int32_t *layers = nullptr;
};

inline V3 v3(const float *p) { return V3(p[0], p[1], p[2]); }

// static_cast<int> of a float as the x86 build the reference runs on
// computes it (cvttss2si): truncation, INT_MIN for NaN and out of range.
inline int to_int(float x) { return (x >= -2147483648.0f && x < 2147483648.0f) ? static_cast<int>(x) : INT_MIN; }

// apply_wrapping_mode (texture_mapper.h).
inline bool apply_wrapping_mode(int *c, int s, int m) {
    switch (m) {
        case 0:  // blank
            if (((*c) < 0.0) || ((*c) >= s)) return false;
            break;
        case 1:  // clamp
            if ((*c) < 0) {
                (*c) = 0;
            } else if ((*c) >= s) {
                (*c) = s - 1;
            }
            break;
        case 2:  // tile
            if ((*c) < 0) {
                (*c) = (*c) % (s - 1);
                (*c) += s - 1;
            } else if ((*c) >= s) {
                (*c) = (*c) % s;
            }
            break;
        default:  // mirror
            if ((*c) < 0) {
                (*c) = (*c) % ((s - 1) << 1);
                (*c) += s - 1;
                (*c) = std::abs((*c));
            } else if ((*c) >= s) {
                (*c) = (*c) % ((s - 1) << 1);
                (*c) -= s - 1;
                (*c) = std::abs((*c));
            }
            break;
    }
    return true;
}

// image_texture_mapper::texel_lookup.
inline Rgb3 texel_lookup(const Tables &T, const oracle_texture_image &im, float u_co, float v_co, int u0, int v0) {
    const int u1 = std::min(u0 + 1, im.width - 1);
    const int v1 = std::min(v0 + 1, im.height - 1);
    const float fu = u_co - std::floor(u_co);
    const float fv = v_co - std::floor(v_co);
    const float w0 = (1.0f - fu) * (1.0f - fv);
    const float w1 = fu * (1.0f - fv);
    const float w2 = fv * (1.0f - fu);
    const float w3 = fu * fv;
    const float *img = T.texels + im.offset;
    const int64_t w = im.width;
    Rgb3 c;
    if (im.cpp != 1) {
        const float *c0 = img + (u0 + (v0 * w)) * im.cpp;
        const float *c1 = img + (u1 + (v0 * w)) * im.cpp;
        const float *c2 = img + (u0 + (v1 * w)) * im.cpp;
        const float *c3 = img + (u1 + (v1 * w)) * im.cpp;
        c.r = (c0[0] * w0) + (c1[0] * w1) + (c2[0] * w2) + (c3[0] * w3);
        c.g = (c0[1] * w0) + (c1[1] * w1) + (c2[1] * w2) + (c3[1] * w3);
        c.b = (c0[2] * w0) + (c1[2] * w1) + (c2[2] * w2) + (c3[2] * w3);
    } else {
        c.r = (img[u0 + (v0 * w)] * w0) + (img[u1 + (v0 * w)] * w1) + (img[u0 + (v1 * w)] * w2) +
              (img[u1 + (v1 * w)] * w3);
        c.g = c.r;
        c.b = c.r;
    }
    return c;
}

// The mappers' texture_coordinates.
inline void texture_coordinates(const oracle_texture_map &t, const oracle_texture_image &im, const V3 &dst,
                                float *u_co, float *v_co) {
    const V3 c = v3(t.c), s = v3(t.s), u = v3(t.u), v = v3(t.v);
    const float w = static_cast<float>(static_cast<unsigned>(im.width));
    const float h = static_cast<float>(static_cast<unsigned>(im.height));
    if (t.kind == 2) {  // cylindrical_mapper.cc:8-20
        const V3 dist(dst - c);
        const float h_off = dot(dist, v);
        const float v_len = dot(s, v);
        *v_co = (h_off + (0.5f * v_len)) * (h / v_len);
        const V3 height_align(c + (v * h_off));
        const float theta =
            std::acos(std::max(-1.0f, std::min(1.0f, dot(normalise_exact(dst - height_align), u))));
        const float phase = theta * (0.5f / kPi);
        *u_co = ((t.cycles * phase) * w) / dot(s, u);
        return;
    }
    const V3 dist((dst - c) + (((s * u) + (s * v)) * 0.5f));
    if (t.kind == 1) {  // cubic_mapper.cc:8-21
        const float dist_u = dot(dist, u) + dot(dist, v3(t.n));
        const float dist_v = dot(dist, v);
        *u_co = dist_u * (w / std::fabs(dot(s, u)));
        *v_co = dist_v * (h / std::fabs(dot(s, v)));
        return;
    }
    // planar_mapper.cc:8-17
    *u_co = (dot(dist, u)) * (w / std::fabs(dot(s, u)));
    *v_co = (dot(dist, v)) * (h / std::fabs(dot(s, v)));
}

// mapper_falloff::falloff.
inline void falloff(const oracle_texture_map &t, Rgb3 *c, const V3 &dst) {
    const V3 cnt = v3(t.falloff_centre), grad = v3(t.falloff_grad);
    switch (t.falloff) {
        case 0: {  // cubic
            const V3 dist(dst - cnt);
            const V3 fall(dist * grad);
            const float all_fall = (std::fabs(fall.x) + std::fabs(fall.y) + std::fabs(fall.z)) * 255.0f;
            c->r = ref_max(0.0f, c->r - all_fall);
            c->g = ref_max(0.0f, c->g - all_fall);
            c->b = ref_max(0.0f, c->b - all_fall);
            break;
        }
        case 1: {  // spherical
            const float dist = magnitude(dst - cnt);
            const V3 fall(dist * grad);
            const float all_fall =
                ref_max(0.0f, 1.0f - fall.x) * ref_max(0.0f, 1.0f - fall.y) * ref_max(0.0f, 1.0f - fall.z);
            c->r *= all_fall;
            c->g *= all_fall;
            c->b *= all_fall;
            break;
        }
        case 2:
        case 3:
        case 4: {  // linear x, y, z
            const int a = t.falloff - 2;
            const float dist = dst[a] - cnt[a];
            const float fall = dist * grad[a];
            const float k = ref_max(0.0f, 1.0f - fall);
            c->r *= k;
            c->g *= k;
            c->b *= k;
            break;
        }
        default:
            break;
    }
}

// checker_board_mapper::run_procedure.
inline float run_procedure(const oracle_texture_map &t, Rgb3 *c, const V3 &dst) {
    const V3 d(dst - v3(t.c));
    const V3 grid_pos(d.x / t.s[0], d.y / t.s[1], d.z / t.s[2]);  // point_t / point_t, per component
    if ((to_int(grid_pos.x) ^ to_int(grid_pos.z) ^ to_int(grid_pos.z)) & 0x1) *c = Rgb3{t.rgb[0], t.rgb[1], t.rgb[2]};
    return t.opacity;
}

// The address an image mapper samples: the interpolated texture coordinates
// (vt.x != MAX_DIST) scaled to the image, else its projection of dst.
inline void address(const oracle_texture_map &t, const oracle_texture_image &im, const V3 &dst, bool has_vt,
                    float vtu, float vtv, float *u_co, float *v_co) {
    if (has_vt && vtu != FLT_MAX) {
        *u_co = vtu * static_cast<float>(static_cast<unsigned>(im.width));
        *v_co = vtv * static_cast<float>(static_cast<unsigned>(im.height));
    } else {
        texture_coordinates(t, im, dst, u_co, v_co);
    }
}

// texture_mapper::texture_map.
inline float texture_map(const Tables &T, int map, Rgb3 *c, const V3 &dst, bool has_vt, float vtu, float vtv) {
    const oracle_texture_map &t = T.maps[map];
    float opac = 1.0f;
    if (t.kind == 3) {
        opac = run_procedure(t, c, dst);
    } else {
        const oracle_texture_image &im = T.images[t.image];
        float u_co, v_co;
        address(t, im, dst, has_vt, vtu, vtv, &u_co, &v_co);
        int u0 = to_int(u_co) + t.u_off;
        int v0 = to_int(v_co) + t.v_off;
        if (!apply_wrapping_mode(&u0, im.width, t.u_wrap)) {
            *c = Rgb3{};
        } else if (!apply_wrapping_mode(&v0, im.height, t.v_wrap)) {
            *c = Rgb3{};
        } else {
            *c = texel_lookup(T, im, u_co, v_co, u0, v0);
        }
    }
    if (t.falloff != 5) falloff(t, c, dst);
    if (t.inverse) *c = Rgb3{255.0f - c->r, 255.0f - c->g, 255.0f - c->b};
    return opac;
}

// texture_mapper::texture_map_monochrome with offsets (0, 0); *z is p.z.
inline float texture_map_monochrome(const Tables &T, int map, float *z, const V3 &dst, bool has_vt, float vtu,
                                    float vtv) {
    const oracle_texture_map &t = T.maps[map];
    float opac = 1.0f;
    if (t.kind == 3) {
        // procedural_texture_mapper: dst + u 0 + v 0, the colour's mean.
        Rgb3 c;
        opac = run_procedure(t, &c, dst);
        *z = (c.r + c.g + c.b) * (1.0f / 3.0f);
    } else {
        const oracle_texture_image &im = T.images[t.image];
        float u_co, v_co;
        address(t, im, dst, has_vt, vtu, vtv, &u_co, &v_co);
        u_co += 0;
        v_co += 0;
        int u0 = to_int(u_co) + t.u_off;
        int v0 = to_int(v_co) + t.v_off;
        if (!apply_wrapping_mode(&u0, im.width, t.u_wrap) || !apply_wrapping_mode(&v0, im.height, t.v_wrap)) {
            *z = 0.0f;  // p = (0, x_off, y_off)
        } else {
            const Rgb3 c = texel_lookup(T, im, u_co, v_co, u0, v0);
            *z = (im.cpp != 1) ? (c.r + c.g + c.b) * (1.0f / 3.0f) : c.r;
        }
    }
    if (t.falloff != 5) {
        Rgb3 c{*z, *z, *z};
        falloff(t, &c, dst);
        *z = c.r;
    }
    if (t.inverse) *z = 255.0f - *z;
    return opac;
}

}  // namespace oracle_tex
