// Oracle — TEST INFRASTRUCTURE ONLY (see vec.hpp for the rules).
//
// Texture maps: texture_mapper::texture_map / texture_map_monochrome
// (rt/materials/texture_mapper.h), image_texture_mapper's sample_texture,
// sample_texture_monochrome and texel_lookup, apply_wrapping_mode, the
// planar, cubic and cylindrical texture_coordinates
// (rt/materials/planar_mapper.cc:8-17, cubic_mapper.cc:8-21,
// cylindrical_mapper.cc:8-20), checker_board_mapper::run_procedure
// (rt/materials/checker_board_mapper.cc:8-21), perlin_noise_3d_mapper
// (rt/materials/perlin_noise_3d_mapper.cc:12-40 over common/perlin_noise_3d.cc),
// gradient_mapper (rt/materials/gradient_mapper.cc:12-105),
// procedural_texture_mapper's samples and mapper_falloff::falloff
// (rt/materials/mapper_falloff.cc:9-61).
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

// perlin_noise_3d::noise (common/perlin_noise_3d.cc:10-15). The reference
// computes in int and its overflow wraps on x86; uint32_t wraps the same
// way without the undefined behaviour.
inline float perlin_noise(uint32_t x, uint32_t y, uint32_t z, uint32_t s) {
    uint32_t n = x + y * 57u + z * 13u + s;
    n = (n << 13) ^ n;
    const int32_t masked = static_cast<int32_t>((n * (n * n * 15731u + 789221u) + 1376312589u) & 0x7fffffffu);
    return static_cast<float>(1.0f - masked / 1073741824.0f);
}

// perlin_noise_3d::smooth_noise (perlin_noise_3d.cc:26-32).
inline float perlin_smooth_noise(uint32_t x, uint32_t y, uint32_t z, uint32_t s) {
    const float sides = (perlin_noise(x - 1, y, z, s) + perlin_noise(x + 1, y, z, s) + perlin_noise(x, y - 1, z, s) +
                         perlin_noise(x, y + 1, z, s) + perlin_noise(x, y, z - 1, s) + perlin_noise(x, y, z + 1, s)) *
                        (1.0f / 16.0f);
    const float center = perlin_noise(x, y, z, s) * (1.0f / 4.0f);
    return sides + center;
}

// perlin_noise_3d::interpolate (perlin_noise_3d.cc:41-46).
inline float perlin_interpolate(float a, float b, float x) {
    const float ft = x * kPi;
    const float f = (1.0f - std::cos(ft)) * 0.5f;
    return a * (1.0f - f) + b * f;
}

// perlin_noise_3d::interpolated_noise (perlin_noise_3d.cc:55-80).
inline float perlin_interpolated_noise(float x, float y, float z, uint32_t s) {
    const int wx = to_int(static_cast<float>(to_int(x)) + 0.5f);
    const float px = x - wx;
    const int wy = to_int(static_cast<float>(to_int(y)) + 0.5f);
    const float py = y - wy;
    const int wz = to_int(static_cast<float>(to_int(z)) + 0.5f);
    const float pz = z - wz;
    const uint32_t X = static_cast<uint32_t>(wx), Y = static_cast<uint32_t>(wy), Z = static_cast<uint32_t>(wz);
    const float v1 = perlin_smooth_noise(X, Y, Z, s);
    const float v2 = perlin_smooth_noise(X + 1, Y, Z, s);
    const float v3 = perlin_smooth_noise(X, Y + 1, Z, s);
    const float v4 = perlin_smooth_noise(X + 1, Y + 1, Z, s);
    const float v5 = perlin_smooth_noise(X, Y, Z + 1, s);
    const float v6 = perlin_smooth_noise(X + 1, Y, Z + 1, s);
    const float v7 = perlin_smooth_noise(X, Y + 1, Z + 1, s);
    const float v8 = perlin_smooth_noise(X + 1, Y + 1, Z + 1, s);
    const float ix0 = perlin_interpolate(v1, v2, px);
    const float ix1 = perlin_interpolate(v3, v4, px);
    const float ix2 = perlin_interpolate(v5, v6, px);
    const float ix3 = perlin_interpolate(v7, v8, px);
    const float iy0 = perlin_interpolate(ix0, ix1, py);
    const float iy1 = perlin_interpolate(ix2, ix3, py);
    return perlin_interpolate(iy0, iy1, pz);
}

// perlin_noise_3d_mapper::run_procedure; its constructor keeps rgb * 0.5.
inline float perlin_procedure(const oracle_texture_map &t, Rgb3 *c, const V3 &dst) {
    const Rgb3 rgb{t.rgb[0] * 0.5f, t.rgb[1] * 0.5f, t.rgb[2] * 0.5f};
    const float x_co = dst.x, y_co = dst.y, z_co = dst.z;
    float total = 0.0f;
    float frequency = 1.0f / t.zoom;
    float amplitude = 1.0f;
    for (int i = 0; i < t.octaves; ++i) {
        total += perlin_interpolated_noise(x_co * frequency, y_co * frequency, z_co * frequency,
                                           static_cast<uint32_t>(t.seed)) *
                 amplitude;
        frequency *= 2.0f;
        amplitude *= t.persistence;
    }
    total++;
    *c = Rgb3{rgb.r * total, rgb.g * total, rgb.b * total};
    if (total > 1.0f) return 0.0f;
    return t.opacity;
}

// gradient_mapper::run_procedure with grad_of incidence_angle; keys of six
// floats {key, r, g, b, alpha, interpolation (0 line, 2 step)}.
inline float gradient_procedure(const Tables &T, const oracle_texture_map &t, Rgb3 *c, const V3 &dir, const V3 &n) {
    const float key = std::acos(std::fabs(dot(n, dir))) * (180.0f / kPi);
    const float *K = T.texels + t.key_offset;
    const unsigned size = static_cast<unsigned>(t.key_count);
    // std::upper_bound over the keys.
    unsigned upper_idx = 0;
    {
        unsigned first = 0, count = size;
        while (count > 0) {
            const unsigned step = count / 2;
            const unsigned it = first + step;
            if (!(key < K[6 * it])) {
                first = it + 1;
                count -= step + 1;
            } else {
                count = step;
            }
        }
        upper_idx = first;
    }
    auto value = [&](unsigned i) { return Rgb3{K[6 * i + 1], K[6 * i + 2], K[6 * i + 3]}; };
    float alpha = 0.0f;
    if (upper_idx == 0) {
        *c = value(0);
        alpha = K[4];
    } else if (upper_idx == size) {
        *c = value(size - 1);
        alpha = K[6 * (size - 1) + 4];
    } else {
        const int interp = static_cast<int>(K[6 * upper_idx + 5]);
        if (interp == 0) {  // line
            const float key_weight = (key - K[6 * (upper_idx - 1)]) / (K[6 * upper_idx] - K[6 * (upper_idx - 1)]);
            const Rgb3 col_0 = value(upper_idx - 1), hi = value(upper_idx);
            const Rgb3 col_dist{hi.r - col_0.r, hi.g - col_0.g, hi.b - col_0.b};
            *c = Rgb3{col_0.r + (col_dist.r * key_weight), col_0.g + (col_dist.g * key_weight),
                      col_0.b + (col_dist.b * key_weight)};
            const float alp_0 = K[6 * (upper_idx - 1) + 4];
            const float alp_dist = alp_0 - K[6 * (upper_idx - 1) + 4];
            alpha = alp_0 + (alp_dist * key_weight);
        } else {  // step
            *c = value(upper_idx);
            alpha = K[6 * upper_idx + 4];
        }
    }
    if (t.inverse) *c = Rgb3{255.0f - c->r, 255.0f - c->g, 255.0f - c->b};
    return alpha;
}

// The run_procedure of a procedural map (kinds 3, 4, 5).
inline float procedure(const Tables &T, const oracle_texture_map &t, Rgb3 *c, const V3 &dst, const V3 &dir,
                       const V3 &n) {
    if (t.kind == 3) return run_procedure(t, c, dst);
    if (t.kind == 4) return perlin_procedure(t, c, dst);
    return gradient_procedure(T, t, c, dir, n);
}

inline bool procedural(int kind) { return kind >= 3 && kind <= 5; }

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

// texture_mapper::texture_map of a hit at dst on a ray of direction dir,
// shading normal n.
inline float texture_map(const Tables &T, int map, Rgb3 *c, const V3 &dst, const V3 &dir, const V3 &n, bool has_vt,
                         float vtu, float vtv) {
    const oracle_texture_map &t = T.maps[map];
    float opac = 1.0f;
    if (procedural(t.kind)) {
        opac = procedure(T, t, c, dst, dir, n);
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
inline float texture_map_monochrome(const Tables &T, int map, float *z, const V3 &dst, const V3 &dir, const V3 &n,
                                    bool has_vt, float vtu, float vtv) {
    const oracle_texture_map &t = T.maps[map];
    float opac = 1.0f;
    if (procedural(t.kind)) {
        // procedural_texture_mapper::sample_texture_monochrome
        // (texture_mapper.h:316-334) at offsets 0: dst + u 0 + v 0, the
        // colour's mean.
        const float x_off_dist = 0 * 0.1f, y_off_dist = 0 * 0.1f;
        const V3 u(n.y + n.z, 0.0f, n.x);
        const V3 v(0.0f, -n.x - n.z, -n.y);
        const V3 q(dst + (u * x_off_dist) + (v * y_off_dist));
        Rgb3 c;
        opac = procedure(T, t, &c, q, dir, n);
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
