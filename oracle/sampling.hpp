// Oracle — TEST INFRASTRUCTURE ONLY (see vec.hpp for the rules).
//
// Project-defined randomness, written independently of the HIP version so
// that the two can be checked against each other and against the published
// SplitMix64 and PCG32 outputs. RAPTor itself has no diffuse bounce (its
// secondary rays are mirror reflect/refract, rt/ray.cc:118-281), so the bounce
// below is this project's definition:
//   key  = mix64(mix64(seed * 0x9E3779B97F4A7C15 + pixel)
//                + ((sample << 8) | bounce) * 0xD1B54A32D192ED03)
//   r1   = (key >> 40) * 2^-24,  r2 = ((key >> 16) & 0xFFFFFF) * 2^-24
//   (s, c) = sin/cos of 2*pi*r1 by quadrant reduction + Taylor polynomials
//   local = (sqrt(r2) * c, sqrt(r2) * s, sqrt(1 - r2))
//   basis = Duff et al. 2017 "Building an Orthonormal Basis, Revisited"
//   dir  = normalise(b1 * lx + b2 * ly + n * lz)
#pragma once

#include <cstdint>

#include "vec.hpp"

namespace oracle {

inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline uint64_t splitmix64_next(uint64_t *state) {
    *state += 0x9E3779B97F4A7C15ull;
    return mix64(*state);
}

// O'Neill's pcg32 (XSH RR 64/32).
struct Pcg32 {
    uint64_t state = 0, inc = 1;
    uint32_t next() {
        const uint64_t old = state;
        state = old * 6364136223846793005ull + inc;
        const uint32_t xorshifted = static_cast<uint32_t>(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = static_cast<uint32_t>(old >> 59u);
        return (xorshifted >> rot) | (xorshifted << ((0u - rot) & 31u));
    }
    void seed(uint64_t initstate, uint64_t initseq) {
        state = 0u;
        inc = (initseq << 1u) | 1u;
        next();
        state += initstate;
        next();
    }
};

inline uint64_t bounce_key(uint64_t seed, uint32_t pixel, uint32_t sample, uint32_t bounce) {
    const uint64_t k0 = mix64(seed * 0x9E3779B97F4A7C15ull + pixel);
    return mix64(k0 + ((static_cast<uint64_t>(sample) << 8) | bounce) * 0xD1B54A32D192ED03ull);
}

inline float unit24(uint32_t x24) { return static_cast<float>(x24) * (1.0f / 16777216.0f); }

// sin and cos of 2*pi*r for r in [0, 1).
inline void sincos_turns(float r, float *s_out, float *c_out) {
    const float x = r * 4.0f;
    const int q = static_cast<int>(x);
    const float f = x - static_cast<float>(q);
    const float th = f * 1.57079632679489662f;
    const float t2 = th * th;
    float s = -7.6471637318198164759e-13f;                 // -1/15!
    s = (s * t2) + 1.6059043836821614599e-10f;              // 1/13!
    s = (s * t2) + -2.5052108385441718775e-8f;              // -1/11!
    s = (s * t2) + 2.7557319223985890653e-6f;               // 1/9!
    s = (s * t2) + -1.9841269841269841270e-4f;              // -1/7!
    s = (s * t2) + 8.3333333333333333333e-3f;               // 1/5!
    s = (s * t2) + -1.6666666666666666667e-1f;              // -1/3!
    s = (s * t2) + 1.0f;
    s = s * th;
    float c = 4.7794773323873852974e-14f;                   // 1/16!
    c = (c * t2) + -1.1470745597729724714e-11f;             // -1/14!
    c = (c * t2) + 2.0876756987868098979e-9f;               // 1/12!
    c = (c * t2) + -2.7557319223985890653e-7f;              // -1/10!
    c = (c * t2) + 2.4801587301587301587e-5f;               // 1/8!
    c = (c * t2) + -1.3888888888888888889e-3f;              // -1/6!
    c = (c * t2) + 4.1666666666666666667e-2f;               // 1/4!
    c = (c * t2) + -0.5f;
    c = (c * t2) + 1.0f;
    switch (q & 3) {
        case 0: *s_out = s; *c_out = c; break;
        case 1: *s_out = c; *c_out = -s; break;
        case 2: *s_out = -s; *c_out = -c; break;
        default: *s_out = -c; *c_out = s; break;
    }
}

inline V3 cosine_direction(uint64_t key, const V3 &n) {
    const float r1 = unit24(static_cast<uint32_t>(key >> 40));
    const float r2 = unit24(static_cast<uint32_t>((key >> 16) & 0xFFFFFFu));
    float s, c;
    sincos_turns(r1, &s, &c);
    const float sr = std::sqrt(r2);
    const float lx = sr * c;
    const float ly = sr * s;
    const float lz = std::sqrt(1.0f - r2);
    const float sign = std::copysign(1.0f, n.z);
    const float a = -1.0f / (sign + n.z);
    const float b = (n.x * n.y) * a;
    const V3 b1(1.0f + (((sign * n.x) * n.x) * a), sign * b, -sign * n.x);
    const V3 b2(b, sign + ((n.y * n.y) * a), -n.y);
    const V3 d(((b1.x * lx) + (b2.x * ly)) + (n.x * lz), ((b1.y * lx) + (b2.y * ly)) + (n.y * lz),
               ((b1.z * lx) + (b2.z * ly)) + (n.z * lz));
    return normalise_exact(d);
}

// Jitter inside a pixel for sample s of spp > 1 (bounce index 255 keys it).
inline void pixel_jitter(uint64_t seed, uint32_t pixel, uint32_t sample, float *jx, float *jy) {
    const uint64_t k = bounce_key(seed, pixel, sample, 255u);
    *jx = unit24(static_cast<uint32_t>(k >> 40));
    *jy = unit24(static_cast<uint32_t>((k >> 16) & 0xFFFFFFu));
}

}  // namespace oracle
