// Device-side arithmetic of the ray-tracing hot path.
//
// The library is compiled with -ffp-contract=off, IEEE division and square
// root (-fhip-fp32-correctly-rounded-divide-sqrt) and without denormal
// flushing, so each float expression here is the same sequence of binary32
// operations as the reference's scalar path: hit ids, t, u and v come out
// bit-identical to rt/triangle.h:190-235 on the same ray bits.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rtd {

struct V3 {
    float x, y, z;
};

__host__ __device__ inline V3 mk(float x, float y, float z) { return V3{x, y, z}; }
__host__ __device__ inline V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ inline V3 add(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ inline V3 mul(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
__host__ __device__ inline V3 neg(V3 a) { return V3{-a.x, -a.y, -a.z}; }
__host__ __device__ inline V3 divs(V3 a, float s) { return V3{a.x / s, a.y / s, a.z / s}; }
// common/point_t.h:498-525 operand order.
__host__ __device__ inline V3 cross(V3 a, V3 b) {
    return V3{(a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x)};
}
__host__ __device__ inline float dot(V3 a, V3 b) { return ((a.x * b.x) + (a.y * b.y)) + (a.z * b.z); }
__host__ __device__ inline float magnitude(V3 a) { return sqrtf(((a.x * a.x) + (a.y * a.y)) + (a.z * a.z)); }
// common/point_t.h:553-568 under EXACT_NORMALISE.
__host__ __device__ inline V3 normalise_exact(V3 a) {
    const float dist = sqrtf(((a.x * a.x) + (a.y * a.y)) + (a.z * a.z));
    return V3{a.x / dist, a.y / dist, a.z / dist};
}
__host__ __device__ inline V3 f4v(float4 f) { return V3{f.x, f.y, f.z}; }

constexpr float kEpsilon = 1.0e-3f;        // common/common.cc:14
constexpr float kFpDelta = 32.0f;          // common/common.cc:15
constexpr float kFpDeltaSmall = 1.0e-4f;   // common/common.cc:16
constexpr float kMaxDist = 3.402823466e+38f;

// Moller-Trumbore in the operand order of rt/triangle.h:190-235, on a
// triangle stored as a, e1 = c - a, e2 = b - a. Returns true with (t, u, v)
// when the triangle accepts the ray with t > tmin.
__device__ __forceinline__ bool intersect_mt(V3 o, V3 d, V3 a, V3 e1, V3 e2, float tmin, float *t_out, float *u_out,
                                             float *v_out) {
    const V3 P = cross(d, e2);
    const float det = dot(e1, P);
    if (det > -0.000001f && det < 0.000001f) return false;
    const float inv_det = 1.0f / det;
    const V3 T = sub(o, a);
    const float u = dot(T, P) * inv_det;
    if (u < 0.0f) return false;
    const V3 Q = cross(T, e1);
    const float v = dot(d, Q) * inv_det;
    if ((v < 0.0f) || ((u + v) > 1.0f)) return false;
    const float t = dot(e2, Q) * inv_det;
    if (!(t > tmin)) return false;
    *t_out = t;
    *u_out = u;
    *v_out = v;
    return true;
}

// rt/ray.h:111-157: shadow-ray origin moved int(+-32 * n * d) ulps in float
// bit space; coordinates with |x| < 1e-4 become 0 + 1e-4 * n.
__device__ __forceinline__ V3 offset_start_point(V3 dst, V3 g, int dd) {
    const float df = static_cast<float>(dd);
    const uint32_t small = __float_as_uint(kFpDeltaSmall);
    const float p[3] = {dst.x, dst.y, dst.z};
    const float gn[3] = {g.x, g.y, g.z};
    float out[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        uint32_t bits = __float_as_uint(p[k]);
        if ((bits & 0x7fffffffu) < small) {
            out[k] = 0.0f + (kFpDeltaSmall * gn[k]);
        } else {
            const float off = (bits >> 31) ? -kFpDelta : kFpDelta;
            bits += static_cast<uint32_t>(static_cast<int>((off * gn[k]) * df));
            out[k] = __uint_as_float(bits);
        }
    }
    return V3{out[0], out[1], out[2]};
}

// ---- counter-based sampling for the diffuse bounce (project-defined) ----
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ uint64_t sample_key(uint64_t seed, uint32_t pixel, uint32_t sample, uint32_t bounce) {
    const uint64_t k0 = mix64(seed * 0x9E3779B97F4A7C15ull + pixel);
    return mix64(k0 + ((static_cast<uint64_t>(sample) << 8) | bounce) * 0xD1B54A32D192ED03ull);
}
__device__ __forceinline__ float unit24(uint32_t x) { return static_cast<float>(x) * (1.0f / 16777216.0f); }

// sin/cos of 2*pi*r, r in [0,1): quadrant reduction, Horner-evaluated Taylor
// series to x^15 / x^16 on [0, pi/2).
__device__ __forceinline__ void sincos_turns(float r, float *so, float *co) {
    const float x = r * 4.0f;
    const int q = static_cast<int>(x);
    const float f = x - static_cast<float>(q);
    const float th = f * 1.57079632679489662f;
    const float t2 = th * th;
    float s = -7.6471637318198164759e-13f;
    s = (s * t2) + 1.6059043836821614599e-10f;
    s = (s * t2) + -2.5052108385441718775e-8f;
    s = (s * t2) + 2.7557319223985890653e-6f;
    s = (s * t2) + -1.9841269841269841270e-4f;
    s = (s * t2) + 8.3333333333333333333e-3f;
    s = (s * t2) + -1.6666666666666666667e-1f;
    s = (s * t2) + 1.0f;
    s = s * th;
    float c = 4.7794773323873852974e-14f;
    c = (c * t2) + -1.1470745597729724714e-11f;
    c = (c * t2) + 2.0876756987868098979e-9f;
    c = (c * t2) + -2.7557319223985890653e-7f;
    c = (c * t2) + 2.4801587301587301587e-5f;
    c = (c * t2) + -1.3888888888888888889e-3f;
    c = (c * t2) + 4.1666666666666666667e-2f;
    c = (c * t2) + -0.5f;
    c = (c * t2) + 1.0f;
    const int qq = q & 3;
    *so = (qq == 0) ? s : ((qq == 1) ? c : ((qq == 2) ? -s : -c));
    *co = (qq == 0) ? c : ((qq == 1) ? -s : ((qq == 2) ? -c : s));
}

// Cosine-weighted direction about n (Duff et al. 2017 orthonormal basis).
__device__ __forceinline__ V3 cosine_direction(uint64_t key, V3 n) {
    const float r1 = unit24(static_cast<uint32_t>(key >> 40));
    const float r2 = unit24(static_cast<uint32_t>((key >> 16) & 0xFFFFFFu));
    float s, c;
    sincos_turns(r1, &s, &c);
    const float sr = sqrtf(r2);
    const float lx = sr * c;
    const float ly = sr * s;
    const float lz = sqrtf(1.0f - r2);
    const float sign = copysignf(1.0f, n.z);
    const float a = -1.0f / (sign + n.z);
    const float b = (n.x * n.y) * a;
    const V3 b1{1.0f + (((sign * n.x) * n.x) * a), sign * b, -sign * n.x};
    const V3 b2{b, sign + ((n.y * n.y) * a), -n.y};
    const V3 dv{((b1.x * lx) + (b2.x * ly)) + (n.x * lz), ((b1.y * lx) + (b2.y * ly)) + (n.y * lz),
                ((b1.z * lx) + (b2.z * ly)) + (n.z * lz)};
    return normalise_exact(dv);
}

}  // namespace rtd
