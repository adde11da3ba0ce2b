// Oracle — TEST INFRASTRUCTURE ONLY.
//
// This directory is a CPU restatement of RAPTor's ray-tracing hot path. It is
// the checker for the HIP product path: only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg may load it. The product never links it.
//
// Build flags matter for bit-exactness: -O2 -msse4.2 -ffp-contract=off, no
// -ffast-math, no -march=native, so every float expression below is evaluated
// as written, one IEEE-754 binary32 operation at a time, with no FMA.
//
// vec.hpp restates the 3-vector arithmetic of common/point_t.h the hot path
// uses: component-wise +,-,*,/, the cross product and dot product with their
// exact operand order, magnitude, and the EXACT_NORMALISE normalise().
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>

namespace oracle {

// common/common.cc:14-17
constexpr float kEpsilon = 1.0e-3f;
constexpr float kFpDelta = 32.0f;
constexpr float kFpDeltaSmall = 1.0e-4f;
constexpr float kMaxDist = std::numeric_limits<float>::max();

// std::min / std::max as libstdc++ defines them: min(a,b) = (b < a) ? b : a,
// max(a,b) = (a < b) ? b : a. The NaN behaviour of the slab test depends on it.
inline float ref_min(float a, float b) { return (b < a) ? b : a; }
inline float ref_max(float a, float b) { return (a < b) ? b : a; }

struct V3 {
    float x, y, z;
    V3() : x(0.0f), y(0.0f), z(0.0f) {}  // point_t<> default ctor zero-fills (common/point_t.h)
    V3(float a, float b, float c) : x(a), y(b), z(c) {}
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};

inline V3 operator+(const V3 &a, const V3 &b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator-(const V3 &a, const V3 &b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator*(const V3 &a, const V3 &b) { return V3(a.x * b.x, a.y * b.y, a.z * b.z); }
inline V3 operator*(const V3 &a, float s) { return V3(a.x * s, a.y * s, a.z * s); }
inline V3 operator*(float s, const V3 &a) { return V3(s * a.x, s * a.y, s * a.z); }
inline V3 operator/(const V3 &a, float s) { return V3(a.x / s, a.y / s, a.z / s); }
// common/point_t.h: T / point_t<T>, used for the inverse ray direction.
inline V3 operator/(float s, const V3 &a) { return V3(s / a.x, s / a.y, s / a.z); }
inline V3 operator-(const V3 &a) { return V3(-a.x, -a.y, -a.z); }

inline V3 vmin(const V3 &a, const V3 &b) { return V3(ref_min(a.x, b.x), ref_min(a.y, b.y), ref_min(a.z, b.z)); }
inline V3 vmax(const V3 &a, const V3 &b) { return V3(ref_max(a.x, b.x), ref_max(a.y, b.y), ref_max(a.z, b.z)); }

// common/point_t.h cross_product: c = a x b, each component (p*q) - (r*s).
inline V3 cross(const V3 &a, const V3 &b) {
    return V3((a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x));
}
// common/point_t.h dot_product: ((a.x*b.x) + (a.y*b.y)) + (a.z*b.z).
inline float dot(const V3 &a, const V3 &b) { return ((a.x * b.x) + (a.y * b.y)) + (a.z * b.z); }
// common/point_t.h magnitude.
inline float magnitude(const V3 &a) { return std::sqrt(((a.x * a.x) + (a.y * a.y)) + (a.z * a.z)); }
// common/point_t.h normalise<float> under EXACT_NORMALISE: divide by the IEEE sqrt.
inline V3 normalise_exact(const V3 &a) {
    const float dist = std::sqrt(((a.x * a.x) + (a.y * a.y)) + (a.z * a.z));
    return V3(a.x / dist, a.y / dist, a.z / dist);
}

inline uint32_t f2u(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float u2f(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

}  // namespace oracle
