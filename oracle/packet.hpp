// Oracle — TEST INFRASTRUCTURE ONLY (see vec.hpp for the rules).
//
// SSE packet path of the reference (rt/packet_ray.h:159-194, four rays per
// packet). See packet.cpp.
#pragma once

#include <emmintrin.h>
#include <xmmintrin.h>

#include "oracle.hpp"

namespace oracle {

// rt/packet_ray.h:159-166 — origins and directions of four rays.
struct PacketRay {
    __m128 ox, oy, oz, dx, dy, dz;
};

// rt/packet_ray.h:171-194 plus the vint_t of hit triangles.
struct PacketHit {
    __m128 d = _mm_set1_ps(kMaxDist);
    __m128 u = _mm_setzero_ps();
    __m128 v = _mm_setzero_ps();
    __m128i id = _mm_set1_epi32(-1);
};

__m128 packet_inverse(__m128 x);        // common/simd.h:324-329
__m128 packet_inverse_sqrt(__m128 x);   // common/simd.h:332-337
__m128 packet_slab(const BvhNode &n, const PacketRay &r, const __m128 *i_rd);
void packet_intersect(const Triangle &tri, const PacketRay &r, PacketHit *h, int tri_idx);
void packet_nearest(const Scene &s, const PacketRay &r, PacketHit *h, Counters *c);
__m128 packet_nearer(const Scene &s, const PacketRay &r, __m128 t, Counters *c);
void packet_brute_nearest(const Scene &s, const PacketRay &r, PacketHit *h);
__m128 packet_brute_nearer(const Scene &s, const PacketRay &r, __m128 t);

}  // namespace oracle
