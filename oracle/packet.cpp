// Oracle — TEST INFRASTRUCTURE ONLY (see vec.hpp for the rules).
//
// The reference's SSE packet path (#ifdef SIMD_PACKET_TRACING): four rays per
// packet, restated with the same SSE operations so that its results follow
// the reference's packet semantics rather than the scalar ones:
//   packet slab test     ssd/bvh_node.h:101-132 (minps / maxps NaN rules)
//   packet MT            rt/triangle.h:242-305: inverse(det) = rcpps + one
//                        Newton step (common/simd.h:324-329), no |det|
//                        reject, accepted when 0 < t < best and u, v, u + v
//                        in range (no EPSILON)
//   packet traversal     ssd/bvh.cc:341-500: the left child is always "near";
//                        the right one is pushed when both are needed; no
//                        root pre-test and no 100 * EPSILON slack
//   packet leaf loops    ssd/bvh_node.h:193-258
// This path is not bit-portable (rcpps/rsqrtps differ between CPU vendors);
// it times the CPU baseline. The scalar path is the parity oracle.
#include <emmintrin.h>
#include <xmmintrin.h>

#include <cstdlib>

#include "oracle.hpp"
#include "packet.hpp"

namespace oracle {

__m128 packet_inverse(__m128 x) {
    // common/simd.h:324-329: rcp, then r' = (r + r) - (x * r) * r.
    const __m128 r = _mm_rcp_ps(x);
    return _mm_sub_ps(_mm_add_ps(r, r), _mm_mul_ps(_mm_mul_ps(x, r), r));
}

__m128 packet_inverse_sqrt(__m128 x) {
    // common/simd.h:332-337: rsqrt, then r' = (0.5 * r) * (3 - (x * r) * r).
    const __m128 r = _mm_rsqrt_ps(x);
    return _mm_mul_ps(_mm_mul_ps(_mm_set1_ps(0.5f), r),
                      _mm_sub_ps(_mm_set1_ps(3.0f), _mm_mul_ps(_mm_mul_ps(x, r), r)));
}

static inline __m128 mov_p(__m128 p, __m128 a, __m128 b) {
    // common/simd.h:398-401: andnot(p, b) | (p & a).
    return _mm_or_ps(_mm_andnot_ps(p, b), _mm_and_ps(p, a));
}

__m128 packet_slab(const BvhNode &n, const PacketRay &r, const __m128 *i_rd) {
    const __m128 dist_x0 = _mm_mul_ps(_mm_sub_ps(_mm_set1_ps(n.low.x), r.ox), i_rd[0]);
    const __m128 dist_x1 = _mm_mul_ps(_mm_sub_ps(_mm_set1_ps(n.high.x), r.ox), i_rd[0]);
    const __m128 low_x = _mm_min_ps(dist_x0, dist_x1);
    const __m128 high_x = _mm_max_ps(dist_x0, dist_x1);
    const __m128 dist_y0 = _mm_mul_ps(_mm_sub_ps(_mm_set1_ps(n.low.y), r.oy), i_rd[1]);
    const __m128 dist_y1 = _mm_mul_ps(_mm_sub_ps(_mm_set1_ps(n.high.y), r.oy), i_rd[1]);
    const __m128 low_y = _mm_min_ps(dist_y0, dist_y1);
    const __m128 high_y = _mm_max_ps(dist_y0, dist_y1);
    const __m128 dist_z0 = _mm_mul_ps(_mm_sub_ps(_mm_set1_ps(n.low.z), r.oz), i_rd[2]);
    const __m128 dist_z1 = _mm_mul_ps(_mm_sub_ps(_mm_set1_ps(n.high.z), r.oz), i_rd[2]);
    const __m128 low_z = _mm_min_ps(dist_z0, dist_z1);
    const __m128 high_z = _mm_max_ps(dist_z0, dist_z1);
    const __m128 enter_t = _mm_max_ps(_mm_max_ps(low_x, low_y), _mm_max_ps(low_z, _mm_setzero_ps()));
    const __m128 exit_t = _mm_min_ps(_mm_min_ps(high_x, high_y), high_z);
    return mov_p(_mm_cmpgt_ps(enter_t, exit_t), _mm_set1_ps(kMaxDist), enter_t);
}

void packet_intersect(const Triangle &tri, const PacketRay &r, PacketHit *h, int tri_idx) {
    const __m128 e1_x = _mm_set1_ps(tri.c.x - tri.a.x), e1_y = _mm_set1_ps(tri.c.y - tri.a.y),
                 e1_z = _mm_set1_ps(tri.c.z - tri.a.z);
    const __m128 e2_x = _mm_set1_ps(tri.b.x - tri.a.x), e2_y = _mm_set1_ps(tri.b.y - tri.a.y),
                 e2_z = _mm_set1_ps(tri.b.z - tri.a.z);
    const __m128 p_x = _mm_sub_ps(_mm_mul_ps(r.dy, e2_z), _mm_mul_ps(e2_y, r.dz));
    const __m128 p_y = _mm_sub_ps(_mm_mul_ps(r.dz, e2_x), _mm_mul_ps(e2_z, r.dx));
    const __m128 p_z = _mm_sub_ps(_mm_mul_ps(r.dx, e2_y), _mm_mul_ps(e2_x, r.dy));
    const __m128 det = _mm_add_ps(_mm_add_ps(_mm_mul_ps(e1_x, p_x), _mm_mul_ps(e1_y, p_y)), _mm_mul_ps(e1_z, p_z));
    const __m128 inv_det = packet_inverse(det);
    const __m128 t_x = _mm_sub_ps(r.ox, _mm_set1_ps(tri.a.x));
    const __m128 t_y = _mm_sub_ps(r.oy, _mm_set1_ps(tri.a.y));
    const __m128 t_z = _mm_sub_ps(r.oz, _mm_set1_ps(tri.a.z));
    const __m128 u = _mm_mul_ps(
        _mm_add_ps(_mm_add_ps(_mm_mul_ps(t_x, p_x), _mm_mul_ps(t_y, p_y)), _mm_mul_ps(t_z, p_z)), inv_det);
    const __m128 q_x = _mm_sub_ps(_mm_mul_ps(t_y, e1_z), _mm_mul_ps(e1_y, t_z));
    const __m128 q_y = _mm_sub_ps(_mm_mul_ps(t_z, e1_x), _mm_mul_ps(e1_z, t_x));
    const __m128 q_z = _mm_sub_ps(_mm_mul_ps(t_x, e1_y), _mm_mul_ps(e1_x, t_y));
    const __m128 v = _mm_mul_ps(
        _mm_add_ps(_mm_add_ps(_mm_mul_ps(r.dx, q_x), _mm_mul_ps(r.dy, q_y)), _mm_mul_ps(r.dz, q_z)), inv_det);
    const __m128 dist = _mm_mul_ps(
        _mm_add_ps(_mm_add_ps(_mm_mul_ps(e2_x, q_x), _mm_mul_ps(e2_y, q_y)), _mm_mul_ps(e2_z, q_z)), inv_det);
    __m128 mask = _mm_cmpgt_ps(dist, _mm_setzero_ps());
    mask = _mm_and_ps(mask, _mm_cmpgt_ps(h->d, dist));
    mask = _mm_and_ps(mask, _mm_cmpge_ps(u, _mm_setzero_ps()));
    mask = _mm_and_ps(mask, _mm_cmpge_ps(v, _mm_setzero_ps()));
    mask = _mm_and_ps(mask, _mm_cmple_ps(_mm_add_ps(u, v), _mm_set1_ps(1.0f)));
    if (_mm_movemask_ps(mask) > 0) {
        h->d = mov_p(mask, dist, h->d);
        h->u = mov_p(mask, u, h->u);
        h->v = mov_p(mask, v, h->v);
        const __m128i mi = _mm_castps_si128(mask);
        h->id = _mm_or_si128(_mm_andnot_si128(mi, h->id), _mm_and_si128(mi, _mm_set1_epi32(tri_idx)));
    }
}

namespace {

// ssd/bvh.h:56-69 (one t_min vector per entry).
struct PStackElement {
    __m128 t_min;
    int idx;
};
constexpr int kPStackHeight = 128;
struct PStack {
    PStackElement e[kPStackHeight];
};

// ssd/bvh.cc:341-404.
bool find_leaf_node(const Scene &s, const PacketRay &r, int *entry_idx, PStack &stack, int *sp, const __m128 *i_rd,
                    __m128 t_max, Counters *c) {
    int cur = *entry_idx;
    while (true) {
        const BvhNode &node = s.nodes[cur];
        if (node.leaf) {
            *entry_idx = cur;
            return true;
        }
        const int near_idx = node.left_b;
        const int far_idx = node.right_e;
        const __m128 dist_0 = packet_slab(s.nodes[near_idx], r, i_rd);
        const __m128 dist_1 = packet_slab(s.nodes[far_idx], r, i_rd);
        if (c) c->box += 2;
        const __m128 far_dist = _mm_max_ps(dist_0, dist_1);
        const __m128 near_dist = _mm_min_ps(dist_0, dist_1);
        if (!_mm_movemask_ps(_mm_cmplt_ps(near_dist, t_max))) return false;
        const int o_same_order = _mm_movemask_ps(_mm_cmplt_ps(dist_0, dist_1));
        const int o_far = _mm_movemask_ps(_mm_cmplt_ps(far_dist, t_max));
        if (!o_far && o_same_order == 0xf) {
            cur = near_idx;
        } else if (!o_far && o_same_order == 0x0) {
            cur = far_idx;
        } else {
            ++*sp;
            if (*sp >= kPStackHeight) std::abort();  // the reference asserts here too
            stack.e[*sp].idx = far_idx;
            stack.e[*sp].t_min = dist_1;
            cur = near_idx;
        }
    }
}

}  // namespace

// ssd/bvh.cc:406-447 with the leaf loop of ssd/bvh_node.h:193-213.
void packet_nearest(const Scene &s, const PacketRay &r, PacketHit *h, Counters *c) {
    h->d = _mm_set1_ps(kMaxDist);
    if (s.root < 0) return;
    PStack stack;
    int sp = 0;
    stack.e[0].idx = s.root;
    stack.e[0].t_min = _mm_set1_ps(kMaxDist);
    int entry_idx = s.root;
    __m128 entry_t_min = _mm_set1_ps(kMaxDist);
    const __m128 one = _mm_set1_ps(1.0f);
    const __m128 i_rd[3] = {_mm_div_ps(one, r.dx), _mm_div_ps(one, r.dy), _mm_div_ps(one, r.dz)};
    while (true) {
        if (find_leaf_node(s, r, &entry_idx, stack, &sp, i_rd, h->d, c)) {
            const BvhNode &n = s.nodes[entry_idx];
            for (int i = n.left_b; i < n.right_e; ++i) {
                packet_intersect(s.tris[s.indirect[i]], r, h, s.indirect[i]);
                if (c) ++c->tri;
            }
        }
        do {
            if (sp == 0) return;
            entry_idx = stack.e[sp].idx;
            entry_t_min = stack.e[sp].t_min;
            --sp;
        } while (!_mm_movemask_ps(_mm_cmplt_ps(entry_t_min, h->d)));
    }
}

// ssd/bvh.cc:449-500 with the leaf loop of ssd/bvh_node.h:237-258. Returns
// the occluded lanes as an all-ones mask.
__m128 packet_nearer(const Scene &s, const PacketRay &r, __m128 t, Counters *c) {
    __m128 closer = _mm_setzero_ps();
    if (s.root < 0) return closer;
    PStack stack;
    int sp = 0;
    stack.e[0].idx = s.root;
    stack.e[0].t_min = _mm_set1_ps(kMaxDist);
    int entry_idx = s.root;
    __m128 entry_t_min = _mm_set1_ps(kMaxDist);
    PacketHit h;
    h.d = t;
    const __m128 one = _mm_set1_ps(1.0f);
    const __m128 i_rd[3] = {_mm_div_ps(one, r.dx), _mm_div_ps(one, r.dy), _mm_div_ps(one, r.dz)};
    while (true) {
        if (find_leaf_node(s, r, &entry_idx, stack, &sp, i_rd, t, c)) {
            const BvhNode &n = s.nodes[entry_idx];
            for (int i = n.left_b; i < n.right_e; ++i) {
                const Triangle &tri = s.tris[s.indirect[i]];
                if (tri.light || tri.transparent) continue;
                packet_intersect(tri, r, &h, s.indirect[i]);
                if (c) ++c->tri;
                closer = _mm_or_ps(closer, _mm_cmplt_ps(h.d, t));
                if (_mm_movemask_ps(closer) == 0xf) return closer;
            }
            if (_mm_movemask_ps(closer) == 0xf) return closer;
        }
        do {
            if (sp == 0) return closer;
            entry_idx = stack.e[sp].idx;
            entry_t_min = stack.e[sp].t_min;
            --sp;
        } while (!_mm_movemask_ps(_mm_cmplt_ps(entry_t_min, h.d)));
    }
}

void packet_brute_nearest(const Scene &s, const PacketRay &r, PacketHit *h) {
    h->d = _mm_set1_ps(kMaxDist);
    for (int i = 0; i < static_cast<int>(s.tris.size()); ++i) packet_intersect(s.tris[i], r, h, i);
}

__m128 packet_brute_nearer(const Scene &s, const PacketRay &r, __m128 t) {
    PacketHit h;
    h.d = t;
    __m128 closer = _mm_setzero_ps();
    for (int i = 0; i < static_cast<int>(s.tris.size()); ++i) {
        if (s.tris[i].light || s.tris[i].transparent) continue;
        packet_intersect(s.tris[i], r, &h, i);
        closer = _mm_or_ps(closer, _mm_cmplt_ps(h.d, t));
    }
    return closer;
}

}  // namespace oracle
