// Oracle — TEST INFRASTRUCTURE ONLY (see vec.hpp for the rules).
//
// Triangle, slab test, Morton code, min_element and the shadow-ray offset.
#include "oracle.hpp"

namespace oracle {

// rt/triangle.h:124-176
Triangle make_triangle(const V3 &a, const V3 &b, const V3 &c, bool light, bool transparent) {
    Triangle t;
    t.a = a;
    t.b = b;
    t.c = c;
    t.light = light;
    t.transparent = transparent;
    const V3 dir_b = b - a;
    const V3 dir_c = c - a;
    t.n = normalise_exact(cross(dir_b, dir_c));
    return t;
}

// rt/triangle.h:190-235. e1 runs to c and e2 to b, so u weights c and v weights b.
void intersect_scalar(const Triangle &tri, const Ray &r, Hit *h) {
    const V3 e1 = tri.c - tri.a;
    const V3 e2 = tri.b - tri.a;
    const V3 P = cross(r.dir, e2);
    const float det = dot(e1, P);
    if (det > -0.000001f && det < 0.000001f) {
        h->d = kMaxDist;
        return;
    }
    const float inv_det = 1.0f / det;
    const V3 T = r.ogn - tri.a;
    const float u = dot(T, P) * inv_det;
    if (u < 0.0f) {
        h->d = kMaxDist;
        return;
    }
    const V3 Q = cross(T, e1);
    const float v = dot(r.dir, Q) * inv_det;
    if ((v < 0.0f) || ((u + v) > 1.0f)) {
        h->d = kMaxDist;
        return;
    }
    const float t = dot(e2, Q) * inv_det;
    if (t > kEpsilon) {
        h->d = t;
        h->u = u;
        h->v = v;
    }
}

// ssd/bvh_node.h:61-98
float slab_distance(const BvhNode &n, const V3 &ogn, const V3 &i_rd) {
    const V3 low_range = n.low - ogn;
    const V3 high_range = n.high - ogn;
    V3 lo = low_range * i_rd;
    V3 hi = high_range * i_rd;
    if (lo.x > hi.x) std::swap(lo.x, hi.x);
    if (lo.y > hi.y) std::swap(lo.y, hi.y);
    if (lo.z > hi.z) std::swap(lo.z, hi.z);
    const float exit_t = ref_min(ref_min(hi.x, hi.y), hi.z);
    const float enter_t = ref_max(ref_max(lo.x, lo.y), lo.z);
    if (ref_max(0.0f, enter_t) > exit_t) return kMaxDist;
    return enter_t;
}

// ssd/bvh_node.h:48-58
float combined_surface_area(const BvhNode &l, const BvhNode &r) {
    const V3 bl = vmin(l.low, r.low);
    const V3 tr = vmax(l.high, r.high);
    const V3 dist = tr - bl;
    return ((dist.x * dist.y) + (dist.x * dist.z)) + (dist.y * dist.z);
}

// common/simd.h:798-811
static int split_by_three(int a) {
    a |= (a << 16);
    a &= 0x030000FF;
    a |= (a << 8);
    a &= 0x0300F00F;
    a |= (a << 4);
    a &= 0x030C30C3;
    a |= (a << 2);
    a &= 0x09249249;
    return a;
}

// common/simd.h:813-823 (and the 4-wide form at :842-854, which converts with
// cvttps, i.e. the same truncation).
int morton_code(float x, float y, float z, float x_mul, float y_mul, float z_mul) {
    const int xi = static_cast<int>(x * x_mul);
    const int yi = static_cast<int>(y * y_mul);
    const int zi = static_cast<int>(z * z_mul);
    return split_by_three(xi) | (split_by_three(yi) << 1) | (split_by_three(zi) << 2);
}

// common/simd.h:857-894. Four lanes keep a running minimum with a strict <,
// recording the 4-aligned block base; the lowest lane holding the overall
// minimum wins; trailing elements replace it only on a strict <.
int min_element_n(const float *d, float *m, int n) {
    const float inf = std::numeric_limits<float>::infinity();
    float lane_v[4] = {inf, inf, inf, inf};
    int lane_i[4] = {-1, -1, -1, -1};
    for (int i = 0; i <= n - 4; i += 4) {
        for (int l = 0; l < 4; ++l) {
            if (d[i + l] < lane_v[l]) lane_i[l] = i;
            // _mm_min_ps(data, min_v) returns data where data < min_v.
            lane_v[l] = (d[i + l] < lane_v[l]) ? d[i + l] : lane_v[l];
        }
    }
    // Horizontal reduction exactly as the two shuffles + mins do it.
    auto mn = [](float a, float b) { return (a < b) ? a : b; };
    const float m2[4] = {mn(lane_v[0], lane_v[2]), mn(lane_v[1], lane_v[3]), mn(lane_v[2], lane_v[0]),
                         mn(lane_v[3], lane_v[1])};
    const float m4[4] = {mn(m2[0], m2[1]), mn(m2[1], m2[0]), mn(m2[2], m2[1]), mn(m2[3], m2[0])};
    float min_d = m4[0];
    int min_idx = 0;
    for (int l = 0; l < 4; ++l) {
        if (lane_v[l] == m4[l]) {
            min_idx = l;
            break;
        }
    }
    int ret = lane_i[min_idx] + min_idx;
    for (int i = (n & ~3); i < n; ++i) {
        if (d[i] < min_d) {
            min_d = d[i];
            ret = i;
        }
    }
    *m = min_d;
    return ret;
}

// common/simd.h:896-927
int min_element_range(const float *d, float *m, int begin, int end) {
    int min_i = -1;
    float min_d = std::numeric_limits<float>::infinity();
    const int aligned = (begin & ~3) + ((begin & 3) ? 4 : 0);
    for (int i = begin; i < std::min(end, aligned); ++i) {
        if (d[i] < min_d) {
            min_d = d[i];
            min_i = i;
        }
    }
    float r_m = std::numeric_limits<float>::infinity();
    const int n = end - aligned;
    const int r_i = (n > 0) ? min_element_n(&d[aligned], &r_m, n) : 0;
    if (r_m < min_d) {
        *m = r_m;
        return r_i + aligned;
    }
    *m = min_d;
    return min_i;
}

// rt/ray.h:111-157. Each coordinate's float bit pattern moves by
// int(+-32 * geo_norm * d) ulps; a coordinate with |x| < 1e-4 is replaced by
// 0 + 1e-4 * geo_norm (the start point is zero-initialised).
V3 offset_start_point(const V3 &dst, const V3 &geo_norm, int d) {
    static const float offset[2] = {kFpDelta, -kFpDelta};
    const float df = static_cast<float>(d);
    const uint32_t small = f2u(kFpDeltaSmall);
    float out[3] = {0.0f, 0.0f, 0.0f};
    const float p[3] = {dst.x, dst.y, dst.z};
    const float g[3] = {geo_norm.x, geo_norm.y, geo_norm.z};
    for (int k = 0; k < 3; ++k) {
        uint32_t bits = f2u(p[k]);
        if ((bits & 0x7fffffffu) < small) {
            out[k] += kFpDeltaSmall * g[k];
        } else {
            bits += static_cast<uint32_t>(static_cast<int>((offset[(bits >> 31) & 0x1] * g[k]) * df));
            out[k] = u2f(bits);
        }
    }
    return V3(out[0], out[1], out[2]);
}

}  // namespace oracle
