// Oracle — TEST INFRASTRUCTURE ONLY (see vec.hpp for the rules).
//
// Internal C++ interface of the CPU restatement of RAPTor's ray tracer hot
// path. The C entry points in rt_oracle.h wrap these for ctypes.
#pragma once

#include <cstdint>
#include <vector>

#include "vec.hpp"

namespace oracle {

// rt/triangle.h:106-111 — vertices a, b, c and the normalised face normal;
// rt/triangle.h:40-41 — the light and transparent flags the any-hit test skips.
struct Triangle {
    V3 a, b, c, n;
    bool light = false;
    bool transparent = false;
};

// rt/triangle.h:124-176 — the constructor computes n = normalise(cross(b-a, c-a)).
Triangle make_triangle(const V3 &a, const V3 &b, const V3 &c, bool light, bool transparent);

// common/common.h:230-239 hit_description (hit_t is tracked separately).
struct Hit {
    float u = 0.0f, v = 0.0f, d = kMaxDist;
};

// rt/ray.h:11-174: origin, destination, direction, geometry normal, length,
// magnitude. Only the members the hot path reads are kept.
struct Ray {
    V3 ogn, dst, dir, geo_norm;
    float length = 0.0f, magn = 1.0f;
};

// rt/triangle.h:190-235 — scalar Moller-Trumbore. On a reject h->d = MAX_DIST;
// on accept with t > EPSILON (d, u, v) are written; otherwise h is unchanged.
void intersect_scalar(const Triangle &tri, const Ray &r, Hit *h);

// ssd/bvh_node.h:22-46, 294-311 — binary node. For a leaf (b, e) index the
// primitive indirection; for an interior node (left, right) index nodes.
struct BvhNode {
    int left_b = 0;
    int right_e = 0;
    V3 low, high;
    bool leaf = false;
};

// ssd/bvh_node.h:61-98 — scalar slab test. Returns MAX_DIST on a miss, else
// the (possibly negative) entry distance.
float slab_distance(const BvhNode &n, const V3 &ogn, const V3 &i_rd);
// ssd/bvh_node.h:48-58 — half surface area of the union of two boxes.
float combined_surface_area(const BvhNode &l, const BvhNode &r);

// common/simd.h:798-823 — 10-bit-per-axis Morton code with int truncation.
int morton_code(float x, float y, float z, float x_mul, float y_mul, float z_mul);
// common/simd.h:857-927 — SSE min_element with lowest-lane tie-breaking,
// restated lane by lane.
int min_element_n(const float *d, float *m, int n);
int min_element_range(const float *d, float *m, int begin, int end);

// rt/ray.h:111-157 — shadow-ray origin offset in float bit space.
V3 offset_start_point(const V3 &dst, const V3 &geo_norm, int d);

struct Scene {
    std::vector<Triangle> tris;   // emplace order: index = hit id
    std::vector<int> indirect;    // rt/primitive_store.h:17-110 indirection
    std::vector<BvhNode> nodes;
    int root = -1;
    V3 scene_lo, scene_hi;
};

// ssd/bvh_builder.cc:18-178 — AAC build. Fills scene->nodes, permutes
// scene->indirect, returns the root index (the last node made).
struct AacParams {
    float alpha = 0.4f, delta = 20.0f, max_leaf_sah_factor = 0.00000000009f;
    int max_down_phase_depth = 35;
};
int build_aac(Scene *scene, const AacParams &p = AacParams());
// The builder's Morton codes of the triangles (emplace order) and the order
// its radix sort leaves them in (ssd/bvh_builder.cc:18-158).
void aac_morton(const Scene &s, std::vector<int> *codes, std::vector<int> *order);
// The three-pass 10-bit LSD counting sort of ssd/bvh_builder.cc:50-158.
void morton_radix_sort(const std::vector<int> &codes, std::vector<int> *sorted_codes, std::vector<int> *order);

// Traversal counters: child slab tests and triangle tests.
struct Counters {
    uint64_t box = 0, tri = 0;
};

// ssd/bvh.cc:586-709 (closest hit) and 711-759 (any hit). Re-entrant: the
// stack lives in the caller's frame, unlike ssd/bvh.h:84.
int find_nearest_object(const Scene &s, const Ray &r, Hit *h, Counters *c);
bool found_nearer_object(const Scene &s, const Ray &r, float t, Counters *c);

// Brute force over every triangle with the same leaf rules (strict < keeps
// the first of equal distances in emplace order).
int brute_nearest(const Scene &s, const Ray &r, Hit *h);
bool brute_nearer(const Scene &s, const Ray &r, float t);

}  // namespace oracle
