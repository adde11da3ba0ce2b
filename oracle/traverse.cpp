// Oracle — TEST INFRASTRUCTURE ONLY (see vec.hpp for the rules).
//
// Scalar BVH traversal, ssd/bvh.cc:586-759, with the leaf loops of
// ssd/bvh_node.h:152-190, plus a brute-force checker.
#include <cstdlib>

#include "oracle.hpp"

namespace oracle {

namespace {

// ssd/bvh.h:56-69
struct StackElement {
    float t_min;
    int idx;
};

// 100 * EPSILON as the reference forms it (a float product).
const float kSlack = 100.0f * kEpsilon;

// common/common.h:63-65 MAX_BVH_STACK_HEIGHT is 100; the reference asserts on
// overflow (ssd/bvh.cc:636). A fixed array in the caller's frame keeps each
// query allocation-free and re-entrant.
constexpr int kStackHeight = 128;
struct Stack {
    StackElement e[kStackHeight];
};

// ssd/bvh.cc:586-647 — near-first descent to the next leaf. Returns false when
// both children are culled; the far child is pushed when both are needed.
bool find_leaf_node(const Scene &s, const Ray &r, int *entry_idx, Stack &stack, int *sp, const V3 &i_rd,
                    float t_max, Counters *c) {
    int cur = *entry_idx;
    while (true) {
        const BvhNode &node = s.nodes[cur];
        if (node.leaf) {
            *entry_idx = cur;
            return true;
        }
        int near_idx = node.left_b;
        int far_idx = node.right_e;
        float near_dist = slab_distance(s.nodes[near_idx], r.ogn, i_rd);
        float far_dist = slab_distance(s.nodes[far_idx], r.ogn, i_rd);
        if (c) c->box += 2;
        if (near_dist > far_dist) {
            std::swap(near_idx, far_idx);
            std::swap(near_dist, far_dist);
        }
        if (near_dist >= (t_max + kSlack)) return false;
        if (far_dist >= (t_max + kSlack)) {
            cur = near_idx;
        } else {
            ++*sp;
            if (*sp >= kStackHeight) std::abort();  // the reference asserts here too
            stack.e[*sp].idx = far_idx;
            stack.e[*sp].t_min = far_dist;
            cur = near_idx;
        }
    }
}

// ssd/bvh_node.h:152-168 — replace only on a strictly smaller distance.
int test_leaf_node_nearest(const Scene &s, const BvhNode &n, const Ray &r, Hit *h, Counters *c) {
    int intersecting_object = -1;
    for (int i = n.left_b; i < n.right_e; ++i) {
        Hit hit_type;
        hit_type.d = h->d;
        intersect_scalar(s.tris[s.indirect[i]], r, &hit_type);
        if (c) ++c->tri;
        if (hit_type.d < h->d) {
            *h = hit_type;
            intersecting_object = s.indirect[i];
        }
    }
    return intersecting_object;
}

// ssd/bvh_node.h:170-190 — lights and transparent triangles never occlude.
bool test_leaf_node_nearer(const Scene &s, const BvhNode &n, const Ray &r, float max, Counters *c) {
    for (int i = n.left_b; i < n.right_e; ++i) {
        const Triangle &tri = s.tris[s.indirect[i]];
        if (tri.light || tri.transparent) continue;
        Hit hit_type;
        hit_type.d = max;
        intersect_scalar(tri, r, &hit_type);
        if (c) ++c->tri;
        if (hit_type.d < max) return true;
    }
    return false;
}

}  // namespace

// ssd/bvh.cc:649-709
int find_nearest_object(const Scene &s, const Ray &r, Hit *h, Counters *c) {
    if (s.root < 0) {
        *h = Hit();
        return -1;
    }
    const V3 ray_dir_inv = 1.0f / r.dir;
    if (c) ++c->box;
    if (slab_distance(s.nodes[s.root], r.ogn, ray_dir_inv) == kMaxDist) {
        return -1;
    }
    Stack stack;
    int sp = 0;
    stack.e[0].idx = s.root;
    stack.e[0].t_min = kMaxDist;
    int entry_idx = s.root;
    float entry_t_min = kMaxDist;
    Hit nearest;
    int hit_object = -1;
    while (true) {
        const bool leaf = find_leaf_node(s, r, &entry_idx, stack, &sp, ray_dir_inv, nearest.d, c);
        if (leaf) {
            const int io = test_leaf_node_nearest(s, s.nodes[entry_idx], r, &nearest, c);
            if (io != -1) hit_object = io;
        }
        do {
            if (sp == 0) {
                *h = nearest;
                return hit_object;
            }
            entry_idx = stack.e[sp].idx;
            entry_t_min = stack.e[sp].t_min;
            --sp;
        } while (entry_t_min > (nearest.d + kSlack));
    }
}

// ssd/bvh.cc:711-759
bool found_nearer_object(const Scene &s, const Ray &r, float t, Counters *c) {
    if (s.root < 0) return false;
    Stack stack;
    int sp = 0;
    stack.e[0].idx = s.root;
    stack.e[0].t_min = kMaxDist;
    int entry_idx = s.root;
    const V3 ray_dir_inv = 1.0f / r.dir;
    while (true) {
        const bool leaf = find_leaf_node(s, r, &entry_idx, stack, &sp, ray_dir_inv, t, c);
        if (leaf && test_leaf_node_nearer(s, s.nodes[entry_idx], r, t, c)) return true;
        if (sp == 0) return false;
        entry_idx = stack.e[sp].idx;
        --sp;
    }
}

int brute_nearest(const Scene &s, const Ray &r, Hit *h) {
    Hit nearest;
    int hit_object = -1;
    for (int i = 0; i < static_cast<int>(s.tris.size()); ++i) {
        Hit hit_type;
        hit_type.d = nearest.d;
        intersect_scalar(s.tris[i], r, &hit_type);
        if (hit_type.d < nearest.d) {
            nearest = hit_type;
            hit_object = i;
        }
    }
    *h = nearest;
    return hit_object;
}

bool brute_nearer(const Scene &s, const Ray &r, float t) {
    for (int i = 0; i < static_cast<int>(s.tris.size()); ++i) {
        const Triangle &tri = s.tris[i];
        if (tri.light || tri.transparent) continue;
        Hit hit_type;
        hit_type.d = t;
        intersect_scalar(tri, r, &hit_type);
        if (hit_type.d < t) return true;
    }
    return false;
}

}  // namespace oracle
