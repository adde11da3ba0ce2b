// Wide-BVH traversal of one ray per lane, shared by the per-launch kernels
// (trace.hip) and the persistent batched kernel (batch.hip).
//
// Replaces the scalar and 4-wide packet traversals of ssd/bvh.cc:341-759.
// A lane visits a node by loading the W child boxes (structure of arrays,
// 16-byte loads), testing them, sorting the hits near-to-far with a sorting
// network and keeping the nearest as the next node; the others go on a
// per-lane stack whose first kLdsDepth entries live in LDS (one column per
// lane, so a wave's accesses hit 64 different banks) and whose remainder
// spills to private memory. Stack entries carry their entry distance so that
// closest-hit pops skip subtrees beyond the current hit. Leaves are tested
// with the reference's Moller-Trumbore (rt_device.hpp); an exact tie in t
// keeps the lowest emplace id.
//
// Loop structure (Aila & Laine 2009, "while-while" with speculative
// traversal): lanes first walk interior nodes until all but kSearchCut lanes
// of the wavefront hold a leaf or have nothing left; a lane that finds a leaf
// early parks it and keeps traversing while others still search. Then the
// lanes holding leaves test their triangles together until all but kLeafCut
// lanes are done with theirs; the few lanes still searching, and those with
// a leaf left, go on in the next round.
//
// The box test is conservative: boxes are padded at build time, the inverse
// direction never divides by zero, t_far is stretched by 2^-20 and the cull
// against the current hit by 2^-18, so no box holding an accepted triangle is
// skipped (the reference's own slack is the absolute +100*EPSILON of
// ssd/bvh.cc:620,627,705).
#pragma once

#include "rt_device.hpp"
#include "rt_internal.hpp"

namespace rt {

constexpr int kStackDepth = 128;  // LDS part + private spill part
constexpr float kFarStretch = 1.00000095f;   // 1 + 2^-20
constexpr float kHitStretch = 1.0000038f;    // 1 + 2^-18
// The interior phase ends once at most this many lanes are still looking for
// their first leaf; they keep looking in the next round, after the others'
// leaf tests. Waiting for the last few searching lanes held the rest of the
// wave idle: on the 1M-triangle terrain frame a cut of 4 is 8% faster than
// waiting for every lane (0 would be the classic while-while loop).
constexpr int kSearchCut = 4;
// Likewise the leaf phase ends once at most this many lanes still hold a leaf
// to test (a lane reaches a further leaf when the next node it pops is one);
// they test it in the next round. 3% faster than draining every lane's leaves.
constexpr int kLeafCut = 12;

__device__ __forceinline__ float safe_inv(float d) {
    return 1.0f / ((fabsf(d) > 1e-20f) ? d : copysignf(1e-20f, d));
}

// Byte offsets, inside a node, of each axis's near and far plane arrays for
// one ray: the near plane of an axis is lo when the ray's inverse direction on
// that axis is positive (sign bit clear) and hi otherwise. Computed once per
// ray, so a node visit loads the near and far planes directly and the slab
// test needs no per-axis min/max. Because fma(p, 1/d, -o/d) and (p - o) / d
// are monotonic in p, the near plane's distance equals the min of the two
// planes' distances, bit for bit, and the far plane's the max: the result is
// that of the symmetric slab test.
struct PlaneOffsets {
    uint32_t n[3], f[3];
};

template <int W>
__device__ __forceinline__ PlaneOffsets plane_offsets(rtd::V3 id) {
    constexpr uint32_t A = 4u * W;  // bytes of one plane array
    const uint32_t sx = __float_as_uint(id.x) >> 31, sy = __float_as_uint(id.y) >> 31,
                   sz = __float_as_uint(id.z) >> 31;
    PlaneOffsets p;
    p.n[0] = sx * A;
    p.f[0] = A - sx * A;
    p.n[1] = 2u * A + sy * A;
    p.f[1] = 3u * A - sy * A;
    p.n[2] = 4u * A + sz * A;
    p.f[2] = 5u * A - sz * A;
    return p;
}

template <int W>
struct NodeView {
    float nx[W], fx[W], ny[W], fy[W], nz[W], fz[W];  // near and far planes per axis
    int32_t c[W];
};

// Node addressing uses 32-bit byte offsets from the node array's base (the
// array is limited to 4 GiB at build time), so a lane's loads take the
// base-plus-offset form and each costs one 32-bit add.
template <int W>
__device__ __forceinline__ void load_planes(const char *base, uint32_t off, float *dst) {
#pragma unroll
    for (int q = 0; q < W / 4; ++q) {
        const float4 x = *reinterpret_cast<const float4 *>(base + static_cast<uint32_t>(off + 16u * q));
        dst[4 * q + 0] = x.x;
        dst[4 * q + 1] = x.y;
        dst[4 * q + 2] = x.z;
        dst[4 * q + 3] = x.w;
    }
}

template <int W>
__device__ __forceinline__ void load_node(const WideNode<W> *__restrict__ nodes, int32_t idx, const PlaneOffsets &po,
                                          NodeView<W> &v) {
    const char *base = reinterpret_cast<const char *>(nodes);
    const uint32_t o = static_cast<uint32_t>(idx) * static_cast<uint32_t>(sizeof(WideNode<W>));
    load_planes<W>(base, o + po.n[0], v.nx);
    load_planes<W>(base, o + po.f[0], v.fx);
    load_planes<W>(base, o + po.n[1], v.ny);
    load_planes<W>(base, o + po.f[1], v.fy);
    load_planes<W>(base, o + po.n[2], v.nz);
    load_planes<W>(base, o + po.f[2], v.fz);
#pragma unroll
    for (int q = 0; q < W / 4; ++q) {
        const int4 g = *reinterpret_cast<const int4 *>(base + static_cast<uint32_t>(o + 24u * W + 16u * q));
        v.c[4 * q + 0] = g.x;
        v.c[4 * q + 1] = g.y;
        v.c[4 * q + 2] = g.z;
        v.c[4 * q + 3] = g.w;
    }
}

// Slab test of W children. With FUSED the plane distances are
// fma(plane, 1/d, -o/d): one operation instead of two, with an extra error of
// about |o| * 2^-24 in position, which the build's box padding (2^-15 of the
// largest scene coordinate) covers while |o| stays within 256x the scene;
// rays from farther away use the subtract-then-multiply form. An empty child
// slot holds the inverted box lo = +inf, hi = -inf, whose near distance is
// +inf and far distance -inf on every axis, so it never passes.
template <int W, bool FUSED>
__device__ __forceinline__ void box_test(const NodeView<W> &v, rtd::V3 o, rtd::V3 id, rtd::V3 oid, float tlimit,
                                         float *tn) {
#pragma unroll
    for (int k = 0; k < W; ++k) {
        float xn, xf, yn, yf, zn, zf;
        if (FUSED) {
            xn = fmaf(v.nx[k], id.x, oid.x);
            xf = fmaf(v.fx[k], id.x, oid.x);
            yn = fmaf(v.ny[k], id.y, oid.y);
            yf = fmaf(v.fy[k], id.y, oid.y);
            zn = fmaf(v.nz[k], id.z, oid.z);
            zf = fmaf(v.fz[k], id.z, oid.z);
        } else {
            xn = (v.nx[k] - o.x) * id.x;
            xf = (v.fx[k] - o.x) * id.x;
            yn = (v.ny[k] - o.y) * id.y;
            yf = (v.fy[k] - o.y) * id.y;
            zn = (v.nz[k] - o.z) * id.z;
            zf = (v.fz[k] - o.z) * id.z;
        }
        const float tnear = fmaxf(fmaxf(fmaxf(xn, yn), zn), 0.0f);
        const float tfar = fminf(fminf(xf, yf), zf);
        // Both bounds are never NaN (finite planes or the infinite empty box).
        tn[k] = (tnear <= fminf(tfar * kFarStretch, tlimit)) ? tnear : INFINITY;
    }
}

__device__ __forceinline__ void cswap(float &ta, int32_t &ca, float &tb, int32_t &cb) {
    const bool sw = tb < ta;
    const float t = sw ? tb : ta;
    const float u = sw ? ta : tb;
    const int32_t c = sw ? cb : ca;
    const int32_t d = sw ? ca : cb;
    ta = t;
    tb = u;
    ca = c;
    cb = d;
}

template <int W>
__device__ __forceinline__ void sort_children(float *t, int32_t *c);

template <>
__device__ __forceinline__ void sort_children<4>(float *t, int32_t *c) {
    cswap(t[0], c[0], t[1], c[1]);
    cswap(t[2], c[2], t[3], c[3]);
    cswap(t[0], c[0], t[2], c[2]);
    cswap(t[1], c[1], t[3], c[3]);
    cswap(t[1], c[1], t[2], c[2]);
}

template <>
__device__ __forceinline__ void sort_children<8>(float *t, int32_t *c) {
    // Batcher's odd-even merge sort network for 8 keys (19 comparators).
    cswap(t[0], c[0], t[1], c[1]); cswap(t[2], c[2], t[3], c[3]);
    cswap(t[4], c[4], t[5], c[5]); cswap(t[6], c[6], t[7], c[7]);
    cswap(t[0], c[0], t[2], c[2]); cswap(t[1], c[1], t[3], c[3]);
    cswap(t[4], c[4], t[6], c[6]); cswap(t[5], c[5], t[7], c[7]);
    cswap(t[1], c[1], t[2], c[2]); cswap(t[5], c[5], t[6], c[6]);
    cswap(t[0], c[0], t[4], c[4]); cswap(t[1], c[1], t[5], c[5]);
    cswap(t[2], c[2], t[6], c[6]); cswap(t[3], c[3], t[7], c[7]);
    cswap(t[2], c[2], t[4], c[4]); cswap(t[3], c[3], t[5], c[5]);
    cswap(t[1], c[1], t[2], c[2]); cswap(t[3], c[3], t[4], c[4]);
    cswap(t[5], c[5], t[6], c[6]);
}

// Traversal of one ray (one lane of a block of kCols threads; lane is the
// thread's index in the block, its LDS stack column). On return hit holds
// (t, u, v, id) for a closest-hit query (id -1 on a miss) or id >= 0 when an
// any-hit query found an occluder. The LDS stack is declared here, so its
// accesses stay LDS instructions after inlining (a pointer parameter loses the
// address space and lets the compiler fold LDS and scratch reads into flat
// loads).
template <int W, bool CLOSEST, bool COUNT, int kLdsDepth, int kCols = 64>
__device__ __forceinline__ void trace_ray(const WideNode<W> *__restrict__ nodes, const LeafTri *__restrict__ tris,
                                          int32_t root, float near_limit, rtd::V3 o, rtd::V3 d, float tmin, float tmax,
                                          int lane, float4 &hit, uint32_t &visits, uint32_t &tests, bool &overflow) {
    constexpr int kSpillDepth = kStackDepth - kLdsDepth;
    __shared__ int2 lds_stack[kLdsDepth][kCols];
    const rtd::V3 id = rtd::mk(safe_inv(d.x), safe_inv(d.y), safe_inv(d.z));
    const rtd::V3 oid = rtd::mk(-(o.x * id.x), -(o.y * id.y), -(o.z * id.z));
    const PlaneOffsets po = plane_offsets<W>(id);
    const bool near = fmaxf(fmaxf(fabsf(o.x), fabsf(o.y)), fabsf(o.z)) <= near_limit;

    // Two separate arrays so that a spill read never has the type of an LDS
    // entry read (the compiler would otherwise merge both into a flat load).
    int32_t spill_code[kSpillDepth];
    float spill_t[kSpillDepth];
    int sp = 0;

    float best_t = tmax, best_u = 0.0f, best_v = 0.0f;
    int32_t best_id = -1;

    // Pop the nearest stacked subtree that can still hold a closer hit.
    auto pop = [&]() -> int32_t {
        while (sp > 0) {
            --sp;
            int2 e;
            if (sp < kLdsDepth) {
                e = lds_stack[sp][lane];
            } else {
                e.x = spill_code[sp - kLdsDepth];
                e.y = __float_as_int(spill_t[sp - kLdsDepth]);
            }
            if (!CLOSEST || __int_as_float(e.y) <= best_t * kHitStretch) return e.x;
        }
        return kEmpty;
    };
    auto push = [&](int32_t code, float t) {
        if (sp < kLdsDepth) {
            lds_stack[sp][lane] = make_int2(code, __float_as_int(t));
            ++sp;
        } else if (sp < kLdsDepth + kSpillDepth) {
            spill_code[sp - kLdsDepth] = code;
            spill_t[sp - kLdsDepth] = t;
            ++sp;
        } else {
            overflow = true;
        }
    };

    int32_t cur = root;      // interior node, leaf code or kEmpty (done)
    int32_t leaf = kEmpty;
    while (true) {
        // ---- interior phase
        while (true) {
            if (cur >= 0) {
                NodeView<W> v;
                load_node<W>(nodes, cur, po, v);
                if (COUNT) ++visits;
                float tn[W];
                if (near)
                    box_test<W, true>(v, o, id, oid, best_t * kHitStretch, tn);
                else
                    box_test<W, false>(v, o, id, oid, best_t * kHitStretch, tn);
                int32_t cc[W];
#pragma unroll
                for (int k = 0; k < W; ++k) cc[k] = v.c[k];
                if (CLOSEST) sort_children<W>(tn, cc);
                int32_t next = kEmpty;
                float next_t = INFINITY;
#pragma unroll
                for (int k = W - 1; k >= 0; --k) {
                    if (tn[k] < INFINITY) {
                        if (next != kEmpty) push(next, next_t);
                        next = cc[k];
                        next_t = tn[k];
                    }
                }
                cur = (next != kEmpty) ? next : pop();
            }
            // Park the first leaf found and keep looking (speculation).
            if (cur != kEmpty && cur < 0 && leaf == kEmpty) {
                leaf = cur;
                cur = pop();
            }
            // Leave once at most kSearchCut lanes are still looking for their
            // first leaf.
            const bool searching = (cur >= 0) && (leaf == kEmpty);
            if (__popcll(__ballot(searching)) <= kSearchCut) break;
            // Lanes holding a leaf whose next node is also a leaf wait here.
        }
        // ---- leaf phase
        bool occluded = false;
        while (true) {
            if (leaf != kEmpty) {
                const int32_t first = leaf_first(leaf);
                const int32_t count = leaf_count(leaf);
                for (int k = 0; k < count; ++k) {
                    const LeafTri lt = tris[first + k];
                    const int32_t pid = __float_as_int(lt.a_id.w);
                    if (COUNT) ++tests;
                    if (!CLOSEST && (__float_as_int(lt.e1.w) & 3)) continue;  // lights, transparent
                    float t, u, vv;
                    if (rtd::intersect_mt(o, d, rtd::f4v(lt.a_id), rtd::f4v(lt.e1), rtd::f4v(lt.e2), tmin, &t, &u,
                                          &vv)) {
                        if (CLOSEST) {
                            if (t < best_t || (t == best_t && best_id >= 0 && pid < best_id)) {
                                best_t = t;
                                best_u = u;
                                best_v = vv;
                                best_id = pid;
                            }
                        } else if (t < tmax) {
                            occluded = true;
                            break;
                        }
                    }
                }
                leaf = kEmpty;
                if (!CLOSEST && occluded) break;
                // A leaf that was reached after the parked one is tested now.
                if (cur != kEmpty && cur < 0) {
                    leaf = cur;
                    cur = pop();
                }
            }
            if (__popcll(__ballot(leaf != kEmpty)) <= kLeafCut) break;
        }
        if (!CLOSEST && occluded) {
            best_id = 0;  // marks "occluded" for the store below
            break;
        }
        if (cur == kEmpty && leaf == kEmpty) break;
    }
    hit = make_float4(best_t, best_u, best_v, __int_as_float(best_id));
}

}  // namespace rt
