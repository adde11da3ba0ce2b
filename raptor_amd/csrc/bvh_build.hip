// GPU BVH build: Morton-code LBVH, collapsed to a wide (4- or 8-child) BVH.
//
// Replaces the CPU AAC builder of ssd/bvh_builder.cc:18-178. The reference
// sorts 30-bit Morton codes (ssd/bvh_builder.cc:44-158) and clusters
// bottom-up on one core; here:
//   1. k_tri_bounds   per-triangle AABBs, centroid and scene bounds
//   2. k_morton       30-bit Morton codes of the centroids
//   3. radix sort     stable (code, index) sort (radix_sort.hpp)
//   4. k_karras       binary radix tree over the sorted codes (Karras 2012),
//                     duplicates split by sorted position
//   5. refit          bottom-up boxes, one launch per tree level
//   6. k_collapse     top-down, one level per launch: each wide node takes its
//                     binary node's children and keeps opening the child of
//                     largest surface area until it has W children; subtrees
//                     of at most leaf_size triangles become leaves
//   7. k_leaf_tris    triangles copied into BVH order as (a, id, c-a, b-a)
// Child boxes are padded outward by 2^-15 of the largest coordinate magnitude
// so that traversal never rejects a box holding a triangle the Moller-Trumbore
// test accepts; hit ids therefore do not depend on the tree's shape.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "radix_sort.hpp"
#include "rt_device.hpp"
#include "rt_internal.hpp"

namespace rt {

namespace {

// Order-preserving float <-> uint32 map for atomic min/max on floats.
__device__ __forceinline__ uint32_t f2ord(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ inline float ord2f(uint32_t u) {
    const uint32_t b = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
    float f;
    __builtin_memcpy(&f, &b, 4);
    return f;
}

// bounds[0..2] min of centroid, [3..5] max of centroid, [6..8] min of AABB,
// [9..11] max of AABB, all order-mapped. Each block writes its 12 partial
// bounds to parts[k * gridDim.x + block] and k_bounds_final reduces them: one
// atomic per block on 12 words took 97 us for 1M triangles (the updates
// serialise across the XCDs' L2s), the second pass takes a few.
__global__ void __launch_bounds__(256) k_tri_bounds(const float *__restrict__ abc, int64_t n,
                                                    float4 *__restrict__ aabb_lo, float4 *__restrict__ aabb_hi,
                                                    uint32_t *__restrict__ parts) {
    __shared__ uint32_t red[12][4];
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    float v[12];
    if (i < n) {
        const float *t = abc + 9 * i;
        const float lx = fminf(t[0], fminf(t[3], t[6])), ly = fminf(t[1], fminf(t[4], t[7])),
                    lz = fminf(t[2], fminf(t[5], t[8]));
        const float hx = fmaxf(t[0], fmaxf(t[3], t[6])), hy = fmaxf(t[1], fmaxf(t[4], t[7])),
                    hz = fmaxf(t[2], fmaxf(t[5], t[8]));
        aabb_lo[i] = make_float4(lx, ly, lz, 0.0f);
        aabb_hi[i] = make_float4(hx, hy, hz, 0.0f);
        const float cx = (lx + hx) * 0.5f, cy = (ly + hy) * 0.5f, cz = (lz + hz) * 0.5f;
        v[0] = cx; v[1] = cy; v[2] = cz; v[3] = cx; v[4] = cy; v[5] = cz;
        v[6] = lx; v[7] = ly; v[8] = lz; v[9] = hx; v[10] = hy; v[11] = hz;
    } else {
        for (int k = 0; k < 12; ++k) v[k] = (k < 3 || (k >= 6 && k < 9)) ? INFINITY : -INFINITY;
    }
    // Wave reduction, then one atomic per wave per bound.
    for (int k = 0; k < 12; ++k) {
        float x = v[k];
        const bool is_min = (k < 3) || (k >= 6 && k < 9);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const float y = __shfl_xor(x, o, 64);
            x = is_min ? fminf(x, y) : fmaxf(x, y);
        }
        v[k] = x;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
        for (int k = 0; k < 12; ++k) red[k][wave] = f2ord(v[k]);
    __syncthreads();
    if (threadIdx.x < 12) {
        const int k = threadIdx.x;
        const bool is_min = (k < 3) || (k >= 6 && k < 9);
        uint32_t r = red[k][0];
        for (int w = 1; w < 4; ++w) r = is_min ? min(r, red[k][w]) : max(r, red[k][w]);
        parts[static_cast<int64_t>(k) * gridDim.x + blockIdx.x] = r;
    }
}

__global__ void __launch_bounds__(256) k_bounds_final(const uint32_t *__restrict__ parts, int nparts,
                                                      uint32_t *__restrict__ bounds) {
    __shared__ uint32_t red[12][4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int k = 0; k < 12; ++k) {
        const bool is_min = (k < 3) || (k >= 6 && k < 9);
        uint32_t r = is_min ? 0xffffffffu : 0u;
        for (int b = threadIdx.x; b < nparts; b += blockDim.x) {
            const uint32_t y = parts[static_cast<int64_t>(k) * nparts + b];
            r = is_min ? min(r, y) : max(r, y);
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const uint32_t y = static_cast<uint32_t>(__shfl_xor(static_cast<int>(r), o, 64));
            r = is_min ? min(r, y) : max(r, y);
        }
        if (lane == 0) red[k][wave] = r;
    }
    __syncthreads();
    if (threadIdx.x < 12) {
        const int k = threadIdx.x;
        const bool is_min = (k < 3) || (k >= 6 && k < 9);
        uint32_t r = red[k][0];
        for (int w = 1; w < 4; ++w) r = is_min ? min(r, red[k][w]) : max(r, red[k][w]);
        bounds[k] = r;
    }
}

__device__ __forceinline__ uint32_t expand_bits10(uint32_t v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__global__ void __launch_bounds__(256) k_morton(const float4 *__restrict__ aabb_lo, const float4 *__restrict__ aabb_hi,
                                                int64_t n, const uint32_t *__restrict__ bounds,
                                                uint32_t *__restrict__ codes, uint32_t *__restrict__ idx) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 lo = aabb_lo[i], hi = aabb_hi[i];
    const float c[3] = {(lo.x + hi.x) * 0.5f, (lo.y + hi.y) * 0.5f, (lo.z + hi.z) * 0.5f};
    // One cell size for all axes (the largest centroid extent): a flat scene
    // such as a heightfield then spends few code bits on its thin axis.
    float ext = 0.0f;
#pragma unroll
    for (int k = 0; k < 3; ++k) ext = fmaxf(ext, ord2f(bounds[3 + k]) - ord2f(bounds[k]));
    uint32_t q[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float bmin = ord2f(bounds[k]);
        const float f = (ext > 0.0f) ? (c[k] - bmin) / ext : 0.0f;
        q[k] = static_cast<uint32_t>(fminf(fmaxf(f * 1024.0f, 0.0f), 1023.0f));
    }
    codes[i] = expand_bits10(q[0]) | (expand_bits10(q[1]) << 1) | (expand_bits10(q[2]) << 2);
    idx[i] = static_cast<uint32_t>(i);
}

__device__ __forceinline__ int delta(const uint32_t *codes, int64_t n, int64_t i, int64_t j) {
    if (j < 0 || j >= n) return -1;
    const uint32_t a = codes[i], b = codes[j];
    if (a != b) return __clz(a ^ b);
    return 32 + __clz(static_cast<uint32_t>(i) ^ static_cast<uint32_t>(j));
}

// Binary node children: c >= 0 internal node c, c < 0 leaf at sorted position ~c.
__global__ void __launch_bounds__(256) k_karras(const uint32_t *__restrict__ codes, int64_t n, int2 *__restrict__ child,
                                                int2 *__restrict__ range, int32_t *__restrict__ parent,
                                                int32_t *__restrict__ leaf_parent) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    const int d = (delta(codes, n, i, i + 1) - delta(codes, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = delta(codes, n, i, i - d);
    int64_t lmax = 2;
    while (delta(codes, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int64_t l = 0;
    for (int64_t t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(codes, n, i, i + (l + t) * d) > dmin) l += t;
    const int64_t j = i + l * d;
    const int dnode = delta(codes, n, i, j);
    int64_t s = 0;
    int64_t t = l;
    do {
        t = (t + 1) >> 1;
        if (delta(codes, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int64_t gamma = i + s * d + min(d, 0);
    const int64_t first = min(i, j), last = max(i, j);
    int32_t left, right;
    if (first == gamma) {
        left = ~static_cast<int32_t>(gamma);
        leaf_parent[gamma] = static_cast<int32_t>(i);
    } else {
        left = static_cast<int32_t>(gamma);
        parent[gamma] = static_cast<int32_t>(i);
    }
    if (last == gamma + 1) {
        right = ~static_cast<int32_t>(gamma + 1);
        leaf_parent[gamma + 1] = static_cast<int32_t>(i);
    } else {
        right = static_cast<int32_t>(gamma + 1);
        parent[gamma + 1] = static_cast<int32_t>(i);
    }
    child[i] = make_int2(left, right);
    range[i] = make_int2(static_cast<int32_t>(first), static_cast<int32_t>(last));
    if (i == 0) parent[0] = -1;
}

__global__ void __launch_bounds__(256) k_gather_leaf_boxes(const uint32_t *__restrict__ sorted_idx, int64_t n,
                                                           const float4 *__restrict__ aabb_lo,
                                                           const float4 *__restrict__ aabb_hi,
                                                           float4 *__restrict__ leaf_lo, float4 *__restrict__ leaf_hi) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t i = sorted_idx[k];
    leaf_lo[k] = aabb_lo[i];
    leaf_hi[k] = aabb_hi[i];
}

constexpr int kMaxLevels = 256;
constexpr int kLevelGroup = 8;  // level launches per host read-back of the level sizes

__device__ __forceinline__ void child_box(int32_t c, const float4 *node_lo, const float4 *node_hi,
                                          const float4 *leaf_lo, const float4 *leaf_hi, float4 *lo, float4 *hi) {
    if (c >= 0) {
        *lo = node_lo[c];
        *hi = node_hi[c];
    } else {
        *lo = leaf_lo[~c];
        *hi = leaf_hi[~c];
    }
}

// Refit without cross-workgroup fences: MI355X's L2 is per XCD, so a device-
// scope fence per tree level (the usual atomic bottom-up refit) writes back L2
// every step. Instead a breadth-first sweep from the root lists the internal
// nodes level by level (one launch per level, wave-aggregated appends), and
// one launch per level, deepest first, unions the two child boxes; kernel
// boundaries order the levels.
// Top-down passes keep their levels on the device: level l occupies
// positions [lvl[2l], lvl[2l + 1]) of the pass's order array, and *total
// counts what has been enqueued. A level kernel reads its range from lvl, and
// k_level_close (one thread, launched after it) records the next level's
// range, so the host reads the level sizes back once per group of levels
// instead of once per level.
__global__ void __launch_bounds__(256) k_bfs_level(int32_t *__restrict__ order, const uint32_t *__restrict__ lvl, int l,
                                                   const int2 *__restrict__ child, uint32_t *__restrict__ total) {
    const uint32_t lo = lvl[2 * l], hi = lvl[2 * l + 1];
    const int64_t k = static_cast<int64_t>(lo) + static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k - (threadIdx.x & 63) >= hi) return;  // whole wave past the level
    int2 ch = make_int2(-1, -1);
    if (k < hi) ch = child[order[k]];
    const bool a = ch.x >= 0, b = ch.y >= 0;
    const uint64_t ma = __ballot(a), mb = __ballot(b);
    const int lane = threadIdx.x & 63;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint32_t cnt = static_cast<uint32_t>(__popcll(ma) + __popcll(mb));
    uint32_t base = 0;
    if (lane == 0 && cnt) base = atomicAdd(total, cnt);
    base = __shfl(base, 0, 64);
    if (a) order[base + __popcll(ma & lt)] = ch.x;
    if (b) order[base + __popcll(ma) + __popcll(mb & lt)] = ch.y;
}

__global__ void k_level_init(int32_t *__restrict__ order, int32_t root, uint32_t *__restrict__ lvl,
                             uint32_t *__restrict__ total) {
    if (threadIdx.x == 0) {
        if (order) order[0] = root;
        lvl[0] = 0;
        lvl[1] = 1;
        *total = 1;
    }
}

__global__ void k_level_close(uint32_t *__restrict__ lvl, int l, const uint32_t *__restrict__ total) {
    if (threadIdx.x == 0) {
        lvl[2 * l + 2] = lvl[2 * l + 1];
        lvl[2 * l + 3] = *total;
    }
}

__global__ void __launch_bounds__(256) k_refit_level(const int32_t *__restrict__ nodes_at_level, int count,
                                                     const int2 *__restrict__ child, const float4 *leaf_lo,
                                                     const float4 *leaf_hi, float4 *node_lo, float4 *node_hi) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const int32_t p = nodes_at_level[k];
    const int2 ch = child[p];
    float4 l0, h0, l1, h1;
    child_box(ch.x, node_lo, node_hi, leaf_lo, leaf_hi, &l0, &h0);
    child_box(ch.y, node_lo, node_hi, leaf_lo, leaf_hi, &l1, &h1);
    node_lo[p] = make_float4(fminf(l0.x, l1.x), fminf(l0.y, l1.y), fminf(l0.z, l1.z), 0.0f);
    node_hi[p] = make_float4(fmaxf(h0.x, h1.x), fmaxf(h0.y, h1.y), fmaxf(h0.z, h1.z), 0.0f);
}

__device__ __forceinline__ float half_area(float4 lo, float4 hi) {
    const float dx = hi.x - lo.x, dy = hi.y - lo.y, dz = hi.z - lo.z;
    return dx * dy + dx * dz + dy * dz;
}

struct CollapseArgs {
    const int2 *child;
    const int2 *range;
    const float4 *node_lo, *node_hi, *leaf_lo, *leaf_hi;
    int32_t *q;              // binary node of each wide node, by wide node index
    const uint32_t *lvl;     // wide nodes of tree level d: [lvl[2d], lvl[2d + 1])
    int level;
    uint32_t *node_counter;
    void *nodes;
    int leaf_size;
    float pad;
    unsigned long long *leaf_counter;
};

template <int W>
__global__ void __launch_bounds__(256) k_collapse(CollapseArgs a) {
    const uint32_t lo = a.lvl[2 * a.level], hi = a.lvl[2 * a.level + 1];
    const int64_t idx = static_cast<int64_t>(lo) + static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= hi) return;
    const int2 job = make_int2(a.q[idx], static_cast<int32_t>(idx));
    int32_t c[W];
    float area[W];
    int cnt = 2;
    const int2 ch = a.child[job.x];
    c[0] = ch.x;
    c[1] = ch.y;
    auto count_of = [&](int32_t x) -> int {
        if (x < 0) return 1;
        const int2 r = a.range[x];
        return r.y - r.x + 1;
    };
    auto area_of = [&](int32_t x) -> float {
        if (x < 0) return -1.0f;  // a single triangle is never opened
        if (count_of(x) <= a.leaf_size) return -1.0f;
        return half_area(a.node_lo[x], a.node_hi[x]);
    };
    area[0] = area_of(c[0]);
    area[1] = area_of(c[1]);
    while (cnt < W) {
        int best = -1;
        float ba = -1.0f;
        for (int k = 0; k < cnt; ++k)
            if (area[k] > ba) {
                ba = area[k];
                best = k;
            }
        if (best < 0) break;
        const int2 cc = a.child[c[best]];
        c[best] = cc.x;
        area[best] = area_of(cc.x);
        c[cnt] = cc.y;
        area[cnt] = area_of(cc.y);
        ++cnt;
    }
    WideNode<W> *node = reinterpret_cast<WideNode<W> *>(a.nodes) + job.y;
    WideNode<W> out;
    int leaves = 0;
    for (int k = 0; k < W; ++k) {
        if (k >= cnt) {
            out.lo_x[k] = out.lo_y[k] = out.lo_z[k] = INFINITY;  // inverted: never hit
            out.hi_x[k] = out.hi_y[k] = out.hi_z[k] = -INFINITY;
            out.child[k] = kEmpty;
            continue;
        }
        const int32_t x = c[k];
        float4 lo, hi;
        child_box(x, a.node_lo, a.node_hi, a.leaf_lo, a.leaf_hi, &lo, &hi);
        out.lo_x[k] = lo.x - a.pad;
        out.lo_y[k] = lo.y - a.pad;
        out.lo_z[k] = lo.z - a.pad;
        out.hi_x[k] = hi.x + a.pad;
        out.hi_y[k] = hi.y + a.pad;
        out.hi_z[k] = hi.z + a.pad;
        if (x < 0) {
            out.child[k] = leaf_code(~x, 1);
            ++leaves;
        } else if (count_of(x) <= a.leaf_size) {
            const int2 r = a.range[x];
            out.child[k] = leaf_code(r.x, r.y - r.x + 1);
            ++leaves;
        } else {
            const uint32_t nidx = atomicAdd(a.node_counter, 1u);
            out.child[k] = static_cast<int32_t>(nidx);
            a.q[nidx] = x;
        }
    }
    for (int k = 0; k < (W == 4 ? 4 : 8); ++k) out.pad[k] = 0;
    *node = out;
    if (leaves) atomicAdd(a.leaf_counter, static_cast<unsigned long long>(leaves));
}

__global__ void __launch_bounds__(256) k_leaf_tris(const float *__restrict__ abc, const uint8_t *__restrict__ flags,
                                                   const uint32_t *__restrict__ sorted_idx, int64_t n,
                                                   LeafTri *__restrict__ out) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t i = sorted_idx[k];
    const float *t = abc + 9 * static_cast<int64_t>(i);
    const rtd::V3 a = rtd::mk(t[0], t[1], t[2]);
    const rtd::V3 b = rtd::mk(t[3], t[4], t[5]);
    const rtd::V3 c = rtd::mk(t[6], t[7], t[8]);
    // rt/triangle.h:193-194: e1 = c - a, e2 = b - a.
    const rtd::V3 e1 = rtd::sub(c, a);
    const rtd::V3 e2 = rtd::sub(b, a);
    LeafTri lt;
    lt.a_id = make_float4(a.x, a.y, a.z, __int_as_float(static_cast<int32_t>(i)));
    // Flags ride in e1.w: bit0 light, bit1 transparent (skipped by any hit).
    lt.e1 = make_float4(e1.x, e1.y, e1.z, __int_as_float(flags ? static_cast<int32_t>(flags[i]) : 0));
    lt.e2 = make_float4(e2.x, e2.y, e2.z, 0.0f);
    out[k] = lt;
}


// ---------------------------------------------------------------------------
// Binned-SAH refinement of the LBVH's upper levels.
//
// The LBVH is cut into clusters: subtrees of at most kSahCluster triangles
// whose parent holds more. The levels above the cut are rebuilt top-down with
// a binned SAH over the clusters (one pass per level over all clusters: node
// centroid bounds, bins of count / triangles / box per axis, best plane by
// A_L * T_L + A_R * T_R, stable partition by a scan), so that splits follow
// surface area instead of Morton-code midpoints. Nodes of at most
// kSahSweepItems clusters are split by an exact sweep instead (every cluster
// centroid is a candidate plane on every axis, one wave per node and one
// candidate per lane). The
// triangles are then laid out in the new cluster order, so
// every node of the combined tree covers a contiguous range and the collapse
// can still make leaves of ranges. SAH nodes are appended after the LBVH's
// internal nodes in the binary arrays; the tree's root becomes the SAH root.
//
// Cut size, measured on the 1M-triangle terrain frame (BVH4, leaf 2): the
// LBVH alone traces 3190 Mrays/s; clusters of 8 triangles 3470; single
// triangles, i.e. the whole tree rebuilt by SAH over the Morton-ordered
// triangles, 3750-3780, for a build 6-7x the LBVH's (30-31 ms against 4.7 ms
// then; 8.7 ms against 1.8 ms since the level loops stay on the device).
constexpr int kSahCluster = 1;
constexpr int kSahSweepItems = 64;  // one item per lane of the sweeping wave
constexpr int kSahBinsMax = 32;
constexpr int kSahBinWords = 8;            // count, triangles, lo xyz, hi xyz (lo/hi order-mapped)
constexpr int32_t kSahRef = 1 << 30;       // child ref of SAH node s: kSahRef | s
constexpr int32_t kNoCluster = INT32_MIN;

struct SahNode {
    int32_t begin, end;   // item range
    int32_t id;           // SAH node index
    int32_t slot;         // bin slot (binned split) or -1
};

struct SahSplit {
    int32_t axis;         // 0..2 a plane, -1 middle of the items
    int32_t bin;          // binned plane: items with bin <= this go left; -1: plane below
    int32_t nl;           // items on the left
    float plane;          // swept plane: items with centroid < plane go left
    int32_t next[2];      // next-level node of the left / right side, -1 for a single item
};

// Items per block of the binning passes: a block walks kSahChunk tiles of 256
// items and accumulates runs of tiles that belong to one node in LDS, so the
// top levels (a few nodes of up to n items) make one set of global atomics
// per block instead of per wave. Updates of a few words by every wave
// serialise across the eight XCDs' L2s: the per-wave form took 0.43 ms
// (bins) and 0.29 ms (centroid bounds) per level on average for 1M items.
constexpr int kSahChunk = 8;

__global__ void __launch_bounds__(256) k_fill_i32(int32_t *p, int64_t n, int32_t v) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void __launch_bounds__(256) k_fill_u32(uint32_t *p, int64_t n, uint32_t v) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__device__ __forceinline__ int range_count(const int2 *range, int32_t x) {
    if (x < 0) return 1;
    const int2 r = range[x];
    return r.y - r.x + 1;
}

// Cluster roots: children of nodes holding more than kSahCluster triangles
// that hold at most kSahCluster themselves; recorded at their first sorted
// position so that the clusters come out in Morton order.
__global__ void __launch_bounds__(256) k_find_clusters(const int2 *__restrict__ child, const int2 *__restrict__ range,
                                                       int64_t n, int32_t *__restrict__ cluster_at) {
    const int64_t x = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (x >= n - 1) return;
    if (range_count(range, static_cast<int32_t>(x)) <= kSahCluster) return;
    const int2 ch = child[x];
    const int32_t c2[2] = {ch.x, ch.y};
    for (int k = 0; k < 2; ++k) {
        const int32_t c = c2[k];
        if (c < 0)
            cluster_at[~c] = c;
        else if (range_count(range, c) <= kSahCluster)
            cluster_at[range[c].x] = c;
    }
}

__global__ void __launch_bounds__(256) k_cluster_flags(const int32_t *__restrict__ cluster_at, int64_t n,
                                                       uint32_t *__restrict__ flags) {
    const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p < n) flags[p] = cluster_at[p] != kNoCluster ? 1u : 0u;
}

// Items: lo (xyz, w = cluster ref), hi (xyz, w = triangle count).
__global__ void __launch_bounds__(256) k_make_items(const int32_t *__restrict__ cluster_at,
                                                    const uint32_t *__restrict__ pos, int64_t n,
                                                    const int2 *__restrict__ range, const float4 *__restrict__ node_lo,
                                                    const float4 *__restrict__ node_hi, const float4 *__restrict__ leaf_lo,
                                                    const float4 *__restrict__ leaf_hi, float4 *__restrict__ it_lo,
                                                    float4 *__restrict__ it_hi, int32_t *__restrict__ it_node) {
    const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int32_t c = cluster_at[p];
    if (c == kNoCluster) return;
    const uint32_t j = pos[p];
    float4 lo, hi;
    child_box(c, node_lo, node_hi, leaf_lo, leaf_hi, &lo, &hi);
    it_lo[j] = make_float4(lo.x, lo.y, lo.z, __int_as_float(c));
    it_hi[j] = make_float4(hi.x, hi.y, hi.z, __int_as_float(range_count(range, c)));
    it_node[j] = 0;
}

// Counters of a SAH level: ctr[0] bin slots handed out, ctr[1] nodes of the
// next level, ctr[2] SAH nodes, ctr[3] partition error (every level kernel
// returns at once when it is set). count: this level's node count.
__global__ void __launch_bounds__(256) k_sah_classify(SahNode *__restrict__ nodes, const uint32_t *__restrict__ count,
                                                      uint32_t *__restrict__ ctr, uint32_t *__restrict__ cb) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (ctr[3] || k >= static_cast<int>(*count)) return;
    SahNode nd = nodes[k];
    nd.slot = (nd.end - nd.begin > kSahSweepItems) ? static_cast<int32_t>(atomicAdd(&ctr[0], 1u)) : -1;
    nodes[k] = nd;
    if (nd.slot >= 0)
        for (int w = 0; w < 6; ++w) cb[6 * static_cast<int64_t>(nd.slot) + w] = w < 3 ? 0xffffffffu : 0u;
}

__device__ __forceinline__ float3 item_centroid(float4 lo, float4 hi) {
    return make_float3((lo.x + hi.x) * 0.5f, (lo.y + hi.y) * 0.5f, (lo.z + hi.z) * 0.5f);
}

// bin_add on an LDS copy of the bins (the builtins take an LDS pointer, so
// the updates are ds atomics that the workgroup barrier waits for).
__device__ __forceinline__ void bin_add_lds(uint32_t *w, uint32_t count, uint32_t tris, float lx, float ly, float lz,
                                            float hx, float hy, float hz) {
    typedef __attribute__((address_space(3))) uint32_t lds_u32;
    lds_u32 *l = (lds_u32 *)(w);
    __hip_atomic_fetch_add(&l[0], count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(&l[1], tris, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_min(&l[2], f2ord(lx), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_min(&l[3], f2ord(ly), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_min(&l[4], f2ord(lz), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_max(&l[5], f2ord(hx), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_max(&l[6], f2ord(hy), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_max(&l[7], f2ord(hz), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Runs of equal key in a wave: after the call the first lane of each run
// holds the run's minima (v[0..2]), maxima (v[3..5]) and sums (c[0..1]), and
// the result says whether this lane is such a first lane. Items arrive
// grouped by node and in Morton order within a node, so a wave's 64 items
// fall into a few runs of one (node, bin) and one atomic update per run
// replaces one per item (per-item updates of a node's few bins took up to
// 1.3 ms per level of 1M items).
// Lanes are matched by run index, not by key: one key can make several runs
// in a wave (bins alternate along a node's items), and matching keys across
// runs would count items twice.
__device__ __forceinline__ bool run_reduce(int32_t key, float (&v)[6], uint32_t (&c)[2]) {
    const int lane = threadIdx.x & 63;
    // Every lane takes part in the shuffle: written inside lane == 0 || ...,
    // it ran with lane 0 masked off, and lane 1 compared its key with a stale
    // register of the inactive lane 0 (runs then merged across nodes).
    const int32_t prev = __shfl_up(key, 1, 64);
    const bool head = (lane == 0) | (prev != key);
    const int run = __popcll(__ballot(head) & (~0ull >> (63 - lane)));  // heads at lanes <= this one
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int ro = __shfl_down(run, o, 64);
        const bool take = lane + o < 64 && ro == run;
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const float y = __shfl_down(v[k], o, 64);
            if (take) v[k] = k < 3 ? fminf(v[k], y) : fmaxf(v[k], y);
        }
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const uint32_t y = __shfl_down(c[k], o, 64);
            if (take) c[k] += y;
        }
    }
    return head;
}

// Centroid bounds of each binned node (6 order-mapped words per slot). Each
// wave reduces its runs of one node's items in registers; runs of tiles that
// all belong to one node then accumulate in LDS, others update the node's
// words directly.
__global__ void __launch_bounds__(256) k_sah_cbounds(const float4 *__restrict__ it_lo, const float4 *__restrict__ it_hi,
                                                     const int32_t *__restrict__ it_node, int64_t m,
                                                     const SahNode *__restrict__ nodes, uint32_t *__restrict__ cb,
                                                     const uint32_t *__restrict__ ctr) {
    __shared__ uint32_t lcb[6];
    if (ctr[3]) return;  // the same in every thread
    __shared__ int32_t tile_slot;
    int32_t cur = -1;  // node slot accumulated in lcb (the same in every thread)
    if (threadIdx.x < 6) lcb[threadIdx.x] = threadIdx.x < 3 ? 0xffffffffu : 0u;
    auto flush = [&](int32_t sl) {
        if (threadIdx.x < 6) {
            uint32_t *b = cb + 6 * static_cast<int64_t>(sl) + threadIdx.x;
            if (threadIdx.x < 3)
                atomicMin(b, lcb[threadIdx.x]);
            else
                atomicMax(b, lcb[threadIdx.x]);
            lcb[threadIdx.x] = threadIdx.x < 3 ? 0xffffffffu : 0u;
        }
    };
    for (int t = 0; t < kSahChunk; ++t) {
        const int64_t base = (static_cast<int64_t>(blockIdx.x) * kSahChunk + t) * blockDim.x;
        if (base >= m) break;
        const int64_t i = base + threadIdx.x;
        int32_t slot = -1;
        float3 c = make_float3(0.0f, 0.0f, 0.0f);
        if (i < m) {
            const int32_t nd = it_node[i];
            if (nd >= 0) {
                slot = nodes[nd].slot;
                c = item_centroid(it_lo[i], it_hi[i]);
            }
        }
        if (threadIdx.x == 0) tile_slot = slot;
        __syncthreads();
        const int32_t ts = tile_slot;
        const bool block_uniform = __syncthreads_and(slot == ts) && ts >= 0;
        if (block_uniform && ts != cur) {
            if (cur >= 0) flush(cur);
            __syncthreads();
            cur = ts;
        }
        float v[6] = {c.x, c.y, c.z, c.x, c.y, c.z};
        uint32_t cnt[2] = {0u, 0u};
        if (run_reduce(slot, v, cnt) && slot >= 0) {
            if (block_uniform) {  // separate from the global path: see k_sah_bin
                typedef __attribute__((address_space(3))) uint32_t lds_u32;
                lds_u32 *b = (lds_u32 *)(lcb);
                for (int k = 0; k < 3; ++k)
                    __hip_atomic_fetch_min(&b[k], f2ord(v[k]), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                for (int k = 3; k < 6; ++k)
                    __hip_atomic_fetch_max(&b[k], f2ord(v[k]), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            } else {
                uint32_t *b = cb + 6 * static_cast<int64_t>(slot);
                for (int k = 0; k < 3; ++k) atomicMin(&b[k], f2ord(v[k]));
                for (int k = 3; k < 6; ++k) atomicMax(&b[k], f2ord(v[k]));
            }
        }
        __syncthreads();
    }
    if (cur >= 0) flush(cur);
}

__device__ __forceinline__ int sah_bin_of(float c, float lo, float hi, int nbins) {
    const float ext = hi - lo;
    if (!(ext > 0.0f)) return 0;
    const int b = static_cast<int>((c - lo) * (static_cast<float>(nbins) / ext));
    return min(max(b, 0), nbins - 1);
}

// Bins per binned node: [slot][axis][bin][count, tris, lo xyz, hi xyz]. Runs
// of tiles whose items all belong to one node bin into LDS first and merge
// their non-empty bins with one set of global atomics. A node's bins are
// cleared by k_sah_split once read, so they start empty on every level.
__device__ __forceinline__ void bin_add(uint32_t *w, uint32_t count, uint32_t tris, float lx, float ly, float lz,
                                        float hx, float hy, float hz) {
    atomicAdd(&w[0], count);
    atomicAdd(&w[1], tris);
    atomicMin(&w[2], f2ord(lx));
    atomicMin(&w[3], f2ord(ly));
    atomicMin(&w[4], f2ord(lz));
    atomicMax(&w[5], f2ord(hx));
    atomicMax(&w[6], f2ord(hy));
    atomicMax(&w[7], f2ord(hz));
}

__global__ void __launch_bounds__(256) k_sah_bin(const float4 *__restrict__ it_lo, const float4 *__restrict__ it_hi,
                                                 const int32_t *__restrict__ it_node, int64_t m,
                                                 const SahNode *__restrict__ nodes, const uint32_t *__restrict__ cb,
                                                 int nbins, uint32_t *__restrict__ bins, const uint32_t *__restrict__ ctr) {
    __shared__ uint32_t lbins[3 * kSahBinsMax * kSahBinWords];
    if (ctr[3]) return;  // the same in every thread
    __shared__ int32_t tile_slot;
    const int words = 3 * nbins * kSahBinWords;
    for (int k = threadIdx.x; k < words; k += blockDim.x) {
        const int w = k % kSahBinWords;
        lbins[k] = (w >= 2 && w < 5) ? 0xffffffffu : 0u;
    }
    int32_t cur = -1;  // node slot binned in lbins (the same in every thread)
    auto flush = [&](int32_t sl) {  // LDS bins -> global, then cleared
        for (int k = threadIdx.x; k < 3 * nbins; k += blockDim.x) {
            uint32_t *l = lbins + k * kSahBinWords;
            if (l[0] == 0) continue;
            uint32_t *g = bins + (static_cast<int64_t>(sl) * 3 * nbins + k) * kSahBinWords;
            atomicAdd(&g[0], l[0]);
            atomicAdd(&g[1], l[1]);
            for (int w = 2; w < 5; ++w) atomicMin(&g[w], l[w]);
            for (int w = 5; w < 8; ++w) atomicMax(&g[w], l[w]);
            l[0] = l[1] = 0u;
            l[2] = l[3] = l[4] = 0xffffffffu;
            l[5] = l[6] = l[7] = 0u;
        }
    };
    for (int t = 0; t < kSahChunk; ++t) {
        const int64_t base = (static_cast<int64_t>(blockIdx.x) * kSahChunk + t) * blockDim.x;
        if (base >= m) break;
        const int64_t i = base + threadIdx.x;
        int32_t slot = -1;
        if (i < m) {
            const int32_t nd = it_node[i];
            if (nd >= 0) slot = nodes[nd].slot;
        }
        if (threadIdx.x == 0) tile_slot = slot;
        __syncthreads();
        const int32_t ts = tile_slot;
        const bool uniform = __syncthreads_and(slot == ts) && ts >= 0;
        if (uniform && ts != cur) {
            if (cur >= 0) flush(cur);
            __syncthreads();
            cur = ts;
        }
        float4 lo = make_float4(0.0f, 0.0f, 0.0f, 0.0f), hi = lo;
        float cc[3] = {0.0f, 0.0f, 0.0f};
        const uint32_t *b = cb;
        if (slot >= 0) {
            lo = it_lo[i];
            hi = it_hi[i];
            const float3 c = item_centroid(lo, hi);
            cc[0] = c.x;
            cc[1] = c.y;
            cc[2] = c.z;
            b = cb + 6 * static_cast<int64_t>(slot);
        }
        for (int a = 0; a < 3; ++a) {
            const int k = slot >= 0 ? sah_bin_of(cc[a], ord2f(b[a]), ord2f(b[3 + a]), nbins) : 0;
            float v[6] = {lo.x, lo.y, lo.z, hi.x, hi.y, hi.z};
            uint32_t c[2] = {1u, static_cast<uint32_t>(__float_as_int(hi.w))};
            const bool head = run_reduce(slot >= 0 ? slot * kSahBinsMax + k : -1, v, c);
            if (!head || slot < 0) continue;
            // LDS and global updates stay separate calls: through one generic
            // pointer they become flat atomics, and flat atomics that land in
            // LDS were not covered by the barrier's wait (items moved between
            // neighbouring nodes' bins across a flush).
            if (uniform)
                bin_add_lds(lbins + (a * nbins + k) * kSahBinWords, c[0], c[1], v[0], v[1], v[2], v[3], v[4], v[5]);
            else
                bin_add(bins + ((static_cast<int64_t>(slot) * 3 + a) * nbins + k) * kSahBinWords, c[0], c[1], v[0],
                        v[1], v[2], v[3], v[4], v[5]);
        }
        __syncthreads();
    }
    if (cur >= 0) flush(cur);
}

// One wave per node. A node of at most kSahSweepItems items is split by an
// exact sweep: lane j holds item j, its centroid is the candidate plane of the
// lane on each axis (items with centroid < plane go left), and every lane
// walks the node's items from LDS (broadcast reads) to price its candidates.
// The winner is the least cost, ties to the lowest (axis, item) pair, as a
// sequential scan over axes then items with a strict < would pick. Larger
// nodes take the best binned plane (lane 0 scans the bins, then clears them
// for the next level). With neither, the middle of the items. Sides with more
// than one item get an SAH node index and a place on the next level;
// single-item sides are linked by k_sah_scatter.
__device__ SahSplit decide_split(const SahNode &nd, const float4 *__restrict__ it_lo, const float4 *__restrict__ it_hi,
                                 uint32_t *__restrict__ bins, int nbins, float4 (*s_lo)[kSahSweepItems],
                                 float4 (*s_hi)[kSahSweepItems], int wave, int lane) {
    const int32_t size = nd.end - nd.begin;
    SahSplit sp{-1, 0, size / 2, 0.0f, {-1, -1}};
    if (nd.slot < 0) {
        float4 mylo = make_float4(0.0f, 0.0f, 0.0f, 0.0f), myhi = mylo;
        if (lane < size) {
            mylo = it_lo[nd.begin + lane];
            myhi = it_hi[nd.begin + lane];
            s_lo[wave][lane] = mylo;
            s_hi[wave][lane] = myhi;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        const float3 mc = item_centroid(mylo, myhi);
        float best = INFINITY;
        int best_key = 0x7fffffff;  // axis * 64 + item
        int32_t best_nl = 0;
        for (int a = 0; a < 3; ++a) {
            const float v = a == 0 ? mc.x : (a == 1 ? mc.y : mc.z);
            float llx = INFINITY, lly = INFINITY, llz = INFINITY, lhx = -INFINITY, lhy = -INFINITY, lhz = -INFINITY;
            float rlx = INFINITY, rly = INFINITY, rlz = INFINITY, rhx = -INFINITY, rhy = -INFINITY, rhz = -INFINITY;
            uint32_t lt = 0, rt_ = 0;
            int32_t nl = 0;
            for (int32_t q = 0; q < size; ++q) {
                const float4 lo = s_lo[wave][q], hi = s_hi[wave][q];
                const float3 cq = item_centroid(lo, hi);
                const float ca = a == 0 ? cq.x : (a == 1 ? cq.y : cq.z);
                const uint32_t t = static_cast<uint32_t>(__float_as_int(hi.w));
                if (ca < v) {
                    llx = fminf(llx, lo.x); lly = fminf(lly, lo.y); llz = fminf(llz, lo.z);
                    lhx = fmaxf(lhx, hi.x); lhy = fmaxf(lhy, hi.y); lhz = fmaxf(lhz, hi.z);
                    lt += t;
                    ++nl;
                } else {
                    rlx = fminf(rlx, lo.x); rly = fminf(rly, lo.y); rlz = fminf(rlz, lo.z);
                    rhx = fmaxf(rhx, hi.x); rhy = fmaxf(rhy, hi.y); rhz = fmaxf(rhz, hi.z);
                    rt_ += t;
                }
            }
            if (lane >= size || nl == 0 || nl == size) continue;
            const float ldx = lhx - llx, ldy = lhy - lly, ldz = lhz - llz;
            const float rdx = rhx - rlx, rdy = rhy - rly, rdz = rhz - rlz;
            const float cost = (ldx * ldy + ldx * ldz + ldy * ldz) * static_cast<float>(lt) +
                               (rdx * rdy + rdx * rdz + rdy * rdz) * static_cast<float>(rt_);
            if (cost < best) {  // per lane the axes come in order, so the first win stays
                best = cost;
                best_key = a * 64 + lane;
                best_nl = nl;
            }
        }
        // Least (cost, key) over the wave.
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const float oc = __shfl_xor(best, o, 64);
            const int ok = __shfl_xor(best_key, o, 64);
            const int on = __shfl_xor(best_nl, o, 64);
            if (oc < best || (oc == best && ok < best_key)) {
                best = oc;
                best_key = ok;
                best_nl = on;
            }
        }
        if (best_key != 0x7fffffff) {
            const int a = best_key >> 6, j = best_key & 63;
            const float4 lo = s_lo[wave][j], hi = s_hi[wave][j];
            const float3 cj = item_centroid(lo, hi);
            sp.axis = a;
            sp.bin = -1;
            sp.nl = best_nl;
            sp.plane = a == 0 ? cj.x : (a == 1 ? cj.y : cj.z);
        }
    } else {
        // Lane b holds bin b; inclusive prefix and suffix scans over the lanes
        // give every plane's left and right count, triangles and box.
        float best = INFINITY;
        int best_key = 0x7fffffff;  // axis * 64 + bin
        int32_t best_nl = 0;
        for (int a = 0; a < 3; ++a) {
            uint32_t c = 0, t = 0;
            float b6[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
            if (lane < nbins) {
                uint32_t *w = bins + ((static_cast<int64_t>(nd.slot) * 3 + a) * nbins + lane) * kSahBinWords;
                c = w[0];
                t = w[1];
                if (c)
                    for (int q = 0; q < 6; ++q) b6[q] = ord2f(w[2 + q]);
                w[0] = w[1] = 0u;  // cleared for the next level
                w[2] = w[3] = w[4] = 0xffffffffu;
                w[5] = w[6] = w[7] = 0u;
            }
            uint32_t lc = c, lt = t, rc = c, rt_ = t;
            float l6[6], r6[6];
            for (int q = 0; q < 6; ++q) l6[q] = r6[q] = b6[q];
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t yc = __shfl_up(lc, o, 64), yt = __shfl_up(lt, o, 64);
                const uint32_t zc = __shfl_down(rc, o, 64), zt = __shfl_down(rt_, o, 64);
                const bool up = lane >= o, down = lane + o < 64;
                if (up) {
                    lc += yc;
                    lt += yt;
                }
                if (down) {
                    rc += zc;
                    rt_ += zt;
                }
#pragma unroll
                for (int q = 0; q < 6; ++q) {
                    const float yl = __shfl_up(l6[q], o, 64), zr = __shfl_down(r6[q], o, 64);
                    if (up) l6[q] = q < 3 ? fminf(l6[q], yl) : fmaxf(l6[q], yl);
                    if (down) r6[q] = q < 3 ? fminf(r6[q], zr) : fmaxf(r6[q], zr);
                }
            }
            // The right side of plane b is the suffix from bin b + 1.
            const uint32_t nrt = __shfl_down(rt_, 1, 64);
            float n6[6];
#pragma unroll
            for (int q = 0; q < 6; ++q) n6[q] = __shfl_down(r6[q], 1, 64);
            if (lane >= nbins - 1 || lc == 0 || lc == static_cast<uint32_t>(size)) continue;
            const float ldx = l6[3] - l6[0], ldy = l6[4] - l6[1], ldz = l6[5] - l6[2];
            const float rdx = n6[3] - n6[0], rdy = n6[4] - n6[1], rdz = n6[5] - n6[2];
            const float cost = (ldx * ldy + ldx * ldz + ldy * ldz) * static_cast<float>(lt) +
                               (rdx * rdy + rdx * rdz + rdy * rdz) * static_cast<float>(nrt);
            if (cost < best) {
                best = cost;
                best_key = a * 64 + lane;
                best_nl = static_cast<int32_t>(lc);
            }
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const float oc = __shfl_xor(best, o, 64);
            const int ok = __shfl_xor(best_key, o, 64);
            const int on = __shfl_xor(best_nl, o, 64);
            if (oc < best || (oc == best && ok < best_key)) {
                best = oc;
                best_key = ok;
                best_nl = on;
            }
        }
        if (best_key != 0x7fffffff) {
            sp.axis = best_key >> 6;
            sp.bin = best_key & 63;
            sp.nl = best_nl;
        }
    }
    return sp;
}

// The level's node count is only known on the device, so the grid is fixed
// and every wave walks the nodes with a stride (a grid sized for the upper
// bound of 2^level nodes launched up to a million mostly idle waves). Each
// node's children of more than one item are counted into kids[k] (zeros up to
// the bound, which the scan covers); their SAH node indices and next-level
// places come from the scan of kids in k_sah_link. A pair of atomics per node
// on two shared counters serialised across the XCDs and cost 7 ms per level
// on the deep levels.
constexpr int kSplitBlocks = 2048;

__global__ void __launch_bounds__(256) k_sah_split(const SahNode *__restrict__ nodes, const uint32_t *__restrict__ count,
                                                   int64_t bound, const float4 *__restrict__ it_lo,
                                                   const float4 *__restrict__ it_hi, uint32_t *__restrict__ bins,
                                                   int nbins, SahSplit *__restrict__ splits,
                                                   uint32_t *__restrict__ kids, const uint32_t *__restrict__ ctr) {
    __shared__ float4 s_lo[4][kSahSweepItems], s_hi[4][kSahSweepItems];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (ctr[3]) return;
    const int n = static_cast<int>(*count);
    for (int k = blockIdx.x * 4 + wave; k < n; k += gridDim.x * 4) {  // the same in every lane of the wave
        const SahNode nd = nodes[k];
        const SahSplit sp = decide_split(nd, it_lo, it_hi, bins, nbins, s_lo, s_hi, wave, lane);
        if (lane == 0) {
            const int32_t size = nd.end - nd.begin;
            splits[k] = sp;
            kids[k] = static_cast<uint32_t>((sp.nl > 1) + (size - sp.nl > 1));
        }
    }
    for (int64_t k = n + static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; k < bound;
         k += static_cast<int64_t>(gridDim.x) * blockDim.x)
        kids[k] = 0u;
}

// Links each node to its children (kids holds the exclusive scan): side s of
// node k with more than one item becomes SAH node ctr[2] + kids[k] + s' and
// next-level node kids[k] + s' (s' counts such sides before it). The last
// node leaves the next level's size in ctr[1].
__global__ void __launch_bounds__(256) k_sah_link(const SahNode *__restrict__ nodes, const uint32_t *__restrict__ count,
                                                  SahSplit *__restrict__ splits, const uint32_t *__restrict__ kids,
                                                  SahNode *__restrict__ next, uint32_t *__restrict__ ctr,
                                                  int2 *__restrict__ sah_child, int2 *__restrict__ sah_items) {
    if (ctr[3]) return;
    const int n = static_cast<int>(*count);
    const uint32_t id_base = ctr[2];
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const SahNode nd = nodes[k];
        SahSplit sp = splits[k];
        const int32_t size = nd.end - nd.begin;
        const int32_t sizes[2] = {sp.nl, size - sp.nl};
        const int32_t begins[2] = {nd.begin, nd.begin + sp.nl};
        uint32_t q = kids[k];
        sah_items[nd.id] = make_int2(nd.begin, nd.end);
        int2 ch = make_int2(0, 0);
        for (int side = 0; side < 2; ++side) {
            if (sizes[side] <= 1) continue;  // linked by the scatter
            const int32_t id = static_cast<int32_t>(id_base + q);
            next[q] = SahNode{begins[side], begins[side] + sizes[side], id, -1};
            sp.next[side] = static_cast<int32_t>(q);
            if (side == 0)
                ch.x = kSahRef | id;
            else
                ch.y = kSahRef | id;
            ++q;
        }
        splits[k] = sp;
        // Sides of one item are written by the scatter; keep the other half.
        if (sizes[0] > 1) sah_child[nd.id].x = ch.x;
        if (sizes[1] > 1) sah_child[nd.id].y = ch.y;
        if (k == n - 1) ctr[1] = q;
    }
}

__device__ __forceinline__ bool sah_goes_left(const SahNode &nd, const SahSplit &sp, const uint32_t *cb, int nbins,
                                              float4 lo, float4 hi, int64_t i) {
    if (sp.axis < 0) return (i - nd.begin) < sp.nl;
    const float3 c = item_centroid(lo, hi);
    const float cc = sp.axis == 0 ? c.x : (sp.axis == 1 ? c.y : c.z);
    if (sp.bin < 0) return cc < sp.plane;
    const uint32_t *b = cb + 6 * static_cast<int64_t>(nd.slot);
    return sah_bin_of(cc, ord2f(b[sp.axis]), ord2f(b[3 + sp.axis]), nbins) <= sp.bin;
}

// Node index of each item on this level: it_node (items of finished nodes -1).
__global__ void __launch_bounds__(256) k_sah_flags(const float4 *__restrict__ it_lo, const float4 *__restrict__ it_hi,
                                                   const int32_t *__restrict__ it_node, int64_t m,
                                                   const SahNode *__restrict__ nodes, const SahSplit *__restrict__ splits,
                                                   const uint32_t *__restrict__ cb, int nbins,
                                                   uint32_t *__restrict__ flags, const uint32_t *__restrict__ ctr) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (ctr[3] || i >= m) return;
    const int32_t nd = it_node[i];
    bool left = false;
    if (nd >= 0) left = sah_goes_left(nodes[nd], splits[nd], cb, nbins, it_lo[i], it_hi[i], i);
    flags[i] = left ? 1u : 0u;
}

// Stable partition of each node's items (flags holds the exclusive scan of
// the left flags); sides of one item are linked to their parent here.
__global__ void __launch_bounds__(256) k_sah_scatter(const float4 *__restrict__ it_lo, const float4 *__restrict__ it_hi,
                                                     const int32_t *__restrict__ it_node, int64_t m,
                                                     const SahNode *__restrict__ nodes,
                                                     const SahSplit *__restrict__ splits, const uint32_t *__restrict__ cb,
                                                     int nbins, const uint32_t *__restrict__ scanned,
                                                     float4 *__restrict__ out_lo, float4 *__restrict__ out_hi,
                                                     int32_t *__restrict__ out_node, int2 *__restrict__ sah_child,
                                                     uint32_t *__restrict__ ctr) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (ctr[3] || i >= m) return;
    const float4 lo = it_lo[i], hi = it_hi[i];
    const int32_t ndi = it_node[i];
    if (ndi < 0) {
        out_lo[i] = lo;
        out_hi[i] = hi;
        out_node[i] = -1;
        return;
    }
    const SahNode nd = nodes[ndi];
    const SahSplit sp = splits[ndi];
    const bool left = sah_goes_left(nd, sp, cb, nbins, lo, hi, i);
    const int64_t rl = static_cast<int64_t>(scanned[i]) - static_cast<int64_t>(scanned[nd.begin]);
    const int64_t dest = left ? nd.begin + rl : nd.begin + sp.nl + (i - nd.begin - rl);
    // The split's left count comes from the bins and the sides from the
    // flags; a disagreement would move items out of their node's range.
    if (left ? dest >= nd.begin + sp.nl : (dest < nd.begin + sp.nl || dest >= nd.end)) {
        atomicOr(&ctr[3], 1u);
        return;
    }
    out_lo[dest] = lo;
    out_hi[dest] = hi;
    const int32_t side_size = left ? sp.nl : (nd.end - nd.begin - sp.nl);
    if (side_size == 1) {
        out_node[dest] = -1;
        const int32_t ref = __float_as_int(lo.w);
        if (left)
            sah_child[nd.id].x = ref;
        else
            sah_child[nd.id].y = ref;
    } else {
        out_node[dest] = sp.next[left ? 0 : 1];
    }
}

// Empty bins for every slot, once per build (k_sah_split clears what it reads).
__global__ void __launch_bounds__(256) k_sah_reset(uint32_t *__restrict__ bins, int64_t words) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= words) return;
    const int w = static_cast<int>(i % kSahBinWords);
    bins[i] = (w >= 2 && w < 5) ? 0xffffffffu : 0u;
}

__global__ void k_sah_close(uint32_t *__restrict__ ctr, uint32_t *__restrict__ count, int l) {
    if (threadIdx.x == 0) {
        count[l + 1] = ctr[1];
        ctr[2] += ctr[1];  // SAH node indices handed out on this level
        ctr[0] = 0;
        ctr[1] = 0;
    }
}

__global__ void __launch_bounds__(256) k_item_tris(const float4 *__restrict__ it_hi, int64_t m,
                                                   uint32_t *__restrict__ out) {
    const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j < m) out[j] = static_cast<uint32_t>(__float_as_int(it_hi[j].w));
}

// New sorted position of every triangle: clusters in their final order.
__global__ void __launch_bounds__(256) k_cluster_newpos(const float4 *__restrict__ it_lo, const float4 *__restrict__ it_hi,
                                                        const uint32_t *__restrict__ start, int64_t m,
                                                        const int2 *__restrict__ range, int32_t *__restrict__ newpos) {
    const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= m) return;
    const int32_t c = __float_as_int(it_lo[j].w);
    const int32_t t = __float_as_int(it_hi[j].w);
    const int32_t b = c < 0 ? ~c : range[c].x;
    for (int32_t k = 0; k < t; ++k) newpos[b + k] = static_cast<int32_t>(start[j]) + k;
}

__global__ void __launch_bounds__(256) k_permute_leaves(const int32_t *__restrict__ newpos, int64_t n,
                                                        const uint32_t *__restrict__ idx, const float4 *__restrict__ llo,
                                                        const float4 *__restrict__ lhi, uint32_t *__restrict__ idx_out,
                                                        float4 *__restrict__ llo_out, float4 *__restrict__ lhi_out) {
    const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int32_t q = newpos[p];
    idx_out[q] = idx[p];
    llo_out[q] = llo[p];
    lhi_out[q] = lhi[p];
}

// LBVH nodes inside clusters: leaf children and ranges in the new order
// (nodes above the cut are no longer referenced).
__global__ void __launch_bounds__(256) k_fix_lbvh(int2 *__restrict__ child, int2 *__restrict__ range, int64_t n,
                                                  const int32_t *__restrict__ newpos) {
    const int64_t x = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (x >= n - 1) return;
    int2 ch = child[x];
    if (ch.x < 0) ch.x = ~newpos[~ch.x];
    if (ch.y < 0) ch.y = ~newpos[~ch.y];
    child[x] = ch;
    const int2 r = range[x];
    range[x] = make_int2(newpos[r.x], newpos[r.y]);
}

__device__ __forceinline__ int32_t sah_resolve(int32_t ref, int64_t n, const int32_t *newpos) {
    if (ref < 0) return ~newpos[~ref];
    if (ref & kSahRef) return static_cast<int32_t>(n - 1) + (ref & ~kSahRef);
    return ref;
}

__global__ void __launch_bounds__(256) k_write_sah(const int2 *__restrict__ sah_child, const int2 *__restrict__ sah_items,
                                                   int count, int64_t n, const int32_t *__restrict__ newpos,
                                                   const uint32_t *__restrict__ start, const float4 *__restrict__ it_hi,
                                                   int2 *__restrict__ child, int2 *__restrict__ range) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const int2 sc = sah_child[k];
    const int2 it = sah_items[k];
    const int64_t x = n - 1 + k;
    child[x] = make_int2(sah_resolve(sc.x, n, newpos), sah_resolve(sc.y, n, newpos));
    const int32_t last = it.y - 1;
    range[x] = make_int2(static_cast<int32_t>(start[it.x]),
                         static_cast<int32_t>(start[last]) + __float_as_int(it_hi[last].w) - 1);
}

hipError_t exclusive_scan_u32(uint32_t *data, int64_t n, uint32_t *tile_sums, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const int64_t blocks = (n + kScanTile - 1) / kScanTile;
    if (blocks > kScanTile) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_scan_tiles, dim3(static_cast<unsigned>(blocks)), dim3(kScanThreads), 0, s, data, n, tile_sums);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanThreads), 0, s, tile_sums, static_cast<int>(blocks));
    hipLaunchKernelGGL(k_scan_add, dim3(static_cast<unsigned>(blocks)), dim3(kScanThreads), 0, s, data, n, tile_sums);
    return hipGetLastError();
}

// A build's temporary buffer, carved from the context's arena.
struct DevBuf {
    void *p = nullptr;
};

}  // namespace

hipError_t BuildArena::begin() {
    if (chunks.size() > 1 || (!chunks.empty() && chunks[0].second < high)) {
        release();
    }
    if (chunks.empty() && high > 0) {
        void *p = nullptr;
        const hipError_t e = hipMalloc(&p, high);
        if (e != hipSuccess) return e;
        chunks.push_back({p, high});
    }
    cur = 0;
    used = 0;
    taken = 0;
    return hipSuccess;
}

hipError_t BuildArena::take(void **p, size_t bytes) {
    bytes = (std::max<size_t>(1, bytes) + 255) & ~size_t(255);
    taken += bytes;
    high = std::max(high, taken);
    while (cur < chunks.size() && used + bytes > chunks[cur].second) {
        ++cur;
        used = 0;
    }
    if (cur == chunks.size()) {
        void *q = nullptr;
        const size_t size = std::max<size_t>(bytes, size_t(64) << 20);
        const hipError_t e = hipMalloc(&q, size);
        if (e != hipSuccess) return e;
        chunks.push_back({q, size});
        used = 0;
    }
    *p = static_cast<char *>(chunks[cur].first) + used;
    used += bytes;
    return hipSuccess;
}

void BuildArena::release() {
    for (auto &c : chunks) (void)hipFree(c.first);
    chunks.clear();
    cur = used = 0;
}

namespace {

template <class T>
hipError_t dalloc_in(BuildArena &ar, DevBuf &b, size_t count) {
    return ar.take(&b.p, count * sizeof(T));
}

}  // namespace

hipError_t make_leaf_tris(const float *abc, const uint8_t *flags, const uint32_t *order, int64_t n, LeafTri *out,
                          hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_leaf_tris, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, s, abc, flags, order,
                       n, out);
    return hipGetLastError();
}

int32_t build_accel(rt_scene *scene, const rt_build_params &p, rt_accel **out) {
    rt_ctx *ctx = scene->ctx;
    if (p.builder == RT_BUILDER_AAC) return build_accel_aac(scene, p, out);
    if (p.builder != RT_BUILDER_LBVH) return fail(ctx, RT_ERR_INVALID, "unknown builder");
    const int W = p.width;
    if (W != 4 && W != 8) return fail(ctx, RT_ERR_INVALID, "width must be 4 or 8");
    if (p.leaf_size < 1 || p.leaf_size > kMaxLeafSize) return fail(ctx, RT_ERR_INVALID, "leaf_size must be 1..16");
    const int64_t n = scene->n;
    if (n >= (int64_t(1) << 27)) return fail(ctx, RT_ERR_INVALID, "too many triangles (limit 2^27)");
    hipStream_t s = ctx->stream;
    const auto t0 = std::chrono::steady_clock::now();

    {
        const hipError_t e = ctx->arena.begin();
        if (e != hipSuccess) return fail(ctx, RT_ERR_OOM, std::string("build scratch: ") + hipGetErrorString(e));
    }
    rt_accel *a = new rt_accel();
    a->scene = scene;
    a->triangles = n;
    a->width = W;
    a->leaf_size = p.leaf_size;
    a->sah_bins = p.sah_bins;
    a->lds_stack = (p.lds_stack <= 0) ? 16 : (p.lds_stack <= 4 ? 4 : (p.lds_stack <= 8 ? 8 : 16));
    auto cleanup = [&]() {
        (void)hipStreamSynchronize(s);  // the arena's buffers may be reused once this returns
        if (a->d_nodes) (void)hipFree(a->d_nodes);
        if (a->d_tris) (void)hipFree(a->d_tris);
        delete a;
    };
#define BUILD_CHECK(expr)                                                                        \
    do {                                                                                         \
        hipError_t _e = (expr);                                                                  \
        if (_e != hipSuccess) {                                                                  \
            cleanup();                                                                           \
            return fail(ctx, _e == hipErrorOutOfMemory ? RT_ERR_OOM : RT_ERR_HIP,               \
                        std::string(#expr) + ": " + hipGetErrorString(_e));                     \
        }                                                                                        \
    } while (0)

    if (n == 0) {
        a->root = kEmpty;
        *out = a;
        return RT_OK;
    }

    DevBuf b_aabb_lo, b_aabb_hi, b_bounds, b_codes, b_idx, b_codes_t, b_idx_t, b_hist, b_sums;
    const int64_t tiles = (n + kSortTile - 1) / kSortTile;
    BUILD_CHECK(dalloc_in<float4>(ctx->arena, b_aabb_lo, n));
    BUILD_CHECK(dalloc_in<float4>(ctx->arena, b_aabb_hi, n));
    BUILD_CHECK(dalloc_in<uint32_t>(ctx->arena, b_bounds, 12));
    BUILD_CHECK(dalloc_in<uint32_t>(ctx->arena, b_codes, n));
    BUILD_CHECK(dalloc_in<uint32_t>(ctx->arena, b_idx, n));
    BUILD_CHECK(dalloc_in<uint32_t>(ctx->arena, b_codes_t, n));
    BUILD_CHECK(dalloc_in<uint32_t>(ctx->arena, b_idx_t, n));
    BUILD_CHECK(dalloc_in<uint32_t>(ctx->arena, b_hist, 256 * tiles));
    BUILD_CHECK(dalloc_in<uint32_t>(ctx->arena, b_sums, kScanTile));
    DevBuf b_parts;
    BUILD_CHECK(dalloc_in<uint32_t>(ctx->arena, b_parts, 12 * static_cast<int64_t>((n + 255) / 256)));
    const int blocks = static_cast<int>((n + 255) / 256);
    hipLaunchKernelGGL(k_tri_bounds, dim3(blocks), dim3(256), 0, s, scene->d_abc, n, (float4 *)b_aabb_lo.p,
                       (float4 *)b_aabb_hi.p, (uint32_t *)b_parts.p);
    hipLaunchKernelGGL(k_bounds_final, dim3(1), dim3(256), 0, s, (const uint32_t *)b_parts.p, blocks,
                       (uint32_t *)b_bounds.p);
    hipLaunchKernelGGL(k_morton, dim3(blocks), dim3(256), 0, s, (const float4 *)b_aabb_lo.p,
                       (const float4 *)b_aabb_hi.p, n, (const uint32_t *)b_bounds.p, (uint32_t *)b_codes.p,
                       (uint32_t *)b_idx.p);
    BUILD_CHECK(hipGetLastError());
    BUILD_CHECK(radix_sort_pairs((uint32_t *)b_codes.p, (uint32_t *)b_idx.p, (uint32_t *)b_codes_t.p,
                                 (uint32_t *)b_idx_t.p, n, 30, (uint32_t *)b_hist.p, (uint32_t *)b_sums.p, s));
    uint32_t hb[12];
    BUILD_CHECK(hipMemcpyAsync(hb, b_bounds.p, sizeof(hb), hipMemcpyDeviceToHost, s));
    BUILD_CHECK(hipStreamSynchronize(s));
    float maxabs = 0.0f;
    for (int k = 6; k < 12; ++k) maxabs = std::max(maxabs, std::fabs(ord2f(hb[k])));
    a->box_pad = std::ldexp(std::max(maxabs, 1e-30f), -15);
    a->near_limit = 256.0f * maxabs;
    for (int k = 0; k < 3; ++k) {
        a->scene_lo[k] = ord2f(hb[6 + k]);
        const float ext = ord2f(hb[9 + k]) - a->scene_lo[k];
        a->scene_inv[k] = ext > 0.0f ? 1.0f / ext : 0.0f;
    }

    if (n == 1) {
        // A single triangle: the root is a leaf, no wide node is needed.
        BUILD_CHECK(hipMalloc(&a->d_tris, n * sizeof(LeafTri)));
        hipLaunchKernelGGL(k_leaf_tris, dim3(blocks), dim3(256), 0, s, scene->d_abc, scene->d_flags,
                           (const uint32_t *)b_idx.p, n, a->d_tris);
        BUILD_CHECK(hipGetLastError());
        a->root = leaf_code(0, 1);
        a->nodes = 0;
        a->leaves = 1;
        a->depth = 0;
        BUILD_CHECK(hipStreamSynchronize(s));
        a->build_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
        *out = a;
        return RT_OK;
    }

    const bool refine = p.sah_bins > 0 && n > kSahCluster;
    // Binary nodes: the LBVH's n - 1, then up to n - 1 SAH nodes.
    const int64_t bcap = (n - 1) + (refine ? n : 0);
    DevBuf b_child, b_range, b_parent, b_leaf_parent, b_leaf_lo, b_leaf_hi, b_node_lo, b_node_hi;
    BUILD_CHECK(dalloc_in<int2>(ctx->arena, b_child, bcap));
    BUILD_CHECK(dalloc_in<int2>(ctx->arena, b_range, bcap));
    BUILD_CHECK(dalloc_in<int32_t>(ctx->arena, b_parent, n - 1));
    BUILD_CHECK(dalloc_in<int32_t>(ctx->arena, b_leaf_parent, n));
    BUILD_CHECK(dalloc_in<float4>(ctx->arena, b_leaf_lo, n));
    BUILD_CHECK(dalloc_in<float4>(ctx->arena, b_leaf_hi, n));
    BUILD_CHECK(dalloc_in<float4>(ctx->arena, b_node_lo, bcap));
    BUILD_CHECK(dalloc_in<float4>(ctx->arena, b_node_hi, bcap));
    const int iblocks = static_cast<int>((n - 1 + 255) / 256);
    hipLaunchKernelGGL(k_karras, dim3(iblocks), dim3(256), 0, s, (const uint32_t *)b_codes.p, n, (int2 *)b_child.p,
                       (int2 *)b_range.p, (int32_t *)b_parent.p, (int32_t *)b_leaf_parent.p);
    hipLaunchKernelGGL(k_gather_leaf_boxes, dim3(blocks), dim3(256), 0, s, (const uint32_t *)b_idx.p, n,
                       (const float4 *)b_aabb_lo.p, (const float4 *)b_aabb_hi.p, (float4 *)b_leaf_lo.p,
                       (float4 *)b_leaf_hi.p);
    // Level-synchronous refit from a root (see k_bfs_level): the levels are
    // found top-down in groups of kLevelGroup launches per host read-back,
    // then refit bottom-up.
    DevBuf b_order, b_lvl;
    BUILD_CHECK(dalloc_in<int32_t>(ctx->arena, b_order, n - 1));
    BUILD_CHECK(dalloc_in<uint32_t>(ctx->arena, b_lvl, 2 * (kMaxLevels + kLevelGroup + 2) + 1));
    uint32_t *lvl = (uint32_t *)b_lvl.p;
    uint32_t *lvl_total = lvl + 2 * (kMaxLevels + kLevelGroup + 2);
    auto refit = [&](int32_t root) -> hipError_t {
        int32_t *order = (int32_t *)b_order.p;
        hipLaunchKernelGGL(k_level_init, dim3(1), dim3(64), 0, s, order, root, lvl, lvl_total);
        std::vector<uint32_t> hl;
        int levels = -1;
        for (int l = 0; levels < 0;) {
            if (l > kMaxLevels) return hipErrorInvalidValue;
            for (int g = 0; g < kLevelGroup; ++g, ++l) {
                const int64_t bound = std::min<int64_t>(l < 30 ? (int64_t(1) << l) : n, n - 1);
                hipLaunchKernelGGL(k_bfs_level, dim3(static_cast<unsigned>((bound + 255) / 256)), dim3(256), 0, s,
                                   order, lvl, l, (const int2 *)b_child.p, lvl_total);
                hipLaunchKernelGGL(k_level_close, dim3(1), dim3(64), 0, s, lvl, l, lvl_total);
            }
            hl.resize(2 * (l + 1));
            hipError_t e = hipMemcpyAsync(hl.data(), lvl, hl.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            if (e != hipSuccess) return e;
            for (int k = 0; k <= l; ++k)
                if (hl[2 * k] == hl[2 * k + 1]) {
                    levels = k;
                    break;
                }
        }
        for (int l = levels - 1; l >= 0; --l) {
            const uint32_t off = hl[2 * l], count = hl[2 * l + 1] - hl[2 * l];
            hipLaunchKernelGGL(k_refit_level, dim3((count + 255) / 256), dim3(256), 0, s, order + off,
                               static_cast<int>(count), (const int2 *)b_child.p, (const float4 *)b_leaf_lo.p,
                               (const float4 *)b_leaf_hi.p, (float4 *)b_node_lo.p, (float4 *)b_node_hi.p);
        }
        a->binary_depth = levels;
        return hipGetLastError();
    };
    {
        const hipError_t e = refit(0);
        if (e == hipErrorInvalidValue) {
            cleanup();
            return fail(ctx, RT_ERR_INVALID, "binary tree deeper than 256 levels");
        }
        BUILD_CHECK(e);
    }
    int32_t root = 0;
    uint32_t *tri_order = (uint32_t *)b_idx.p;  // sorted position -> triangle
    DevBuf b_idx2, b_llo2, b_lhi2;
    if (refine) {
        const int nbins = std::min(std::max(p.sah_bins, 4), kSahBinsMax);
        DevBuf b_cat, b_pos;
        BUILD_CHECK(dalloc_in<int32_t>(ctx->arena, b_cat, n));
        BUILD_CHECK(dalloc_in<uint32_t>(ctx->arena, b_pos, n));
        int32_t *cat = (int32_t *)b_cat.p;
        uint32_t *pos = (uint32_t *)b_pos.p;
        hipLaunchKernelGGL(k_fill_i32, dim3(blocks), dim3(256), 0, s, cat, n, kNoCluster);
        hipLaunchKernelGGL(k_find_clusters, dim3(iblocks), dim3(256), 0, s, (const int2 *)b_child.p,
                           (const int2 *)b_range.p, n, cat);
        hipLaunchKernelGGL(k_cluster_flags, dim3(blocks), dim3(256), 0, s, cat, n, pos);
        BUILD_CHECK(exclusive_scan_u32(pos, n, (uint32_t *)b_sums.p, s));
        int32_t last_cat = 0;
        uint32_t last_pos = 0;
        BUILD_CHECK(hipMemcpyAsync(&last_cat, cat + n - 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        BUILD_CHECK(hipMemcpyAsync(&last_pos, pos + n - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        BUILD_CHECK(hipStreamSynchronize(s));
        const int64_t m = static_cast<int64_t>(last_pos) + (last_cat != kNoCluster ? 1 : 0);
        if (m >= 2) {
            const int mblocks = static_cast<int>((m + 255) / 256);
            DevBuf b_lo[2], b_hi[2], b_nd[2], b_flags, b_nodes[2], b_splits, b_sc, b_si, b_ctr, b_cb, b_bins;
            for (int k = 0; k < 2; ++k) {
                BUILD_CHECK(dalloc_in<float4>(ctx->arena, b_lo[k], m));
                BUILD_CHECK(dalloc_in<float4>(ctx->arena, b_hi[k], m));
                BUILD_CHECK(dalloc_in<int32_t>(ctx->arena, b_nd[k], m));
                BUILD_CHECK(dalloc_in<SahNode>(ctx->arena, b_nodes[k], m));
            }
            BUILD_CHECK(dalloc_in<uint32_t>(ctx->arena, b_flags, m));
            BUILD_CHECK(dalloc_in<SahSplit>(ctx->arena, b_splits, m));
            DevBuf b_kids;  // children per node of a level, then their exclusive scan
            BUILD_CHECK(dalloc_in<uint32_t>(ctx->arena, b_kids, m));
            BUILD_CHECK(dalloc_in<int2>(ctx->arena, b_sc, m - 1));
            BUILD_CHECK(dalloc_in<int2>(ctx->arena, b_si, m - 1));
            BUILD_CHECK(dalloc_in<uint32_t>(ctx->arena, b_ctr, 4));  // [0] bin slots, [1] next-level nodes, [2] SAH nodes
            const int64_t max_slots = m / (kSahSweepItems + 1) + 1;
            BUILD_CHECK(dalloc_in<uint32_t>(ctx->arena, b_cb, 6 * max_slots));
            BUILD_CHECK(dalloc_in<uint32_t>(ctx->arena, b_bins, max_slots * 3 * nbins * kSahBinWords));
            hipLaunchKernelGGL(k_make_items, dim3(blocks), dim3(256), 0, s, cat, pos, n, (const int2 *)b_range.p,
                               (const float4 *)b_node_lo.p, (const float4 *)b_node_hi.p, (const float4 *)b_leaf_lo.p,
                               (const float4 *)b_leaf_hi.p, (float4 *)b_lo[0].p, (float4 *)b_hi[0].p,
                               (int32_t *)b_nd[0].p);
            const SahNode root_node{0, static_cast<int32_t>(m), 0, -1};
            BUILD_CHECK(hipMemcpyAsync(b_nodes[0].p, &root_node, sizeof(SahNode), hipMemcpyHostToDevice, s));
            uint32_t *ctr = (uint32_t *)b_ctr.p;
            const uint32_t init_ctr[4] = {0, 0, 1, 0};
            BUILD_CHECK(hipMemcpyAsync(ctr, init_ctr, sizeof(init_ctr), hipMemcpyHostToDevice, s));
            const int64_t bin_words = max_slots * 3 * nbins * kSahBinWords;
            hipLaunchKernelGGL(k_sah_reset, dim3(static_cast<unsigned>((bin_words + 255) / 256)), dim3(256), 0, s,
                               (uint32_t *)b_bins.p, bin_words);
            const int cblocks = static_cast<int>((m + 256 * kSahChunk - 1) / (256 * kSahChunk));
            // Level L reads nodes from b_nodes[L & 1] and items from b_lo/hi/nd[L & 1]
            // and writes the next level to the other buffers; its node count is
            // scount[L] on the device. Levels are launched in groups with one
            // read-back of the counts per group; levels after the last one
            // only carry the finished items over to the other buffers.
            constexpr int kMaxSahLevels = 4096;
            DevBuf b_scount;
            BUILD_CHECK(dalloc_in<uint32_t>(ctx->arena, b_scount, kMaxSahLevels + kLevelGroup + 2));
            uint32_t *scount = (uint32_t *)b_scount.p;
            const uint32_t one_node = 1;
            BUILD_CHECK(hipMemcpyAsync(scount, &one_node, sizeof(uint32_t), hipMemcpyHostToDevice, s));
            uint32_t *cb = (uint32_t *)b_cb.p, *bins = (uint32_t *)b_bins.p;
            int launched = 0;
            for (bool done = false; !done;) {
                if (launched > kMaxSahLevels) {
                    cleanup();
                    return fail(ctx, RT_ERR_INVALID, "SAH refinement did not converge");
                }
                for (int g = 0; g < kLevelGroup; ++g, ++launched) {
                    const int cur = launched & 1;
                    SahNode *nodes = (SahNode *)b_nodes[cur].p;
                    SahNode *next = (SahNode *)b_nodes[cur ^ 1].p;
                    const float4 *ilo = (const float4 *)b_lo[cur].p, *ihi = (const float4 *)b_hi[cur].p;
                    const int32_t *ind = (const int32_t *)b_nd[cur].p;
                    const uint32_t *count = scount + launched;
                    const int64_t bound = std::min<int64_t>(launched < 30 ? (int64_t(1) << launched) : m, m);
                    const unsigned nb256 = static_cast<unsigned>((bound + 255) / 256);
                    const unsigned nb4 = static_cast<unsigned>((bound + 3) / 4);
                    hipLaunchKernelGGL(k_sah_classify, dim3(nb256), dim3(256), 0, s, nodes, count, ctr, cb);
                    hipLaunchKernelGGL(k_sah_cbounds, dim3(cblocks), dim3(256), 0, s, ilo, ihi, ind, m, nodes, cb,
                                       (const uint32_t *)ctr);
                    hipLaunchKernelGGL(k_sah_bin, dim3(cblocks), dim3(256), 0, s, ilo, ihi, ind, m, nodes, cb, nbins,
                                       bins, (const uint32_t *)ctr);
                    hipLaunchKernelGGL(k_sah_split, dim3(std::min<unsigned>(nb4, kSplitBlocks)), dim3(256), 0, s,
                                       nodes, count, bound, ilo, ihi, bins, nbins, (SahSplit *)b_splits.p,
                                       (uint32_t *)b_kids.p, (const uint32_t *)ctr);
                    BUILD_CHECK(exclusive_scan_u32((uint32_t *)b_kids.p, bound, (uint32_t *)b_sums.p, s));
                    hipLaunchKernelGGL(k_sah_link, dim3(std::min<unsigned>(nb256, kSplitBlocks)), dim3(256), 0, s,
                                       (const SahNode *)nodes, count, (SahSplit *)b_splits.p,
                                       (const uint32_t *)b_kids.p, next, ctr, (int2 *)b_sc.p, (int2 *)b_si.p);
                    hipLaunchKernelGGL(k_sah_flags, dim3(mblocks), dim3(256), 0, s, ilo, ihi, ind, m, nodes,
                                       (const SahSplit *)b_splits.p, cb, nbins, (uint32_t *)b_flags.p,
                                       (const uint32_t *)ctr);
                    BUILD_CHECK(exclusive_scan_u32((uint32_t *)b_flags.p, m, (uint32_t *)b_sums.p, s));
                    hipLaunchKernelGGL(k_sah_scatter, dim3(mblocks), dim3(256), 0, s, ilo, ihi, ind, m, nodes,
                                       (const SahSplit *)b_splits.p, cb, nbins, (const uint32_t *)b_flags.p,
                                       (float4 *)b_lo[cur ^ 1].p, (float4 *)b_hi[cur ^ 1].p,
                                       (int32_t *)b_nd[cur ^ 1].p, (int2 *)b_sc.p, ctr);
                    hipLaunchKernelGGL(k_sah_close, dim3(1), dim3(64), 0, s, ctr, scount, launched);
                }
                BUILD_CHECK(hipGetLastError());
                std::vector<uint32_t> hs(launched + 1);
                uint32_t hc[4] = {0, 0, 0, 0};
                BUILD_CHECK(hipMemcpyAsync(hs.data(), scount, hs.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
                BUILD_CHECK(hipMemcpyAsync(hc, ctr, sizeof(hc), hipMemcpyDeviceToHost, s));
                BUILD_CHECK(hipStreamSynchronize(s));
                if (hc[3]) {
                    cleanup();
                    return fail(ctx, RT_ERR_HIP, "SAH refinement: inconsistent partition (internal error)");
                }
                for (uint32_t c : hs) done = done || c == 0;
            }
            const int cur = launched & 1;  // where the last level left the items
            // Triangles in the final cluster order.
            const float4 *flo = (const float4 *)b_lo[cur].p, *fhi = (const float4 *)b_hi[cur].p;
            (void)flo;
            DevBuf b_start, b_newpos;
            BUILD_CHECK(dalloc_in<uint32_t>(ctx->arena, b_start, m));
            BUILD_CHECK(dalloc_in<int32_t>(ctx->arena, b_newpos, n));
            uint32_t *start = (uint32_t *)b_start.p;
            int32_t *newpos = (int32_t *)b_newpos.p;
            hipLaunchKernelGGL(k_item_tris, dim3(mblocks), dim3(256), 0, s, fhi, m, start);
            BUILD_CHECK(exclusive_scan_u32(start, m, (uint32_t *)b_sums.p, s));
            hipLaunchKernelGGL(k_cluster_newpos, dim3(mblocks), dim3(256), 0, s, (const float4 *)b_lo[cur].p, fhi,
                               start, m, (const int2 *)b_range.p, newpos);
            BUILD_CHECK(dalloc_in<uint32_t>(ctx->arena, b_idx2, n));
            BUILD_CHECK(dalloc_in<float4>(ctx->arena, b_llo2, n));
            BUILD_CHECK(dalloc_in<float4>(ctx->arena, b_lhi2, n));
            hipLaunchKernelGGL(k_permute_leaves, dim3(blocks), dim3(256), 0, s, newpos, n, (const uint32_t *)b_idx.p,
                               (const float4 *)b_leaf_lo.p, (const float4 *)b_leaf_hi.p, (uint32_t *)b_idx2.p,
                               (float4 *)b_llo2.p, (float4 *)b_lhi2.p);
            BUILD_CHECK(hipMemcpyAsync(b_leaf_lo.p, b_llo2.p, n * sizeof(float4), hipMemcpyDeviceToDevice, s));
            BUILD_CHECK(hipMemcpyAsync(b_leaf_hi.p, b_lhi2.p, n * sizeof(float4), hipMemcpyDeviceToDevice, s));
            hipLaunchKernelGGL(k_fix_lbvh, dim3(iblocks), dim3(256), 0, s, (int2 *)b_child.p, (int2 *)b_range.p, n,
                               newpos);
            hipLaunchKernelGGL(k_write_sah, dim3(static_cast<unsigned>((m - 1 + 255) / 256)), dim3(256), 0, s,
                               (const int2 *)b_sc.p, (const int2 *)b_si.p, static_cast<int>(m - 1), n, newpos, start,
                               fhi, (int2 *)b_child.p, (int2 *)b_range.p);
            BUILD_CHECK(hipGetLastError());
            tri_order = (uint32_t *)b_idx2.p;
            root = static_cast<int32_t>(n - 1);
            const hipError_t e = refit(root);
            if (e == hipErrorInvalidValue) {
                cleanup();
                return fail(ctx, RT_ERR_INVALID, "binary tree deeper than 256 levels");
            }
            BUILD_CHECK(e);
        }
    }

    // Triangles in BVH order.
    BUILD_CHECK(hipMalloc(&a->d_tris, n * sizeof(LeafTri)));
    hipLaunchKernelGGL(k_leaf_tris, dim3(blocks), dim3(256), 0, s, scene->d_abc, scene->d_flags, tri_order, n,
                       a->d_tris);
    BUILD_CHECK(hipGetLastError());

    // Wide nodes: at most n - 1 (each wide node consumes >= 1 binary node).
    const size_t node_bytes = (W == 4) ? sizeof(WideNode<4>) : sizeof(WideNode<8>);
    const int64_t max_nodes = std::max<int64_t>(1, n - 1);
    BUILD_CHECK(hipMalloc(&a->d_nodes, max_nodes * node_bytes));
    // Wide nodes are numbered level by level (a level's nodes are handed out
    // while the level above is collapsed), so q[wide index] = binary node is
    // the work list and the levels are ranges of wide indices.
    DevBuf b_q, b_counters;
    BUILD_CHECK(dalloc_in<int32_t>(ctx->arena, b_q, max_nodes));
    BUILD_CHECK(dalloc_in<unsigned long long>(ctx->arena, b_counters, 4));
    BUILD_CHECK(hipMemsetAsync(b_counters.p, 0, 4 * sizeof(unsigned long long), s));
    uint32_t *node_counter = reinterpret_cast<uint32_t *>(b_counters.p);
    unsigned long long *leaf_counter = reinterpret_cast<unsigned long long *>(b_counters.p) + 1;
    hipLaunchKernelGGL(k_level_init, dim3(1), dim3(64), 0, s, (int32_t *)b_q.p, root, lvl, node_counter);
    int depth = -1;
    std::vector<uint32_t> hl;
    for (int d = 0; depth < 0;) {
        if (d > kMaxLevels) {
            cleanup();
            return fail(ctx, RT_ERR_INVALID, "wide tree deeper than 256 levels");
        }
        for (int g = 0; g < kLevelGroup; ++g, ++d) {
            CollapseArgs ca;
            ca.child = (const int2 *)b_child.p;
            ca.range = (const int2 *)b_range.p;
            ca.node_lo = (const float4 *)b_node_lo.p;
            ca.node_hi = (const float4 *)b_node_hi.p;
            ca.leaf_lo = (const float4 *)b_leaf_lo.p;
            ca.leaf_hi = (const float4 *)b_leaf_hi.p;
            ca.q = (int32_t *)b_q.p;
            ca.lvl = lvl;
            ca.level = d;
            ca.node_counter = node_counter;
            ca.nodes = a->d_nodes;
            ca.leaf_size = p.leaf_size;
            ca.pad = a->box_pad;
            ca.leaf_counter = leaf_counter;
            int64_t bound = 1;
            for (int k = 0; k < d && bound < max_nodes; ++k) bound *= W;
            bound = std::min<int64_t>(bound, max_nodes);
            const dim3 g_dim(static_cast<unsigned>((bound + 255) / 256));
            if (W == 4)
                hipLaunchKernelGGL(k_collapse<4>, g_dim, dim3(256), 0, s, ca);
            else
                hipLaunchKernelGGL(k_collapse<8>, g_dim, dim3(256), 0, s, ca);
            hipLaunchKernelGGL(k_level_close, dim3(1), dim3(64), 0, s, lvl, d, node_counter);
        }
        BUILD_CHECK(hipGetLastError());
        hl.resize(2 * (d + 1));
        BUILD_CHECK(hipMemcpyAsync(hl.data(), lvl, hl.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        BUILD_CHECK(hipStreamSynchronize(s));
        for (int k = 0; k <= d; ++k)
            if (hl[2 * k] == hl[2 * k + 1]) {
                depth = k;
                break;
            }
    }
    a->level_start.assign(depth + 1, 0);  // level L = wide nodes [level_start[L], level_start[L + 1])
    for (int k = 0; k <= depth; ++k) a->level_start[k] = hl[2 * k];
    uint64_t counters[2];
    BUILD_CHECK(hipMemcpyAsync(counters, b_counters.p, 2 * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    BUILD_CHECK(hipStreamSynchronize(s));
    a->nodes = static_cast<int64_t>(counters[0] & 0xffffffffull);
    a->leaves = static_cast<int64_t>(counters[1]);
    if (static_cast<uint64_t>(a->nodes) * node_bytes > 0xffffffffull) {
        // Traversal addresses a node's planes with 32-bit byte offsets.
        cleanup();
        return fail(ctx, RT_ERR_UNSUPPORTED, "BVH node array beyond 4 GiB: use a larger leaf size or width 4");
    }
    a->depth = depth;
    a->root = 0;
    a->build_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    *out = a;
#undef BUILD_CHECK
    return RT_OK;
}

namespace {

// Refit scratch: [0, 6) scene bounds lo xyz, hi xyz, [6] padding (floats),
// then kBoundsBlocks partial bounds per coordinate. The bounds are reduced in
// two passes with no atomics: a few words updated by every wave serialise
// across the chip's L2s (8192 waves x 6 atomics took 561 us for 1M triangles).
constexpr int kBoundsBlocks = 1024;
constexpr int kRefitWords = 8 + 6 * kBoundsBlocks;

__global__ void __launch_bounds__(256) k_scene_bounds(const float *__restrict__ abc, int64_t n, float *__restrict__ w) {
    __shared__ float red[6][4];
    float v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const float *t = abc + 9 * i;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            v[c] = fminf(v[c], fminf(t[c], fminf(t[3 + c], t[6 + c])));
            v[3 + c] = fmaxf(v[3 + c], fmaxf(t[c], fmaxf(t[3 + c], t[6 + c])));
        }
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        float x = v[k];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const float y = __shfl_xor(x, o, 64);
            x = k < 3 ? fminf(x, y) : fmaxf(x, y);
        }
        if ((threadIdx.x & 63) == 0) red[k][threadIdx.x >> 6] = x;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        float x = red[k][0];
        for (int q = 1; q < 4; ++q) x = k < 3 ? fminf(x, red[k][q]) : fmaxf(x, red[k][q]);
        w[8 + k * kBoundsBlocks + blockIdx.x] = x;
    }
}

// Second pass over the block partials (one block of 256), then the box
// padding for the new bounds, never below the one in use.
__global__ void __launch_bounds__(256) k_refit_pad(float *__restrict__ w, int parts, float old_pad) {
    __shared__ float red[6][4];
    for (int k = 0; k < 6; ++k) {
        float x = k < 3 ? INFINITY : -INFINITY;
        for (int b = threadIdx.x; b < parts; b += 256) {
            const float y = w[8 + k * kBoundsBlocks + b];
            x = k < 3 ? fminf(x, y) : fmaxf(x, y);
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const float y = __shfl_xor(x, o, 64);
            x = k < 3 ? fminf(x, y) : fmaxf(x, y);
        }
        if ((threadIdx.x & 63) == 0) red[k][threadIdx.x >> 6] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float m = 0.0f;
        for (int k = 0; k < 6; ++k) {
            float x = red[k][0];
            for (int q = 1; q < 4; ++q) x = k < 3 ? fminf(x, red[k][q]) : fmaxf(x, red[k][q]);
            w[k] = x;
            m = fmaxf(m, fabsf(x));
        }
        w[6] = fmaxf(old_pad, ldexpf(fmaxf(m, 1e-30f), -15));
    }
}

// Leaf triangles from the scene's current vertices (the emplace id rides in
// a.w, the flags in e1.w).
__global__ void __launch_bounds__(256) k_refit_tris(const float *__restrict__ abc, int64_t n, LeafTri *__restrict__ tris) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    LeafTri lt = tris[k];
    const int32_t i = __float_as_int(lt.a_id.w);
    const float *t = abc + 9 * static_cast<int64_t>(i);
    const rtd::V3 a = rtd::mk(t[0], t[1], t[2]);
    const rtd::V3 b = rtd::mk(t[3], t[4], t[5]);
    const rtd::V3 c = rtd::mk(t[6], t[7], t[8]);
    // rt/triangle.h:193-194: e1 = c - a, e2 = b - a.
    const rtd::V3 e1 = rtd::sub(c, a);
    const rtd::V3 e2 = rtd::sub(b, a);
    lt.a_id = make_float4(a.x, a.y, a.z, lt.a_id.w);
    lt.e1 = make_float4(e1.x, e1.y, e1.z, lt.e1.w);
    lt.e2 = make_float4(e2.x, e2.y, e2.z, 0.0f);
    tris[k] = lt;
}

// One level of wide nodes: a leaf child's box from its triangles' vertices
// (padded), an interior child's box the union of that node's (padded) child
// boxes, refit one level deeper before.
template <int W>
__global__ void __launch_bounds__(256) k_refit_wide(void *nodes_v, int64_t lo, int64_t hi, const LeafTri *__restrict__ tris,
                                                    const float *__restrict__ abc, const float *__restrict__ w) {
    const int64_t x = lo + static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (x >= hi) return;
    const float pad = w[6];
    WideNode<W> *nodes = reinterpret_cast<WideNode<W> *>(nodes_v);
    WideNode<W> nd = nodes[x];
    for (int k = 0; k < W; ++k) {
        const int32_t c = nd.child[k];
        if (c == kEmpty) continue;
        float lx = INFINITY, ly = INFINITY, lz = INFINITY, hx = -INFINITY, hy = -INFINITY, hz = -INFINITY;
        if (is_leaf(c)) {
            const int32_t first = leaf_first(c), count = leaf_count(c);
            for (int j = 0; j < count; ++j) {
                const float *t = abc + 9 * static_cast<int64_t>(__float_as_int(tris[first + j].a_id.w));
                for (int v = 0; v < 3; ++v) {
                    lx = fminf(lx, t[3 * v]);
                    ly = fminf(ly, t[3 * v + 1]);
                    lz = fminf(lz, t[3 * v + 2]);
                    hx = fmaxf(hx, t[3 * v]);
                    hy = fmaxf(hy, t[3 * v + 1]);
                    hz = fmaxf(hz, t[3 * v + 2]);
                }
            }
            lx -= pad; ly -= pad; lz -= pad;
            hx += pad; hy += pad; hz += pad;
        } else {
            const WideNode<W> &ch = nodes[c];
            for (int j = 0; j < W; ++j) {
                if (ch.child[j] == kEmpty) continue;
                lx = fminf(lx, ch.lo_x[j]);
                ly = fminf(ly, ch.lo_y[j]);
                lz = fminf(lz, ch.lo_z[j]);
                hx = fmaxf(hx, ch.hi_x[j]);
                hy = fmaxf(hy, ch.hi_y[j]);
                hz = fmaxf(hz, ch.hi_z[j]);
            }
        }
        nd.lo_x[k] = lx;
        nd.lo_y[k] = ly;
        nd.lo_z[k] = lz;
        nd.hi_x[k] = hx;
        nd.hi_y[k] = hy;
        nd.hi_z[k] = hz;
    }
    nodes[x] = nd;
}

}  // namespace

int32_t refit_accel(rt_accel *a) {
    rt_ctx *ctx = a->scene->ctx;
    hipStream_t s = ctx->stream;
    const int64_t n = a->scene->n;
    if (n != a->triangles) return fail(ctx, RT_ERR_INVALID, "rt_accel_refit: the scene's triangle count changed");
    if (n == 0) return RT_OK;
    const auto t0 = std::chrono::steady_clock::now();
    if (!a->d_refit) RT_HIP_CHECK(ctx, hipMalloc(&a->d_refit, kRefitWords * sizeof(float)));
    const float *abc = a->scene->d_abc;
    const int blocks = static_cast<int>((n + 255) / 256);
    const int parts = std::min(blocks, kBoundsBlocks);
    hipLaunchKernelGGL(k_scene_bounds, dim3(parts), dim3(256), 0, s, abc, n, a->d_refit);
    hipLaunchKernelGGL(k_refit_pad, dim3(1), dim3(256), 0, s, a->d_refit, parts, a->box_pad);
    hipLaunchKernelGGL(k_refit_tris, dim3(blocks), dim3(256), 0, s, abc, n, a->d_tris);
    for (int l = static_cast<int>(a->level_start.size()) - 2; l >= 0; --l) {
        const int64_t lo = a->level_start[l], hi = a->level_start[l + 1];
        if (hi <= lo) continue;
        const dim3 g(static_cast<unsigned>((hi - lo + 255) / 256));
        if (a->width == 4)
            hipLaunchKernelGGL(k_refit_wide<4>, g, dim3(256), 0, s, a->d_nodes, lo, hi, a->d_tris, abc, a->d_refit);
        else
            hipLaunchKernelGGL(k_refit_wide<8>, g, dim3(256), 0, s, a->d_nodes, lo, hi, a->d_tris, abc, a->d_refit);
    }
    RT_HIP_CHECK(ctx, hipGetLastError());
    float hw[7];
    RT_HIP_CHECK(ctx, hipMemcpyAsync(hw, a->d_refit, sizeof(hw), hipMemcpyDeviceToHost, s));
    RT_HIP_CHECK(ctx, hipStreamSynchronize(s));
    float maxabs = 0.0f;
    for (int k = 0; k < 6; ++k) maxabs = std::max(maxabs, std::fabs(hw[k]));
    // The padding never shrinks and the fused-test limit never grows, so the
    // copies of both that batches and inverse-rendering sets took earlier
    // stay conservative: near_limit * 2^-24 <= box_pad / 2 still holds.
    a->box_pad = hw[6];
    a->near_limit = std::min(a->near_limit, 256.0f * maxabs);
    for (int k = 0; k < 3; ++k) {
        a->scene_lo[k] = hw[k];
        const float ext = hw[3 + k] - a->scene_lo[k];
        a->scene_inv[k] = ext > 0.0f ? 1.0f / ext : 0.0f;
    }
    a->refit_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return RT_OK;
}

}  // namespace rt

// ---------------------------------------------------------------------------
// Inspection of the build (tests and tools): the Morton stages on their own,
// and a download of the finished tree.

extern "C" {

int32_t rt_build_morton(rt_scene *scene, uint32_t *codes, uint32_t *order) {
    if (!scene || (scene->n > 0 && (!codes || !order))) return RT_ERR_INVALID;
    rt_ctx *ctx = scene->ctx;
    RT_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    const int64_t n = scene->n;
    if (n == 0) return RT_OK;
    if (n >= (int64_t(1) << 27)) return rt::fail(ctx, RT_ERR_INVALID, "too many triangles (limit 2^27)");
    hipStream_t s = ctx->stream;
    rt::BuildArena &ar = ctx->arena;
    {
        const hipError_t e = ar.begin();
        if (e != hipSuccess) return rt::fail(ctx, RT_ERR_OOM, std::string("build scratch: ") + hipGetErrorString(e));
    }
    const int64_t tiles = (n + rt::kSortTile - 1) / rt::kSortTile;
    const int blocks = static_cast<int>((n + 255) / 256);
    rt::DevBuf lo, hi, bounds, c, idx, c_t, idx_t, hist, sums, parts;
    hipError_t e = hipSuccess;
    if (e == hipSuccess) e = rt::dalloc_in<float4>(ar, lo, n);
    if (e == hipSuccess) e = rt::dalloc_in<float4>(ar, hi, n);
    if (e == hipSuccess) e = rt::dalloc_in<uint32_t>(ar, bounds, 12);
    if (e == hipSuccess) e = rt::dalloc_in<uint32_t>(ar, c, n);
    if (e == hipSuccess) e = rt::dalloc_in<uint32_t>(ar, idx, n);
    if (e == hipSuccess) e = rt::dalloc_in<uint32_t>(ar, c_t, n);
    if (e == hipSuccess) e = rt::dalloc_in<uint32_t>(ar, idx_t, n);
    if (e == hipSuccess) e = rt::dalloc_in<uint32_t>(ar, hist, 256 * tiles);
    if (e == hipSuccess) e = rt::dalloc_in<uint32_t>(ar, sums, rt::kScanTile);
    if (e == hipSuccess) e = rt::dalloc_in<uint32_t>(ar, parts, 12 * static_cast<int64_t>(blocks));
    if (e != hipSuccess) return rt::fail(ctx, RT_ERR_OOM, std::string("rt_build_morton: ") + hipGetErrorString(e));
    hipLaunchKernelGGL(rt::k_tri_bounds, dim3(blocks), dim3(256), 0, s, scene->d_abc, n, (float4 *)lo.p,
                       (float4 *)hi.p, (uint32_t *)parts.p);
    hipLaunchKernelGGL(rt::k_bounds_final, dim3(1), dim3(256), 0, s, (const uint32_t *)parts.p, blocks,
                       (uint32_t *)bounds.p);
    hipLaunchKernelGGL(rt::k_morton, dim3(blocks), dim3(256), 0, s, (const float4 *)lo.p, (const float4 *)hi.p, n,
                       (const uint32_t *)bounds.p, (uint32_t *)c.p, (uint32_t *)idx.p);
    RT_HIP_CHECK(ctx, hipGetLastError());
    RT_HIP_CHECK(ctx, hipMemcpyAsync(codes, c.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    RT_HIP_CHECK(ctx, rt::radix_sort_pairs((uint32_t *)c.p, (uint32_t *)idx.p, (uint32_t *)c_t.p, (uint32_t *)idx_t.p,
                                           n, 30, (uint32_t *)hist.p, (uint32_t *)sums.p, s));
    RT_HIP_CHECK(ctx, hipMemcpyAsync(order, idx.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    RT_HIP_CHECK(ctx, hipStreamSynchronize(s));
    return RT_OK;
}

int32_t rt_accel_download(const rt_accel *accel, void *nodes, int32_t *tri_ids) {
    if (!accel || (accel->nodes > 0 && !nodes) || (accel->scene->n > 0 && !tri_ids)) return RT_ERR_INVALID;
    rt_ctx *ctx = accel->scene->ctx;
    RT_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    const int64_t node_bytes = (accel->width == 4) ? sizeof(rt::WideNode<4>) : sizeof(rt::WideNode<8>);
    if (accel->nodes > 0)
        RT_HIP_CHECK(ctx, hipMemcpyAsync(nodes, accel->d_nodes, accel->nodes * node_bytes, hipMemcpyDeviceToHost, s));
    const int64_t n = accel->scene->n;
    std::vector<rt::LeafTri> tris(accel->d_tris ? n : 0);
    if (!tris.empty())
        RT_HIP_CHECK(ctx, hipMemcpyAsync(tris.data(), accel->d_tris, n * sizeof(rt::LeafTri), hipMemcpyDeviceToHost, s));
    RT_HIP_CHECK(ctx, hipStreamSynchronize(s));
    // The emplace id is the w word of each leaf triangle's first float4.
    for (size_t i = 0; i < tris.size(); ++i) {
        const float w = tris[i].a_id.w;
        std::memcpy(&tri_ids[i], &w, sizeof(int32_t));
    }
    return RT_OK;
}

int32_t rt_build_aac(rt_scene *scene, float alpha, float delta, uint32_t *codes, uint32_t *order, int32_t *indirect,
                     rt_aac_node *nodes, int64_t capacity, int64_t *nnodes, int32_t *root, float *build_ms) {
    if (!scene) return RT_ERR_INVALID;
    rt_ctx *ctx = scene->ctx;
    RT_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    const int64_t n = scene->n;
    if ((n > 0 && (!indirect || !nodes)) || capacity < 2 * n - 1)
        return rt::fail(ctx, RT_ERR_INVALID, "rt_build_aac: indirect (n) and nodes (at least 2n - 1) are required");
    rt::AacTree t;
    if (int32_t rc = rt::build_aac_tree(scene, alpha, delta, &t)) return rc;
    if (t.nodes > capacity) return rt::fail(ctx, RT_ERR_INVALID, "rt_build_aac: more nodes than capacity");
    for (int64_t i = 0; i < n; ++i) {
        if (codes) codes[i] = t.codes[i];
        if (order) order[i] = t.order[i];
        indirect[i] = t.indirect[i];
    }
    for (int64_t i = 0; i < t.nodes; ++i) {
        rt_aac_node &o = nodes[i];
        o.lo[0] = t.lo[i].x;
        o.lo[1] = t.lo[i].y;
        o.lo[2] = t.lo[i].z;
        o.hi[0] = t.hi[i].x;
        o.hi[1] = t.hi[i].y;
        o.hi[2] = t.hi[i].z;
        o.child[0] = t.kid[i].x;
        o.child[1] = t.kid[i].y;
        o.leaf = t.leaf[i];
    }
    if (nnodes) *nnodes = t.nodes;
    if (root) *root = t.root;
    if (build_ms) *build_ms = t.build_ms;
    return RT_OK;
}

int32_t rt_accel_root(const rt_accel *accel, int32_t *root, float *box_pad) {
    if (!accel || !root) return RT_ERR_INVALID;
    *root = accel->root;
    if (box_pad) *box_pad = accel->box_pad;
    return RT_OK;
}

}  // extern "C"
