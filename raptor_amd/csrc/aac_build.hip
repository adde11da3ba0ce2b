// The reference builder's tree on the GPU: RAPTor's approximate agglomerative
// clustering (ssd/bvh_builder.cc:18-642), reproduced node for node.
//
// The reference builds depth first on one thread: a top-down split of the
// Morton-sorted triangles (build_layer, build_layer_primitive) down to ranges
// of fewer than delta triangles, greedy leaf packing (build_leaf_node), then
// on the way back up an agglomerative merge of each node's clusters
// (combine_nodes) through one packed cost matrix shared by the whole
// recursion. Here the same tree comes from level-parallel passes:
//
//   k_aac_bounds / k_aac_box   triangle boxes; the scene box with the
//                              reference's fold (first occurrence of the
//                              extreme value, so the sign of a zero matches)
//   k_aac_codes                the builder's Morton code: per-axis cells of
//                              width scene extent * 1.00001 / 1024
//   radix sort                 stable, so the same order as the reference's
//                              three-pass LSD sort
//   k_aac_split (per level)    the top-down split: Morton-range midpoint by
//                              binary search, or the spatial-midpoint
//                              partition once the code range is exhausted
//   k_aac_pack                 greedy leaf packing of each small range
//   host                       per node: clusters left after its merge, the
//                              position its cost triangle would have in the
//                              reference's matrix (only its residue mod 4
//                              matters, below) and the ids of its new nodes
//   k_aac_merge (per level,    each node's merge, deepest level first: one
//   deepest first)             wave per node of at most 64 clusters, one
//                              1024-lane workgroup above, which keeps each
//                              row's least entry between merges
//
// No cost matrix is stored: every entry the reference reads equals the
// combined surface area of the two clusters at its row and column in the
// current cluster order (the reference's compaction keeps that so), and the
// order is replayed exactly, so costs are recomputed from the cluster boxes.
// The reference's min_element (common/simd.h:857-927) decides ties by the
// entry's absolute address in the matrix: entries before the first
// 16-byte-aligned one win ties in address order, then the aligned body,
// where the lowest of four SSE lanes (address mod 4) wins and then the
// lowest address, then the trailing entries. The argmin here uses that key,
// (cost, class, lane, address), with each node's matrix start offset taken
// from the sizes of the cost triangles the reference keeps below it.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "radix_sort.hpp"
#include "rt_internal.hpp"

namespace rt {

namespace {

// point_t min / max as the reference's helpers evaluate them.
__host__ __device__ __forceinline__ float rmin(float a, float b) { return (b < a) ? b : a; }
__host__ __device__ __forceinline__ float rmax(float a, float b) { return (a < b) ? b : a; }

// A node of the top-down split.
struct AacNode {
    int32_t b, e;       // triangle range [b, e) of the sorted order
    int32_t clo, chi;   // Morton code range (the reference's *cost_b, *cost_e on the way down)
    int32_t depth;      // recursion depth after the reference's ++depth_
    int32_t axis;       // axis this node halves: 0 x, 1 y, 2 z
    int32_t prim;       // 1: spatial-midpoint partition (build_layer_primitive)
    int32_t leaf;       // 1: a leaf range (build_leaf_node)
    float bl[3], tr[3]; // the recursion's spatial bounds
    int32_t kid[2];     // left / right child node, -1 none
};

struct SplitArgs {
    AacNode *nodes;
    int32_t lo, hi;             // the level: nodes [lo, hi)
    uint32_t *counter;          // next free node
    int32_t cap;                // node capacity
    uint32_t *overflow;
    const uint32_t *codes;      // sorted codes
    int32_t *ind;               // the triangle order (partitions permute it)
    const float4 *tlo, *thi;    // triangle boxes, emplace order
    float delta;
    int32_t max_depth;
};

// ssd/bvh_builder.cc:208-239: halve the axis and move on z -> y -> x -> z.
__device__ __forceinline__ int next_axis(int axis) { return axis == 0 ? 2 : (axis == 1 ? 0 : 1); }

__global__ void __launch_bounds__(256) k_aac_split(SplitArgs a) {
    const int32_t i = a.lo + static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= a.hi) return;
    AacNode nd = a.nodes[i];
    nd.kid[0] = nd.kid[1] = -1;
    const int32_t size = nd.e - nd.b;
    auto small = [&](int32_t depth) {
        return static_cast<float>(size) < a.delta || depth == a.max_depth;
    };
    if (small(nd.depth)) {
        nd.leaf = 1;
        a.nodes[i] = nd;
        return;
    }
    // build_layer hands a range whose codes are exhausted to
    // build_layer_primitive, one level deeper (ssd/bvh_builder.cc:415-418).
    if (!nd.prim && nd.clo == nd.chi - 1) {
        nd.prim = 1;
        nd.depth += 1;
        if (small(nd.depth)) {
            nd.leaf = 1;
            a.nodes[i] = nd;
            return;
        }
    }
    int32_t m;
    int32_t mid = 0;
    const int ax = nd.axis;
    if (!nd.prim) {
        // ssd/bvh_builder.cc:422-423: lower_bound of the code midpoint.
        mid = nd.clo + ((nd.chi - nd.clo) >> 1);
        int32_t lo = nd.b, hi = nd.e;
        while (lo < hi) {
            const int32_t h = lo + ((hi - lo) >> 1);
            if (static_cast<int32_t>(a.codes[h]) < mid) {
                lo = h + 1;
            } else {
                hi = h;
            }
        }
        m = lo;
    } else {
        // ssd/bvh_builder.cc:492-531: partition about the spatial midpoint of
        // the axis, on box centre times two, in place in the triangle order.
        auto centre2 = [&](int32_t k) {
            const int32_t t = a.ind[k];
            const float h = ax == 0 ? a.thi[t].x : (ax == 1 ? a.thi[t].y : a.thi[t].z);
            const float l = ax == 0 ? a.tlo[t].x : (ax == 1 ? a.tlo[t].y : a.tlo[t].z);
            return h + l;
        };
        const float dbl_split = nd.tr[ax] + nd.bl[ax];
        int32_t top = nd.e - 1, bottom = nd.b;
        while (bottom < top) {
            while (dbl_split < centre2(top) && bottom < top) --top;
            while (dbl_split >= centre2(bottom) && bottom < top) ++bottom;
            if (bottom < top) {
                const int32_t t = a.ind[bottom];
                a.ind[bottom] = a.ind[top];
                a.ind[top] = t;
                --top;
                ++bottom;
            }
        }
        m = bottom;
        if ((top == bottom) && (dbl_split >= centre2(top))) ++m;
    }
    const float split = (nd.tr[ax] + nd.bl[ax]) * 0.5f;
    const int nxt = next_axis(ax);
    const int nk = (nd.b != m ? 1 : 0) + (m != nd.e ? 1 : 0);
    const uint32_t first = atomicAdd(a.counter, static_cast<uint32_t>(nk));
    if (static_cast<int64_t>(first) + nk > a.cap) {
        atomicOr(a.overflow, 1u);
        nd.leaf = 1;  // keeps the node array consistent; the build reports the overflow
        a.nodes[i] = nd;
        return;
    }
    int32_t slot = static_cast<int32_t>(first);
    for (int side = 0; side < 2; ++side) {
        const int32_t cb = side == 0 ? nd.b : m, ce = side == 0 ? m : nd.e;
        if (cb == ce) continue;
        AacNode c;
        c.b = cb;
        c.e = ce;
        c.clo = nd.prim ? 0 : (side == 0 ? nd.clo : mid);
        c.chi = nd.prim ? 0 : (side == 0 ? mid : nd.chi);
        c.depth = nd.depth + 1;
        c.axis = nxt;
        c.prim = nd.prim;
        c.leaf = 0;
        for (int k = 0; k < 3; ++k) {
            c.bl[k] = nd.bl[k];
            c.tr[k] = nd.tr[k];
        }
        if (side == 0) {
            c.tr[ax] = split;  // tm: the top of the lower half
        } else {
            c.bl[ax] = split;  // bm: the bottom of the upper half
        }
        c.kid[0] = c.kid[1] = -1;
        a.nodes[slot] = c;
        nd.kid[side] = slot;
        ++slot;
    }
    a.nodes[i] = nd;
}

// Triangle boxes (rt/triangle.h:56-57) and, per block, the first occurrence
// of the smallest low and the largest high on each axis.
__global__ void __launch_bounds__(256) k_aac_bounds(const float *__restrict__ abc, int64_t n, float4 *__restrict__ tlo,
                                                    float4 *__restrict__ thi, float *__restrict__ pv,
                                                    int64_t *__restrict__ pi) {
    __shared__ float sv[6][256];
    __shared__ int64_t si[6][256];
    const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    float v[6];
    int64_t idx[6];
    if (i < n) {
        const float *t = abc + 9 * i;
        float lo[3], hi[3];
        for (int k = 0; k < 3; ++k) {
            lo[k] = rmin(t[k], rmin(t[3 + k], t[6 + k]));
            hi[k] = rmax(t[k], rmax(t[3 + k], t[6 + k]));
        }
        tlo[i] = make_float4(lo[0], lo[1], lo[2], 0.0f);
        thi[i] = make_float4(hi[0], hi[1], hi[2], 0.0f);
        for (int k = 0; k < 3; ++k) {
            v[k] = lo[k];
            v[3 + k] = hi[k];
            idx[k] = idx[3 + k] = i;
        }
    } else {
        for (int k = 0; k < 6; ++k) {
            v[k] = k < 3 ? INFINITY : -INFINITY;
            idx[k] = INT64_MAX;
        }
    }
    for (int k = 0; k < 6; ++k) {
        sv[k][threadIdx.x] = v[k];
        si[k][threadIdx.x] = idx[k];
    }
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            for (int k = 0; k < 6; ++k) {
                const float a = sv[k][threadIdx.x], b = sv[k][threadIdx.x + s];
                const int64_t ia = si[k][threadIdx.x], ib = si[k][threadIdx.x + s];
                const bool take = (k < 3) ? (b < a || (b == a && ib < ia)) : (b > a || (b == a && ib < ia));
                if (take) {
                    sv[k][threadIdx.x] = b;
                    si[k][threadIdx.x] = ib;
                }
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) {
        pv[6 * static_cast<int64_t>(blockIdx.x) + threadIdx.x] = sv[threadIdx.x][0];
        pi[6 * static_cast<int64_t>(blockIdx.x) + threadIdx.x] = si[threadIdx.x][0];
    }
}

// The per-block extremes folded into the scene box, earliest triangle first
// on equal values, as the reference's sequential fold keeps them.
__global__ void __launch_bounds__(256) k_aac_box(const float *__restrict__ pv, const int64_t *__restrict__ pi,
                                                 int64_t nparts, float *__restrict__ box) {
    __shared__ float sv[6][256];
    __shared__ int64_t si[6][256];
    float v[6];
    int64_t idx[6];
    for (int k = 0; k < 6; ++k) {
        v[k] = k < 3 ? INFINITY : -INFINITY;
        idx[k] = INT64_MAX;
    }
    for (int64_t p = threadIdx.x; p < nparts; p += 256) {
        for (int k = 0; k < 6; ++k) {
            const float b = pv[6 * p + k];
            const int64_t ib = pi[6 * p + k];
            const bool take = (k < 3) ? (b < v[k] || (b == v[k] && ib < idx[k])) : (b > v[k] || (b == v[k] && ib < idx[k]));
            if (take) {
                v[k] = b;
                idx[k] = ib;
            }
        }
    }
    for (int k = 0; k < 6; ++k) {
        sv[k][threadIdx.x] = v[k];
        si[k][threadIdx.x] = idx[k];
    }
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            for (int k = 0; k < 6; ++k) {
                const float a = sv[k][threadIdx.x], b = sv[k][threadIdx.x + s];
                const int64_t ia = si[k][threadIdx.x], ib = si[k][threadIdx.x + s];
                const bool take = (k < 3) ? (b < a || (b == a && ib < ia)) : (b > a || (b == a && ib < ia));
                if (take) {
                    sv[k][threadIdx.x] = b;
                    si[k][threadIdx.x] = ib;
                }
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) box[threadIdx.x] = sv[threadIdx.x][0];
}

// common/simd.h:798-811
__host__ __device__ __forceinline__ uint32_t split_by_three(uint32_t a) {
    a |= (a << 16);
    a &= 0x030000FFu;
    a |= (a << 8);
    a &= 0x0300F00Fu;
    a |= (a << 4);
    a &= 0x030C30C3u;
    a |= (a << 2);
    a &= 0x09249249u;
    return a;
}

// ssd/bvh_builder.cc:33-49 with common/simd.h:813-823: the box centre
// relative to the scene's low corner, scaled by each axis's cell count.
__global__ void __launch_bounds__(256) k_aac_codes(const float4 *__restrict__ tlo, const float4 *__restrict__ thi,
                                                   int64_t n, float bx, float by, float bz, float ix, float iy,
                                                   float iz, uint32_t *__restrict__ codes,
                                                   uint32_t *__restrict__ order) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n) return;
    const float4 l = tlo[i], h = thi[i];
    const float cx = ((h.x + l.x) * 0.5f) - bx;
    const float cy = ((h.y + l.y) * 0.5f) - by;
    const float cz = ((h.z + l.z) * 0.5f) - bz;
    const uint32_t qx = static_cast<uint32_t>(static_cast<int>(cx * ix));
    const uint32_t qy = static_cast<uint32_t>(static_cast<int>(cy * iy));
    const uint32_t qz = static_cast<uint32_t>(static_cast<int>(cz * iz));
    codes[i] = split_by_three(qx) | (split_by_three(qy) << 1) | (split_by_three(qz) << 2);
    order[i] = static_cast<uint32_t>(i);
}

struct PackArgs {
    const AacNode *nodes;
    const int32_t *leaf_nodes;  // the leaf ranges
    int32_t count;
    int32_t *ind;
    const float4 *tlo, *thi;
    float max_leaf_sah;
    float4 *cl_lo, *cl_hi;      // clusters of a range at positions [b, b + c)
    int2 *cl_range;
    int32_t *nc;                // clusters per leaf range, by its index in leaf_nodes
};

// ssd/bvh_builder.cc:241-275: pack consecutive triangles into a leaf while
// the leaf's surface area times its size stays within max_leaf_sah; a
// triangle that would break it is swapped to the back of the range.
__global__ void __launch_bounds__(256) k_aac_pack(PackArgs a) {
    const int32_t q = static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x);
    if (q >= a.count) return;
    const AacNode nd = a.nodes[a.leaf_nodes[q]];
    int32_t c = 0;
    for (int32_t i = nd.b; i < nd.e;) {
        int32_t node_end = i + 1;
        int32_t layer_top = nd.e;
        float4 lo = a.tlo[a.ind[i]], hi = a.thi[a.ind[i]];
        while (node_end < layer_top) {
            const float4 tl = a.tlo[a.ind[node_end]], th = a.thi[a.ind[node_end]];
            const float blx = rmin(lo.x, tl.x), bly = rmin(lo.y, tl.y), blz = rmin(lo.z, tl.z);
            const float trx = rmax(hi.x, th.x), try_ = rmax(hi.y, th.y), trz = rmax(hi.z, th.z);
            const float dx = trx - blx, dy = try_ - bly, dz = trz - blz;
            const float cost = (((dx * dy) + (dx * dz)) + (dy * dz)) * static_cast<float>(node_end - i + 1);
            if (cost > a.max_leaf_sah) {
                --layer_top;
                const int32_t t = a.ind[node_end];
                a.ind[node_end] = a.ind[layer_top];
                a.ind[layer_top] = t;
            } else {
                lo = make_float4(blx, bly, blz, 0.0f);
                hi = make_float4(trx, try_, trz, 0.0f);
                ++node_end;
            }
        }
        a.cl_lo[nd.b + c] = lo;
        a.cl_hi[nd.b + c] = hi;
        a.cl_range[nd.b + c] = make_int2(i, node_end);
        ++c;
        i = node_end;
    }
    a.nc[q] = c;
}

// One merge job: a node of the split whose clusters are merged.
struct MergeJob {
    int32_t node;      // split node
    int32_t src;       // list source: -1 the node's own leaves, else the number of children (1 or 2)
    int32_t k0, k;     // clusters before and after the merge
    int32_t leaf_id;   // leaf ranges: the id of the first leaf cluster
    int32_t node_id;   // the id of the first node the merge makes
    int64_t off;       // the start of the node's cost triangle in the reference's matrix
    int32_t b;         // the node's list position (its triangle range's start)
    int32_t kb[2];     // children's list positions
    int32_t kk[2];     // children's final cluster counts
};

struct MergeArgs {
    const MergeJob *jobs;
    int32_t count;
    int32_t *lists;            // cluster ids by list position
    float4 *bn_lo, *bn_hi;     // binary node boxes, by id
    int2 *bn_kid;              // binary node children (leaf: the triangle range, flagged below)
    uint8_t *bn_leaf;
    const float4 *cl_lo, *cl_hi;
    const int2 *cl_range;
};

// Ordering key of one cost-matrix entry for the reference's min_element.
struct Key {
    float cost;
    int32_t cls;   // 0 head, 1 aligned body, 2 tail
    int32_t lane;  // body: the SSE lane (address mod 4)
    int64_t addr;
    int32_t i, j;
};

__device__ __forceinline__ bool key_less(const Key &a, const Key &b) {
    if (a.cost != b.cost) return a.cost < b.cost;
    if (a.cls != b.cls) return a.cls < b.cls;
    if (a.lane != b.lane) return a.lane < b.lane;
    return a.addr < b.addr;
}

__device__ __forceinline__ float combined_area(float4 l1, float4 h1, float4 l2, float4 h2) {
    const float dx = rmax(h1.x, h2.x) - rmin(l1.x, l2.x);
    const float dy = rmax(h1.y, h2.y) - rmin(l1.y, l2.y);
    const float dz = rmax(h1.z, h2.z) - rmin(l1.z, l2.z);
    return ((dx * dy) + (dx * dz)) + (dy * dz);
}

__device__ __forceinline__ Key shfl_key(const Key &k, int lane_mask) {
    Key o;
    o.cost = __shfl_xor(k.cost, lane_mask, 64);
    o.cls = __shfl_xor(k.cls, lane_mask, 64);
    o.lane = __shfl_xor(k.lane, lane_mask, 64);
    o.addr = static_cast<int64_t>(__shfl_xor(static_cast<long long>(k.addr), lane_mask, 64));
    o.i = __shfl_xor(k.i, lane_mask, 64);
    o.j = __shfl_xor(k.j, lane_mask, 64);
    return o;
}

__device__ __forceinline__ Key wave_min(Key k) {
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) {
        const Key o = shfl_key(k, m);
        if (key_less(o, k)) k = o;
    }
    return k;
}

// This thread's best entry among the packed positions p = t, t + stride, ...
// of the current k clusters (row i >= 1, column j < i; address S + p).
__device__ __forceinline__ Key scan_pairs(const int32_t *ids, const float4 *lo, const float4 *hi, int32_t k,
                                          int64_t S, int t, int stride) {
    Key best;
    best.cost = INFINITY;
    best.cls = 3;
    best.lane = 0;
    best.addr = INT64_MAX;
    best.i = best.j = -1;
    const int64_t T = (static_cast<int64_t>(k) * (k - 1)) >> 1;
    const int64_t E = S + T;
    const int64_t A = (S & ~int64_t(3)) + ((S & 3) ? 4 : 0);
    const int64_t body_end = (E > A) ? A + (((E - A) >> 2) << 2) : A;
    if (t >= T) return best;
    // row of packed position t: the largest i with i (i - 1) / 2 <= t
    int32_t i = static_cast<int32_t>((1.0 + sqrt(1.0 + 8.0 * static_cast<double>(t))) * 0.5);
    while (static_cast<int64_t>(i) * (i - 1) / 2 > t) --i;
    while (static_cast<int64_t>(i + 1) * i / 2 <= t) ++i;
    int32_t j = static_cast<int32_t>(t - static_cast<int64_t>(i) * (i - 1) / 2);
    for (int64_t p = t; p < T; p += stride) {
        const float c = combined_area(lo[i], hi[i], lo[j], hi[j]);
        const int64_t P = S + p;
        Key key;
        key.cost = c;
        key.addr = P;
        if (P < A) {
            key.cls = 0;
            key.lane = 0;
        } else if (P < body_end) {
            key.cls = 1;
            key.lane = static_cast<int32_t>(P & 3);
        } else {
            key.cls = 2;
            key.lane = 0;
        }
        key.i = i;
        key.j = j;
        if (key_less(key, best)) best = key;
        j += stride;
        while (j >= i) {
            j -= i;
            ++i;
        }
    }
    (void)ids;
    return best;
}

// Leaf clusters of a leaf range become binary leaves with their ids.
__device__ __forceinline__ void write_leaves(const MergeArgs &a, const MergeJob &jb, int t, int stride) {
    if (jb.src >= 0) return;
    for (int32_t q = t; q < jb.k0; q += stride) {
        const int32_t id = jb.leaf_id + q;
        a.bn_lo[id] = a.cl_lo[jb.b + q];
        a.bn_hi[id] = a.cl_hi[jb.b + q];
        a.bn_kid[id] = a.cl_range[jb.b + q];
        a.bn_leaf[id] = 1;
    }
}

// ssd/bvh_builder.cc:352-357 after the argmin: the new node joins the
// clusters at rows i and j (left i, right j); the last cluster moves to
// row i and the new node takes row j.
__device__ __forceinline__ void merge_step(const MergeArgs &a, int32_t *ids, float4 *lo, float4 *hi, int32_t &k,
                                           int32_t node_id, int32_t i, int32_t j) {
    // No least entry (every cost NaN, from non-finite vertices): merge the
    // first two rows rather than index outside the list.
    if (i < 1 || i >= k || j < 0 || j >= i) {
        i = 1;
        j = 0;
    }
    const float4 nlo = make_float4(rmin(lo[i].x, lo[j].x), rmin(lo[i].y, lo[j].y), rmin(lo[i].z, lo[j].z), 0.0f);
    const float4 nhi = make_float4(rmax(hi[i].x, hi[j].x), rmax(hi[i].y, hi[j].y), rmax(hi[i].z, hi[j].z), 0.0f);
    a.bn_lo[node_id] = nlo;
    a.bn_hi[node_id] = nhi;
    a.bn_kid[node_id] = make_int2(ids[i], ids[j]);
    a.bn_leaf[node_id] = 0;
    ids[i] = ids[k - 1];
    lo[i] = lo[k - 1];
    hi[i] = hi[k - 1];
    ids[j] = node_id;
    lo[j] = nlo;
    hi[j] = nhi;
    --k;
}

constexpr int kWaveList = 64;      // clusters a wave job holds in LDS
constexpr int kBlockList = 2048;   // clusters a workgroup job holds in LDS

// One wave (workgroup) per job of at most kWaveList clusters.
__global__ void __launch_bounds__(64) k_aac_merge_wave(MergeArgs a) {
    __shared__ int32_t ids[kWaveList];
    __shared__ float4 lo[kWaveList], hi[kWaveList];
    const int lane = threadIdx.x;
    const MergeJob jb = a.jobs[blockIdx.x];
    write_leaves(a, jb, lane, 64);
    for (int32_t q = lane; q < jb.k0; q += 64) {
        int32_t id;
        float4 l, h;
        if (jb.src < 0) {
            id = jb.leaf_id + q;
            l = a.cl_lo[jb.b + q];
            h = a.cl_hi[jb.b + q];
        } else {
            id = (jb.src == 1 || q < jb.kk[0]) ? a.lists[jb.kb[0] + q] : a.lists[jb.kb[1] + (q - jb.kk[0])];
            l = a.bn_lo[id];
            h = a.bn_hi[id];
        }
        ids[q] = id;
        lo[q] = l;
        hi[q] = h;
    }
    __syncthreads();
    int32_t k = jb.k0;
    int32_t made = 0;
    while (k > jb.k) {
        const Key best = wave_min(scan_pairs(ids, lo, hi, k, jb.off, lane, 64));
        __syncthreads();
        if (lane == 0) merge_step(a, ids, lo, hi, k, jb.node_id + made, best.i, best.j);
        k = __shfl(k, 0, 64);
        ++made;
        __syncthreads();
    }
    for (int32_t q = lane; q < k; q += 64) a.lists[jb.b + q] = ids[q];
}

// Tie part of an entry's key, below its cost: class (head 0, aligned body 1,
// tail 2), SSE lane in the body, then the address relative to the node's
// matrix start. Orders entries of equal cost as key_less does.
__device__ __forceinline__ uint64_t entry_tie(int64_t P, int64_t S, int64_t A, int64_t body_end) {
    uint64_t cls = 2, lane = 0;
    if (P < A) {
        cls = 0;
    } else if (P < body_end) {
        cls = 1;
        lane = static_cast<uint64_t>(P & 3);
    }
    return (cls << 62) | (lane << 60) | static_cast<uint64_t>(P - S);
}

__device__ __forceinline__ bool entry_less(float c1, uint64_t t1, float c2, uint64_t t2) {
    return c1 < c2 || (c1 == c2 && t1 < t2);
}

// The region boundaries of the current packed triangle of k clusters that
// starts at S: head entries lie before A, the aligned body before body_end.
__device__ __forceinline__ void matrix_regions(int64_t S, int32_t k, int64_t *A, int64_t *body_end) {
    const int64_t E = S + ((static_cast<int64_t>(k) * (k - 1)) >> 1);
    *A = (S & ~int64_t(3)) + ((S & 3) ? 4 : 0);
    *body_end = (E > *A) ? *A + (((E - *A) >> 2) << 2) : *A;
}

// Cluster rows of a workgroup job, in LDS or (long lists) in global memory.
struct Rows {
    int32_t *ids;
    float4 *lo, *hi;
    float *rc;        // per row r >= 1: the least cost among its entries (r, c < r)
    uint64_t *rt;     // and that entry's tie key
    int32_t *todo;    // rows to rescan in full
};

// Rows of more than kWaveList and at most kBlockList clusters stay in LDS.
struct RowsLds {
    int32_t ids[kBlockList];
    float4 lo[kBlockList], hi[kBlockList];
    float rc[kBlockList];
    uint64_t rt[kBlockList];
    int32_t todo[kBlockList];
};

// Full rescans of the listed rows, one wave per row.
__device__ __forceinline__ void rescan_rows(const Rows &R, int32_t nrows, int64_t S, int64_t A, int64_t body_end,
                                            int w, int lane) {
    for (int32_t q = w; q < nrows; q += 16) {
        const int32_t r = R.todo[q];
        const int64_t row0 = S + ((static_cast<int64_t>(r) * (r - 1)) >> 1);
        const float4 lr = R.lo[r], hr = R.hi[r];
        float bc = INFINITY;
        uint64_t bt = ~0ull;
        for (int32_t c = lane; c < r; c += 64) {
            const float cost = combined_area(lr, hr, R.lo[c], R.hi[c]);
            const uint64_t tie = entry_tie(row0 + c, S, A, body_end);
            if (entry_less(cost, tie, bc, bt)) {
                bc = cost;
                bt = tie;
            }
        }
#pragma unroll
        for (int m = 32; m > 0; m >>= 1) {
            const float oc = __shfl_xor(bc, m, 64);
            const uint64_t ot = static_cast<uint64_t>(__shfl_xor(static_cast<long long>(bt), m, 64));
            if (entry_less(oc, ot, bc, bt)) {
                bc = oc;
                bt = ot;
            }
        }
        if (lane == 0) {
            R.rc[r] = bc;
            R.rt[r] = bt;
        }
    }
}

// One 1024-lane workgroup per job of more than kWaveList clusters. Each row
// keeps its least entry under the reference's min_element order; after a
// merge of rows i > j (the new node at row j, the last row moved to row i)
// only the entries in columns i and j change in the other rows, so a row is
// updated from those two entries, and rescanned in full only when its least
// entry was in one of those columns, when it is row i or j itself, or when
// it is one of the last two rows, the only rows whose entries can move from
// the aligned body to the tail as the triangle shrinks. Lists of more than
// kBlockList clusters stay in global memory (the top of scenes of several
// million triangles).
__global__ void __launch_bounds__(1024) k_aac_merge_block(MergeArgs a, Rows g, int32_t g_stride) {
    __shared__ RowsLds L;
    __shared__ float s_bc[16];
    __shared__ uint64_t s_bt[16];
    __shared__ int32_t s_br[16];
    __shared__ int32_t s_k, s_i, s_j, s_ntodo;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const MergeJob jb = a.jobs[blockIdx.x];
    Rows R;
    if (jb.k0 <= kBlockList) {
        R = Rows{L.ids, L.lo, L.hi, L.rc, L.rt, L.todo};
    } else {
        const int64_t o = static_cast<int64_t>(blockIdx.x) * g_stride;
        R = Rows{g.ids + o, g.lo + o, g.hi + o, g.rc + o, g.rt + o, g.todo + o};
    }
    const int64_t S = jb.off;
    write_leaves(a, jb, t, 1024);
    for (int32_t qq = t; qq < jb.k0; qq += 1024) {
        int32_t id;
        float4 l, h;
        if (jb.src < 0) {
            id = jb.leaf_id + qq;
            l = a.cl_lo[jb.b + qq];
            h = a.cl_hi[jb.b + qq];
        } else {
            id = (jb.src == 1 || qq < jb.kk[0]) ? a.lists[jb.kb[0] + qq] : a.lists[jb.kb[1] + (qq - jb.kk[0])];
            l = a.bn_lo[id];
            h = a.bn_hi[id];
        }
        R.ids[qq] = id;
        R.lo[qq] = l;
        R.hi[qq] = h;
        if (qq >= 1) R.todo[qq - 1] = qq;
    }
    if (t == 0) s_k = jb.k0;
    __syncthreads();
    {
        int64_t A, body_end;
        matrix_regions(S, jb.k0, &A, &body_end);
        rescan_rows(R, jb.k0 - 1, S, A, body_end, w, lane);
    }
    __syncthreads();
    int32_t made = 0;
    while (true) {
        const int32_t k = s_k;
        if (k <= jb.k) break;
        // The least entry over all rows.
        float bc = INFINITY;
        uint64_t bt = ~0ull;
        int32_t br = -1;
        for (int32_t r = 1 + t; r < k; r += 1024) {
            if (entry_less(R.rc[r], R.rt[r], bc, bt)) {
                bc = R.rc[r];
                bt = R.rt[r];
                br = r;
            }
        }
#pragma unroll
        for (int m = 32; m > 0; m >>= 1) {
            const float oc = __shfl_xor(bc, m, 64);
            const uint64_t ot = static_cast<uint64_t>(__shfl_xor(static_cast<long long>(bt), m, 64));
            const int32_t orr = __shfl_xor(br, m, 64);
            if (entry_less(oc, ot, bc, bt)) {
                bc = oc;
                bt = ot;
                br = orr;
            }
        }
        if (lane == 0) {
            s_bc[w] = bc;
            s_bt[w] = bt;
            s_br[w] = br;
        }
        __syncthreads();
        if (t == 0) {
            float c0 = s_bc[0];
            uint64_t t0 = s_bt[0];
            int32_t r0 = s_br[0];
            for (int q = 1; q < 16; ++q) {
                if (entry_less(s_bc[q], s_bt[q], c0, t0)) {
                    c0 = s_bc[q];
                    t0 = s_bt[q];
                    r0 = s_br[q];
                }
            }
            int32_t i = r0;
            int32_t j = (i >= 1) ? static_cast<int32_t>(static_cast<int64_t>(t0 & ((1ull << 60) - 1)) -
                                                        ((static_cast<int64_t>(i) * (i - 1)) >> 1))
                                 : -1;
            if (i < 1 || i >= k || j < 0 || j >= i) {  // as merge_step: no least entry
                i = 1;
                j = 0;
            }
            int32_t kk = k;
            merge_step(a, R.ids, R.lo, R.hi, kk, jb.node_id + made, i, j);
            s_k = kk;
            s_i = i;
            s_j = j;
            s_ntodo = 0;
        }
        ++made;
        __syncthreads();
        const int32_t k1 = s_k, i = s_i, j = s_j;
        int64_t A, body_end;
        matrix_regions(S, k1, &A, &body_end);
        const bool i_live = i < k1;
        for (int32_t r = 1 + t; r < k1; r += 1024) {
            bool full = (r == j) || (r == i) || (r >= k1 - 2);
            if (!full) {
                const float mc = R.rc[r];
                const uint64_t mt = R.rt[r];
                const int32_t col = static_cast<int32_t>(static_cast<int64_t>(mt & ((1ull << 60) - 1)) -
                                                         ((static_cast<int64_t>(r) * (r - 1)) >> 1));
                if (col == j || (i_live && col == i)) {
                    full = true;
                } else {
                    float nc = mc;
                    uint64_t nt = mt;
                    const int64_t row0 = S + ((static_cast<int64_t>(r) * (r - 1)) >> 1);
                    if (j < r) {
                        const float c = combined_area(R.lo[r], R.hi[r], R.lo[j], R.hi[j]);
                        const uint64_t tie = entry_tie(row0 + j, S, A, body_end);
                        if (entry_less(c, tie, nc, nt)) {
                            nc = c;
                            nt = tie;
                        }
                    }
                    if (i_live && i < r) {
                        const float c = combined_area(R.lo[r], R.hi[r], R.lo[i], R.hi[i]);
                        const uint64_t tie = entry_tie(row0 + i, S, A, body_end);
                        if (entry_less(c, tie, nc, nt)) {
                            nc = c;
                            nt = tie;
                        }
                    }
                    R.rc[r] = nc;
                    R.rt[r] = nt;
                }
            }
            if (full) R.todo[atomicAdd(&s_ntodo, 1)] = r;
        }
        __syncthreads();
        rescan_rows(R, s_ntodo, S, A, body_end, w, lane);
        __syncthreads();
    }
    const int32_t k = s_k;
    for (int32_t qq = t; qq < k; qq += 1024) a.lists[jb.b + qq] = R.ids[qq];
}

// ssd/bvh_builder.cc:181-185: clusters a node of n triangles keeps, with the
// host's powf (the reference's arithmetic).
int32_t reduction(float delta, float alpha, int64_t n) {
    const float eps = 1e-8f;
    const float r = std::max(1.0f, (std::pow(delta, 0.5f + eps) * std::pow(static_cast<float>(n), alpha) * 0.5f));
    return static_cast<int32_t>(r);
}

}  // namespace

int32_t build_aac_tree(rt_scene *scene, float alpha, float delta, AacTree *out) {
    rt_ctx *ctx = scene->ctx;
    const int64_t n = scene->n;
    hipStream_t s = ctx->stream;
    *out = AacTree();
    if (!(delta >= 2.0f) || !(alpha > 0.0f)) return fail(ctx, RT_ERR_INVALID, "rt_build_aac: delta >= 2, alpha > 0");
    if (n >= (int64_t(1) << 28)) return fail(ctx, RT_ERR_INVALID, "rt_build_aac: too many triangles (limit 2^28)");
    out->n = n;
    if (n == 0) return RT_OK;
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<void *> bufs;
    auto cleanup = [&]() {
        (void)hipStreamSynchronize(s);
        for (void *p : bufs) (void)hipFree(p);
        bufs.clear();
    };
    auto dmalloc = [&](void **p, size_t bytes) -> hipError_t {
        const hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 16));
        if (e == hipSuccess) bufs.push_back(*p);
        return e;
    };
#define AAC_CHECK(expr)                                                                                  \
    do {                                                                                                 \
        hipError_t _e = (expr);                                                                          \
        if (_e != hipSuccess) {                                                                          \
            cleanup();                                                                                   \
            return fail(ctx, _e == hipErrorOutOfMemory ? RT_ERR_OOM : RT_ERR_HIP,                       \
                        std::string("rt_build_aac: ") + #expr + ": " + hipGetErrorString(_e));         \
        }                                                                                                \
    } while (0)

    float4 *tlo, *thi;
    AAC_CHECK(dmalloc((void **)&tlo, n * sizeof(float4)));
    AAC_CHECK(dmalloc((void **)&thi, n * sizeof(float4)));
    const int64_t blocks = (n + 255) / 256;
    float *pv, *box;
    int64_t *pi;
    AAC_CHECK(dmalloc((void **)&pv, 6 * blocks * sizeof(float)));
    AAC_CHECK(dmalloc((void **)&pi, 6 * blocks * sizeof(int64_t)));
    AAC_CHECK(dmalloc((void **)&box, 6 * sizeof(float)));
    hipLaunchKernelGGL(k_aac_bounds, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, scene->d_abc, n, tlo, thi,
                       pv, pi);
    hipLaunchKernelGGL(k_aac_box, dim3(1), dim3(256), 0, s, pv, pi, blocks, box);
    AAC_CHECK(hipGetLastError());
    float hb[6];
    AAC_CHECK(hipMemcpyAsync(hb, box, sizeof(hb), hipMemcpyDeviceToHost, s));
    AAC_CHECK(hipStreamSynchronize(s));
    // ssd/bvh_builder.cc:40-48 on the host, in its float arithmetic.
    float sw[3], winv[3];
    for (int k = 0; k < 3; ++k) {
        sw[k] = hb[3 + k] - hb[k];
        const float width = sw[k] * (1.00001f / 1024.0f);
        winv[k] = 1.0f / width;
    }
    const float max_leaf_sah =
        ((((sw[0] * sw[1]) + (sw[0] * sw[2])) + (sw[1] * sw[2])) * static_cast<float>(n)) * 0.00000000009f;

    uint32_t *codes, *order, *codes_t, *order_t, *hist, *sums;
    const int64_t tiles = (n + kSortTile - 1) / kSortTile;
    AAC_CHECK(dmalloc((void **)&codes, n * sizeof(uint32_t)));
    AAC_CHECK(dmalloc((void **)&order, n * sizeof(uint32_t)));
    AAC_CHECK(dmalloc((void **)&codes_t, n * sizeof(uint32_t)));
    AAC_CHECK(dmalloc((void **)&order_t, n * sizeof(uint32_t)));
    AAC_CHECK(dmalloc((void **)&hist, 256 * tiles * sizeof(uint32_t)));
    AAC_CHECK(dmalloc((void **)&sums, kScanTile * sizeof(uint32_t)));
    hipLaunchKernelGGL(k_aac_codes, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, tlo, thi, n, hb[0], hb[1],
                       hb[2], winv[0], winv[1], winv[2], codes, order);
    AAC_CHECK(hipGetLastError());
    out->codes.resize(n);
    AAC_CHECK(hipMemcpyAsync(out->codes.data(), codes, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    AAC_CHECK(radix_sort_pairs(codes, order, codes_t, order_t, n, 30, hist, sums, s));
    out->order.resize(n);
    AAC_CHECK(hipMemcpyAsync(out->order.data(), order, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s));

    // ---- top-down split, one launch per level
    const int32_t cap = static_cast<int32_t>(std::min<int64_t>(4 * n + 64, INT32_MAX / 2));
    AacNode *nodes;
    uint32_t *ctr;
    AAC_CHECK(dmalloc((void **)&nodes, cap * sizeof(AacNode)));
    AAC_CHECK(dmalloc((void **)&ctr, 2 * sizeof(uint32_t)));
    const int32_t max_depth = 35;  // ssd/bvh_builder.h:23
    {
        AacNode root{};
        root.b = 0;
        root.e = static_cast<int32_t>(n);
        root.clo = 0;
        root.chi = 0x40000000;
        root.depth = 1;
        root.axis = 2;
        root.prim = 0;
        root.leaf = 0;
        for (int k = 0; k < 3; ++k) {
            root.bl[k] = hb[k];
            root.tr[k] = hb[3 + k];
        }
        root.kid[0] = root.kid[1] = -1;
        const uint32_t c0[2] = {1u, 0u};
        AAC_CHECK(hipMemcpyAsync(nodes, &root, sizeof(root), hipMemcpyHostToDevice, s));
        AAC_CHECK(hipMemcpyAsync(ctr, c0, sizeof(c0), hipMemcpyHostToDevice, s));
    }
    int32_t lo = 0, hi = 1;
    while (lo < hi) {
        SplitArgs sa{nodes, lo, hi, ctr, cap, ctr + 1, codes, reinterpret_cast<int32_t *>(order), tlo, thi, delta,
                     max_depth};
        hipLaunchKernelGGL(k_aac_split, dim3(static_cast<unsigned>((hi - lo + 255) / 256)), dim3(256), 0, s, sa);
        AAC_CHECK(hipGetLastError());
        uint32_t c[2];
        AAC_CHECK(hipMemcpyAsync(c, ctr, sizeof(c), hipMemcpyDeviceToHost, s));
        AAC_CHECK(hipStreamSynchronize(s));
        if (c[1]) {
            cleanup();
            return fail(ctx, RT_ERR_INVALID, "rt_build_aac: split tree larger than 4n nodes");
        }
        lo = hi;
        hi = static_cast<int32_t>(c[0]);
    }
    const int32_t nn = hi;
    std::vector<AacNode> hn(nn);
    AAC_CHECK(hipMemcpyAsync(hn.data(), nodes, nn * sizeof(AacNode), hipMemcpyDeviceToHost, s));
    AAC_CHECK(hipStreamSynchronize(s));

    // ---- leaf packing
    std::vector<int32_t> leaf_nodes;
    for (int32_t i = 0; i < nn; ++i)
        if (hn[i].leaf) leaf_nodes.push_back(i);
    int32_t *d_leaf_nodes, *d_nc;
    float4 *cl_lo, *cl_hi;
    int2 *cl_range;
    AAC_CHECK(dmalloc((void **)&d_leaf_nodes, leaf_nodes.size() * sizeof(int32_t)));
    AAC_CHECK(dmalloc((void **)&d_nc, leaf_nodes.size() * sizeof(int32_t)));
    AAC_CHECK(dmalloc((void **)&cl_lo, n * sizeof(float4)));
    AAC_CHECK(dmalloc((void **)&cl_hi, n * sizeof(float4)));
    AAC_CHECK(dmalloc((void **)&cl_range, n * sizeof(int2)));
    AAC_CHECK(hipMemcpyAsync(d_leaf_nodes, leaf_nodes.data(), leaf_nodes.size() * sizeof(int32_t),
                             hipMemcpyHostToDevice, s));
    PackArgs pk{nodes, d_leaf_nodes, static_cast<int32_t>(leaf_nodes.size()), reinterpret_cast<int32_t *>(order),
                tlo, thi, max_leaf_sah, cl_lo, cl_hi, cl_range, d_nc};
    hipLaunchKernelGGL(k_aac_pack, dim3(static_cast<unsigned>((leaf_nodes.size() + 255) / 256)), dim3(256), 0, s, pk);
    AAC_CHECK(hipGetLastError());
    std::vector<int32_t> nc(leaf_nodes.size());
    AAC_CHECK(hipMemcpyAsync(nc.data(), d_nc, nc.size() * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    AAC_CHECK(hipStreamSynchronize(s));

    // ---- per node: clusters before and after its merge, its cost triangle's
    // start, the ids of its leaves and of the nodes it makes.
    std::unordered_map<int64_t, int32_t> red_cache;
    auto red = [&](int64_t size) {
        auto it = red_cache.find(size);
        if (it != red_cache.end()) return it->second;
        const int32_t r = reduction(delta, alpha, size);
        red_cache.emplace(size, r);
        return r;
    };
    std::vector<int32_t> k0(nn, 0), kf(nn, 0), leaf_pos(nn, -1);
    for (size_t q = 0; q < leaf_nodes.size(); ++q) leaf_pos[leaf_nodes[q]] = static_cast<int32_t>(q);
    // Children are made after their parents, so descending index is bottom-up.
    for (int32_t i = nn - 1; i >= 0; --i) {
        const AacNode &nd = hn[i];
        if (nd.leaf) {
            const int32_t c = nc[leaf_pos[i]];
            k0[i] = c;
            kf[i] = std::min(c, red(std::max<int64_t>(c, static_cast<int64_t>(delta))));
        } else {
            int32_t sum = 0;
            for (int side = 0; side < 2; ++side)
                if (nd.kid[side] >= 0) sum += kf[nd.kid[side]];
            k0[i] = sum;
            kf[i] = std::min(sum, red(nd.e - nd.b));
        }
    }
    std::vector<int64_t> off(nn, 0);
    for (int32_t i = 0; i < nn; ++i) {
        const AacNode &nd = hn[i];
        if (nd.leaf) continue;
        if (nd.kid[0] >= 0) {
            off[nd.kid[0]] = off[i];
            if (nd.kid[1] >= 0) {
                const int64_t kl = kf[nd.kid[0]];
                off[nd.kid[1]] = off[i] + ((kl * (kl - 1)) >> 1);
            }
        } else if (nd.kid[1] >= 0) {
            off[nd.kid[1]] = off[i];
        }
    }
    int32_t leaves = 0;
    std::vector<int32_t> leaf_id(nn, 0);
    for (size_t q = 0; q < leaf_nodes.size(); ++q) {
        leaf_id[leaf_nodes[q]] = leaves;
        leaves += nc[q];
    }
    int64_t next_id = leaves;
    std::vector<int32_t> node_id(nn, 0);
    for (int32_t i = 0; i < nn; ++i) {
        node_id[i] = static_cast<int32_t>(next_id);
        next_id += k0[i] - kf[i];
    }
    const int64_t root_merges = kf[0] - 1;
    const int64_t total_nodes = next_id + root_merges;
    if (total_nodes > INT32_MAX) {
        cleanup();
        return fail(ctx, RT_ERR_INVALID, "rt_build_aac: too many nodes");
    }
    // Jobs grouped by split depth, deepest first; nodes with one child keep
    // their child's clusters in place (same list position, no merge) and need
    // no job.
    int32_t max_d = 0;
    for (const AacNode &nd : hn) max_d = std::max(max_d, nd.depth);
    std::vector<std::vector<MergeJob>> by_depth(max_d + 1);
    for (int32_t i = 0; i < nn; ++i) {
        const AacNode &nd = hn[i];
        MergeJob jb{};
        jb.node = i;
        jb.k0 = k0[i];
        jb.k = kf[i];
        jb.off = off[i];
        jb.b = nd.b;
        jb.node_id = node_id[i];
        if (nd.leaf) {
            jb.src = -1;
            jb.leaf_id = leaf_id[i];
        } else {
            const int nk = (nd.kid[0] >= 0 ? 1 : 0) + (nd.kid[1] >= 0 ? 1 : 0);
            if (nk == 1) continue;  // the child's list is already at nd.b, and no merge is due
            jb.src = 2;
            for (int side = 0; side < 2; ++side) {
                jb.kb[side] = hn[nd.kid[side]].b;
                jb.kk[side] = kf[nd.kid[side]];
            }
        }
        by_depth[nd.depth].push_back(jb);
    }
    int32_t *lists;
    float4 *bn_lo, *bn_hi;
    int2 *bn_kid;
    uint8_t *bn_leaf;
    MergeJob *d_jobs;
    AAC_CHECK(dmalloc((void **)&lists, n * sizeof(int32_t)));
    AAC_CHECK(dmalloc((void **)&bn_lo, total_nodes * sizeof(float4)));
    AAC_CHECK(dmalloc((void **)&bn_hi, total_nodes * sizeof(float4)));
    AAC_CHECK(dmalloc((void **)&bn_kid, total_nodes * sizeof(int2)));
    AAC_CHECK(dmalloc((void **)&bn_leaf, total_nodes));
    size_t max_jobs = 1;
    for (auto &v : by_depth) max_jobs = std::max(max_jobs, v.size());
    AAC_CHECK(dmalloc((void **)&d_jobs, max_jobs * sizeof(MergeJob)));
    Rows g{};
    int32_t g_stride = 0;
    int64_t g_cap = 0;
    auto run = [&](const std::vector<MergeJob> &jobs) -> hipError_t {
        if (jobs.empty()) return hipSuccess;
        std::vector<MergeJob> small, big;
        for (const MergeJob &jb : jobs) (jb.k0 <= kWaveList ? small : big).push_back(jb);
        MergeArgs ma{d_jobs, 0, lists, bn_lo, bn_hi, bn_kid, bn_leaf, cl_lo, cl_hi, cl_range};
        hipError_t e = hipSuccess;
        if (!small.empty()) {
            e = hipMemcpyAsync(d_jobs, small.data(), small.size() * sizeof(MergeJob), hipMemcpyHostToDevice, s);
            if (e != hipSuccess) return e;
            ma.count = static_cast<int32_t>(small.size());
            hipLaunchKernelGGL(k_aac_merge_wave, dim3(static_cast<unsigned>(small.size())), dim3(64), 0, s, ma);
            e = hipGetLastError();
            if (e != hipSuccess) return e;
            e = hipStreamSynchronize(s);  // d_jobs is reused below
            if (e != hipSuccess) return e;
        }
        if (!big.empty()) {
            int32_t kmax = 0;
            for (const MergeJob &jb : big) kmax = std::max(kmax, jb.k0);
            if (kmax > kBlockList) {
                // Lists too long for LDS: one row of kmax clusters per job.
                const int64_t need = static_cast<int64_t>(kmax) * static_cast<int64_t>(big.size());
                if (need > g_cap) {
                    for (void *p : {static_cast<void *>(g.ids), static_cast<void *>(g.lo), static_cast<void *>(g.hi),
                                    static_cast<void *>(g.rc), static_cast<void *>(g.rt), static_cast<void *>(g.todo)}) {
                        if (!p) continue;
                        bufs.erase(std::remove(bufs.begin(), bufs.end(), p), bufs.end());
                        (void)hipFree(p);
                    }
                    g = Rows{};
                    g_cap = 0;
                    e = dmalloc((void **)&g.ids, need * sizeof(int32_t));
                    if (e == hipSuccess) e = dmalloc((void **)&g.lo, need * sizeof(float4));
                    if (e == hipSuccess) e = dmalloc((void **)&g.hi, need * sizeof(float4));
                    if (e == hipSuccess) e = dmalloc((void **)&g.rc, need * sizeof(float));
                    if (e == hipSuccess) e = dmalloc((void **)&g.rt, need * sizeof(uint64_t));
                    if (e == hipSuccess) e = dmalloc((void **)&g.todo, need * sizeof(int32_t));
                    if (e != hipSuccess) return e;
                    g_cap = need;
                }
                g_stride = kmax;
            }
            e = hipMemcpyAsync(d_jobs, big.data(), big.size() * sizeof(MergeJob), hipMemcpyHostToDevice, s);
            if (e != hipSuccess) return e;
            ma.count = static_cast<int32_t>(big.size());
            hipLaunchKernelGGL(k_aac_merge_block, dim3(static_cast<unsigned>(big.size())), dim3(1024), 0, s, ma, g,
                               g_stride);
            e = hipGetLastError();
            if (e != hipSuccess) return e;
            e = hipStreamSynchronize(s);
        }
        return e;
    };
    for (int32_t d = max_d; d >= 1; --d) AAC_CHECK(run(by_depth[d]));
    // ssd/bvh_builder.cc:173-175: the root's clusters merged into one, its
    // cost triangle at the start of the matrix.
    if (kf[0] > 1) {
        MergeJob jb{};
        jb.node = 0;
        jb.src = 1;
        jb.k0 = kf[0];
        jb.k = 1;
        jb.off = 0;
        jb.b = 0;
        jb.kb[0] = 0;
        jb.kk[0] = kf[0];
        jb.node_id = static_cast<int32_t>(next_id);
        AAC_CHECK(run(std::vector<MergeJob>{jb}));
    }
    int32_t root = 0;
    AAC_CHECK(hipMemcpyAsync(&root, lists, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    out->nodes = total_nodes;
    out->root = root;
    out->lo.resize(total_nodes);
    out->hi.resize(total_nodes);
    out->kid.resize(total_nodes);
    out->leaf.resize(total_nodes);
    out->indirect.resize(n);
    AAC_CHECK(hipMemcpyAsync(out->lo.data(), bn_lo, total_nodes * sizeof(float4), hipMemcpyDeviceToHost, s));
    AAC_CHECK(hipMemcpyAsync(out->hi.data(), bn_hi, total_nodes * sizeof(float4), hipMemcpyDeviceToHost, s));
    AAC_CHECK(hipMemcpyAsync(out->kid.data(), bn_kid, total_nodes * sizeof(int2), hipMemcpyDeviceToHost, s));
    AAC_CHECK(hipMemcpyAsync(out->leaf.data(), bn_leaf, total_nodes, hipMemcpyDeviceToHost, s));
    AAC_CHECK(hipMemcpyAsync(out->indirect.data(), order, n * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    out->tri_lo.resize(n);
    out->tri_hi.resize(n);
    AAC_CHECK(hipMemcpyAsync(out->tri_lo.data(), tlo, n * sizeof(float4), hipMemcpyDeviceToHost, s));
    AAC_CHECK(hipMemcpyAsync(out->tri_hi.data(), thi, n * sizeof(float4), hipMemcpyDeviceToHost, s));
    AAC_CHECK(hipStreamSynchronize(s));
    out->split_nodes = nn;
    out->build_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    cleanup();
#undef AAC_CHECK
    return RT_OK;
}

namespace {

// One child slot of a wide node under construction: a binary node of the
// tree, or a run of a leaf's triangles (a leaf of more than kMaxLeafSize
// triangles is halved until each run fits one leaf code).
struct Slot {
    int32_t bin;         // binary node id, or -1 for a run
    int32_t b, e;        // the run (and a leaf node's range)
    float4 lo, hi;
    bool interior;       // an interior binary node
};

float half_area_host(const float4 &lo, const float4 &hi) {
    const float dx = hi.x - lo.x, dy = hi.y - lo.y, dz = hi.z - lo.z;
    return (dx * dy) + (dx * dz) + (dy * dz);
}

// Opens a slot into its two halves: an interior node's children, or a long
// run's halves (boxes from the triangles').
void open_slot(const AacTree &t, const Slot &x, Slot out[2]) {
    if (x.interior) {
        for (int side = 0; side < 2; ++side) {
            const int32_t c = side == 0 ? t.kid[x.bin].x : t.kid[x.bin].y;
            Slot &o = out[side];
            o.bin = c;
            o.interior = !t.leaf[c];
            o.lo = t.lo[c];
            o.hi = t.hi[c];
            o.b = o.interior ? 0 : t.kid[c].x;
            o.e = o.interior ? 0 : t.kid[c].y;
        }
        return;
    }
    const int32_t m = x.b + (x.e - x.b + 1) / 2;
    for (int side = 0; side < 2; ++side) {
        Slot &o = out[side];
        o.bin = -1;
        o.interior = false;
        o.b = side == 0 ? x.b : m;
        o.e = side == 0 ? m : x.e;
        o.lo = make_float4(INFINITY, INFINITY, INFINITY, 0.0f);
        o.hi = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.0f);
        for (int32_t k = o.b; k < o.e; ++k) {
            const float4 l = t.tri_lo[t.indirect[k]], h = t.tri_hi[t.indirect[k]];
            o.lo = make_float4(std::min(o.lo.x, l.x), std::min(o.lo.y, l.y), std::min(o.lo.z, l.z), 0.0f);
            o.hi = make_float4(std::max(o.hi.x, h.x), std::max(o.hi.y, h.y), std::max(o.hi.z, h.z), 0.0f);
        }
    }
}

bool openable(const Slot &x) { return x.interior || (x.e - x.b) > kMaxLeafSize; }

// Breadth-first collapse of the binary tree into W-wide nodes (so each
// level's nodes are contiguous, as refit expects): a node opens its
// largest-area openable child until it holds W children.
template <int W>
void collapse_host(const AacTree &t, const Slot &root, float pad, std::vector<WideNode<W>> *nodes,
                   std::vector<int64_t> *level_start, int64_t *leaves) {
    std::vector<Slot> cur{root}, next;
    *leaves = 0;
    while (!cur.empty()) {
        level_start->push_back(static_cast<int64_t>(nodes->size()));
        const int64_t base_next = static_cast<int64_t>(nodes->size() + cur.size());
        next.clear();
        for (const Slot &job : cur) {
            Slot c[W];
            int cnt = 0;
            open_slot(t, job, c);
            cnt = 2;
            while (cnt < W) {
                int best = -1;
                float ba = -1.0f;
                for (int k = 0; k < cnt; ++k)
                    if (openable(c[k]) && half_area_host(c[k].lo, c[k].hi) > ba) {
                        ba = half_area_host(c[k].lo, c[k].hi);
                        best = k;
                    }
                if (best < 0) break;
                Slot two[2];
                open_slot(t, c[best], two);
                c[best] = two[0];
                c[cnt++] = two[1];
            }
            WideNode<W> out{};
            for (int k = 0; k < W; ++k) {
                if (k >= cnt) {
                    out.lo_x[k] = out.lo_y[k] = out.lo_z[k] = INFINITY;
                    out.hi_x[k] = out.hi_y[k] = out.hi_z[k] = -INFINITY;
                    out.child[k] = kEmpty;
                    continue;
                }
                out.lo_x[k] = c[k].lo.x - pad;
                out.lo_y[k] = c[k].lo.y - pad;
                out.lo_z[k] = c[k].lo.z - pad;
                out.hi_x[k] = c[k].hi.x + pad;
                out.hi_y[k] = c[k].hi.y + pad;
                out.hi_z[k] = c[k].hi.z + pad;
                if (openable(c[k])) {
                    out.child[k] = static_cast<int32_t>(base_next + static_cast<int64_t>(next.size()));
                    next.push_back(c[k]);
                } else {
                    out.child[k] = leaf_code(c[k].b, c[k].e - c[k].b);
                    ++*leaves;
                }
            }
            nodes->push_back(out);
        }
        std::swap(cur, next);
    }
}

}  // namespace

int32_t build_accel_aac(rt_scene *scene, const rt_build_params &p, rt_accel **out) {
    rt_ctx *ctx = scene->ctx;
    const int W = p.width;
    if (W != 4 && W != 8) return fail(ctx, RT_ERR_INVALID, "width must be 4 or 8");
    const auto t0 = std::chrono::steady_clock::now();
    AacTree t;
    if (int32_t rc = build_aac_tree(scene, 0.4f, 20.0f, &t)) return rc;
    const int64_t n = scene->n;
    hipStream_t s = ctx->stream;
    rt_accel *a = new rt_accel();
    a->scene = scene;
    a->triangles = n;
    a->width = W;
    a->leaf_size = kMaxLeafSize;
    a->sah_bins = 0;
    a->lds_stack = (p.lds_stack <= 0) ? 16 : (p.lds_stack <= 4 ? 4 : (p.lds_stack <= 8 ? 8 : 16));
    auto bad = [&](hipError_t e, const char *what) {
        (void)hipStreamSynchronize(s);
        if (a->d_nodes) (void)hipFree(a->d_nodes);
        if (a->d_tris) (void)hipFree(a->d_tris);
        delete a;
        return fail(ctx, e == hipErrorOutOfMemory ? RT_ERR_OOM : RT_ERR_HIP,
                    std::string("AAC accel: ") + what + ": " + hipGetErrorString(e));
    };
    if (n == 0) {
        a->root = kEmpty;
        *out = a;
        return RT_OK;
    }
    const float4 rlo = t.lo[t.root], rhi = t.hi[t.root];
    float maxabs = 0.0f;
    for (const float v : {rlo.x, rlo.y, rlo.z, rhi.x, rhi.y, rhi.z}) maxabs = std::max(maxabs, std::fabs(v));
    a->box_pad = std::ldexp(std::max(maxabs, 1e-30f), -15);
    a->near_limit = 256.0f * maxabs;
    const float ext[3] = {rhi.x - rlo.x, rhi.y - rlo.y, rhi.z - rlo.z};
    a->scene_lo[0] = rlo.x;
    a->scene_lo[1] = rlo.y;
    a->scene_lo[2] = rlo.z;
    for (int k = 0; k < 3; ++k) a->scene_inv[k] = ext[k] > 0.0f ? 1.0f / ext[k] : 0.0f;
    // The triangles in the builder's order.
    uint32_t *d_order = nullptr;
    hipError_t e = hipMalloc(&d_order, n * sizeof(uint32_t));
    if (e != hipSuccess) return bad(e, "order");
    e = hipMemcpyAsync(d_order, t.indirect.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMalloc(&a->d_tris, n * sizeof(LeafTri));
    if (e == hipSuccess) e = make_leaf_tris(scene->d_abc, scene->d_flags, d_order, n, a->d_tris, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(d_order);
    if (e != hipSuccess) return bad(e, "triangles");
    Slot root;
    root.bin = t.root;
    root.interior = !t.leaf[t.root];
    root.lo = rlo;
    root.hi = rhi;
    root.b = root.interior ? 0 : t.kid[t.root].x;
    root.e = root.interior ? 0 : t.kid[t.root].y;
    if (!openable(root)) {
        a->root = leaf_code(root.b, root.e - root.b);
        a->nodes = 0;
        a->leaves = 1;
        a->depth = 0;
    } else {
        const float pad = a->box_pad;
        int64_t leaves = 0;
        size_t bytes = 0;
        std::vector<char> host;
        if (W == 4) {
            std::vector<WideNode<4>> nodes;
            collapse_host<4>(t, root, pad, &nodes, &a->level_start, &leaves);
            a->nodes = static_cast<int64_t>(nodes.size());
            bytes = nodes.size() * sizeof(WideNode<4>);
            host.resize(bytes);
            std::memcpy(host.data(), nodes.data(), bytes);
        } else {
            std::vector<WideNode<8>> nodes;
            collapse_host<8>(t, root, pad, &nodes, &a->level_start, &leaves);
            a->nodes = static_cast<int64_t>(nodes.size());
            bytes = nodes.size() * sizeof(WideNode<8>);
            host.resize(bytes);
            std::memcpy(host.data(), nodes.data(), bytes);
        }
        if (bytes > 0xffffffffull) return bad(hipErrorInvalidValue, "node array beyond 4 GiB");
        e = hipMalloc(&a->d_nodes, bytes);
        if (e == hipSuccess) e = hipMemcpyAsync(a->d_nodes, host.data(), bytes, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return bad(e, "nodes");
        a->leaves = leaves;
        a->depth = static_cast<int>(a->level_start.size());
        a->level_start.push_back(a->nodes);
        a->root = 0;
    }
    a->build_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    *out = a;
    return RT_OK;
}

}  // namespace rt
