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
// [9..11] max of AABB, all order-mapped.
__global__ void __launch_bounds__(256) k_tri_bounds(const float *__restrict__ abc, int64_t n,
                                                    float4 *__restrict__ aabb_lo, float4 *__restrict__ aabb_hi,
                                                    uint32_t *__restrict__ bounds) {
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
        if (is_min)
            atomicMin(&bounds[k], r);
        else
            atomicMax(&bounds[k], r);
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
__global__ void __launch_bounds__(256) k_bfs_level(const int32_t *__restrict__ level, int count,
                                                   const int2 *__restrict__ child, int32_t *__restrict__ next,
                                                   uint32_t *__restrict__ next_count) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    int2 ch = make_int2(-1, -1);
    if (k < count) ch = child[level[k]];
    const bool a = ch.x >= 0, b = ch.y >= 0;
    const uint64_t ma = __ballot(a), mb = __ballot(b);
    const int lane = threadIdx.x & 63;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint32_t total = static_cast<uint32_t>(__popcll(ma) + __popcll(mb));
    uint32_t base = 0;
    if (lane == 0 && total) base = atomicAdd(next_count, total);
    base = __shfl(base, 0, 64);
    if (a) next[base + __popcll(ma & lt)] = ch.x;
    if (b) next[base + __popcll(ma) + __popcll(mb & lt)] = ch.y;
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
    const int2 *q_in;        // (binary internal node, wide node index)
    int q_in_n;
    int2 *q_out;
    uint32_t *q_out_n;
    uint32_t *node_counter;
    void *nodes;
    int leaf_size;
    float pad;
    unsigned long long *leaf_counter;
};

template <int W>
__global__ void __launch_bounds__(256) k_collapse(CollapseArgs a) {
    const int qi = blockIdx.x * blockDim.x + threadIdx.x;
    if (qi >= a.q_in_n) return;
    const int2 job = a.q_in[qi];
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
            out.lo_x[k] = out.lo_y[k] = out.lo_z[k] = 0.0f;
            out.hi_x[k] = out.hi_y[k] = out.hi_z[k] = 0.0f;
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
            const uint32_t idx = atomicAdd(a.node_counter, 1u);
            out.child[k] = static_cast<int32_t>(idx);
            const uint32_t slot = atomicAdd(a.q_out_n, 1u);
            a.q_out[slot] = make_int2(x, static_cast<int32_t>(idx));
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

struct DevBuf {
    void *p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

template <class T>
hipError_t dalloc(DevBuf &b, size_t count) {
    return hipMalloc(&b.p, std::max<size_t>(1, count) * sizeof(T));
}

}  // namespace

int32_t build_accel(rt_scene *scene, const rt_build_params &p, rt_accel **out) {
    rt_ctx *ctx = scene->ctx;
    const int W = p.width;
    if (W != 4 && W != 8) return fail(ctx, RT_ERR_INVALID, "width must be 4 or 8");
    if (p.leaf_size < 1 || p.leaf_size > kMaxLeafSize) return fail(ctx, RT_ERR_INVALID, "leaf_size must be 1..16");
    const int64_t n = scene->n;
    if (n >= (int64_t(1) << 27)) return fail(ctx, RT_ERR_INVALID, "too many triangles (limit 2^27)");
    hipStream_t s = ctx->stream;
    const auto t0 = std::chrono::steady_clock::now();

    rt_accel *a = new rt_accel();
    a->scene = scene;
    a->width = W;
    a->leaf_size = p.leaf_size;
    a->lds_stack = (p.lds_stack <= 0) ? 16 : (p.lds_stack <= 4 ? 4 : (p.lds_stack <= 8 ? 8 : 16));
    auto cleanup = [&]() {
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
    BUILD_CHECK(dalloc<float4>(b_aabb_lo, n));
    BUILD_CHECK(dalloc<float4>(b_aabb_hi, n));
    BUILD_CHECK(dalloc<uint32_t>(b_bounds, 12));
    BUILD_CHECK(dalloc<uint32_t>(b_codes, n));
    BUILD_CHECK(dalloc<uint32_t>(b_idx, n));
    BUILD_CHECK(dalloc<uint32_t>(b_codes_t, n));
    BUILD_CHECK(dalloc<uint32_t>(b_idx_t, n));
    BUILD_CHECK(dalloc<uint32_t>(b_hist, 256 * tiles));
    BUILD_CHECK(dalloc<uint32_t>(b_sums, kScanTile));
    uint32_t init_bounds[12];
    for (int k = 0; k < 12; ++k) {
        const bool is_min = (k < 3) || (k >= 6 && k < 9);
        init_bounds[k] = is_min ? 0xffffffffu : 0u;
    }
    BUILD_CHECK(hipMemcpyAsync(b_bounds.p, init_bounds, sizeof(init_bounds), hipMemcpyHostToDevice, s));
    const int blocks = static_cast<int>((n + 255) / 256);
    hipLaunchKernelGGL(k_tri_bounds, dim3(blocks), dim3(256), 0, s, scene->d_abc, n, (float4 *)b_aabb_lo.p,
                       (float4 *)b_aabb_hi.p, (uint32_t *)b_bounds.p);
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

    // Triangles in BVH (sorted) order.
    BUILD_CHECK(hipMalloc(&a->d_tris, n * sizeof(LeafTri)));
    hipLaunchKernelGGL(k_leaf_tris, dim3(blocks), dim3(256), 0, s, scene->d_abc, scene->d_flags,
                       (const uint32_t *)b_idx.p, n, a->d_tris);
    BUILD_CHECK(hipGetLastError());

    if (n == 1) {
        // A single triangle: the root is a leaf, no wide node is needed.
        a->root = leaf_code(0, 1);
        a->nodes = 0;
        a->leaves = 1;
        a->depth = 0;
        BUILD_CHECK(hipStreamSynchronize(s));
        a->build_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
        *out = a;
        return RT_OK;
    }

    DevBuf b_child, b_range, b_parent, b_leaf_parent, b_leaf_lo, b_leaf_hi, b_node_lo, b_node_hi;
    BUILD_CHECK(dalloc<int2>(b_child, n - 1));
    BUILD_CHECK(dalloc<int2>(b_range, n - 1));
    BUILD_CHECK(dalloc<int32_t>(b_parent, n - 1));
    BUILD_CHECK(dalloc<int32_t>(b_leaf_parent, n));
    BUILD_CHECK(dalloc<float4>(b_leaf_lo, n));
    BUILD_CHECK(dalloc<float4>(b_leaf_hi, n));
    BUILD_CHECK(dalloc<float4>(b_node_lo, n - 1));
    BUILD_CHECK(dalloc<float4>(b_node_hi, n - 1));
    const int iblocks = static_cast<int>((n - 1 + 255) / 256);
    hipLaunchKernelGGL(k_karras, dim3(iblocks), dim3(256), 0, s, (const uint32_t *)b_codes.p, n, (int2 *)b_child.p,
                       (int2 *)b_range.p, (int32_t *)b_parent.p, (int32_t *)b_leaf_parent.p);
    hipLaunchKernelGGL(k_gather_leaf_boxes, dim3(blocks), dim3(256), 0, s, (const uint32_t *)b_idx.p, n,
                       (const float4 *)b_aabb_lo.p, (const float4 *)b_aabb_hi.p, (float4 *)b_leaf_lo.p,
                       (float4 *)b_leaf_hi.p);
    // Level-synchronous refit (see k_bfs_level).
    {
        DevBuf b_order, b_cnt;
        BUILD_CHECK(dalloc<int32_t>(b_order, n - 1));
        BUILD_CHECK(dalloc<uint32_t>(b_cnt, 1));
        int32_t *order = (int32_t *)b_order.p;
        uint32_t *cnt = (uint32_t *)b_cnt.p;
        const int32_t zero = 0;
        BUILD_CHECK(hipMemcpyAsync(order, &zero, sizeof(int32_t), hipMemcpyHostToDevice, s));
        std::vector<std::pair<uint32_t, uint32_t>> levels;  // (offset, count)
        uint32_t off = 0, count = 1;
        while (count > 0) {
            levels.push_back({off, count});
            if (static_cast<int>(levels.size()) > kMaxLevels) {
                cleanup();
                return fail(ctx, RT_ERR_INVALID, "binary tree deeper than 256 levels");
            }
            BUILD_CHECK(hipMemsetAsync(cnt, 0, sizeof(uint32_t), s));
            hipLaunchKernelGGL(k_bfs_level, dim3((count + 255) / 256), dim3(256), 0, s, order + off,
                               static_cast<int>(count), (const int2 *)b_child.p, order + off + count, cnt);
            uint32_t next = 0;
            BUILD_CHECK(hipMemcpyAsync(&next, cnt, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
            BUILD_CHECK(hipStreamSynchronize(s));
            off += count;
            count = next;
        }
        for (int l = static_cast<int>(levels.size()) - 1; l >= 0; --l) {
            hipLaunchKernelGGL(k_refit_level, dim3((levels[l].second + 255) / 256), dim3(256), 0, s,
                               order + levels[l].first, static_cast<int>(levels[l].second), (const int2 *)b_child.p,
                               (const float4 *)b_leaf_lo.p, (const float4 *)b_leaf_hi.p, (float4 *)b_node_lo.p,
                               (float4 *)b_node_hi.p);
        }
        BUILD_CHECK(hipGetLastError());
        a->binary_depth = static_cast<int>(levels.size());
    }

    // Wide nodes: at most n - 1 (each wide node consumes >= 1 binary node).
    const size_t node_bytes = (W == 4) ? sizeof(WideNode<4>) : sizeof(WideNode<8>);
    const int64_t max_nodes = std::max<int64_t>(1, n - 1);
    BUILD_CHECK(hipMalloc(&a->d_nodes, max_nodes * node_bytes));
    DevBuf b_qa, b_qb, b_counters;
    BUILD_CHECK(dalloc<int2>(b_qa, max_nodes));
    BUILD_CHECK(dalloc<int2>(b_qb, max_nodes));
    BUILD_CHECK(dalloc<unsigned long long>(b_counters, 4));
    BUILD_CHECK(hipMemsetAsync(b_counters.p, 0, 4 * sizeof(unsigned long long), s));
    const int2 root_job = make_int2(0, 0);
    BUILD_CHECK(hipMemcpyAsync(b_qa.p, &root_job, sizeof(int2), hipMemcpyHostToDevice, s));
    const uint32_t one = 1;
    uint32_t *node_counter = reinterpret_cast<uint32_t *>(b_counters.p);
    uint32_t *qout_n = node_counter + 1;
    unsigned long long *leaf_counter = reinterpret_cast<unsigned long long *>(b_counters.p) + 1;
    BUILD_CHECK(hipMemcpyAsync(node_counter, &one, sizeof(uint32_t), hipMemcpyHostToDevice, s));
    int qn = 1;
    int depth = 0;
    int2 *qin = (int2 *)b_qa.p, *qout = (int2 *)b_qb.p;
    while (qn > 0) {
        BUILD_CHECK(hipMemsetAsync(qout_n, 0, sizeof(uint32_t), s));
        CollapseArgs ca;
        ca.child = (const int2 *)b_child.p;
        ca.range = (const int2 *)b_range.p;
        ca.node_lo = (const float4 *)b_node_lo.p;
        ca.node_hi = (const float4 *)b_node_hi.p;
        ca.leaf_lo = (const float4 *)b_leaf_lo.p;
        ca.leaf_hi = (const float4 *)b_leaf_hi.p;
        ca.q_in = qin;
        ca.q_in_n = qn;
        ca.q_out = qout;
        ca.q_out_n = qout_n;
        ca.node_counter = node_counter;
        ca.nodes = a->d_nodes;
        ca.leaf_size = p.leaf_size;
        ca.pad = a->box_pad;
        ca.leaf_counter = leaf_counter;
        const int cb = (qn + 255) / 256;
        if (W == 4)
            hipLaunchKernelGGL(k_collapse<4>, dim3(cb), dim3(256), 0, s, ca);
        else
            hipLaunchKernelGGL(k_collapse<8>, dim3(cb), dim3(256), 0, s, ca);
        BUILD_CHECK(hipGetLastError());
        uint32_t next = 0;
        BUILD_CHECK(hipMemcpyAsync(&next, qout_n, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        BUILD_CHECK(hipStreamSynchronize(s));
        qn = static_cast<int>(next);
        std::swap(qin, qout);
        ++depth;
    }
    uint64_t counters[2];
    BUILD_CHECK(hipMemcpyAsync(counters, b_counters.p, 2 * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    BUILD_CHECK(hipStreamSynchronize(s));
    a->nodes = static_cast<int64_t>(counters[0] & 0xffffffffull);
    a->leaves = static_cast<int64_t>(counters[1]);
    a->depth = depth;
    a->root = 0;
    a->build_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    *out = a;
#undef BUILD_CHECK
    return RT_OK;
}

}  // namespace rt
