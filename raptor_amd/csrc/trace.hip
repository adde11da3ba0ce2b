// Per-launch traversal kernels (closest hit and any hit) and their launchers.
// The traversal itself is in traverse.hpp.
#include "traverse.hpp"

namespace rt {

namespace {

struct TraceArgs {
    const uint32_t *order;   // if non-null, lane i traces ray order[i] (regrouped rays)
    float near_limit;        // origins with all |o_k| <= near_limit use the fused box test
    const void *nodes;
    const LeafTri *tris;
    int32_t root;
    const rt_ray *rays;
    const uint32_t *count;   // if non-null, n = count[0] * count_mul
    uint32_t count_mul;
    int64_t n;
    rt_hit *hits;
    uint8_t *occ;
    unsigned long long *counters;  // [0] node visits, [1] triangle tests
    uint32_t *overflow;            // set to 1 if a stack overflowed
};

// STAGE names the launch (0 primary, 2 bounce, 4 API: closest hit; 1 shadow,
// 3 bounce shadow, 5 API: any hit) so each has its own kernel symbol in
// profiles; even stages are closest-hit queries.
template <int W, int STAGE, bool COUNT, int kLdsDepth>
__global__ void __launch_bounds__(64) k_trace(TraceArgs a) {
    constexpr bool CLOSEST = (STAGE % 2) == 0;
    const int lane = threadIdx.x;
    const int64_t i = static_cast<int64_t>(blockIdx.x) * 64 + lane;
    const int64_t n = a.count ? static_cast<int64_t>(a.count[0]) * a.count_mul : a.n;
    if (i >= n) return;
    const int64_t ri = a.order ? static_cast<int64_t>(a.order[i]) : i;
    const float4 r0 = reinterpret_cast<const float4 *>(a.rays)[2 * ri];
    const float4 r1 = reinterpret_cast<const float4 *>(a.rays)[2 * ri + 1];
    float4 hit;
    uint32_t visits = 0, tests = 0;
    bool overflow = false;
    trace_ray<W, CLOSEST, COUNT, kLdsDepth>(reinterpret_cast<const WideNode<W> *>(a.nodes), a.tris, a.root,
                                            a.near_limit, rtd::f4v(r0), rtd::f4v(r1), r0.w, r1.w, lane,
                                            hit, visits, tests, overflow);
    const int32_t best_id = __float_as_int(hit.w);
    if (CLOSEST) {
        const float ht = (best_id >= 0) ? hit.x : rtd::kMaxDist;
        const float hu = (best_id >= 0) ? hit.y : 0.0f;
        const float hv = (best_id >= 0) ? hit.z : 0.0f;
        reinterpret_cast<float4 *>(a.hits)[ri] = make_float4(ht, hu, hv, __int_as_float(best_id));
    } else {
        a.occ[ri] = (best_id >= 0) ? 1 : 0;
    }
    if (COUNT) {
        unsigned long long vs = visits, ts = tests;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            vs += __shfl_xor(vs, off, 64);
            ts += __shfl_xor(ts, off, 64);
        }
        if (lane == 0) {
            atomicAdd(&a.counters[0], vs);
            atomicAdd(&a.counters[1], ts);
        }
    }
    if (overflow && a.overflow) atomicOr(a.overflow, 1u);
}

template <int W, int STAGE, int D>
void launch_wd(const TraceArgs &ta, int64_t blocks, hipStream_t s) {
    if (ta.counters)
        hipLaunchKernelGGL((k_trace<W, STAGE, true, D>), dim3(blocks), dim3(64), 0, s, ta);
    else
        hipLaunchKernelGGL((k_trace<W, STAGE, false, D>), dim3(blocks), dim3(64), 0, s, ta);
}

template <int W, int STAGE>
void launch_w(int depth, const TraceArgs &ta, int64_t blocks, hipStream_t s) {
    if (depth <= 4)
        launch_wd<W, STAGE, 4>(ta, blocks, s);
    else if (depth <= 8)
        launch_wd<W, STAGE, 8>(ta, blocks, s);
    else
        launch_wd<W, STAGE, 16>(ta, blocks, s);
}

template <int STAGE>
void launch_stage(int width, int depth, const TraceArgs &ta, int64_t blocks, hipStream_t s) {
    if (width == 4)
        launch_w<4, STAGE>(depth, ta, blocks, s);
    else
        launch_w<8, STAGE>(depth, ta, blocks, s);
}

int32_t launch(rt_accel *acc, TraceArgs &ta, int64_t max_n, int stage, hipStream_t s) {
    rt_ctx *ctx = acc->scene->ctx;
    if (max_n <= 0) return RT_OK;
    ta.nodes = acc->d_nodes;
    ta.tris = acc->d_tris;
    ta.root = acc->root;
    ta.near_limit = acc->near_limit;
    ta.counters = ctx->counting ? ctx->d_counters + 2 * ctx->count_slot : nullptr;
    ta.overflow = ctx->counting ? reinterpret_cast<uint32_t *>(ctx->d_counters + 15) : nullptr;
    const int64_t blocks = (max_n + 63) / 64;
    switch (stage) {
        case 0: launch_stage<0>(acc->width, acc->lds_stack, ta, blocks, s); break;
        case 1: launch_stage<1>(acc->width, acc->lds_stack, ta, blocks, s); break;
        case 2: launch_stage<2>(acc->width, acc->lds_stack, ta, blocks, s); break;
        case 3: launch_stage<3>(acc->width, acc->lds_stack, ta, blocks, s); break;
        case 4: launch_stage<4>(acc->width, acc->lds_stack, ta, blocks, s); break;
        default: launch_stage<5>(acc->width, acc->lds_stack, ta, blocks, s); break;
    }
    RT_HIP_CHECK(ctx, hipGetLastError());
    return RT_OK;
}

__global__ void k_fill_miss(rt_hit *hits, uint8_t *occ, const uint32_t *count, uint32_t mul, int64_t n_static) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t n = count ? static_cast<int64_t>(count[0]) * mul : n_static;
    if (i >= n) return;
    if (hits) reinterpret_cast<float4 *>(hits)[i] = make_float4(rtd::kMaxDist, 0.0f, 0.0f, __int_as_float(-1));
    if (occ) occ[i] = 0;
}

int32_t empty_scene(rt_accel *acc, rt_hit *hits, uint8_t *occ, const uint32_t *count, uint32_t mul, int64_t max_n,
                    hipStream_t s) {
    if (max_n <= 0) return RT_OK;
    hipLaunchKernelGGL(k_fill_miss, dim3((max_n + 255) / 256), dim3(256), 0, s, hits, occ, count, mul, max_n);
    RT_HIP_CHECK(acc->scene->ctx, hipGetLastError());
    return RT_OK;
}

}  // namespace

int32_t launch_closest(rt_accel *acc, const rt_ray *d_rays, int64_t n, rt_hit *d_hits, hipStream_t s, int stage) {
    if (acc->root == kEmpty) return empty_scene(acc, d_hits, nullptr, nullptr, 1, n, s);
    TraceArgs ta{};
    ta.rays = d_rays;
    ta.n = n;
    ta.hits = d_hits;
    return launch(acc, ta, n, stage, s);
}

int32_t launch_occluded(rt_accel *acc, const rt_ray *d_rays, int64_t n, uint8_t *d_occ, hipStream_t s, int stage) {
    if (acc->root == kEmpty) return empty_scene(acc, nullptr, d_occ, nullptr, 1, n, s);
    TraceArgs ta{};
    ta.rays = d_rays;
    ta.n = n;
    ta.occ = d_occ;
    return launch(acc, ta, n, stage, s);
}

int32_t launch_closest_q(rt_accel *acc, const rt_ray *d_rays, const uint32_t *d_count, uint32_t mul, int64_t max_n,
                         rt_hit *d_hits, hipStream_t s, int stage, const uint32_t *order) {
    if (acc->root == kEmpty) return empty_scene(acc, d_hits, nullptr, d_count, mul, max_n, s);
    TraceArgs ta{};
    ta.order = order;
    ta.rays = d_rays;
    ta.count = d_count;
    ta.count_mul = mul;
    ta.n = max_n;
    ta.hits = d_hits;
    return launch(acc, ta, max_n, stage, s);
}

int32_t launch_occluded_q(rt_accel *acc, const rt_ray *d_rays, const uint32_t *d_count, uint32_t mul, int64_t max_n,
                          uint8_t *d_occ, hipStream_t s, int stage) {
    if (acc->root == kEmpty) return empty_scene(acc, nullptr, d_occ, d_count, mul, max_n, s);
    TraceArgs ta{};
    ta.rays = d_rays;
    ta.count = d_count;
    ta.count_mul = mul;
    ta.n = max_n;
    ta.occ = d_occ;
    return launch(acc, ta, max_n, stage, s);
}

}  // namespace rt
