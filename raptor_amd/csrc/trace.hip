// Per-launch traversal kernels (closest hit and any hit) and their launchers.
// The traversal itself is in traverse.hpp.
#include <algorithm>

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
    uint32_t *queue;               // non-null: the block-local form with ray refills
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
    // Lanes past the end leave at once, except when counting: the count
    // reduction below shuffles across the whole wave, and a lane that has
    // returned would lend it a stale register.
    if (!COUNT && i >= n) return;
    uint32_t visits = 0, tests = 0;
    bool overflow = false;
    if (i < n) {
        const int64_t ri = a.order ? static_cast<int64_t>(a.order[i]) : i;
        const float4 r0 = reinterpret_cast<const float4 *>(a.rays)[2 * ri];
        const float4 r1 = reinterpret_cast<const float4 *>(a.rays)[2 * ri + 1];
        float4 hit;
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


// Block-local form with dynamic ray fetch (after Aila and Laine's
// while-while traversal with replacement of terminated rays): a 256-thread
// block owns kBlockRays rays, 16 tiles of 64 consecutive rays spread over the
// whole queue (tile j * G + block, G working blocks); each lane holds one ray
// and, after every leaf phase, once at least kRefill lanes of a wave have
// finished, the wave takes that many new rays from the block's LDS cursor.
// Incoherent rays end at very different depths; without refills the finished
// lanes ride along idle until the wave's slowest ray is done. All blocks are
// resident at once, so the launch lasts as long as its slowest block: spread
// tiles even the blocks' work out (1-2% on the bounce launch against 16
// consecutive tiles). (A single global queue instead serialises its atomics
// across the XCDs and was 2x slower; waves taking 256-ray chunks from a
// global counter lose the tile locality of a block's waves and were 20%
// slower.)
// Arithmetic and tie rules are those of trace_ray, so the hits are identical.
constexpr int kRefill = 16;
constexpr int kBlockThreads = 256;
constexpr int kBlockRays = 1024;
// Launch stages (bit = STAGE) that use the block-local form. Measured on the
// 1M-triangle terrain frame: the incoherent bounce closest-hit launch gains
// 10% (0.593 -> 0.534 ms); coherent primary and shadow rays lose (0.40 ->
// 0.54 ms primary) because refills break the 8x8-tile order of a wave; the
// bounce-shadow any-hit launch (494k rays from scattered bounce hits, one
// round of waves) loses too (0.092 -> 0.108 ms).
constexpr unsigned kPersistentStages = 1u << 2;

template <int W, int STAGE, bool COUNT, int kLdsDepth>
__global__ void __launch_bounds__(kBlockThreads) k_trace_pers(TraceArgs a) {
    constexpr bool CLOSEST = (STAGE % 2) == 0;
    constexpr int kSpillDepth = kStackDepth - kLdsDepth;
    __shared__ int2 lds_stack[kLdsDepth][kBlockThreads];
    __shared__ uint32_t cursor;
    const int lane = threadIdx.x & 63;
    const int slot = threadIdx.x;  // LDS stack column
    const int64_t n = a.count ? static_cast<int64_t>(a.count[0]) * a.count_mul : a.n;
    const int64_t nblk = (n + kBlockRays - 1) / kBlockRays;  // working blocks
    if (threadIdx.x == 0) cursor = 0u;
    __syncthreads();
    const WideNode<W> *nodes = reinterpret_cast<const WideNode<W> *>(a.nodes);
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));

    int64_t ri = -1;  // ray of this lane, -1: none
    rtd::V3 o = rtd::mk(0.0f, 0.0f, 0.0f), d = o, id = o, oid = o;
    PlaneOffsets po{};
    float tmin = 0.0f, tmax = 0.0f;
    bool near = false;
    int32_t spill_code[kSpillDepth];
    float spill_t[kSpillDepth];
    int sp = 0;
    float best_t = 0.0f, best_u = 0.0f, best_v = 0.0f;
    int32_t best_id = -1;
    int32_t cur = kEmpty;
    uint32_t visits = 0, tests = 0;
    bool overflow = false;

    auto pop = [&]() -> int32_t {
        while (sp > 0) {
            --sp;
            int2 e;
            if (sp < kLdsDepth) {
                e = lds_stack[sp][slot];
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
            lds_stack[sp][slot] = make_int2(code, __float_as_int(t));
            ++sp;
        } else if (sp < kLdsDepth + kSpillDepth) {
            spill_code[sp - kLdsDepth] = code;
            spill_t[sp - kLdsDepth] = t;
            ++sp;
        } else {
            overflow = true;
        }
    };

    bool drained = static_cast<int64_t>(blockIdx.x) >= nblk;  // the block's rays are all handed out
    int32_t leaf = kEmpty;
    while (true) {
        // ---- refill finished lanes from the block's cursor
        const uint64_t idle = __ballot(ri < 0);
        const int nidle = __popcll(idle);
        if (!drained && nidle >= kRefill) {
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(&cursor, static_cast<uint32_t>(nidle));
            base = __builtin_amdgcn_readfirstlane(__shfl(base, 0, 64));
            if (static_cast<int64_t>(base) + nidle >= kBlockRays) drained = true;
            if (ri < 0) {
                // The block's k-th ray lies in tile (k / 64) * G + block of 64
                // rays, G the number of working blocks: a block's 16 tiles are
                // spread over the whole queue, which evens out the blocks' work.
                const int64_t k = static_cast<int64_t>(base) + __popcll(idle & lt_mask);
                const int64_t idx = ((k >> 6) * nblk + blockIdx.x) * 64 + (k & 63);
                if (k < kBlockRays && idx < n) {
                    ri = a.order ? static_cast<int64_t>(a.order[idx]) : idx;
                    const float4 r0 = reinterpret_cast<const float4 *>(a.rays)[2 * ri];
                    const float4 r1 = reinterpret_cast<const float4 *>(a.rays)[2 * ri + 1];
                    o = rtd::f4v(r0);
                    d = rtd::f4v(r1);
                    tmin = r0.w;
                    tmax = r1.w;
                    id = rtd::mk(safe_inv(d.x), safe_inv(d.y), safe_inv(d.z));
                    oid = rtd::mk(-(o.x * id.x), -(o.y * id.y), -(o.z * id.z));
                    po = plane_offsets<W>(id);
                    near = fmaxf(fmaxf(fabsf(o.x), fabsf(o.y)), fabsf(o.z)) <= a.near_limit;
                    best_t = tmax;
                    best_u = 0.0f;
                    best_v = 0.0f;
                    best_id = -1;
                    sp = 0;
                    cur = a.root;
                }
            }
        }
        if (!__any(ri >= 0)) break;

        // ---- interior phase (trace_ray's, per lane)
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
            if (cur != kEmpty && cur < 0 && leaf == kEmpty) {
                leaf = cur;
                cur = pop();
            }
            const bool searching = (cur >= 0) && (leaf == kEmpty);
            if (__popcll(__ballot(searching)) <= kSearchCut) break;
        }
        // ---- leaf phase
        bool occluded = false;
        while (true) {
            if (leaf != kEmpty) {
                const int32_t first = leaf_first(leaf);
                const int32_t count = leaf_count(leaf);
                for (int k = 0; k < count; ++k) {
                    const LeafTri lt = a.tris[first + k];
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
                if (cur != kEmpty && cur < 0) {
                    leaf = cur;
                    cur = pop();
                }
            }
            if (__popcll(__ballot(leaf != kEmpty)) <= kLeafCut) break;
        }
        // ---- finished rays write their result and free the lane
        if (ri >= 0 && ((cur == kEmpty && leaf == kEmpty) || (!CLOSEST && occluded))) {
            if (CLOSEST) {
                const float ht = (best_id >= 0) ? best_t : rtd::kMaxDist;
                const float hu = (best_id >= 0) ? best_u : 0.0f;
                const float hv = (best_id >= 0) ? best_v : 0.0f;
                reinterpret_cast<float4 *>(a.hits)[ri] = make_float4(ht, hu, hv, __int_as_float(best_id));
            } else {
                a.occ[ri] = occluded ? 1 : 0;
            }
            ri = -1;
            cur = kEmpty;
            leaf = kEmpty;
        }
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

// Waves of the persistent form: as many as are resident at once.
template <int W, int STAGE, int D>
void launch_wd(const TraceArgs &ta, int64_t blocks, hipStream_t s) {
    if (ta.queue) {
        const dim3 grid(static_cast<unsigned>((blocks * 64 + kBlockRays - 1) / kBlockRays));
        if (ta.counters)
            hipLaunchKernelGGL((k_trace_pers<W, STAGE, true, D>), grid, dim3(kBlockThreads), 0, s, ta);
        else
            hipLaunchKernelGGL((k_trace_pers<W, STAGE, false, D>), grid, dim3(kBlockThreads), 0, s, ta);
        return;
    }
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
    ta.queue = ((kPersistentStages >> stage) & 1) ? ctx->d_queue : nullptr;
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
