// Batched tracing of many scenes' ray batches in one launch.
//
// The reference traces one scene per ray_tracer() call (rt/raytracer.cc:563-598)
// and has no batching. Here a batch is a table of (BVH, rays, hits) items that
// share a node width. Two ways to run it:
//   persistent  one launch of k_trace_batch: a fixed grid of wavefronts pulls
//               chunks of 64 rays from a global counter (grab chunks per
//               atomic), finds the chunk's scene by binary search over the
//               per-scene chunk offsets and traces it with that scene's BVH.
//               Small scenes and small ray batches no longer leave the GPU
//               idle between launches. The counter reset and the launch are
//               captured once in a hipGraph and replayed per frame.
//   plain       one closest-hit launch per scene (the baseline).
// The persistent form comes in two shapes: chunk at a time (a wavefront traces
// its 64-ray chunk to the last ray) and refill (k_trace_batch_refill: a
// wavefront keeps a pool of ray slots from the chunks it took and hands new
// rays to its finished lanes once `refill` of them are idle), each with 1, 2
// or 4 wavefronts per workgroup. rt_batch_autotune times node width, LDS
// stack depth, resident wavefronts per CU, wavefronts per workgroup, the
// refill threshold and chunks per grab, and returns the fastest parameters.
#include <algorithm>
#include <chrono>
#include <vector>

#include "traverse.hpp"

struct rt_batch {
    rt_ctx *ctx = nullptr;
    int width = 4;
    rt_batch_params params{};
    std::vector<rt_batch_item> items;
    std::vector<rt_accel *> owned;  // BVHs the batch built at another node width
    std::vector<rt_accel *> traced; // per item: the BVH traced (the item's or an owned one)
    void *d_scenes = nullptr;      // BatchScene[nscenes]
    uint32_t *d_counter = nullptr;  // [0] chunk queue head, [1] stack overflow flag
    uint32_t total_chunks = 0;
    int nscenes = 0;
    int grid = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

namespace rt {

namespace {

struct BatchScene {
    const void *nodes;
    const LeafTri *tris;
    const float4 *rays;
    float4 *hits;
    int32_t root;
    float near_limit;
    uint32_t chunk_begin;
    uint32_t n_rays;
};

struct BatchArgs {
    const BatchScene *scenes;
    int nscenes;
    uint32_t total_chunks;
    uint32_t grab;
    uint32_t *counter;
    uint32_t *overflow;
};

// A fill kernel rather than a memset so the reset can sit inside a graph.
__global__ void k_reset_counter(uint32_t *c) {
    if (threadIdx.x < 2) c[threadIdx.x] = 0u;
}

// The scene of global chunk c: the last scene whose first chunk is <= c.
__device__ __forceinline__ int scene_of_chunk(const BatchArgs &a, uint32_t c) {
    int lo = 0, hi = a.nscenes - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (a.scenes[mid].chunk_begin <= c)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

template <int W, int kLdsDepth, int kWaves>
__global__ void __launch_bounds__(64 * kWaves) k_trace_batch(BatchArgs a) {
    const int lane = threadIdx.x & 63;
    while (true) {
        uint32_t g = 0;
        if (lane == 0) g = atomicAdd(a.counter, a.grab);
        g = __builtin_amdgcn_readfirstlane(__shfl(g, 0, 64));
        if (g >= a.total_chunks) break;
        const uint32_t g_end = min(g + a.grab, a.total_chunks);
        for (uint32_t c = g; c < g_end; ++c) {
            const BatchScene sc = a.scenes[scene_of_chunk(a, c)];
            const uint32_t i = (c - sc.chunk_begin) * 64u + static_cast<uint32_t>(lane);
            if (i < sc.n_rays) {
                const float4 r0 = sc.rays[2 * static_cast<int64_t>(i)];
                const float4 r1 = sc.rays[2 * static_cast<int64_t>(i) + 1];
                float4 hit;
                uint32_t visits = 0, tests = 0;
                bool overflow = false;
                trace_ray<W, true, false, kLdsDepth, 64 * kWaves>(reinterpret_cast<const WideNode<W> *>(sc.nodes),
                                                                  sc.tris, sc.root, sc.near_limit, rtd::f4v(r0),
                                                                  rtd::f4v(r1), r0.w, r1.w, static_cast<int>(threadIdx.x),
                                                                  hit, visits, tests, overflow);
                const int32_t id = __float_as_int(hit.w);
                sc.hits[i] = (id >= 0) ? hit : make_float4(rtd::kMaxDist, 0.0f, 0.0f, __int_as_float(-1));
                if (overflow) atomicOr(a.overflow, 1u);
            }
        }
    }
}

// The refill form: Aila and Laine's while-while traversal with replacement of
// finished rays, as the frame's block-local kernel runs it, over a
// wavefront's pool of ray slots (64 per chunk taken from the global queue).
// A slot past its scene's last ray is skipped. Every lane carries its own
// scene's BVH, triangles and hit buffer; arithmetic and tie rules are
// trace_ray's, so the hits are those of the chunk-at-a-time form.
template <int W, int kLdsDepth, int kWaves>
__global__ void __launch_bounds__(64 * kWaves) k_trace_batch_refill(BatchArgs a, int refill) {
    constexpr int kSpillDepth = kStackDepth - kLdsDepth;
    __shared__ int2 lds_stack[kLdsDepth][64 * kWaves];
    const int lane = threadIdx.x & 63;
    const int slot = threadIdx.x;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint32_t pool = 0, pool_end = 0;  // the wavefront's ray slots [pool, pool_end)
    int32_t leaf = kEmpty;            // a leaf still to test (it may wait a round)
    bool drained = false;             // the global queue is empty

    int32_t scene = -1;  // scene of this lane's ray, -1: idle
    uint32_t ri = 0;
    const WideNode<W> *nodes = nullptr;
    const LeafTri *tris = nullptr;
    rtd::V3 o = rtd::mk(0.0f, 0.0f, 0.0f), d = o, id = o, oid = o;
    PlaneOffsets po{};
    float tmin = 0.0f;
    bool near = false;
    int32_t spill_code[kSpillDepth];
    float spill_t[kSpillDepth];
    int sp = 0;
    float best_t = 0.0f, best_u = 0.0f, best_v = 0.0f;
    int32_t best_id = -1;
    int32_t cur = kEmpty;
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
            if (__int_as_float(e.y) <= best_t * kHitStretch) return e.x;
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

    while (true) {
        // ---- hand new rays to idle lanes
        const uint64_t idle = __ballot(scene < 0);
        const int nidle = __popcll(idle);
        if (nidle == 64 || (nidle > 0 && nidle >= refill)) {
            if (pool >= pool_end && !drained) {
                uint32_t g = 0;
                if (lane == 0) g = atomicAdd(a.counter, a.grab);
                g = __builtin_amdgcn_readfirstlane(__shfl(g, 0, 64));
                if (g >= a.total_chunks) {
                    drained = true;
                } else {
                    pool = g * 64u;
                    pool_end = min(g + a.grab, a.total_chunks) * 64u;
                }
            }
            const uint32_t take = min(static_cast<uint32_t>(nidle), pool_end - pool);
            if (scene < 0) {
                const uint32_t rank = static_cast<uint32_t>(__popcll(idle & lt_mask));
                if (rank < take) {
                    const uint32_t s_slot = pool + rank;
                    const int sc_i = scene_of_chunk(a, s_slot >> 6);
                    const BatchScene &sc = a.scenes[sc_i];
                    const uint32_t i = ((s_slot >> 6) - sc.chunk_begin) * 64u + (s_slot & 63u);
                    if (i < sc.n_rays) {
                        scene = sc_i;
                        ri = i;
                        nodes = reinterpret_cast<const WideNode<W> *>(sc.nodes);
                        tris = sc.tris;
                        const float4 r0 = sc.rays[2 * static_cast<int64_t>(i)];
                        const float4 r1 = sc.rays[2 * static_cast<int64_t>(i) + 1];
                        o = rtd::f4v(r0);
                        d = rtd::f4v(r1);
                        tmin = r0.w;
                        id = rtd::mk(safe_inv(d.x), safe_inv(d.y), safe_inv(d.z));
                        oid = rtd::mk(-(o.x * id.x), -(o.y * id.y), -(o.z * id.z));
                        po = plane_offsets<W>(id);
                        near = fmaxf(fmaxf(fabsf(o.x), fabsf(o.y)), fabsf(o.z)) <= sc.near_limit;
                        best_t = r1.w;
                        best_u = 0.0f;
                        best_v = 0.0f;
                        best_id = -1;
                        sp = 0;
                        cur = sc.root;
                    }
                }
            }
            pool += take;
        }
        if (!__any(scene >= 0)) {
            if (drained && pool >= pool_end) break;
            continue;
        }

        // ---- interior phase (trace_ray's, per lane)
        while (true) {
            if (scene >= 0 && cur >= 0) {
                NodeView<W> v;
                load_node<W>(nodes, cur, po, v);
                float tn[W];
                if (near)
                    box_test<W, true>(v, o, id, oid, best_t * kHitStretch, tn);
                else
                    box_test<W, false>(v, o, id, oid, best_t * kHitStretch, tn);
                int32_t cc[W];
#pragma unroll
                for (int k = 0; k < W; ++k) cc[k] = v.c[k];
                sort_children<W>(tn, cc);
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
            if (scene >= 0 && cur != kEmpty && cur < 0 && leaf == kEmpty) {
                leaf = cur;
                cur = pop();
            }
            const bool searching = scene >= 0 && (cur >= 0) && (leaf == kEmpty);
            if (__popcll(__ballot(searching)) <= kSearchCut) break;
        }
        // ---- leaf phase
        while (true) {
            if (leaf != kEmpty) {
                const int32_t first = leaf_first(leaf);
                const int32_t count = leaf_count(leaf);
                for (int k = 0; k < count; ++k) {
                    const LeafTri lt = tris[first + k];
                    const int32_t pid = __float_as_int(lt.a_id.w);
                    float t, u, vv;
                    if (rtd::intersect_mt(o, d, rtd::f4v(lt.a_id), rtd::f4v(lt.e1), rtd::f4v(lt.e2), tmin, &t, &u,
                                          &vv)) {
                        if (t < best_t || (t == best_t && best_id >= 0 && pid < best_id)) {
                            best_t = t;
                            best_u = u;
                            best_v = vv;
                            best_id = pid;
                        }
                    }
                }
                leaf = kEmpty;
                if (cur != kEmpty && cur < 0) {
                    leaf = cur;
                    cur = pop();
                }
            }
            if (__popcll(__ballot(leaf != kEmpty)) <= kLeafCut) break;
        }
        // ---- finished rays write their hit and free the lane
        if (scene >= 0 && cur == kEmpty && leaf == kEmpty) {
            a.scenes[scene].hits[ri] = (best_id >= 0) ? make_float4(best_t, best_u, best_v, __int_as_float(best_id))
                                                      : make_float4(rtd::kMaxDist, 0.0f, 0.0f, __int_as_float(-1));
            scene = -1;
        }
    }
    if (overflow) atomicOr(a.overflow, 1u);
}

template <int W, int D, int G>
void launch_batch_g(int refill, int blocks, const BatchArgs &ba, hipStream_t s) {
    if (refill > 0)
        hipLaunchKernelGGL((k_trace_batch_refill<W, D, G>), dim3(blocks), dim3(64 * G), 0, s, ba, refill);
    else
        hipLaunchKernelGGL((k_trace_batch<W, D, G>), dim3(blocks), dim3(64 * G), 0, s, ba);
}

template <int W, int D>
void launch_batch_d(int group, int refill, int waves, const BatchArgs &ba, hipStream_t s) {
    const int g = (group >= 4) ? 4 : (group >= 2 ? 2 : 1);
    const int blocks = std::max(1, waves / g);
    if (g == 4)
        launch_batch_g<W, D, 4>(refill, blocks, ba, s);
    else if (g == 2)
        launch_batch_g<W, D, 2>(refill, blocks, ba, s);
    else
        launch_batch_g<W, D, 1>(refill, blocks, ba, s);
}

template <int W>
void launch_batch_w(const rt_batch_params &p, int waves, const BatchArgs &ba, hipStream_t s) {
    if (p.lds_stack <= 4)
        launch_batch_d<W, 4>(p.group_waves, p.refill, waves, ba, s);
    else if (p.lds_stack <= 8)
        launch_batch_d<W, 8>(p.group_waves, p.refill, waves, ba, s);
    else
        launch_batch_d<W, 16>(p.group_waves, p.refill, waves, ba, s);
}

void enqueue_persistent(rt_batch *b, hipStream_t s) {
    BatchArgs ba;
    ba.scenes = reinterpret_cast<const BatchScene *>(b->d_scenes);
    ba.nscenes = b->nscenes;
    ba.total_chunks = b->total_chunks;
    ba.grab = static_cast<uint32_t>(std::max(1, b->params.grab));
    ba.counter = b->d_counter;
    ba.overflow = b->d_counter + 1;
    hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(64), 0, s, b->d_counter);
    if (b->width == 4)
        launch_batch_w<4>(b->params, b->grid, ba, s);
    else
        launch_batch_w<8>(b->params, b->grid, ba, s);
}

int32_t destroy_graph(rt_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    return RT_OK;
}

void free_batch(rt_batch *b) {
    if (!b) return;
    destroy_graph(b);
    for (rt_accel *a : b->owned) rt_accel_destroy(a);
    b->owned.clear();
    if (b->d_scenes) (void)hipFree(b->d_scenes);
    if (b->d_counter) (void)hipFree(b->d_counter);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    delete b;
}

int32_t create_batch(rt_ctx *ctx, const rt_batch_item *items, int32_t nitems, const rt_batch_params &p,
                     rt_batch **out) {
    *out = nullptr;
    if (nitems <= 0 || !items) return fail(ctx, RT_ERR_INVALID, "rt_batch_create: no items");
    if (!items[0].accel) return fail(ctx, RT_ERR_INVALID, "rt_batch_create: null accel");
    const int item_width = items[0].accel->width;
    if (p.width != 0 && p.width != 4 && p.width != 8)
        return fail(ctx, RT_ERR_INVALID, "rt_batch_create: width 0, 4 or 8");
    if (p.group_waves < 0 || p.group_waves > 4 || p.group_waves == 3 || p.refill < 0 || p.refill > 64)
        return fail(ctx, RT_ERR_INVALID, "rt_batch_create: group_waves 0, 1, 2 or 4 and refill 0..64");
    const int width = (p.width == 0) ? item_width : p.width;
    std::vector<rt_accel *> owned, traced(nitems, nullptr);
    auto drop_owned = [&]() {
        for (rt_accel *a : owned) rt_accel_destroy(a);
    };
    std::vector<BatchScene> host(nitems);
    uint64_t chunks = 0;
    for (int k = 0; k < nitems; ++k) {
        const rt_accel *ia = items[k].accel;
        if (!ia || ia->width != item_width) {
            drop_owned();
            return fail(ctx, RT_ERR_INVALID, "rt_batch_create: items must share a width");
        }
        if (width != item_width) {
            // The same scene at the other node width, with the item's leaf size and SAH bins.
            rt_build_params bp{width, ia->leaf_size, ia->sah_bins, ia->lds_stack};
            rt_accel *na = nullptr;
            const int32_t rc = rt_accel_build(ia->scene, &bp, &na);
            if (rc != RT_OK) {
                drop_owned();
                return rc;
            }
            owned.push_back(na);
            traced[k] = na;
        } else {
            traced[k] = items[k].accel;
        }
        const rt_accel *a = traced[k];
        int32_t bad = RT_OK;
        if (a->scene->ctx->device != ctx->device)
            bad = fail(ctx, RT_ERR_INVALID, "rt_batch_create: items must live on the context's device");
        else if (items[k].n > 0 && (!items[k].rays || !items[k].hits))
            bad = fail(ctx, RT_ERR_INVALID, "rt_batch_create: null ray or hit buffer");
        else if (a->root == kEmpty || a->root < 0)
            bad = fail(ctx, RT_ERR_UNSUPPORTED, "rt_batch_create: scenes need at least two triangles");
        else if (items[k].n < 0 || items[k].n >= (int64_t(1) << 32))
            bad = fail(ctx, RT_ERR_INVALID, "bad ray count");
        if (bad != RT_OK) {
            drop_owned();
            return bad;
        }
        host[k].nodes = a->d_nodes;
        host[k].tris = a->d_tris;
        host[k].rays = reinterpret_cast<const float4 *>(items[k].rays);
        host[k].hits = reinterpret_cast<float4 *>(items[k].hits);
        host[k].root = a->root;
        host[k].near_limit = a->near_limit;
        host[k].chunk_begin = static_cast<uint32_t>(chunks);
        host[k].n_rays = static_cast<uint32_t>(items[k].n);
        chunks += (items[k].n + 63) / 64;
    }
    if (chunks >= (uint64_t(1) << 25)) {  // the refill form numbers 64 ray slots per chunk in 32 bits
        drop_owned();
        return fail(ctx, RT_ERR_INVALID, "rt_batch_create: too many rays");
    }
    rt_batch *b = new rt_batch();
    b->ctx = ctx;
    b->width = width;
    b->owned = owned;
    b->traced = traced;
    b->params = p;
    b->params.lds_stack = (p.lds_stack <= 0) ? 8 : p.lds_stack;
    b->params.waves_per_cu = (p.waves_per_cu <= 0) ? 16 : p.waves_per_cu;
    b->params.grab = (p.grab <= 0) ? 1 : p.grab;
    b->params.group_waves = (p.group_waves <= 0) ? 1 : p.group_waves;
    b->params.width = width;
    b->items.assign(items, items + nitems);
    b->nscenes = nitems;
    b->total_chunks = static_cast<uint32_t>(chunks);
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess || cus <= 0)
        cus = 256;
    b->grid = cus * b->params.waves_per_cu;
    hipError_t e = hipMalloc(&b->d_scenes, nitems * sizeof(BatchScene));
    if (e == hipSuccess) e = hipMalloc(&b->d_counter, 2 * sizeof(uint32_t));
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->d_scenes, host.data(), nitems * sizeof(BatchScene), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipEventCreate(&b->ev0);
    if (e == hipSuccess) e = hipEventCreate(&b->ev1);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        free_batch(b);
        return fail(ctx, e == hipErrorOutOfMemory ? RT_ERR_OOM : RT_ERR_HIP,
                    std::string("rt_batch_create: ") + hipGetErrorString(e));
    }
    if (b->params.use_graph) {
        // Capture reset + persistent launch once; rt_batch_run replays it.
        hipStream_t s = ctx->stream;
        e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            enqueue_persistent(b, s);
            e = hipStreamEndCapture(s, &b->graph);
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            free_batch(b);
            return fail(ctx, RT_ERR_HIP, std::string("rt_batch_create: graph capture: ") + hipGetErrorString(e));
        }
    }
    *out = b;
    return RT_OK;
}

int32_t run_batch(rt_batch *b, int32_t mode, float *ms) {
    rt_ctx *ctx = b->ctx;
    hipStream_t s = ctx->stream;
    RT_HIP_CHECK(ctx, hipEventRecord(b->ev0, s));
    if (mode == 0) {
        if (b->exec)
            RT_HIP_CHECK(ctx, hipGraphLaunch(b->exec, s));
        else
            enqueue_persistent(b, s);
    } else {
        for (size_t k = 0; k < b->items.size(); ++k) {
            const rt_batch_item &it = b->items[k];
            int32_t rc = launch_closest(b->traced[k], it.rays, it.n, it.hits, s, 4);
            if (rc) return rc;
        }
    }
    RT_HIP_CHECK(ctx, hipGetLastError());
    RT_HIP_CHECK(ctx, hipEventRecord(b->ev1, s));
    RT_HIP_CHECK(ctx, hipEventSynchronize(b->ev1));
    if (ms) RT_HIP_CHECK(ctx, hipEventElapsedTime(ms, b->ev0, b->ev1));
    if (mode == 0) {
        uint32_t flag = 0;
        RT_HIP_CHECK(ctx, hipMemcpy(&flag, b->d_counter + 1, sizeof(flag), hipMemcpyDeviceToHost));
        if (flag) return fail(ctx, RT_ERR_INVALID, "rt_batch_run: traversal stack overflow");
    }
    return RT_OK;
}

}  // namespace
}  // namespace rt

using rt::fail;

extern "C" {

int32_t rt_batch_create(rt_ctx *ctx, const rt_batch_item *items, int32_t nitems, const rt_batch_params *params,
                        rt_batch **out) {
    if (!ctx || !out) return RT_ERR_INVALID;
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, RT_ERR_HIP, "hipSetDevice");
    rt_batch_params p{8, 16, 1, 1, 1, 0, 0};
    if (params) p = *params;
    const int32_t rc = rt::create_batch(ctx, items, nitems, p, out);
    if (rc == RT_OK) {
        ctx->refs.fetch_add(1);
        for (const rt_batch_item &it : (*out)->items) it.accel->refs.fetch_add(1);
    }
    return rc;
}

int32_t rt_batch_run(rt_batch *batch, int32_t mode, float *ms) {
    if (!batch || (mode != 0 && mode != 1)) return RT_ERR_INVALID;
    if (hipSetDevice(batch->ctx->device) != hipSuccess) return fail(batch->ctx, RT_ERR_HIP, "hipSetDevice");
    return rt::run_batch(batch, mode, ms);
}

void rt_batch_destroy(rt_batch *batch) {
    if (!batch) return;
    (void)hipSetDevice(batch->ctx->device);
    (void)hipStreamSynchronize(batch->ctx->stream);
    std::vector<rt_batch_item> items;
    items.swap(batch->items);
    rt_ctx *ctx = batch->ctx;
    rt::free_batch(batch);
    for (const rt_batch_item &it : items) rt::release_accel(it.accel);
    rt::release_ctx(ctx);
}

int32_t rt_batch_autotune(rt_ctx *ctx, const rt_batch_item *items, int32_t nitems, int32_t reps,
                          rt_batch_params *best, float *best_ms) {
    if (!ctx || !best || nitems <= 0 || !items || !items[0].accel) return RT_ERR_INVALID;
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, RT_ERR_HIP, "hipSetDevice");
    reps = std::max(1, reps);
    const int own = items[0].accel->width;
    const int widths[] = {own, own == 4 ? 8 : 4};
    const int depths[] = {4, 8, 16};
    const int waves[] = {8, 16, 32};
    const int groups[] = {1, 2, 4};
    const int refills[] = {0, 16, 32};
    float best_t = 1e30f;
    rt_batch_params bp{8, 16, 1, 1, 1, 0, own};
    // One timed configuration (a batch built and freed per candidate; the
    // other width's BVHs are built once and shared through items2).
    auto time_one = [&](const rt_batch_item *its, const rt_batch_params &p, float *avg) -> int32_t {
        rt_batch *b = nullptr;
        int32_t rc = rt::create_batch(ctx, its, nitems, p, &b);
        if (rc) return rc;
        float t = 0.0f, sum = 0.0f;
        rc = rt::run_batch(b, 0, &t);  // warm-up
        for (int r = 0; r < reps && rc == RT_OK; ++r) {
            rc = rt::run_batch(b, 0, &t);
            sum += t;
        }
        rt::free_batch(b);
        *avg = sum / reps;
        return rc;
    };
    std::vector<rt_accel *> other;
    std::vector<rt_batch_item> items2(items, items + nitems);
    auto drop = [&]() {
        for (rt_accel *a : other) rt_accel_destroy(a);
    };
    for (int k = 0; k < nitems; ++k) {
        const rt_accel *ia = items[k].accel;
        if (!ia) {
            drop();
            return fail(ctx, RT_ERR_INVALID, "rt_batch_autotune: null accel");
        }
        rt_build_params p{widths[1], ia->leaf_size, ia->sah_bins, ia->lds_stack};
        rt_accel *na = nullptr;
        if (int32_t rc = rt_accel_build(ia->scene, &p, &na)) {
            drop();
            return rc;
        }
        other.push_back(na);
        items2[k].accel = na;
    }
    for (int wi = 0; wi < 2; ++wi) {
        const rt_batch_item *its = (wi == 0) ? items : items2.data();
        for (int d : depths)
            for (int w : waves)
                for (int g : groups)
                    for (int rf : refills) {
                        rt_batch_params p{d, w, 1, 0, g, rf, 0};
                        float avg = 0.0f;
                        if (int32_t rc = time_one(its, p, &avg)) {
                            drop();
                            return rc;
                        }
                        if (avg < best_t) {
                            best_t = avg;
                            bp = p;
                            bp.width = widths[wi];
                        }
                    }
    }
    // Chunks per grab for the best shape.
    const rt_batch_item *best_items = (bp.width == own) ? items : items2.data();
    for (int gr : {2, 4}) {
        rt_batch_params p = bp;
        p.grab = gr;
        p.width = 0;
        float avg = 0.0f;
        if (int32_t rc = time_one(best_items, p, &avg)) {
            drop();
            return rc;
        }
        if (avg < best_t) {
            best_t = avg;
            bp.grab = gr;
        }
    }
    drop();
    bp.use_graph = 1;
    *best = bp;
    if (best_ms) *best_ms = best_t;
    return RT_OK;
}

}  // extern "C"
