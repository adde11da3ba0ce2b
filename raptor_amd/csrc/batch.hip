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
// rt_batch_autotune times the persistent form over LDS stack depth, resident
// wavefronts per CU and chunks per grab, and returns the fastest parameters.
#include <algorithm>
#include <chrono>
#include <vector>

#include "traverse.hpp"

struct rt_batch {
    rt_ctx *ctx = nullptr;
    int width = 4;
    rt_batch_params params{};
    std::vector<rt_batch_item> items;
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

template <int W, int kLdsDepth>
__global__ void __launch_bounds__(64) k_trace_batch(BatchArgs a) {
    const int lane = threadIdx.x;
    while (true) {
        uint32_t g = 0;
        if (lane == 0) g = atomicAdd(a.counter, a.grab);
        g = __builtin_amdgcn_readfirstlane(__shfl(g, 0, 64));
        if (g >= a.total_chunks) break;
        const uint32_t g_end = min(g + a.grab, a.total_chunks);
        for (uint32_t c = g; c < g_end; ++c) {
            // Scene of chunk c: the last scene whose first chunk is <= c.
            int lo = 0, hi = a.nscenes - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (a.scenes[mid].chunk_begin <= c)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            const BatchScene sc = a.scenes[lo];
            const uint32_t i = (c - sc.chunk_begin) * 64u + static_cast<uint32_t>(lane);
            if (i < sc.n_rays) {
                const float4 r0 = sc.rays[2 * static_cast<int64_t>(i)];
                const float4 r1 = sc.rays[2 * static_cast<int64_t>(i) + 1];
                float4 hit;
                uint32_t visits = 0, tests = 0;
                bool overflow = false;
                trace_ray<W, true, false, kLdsDepth>(reinterpret_cast<const WideNode<W> *>(sc.nodes), sc.tris,
                                                     sc.root, sc.near_limit, rtd::f4v(r0), rtd::f4v(r1), r0.w, r1.w, lane,
                                                     hit, visits, tests, overflow);
                const int32_t id = __float_as_int(hit.w);
                sc.hits[i] = (id >= 0) ? hit : make_float4(rtd::kMaxDist, 0.0f, 0.0f, __int_as_float(-1));
                if (overflow) atomicOr(a.overflow, 1u);
            }
        }
    }
}

template <int W>
void launch_batch_w(int depth, dim3 grid, const BatchArgs &ba, hipStream_t s) {
    if (depth <= 4)
        hipLaunchKernelGGL((k_trace_batch<W, 4>), grid, dim3(64), 0, s, ba);
    else if (depth <= 8)
        hipLaunchKernelGGL((k_trace_batch<W, 8>), grid, dim3(64), 0, s, ba);
    else
        hipLaunchKernelGGL((k_trace_batch<W, 16>), grid, dim3(64), 0, s, ba);
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
        launch_batch_w<4>(b->params.lds_stack, dim3(b->grid), ba, s);
    else
        launch_batch_w<8>(b->params.lds_stack, dim3(b->grid), ba, s);
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
    const int width = items[0].accel->width;
    std::vector<BatchScene> host(nitems);
    uint64_t chunks = 0;
    for (int k = 0; k < nitems; ++k) {
        const rt_accel *a = items[k].accel;
        if (!a || a->width != width) return fail(ctx, RT_ERR_INVALID, "rt_batch_create: items must share a width");
        if (a->scene->ctx->device != ctx->device)
            return fail(ctx, RT_ERR_INVALID, "rt_batch_create: items must live on the context's device");
        if (items[k].n > 0 && (!items[k].rays || !items[k].hits))
            return fail(ctx, RT_ERR_INVALID, "rt_batch_create: null ray or hit buffer");
        if (a->root == kEmpty || a->root < 0)
            return fail(ctx, RT_ERR_UNSUPPORTED, "rt_batch_create: scenes need at least two triangles");
        if (items[k].n < 0 || items[k].n >= (int64_t(1) << 32)) return fail(ctx, RT_ERR_INVALID, "bad ray count");
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
    if (chunks >= (uint64_t(1) << 31)) return fail(ctx, RT_ERR_INVALID, "rt_batch_create: too many rays");
    rt_batch *b = new rt_batch();
    b->ctx = ctx;
    b->width = width;
    b->params = p;
    b->params.lds_stack = (p.lds_stack <= 0) ? 8 : p.lds_stack;
    b->params.waves_per_cu = (p.waves_per_cu <= 0) ? 16 : p.waves_per_cu;
    b->params.grab = (p.grab <= 0) ? 1 : p.grab;
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
        for (const rt_batch_item &it : b->items) {
            int32_t rc = launch_closest(it.accel, it.rays, it.n, it.hits, s, 4);
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
    rt_batch_params p{8, 16, 1, 1};
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
    if (!ctx || !best || nitems <= 0) return RT_ERR_INVALID;
    reps = std::max(1, reps);
    const int depths[] = {4, 8, 16};
    const int waves[] = {8, 16, 32};
    const int grabs[] = {1, 4};
    float best_t = 1e30f;
    rt_batch_params bp{8, 16, 1, 1};
    for (int d : depths)
        for (int w : waves)
            for (int g : grabs) {
                rt_batch_params p{d, w, g, 0};
                rt_batch *b = nullptr;
                int32_t rc = rt::create_batch(ctx, items, nitems, p, &b);
                if (rc) return rc;
                float t = 0.0f, sum = 0.0f;
                rc = rt::run_batch(b, 0, &t);  // warm-up
                for (int r = 0; r < reps && rc == RT_OK; ++r) {
                    rc = rt::run_batch(b, 0, &t);
                    sum += t;
                }
                rt::free_batch(b);
                if (rc) return rc;
                const float avg = sum / reps;
                if (avg < best_t) {
                    best_t = avg;
                    bp = p;
                }
            }
    bp.use_graph = 1;
    *best = bp;
    if (best_ms) *best_ms = best_t;
    return RT_OK;
}

}  // extern "C"
