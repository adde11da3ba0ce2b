// Oracle — TEST INFRASTRUCTURE ONLY (see vec.hpp for the rules).
//
// extern "C" wrappers for ctypes.
#include <atomic>
#include <thread>
#include <vector>

#include "oracle.hpp"
#include "packet.hpp"
#include "rt_oracle.h"
#include "sampling.hpp"

namespace oracle {
int render_frame(const Scene &s, const float *albedo, const oracle_camera &cam, const oracle_light *lights,
                 int nlights, const oracle_render_params &p, const oracle_render_out &out);
struct CameraState;
}  // namespace oracle

struct oracle_scene {
    oracle::Scene s;
};

using namespace oracle;

namespace {

template <class F>
void parallel_for(int64_t n, int nthreads, F &&f) {
    nthreads = std::max(1, nthreads);
    std::atomic<int64_t> next{0};
    auto worker = [&](int tid) {
        const int64_t chunk = 256;
        while (true) {
            const int64_t b = next.fetch_add(chunk);
            if (b >= n) break;
            const int64_t e = std::min(n, b + chunk);
            for (int64_t i = b; i < e; ++i) f(tid, i);
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nthreads; ++t) pool.emplace_back(worker, t);
    worker(0);
    for (auto &th : pool) th.join();
}

Ray ray_from(const float *r8) {
    Ray r;
    r.ogn = V3(r8[0], r8[1], r8[2]);
    r.dir = V3(r8[4], r8[5], r8[6]);
    r.length = r8[7];
    return r;
}

}  // namespace

extern "C" {

oracle_scene *oracle_scene_create(const float *abc, int64_t n, const uint8_t *flags) {
    auto *o = new oracle_scene();
    o->s.tris.reserve(n);
    for (int64_t i = 0; i < n; ++i) {
        const float *t = abc + 9 * i;
        const uint8_t f = flags ? flags[i] : 0;
        o->s.tris.push_back(make_triangle(V3(t[0], t[1], t[2]), V3(t[3], t[4], t[5]), V3(t[6], t[7], t[8]),
                                          (f & 1) != 0, (f & 2) != 0));
    }
    o->s.indirect.resize(n);
    for (int64_t i = 0; i < n; ++i) o->s.indirect[i] = static_cast<int>(i);
    return o;
}

void oracle_scene_destroy(oracle_scene *s) { delete s; }

int32_t oracle_scene_build(oracle_scene *s, float alpha, float delta) {
    AacParams p;
    p.alpha = alpha;
    p.delta = delta;
    return build_aac(&s->s, p);
}

int32_t oracle_scene_node_count(const oracle_scene *s) { return s->s.root + 1; }

int32_t oracle_scene_tree(const oracle_scene *so, float *lohi, int32_t *child, uint8_t *leaf, int32_t *indirect) {
    const Scene &s = so->s;
    const int count = s.root + 1;
    for (int i = 0; i < count; ++i) {
        const BvhNode &nd = s.nodes[i];
        const float v[6] = {nd.low.x, nd.low.y, nd.low.z, nd.high.x, nd.high.y, nd.high.z};
        for (int k = 0; k < 6; ++k) lohi[6 * i + k] = v[k];
        child[2 * i] = nd.left_b;
        child[2 * i + 1] = nd.right_e;
        leaf[i] = nd.leaf ? 1 : 0;
    }
    for (size_t i = 0; i < s.indirect.size(); ++i) indirect[i] = s.indirect[i];
    return count;
}

void oracle_aac_morton(const oracle_scene *so, int32_t *codes, int32_t *order) {
    std::vector<int> c, o;
    aac_morton(so->s, &c, &o);
    for (size_t i = 0; i < c.size(); ++i) {
        codes[i] = c[i];
        order[i] = o[i];
    }
}

int32_t oracle_bvh_check(const oracle_scene *so, int32_t *stats) {
    const Scene &s = so->s;
    const int n = static_cast<int>(s.tris.size());
    if (n == 0) return s.root == -1 ? 0 : 1;
    std::vector<int> seen(n, 0);
    int leaves = 0, interior = 0, max_depth = 0, max_leaf = 0;
    std::vector<std::pair<int, int>> stack{{s.root, 1}};
    int err = 0;
    while (!stack.empty()) {
        auto [idx, depth] = stack.back();
        stack.pop_back();
        max_depth = std::max(max_depth, depth);
        const BvhNode &nd = s.nodes[idx];
        if (nd.leaf) {
            ++leaves;
            max_leaf = std::max(max_leaf, nd.right_e - nd.left_b);
            for (int i = nd.left_b; i < nd.right_e; ++i) {
                const Triangle &t = s.tris[s.indirect[i]];
                ++seen[s.indirect[i]];
                const V3 lo = vmin(t.a, vmin(t.b, t.c)), hi = vmax(t.a, vmax(t.b, t.c));
                if (lo.x < nd.low.x || lo.y < nd.low.y || lo.z < nd.low.z) err |= 2;
                if (hi.x > nd.high.x || hi.y > nd.high.y || hi.z > nd.high.z) err |= 2;
            }
        } else {
            ++interior;
            const BvhNode &l = s.nodes[nd.left_b], &r = s.nodes[nd.right_e];
            const V3 lo = vmin(l.low, r.low), hi = vmax(l.high, r.high);
            if (lo.x != nd.low.x || lo.y != nd.low.y || lo.z != nd.low.z) err |= 4;
            if (hi.x != nd.high.x || hi.y != nd.high.y || hi.z != nd.high.z) err |= 4;
            stack.push_back({nd.left_b, depth + 1});
            stack.push_back({nd.right_e, depth + 1});
        }
    }
    for (int i = 0; i < n; ++i)
        if (seen[i] != 1) err |= 1;
    const BvhNode &root = s.nodes[s.root];
    if (root.low.x != s.scene_lo.x || root.low.y != s.scene_lo.y || root.low.z != s.scene_lo.z) err |= 8;
    if (root.high.x != s.scene_hi.x || root.high.y != s.scene_hi.y || root.high.z != s.scene_hi.z) err |= 8;
    if (interior != leaves - 1) err |= 16;
    if (stats) {
        stats[0] = leaves;
        stats[1] = interior;
        stats[2] = max_depth;
        stats[3] = max_leaf;
    }
    return err;
}

void oracle_scene_normals(const oracle_scene *s, float *n3) {
    for (size_t i = 0; i < s->s.tris.size(); ++i) {
        n3[3 * i + 0] = s->s.tris[i].n.x;
        n3[3 * i + 1] = s->s.tris[i].n.y;
        n3[3 * i + 2] = s->s.tris[i].n.z;
    }
}

void oracle_trace_closest(const oracle_scene *s, const float *rays, int64_t n, int32_t *prim, float *tuv,
                          int32_t use_bvh, int32_t nthreads, uint64_t *counters) {
    std::vector<Counters> cs(std::max(1, nthreads));
    parallel_for(n, nthreads, [&](int tid, int64_t i) {
        const Ray r = ray_from(rays + 8 * i);
        Hit h;
        const int id = use_bvh ? find_nearest_object(s->s, r, &h, &cs[tid]) : brute_nearest(s->s, r, &h);
        if (prim) prim[i] = id;
        if (tuv) {
            tuv[3 * i + 0] = h.d;
            tuv[3 * i + 1] = h.u;
            tuv[3 * i + 2] = h.v;
        }
    });
    if (counters) {
        counters[0] = counters[1] = 0;
        for (auto &c : cs) {
            counters[0] += c.box;
            counters[1] += c.tri;
        }
    }
}

void oracle_trace_occluded(const oracle_scene *s, const float *rays, int64_t n, uint8_t *occluded, int32_t use_bvh,
                           int32_t nthreads, uint64_t *counters) {
    std::vector<Counters> cs(std::max(1, nthreads));
    parallel_for(n, nthreads, [&](int tid, int64_t i) {
        const Ray r = ray_from(rays + 8 * i);
        const bool occ = use_bvh ? found_nearer_object(s->s, r, r.length, &cs[tid]) : brute_nearer(s->s, r, r.length);
        occluded[i] = occ ? 1 : 0;
    });
    if (counters) {
        counters[0] = counters[1] = 0;
        for (auto &c : cs) {
            counters[0] += c.box;
            counters[1] += c.tri;
        }
    }
}

int32_t oracle_render(const oracle_scene *s, const float *albedo, const oracle_camera *cam,
                      const oracle_light *lights, int32_t nlights, const oracle_render_params *p,
                      oracle_render_out *out) {
    // The packet frame: one sample per pixel, at most one bounce (the
    // reference's packet camera has no jitter); pixel_stride strides tiles.
    if (p->packet && (p->spp > 1 || p->bounce > 1)) return -1;
    if (nlights < 0 || nlights > 8) return -2;
    if ((p->materials || p->vertex_normals) && (p->bounce != 0 || p->packet || out->rec_levels > 0)) return -3;
    return render_frame(s->s, albedo, *cam, lights, nlights, *p, *out);
}

int32_t oracle_min_element_n(const float *d, float *m, int32_t n) { return min_element_n(d, m, n); }
int32_t oracle_min_element_range(const float *d, float *m, int32_t begin, int32_t end) {
    return min_element_range(d, m, begin, end);
}
int32_t oracle_morton(float x, float y, float z, float xm, float ym, float zm) {
    return morton_code(x, y, z, xm, ym, zm);
}
void oracle_morton_sort(const int32_t *codes, int64_t n, int32_t *sorted, int32_t *order) {
    std::vector<int> c(codes, codes + n), sc, o;
    morton_radix_sort(c, &sc, &o);
    for (int64_t i = 0; i < n; ++i) {
        sorted[i] = sc[i];
        order[i] = o[i];
    }
}
float oracle_slab(const float *low, const float *high, const float *ogn, const float *i_rd) {
    BvhNode nd;
    nd.low = V3(low[0], low[1], low[2]);
    nd.high = V3(high[0], high[1], high[2]);
    return slab_distance(nd, V3(ogn[0], ogn[1], ogn[2]), V3(i_rd[0], i_rd[1], i_rd[2]));
}
float oracle_combined_sa(const float *lo1, const float *hi1, const float *lo2, const float *hi2) {
    BvhNode a, b;
    a.low = V3(lo1[0], lo1[1], lo1[2]);
    a.high = V3(hi1[0], hi1[1], hi1[2]);
    b.low = V3(lo2[0], lo2[1], lo2[2]);
    b.high = V3(hi2[0], hi2[1], hi2[2]);
    return combined_surface_area(a, b);
}
int32_t oracle_intersect(const float *abc, const float *ogn, const float *dir, float *tuv) {
    const Triangle t = make_triangle(V3(abc[0], abc[1], abc[2]), V3(abc[3], abc[4], abc[5]),
                                     V3(abc[6], abc[7], abc[8]), false, false);
    Ray r;
    r.ogn = V3(ogn[0], ogn[1], ogn[2]);
    r.dir = V3(dir[0], dir[1], dir[2]);
    Hit h;
    intersect_scalar(t, r, &h);
    tuv[0] = h.d;
    tuv[1] = h.u;
    tuv[2] = h.v;
    return h.d < kMaxDist ? 1 : 0;
}
void oracle_offset_start_point(const float *dst, const float *geo_norm, int32_t d, float *out) {
    const V3 o = offset_start_point(V3(dst[0], dst[1], dst[2]), V3(geo_norm[0], geo_norm[1], geo_norm[2]), d);
    out[0] = o.x;
    out[1] = o.y;
    out[2] = o.z;
}
void oracle_bounce_dir(uint64_t seed, uint32_t pixel, uint32_t sample, uint32_t bounce, const float *n, float *dir) {
    const V3 d = cosine_direction(bounce_key(seed, pixel, sample, bounce), V3(n[0], n[1], n[2]));
    dir[0] = d.x;
    dir[1] = d.y;
    dir[2] = d.z;
}
void oracle_sincos_turns(float r, float *s, float *c) { sincos_turns(r, s, c); }
uint64_t oracle_pcg32_stream(uint64_t initstate, uint64_t initseq, int32_t count, uint32_t *out) {
    Pcg32 g;
    g.seed(initstate, initseq);
    for (int i = 0; i < count; ++i) out[i] = g.next();
    return g.state;
}
uint64_t oracle_splitmix64_stream(uint64_t seed, int32_t count, uint64_t *out) {
    uint64_t s = seed;
    for (int i = 0; i < count; ++i) out[i] = splitmix64_next(&s);
    return s;
}
void oracle_primary_ray(const oracle_camera *cam, int32_t x, int32_t y, float *ogn, float *dir) {
    // Same arithmetic as render.cpp primary_ray (rt/camera.h:127-145).
    const float x_m = -cam->w, y_m = -cam->h;
    const float x_inc = (cam->w * 2.0f) / static_cast<float>(cam->x_res);
    const float y_inc = (cam->h * 2.0f) / static_cast<float>(cam->y_res);
    const float x_t = (static_cast<float>(x) * x_inc) + x_m;
    const float y_t = (static_cast<float>(y) * y_inc) + y_m;
    const V3 X(cam->x_axis[0], cam->x_axis[1], cam->x_axis[2]);
    const V3 Y(cam->y_axis[0], cam->y_axis[1], cam->y_axis[2]);
    const V3 Z(cam->z_axis[0], cam->z_axis[1], cam->z_axis[2]);
    const V3 d = normalise_exact(((X * x_t) + (Y * y_t)) + (Z * cam->t));
    for (int k = 0; k < 3; ++k) {
        ogn[k] = cam->pos[k];
        dir[k] = d[k];
    }
}

}  // extern "C"

extern "C" {

// Packet slab test of one box for four rays: ogn4 / i_rd4 are x[4], y[4], z[4].
void oracle_packet_slab(const float *low, const float *high, const float *ogn4, const float *i_rd4, float *out4) {
    oracle::BvhNode n;
    n.low = oracle::V3(low[0], low[1], low[2]);
    n.high = oracle::V3(high[0], high[1], high[2]);
    oracle::PacketRay r;
    r.ox = _mm_loadu_ps(ogn4);
    r.oy = _mm_loadu_ps(ogn4 + 4);
    r.oz = _mm_loadu_ps(ogn4 + 8);
    r.dx = r.dy = r.dz = _mm_setzero_ps();
    const __m128 i_rd[3] = {_mm_loadu_ps(i_rd4), _mm_loadu_ps(i_rd4 + 4), _mm_loadu_ps(i_rd4 + 8)};
    _mm_storeu_ps(out4, oracle::packet_slab(n, r, i_rd));
}

void oracle_packet_inverse(const float *x4, float *out4) { _mm_storeu_ps(out4, oracle::packet_inverse(_mm_loadu_ps(x4))); }
void oracle_packet_inverse_sqrt(const float *x4, float *out4) {
    _mm_storeu_ps(out4, oracle::packet_inverse_sqrt(_mm_loadu_ps(x4)));
}

static oracle::PacketRay load_packet(const float *rays) {
    // four consecutive rays of 8 floats: o.xyz, tmin, d.xyz, tmax
    oracle::PacketRay p;
    p.ox = _mm_setr_ps(rays[0], rays[8], rays[16], rays[24]);
    p.oy = _mm_setr_ps(rays[1], rays[9], rays[17], rays[25]);
    p.oz = _mm_setr_ps(rays[2], rays[10], rays[18], rays[26]);
    p.dx = _mm_setr_ps(rays[4], rays[12], rays[20], rays[28]);
    p.dy = _mm_setr_ps(rays[5], rays[13], rays[21], rays[29]);
    p.dz = _mm_setr_ps(rays[6], rays[14], rays[22], rays[30]);
    return p;
}

// Packets of four consecutive rays (n a multiple of 4) through the packet
// traversal (use_bvh) or packet brute force.
void oracle_packet_trace_closest(const oracle_scene *s, const float *rays, int64_t n, int32_t *prim, float *tuv,
                                 int32_t use_bvh) {
    for (int64_t i = 0; i + 4 <= n; i += 4) {
        const oracle::PacketRay p = load_packet(rays + 8 * i);
        oracle::PacketHit h;
        if (use_bvh)
            oracle::packet_nearest(s->s, p, &h, nullptr);
        else
            oracle::packet_brute_nearest(s->s, p, &h);
        alignas(16) float d[4], u[4], v[4];
        alignas(16) int id[4];
        _mm_store_ps(d, h.d);
        _mm_store_ps(u, h.u);
        _mm_store_ps(v, h.v);
        _mm_store_si128(reinterpret_cast<__m128i *>(id), h.id);
        for (int k = 0; k < 4; ++k) {
            prim[i + k] = d[k] < oracle::kMaxDist ? id[k] : -1;
            tuv[3 * (i + k) + 0] = d[k];
            tuv[3 * (i + k) + 1] = u[k];
            tuv[3 * (i + k) + 2] = v[k];
        }
    }
}

void oracle_packet_trace_occluded(const oracle_scene *s, const float *rays, int64_t n, uint8_t *occ, int32_t use_bvh) {
    for (int64_t i = 0; i + 4 <= n; i += 4) {
        const oracle::PacketRay p = load_packet(rays + 8 * i);
        const float *r = rays + 8 * i;
        const __m128 t = _mm_setr_ps(r[7], r[15], r[23], r[31]);
        const __m128 c = use_bvh ? oracle::packet_nearer(s->s, p, t, nullptr) : oracle::packet_brute_nearer(s->s, p, t);
        const int m = _mm_movemask_ps(c);
        for (int k = 0; k < 4; ++k) occ[i + k] = (m >> k) & 1;
    }
}

}  // extern "C"
