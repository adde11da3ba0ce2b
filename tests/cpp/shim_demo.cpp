// Drives include/raptor_gpu.hpp with stand-in types that expose the same
// public accessors as the reference's primitive_store / triangle / camera
// (rt/primitive_store.h:17-110, rt/triangle.h:40-49, rt/camera.h:281-285),
// and checks the shim's frame against a direct rt_render call.
// Usage: shim_demo            (needs a GPU; prints "shim ok" on success)
#include <cmath>
#include <cstdio>
#include <vector>

#include "raptor_gpu.hpp"

struct point {
    float x, y, z;
};

struct triangle {
    point a, b, c;
    point get_vertex_a() const { return a; }
    point get_vertex_b() const { return b; }
    point get_vertex_c() const { return c; }
    bool get_light() const { return false; }
    bool is_transparent() const { return false; }
};

struct primitive_store {
    std::vector<triangle> tris;
    int size() const { return static_cast<int>(tris.size()); }
    const triangle *primitive(int i) const { return &tris[i]; }
};

struct colour {
    float r, g, b;
    colour(float r_, float g_, float b_) : r(r_), g(g_), b(b_) {}
};

struct camera {
    using colour_type = colour;
    int x_res;
    std::vector<colour> image;
    void set_pixel(const colour &p, int x, int y) { image[x + y * x_res] = p; }
};

int main() {
    const int64_t n = 4096;
    std::vector<float> abc(9 * n), alb(3 * n);
    rt_gen_random_triangles(1, n, 0.02f, 0.0f, 1.0f, abc.data(), alb.data());
    primitive_store prims;
    for (int64_t i = 0; i < n; ++i) {
        const float *t = &abc[9 * i];
        prims.tris.push_back({{t[0], t[1], t[2]}, {t[3], t[4], t[5]}, {t[6], t[7], t[8]}});
    }
    raptor_gpu::context ctx(0);
    raptor_gpu::gpu_bvh<primitive_store> accel(ctx, prims, [&](int i, float *rgb) {
        for (int k = 0; k < 3; ++k) rgb[k] = alb[3 * i + k];
    });
    rt_camera cam{{0.5f, 0.5f, -1.5f}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, 0.35f, 0.35f, 1.0f, 64, 64};
    std::vector<rt_light> lights{{{0.5f, 2.0f, 0.5f}, {255.0f, 255.0f, 255.0f}, 0.0f}};
    camera c{64, std::vector<colour>(64 * 64, colour(-1, -1, -1))};
    raptor_gpu::ray_tracer(accel, lights, cam, c);

    std::vector<float> ref(3 * 64 * 64);
    if (rt_render(accel.get(), &cam, lights.data(), 1, nullptr, ref.data(), nullptr, nullptr, nullptr) != RT_OK) {
        std::printf("rt_render failed\n");
        return 1;
    }
    int bad = 0, lit = 0;
    for (int p = 0; p < 64 * 64; ++p) {
        if (c.image[p].r != ref[3 * p] || c.image[p].g != ref[3 * p + 1] || c.image[p].b != ref[3 * p + 2]) ++bad;
        if (ref[3 * p] > 0.0f) ++lit;
    }
    // One-ray legacy forms against the batched call.
    struct ray {
        point o, d;
        point get_ogn() const { return o; }
        point get_dir() const { return d; }
    };
    struct hit {
        float u, v, d;
    };
    int id_bad = 0;
    for (int k = 0; k < 32; ++k) {
        const float x = 0.2f + 0.02f * k;
        const float len = std::sqrt(x * x + 1.0f);
        ray r{{x, 0.5f, -1.0f}, {0.0f, 0.0f, 1.0f}};
        (void)len;
        hit h;
        const int id = accel.find_nearest_object(&r, &h);
        rt_ray q{{x, 0.5f, -1.0f}, 1e-3f, {0.0f, 0.0f, 1.0f}, 3.402823466e+38f};
        rt_hit out;
        accel.trace_closest(&q, 1, &out);
        if (id != out.prim || h.d != out.t) ++id_bad;
        if (accel.found_nearer_object(&r, 0.01f)) ++id_bad;  // nothing within 0.01 of z = -1
    }
    // The physics loop: move every triangle, refit, and compare with a new
    // accelerator over the moved store.
    for (auto &t : prims.tris)
        for (point *v : {&t.a, &t.b, &t.c}) {
            v->x += 0.05f * std::sin(7.0f * v->y);
            v->z += 0.03f;
        }
    accel.refit(prims);
    raptor_gpu::gpu_bvh<primitive_store> fresh(ctx, prims, [&](int i, float *rgb) {
        for (int k = 0; k < 3; ++k) rgb[k] = alb[3 * i + k];
    });
    camera c1{64, std::vector<colour>(64 * 64, colour(-1, -1, -1))};
    camera c2{64, std::vector<colour>(64 * 64, colour(-1, -1, -1))};
    raptor_gpu::ray_tracer(accel, lights, cam, c1);
    raptor_gpu::ray_tracer(fresh, lights, cam, c2);
    int moved_bad = 0, changed = 0;
    for (int p = 0; p < 64 * 64; ++p) {
        if (c1.image[p].r != c2.image[p].r || c1.image[p].g != c2.image[p].g || c1.image[p].b != c2.image[p].b)
            ++moved_bad;
        if (c1.image[p].r != c.image[p].r) ++changed;
    }
    // Materials through the shim: every other triangle a shiny mirror,
    // against the same table set directly on a second scene.
    const rt_material shiny{{2.0f, 2.0f, 2.0f}, {120.0f, 90.0f, 60.0f}, {80.0f, 80.0f, 80.0f}, 16.0f, 0.0f, 1.0f, 0.5f};
    accel.set_materials([&](int i) { return (i % 2) ? &shiny : nullptr; });
    camera c3{64, std::vector<colour>(64 * 64, colour(-1, -1, -1))};
    raptor_gpu::ray_tracer(accel, lights, cam, c3);
    std::vector<int32_t> idx(n);
    for (int64_t i = 0; i < n; ++i) idx[i] = (i % 2) ? 0 : -1;
    if (rt_scene_set_materials(fresh.scene(), &shiny, 1, idx.data()) != RT_OK) {
        std::printf("rt_scene_set_materials failed\n");
        return 1;
    }
    std::vector<float> ref3(3 * 64 * 64);
    rt_stats st{};
    if (rt_render(fresh.get(), &cam, lights.data(), 1, nullptr, ref3.data(), nullptr, nullptr, &st) != RT_OK) {
        std::printf("rt_render with materials failed\n");
        return 1;
    }
    int mat_bad = 0;
    for (int p = 0; p < 64 * 64; ++p)
        if (c3.image[p].r != ref3[3 * p] || c3.image[p].g != ref3[3 * p + 1] || c3.image[p].b != ref3[3 * p + 2])
            ++mat_bad;
    // A model file: an OFF pentagon-and-quad in front of the camera, on a
    // white background.
    const char *off_path = "/tmp/raptor_shim_demo.off";
    if (FILE *f = std::fopen(off_path, "w")) {
        std::fputs("OFF\n9 2 0\n0.3 0.3 0.5\n0.7 0.3 0.5\n0.8 0.6 0.5\n0.5 0.8 0.5\n0.2 0.6 0.5\n"
                   "0 0 0.9\n1 0 0.9\n1 1 0.9\n0 1 0.9\n5 0 1 2 3 4\n4 5 6 7 8\n", f);
        std::fclose(f);
    }
    raptor_gpu::mesh_scene model(ctx, off_path);
    rt_render_params bgp{};
    bgp.spp = 1;
    bgp.background[0] = bgp.background[1] = bgp.background[2] = 255.0f;
    std::vector<float> img(3 * 64 * 64);
    std::vector<rt_hit> hits(64 * 64);
    if (rt_render(model.get(), &cam, lights.data(), 1, &bgp, img.data(), hits.data(), nullptr, nullptr) != RT_OK) {
        std::printf("rt_render of the model failed\n");
        return 1;
    }
    int front = 0, back = 0, white = 0;
    for (int p = 0; p < 64 * 64; ++p) {
        if (hits[p].prim >= 0 && hits[p].prim < 3) ++front;  // the pentagon: 3 triangles, emplaced first
        if (hits[p].prim >= 3) ++back;                      // the quad: 2 triangles
        if (hits[p].prim < 0 && img[3 * p] == 255.0f) ++white;
    }
    const bool model_ok = model.triangles() == 5 && front > 100 && back > 100 && white > 0;
    std::printf("%s: %lld triangles, %d pentagon, %d quad, %d background pixels\n",
                model_ok ? "model ok" : "model FAILED", static_cast<long long>(model.triangles()), front, back,
                white);
    const bool ok = bad == 0 && id_bad == 0 && lit > 100 && moved_bad == 0 && changed > 100 && mat_bad == 0 &&
                    st.rays[2] > 0 && model_ok;
    std::printf("%s: %d pixel mismatches, %d lit pixels, %d ray mismatches, %d refit/rebuild mismatches, "
                "%d pixels changed by the move, %d material mismatches, %llu mirror rays\n",
                ok ? "shim ok" : "shim FAILED", bad, lit, id_bad, moved_bad, changed, mat_bad,
                static_cast<unsigned long long>(st.rays[2]));
    return ok ? 0 : 1;
}
