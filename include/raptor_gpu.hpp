// raptor_gpu.hpp — header-only C++ shim over the C ABI of libraptor_rt, shaped
// like the reference's own entry points so that a RAPTor caller can switch to
// the GPU path with few changes:
//
//   reference                                     this shim
//   bvh accel(prims);                             raptor_gpu::gpu_bvh<primitive_store> accel(ctx, prims, albedo);
//   accel.find_nearest_object(&r, &h)             accel.trace_closest(rays, n, hits)   (batched)
//   accel.found_nearer_object(&r, t)              accel.trace_occluded(rays, n, flags) (batched)
//   ray_tracer(&accel, lights, prims, cam);       raptor_gpu::ray_tracer(accel, lights, cam_params, cam);
//   new bih(*tris) every frame                    accel.refit(*tris)   (same triangles, moved)
//
// (ssd/bvh.h:27-32, ssd/bvh.cc:649-759, rt/raytracer.h:105.) The store is read
// through the public triangle accessors get_vertex_a/b/c(), get_light() and
// is_transparent() (rt/triangle.h:40-49); the caller's indirection is not
// touched (the reference builder permutes it, ssd/bvh_builder.cc:146-158).
// Errors throw raptor_gpu::error carrying rt_last_error().
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "raptor_rt.h"

namespace raptor_gpu {

class error : public std::runtime_error {
  public:
    explicit error(const std::string &what) : std::runtime_error(what) {}
};

class context {
  public:
    explicit context(int device = 0) {
        if (rt_ctx_create(device, &ctx_) != RT_OK) throw error("rt_ctx_create failed");
    }
    ~context() { rt_ctx_destroy(ctx_); }
    context(const context &) = delete;
    context &operator=(const context &) = delete;
    rt_ctx *get() const { return ctx_; }
    void check(int32_t rc, const char *what) const {
        if (rc != RT_OK) throw error(std::string(what) + ": " + rt_last_error(ctx_));
    }

  private:
    rt_ctx *ctx_ = nullptr;
};

// A model file as the reference's regression fixture loads it
// (rt/regression_tests/regression_fixture.h:86-168: off_parser, ply_parser or
// obj_parser by extension): the parsed triangles, materials and vertex
// normals as a scene, and its accelerator.
class mesh_scene {
  public:
    mesh_scene(const context &ctx, const std::string &path, const rt_build_params *params = nullptr) {
        rt_mesh *mesh = nullptr;
        if (rt_mesh_load(path.c_str(), -1, &mesh) != RT_OK) throw error(rt_mesh_last_error());
        rt_mesh_info info{};
        rt_mesh_info_get(mesh, &info);
        triangles_ = info.triangles;
        const int32_t rc = rt_scene_create_from_mesh(ctx.get(), mesh, &scene_);
        rt_mesh_destroy(mesh);
        ctx.check(rc, "rt_scene_create_from_mesh");
        if (rt_accel_build(scene_, params, &accel_) != RT_OK) {
            const std::string msg = std::string("rt_accel_build: ") + rt_last_error(ctx.get());
            rt_scene_destroy(scene_);
            throw error(msg);
        }
    }
    ~mesh_scene() {
        rt_accel_destroy(accel_);
        rt_scene_destroy(scene_);
    }
    mesh_scene(const mesh_scene &) = delete;
    mesh_scene &operator=(const mesh_scene &) = delete;
    rt_accel *get() const { return accel_; }
    rt_scene *scene() const { return scene_; }
    int64_t triangles() const { return triangles_; }

  private:
    rt_scene *scene_ = nullptr;
    rt_accel *accel_ = nullptr;
    int64_t triangles_ = 0;
};

// A scene built from any store with size() and primitive(i) returning a
// triangle with get_vertex_a/b/c() (x, y, z members), get_light() and
// is_transparent(). AlbedoFn maps a triangle index to three floats in [0, 1].
template <class PrimitiveStore>
class gpu_bvh {
  public:
    template <class AlbedoFn>
    gpu_bvh(const context &ctx, const PrimitiveStore &prims, AlbedoFn albedo_of,
            const rt_build_params *params = nullptr)
        : ctx_(ctx), n_(static_cast<int64_t>(prims.size())) {
        const int64_t n = n_;
        const std::vector<float> abc = vertices(prims);
        std::vector<float> albedo(3 * n);
        std::vector<uint8_t> flags(n);
        for (int64_t i = 0; i < n; ++i) {
            const auto *t = prims.primitive(static_cast<int>(i));
            float rgb[3];
            albedo_of(static_cast<int>(i), rgb);
            for (int k = 0; k < 3; ++k) albedo[3 * i + k] = rgb[k];
            flags[i] = static_cast<uint8_t>((t->get_light() ? 1 : 0) | (t->is_transparent() ? 2 : 0));
        }
        ctx_.check(rt_scene_create(ctx_.get(), abc.data(), albedo.data(), flags.data(), n, &scene_),
                   "rt_scene_create");
        ctx_.check(rt_accel_build(scene_, params, &accel_), "rt_accel_build");
    }
    ~gpu_bvh() {
        rt_accel_destroy(accel_);
        rt_scene_destroy(scene_);
    }
    gpu_bvh(const gpu_bvh &) = delete;
    gpu_bvh &operator=(const gpu_bvh &) = delete;

    // The physics render loop builds a new accelerator over the moved
    // triangles every frame (RAPTor/physics_engine/simulation_environment.cc:53-58).
    // refit takes the same store after the move (same triangles, same order)
    // and refits the boxes in place; hits equal those of a new gpu_bvh.
    void refit(const PrimitiveStore &prims) {
        const std::vector<float> abc = vertices(prims);
        ctx_.check(rt_scene_update(scene_, abc.data(), 0), "rt_scene_update");
        ctx_.check(rt_accel_refit(accel_), "rt_accel_refit");
    }

    // Phong materials for ray_tracer: material_of(i) returns the rt_material
    // of triangle i (the state of its phong_shader,
    // rt/materials/phong_shader.h:34-53, 67-73) or nullptr to keep the
    // Lambert albedo. Triangles returning the same pointer share a table
    // entry. A transparent material (transmittance > 0) needs a triangle whose
    // is_transparent() is true, as material(tran > 0) makes it in the
    // reference (rt/materials/material.h:50-66).
    template <class MaterialOf>
    void set_materials(MaterialOf material_of) {
        const int64_t n = n_;
        std::vector<rt_material> table;
        std::vector<int32_t> index(static_cast<size_t>(n), -1);
        std::unordered_map<const rt_material *, int32_t> seen;
        for (int64_t i = 0; i < n; ++i) {
            const rt_material *m = material_of(static_cast<int>(i));
            if (!m) continue;
            auto it = seen.find(m);
            if (it == seen.end()) {
                it = seen.emplace(m, static_cast<int32_t>(table.size())).first;
                table.push_back(*m);
            }
            index[static_cast<size_t>(i)] = it->second;
        }
        ctx_.check(rt_scene_set_materials(scene_, table.data(), static_cast<int32_t>(table.size()), index.data()),
                   "rt_scene_set_materials");
    }

    // Batched closest hit / any hit; pointers are host memory unless flags
    // carries RT_DEVICE_PTRS.
    void trace_closest(const rt_ray *rays, int64_t n, rt_hit *out, uint32_t flags = 0) const {
        ctx_.check(rt_trace_closest(accel_, rays, n, out, flags), "rt_trace_closest");
    }
    void trace_occluded(const rt_ray *rays, int64_t n, uint8_t *occluded, uint32_t flags = 0) const {
        ctx_.check(rt_trace_occluded(accel_, rays, n, occluded, flags), "rt_trace_occluded");
    }

    // One-ray forms with the reference's signatures (slow: one launch each;
    // for legacy callers and tests). Ray must offer get_ogn(), get_dir(),
    // get_length(); Hit must have u, v, d members.
    template <class Ray, class Hit>
    int find_nearest_object(const Ray *r, Hit *h) const {
        rt_ray q = make_ray(*r, 3.402823466e+38f);
        rt_hit out;
        trace_closest(&q, 1, &out);
        h->d = out.t;
        h->u = out.u;
        h->v = out.v;
        return out.prim;
    }
    template <class Ray>
    bool found_nearer_object(const Ray *r, float t) const {
        rt_ray q = make_ray(*r, t);
        uint8_t occ = 0;
        trace_occluded(&q, 1, &occ);
        return occ != 0;
    }

    rt_accel *get() const { return accel_; }
    rt_scene *scene() const { return scene_; }

  private:
    static std::vector<float> vertices(const PrimitiveStore &prims) {
        const int64_t n = static_cast<int64_t>(prims.size());
        std::vector<float> abc(9 * n);
        for (int64_t i = 0; i < n; ++i) {
            const auto *t = prims.primitive(static_cast<int>(i));
            const auto a = t->get_vertex_a(), b = t->get_vertex_b(), c = t->get_vertex_c();
            const float v[9] = {a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z};
            for (int k = 0; k < 9; ++k) abc[9 * i + k] = v[k];
        }
        return abc;
    }

    template <class Ray>
    static rt_ray make_ray(const Ray &r, float tmax) {
        const auto o = r.get_ogn();
        const auto d = r.get_dir();
        rt_ray q;
        q.o[0] = o.x;
        q.o[1] = o.y;
        q.o[2] = o.z;
        q.tmin = 1.0e-3f;  // EPSILON, common/common.cc:14
        q.d[0] = d.x;
        q.d[1] = d.y;
        q.d[2] = d.z;
        q.tmax = tmax;
        return q;
    }

    const context &ctx_;
    int64_t n_ = 0;
    rt_scene *scene_ = nullptr;
    rt_accel *accel_ = nullptr;
};

// Frame into a caller camera: fills it pixel by pixel through set_pixel(colour,
// x, y) as rt/raytracer.cc does (Colour is constructed from three floats on
// the 0-255 scale). cam describes the pinhole (rt/camera.h:52-62).
template <class Store, class Camera, class Colour = typename Camera::colour_type>
void ray_tracer(const gpu_bvh<Store> &accel, const std::vector<rt_light> &lights, const rt_camera &cam,
                Camera &camera, const rt_render_params *params = nullptr) {
    const int64_t npix = static_cast<int64_t>(cam.x_res) * cam.y_res;
    std::vector<float> rgb(3 * npix);
    const int32_t rc = rt_render(accel.get(), &cam, lights.data(), static_cast<int32_t>(lights.size()), params,
                                 rgb.data(), nullptr, nullptr, nullptr);
    if (rc != RT_OK) throw error("rt_render failed");
    for (int y = 0; y < cam.y_res; ++y)
        for (int x = 0; x < cam.x_res; ++x) {
            const float *p = &rgb[3 * (static_cast<int64_t>(y) * cam.x_res + x)];
            camera.set_pixel(Colour(p[0], p[1], p[2]), x, y);
        }
}

}  // namespace raptor_gpu
