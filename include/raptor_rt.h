/* raptor_rt.h — C-ABI of libraptor_rt.so, the MI355X (gfx950) ray-tracing
 * hot path of RAPTor: BVH build, closest-hit and any-hit traversal,
 * Moller-Trumbore intersection and Lambert shading of primary, shadow and
 * diffuse-bounce rays.
 *
 * The ABI is plain C: opaque handles, plain pointers and sizes. Every call
 * returns RT_OK (0) or a negative RT_ERR_* code; rt_last_error() gives the
 * text. No C++ exception crosses it. Input pointers are borrowed for the
 * duration of a call; device memory is owned by the handles; outputs go to
 * caller memory. One rt_ctx is used by one host thread at a time and owns one
 * HIP stream; calls are synchronous.
 *
 * Semantics follow the reference's scalar (IEEE) path: a triangle is hit when
 * rt/triangle.h:190-235 accepts it, hit ids are emplace-order indices
 * (rt/primitive_store.h:17-110), u weights vertex c and v weights vertex b.
 */
#ifndef RAPTOR_RT_H
#define RAPTOR_RT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RT_OK 0
#define RT_ERR_INVALID (-1)
#define RT_ERR_HIP (-2)
#define RT_ERR_OOM (-3)
#define RT_ERR_UNSUPPORTED (-4)

/* Pointer-location flag for trace calls: inputs/outputs are device pointers. */
#define RT_DEVICE_PTRS 1u
/* rt_render_params.flags: regroup bounce rays by direction octant and origin
 * before tracing them. */
#define RT_REGROUP_BOUNCE 2u
/* rt_render_params.flags: enqueue the frame on the context's stream and
 * return without waiting for it (needs RT_DEVICE_PTRS and stats == NULL;
 * frames without materials or vertex normals). Frames so enqueued back to
 * back keep the GPU busy between frames. Their ray counts are summed on the
 * device and read by rt_ctx_frame_rays. No counterpart in the reference,
 * whose ray_tracer() call returns with the frame done. */
#define RT_ASYNC 4u

typedef struct rt_ctx rt_ctx;
typedef struct rt_scene rt_scene;
typedef struct rt_accel rt_accel;

/* A ray: origin, tmin, direction, tmax (32 bytes). A triangle is accepted when
 * tmin < t; closest hit also requires t < tmax; any hit requires t < tmax. The
 * reference fixes tmin at EPSILON = 1e-3 (common/common.cc:14), closest-hit
 * tmax at MAX_DIST (ssd/bvh.cc:649-709) and any-hit tmax at the shadow-ray
 * length (ssd/bvh.cc:711-759). */
typedef struct {
    float o[3], tmin, d[3], tmax;
} rt_ray;

/* A hit (16 bytes): distance, barycentrics, emplace-order triangle id or -1. */
typedef struct {
    float t, u, v;
    int32_t prim;
} rt_hit;

/* Pinhole camera as rt/camera.h:52-62 holds it (x_a_res = y_a_res = 1). */
typedef struct {
    float pos[3];
    float x_axis[3], y_axis[3], z_axis[3];
    float w, h, t;
    int32_t x_res, y_res;
} rt_camera;

/* Point light, rt/light.h:17-18: rgb in [0, 255], distance falloff d. */
typedef struct {
    float pos[3];
    float rgb[3];
    float d;
} rt_light;

/* BVH build and traversal options. width: 4 or 8 children per node;
 * leaf_size: most triangles per leaf (1..16); sah_bins: bins per axis of the
 * binned-SAH rebuild of the Morton LBVH (4..32; 0 = keep the LBVH: faster
 * build, slower traversal); lds_stack: traversal-stack entries per lane kept
 * in LDS before spilling to private memory (4, 8 or 16; 0 = 16); builder:
 * RT_BUILDER_LBVH (the Morton LBVH with the binned-SAH rebuild) or
 * RT_BUILDER_AAC (the reference builder's own tree, rt_build_aac, collapsed
 * to width-wide nodes; its leaves are the reference's, so leaf_size and
 * sah_bins do not apply, and a leaf of more than 16 triangles takes more
 * than one child slot). params == NULL means width 4, leaf size 2, 16 SAH
 * bins, 8 LDS entries, RT_BUILDER_LBVH. */
#define RT_BUILDER_LBVH 0
#define RT_BUILDER_AAC 1
typedef struct {
    int32_t width;
    int32_t leaf_size;
    int32_t sah_bins;
    int32_t lds_stack;
    int32_t builder;
} rt_build_params;

/* Frame options: bounce = 1 adds one cosine-weighted diffuse bounce with
 * direct light at the bounce hit; spp samples per pixel; seed keys the
 * counter-based sampler. background is the camera's background colour
 * (rt/camera.h:234-240, 0-255 scale): the colour of a primary ray, or of a
 * mirror or refraction ray, that hits nothing (rt/raytracer.cc:82-86); a
 * diffuse bounce that hits nothing adds nothing. */
typedef struct {
    int32_t bounce;
    int32_t spp;
    uint64_t seed;
    int32_t flags;
    int32_t reserved;
    float background[3];
    int32_t reserved2;
} rt_render_params;

/* Frame statistics. rays: primary, shadow, bounce, bounce-shadow.
 * ms: device time of the frame's kernels (HIP events): [0] whole frame,
 * [1] primary closest hit, [2] shadow any hit, [3] bounce closest hit,
 * [4] bounce-shadow any hit, [5] ray generation and shading. With a bounce
 * the shadow any-hit launch runs on the context's second stream beside the
 * bounce closest-hit launch, so [1]..[5] can sum to more than [0]. visits: node visits and triangle tests of each of the four trace
 * launches, [2*k] and [2*k+1] for k = primary, shadow, bounce, bounce-shadow
 * (all zero unless counting is on). */
typedef struct {
    uint64_t rays[4];
    float ms[6];
    uint64_t visits[8];
} rt_stats;

/* BVH statistics. */
typedef struct {
    int64_t nodes, leaves, triangles;
    int32_t width, depth;
    int64_t bytes;
    float build_ms;
    float sah_cost;
    float refit_ms;  /* the last rt_accel_refit, 0 before any */
} rt_accel_info;

/* Destroy calls release the caller's handle and may come in any order: a
 * context, scene or accel is freed once no scene, accel or batch built on it
 * remains. */
int32_t rt_ctx_create(int32_t hip_device, rt_ctx **out);
void rt_ctx_destroy(rt_ctx *ctx);
const char *rt_last_error(const rt_ctx *ctx);
/* Waits for the context's frames and returns the rays traced by its
 * RT_ASYNC frames since the last call, by stage (primary, shadow, bounce,
 * bounce shadow), then sets the sums to zero. */
int32_t rt_ctx_frame_rays(rt_ctx *ctx, uint64_t rays[4]);

/* Scene in emplace order: abc is n x 9 floats (a, b, c), albedo n x 3 in
 * [0, 1] or NULL (treated as 1), flags one byte per triangle (bit0 light,
 * bit1 transparent; both skipped by any-hit queries) or NULL. Replaces
 * primitive_store + triangle construction (rt/primitive_store.h:17-110,
 * rt/triangle.h:124-176). */
int32_t rt_scene_create(rt_ctx *ctx, const float *abc, const float *albedo, const uint8_t *flags, int64_t n,
                        rt_scene **out);
void rt_scene_destroy(rt_scene *scene);
/* Per-triangle normals as the reference computes them (n x 3 floats, host). */
int32_t rt_scene_normals(rt_scene *scene, float *n3);

/* Build the acceleration structure. Replaces bvh::bvh(primitive_store&)
 * (ssd/bvh.h:27-32, ssd/bvh_builder.cc:18-178); the caller's triangle order is
 * not permuted. params may be NULL (see rt_build_params). */
int32_t rt_accel_build(rt_scene *scene, const rt_build_params *params, rt_accel **out);
void rt_accel_destroy(rt_accel *accel);
int32_t rt_accel_info_get(const rt_accel *accel, rt_accel_info *info);

/* ---- Dynamic scenes. The physics engine's render loop builds a new
 * accelerator over the moved triangles for every frame
 * (RAPTor/physics_engine/simulation_environment.cc:53-58); these two calls
 * serve that loop without a rebuild. ----
 *
 * rt_scene_update replaces the scene's vertex positions: abc is n x 9 floats
 * with the scene's own n, on the host, or on the device with RT_DEVICE_PTRS.
 * Normals follow the new vertices; albedo and flags stay. Every accel built
 * on the scene must be refit (or rebuilt) before it is traced again.
 *
 * rt_accel_refit fits the accel to its scene's current vertices and keeps
 * its tree: leaf triangles are re-read and every box is recomputed from the
 * deepest level to the root. Hits are those of a fresh build over the same
 * triangles (up to exact-t ties); traversal slows as the motion departs from
 * the one the tree was built for, when a rebuild (rt_accel_build, sah_bins 0
 * for the fastest) pays again. Batches and inverse-rendering sets built on
 * the accel see the refit boxes without being recreated. */
int32_t rt_scene_update(rt_scene *scene, const float *abc, uint32_t flags);
int32_t rt_accel_refit(rt_accel *accel);

/* ---- Materials. ----
 * kind RT_MATERIAL_PHONG: the state of a phong_shader
 * (rt/materials/phong_shader.h:34-53, 67-73): ka, kd, ks colours on the
 * 0-255 scale, shininess s, transmittance tran, refractive index ri,
 * reflectance rf; chroma, diffuse, roughness and extinction are unused.
 * kind RT_MATERIAL_COOK_TORRANCE: the state of a cook_torrance_cxy
 * (rt/materials/cook_torrance_cxy.h:14-30, 45-54): chroma (x, y), diffuse
 * reflectance rd (diffuse), surface roughness sr (roughness), imaginary
 * refractive index ri_i (extinction), and reflectance, transmittance,
 * refractive_index holding its specular reflectance rs, specular
 * transmittance ts and real index ri_r; ka, kd, ks and shininess are
 * unused. */
#define RT_MATERIAL_PHONG 0
#define RT_MATERIAL_COOK_TORRANCE 1
/* kind RT_MATERIAL_COLOURED: a coloured_mapper_shader
 * (rt/materials/coloured_mapper_shader.{h,cc}), the material an OBJ/MTL file
 * gives: the fields and shading of RT_MATERIAL_PHONG, with the wider ranges
 * its constructor asserts (tran and rf in [0, 1]), and optional texture maps
 * tex[] = 1 + the index of a map of rt_scene_set_textures (0: none) for, in
 * order, t_ka, t_kd, t_ks, t_ns, t_refl and t_d: a colour map replaces the
 * colour by the texture times colour / 255 (a checker board leaves the
 * colour where it does not paint); t_ns makes the shininess
 * ns / 255 times the map's opacity (the reference multiplies the returned
 * alpha, not the texel); t_refl and t_d scale the mirror and refraction
 * rays' coefficients by red / 255 of their map while the colours those rays
 * bring back stay weighted by rf and tran (rt/materials/coloured_mapper_shader.cc:8-101).
 * kind RT_MATERIAL_MAPPER: a mapper_shader (rt/materials/mapper_shader.{h,cc}):
 * kd[] the base colour rgb (0-255), mapper_kd and mapper_ks its kd (>= 0)
 * and ks (0..1), shininess s, transmittance, refractive_index, reflectance,
 * clrh (specular colour blend) and clrf (transmitted colour blend), and five
 * layer lists of maps, layers[layer_first[k] .. + layer_count[k]) of
 * rt_scene_set_textures, for k = 0 colour, 1 kd, 2 s, 3 reflectance,
 * 4 transmittance. Layers apply in order and stop after one whose opacity
 * makes the product 0; a monochrome layer (kd, s, reflectance,
 * transmittance) sets the value to its texel's mean / 255. No ambient term;
 * per light: rgb x (n.l kd) x light, plus (clrh rgb + (1 - clrh) light) x
 * (r.v)^s ks, where r reflects the light about n scaled by 2 (n.l) kd as
 * the reference writes it; the refracted colour adds as
 * c += clrf c + (1 - clrf) refracted x transmittance. */
#define RT_MATERIAL_COLOURED 2
#define RT_MATERIAL_MAPPER 3
/* kind RT_MATERIAL_MANDELBROT: a mandelbrot_shader
 * (rt/materials/mandelbrot_shader.{h,cc}) with iteration limit iterations
 * (>= 1); the constructor's x and y widths do not enter its shading. One
 * shadow ray per light and no secondary rays; colour = the escape-time
 * colour of the hit point's fractal coordinates (taken on the y-z plane
 * where the shading normal is exactly +-x, on x-y where it is exactly +-z,
 * else on x-z; a point inside the set is tried once more on shifted,
 * 1.5x scaled coordinates, then black) times the sum over lit lights of
 * n.l where n.l >= 0; no light colour, falloff or ambient term. */
#define RT_MATERIAL_MANDELBROT 4
typedef struct {
    float ka[3], kd[3], ks[3];
    float shininess, transmittance, refractive_index, reflectance;
    int32_t kind;
    float chroma[2];
    float diffuse, roughness, extinction;
    int32_t tex[6];
    float mapper_kd, mapper_ks, clrh, clrf;
    int32_t layer_first[5], layer_count[5];
    int32_t iterations;  /* RT_MATERIAL_MANDELBROT */
} rt_material;

/* ---- Texture maps (rt/materials/texture_mapper.h) ----
 * A map's final state, as the reference's constructors leave it:
 *   RT_MAP_PLANAR (planar_mapper), RT_MAP_CUBIC (cubic_mapper),
 *   RT_MAP_CYLINDRICAL (cylindrical_mapper): an image, centre c, size s,
 *   texture axes u and v (the first constructors derive u = (n.y + n.z, 0,
 *   n.x), v = (0, n.x + n.z, n.y), cylindrical v = n), the cubic mapper's
 *   normal n, the cylindrical mapper's cycles, wrapping modes and texel
 *   offsets. A hit with texture coordinates samples at (vt.u w, vt.v h);
 *   otherwise the mapper projects the hit point
 *   (rt/materials/{planar,cubic,cylindrical}_mapper.cc).
 *   The lookup is bilinear over texels (u0, v0) .. (u0 + 1, v0 + 1), u0 =
 *   int(u_co) + u_off wrapped by the mode, the upper neighbour clamped to
 *   the image (texture_mapper.h image_texture_mapper).
 *   RT_MAP_CHECKER_BOARD (checker_board_mapper): colour rgb where
 *   int((x - c.x) / s.x) is odd (the reference's test xors the z cell with
 *   itself, so only x decides), otherwise the colour is left alone;
 *   opacity is the map's alpha.
 *   RT_MAP_PERLIN_3D (perlin_noise_3d_mapper, rt/materials/
 *   perlin_noise_3d_mapper.{h,cc}): octaves octaves of the seeded value
 *   noise of common/perlin_noise_3d.cc at the hit point times frequency
 *   1 / zoom, doubling, amplitudes persistence^k; total + 1 scales rgb / 2;
 *   alpha 0 where total + 1 > 1, else opacity.
 *   RT_MAP_GRADIENT (gradient_mapper, rt/materials/gradient_mapper.{h,cc}):
 *   the key is the incidence angle acos(|n.dir|) in degrees; key_count
 *   keys of 6 floats at texels + key_offset, {key, r, g, b, alpha,
 *   interpolation}, keys non-decreasing; below the first key the first
 *   value, from the last key on the last value, else the value interpolated
 *   by the upper key's interpolation (RT_GRAD_LINE: colour linear in the
 *   key, alpha the lower key's; RT_GRAD_STEP: the upper value). grad_of must
 *   be RT_GRAD_OF_INCIDENCE_ANGLE and spline interpolation is refused: the
 *   reference asserts on the rest. The mapper inverts its own colour when
 *   inverse is set, before the falloff, and the map inverts it again after.
 * Then, per map, a falloff (mapper_falloff: cubic, spherical, linear x, y,
 * z around falloff_centre with gradient falloff_grad; RT_FALLOFF_NONE
 * skips it) and inverse (255 - colour). The monochrome value of a
 * procedural map (checker board, Perlin, gradient) is its colour's mean. */
#define RT_MAP_PLANAR 0
#define RT_MAP_CUBIC 1
#define RT_MAP_CYLINDRICAL 2
#define RT_MAP_CHECKER_BOARD 3
#define RT_MAP_PERLIN_3D 4
#define RT_MAP_GRADIENT 5
/* grad_interpolation_t and grad_of_t (rt/materials/gradient_mapper.h). */
#define RT_GRAD_LINE 0
#define RT_GRAD_SPLINE 1
#define RT_GRAD_STEP 2
#define RT_GRAD_OF_INCIDENCE_ANGLE 3
/* texture_wrapping_mode_t (texture_mapper.h apply_wrapping_mode): blank
 * (black outside), clamp, tile (negative texels wrap by size - 1), mirror. */
#define RT_WRAP_BLANK 0
#define RT_WRAP_CLAMP 1
#define RT_WRAP_TILE 2
#define RT_WRAP_MIRROR 3
#define RT_FALLOFF_CUBIC 0
#define RT_FALLOFF_SPHERICAL 1
#define RT_FALLOFF_LINEAR_X 2
#define RT_FALLOFF_LINEAR_Y 3
#define RT_FALLOFF_LINEAR_Z 4
#define RT_FALLOFF_NONE 5
typedef struct {
    int32_t kind;
    int32_t image;              /* image maps: index into the images */
    float c[3], n[3], s[3];     /* checker board: c centre, s grid size */
    float u[3], v[3];
    int32_t u_wrap, v_wrap;
    int32_t u_off, v_off;
    int32_t cycles;             /* cylindrical */
    float rgb[3], opacity;      /* checker board */
    int32_t inverse;
    int32_t falloff;            /* RT_FALLOFF_* */
    float falloff_centre[3], falloff_grad[3];
    float persistence, zoom;    /* Perlin */
    int32_t octaves, seed;      /* Perlin */
    int32_t key_count, grad_of; /* gradient */
    int64_t key_offset;         /* gradient: first key's float in the texels */
} rt_texture_map;
/* An image: width x height texels of cpp (1, 3 or 4) floats, 0-255, row 0
 * first (the readers' order: the image's bottom row for TGA/PNG/JPEG),
 * at texels + offset. */
typedef struct {
    int32_t width, height, cpp, reserved;
    int64_t offset;
} rt_texture_image;
/* Gives the scene its maps, images, texels (host arrays) and the layer
 * lists of RT_MATERIAL_MAPPER materials (map indices). nmaps 0 removes
 * them. Materials (rt_scene_set_materials) may reference maps only once
 * the scene has them; calling this again re-checks nothing already set, so
 * set textures first. Tile and mirror wrapping need images of at least 2
 * texels along that axis (the reference divides by size - 1). */
int32_t rt_scene_set_textures(rt_scene *scene, const rt_texture_map *maps, int32_t nmaps,
                              const rt_texture_image *images, int32_t nimages, const float *texels,
                              int64_t ntexels, const int32_t *layers, int32_t nlayers);
/* Per-triangle texture coordinates: vt holds 6 floats per triangle, (u, v)
 * at vertices a, b, c (host array). A triangle whose first value is FLT_MAX
 * has none (rt/triangle.h:139-167). A hit interpolates u vt_c + v vt_b +
 * (1 - (u + v)) vt_a (rt/triangle.h:72-79). vt == NULL removes them. */
int32_t rt_scene_set_texture_coords(rt_scene *scene, const float *vt);
/* Reads a TGA image (uncompressed or run-length encoded true-colour or
 * grey, 8/24/32 bits) as the reference's read_tga does through libtga
 * (rt/materials/texture_mapper.h read_tga): rows in file order, channels
 * swapped to RGB, one float 0-255 per channel. dims receives width,
 * height and channels; texels (capacity floats) may be NULL to query the
 * dimensions. Host only. */
int32_t rt_texture_load_tga(const char *path, float *texels, int64_t capacity, int32_t *dims);

/* Gives the scene's triangles materials for rt_render: tri_material[i]
 * indexes materials (host arrays), -1 keeps triangle i's Lambert shading
 * phong_shader(255 * albedo, 1.0). rt_render then shades every hit as the
 * material's shade does (rt/materials/phong_shader.cc:21-111,
 * rt/materials/cook_torrance_cxy.cc:12-90): light per light, ambient
 * (phong), and the mirror (reflectance > 0) and refraction
 * (transmittance > 0) rays traced recursively while the ray power stays
 * above MIN_REFLECTIVE_POWER (common/common.h:30-32). The ranges the
 * reference asserts are errors here: phong colours in [0, 255], tran and rf
 * in [0, 0.95], ri in [1, 5] when tran > 0; Cook-Torrance x, y, rd, rs, ts,
 * ri_i in [0, 1], sr in [0, 0.5], ri_r in [0, 5]. A transparent material
 * makes its triangles transparent to shadow rays (material(tran > 0),
 * rt/materials/material.h:50-66), so a triangle's flags bit1 must be set
 * exactly when its material's transmittance is > 0. nmaterials 0 removes
 * the materials. Only rt_render reads them; the diffuse bounce and the
 * differentiable frames keep Lambert shading and reject a scene with
 * materials. */
int32_t rt_scene_set_materials(rt_scene *scene, const rt_material *materials, int32_t nmaterials,
                               const int32_t *tri_material);

/* Per-vertex normals for smooth shading: vn holds 9 floats per triangle
 * (the normals at vertices a, b, c; host array). A triangle whose first
 * value is FLT_MAX keeps its geometric normal, as a triangle built without
 * vertex normals does (rt/triangle.h:139-167). rt_render then shades a hit
 * with u * n_c + v * n_b + (1 - (u + v)) * n_a, unnormalised and turned to
 * face the ray (rt/triangle.h:464-498), for light and secondary rays; shadow
 * and secondary ray origins keep the geometric normal. vn == NULL removes
 * them. Like materials, only rt_render reads them. */
int32_t rt_scene_set_vertex_normals(rt_scene *scene, const float *vn);

/* ---- Scene ingestion (host): OFF, PLY, OBJ + MTL files ----
 * Replaces off_parser, ply_parser and obj_parser
 * (rt/parsers/off_parser.cc:14-99, rt/parsers/ply_parser.cc:399-583,
 * rt/parsers/obj_parser.cc:762-857) with polygon triangulation
 * (rt/parsers/polygon_to_triangles.h:141-386): the triangles come out in the
 * order those parsers emplace them (that order is the hit id), with one
 * material per triangle (OFF: phong kd 170; PLY: phong kd 200, or the face
 * colour; OBJ: phong kd 200 until a usemtl or mtllib, MTL materials as
 * RT_MATERIAL_COLOURED) and, for OBJ faces with normals, vertex normals.
 * OBJ texture coordinates (vt) and MTL texture maps (map_Ka, map_Kd, map_Ks,
 * map_refl, map_d: rt/parsers/obj_parser.cc:218-242, 507-561) come with the
 * mesh when the image is a TGA file: each map statement becomes a planar map
 * (centre 0, normal (1, 0, 0), size 0, mirror wrapping) sampled through the
 * triangles' texture coordinates, map_d with the image inverted as the
 * reference's negative() does (1 - texel); the illumination models drop maps
 * as they drop terms. Map files in other formats are counted in
 * texture_maps and not applied. PLY may be ASCII or binary, big- or
 * little-endian. Errors (malformed input, indices out of range) return a
 * negative code with text in rt_mesh_last_error() (per thread). These calls
 * need no GPU. */
#define RT_FORMAT_OFF 0
#define RT_FORMAT_PLY 1
#define RT_FORMAT_OBJ 2
/* NFF (rt/parsers/nff_parser.cc:16-224): the viewpoint, background, point
 * lights, fill materials (phong_shader(255 rgb, kd, ks, shine, T, ri, ks))
 * and polygons / polygon patches, each triangulated; other primitives
 * (spheres, cones) are skipped as the reference skips them. */
#define RT_FORMAT_NFF 3
typedef struct rt_mesh rt_mesh;
typedef struct {
    int64_t triangles;
    int32_t materials;
    int32_t has_vertex_normals;
    int64_t vertices, faces;
    int64_t skipped_faces;       /* degenerate triangles and faces below three vertices */
    int64_t skipped_statements;  /* OBJ/MTL/NFF statements with nothing to shade */
    int32_t texture_maps;        /* MTL map statements whose file cannot be decoded here (not TGA) */
    int32_t lights;              /* point lights the file defines (NFF) */
    int32_t has_camera;          /* 1 when the file defines the camera (NFF) */
    int32_t has_texture_coords;  /* 1 when some triangle has texture coordinates (OBJ vt) */
    int32_t maps, images;        /* texture maps (OBJ/MTL map_* statements) and their distinct images */
    int64_t texels;
} rt_mesh_info;
/* format: RT_FORMAT_*, or -1 to take it from the file extension. */
int32_t rt_mesh_load(const char *path, int32_t format, rt_mesh **out);
/* Parses a file held in memory; dir is where an OBJ's mtllib files are
 * (NULL: the working directory). */
int32_t rt_mesh_parse(const char *data, int64_t len, int32_t format, const char *dir, rt_mesh **out);
int32_t rt_mesh_info_get(const rt_mesh *mesh, rt_mesh_info *info);
/* Copies out any of: abc (triangles x 9), tri_material (triangles),
 * materials (materials), vertex_normals (triangles x 9, only when
 * has_vertex_normals). NULL skips an array. */
int32_t rt_mesh_copy(const rt_mesh *mesh, float *abc, int32_t *tri_material, rt_material *materials,
                     float *vertex_normals);
/* Copies out any of: vt (triangles x 6, only when has_texture_coords),
 * maps (maps), images (images), texels (texels floats); the map indices in
 * materials' tex[] refer to these maps. NULL skips an array. */
int32_t rt_mesh_copy_textures(const rt_mesh *mesh, float *vt, rt_texture_map *maps, rt_texture_image *images,
                              float *texels);
/* The camera and background colour a file defines (has_camera), and its
 * point lights (info.lights of them). */
int32_t rt_mesh_camera(const rt_mesh *mesh, rt_camera *cam, float *background);
int32_t rt_mesh_lights(const rt_mesh *mesh, rt_light *lights);
const char *rt_mesh_last_error(void);
void rt_mesh_destroy(rt_mesh *mesh);
/* A scene from a parsed mesh: rt_scene_create (albedo = kd / 255, flags bit1
 * for transparent materials), rt_scene_set_textures (when the mesh has
 * maps), rt_scene_set_materials and, when the mesh has them,
 * rt_scene_set_vertex_normals and rt_scene_set_texture_coords. The mesh may
 * be destroyed afterwards. */
int32_t rt_scene_create_from_mesh(rt_ctx *ctx, const rt_mesh *mesh, rt_scene **out);

/* ---- Camera post-process: the step after the image buffer ----
 * Images are camera::image as rt_render writes it: x_res * y_res * 3 floats
 * on the 0-255 scale, pixel (x, y) at 3 * (x + y * x_res), on the host or,
 * with RT_DEVICE_PTRS, on the device. ms (may be NULL) receives the device
 * time of the call's kernels.
 *
 * rt_image_tone_map replaces camera::tone_map (rt/camera.cc:1124-1461):
 * mode is tone_mapping_mode_t (rt/camera.h:43-45): 0 none,
 * 1 global_contrast, 4 global_non_linear, 5 global_exposure,
 * 6 global_avg_luminance, 7 global_max_luminance, 8 global_bilateral_filter
 * (camera::bilateral_filter_tone_map, rt/camera.cc:1632-1775: a bilateral
 * grid blurred by a circular Gaussian, evaluated directly in double where
 * the reference uses single-precision FFTs), 9 global_ferwerda, and the
 * histogram modes 2 local_histogram and 3 local_human_histogram
 * (camera::histogram_tone_map, rt/camera.cc:1843-2062), which need x_angle
 * and y_angle of at least one degree.
 * key, xw, yw, desaturate (ds), stretch (sc) and adapt (ta) with time_step
 * are its arguments; adaptation_level is the camera's adaptation state (read
 * when adapting, always written), log_avg_luminance receives the adapted
 * log-average luminance. The log-average is summed in double (the
 * reference sums in float, pixel by pixel).
 * glare (gf) adds camera::perform_glare_filter (rt/camera.cc:1504-1630) after
 * the mapping: the pixels with a channel above 255 convolved circularly with
 * the glare filter of the light level of the adapted luminance (scotopic
 * below 10^-2, mesopic below 10^0.6, photopic above; rt/camera.cc:145-462:
 * 2 x 200 random flare lines and a bloom, spread by Spencer's functions,
 * scaled to 1e-4 per pixel), divided by 255 and added. It needs x_angle and
 * a frame of at most 4096 x 4096. The filter is drawn from the reference's
 * generator as first used in a process (MT19937 seeded 5489) and kept on the
 * context per light level; the convolution is exact through zero-padded FFTs
 * in single precision (the reference: fftw3f). */
typedef struct {
    int32_t mode;
    float key, xw, yw;
    int32_t desaturate, stretch, adapt;
    float time_step;
    float adaptation_level;
    float log_avg_luminance;
    /* The camera's field of view in radians, 2 w / t and 2 h / t (modes 2
     * and 3 average the frame to one sample per degree of view; glare uses
     * x_angle for the spreading functions' degrees per pixel). */
    float x_angle, y_angle;
    int32_t glare;
} rt_tone_params;
int32_t rt_image_tone_map(rt_ctx *ctx, float *rgb, int32_t x_res, int32_t y_res, rt_tone_params *params,
                          uint32_t flags, float *ms);
/* The glare filter of an x_res x y_res camera with field of view x_angle
 * (2 w / t) at light level 0 scotopic, 1 mesopic, 2 photopic: x_res * y_res
 * rgb floats into out (host, or device with RT_DEVICE_PTRS), pixel (x, y) at
 * 3 * (x + y * x_res). camera::generate_glare_filter (rt/camera.cc:325-462). */
int32_t rt_glare_filter(rt_ctx *ctx, int32_t x_res, int32_t y_res, float x_angle, int32_t level, float *out,
                        uint32_t flags);
/* camera::gamma_correct (rt/camera.cc:1466-1501) of npix pixels, in place. */
int32_t rt_image_gamma(rt_ctx *ctx, float *rgb, int64_t npix, float gamma, uint32_t flags, float *ms);
/* clip_image_to_rgb (bgr = 0) or clip_image_to_bgr (bgr = 1)
 * (rt/camera.h:289-351): an x_res x y_res image with x_samples x y_samples
 * samples per output pixel box-filtered to (x_res / x_samples) x
 * (y_res / y_samples) pixels, clamped to [0, 255] and truncated to bytes. */
int32_t rt_image_clip(rt_ctx *ctx, const float *rgb, int32_t x_res, int32_t y_res, int32_t x_samples,
                      int32_t y_samples, int32_t bgr, uint8_t *out, uint32_t flags, float *ms);
/* Host writers for clipped images: write_png_file (rt/camera.cc:2158-2210,
 * rgb bytes, last image row first, stored deflate blocks) and
 * write_tga_file (rt/camera.cc:2102-2155, bgr bytes). */
int32_t rt_write_png(const char *path, const uint8_t *rgb, int32_t x_res, int32_t y_res);
int32_t rt_write_tga(const char *path, const uint8_t *bgr, int32_t x_res, int32_t y_res);
/* camera::write_jpeg_file (rt/camera.cc:2305-2366): rgb bytes (host, or
 * device with RT_DEVICE_PTRS), last image row first, as a baseline JFIF with
 * libjpeg's defaults at quality 1..100 (4:2:0 chroma, the T.81 Annex K
 * tables scaled by the IJG quality rule, Annex K Huffman codes). Colour
 * conversion, subsampling, DCT and quantisation run on the device; ms (may
 * be NULL) receives their time. */
int32_t rt_write_jpeg(rt_ctx *ctx, const char *path, const uint8_t *rgb, int32_t x_res, int32_t y_res,
                      int32_t quality, uint32_t flags, float *ms);

/* Inspection of the build (tests and tools; host outputs).
 * rt_build_morton runs the first stages of rt_accel_build on the scene:
 * per-triangle bounds, 30-bit Morton codes of the box centroids and the
 * stable LSD radix sort by code. codes[i]: triangle i's code (emplace order);
 * order[k]: the emplace id at position k of the sorted order. Counterpart of
 * the reference builder's Morton stages (ssd/bvh_builder.cc:33-158), with one
 * cell size for all axes (the largest centroid extent / 1024) where the
 * reference quantises each axis by its own extent (ssd/bvh_builder.cc:45-48).
 * rt_accel_download copies the tree: rt_accel_info.nodes wide nodes of width
 * W (float lo_x[W], hi_x[W], lo_y[W], hi_y[W], lo_z[W], hi_z[W]; int32
 * child[W]; padding to 128 bytes for W = 4, 256 for W = 8; child >= 0 an
 * interior node, 0x80000000 | first << 4 | (count - 1) a leaf of count
 * triangles at BVH positions [first, first + count), -1 empty) and tri_ids[n],
 * the emplace id at each BVH position. rt_accel_root gives the root's child
 * code (0, or a leaf code for a one-triangle scene, -1 for an empty one) and
 * the padding the build added to every box. */
int32_t rt_build_morton(rt_scene *scene, uint32_t *codes, uint32_t *order);
/* The reference builder's own tree, built on the GPU: RAPTor's approximate
 * agglomerative clustering (ssd/bvh_builder.cc:18-642, class bvh_builder),
 * node for node: Morton codes of the box centroids on per-axis cells of the
 * scene extent * 1.00001 / 1024 (ssd/bvh_builder.cc:33-49), the stable radix
 * order, the top-down split at the Morton-range midpoint with the
 * spatial-midpoint partition once codes are exhausted (build_layer,
 * build_layer_primitive), greedy leaf packing under max_leaf_sah
 * (build_leaf_node) and the agglomerative merge of each node's clusters by
 * least combined surface area down to delta^0.5 n^alpha / 2 clusters, ties
 * decided as the reference's SSE min_element decides them (combine_nodes,
 * reduction_function; common/simd.h:857-927). alpha and delta as
 * ssd/bvh_builder.h:23 (0.4, 20). Host outputs: codes[n] (emplace order,
 * may be NULL), order[n] (the radix sort's order, may be NULL),
 * indirect[n] (the triangle order after the builder's partitions and leaf
 * packing), nodes[capacity >= 2n - 1], *nnodes, *root, *build_ms (may be
 * NULL). A leaf covers indirect[child[0] .. child[1]); an interior node's
 * child[0] and child[1] are the node ids the reference makes its left_b and
 * right_e. Node ids differ from the reference's creation order; the tree is
 * the same. */
typedef struct {
    float lo[3], hi[3];
    int32_t child[2];
    int32_t leaf;
} rt_aac_node;
int32_t rt_build_aac(rt_scene *scene, float alpha, float delta, uint32_t *codes, uint32_t *order, int32_t *indirect,
                     rt_aac_node *nodes, int64_t capacity, int64_t *nnodes, int32_t *root, float *build_ms);
int32_t rt_accel_download(const rt_accel *accel, void *nodes, int32_t *tri_ids);
int32_t rt_accel_root(const rt_accel *accel, int32_t *root, float *box_pad);

/* Closest hit for n rays. Replaces bvh::find_nearest_object(ray*)
 * (ssd/bvh.cc:649-709) and its packet form (ssd/bvh.cc:406-447). */
int32_t rt_trace_closest(rt_accel *accel, const rt_ray *rays, int64_t n, rt_hit *out, uint32_t flags);
/* Any hit (occlusion) for n rays; occluded[i] = 1 when blocked. Replaces
 * bvh::found_nearer_object(ray*, t) (ssd/bvh.cc:711-759) and its packet form
 * (ssd/bvh.cc:449-500). */
int32_t rt_trace_occluded(rt_accel *accel, const rt_ray *rays, int64_t n, uint8_t *occluded, uint32_t flags);

/* One frame: primary rays, one shadow ray per hit per light, Lambert shading
 * (phong_shader(255 * albedo, 1.0)), optional diffuse bounce. Replaces
 * ray_tracer(ssd*, light_list&, primitive_store&, camera&)
 * (rt/raytracer.h:105, rt/raytracer.cc:563-598). rgb: x_res*y_res*3 floats on
 * the 0-255 scale of camera::image (rt/camera.h:281-285), row y at
 * rgb[3*(y*x_res + x)]. primary, rgb, bounce may each be NULL. With
 * RT_DEVICE_PTRS in params->flags the output pointers are device pointers. */
int32_t rt_render(rt_accel *accel, const rt_camera *cam, const rt_light *lights, int32_t nlights,
                  const rt_render_params *params, float *rgb, rt_hit *primary, rt_hit *bounce, rt_stats *stats);

/* rt_render without any device allocation. No counterpart in the reference:
 * for callers that cannot allocate (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc). The frame's workspace (rays, hits,
 * queues, shadow rays, compaction counts) lives in a static device array of
 * 64 MiB per GPU, held from the library's load on that device until the
 * process ends; frames whose workspace exceeds it run in chunks of 8x8 tiles
 * that fit, and a frame whose single tile does not fit (a very high spp)
 * returns RT_ERR_INVALID. rgb (required), primary and bounce are device
 * pointers (RT_DEVICE_PTRS is implied). Scenes with materials or vertex
 * normals return RT_ERR_UNSUPPORTED. Every call on one GPU shares that GPU's
 * array, from any context: calls that overlap on one GPU (other host
 * threads, other contexts) race and must be serialised by the caller. */
int32_t rt_render_static(rt_accel *accel, const rt_camera *cam, const rt_light *lights, int32_t nlights,
                         const rt_render_params *params, float *rgb, rt_hit *primary, rt_hit *bounce,
                         rt_stats *stats);

/* ---- Batched tracing (many scenes in one launch) ----
 * An item is one scene's BVH with a batch of rays and room for the hits, all
 * device pointers. Items of one batch share a node width. The reference has no
 * batching; each item is the work of one bvh::find_nearest_object loop. */
typedef struct {
    rt_accel *accel;
    const rt_ray *rays;
    int64_t n;
    rt_hit *hits;
} rt_batch_item;

/* Persistent-kernel parameters: lds_stack (4, 8, 16), resident wavefronts per
 * CU, 64-ray chunks taken per atomic, and whether to replay a captured
 * hipGraph (1) or launch directly (0). Zeros select 8, 16, 1.
 * group_waves: wavefronts per workgroup (1, 2 or 4; 0 selects 1).
 * refill: the compaction threshold. 0 traces each 64-ray chunk to its last
 * ray (finished lanes idle); 1..64 lets a wavefront keep a pool of ray slots
 * from the chunks it took and hand new rays to its finished lanes once at
 * least refill of them are idle (lanes of one wavefront may then trace
 * different scenes). width: 0 traces the items' BVHs; 4 or 8 traces BVHs of
 * that node width, which the batch builds from the items' scenes (with the
 * items' leaf size and SAH bins) when the items' width differs, and owns. */
typedef struct {
    int32_t lds_stack;
    int32_t waves_per_cu;
    int32_t grab;
    int32_t use_graph;
    int32_t group_waves;
    int32_t refill;
    int32_t width;
} rt_batch_params;

typedef struct rt_batch rt_batch;

int32_t rt_batch_create(rt_ctx *ctx, const rt_batch_item *items, int32_t nitems, const rt_batch_params *params,
                        rt_batch **out);
/* mode 0: the persistent work-queue launch (hipGraph replay when captured);
 * mode 1: one closest-hit launch per item (baseline). ms: device time. */
int32_t rt_batch_run(rt_batch *batch, int32_t mode, float *ms);
void rt_batch_destroy(rt_batch *batch);
/* Times the persistent launch (reps runs each after a warm-up) over node
 * width (the items' and the other, built from the items' scenes) x
 * lds_stack x waves_per_cu x group_waves x refill, then grab for the best,
 * and returns the fastest parameters (use_graph = 1, width set). */
int32_t rt_batch_autotune(rt_ctx *ctx, const rt_batch_item *items, int32_t nitems, int32_t reps,
                          rt_batch_params *best, float *best_ms);

/* Incoherent rays leaving the scene's surfaces, on the device: a uniformly
 * chosen triangle, a uniform point on it, the origin offset as shadow-ray
 * origins are (rt/ray.h:111-157) along the triangle normal, and a
 * cosine-weighted direction about that normal (counter-based sampler keyed by
 * seed and ray index). rays: n device rt_ray. */
int32_t rt_gen_surface_rays(rt_scene *scene, uint64_t seed, int64_t n, rt_ray *rays);

/* ---- Differentiable frame (no counterpart in the reference, which renders
 * forward only: rt/raytracer.cc:20-87, 563-598). ----
 *
 * Replaces the scene's per-triangle albedo (n x 3 in [0, 1]); kd becomes
 * 255 * albedo as at creation. Device pointer with RT_DEVICE_PTRS. */
int32_t rt_scene_set_albedo(rt_scene *scene, const float *albedo, uint32_t flags);

/* A frame rendered for differentiation. rt_render_tape renders as rt_render
 * does (rgb: device, y_res x x_res x 3) and keeps what the adjoint needs:
 * every path record (hits, normals, shadow rays, visibility) when they fit in
 * max_record_bytes (<= 0: no limit); otherwise only the frame's inputs, and
 * the adjoint re-traces the frame chunk by chunk with each chunk's records
 * within max_record_bytes (path replay). The scene's albedo must not change
 * between the tape and its adjoint. */
typedef struct rt_tape rt_tape;
int32_t rt_render_tape(rt_accel *accel, const rt_camera *cam, const rt_light *lights, int32_t nlights,
                       const rt_render_params *params, float *rgb, int64_t max_record_bytes, rt_tape **tape,
                       rt_stats *stats);
/* replay: 1 if the adjoint re-traces; record_bytes: the whole frame's path
 * records; chunk_tiles: 8x8 tiles per chunk. Any output may be NULL. */
int32_t rt_tape_info(const rt_tape *tape, int32_t *replay, int64_t *record_bytes, int64_t *chunk_tiles);
/* Vector-Jacobian product of the taped frame. d_rgb (device, like rgb) is
 * dL/d(pixel radiance). Adds dL/d(albedo) into d_albedo (device, n x 3) and
 * dL/d(light) into d_lights (device, nlights x 6: position xyz, then rgb on
 * the 0-255 scale). Hit points and visibility are held at the forward's.
 * ms: device time of the adjoint (re-tracing included). */
int32_t rt_render_adjoint(rt_tape *tape, const float *d_rgb, float *d_albedo, float *d_lights, float *ms);
void rt_tape_destroy(rt_tape *tape);

/* ---- Batched inverse rendering (no counterpart in the reference). ----
 *
 * Many scenes, each rendered at spp samples per pixel with up to `bounces`
 * diffuse bounces and the common lights, in one persistent launch: a path
 * megakernel takes (scene, 8x8 tile) work items from a queue; each lane
 * traces its pixel's samples, and with a target image computes the gradient
 * of L = sum (I - target)^2 with respect to every triangle's albedo from the
 * path vertices it kept (nothing stored in memory, nothing re-traced). A step
 * of Adam (lr, beta1, beta2, eps; albedo clamped to [0, 1]) follows; one step
 * is captured as a hipGraph when use_graph is set. All buffers are device
 * memory: albedo (n x 3, the parameters, updated in place), d_albedo (n x 3,
 * gradient), target and image (y_res x x_res x 3). Items share a node width. */
typedef struct {
    rt_accel *accel;
    rt_camera cam;
    float *albedo;
    float *d_albedo;
    const float *target;
    float *image;
} rt_invrender_item;
typedef struct {
    int32_t spp;          /* 1..8 */
    int32_t bounces;      /* 0..3 */
    uint64_t seed;
    float lr, beta1, beta2, eps;
    int32_t use_graph;
    int32_t waves_per_cu; /* wavefronts per CU of the persistent launch (0: as many as fit) */
} rt_invrender_params;
typedef struct rt_invrender rt_invrender;
int32_t rt_invrender_create(rt_ctx *ctx, const rt_invrender_item *items, int32_t nitems, const rt_light *lights,
                            int32_t nlights, const rt_invrender_params *params, rt_invrender **out);
/* mode 0: render the images; 1: images and the albedo gradients (cleared
 * first, not applied); 2: `steps` Adam steps (image, gradient, update each).
 * plain != 0 launches once per scene instead of once for all. loss (nitems,
 * may be NULL): each scene's L at the last frame; rays: rays traced over the
 * call; ms: device time of the call. */
int32_t rt_invrender_run(rt_invrender *r, int32_t mode, int32_t steps, int32_t plain, double *loss, uint64_t *rays,
                         float *ms);
void rt_invrender_destroy(rt_invrender *r);

/* Count node visits and triangle tests in later calls (0/1). */
int32_t rt_set_counting(rt_ctx *ctx, int32_t on);

/* Scene generators (host only). Each returns the triangle count; with
 * abc == NULL only the count is computed.
 * Random triangles: centroids uniform in [0,1)^3, each vertex the centroid plus
 * a uniform offset in [-spread, spread)^3; albedo uniform in [lo, hi)^3.
 * Terrain: RAPTor/terrain_generator/height_map.h:16-43 over
 * common/perlin_noise_2d.cc:3-47 on a grid x grid vertex lattice spanning
 * [0,1] in x and z, two triangles per cell. */
int64_t rt_gen_random_triangles(uint64_t seed, int64_t n, float spread, float albedo_lo, float albedo_hi,
                                float *abc, float *albedo);
int64_t rt_gen_terrain(int32_t grid, float amplitude, float zoom, float persistence, int32_t octaves,
                       int32_t noise_seed, uint64_t albedo_seed, float albedo_lo, float albedo_hi, float *abc,
                       float *albedo);

#ifdef __cplusplus
}
#endif

#endif /* RAPTOR_RT_H */
