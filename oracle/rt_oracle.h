/* Oracle — TEST INFRASTRUCTURE ONLY.
 *
 * C entry points of the CPU restatement of RAPTor's ray tracer hot path, for
 * ctypes. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * load liboracle_rt.so; the product library never links it.
 */
#ifndef RT_ORACLE_H
#define RT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_scene oracle_scene;

/* Pinhole camera as rt/camera.h:52-62 stores it: position, basis, image-plane
 * half widths w and h at distance t, resolution. */
typedef struct {
    float pos[3];
    float x_axis[3], y_axis[3], z_axis[3];
    float w, h, t;
    int32_t x_res, y_res;
} oracle_camera;

/* Point light, rt/light.h:17-18: rgb in [0,255], falloff d. */
typedef struct {
    float pos[3];
    float rgb[3];
    float d;
} oracle_light;

/* A texture map's state (rt/materials/texture_mapper.h and the mappers'
 * constructors): kind 0 planar, 1 cubic, 2 cylindrical, 3 checker board,
 * 4 Perlin 3-D noise, 5 gradient over the incidence angle;
 * wrapping 0 blank, 1 clamp, 2 tile, 3 mirror; falloff 0 cubic, 1
 * spherical, 2-4 linear x/y/z, 5 none. Same layout as the library's
 * rt_texture_map. */
typedef struct {
    int32_t kind;
    int32_t image;
    float c[3], n[3], s[3];
    float u[3], v[3];
    int32_t u_wrap, v_wrap;
    int32_t u_off, v_off;
    int32_t cycles;
    float rgb[3], opacity;
    int32_t inverse;
    int32_t falloff;
    float falloff_centre[3], falloff_grad[3];
    float persistence, zoom; /* kind 4: perlin_noise_3d_mapper */
    int32_t octaves, seed;
    int32_t key_count, grad_of; /* kind 5: gradient_mapper, keys of 6 floats */
    int64_t key_offset;         /* {key, r, g, b, alpha, interpolation} in the texels */
} oracle_texture_map;

/* width x height texels of cpp floats at texels + offset, row 0 first. */
typedef struct {
    int32_t width, height, cpp, reserved;
    int64_t offset;
} oracle_texture_image;

/* Render options. bounce: number of cosine-weighted diffuse bounces, each
 * with direct light at its hit (radiance c = sum_k T_k direct(x_k), T_k the
 * product of kd / 255 over the earlier vertices). spp > 1 jitters the sample
 * inside the pixel. */
typedef struct {
    int32_t bounce;
    int32_t spp;
    uint64_t seed;
    int32_t pixel_stride; /* trace pixels p with p % stride == 0 (packet frame: 8x8 tiles) */
    int32_t nthreads;
    int32_t packet;       /* 1: the SSE packet frame (spp 1, bounce <= 1) */
    /* Materials: nmaterials x 40 floats ka rgb, kd rgb, ks rgb, s, tran,
     * ri, rf, kind, cx, cy, rd, sr, ri_i, six texture slots (1 + map, 0
     * none: t_ka, t_kd, t_ks, t_ns, t_refl, t_d of a coloured_mapper_shader,
     * kind 2), mapper kd, mapper ks, clrh, clrf, and five layer lists
     * (first x 5, count x 5) of a mapper_shader (kind 3, base colour in kd
     * rgb), the iteration limit of a mandelbrot_shader (kind 4), and one
     * index per triangle (-1:
     * phong_shader(255 * albedo, 1), the Lambert default). kind 0 is
     * phong_shader (rt/materials/phong_shader.h:34-53, coloured form; the
     * last five floats unused); kind 1 is cook_torrance_cxy
     * (rt/materials/cook_torrance_cxy.h:14-30) with chroma (cx, cy), diffuse
     * reflectance rd, roughness sr, imaginary index ri_i, and rf, tran, ri
     * holding its rs, ts, ri_r (ka, kd, ks, s unused). NULL materials: every
     * triangle takes the default. With materials the frame recurses into
     * mirror and refraction rays (rt/ray.cc:118-258) and needs bounce 0,
     * packet 0 and no records. */
    const float *materials;
    int32_t nmaterials;
    const int32_t *tri_material;
    /* Vertex normals, 9 floats per triangle (a, b, c); FLT_MAX as the first
     * keeps the geometric normal (rt/triangle.h:139-167). Shading then uses
     * u n_c + v n_b + (1 - (u + v)) n_a turned to face the ray
     * (rt/triangle.h:464-498). Non-NULL selects the material frame. */
    const float *vertex_normals;
    /* The camera's background colour: a primary, mirror or refraction ray
     * that hits nothing takes it (rt/camera.h:234-240). */
    float background[3];
    /* Texture maps, their images and texels, the mapper_shader layer lists
     * (map indices) and per-triangle texture coordinates (6 floats: u, v
     * at a, b, c; FLT_MAX first: none). Any may be NULL. */
    const oracle_texture_map *maps;
    const oracle_texture_image *images;
    const float *texels;
    const int32_t *layers;
    const float *texture_coords;
} oracle_render_params;

/* Outputs; any pointer may be NULL. Arrays are indexed by pixel p = y*W + x.
 * tuv are (t, u, v) triples, vis is one byte per light (1 = lit). */
typedef struct {
    float *rgb;
    int32_t *prim;
    float *tuv;
    uint8_t *vis;
    int32_t *bounce_prim;
    float *bounce_tuv;
    uint8_t *bounce_vis;
    /* [0] primary, [1] shadow, [2] bounce, [3] bounce-shadow rays,
     * [4] child box tests, [5] triangle tests */
    uint64_t *counters;
    /* Per-sample path records of rec_levels path vertices (0 = primary hit,
     * k = k-th bounce hit), level-major: index level * (pixels * spp) +
     * p * spp + s. prim -1: no surface. o: the shadow-ray origin (offset hit
     * point), n: the shading normal, vis: one byte per light (1 = lit).
     * rec_levels = 0 records nothing. */
    int32_t rec_levels;
    int32_t *rec_prim;
    float *rec_o, *rec_n;
    uint8_t *rec_vis;
    /* With materials: [0] reflected + refracted rays, [1] their shadow rays. */
    uint64_t *sec_counters;
} oracle_render_out;

/* One texture_map (mono 0: rgb in and out) or texture_map_monochrome (mono
 * 1: the value in rgb[0]) of maps[map] at hit point dst of a ray with
 * direction dir on a surface of shading normal n, with texture coordinates
 * (vtu, vtv) when has_vt; returns the alpha. */
float oracle_texture_sample(const oracle_texture_map *maps, const oracle_texture_image *images, const float *texels,
                            int32_t map, const float *dst, const float *dir, const float *n, int32_t has_vt,
                            float vtu, float vtv, int32_t mono, float *rgb);
/* perlin_noise_3d of seed: lattice 1 gives noise(int x, int y, int z),
 * lattice 0 interpolated_noise(x, y, z) (common/perlin_noise_3d.cc). */
float oracle_perlin_noise(float x, float y, float z, int32_t seed, int32_t lattice);
/* mandelbrot_shader's colour at hit point dst with shading normal n, before
 * the light term (rt/materials/mandelbrot_shader.cc:41-123). */
void oracle_mandelbrot_colour(const float *dst, const float *n, int32_t iterations, float *rgb);
oracle_scene *oracle_scene_create(const float *abc, int64_t n, const uint8_t *flags);
void oracle_scene_destroy(oracle_scene *s);
/* AAC build (alpha, delta as ssd/bvh_builder.h:23). Returns the root. */
int32_t oracle_scene_build(oracle_scene *s, float alpha, float delta);
int32_t oracle_scene_node_count(const oracle_scene *s);
/* The AAC tree as build_aac numbers it (root = count - 1): per node the box
 * (lo xyz, hi xyz), child[2] (interior: left and right node; leaf: the
 * indirect range [b, e)), leaf flag; and the indirect order. Returns the
 * node count. */
int32_t oracle_scene_tree(const oracle_scene *s, float *lohi, int32_t *child, uint8_t *leaf, int32_t *indirect);
/* The AAC builder's Morton code of each triangle (emplace order) and its
 * radix sort's order. */
void oracle_aac_morton(const oracle_scene *s, int32_t *codes, int32_t *order);
/* Checks BVH invariants; returns 0 when they hold. stats: [leaves, interior,
 * max depth, max leaf size]. */
int32_t oracle_bvh_check(const oracle_scene *s, int32_t *stats);
void oracle_scene_normals(const oracle_scene *s, float *n3);

/* rays: n x 8 floats (o.xyz, tmin, d.xyz, tmax); tmin is EPSILON as the
 * reference fixes it. Closest hit ignores tmax (the reference starts from
 * MAX_DIST); any hit uses tmax as the shadow-ray length. use_bvh = 0 runs the
 * brute-force checker instead of the AAC BVH. */
void oracle_trace_closest(const oracle_scene *s, const float *rays, int64_t n, int32_t *prim, float *tuv,
                          int32_t use_bvh, int32_t nthreads, uint64_t *counters);
void oracle_trace_occluded(const oracle_scene *s, const float *rays, int64_t n, uint8_t *occluded,
                           int32_t use_bvh, int32_t nthreads, uint64_t *counters);

int32_t oracle_render(const oracle_scene *s, const float *albedo, const oracle_camera *cam,
                      const oracle_light *lights, int32_t nlights, const oracle_render_params *p,
                      oracle_render_out *out);

/* Known-answer hooks. */
int32_t oracle_min_element_n(const float *d, float *m, int32_t n);
int32_t oracle_min_element_range(const float *d, float *m, int32_t begin, int32_t end);
int32_t oracle_morton(float x, float y, float z, float xm, float ym, float zm);
void oracle_morton_sort(const int32_t *codes, int64_t n, int32_t *sorted, int32_t *order);
float oracle_slab(const float *low, const float *high, const float *ogn, const float *i_rd);
float oracle_combined_sa(const float *lo1, const float *hi1, const float *lo2, const float *hi2);
/* Scalar Moller-Trumbore on one triangle; returns 1 and writes t,u,v on a hit. */
int32_t oracle_intersect(const float *abc, const float *ogn, const float *dir, float *tuv);
void oracle_offset_start_point(const float *dst, const float *geo_norm, int32_t d, float *out);
/* Counter-based bounce sampling, for checking the HIP version. */
void oracle_bounce_dir(uint64_t seed, uint32_t pixel, uint32_t sample, uint32_t bounce, const float *n, float *dir);
void oracle_sincos_turns(float r, float *s, float *c);
uint64_t oracle_pcg32_stream(uint64_t initstate, uint64_t initseq, int32_t count, uint32_t *out);
uint64_t oracle_splitmix64_stream(uint64_t seed, int32_t count, uint64_t *out);
/* SSE packet path (rt/triangle.h:242-305, ssd/bvh_node.h:101-132,
 * ssd/bvh.cc:341-500): slab test of one box for four rays (ogn4, i_rd4 as
 * x[4], y[4], z[4]); rcp / rsqrt with one Newton step; packets of four
 * consecutive rays through the packet traversal or packet brute force. */
void oracle_packet_slab(const float *low, const float *high, const float *ogn4, const float *i_rd4, float *out4);
void oracle_packet_inverse(const float *x4, float *out4);
void oracle_packet_inverse_sqrt(const float *x4, float *out4);
void oracle_packet_trace_closest(const oracle_scene *s, const float *rays, int64_t n, int32_t *prim, float *tuv,
                                 int32_t use_bvh);
void oracle_packet_trace_occluded(const oracle_scene *s, const float *rays, int64_t n, uint8_t *occ, int32_t use_bvh);
/* Secondary-ray directions of rt/ray.cc:121-131 (reflect) and :200-227
 * (refract; n faces the ray, leaving = the ray exits through the back face;
 * returns 0 on total internal reflection). */
void oracle_reflect_dir(const float *d, const float *n, float *out);
int32_t oracle_refract_dir(const float *d, const float *n, float ri, int32_t leaving, float *out);
/* Primary ray of rt/camera.h:127-158 (EXACT_NORMALISE). */
void oracle_primary_ray(const oracle_camera *cam, int32_t x, int32_t y, float *ogn, float *dir);

/* Camera post-process on an image of 0-255 floats, pixel (x, y) at
 * 3 * (x + y * x_res) (postprocess.cpp). tone_map: camera::tone_map
 * (rt/camera.cc:1124-1461) for mode 0 (none) and the global modes 1, 4, 5,
 * 6, 7, 9 of tone_mapping_mode_t (rt/camera.h:43-45), the local modes 2, 3
 * and 8, and the glare filter (glare, which needs x_angle). adaptation_level is the camera's
 * adaptation state (read with adapt, always written); yi_out the adapted
 * log-average luminance. gamma: camera::gamma_correct
 * (rt/camera.cc:1466-1501). clip: clip_image_to_rgb / _bgr
 * (rt/camera.h:289-351) of an x_res x y_res image with xs x ys samples per
 * output pixel. */
typedef struct {
    int32_t mode;
    float key, xw, yw;
    int32_t desaturate, stretch, adapt, glare;
    float time_step;
    /* The camera's field of view, 2 w / t and 2 h / t in radians (modes 2, 3). */
    float x_angle, y_angle;
} oracle_tone_params;
int32_t oracle_tone_map(float *img, int32_t x_res, int32_t y_res, const oracle_tone_params *tp,
                        float *adaptation_level, float *yi_out);
void oracle_gamma(float *img, int64_t npix, float gamma);
/* The bilateral grid's circular Gaussian convolution (rt/camera.cc:1697-1734)
 * of both channels of a W x H x D grid, in place. */
void oracle_gauss_circular3d(float *grid, int32_t W, int32_t H, int32_t D, float ss, float sr);
void oracle_clip(const float *img, int32_t x_res, int32_t y_res, int32_t xs, int32_t ys, int32_t bgr, uint8_t *out);

/* The glare filter (glare.cpp): the reference's MT19937 stream (raw tempered
 * words), the filter of an x_res x y_res camera with field of view x_angle
 * = 2 w / t at light level 0 scotopic, 1 mesopic, 2 photopic (x_res * y_res
 * rgb floats), and perform_glare_filter on a tone-mapped image at adapted
 * luminance yi (level_out: the light level used). tone_map applies it when
 * glare is set. */
void oracle_mt_stream(int32_t count, uint32_t *out);
void oracle_glare_filter(int32_t x_res, int32_t y_res, float x_angle, int32_t level, float *out);
int32_t oracle_glare(float *img, int32_t x_res, int32_t y_res, float yi, float x_angle, int32_t *level_out);

#ifdef __cplusplus
}
#endif

#endif /* RT_ORACLE_H */
