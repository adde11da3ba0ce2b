"""Oracle — TEST INFRASTRUCTURE ONLY.

ctypes binding of liboracle_rt.so, the CPU restatement of RAPTor's ray tracer
hot path (see vec.hpp). Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_rt.so")

HIT_DTYPE = np.dtype([("t", "<f4"), ("u", "<f4"), ("v", "<f4"), ("prim", "<i4")])


class oracle_camera(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("x_axis", C.c_float * 3), ("y_axis", C.c_float * 3),
                ("z_axis", C.c_float * 3), ("w", C.c_float), ("h", C.c_float), ("t", C.c_float),
                ("x_res", C.c_int32), ("y_res", C.c_int32)]


class oracle_light(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("rgb", C.c_float * 3), ("d", C.c_float)]


class oracle_render_params(C.Structure):
    _fields_ = [("bounce", C.c_int32), ("spp", C.c_int32), ("seed", C.c_uint64), ("pixel_stride", C.c_int32),
                ("nthreads", C.c_int32), ("packet", C.c_int32), ("materials", C.c_void_p),
                ("nmaterials", C.c_int32), ("tri_material", C.c_void_p), ("vertex_normals", C.c_void_p),
                ("background", C.c_float * 3), ("maps", C.c_void_p), ("images", C.c_void_p),
                ("texels", C.c_void_p), ("layers", C.c_void_p), ("texture_coords", C.c_void_p)]


# oracle_texture_map and oracle_texture_image (rt_oracle.h).
MAP_DTYPE = np.dtype([("kind", "<i4"), ("image", "<i4"), ("c", "<f4", 3), ("n", "<f4", 3), ("s", "<f4", 3),
                      ("u", "<f4", 3), ("v", "<f4", 3), ("u_wrap", "<i4"), ("v_wrap", "<i4"), ("u_off", "<i4"),
                      ("v_off", "<i4"), ("cycles", "<i4"), ("rgb", "<f4", 3), ("opacity", "<f4"),
                      ("inverse", "<i4"), ("falloff", "<i4"), ("falloff_centre", "<f4", 3),
                      ("falloff_grad", "<f4", 3), ("persistence", "<f4"), ("zoom", "<f4"), ("octaves", "<i4"),
                      ("seed", "<i4"), ("key_count", "<i4"), ("grad_of", "<i4"), ("key_offset", "<i8")])
IMAGE_DTYPE = np.dtype([("width", "<i4"), ("height", "<i4"), ("cpp", "<i4"), ("reserved", "<i4"),
                        ("offset", "<i8")])


class oracle_render_out(C.Structure):
    _fields_ = [("rgb", C.c_void_p), ("prim", C.c_void_p), ("tuv", C.c_void_p), ("vis", C.c_void_p),
                ("bounce_prim", C.c_void_p), ("bounce_tuv", C.c_void_p), ("bounce_vis", C.c_void_p),
                ("counters", C.c_void_p), ("rec_levels", C.c_int32), ("rec_prim", C.c_void_p),
                ("rec_o", C.c_void_p), ("rec_n", C.c_void_p), ("rec_vis", C.c_void_p),
                ("sec_counters", C.c_void_p)]


# Phong material fields, in the order the oracle (and include/raptor_rt.h's
# rt_material) stores them: rt/materials/phong_shader.h:38-60.
class oracle_tone_params(C.Structure):
    _fields_ = [("mode", C.c_int32), ("key", C.c_float), ("xw", C.c_float), ("yw", C.c_float),
                ("desaturate", C.c_int32), ("stretch", C.c_int32), ("adapt", C.c_int32), ("glare", C.c_int32),
                ("time_step", C.c_float), ("x_angle", C.c_float), ("y_angle", C.c_float)]


def material_array(materials):
    """materials: list of dicts. Phong (kind 0, the default): ka, kd, ks (rgb
    in [0, 255]), s, tran, ri, rf. Cook-Torrance (kind 1, cook_torrance_cxy):
    x, y (chroma), rd, rs, ts, sr, ri_r, ri_i. Coloured mapper (kind 2): the
    phong fields and tex (map indices or None for t_ka, t_kd, t_ks, t_ns,
    t_refl, t_d). Mapper shader (kind 3): rgb, kd, ks (scalars), s, tran,
    ri, rf, clrh, clrf, layers ((first, count) x 5). Mandelbrot shader
    (kind 4): iterations. Returns an (n, 40) float32 array in the oracle's
    row layout."""
    rows = []
    for m in materials:
        kind = m.get("kind", 0)
        tail = [0.0] * 21
        if kind == 1:
            row = [0.0] * 10 + [m.get("ts", 0.0), m.get("ri_r", 1.0), m.get("rs", 0.0), 1.0,
                                m.get("x", 0.33), m.get("y", 0.33), m.get("rd", 0.5), m.get("sr", 0.0),
                                m.get("ri_i", 0.0)]
        elif kind == 4:
            row = [0.0] * 13 + [4.0, 0.33, 0.33, 0.5, 0.0, 0.0]
            tail[20] = float(m.get("iterations", 100))
        elif kind == 3:
            row = [0.0] * 3 + list(m.get("rgb", (0.0, 0.0, 0.0))) + [0.0] * 3 + \
                [m.get("s", 0.0), m.get("tran", 0.0), m.get("ri", 1.0), m.get("rf", 0.0)] + \
                [3.0, 0.33, 0.33, 0.5, 0.0, 0.0]
            tail[6:10] = [m.get("kd", 0.0), m.get("ks", 0.0), m.get("clrh", 0.0), m.get("clrf", 0.0)]
            for k, (first, count) in enumerate(m.get("layers", [(0, 0)] * 5)):
                tail[10 + k], tail[15 + k] = float(first), float(count)
        else:
            row = list(m.get("ka", (0.0, 0.0, 0.0))) + list(m.get("kd", (0.0, 0.0, 0.0))) + \
                list(m.get("ks", (0.0, 0.0, 0.0))) + \
                [m.get("s", 0.0), m.get("tran", 0.0), m.get("ri", 1.0), m.get("rf", 0.0)] + \
                [float(kind), 0.33, 0.33, 0.5, 0.0, 0.0]
            for k, t in enumerate(m.get("tex", ())):
                tail[k] = 0.0 if t is None or t < 0 else float(t + 1)
        rows.append(row + tail)
    return np.asarray(rows, np.float32).reshape(-1, 40)


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _load():
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    vp, i32, i64, u32, u64, f = C.c_void_p, C.c_int32, C.c_int64, C.c_uint32, C.c_uint64, C.c_float
    sig = {
        "oracle_scene_create": (vp, [vp, i64, vp]),
        "oracle_scene_destroy": (None, [vp]),
        "oracle_scene_build": (i32, [vp, f, f]),
        "oracle_scene_node_count": (i32, [vp]),
        "oracle_scene_tree": (i32, [vp, vp, vp, vp, vp]),
        "oracle_aac_morton": (None, [vp, vp, vp]),
        "oracle_bvh_check": (i32, [vp, vp]),
        "oracle_scene_normals": (None, [vp, vp]),
        "oracle_trace_closest": (None, [vp, vp, i64, vp, vp, i32, i32, vp]),
        "oracle_trace_occluded": (None, [vp, vp, i64, vp, i32, i32, vp]),
        "oracle_render": (i32, [vp, vp, C.POINTER(oracle_camera), C.POINTER(oracle_light), i32,
                                C.POINTER(oracle_render_params), C.POINTER(oracle_render_out)]),
        "oracle_reflect_dir": (None, [vp, vp, vp]),
        "oracle_refract_dir": (i32, [vp, vp, f, i32, vp]),
        "oracle_min_element_n": (i32, [vp, vp, i32]),
        "oracle_min_element_range": (i32, [vp, vp, i32, i32]),
        "oracle_morton": (i32, [f, f, f, f, f, f]),
        "oracle_morton_sort": (None, [vp, i64, vp, vp]),
        "oracle_slab": (f, [vp, vp, vp, vp]),
        "oracle_combined_sa": (f, [vp, vp, vp, vp]),
        "oracle_intersect": (i32, [vp, vp, vp, vp]),
        "oracle_offset_start_point": (None, [vp, vp, i32, vp]),
        "oracle_bounce_dir": (None, [u64, u32, u32, u32, vp, vp]),
        "oracle_sincos_turns": (None, [f, vp, vp]),
        "oracle_pcg32_stream": (u64, [u64, u64, i32, vp]),
        "oracle_splitmix64_stream": (u64, [u64, i32, vp]),
        "oracle_primary_ray": (None, [C.POINTER(oracle_camera), i32, i32, vp, vp]),
        "oracle_packet_slab": (None, [vp, vp, vp, vp, vp]),
        "oracle_packet_inverse": (None, [vp, vp]),
        "oracle_packet_inverse_sqrt": (None, [vp, vp]),
        "oracle_packet_trace_closest": (None, [vp, vp, i64, vp, vp, i32]),
        "oracle_packet_trace_occluded": (None, [vp, vp, i64, vp, i32]),
        "oracle_tone_map": (i32, [vp, i32, i32, C.POINTER(oracle_tone_params), vp, vp]),
        "oracle_gamma": (None, [vp, i64, f]),
        "oracle_gauss_circular3d": (None, [vp, i32, i32, i32, f, f]),
        "oracle_clip": (None, [vp, i32, i32, i32, i32, i32, vp]),
        "oracle_mt_stream": (None, [i32, vp]),
        "oracle_texture_sample": (f, [vp, vp, vp, i32, vp, vp, vp, i32, f, f, i32, vp]),
        "oracle_perlin_noise": (f, [f, f, f, i32, i32]),
        "oracle_mandelbrot_colour": (None, [vp, vp, i32, vp]),
        "oracle_glare_filter": (None, [i32, i32, f, i32, vp]),
        "oracle_glare": (i32, [vp, i32, i32, f, f, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


LIB = _load()


def _p(a):
    return None if a is None else a.ctypes.data


def f32a(*v):
    return np.asarray(v, np.float32)


def camera_c(cam):
    c = oracle_camera()
    for k in range(3):
        c.pos[k], c.x_axis[k], c.y_axis[k], c.z_axis[k] = cam.pos[k], cam.x_axis[k], cam.y_axis[k], cam.z_axis[k]
    c.w, c.h, c.t, c.x_res, c.y_res = cam.w, cam.h, cam.t, cam.x_res, cam.y_res
    return c


def lights_c(lights):
    arr = (oracle_light * max(1, len(lights)))()
    for i, l in enumerate(lights):
        for k in range(3):
            arr[i].pos[k] = l.pos[k]
            arr[i].rgb[k] = l.rgb[k]
        arr[i].d = l.d
    return arr


class OracleScene:
    """A triangle store with the reference's AAC BVH (built on construction)."""

    def __init__(self, abc, flags=None, build=True, alpha=0.4, delta=20.0):
        self.abc = np.ascontiguousarray(abc, np.float32).reshape(-1, 9)
        self.n = self.abc.shape[0]
        self.flags = None if flags is None else np.ascontiguousarray(flags, np.uint8)
        self.h = LIB.oracle_scene_create(_p(self.abc), self.n, _p(self.flags))
        self.root = LIB.oracle_scene_build(self.h, alpha, delta) if build else -1

    def bvh_check(self):
        st = np.zeros(4, np.int32)
        err = LIB.oracle_bvh_check(self.h, _p(st))
        return err, dict(leaves=int(st[0]), interior=int(st[1]), depth=int(st[2]), max_leaf=int(st[3]))

    def normals(self):
        out = np.empty((self.n, 3), np.float32)
        LIB.oracle_scene_normals(self.h, _p(out))
        return out

    def tree(self):
        """The AAC tree: dict(lohi (m, 6), child (m, 2), leaf (m,) bool,
        indirect (n,), root) with nodes numbered as the builder made them."""
        m = max(1, 2 * self.n)
        lohi = np.empty((m, 6), np.float32)
        child = np.empty((m, 2), np.int32)
        leaf = np.empty(m, np.uint8)
        ind = np.empty(max(1, self.n), np.int32)
        cnt = LIB.oracle_scene_tree(self.h, _p(lohi), _p(child), _p(leaf), _p(ind))
        return dict(lohi=lohi[:cnt], child=child[:cnt], leaf=leaf[:cnt].astype(bool), indirect=ind[:self.n],
                    root=cnt - 1)

    def aac_morton(self):
        """The AAC builder's Morton codes (emplace order) and sort order."""
        codes = np.empty(max(1, self.n), np.int32)
        order = np.empty(max(1, self.n), np.int32)
        LIB.oracle_aac_morton(self.h, _p(codes), _p(order))
        return codes[:self.n], order[:self.n]

    def trace_closest(self, rays, use_bvh=True, nthreads=8):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = rays.shape[0]
        prim = np.empty(n, np.int32)
        tuv = np.empty((n, 3), np.float32)
        cnt = np.zeros(2, np.uint64)
        LIB.oracle_trace_closest(self.h, _p(rays), n, _p(prim), _p(tuv), 1 if use_bvh else 0, nthreads, _p(cnt))
        out = np.empty(n, HIT_DTYPE)
        out["t"], out["u"], out["v"], out["prim"] = tuv[:, 0], tuv[:, 1], tuv[:, 2], prim
        return out, cnt

    def packet_trace_closest(self, rays, use_bvh=True):
        """The SSE packet path on packets of four consecutive rays (n % 4 == 0)."""
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = rays.shape[0] - rays.shape[0] % 4
        prim = np.full(rays.shape[0], -2, np.int32)
        tuv = np.zeros((rays.shape[0], 3), np.float32)
        LIB.oracle_packet_trace_closest(self.h, _p(rays), n, _p(prim), _p(tuv), 1 if use_bvh else 0)
        out = np.empty(rays.shape[0], HIT_DTYPE)
        out["t"], out["u"], out["v"], out["prim"] = tuv[:, 0], tuv[:, 1], tuv[:, 2], prim
        return out

    def packet_trace_occluded(self, rays, use_bvh=True):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = rays.shape[0] - rays.shape[0] % 4
        occ = np.zeros(rays.shape[0], np.uint8)
        LIB.oracle_packet_trace_occluded(self.h, _p(rays), n, _p(occ), 1 if use_bvh else 0)
        return occ

    def trace_occluded(self, rays, use_bvh=True, nthreads=8):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = rays.shape[0]
        occ = np.empty(n, np.uint8)
        cnt = np.zeros(2, np.uint64)
        LIB.oracle_trace_occluded(self.h, _p(rays), n, _p(occ), 1 if use_bvh else 0, nthreads, _p(cnt))
        return occ, cnt

    def render(self, albedo, camera, lights, bounce=False, spp=1, seed=0, pixel_stride=1, nthreads=8, packet=False,
               records=False, materials=None, tri_material=None, vertex_normals=None, background=(0.0, 0.0, 0.0),
               textures=None, texture_coords=None):
        """One frame; bounce is the number of diffuse bounces (True = 1).
        records=True adds per-sample path records over the bounce + 1 path
        vertices, level-major (vertex k, sample p * spp + s): prim (L, S)
        (-1: none), o (L, S, 3) shadow-ray origins, n (L, S, 3) shading
        normals, vis (L, S, nlights) 1 = lit. materials (list of dicts, see
        material_array) with tri_material (one index per triangle, -1: the
        Lambert albedo default) renders phong materials with mirror and
        refraction rays; counters then gain sec_counters (secondary rays,
        their shadow rays). vertex_normals (n, 9) float32 (FLT_MAX first:
        none) shade with interpolated normals and select that frame too.
        textures = (maps, images, texels, layers): MAP_DTYPE and IMAGE_DTYPE
        arrays (the library's rt_texture_map / rt_texture_image layout),
        float32 texels, int32 layer lists (or None); texture_coords (n, 6)
        float32 (FLT_MAX first: none)."""
        npix = camera.x_res * camera.y_res
        nl = len(lights)
        rgb = np.full((npix, 3), np.nan, np.float32)
        prim = np.full(npix, -2, np.int32)
        tuv = np.full((npix, 3), np.nan, np.float32)
        vis = np.full((npix, max(1, nl)), 255, np.uint8)
        bprim = np.full(npix, -2, np.int32)
        btuv = np.full((npix, 3), np.nan, np.float32)
        bvis = np.full((npix, max(1, nl)), 255, np.uint8)
        cnt = np.zeros(6, np.uint64)
        out = oracle_render_out(_p(rgb), _p(prim), _p(tuv), _p(vis), _p(bprim), _p(btuv), _p(bvis), _p(cnt))
        rec = None
        nb = int(bounce)
        if records:
            ns = npix * max(1, spp)
            lv = nb + 1
            rec = dict(prim=np.full((lv, ns), -2, np.int32), o=np.zeros((lv, ns, 3), np.float32),
                       n=np.zeros((lv, ns, 3), np.float32), vis=np.zeros((lv, ns, max(1, nl)), np.uint8))
            out.rec_levels = lv
            out.rec_prim, out.rec_o, out.rec_n, out.rec_vis = (_p(rec["prim"]), _p(rec["o"]), _p(rec["n"]),
                                                               _p(rec["vis"]))
        p = oracle_render_params(nb, spp, seed, pixel_stride, nthreads, 1 if packet else 0)
        p.background = (C.c_float * 3)(*background)
        sec = np.zeros(2, np.uint64)
        out.sec_counters = _p(sec)
        if materials is not None:
            if isinstance(materials, np.ndarray):
                mats = np.asarray(materials, np.float32)
                mats = mats.reshape(-1, mats.shape[-1] if mats.ndim > 1 else mats.size)
                if mats.shape[1] == 19:  # rows without textures: no texture slots, no layers
                    mats = np.concatenate([mats, np.zeros((mats.shape[0], 20), np.float32)], axis=1)
                mats = np.ascontiguousarray(mats).reshape(-1, 40)
            else:
                mats = material_array(materials)
            tm = np.ascontiguousarray(tri_material, np.int32)
            p.materials, p.nmaterials, p.tri_material = _p(mats), mats.shape[0], _p(tm)
        if vertex_normals is not None:
            vn = np.ascontiguousarray(vertex_normals, np.float32).reshape(-1, 9)
            p.vertex_normals = _p(vn)
        keep = []
        if textures is not None:
            maps, images, texels, layers = textures
            maps = np.ascontiguousarray(maps)
            images = np.ascontiguousarray(images)
            if maps.dtype.itemsize != MAP_DTYPE.itemsize or (images.size and images.dtype.itemsize != 24):
                raise ValueError("texture maps and images must have the rt_texture_map / rt_texture_image layout")
            texels = np.ascontiguousarray(texels, np.float32)
            layers = np.ascontiguousarray(layers if layers is not None else np.zeros(1, np.int32), np.int32)
            keep += [maps, images, texels, layers]
            p.maps, p.images, p.texels, p.layers = _p(maps), _p(images), _p(texels), _p(layers)
        if texture_coords is not None:
            vt = np.ascontiguousarray(texture_coords, np.float32).reshape(-1, 6)
            keep.append(vt)
            p.texture_coords = _p(vt)
        alb = np.ascontiguousarray(albedo, np.float32)
        rc = LIB.oracle_render(self.h, _p(alb), C.byref(camera_c(camera)), lights_c(lights), nl, C.byref(p),
                               C.byref(out))
        if rc != 0:
            raise RuntimeError(f"oracle_render failed ({rc})")
        res = dict(rgb=rgb, prim=prim, tuv=tuv, vis=vis, bounce_prim=bprim, bounce_tuv=btuv, bounce_vis=bvis,
                   counters=cnt, sec_counters=sec)
        if rec is not None:
            res["records"] = rec
        return res

    def close(self):
        if self.h:
            LIB.oracle_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# Arrays are bound to names before their pointers are taken so that they stay
# alive for the duration of each call.

def intersect(abc9, ogn, d):
    tuv = np.empty(3, np.float32)
    a, o, dd = np.asarray(abc9, np.float32), np.asarray(ogn, np.float32), np.asarray(d, np.float32)
    hit = LIB.oracle_intersect(_p(a), _p(o), _p(dd), _p(tuv))
    return bool(hit), tuv


def packet_slab(low, high, ogn4, i_rd4):
    """ogn4, i_rd4: (4, 3) per-ray origins and inverse directions."""
    lo, hi = f32a(*low), f32a(*high)
    o = np.ascontiguousarray(np.asarray(ogn4, np.float32).T)
    i = np.ascontiguousarray(np.asarray(i_rd4, np.float32).T)
    out = np.empty(4, np.float32)
    LIB.oracle_packet_slab(_p(lo), _p(hi), _p(o), _p(i), _p(out))
    return out


def packet_inverse(x4, sqrt=False):
    x = np.ascontiguousarray(x4, np.float32)
    out = np.empty(4, np.float32)
    (LIB.oracle_packet_inverse_sqrt if sqrt else LIB.oracle_packet_inverse)(_p(x), _p(out))
    return out


def slab(low, high, ogn, i_rd):
    a, b, c, d = f32a(*low), f32a(*high), f32a(*ogn), f32a(*i_rd)
    return LIB.oracle_slab(_p(a), _p(b), _p(c), _p(d))


def combined_sa(lo1, hi1, lo2, hi2):
    a, b, c, d = f32a(*lo1), f32a(*hi1), f32a(*lo2), f32a(*hi2)
    return LIB.oracle_combined_sa(_p(a), _p(b), _p(c), _p(d))


def min_element_n(d):
    d = np.asarray(d, np.float32)
    m = np.zeros(1, np.float32)
    i = LIB.oracle_min_element_n(_p(d), _p(m), d.size)
    return i, float(m[0])


def min_element_range(d, begin, end):
    d = np.asarray(d, np.float32)
    m = np.zeros(1, np.float32)
    i = LIB.oracle_min_element_range(_p(d), _p(m), begin, end)
    return i, float(m[0])


def morton(x, y, z, xm, ym, zm):
    return LIB.oracle_morton(x, y, z, xm, ym, zm)


def morton_sort(codes):
    codes = np.ascontiguousarray(codes, np.int32)
    s = np.empty_like(codes)
    o = np.empty_like(codes)
    LIB.oracle_morton_sort(_p(codes), codes.size, _p(s), _p(o))
    return s, o


def offset_start_point(dst, geo, d=1):
    out = np.empty(3, np.float32)
    a, b = f32a(*dst), f32a(*geo)
    LIB.oracle_offset_start_point(_p(a), _p(b), d, _p(out))
    return out


def bounce_dir(seed, pixel, sample, bounce, n):
    out = np.empty(3, np.float32)
    nn = np.asarray(n, np.float32)
    LIB.oracle_bounce_dir(seed, pixel, sample, bounce, _p(nn), _p(out))
    return out


def sincos_turns(r):
    s = np.zeros(1, np.float32)
    c = np.zeros(1, np.float32)
    LIB.oracle_sincos_turns(r, _p(s), _p(c))
    return float(s[0]), float(c[0])


def pcg32_stream(initstate, initseq, count):
    out = np.empty(count, np.uint32)
    LIB.oracle_pcg32_stream(initstate, initseq, count, _p(out))
    return out


def splitmix64_stream(seed, count):
    out = np.empty(count, np.uint64)
    LIB.oracle_splitmix64_stream(seed, count, _p(out))
    return out


def primary_rays(camera):
    """All primary rays of a camera as (n, 8) float32 (o, tmin=EPSILON, d, tmax=MAX_DIST)."""
    n = camera.x_res * camera.y_res
    rays = np.empty((n, 8), np.float32)
    c = camera_c(camera)
    o = np.empty(3, np.float32)
    d = np.empty(3, np.float32)
    for y in range(camera.y_res):
        for x in range(camera.x_res):
            LIB.oracle_primary_ray(C.byref(c), x, y, _p(o), _p(d))
            i = y * camera.x_res + x
            rays[i, 0:3] = o
            rays[i, 3] = np.float32(1e-3)
            rays[i, 4:7] = d
            rays[i, 7] = np.finfo(np.float32).max
    return rays


def reflect_dir(d, n):
    """rt/ray.cc:121-131 direction of the mirror ray (float32)."""
    d, n = f32a(*d), f32a(*n)
    out = np.zeros(3, np.float32)
    LIB.oracle_reflect_dir(_p(d), _p(n), _p(out))
    return out


def refract_dir(d, n, ri, leaving):
    """rt/ray.cc:200-227 direction of the refracted ray, or None on total
    internal reflection (n faces the ray)."""
    d, n = f32a(*d), f32a(*n)
    out = np.zeros(3, np.float32)
    ok = LIB.oracle_refract_dir(_p(d), _p(n), float(ri), 1 if leaving else 0, _p(out))
    return out if ok else None


def tone_map(rgb, mode=0, key=0.18, xw=1.0 / 3.0, yw=1.0 / 3.0, desaturate=False, stretch=False, adapt=False,
             time_step=0.0, adaptation_level=0.0, x_angle=0.0, y_angle=0.0, glare=False):
    """camera::tone_map on a copy of rgb ((H, W, 3) float32, 0-255); returns
    (image, adaptation level, adapted log-average luminance). glare needs
    x_angle (2 w / t)."""
    img = np.array(rgb, np.float32, order="C", copy=True)
    h, w = img.shape[0], img.shape[1]
    tp = oracle_tone_params(int(mode), key, xw, yw, int(desaturate), int(stretch), int(adapt), int(glare), time_step,
                            x_angle, y_angle)
    lvl = np.array([adaptation_level], np.float32)
    yi = np.zeros(1, np.float32)
    if LIB.oracle_tone_map(_p(img), w, h, C.byref(tp), _p(lvl), _p(yi)) != 0:
        raise ValueError(f"tone mapping mode {mode} is not restated")
    return img, float(lvl[0]), float(yi[0])


def texture_sample(maps, images, texels, index, dst, vt=None, mono=False, rgb=(0.0, 0.0, 0.0), dir=(0.0, 0.0, 1.0),
                   n=(0.0, 0.0, -1.0)):
    """texture_map (rgb in, rgb out) or texture_map_monochrome (the value) of
    maps[index] at hit point dst of a ray of direction dir on a surface of
    shading normal n, through texture coordinates vt = (u, v) when given.
    Returns (colour or value, alpha)."""
    maps, images = np.ascontiguousarray(maps), np.ascontiguousarray(images)
    texels = np.ascontiguousarray(texels, np.float32)
    d, rd, nn = f32a(*dst), f32a(*dir), f32a(*n)
    c = np.asarray(rgb, np.float32).copy()
    a = LIB.oracle_texture_sample(_p(maps), _p(images) if images.size else None, _p(texels) if texels.size else None,
                                  index, _p(d), _p(rd), _p(nn), 0 if vt is None else 1,
                                  0.0 if vt is None else float(vt[0]),
                                  0.0 if vt is None else float(vt[1]), 1 if mono else 0, _p(c))
    return (float(c[0]) if mono else c), float(a)


def perlin_noise(x, y, z, seed, lattice=False):
    """perlin_noise_3d::noise at integer (x, y, z) (lattice) or
    interpolated_noise at (x, y, z)."""
    return float(LIB.oracle_perlin_noise(float(x), float(y), float(z), int(seed), 1 if lattice else 0))


def mandelbrot_colour(dst, n, iterations):
    """mandelbrot_shader's colour at hit point dst, shading normal n, before
    the light term."""
    out = np.zeros(3, np.float32)
    d, nn = f32a(*dst), f32a(*n)
    LIB.oracle_mandelbrot_colour(_p(d), _p(nn), int(iterations), _p(out))
    return out


def mt_stream(count):
    """The reference's random generator (common/common.cc:30-105): the first
    count tempered 32-bit words after its seeding with 5489."""
    out = np.empty(count, np.uint32)
    LIB.oracle_mt_stream(count, _p(out))
    return out


def glare_filter(x_res, y_res, x_angle, level):
    """The glare filter (y_res, x_res, 3) float32 of a camera with field of
    view x_angle = 2 w / t at light level 0 scotopic, 1 mesopic, 2 photopic."""
    out = np.empty((y_res, x_res, 3), np.float32)
    LIB.oracle_glare_filter(x_res, y_res, x_angle, level, _p(out))
    return out


def glare(rgb, yi, x_angle):
    """perform_glare_filter on a copy of a tone-mapped image; returns (image, light level)."""
    img = np.array(rgb, np.float32, order="C", copy=True)
    lvl = np.zeros(1, np.int32)
    if LIB.oracle_glare(_p(img), img.shape[1], img.shape[0], yi, x_angle, _p(lvl)) != 0:
        raise ValueError("glare needs a positive x_angle")
    return img, int(lvl[0])


def gauss_circular3d(grid, ss=1.0, sr=1.0):
    """The bilateral grid's circular Gaussian convolution of a (2, D, H, W)
    float32 grid (the two channels the reference convolves)."""
    g = np.array(grid, np.float32, order="C", copy=True)
    LIB.oracle_gauss_circular3d(_p(g), g.shape[3], g.shape[2], g.shape[1], ss, sr)
    return g


def gamma_correct(rgb, gamma):
    img = np.array(rgb, np.float32, order="C", copy=True)
    LIB.oracle_gamma(_p(img), img.size // 3, gamma)
    return img


def clip(rgb, xs=1, ys=1, bgr=False):
    """Box down-sampling by (xs, ys), clamp and truncation to bytes:
    (H / ys, W / xs, 3) uint8."""
    img = np.ascontiguousarray(rgb, np.float32)
    h, w = img.shape[0], img.shape[1]
    out = np.zeros((h // ys, w // xs, 3), np.uint8)
    LIB.oracle_clip(_p(img), w, h, xs, ys, int(bgr), _p(out))
    return out
