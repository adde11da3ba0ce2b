"""ctypes binding of libraptor_rt.so (include/raptor_rt.h) for tests and benchmarks.

The library is the product: HIP kernels for gfx950 behind a C-ABI. This module
only marshals numpy arrays and torch device pointers across that ABI. It has no
fallback: if the shared library is missing, importing this module raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libraptor_rt.so")

RT_OK = 0
RT_ERR_INVALID = -1
RT_ERR_HIP = -2
RT_ERR_OOM = -3
RT_ERR_UNSUPPORTED = -4
RT_DEVICE_PTRS = 1
RT_REGROUP_BOUNCE = 2
RT_ASYNC = 4

EPSILON = np.float32(1.0e-3)  # common/common.cc:14
MAX_DIST = np.float32(np.finfo(np.float32).max)  # common/common.cc:17

HIT_DTYPE = np.dtype([("t", "<f4"), ("u", "<f4"), ("v", "<f4"), ("prim", "<i4")])

# Symbols include/raptor_rt.h declares.
EXPORTS = [
    "rt_ctx_create", "rt_ctx_destroy", "rt_last_error", "rt_scene_create", "rt_scene_destroy",
    "rt_scene_normals", "rt_accel_build", "rt_accel_destroy", "rt_accel_info_get", "rt_trace_closest",
    "rt_trace_occluded", "rt_render", "rt_render_static", "rt_set_counting", "rt_ctx_frame_rays", "rt_gen_random_triangles", "rt_gen_terrain",
    "rt_batch_create", "rt_batch_run", "rt_batch_destroy", "rt_batch_autotune", "rt_gen_surface_rays",
    "rt_scene_set_albedo", "rt_render_tape", "rt_tape_info", "rt_render_adjoint", "rt_tape_destroy",
    "rt_invrender_create", "rt_invrender_run", "rt_invrender_destroy", "rt_scene_update", "rt_accel_refit",
    "rt_scene_set_materials", "rt_image_tone_map", "rt_image_gamma", "rt_image_clip", "rt_write_png",
    "rt_write_tga", "rt_scene_set_vertex_normals", "rt_mesh_load", "rt_mesh_parse", "rt_mesh_info_get",
    "rt_mesh_copy", "rt_mesh_last_error", "rt_mesh_destroy", "rt_scene_create_from_mesh", "rt_write_jpeg",
    "rt_mesh_camera", "rt_mesh_lights", "rt_build_morton", "rt_accel_download", "rt_accel_root", "rt_glare_filter",
    "rt_scene_set_textures", "rt_scene_set_texture_coords", "rt_texture_load_tga", "rt_mesh_copy_textures",
    "rt_build_aac",
]


class rt_camera(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("x_axis", C.c_float * 3), ("y_axis", C.c_float * 3),
                ("z_axis", C.c_float * 3), ("w", C.c_float), ("h", C.c_float), ("t", C.c_float),
                ("x_res", C.c_int32), ("y_res", C.c_int32)]


class rt_light(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("rgb", C.c_float * 3), ("d", C.c_float)]


BUILDER_LBVH, BUILDER_AAC = 0, 1


class rt_build_params(C.Structure):
    _fields_ = [("width", C.c_int32), ("leaf_size", C.c_int32), ("sah_bins", C.c_int32), ("lds_stack", C.c_int32),
                ("builder", C.c_int32)]


class rt_render_params(C.Structure):
    _fields_ = [("bounce", C.c_int32), ("spp", C.c_int32), ("seed", C.c_uint64), ("flags", C.c_int32),
                ("reserved", C.c_int32), ("background", C.c_float * 3), ("reserved2", C.c_int32)]


class rt_stats(C.Structure):
    _fields_ = [("rays", C.c_uint64 * 4), ("ms", C.c_float * 6), ("visits", C.c_uint64 * 8)]


class rt_batch_item(C.Structure):
    _fields_ = [("accel", C.c_void_p), ("rays", C.c_void_p), ("n", C.c_int64), ("hits", C.c_void_p)]


class rt_batch_params(C.Structure):
    _fields_ = [("lds_stack", C.c_int32), ("waves_per_cu", C.c_int32), ("grab", C.c_int32), ("use_graph", C.c_int32),
                ("group_waves", C.c_int32), ("refill", C.c_int32), ("width", C.c_int32)]


class rt_invrender_item(C.Structure):
    _fields_ = [("accel", C.c_void_p), ("cam", rt_camera), ("albedo", C.c_void_p), ("d_albedo", C.c_void_p),
                ("target", C.c_void_p), ("image", C.c_void_p)]


class rt_invrender_params(C.Structure):
    _fields_ = [("spp", C.c_int32), ("bounces", C.c_int32), ("seed", C.c_uint64), ("lr", C.c_float),
                ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float), ("use_graph", C.c_int32),
                ("waves_per_cu", C.c_int32)]


class rt_tone_params(C.Structure):
    _fields_ = [("mode", C.c_int32), ("key", C.c_float), ("xw", C.c_float), ("yw", C.c_float),
                ("desaturate", C.c_int32), ("stretch", C.c_int32), ("adapt", C.c_int32), ("time_step", C.c_float),
                ("adaptation_level", C.c_float), ("log_avg_luminance", C.c_float), ("x_angle", C.c_float),
                ("y_angle", C.c_float), ("glare", C.c_int32)]


class rt_accel_info(C.Structure):
    _fields_ = [("nodes", C.c_int64), ("leaves", C.c_int64), ("triangles", C.c_int64), ("width", C.c_int32),
                ("depth", C.c_int32), ("bytes", C.c_int64), ("build_ms", C.c_float), ("sah_cost", C.c_float),
                ("refit_ms", C.c_float)]


class rt_mesh_info(C.Structure):
    _fields_ = [("triangles", C.c_int64), ("materials", C.c_int32), ("has_vertex_normals", C.c_int32),
                ("vertices", C.c_int64), ("faces", C.c_int64), ("skipped_faces", C.c_int64),
                ("skipped_statements", C.c_int64), ("texture_maps", C.c_int32), ("lights", C.c_int32),
                ("has_camera", C.c_int32), ("has_texture_coords", C.c_int32), ("maps", C.c_int32),
                ("images", C.c_int32), ("texels", C.c_int64)]


def load_library(path=LIB_PATH):
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python raptor_amd/build.py` (no fallback exists)")
    # torch ships its own libamdhip64 under the same soname. Loading torch
    # first lets this library bind to that runtime, so a process that uses
    # both holds one HIP runtime; loaded the other way round, two runtimes
    # compete for the device and torch finds no GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(path)
    vp, i32, i64, u32, u64, fp = C.c_void_p, C.c_int32, C.c_int64, C.c_uint32, C.c_uint64, C.c_void_p
    sig = {
        "rt_ctx_create": (i32, [i32, C.POINTER(vp)]),
        "rt_ctx_destroy": (None, [vp]),
        "rt_last_error": (C.c_char_p, [vp]),
        "rt_scene_create": (i32, [vp, fp, fp, fp, i64, C.POINTER(vp)]),
        "rt_scene_destroy": (None, [vp]),
        "rt_scene_normals": (i32, [vp, fp]),
        "rt_accel_build": (i32, [vp, C.POINTER(rt_build_params), C.POINTER(vp)]),
        "rt_accel_destroy": (None, [vp]),
        "rt_accel_info_get": (i32, [vp, C.POINTER(rt_accel_info)]),
        "rt_trace_closest": (i32, [vp, fp, i64, fp, u32]),
        "rt_trace_occluded": (i32, [vp, fp, i64, fp, u32]),
        "rt_render": (i32, [vp, C.POINTER(rt_camera), C.POINTER(rt_light), i32, C.POINTER(rt_render_params),
                            fp, fp, fp, C.POINTER(rt_stats)]),
        "rt_render_static": (i32, [vp, C.POINTER(rt_camera), C.POINTER(rt_light), i32, C.POINTER(rt_render_params),
                                   fp, fp, fp, C.POINTER(rt_stats)]),
        "rt_build_morton": (i32, [vp, fp, fp]),
        "rt_build_aac": (i32, [vp, C.c_float, C.c_float, fp, fp, fp, fp, i64, C.POINTER(C.c_int64),
                               C.POINTER(C.c_int32), C.POINTER(C.c_float)]),
        "rt_glare_filter": (i32, [vp, i32, i32, C.c_float, i32, fp, u32]),
        "rt_accel_download": (i32, [vp, fp, fp]),
        "rt_accel_root": (i32, [vp, C.POINTER(C.c_int32), C.POINTER(C.c_float)]),
        "rt_set_counting": (i32, [vp, i32]),
        "rt_ctx_frame_rays": (i32, [vp, C.POINTER(C.c_uint64)]),
        "rt_batch_create": (i32, [vp, C.POINTER(rt_batch_item), i32, C.POINTER(rt_batch_params), C.POINTER(vp)]),
        "rt_batch_run": (i32, [vp, i32, C.POINTER(C.c_float)]),
        "rt_batch_destroy": (None, [vp]),
        "rt_batch_autotune": (i32, [vp, C.POINTER(rt_batch_item), i32, i32, C.POINTER(rt_batch_params),
                                    C.POINTER(C.c_float)]),
        "rt_gen_surface_rays": (i32, [vp, u64, i64, fp]),
        "rt_scene_set_albedo": (i32, [vp, fp, u32]),
        "rt_scene_update": (i32, [vp, fp, u32]),
        "rt_accel_refit": (i32, [vp]),
        "rt_scene_set_materials": (i32, [vp, fp, i32, fp]),
        "rt_image_tone_map": (i32, [vp, fp, i32, i32, C.POINTER(rt_tone_params), u32, C.POINTER(C.c_float)]),
        "rt_image_gamma": (i32, [vp, fp, i64, C.c_float, u32, C.POINTER(C.c_float)]),
        "rt_image_clip": (i32, [vp, fp, i32, i32, i32, i32, i32, fp, u32, C.POINTER(C.c_float)]),
        "rt_write_png": (i32, [C.c_char_p, fp, i32, i32]),
        "rt_write_tga": (i32, [C.c_char_p, fp, i32, i32]),
        "rt_write_jpeg": (i32, [vp, C.c_char_p, fp, i32, i32, i32, u32, C.POINTER(C.c_float)]),
        "rt_render_tape": (i32, [vp, C.POINTER(rt_camera), C.POINTER(rt_light), i32, C.POINTER(rt_render_params), fp,
                                 i64, C.POINTER(vp), C.POINTER(rt_stats)]),
        "rt_tape_info": (i32, [vp, C.POINTER(i32), C.POINTER(i64), C.POINTER(i64)]),
        "rt_render_adjoint": (i32, [vp, fp, fp, fp, C.POINTER(C.c_float)]),
        "rt_tape_destroy": (None, [vp]),
        "rt_invrender_create": (i32, [vp, C.POINTER(rt_invrender_item), i32, C.POINTER(rt_light), i32,
                                      C.POINTER(rt_invrender_params), C.POINTER(vp)]),
        "rt_invrender_run": (i32, [vp, i32, i32, i32, C.POINTER(C.c_double), C.POINTER(C.c_uint64),
                                   C.POINTER(C.c_float)]),
        "rt_invrender_destroy": (None, [vp]),
        "rt_scene_set_vertex_normals": (i32, [vp, fp]),
        "rt_scene_set_textures": (i32, [vp, vp, i32, vp, i32, fp, i64, vp, i32]),
        "rt_scene_set_texture_coords": (i32, [vp, fp]),
        "rt_texture_load_tga": (i32, [C.c_char_p, fp, i64, vp]),
        "rt_mesh_copy_textures": (i32, [vp, fp, vp, vp, fp]),
        "rt_mesh_load": (i32, [C.c_char_p, i32, C.POINTER(vp)]),
        "rt_mesh_parse": (i32, [C.c_char_p, i64, i32, C.c_char_p, C.POINTER(vp)]),
        "rt_mesh_info_get": (i32, [vp, C.POINTER(rt_mesh_info)]),
        "rt_mesh_copy": (i32, [vp, fp, fp, fp, fp]),
        "rt_mesh_last_error": (C.c_char_p, []),
        "rt_mesh_camera": (i32, [vp, C.POINTER(rt_camera), fp]),
        "rt_mesh_lights": (i32, [vp, fp]),
        "rt_mesh_destroy": (None, [vp]),
        "rt_scene_create_from_mesh": (i32, [vp, vp, C.POINTER(vp)]),
        "rt_gen_random_triangles": (i64, [u64, i64, C.c_float, C.c_float, C.c_float, fp, fp]),
        "rt_gen_terrain": (i64, [i32, C.c_float, C.c_float, C.c_float, i32, i32, u64, C.c_float, C.c_float, fp, fp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


_LIB = load_library()


def lib():
    return _LIB


def _ptr(a):
    """Address of a host array (numpy) for a call without RT_DEVICE_PTRS, or
    of a device tensor already checked by _dev."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError("host arrays passed to the library must be C-contiguous")
        return a.ctypes.data
    return _dev(a, a.dtype, None, "tensor")


def _dev(t, dtype, numel, name):
    """Device address of a torch tensor the library reads or writes under
    RT_DEVICE_PTRS. The tensor must be a contiguous GPU tensor of dtype
    holding numel elements (at least numel when numel is negative, none
    checked when None); anything else raises ValueError before the library
    sees it, so a kernel never reads host memory or past the end of a
    buffer. Torch's current stream is synchronised first: the library works
    on its own stream, and its calls return only when their kernels are done,
    so torch may use an output as soon as the call returns."""
    import torch

    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{name}: expected a torch tensor, got {type(t).__name__}")
    if t.dtype != dtype:
        raise ValueError(f"{name}: dtype {t.dtype}, expected {dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name}: the tensor must be contiguous")
    if numel is not None and ((numel >= 0 and t.numel() != numel) or (numel < 0 and t.numel() < -numel)):
        raise ValueError(f"{name}: {t.numel()} elements, expected {'at least ' if numel < 0 else ''}{abs(numel)}")
    if not t.is_cuda:
        raise ValueError(f"{name}: a GPU tensor is required (got one on {t.device})")
    torch.cuda.current_stream(t.device).synchronize()
    return int(t.data_ptr())


class RTError(RuntimeError):
    pass


# ---------------------------------------------------------------- generators

def gen_random_triangles(seed, n, spread=0.01, albedo_lo=0.0, albedo_hi=1.0):
    """Triangles with centroids uniform in [0,1)^3, vertices centroid + U[-spread, spread)^3."""
    abc = np.empty((n, 9), np.float32)
    alb = np.empty((n, 3), np.float32)
    r = _LIB.rt_gen_random_triangles(seed, n, spread, albedo_lo, albedo_hi, _ptr(abc), _ptr(alb))
    if r != n:
        raise RTError("rt_gen_random_triangles failed")
    return abc, alb


def gen_terrain(grid, amplitude=0.15, zoom=0.02, persistence=0.5, octaves=8, noise_seed=57, albedo_seed=2,
                albedo_lo=0.2, albedo_hi=0.8):
    """fBm value-noise heightfield of grid x grid vertices spanning [0,1] in x and z."""
    n = _LIB.rt_gen_terrain(grid, amplitude, zoom, persistence, octaves, noise_seed, albedo_seed, albedo_lo,
                            albedo_hi, None, None)
    if n < 0:
        raise RTError("rt_gen_terrain: bad grid")
    abc = np.empty((n, 9), np.float32)
    alb = np.empty((n, 3), np.float32)
    _LIB.rt_gen_terrain(grid, amplitude, zoom, persistence, octaves, noise_seed, albedo_seed, albedo_lo, albedo_hi,
                        _ptr(abc), _ptr(alb))
    return abc, alb


# ---------------------------------------------------------------- handles

MATERIAL_PHONG, MATERIAL_COOK_TORRANCE, MATERIAL_COLOURED, MATERIAL_MAPPER, MATERIAL_MANDELBROT = 0, 1, 2, 3, 4
# rt_material (include/raptor_rt.h), 160 bytes.
MATERIAL_DTYPE = np.dtype([("ka", "<f4", 3), ("kd", "<f4", 3), ("ks", "<f4", 3), ("shininess", "<f4"),
                           ("transmittance", "<f4"), ("refractive_index", "<f4"), ("reflectance", "<f4"),
                           ("kind", "<i4"), ("chroma", "<f4", 2), ("diffuse", "<f4"), ("roughness", "<f4"),
                           ("extinction", "<f4"), ("tex", "<i4", 6), ("mapper_kd", "<f4"), ("mapper_ks", "<f4"),
                           ("clrh", "<f4"), ("clrf", "<f4"), ("layer_first", "<i4", 5), ("layer_count", "<i4", 5),
                           ("iterations", "<i4")])
assert MATERIAL_DTYPE.itemsize == 160

# rt_aac_node (include/raptor_rt.h).
AAC_NODE_DTYPE = np.dtype([("lo", "<f4", 3), ("hi", "<f4", 3), ("child", "<i4", 2), ("leaf", "<i4")])
assert AAC_NODE_DTYPE.itemsize == 36

MAP_PLANAR, MAP_CUBIC, MAP_CYLINDRICAL, MAP_CHECKER_BOARD, MAP_PERLIN_3D, MAP_GRADIENT = 0, 1, 2, 3, 4, 5
GRAD_LINE, GRAD_SPLINE, GRAD_STEP = 0, 1, 2
GRAD_OF_INCIDENCE_ANGLE = 3
WRAP_BLANK, WRAP_CLAMP, WRAP_TILE, WRAP_MIRROR = 0, 1, 2, 3
FALLOFF_CUBIC, FALLOFF_SPHERICAL, FALLOFF_LINEAR_X, FALLOFF_LINEAR_Y, FALLOFF_LINEAR_Z, FALLOFF_NONE = range(6)
# rt_texture_map and rt_texture_image (include/raptor_rt.h).
TEXTURE_MAP_DTYPE = np.dtype([("kind", "<i4"), ("image", "<i4"), ("c", "<f4", 3), ("n", "<f4", 3),
                              ("s", "<f4", 3), ("u", "<f4", 3), ("v", "<f4", 3), ("u_wrap", "<i4"),
                              ("v_wrap", "<i4"), ("u_off", "<i4"), ("v_off", "<i4"), ("cycles", "<i4"),
                              ("rgb", "<f4", 3), ("opacity", "<f4"), ("inverse", "<i4"), ("falloff", "<i4"),
                              ("falloff_centre", "<f4", 3), ("falloff_grad", "<f4", 3), ("persistence", "<f4"),
                              ("zoom", "<f4"), ("octaves", "<i4"), ("seed", "<i4"), ("key_count", "<i4"),
                              ("grad_of", "<i4"), ("key_offset", "<i8")])
TEXTURE_IMAGE_DTYPE = np.dtype([("width", "<i4"), ("height", "<i4"), ("cpp", "<i4"), ("reserved", "<i4"),
                                ("offset", "<i8")])
assert TEXTURE_MAP_DTYPE.itemsize == 168 and TEXTURE_IMAGE_DTYPE.itemsize == 24


def _f3(v):
    return np.asarray(v, np.float32).reshape(3)


def _default_axes(n):
    """u = (n.y + n.z, 0, n.x), v = (0, n.x + n.z, n.y) in float32, as the
    image mappers' first constructors derive them
    (rt/materials/planar_mapper.h:23-29, cubic_mapper.h:24-34)."""
    n = _f3(n)
    return (np.array([n[1] + n[2], 0.0, n[0]], np.float32), np.array([0.0, n[0] + n[2], n[1]], np.float32))


def _map(kind):
    m = np.zeros((), TEXTURE_MAP_DTYPE)
    m["kind"] = kind
    m["falloff"] = FALLOFF_NONE
    m["opacity"] = 1.0
    return m


def image_map(kind, image, c, n, s, u_wrap=WRAP_TILE, v_wrap=WRAP_TILE, u_off=0, v_off=0, u=None, v=None,
              cycles=1):
    """A planar, cubic or cylindrical map of image index `image` as the
    reference's constructors leave it. Without u, v: planar and cubic take
    u = (n.y + n.z, 0, n.x), v = (0, n.x + n.z, n.y) and size s; the
    cylindrical mapper takes that u, v = n, size s, tile / blank wrapping
    (rt/materials/cylindrical_mapper.h:20-24). With u and v (the explicit
    constructors): size = u s.x + v s.y + n s.z for planar and cubic."""
    m = _map(kind)
    m["image"] = image
    m["c"], m["n"] = _f3(c), _f3(n)
    if u is None:
        du, dv = _default_axes(n)
        m["u"] = du
        m["v"] = _f3(n) if kind == MAP_CYLINDRICAL else dv
        m["s"] = _f3(s)
    else:
        m["u"], m["v"] = _f3(u), _f3(v)
        sz = _f3(s)
        m["s"] = (_f3(u) * sz[0] + _f3(v) * sz[1]) + _f3(n) * sz[2] if kind != MAP_CYLINDRICAL else sz
    if kind == MAP_CYLINDRICAL:
        u_wrap, v_wrap = WRAP_TILE, WRAP_BLANK
    m["u_wrap"], m["v_wrap"], m["u_off"], m["v_off"], m["cycles"] = u_wrap, v_wrap, u_off, v_off, cycles
    return m


def checker_board_map(rgb, centre, grid_size, opacity=1.0):
    """checker_board_mapper(rgb, cnt, grid_size, op)
    (rt/materials/checker_board_mapper.h:17-18)."""
    m = _map(MAP_CHECKER_BOARD)
    m["rgb"], m["c"], m["s"], m["opacity"] = _f3(rgb), _f3(centre), _f3(grid_size), opacity
    return m


def perlin_3d_map(rgb, persistence, zoom, octaves, seed, opacity=1.0):
    """perlin_noise_3d_mapper(rgb, p, z, o, s, op)
    (rt/materials/perlin_noise_3d_mapper.h:20-21): rgb 0-255."""
    m = _map(MAP_PERLIN_3D)
    m["rgb"], m["persistence"], m["zoom"], m["octaves"], m["seed"] = _f3(rgb), persistence, zoom, octaves, seed
    m["opacity"] = opacity
    return m


def gradient_map(key_offset, key_count, inverse=False):
    """gradient_mapper over the incidence angle
    (rt/materials/gradient_mapper.h:31-33) whose key_count keys sit at
    texels[key_offset:] (see gradient_keys)."""
    m = _map(MAP_GRADIENT)
    m["grad_of"], m["key_offset"], m["key_count"], m["inverse"] = GRAD_OF_INCIDENCE_ANGLE, key_offset, key_count, inverse
    return m


def gradient_keys(keys, colours, alphas, interpolations):
    """A gradient's keys as the texel rows its map reads: per key (degrees,
    non-decreasing) {key, r, g, b, alpha, interpolation} (GRAD_LINE or
    GRAD_STEP, the interpolation towards that key from the one below)."""
    k = np.asarray(keys, np.float32).reshape(-1)
    out = np.zeros((k.size, 6), np.float32)
    out[:, 0] = k
    out[:, 1:4] = np.asarray(colours, np.float32).reshape(-1, 3)
    out[:, 4] = np.asarray(alphas, np.float32).reshape(-1)
    out[:, 5] = np.asarray(interpolations, np.float32).reshape(-1)
    return out.reshape(-1)


def with_falloff(m, kind, centre, grad):
    """A copy of map m with a mapper_falloff (rt/materials/mapper_falloff.h)."""
    m = m.copy()
    m["falloff"], m["falloff_centre"], m["falloff_grad"] = kind, _f3(centre), _f3(grad)
    return m


def texture_images(images):
    """(rt_texture_image table, texels) from (h, w, cpp) arrays (cpp 1, 3 or 4;
    a 2-D array is one channel), row 0 first."""
    table = np.zeros(len(images), TEXTURE_IMAGE_DTYPE)
    parts, off = [], 0
    for i, im in enumerate(images):
        a = np.asarray(im, np.float32)
        if a.ndim == 2:
            a = a[:, :, None]
        table[i] = (a.shape[1], a.shape[0], a.shape[2], 0, off)
        parts.append(a.reshape(-1))
        off += a.size
    texels = np.concatenate(parts) if parts else np.zeros(0, np.float32)
    return table, np.ascontiguousarray(texels, np.float32)


def load_tga(path):
    """A TGA image as the reference reads it (rows in file order, RGB):
    (height, width, channels) float32, 0-255."""
    dims = np.zeros(3, np.int32)
    rc = _LIB.rt_texture_load_tga(str(path).encode(), None, 0, _ptr(dims))
    if rc != RT_OK:
        raise RTError(f"rt_texture_load_tga: {rc} {_LIB.rt_mesh_last_error().decode()}")
    out = np.empty((dims[1], dims[0], dims[2]), np.float32)
    rc = _LIB.rt_texture_load_tga(str(path).encode(), _ptr(out), out.size, _ptr(dims))
    if rc != RT_OK:
        raise RTError(f"rt_texture_load_tga: {rc} {_LIB.rt_mesh_last_error().decode()}")
    return out


def layer_lists(materials):
    """mapper_shader layer lists: dict materials of kind MATERIAL_MAPPER with
    ctex, dtex, stex, rtex, ttex (lists of map indices: colour, kd, s,
    reflectance, transmittance) get "layers" ((first, count) x 5) into the
    returned int32 array."""
    out = []
    for m in materials:
        if isinstance(m, dict) and m.get("kind") == MATERIAL_MAPPER:
            lists = []
            for key in ("ctex", "dtex", "stex", "rtex", "ttex"):
                idx = [int(i) for i in m.get(key, ())]
                lists.append((len(out), len(idx)))
                out.extend(idx)
            m["layers"] = lists
    return np.asarray(out, np.int32)


def material_table(materials):
    """An rt_material array from dicts. Phong (the default kind): ka, kd, ks
    (rgb, 0..255; default 0), s (shininess, 0), tran (transmittance, 0), ri
    (refractive index, 1), rf (reflectance, 0). Cook-Torrance (kind 1,
    cook_torrance_cxy): x, y (chroma, 0.33), rd (0.5), rs, ts, sr (0), ri_r
    (1), ri_i (0). Kind 2 (coloured_mapper_shader) takes the phong fields and
    tex (t_ka, t_kd, t_ks, t_ns, t_refl, t_d: map indices or None). Kind 3
    (mapper_shader): rgb (0..255), kd, ks (scalars), s, tran, ri, rf, clrh,
    clrf and layers ((first, count) x 5, see layer_lists). Kind 4
    (mandelbrot_shader): iterations. A row of 13 floats (ka, kd, ks, s,
    tran, ri, rf) is a phong material."""
    out = np.zeros(len(materials), MATERIAL_DTYPE)
    for i, m in enumerate(materials):
        r = out[i]
        if not isinstance(m, dict):
            v = [float(x) for x in m]
            if len(v) != 13:
                raise ValueError("a phong row has 13 floats: ka rgb, kd rgb, ks rgb, s, tran, ri, rf")
            m = dict(ka=v[0:3], kd=v[3:6], ks=v[6:9], s=v[9], tran=v[10], ri=v[11], rf=v[12])
        if m.get("kind", MATERIAL_PHONG) == MATERIAL_COOK_TORRANCE:
            r["kind"] = MATERIAL_COOK_TORRANCE
            r["chroma"] = (m.get("x", 0.33), m.get("y", 0.33))
            r["diffuse"] = m.get("rd", 0.5)
            r["reflectance"] = m.get("rs", 0.0)
            r["transmittance"] = m.get("ts", 0.0)
            r["roughness"] = m.get("sr", 0.0)
            r["refractive_index"] = m.get("ri_r", 1.0)
            r["extinction"] = m.get("ri_i", 0.0)
        elif m.get("kind") == MATERIAL_MANDELBROT:
            r["kind"] = MATERIAL_MANDELBROT
            r["iterations"] = m.get("iterations", 100)
        elif m.get("kind") == MATERIAL_MAPPER:
            r["kind"] = MATERIAL_MAPPER
            r["kd"] = m.get("rgb", (0, 0, 0))
            r["mapper_kd"], r["mapper_ks"] = m.get("kd", 0.0), m.get("ks", 0.0)
            r["shininess"] = m.get("s", 0.0)
            r["transmittance"] = m.get("tran", 0.0)
            r["refractive_index"] = m.get("ri", 1.0)
            r["reflectance"] = m.get("rf", 0.0)
            r["clrh"], r["clrf"] = m.get("clrh", 0.0), m.get("clrf", 0.0)
            for k, (first, count) in enumerate(m.get("layers", [(0, 0)] * 5)):
                r["layer_first"][k], r["layer_count"][k] = first, count
        else:
            r["kind"] = m.get("kind", MATERIAL_PHONG)
            r["ka"], r["kd"], r["ks"] = m.get("ka", (0, 0, 0)), m.get("kd", (0, 0, 0)), m.get("ks", (0, 0, 0))
            r["shininess"] = m.get("s", 0.0)
            r["transmittance"] = m.get("tran", 0.0)
            r["refractive_index"] = m.get("ri", 1.0)
            r["reflectance"] = m.get("rf", 0.0)
            for k, t in enumerate(m.get("tex", ())):
                r["tex"][k] = 0 if t is None or t < 0 else int(t) + 1
    return out


class Context:
    def __init__(self, device=0):
        h = C.c_void_p()
        rc = _LIB.rt_ctx_create(device, C.byref(h))
        if rc != RT_OK:
            raise RTError(f"rt_ctx_create failed ({rc})")
        self.h = h

    def check(self, rc, what):
        if rc != RT_OK:
            raise RTError(f"{what}: {rc} {_LIB.rt_last_error(self.h).decode()}")

    def set_counting(self, on):
        self.check(_LIB.rt_set_counting(self.h, 1 if on else 0), "rt_set_counting")

    def frame_rays(self):
        """Waits for the context's frames; rays of its asynchronous frames
        since the last call by stage (primary, shadow, bounce, bounce shadow)."""
        v = (C.c_uint64 * 4)()
        self.check(_LIB.rt_ctx_frame_rays(self.h, v), "rt_ctx_frame_rays")
        return [int(x) for x in v]

    def close(self):
        if self.h:
            _LIB.rt_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Scene:
    @classmethod
    def from_mesh(cls, ctx, mesh):
        """The scene of a parsed Mesh: triangles, albedo kd / 255,
        transparency flags, materials and vertex normals."""
        self = cls.__new__(cls)
        self.ctx = ctx
        self.n = mesh.n
        h = C.c_void_p()
        ctx.check(_LIB.rt_scene_create_from_mesh(ctx.h, mesh.h, C.byref(h)), "rt_scene_create_from_mesh")
        self.h = h
        return self

    def __init__(self, ctx, abc, albedo=None, flags=None):
        self.ctx = ctx
        abc = np.ascontiguousarray(abc, np.float32).reshape(-1, 9)
        self.n = abc.shape[0]
        alb = None if albedo is None else np.ascontiguousarray(albedo, np.float32).reshape(-1, 3)
        fl = None if flags is None else np.ascontiguousarray(flags, np.uint8)
        h = C.c_void_p()
        ctx.check(_LIB.rt_scene_create(ctx.h, _ptr(abc), _ptr(alb), _ptr(fl), self.n, C.byref(h)), "rt_scene_create")
        self.h = h

    def set_albedo(self, albedo):
        """albedo: (n, 3) float32, a host array or a device tensor."""
        if isinstance(albedo, np.ndarray):
            a = np.ascontiguousarray(albedo, np.float32).reshape(-1, 3)
            self.ctx.check(_LIB.rt_scene_set_albedo(self.h, _ptr(a), 0), "rt_scene_set_albedo")
        else:
            import torch

            a = albedo.detach()
            ptr = _dev(a, torch.float32, 3 * self.n, "rt_scene_set_albedo: albedo")
            self.ctx.check(_LIB.rt_scene_set_albedo(self.h, ptr, RT_DEVICE_PTRS), "rt_scene_set_albedo")

    def update(self, abc):
        """New vertex positions, (n, 9) float32: a host array or a device tensor.
        Accels built on the scene must be refit (Accel.refit) or rebuilt
        before they are traced again."""
        if isinstance(abc, np.ndarray):
            a = np.ascontiguousarray(abc, np.float32).reshape(-1, 9)
            if a.shape[0] != self.n:
                raise ValueError(f"rt_scene_update: {a.shape[0]} triangles, the scene has {self.n}")
            self.ctx.check(_LIB.rt_scene_update(self.h, _ptr(a), 0), "rt_scene_update")
        else:
            import torch

            a = abc.detach()
            ptr = _dev(a, torch.float32, 9 * self.n, "rt_scene_update: vertices")
            self.ctx.check(_LIB.rt_scene_update(self.h, ptr, RT_DEVICE_PTRS), "rt_scene_update")

    def set_materials(self, materials, tri_material):
        """Materials for rt_render: materials is an array of MATERIAL_DTYPE
        or anything material_table takes (dicts, phong rows of 13 floats);
        tri_material (n,) int32 indices, -1 for Lambert shading from the
        albedo. No materials removes them."""
        if not (isinstance(materials, np.ndarray) and materials.dtype == MATERIAL_DTYPE):
            materials = material_table(list(materials))
        m = np.ascontiguousarray(materials)
        k = np.ascontiguousarray(tri_material, np.int32).reshape(-1)
        if m.shape[0] > 0 and k.shape[0] != self.n:
            raise ValueError(f"rt_scene_set_materials: {k.shape[0]} indices, the scene has {self.n} triangles")
        self.ctx.check(_LIB.rt_scene_set_materials(self.h, _ptr(m), m.shape[0], _ptr(k)), "rt_scene_set_materials")

    def set_textures(self, maps, images=None, texels=None, layers=None):
        """Texture maps (TEXTURE_MAP_DTYPE array or list of maps), their images
        (TEXTURE_IMAGE_DTYPE with texels, or a list of (h, w, cpp) arrays)
        and mapper_shader layers (int32 map indices). Empty maps remove them."""
        m = np.ascontiguousarray(np.asarray(list(maps) if not isinstance(maps, np.ndarray) else maps,
                                            TEXTURE_MAP_DTYPE).reshape(-1))
        if images is None:
            images = []
        if isinstance(images, np.ndarray) and images.dtype == TEXTURE_IMAGE_DTYPE:
            table, tx = np.ascontiguousarray(images), np.ascontiguousarray(texels, np.float32).reshape(-1)
        else:
            table, tx = texture_images(images)
        lay = np.ascontiguousarray(layers if layers is not None else np.zeros(0, np.int32), np.int32).reshape(-1)
        self.ctx.check(_LIB.rt_scene_set_textures(self.h, _ptr(m), m.shape[0], _ptr(table), table.shape[0],
                                                  _ptr(tx), tx.size, _ptr(lay), lay.size), "rt_scene_set_textures")

    def set_texture_coords(self, vt):
        """(n, 6) float32 texture coordinates (u, v) at vertices a, b, c; a
        first value of FLT_MAX: none for that triangle; None removes them."""
        if vt is None:
            self.ctx.check(_LIB.rt_scene_set_texture_coords(self.h, None), "rt_scene_set_texture_coords")
            return
        v = np.ascontiguousarray(vt, np.float32).reshape(-1, 6)
        if v.shape[0] != self.n:
            raise ValueError(f"rt_scene_set_texture_coords: {v.shape[0]} rows, the scene has {self.n} triangles")
        self.ctx.check(_LIB.rt_scene_set_texture_coords(self.h, _ptr(v)), "rt_scene_set_texture_coords")

    def set_vertex_normals(self, vn):
        """(n, 9) float32 normals at vertices a, b, c; a first value of
        FLT_MAX keeps that triangle's geometric normal; None removes them."""
        if vn is None:
            self.ctx.check(_LIB.rt_scene_set_vertex_normals(self.h, None), "rt_scene_set_vertex_normals")
            return
        v = np.ascontiguousarray(vn, np.float32).reshape(-1, 9)
        if v.shape[0] != self.n:
            raise ValueError(f"rt_scene_set_vertex_normals: {v.shape[0]} rows, the scene has {self.n} triangles")
        self.ctx.check(_LIB.rt_scene_set_vertex_normals(self.h, _ptr(v)), "rt_scene_set_vertex_normals")

    def normals(self):
        out = np.empty((self.n, 3), np.float32)
        self.ctx.check(_LIB.rt_scene_normals(self.h, _ptr(out)), "rt_scene_normals")
        return out

    def morton(self):
        """The build's Morton stages: (codes in emplace order, emplace ids in
        stable code order), both uint32."""
        codes = np.empty(self.n, np.uint32)
        order = np.empty(self.n, np.uint32)
        self.ctx.check(_LIB.rt_build_morton(self.h, _ptr(codes), _ptr(order)), "rt_build_morton")
        return codes, order

    def build_aac(self, alpha=0.4, delta=20.0):
        """The reference builder's tree built on the GPU (rt_build_aac):
        dict(codes, order, indirect, lohi (m, 6), child (m, 2), leaf (m,)
        bool, root, build_ms)."""
        n = self.n
        cap = max(1, 2 * n - 1)
        codes = np.empty(max(1, n), np.uint32)
        order = np.empty(max(1, n), np.uint32)
        ind = np.empty(max(1, n), np.int32)
        nodes = np.empty(cap, AAC_NODE_DTYPE)
        nn, root, ms = C.c_int64(0), C.c_int32(-1), C.c_float(0.0)
        self.ctx.check(_LIB.rt_build_aac(self.h, alpha, delta, _ptr(codes), _ptr(order), _ptr(ind), _ptr(nodes), cap,
                                         C.byref(nn), C.byref(root), C.byref(ms)), "rt_build_aac")
        m = nn.value
        nd = nodes[:m]
        return dict(codes=codes[:n], order=order[:n], indirect=ind[:n],
                    lohi=np.concatenate([nd["lo"], nd["hi"]], axis=1), child=nd["child"].copy(),
                    leaf=nd["leaf"] != 0, root=root.value, build_ms=ms.value)

    def close(self):
        if self.h:
            _LIB.rt_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


FORMAT_OFF, FORMAT_PLY, FORMAT_OBJ, FORMAT_NFF = 0, 1, 2, 3


class Mesh:
    """A parsed OFF, PLY, OBJ or NFF file (host memory; no GPU needed).
    Fields: abc (n, 9) float32 in emplace order, tri_material (n,) int32,
    materials (MATERIAL_DTYPE), vertex_normals (n, 9) float32 or None, info
    (dict); for a file that defines them (NFF): camera (an rt_camera),
    background (3 floats, 0-255) and lights ((k, 7) float32 rows: position,
    rgb, falloff), else None; texture_coords (n, 6) float32 or None, maps
    (TEXTURE_MAP_DTYPE), images (TEXTURE_IMAGE_DTYPE), texels (float32)."""

    def __init__(self, path=None, data=None, fmt=-1, directory=None):
        h = C.c_void_p()
        if path is not None:
            rc = _LIB.rt_mesh_load(os.fsencode(path), fmt, C.byref(h))
        else:
            buf = bytes(data)
            rc = _LIB.rt_mesh_parse(buf, len(buf), fmt, None if directory is None else os.fsencode(directory),
                                    C.byref(h))
        if rc != RT_OK:
            raise RTError(f"rt_mesh: {rc} {_LIB.rt_mesh_last_error().decode()}")
        self.h = h
        inf = rt_mesh_info()
        _LIB.rt_mesh_info_get(h, C.byref(inf))
        self.info = {k: getattr(inf, k) for k, _ in rt_mesh_info._fields_}
        self.n = inf.triangles
        self.abc = np.empty((self.n, 9), np.float32)
        self.tri_material = np.empty(self.n, np.int32)
        self.materials = np.zeros(inf.materials, MATERIAL_DTYPE)
        self.vertex_normals = np.empty((self.n, 9), np.float32) if inf.has_vertex_normals else None
        _LIB.rt_mesh_copy(h, _ptr(self.abc), _ptr(self.tri_material), _ptr(self.materials),
                          _ptr(self.vertex_normals))
        self.camera = self.background = None
        if inf.has_camera:
            self.camera = rt_camera()
            self.background = np.zeros(3, np.float32)
            _LIB.rt_mesh_camera(h, C.byref(self.camera), _ptr(self.background))
        self.lights = np.zeros((inf.lights, 7), np.float32) if inf.lights else None
        if inf.lights:
            _LIB.rt_mesh_lights(h, _ptr(self.lights))
        self.texture_coords = np.empty((self.n, 6), np.float32) if inf.has_texture_coords else None
        self.maps = np.zeros(inf.maps, TEXTURE_MAP_DTYPE)
        self.images = np.zeros(inf.images, TEXTURE_IMAGE_DTYPE)
        self.texels = np.zeros(inf.texels, np.float32)
        _LIB.rt_mesh_copy_textures(h, _ptr(self.texture_coords), _ptr(self.maps), _ptr(self.images),
                                   _ptr(self.texels))

    def close(self):
        if self.h:
            _LIB.rt_mesh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def surface_rays(scene, seed, n):
    """n incoherent rays leaving the scene's surfaces, as a (n, 8) float32 device tensor."""
    import torch

    rays = torch.empty((n, 8), dtype=torch.float32, device="cuda:0")
    ptr = _dev(rays, torch.float32, 8 * n, "rt_gen_surface_rays: rays")
    scene.ctx.check(_LIB.rt_gen_surface_rays(scene.h, seed, n, ptr), "rt_gen_surface_rays")
    return rays


def _batch_items(accels, rays, hits):
    """rt_batch_item table: per item (n, 8) float32 rays and room for n
    hits, (n, 4) float32, both contiguous device tensors."""
    import torch

    items = []
    for k, (a, r, h) in enumerate(zip(accels, rays, hits)):
        n = int(r.shape[0]) if r.dim() > 0 else 0
        rp = _dev(r, torch.float32, 8 * n, f"batch item {k}: rays")
        hp = _dev(h, torch.float32, 4 * n, f"batch item {k}: hits")
        items.append(rt_batch_item(a.h.value, rp, n, hp))
    return (rt_batch_item * len(items))(*items)


class Batch:
    """Many (BVH, rays, hits) items traced by one persistent launch (device tensors).

    group_waves: wavefronts per workgroup (1, 2 or 4); refill: 0 traces a
    chunk of 64 rays to its last ray, 1..64 hands new rays to a wavefront's
    finished lanes once that many are idle; width: 0 traces the items' BVHs,
    4 or 8 builds the scenes' BVHs at that node width for the batch."""

    def __init__(self, ctx, accels, rays, hits, lds_stack=8, waves_per_cu=16, grab=1, use_graph=True,
                 group_waves=1, refill=0, width=0):
        self.ctx = ctx
        self._keep = (accels, rays, hits)
        self.items = _batch_items(accels, rays, hits)
        self.params = rt_batch_params(lds_stack, waves_per_cu, grab, 1 if use_graph else 0, group_waves, refill,
                                      width)
        h = C.c_void_p()
        ctx.check(_LIB.rt_batch_create(ctx.h, self.items, len(accels), C.byref(self.params), C.byref(h)),
                  "rt_batch_create")
        self.h = h

    def run(self, plain=False):
        ms = C.c_float()
        self.ctx.check(_LIB.rt_batch_run(self.h, 1 if plain else 0, C.byref(ms)), "rt_batch_run")
        return ms.value

    @staticmethod
    def autotune(ctx, accels, rays, hits, reps=3):
        items = _batch_items(accels, rays, hits)
        best = rt_batch_params()
        ms = C.c_float()
        ctx.check(_LIB.rt_batch_autotune(ctx.h, items, len(accels), reps, C.byref(best), C.byref(ms)),
                  "rt_batch_autotune")
        return {"lds_stack": best.lds_stack, "waves_per_cu": best.waves_per_cu, "grab": best.grab,
                "group_waves": best.group_waves, "refill": best.refill, "width": best.width}, ms.value

    def close(self):
        if self.h:
            _LIB.rt_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Accel:
    def __init__(self, scene, width=4, leaf_size=2, sah_bins=16, lds_stack=8, builder=BUILDER_LBVH):
        """builder BUILDER_AAC traverses the reference builder's own tree
        (leaf_size and sah_bins do not apply)."""
        self.scene = scene
        self.ctx = scene.ctx
        p = rt_build_params(width, leaf_size, sah_bins, lds_stack, builder)
        h = C.c_void_p()
        self.ctx.check(_LIB.rt_accel_build(scene.h, C.byref(p), C.byref(h)), "rt_accel_build")
        self.h = h

    def refit(self):
        """Fit the boxes to the scene's current vertices, keeping the tree."""
        self.ctx.check(_LIB.rt_accel_refit(self.h), "rt_accel_refit")

    def download(self):
        """The tree on the host: dict(width, root, box_pad, lo (nodes, 3, W),
        hi (nodes, 3, W), child (nodes, W) int32, tri_ids (n,) int32)."""
        info = self.info()
        W, nn = info["width"], info["nodes"]
        words = 32 if W == 4 else 64
        raw = np.zeros((max(nn, 1), words), np.float32)
        ids = np.empty(self.scene.n, np.int32)
        self.ctx.check(_LIB.rt_accel_download(self.h, _ptr(raw), _ptr(ids)), "rt_accel_download")
        root, pad = C.c_int32(), C.c_float()
        self.ctx.check(_LIB.rt_accel_root(self.h, C.byref(root), C.byref(pad)), "rt_accel_root")
        raw = raw[:nn]
        box = raw[:, :6 * W].reshape(nn, 3, 2, W)  # axis, (lo, hi), child
        return dict(width=W, root=root.value, box_pad=pad.value, lo=box[:, :, 0, :], hi=box[:, :, 1, :],
                    child=raw[:, 6 * W:7 * W].view(np.int32), tri_ids=ids)

    def info(self):
        i = rt_accel_info()
        self.ctx.check(_LIB.rt_accel_info_get(self.h, C.byref(i)), "rt_accel_info_get")
        return {k: getattr(i, k) for k, _ in rt_accel_info._fields_}

    def trace_closest(self, rays):
        """rays: (n, 8) float32 (o, tmin, d, tmax) on the host, or a device tensor."""
        if isinstance(rays, np.ndarray):
            rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
            out = np.empty(rays.shape[0], HIT_DTYPE)
            self.ctx.check(_LIB.rt_trace_closest(self.h, _ptr(rays), rays.shape[0], _ptr(out), 0), "rt_trace_closest")
            return out
        import torch
        n = int(rays.shape[0])
        rp = _dev(rays, torch.float32, 8 * n, "rt_trace_closest: rays")
        out = torch.empty((n, 4), dtype=torch.float32, device=rays.device)
        self.ctx.check(_LIB.rt_trace_closest(self.h, rp, n, _ptr(out), RT_DEVICE_PTRS), "rt_trace_closest")
        return out

    def trace_occluded(self, rays):
        if isinstance(rays, np.ndarray):
            rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
            out = np.empty(rays.shape[0], np.uint8)
            self.ctx.check(_LIB.rt_trace_occluded(self.h, _ptr(rays), rays.shape[0], _ptr(out), 0),
                           "rt_trace_occluded")
            return out
        import torch
        n = int(rays.shape[0])
        rp = _dev(rays, torch.float32, 8 * n, "rt_trace_occluded: rays")
        out = torch.empty(n, dtype=torch.uint8, device=rays.device)
        self.ctx.check(_LIB.rt_trace_occluded(self.h, rp, n, _ptr(out), RT_DEVICE_PTRS), "rt_trace_occluded")
        return out

    def render(self, camera, lights, bounce=False, spp=1, seed=0, want_hits=True, stats=True, regroup=False,
               background=(0.0, 0.0, 0.0)):
        """One frame; returns dict(rgb (H,W,3), primary hits, bounce hits, stats).
        background: the colour of rays that hit nothing (0-255 scale)."""
        cam = camera.to_c()
        nl = len(lights)
        larr = (rt_light * max(1, nl))(*[l.to_c() for l in lights])
        p = rt_render_params(1 if bounce else 0, spp, seed, RT_REGROUP_BOUNCE if regroup else 0, 0,
                             (C.c_float * 3)(*background))
        npix = camera.x_res * camera.y_res
        rgb = np.empty((camera.y_res, camera.x_res, 3), np.float32)
        prim = np.empty(npix * spp, HIT_DTYPE) if want_hits else None
        bnc = np.empty(npix * spp, HIT_DTYPE) if (want_hits and bounce) else None
        st = rt_stats()
        self.ctx.check(_LIB.rt_render(self.h, C.byref(cam), larr, nl, C.byref(p), _ptr(rgb), _ptr(prim), _ptr(bnc),
                                      C.byref(st) if stats else None), "rt_render")
        out = {"rgb": rgb, "primary": prim, "bounce": bnc}
        if stats:
            out["stats"] = {"rays": list(st.rays), "ms": list(st.ms), "visits": list(st.visits)}
        return out

    def render_device(self, camera, lights, rgb_dev, bounce=False, spp=1, seed=0, regroup=False, primary_dev=None,
                      bounce_dev=None, static=False, asynchronous=False):
        """Frame into device buffers (no host copies); returns the stats dict.
        rgb_dev (H, W, 3) float32; primary_dev, bounce_dev: optional (H*W*spp,
        4) float32 hit outputs. static=True calls rt_render_static (no device
        allocation; the frame's workspace is the library's static scratch).
        asynchronous=True enqueues the frame and returns None at once; its rays
        are summed on the device (Context.frame_rays)."""
        import torch

        cam = camera.to_c()
        nl = len(lights)
        larr = (rt_light * max(1, nl))(*[l.to_c() for l in lights])
        p = rt_render_params(1 if bounce else 0, spp, seed, RT_DEVICE_PTRS | (RT_REGROUP_BOUNCE if regroup else 0)
                             | (RT_ASYNC if asynchronous else 0), 0)
        st = rt_stats()
        nout = camera.x_res * camera.y_res * spp
        ptr = _dev(rgb_dev, torch.float32, 3 * camera.x_res * camera.y_res, "rt_render: rgb")
        pp = None if primary_dev is None else _dev(primary_dev, torch.float32, 4 * nout, "rt_render: primary")
        bp = None if bounce_dev is None else _dev(bounce_dev, torch.float32, 4 * nout, "rt_render: bounce")
        fn, name = (_LIB.rt_render_static, "rt_render_static") if static else (_LIB.rt_render, "rt_render")
        self.ctx.check(fn(self.h, C.byref(cam), larr, nl, C.byref(p), ptr, pp, bp, None if asynchronous else C.byref(st)),
                       name)
        if asynchronous:
            return None
        return {"rays": list(st.rays), "ms": list(st.ms), "visits": list(st.visits)}

    def render_tape(self, camera, lights, rgb_dev, bounce=False, spp=1, seed=0, max_record_bytes=0):
        """Frame into rgb_dev that keeps what the adjoint needs; returns (Tape, stats)."""
        return Tape(self, camera, lights, rgb_dev, bounce, spp, seed, max_record_bytes)

    def close(self):
        if self.h:
            _LIB.rt_accel_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Tape:
    """A frame rendered for differentiation (see rt_render_tape)."""

    def __init__(self, accel, camera, lights, rgb_dev, bounce, spp, seed, max_record_bytes):
        self.ctx = accel.ctx
        self.accel = accel
        self.nlights = len(lights)
        cam = camera.to_c()
        larr = (rt_light * max(1, self.nlights))(*[l.to_c() for l in lights])
        p = rt_render_params(1 if bounce else 0, spp, seed, RT_DEVICE_PTRS, 0)
        st = rt_stats()
        h = C.c_void_p()
        import torch

        self.npix = camera.x_res * camera.y_res
        ptr = _dev(rgb_dev, torch.float32, 3 * self.npix, "rt_render_tape: rgb")
        self.ctx.check(_LIB.rt_render_tape(accel.h, C.byref(cam), larr, self.nlights, C.byref(p), ptr,
                                           int(max_record_bytes), C.byref(h), C.byref(st)), "rt_render_tape")
        self.h = h
        self.stats = {"rays": list(st.rays), "ms": list(st.ms)}

    def info(self):
        r, b, t = C.c_int32(), C.c_int64(), C.c_int64()
        self.ctx.check(_LIB.rt_tape_info(self.h, C.byref(r), C.byref(b), C.byref(t)), "rt_tape_info")
        return {"replay": bool(r.value), "record_bytes": b.value, "chunk_tiles": t.value}

    def adjoint(self, d_rgb, d_albedo, d_lights):
        """Adds dL/d albedo and dL/d lights (device tensors) for d_rgb = dL/d image; returns ms."""
        import torch

        ms = C.c_float()
        g = d_rgb.contiguous()
        gp = _dev(g, torch.float32, 3 * self.npix, "rt_render_adjoint: d_rgb")
        ap = _dev(d_albedo, torch.float32, 3 * self.accel.scene.n, "rt_render_adjoint: d_albedo")
        lp = _dev(d_lights, torch.float32, 6 * self.nlights, "rt_render_adjoint: d_lights") if self.nlights else None
        self.ctx.check(_LIB.rt_render_adjoint(self.h, gp, ap, lp, C.byref(ms)), "rt_render_adjoint")
        return ms.value

    def close(self):
        if self.h:
            _LIB.rt_tape_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class InvRender:
    """Batched inverse rendering (see rt_invrender_create). scenes: list of
    (accel, camera, albedo, d_albedo, target, image) with device tensors."""

    def __init__(self, ctx, scenes, lights, spp=2, bounces=2, seed=0, lr=0.02, beta1=0.9, beta2=0.999, eps=1e-8,
                 use_graph=True, waves_per_cu=0):
        self.ctx = ctx
        self._keep = scenes
        self.n = len(scenes)
        items = []
        import torch

        for k, (accel, camera, albedo, d_albedo, target, image) in enumerate(scenes):
            npix, n = camera.x_res * camera.y_res, accel.scene.n
            items.append(rt_invrender_item(
                accel.h.value, camera.to_c(), _dev(albedo, torch.float32, 3 * n, f"scene {k}: albedo"),
                _dev(d_albedo, torch.float32, 3 * n, f"scene {k}: d_albedo"),
                _dev(target, torch.float32, 3 * npix, f"scene {k}: target") if target is not None else None,
                _dev(image, torch.float32, 3 * npix, f"scene {k}: image")))
        self.items = (rt_invrender_item * self.n)(*items)
        larr = (rt_light * max(1, len(lights)))(*[l.to_c() for l in lights])
        self.params = rt_invrender_params(spp, bounces, seed, lr, beta1, beta2, eps, 1 if use_graph else 0,
                                          waves_per_cu)
        h = C.c_void_p()
        ctx.check(_LIB.rt_invrender_create(ctx.h, self.items, self.n, larr, len(lights), C.byref(self.params),
                                           C.byref(h)), "rt_invrender_create")
        self.h = h

    def run(self, mode, steps=1, plain=False):
        """mode 0 render, 1 render + gradients, 2 Adam steps; returns (loss per scene, rays, ms)."""
        loss = (C.c_double * self.n)()
        rays = C.c_uint64()
        ms = C.c_float()
        self.ctx.check(_LIB.rt_invrender_run(self.h, mode, steps, 1 if plain else 0, loss, C.byref(rays),
                                             C.byref(ms)), "rt_invrender_run")
        return list(loss), rays.value, ms.value

    def close(self):
        if self.h:
            _LIB.rt_invrender_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- post-process
# Camera post-process on images of 0-255 floats, (H, W, 3) with row y at
# [y] (camera::image order): numpy arrays are processed through host staging,
# torch tensors on the device in place.

TONE_NONE, TONE_CONTRAST, TONE_NON_LINEAR, TONE_EXPOSURE, TONE_AVG_LUMINANCE, TONE_MAX_LUMINANCE, TONE_FERWERDA = \
    0, 1, 4, 5, 6, 7, 9


def _image_arg(img):
    if isinstance(img, np.ndarray):
        a = np.ascontiguousarray(img, np.float32)
        return a, 0
    import torch

    if img.dim() != 3 or img.shape[2] != 3:
        raise ValueError(f"device images are (H, W, 3), got {tuple(img.shape)}")
    _dev(img, torch.float32, None, "image")
    return img, RT_DEVICE_PTRS


def tone_map(ctx, img, mode=TONE_NONE, key=0.18, xw=1.0 / 3.0, yw=1.0 / 3.0, desaturate=False, stretch=False,
             adapt=False, time_step=0.0, adaptation_level=0.0, x_angle=0.0, y_angle=0.0, glare=False):
    """camera::tone_map. Returns (image, adaptation level, log-average
    luminance, device ms); a numpy image is returned as a new array, a
    device tensor is changed in place and returned. x_angle, y_angle: the
    camera's field of view in radians (2 w / t, 2 h / t), for the histogram
    modes and the glare filter (glare=True)."""
    a, flags = _image_arg(img)
    if flags == 0:
        a = a.copy()
    h, w = a.shape[0], a.shape[1]
    tp = rt_tone_params(int(mode), key, xw, yw, int(desaturate), int(stretch), int(adapt), time_step,
                        adaptation_level, 0.0, x_angle, y_angle, int(glare))
    ms = C.c_float()
    ctx.check(_LIB.rt_image_tone_map(ctx.h, _ptr(a), w, h, C.byref(tp), flags, C.byref(ms)), "rt_image_tone_map")
    return a, tp.adaptation_level, tp.log_avg_luminance, ms.value


def glare_filter(ctx, x_res, y_res, x_angle, level):
    """The glare filter (y_res, x_res, 3) float32 of a camera with field of
    view x_angle = 2 w / t at light level 0 scotopic, 1 mesopic, 2 photopic."""
    out = np.empty((y_res, x_res, 3), np.float32)
    ctx.check(_LIB.rt_glare_filter(ctx.h, x_res, y_res, x_angle, level, _ptr(out), 0), "rt_glare_filter")
    return out


def gamma_correct(ctx, img, gamma):
    """camera::gamma_correct; returns (image, device ms) as tone_map does."""
    a, flags = _image_arg(img)
    if flags == 0:
        a = a.copy()
    ms = C.c_float()
    ctx.check(_LIB.rt_image_gamma(ctx.h, _ptr(a), a.numel() // 3 if flags else a.size // 3, gamma, flags,
                                  C.byref(ms)), "rt_image_gamma")
    return a, ms.value


def clip(ctx, img, x_samples=1, y_samples=1, bgr=False, out=None):
    """Down-sampled, clamped bytes (H / y_samples, W / x_samples, 3): a numpy
    array for a numpy image; for a device image, out (a uint8 tensor) or a
    new one. Returns (bytes, device ms)."""
    a, flags = _image_arg(img)
    h, w = a.shape[0], a.shape[1]
    oh, ow = h // y_samples, w // x_samples
    if flags == 0:
        o = np.zeros((oh, ow, 3), np.uint8)
    else:
        import torch

        o = out if out is not None else torch.empty((oh, ow, 3), dtype=torch.uint8, device=a.device)
        _dev(o, torch.uint8, 3 * oh * ow, "rt_image_clip: out")
    ms = C.c_float()
    ctx.check(_LIB.rt_image_clip(ctx.h, _ptr(a), w, h, x_samples, y_samples, int(bgr), _ptr(o), flags, C.byref(ms)),
              "rt_image_clip")
    return o, ms.value


def write_png(path, rgb_bytes):
    """rgb_bytes: (H, W, 3) uint8 in camera::image row order (row 0 at the
    bottom of the picture)."""
    b = np.ascontiguousarray(rgb_bytes, np.uint8)
    if _LIB.rt_write_png(str(path).encode(), _ptr(b), b.shape[1], b.shape[0]) != RT_OK:
        raise RTError(f"rt_write_png: cannot write {path}")


def write_tga(path, bgr_bytes):
    b = np.ascontiguousarray(bgr_bytes, np.uint8)
    if _LIB.rt_write_tga(str(path).encode(), _ptr(b), b.shape[1], b.shape[0]) != RT_OK:
        raise RTError(f"rt_write_tga: cannot write {path}")


def write_jpeg(ctx, path, rgb_bytes, quality=90):
    """rgb_bytes: (H, W, 3) uint8 in camera::image row order, a numpy array
    or a device tensor; written as a baseline JPEG (last row first, as
    camera::write_jpeg_file). Returns the device ms of the transform."""
    ms = C.c_float()
    if isinstance(rgb_bytes, np.ndarray):
        b = np.ascontiguousarray(rgb_bytes, np.uint8)
        flags = 0
    else:
        import torch

        b = rgb_bytes.contiguous()
        if b.dim() != 3 or b.shape[2] != 3:
            raise ValueError(f"rt_write_jpeg: bytes are (H, W, 3), got {tuple(b.shape)}")
        _dev(b, torch.uint8, None, "rt_write_jpeg: bytes")
        flags = RT_DEVICE_PTRS
    ctx.check(_LIB.rt_write_jpeg(ctx.h, str(path).encode(), _ptr(b), b.shape[1], b.shape[0], int(quality), flags,
                                 C.byref(ms)), "rt_write_jpeg")
    return ms.value
