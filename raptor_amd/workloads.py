"""Synthetic workloads: scene, camera and light of each benchmark configuration.

* random65k  — 65,536 random triangles (seed 1; centroids uniform in [0,1)^3,
  vertices offset within +-0.01, albedo uniform in [0,1)^3), pinhole camera at
  (0.5, 0.5, -1.5) looking +z with half widths 0.35 at distance 1, 256x256,
  point light (0.5, 2, 0.5), one shadow ray per hit.
* terrain1m  — 708 x 708 heightfield (999,698 triangles; fBm value noise,
  8 octaves, lacunarity 2, gain 0.5, amplitude 0.15, zoom 0.02, noise seed 57,
  albedo uniform in [0.2, 0.8)^3 with seed 2), 1920x1080 primary rays + one
  shadow ray per hit + one cosine-weighted bounce with its shadow ray. Camera
  at (0.5, 0.55, -0.35) looking at (0.5, 0, 0.6), half width 0.5.

A light is rt/light.h:17-18's point light: rgb in [0,255], falloff d, radius 0.
"""
from dataclasses import dataclass, field

import numpy as np

from . import rt


@dataclass
class Camera:
    pos: tuple
    x_axis: tuple
    y_axis: tuple
    z_axis: tuple
    w: float
    h: float
    t: float
    x_res: int
    y_res: int

    def to_c(self):
        c = rt.rt_camera()
        for k in range(3):
            c.pos[k] = self.pos[k]
            c.x_axis[k] = self.x_axis[k]
            c.y_axis[k] = self.y_axis[k]
            c.z_axis[k] = self.z_axis[k]
        c.w, c.h, c.t = self.w, self.h, self.t
        c.x_res, c.y_res = self.x_res, self.y_res
        return c

    def as_dict(self):
        return dict(pos=list(self.pos), x_axis=list(self.x_axis), y_axis=list(self.y_axis), z_axis=list(self.z_axis),
                    w=self.w, h=self.h, t=self.t, x_res=self.x_res, y_res=self.y_res)


@dataclass
class Light:
    pos: tuple
    rgb: tuple = (255.0, 255.0, 255.0)
    d: float = 0.0

    def to_c(self):
        l = rt.rt_light()
        for k in range(3):
            l.pos[k] = self.pos[k]
            l.rgb[k] = self.rgb[k]
        l.d = self.d
        return l


def f32(v):
    return tuple(float(np.float32(x)) for x in v)


def look_at(pos, target, up=(0.0, 1.0, 0.0)):
    """Camera basis (x, y, z) for a camera at pos looking at target, in float64, rounded to float32."""
    p = np.asarray(pos, np.float64)
    z = np.asarray(target, np.float64) - p
    z /= np.linalg.norm(z)
    x = np.cross(np.asarray(up, np.float64), z)
    x /= np.linalg.norm(x)
    y = np.cross(z, x)
    return f32(x + 0.0), f32(y + 0.0), f32(z + 0.0)


def fixture_camera(pos, x_axis, y_axis, z_axis, x_res=640, y_res=480):
    """The camera the reference's regression fixture builds for OFF, PLY and
    OBJ scenes (rt/regression_tests/regression_fixture.h:86-100, 150-168):
    screen width 10, height 10 * y_res / x_res, image plane at 20."""
    w = np.float32(10.0)
    h = w * (np.float32(y_res) / np.float32(x_res))
    return Camera(f32(pos), f32(x_axis), f32(y_axis), f32(z_axis), float(w), float(h), 20.0, x_res, y_res)


@dataclass
class Workload:
    name: str
    abc: np.ndarray
    albedo: np.ndarray
    camera: Camera
    lights: list
    bounce: bool
    spp: int = 1
    seed: int = 0
    meta: dict = field(default_factory=dict)


def random65k(n=65536, res=256):
    abc, alb = rt.gen_random_triangles(1, n, 0.01, 0.0, 1.0)
    cam = Camera(f32((0.5, 0.5, -1.5)), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0), float(np.float32(0.35)),
                 float(np.float32(0.35)), 1.0, res, res)
    return Workload("random65k", abc, alb, cam, [Light(f32((0.5, 2.0, 0.5)))], bounce=False,
                    meta=dict(triangles=n, resolution=[res, res]))


def terrain_arrays(grid, amplitude=0.15, zoom=0.02, noise_seed=57, albedo_seed=2, lo=0.2, hi=0.8):
    """fBm terrain of grid x grid vertices (rt.gen_terrain) with at least the
    1M terrain's cell size: a grid beyond 708 covers [0, (grid - 1) / 707]^2
    in x and z instead of the unit square (heights unchanged, zoom per cell
    unchanged, so the relief per unit of distance is the 1M terrain's). The
    reference's Moller-Trumbore rejects |det| < 1e-6 (rt/triangle.h:190-235),
    and det is twice the triangle's area times the cosine to the ray: on the
    unit square a 1414^2 grid has 31% of its triangles below that area and a
    2795^2 grid none above it, so they would be invisible to every ray."""
    abc, alb = rt.gen_terrain(grid, amplitude, zoom, 0.5, 8, noise_seed, albedo_seed, lo, hi)
    ext = max(1.0, (grid - 1) / 707.0)
    if ext > 1.0:
        v = abc.reshape(-1, 3, 3)
        v[:, :, 0] *= np.float32(ext)
        v[:, :, 2] *= np.float32(ext)
        abc = v.reshape(-1, 9)
    return abc, alb


def terrain(grid=708, x_res=1920, y_res=1080, bounce=True, seed=7, zoom=0.02, amplitude=0.15):
    abc, alb = terrain_arrays(grid, amplitude, zoom)
    pos = (0.5, 0.55, -0.35)
    x, y, z = look_at(pos, (0.5, 0.0, 0.6))
    w = 0.5
    h = w * y_res / x_res
    cam = Camera(f32(pos), x, y, z, float(np.float32(w)), float(np.float32(h)), 1.0, x_res, y_res)
    return Workload(f"terrain{grid}", abc, alb, cam, [Light(f32((0.5, 2.0, 0.5)))], bounce=bounce, seed=seed,
                    meta=dict(triangles=int(abc.shape[0]), grid=grid, resolution=[x_res, y_res]))


def terrain_wave(abc, phase, amplitude=0.01, wavelength=0.25):
    """Terrain vertices under a travelling wave: height (y) += amplitude *
    sin(2 pi (x / wavelength - phase)) * cos(2 pi z / wavelength). Works on
    numpy arrays and torch tensors of shape (n, 9)."""
    mod = np if isinstance(abc, np.ndarray) else __import__("torch")
    v = abc.reshape(-1, 3, 3)
    x, z = v[:, :, 0], v[:, :, 2]
    k = 2.0 * np.pi / wavelength
    dy = amplitude * mod.sin(k * x - 2.0 * np.pi * phase) * mod.cos(k * z)
    out = v.clone() if mod is not np else v.copy()
    out[:, :, 1] = v[:, :, 1] + dy.astype(np.float32) if mod is np else v[:, :, 1] + dy
    return out.reshape(abc.shape)


def drift(abc, step, seed=3, speed=0.002):
    """Rigid per-triangle motion (bodies of a physics step): triangle i moves
    by step * speed * u_i with u_i uniform in [-1, 1]^3 (seeded). numpy."""
    rng = np.random.default_rng(seed)
    vel = rng.uniform(-1.0, 1.0, (abc.shape[0], 1, 3)).astype(np.float32)
    v = abc.reshape(-1, 3, 3) + np.float32(step * speed) * vel
    return v.reshape(abc.shape).astype(np.float32)


def quad(p0, p1, p2, p3):
    """Two triangles (p0, p1, p2) and (p0, p2, p3) as a (2, 9) float32 array."""
    p0, p1, p2, p3 = (np.asarray(p, np.float32) for p in (p0, p1, p2, p3))
    return np.stack([np.concatenate([p0, p1, p2]), np.concatenate([p0, p2, p3])]).astype(np.float32)


def material_scene(n_objects=400, res=96, seed=5, mirror=None, glass=None, mirrored=False, light=(0.3, 3.0, -0.2),
                   size=0.08, floor=True):
    """Small scene for the phong materials: a diffuse floor at y = -1 (unless
    floor is False), random triangles of the given size (diffuse, albedo
    seeded) in x, z in [-0.8, 0.8], y in [0.2, 0.7], one light, a camera at
    (0, 1.6, -2.2) looking at (0, 0, 0.3). mirror: reflectance of a mirror wall
    at z = 1 behind the objects (kd = ks = 0); glass: (tran, ri) of a
    transparent wall at z = -0.9 between the camera and the objects.
    mirrored: the objects and the light reflected about z = 1 and no wall, the
    scene that the mirror shows. Returns (Workload, materials, tri_material,
    flags); the floor and the objects take the Lambert default (-1)."""
    rng = np.random.default_rng(seed)
    c = rng.uniform((-0.8, 0.2, -0.8), (0.8, 0.7, 0.8), (n_objects, 1, 3))
    objs = (c + rng.uniform(-size, size, (n_objects, 3, 3))).astype(np.float32).reshape(-1, 9)
    if mirrored:
        v = objs.reshape(-1, 3, 3)
        v[:, :, 2] = np.float32(2.0) - v[:, :, 2]
        light = (light[0], light[1], 2.0 - light[2])
    parts = ([quad((-3, -1, -3), (3, -1, -3), (3, -1, 3), (-3, -1, 3))] if floor else []) + [objs]
    kinds = [-1] * sum(q.shape[0] for q in parts)
    flags = [0] * len(kinds)
    materials = []
    if mirror is not None and not mirrored:
        parts.append(quad((-3, -1, 1), (3, -1, 1), (3, 3, 1), (-3, 3, 1)))
        materials.append(dict(rf=float(mirror)))
        kinds += [len(materials) - 1] * 2
        flags += [0, 0]
    if glass is not None:
        tran, ri = glass
        parts.append(quad((-1.5, -1.2, -0.9), (1.5, -1.2, -0.9), (1.5, 1.6, -0.9), (-1.5, 1.6, -0.9)))
        materials.append(dict(tran=float(tran), ri=float(ri)))
        kinds += [len(materials) - 1] * 2
        flags += [2, 2]  # transparent: shadow rays pass (the reference marks materials with tran > 0)
    abc = np.concatenate(parts).astype(np.float32)
    albedo = rng.uniform(0.2, 0.9, (abc.shape[0], 3)).astype(np.float32)
    pos = (0.0, 1.6, -2.2)
    x, y, z = look_at(pos, (0.0, 0.0, 0.3))
    cam = Camera(f32(pos), x, y, z, 0.6, 0.6, 1.0, res, res)
    w = Workload("materials", abc, albedo, cam, [Light(f32(light), (255.0, 240.0, 230.0), 0.1)], bounce=False,
                 meta=dict(triangles=int(abc.shape[0])))
    return w, materials, np.asarray(kinds, np.int32), np.asarray(flags, np.uint8)


def mesh_camera_lights(mesh):
    """The Camera, background and Lights a parsed file defines (NFF)."""
    c = mesh.camera
    cam = Camera(tuple(c.pos), tuple(c.x_axis), tuple(c.y_axis), tuple(c.z_axis), c.w, c.h, c.t, c.x_res, c.y_res)
    lights = [Light(tuple(float(x) for x in r[0:3]), tuple(float(x) for x in r[3:6]), float(r[6]))
              for r in (mesh.lights if mesh.lights is not None else [])]
    return cam, tuple(float(x) for x in mesh.background), lights


def mesh_view(abc, x_res=320, y_res=240, direction=(0.45, 0.35, -0.82), light_rgb=(255.0, 255.0, 255.0)):
    """A fixture camera that frames a parsed mesh from the given direction
    (the view direction, camera to mesh), and a point light on the camera's
    side of the mesh, above and to the side of the camera."""
    v = np.asarray(abc, np.float64).reshape(-1, 3)
    lo, hi = v.min(0), v.max(0)
    centre, radius = 0.5 * (lo + hi), 0.5 * np.linalg.norm(hi - lo)
    d = np.asarray(direction, np.float64)
    d /= np.linalg.norm(d)
    pos = centre - d * (radius * 3.0)
    x, y, z = look_at(pos, centre)
    cam = fixture_camera(pos, x, y, z, x_res, y_res)
    side = np.cross(d, (0.0, 1.0, 0.0))
    side = side / np.linalg.norm(side) if np.linalg.norm(side) > 0 else np.array([1.0, 0.0, 0.0])
    light = centre - d * (radius * 2.5) + radius * (1.5 * side + np.array([0.0, 2.0, 0.0]))
    return Workload("mesh", np.asarray(abc, np.float32), None, cam, [Light(f32(light), light_rgb, 0.0)],
                    bounce=False, meta=dict(triangles=int(v.shape[0] // 3)))


def material_terrain(grid=708, x_res=1920, y_res=1080):
    """The terrain of terrain1m with phong materials: each grid cell takes one
    of four materials (seeded), one of them wet rock (rf 0.25, shiny); a
    mirror wall (rf 0.8) stands behind the far edge at z = 1.02 and a glass
    pane (tran 0.6, ri 1.5) covers the left of the view at z = -0.1. One
    light. Returns (Workload, materials, tri_material, flags)."""
    w = terrain(grid, x_res, y_res, bounce=False)
    n = w.abc.shape[0]
    mirror = quad((-0.3, -0.2, 1.02), (1.3, -0.2, 1.02), (1.3, 1.2, 1.02), (-0.3, 1.2, 1.02))
    glass = quad((-0.2, -0.3, -0.1), (0.45, -0.3, -0.1), (0.45, 0.9, -0.1), (-0.2, 0.9, -0.1))
    abc = np.concatenate([w.abc, mirror, glass]).astype(np.float32)
    albedo = np.concatenate([w.albedo, np.full((4, 3), 0.5, np.float32)])
    materials = [
        dict(kd=(150.0, 130.0, 100.0), ks=(40.0, 40.0, 40.0), s=8.0),
        dict(kd=(90.0, 140.0, 70.0), ka=(3.0, 5.0, 2.0)),
        dict(kd=(120.0, 120.0, 125.0), ks=(120.0, 120.0, 120.0), s=32.0, rf=0.25),
        dict(kd=(170.0, 150.0, 120.0), ks=(20.0, 20.0, 20.0), s=4.0),
        dict(rf=0.8, ks=(30.0, 30.0, 30.0), s=64.0),
        dict(tran=0.6, ri=1.5, ks=(60.0, 60.0, 60.0), s=48.0),
    ]
    cell = np.arange(n, dtype=np.uint64) // np.uint64(2)
    kinds = ((cell * np.uint64(2654435761)) >> np.uint64(13)) % np.uint64(4)
    kinds = np.concatenate([kinds.astype(np.int32), [4, 4, 5, 5]]).astype(np.int32)
    flags = np.zeros(abc.shape[0], np.uint8)
    flags[-2:] = 2  # the glass is transparent to shadow rays
    out = Workload("material_terrain", abc, albedo, w.camera, w.lights, bounce=False,
                   meta=dict(triangles=int(abc.shape[0]), grid=grid, resolution=[x_res, y_res]))
    return out, materials, kinds, flags


def grass_image(size=256, seed=3):
    """A procedural RGB texture (size x size, 0-255): value noise of a few
    octaves tinted green-brown, seeded."""
    rng = np.random.default_rng(seed)
    acc = np.zeros((size, size), np.float64)
    amp = 1.0
    for cells in (8, 16, 32, 64):
        g = rng.random((cells + 1, cells + 1))
        t = np.linspace(0.0, cells, size, endpoint=False)
        i = t.astype(int)
        f = t - i
        f = f * f * (3.0 - 2.0 * f)
        top = g[i][:, i] * (1 - f)[None, :] + g[i][:, i + 1] * f[None, :]
        bot = g[i + 1][:, i] * (1 - f)[None, :] + g[i + 1][:, i + 1] * f[None, :]
        acc += amp * (top * (1 - f)[:, None] + bot * f[:, None])
        amp *= 0.5
    acc = (acc - acc.min()) / (acc.max() - acc.min())
    rgb = np.stack([60 + 90 * acc, 110 + 120 * acc, 40 + 50 * acc], -1)
    return rgb.astype(np.float32)


def textured_terrain(grid=708, x_res=1920, y_res=1080):
    """The terrain of terrain1m shaded as RAPTor's terrain generator does
    (RAPTor/terrain_generator/main.cc:141-142): a coloured_mapper_shader with
    black ka and ks, white kd and a planar map (normal y, tile / tile) on kd,
    here a procedural 256 x 256 image repeated every 0.1 units. Returns
    (Workload, materials, tri_material, flags, maps, images)."""
    from . import rt as _rt

    w = terrain(grid, x_res, y_res, bounce=False)
    n = w.abc.shape[0]
    maps = [_rt.image_map(_rt.MAP_PLANAR, 0, (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.1, 0.0, 0.1), _rt.WRAP_TILE,
                          _rt.WRAP_TILE)]
    materials = [dict(kind=_rt.MATERIAL_COLOURED, kd=(255.0, 255.0, 255.0), ri=1.0,
                      tex=(None, 0, None, None, None, None))]
    out = Workload("textured_terrain", w.abc, w.albedo, w.camera, w.lights, bounce=False,
                   meta=dict(triangles=int(n), grid=grid, resolution=[x_res, y_res]))
    return out, materials, np.zeros(n, np.int32), np.zeros(n, np.uint8), maps, [grass_image()]


def terrain1m(x_res=1920, y_res=1080):
    return terrain(708, x_res, y_res)


def diff_terrain(grid=1414, x_res=1024, y_res=1024, spp=4, seed=11, bounce=True, zoom=0.02, amplitude=0.15):
    """Differentiable frame: grid x grid fBm terrain (as terrain(), albedo seed
    2), the terrain camera, spp jittered samples per pixel, direct light + one
    diffuse bounce, four point lights on a ring of radius 0.6 at height 2 with
    falloff d = 1. The target is the same frame with albedo * (1 + 0.2 n),
    n ~ N(0, 1) (seeded), clamped to [0, 1]: meta["target_albedo"]. Small
    test grids pass a larger zoom (and amplitude) to keep the relief."""
    w = terrain(grid, x_res, y_res, bounce=bounce, seed=seed, zoom=zoom, amplitude=amplitude)
    lights = []
    colours = [(255.0, 240.0, 220.0), (200.0, 220.0, 255.0), (255.0, 200.0, 200.0), (220.0, 255.0, 210.0)]
    for k in range(4):
        a = 2.0 * np.pi * (k + 0.5) / 4.0
        lights.append(Light(f32((0.5 + 0.6 * np.cos(a), 2.0, 0.5 + 0.6 * np.sin(a))), colours[k], 1.0))
    rng = np.random.default_rng(seed)
    target = np.clip(w.albedo * (1.0 + 0.2 * rng.standard_normal(w.albedo.shape)), 0.0, 1.0).astype(np.float32)
    return Workload(f"diff_terrain{grid}", w.abc, w.albedo, w.camera, lights, bounce=bounce, spp=spp, seed=seed,
                    meta=dict(triangles=int(w.abc.shape[0]), grid=grid, resolution=[x_res, y_res], spp=spp,
                              target_albedo=target))


@dataclass
class BatchScene:
    kind: str
    abc: np.ndarray
    albedo: np.ndarray
    rays: int
    seed: int


def batched(n_scenes=64, tri_lo=1e3, tri_hi=2e6, side_lo=64, side_hi=1024, seed=3):
    """Many independent scenes: triangle counts log-uniform in [tri_lo, tri_hi],
    terrain (grid ceil(sqrt(N/2)) + 1, cells at least the 1M terrain's, see
    terrain_arrays) or random triangles 50/50, ray batches of
    side x side incoherent surface rays with side log-uniform in [side_lo, side_hi]."""
    rng = np.random.default_rng(seed)
    out = []
    for s in range(n_scenes):
        n = int(np.exp(rng.uniform(np.log(tri_lo), np.log(tri_hi))))
        side = int(np.exp(rng.uniform(np.log(side_lo), np.log(side_hi))))
        if rng.uniform() < 0.5:
            g = int(np.ceil(np.sqrt(n / 2.0))) + 1
            abc, alb = terrain_arrays(g, 0.15, 0.02, 57 + s, 100 + s, 0.2, 0.8)
            kind = "terrain"
        else:
            abc, alb = rt.gen_random_triangles(1000 + s, n, 0.01, 0.0, 1.0)
            kind = "random"
        out.append(BatchScene(kind, abc, alb, side * side, 5000 + s))
    return out


def random_layer(seed, n, albedo_lo=0.2, albedo_hi=0.8, thickness=0.1, spread=0.01):
    """n random triangles (rt.gen_random_triangles: vertices within +-spread
    of their centroid) whose centroids are moved from the unit cube into a
    layer y in [0, thickness) over [0, ext]^2 in x and z, ext the side the
    terrain of the same triangle count covers (terrain_arrays). Inside the
    unit cube a cloud of 0.5M-16M such triangles has a mean free path of
    about 1e-3: the terrain camera sees its front face, whose shadow rays to
    lights above run through the cloud, and the frame is black. In the layer
    the camera sees its top, lit from above."""
    abc, alb = rt.gen_random_triangles(seed, n, spread, albedo_lo, albedo_hi)
    g = int(np.ceil(np.sqrt(n / 2.0))) + 1
    ext = np.float32(max(1.0, (g - 1) / 707.0))
    v = abc.reshape(-1, 3, 3)
    c = v.mean(axis=1, keepdims=True, dtype=np.float32)
    scale = np.array([ext, np.float32(thickness), ext], np.float32)
    v += c * (scale - np.float32(1.0))
    return v.reshape(-1, 9), alb


def invrender_scenes(n_scenes=16, tri_lo=5e5, tri_hi=1.6e7, res=512, seed=21, pick=None):
    """Inverse-rendering batch: triangle counts log-uniform in [tri_lo,
    tri_hi], terrain (grid ceil(sqrt(N/2)) + 1 at the 1M terrain's cell size,
    see terrain_arrays) or a layer of random triangles over the same
    footprint (random_layer) 50/50, each seen res x res by the terrain camera
    under the four ring lights of diff_terrain. The generated albedo is the
    ground truth that inverse rendering should recover. pick: indices of the
    scenes to generate (the draws of the others are made and skipped)."""
    rng = np.random.default_rng(seed)
    base = diff_terrain(grid=16, x_res=res, y_res=res, spp=2, seed=seed)
    out = []
    for s in range(n_scenes):
        n = int(np.exp(rng.uniform(np.log(tri_lo), np.log(tri_hi))))
        terrain_kind = rng.uniform() < 0.5
        if pick is not None and s not in pick:
            continue
        if terrain_kind:
            g = int(np.ceil(np.sqrt(n / 2.0))) + 1
            abc, alb = terrain_arrays(g, 0.15, 0.02, 57 + s, 300 + s, 0.2, 0.8)
            kind = "terrain"
        else:
            abc, alb = random_layer(2000 + s, n)
            kind = "random"
        out.append(Workload(f"{kind}{abc.shape[0]}", abc, alb, base.camera, base.lights, bounce=True, spp=2,
                            seed=seed, meta=dict(kind=kind, triangles=int(abc.shape[0]))))
    return out
