"""Pins the oracle's frame arithmetic by closed forms evaluated in float64 from
the scene geometry, independently of the oracle's own code path:

* primary rays (rt/camera.h:54-62, 127-145): direction = normalise(X x_t +
  Y y_t + Z t) with x_t = x (2w / x_res) - w, y_t = y (2h / y_res) - h; the
  pixel corner (no +0.5), row 0 at the bottom;
* Lambert shading with the light falloff (rt/materials/phong_shader.cc:21-64,
  rt/light.h:44-59): c = sum over lights of (rgb / 255) / ((len d)^2 + 1) *
  max(0, dot(L_dir, n_s)) * 255 albedo * visibility, ka = ks = 0;
* the hit normal turned to face the ray (rt/triangle.h:464-498) and the shadow
  ray from the hit to the light centre with length |L - o| (rt/ray.cc:91-104,
  rt/light.h:96-100, rt/ray.h:44-54);
* the project-defined diffuse bounce: c = direct(x1) + (kd1 / 255) * direct(x2)
  at the bounce hit x2.

Each test fails for a plausible slip: a pixel-centre offset, a flipped row
order, a missing +1 or a misplaced square in the falloff, a missing /255, an
unflipped normal, the bounce weighted by the wrong albedo, a shadow ray that
does not stop at the light.
"""
from collections import namedtuple

import numpy as np
import pytest

import oracle.oracle as O

Cam = namedtuple("Cam", "pos x_axis y_axis z_axis w h t x_res y_res")
Lt = namedtuple("Lt", "pos rgb d")


def f32(*v):
    return tuple(float(np.float32(x)) for x in v)


def expected_dirs(cam):
    """Primary directions by the camera formula, in float64."""
    X, Y, Z = (np.asarray(a, np.float64) for a in (cam.x_axis, cam.y_axis, cam.z_axis))
    xs = np.arange(cam.x_res, dtype=np.float64) * (2.0 * cam.w / cam.x_res) - cam.w
    ys = np.arange(cam.y_res, dtype=np.float64) * (2.0 * cam.h / cam.y_res) - cam.h
    xt, yt = np.meshgrid(xs, ys)  # row-major: index y * x_res + x
    d = xt[..., None] * X + yt[..., None] * Y + cam.t * Z
    d = d.reshape(-1, 3)
    return d / np.linalg.norm(d, axis=1, keepdims=True)


def test_primary_rays_follow_the_camera_formula():
    # A rotated orthonormal basis (rounded to float32) and an off-centre position.
    ang = 0.3
    X = f32(np.cos(ang), 0.0, -np.sin(ang))
    Y = f32(0.0, 1.0, 0.0)
    Z = f32(np.sin(ang), 0.0, np.cos(ang))
    cam = Cam(f32(0.25, -0.5, 1.0), X, Y, Z, 0.4, 0.3, 1.0, 8, 6)
    rays = O.primary_rays(cam)
    exp = expected_dirs(cam)
    np.testing.assert_allclose(rays[:, 4:7], exp, rtol=0, atol=3e-7)
    assert np.all(rays[:, 0:3] == np.asarray(cam.pos, np.float32))
    assert np.all(rays[:, 3] == np.float32(1e-3))
    # Pixel (0, 0) is the bottom-left corner: its screen point is exactly
    # (-w, -h), not the pixel centre.
    corner = np.asarray(X) * -0.4 + np.asarray(Y) * -0.3 + np.asarray(Z)
    np.testing.assert_allclose(rays[0, 4:7], corner / np.linalg.norm(corner), atol=3e-7)
    # A pixel centre would move the ray by half a pixel (0.05 of the screen).
    centre = np.asarray(X) * (-0.4 + 0.05) + np.asarray(Y) * (-0.3 + 0.05) + np.asarray(Z)
    assert np.abs(rays[0, 4:7] - centre / np.linalg.norm(centre)).max() > 1e-2
    # Row 0 looks down (negative Y component), the last row up.
    assert np.dot(rays[0, 4:7], Y) < 0 < np.dot(rays[-1, 4:7], Y)


def test_centre_pixel_of_a_symmetric_camera_looks_along_z():
    cam = Cam((0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0), 0.4, 0.3, 1.0, 8, 6)
    rays = O.primary_rays(cam)
    c = 3 * 8 + 4
    assert tuple(rays[c, 4:7]) == (0.0, 0.0, 1.0)


# ---------------------------------------------------------------- single-plane Lambert frames

def wall(z, half=50.0, flip=False):
    """Two triangles spanning the plane z = const (normal along -z, or +z flipped)."""
    a, b, c, d = (-half, -half, z), (half, -half, z), (half, half, z), (-half, half, z)
    t1, t2 = [*a, *b, *c], [*a, *c, *d]
    if flip:
        t1, t2 = [*a, *c, *b], [*a, *d, *c]
    return np.asarray([t1, t2], np.float32)


CAM = Cam((0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0), 0.5, 0.5, 1.0, 16, 16)


def lambert(o, n_s, kd, lights, vis=None):
    """Direct light at a surface point o (the shadow-ray origin) in float64."""
    c = np.zeros(3)
    for k, l in enumerate(lights):
        v = np.asarray(l.pos, np.float64) - o
        ln = np.linalg.norm(v)
        s = np.dot(v / ln, n_s)
        if s < 0:
            continue
        scale = (ln * l.d) ** 2 + 1.0
        w = 1.0 if vis is None else vis[k]
        c += (np.asarray(l.rgb, np.float64) / 255.0) / scale * s * kd * w
    return c


def plane_hits(cam, z):
    """First-hit points on the plane z = const of every primary ray (float64
    geometry on the float32 directions), and the shadow-ray origins: the
    point rounded to float32 and moved off the surface towards the camera."""
    rays = O.primary_rays(cam)
    d = rays[:, 4:7].astype(np.float64)
    o = rays[:, 0:3].astype(np.float64)
    t = (z - o[:, 2]) / d[:, 2]
    p = o + d * t[:, None]
    n_s = np.array([0.0, 0.0, -1.0 if cam.pos[2] < z else 1.0])
    org = np.stack([O.offset_start_point(q.astype(np.float32), n_s.astype(np.float32)) for q in p]).astype(np.float64)
    return p, org, n_s


@pytest.mark.parametrize("falloff", [0.0, 0.25, 1.5])
@pytest.mark.parametrize("flip", [False, True])
def test_single_plane_lambert_radiance(falloff, flip):
    abc = wall(2.0, flip=flip)
    albedo = np.asarray([[0.6, 0.3, 0.9], [0.6, 0.3, 0.9]], np.float32)
    lights = [Lt((0.3, 0.8, 0.5), (255.0, 128.0, 64.0), falloff)]
    res = O.OracleScene(abc).render(albedo, CAM, lights)
    assert np.all(res["prim"] >= 0)
    _, org, n_s = plane_hits(CAM, 2.0)
    kd = 255.0 * albedo[0].astype(np.float64)
    exp = np.stack([lambert(q, n_s, kd, lights) for q in org])
    # Whichever way the triangles wind, the shading normal faces the camera.
    assert exp.min() > 0
    np.testing.assert_allclose(res["rgb"], exp, rtol=2e-6, atol=1e-6)


def test_two_lights_sum_and_a_light_behind_the_surface_adds_nothing():
    abc = wall(2.0)
    albedo = np.full((2, 3), 0.5, np.float32)
    front = Lt((-0.4, 0.2, 0.3), (200.0, 255.0, 10.0), 0.5)
    behind = Lt((0.0, 0.0, 6.0), (255.0, 255.0, 255.0), 0.0)
    res = O.OracleScene(abc).render(albedo, CAM, [front, behind])
    _, org, n_s = plane_hits(CAM, 2.0)
    kd = 255.0 * 0.5
    exp = np.stack([lambert(q, n_s, kd, [front]) for q in org])
    np.testing.assert_allclose(res["rgb"], exp, rtol=2e-6, atol=1e-6)
    alone = O.OracleScene(abc).render(albedo, CAM, [behind])
    assert np.all(alone["rgb"] == 0.0)
    # The light behind the surface is not occluded: it is turned away.
    assert np.all(alone["vis"][:, 0] == 1)


def seg_hits_triangle(o, l, tri):
    """Float64 Moller-Trumbore of the segment o -> l against one triangle,
    with the margin of the decision (distance of the crossing from the
    triangle's edges in barycentric units)."""
    a, b, c = (np.asarray(tri[3 * k:3 * k + 3], np.float64) for k in range(3))
    d = l - o
    e1, e2 = b - a, c - a
    p = np.cross(d, e2)
    det = np.dot(e1, p)
    inv = 1.0 / det
    s = o - a
    u = np.dot(s, p) * inv
    q = np.cross(s, e1)
    v = np.dot(d, q) * inv
    t = np.dot(e2, q) * inv
    margin = min(u, v, 1.0 - u - v)
    return (margin > 0) and (0 < t < 1), abs(margin)


def test_occluder_shadows_exactly_the_pixels_whose_light_segment_it_crosses():
    occ = [-0.3, -0.2, 1.7, 0.4, -0.1, 1.7, 0.0, 0.5, 1.7]  # a triangle in the plane z = 1.7
    abc = np.concatenate([wall(2.0), np.asarray([occ], np.float32)])
    albedo = np.full((3, 3), 0.7, np.float32)
    # Off to the side, so that the shadow falls on wall the camera sees.
    light = Lt((3.0, 0.05, 0.5), (255.0, 255.0, 255.0), 0.0)
    res = O.OracleScene(abc).render(albedo, CAM, [light])
    # Primary rays that hit the occluder itself are not on the wall.
    on_wall = res["prim"] < 2
    _, org, n_s = plane_hits(CAM, 2.0)
    kd = 255.0 * 0.7
    shadowed = lit = 0
    for i in np.flatnonzero(on_wall):
        hit, margin = seg_hits_triangle(org[i], np.asarray(light.pos), occ)
        if margin < 1e-4:
            continue  # too close to an edge to decide in float64
        exp = 0.0 if hit else lambert(org[i], n_s, kd, [light])
        np.testing.assert_allclose(res["rgb"][i], exp, rtol=2e-6, atol=1e-6)
        assert res["vis"][i, 0] == (0 if hit else 1)
        shadowed += hit
        lit += not hit
    assert shadowed > 10 and lit > 10


def test_shadow_ray_stops_at_the_light():
    # An occluder beyond the light (on the far side from the wall) must not
    # shadow: the shadow ray's length is |L - o|.
    beyond = [-5.0, -5.0, -2.0, 5.0, -5.0, -2.0, 0.0, 5.0, -2.0]
    abc = np.concatenate([wall(2.0), np.asarray([beyond], np.float32)])
    albedo = np.full((3, 3), 0.4, np.float32)
    light = Lt((0.0, 0.0, -1.0), (255.0, 255.0, 255.0), 0.0)
    res = O.OracleScene(abc).render(albedo, CAM, [light])
    assert np.all(res["prim"] < 2)
    assert np.all(res["vis"][:, 0] == 1)
    _, org, n_s = plane_hits(CAM, 2.0)
    exp = np.stack([lambert(q, n_s, 255.0 * 0.4, [light]) for q in org])
    np.testing.assert_allclose(res["rgb"], exp, rtol=2e-6, atol=1e-6)


# ---------------------------------------------------------------- bounce composition

def slab(y, half, up):
    """Two triangles spanning the plane y = const."""
    a, b, c, d = (-half, y, -half), (half, y, -half), (half, y, half), (-half, y, half)
    t1, t2 = [*a, *b, *c], [*a, *c, *d]
    if not up:
        t1, t2 = [*a, *c, *b], [*a, *d, *c]
    return [t1, t2]


def test_bounce_adds_kd1_over_255_times_direct_light_at_the_bounce_hit():
    half = 1e4
    abc = np.asarray(slab(0.0, half, True) + slab(1.0, half, False), np.float32)
    alb_floor, alb_ceil = np.array([0.8, 0.5, 0.2]), np.array([0.3, 0.6, 0.9])
    albedo = np.asarray([alb_floor, alb_floor, alb_ceil, alb_ceil], np.float32)
    light = Lt((0.2, 0.5, 2.0), (255.0, 200.0, 150.0), 0.5)
    # Camera between the planes, looking along +z: rows below the horizon see
    # the floor, rows above it the ceiling.
    cam = Cam((0.0, 0.5, 0.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0), 0.5, 0.5, 1.0, 12, 12)
    seed = 77
    res = O.OracleScene(abc).render(albedo, cam, [light], bounce=True, seed=seed)
    rays = O.primary_rays(cam)
    checked = 0
    for p in range(cam.x_res * cam.y_res):
        d = rays[p, 4:7].astype(np.float64)
        if abs(d[1]) < 1e-6:
            continue
        floor = d[1] < 0
        y1 = 0.0 if floor else 1.0
        t1 = (y1 - 0.5) / d[1]
        p1 = np.asarray(cam.pos, np.float64) + d * t1
        n1 = np.array([0.0, 1.0 if floor else -1.0, 0.0])
        o1 = O.offset_start_point(p1.astype(np.float32), n1.astype(np.float32)).astype(np.float64)
        kd1 = 255.0 * (alb_floor if floor else alb_ceil)
        c = lambert(o1, n1, kd1, [light])
        # The bounce direction comes from the (separately pinned) sampler.
        bd = O.bounce_dir(seed, p, 0, 1, n1.astype(np.float32)).astype(np.float64)
        y2 = 1.0 if floor else 0.0
        t2 = (y2 - o1[1]) / bd[1]
        p2 = o1 + bd * t2
        assert res["bounce_prim"][p] >= 0
        n2 = -n1
        o2 = O.offset_start_point(p2.astype(np.float32), n2.astype(np.float32)).astype(np.float64)
        kd2 = 255.0 * (alb_ceil if floor else alb_floor)
        c = c + (kd1 / 255.0) * lambert(o2, n2, kd2, [light])
        np.testing.assert_allclose(res["rgb"][p], c, rtol=1e-5, atol=1e-5)
        checked += 1
    assert checked >= 100


def test_bounce_with_unit_albedo_adds_the_bounce_hits_direct_light_unscaled():
    # kd1 = 255: the throughput kd1 / 255 is exactly 1.
    abc = np.asarray(slab(0.0, 1e4, True) + slab(1.0, 1e4, False), np.float32)
    albedo = np.ones((4, 3), np.float32)
    light = Lt((0.0, 0.5, 1.0), (255.0, 255.0, 255.0), 0.0)
    cam = Cam((0.0, 0.5, 0.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0), 0.5, 0.5, 1.0, 8, 8)
    sc = O.OracleScene(abc)
    with_b = sc.render(albedo, cam, [light], bounce=True, seed=5)
    without = sc.render(albedo, cam, [light], bounce=False)
    extra = with_b["rgb"] - without["rgb"]
    hit = with_b["prim"] >= 0
    # The horizon row is parallel to both planes and misses: no bounce there.
    assert 0 < hit.sum() < hit.size and np.all(extra[~hit] == 0.0)
    # Every bounce hit is lit by a light between the planes, so the bounce
    # term is positive, and at most the light's unattenuated 255.
    assert np.all(with_b["bounce_prim"][hit] >= 0)
    assert np.all(extra[hit] > 0) and np.all(extra[hit] <= 255.0)
