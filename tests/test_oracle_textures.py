"""The oracle's texture maps and textured shaders (rt/materials/
texture_mapper.h, planar_mapper.cc, cubic_mapper.cc, cylindrical_mapper.cc,
checker_board_mapper.cc, mapper_falloff.cc, coloured_mapper_shader.cc,
mapper_shader.cc) pinned by closed forms and hand-evaluated values:

* bilinear lookup: a texture that is a linear function of the texel
  coordinates comes back as that function at (vt.u w, vt.v h), so the
  scale by w and h (not w - 1), the row-major address and the channel order
  are fixed (a transposed operand breaks it when the slopes differ);
* wrapping: tile, mirror, clamp and blank at hand-evaluated texels
  (negative tiles wrap by size - 1, mirror folds by 2 (size - 1));
* planar, cubic and cylindrical projections against their geometry (the
  map's centre offset by half its size, the cubic's extra normal term, the
  cylinder's angle folded into [0, pi]);
* checker board: only the x cell decides (the reference xors the z cell
  with itself), the colour is left where it does not paint, the opacity
  is the alpha;
* falloffs and the inverse by their formulas on one value; monochrome is the
  texel mean, 0 outside a blank image;
* shaders through whole frames: coloured_mapper_shader with uniform maps
  equals the untextured shader with the scaled colours (to the rounding of
  the bilinear weights), t_ns
  gives ns / 255 whatever the texel, t_refl gates the mirror ray while its
  colour keeps the untextured weight; mapper_shader's per-light formula on a
  plane against float64 geometry, and its clrf blend of the refracted
  background;
* the TGA reader: the shipped images' header and pixel bytes read directly,
  and a hand-made run-length file.
"""
import os

import numpy as np
import pytest

from raptor_amd import rt
from raptor_amd.workloads import Camera, Light

import oracle.oracle as O
from oracle import ingest as OI

TEX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "textures")
F = np.float32


def linear_image(w, h, a=(3.0, 5.0, 7.0), b=(11.0, 2.0, 0.5), c=(1.0, 20.0, 40.0)):
    """texel (x, y) channel k = a_k x + b_k y + c_k (row y first)."""
    ys, xs = np.mgrid[0:h, 0:w].astype(np.float64)
    return np.stack([a[k] * xs + b[k] * ys + c[k] for k in range(3)], -1).astype(np.float32)


def sample(maps, images, index, dst=(0.0, 0.0, 0.0), vt=None, mono=False, rgb=(0.0, 0.0, 0.0)):
    table, texels = rt.texture_images(images)
    return O.texture_sample(np.asarray(maps, rt.TEXTURE_MAP_DTYPE).reshape(-1), table, texels, index, dst, vt, mono,
                            rgb)


def test_default_axes_follow_the_constructors():
    m = rt.image_map(rt.MAP_PLANAR, 0, (0, 0, 0), (0, 0, 1), (1, 1, 0))
    assert list(m["u"]) == [1, 0, 0] and list(m["v"]) == [0, 1, 0]
    m = rt.image_map(rt.MAP_CYLINDRICAL, 0, (0, 0, 0), (0, 1, 0), (1, 1, 1))
    assert list(m["u"]) == [1, 0, 0] and list(m["v"]) == [0, 1, 0]
    assert m["u_wrap"] == rt.WRAP_TILE and m["v_wrap"] == rt.WRAP_BLANK


@pytest.mark.parametrize("w,h", [(16, 8), (7, 13)])
def test_bilinear_of_a_linear_texture_is_the_function(w, h):
    img = linear_image(w, h)
    m = rt.image_map(rt.MAP_PLANAR, 0, (0, 0, 0), (1, 0, 0), (0, 0, 0), rt.WRAP_CLAMP, rt.WRAP_CLAMP)
    rng = np.random.default_rng(1)
    for _ in range(200):
        u = F(rng.uniform(0.0, (w - 1.001) / w))
        v = F(rng.uniform(0.0, (h - 1.001) / h))
        got, alpha = sample([m], [img], 0, vt=(u, v))
        uc, vc = float(u * F(w)), float(v * F(h))
        want = [3.0 * uc + 11.0 * vc + 1.0, 5.0 * uc + 2.0 * vc + 20.0, 7.0 * uc + 0.5 * vc + 40.0]
        np.testing.assert_allclose(got, want, rtol=2e-6, atol=2e-4)
        assert alpha == 1.0


def column_image(w=4, h=2):
    """texel (x, y) = 10 x + 100 y in every channel."""
    ys, xs = np.mgrid[0:h, 0:w]
    return np.repeat((10.0 * xs + 100.0 * ys)[:, :, None], 3, -1).astype(np.float32)


# (mode, texel column c, expected column or None for black) for size 4.
WRAP_CASES = [
    (rt.WRAP_TILE, -1, 2), (rt.WRAP_TILE, -3, 3), (rt.WRAP_TILE, -4, 2), (rt.WRAP_TILE, 4, 0), (rt.WRAP_TILE, 9, 1),
    (rt.WRAP_MIRROR, -1, 2), (rt.WRAP_MIRROR, -3, 0), (rt.WRAP_MIRROR, 4, 1), (rt.WRAP_MIRROR, 7, 2),
    (rt.WRAP_MIRROR, 6, 3), (rt.WRAP_CLAMP, -5, 0), (rt.WRAP_CLAMP, 9, 3), (rt.WRAP_CLAMP, 2, 2),
    (rt.WRAP_BLANK, -1, None), (rt.WRAP_BLANK, 4, None), (rt.WRAP_BLANK, 3, 3),
]


@pytest.mark.parametrize("mode,col,want", WRAP_CASES)
def test_wrapping_modes(mode, col, want):
    img = column_image()
    m = rt.image_map(rt.MAP_PLANAR, 0, (0, 0, 0), (1, 0, 0), (0, 0, 0), mode, rt.WRAP_CLAMP)
    # u_co = col exactly (fraction 0): the texel of the wrapped column, row 0.
    got, _ = sample([m], [img], 0, vt=(F(col) / F(4), F(0.0)), rgb=(7.0, 7.0, 7.0))
    assert np.array_equal(got, np.full(3, 0.0 if want is None else 10.0 * want, np.float32))


def test_upper_neighbour_is_clamped_to_the_image():
    img = column_image()
    m = rt.image_map(rt.MAP_PLANAR, 0, (0, 0, 0), (1, 0, 0), (0, 0, 0), rt.WRAP_CLAMP, rt.WRAP_CLAMP)
    got, _ = sample([m], [img], 0, vt=(F(3.5) / F(4), F(0.0)))
    assert np.all(got == F(30.0))  # texel 3 blended with itself


def test_planar_projection():
    w, h = 16, 8
    img = linear_image(w, h)
    # normal z: u = (1, 0, 0), v = (0, 1, 0); size 2 x 4 centred at (0.5, 0.5)
    m = rt.image_map(rt.MAP_PLANAR, 0, (0.5, 0.5, 0.0), (0, 0, 1), (2.0, 4.0, 0.0), rt.WRAP_CLAMP, rt.WRAP_CLAMP)
    for x, y in [(-0.3, -1.2), (0.1, 0.4), (1.2, 1.8)]:  # inside the texels whose upper neighbours exist
        got, _ = sample([m], [img], 0, dst=(x, y, 0.25))
        uc = (x - 0.5 + 1.0) * (w / 2.0)
        vc = (y - 0.5 + 2.0) * (h / 4.0)
        want = [3.0 * uc + 11.0 * vc + 1.0, 5.0 * uc + 2.0 * vc + 20.0, 7.0 * uc + 0.5 * vc + 40.0]
        np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-3)


def test_cubic_adds_the_distance_along_its_normal():
    w, h = 16, 8
    img = linear_image(w, h)
    m = rt.image_map(rt.MAP_CUBIC, 0, (0.0, 0.0, 0.0), (0, 0, 1), (4.0, 4.0, 0.0), rt.WRAP_CLAMP, rt.WRAP_CLAMP)
    x, y, z = 0.3, 0.2, 0.5
    got, _ = sample([m], [img], 0, dst=(x, y, z))
    uc = ((x + 2.0) + z) * (w / 4.0)  # dot(dist, u) + dot(dist, n)
    vc = (y + 2.0) * (h / 4.0)
    np.testing.assert_allclose(got[0], 3.0 * uc + 11.0 * vc + 1.0, rtol=1e-5)


def test_cylindrical_angle_and_height():
    w, h = 32, 16
    img = linear_image(w, h)
    m = rt.image_map(rt.MAP_CYLINDRICAL, 0, (0.0, 0.0, 0.0), (0, 1, 0), (2.0, 3.0, 2.0), cycles=2)
    for theta, y in [(0.4, 0.5), (-0.4, 0.5), (2.5, -1.0)]:
        got, _ = sample([m], [img], 0, dst=(np.cos(theta), y, np.sin(theta)))
        ang = abs(theta)  # acos folds the angle into [0, pi]
        uc = (2 * ang / (2 * np.pi)) * w / 2.0
        vc = (y + 1.5) * (h / 3.0)
        u0 = int(uc)
        if u0 + 1 < w:
            np.testing.assert_allclose(got[0], 3.0 * uc + 11.0 * vc + 1.0, rtol=1e-4)


def test_checker_board_only_x_decides():
    m = rt.checker_board_map((200.0, 100.0, 50.0), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), opacity=0.5)
    base = (1.0, 2.0, 3.0)
    for x, want in [(0.5, base), (1.5, (200.0, 100.0, 50.0)), (2.5, base), (3.9, (200.0, 100.0, 50.0))]:
        for z in (0.5, 1.5, 7.5):
            got, alpha = sample([m], [], 0, dst=(x, 0.3, z), rgb=base)
            assert np.array_equal(got, np.asarray(want, np.float32)) and alpha == 0.5
    val, _ = sample([m], [], 0, dst=(1.5, 0.0, 0.0), mono=True)
    assert val == (F(200.0) + F(100.0) + F(50.0)) * (F(1.0) / F(3.0))
    val, _ = sample([m], [], 0, dst=(0.5, 0.0, 0.0), mono=True, rgb=base)
    assert val == 0.0  # a fresh colour for the monochrome sample


def test_falloffs_and_inverse():
    m = rt.checker_board_map((200.0, 100.0, 50.0), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    d = (1.5, 0.25, -0.5)
    cube = rt.with_falloff(m, rt.FALLOFF_CUBIC, (0.5, 0.0, 0.0), (0.1, 0.2, 0.3))
    got, _ = sample([cube], [], 0, dst=d)
    fall = (abs(1.0 * 0.1) + abs(0.25 * 0.2) + abs(-0.5 * 0.3)) * 255.0
    np.testing.assert_allclose(got, np.maximum(0.0, np.array([200.0, 100.0, 50.0]) - fall), rtol=1e-5)
    sph = rt.with_falloff(m, rt.FALLOFF_SPHERICAL, (0.5, 0.0, 0.0), (0.2, 0.3, 0.5))
    got, _ = sample([sph], [], 0, dst=d)
    r = np.sqrt(1.0 + 0.0625 + 0.25)
    k = max(0, 1 - r * 0.2) * max(0, 1 - r * 0.3) * max(0, 1 - r * 0.5)
    np.testing.assert_allclose(got, np.array([200.0, 100.0, 50.0]) * k, rtol=1e-5)
    lin = rt.with_falloff(m, rt.FALLOFF_LINEAR_Z, (0.0, 0.0, -1.0), (0.0, 0.0, 0.8))
    got, _ = sample([lin], [], 0, dst=d)
    np.testing.assert_allclose(got, np.array([200.0, 100.0, 50.0]) * (1 - 0.5 * 0.8), rtol=1e-5)
    inv = m.copy()
    inv["inverse"] = 1
    got, _ = sample([inv], [], 0, dst=d)
    assert np.array_equal(got, np.float32([55.0, 155.0, 205.0]))


def test_monochrome_is_the_texel_mean_and_zero_outside_blank():
    img = linear_image(8, 8)
    m = rt.image_map(rt.MAP_PLANAR, 0, (0, 0, 0), (1, 0, 0), (0, 0, 0), rt.WRAP_BLANK, rt.WRAP_BLANK)
    col, _ = sample([m], [img], 0, vt=(F(2) / F(8), F(3) / F(8)))
    val, _ = sample([m], [img], 0, vt=(F(2) / F(8), F(3) / F(8)), mono=True)
    assert val == (col[0] + col[1] + col[2]) * (F(1.0) / F(3.0))
    val, _ = sample([m], [img], 0, vt=(F(-1.5), F(0.0)), mono=True)
    assert val == 0.0


# ------------------------------------------------------------- frames

def plane_scene(z=2.0, size=4.0):
    """Two triangles of the plane z (facing -z), spanning [-size, size]^2."""
    s = size
    a, b, c, d = (-s, -s, z), (s, -s, z), (s, s, z), (-s, s, z)
    return np.asarray([[*a, *b, *c], [*a, *c, *d]], np.float32)


def camera(res=24):
    return Camera(pos=(0.0, 0.0, 0.0), x_axis=(1.0, 0.0, 0.0), y_axis=(0.0, 1.0, 0.0), z_axis=(0.0, 0.0, 1.0),
                  w=0.5, h=0.5, t=1.0, x_res=res, y_res=res)


def render(abc, mats, tri_mat, maps=None, images=None, layers=None, vt=None, lights=None, background=(0.0, 0.0, 0.0)):
    sc = O.OracleScene(abc)
    tex = None
    if maps is not None:
        table, texels = rt.texture_images(images or [])
        tex = (np.asarray(maps, rt.TEXTURE_MAP_DTYPE).reshape(-1), table, texels, layers)
    lights = lights if lights is not None else [Light((0.5, 0.7, -1.0), (255.0, 240.0, 200.0), 0.3)]
    r = sc.render(np.full((abc.shape[0], 3), 0.5, np.float32), camera(), lights, materials=mats,
                  tri_material=np.asarray(tri_mat, np.int32), textures=tex, texture_coords=vt, background=background)
    return r


def uniform_image(v, w=4, h=4):
    return np.full((h, w, 3), v, np.float32)


def test_coloured_uniform_maps_equal_scaled_colours():
    abc = plane_scene()
    m = rt.image_map(rt.MAP_PLANAR, 0, (0, 0, 0), (0, 0, 1), (8.0, 8.0, 0.0))
    base = dict(kind=2, ka=(40.0, 30.0, 20.0), kd=(200.0, 150.0, 100.0), ks=(60.0, 60.0, 60.0), s=12.0)
    tex = dict(base, tex=(0, 0, 0, None, None, None))
    t = F(90.0)
    scaled = dict(base, ka=tuple(float(t * (F(x) * (F(1) / F(255)))) for x in base["ka"]),
                  kd=tuple(float(t * (F(x) * (F(1) / F(255)))) for x in base["kd"]),
                  ks=tuple(float(t * (F(x) * (F(1) / F(255)))) for x in base["ks"]))
    a = render(abc, O.material_array([tex]), [0, 0], [m], [uniform_image(90.0)])["rgb"]
    b = render(abc, O.material_array([scaled]), [0, 0])["rgb"]
    # A constant image blends back to its value up to the weights' rounding.
    np.testing.assert_allclose(a, b, rtol=2e-6)
    assert np.abs(a).max() > 10
    # t_ns: the shininess becomes the map's alpha times ns / 255, whatever the texel.
    ns = dict(base, tex=(None, None, None, 0, None, None))
    c = render(abc, O.material_array([ns]), [0, 0], [m], [uniform_image(3.0)])["rgb"]
    d = render(abc, O.material_array([dict(base, s=float(F(1.0) * (F(12.0) * (F(1) / F(255)))))]), [0, 0])["rgb"]
    assert np.array_equal(c, d)
    assert not np.array_equal(c, render(abc, O.material_array([base]), [0, 0])["rgb"])


def mirror_scene():
    """A mirror plane at z = 2 seen by the camera, and a diffuse floor behind the camera."""
    mirror = plane_scene(2.0)
    back = plane_scene(-1.0)[:, [0, 1, 2, 6, 7, 8, 3, 4, 5]]  # the same plane, wound the other way
    return np.concatenate([mirror, back])


def test_reflection_map_gates_the_mirror_ray_and_keeps_the_weight():
    abc = mirror_scene()
    m = rt.image_map(rt.MAP_PLANAR, 0, (0, 0, 0), (0, 0, 1), (8.0, 8.0, 0.0))
    diffuse = dict(kind=2, kd=(180.0, 180.0, 180.0))
    mirror = dict(kind=2, kd=(20.0, 20.0, 20.0), rf=0.9)
    plain = render(abc, O.material_array([mirror, diffuse]), [0, 0, 1, 1])
    full = render(abc, O.material_array([dict(mirror, tex=(None, None, None, None, 0, None)), diffuse]), [0, 0, 1, 1],
                  [m], [uniform_image(255.0)])
    assert np.array_equal(full["rgb"], plain["rgb"]) and full["sec_counters"][0] > 0
    # red 100: the ray still spawns (power 100 x 0.9 / 255 > 0.1) and comes back weighted by rf.
    part = render(abc, O.material_array([dict(mirror, tex=(None, None, None, None, 0, None)), diffuse]),
                  [0, 0, 1, 1], [m], [uniform_image(100.0)])
    assert np.array_equal(part["rgb"], plain["rgb"])
    # red 20: power 0.07 <= 0.1, no mirror ray.
    none = render(abc, O.material_array([dict(mirror, tex=(None, None, None, None, 0, None)), diffuse]),
                  [0, 0, 1, 1], [m], [uniform_image(20.0)])
    assert none["sec_counters"][0] == 0
    local = render(abc, O.material_array([dict(mirror, rf=0.0), diffuse]), [0, 0, 1, 1])
    assert np.array_equal(none["rgb"], local["rgb"])


def test_mapper_shader_against_the_plane_geometry():
    abc = plane_scene()
    cam = camera()
    L = np.array([0.5, 0.7, -1.0])
    lrgb, fall = np.array([255.0, 240.0, 200.0]), 0.3
    rgb, kd, ks, s, clrh = np.array([200.0, 120.0, 60.0]), 0.8, 0.6, 5.0, 0.3
    mat = dict(kind=3, rgb=tuple(rgb), kd=kd, ks=ks, s=s, clrh=clrh)
    r = render(abc, O.material_array([mat]), [0, 0])
    res = cam.x_res
    got = r["rgb"].reshape(res, res, 3)
    n = np.array([0.0, 0.0, -1.0])  # facing the camera
    worst = 0.0
    for py in range(0, res, 3):
        for px in range(0, res, 3):
            xt = px * (2 * 0.5 / res) - 0.5
            yt = py * (2 * 0.5 / res) - 0.5
            d = np.array([xt, yt, 1.0])
            d /= np.linalg.norm(d)
            p = d * (2.0 / d[2])
            ld = L - p
            dist = np.linalg.norm(ld)
            ld /= dist
            li = (lrgb / 255.0) / ((dist * fall) ** 2 + 1.0)  # the light keeps rgb / 255 (rt/light.h:44-59)
            diff = ld @ n * kd
            want = np.zeros(3)
            if diff >= 0:
                ref = ld - n * (2.0 * diff)
                rdr = d @ ref
                spec = rdr ** s * ks if rdr > 0 else 0.0
                want = rgb * diff * li + (clrh * rgb + (1 - clrh) * li) * spec
            worst = max(worst, np.abs(got[py, px] - want).max() / max(1.0, np.abs(want).max()))
    assert worst < 1e-4, worst


def test_mapper_refraction_blends_by_clrf():
    abc = plane_scene()
    mat = dict(kind=3, rgb=(0.0, 0.0, 0.0), kd=0.0, tran=0.8, ri=1.0, clrf=0.25)
    bg = (100.0, 50.0, 10.0)
    sc = O.OracleScene(abc, flags=np.full(2, 2, np.uint8))
    r = sc.render(np.full((2, 3), 0.5, np.float32), camera(), [], materials=O.material_array([mat]),
                  tri_material=np.zeros(2, np.int32), background=bg)
    want = np.float32([(F(1) - F(0.25)) * F(x) * F(0.8) for x in bg])
    # local colour 0: c = 0 + (clrf * 0 + (1 - clrf) * background * tran)
    assert np.allclose(r["rgb"], want, rtol=1e-6), (r["rgb"][:3], want)


def test_mapper_layers_set_kd_and_colour():
    abc = plane_scene()
    m_kd = rt.image_map(rt.MAP_PLANAR, 0, (0, 0, 0), (0, 0, 1), (8.0, 8.0, 0.0))
    m_col = rt.checker_board_map((10.0, 200.0, 30.0), (0.0, 0.0, 0.0), (0.25, 0.25, 0.25))
    base = dict(kind=3, rgb=(200.0, 120.0, 60.0), kd=0.8, ks=0.0)
    mats = [dict(base, ctex=[1], dtex=[0])]
    layers = rt.layer_lists(mats)
    a = render(abc, O.material_array(mats), [0, 0], [m_kd, m_col], [uniform_image(127.5)], layers)["rgb"]
    # kd from the map: 127.5 / 255 = 0.5 (to the bilinear weights' rounding);
    # the colour is the checker's where the x cell is odd.
    kd = float(F(127.5) * (F(1) / F(255)))
    b = render(abc, O.material_array([dict(base, kd=kd)]), [0, 0])["rgb"]
    c = render(abc, O.material_array([dict(base, kd=kd, rgb=(10.0, 200.0, 30.0))]), [0, 0])["rgb"]
    res = 24
    xs = (np.arange(res) * (2 * 0.5 / res) - 0.5).astype(np.float32)
    hit_x = xs * 2.0  # the plane at z = 2
    painted = (np.trunc(hit_x / 0.25).astype(int) & 1).astype(bool)
    assert 4 <= painted.sum() <= 20
    a, b, c = (v.reshape(res, res, 3) for v in (a, b, c))
    # Columns next to a cell edge may fall either side under float rounding.
    edge = np.abs(hit_x / 0.25 - np.round(hit_x / 0.25)) < 1e-3
    np.testing.assert_allclose(a[:, ~painted & ~edge], b[:, ~painted & ~edge], rtol=2e-6)
    np.testing.assert_allclose(a[:, painted & ~edge], c[:, painted & ~edge], rtol=2e-6)


# ------------------------------------------------------------- TGA

def test_tga_reader_against_the_file_bytes():
    for name in ("marble.tga", "wood.tga"):
        data = open(os.path.join(TEX, name), "rb").read()
        img = rt.load_tga(os.path.join(TEX, name))
        assert img.shape == (200, 200, 3)
        assert np.array_equal(img, OI.read_tga(data))
        raw = np.frombuffer(data[18:18 + 200 * 200 * 3], np.uint8).reshape(200, 200, 3)
        assert np.array_equal(img[..., 0], raw[..., 2]) and np.array_equal(img[..., 2], raw[..., 0])
        assert np.array_equal(img[..., 1], raw[..., 1])


def test_tga_run_length_by_hand(tmp_path):
    hdr = bytearray(18)
    hdr[2] = 10  # run-length true colour
    hdr[12], hdr[14], hdr[16] = 3, 2, 24
    body = bytes([0x82, 1, 2, 3,          # a run of 3 pixels (B, G, R) = (1, 2, 3)
                  0x01, 4, 5, 6, 7, 8, 9,  # 2 raw pixels
                  0x80, 10, 11, 12])       # a run of 1
    p = tmp_path / "rle.tga"
    p.write_bytes(bytes(hdr) + body)
    img = rt.load_tga(str(p))
    want = np.float32([[[3, 2, 1], [3, 2, 1], [3, 2, 1]], [[6, 5, 4], [9, 8, 7], [12, 11, 10]]])
    assert np.array_equal(img, want)
    assert np.array_equal(OI.read_tga(p.read_bytes()), want)


def test_tga_errors(tmp_path):
    p = tmp_path / "cmap.tga"
    hdr = bytearray(18)
    hdr[1], hdr[2], hdr[12], hdr[14], hdr[16] = 1, 1, 1, 1, 8
    p.write_bytes(bytes(hdr) + bytes(3 * 256 + 1))
    with pytest.raises(rt.RTError, match="type 1"):
        rt.load_tga(str(p))
    short = tmp_path / "short.tga"
    hdr = bytearray(18)
    hdr[2], hdr[12], hdr[14], hdr[16] = 2, 4, 4, 24
    short.write_bytes(bytes(hdr) + bytes(10))
    with pytest.raises(rt.RTError, match="cut short"):
        rt.load_tga(str(short))
