"""Oracle pins for the phong materials with mirror and refraction rays
(rt/materials/phong_shader.cc:21-111, rt/ray.cc:118-258, rt/raytracer.cc:20-87).

None of these retype the oracle's arithmetic: the default material must give
the Lambert frame bit for bit; an index-matched glass (ri = 1) must scale what
lies behind it by its transmittance; a mirror must show the mirrored scene;
the recursion must stop where the ray power falls to MIN_REFLECTIVE_POWER
(0.1, compared in double); the specular term must follow x^s; refraction must
obey Snell's law and total internal reflection.
"""
import numpy as np
import pytest

from oracle import oracle as O

from helpers import rel_err


def _w(**kw):
    from raptor_amd import workloads

    return workloads.material_scene(**kw)


def test_default_material_is_the_lambert_frame():
    w, _, kinds, flags = _w(res=48)
    sc = O.OracleScene(w.abc, flags)
    a = sc.render(w.albedo, w.camera, w.lights, nthreads=8)
    b = sc.render(w.albedo, w.camera, w.lights, nthreads=8, materials=[dict(kd=(10.0, 20.0, 30.0))],
                  tri_material=np.full(w.abc.shape[0], -1, np.int32))
    assert np.array_equal(a["rgb"].view(np.uint32), b["rgb"].view(np.uint32))
    assert np.array_equal(a["prim"], b["prim"])
    assert int(b["sec_counters"][0]) == 0


def test_index_matched_glass_scales_what_is_behind_it():
    w0, _, _, f0 = _w(res=64)
    w1, m1, k1, f1 = _w(res=64, glass=(0.6, 1.0))
    base = O.OracleScene(w0.abc, f0).render(w0.albedo, w0.camera, w0.lights, nthreads=8,
                                            materials=[dict()], tri_material=np.full(w0.abc.shape[0], -1, np.int32))
    r = O.OracleScene(w1.abc, f1).render(w1.albedo, w1.camera, w1.lights, nthreads=8, materials=m1, tri_material=k1)
    glass = r["prim"] >= w0.abc.shape[0]
    assert glass.mean() > 0.5
    e = rel_err(r["rgb"][glass], np.float32(0.6) * base["rgb"][glass], floor=1e-2)
    assert (e > 1e-5).mean() < 0.01 and np.median(e) < 1e-6
    assert (base["rgb"][glass] > 0).mean() > 0.3
    # One refracted ray per glass pixel, with its shadow ray when it hits.
    assert int(r["sec_counters"][0]) == int(glass.sum())


def test_mirror_shows_the_mirrored_scene():
    """A mirror wall at z = 1 (normal along -z, so reflections are exact)
    shows, times its reflectance, the scene reflected about z = 1."""
    kw = dict(res=80, n_objects=400, size=0.15, light=(0.3, 1.5, 0.9), floor=False)
    w, m, k, f = _w(mirror=0.5, **kw)
    wm, _, _, fm = _w(mirrored=True, **kw)
    r = O.OracleScene(w.abc, f).render(w.albedo, w.camera, w.lights, nthreads=8, materials=m, tri_material=k)
    ref = O.OracleScene(wm.abc, fm).render(w.albedo[:wm.abc.shape[0]], wm.camera, wm.lights, nthreads=8)
    mirror = r["prim"] >= w.abc.shape[0] - 2
    lit = ref["rgb"][mirror].max(1) > 0
    assert mirror.mean() > 0.3 and lit.sum() > 100
    e = rel_err(r["rgb"][mirror], np.float32(0.5) * ref["rgb"][mirror], floor=1e-2)
    assert (e > 1e-4).mean() < 0.01, (e > 1e-4).mean()
    # Whatever the mirror shows, the reflected rays were traced: one per mirror pixel.
    assert int(r["sec_counters"][0]) == int(mirror.sum())


def _corridor(rf):
    from raptor_amd import workloads

    abc = np.concatenate([workloads.quad((1, -1, -50), (1, 1, -50), (1, 1, 50), (1, -1, 50)),
                          workloads.quad((-1, -1, -50), (-1, -1, 50), (-1, 1, 50), (-1, 1, -50))])
    cam = workloads.Camera(workloads.f32((0.0, 0.0, 0.0)), (0.0, 0.0, -1.0), (0.0, 1.0, 0.0), (1.0, 0.0, 0.0),
                           0.05, 0.05, 1.0, 16, 16)
    lights = [workloads.Light(workloads.f32((0.0, 0.5, 0.0)))]
    albedo = np.full((4, 3), 0.5, np.float32)
    return abc, albedo, cam, lights, [dict(rf=rf, kd=(100.0, 100.0, 100.0))], np.zeros(4, np.int32)


@pytest.mark.parametrize("rf,rays", [(0.5, 3), (np.float32(0.1), 1), (0.09, 0)])
def test_recursion_stops_at_min_reflective_power(rf, rays):
    """Two facing mirrors: powers rf^k while above 0.1 (a float 0.1f is above
    the double 0.1, so it still recurses once)."""
    abc, albedo, cam, lights, mats, kinds = _corridor(float(rf))
    r = O.OracleScene(abc).render(albedo, cam, lights, nthreads=2, materials=mats, tri_material=kinds)
    npix = cam.x_res * cam.y_res
    assert (r["prim"] >= 0).all()
    assert int(r["sec_counters"][0]) == rays * npix
    assert int(r["sec_counters"][1]) == rays * npix * len(lights)


def test_specular_term_is_a_power_of_the_cosine():
    """kd = 0: c_s = K x^s per pixel, so c_1 c_3 = c_2^2."""
    from raptor_amd import workloads

    abc = workloads.quad((-2, 0, -2), (2, 0, -2), (2, 0, 2), (-2, 0, 2))
    pos = (0.0, 1.0, -1.0)
    x, y, z = workloads.look_at(pos, (0.0, 0.0, 0.2))
    cam = workloads.Camera(workloads.f32(pos), x, y, z, 0.5, 0.5, 1.0, 32, 32)
    lights = [workloads.Light(workloads.f32((0.0, 1.0, 1.5)), (255.0, 255.0, 255.0), 0.0)]
    albedo = np.full((2, 3), 0.5, np.float32)
    out = {}
    for s in (1.0, 2.0, 3.0):
        m = [dict(ks=(200.0, 150.0, 100.0), s=s)]
        out[s] = O.OracleScene(abc).render(albedo, cam, lights, nthreads=2, materials=m,
                                           tri_material=np.zeros(2, np.int32))["rgb"].astype(np.float64)
    ok = out[2.0] > 1e-3
    assert ok.mean() > 0.3
    assert np.allclose((out[1.0] * out[3.0])[ok], (out[2.0] ** 2)[ok], rtol=2e-6)


def test_refraction_obeys_snell_and_total_internal_reflection():
    rng = np.random.default_rng(3)
    for _ in range(300):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        if d @ n > 0:
            d = -d  # n faces the ray
        ri = float(rng.choice([1.2, 1.5, 2.4]))
        for leaving in (False, True):
            eta = ri if leaving else 1.0 / ri
            sin_i = np.linalg.norm(np.cross(d, n))
            out = O.refract_dir(d, n, ri, leaving)
            if eta * sin_i > 1.0 + 1e-5:
                assert out is None
                continue
            if eta * sin_i > 1.0 - 1e-5:
                continue  # at the critical angle either answer is right
            assert out is not None
            o = out.astype(np.float64)
            assert abs(np.linalg.norm(o) - 1.0) < 1e-5
            assert o @ n < 0  # through the surface
            assert abs(np.linalg.norm(np.cross(o, n)) - eta * sin_i) < 2e-5
            assert abs(np.cross(d, n) @ o) < 2e-5  # in the plane of incidence


def test_reflection_keeps_the_tangent_and_flips_the_normal_part():
    rng = np.random.default_rng(4)
    for _ in range(200):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        o = O.reflect_dir(d, n).astype(np.float64)
        assert abs(o @ n + d @ n) < 2e-6
        assert np.allclose(o - (o @ n) * n, d - (d @ n) * n, atol=2e-6)
    # Axis-aligned normals reflect exactly.
    d = np.array([0.3, -0.7, 0.648074], np.float32)
    assert np.array_equal(O.reflect_dir(d, (0, 1, 0)), np.array([d[0], -d[1], d[2]], np.float32))


def test_ambient_term_is_added_once_per_hit():
    """ka alone (kd = ks = 0): every hit pixel is exactly ka, lit or not;
    with a mirror of reflectance rf facing it, ka + rf ka for pixels whose
    mirror ray comes back to an ambient surface."""
    w, _, _, f = _w(res=40)
    n = w.abc.shape[0]
    ka = (7.0, 11.0, 13.0)
    r = O.OracleScene(w.abc, f).render(w.albedo, w.camera, w.lights, nthreads=4, materials=[dict(ka=ka)],
                                       tri_material=np.zeros(n, np.int32))
    hit = r["prim"] >= 0
    assert hit.mean() > 0.5
    assert np.array_equal(r["rgb"][hit], np.tile(np.float32(ka), (int(hit.sum()), 1)))
    assert np.all(r["rgb"][~hit] == 0)


def _ct(**kw):
    return dict(kind=1, **kw)


def test_cook_torrance_white_diffuse_is_lambert():
    """cook_torrance_cxy with rs = 0 and the white chroma of the colour-space
    conversion (0.3333, 0.3333) gives r = g = b = 255 * shade * rd: the phong
    frame with kd = 255 * rd."""
    w, _, _, f = _w(res=48)
    n = w.abc.shape[0]
    k = np.zeros(n, np.int32)
    sc = O.OracleScene(w.abc, f)
    ct = sc.render(w.albedo, w.camera, w.lights, nthreads=4, materials=[_ct(x=0.3333, y=0.3333, rd=0.5)],
                   tri_material=k)
    ph = sc.render(w.albedo, w.camera, w.lights, nthreads=4, materials=[dict(kd=(127.5, 127.5, 127.5))],
                   tri_material=k)
    lit = ph["rgb"].max(1) > 0
    assert lit.mean() > 0.3
    assert np.array_equal(lit, ct["rgb"].max(1) > 0)
    e = rel_err(ct["rgb"][lit], ph["rgb"][lit], floor=1e-2)
    assert e.max() < 1e-5, e.max()


def test_cook_torrance_primary_chroma_is_a_primary():
    """The red primary's chroma (0.64, 0.33) converts to pure red, the green
    primary's (0.29, 0.60) to pure green (the XYZ -> RGB matrix maps each
    primary to its own channel)."""
    w, _, _, f = _w(res=40)
    n = w.abc.shape[0]
    for (x, y), ch in (((0.64, 0.33), 0), ((0.29, 0.60), 1)):
        r = O.OracleScene(w.abc, f).render(w.albedo, w.camera, w.lights, nthreads=4,
                                           materials=[_ct(x=x, y=y, rd=0.3)], tri_material=np.zeros(n, np.int32))
        rgb = r["rgb"].astype(np.float64)
        lit = rgb[:, ch] > 1.0
        assert lit.mean() > 0.3
        others = [c for c in range(3) if c != ch]
        assert np.all(rgb[lit][:, others] <= 1e-4 * rgb[lit][:, [ch]])


def test_cook_torrance_specular_lobe_vanishes():
    """The shading normal faces the ray, so nv = n.v < 0 < nl = n.l and the
    two geometric terms g0 = 2 nh nv / vh and g1 = 2 nh nl / vh have product
    4 nh^2 nv nl / vh^2 <= 0: min(g0, g1) <= 0 and the clamped g is 0. The
    specular term s = rs g f ro / (nl nv) is then a zero, and below the
    reflection threshold (rs < 0.1) neither rs nor the roughness changes the
    frame, bit for bit."""
    w, _, _, f = _w(res=48)
    n = w.abc.shape[0]
    k = np.zeros(n, np.int32)
    sc = O.OracleScene(w.abc, f)
    frames = [sc.render(w.albedo, w.camera, w.lights, nthreads=4,
                        materials=[_ct(x=0.31, y=0.35, rd=0.3, rs=rs, sr=sr)], tri_material=k)
              for rs, sr in ((0.0, 0.0), (0.09, 0.2), (0.05, 0.5))]
    assert (frames[0]["rgb"].max(1) > 0).mean() > 0.3
    for fr in frames[1:]:
        assert np.array_equal(fr["rgb"].view(np.uint32), frames[0]["rgb"].view(np.uint32))
        assert int(fr["sec_counters"][0]) == 0


def test_cook_torrance_recursion_uses_rs_and_ts():
    abc, albedo, cam, lights, _, kinds = _corridor(0.5)
    mats = [_ct(rs=0.5, rd=0.4, sr=0.1)]
    r = O.OracleScene(abc).render(albedo, cam, lights, nthreads=2, materials=mats, tri_material=kinds)
    npix = cam.x_res * cam.y_res
    assert int(r["sec_counters"][0]) == 3 * npix
    w1, m1, k1, f1 = _w(res=64, glass=(0.6, 1.0))
    m1 = [_ct(ts=0.6, ri_r=1.0, rd=0.0)]
    r1 = O.OracleScene(w1.abc, f1).render(w1.albedo, w1.camera, w1.lights, nthreads=4, materials=m1,
                                          tri_material=np.where(k1 >= 0, 0, -1).astype(np.int32))
    glass = r1["prim"] >= w1.abc.shape[0] - 2
    assert int(r1["sec_counters"][0]) == int(glass.sum())
