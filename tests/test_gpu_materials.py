"""GPU parity of frames with phong materials: mirror and refraction rays traced
recursively (rt/materials/phong_shader.cc:21-111, rt/ray.cc:118-258,
rt/raytracer.cc:20-87) through the C-ABI against the oracle.

Primary hit ids are bit-exact except at exact-t ties, radiance agrees within
1e-5 relative (the GPU folds the ray tree in the reference's order; only powf
may round differently from the host's pow), and the numbers of secondary rays
and their shadow rays equal the oracle's.
"""
import numpy as np
import pytest

from helpers import compare_hits, rel_err

pytestmark = pytest.mark.gpu

RAD_TOL = 1e-5


def _scene(**kw):
    from raptor_amd import workloads

    return workloads.material_scene(**kw)


def _specular_objects(mats, kinds, flags, n_floor=2):
    """Objects (after the floor's two triangles, before any wall) get a
    shiny, slightly ambient material."""
    mats = list(mats) + [dict(ka=(4.0, 3.0, 2.0), kd=(180.0, 120.0, 60.0), ks=(90.0, 90.0, 90.0), s=24.0)]
    kinds = kinds.copy()
    obj = (kinds == -1)
    obj[:n_floor] = False
    kinds[obj] = len(mats) - 1
    return mats, kinds, flags


def _both(rt, O, ctx, w, mats, kinds, flags, spp=1, seed=0, width=4, sah=16):
    scene = rt.Scene(ctx, w.abc, w.albedo, flags)
    scene.set_materials(mats, kinds)
    accel = rt.Accel(scene, width=width, sah_bins=sah)
    out = accel.render(w.camera, w.lights, spp=spp, seed=seed)
    ref = O.OracleScene(w.abc, flags).render(w.albedo, w.camera, w.lights, spp=spp, seed=seed, nthreads=8,
                                             materials=mats, tri_material=kinds)
    return out, ref


def _check(out, ref, spp=1, nl=1):
    g = out["primary"][::spp]
    bad, ties = compare_hits(g, ref["prim"], ref["tuv"])
    assert not bad.any(), f"{bad.sum()} primary hit ids differ"
    e = rel_err(out["rgb"].reshape(-1, 3)[~ties], ref["rgb"][~ties])
    assert e.max() <= RAD_TOL, (e.max(), int((e > RAD_TOL).sum()))
    rays = out["stats"]["rays"]
    assert rays[2] == int(ref["sec_counters"][0])
    assert rays[3] == int(ref["sec_counters"][1])
    if spp == 1:
        assert rays[1] == int((ref["prim"] >= 0).sum()) * nl
    return e


@pytest.mark.parametrize("width,sah", [(4, 16), (8, 0)])
def test_mirror_frame(gpu_ctx, rt, O, width, sah):
    w, m, k, f = _scene(res=96, mirror=0.5)
    out, ref = _both(rt, O, gpu_ctx, w, m, k, f, width=width, sah=sah)
    _check(out, ref)
    assert out["stats"]["rays"][2] > 1000
    assert (out["rgb"] > 0).any(axis=2).mean() > 0.5


def test_glass_and_mirror_with_specular_objects(gpu_ctx, rt, O):
    w, m, k, f = _scene(res=112, mirror=0.9, glass=(0.7, 1.5))
    m, k, f = _specular_objects(m, k, f)
    out, ref = _both(rt, O, gpu_ctx, w, m, k, f)
    _check(out, ref)
    # Refraction through the glass, then the mirror behind the objects.
    assert out["stats"]["rays"][2] > 2 * 112 * 112 * 0.5


def test_multisample_frame(gpu_ctx, rt, O):
    w, m, k, f = _scene(res=64, mirror=0.6, glass=(0.5, 1.3))
    m, k, f = _specular_objects(m, k, f)
    out, ref = _both(rt, O, gpu_ctx, w, m, k, f, spp=3, seed=11)
    _check(out, ref, spp=3)


@pytest.mark.parametrize("rf", [0.5, 0.95, 0.09])
def test_mirror_corridor_recursion(gpu_ctx, rt, O, rf):
    """Two facing mirrors: reflected rays while the power rf^k stays above
    0.1 (44 generations for rf = 0.95, none for 0.09)."""
    from raptor_amd import workloads

    abc = np.concatenate([workloads.quad((1, -50, -50), (1, 50, -50), (1, 50, 50), (1, -50, 50)),
                          workloads.quad((-1, -50, -50), (-1, -50, 50), (-1, 50, 50), (-1, 50, -50))])
    cam = workloads.Camera(workloads.f32((0.0, 0.0, 0.0)), (0.0, 0.0, -1.0), (0.0, 1.0, 0.0), (1.0, 0.0, 0.0),
                           0.05, 0.05, 1.0, 24, 24)
    lights = [workloads.Light(workloads.f32((0.0, 0.5, 0.0)))]
    albedo = np.full((4, 3), 0.5, np.float32)
    mats = [dict(rf=rf, kd=(100.0, 100.0, 100.0), ks=(50.0, 50.0, 50.0), s=5.0)]
    kinds = np.zeros(4, np.int32)
    scene = rt.Scene(gpu_ctx, abc, albedo)
    scene.set_materials(mats, kinds)
    accel = rt.Accel(scene)
    out = accel.render(cam, lights)
    ref = O.OracleScene(abc).render(albedo, cam, lights, nthreads=4, materials=mats, tri_material=kinds)
    e = rel_err(out["rgb"].reshape(-1, 3), ref["rgb"])
    assert e.max() <= RAD_TOL, e.max()
    # Every primary ray bounces between the walls (50 units on each side)
    # until its float power falls to 0.1.
    depth, power = 0, np.float32(1.0)
    while float(power * np.float32(rf)) > 0.1:
        power = power * np.float32(rf)
        depth += 1
    assert depth == {0.5: 3, 0.95: 44, 0.09: 0}[rf]
    assert out["stats"]["rays"][2] == int(ref["sec_counters"][0]) == depth * 24 * 24
    assert out["stats"]["rays"][3] == int(ref["sec_counters"][1]) == depth * 24 * 24


def test_lambert_materials_match_the_lambert_pipeline(gpu_ctx, rt):
    """With a table whose every triangle keeps the default (-1), the materials
    pipeline gives the plain rt_render frame bit for bit; removing the
    materials restores the plain pipeline."""
    w, m, k, f = _scene(res=80)
    scene = rt.Scene(gpu_ctx, w.abc, w.albedo, f)
    accel = rt.Accel(scene)
    plain = accel.render(w.camera, w.lights)
    scene.set_materials([dict(kd=(1.0, 2.0, 3.0))], np.full(w.abc.shape[0], -1, np.int32))
    mat = accel.render(w.camera, w.lights)
    assert np.array_equal(plain["rgb"].view(np.uint32), mat["rgb"].view(np.uint32))
    assert np.array_equal(plain["primary"].view(np.uint32), mat["primary"].view(np.uint32))
    assert mat["stats"]["rays"][:2] == plain["stats"]["rays"][:2]
    assert mat["stats"]["rays"][2] == 0
    scene.set_materials(np.zeros((0, 13), np.float32), np.zeros(0, np.int32))
    again = accel.render(w.camera, w.lights, bounce=True)  # the bounce is back
    assert again["stats"]["rays"][2] > 0


def test_material_frames_are_reproducible(gpu_ctx, rt):
    w, m, k, f = _scene(res=96, mirror=0.8, glass=(0.6, 1.4))
    m, k, f = _specular_objects(m, k, f)
    scene = rt.Scene(gpu_ctx, w.abc, w.albedo, f)
    scene.set_materials(m, k)
    accel = rt.Accel(scene)
    a = accel.render(w.camera, w.lights, spp=2, seed=3)
    b = accel.render(w.camera, w.lights, spp=2, seed=3)
    assert np.array_equal(a["rgb"].view(np.uint32), b["rgb"].view(np.uint32))


def test_material_errors(gpu_ctx, rt):
    w, m, k, f = _scene(res=32, glass=(0.5, 1.5))
    scene = rt.Scene(gpu_ctx, w.abc, w.albedo, f)
    with pytest.raises(rt.RTError, match="out of range"):
        scene.set_materials([dict(rf=0.96)], np.zeros(w.abc.shape[0], np.int32))
    with pytest.raises(rt.RTError, match="out of range"):
        scene.set_materials([dict(tran=0.5, ri=0.9)], k)
    with pytest.raises(rt.RTError, match="no material"):
        scene.set_materials(m, np.full(w.abc.shape[0], 3, np.int32))
    with pytest.raises(rt.RTError, match="transparent"):
        scene.set_materials(m, np.full_like(k, -1))  # the glass is flagged but keeps no material
    unflagged = rt.Scene(gpu_ctx, w.abc, w.albedo)
    with pytest.raises(rt.RTError, match="transparent"):
        unflagged.set_materials(m, k)
    scene.set_materials(m, k)
    accel = rt.Accel(scene)
    with pytest.raises(rt.RTError, match="bounce"):
        accel.render(w.camera, w.lights, bounce=True)
    assert accel.render(w.camera, w.lights)["stats"]["rays"][2] > 0


def test_material_terrain_frame(gpu_ctx, rt, O):
    """The 1M-triangle terrain with a mirror wall, a glass pane and wet rock
    (the benchmarked material workload) at 480 x 270."""
    from raptor_amd import workloads

    w, m, k, f = workloads.material_terrain(x_res=480, y_res=270)
    out, ref = _both(rt, O, gpu_ctx, w, m, k, f)
    e = _check(out, ref)
    rays = out["stats"]["rays"]
    assert rays[2] > 0.5 * rays[0] and rays[3] > 0.25 * rays[0]
    assert np.median(e) < 1e-7


def test_cook_torrance_frame(gpu_ctx, rt, O):
    """cook_torrance_cxy objects and mirror wall next to a phong glass pane:
    chroma conversion, recursion through rs and ts, mixed material kinds."""
    w, m, k, f = _scene(res=96, mirror=0.8, glass=(0.6, 1.5))
    # Material 0 is the mirror wall, 1 the glass; objects take material 2.
    m = [dict(kind=1, x=0.30, y=0.32, rd=0.05, rs=0.8, sr=0.1),
         m[1],
         dict(kind=1, x=0.36, y=0.38, rd=0.7, rs=0.3, sr=0.2)]
    k = k.copy()
    obj = k == -1
    obj[:2] = False
    k[obj] = 2
    out, ref = _both(rt, O, gpu_ctx, w, m, k, f)
    _check(out, ref)
    assert out["stats"]["rays"][2] > 96 * 96 * 0.5
    # A glass with cook_torrance transmittance, multi-sample.
    m2 = [m[0], dict(kind=1, ts=0.5, ri_r=1.3, rd=0.1), m[2]]
    out, ref = _both(rt, O, gpu_ctx, w, m2, k, f, spp=2, seed=5)
    _check(out, ref, spp=2)


def test_cook_torrance_errors(gpu_ctx, rt):
    w, m, k, f = _scene(res=32)
    scene = rt.Scene(gpu_ctx, w.abc, w.albedo, f)
    zeros = np.zeros(w.abc.shape[0], np.int32)
    with pytest.raises(rt.RTError, match="out of range"):
        scene.set_materials([dict(kind=1, sr=0.6)], zeros)
    with pytest.raises(rt.RTError, match="out of range"):
        scene.set_materials([dict(kind=1, x=1.5)], zeros)
    bad = rt.material_table([dict()])
    bad["kind"] = 7
    with pytest.raises(rt.RTError, match="unknown kind"):
        scene.set_materials(bad, zeros)
