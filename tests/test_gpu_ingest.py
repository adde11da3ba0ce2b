"""GPU frames of parsed scenes and of smooth-shaded (vertex-normal) scenes,
through the C-ABI, against the oracle.

Scenes: the reference's own meshes (tests/golden/meshes, OFF, phong kd 170),
an OBJ/MTL scene with polygons, vertex normals, a mirror and a glass
material, a binary PLY with face colours, and a terrain with per-vertex
normals averaged from its faces (with and without materials). The camera is
the reference regression fixture's (screen width 10 at distance 20).
Primary hit ids are bit-exact except at exact-t ties; radiance agrees within
1e-5 relative; secondary-ray counts equal the oracle's.
"""
import os

import numpy as np
import pytest

from helpers import compare_hits, oracle_material_rows, rel_err

pytestmark = pytest.mark.gpu

RAD_TOL = 1e-5
HERE = os.path.dirname(os.path.abspath(__file__))
MESHES = os.path.join(HERE, "golden", "meshes")


def _flags(materials, tri_material):
    return np.where(materials["transmittance"][tri_material] > 0, 2, 0).astype(np.uint8)


def _render_mesh(rt, O, ctx, mesh, w, width=4):
    scene = rt.Scene.from_mesh(ctx, mesh)
    accel = rt.Accel(scene, width=width)
    out = accel.render(w.camera, w.lights)
    albedo = (mesh.materials["kd"][mesh.tri_material] / np.float32(255.0)).astype(np.float32)
    ref = O.OracleScene(mesh.abc, _flags(mesh.materials, mesh.tri_material)).render(
        albedo, w.camera, w.lights, nthreads=8, materials=oracle_material_rows(mesh.materials),
        tri_material=mesh.tri_material, vertex_normals=mesh.vertex_normals)
    return out, ref


def _check(out, ref, nl=1, min_hits=100):
    bad, ties = compare_hits(out["primary"], ref["prim"], ref["tuv"])
    assert not bad.any(), f"{bad.sum()} primary hit ids differ"
    assert int((ref["prim"] >= 0).sum()) >= min_hits
    e = rel_err(out["rgb"].reshape(-1, 3)[~ties], ref["rgb"][~ties])
    assert e.max() <= RAD_TOL, (e.max(), int((e > RAD_TOL).sum()))
    rays = out["stats"]["rays"]
    assert rays[1] == int((ref["prim"] >= 0).sum()) * nl
    assert rays[2] == int(ref["sec_counters"][0])
    assert rays[3] == int(ref["sec_counters"][1])


@pytest.mark.parametrize("name,width", [("tstTorusModel", 4), ("helmet", 8), ("hornbug", 4), ("helix", 4)])
def test_parsed_off_mesh_frame(gpu_ctx, rt, O, name, width):
    from raptor_amd import workloads

    mesh = rt.Mesh(os.path.join(MESHES, name + ".off"))
    w = workloads.mesh_view(mesh.abc, 192, 144)
    out, ref = _render_mesh(rt, O, gpu_ctx, mesh, w, width)
    _check(out, ref)


OBJ = """mtllib scene.mtl
v -3 -1 -3
v 3 -1 -3
v 3 -1 3
v -3 -1 3
v -1 -1 0.5
v 1 -1 0.5
v 1.2 0.5 0.5
v 0 1.4 0.6
v -1.2 0.5 0.5
v -0.8 -0.6 -0.9
v 0.8 -0.6 -0.9
v 0.8 0.9 -0.9
v -0.8 0.9 -0.9
v -2 -1 1.5
v 2 -1 1.5
v 2 2 1.5
v -2 2 1.5
vn 0 1 0
vn 0.2 1 -0.1
vn -0.3 1 0.2
vn 0.5 0.2 -1
vn -0.4 0.6 -1
vn 0 -0.3 -1
usemtl floor
f 1//1 2//2 3//1 4//3
usemtl gold
f 5//4 6//5 7//6 8//4 9//5
usemtl glass
f 10 11 12 13
usemtl mirror
f 14 15 16 17
"""

MTL = """newmtl floor
Kd 0.7 0.7 0.6
Ka 0.1 0.1 0.1
illum 1
newmtl gold
Kd 0.8 0.6 0.2
Ks 0.6 0.5 0.3
Ns 20
illum 2
newmtl glass
Kd 0.05 0.05 0.05
Ni 1.45
d 0.3
illum 4
newmtl mirror
Kd 0.1 0.1 0.1
illum 3
"""


def test_obj_scene_with_normals_mirror_and_glass(gpu_ctx, rt, O, tmp_path):
    from raptor_amd import workloads

    (tmp_path / "scene.mtl").write_text(MTL)
    (tmp_path / "scene.obj").write_text(OBJ)
    mesh = rt.Mesh(str(tmp_path / "scene.obj"))
    assert mesh.vertex_normals is not None and mesh.n == 2 + 3 + 2 + 2
    w = workloads.mesh_view(mesh.abc, 160, 120, direction=(0.2, -0.35, 1.0))
    out, ref = _render_mesh(rt, O, gpu_ctx, mesh, w)
    _check(out, ref)
    assert out["stats"]["rays"][2] > 1000  # mirror and glass rays


def test_binary_ply_with_face_colours(gpu_ctx, rt, O):
    import struct

    from raptor_amd import workloads

    rng = np.random.default_rng(9)
    g = 24
    xs, zs = np.meshgrid(np.linspace(-1, 1, g), np.linspace(-1, 1, g))
    ys = 0.2 * np.sin(3 * xs) * np.cos(2 * zs) + rng.uniform(-0.02, 0.02, xs.shape)
    verts = np.stack([xs, ys, zs], -1).reshape(-1, 3).astype(np.float32)
    faces = [[r * g + c, r * g + c + 1, (r + 1) * g + c + 1, (r + 1) * g + c] for r in range(g - 1)
             for c in range(g - 1)]
    head = ("ply\nformat binary_big_endian 1.0\nelement vertex %d\nproperty float x\nproperty float y\n"
            "property float z\nelement face %d\nproperty list uchar int vertex_indices\nproperty uchar red\n"
            "property uchar green\nproperty uchar blue\nend_header\n" % (len(verts), len(faces))).encode()
    body = b"".join(struct.pack(">3f", *v) for v in verts)
    for f in faces:
        body += struct.pack(">B4i", 4, *f) + bytes(int(x) for x in rng.integers(40, 256, 3))
    mesh = rt.Mesh(data=head + body, fmt=rt.FORMAT_PLY)
    assert mesh.n == 2 * len(faces) and mesh.info["materials"] > 100
    w = workloads.mesh_view(mesh.abc, 160, 120, direction=(0.3, -0.8, 0.5))
    out, ref = _render_mesh(rt, O, gpu_ctx, mesh, w)
    _check(out, ref)


def smooth_normals(abc):
    """Per-vertex normals of a triangle soup: the area-weighted mean of the
    face normals sharing each vertex position."""
    v = np.asarray(abc, np.float32).reshape(-1, 3, 3)
    fn = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]).astype(np.float64)
    keys, inv = np.unique(v.reshape(-1, 3), axis=0, return_inverse=True)
    acc = np.zeros((keys.shape[0], 3))
    np.add.at(acc, inv.reshape(-1), np.repeat(fn, 3, axis=0))
    acc /= np.linalg.norm(acc, axis=1, keepdims=True)
    return acc[inv.reshape(-1)].reshape(-1, 9).astype(np.float32)


@pytest.mark.parametrize("with_materials", [False, True])
def test_smooth_terrain(gpu_ctx, rt, O, with_materials):
    from raptor_amd import workloads

    w = workloads.terrain(grid=96, x_res=160, y_res=96, bounce=False)
    vn = smooth_normals(w.abc)
    vn[::7, 0] = np.finfo(np.float32).max  # every seventh triangle keeps its geometric normal
    scene = rt.Scene(gpu_ctx, w.abc, w.albedo)
    mats = kinds = None
    if with_materials:
        mats = [dict(ka=(3.0, 3.0, 3.0), kd=(150.0, 170.0, 120.0), ks=(80.0, 80.0, 80.0), s=16.0)]
        kinds = np.where(np.arange(w.abc.shape[0]) % 3 == 0, 0, -1).astype(np.int32)
        scene.set_materials(mats, kinds)
    scene.set_vertex_normals(vn)
    accel = rt.Accel(scene)
    out = accel.render(w.camera, w.lights)
    ref = O.OracleScene(w.abc).render(w.albedo, w.camera, w.lights, nthreads=8, materials=mats, tri_material=kinds,
                                      vertex_normals=vn)
    _check(out, ref, min_hits=1000)
    # The normals change the picture: the geometric-normal frame differs.
    flat = O.OracleScene(w.abc).render(w.albedo, w.camera, w.lights, nthreads=8, materials=mats, tri_material=kinds)
    hit = ref["prim"] >= 0
    assert np.abs(flat["rgb"][hit] - ref["rgb"][hit]).max() > 1.0
    # Removing them restores the flat frame; the bounce refuses them.
    with pytest.raises(rt.RTError, match="vertex normals"):
        accel.render(w.camera, w.lights, bounce=True)
    scene.set_vertex_normals(None)
    flat_gpu = accel.render(w.camera, w.lights)
    bad, ties = compare_hits(flat_gpu["primary"], flat["prim"], flat["tuv"])
    assert not bad.any()
    assert rel_err(flat_gpu["rgb"].reshape(-1, 3)[~ties], flat["rgb"][~ties]).max() <= RAD_TOL


@pytest.mark.parametrize("with_materials", [False, True])
def test_background_colour(gpu_ctx, rt, O, with_materials):
    # Primary misses, and mirror or refraction rays that escape, take the
    # camera's background colour (rt/camera.h:234-240, rt/raytracer.cc:82-86).
    from raptor_amd import workloads

    w, mats, kinds, flags = workloads.material_scene(res=96, mirror=0.6 if with_materials else None, floor=False)
    if not with_materials:
        mats = kinds = flags = None
    bg = (40.0, 90.0, 200.0)
    scene = rt.Scene(gpu_ctx, w.abc, w.albedo, flags)
    if with_materials:
        scene.set_materials(mats, kinds)
    accel = rt.Accel(scene)
    out = accel.render(w.camera, w.lights, background=bg)
    ref = O.OracleScene(w.abc, flags).render(w.albedo, w.camera, w.lights, nthreads=8, materials=mats,
                                             tri_material=kinds, background=bg)
    bad, ties = compare_hits(out["primary"], ref["prim"], ref["tuv"])
    assert not bad.any()
    e = rel_err(out["rgb"].reshape(-1, 3)[~ties], ref["rgb"][~ties])
    assert e.max() <= RAD_TOL, e.max()
    miss = ref["prim"] < 0
    assert miss.sum() > 100 and np.all(out["rgb"].reshape(-1, 3)[miss] == np.asarray(bg, np.float32))
    black = accel.render(w.camera, w.lights)
    hit = ~miss
    if with_materials:  # escaping secondary rays brighten some hit pixels
        assert (out["rgb"].reshape(-1, 3)[hit] > black["rgb"].reshape(-1, 3)[hit]).any()
    else:
        assert np.array_equal(out["rgb"].reshape(-1, 3)[hit], black["rgb"].reshape(-1, 3)[hit])


def test_nff_scene_with_its_camera_and_lights(gpu_ctx, rt, O):
    # The viewpoint, background, lights and fill materials of an NFF file
    # (rt/parsers/nff_parser.cc:16-224), rendered as the file defines them.
    import sys

    from raptor_amd import workloads

    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from test_ingest import NFF

    mesh = rt.Mesh(data=NFF.encode(), fmt=rt.FORMAT_NFF)
    cam, bg, lights = workloads.mesh_camera_lights(mesh)
    scene = rt.Scene.from_mesh(gpu_ctx, mesh)
    accel = rt.Accel(scene)
    out = accel.render(cam, lights, background=bg)
    albedo = (mesh.materials["kd"][mesh.tri_material] / np.float32(255.0)).astype(np.float32)
    ref = O.OracleScene(mesh.abc, _flags(mesh.materials, mesh.tri_material)).render(
        albedo, cam, lights, nthreads=8, materials=oracle_material_rows(mesh.materials),
        tri_material=mesh.tri_material, vertex_normals=mesh.vertex_normals, background=bg)
    _check(out, ref, nl=2, min_hits=1000)
    miss = ref["prim"] < 0
    assert miss.any() and np.all(out["rgb"].reshape(-1, 3)[miss] == np.asarray(bg, np.float32))
