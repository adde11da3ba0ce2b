"""Scene ingestion on the host: the library's OFF, PLY and OBJ/MTL parsers
(rt_mesh_*) against the oracle's restatement (oracle/ingest.py), and the
oracle pinned by what the geometry fixes: a triangle mesh comes back as its
face table, an ear-clipped simple polygon keeps its area and orientation, a
binary file equals its ASCII form, negative OBJ indices equal positive ones,
MTL illumination models set the terms the reference's table names.

Real inputs: tests/golden/meshes/*.off are meshes from the reference's own
regression data (tests/golden/meshes/SOURCES.txt). No GPU is needed."""
import math
import os
import struct

import numpy as np
import pytest

from oracle import ingest as OI

HERE = os.path.dirname(os.path.abspath(__file__))
MESHES = os.path.join(HERE, "golden", "meshes")
FIXTURES = ["test", "table", "tstTorusModel", "helmet", "helix", "hornbug"]


def mat_rows(m):
    """The library mesh's per-triangle material rows in the oracle's layout."""
    q = m.materials[m.tri_material]
    return np.concatenate([q["ka"], q["kd"], q["ks"], q["shininess"][:, None], q["transmittance"][:, None],
                           q["refractive_index"][:, None], q["reflectance"][:, None],
                           q["kind"][:, None].astype(np.float32), q["tex"].astype(np.float32)],
                          axis=1).astype(np.float32)


def same_bits(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def assert_matches_oracle(m, ref):
    assert m.n == ref["abc"].shape[0]
    assert same_bits(m.abc, ref["abc"])
    assert same_bits(mat_rows(m), ref["mats"])
    if ref["vn"] is None:
        assert m.vertex_normals is None
    else:
        assert same_bits(m.vertex_normals, ref["vn"])
    if ref.get("vt") is None:
        assert m.texture_coords is None
    else:
        assert same_bits(m.texture_coords, ref["vt"])
    if "maps" in ref:
        # every MTL map: a planar map (centre 0, normal (1, 0, 0), size 0,
        # mirror wrapping) of its image (obj_parser.cc:241)
        assert [int(x) for x in m.maps["image"]] == ref["maps"]
        assert np.all(m.maps["kind"] == 0) and np.all(m.maps["u_wrap"] == 3) and np.all(m.maps["v_wrap"] == 3)
        assert np.all(m.maps["n"] == np.float32([1, 0, 0])) and np.all(m.maps["s"] == 0)
        assert np.all(m.maps["u"] == np.float32([0, 0, 1])) and np.all(m.maps["v"] == np.float32([0, 1, 0]))
        assert m.images.shape[0] == len(ref["images"])
        for im, want in zip(m.images, ref["images"]):
            h, w, cpp = want.shape
            assert (im["height"], im["width"], im["cpp"]) == (h, w, cpp)
            got = m.texels[im["offset"]:im["offset"] + want.size].reshape(want.shape)
            assert same_bits(got, want)
        assert m.info["texture_maps"] == ref["unread_maps"]
    assert m.info["faces"] == ref["faces"]
    assert m.info["skipped_faces"] == ref["skipped"]


def off_text(verts, faces):
    out = ["OFF", f"{len(verts)} {len(faces)} 0"]
    out += [" ".join(repr(float(x)) for x in v) for v in verts]
    out += [" ".join(str(x) for x in [len(f)] + list(f)) for f in faces]
    return "\n".join(out) + "\n"


# ----------------------------------------------------------------- OFF

@pytest.mark.parametrize("name", FIXTURES)
def test_off_fixture_matches_oracle(rt, name):
    path = os.path.join(MESHES, name + ".off")
    m = rt.Mesh(path)
    ref = OI.parse_off(open(path).read())
    assert_matches_oracle(m, ref)
    assert m.info["materials"] == 1
    assert np.all(m.materials["kd"] == 170.0)


@pytest.mark.parametrize("name", FIXTURES)
def test_off_triangle_mesh_is_its_face_table(rt, name):
    # Closed form: a mesh of triangles comes back face by face, in file order.
    path = os.path.join(MESHES, name + ".off")
    lines = open(path).read().split("\n")
    nv, nf = (int(x) for x in lines[1].split()[:2])
    verts = np.array([[float(x) for x in ln.split()[:3]] for ln in lines[2:2 + nv]], np.float64).astype(np.float32)
    faces = np.array([[int(x) for x in ln.split()[1:4]] for ln in lines[2 + nv:2 + nv + nf]], np.int64)
    m = rt.Mesh(path)
    assert m.n == nf
    assert same_bits(m.abc, verts[faces].reshape(-1, 9))


# ----------------------------------------------------------- polygons

def polygon_area_normal(p):
    p = np.asarray(p, np.float64)
    s = np.zeros(3)
    for i in range(len(p)):
        s += np.cross(p[i], p[(i + 1) % len(p)])
    return 0.5 * np.linalg.norm(s), s / np.linalg.norm(s)


def tri_area_normal(t):
    t = np.asarray(t, np.float64).reshape(3, 3)
    c = np.cross(t[1] - t[0], t[2] - t[0])
    return 0.5 * np.linalg.norm(c), c


def frame(rng):
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    return q


def place(pts2d, rng):
    q = frame(rng)
    off = rng.uniform(-2, 2, 3)
    return [tuple(np.float32(x) for x in (q[:, 0] * a + q[:, 1] * b + off)) for a, b in pts2d]


def star(n, r0, r1, rng):
    return [((r0 if i % 2 == 0 else r1) * math.cos(2 * math.pi * i / n),
             (r0 if i % 2 == 0 else r1) * math.sin(2 * math.pi * i / n)) for i in range(n)]


SHAPES = {
    "triangle": lambda rng: [(0, 0), (1, 0), (0.2, 0.9)],
    "square": lambda rng: [(0, 0), (1, 0), (1, 1), (0, 1)],
    "hexagon": lambda rng: [(math.cos(k * math.pi / 3), math.sin(k * math.pi / 3)) for k in range(6)],
    "ngon13": lambda rng: [(math.cos(k * 2 * math.pi / 13), math.sin(k * 2 * math.pi / 13)) for k in range(13)],
    "L": lambda rng: [(0, 0), (2, 0), (2, 1), (1, 1), (1, 3), (0, 3)],
    "star10": lambda rng: star(10, 1.0, 0.4, rng),
    "star16": lambda rng: star(16, 1.0, 0.55, rng),
    "comb": lambda rng: [(0, 0), (5, 0), (5, 2), (4, 2), (4, 0.5), (3, 0.5), (3, 2), (2, 2), (2, 0.5), (1, 0.5),
                         (1, 2), (0, 2)],
    "random_star": lambda rng: [(r * math.cos(a), r * math.sin(a)) for a, r in
                                zip(np.sort(rng.uniform(0, 2 * math.pi, 24)), rng.uniform(0.3, 1.0, 24))],
}


@pytest.mark.parametrize("shape", sorted(SHAPES))
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_polygon_triangulation(rt, shape, seed):
    rng = np.random.default_rng(seed + 17 * len(shape))
    pts = place(SHAPES[shape](rng), rng)
    n = len(pts)
    text = off_text(pts, [list(range(n))])
    m = rt.Mesh(data=text.encode(), fmt=rt.FORMAT_OFF)
    ref = OI.parse_off(text)
    assert_matches_oracle(m, ref)
    if n == 3:
        return
    # The oracle's triangulation: n - 2 ears covering the polygon exactly,
    # each facing the polygon's normal.
    area, normal = polygon_area_normal(pts)
    tris = ref["abc"]
    assert tris.shape[0] == n - 2
    tot = 0.0
    for t in tris:
        a, c = tri_area_normal(t)
        tot += a
        assert np.dot(c, normal) > 0
    assert abs(tot - area) <= 1e-5 * area


@pytest.mark.parametrize("pts,want", [
    # Square: every corner is sqrt(0.5) from the centre, so the walk starts
    # at corner 0 (the first strictly furthest); ears (3, 0, 1), (3, 1, 2).
    ([(0, 0), (1, 0), (1, 1), (0, 1)], [(3, 0, 1), (3, 1, 2)]),
    # House: corner 3 is the unique furthest from the centre (1, 1); ears
    # (2, 3, 4), then (2, 4, 0), then (2, 0, 1).
    ([(0, 0), (2, 0), (2, 1), (1, 3), (0, 1)], [(2, 3, 4), (2, 4, 0), (2, 0, 1)]),
])
def test_triangulation_known_answers(rt, pts, want):
    # Hand-traced from rt/parsers/polygon_to_triangles.h:141-386.
    p3 = [(np.float32(x), np.float32(y), np.float32(0)) for x, y in pts]
    assert OI.face_to_triangles(p3) == want
    m = rt.Mesh(data=off_text(p3, [list(range(len(p3)))]).encode(), fmt=rt.FORMAT_OFF)
    assert same_bits(m.abc, np.asarray([[*p3[i], *p3[j], *p3[k]] for i, j, k in want], np.float32))


def test_polygon_with_straight_corners(rt):
    # Points in the middle of a side are dropped without a triangle.
    rng = np.random.default_rng(5)
    pts = place([(0, 0), (0.5, 0), (1, 0), (1, 0.5), (1, 1), (0, 1)], rng)
    text = off_text(pts, [list(range(6))])
    m = rt.Mesh(data=text.encode(), fmt=rt.FORMAT_OFF)
    ref = OI.parse_off(text)
    assert_matches_oracle(m, ref)
    area, _ = polygon_area_normal(pts)
    tot = sum(tri_area_normal(t)[0] for t in ref["abc"])
    assert all(tri_area_normal(t)[0] > 1e-6 for t in ref["abc"])
    assert abs(tot - area) <= 1e-5 * area


def test_short_faces_give_no_triangle(rt):
    text = off_text([(0, 0, 0), (1, 0, 0), (0, 1, 0)], [[0, 1], [0], [0, 1, 2]])
    m = rt.Mesh(data=text.encode(), fmt=rt.FORMAT_OFF)
    assert m.n == 1 and m.info["skipped_faces"] == 2 and m.info["faces"] == 3


# ---------------------------------------------------------------- PLY

def ply_bytes(verts, faces, fmt="ascii", extra_vertex=(), colours=None, intensity=False, extra_face=()):
    """A PLY file. extra_vertex: (type, name) properties after x, y, z;
    colours: per-face rgb bytes (binary only)."""
    head = ["ply", f"format {fmt} 1.0", "comment made by the ingestion tests", f"element vertex {len(verts)}",
            "property float x", "property float y", "property float z"]
    head += [f"property {t} {n}" for t, n in extra_vertex]
    head += [f"element face {len(faces)}"]
    if intensity:
        head += ["property uchar intensity"]
    head += ["property list uchar int vertex_indices"]
    if colours is not None:
        head += ["property uchar red", "property uchar green", "property uchar blue"]
    head += [f"property uchar {n}" for n in extra_face]
    head += ["end_header"]
    out = ("\n".join(head) + "\n").encode()
    if fmt == "ascii":
        rows = [" ".join(repr(float(x)) for x in v) + "".join(" 0" for _ in extra_vertex) for v in verts]
        rows += [" ".join(str(x) for x in [len(f)] + list(f)) for f in faces]
        return out + ("\n".join(rows) + "\n").encode()
    e = ">" if fmt == "binary_big_endian" else "<"
    for v in verts:
        out += struct.pack(e + "3f", *v)
        for t, _ in extra_vertex:
            out += b"\x07" if t in ("uchar", "uint8") else struct.pack(e + "f", 0.25)
    for i, f in enumerate(faces):
        if intensity:
            out += b"\x09"
        out += struct.pack(e + "B" + "i" * len(f), len(f), *f)
        if colours is not None:
            out += bytes(colours[i])
        out += b"\x05" * len(extra_face)
    return out


def mesh_soup(seed, nv=40, nf=60):
    rng = np.random.default_rng(seed)
    verts = rng.uniform(-1, 1, (nv, 3)).astype(np.float32)
    faces = [list(rng.choice(nv, 3, replace=False)) for _ in range(nf)]
    faces[3] = [faces[3][0], faces[3][0], faces[3][1]]  # repeated corner: filtered
    verts[5] = verts[6]
    faces[4] = [5, 6, 7]  # coincident corners: filtered
    faces[7] = [0, 1, 2, 3]  # a quad
    return [tuple(v) for v in verts], faces


@pytest.mark.parametrize("fmt", ["ascii", "binary_big_endian", "binary_little_endian"])
def test_ply_matches_oracle(rt, fmt):
    verts, faces = mesh_soup(3)
    extra = (("float", "nx"), ("float", "ny"), ("float", "nz"), ("uchar", "red"), ("uchar", "green"),
             ("uchar", "blue"), ("uchar", "alpha"))
    data = ply_bytes(verts, faces, fmt, extra_vertex=extra)
    m = rt.Mesh(data=data, fmt=rt.FORMAT_PLY)
    ref = OI.parse_ply(data)
    assert_matches_oracle(m, ref)
    assert m.info["skipped_faces"] == 2
    assert np.all(m.materials["kd"] == 200.0)


def test_ply_binary_equals_ascii(rt):
    verts, faces = mesh_soup(4)
    a = rt.Mesh(data=ply_bytes(verts, faces, "ascii"), fmt=rt.FORMAT_PLY)
    for fmt in ("binary_big_endian", "binary_little_endian"):
        b = rt.Mesh(data=ply_bytes(verts, faces, fmt), fmt=rt.FORMAT_PLY)
        assert same_bits(a.abc, b.abc) and np.array_equal(a.tri_material, b.tri_material)


def test_ply_face_colours(rt):
    verts, faces = mesh_soup(6)
    rng = np.random.default_rng(6)
    colours = [tuple(int(x) for x in rng.integers(0, 256, 3)) for _ in faces]
    colours[10] = colours[11] = (12, 200, 255)
    colours[12] = (250, 223, 1)  # keyed 0xdefa by the signed-char key: takes the default material
    data = ply_bytes(verts, faces, "binary_big_endian", colours=colours, intensity=True, extra_face=("flags",))
    m = rt.Mesh(data=data, fmt=rt.FORMAT_PLY)
    ref = OI.parse_ply(data)
    assert_matches_oracle(m, ref)
    kept = [i for i, f in enumerate(faces) if i not in (3, 4)]
    kd = m.materials[m.tri_material]["kd"]
    row, first = 0, {}
    for i in kept:
        nt = 1 if len(faces[i]) == 3 else 2
        want = (200.0, 200.0, 200.0) if i == 12 else colours[i]
        assert np.all(kd[row:row + nt] == np.asarray(want, np.float32)), i
        first[i] = row
        row += nt
    assert m.tri_material[first[10]] == m.tri_material[first[11]]  # one material per colour
    assert m.tri_material[first[12]] == 0


# ---------------------------------------------------------------- OBJ

MTL = """# materials
newmtl matte
Ka 0.2 0.2 0.2
Kd 0.8 0.4 0.2
Ks 0.5
Ns 12
illum 1
newmtl shiny
Kd 0.1 0.2 0.3
Ks 0.9 0.8 0.7
Ns 40
Ni 1.3
illum 2
newmtl mirror
Kd 0.5 0.5 0.5
illum 3
newmtl glass
Kd 0.2 0.2 0.2
Ni 1.5
d 0.25
illum 4
newmtl inherit
Kd 0.3 0.3 0.3
map_Kd wood.tga
newmtl flat
Kd 0.6 0.7 0.8
Ks 0.3
Tr 0.5
illum 0
"""

OBJ = """# a small scene
mtllib scene.mtl
v 0 0 0
v 2 0 0
v 2 2 0
v 0 2 0
v 1 1 1.5
v 3 0 0
v 3 1 0.5
vn 0 0 3
vn 0 1 1
vn 1 0 1
vt 0 0
vt 1 0
vt 1 1
g base
f 1/1/1 2/2/1 3/3/1 4/1/1
usemtl matte
f 1//2 2//3 5//1
usemtl shiny
f -7 -6 -3
usemtl unknown
f 2 6 7 3
usemtl mirror
f 3/1 4/2 5/3
usemtl glass
f 1 1 2
f 4 1 5
usemtl inherit
f 6 7 5
usemtl flat
f 2 6 5
s off
o thing
"""


TEXTURES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "textures")


def write_obj(tmp_path, obj=OBJ, mtl=MTL):
    import shutil

    for name in ("wood.tga", "marble.tga"):
        shutil.copyfile(os.path.join(TEXTURES, name), tmp_path / name)
    (tmp_path / "scene.mtl").write_text(mtl)
    p = tmp_path / "scene.obj"
    p.write_text(obj)
    return p


def test_obj_matches_oracle(rt, tmp_path):
    p = write_obj(tmp_path)
    m = rt.Mesh(str(p))
    ref = OI.parse_obj(OBJ, lambda name: (tmp_path / name).read_text(),
                       lambda name: (tmp_path / name).read_bytes())
    assert_matches_oracle(m, ref)
    assert m.info["texture_maps"] == 0 and m.info["maps"] == 1 and m.info["images"] == 1
    assert m.info["skipped_faces"] == 1  # f 1 1 2
    # vn are normalised; triangles without normals are marked.
    vn = m.vertex_normals
    assert vn is not None and vn[0, 0] == np.float32(0.0) and vn[0, 2] == np.float32(1.0)
    marked = vn[:, 0] == np.finfo(np.float32).max
    assert marked.sum() == m.n - 3


def test_obj_negative_indices_equal_positive(rt):
    body = "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nv 0.5 0.5 1\n"
    pos = rt.Mesh(data=(body + "f 1 2 3 4\nf 2 3 5\n").encode(), fmt=rt.FORMAT_OBJ)
    neg = rt.Mesh(data=(body + "f -5 -4 -3 -2\nf -4 -3 -1\n").encode(), fmt=rt.FORMAT_OBJ)
    assert pos.n == 3 and same_bits(pos.abc, neg.abc)


def test_mtl_illumination_table(rt, tmp_path):
    # obj_parser.cc:294-378: the terms each illumination model keeps, the
    # clamp of the transmittance and the halved ambient colour.
    p = write_obj(tmp_path)
    m = rt.Mesh(str(p))
    by_kd = {tuple(np.round(r["kd"].astype(np.float64) / 255.0, 4).tolist()): r for r in m.materials
             if r["kind"] == 2}
    f32 = np.float32
    matte = by_kd[(0.8, 0.4, 0.2)]
    assert np.all(matte["ka"] == f32(f32(0.2) * f32(255.0)) * f32(0.5))
    assert np.all(matte["ks"] == 0) and matte["shininess"] == 0 and matte["reflectance"] == 0
    shiny = by_kd[(0.1, 0.2, 0.3)]
    assert shiny["shininess"] == 40 and shiny["refractive_index"] == 1 and shiny["ks"][0] == f32(0.9) * f32(255.0)
    mirror = by_kd[(0.5, 0.5, 0.5)]
    assert mirror["reflectance"] == f32(0.75) and mirror["refractive_index"] == 1
    glass = by_kd[(0.2, 0.2, 0.2)]
    assert glass["transmittance"] == f32(1.0) - f32(0.25) and glass["refractive_index"] == f32(1.5)
    inherit = by_kd[(0.3, 0.3, 0.3)]  # no illum line: keeps illum 4 of the material before
    assert inherit["reflectance"] == 0 and inherit["refractive_index"] == 1
    flat = [r for r in m.materials if r["kind"] == 2 and r["ka"][0] > 0 and np.all(r["kd"] == 0)]
    assert len(flat) == 1  # illum 0: the diffuse colour becomes the ambient one, halved
    assert np.all(flat[0]["ka"] == np.asarray([f32(f32(v) * f32(255.0)) * f32(0.5) for v in (0.6, 0.7, 0.8)]))
    assert flat[0]["transmittance"] == f32(0.5) and np.all(flat[0]["ks"] == 0)


def test_mtllib_makes_its_last_material_current(rt, tmp_path):
    p = write_obj(tmp_path, obj="mtllib scene.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")
    m = rt.Mesh(str(p))
    assert m.n == 1 and np.all(m.materials[m.tri_material[0]]["kd"] == 0)  # "flat": illum 0


# ------------------------------------------------------------- errors

@pytest.mark.parametrize("fmt,data,what", [
    (0, b"OFX\n3 1 0\n0 0 0\n1 0 0\n0 1 0\n3 0 1 2\n", "not an OFF"),
    (0, b"OFF\n3 1 0\n0 0 0\n1 0 0\n0 1 0\n3 0 1 7\n", "out of range"),
    (0, b"OFF\n3 2 0\n0 0 0\n1 0 0\n0 1 0\n3 0 1 2\n", "ends inside"),
    (1, b"ply\nformat ascii 1.0\nelement vertex 1\nproperty double x\nend_header\n", "start with float x"),
    (1, b"ply\nformat ascii 1.0\nelement vertex 1\nproperty float x\nproperty float y\nproperty float z\n"
        b"property int q\nelement face 0\nproperty list uchar int vertex_indices\nend_header\n0 0 0 0\n",
     "unsupported PLY vertex property"),
    (2, b"v 0 0 0\nv 1 0 0\nf 1 2 9\n", "out of range"),
])
def test_malformed_input_is_an_error(rt, fmt, data, what):
    with pytest.raises(rt.RTError, match=what):
        rt.Mesh(data=data, fmt=fmt)


def test_format_from_extension_and_missing_file(rt, tmp_path):
    p = tmp_path / "tri.ply"
    p.write_bytes(ply_bytes([(0, 0, 0), (1, 0, 0), (0, 1, 0)], [[0, 1, 2]]))
    assert rt.Mesh(str(p)).n == 1
    with pytest.raises(rt.RTError, match="cannot open"):
        rt.Mesh(str(tmp_path / "absent.off"))
    with pytest.raises(rt.RTError, match="no format"):
        rt.Mesh(str(tmp_path / "scene.xyz"))


NFF = """b 0.1 0.2 0.3
v
from 1.5 2.0 -6.0
at 0.2 0.4 0.5
up 0.0 1.0 0.0
angle 40
hither 1
resolution 160 120
l 4 6 -3
l -3 5 -2 0.5 0.6 0.7
f 0.8 0.6 0.4 0.7 0.3 12 0 1
p 4
-2 -1 -2
2 -1 -2
2 -1 2
-2 -1 2
s 0 0 0 1
f 0.2 0.5 0.9 0.8 0.1 30 0.4 1.4
pp 3
-1 0 0 0 0.2 -1
1 0 0 0.1 0 -1
0 1.5 0.2 -0.1 0.1 -1
p 5
-1.5 0.5 1
-0.5 0.5 1
-0.3 1.5 1
-1 2 1
-1.7 1.4 1
"""


def test_nff_matches_oracle(rt):
    m = rt.Mesh(data=NFF.encode(), fmt=rt.FORMAT_NFF)
    ref = OI.parse_nff(NFF)
    assert_matches_oracle(m, ref)
    assert m.n == 2 + 1 + 3 and m.info["skipped_statements"] == 1  # the sphere
    c = m.camera
    got = np.concatenate([c.pos, c.x_axis, c.y_axis, c.z_axis, [c.w, c.h, c.t]]).astype(np.float32)
    want = np.concatenate([ref["camera"][k] for k in ("pos", "x", "y", "z")] +
                          [[ref["camera"]["w"], ref["camera"]["h"], ref["camera"]["t"]]]).astype(np.float32)
    assert same_bits(got, want) and (c.x_res, c.y_res) == (160, 120)
    assert same_bits(m.background, ref["background"]) and same_bits(m.lights, ref["lights"])


def test_nff_camera_closed_form(rt):
    # rt/parsers/nff_parser.cc:58-110: an orthonormal basis with x = up x at,
    # half width tan(angle / 2) * hither, image plane at hither.
    m = rt.Mesh(data=NFF.encode(), fmt=rt.FORMAT_NFF)
    c = m.camera
    x, y, z = (np.asarray(v, np.float64) for v in (c.x_axis, c.y_axis, c.z_axis))
    at = np.array([0.2 - 1.5, 0.4 - 2.0, 0.5 + 6.0])
    assert np.allclose(z, at / np.linalg.norm(at), atol=1e-6)
    assert np.allclose([x @ y, y @ z, x @ z], 0.0, atol=1e-6) and np.allclose(np.cross(y, z), x, atol=1e-6)
    assert y[1] > 0 and abs(c.w - np.tan(np.radians(20.0))) < 1e-6 and c.t == 1.0
    assert np.allclose(m.background, [25.5, 51.0, 76.5]) and np.allclose(m.lights[1, 3:6], [0.5, 0.6, 0.7])
    # fill material: kd = rgb kd, ks = rgb ks, reflectance ks (rt/materials/phong_shader.h:17-19)
    mat = m.materials[m.tri_material[0]]
    assert np.allclose(mat["kd"], np.array([0.8, 0.6, 0.4]) * 255 * 0.7, rtol=1e-6)
    assert np.allclose(mat["ks"], np.array([0.8, 0.6, 0.4]) * 255 * 0.3, rtol=1e-6) and mat["reflectance"] == np.float32(0.3)


@pytest.mark.parametrize("seed", range(6))
def test_random_polygon_soup_matches_oracle(rt, seed):
    # Random faces of 3 to 14 corners on a shared vertex table: star-shaped
    # and arbitrary (self-crossing) rings, repeated corners and collinear
    # runs, in random planes. The library and the restatement must agree bit
    # for bit on every triangle, whatever the ear clipping makes of them.
    rng = np.random.default_rng(100 + seed)
    verts, faces = [], []
    for f in range(60):
        n = int(rng.integers(3, 15))
        kind = f % 4
        if kind == 0:  # star-shaped
            ang = np.sort(rng.uniform(0, 2 * np.pi, n))
            pts = [(r * np.cos(a), r * np.sin(a)) for a, r in zip(ang, rng.uniform(0.2, 1.0, n))]
        elif kind == 1:  # arbitrary ring
            pts = [tuple(p) for p in rng.uniform(-1, 1, (n, 2))]
        elif kind == 2:  # a repeated corner
            ang = np.sort(rng.uniform(0, 2 * np.pi, n))
            pts = [(np.cos(a), np.sin(a)) for a in ang]
            pts[n // 2] = pts[n // 2 - 1]
        else:  # collinear runs along the sides of a square
            m = max(1, n // 4)
            side = [(t, 0.0) for t in np.linspace(0, 1, m, endpoint=False)]
            pts = side + [(1.0, t) for t, _ in side] + [(1.0 - t, 1.0) for t, _ in side] + \
                [(0.0, 1.0 - t) for t, _ in side]
        base = len(verts)
        verts += place(pts, rng)
        faces.append(list(range(base, base + len(pts))))
    text = off_text(verts, faces)
    m = rt.Mesh(data=text.encode(), fmt=rt.FORMAT_OFF)
    assert_matches_oracle(m, OI.parse_off(text))
    ply = ply_bytes(verts, faces, "binary_little_endian")
    assert_matches_oracle(rt.Mesh(data=ply, fmt=rt.FORMAT_PLY), OI.parse_ply(ply))
    obj = "".join(f"v {x!r} {y!r} {z!r}\n" for x, y, z in (map(float, v) for v in verts))
    obj += "".join("f " + " ".join(str(i + 1) for i in f) + "\n" for f in faces)
    assert_matches_oracle(rt.Mesh(data=obj.encode(), fmt=rt.FORMAT_OBJ), OI.parse_obj(obj))
