"""Oracle for scene ingestion — TEST INFRASTRUCTURE ONLY.

A pure-Python restatement of the reference's OFF, PLY, OBJ/MTL and NFF
parsers and of its polygon triangulation, written for the parity tests of the
library's parsers (only tests/ import this module). Each function cites the
reference code it restates. Float arithmetic runs on numpy float32 scalars,
one rounding per operation as the reference's float code has (no FMA), so
the emitted triangles match bit for bit.

Returns per file: abc (n, 9) float32 in emplace order, mats (n, 20) float32
rows of the material each triangle shades with (ka rgb, kd rgb, ks rgb, s,
tran, ri, rf, kind, and the six texture slots t_ka, t_kd, t_ks, t_ns,
t_refl, t_d as 1 + map index, 0 for none), vn (n, 9) float32 or None, vt
(n, 6) float32 or None, the texture maps (image index, inverted) in
creation order, the images ((h, w, cpp) float32, from read_tga), the count
of map files that are not TGA, and counts.
"""
import numpy as np

F = np.float32
FLT_MAX = float(np.finfo(np.float32).max)


def _v(x, y, z):
    return (F(x), F(y), F(z))


def _sub(a, b):
    return (a[0] - b[0], a[1] - b[1], a[2] - b[2])


def _cross(a, b):
    # common/point_t.h:498-503
    return ((a[1] * b[2]) - (a[2] * b[1]), (a[2] * b[0]) - (a[0] * b[2]), (a[0] * b[1]) - (a[1] * b[0]))


def _dot(a, b):
    # common/point_t.h:519-523
    return ((a[0] * b[0]) + (a[1] * b[1])) + (a[2] * b[2])


def _norm(a):
    # normalise<float> with EXACT_NORMALISE, common/point_t.h:553-569
    d = np.sqrt(((a[0] * a[0]) + (a[1] * a[1])) + (a[2] * a[2]))
    return (a[0] / d, a[1] / d, a[2] / d)


def _mag(a):
    return np.sqrt(((a[0] * a[0]) + (a[1] * a[1])) + (a[2] * a[2]))


def _zero(a):
    return a[0] == 0 and a[1] == 0 and a[2] == 0


def _phong(ka, kd, ks, s=0.0, tran=0.0, ri=1.0, rf=0.0, kind=0, tex=(0, 0, 0, 0, 0, 0)):
    return [F(ka[0]), F(ka[1]), F(ka[2]), F(kd[0]), F(kd[1]), F(kd[2]), F(ks[0]), F(ks[1]), F(ks[2]),
            F(s), F(tran), F(ri), F(rf), F(kind)] + [F(t) for t in tex]


def read_tga(data):
    """A TGA file as libtga's TGAReadImage with TGA_RGB reads it
    (rt/materials/texture_mapper.h read_tga): pixels in file order, BGR(A)
    swapped to RGB(A). Types 2, 3 (raw) and 10, 11 (run-length encoded).
    Returns (height, width, channels) float32."""
    d = np.frombuffer(data, np.uint8)
    id_len, cmap_type, typ = int(d[0]), int(d[1]), int(d[2])
    cmap_len, cmap_depth = int(d[5]) | (int(d[6]) << 8), int(d[7])
    w, h, depth = int(d[12]) | (int(d[13]) << 8), int(d[14]) | (int(d[15]) << 8), int(d[16])
    cpp = depth // 8
    at = 18 + id_len + (cmap_len * ((cmap_depth + 7) // 8) if cmap_type == 1 else 0)
    n = w * h
    if typ in (2, 3):
        px = d[at:at + n * cpp].reshape(n, cpp)
    else:
        out, p = [], 0
        while p < n:
            hdr = int(d[at])
            at += 1
            cnt = (hdr & 0x7F) + 1
            if hdr & 0x80:
                out.append(np.repeat(d[at:at + cpp][None, :], cnt, axis=0))
                at += cpp
            else:
                out.append(d[at:at + cnt * cpp].reshape(cnt, cpp))
                at += cnt * cpp
            p += cnt
        px = np.concatenate(out)
    px = px.astype(np.float32)
    if cpp >= 3:
        px = px[:, [2, 1, 0] + ([3] if cpp == 4 else [])]
    return px.reshape(h, w, cpp)


def face_to_triangles(p):
    """rt/parsers/polygon_to_triangles.h:141-386 (face_to_triangle_edges
    :136-229): index triples of the ear-clipped polygon p (list of float32
    3-tuples)."""
    n = len(p)
    if n < 3:
        return []
    with np.errstate(all="ignore"):
        # find_centre (:50-62): sum, then divide by the count.
        tot = (F(0), F(0), F(0))
        for q in p:
            tot = (tot[0] + q[0], tot[1] + q[1], tot[2] + q[2])
        cnt = F(n)
        com = (tot[0] / cnt, tot[1] / cnt, tot[2] / cnt)
        # find_furthest (:64-80): first strictly greater distance.
        mx, best = 0, F(0)
        for i in range(n):
            d = _mag(_sub(com, p[i]))
            if d > best:
                best, mx = d, i
        m1 = n - 1 if mx == 0 else mx - 1
        p1 = 0 if mx == n - 1 else mx + 1
        fn = _norm(_cross(_sub(p[m1], p[mx]), _sub(p[mx], p[p1])))
        invalid = [False] * n
        size = n
        no_success = 0
        out = []
        while True:
            success = True
            a, b, c = p[m1], p[mx], p[p1]
            # is_straight_line (:83-91)
            if not _zero(_cross(_sub(b, a), _sub(c, a))):
                tn = _norm(_cross(_sub(a, b), _sub(b, c)))
                if _dot(tn, fn) > 0 and not _in_triangle(p, a, b, c, fn):
                    out.append((m1, mx, p1))
                else:
                    success = False
            if success:
                size -= 1
                no_success = 0
                invalid[mx] = True
            else:
                m1 = mx
                no_success += 1
            mx = p1
            while True:
                p1 += 1
                if p1 == n:
                    p1 = 0
                if not invalid[p1]:
                    break
            if size == 2 or no_success == size:
                break
        return out


def _in_triangle(p, a, b, c, fn):
    # is_in_triangle (:111-131) with same_side (:94-108)
    ab, ca, bc = _sub(a, b), _sub(c, a), _sub(b, c)

    def same_side(q, s, e):
        return _dot(_cross(_sub(q, s), e), fn) > 0

    return any(same_side(q, a, ab) and same_side(q, b, bc) and same_side(q, c, ca) for q in p)


class _Out:
    def __init__(self):
        self.abc, self.mats, self.vn, self.vt = [], [], [], []
        self.any_vn = self.any_vt = False
        self.faces = self.skipped = 0
        self.maps, self.images, self.image_ids, self.unread_maps = [], [], {}, 0

    def add(self, a, b, c, mat, n=None, t=None):
        self.abc.append([*a, *b, *c])
        self.mats.append(mat)
        self.vn.append([*n[0], *n[1], *n[2]] if n is not None else [FLT_MAX] + [0.0] * 8)
        self.any_vn = self.any_vn or n is not None
        # rt/triangle.h:139-167: a triangle without texture coordinates marks x with MAX_DIST
        self.vt.append([*t[0], *t[1], *t[2]] if t is not None else [FLT_MAX] + [0.0] * 5)
        self.any_vt = self.any_vt or t is not None

    def polygon(self, face, mat, normals=None, texcoords=None):
        if len(face) < 3:
            self.skipped += 1
            return
        for i, j, k in face_to_triangles(face):
            self.add(face[i], face[j], face[k], mat, None if not normals else (normals[i], normals[j], normals[k]),
                     None if not texcoords else (texcoords[i], texcoords[j], texcoords[k]))

    def result(self):
        return dict(abc=np.asarray(self.abc, np.float32).reshape(-1, 9),
                    mats=np.asarray(self.mats, np.float32).reshape(-1, 20),
                    vn=np.asarray(self.vn, np.float32).reshape(-1, 9) if self.any_vn else None,
                    vt=np.asarray(self.vt, np.float32).reshape(-1, 6) if self.any_vt else None,
                    maps=list(self.maps), images=list(self.images), unread_maps=self.unread_maps,
                    faces=self.faces, skipped=self.skipped)


def parse_off(text):
    """rt/parsers/off_parser.cc:14-99: header OFF, counts, vertex lines,
    face lines; triangles kept as they are, larger faces triangulated; one
    phong_shader(0, 170, 0, 0, 0, ri 0, 0)."""
    words = text.split()
    assert words[0] == "OFF"
    nv, nf = int(words[1]), int(words[2])
    lines = text.splitlines()
    # skip the header lines holding the four header words
    k, seen = 0, 0
    while seen < 4:
        seen += len(lines[k].split())
        k += 1
    verts = []
    for line in lines[k:k + nv]:
        x, y, z = line.split()[:3]
        verts.append(_v(float(x), float(y), float(z)))
    mat = _phong((0, 0, 0), (170, 170, 170), (0, 0, 0), ri=0.0)
    o = _Out()
    for line in lines[k + nv:k + nv + nf]:
        t = line.split()
        cnt = int(t[0])
        face = [verts[int(i)] for i in t[1:1 + cnt]]
        o.faces += 1
        if len(face) == 3:
            o.add(face[0], face[1], face[2], mat)
        else:
            o.polygon(face, mat)
    return o.result()


def _ply_tri_ok(a0, b0, c0):
    # rt/parsers/ply_parser.cc:288-309 (binary) and :373-394 (ascii)
    if a0 == b0 or a0 == c0 or b0 == c0:
        return False
    with np.errstate(all="ignore"):
        a, b = _norm(_sub(a0, b0)), _norm(_sub(b0, c0))
        return not _zero(_sub(a, b)) and not _zero((a[0] + b[0], a[1] + b[1], a[2] + b[2]))


def _obj_tri_ok(a0, b0, c0):
    # rt/parsers/obj_parser.cc:152-163
    if a0 == b0 or a0 == c0 or b0 == c0:
        return False
    with np.errstate(all="ignore"):
        a, b = _norm(_sub(a0, b0)), _norm(_sub(b0, c0))
        n = _cross(a, b)
        return _dot(n, n) > 0


def parse_ply(data):
    """rt/parsers/ply_parser.cc:399-583 for the header forms the library
    accepts: vertex x y z floats then uchar/float extras; faces
    [uchar intensity] list uchar int [uchar props, red green blue run].
    binary big- or little-endian, or ascii without face colour."""
    end = data.index(b"end_header")
    end = data.index(b"\n", end) + 1
    hdr = [ln.split() for ln in data[:end].decode("ascii").splitlines()]
    hdr = [h for h in hdr if h and h[0] not in ("comment", "obj_info")]
    fmt = hdr[1][1]
    nv = int(hdr[2][2])
    i = 6
    vbytes = 0
    while hdr[i][0] == "property":
        vbytes += 1 if hdr[i][1] in ("uchar", "uint8") else 4
        i += 1
    nf = int(hdr[i][2])
    i += 1
    pre = 0
    if hdr[i][-1] == "intensity":
        pre, i = 1, i + 1
    i += 1  # the index list
    props = [h[2] for h in hdr[i:] if h[0] == "property"]
    colour = any(props[k:k + 3] == ["red", "green", "blue"] for k in range(len(props)))
    post = len(props) - (3 if colour else 0)
    default = _phong((0, 0, 0), (200, 200, 200), (0, 0, 0))
    keyed = {0xdefa: default}
    o = _Out()
    body = data[end:]
    verts = []
    if fmt == "ascii":
        lines = body.decode("ascii").splitlines()
        for ln in lines[:nv]:
            x, y, z = ln.split()[:3]
            verts.append(_v(float(x), float(y), float(z)))
        faces = [[int(t) for t in ln.split()] for ln in lines[nv:nv + nf]]
        faces = [(f[1:1 + f[0]], None) for f in faces]
    else:
        e = ">" if fmt == "binary_big_endian" else "<"
        off = 0
        for _ in range(nv):
            x, y, z = np.frombuffer(body, e + "f4", 3, off)
            verts.append((F(x), F(y), F(z)))
            off += 12 + vbytes
        faces = []
        for _ in range(nf):
            off += pre
            cnt = body[off]
            idx = np.frombuffer(body, e + "i4", cnt, off + 1).tolist()
            off += 1 + 4 * cnt
            rgb = None
            if colour:
                rgb = body[off:off + 3]
                off += 3
            off += post
            faces.append((idx, rgb))
    for idx, rgb in faces:
        mat = default
        if rgb is not None:
            # the key reads the bytes as signed chars (ply_parser.cc:262-272)
            s = [b - 256 if b > 127 else b for b in rgb]
            key = s[0] + (s[1] << 8) + (s[2] << 16)
            if key not in keyed:
                keyed[key] = _phong((0, 0, 0), tuple(rgb), (0, 0, 0))
            mat = keyed[key]
        face = [verts[k] for k in idx]
        o.faces += 1
        if len(face) == 3:
            if _ply_tri_ok(*face):
                o.add(face[0], face[1], face[2], mat)
            else:
                o.skipped += 1
        else:
            o.polygon(face, mat)
    return o.result()


def _mtl(text, o=None, read_image=None):
    """rt/parsers/obj_parser.cc:245-666: named materials, illum rules (on
    colours and texture maps), the ambient halved; coloured_mapper_shader
    rows (kind 2). Map statements (map_Ka, map_Kd, map_Ks, map_refl, map_d,
    :507-561) make maps through load_image (:218-242): read_image(name)
    gives the TGA's bytes (None: a file that is not TGA, counted in
    o.unread_maps); images are shared per file, map_d's inverted (1 - x,
    picture_functions' negative). Returns (dict, last name)."""
    out, last = {}, None
    name, st, illum = None, None, 0

    def load(file, invert):
        if not file.lower().endswith(".tga"):
            o.unread_maps += 1
            return 0
        key = file + ("_inv" if invert else "")
        if key not in o.image_ids:
            img = read_tga(read_image(file))
            if invert:
                img = F(1) - img
            o.images.append(img)
            o.image_ids[key] = len(o.images) - 1
        o.maps.append(o.image_ids[key])
        return len(o.maps)

    def close():
        nonlocal last
        if name is None:
            return
        ka, kd, ks = st["ka"], st["kd"], st["ks"]
        ns, ri, rf, tr = st["ns"], st["ri"], st["rf"], st["tr"]
        maps = list(st["maps"])
        z = (F(0), F(0), F(0))
        if illum == 0:
            ka, kd, ks, ns, ri, rf = kd, z, z, F(0), F(1), F(0)
            maps = [maps[1], None, None, None, None, maps[5]]
        elif illum == 1:
            ks, ns, ri, rf = z, F(0), F(1), F(0)
            maps[2] = maps[3] = maps[4] = None
        elif illum == 2:
            ri, rf = F(1), F(0)
            maps[4] = None
        elif illum in (3, 5, 7):
            ri, rf = F(1), F(0.75)
        elif illum == 4:
            rf = F(0)
        tr = max(F(0), min(F(1), tr))
        half = F(0.5)
        tex = [0 if m is None else load(m, k == 5) for k, m in enumerate(maps)]
        out[name] = _phong(tuple(x * half for x in ka), kd, ks, ns, tr, ri, rf, kind=2, tex=tex)
        last = name

    def colour(t):
        v = [F(float(x)) * F(255) for x in t[1:4]]
        return (v[0], v[0], v[0]) if len(v) < 3 else tuple(v)

    for ln in text.splitlines():
        t = ln.split()
        if not t or t[0].startswith("#"):
            continue
        w = t[0]
        if w.startswith("newmtl"):
            close()
            name = t[1]
            z = (F(0), F(0), F(0))
            st = dict(ka=z, kd=z, ks=z, ns=F(0), ri=F(1), rf=F(0), tr=F(0), maps=[None] * 6)
        elif w.startswith("Ka"):
            st["ka"] = colour(t)
        elif w.startswith("Kd"):
            st["kd"] = colour(t)
        elif w.startswith("Ks"):
            st["ks"] = colour(t)
        elif w.startswith("Ns"):
            st["ns"] = F(float(t[1]))
        elif w.startswith("Ni"):
            st["ri"] = F(float(t[1]))
        elif w[0] == "d":
            st["tr"] = F(1) - F(float(t[1]))
        elif w.startswith("Tr"):
            st["tr"] = F(float(t[1]))
        elif w.startswith("illum"):
            illum = int(t[1])
        elif w.startswith(("map_Ka", "map_Kd", "map_Ks", "map_refl", "map_d")):
            slot = {"map_Ka": 0, "map_Kd": 1, "map_Ks": 2}.get(w[:6], 4 if w.startswith("map_refl") else 5)
            st["maps"][slot] = t[1]
    close()
    return out, last


def parse_obj(text, read=None, read_image=None):
    """rt/parsers/obj_parser.cc:762-857 with parse_f_statement (:78-206):
    v, vn (normalised), vt (u, v), f (1-based or negative indices,
    v/vt/vn), usemtl, mtllib (read(name) gives the MTL text, read_image(name)
    a map file's bytes; the last material becomes current); default material
    phong kd 200."""
    v, vn, vt = [], [], []
    named = {}
    default = _phong((0, 0, 0), (200, 200, 200), (0, 0, 0))
    cur = default
    o = _Out()
    for ln in text.splitlines():
        t = ln.split()
        if not t or t[0].startswith("#"):
            continue
        w = t[0]
        if w.startswith("mtllib"):
            for f in t[1:]:
                mats, last = _mtl(read(f), o, read_image)
                named.update(mats)
                cur = mats[last] if last is not None else None
        elif w.startswith("usemtl"):
            if t[1] not in named:
                named[t[1]] = _phong((0, 0, 0), (200, 200, 200), (0, 0, 0))
            cur = named[t[1]]
        elif w.startswith("vn"):
            with np.errstate(all="ignore"):
                vn.append(_norm(_v(float(t[1]), float(t[2]), float(t[3]))))
        elif w.startswith("vt"):
            vt.append((F(float(t[1])), F(float(t[2]))))
        elif w[0] == "f":
            face, fn, ft = [], [], []
            for c in t[1:]:
                parts = c.split("/")
                i = int(parts[0])
                face.append(v[i - 1 if i > 0 else len(v) + i])
                if len(parts) >= 2 and parts[1]:
                    k = int(parts[1])
                    ft.append(vt[k - 1 if k > 0 else len(vt) + k])
                if len(parts) == 3:
                    k = int(parts[2])
                    fn.append(vn[k - 1 if k > 0 else len(vn) + k])
            o.faces += 1
            if len(face) == 3:
                if _obj_tri_ok(*face):
                    o.add(face[0], face[1], face[2], cur, tuple(fn) if fn else None, tuple(ft) if ft else None)
                else:
                    o.skipped += 1
            else:
                o.polygon(face, cur, fn, ft)
        elif w[0] == "v":
            v.append(_v(float(t[1]), float(t[2]), float(t[3])))
    return o.result()


def parse_nff(text):
    """rt/parsers/nff_parser.cc:16-224: b, v (six lines), l, f, p / pp;
    anything else skipped. Returns the parse result plus camera (dict of
    float32: pos, x, y, z, w, h, t, x_res, y_res), background and lights
    ((k, 7): position, rgb, falloff 0)."""
    o = _Out()
    lines = text.splitlines()
    bg = (F(0), F(0), F(0))
    cam, cam_bg, lights, cur = None, None, [], None
    k = 0

    def vals(ln):
        return [float(x) for x in ln.split()[1:]]

    while k < len(lines):
        t = lines[k].split()
        k += 1
        if not t:
            continue
        w = t[0]
        if w == "b":
            bg = tuple(F(float(x)) * F(255) for x in t[1:4])
        elif w[0] == "v":
            fr = _v(*vals(lines[k])[:3])
            to = _v(*vals(lines[k + 1])[:3])
            up = _v(*vals(lines[k + 2])[:3])
            angle = F(vals(lines[k + 3])[0]) * (F(np.pi) / F(360))
            ratio = F(np.sin(float(angle)) / np.cos(float(angle)))
            h = F(vals(lines[k + 4])[0])
            if h == 0:
                h = F(1)
            hw = ratio * h
            res = vals(lines[k + 5])
            k += 6
            at = _sub(to, fr)
            x = _cross(up, at)
            up = _cross(at, x)
            cam = dict(pos=fr, x=_norm(x), y=_norm(up), z=_norm(at), w=hw, h=hw, t=h, x_res=int(res[0]),
                       y_res=int(res[1]))
            cam_bg = bg
        elif w == "l":
            v = vals(lines[k - 1])
            rgb = v[3:6] if len(v) >= 6 else [255.0] * 3
            lights.append([F(x) for x in v[:3]] + [F(x) for x in rgb] + [F(0)])
        elif w == "f":
            r, g, b, kd, ks, s, tr, ri = [F(x) for x in vals(lines[k - 1])[:8]]
            rgb = (r * F(255), g * F(255), b * F(255))
            cur = _phong((0, 0, 0), tuple(c * kd for c in rgb), tuple(c * ks for c in rgb), s, tr, ri, ks)
        elif w in ("p", "pp"):
            n = int(F(float(t[1])))
            face, normals = [], []
            for i in range(n):
                v = [float(x) for x in lines[k + i].split()]
                face.append(_v(*v[:3]))
                if w == "pp":
                    with np.errstate(all="ignore"):
                        normals.append(_norm(_v(*v[3:6])))
            k += n
            o.faces += 1
            o.polygon(face, cur, normals if w == "pp" else None)
    res = o.result()
    res.update(camera=cam, background=cam_bg, lights=np.asarray(lights, np.float32).reshape(-1, 7))
    return res
