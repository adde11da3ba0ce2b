// Scene ingestion: OFF, PLY and OBJ (+ MTL) files into triangles in emplace
// order, one material per triangle and optional per-vertex normals, ready for
// rt_scene_create_from_mesh.
//
// The triangles each parser emits, and their order, follow the reference's
// parsers (rt/parsers/off_parser.cc:14-99, rt/parsers/ply_parser.cc:14-583,
// rt/parsers/obj_parser.cc:19-857) and polygon triangulation
// (rt/parsers/polygon_to_triangles.h:50-386), because the emplace order is the
// hit id. Numbers are read as the reference's atof/atoi read them (a double
// narrowed to float; a decimal integer), and every float expression of the
// triangulation is evaluated in the reference's operand order (this file is
// built without FMA contraction).
//
// Input the reference asserts on, or misreads, is an error here with a
// message (rt_mesh_last_error) instead: a header that is not OFF/PLY, a PLY
// property type other than uchar/uint8/float/float32, an ASCII PLY face with
// colour or intensity, an index out of range, a face with a material-less
// OBJ material library. Faces of fewer than three vertices give no triangle
// (the reference loops forever or reads out of range on them). Tokens may be
// separated by any blanks.
#include <cerrno>
#include <charconv>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <new>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/raptor_rt.h"

namespace {

thread_local std::string g_error;

struct P3 {
    float x = 0.0f, y = 0.0f, z = 0.0f;
};

inline P3 sub(const P3 &a, const P3 &b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline P3 cross(const P3 &a, const P3 &b) {
    return {(a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x)};
}
inline float dot(const P3 &a, const P3 &b) { return ((a.x * b.x) + (a.y * b.y)) + (a.z * b.z); }
inline float magnitude(const P3 &a) { return std::sqrt(((a.x * a.x) + (a.y * a.y)) + (a.z * a.z)); }
// normalise<float> under EXACT_NORMALISE (common/point_t.h:553-569).
inline P3 normalised(const P3 &a) {
    const float d = std::sqrt(((a.x * a.x) + (a.y * a.y)) + (a.z * a.z));
    return {a.x / d, a.y / d, a.z / d};
}
inline bool same_point(const P3 &a, const P3 &b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
inline bool all_zero(const P3 &a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }

rt_material phong(float ka, float kd, float ks) {
    rt_material m{};
    for (int k = 0; k < 3; ++k) {
        m.ka[k] = ka;
        m.kd[k] = kd;
        m.ks[k] = ks;
    }
    m.refractive_index = 1.0f;
    m.kind = RT_MATERIAL_PHONG;
    m.chroma[0] = m.chroma[1] = 0.33f;
    return m;
}

struct Mesh {
    std::vector<float> abc;        // 9 floats per triangle, emplace order
    std::vector<int32_t> tri_mat;  // material of each triangle
    std::vector<rt_material> mats;
    std::vector<float> vn;         // 9 floats per triangle once any triangle has vertex normals
    bool any_vn = false;
    std::vector<float> vt;         // 6 floats per triangle once any triangle has texture coordinates
    bool any_vt = false;
    std::vector<rt_texture_map> maps;
    std::vector<rt_texture_image> images;
    std::vector<float> texels;
    std::map<std::string, int32_t> image_ids;  // file (+ "_inv") -> image
    int64_t vertices = 0, faces = 0, skipped_faces = 0, skipped_statements = 0;
    int32_t texture_maps = 0;
    bool has_camera = false;
    rt_camera cam{};
    float background[3] = {0.0f, 0.0f, 0.0f};
    std::vector<rt_light> lights;

    void add(const P3 &a, const P3 &b, const P3 &c, int32_t mat, const P3 *n, const float *t6 = nullptr) {
        const int64_t i = static_cast<int64_t>(tri_mat.size());
        if (t6 && !any_vt) {
            // Earlier triangles carry none: the reference marks them with
            // x = MAX_DIST (rt/triangle.h:158-167).
            any_vt = true;
            vt.assign(static_cast<size_t>(6 * i), 0.0f);
            for (int64_t k = 0; k < i; ++k) vt[6 * k] = FLT_MAX;
        }
        if (any_vt)
            for (int k = 0; k < 6; ++k) vt.push_back(t6 ? t6[k] : (k == 0 ? FLT_MAX : 0.0f));
        for (const P3 *p : {&a, &b, &c}) {
            abc.push_back(p->x);
            abc.push_back(p->y);
            abc.push_back(p->z);
        }
        tri_mat.push_back(mat);
        if (n && !any_vn) {
            // Earlier triangles carry no normals: the reference marks them
            // with x = MAX_DIST (rt/triangle.h:153-156).
            any_vn = true;
            vn.assign(static_cast<size_t>(9 * i), 0.0f);
            for (int64_t k = 0; k < i; ++k) vn[9 * k] = FLT_MAX;
        }
        if (any_vn) {
            for (int k = 0; k < 3; ++k) {
                vn.push_back(n ? n[k].x : (k == 0 ? FLT_MAX : 0.0f));
                vn.push_back(n ? n[k].y : 0.0f);
                vn.push_back(n ? n[k].z : 0.0f);
            }
        }
    }
};

struct ParseError {
    std::string what;
};
[[noreturn]] void bad(const std::string &what) { throw ParseError{what}; }

// ---- Polygon triangulation (ear clipping) ----
//
// rt/parsers/polygon_to_triangles.h:50-386. Starting at the vertex furthest
// from the centroid, the face normal is taken from that corner; the walk then
// tries the ear (prev, cur, next): a straight corner is dropped without a
// triangle, an ear facing the face normal with no polygon vertex strictly
// inside becomes a triangle and its middle vertex is dropped, any other
// corner is passed over. The walk ends when two vertices remain or when it
// has passed over as many corners in a row as vertices remain.
bool point_in_ear(const std::vector<P3> &p, const P3 &a, const P3 &b, const P3 &c, const P3 &fn) {
    const P3 ab = sub(a, b), ca = sub(c, a), bc = sub(b, c);
    auto on_inner_side = [&](const P3 &q, const P3 &corner, const P3 &edge) {
        return dot(cross(sub(q, corner), edge), fn) > 0.0f;
    };
    for (const P3 &q : p)
        if (on_inner_side(q, a, ab) && on_inner_side(q, b, bc) && on_inner_side(q, c, ca)) return true;
    return false;
}

// Emits index triples (prev, cur, next) into tris.
void triangulate(const std::vector<P3> &p, std::vector<uint32_t> *tris) {
    const uint32_t n = static_cast<uint32_t>(p.size());
    if (n < 3) return;
    P3 centre;
    for (const P3 &q : p) {
        centre.x += q.x;
        centre.y += q.y;
        centre.z += q.z;
    }
    const float fn_count = static_cast<float>(n);
    centre = {centre.x / fn_count, centre.y / fn_count, centre.z / fn_count};
    uint32_t cur = 0;
    float far = 0.0f;
    for (uint32_t i = 0; i < n; ++i) {
        const float d = magnitude(sub(centre, p[i]));
        if (d > far) {
            far = d;
            cur = i;
        }
    }
    uint32_t prev = (cur == 0) ? n - 1 : cur - 1;
    uint32_t next = (cur == n - 1) ? 0 : cur + 1;
    const P3 fn = normalised(cross(sub(p[prev], p[cur]), sub(p[cur], p[next])));
    std::vector<char> gone(n, 0);
    uint32_t left = n, passed = 0;
    for (;;) {
        const P3 &a = p[prev], &b = p[cur], &c = p[next];
        bool drop = true;
        if (!all_zero(cross(sub(b, a), sub(c, a)))) {
            const P3 tn = normalised(cross(sub(a, b), sub(b, c)));
            if (dot(tn, fn) > 0.0f && !point_in_ear(p, a, b, c, fn)) {
                tris->push_back(prev);
                tris->push_back(cur);
                tris->push_back(next);
            } else {
                drop = false;
            }
        }
        if (drop) {
            --left;
            passed = 0;
            gone[cur] = 1;
        } else {
            prev = cur;
            ++passed;
        }
        cur = next;
        do {
            next = (next + 1 == n) ? 0 : next + 1;
        } while (gone[next]);
        if (left == 2 || passed == left) break;
    }
}

// A polygon of the face's vertices (and normals and texture coordinates,
// when given) as triangles; the corners of each triangle index all three
// (rt/parsers/polygon_to_triangles.h:329-344).
void add_polygon(Mesh *m, const std::vector<P3> &face, const std::vector<P3> *normals, int32_t mat,
                 const std::vector<float> *texcoords = nullptr) {
    if (face.size() < 3) {
        ++m->skipped_faces;
        return;
    }
    std::vector<uint32_t> tris;
    triangulate(face, &tris);
    for (size_t k = 0; k < tris.size(); k += 3) {
        P3 n3[3];
        const bool with_n = normals && !normals->empty();
        if (with_n)
            for (int j = 0; j < 3; ++j) n3[j] = (*normals)[tris[k + j]];
        float t6[6];
        const bool with_t = texcoords && !texcoords->empty();
        if (with_t)
            for (int j = 0; j < 3; ++j) {
                t6[2 * j] = (*texcoords)[2 * tris[k + j]];
                t6[2 * j + 1] = (*texcoords)[2 * tris[k + j] + 1];
            }
        m->add(face[tris[k]], face[tris[k + 1]], face[tris[k + 2]], mat, with_n ? n3 : nullptr, with_t ? t6 : nullptr);
    }
}

// ---- Lexing ----

struct Cursor {
    const char *p, *end;
    bool at_end() const { return p >= end; }
};

inline bool blank(char c) { return c == ' ' || c == '\t' || c == '\r'; }

// The next line, without its terminator; false at the end of the input.
bool next_line(Cursor *c, const char **b, const char **e) {
    if (c->at_end()) return false;
    const char *s = c->p;
    const char *nl = static_cast<const char *>(std::memchr(s, '\n', static_cast<size_t>(c->end - s)));
    const char *stop = nl ? nl : c->end;
    c->p = nl ? nl + 1 : c->end;
    *b = s;
    *e = stop;
    return true;
}

// Whitespace-separated tokens of one line.
struct Tokens {
    const char *p, *e;
    bool next_view(const char **b, const char **end) {
        while (p < e && blank(*p)) ++p;
        if (p >= e) return false;
        *b = p;
        while (p < e && !blank(*p)) ++p;
        *end = p;
        return true;
    }
    bool next(std::string *t) {
        const char *b, *end;
        if (!next_view(&b, &end)) return false;
        t->assign(b, static_cast<size_t>(end - b));
        return true;
    }
};

// atof: the decimal read as a double, then narrowed to float.
float to_float(const std::string &t, const char *what) {
    char *endp = nullptr;
    const double v = std::strtod(t.c_str(), &endp);
    if (endp == t.c_str()) bad(std::string("expected a number for ") + what + ", found '" + t + "'");
    return static_cast<float>(v);
}

long to_int(const std::string &t, const char *what) {
    char *endp = nullptr;
    errno = 0;
    const long v = std::strtol(t.c_str(), &endp, 10);
    if (endp == t.c_str() || errno == ERANGE) bad(std::string("expected an integer for ") + what + ", found '" + t + "'");
    return v;
}

// The common case without a copy: std::from_chars reads the same correctly
// rounded double as strtod; signs, hex and out-of-range values go the strtod
// way.
float next_float(Tokens *tk, const char *what) {
    const char *b, *e;
    if (!tk->next_view(&b, &e)) bad(std::string("missing ") + what);
    double v;
    if (*b != '+') {
        const auto r = std::from_chars(b, e, v);
        if (r.ec == std::errc() && (r.ptr == e || (*r.ptr != 'x' && *r.ptr != 'X'))) return static_cast<float>(v);
    }
    return to_float(std::string(b, e), what);
}

long next_int(Tokens *tk, const char *what) {
    const char *b, *e;
    if (!tk->next_view(&b, &e)) bad(std::string("missing ") + what);
    long v;
    const auto r = std::from_chars(b, e, v);
    if (r.ec == std::errc() && r.ptr != b) return v;
    return to_int(std::string(b, e), what);
}

// ---- OFF (rt/parsers/off_parser.cc:14-99) ----
// Material phong_shader(ka 0, kd 170, ks 0): every triangle shares it.
void parse_off(const char *data, int64_t len, Mesh *m) {
    Cursor c{data, data + len};
    const char *b, *e;
    std::vector<std::string> head;
    std::string t;
    // "OFF", then the vertex, face and edge counts (possibly on the same line).
    while (head.size() < 4 && next_line(&c, &b, &e)) {
        Tokens tk{b, e};
        while (head.size() < 4 && tk.next(&t)) head.push_back(t);
    }
    if (head.empty() || head[0] != "OFF") bad("not an OFF file (header '" + (head.empty() ? std::string() : head[0]) + "')");
    if (head.size() < 4) bad("OFF header lacks the vertex, face and edge counts");
    const long nv = to_int(head[1], "the vertex count"), nf = to_int(head[2], "the face count");
    if (nv < 0 || nf < 0) bad("negative OFF counts");
    m->mats.push_back(phong(0.0f, 170.0f, 0.0f));
    m->mats.back().refractive_index = 0.0f;
    std::vector<P3> v;
    v.reserve(static_cast<size_t>(nv));
    for (long i = 0; i < nv; ++i) {
        if (!next_line(&c, &b, &e)) bad("OFF file ends inside the vertex list");
        Tokens tk{b, e};
        P3 q;
        q.x = next_float(&tk, "a vertex x");
        q.y = next_float(&tk, "a vertex y");
        q.z = next_float(&tk, "a vertex z");
        v.push_back(q);
    }
    m->vertices = nv;
    m->abc.reserve(static_cast<size_t>(9 * nf));
    m->tri_mat.reserve(static_cast<size_t>(nf));
    std::vector<P3> face;
    for (long i = 0; i < nf; ++i) {
        if (!next_line(&c, &b, &e)) bad("OFF file ends inside the face list");
        Tokens tk{b, e};
        if (!tk.next(&t)) bad("empty OFF face line");
        const long k = to_int(t, "a face size");
        face.clear();
        for (long j = 0; j < k; ++j) {
            if (tk.p >= tk.e) bad("OFF face line shorter than its vertex count");
            const long idx = next_int(&tk, "a vertex index");
            if (idx < 0 || idx >= nv) bad("OFF vertex index " + std::to_string(idx) + " out of range");
            face.push_back(v[static_cast<size_t>(idx)]);
        }
        ++m->faces;
        // A triangle is taken as it is; larger faces are triangulated.
        if (face.size() == 3) {
            m->add(face[0], face[1], face[2], 0, nullptr);
        } else {
            add_polygon(m, face, nullptr, 0);
        }
    }
}

// The PLY and OBJ triangle filters: a triangle whose corners coincide or
// whose edges are parallel gives no triangle.
bool ply_triangle_ok(const P3 &a0, const P3 &b0, const P3 &c0) {
    if (same_point(a0, b0) || same_point(a0, c0) || same_point(b0, c0)) return false;
    const P3 a = normalised(sub(a0, b0)), b = normalised(sub(b0, c0));
    const P3 diff = sub(a, b);
    const P3 sum{a.x + b.x, a.y + b.y, a.z + b.z};
    return !all_zero(diff) && !all_zero(sum);
}

bool obj_triangle_ok(const P3 &a0, const P3 &b0, const P3 &c0) {
    if (same_point(a0, b0) || same_point(a0, c0) || same_point(b0, c0)) return false;
    const P3 a = normalised(sub(a0, b0)), b = normalised(sub(b0, c0));
    const P3 n = cross(a, b);
    return dot(n, n) > 0.0f;
}

// ---- PLY (rt/parsers/ply_parser.cc:14-583) ----
// ASCII or binary (big-endian as in the reference; little-endian accepted
// too). Vertices: float x, y, z first, then any uchar/float properties.
// Faces: an optional uchar intensity, the uchar-counted int index list, then
// uchar properties, of which a red, green, blue run gives the face a colour
// read from the three bytes after the indices. A colour becomes
// phong_shader(0, rgb, 0, 0); faces without one share kd 200.
struct ByteReader {
    const unsigned char *p, *end;
    bool big;
    void need(size_t n) const {
        if (static_cast<size_t>(end - p) < n) bad("PLY binary data ends early");
    }
    uint32_t u32() {
        need(4);
        const uint32_t v = big ? (uint32_t(p[0]) << 24) | (uint32_t(p[1]) << 16) | (uint32_t(p[2]) << 8) | p[3]
                               : (uint32_t(p[3]) << 24) | (uint32_t(p[2]) << 16) | (uint32_t(p[1]) << 8) | p[0];
        p += 4;
        return v;
    }
    float f32() {
        const uint32_t u = u32();
        float f;
        std::memcpy(&f, &u, 4);
        return f;
    }
    unsigned char u8() {
        need(1);
        return *p++;
    }
    void skip(size_t n) {
        need(n);
        p += n;
    }
};

bool starts(const std::string &s, const char *pre) { return s.compare(0, std::strlen(pre), pre) == 0; }

int prop_size(const std::string &type) {
    if (type == "uchar" || type == "uint8") return 1;
    if (type == "float" || type == "float32") return 4;
    return 0;
}

void parse_ply(const char *data, int64_t len, Mesh *m) {
    Cursor c{data, data + len};
    const char *b, *e;
    std::vector<std::vector<std::string>> header;
    bool ended = false;
    while (next_line(&c, &b, &e)) {
        Tokens tk{b, e};
        std::vector<std::string> words;
        std::string t;
        while (tk.next(&t)) words.push_back(t);
        if (words.empty()) continue;
        if (words[0] == "comment" || words[0] == "obj_info") continue;
        if (words[0] == "end_header") {
            ended = true;
            break;
        }
        header.push_back(words);
    }
    if (header.empty() || header[0][0] != "ply") bad("not a PLY file");
    if (!ended) bad("PLY header has no end_header");
    size_t h = 1;
    if (h >= header.size() || header[h][0] != "format" || header[h].size() < 3) bad("PLY header lacks its format");
    const std::string fmt = header[h][1];
    const bool binary = fmt != "ascii";
    if (fmt != "ascii" && fmt != "binary_big_endian" && fmt != "binary_little_endian")
        bad("unknown PLY format " + fmt);
    ++h;
    // element vertex N: float x, y, z, then extra bytes.
    if (h >= header.size() || header[h].size() < 3 || header[h][0] != "element" || header[h][1] != "vertex")
        bad("PLY header must declare the vertex element first");
    const long nv = to_int(header[h][2], "the vertex count");
    ++h;
    for (const char *axis : {"x", "y", "z"}) {
        if (h >= header.size() || header[h].size() != 3 || header[h][0] != "property" ||
            (header[h][1] != "float" && header[h][1] != "float32") || header[h][2] != axis)
            bad(std::string("PLY vertex element must start with float ") + axis);
        ++h;
    }
    size_t v_extra = 0;
    for (; h < header.size() && header[h][0] == "property"; ++h) {
        const int s = header[h].size() == 3 ? prop_size(header[h][1]) : 0;
        if (!s) bad("unsupported PLY vertex property '" + header[h][1] + "'");
        v_extra += static_cast<size_t>(s);
    }
    // element face N: [uchar intensity] list uchar int vertex_indices [uchar ...]
    if (h >= header.size() || header[h].size() < 3 || header[h][0] != "element" || header[h][1] != "face")
        bad("PLY header must declare the face element after the vertices");
    const long nf = to_int(header[h][2], "the face count");
    ++h;
    size_t pre = 0;
    if (h < header.size() && header[h].size() == 3 && header[h][0] == "property" && prop_size(header[h][1]) == 1 &&
        header[h][2] == "intensity") {
        pre = 1;
        ++h;
    }
    if (h >= header.size() || header[h].size() != 5 || header[h][0] != "property" || header[h][1] != "list" ||
        prop_size(header[h][2]) != 1 || (header[h][3] != "int" && header[h][3] != "int32") ||
        header[h][4] != "vertex_indices")
        bad("PLY faces must be a 'list uchar int vertex_indices'");
    ++h;
    std::vector<std::string> face_props;
    for (; h < header.size() && header[h][0] == "property"; ++h) {
        if (header[h].size() != 3 || prop_size(header[h][1]) != 1)
            bad("unsupported PLY face property '" + header[h][1] + "'");
        face_props.push_back(header[h][2]);
    }
    if (h != header.size()) bad("unsupported PLY element '" + header[h][0] + " " + (header[h].size() > 1 ? header[h][1] : "") + "'");
    bool colour = false;
    size_t post = face_props.size();
    for (size_t k = 0; k + 2 < face_props.size(); ++k)
        if (face_props[k] == "red" && face_props[k + 1] == "green" && face_props[k + 2] == "blue") {
            colour = true;
            post -= 3;
            break;
        }
    if (nv < 0 || nf < 0) bad("negative PLY counts");

    m->mats.push_back(phong(0.0f, 200.0f, 0.0f));
    // Face colours keyed as the reference keys them: the three bytes read as
    // signed chars, c0 + (c1 << 8) + (c2 << 16); the default material has key
    // 0xdefa, so a colour with that key shades with the default.
    std::map<int, int32_t> by_key;
    by_key[0xdefa] = 0;
    auto colour_material = [&](const unsigned char rgb[3]) -> int32_t {
        const int key = static_cast<int>(static_cast<signed char>(rgb[0])) +
                        (static_cast<int>(static_cast<signed char>(rgb[1])) * 256) +
                        (static_cast<int>(static_cast<signed char>(rgb[2])) * 65536);
        auto it = by_key.find(key);
        if (it != by_key.end()) return it->second;
        rt_material q = phong(0.0f, 0.0f, 0.0f);
        for (int k = 0; k < 3; ++k) q.kd[k] = static_cast<float>(rgb[k]);
        m->mats.push_back(q);
        const int32_t id = static_cast<int32_t>(m->mats.size() - 1);
        by_key[key] = id;
        return id;
    };

    std::vector<P3> v;
    v.reserve(static_cast<size_t>(nv));
    m->abc.reserve(static_cast<size_t>(9 * nf));
    m->tri_mat.reserve(static_cast<size_t>(nf));
    std::vector<P3> face;
    auto emit = [&](int32_t mat) {
        ++m->faces;
        if (face.size() == 3) {
            if (ply_triangle_ok(face[0], face[1], face[2])) {
                m->add(face[0], face[1], face[2], mat, nullptr);
            } else {
                ++m->skipped_faces;
            }
        } else {
            add_polygon(m, face, nullptr, mat);
        }
    };
    if (binary) {
        ByteReader r{reinterpret_cast<const unsigned char *>(c.p), reinterpret_cast<const unsigned char *>(c.end),
                     fmt == "binary_big_endian"};
        for (long i = 0; i < nv; ++i) {
            P3 q;
            q.x = r.f32();
            q.y = r.f32();
            q.z = r.f32();
            r.skip(v_extra);
            v.push_back(q);
        }
        for (long i = 0; i < nf; ++i) {
            r.skip(pre);
            const unsigned k = r.u8();
            face.clear();
            for (unsigned j = 0; j < k; ++j) {
                const int32_t idx = static_cast<int32_t>(r.u32());
                if (idx < 0 || idx >= nv) bad("PLY vertex index " + std::to_string(idx) + " out of range");
                face.push_back(v[static_cast<size_t>(idx)]);
            }
            int32_t mat = 0;
            if (colour) {
                unsigned char rgb[3];
                for (auto &x : rgb) x = r.u8();
                mat = colour_material(rgb);
            }
            r.skip(post);
            emit(mat);
        }
    } else {
        if (colour || pre) bad("ASCII PLY faces with colour or intensity are not supported");
        std::string t;
        for (long i = 0; i < nv; ++i) {
            if (!next_line(&c, &b, &e)) bad("PLY file ends inside the vertex list");
            Tokens tk{b, e};
            P3 q;
            q.x = next_float(&tk, "a vertex x");
            q.y = next_float(&tk, "a vertex y");
            q.z = next_float(&tk, "a vertex z");
            v.push_back(q);
        }
        for (long i = 0; i < nf; ++i) {
            if (!next_line(&c, &b, &e)) bad("PLY file ends inside the face list");
            Tokens tk{b, e};
            if (!tk.next(&t)) bad("empty PLY face line");
            const long k = to_int(t, "a face size");
            face.clear();
            for (long j = 0; j < k; ++j) {
                if (!tk.next(&t)) bad("PLY face line shorter than its vertex count");
                const long idx = to_int(t, "a vertex index");
                if (idx < 0 || idx >= nv) bad("PLY vertex index " + std::to_string(idx) + " out of range");
                face.push_back(v[static_cast<size_t>(idx)]);
            }
            emit(0);
        }
    }
    m->vertices = nv;
}

// ---- OBJ and MTL (rt/parsers/obj_parser.cc:19-857) ----

struct MtlState {
    P3 ka, kd, ks;
    float ns = 0.0f, ri = 1.0f, rf = 0.0f, trans = 0.0f;
    int illum = 0;  // not reset by newmtl (obj_parser.cc:380-396)
    // Texture map files for t_ka, t_kd, t_ks, t_ns, t_refl, t_d
    // (obj_parser.cc:263-268; the MTL statements never set t_ns).
    std::string maps[6];
};

// A TGA file as read_tga reads it through libtga with TGA_IMAGE_DATA |
// TGA_RGB (rt/materials/texture_mapper.h read_tga): true-colour or grey,
// uncompressed or run-length encoded, 8, 24 or 32 bits; rows in file order;
// blue and red swapped to RGB(A); one float 0-255 per channel.
void decode_tga(const std::string &f, int *w, int *h, int *cpp, std::vector<float> *out) {
    const unsigned char *d = reinterpret_cast<const unsigned char *>(f.data());
    const size_t n = f.size();
    if (n < 18) bad("TGA shorter than its header");
    const int id_len = d[0], cmap_type = d[1], type = d[2];
    const int cmap_len = d[5] | (d[6] << 8), cmap_depth = d[7];
    *w = d[12] | (d[13] << 8);
    *h = d[14] | (d[15] << 8);
    const int depth = d[16];
    if (type != 2 && type != 3 && type != 10 && type != 11)
        bad("TGA image type " + std::to_string(type) + " (only true-colour and grey, raw or run-length)");
    if (depth != 8 && depth != 24 && depth != 32) bad("TGA depth " + std::to_string(depth) + " (8, 24 or 32)");
    if (*w <= 0 || *h <= 0) bad("TGA without pixels");
    *cpp = depth >> 3;
    size_t at = 18 + static_cast<size_t>(id_len);
    if (cmap_type == 1) at += static_cast<size_t>(cmap_len) * ((cmap_depth + 7) / 8);
    const size_t npix = static_cast<size_t>(*w) * static_cast<size_t>(*h);
    const int bpp = *cpp;
    std::vector<unsigned char> raw(npix * bpp);
    if (type == 2 || type == 3) {
        if (at + raw.size() > n) bad("TGA pixel data cut short");
        std::memcpy(raw.data(), d + at, raw.size());
    } else {
        size_t p = 0;
        while (p < npix) {
            if (at >= n) bad("TGA run-length data cut short");
            const int hdr = d[at++];
            const size_t count = static_cast<size_t>(hdr & 0x7f) + 1;
            if (p + count > npix) bad("TGA run-length packet past the image");
            if (hdr & 0x80) {
                if (at + bpp > n) bad("TGA run-length data cut short");
                for (size_t k = 0; k < count; ++k) std::memcpy(&raw[(p + k) * bpp], d + at, bpp);
                at += bpp;
            } else {
                if (at + count * bpp > n) bad("TGA run-length data cut short");
                std::memcpy(&raw[p * bpp], d + at, count * bpp);
                at += count * bpp;
            }
            p += count;
        }
    }
    out->resize(raw.size());
    for (size_t i = 0; i < npix; ++i) {
        const unsigned char *px = &raw[i * bpp];
        float *o = out->data() + i * bpp;
        if (bpp >= 3) {
            o[0] = px[2];
            o[1] = px[1];
            o[2] = px[0];
            if (bpp == 4) o[3] = px[3];
        } else {
            o[0] = px[0];
        }
    }
}

// An MTL colour: one value (grey) or three, scaled to 0..255.
P3 mtl_colour(Tokens *tk) {
    std::vector<float> v;
    std::string t;
    while (v.size() < 3 && tk->next(&t)) v.push_back(to_float(t, "a colour") * 255.0f);
    if (v.empty()) bad("MTL colour without a value");
    if (v.size() < 3) return {v[0], v[0], v[0]};
    return {v[0], v[1], v[2]};
}

std::string read_file(const std::string &path);

// load_image (obj_parser.cc:218-242): the image of a map file, read once per
// file (and once more inverted, image = 1 - texel as picture_functions'
// negative computes it, for map_d), under a planar_mapper centred at 0 with
// normal (1, 0, 0), size 0 and mirror wrapping: it samples only through the
// triangles' texture coordinates. Returns 1 + the map's index, or 0 for a
// file this library cannot decode (only TGA), which m->texture_maps counts.
int32_t mtl_map(Mesh *m, const std::string &file, bool invert) {
    std::string ext = file.size() >= 4 ? file.substr(file.size() - 4) : std::string();
    for (auto &ch : ext) ch = static_cast<char>(std::tolower(static_cast<unsigned char>(ch)));
    if (ext != ".tga") {
        ++m->texture_maps;
        return 0;
    }
    const std::string key = file + (invert ? "_inv" : "");
    auto it = m->image_ids.find(key);
    if (it == m->image_ids.end()) {
        int w, h, cpp;
        std::vector<float> px;
        decode_tga(read_file(file), &w, &h, &cpp, &px);
        if (invert)
            for (auto &v : px) v = 1.0f - v;
        rt_texture_image im{};
        im.width = w;
        im.height = h;
        im.cpp = cpp;
        im.offset = static_cast<int64_t>(m->texels.size());
        m->texels.insert(m->texels.end(), px.begin(), px.end());
        m->images.push_back(im);
        it = m->image_ids.emplace(key, static_cast<int32_t>(m->images.size() - 1)).first;
    }
    rt_texture_map t{};
    t.kind = RT_MAP_PLANAR;
    t.image = it->second;
    t.n[0] = 1.0f;  // u = (n.y + n.z, 0, n.x), v = (0, n.x + n.z, n.y)
    t.u[2] = 1.0f;
    t.v[1] = 1.0f;
    t.u_wrap = t.v_wrap = RT_WRAP_MIRROR;
    t.opacity = 1.0f;
    t.falloff = RT_FALLOFF_NONE;
    m->maps.push_back(t);
    return static_cast<int32_t>(m->maps.size());
}

// The material a finished newmtl block describes (obj_parser.cc:290-378):
// the illumination model zeroes or fixes terms and texture maps,
// transmittance is clamped to [0, 1], the ambient colour is halved; an
// untextured coloured_mapper_shader shades as phong_shader.
rt_material finish_mtl(MtlState s, Mesh *mesh) {
    switch (s.illum) {
        case 0:
            s.maps[0] = s.maps[1];
            for (int k = 1; k < 5; ++k) s.maps[k].clear();
            s.ka = s.kd;
            s.kd = P3{};
            s.ks = P3{};
            s.ns = 0.0f;
            s.ri = 1.0f;
            s.rf = 0.0f;
            break;
        case 1:
            s.maps[2].clear();
            s.maps[3].clear();
            s.maps[4].clear();
            s.ks = P3{};
            s.ns = 0.0f;
            s.ri = 1.0f;
            s.rf = 0.0f;
            break;
        case 2:
            s.maps[4].clear();
            s.ri = 1.0f;
            s.rf = 0.0f;
            break;
        case 3:
        case 5:
        case 7:
            s.ri = 1.0f;
            s.rf = 0.75f;
            break;
        case 4:
            s.rf = 0.0f;
            break;
        default:
            break;
    }
    s.trans = std::max(0.0f, std::min(1.0f, s.trans));
    rt_material m{};
    const float ambient = 0.5f;
    m.ka[0] = s.ka.x * ambient;
    m.ka[1] = s.ka.y * ambient;
    m.ka[2] = s.ka.z * ambient;
    m.kd[0] = s.kd.x;
    m.kd[1] = s.kd.y;
    m.kd[2] = s.kd.z;
    m.ks[0] = s.ks.x;
    m.ks[1] = s.ks.y;
    m.ks[2] = s.ks.z;
    m.shininess = s.ns;
    m.transmittance = s.trans;
    m.refractive_index = s.ri;
    m.reflectance = s.rf;
    m.kind = RT_MATERIAL_COLOURED;
    m.chroma[0] = m.chroma[1] = 0.33f;
    for (int k = 0; k < 6; ++k) m.tex[k] = s.maps[k].empty() ? 0 : mtl_map(mesh, s.maps[k], k == 5);
    return m;
}

std::string read_file(const std::string &path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) bad("cannot open " + path);
    std::ostringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

// Parses one MTL file into named materials; returns the id of the last one
// defined (which the reference makes current for the faces that follow), or
// -1 when the file defines none.
int32_t parse_mtl(const std::string &text, const std::string &dir, Mesh *m, std::map<std::string, int32_t> *named) {
    Cursor c{text.data(), text.data() + text.size()};
    const char *b, *e;
    std::string name, t;
    MtlState s;
    int illum = 0;
    int32_t last = -1;
    auto close = [&]() {
        if (name.empty()) return;
        s.illum = illum;
        m->mats.push_back(finish_mtl(s, m));
        last = static_cast<int32_t>(m->mats.size() - 1);
        (*named)[name] = last;
    };
    while (next_line(&c, &b, &e)) {
        Tokens tk{b, e};
        if (!tk.next(&t) || t[0] == '#') continue;
        if (starts(t, "newmtl")) {
            close();
            if (!tk.next(&name)) bad("newmtl without a name");
            s = MtlState{};
        } else if (starts(t, "Ka")) {
            s.ka = mtl_colour(&tk);
        } else if (starts(t, "Kd")) {
            s.kd = mtl_colour(&tk);
        } else if (starts(t, "Ks")) {
            s.ks = mtl_colour(&tk);
        } else if (starts(t, "Tf") || starts(t, "Ke")) {
            // transmitted and emitted colours are not shaded
        } else if (starts(t, "Ns")) {
            s.ns = next_float(&tk, "Ns");
        } else if (starts(t, "Ni")) {
            s.ri = next_float(&tk, "Ni");
        } else if (t[0] == 'd') {
            // Every statement starting with 'd' is read as the dissolve.
            std::string v;
            s.trans = 1.0f - (tk.next(&v) ? static_cast<float>(std::atof(v.c_str())) : 0.0f);
        } else if (starts(t, "Tr")) {
            s.trans = next_float(&tk, "Tr");
        } else if (starts(t, "illum")) {
            std::string v;
            illum = tk.next(&v) ? std::atoi(v.c_str()) : 0;
        } else if (starts(t, "map_Ka") || starts(t, "map_Kd") || starts(t, "map_Ks") || starts(t, "map_refl") ||
                   starts(t, "map_d")) {
            // obj_parser.cc:507-561: the file, relative to the MTL's directory.
            std::string file;
            if (!tk.next(&file)) bad(t + " without a file");
            const int slot = starts(t, "map_Ka") ? 0 : starts(t, "map_Kd") ? 1 : starts(t, "map_Ks") ? 2
                           : starts(t, "map_refl") ? 4 : 5;
            s.maps[slot] = dir + file;
        } else if (starts(t, "map_bump") || starts(t, "bump")) {
            // bump maps are ignored by the reference too
        } else {
            ++m->skipped_statements;
        }
    }
    close();
    return last;
}

// Resolves an OBJ index (1-based, or negative from the end).
size_t obj_index(long i, size_t n, const char *what) {
    const long r = (i < 0) ? static_cast<long>(n) + i : i - 1;
    if (r < 0 || static_cast<size_t>(r) >= n) bad(std::string("OBJ ") + what + " index " + std::to_string(i) + " out of range");
    return static_cast<size_t>(r);
}

void parse_obj(const char *data, int64_t len, const std::string &dir, Mesh *m) {
    Cursor c{data, data + len};
    const char *b, *e;
    std::vector<P3> v, vn;
    std::vector<float> vts;  // u, v of every vt statement
    std::map<std::string, int32_t> named;
    m->mats.push_back(phong(0.0f, 200.0f, 0.0f));
    int32_t cur = 0;
    std::string t;
    std::vector<P3> face, face_n;
    std::vector<float> face_t;
    while (next_line(&c, &b, &e)) {
        Tokens tk{b, e};
        if (!tk.next(&t) || t[0] == '#') continue;
        if (starts(t, "mtllib")) {
            // The current material becomes the last one the last library
            // defines (none: faces need a usemtl before they shade).
            std::string file;
            while (tk.next(&file)) cur = parse_mtl(read_file(dir + file), dir, m, &named);
        } else if (starts(t, "usemtl")) {
            std::string name;
            if (!tk.next(&name)) bad("usemtl without a name");
            auto it = named.find(name);
            if (it == named.end()) {
                m->mats.push_back(phong(0.0f, 200.0f, 0.0f));
                it = named.emplace(name, static_cast<int32_t>(m->mats.size() - 1)).first;
            }
            cur = it->second;
        } else if (starts(t, "vn")) {
            P3 q;
            q.x = next_float(&tk, "vn x");
            q.y = next_float(&tk, "vn y");
            q.z = next_float(&tk, "vn z");
            vn.push_back(normalised(q));
        } else if (starts(t, "vt")) {
            // parse_vt_statement (obj_parser.cc:732-743): u and v.
            vts.push_back(next_float(&tk, "vt u"));
            vts.push_back(next_float(&tk, "vt v"));
        } else if (t[0] == 'f') {
            face.clear();
            face_n.clear();
            face_t.clear();
            std::string corner;
            while (tk.next(&corner)) {
                // v, v/vt, v/vt/vn or v//vn
                const size_t s1 = corner.find('/');
                face.push_back(v[obj_index(to_int(corner.substr(0, s1), "a vertex index"), v.size(), "vertex")]);
                if (s1 == std::string::npos) continue;
                const size_t s2 = corner.find('/', s1 + 1);
                const std::string vt_s = corner.substr(s1 + 1, s2 == std::string::npos ? std::string::npos : s2 - s1 - 1);
                if (!vt_s.empty()) {
                    const size_t k = obj_index(to_int(vt_s, "a texture index"), vts.size() / 2, "texture");
                    face_t.push_back(vts[2 * k]);
                    face_t.push_back(vts[2 * k + 1]);
                }
                if (s2 != std::string::npos)
                    face_n.push_back(vn[obj_index(to_int(corner.substr(s2 + 1), "a normal index"), vn.size(), "normal")]);
            }
            if (!face_n.empty() && face_n.size() != face.size()) bad("OBJ face gives normals for some corners only");
            if (!face_t.empty() && face_t.size() != 2 * face.size())
                bad("OBJ face gives texture coordinates for some corners only");
            ++m->faces;
            if (cur < 0) bad("OBJ face without a material (its material library defines none)");
            if (face.size() == 3) {
                if (obj_triangle_ok(face[0], face[1], face[2])) {
                    m->add(face[0], face[1], face[2], cur, face_n.empty() ? nullptr : face_n.data(),
                           face_t.empty() ? nullptr : face_t.data());
                } else {
                    ++m->skipped_faces;
                }
            } else {
                add_polygon(m, face, &face_n, cur, &face_t);
            }
        } else if (t[0] == 'v') {
            P3 q;
            q.x = next_float(&tk, "v x");
            q.y = next_float(&tk, "v y");
            q.z = next_float(&tk, "v z");
            v.push_back(q);
        } else if (t[0] == 'g' || t[0] == 'o' || t[0] == 's') {
            // groups, objects and smoothing groups carry nothing to shade
        } else {
            ++m->skipped_statements;
        }
    }
    m->vertices = static_cast<int64_t>(v.size());
}

// ---- NFF (rt/parsers/nff_parser.cc:16-224) ----
// Line-based: b (background, 0..1), v (a six-line viewpoint block: from,
// at, up, angle, hither, resolution), l (point light, white unless a colour
// follows), f (fill material), p n / pp n (a polygon of n vertex lines,
// with normals for pp). Other statements (spheres, cones, comments) are
// skipped, as the reference skips them.
void parse_nff(const char *data, int64_t len, Mesh *m) {
    Cursor c{data, data + len};
    const char *b, *e;
    std::string t;
    int32_t cur = -1;
    float bg[3] = {0.0f, 0.0f, 0.0f};
    auto line_tokens = [&](const char *what) {
        if (!next_line(&c, &b, &e)) bad(std::string("NFF file ends inside ") + what);
        return Tokens{b, e};
    };
    auto vec3 = [&](Tokens *tk, const char *what) {
        P3 q;
        q.x = next_float(tk, what);
        q.y = next_float(tk, what);
        q.z = next_float(tk, what);
        return q;
    };
    std::vector<P3> face, normals;
    while (next_line(&c, &b, &e)) {
        Tokens tk{b, e};
        if (!tk.next(&t)) continue;
        if (t == "b") {
            const P3 q = vec3(&tk, "the background");
            bg[0] = q.x * 255.0f;
            bg[1] = q.y * 255.0f;
            bg[2] = q.z * 255.0f;
        } else if (t[0] == 'v') {
            Tokens f = line_tokens("the viewpoint");
            f.next(&t);
            const P3 from = vec3(&f, "from");
            Tokens a = line_tokens("the viewpoint");
            a.next(&t);
            const P3 to = vec3(&a, "at");
            P3 at_vec = sub(to, from);
            Tokens u = line_tokens("the viewpoint");
            u.next(&t);
            P3 up = vec3(&u, "up");
            Tokens an = line_tokens("the viewpoint");
            an.next(&t);
            const float angle = next_float(&an, "angle") * (3.1415926535897932384626433f / 360.0f);
            // Unqualified sin / cos of a float in the reference: the double ones.
            const float dist_ratio = static_cast<float>(std::sin(static_cast<double>(angle)) /
                                                        std::cos(static_cast<double>(angle)));
            Tokens hi = line_tokens("the viewpoint");
            hi.next(&t);
            float h = next_float(&hi, "hither");
            if (h == 0.0f) h = 1.0f;
            const float hw = dist_ratio * h;
            Tokens r = line_tokens("the viewpoint");
            r.next(&t);
            const float xr = next_float(&r, "resolution"), yr = next_float(&r, "resolution");
            P3 x_vec = cross(up, at_vec);
            up = cross(at_vec, x_vec);
            x_vec = normalised(x_vec);
            up = normalised(up);
            at_vec = normalised(at_vec);
            rt_camera cam{};
            const P3 *axes[4] = {&from, &x_vec, &up, &at_vec};
            float *dst[4] = {cam.pos, cam.x_axis, cam.y_axis, cam.z_axis};
            for (int k = 0; k < 4; ++k) {
                dst[k][0] = axes[k]->x;
                dst[k][1] = axes[k]->y;
                dst[k][2] = axes[k]->z;
            }
            cam.w = hw;
            cam.h = hw;
            cam.t = h;
            cam.x_res = static_cast<int32_t>(static_cast<unsigned int>(xr));
            cam.y_res = static_cast<int32_t>(static_cast<unsigned int>(yr));
            m->cam = cam;
            m->has_camera = true;
            for (int k = 0; k < 3; ++k) m->background[k] = bg[k];  // the background so far
        } else if (t == "l") {
            rt_light l{};
            const P3 q = vec3(&tk, "a light");
            l.pos[0] = q.x;
            l.pos[1] = q.y;
            l.pos[2] = q.z;
            l.rgb[0] = l.rgb[1] = l.rgb[2] = 255.0f;
            std::string v;
            if (tk.next(&v)) {
                l.rgb[0] = to_float(v, "a light colour");
                l.rgb[1] = next_float(&tk, "a light colour");
                l.rgb[2] = next_float(&tk, "a light colour");
            }
            l.d = 0.0f;
            m->lights.push_back(l);
        } else if (t == "f") {
            const P3 rgb0 = vec3(&tk, "a fill colour");
            const float kd = next_float(&tk, "kd"), ks = next_float(&tk, "ks"), s = next_float(&tk, "shine");
            const float tr = next_float(&tk, "transmittance"), ri = next_float(&tk, "refractive index");
            const float rgb[3] = {rgb0.x * 255.0f, rgb0.y * 255.0f, rgb0.z * 255.0f};
            // phong_shader(rgb, kd, ks, s, tran, ri, rf = ks) (rt/materials/phong_shader.h:17-19)
            rt_material q = phong(0.0f, 0.0f, 0.0f);
            for (int k = 0; k < 3; ++k) {
                q.kd[k] = rgb[k] * kd;
                q.ks[k] = rgb[k] * ks;
            }
            q.shininess = s;
            q.transmittance = tr;
            q.refractive_index = ri;
            q.reflectance = ks;
            m->mats.push_back(q);
            cur = static_cast<int32_t>(m->mats.size() - 1);
        } else if (t == "p" || t == "pp") {
            const bool patch = t == "pp";
            std::string v;
            if (!tk.next(&v)) bad("NFF polygon without a vertex count");
            const long n = static_cast<long>(to_float(v, "a vertex count"));
            if (cur < 0) bad("NFF polygon before any fill material");
            face.clear();
            normals.clear();
            for (long i = 0; i < n; ++i) {
                Tokens pv = line_tokens("a polygon");
                face.push_back(vec3(&pv, "a polygon vertex"));
                if (patch) normals.push_back(normalised(vec3(&pv, "a patch normal")));
            }
            ++m->faces;
            m->vertices += n;
            add_polygon(m, face, patch ? &normals : nullptr, cur);
        } else {
            ++m->skipped_statements;
        }
    }
}

int32_t format_of(const std::string &path) {
    const size_t dot = path.find_last_of('.');
    if (dot == std::string::npos) return -1;
    std::string ext = path.substr(dot + 1);
    for (char &ch : ext) ch = static_cast<char>(std::tolower(static_cast<unsigned char>(ch)));
    if (ext == "off") return RT_FORMAT_OFF;
    if (ext == "ply") return RT_FORMAT_PLY;
    if (ext == "obj") return RT_FORMAT_OBJ;
    if (ext == "nff") return RT_FORMAT_NFF;
    return -1;
}

}  // namespace

struct rt_mesh {
    Mesh m;
};

extern "C" {

const char *rt_mesh_last_error(void) { return g_error.c_str(); }

int32_t rt_mesh_parse(const char *data, int64_t len, int32_t format, const char *dir, rt_mesh **out) {
    if (!out) return RT_ERR_INVALID;
    *out = nullptr;
    g_error.clear();
    if (len < 0 || (len > 0 && !data)) {
        g_error = "rt_mesh_parse: bad buffer";
        return RT_ERR_INVALID;
    }
    std::unique_ptr<rt_mesh> mesh(new (std::nothrow) rt_mesh());
    if (!mesh) {
        g_error = "rt_mesh_parse: host allocation";
        return RT_ERR_OOM;
    }
    try {
        switch (format) {
            case RT_FORMAT_OFF:
                parse_off(data, len, &mesh->m);
                break;
            case RT_FORMAT_PLY:
                parse_ply(data, len, &mesh->m);
                break;
            case RT_FORMAT_NFF:
                parse_nff(data, len, &mesh->m);
                break;
            case RT_FORMAT_OBJ: {
                std::string d = dir ? dir : "";
                if (!d.empty() && d.back() != '/') d += '/';
                parse_obj(data, len, d, &mesh->m);
                break;
            }
            default:
                g_error = "rt_mesh_parse: unknown format " + std::to_string(format);
                return RT_ERR_UNSUPPORTED;
        }
    } catch (const ParseError &pe) {
        g_error = "rt_mesh_parse: " + pe.what;
        return RT_ERR_INVALID;
    } catch (const std::bad_alloc &) {
        g_error = "rt_mesh_parse: host allocation";
        return RT_ERR_OOM;
    } catch (...) {
        g_error = "rt_mesh_parse: unexpected failure";
        return RT_ERR_INVALID;
    }
    *out = mesh.release();
    return RT_OK;
}

int32_t rt_mesh_load(const char *path, int32_t format, rt_mesh **out) {
    if (!out || !path) return RT_ERR_INVALID;
    *out = nullptr;
    g_error.clear();
    const std::string p(path);
    if (format < 0) format = format_of(p);
    if (format < 0) {
        g_error = "rt_mesh_load: no format given and none known for " + p;
        return RT_ERR_UNSUPPORTED;
    }
    std::string text;
    try {
        text = read_file(p);
    } catch (const ParseError &pe) {
        g_error = "rt_mesh_load: " + pe.what;
        return RT_ERR_INVALID;
    } catch (const std::bad_alloc &) {
        g_error = "rt_mesh_load: host allocation";
        return RT_ERR_OOM;
    }
    const size_t slash = p.find_last_of('/');
    const std::string dir = (slash == std::string::npos) ? std::string() : p.substr(0, slash + 1);
    return rt_mesh_parse(text.data(), static_cast<int64_t>(text.size()), format, dir.c_str(), out);
}

int32_t rt_mesh_info_get(const rt_mesh *mesh, rt_mesh_info *info) {
    if (!mesh || !info) return RT_ERR_INVALID;
    const Mesh &m = mesh->m;
    info->triangles = static_cast<int64_t>(m.tri_mat.size());
    info->materials = static_cast<int32_t>(m.mats.size());
    info->has_vertex_normals = m.any_vn ? 1 : 0;
    info->vertices = m.vertices;
    info->faces = m.faces;
    info->skipped_faces = m.skipped_faces;
    info->skipped_statements = m.skipped_statements;
    info->texture_maps = m.texture_maps;
    info->lights = static_cast<int32_t>(m.lights.size());
    info->has_camera = m.has_camera ? 1 : 0;
    info->has_texture_coords = m.any_vt ? 1 : 0;
    info->maps = static_cast<int32_t>(m.maps.size());
    info->images = static_cast<int32_t>(m.images.size());
    info->texels = static_cast<int64_t>(m.texels.size());
    return RT_OK;
}

int32_t rt_mesh_copy_textures(const rt_mesh *mesh, float *vt, rt_texture_map *maps, rt_texture_image *images,
                              float *texels) {
    if (!mesh) return RT_ERR_INVALID;
    const Mesh &m = mesh->m;
    if (vt && m.any_vt) std::memcpy(vt, m.vt.data(), m.vt.size() * sizeof(float));
    if (maps && !m.maps.empty()) std::memcpy(maps, m.maps.data(), m.maps.size() * sizeof(rt_texture_map));
    if (images && !m.images.empty())
        std::memcpy(images, m.images.data(), m.images.size() * sizeof(rt_texture_image));
    if (texels && !m.texels.empty()) std::memcpy(texels, m.texels.data(), m.texels.size() * sizeof(float));
    return RT_OK;
}

int32_t rt_texture_load_tga(const char *path, float *texels, int64_t capacity, int32_t *dims) {
    if (!path || !dims) return RT_ERR_INVALID;
    g_error.clear();
    try {
        int w, h, cpp;
        std::vector<float> px;
        decode_tga(read_file(path), &w, &h, &cpp, &px);
        dims[0] = w;
        dims[1] = h;
        dims[2] = cpp;
        if (!texels) return RT_OK;
        if (capacity < static_cast<int64_t>(px.size())) {
            g_error = "rt_texture_load_tga: " + std::to_string(px.size()) + " texels do not fit in " +
                      std::to_string(capacity);
            return RT_ERR_INVALID;
        }
        std::memcpy(texels, px.data(), px.size() * sizeof(float));
    } catch (const ParseError &pe) {
        g_error = "rt_texture_load_tga: " + pe.what;
        return RT_ERR_INVALID;
    } catch (const std::bad_alloc &) {
        g_error = "rt_texture_load_tga: host allocation";
        return RT_ERR_OOM;
    }
    return RT_OK;
}

int32_t rt_mesh_camera(const rt_mesh *mesh, rt_camera *cam, float *background) {
    if (!mesh) return RT_ERR_INVALID;
    if (!mesh->m.has_camera) {
        g_error = "rt_mesh_camera: the file defines no camera";
        return RT_ERR_UNSUPPORTED;
    }
    if (cam) *cam = mesh->m.cam;
    if (background)
        for (int k = 0; k < 3; ++k) background[k] = mesh->m.background[k];
    return RT_OK;
}

int32_t rt_mesh_lights(const rt_mesh *mesh, rt_light *lights) {
    if (!mesh) return RT_ERR_INVALID;
    if (lights && !mesh->m.lights.empty())
        std::memcpy(lights, mesh->m.lights.data(), mesh->m.lights.size() * sizeof(rt_light));
    return RT_OK;
}

int32_t rt_mesh_copy(const rt_mesh *mesh, float *abc, int32_t *tri_material, rt_material *materials,
                     float *vertex_normals) {
    if (!mesh) return RT_ERR_INVALID;
    const Mesh &m = mesh->m;
    if (abc && !m.abc.empty()) std::memcpy(abc, m.abc.data(), m.abc.size() * sizeof(float));
    if (tri_material && !m.tri_mat.empty())
        std::memcpy(tri_material, m.tri_mat.data(), m.tri_mat.size() * sizeof(int32_t));
    if (materials && !m.mats.empty()) std::memcpy(materials, m.mats.data(), m.mats.size() * sizeof(rt_material));
    if (vertex_normals && m.any_vn) std::memcpy(vertex_normals, m.vn.data(), m.vn.size() * sizeof(float));
    return RT_OK;
}

void rt_mesh_destroy(rt_mesh *mesh) { delete mesh; }

int32_t rt_scene_create_from_mesh(rt_ctx *ctx, const rt_mesh *mesh, rt_scene **out) {
    if (!ctx || !mesh || !out) return RT_ERR_INVALID;
    *out = nullptr;
    const Mesh &m = mesh->m;
    const int64_t n = static_cast<int64_t>(m.tri_mat.size());
    // Albedo kd / 255 keeps the Lambert-only frames meaningful; the flags mark
    // triangles of transparent materials (material(tran > 0)).
    std::vector<float> albedo(static_cast<size_t>(3 * std::max<int64_t>(n, 1)));
    std::vector<uint8_t> flags(static_cast<size_t>(std::max<int64_t>(n, 1)), 0);
    for (int64_t i = 0; i < n; ++i) {
        const rt_material &q = m.mats[static_cast<size_t>(m.tri_mat[static_cast<size_t>(i)])];
        for (int k = 0; k < 3; ++k) albedo[static_cast<size_t>(3 * i + k)] = q.kd[k] / 255.0f;
        flags[static_cast<size_t>(i)] = q.transmittance > 0.0f ? 2 : 0;
    }
    rt_scene *sc = nullptr;
    int32_t rc = rt_scene_create(ctx, m.abc.data(), albedo.data(), flags.data(), n, &sc);
    if (rc != RT_OK) return rc;
    if (!m.maps.empty())
        rc = rt_scene_set_textures(sc, m.maps.data(), static_cast<int32_t>(m.maps.size()), m.images.data(),
                                   static_cast<int32_t>(m.images.size()), m.texels.data(),
                                   static_cast<int64_t>(m.texels.size()), nullptr, 0);
    if (rc == RT_OK)
        rc = rt_scene_set_materials(sc, m.mats.data(), static_cast<int32_t>(m.mats.size()), m.tri_mat.data());
    if (rc == RT_OK && m.any_vn) rc = rt_scene_set_vertex_normals(sc, m.vn.data());
    if (rc == RT_OK && m.any_vt) rc = rt_scene_set_texture_coords(sc, m.vt.data());
    if (rc != RT_OK) {
        rt_scene_destroy(sc);
        return rc;
    }
    *out = sc;
    return RT_OK;
}

}  // extern "C"
