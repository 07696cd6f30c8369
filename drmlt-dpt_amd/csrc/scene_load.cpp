// Scene files on the host: the reference's Mitsuba-style XML
// (ParseScene / ParseSceneNode, src/parsescene.cpp:136-613) with its OBJ
// (src/parseobj.cpp) and Mitsuba .serialized (src/loadserialized.cpp) mesh
// loaders, reduced to what the device path renders and flattened into the
// arrays of drmlt_scene_desc: triangle meshes with diffuse, phong,
// roughdielectric (optionally twosided) BSDFs of constant parameters, area
// emitters on shapes, a perspective sensor whose toWorld is a look-at, and
// the <dpt> options. Anything else the reference parses (bitmap textures,
// point and environment emitters, moving shapes) is rejected with an error
// naming it, since the device path has no counterpart for it.
#include "../../include/drmlt_amd.h"

#include <zlib.h>

#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

int drmlt_set_error(int code, const char *msg);  // drmlt_capi.hip

struct drmlt_scene {
    std::vector<float> triP, triN;
    std::vector<int32_t> triMat, triLight;
    std::vector<float> mats;
    std::vector<int32_t> lightTris;
    std::vector<float> lightEmit;
    float cam[14];
    drmlt_options opt;
    std::string output, integrator;
    int32_t reportIntervalSeconds = 0;
};

namespace {

struct Err : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// ------------------------------------------------------------- XML (subset)
// Elements, attributes (single or double quoted, the five predefined
// entities and numeric character references), self-closing tags; the XML
// declaration, comments, DOCTYPE and CDATA are skipped, text is ignored.
struct XNode {
    std::string name;
    std::vector<std::pair<std::string, std::string>> attrs;
    std::vector<XNode> kids;
    const char *attr(const char *n) const {
        for (auto &a : attrs)
            if (a.first == n) return a.second.c_str();
        return nullptr;
    }
    std::string get(const char *n) const {
        const char *v = attr(n);
        return v ? std::string(v) : std::string();
    }
};

struct XParser {
    const std::string &s;
    size_t i = 0;
    explicit XParser(const std::string &src) : s(src) {}
    [[noreturn]] void fail(const char *what) const {
        throw Err(std::string("XML parse error at offset ") + std::to_string(i) + ": " + what);
    }
    bool starts(const char *t) const { return s.compare(i, std::strlen(t), t) == 0; }
    void skip_ws() {
        while (i < s.size() && std::isspace((unsigned char)s[i])) i++;
    }
    void skip_past(const char *t) {
        const size_t j = s.find(t, i);
        if (j == std::string::npos) fail("unterminated construct");
        i = j + std::strlen(t);
    }
    // skips text, comments, declarations up to the next element or end tag
    void skip_misc() {
        for (;;) {
            while (i < s.size() && s[i] != '<') i++;
            if (i >= s.size()) return;
            if (starts("<?")) skip_past("?>");
            else if (starts("<!--")) skip_past("-->");
            else if (starts("<![CDATA[")) skip_past("]]>");
            else if (starts("<!")) skip_past(">");
            else return;
        }
    }
    std::string name() {
        const size_t j = i;
        while (i < s.size() && (std::isalnum((unsigned char)s[i]) || s[i] == '_' || s[i] == '-' || s[i] == ':' || s[i] == '.')) i++;
        if (i == j) fail("expected a name");
        return s.substr(j, i - j);
    }
    static std::string decode(const std::string &v) {
        std::string o;
        for (size_t k = 0; k < v.size(); k++) {
            if (v[k] != '&') { o.push_back(v[k]); continue; }
            const size_t e = v.find(';', k);
            if (e == std::string::npos) throw Err("bad entity in attribute");
            const std::string ent = v.substr(k + 1, e - k - 1);
            if (ent == "amp") o.push_back('&');
            else if (ent == "lt") o.push_back('<');
            else if (ent == "gt") o.push_back('>');
            else if (ent == "quot") o.push_back('"');
            else if (ent == "apos") o.push_back('\'');
            else if (!ent.empty() && ent[0] == '#') {
                const long c = ent.size() > 1 && (ent[1] == 'x' || ent[1] == 'X') ? std::strtol(ent.c_str() + 2, nullptr, 16)
                                                                                  : std::strtol(ent.c_str() + 1, nullptr, 10);
                if (c <= 0 || c > 127) throw Err("non-ASCII character reference in attribute");
                o.push_back(char(c));
            } else throw Err("unknown entity &" + ent + ";");
            k = e;
        }
        return o;
    }
    XNode element() {
        if (i >= s.size() || s[i] != '<') fail("expected an element");
        i++;
        XNode n;
        n.name = name();
        for (;;) {
            skip_ws();
            if (i >= s.size()) fail("unterminated tag");
            if (starts("/>")) { i += 2; return n; }
            if (s[i] == '>') { i++; break; }
            std::string a = name();
            skip_ws();
            if (i >= s.size() || s[i] != '=') fail("expected '='");
            i++;
            skip_ws();
            if (i >= s.size() || (s[i] != '"' && s[i] != '\'')) fail("expected a quoted value");
            const char q = s[i++];
            const size_t e = s.find(q, i);
            if (e == std::string::npos) fail("unterminated attribute value");
            n.attrs.emplace_back(a, decode(s.substr(i, e - i)));
            i = e + 1;
        }
        for (;;) {
            skip_misc();
            if (i >= s.size()) fail("unterminated element");
            if (starts("</")) {
                i += 2;
                const std::string c = name();
                if (c != n.name) fail("mismatched end tag");
                skip_ws();
                if (i >= s.size() || s[i] != '>') fail("expected '>'");
                i++;
                return n;
            }
            n.kids.push_back(element());
        }
    }
    XNode document() {
        skip_misc();
        XNode root = element();
        return root;
    }
};

// ------------------------------------------------------------- numbers
float to_float(const std::string &v, const char *what) {  // std::stof semantics: a leading number
    char *end = nullptr;
    const float f = std::strtof(v.c_str(), &end);
    if (end == v.c_str()) throw Err(std::string("not a number in ") + what + ": '" + v + "'");
    return f;
}
int to_int(const std::string &v, const char *what) {  // std::stoi / atoi semantics
    char *end = nullptr;
    const long x = std::strtol(v.c_str(), &end, 10);
    if (end == v.c_str()) throw Err(std::string("not an integer in ") + what + ": '" + v + "'");
    return int(x);
}
// ParseVector3 (parsescene.cpp:42-60): one value or three, split on commas and spaces
void vec3(const std::string &v, float out[3], const char *what) {
    std::vector<std::string> parts;
    std::string cur;
    for (char c : v) {
        if (c == ',' || c == ' ') {
            if (!cur.empty()) parts.push_back(cur);
            cur.clear();
        } else cur.push_back(c);
    }
    if (!cur.empty()) parts.push_back(cur);
    if (parts.size() == 1) out[0] = out[1] = out[2] = to_float(parts[0], what);
    else if (parts.size() == 3) for (int k = 0; k < 3; k++) out[k] = to_float(parts[k], what);
    else throw Err(std::string("expected 1 or 3 values in ") + what);
}

// ------------------------------------------------------------- 4x4 (row-major, float as the reference's Float)
struct M4 {
    float m[16];
    static M4 id() {
        M4 r{};
        r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1;
        return r;
    }
    float &operator()(int i, int j) { return m[4 * i + j]; }
    float operator()(int i, int j) const { return m[4 * i + j]; }
};
M4 mul(const M4 &a, const M4 &b) {
    M4 r{};
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            float s = 0;
            for (int k = 0; k < 4; k++) s += a(i, k) * b(k, j);
            r(i, j) = s;
        }
    return r;
}
M4 inverse(const M4 &m) {
    double a[4][8];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 8; j++) a[i][j] = j < 4 ? m(i, j) : (j - 4 == i ? 1.0 : 0.0);
    for (int c = 0; c < 4; c++) {
        int p = c;
        for (int r = c + 1; r < 4; r++)
            if (std::fabs(a[r][c]) > std::fabs(a[p][c])) p = r;
        if (a[p][c] == 0) throw Err("singular toWorld transform");
        for (int j = 0; j < 8; j++) std::swap(a[c][j], a[p][j]);
        const double d = a[c][c];
        for (int j = 0; j < 8; j++) a[c][j] /= d;
        for (int r = 0; r < 4; r++)
            if (r != c) {
                const double f = a[r][c];
                for (int j = 0; j < 8; j++) a[r][j] -= f * a[c][j];
            }
    }
    M4 o;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) o(i, j) = float(a[i][j + 4]);
    return o;
}
void normalize3(float v[3]) {
    const float l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    for (int k = 0; k < 3; k++) v[k] /= l;
}
void cross3(const float a[3], const float b[3], float r[3]) {
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}
// transform.cpp:4-90
M4 scale(float x, float y, float z) {
    M4 r = M4::id();
    r(0, 0) = x; r(1, 1) = y; r(2, 2) = z;
    return r;
}
M4 translate(float x, float y, float z) {
    M4 r = M4::id();
    r(0, 3) = x; r(1, 3) = y; r(2, 3) = z;
    return r;
}
M4 rotate(float angle, float ax, float ay, float az) {
    float a[3] = {ax, ay, az};
    normalize3(a);
    const float rad = (float(M_PI) / 180.0f) * angle;
    const float s = std::sin(rad), c = std::cos(rad);
    M4 r = M4::id();
    r(0, 0) = a[0] * a[0] + (1.0f - a[0] * a[0]) * c;
    r(0, 1) = a[0] * a[1] * (1.0f - c) - a[2] * s;
    r(0, 2) = a[0] * a[2] * (1.0f - c) + a[1] * s;
    r(1, 0) = a[0] * a[1] * (1.0f - c) + a[2] * s;
    r(1, 1) = a[1] * a[1] + (1.0f - a[1] * a[1]) * c;
    r(1, 2) = a[1] * a[2] * (1.0f - c) - a[0] * s;
    r(2, 0) = a[0] * a[2] * (1.0f - c) - a[1] * s;
    r(2, 1) = a[1] * a[2] * (1.0f - c) + a[0] * s;
    r(2, 2) = a[2] * a[2] + (1.0f - a[2] * a[2]) * c;
    return r;
}
M4 look_at(const float pos[3], const float look[3], const float up[3]) {
    float dir[3] = {look[0] - pos[0], look[1] - pos[1], look[2] - pos[2]};
    normalize3(dir);
    float un[3] = {up[0], up[1], up[2]};
    normalize3(un);
    float left[3];
    cross3(un, dir, left);
    if (std::sqrt(left[0] * left[0] + left[1] * left[1] + left[2] * left[2]) == 0)
        throw Err("lookat: up vector and viewing direction point in the same direction");
    normalize3(left);
    float nu[3];
    cross3(dir, left, nu);
    M4 r = M4::id();
    for (int k = 0; k < 3; k++) {
        r(k, 0) = left[k];
        r(k, 1) = nu[k];
        r(k, 2) = dir[k];
        r(k, 3) = pos[k];
    }
    return r;
}
// ParseTransform (parsescene.cpp:83-134): each child applies after the previous ones
M4 parse_transform(const XNode &t) {
    M4 r = M4::id();
    for (const XNode &c : t.kids) {
        std::string n = c.name;
        for (char &ch : n) ch = char(std::tolower((unsigned char)ch));
        auto f = [&](const char *a, float d) { const char *v = c.attr(a); return v ? to_float(v, a) : d; };
        if (n == "scale") r = mul(scale(f("x", 1), f("y", 1), f("z", 1)), r);
        else if (n == "translate") r = mul(translate(f("x", 0), f("y", 0), f("z", 0)), r);
        else if (n == "rotate") r = mul(rotate(f("angle", 0), f("x", 0), f("y", 0), f("z", 0)), r);
        else if (n == "lookat") {
            float o[3], g[3], u[3];
            vec3(c.get("origin"), o, "lookat origin");
            vec3(c.get("target"), g, "lookat target");
            vec3(c.get("up"), u, "lookat up");
            r = mul(look_at(o, g, u), r);
        } else if (n == "matrix") {
            const std::string v = c.get("value");
            std::vector<float> x;
            std::string cur;
            for (char ch : v + " ") {
                if (ch == ',' || ch == ' ') {
                    if (!cur.empty()) x.push_back(to_float(cur, "matrix"));
                    cur.clear();
                } else cur.push_back(ch);
            }
            if (x.size() != 16) throw Err("matrix needs 16 values");
            M4 m;
            for (int k = 0; k < 16; k++) m.m[k] = x[k];
            r = mul(m, r);
        }
    }
    return r;
}

// ------------------------------------------------------------- meshes
struct Mesh {
    std::vector<float> p, n;  // per vertex, world space
    std::vector<uint32_t> idx;
};
void xform_point(const M4 &m, const float in[3], float *out) {
    float t[4];
    for (int i = 0; i < 4; i++) t[i] = m(i, 0) * in[0] + m(i, 1) * in[1] + m(i, 2) * in[2] + m(i, 3);
    const float iw = 1.0f / t[3];
    for (int i = 0; i < 3; i++) out[i] = t[i] * iw;
}
void xform_normal(const M4 &inv, const float in[3], float *out) {  // transposed inverse, not renormalised
    for (int i = 0; i < 3; i++) out[i] = inv(0, i) * in[0] + inv(1, i) * in[1] + inv(2, i) * in[2];
}
// ComputeNormal (parseobj.cpp:53-92, loadserialized.cpp:113-152): angle-weighted
// face normals. The reference's angle for obtuse corners is
// (pi - 2) asin(|u + v| / 2) (its UnitAngle, parseobj.cpp:44-51), kept here.
void compute_normals(Mesh &m) {
    const size_t nv = m.p.size() / 3;
    m.n.assign(3 * nv, 0.0f);
    for (size_t t = 0; t + 2 < m.idx.size(); t += 3) {
        float n[3] = {0, 0, 0};
        for (int i = 0; i < 3; i++) {
            const float *v0 = &m.p[3 * m.idx[t + i]], *v1 = &m.p[3 * m.idx[t + (i + 1) % 3]], *v2 = &m.p[3 * m.idx[t + (i + 2) % 3]];
            float a[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]}, b[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
            if (i == 0) {
                cross3(a, b, n);
                const float l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
                if (l == 0) break;
                for (float &x : n) x /= l;
            }
            normalize3(a);
            normalize3(b);
            const float dot = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
            float angle;
            if (dot < 0) {
                const float s[3] = {b[0] + a[0], b[1] + a[1], b[2] + a[2]};
                angle = (float(M_PI) - 2.0f) * std::asin(0.5f * std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]));
            } else {
                const float d[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
                angle = 2.0f * std::asin(0.5f * std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]));
            }
            float *acc = &m.n[3 * m.idx[t + i]];
            for (int k = 0; k < 3; k++) acc[k] += n[k] * angle;
        }
    }
    for (size_t v = 0; v < nv; v++) {
        float *x = &m.n[3 * v];
        const float l = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
        if (l != 0) for (int k = 0; k < 3; k++) x[k] /= l;
        else x[0] = x[1] = x[2] = 0;
    }
}

std::string trim(const std::string &s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) a++;
    while (b > a && std::isspace((unsigned char)s[b - 1])) b--;
    return s.substr(a, b - a);
}

// ParseObj (parseobj.cpp:133-264): v (with optional w), vt, vn (normalised on
// read), triangle and quad faces v, v/vt, v//vn, v/vt/vn (1-based); a vertex
// is one (v, vt, vn) triple.
Mesh load_obj(const std::string &file, const M4 &toWorld) {
    std::ifstream in(file);
    if (!in) throw Err("cannot open OBJ file " + file);
    std::vector<float> pos, nor;
    std::map<std::array<int, 3>, uint32_t> vmap;
    Mesh m;
    bool anyNormal = false, noNormal = false;
    const M4 inv = inverse(toWorld);
    auto vid = [&](const std::string &tok) -> uint32_t {
        std::array<int, 3> id = {0, 0, 0};
        size_t k = 0, f = 0;
        std::string cur;
        for (;; k++) {
            if (k == tok.size() || tok[k] == '/') {
                if (f < 3) id[f] = cur.empty() ? 0 : to_int(cur, "face");
                f++;
                cur.clear();
                if (k == tok.size()) break;
            } else cur.push_back(tok[k]);
        }
        const int v = id[0] - 1, vn = id[2] - 1;
        if (v < 0 || size_t(v) >= pos.size() / 3) throw Err("OBJ face index out of range in " + file);
        auto it = vmap.find({id[0], id[1], id[2]});
        if (it != vmap.end()) return it->second;
        const uint32_t n = uint32_t(m.p.size() / 3);
        float w[3];
        xform_point(toWorld, &pos[3 * v], w);
        m.p.insert(m.p.end(), w, w + 3);
        if (vn >= 0) {
            if (size_t(vn) >= nor.size() / 3) throw Err("OBJ normal index out of range in " + file);
            float t[3];
            xform_normal(inv, &nor[3 * vn], t);
            m.n.insert(m.n.end(), t, t + 3);
            anyNormal = true;
        } else noNormal = true;
        vmap[{id[0], id[1], id[2]}] = n;
        return n;
    };
    std::string line;
    while (std::getline(in, line)) {
        line = trim(line);
        if (line.empty() || line[0] == '#') continue;
        std::stringstream ss(line);
        std::string tok;
        ss >> tok;
        if (tok == "v") {
            float x = 0, y = 0, z = 0, w = 1.f;
            ss >> x >> y >> z >> w;
            const float iw = 1.f / w;
            pos.push_back(x * iw);
            pos.push_back(y * iw);
            pos.push_back(z * iw);
        } else if (tok == "vn") {
            float v[3] = {0, 0, 0};
            ss >> v[0] >> v[1] >> v[2];
            normalize3(v);
            nor.insert(nor.end(), v, v + 3);
        } else if (tok == "f") {
            std::string a, b, c, d, e;
            ss >> a >> b >> c;
            if (c.empty()) throw Err("OBJ face with fewer than three vertices in " + file);
            const uint32_t i0 = vid(a), i1 = vid(b), i2 = vid(c);
            m.idx.insert(m.idx.end(), {i0, i1, i2});
            if (ss >> d) {
                const uint32_t i3 = vid(d);
                m.idx.insert(m.idx.end(), {i0, i2, i3});
            }
            if (ss >> e) throw Err("The object file contains n-gon (n>4) that we do not support: " + file);
        }
    }
    if (anyNormal && noNormal) throw Err("OBJ file mixes vertices with and without normals: " + file);
    if (!anyNormal) compute_normals(m);
    return m;
}

// LoadSerialized (loadserialized.cpp:190-317): Mitsuba's serialized mesh
// format: per shape, the 2-byte format id 0x041C and 2-byte version (3 or 4),
// then a zlib stream of flags (u32), [v4: name, NUL-terminated], vertex and
// triangle counts (u64), positions, [normals], [uv], [colors, read as
// doubles whatever the precision, as the reference does], u32 indices;
// shape k > 0 is found through the offset table at the end of the file.
Mesh load_serialized(const std::string &file, int shapeIndex, const M4 &toWorld) {
    std::ifstream in(file, std::ios::binary);
    if (!in) throw Err("cannot open serialized file " + file);
    std::vector<unsigned char> all((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    if (all.size() < 8) throw Err("serialized file too short: " + file);
    uint16_t fmt, version;
    std::memcpy(&fmt, &all[0], 2);
    std::memcpy(&version, &all[2], 2);
    if (fmt != 0x041C || (version != 3 && version != 4)) throw Err("not a Mitsuba serialized mesh (v3/v4): " + file);
    size_t off = 4;
    if (shapeIndex > 0) {
        uint32_t count;
        std::memcpy(&count, &all[all.size() - 4], 4);
        if (uint32_t(shapeIndex) >= count) throw Err("shapeIndex beyond the shapes of " + file);
        const size_t table = size_t(count) * (version == 4 ? 8 : 4) + 4;
        if (table > all.size()) throw Err("offset table larger than the file " + file);
        if (version == 4) {
            uint64_t o;
            std::memcpy(&o, &all[all.size() - 4 - 8 * size_t(count - shapeIndex)], 8);
            off = size_t(o);
        } else {
            uint32_t o;
            std::memcpy(&o, &all[all.size() - 4 * size_t(count - shapeIndex + 1)], 4);
            off = o;
        }
        off += 4;
    }
    if (off >= all.size()) throw Err("bad shape offset in " + file);
    z_stream z{};
    if (inflateInit2(&z, 15) != Z_OK) throw Err("zlib init failed");
    std::unique_ptr<z_stream, int (*)(z_stream *)> guard(&z, inflateEnd);
    z.next_in = &all[off];
    z.avail_in = uInt(all.size() - off);
    auto read = [&](void *p, size_t n) {  // in pieces: zlib counts bytes in 32 bits
        Bytef *dst = static_cast<Bytef *>(p);
        while (n > 0) {
            const size_t piece = std::min(n, size_t(1) << 30);
            z.next_out = dst;
            z.avail_out = uInt(piece);
            while (z.avail_out > 0) {
                const int r = inflate(&z, Z_NO_FLUSH);
                if (r == Z_STREAM_END && z.avail_out > 0) throw Err("serialized stream ends early in " + file);
                if (r != Z_OK && r != Z_STREAM_END) throw Err("serialized stream is corrupt in " + file);
            }
            dst += piece;
            n -= piece;
        }
    };
    uint32_t flags;
    read(&flags, 4);
    if (version == 4) {
        char c;
        do read(&c, 1);
        while (c != '\0');
    }
    uint64_t nv, nt;
    read(&nv, 8);
    read(&nt, 8);
    if (nv > (1ull << 31) || nt > (1ull << 31)) throw Err("serialized mesh too large: " + file);
    // deflate expands at most 1032:1, so larger arrays than the stream can hold mean a corrupt header
    // (checked before anything is allocated)
    if ((nv * 12 + nt * 12) / 1032 > all.size() - off) throw Err("serialized mesh header exceeds the stream in " + file);
    const bool dbl = (flags & 0x2000) != 0;
    auto vecs = [&](size_t count, int dim, std::vector<float> &out) {
        out.resize(count * dim);
        for (size_t k = 0; k < count * dim; k++) {
            if (dbl) { double d; read(&d, 8); out[k] = float(d); }
            else read(&out[k], 4);
        }
    };
    std::vector<float> P, N, UV, Cl;
    vecs(nv, 3, P);
    if (flags & 0x0001) vecs(nv, 3, N);
    if (flags & 0x0002) vecs(nv, 2, UV);
    if (flags & 0x0008) {
        Cl.resize(nv * 3);
        for (auto &x : Cl) { double d; read(&d, 8); x = float(d); }
    }
    Mesh m;
    m.idx.resize(nt * 3);
    read(m.idx.data(), nt * 12);
    for (uint32_t i : m.idx)
        if (i >= nv) throw Err("serialized face index out of range in " + file);
    m.p.resize(nv * 3);
    for (size_t v = 0; v < nv; v++) xform_point(toWorld, &P[3 * v], &m.p[3 * v]);
    if (!N.empty()) {
        const M4 inv = inverse(toWorld);
        m.n.resize(nv * 3);
        for (size_t v = 0; v < nv; v++) xform_normal(inv, &N[3 * v], &m.n[3 * v]);
    } else compute_normals(m);
    return m;
}

// ------------------------------------------------------------- materials
// Material record of the device path (16 floats): type (0 Lambertian,
// 1 Phong, 2 rough dielectric), Kd|Ks, Ks|Kt, exponent, alpha, eta, 1/eta,
// Phong KsWeight, twoSided.
struct Mat {
    float r[16] = {0};
};
const float kLum[3] = {0.212671f, 0.715160f, 0.072169f};

void value3(const XNode &c, float out[3]) {
    if (c.name == "texture" || c.name == "ref") throw Err("textures are not supported (" + c.get("name") + ")");
    vec3(c.get("value"), out, c.get("name").c_str());
}
float value1(const XNode &c) {
    if (c.name == "texture" || c.name == "ref") throw Err("textures are not supported (" + c.get("name") + ")");
    return to_float(c.get("value"), c.get("name").c_str());
}

// ParseBSDFNode (parsescene.cpp:315-381)
Mat parse_bsdf(const XNode &n, bool twoSided) {
    const std::string type = n.get("type");
    Mat m;
    m.r[12] = twoSided ? 1.0f : 0.0f;
    if (type == "diffuse") {
        float kd[3] = {0.5f, 0.5f, 0.5f};
        for (const XNode &c : n.kids)
            if (c.get("name") == "reflectance") value3(c, kd);
        m.r[0] = 0;
        std::memcpy(&m.r[1], kd, 12);
    } else if (type == "phong") {
        float kd[3] = {0.5f, 0.5f, 0.5f}, ks[3] = {0.2f, 0.2f, 0.2f}, e = 30.0f;
        for (const XNode &c : n.kids) {
            const std::string nm = c.get("name");
            if (nm == "diffuseReflectance") value3(c, kd);
            else if (nm == "specularReflectance") value3(c, ks);
            else if (nm == "exponent") e = value1(c);
        }
        m.r[0] = 1;
        std::memcpy(&m.r[1], kd, 12);
        std::memcpy(&m.r[4], ks, 12);
        m.r[7] = e;
        const float lks = kLum[0] * ks[0] + kLum[1] * ks[1] + kLum[2] * ks[2];
        const float lkd = kLum[0] * kd[0] + kLum[1] * kd[1] + kLum[2] * kd[2];
        m.r[11] = lks + lkd <= 0 ? 0.0f : lks / (lks + lkd);  // phong.cpp:158-166
    } else if (type == "roughdielectric") {
        float ks[3] = {1, 1, 1}, kt[3] = {1, 1, 1}, intIOR = 1.5046f, extIOR = 1.000277f, alpha = 0.1f;
        for (const XNode &c : n.kids) {
            const std::string nm = c.get("name");
            if (nm == "intIOR") intIOR = to_float(c.get("value"), "intIOR");
            else if (nm == "extIOR") extIOR = to_float(c.get("value"), "extIOR");
            else if (nm == "alpha") alpha = value1(c);
            else if (nm == "specularReflectance") value3(c, ks);
            else if (nm == "specularTransmittance") value3(c, kt);
        }
        m.r[0] = 2;
        std::memcpy(&m.r[1], ks, 12);
        std::memcpy(&m.r[4], kt, 12);
        m.r[8] = alpha;
        m.r[9] = intIOR / extIOR;  // roughdielectric.h:17
        m.r[10] = 1.0f / m.r[9];
    } else if (type == "twosided") {
        for (const XNode &c : n.kids)
            if (c.name == "bsdf") return parse_bsdf(c, true);
        throw Err("twosided BSDF without a nested bsdf");
    } else throw Err("Unknown BSDF '" + type + "'");
    return m;
}

// ------------------------------------------------------------- the scene
struct Loader {
    drmlt_scene &S;
    std::string dir;
    std::map<std::string, int> bsdfIds;
    int add_mat(const Mat &m) {
        S.mats.insert(S.mats.end(), m.r, m.r + 16);
        return int(S.mats.size() / 16) - 1;
    }
    std::string path_of(const std::string &f) const { return f.empty() || f[0] == '/' ? f : dir + "/" + f; }

    // ParseSensorNode / ParseFilmNode (parsescene.cpp:148-202)
    void sensor(const XNode &n) {
        float nearC = 1e-2f, farC = 1000.0f, fov = 45.0f;
        int W = 512, H = 512;
        M4 toWorld = M4::id();
        for (const XNode &c : n.kids) {
            const std::string nm = c.get("name");
            if (nm == "nearClip") nearC = to_float(c.get("value"), "nearClip");
            else if (nm == "farClip") farC = to_float(c.get("value"), "farClip");
            else if (nm == "fov") fov = to_float(c.get("value"), "fov");
            else if (nm == "toWorld") {
                if (c.name == "transform") toWorld = parse_transform(c);
                else throw Err("animated sensors are not supported");
            } else if (c.name == "film") {
                for (const XNode &f : c.kids) {
                    const std::string fn = f.get("name");
                    if (fn == "width") W = to_int(f.get("value"), "width");
                    else if (fn == "height") H = to_int(f.get("value"), "height");
                    else if (fn == "filename") S.output = path_of(f.get("value"));
                }
            }
        }
        if (W <= 0 || H <= 0) throw Err("film size must be positive");
        // the device camera is a look-at: recover it from toWorld and check
        const float pos[3] = {toWorld(0, 3), toWorld(1, 3), toWorld(2, 3)};
        const float d[3] = {toWorld(0, 2), toWorld(1, 2), toWorld(2, 2)};
        const float up[3] = {toWorld(0, 1), toWorld(1, 1), toWorld(2, 1)};
        const float tgt[3] = {pos[0] + d[0], pos[1] + d[1], pos[2] + d[2]};
        const M4 back = look_at(pos, tgt, up);
        for (int k = 0; k < 16; k++)
            if (std::fabs(back.m[k] - toWorld.m[k]) > 1e-4f * (1.0f + std::fabs(toWorld.m[k])))
                throw Err("sensor toWorld is not a look-at transform (scales, reflections and shears are not supported)");
        const float cam[14] = {pos[0], pos[1], pos[2], tgt[0], tgt[1], tgt[2], up[0], up[1], up[2], fov, float(W), float(H), nearC, farC};
        std::memcpy(S.cam, cam, sizeof(cam));
    }

    // ParseShapeNode (parsescene.cpp:204-313)
    void shape(const XNode &n) {
        int mat = -1;
        for (const XNode &c : n.kids) {
            if (c.name == "bsdf") { mat = add_mat(parse_bsdf(c, false)); break; }
            if (c.name == "ref") {
                auto it = bsdfIds.find(c.get("id"));
                if (it == bsdfIds.end()) throw Err("ref not found: '" + c.get("id") + "'");
                mat = it->second;
                break;
            }
        }
        const std::string type = n.get("type");
        std::string file;
        int shapeIndex = 0;
        M4 toWorld = M4::id();
        for (const XNode &c : n.kids) {
            const std::string nm = c.get("name");
            if (nm == "filename") file = c.get("value");
            else if (nm == "shapeIndex") shapeIndex = to_int(c.get("value"), "shapeIndex");
            else if (nm == "toWorld") {
                if (c.name == "transform") toWorld = parse_transform(c);
                else throw Err("moving shapes are not supported (static scenes only)");
            }
        }
        Mesh m;
        if (type == "obj") m = load_obj(path_of(file), toWorld);
        else if (type == "serialized") m = load_serialized(path_of(file), shapeIndex, toWorld);
        else throw Err("Invalid shape '" + type + "'");
        int light = -1;
        float radiance[3] = {1, 1, 1};
        for (const XNode &c : n.kids) {
            if (c.name != "emitter") continue;
            for (const XNode &g : c.kids)
                if (g.get("name") == "radiance") vec3(g.get("value"), radiance, "radiance");
            light = int(S.lightTris.size() / 2);
        }
        if (mat < 0) {  // no BSDF on the shape (typically an emitter): black Lambertian
            Mat black;
            mat = add_mat(black);
        }
        const int first = int(S.triMat.size());
        const size_t nt = m.idx.size() / 3;
        for (size_t t = 0; t < nt; t++) {
            for (int k = 0; k < 3; k++) {
                const uint32_t v = m.idx[3 * t + k];
                S.triP.insert(S.triP.end(), &m.p[3 * v], &m.p[3 * v] + 3);
                S.triN.insert(S.triN.end(), &m.n[3 * v], &m.n[3 * v] + 3);
            }
            S.triMat.push_back(mat);
            S.triLight.push_back(light);
        }
        if (light >= 0) {
            S.lightTris.push_back(first);
            S.lightTris.push_back(int32_t(nt));
            S.lightEmit.insert(S.lightEmit.end(), {radiance[0], radiance[1], radiance[2], 1.0f});  // AreaLight weight 1
        }
    }

    // ParseDptOptions (parsescene.cpp:505-564)
    void dpt(const XNode &n) {
        drmlt_options &o = S.opt;
        for (const XNode &c : n.kids) {
            const std::string nm = c.get("name"), v = c.get("value");
            const bool t = v == "true";
            if (nm == "integrator") {
                S.integrator = v;
                o.integrator = v == "drmlt" ? 0 : v == "mcmc" ? 1 : v == "h2mc" ? 2 : -1;  // others: drmlt_ctx_create refuses them
            }
            else if (nm == "spp") o.spp = to_int(v, "spp");
            else if (nm == "numinitsamples") o.num_init_samples = to_int(v, "numinitsamples");
            else if (nm == "largestepprob") o.large_step_prob = to_float(v, "largestepprob");
            else if (nm == "mindepth") o.min_depth = to_int(v, "mindepth");
            else if (nm == "maxdepth") o.max_depth = to_int(v, "maxdepth");
            else if (nm == "directspp") o.direct_spp = to_int(v, "directspp");
            else if (nm == "perturbstddev") o.perturb_std_dev = to_float(v, "perturbstddev");
            else if (nm == "roughnessthreshold") o.roughness_threshold = to_float(v, "roughnessthreshold");
            else if (nm == "numchains") o.num_chains = to_int(v, "numchains");
            else if (nm == "seedoffset") o.seed_offset = to_int(v, "seedoffset");
            else if (nm == "reportintervalseconds") S.reportIntervalSeconds = to_int(v, "reportintervalseconds");
            else if (nm == "uselightcoordinatesampling") o.use_light_coordinate_sampling = t;
            else if (nm == "largestepmultiplexed") o.large_step_multiplexed = t;
            else if (nm == "type") o.type = v == "mira" ? 1 : 0;
            else if (nm == "anisoperturbstddev") o.aniso_perturb_std_dev = o.lens_perturb_std_dev = to_float(v, "anisoperturbstddev");
            else if (nm == "anisoperturbprob") o.aniso_perturb_prob = to_float(v, "anisoperturbprob");
            else if (nm == "acceptanceMap") o.acceptance_map = t;
            else if (nm == "uniformmixingprobability") o.uniform_mixing_prob = to_float(v, "uniformmixingprobability");
            else if (nm == "lensperturbprob") o.lens_perturb_prob = to_float(v, "lensperturbprob");
            else if (nm == "lensperturbstddev") o.lens_perturb_std_dev = to_float(v, "lensperturbstddev");
            else if (nm == "bidirectional" || nm == "reportintervalspp") {
                // parsed by the reference, used by its other integrators only
            } else std::fprintf(stderr, "Unknown dpt option:%s\n", nm.c_str());
        }
    }

    void scene(const XNode &root) {
        if (root.name != "scene") throw Err("the document's root element is not <scene>");
        bool haveSensor = false;
        for (const XNode &c : root.kids) {
            if (c.name == "sensor") { sensor(c); haveSensor = true; }
            else if (c.name == "shape") shape(c);
            else if (c.name == "bsdf") bsdfIds[c.get("id")] = add_mat(parse_bsdf(c, false));
            else if (c.name == "emitter") throw Err("emitter '" + c.get("type") + "' outside a shape is not supported (area lights only)");
            else if (c.name == "texture") throw Err("textures are not supported");
            else if (c.name == "dpt") dpt(c);
        }
        if (!haveSensor) throw Err("the scene has no sensor");
        if (S.triMat.empty()) throw Err("the scene has no triangles");
        if (S.lightTris.empty()) throw Err("the scene has no area light");
    }
};

}  // namespace

extern "C" int drmlt_scene_load(const char *xml_path, drmlt_scene **out) {
    if (!xml_path || !out) return drmlt_set_error(DRMLT_ERR_ARG, "bad argument");
    *out = nullptr;
    try {
        std::ifstream in(xml_path, std::ios::binary);
        if (!in) throw Err(std::string("cannot open ") + xml_path);
        std::stringstream ss;
        ss << in.rdbuf();
        const std::string text = ss.str();
        XParser px(text);
        const XNode root = px.document();
        std::unique_ptr<drmlt_scene> s(new drmlt_scene());
        drmlt_options_default(&s->opt);
        s->integrator = "drmlt";
        std::string p(xml_path);
        const size_t slash = p.find_last_of('/');
        const std::string dir = slash == std::string::npos ? std::string(".") : p.substr(0, slash == 0 ? 1 : slash);
        s->output = dir + "/image.exr";
        Loader L{*s, dir, {}};
        L.scene(root);
        *out = s.release();
        return DRMLT_OK;
    } catch (const std::exception &e) {
        return drmlt_set_error(DRMLT_ERR_ARG, e.what());
    }
}

extern "C" int drmlt_scene_get(const drmlt_scene *s, drmlt_scene_desc *desc, drmlt_options *opt) {
    if (!s) return drmlt_set_error(DRMLT_ERR_ARG, "bad argument");
    if (desc) {
        desc->num_triangles = int32_t(s->triMat.size());
        desc->tri_p = s->triP.data();
        desc->tri_n = s->triN.data();
        desc->tri_mat = s->triMat.data();
        desc->tri_light = s->triLight.data();
        desc->num_materials = int32_t(s->mats.size() / 16);
        desc->materials = s->mats.data();
        desc->num_lights = int32_t(s->lightTris.size() / 2);
        desc->light_tris = s->lightTris.data();
        desc->light_emit = s->lightEmit.data();
        desc->camera = s->cam;
    }
    if (opt) *opt = s->opt;
    return DRMLT_OK;
}

extern "C" int drmlt_scene_output(const drmlt_scene *s, char *output_name, int32_t cap, char *integrator, int32_t icap,
                                  int32_t *report_interval_seconds) {
    if (!s || cap < 0 || icap < 0) return drmlt_set_error(DRMLT_ERR_ARG, "bad argument");
    auto put = [](const std::string &v, char *buf, int32_t n) {
        if (!buf || n <= 0) return;
        const size_t k = std::min(v.size(), size_t(n - 1));
        std::memcpy(buf, v.data(), k);
        buf[k] = '\0';
    };
    put(s->output, output_name, cap);
    put(s->integrator, integrator, icap);
    if (report_interval_seconds) *report_interval_seconds = s->reportIntervalSeconds;
    return int(std::max(s->output.size(), s->integrator.size()));
}

extern "C" int drmlt_scene_free(drmlt_scene *s) {
    delete s;
    return DRMLT_OK;
}
