// scene_loader.cpp -- reference scene JSON + OBJ / PLY import -> kr_scene_desc arrays.
#include "scene_loader.h"

#include <algorithm>
#include <cctype>
#include <tuple>
#include <array>
#include <cstdint>
#include <cstring>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <unordered_map>

#include "log.h"

namespace koishi {

namespace {

std::string read_file(const std::string &path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::logic_error("cannot open '" + path + "'");
    std::ostringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

std::string dir_of(const std::string &path) {
    size_t k = path.find_last_of('/');
    return k == std::string::npos ? std::string(".") : path.substr(0, k);
}

void f3(const Json &j, float out[3], const char *what) {
    if (!j.is_array() || j.size() != 3) throw std::logic_error(std::string(what) + ": expected [x, y, z]");
    for (int k = 0; k < 3; ++k) out[k] = float(j[k].number());
}

// get(props, key, default) of src/util/jsel.hpp:264-279
float getf(const Json &props, const char *key, float dflt) {
    const Json *j = props.find(key);
    return j ? float(j->number()) : dflt;
}
void get3(const Json &props, const char *key, const float dflt[3], float out[3]) {
    const Json *j = props.find(key);
    if (j)
        f3(*j, out, key);
    else
        for (int k = 0; k < 3; ++k) out[k] = dflt[k];
}

// OBJ index: 1-based, negative = relative to the end
long resolve(long i, size_t n) {
    if (i > 0) return i - 1;
    if (i < 0) return long(n) + i;
    return -1;
}

struct Key {
    long v, n;
    bool operator==(const Key &o) const { return v == o.v && n == o.n; }
};
struct KeyHash {
    size_t operator()(const Key &k) const { return std::hash<long>()(k.v * 1000003L ^ k.n); }
};

// Normals for the vertices a file gives none, as assimp's GenSmoothNormals
// step does (the reference imports with aiProcess_GenSmoothNormals,
// src/core/meta/scene.cu:39-44): every face has its unit geometric normal
// (e1 x e2, normalised); a vertex gets the normalised sum of the unit normals
// of all faces with a corner at its position (positions compared exactly),
// each face counted once per corner -- not weighted by area.
void smooth_missing_normals(ObjMesh &m, const std::vector<char> &needs) {
    bool any = false;
    for (char c : needs) any |= c != 0;
    if (!any) return;
    struct PosKey {
        uint32_t b[3];
        bool operator==(const PosKey &o) const { return b[0] == o.b[0] && b[1] == o.b[1] && b[2] == o.b[2]; }
    };
    struct PosHash {
        size_t operator()(const PosKey &k) const {
            return (size_t(k.b[0]) * 0x9E3779B1u) ^ (size_t(k.b[1]) * 0x85EBCA77u) ^ (size_t(k.b[2]) * 0xC2B2AE3Du);
        }
    };
    auto key = [&](uint32_t v) {
        PosKey k;
        for (int c = 0; c < 3; ++c) {
            float x = m.positions[3 * size_t(v) + c];
            if (x == 0.f) x = 0.f;  // +0 and -0 are one position
            std::memcpy(&k.b[c], &x, 4);
        }
        return k;
    };
    std::unordered_map<PosKey, std::array<double, 3>, PosHash> acc;
    for (size_t t = 0; t < m.indices.size(); t += 3) {
        const float *a = &m.positions[3 * size_t(m.indices[t])], *b = &m.positions[3 * size_t(m.indices[t + 1])],
                    *c = &m.positions[3 * size_t(m.indices[t + 2])];
        double e1[3] = {double(b[0]) - a[0], double(b[1]) - a[1], double(b[2]) - a[2]};
        double e2[3] = {double(c[0]) - a[0], double(c[1]) - a[1], double(c[2]) - a[2]};
        double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        double l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        if (!(l > 0)) continue;  // a degenerate face has no normal
        for (int k = 0; k < 3; ++k) {
            auto &s3 = acc.emplace(key(m.indices[t + k]), std::array<double, 3>{0, 0, 0}).first->second;
            for (int c3 = 0; c3 < 3; ++c3) s3[c3] += n[c3] / l;
        }
    }
    for (size_t v = 0; v < needs.size(); ++v) {
        if (!needs[v]) continue;
        auto it = acc.find(key(uint32_t(v)));
        double s3[3] = {0, 0, 0};
        if (it != acc.end())
            for (int c3 = 0; c3 < 3; ++c3) s3[c3] = it->second[c3];
        double l = std::sqrt(s3[0] * s3[0] + s3[1] * s3[1] + s3[2] * s3[2]);
        for (int c3 = 0; c3 < 3; ++c3) m.normals[3 * v + c3] = l > 0 ? float(s3[c3] / l) : 0.f;
    }
}

}  // namespace

std::vector<ObjMesh> load_obj(const std::string &path, std::vector<std::string> &material_names) {
    std::ifstream f(path);
    if (!f) throw std::logic_error("cannot open '" + path + "'");
    std::vector<float> V, N;
    std::vector<ObjMesh> meshes;
    std::vector<std::unordered_map<Key, uint32_t, KeyHash>> remap;
    std::vector<std::vector<char>> needs_normal;  // per mesh: vertex lacks vn
    std::string object = "", material = "DefaultMaterial";
    bool used_default = false;
    std::vector<std::string> used;  // usemtl names in first-use order
    int cur = -1;
    auto open_mesh = [&]() {
        meshes.push_back(ObjMesh{object, material, {}, {}, {}});
        remap.emplace_back();
        needs_normal.emplace_back();
        cur = int(meshes.size()) - 1;
    };
    std::string line;
    size_t lineno = 0;
    while (std::getline(f, line)) {
        ++lineno;
        if (!line.empty() && line.back() == '\r') line.pop_back();
        std::istringstream ls(line);
        std::string tag;
        if (!(ls >> tag) || tag[0] == '#') continue;
        if (tag == "v") {
            float x, y, z;
            if (!(ls >> x >> y >> z)) throw std::logic_error(path + ":" + std::to_string(lineno) + ": bad vertex");
            V.insert(V.end(), {x, y, z});
        } else if (tag == "vn") {
            float x, y, z;
            if (!(ls >> x >> y >> z)) throw std::logic_error(path + ":" + std::to_string(lineno) + ": bad normal");
            N.insert(N.end(), {x, y, z});
        } else if (tag == "o" || tag == "g") {
            std::string name;
            std::getline(ls >> std::ws, name);
            object = name;
            cur = -1;
        } else if (tag == "usemtl") {
            std::string name;
            ls >> name;
            // assimp creates a named material on first mention, faces or not
            if (name != "DefaultMaterial" && std::find(used.begin(), used.end(), name) == used.end())
                used.push_back(name);
            if (name != material) {
                material = name;
                cur = -1;
            }
        } else if (tag == "f") {
            if (cur < 0 || meshes[cur].material != material || meshes[cur].object != object) open_mesh();
            if (material == "DefaultMaterial") used_default = true;
            std::vector<uint32_t> poly;
            std::string tok;
            while (ls >> tok) {
                long vi = 0, ni = 0;
                size_t s1 = tok.find('/');
                vi = std::strtol(tok.substr(0, s1).c_str(), nullptr, 10);
                if (s1 != std::string::npos) {
                    size_t s2 = tok.find('/', s1 + 1);
                    if (s2 != std::string::npos && s2 + 1 < tok.size())
                        ni = std::strtol(tok.substr(s2 + 1).c_str(), nullptr, 10);
                }
                long v = resolve(vi, V.size() / 3), n = resolve(ni, N.size() / 3);
                if (v < 0 || size_t(v) >= V.size() / 3)
                    throw std::logic_error(path + ":" + std::to_string(lineno) + ": vertex index out of range");
                if (n >= 0 && size_t(n) >= N.size() / 3)
                    throw std::logic_error(path + ":" + std::to_string(lineno) + ": normal index out of range");
                Key k{v, n};
                ObjMesh &m = meshes[cur];
                auto it = remap[cur].find(k);
                uint32_t id;
                if (it == remap[cur].end()) {
                    id = uint32_t(m.positions.size() / 3);
                    m.positions.insert(m.positions.end(), {V[3 * v], V[3 * v + 1], V[3 * v + 2]});
                    if (n >= 0)
                        m.normals.insert(m.normals.end(), {N[3 * n], N[3 * n + 1], N[3 * n + 2]});
                    else
                        m.normals.insert(m.normals.end(), {0.f, 0.f, 0.f});
                    needs_normal[cur].push_back(n < 0);
                    remap[cur].emplace(k, id);
                } else {
                    id = it->second;
                }
                poly.push_back(id);
            }
            if (poly.size() < 3) throw std::logic_error(path + ":" + std::to_string(lineno) + ": face with < 3 vertices");
            for (size_t k = 1; k + 1 < poly.size(); ++k)
                meshes[cur].indices.insert(meshes[cur].indices.end(), {poly[0], poly[k], poly[k + 1]});
        }
        // vt, s, mtllib, l, p: not used by the Radiance path
    }
    for (size_t mi = 0; mi < meshes.size(); ++mi) smooth_missing_normals(meshes[mi], needs_normal[mi]);
    material_names.clear();
    if (used_default) material_names.push_back("DefaultMaterial");
    for (auto &u : used) material_names.push_back(u);
    std::vector<ObjMesh> out;
    for (auto &m : meshes)
        if (!m.indices.empty()) out.push_back(std::move(m));
    return out;
}

// PLY (Turk 1994; the format the reference's assimp import also reads):
// ascii, binary_little_endian or binary_big_endian 1.0. The "vertex" element
// gives x, y, z and, when present, nx, ny, nz; the "face" element a list of
// vertex indices ("vertex_indices" or "vertex_index") per polygon, split into
// a fan of triangles like OBJ faces (aiProcess_Triangulate on convex
// polygons). Other elements and properties are read and skipped. A PLY file
// has no materials: assimp gives its one mesh the material
// "DefaultMaterial". Missing normals are generated as for OBJ.
namespace {
struct PlyProp {
    std::string name;
    int type = 0;       // scalar or list item type (bytes: 1, 2, 4, 8; float flag below)
    bool is_float = false, is_signed = false;
    bool list = false;
    int count_type = 0; // list count: bytes
    bool count_signed = false;
};
struct PlyElem {
    std::string name;
    size_t count = 0;
    std::vector<PlyProp> props;
};

bool ply_type(const std::string &t, int &bytes, bool &is_float, bool &is_signed) {
    static const std::map<std::string, std::tuple<int, bool, bool>> types = {
        {"char", {1, false, true}},    {"int8", {1, false, true}},   {"uchar", {1, false, false}},
        {"uint8", {1, false, false}},  {"short", {2, false, true}},  {"int16", {2, false, true}},
        {"ushort", {2, false, false}}, {"uint16", {2, false, false}}, {"int", {4, false, true}},
        {"int32", {4, false, true}},   {"uint", {4, false, false}},  {"uint32", {4, false, false}},
        {"float", {4, true, true}},    {"float32", {4, true, true}}, {"double", {8, true, true}},
        {"float64", {8, true, true}}};
    auto it = types.find(t);
    if (it == types.end()) return false;
    std::tie(bytes, is_float, is_signed) = it->second;
    return true;
}

struct PlyReader {
    const std::string &data, &path;
    size_t pos;
    int format;  // 0 ascii, 1 little endian, 2 big endian
    [[noreturn]] void bad(const std::string &what) const { throw std::logic_error(path + ": PLY: " + what); }
    double binary(int bytes, bool is_float, bool is_signed) {
        if (pos + size_t(bytes) > data.size()) bad("unexpected end of data");
        unsigned char b[8];
        std::memcpy(b, data.data() + pos, size_t(bytes));
        pos += size_t(bytes);
        if (format == 2) std::reverse(b, b + bytes);  // the host is little endian
        if (is_float) {
            if (bytes == 4) {
                float f;
                std::memcpy(&f, b, 4);
                return f;
            }
            double d;
            std::memcpy(&d, b, 8);
            return d;
        }
        uint64_t u = 0;
        std::memcpy(&u, b, size_t(bytes));
        if (is_signed && bytes < 8 && (u >> (8 * bytes - 1)) & 1u) u |= ~uint64_t(0) << (8 * bytes);
        return is_signed ? double(int64_t(u)) : double(u);
    }
    double ascii() {
        while (pos < data.size() && std::isspace(static_cast<unsigned char>(data[pos]))) ++pos;
        if (pos >= data.size()) bad("unexpected end of data");
        const char *b = data.c_str() + pos;
        char *e = nullptr;
        double v = std::strtod(b, &e);
        if (e == b) bad("bad number");
        pos += size_t(e - b);
        return v;
    }
    double value(int bytes, bool is_float, bool is_signed) {
        return format == 0 ? ascii() : binary(bytes, is_float, is_signed);
    }
};
}  // namespace

std::vector<ObjMesh> load_ply(const std::string &path, std::vector<std::string> &material_names) {
    const std::string data = read_file(path);
    // header: lines up to "end_header"
    size_t pos = 0;
    auto next_line = [&](std::string &line) {
        if (pos >= data.size()) return false;
        size_t e = data.find('\n', pos);
        if (e == std::string::npos) e = data.size();
        line = data.substr(pos, e - pos);
        if (!line.empty() && line.back() == '\r') line.pop_back();
        pos = e + 1;
        return true;
    };
    std::string line;
    if (!next_line(line) || line != "ply") throw std::logic_error(path + ": not a PLY file");
    int format = -1;
    std::vector<PlyElem> elems;
    bool ended = false;
    while (next_line(line)) {
        std::istringstream ls(line);
        std::string kw;
        if (!(ls >> kw)) continue;
        if (kw == "format") {
            std::string f, ver;
            ls >> f >> ver;
            format = f == "ascii" ? 0 : f == "binary_little_endian" ? 1 : f == "binary_big_endian" ? 2 : -1;
            if (format < 0) throw std::logic_error(path + ": PLY: unknown format '" + f + "'");
        } else if (kw == "element") {
            PlyElem el;
            long long n = -1;
            ls >> el.name >> n;
            if (n < 0) throw std::logic_error(path + ": PLY: bad element line");
            el.count = size_t(n);
            elems.push_back(el);
        } else if (kw == "property") {
            if (elems.empty()) throw std::logic_error(path + ": PLY: property before any element");
            PlyProp pr;
            std::string t;
            ls >> t;
            if (t == "list") {
                std::string ct, it;
                ls >> ct >> it >> pr.name;
                bool cf;
                if (!ply_type(ct, pr.count_type, cf, pr.count_signed) || cf ||
                    !ply_type(it, pr.type, pr.is_float, pr.is_signed))
                    throw std::logic_error(path + ": PLY: bad list property");
                pr.list = true;
            } else {
                if (!ply_type(t, pr.type, pr.is_float, pr.is_signed))
                    throw std::logic_error(path + ": PLY: unknown property type '" + t + "'");
                ls >> pr.name;
            }
            elems.back().props.push_back(pr);
        } else if (kw == "end_header") {
            ended = true;
            break;
        }
        // comment, obj_info: ignored
    }
    if (!ended || format < 0) throw std::logic_error(path + ": PLY: incomplete header");
    PlyReader rd{data, path, pos, format};
    std::vector<float> V, N;
    bool has_n = false;
    ObjMesh m{"", "DefaultMaterial", {}, {}, {}};
    for (const PlyElem &el : elems) {
        int ix[6] = {-1, -1, -1, -1, -1, -1};
        int fl = -1;
        for (size_t k = 0; k < el.props.size(); ++k) {
            static const char *names[6] = {"x", "y", "z", "nx", "ny", "nz"};
            for (int c = 0; c < 6; ++c)
                if (el.props[k].name == names[c]) ix[c] = int(k);
            if (el.props[k].list && (el.props[k].name == "vertex_indices" || el.props[k].name == "vertex_index"))
                fl = int(k);
        }
        bool is_vertex = el.name == "vertex", is_face = el.name == "face";
        if (is_vertex) {
            if (ix[0] < 0 || ix[1] < 0 || ix[2] < 0) throw std::logic_error(path + ": PLY: vertex without x, y, z");
            has_n = ix[3] >= 0 && ix[4] >= 0 && ix[5] >= 0;
            V.reserve(3 * el.count);
            if (has_n) N.reserve(3 * el.count);
        }
        if (is_face && fl < 0) throw std::logic_error(path + ": PLY: face without vertex_indices");
        std::vector<double> vals(el.props.size());
        std::vector<uint32_t> poly;
        for (size_t i = 0; i < el.count; ++i) {
            for (size_t k = 0; k < el.props.size(); ++k) {
                const PlyProp &pr = el.props[k];
                if (!pr.list) {
                    vals[k] = rd.value(pr.type, pr.is_float, pr.is_signed);
                    continue;
                }
                double cnt = rd.value(pr.count_type, false, pr.count_signed);
                if (cnt < 0) rd.bad("negative list length");
                bool keep = is_face && int(k) == fl;
                if (keep) poly.clear();
                for (size_t j = 0; j < size_t(cnt); ++j) {
                    double x = rd.value(pr.type, pr.is_float, pr.is_signed);
                    if (keep) {
                        if (x < 0 || x >= double(V.size() / 3)) rd.bad("vertex index out of range");
                        poly.push_back(uint32_t(x));
                    }
                }
            }
            if (is_vertex) {
                for (int c = 0; c < 3; ++c) V.push_back(float(vals[size_t(ix[c])]));
                if (has_n)
                    for (int c = 3; c < 6; ++c) N.push_back(float(vals[size_t(ix[c])]));
            } else if (is_face) {
                if (poly.size() < 3) rd.bad("face with < 3 vertices");
                for (size_t k = 1; k + 1 < poly.size(); ++k) m.indices.insert(m.indices.end(), {poly[0], poly[k], poly[k + 1]});
            }
        }
    }
    m.positions = V;
    std::vector<char> needs(V.size() / 3, has_n ? 0 : 1);
    m.normals = has_n ? N : std::vector<float>(V.size(), 0.f);
    smooth_missing_normals(m, needs);
    material_names.assign(1, "DefaultMaterial");
    std::vector<ObjMesh> out;
    if (!m.indices.empty()) out.push_back(std::move(m));
    return out;
}

std::vector<ObjMesh> load_mesh_file(const std::string &path, std::vector<std::string> &material_names) {
    std::string ext;
    size_t d = path.find_last_of('.');
    if (d != std::string::npos)
        for (char c : path.substr(d + 1)) ext.push_back(char(std::tolower(static_cast<unsigned char>(c))));
    if (ext == "ply") return load_ply(path, material_names);
    return load_obj(path, material_names);
}

namespace {
// "distribution": Config of a SphericalDistribution (distribution.impl.cu:8-11)
void parse_distribution(const Json &props, kr_shader_desc &d) {
    const Json &dist = props.at("distribution");  // required (glossy.hpp:13-14)
    std::string dn;
    const Json *dp = nullptr;
    if (dist.is_string()) {
        dn = dist.str();
    } else {
        if (!dist.is_array() || dist.size() < 1) throw std::logic_error("distribution must be [type, {props}]");
        dn = dist[0].str();
        if (dist.size() > 1) dp = &dist[1];
    }
    if (dn == "Beckmann") {
        d.distribution = KR_DIST_BECKMANN;
        d.roughness = dp ? getf(*dp, "roughness", .5f) : .5f;
    } else if (dn == "Cosine") {
        d.distribution = KR_DIST_COSINE;
    } else {
        throw std::logic_error("no such type: " + dn);
    }
}

float required_ior(const Json &props) { return float(props.at("ior").number()); }  // get<float>(props, "ior")

constexpr int kMaxMixedDepth = 3;

// Fills table[slot] from cfg; a Mixed's two inline shaders are appended to
// the table (Factory<Shader>::create of "shaders"[0], [1], mixed.hpp:12-18).
void fill_shader(const Json &cfg, std::vector<kr_shader_desc> &table, size_t slot, int depth) {
    if (cfg.is_string()) {  // a bare name, e.g. "Blackhole"
        Json arr = Json::array();
        arr.push(cfg);
        fill_shader(arr, table, slot, depth);
        return;
    }
    if (!cfg.is_array() || cfg.size() < 1 || !cfg[0].is_string())
        throw std::logic_error("shader config must be [type, {props}]");
    const std::string type = cfg[0].str();
    Json empty = Json::object();
    const Json &props = cfg.size() > 1 ? cfg[1] : empty;
    kr_shader_desc d{};
    const float one[3] = {1, 1, 1};
    get3(props, "color", one, d.color);
    get3(props, "R", one, d.R);
    d.roughness = 0.5f;
    d.strength = getf(props, "strength", 2.f);
    if (type == "Lambert") {
        d.type = KR_SHADER_LAMBERT;
    } else if (type == "GlossyBSDF") {
        d.type = KR_SHADER_GLOSSY;
        parse_distribution(props, d);
    } else if (type == "Emission") {
        d.type = KR_SHADER_EMISSION;
    } else if (type == "GlassBSDF") {
        d.type = KR_SHADER_GLASS;
        d.ior = required_ior(props);
        parse_distribution(props, d);
    } else if (type == "RefractionBSDF") {
        d.type = KR_SHADER_REFRACTION;
        d.ior = required_ior(props);
        parse_distribution(props, d);
    } else if (type == "Transparent") {
        d.type = KR_SHADER_TRANSPARENT;
    } else if (type == "Blackhole") {
        d.type = KR_SHADER_BLACKHOLE;
    } else if (type == "Mixed") {
        d.type = KR_SHADER_MIXED;
        if (depth >= kMaxMixedDepth)
            throw std::logic_error("Mixed shaders nest deeper than " + std::to_string(kMaxMixedDepth) +
                                   " levels; this build does not provide that");
        const Json &fac = props.at("fac");
        if (!fac.is_array() || fac.size() < 1) throw std::logic_error("fac must be [type, {props}]");
        const std::string fn = fac[0].str();
        const Json &fp = fac.size() > 1 ? fac[1] : empty;
        if (fn == "Constant") {
            d.fac_type = KR_SCALA_CONSTANT;
            d.fac_value = float(fp.at("value").number());
        } else if (fn == "Fresnel") {
            d.fac_type = KR_SCALA_FRESNEL;
            d.fac_value = required_ior(fp);
        } else {
            throw std::logic_error("no such type: " + fn);
        }
        const Json &sh = props.at("shaders");
        if (!sh.is_array() || sh.size() < 2) throw std::logic_error("Mixed needs \"shaders\": [a, b]");
        table[slot] = d;
        for (int k = 0; k < 2; ++k) {
            size_t c = table.size();
            table.push_back(kr_shader_desc{});
            table[slot].child[k] = uint32_t(c);
            fill_shader(sh[k], table, c, depth + 1);
        }
        return;
    } else {
        throw std::logic_error("no such type: " + type);
    }
    table[slot] = d;
}
}  // namespace

kr_shader_desc shader_from_config(const Json &cfg) {
    std::vector<kr_shader_desc> table(1);
    fill_shader(cfg, table, 0, 0);
    if (table.size() > 1) throw std::logic_error("Mixed shaders need a shader table (load the scene instead)");
    return table[0];
}

namespace {
kr_camera camera_from_json(const Json &c) {
    kr_camera cam{};
    cam.lens = KR_LENS_PINHOLE;
    if (const Json *l = c.find("lens")) {
        const std::string &s = l->str();
        if (s == "pinhole")
            cam.lens = KR_LENS_PINHOLE;
        else if (s == "orthographic")
            cam.lens = KR_LENS_ORTHOGRAPHIC;
        else
            throw std::logic_error("unknown lens type: " + s);
    }
    cam.fovx = getf(c, "fovx", 90.f);
    const float pos[3] = {1, 0, 0}, tgt[3] = {-1, 0, 0}, up[3] = {0, 0, 1};
    get3(c, "position", pos, cam.position);
    get3(c, "target", tgt, cam.target);
    get3(c, "upaxis", up, cam.upaxis);
    return cam;
}
}  // namespace

LoadedScene load_scene_json(const Json &doc, const std::string &base_dir) {
    LoadedScene out;
    const Json &scene = doc.at("scene");
    // tracer: Config [name, props], default ["CPUMulticore", {}] (renderer.hpp:20-21)
    if (const Json *t = doc.find("tracer")) {
        out.tracer = (*t)[0].str();
        if (t->size() > 1) out.tracer_props = (*t)[1];
    }
    // kernel: tracer props "kernel", default ["Radiance", {}] (tracer.hpp:14-17)
    if (const Json *k = out.tracer_props.find("kernel")) {
        out.kernel = (*k)[0].str();
        if (k->size() > 1) out.max_bounce = uint32_t(getf((*k)[1], "maxBounce", 16.f));
    }
    // profiler (renderer.hpp:19-20, kernel.hpp:37-43): present means enabled
    // unless "enable" says otherwise; area is a uint4, so -1 reads as 2^32 - 1
    if (const Json *pf = doc.find("profiler")) {
        out.profiler_enable = true;
        if (const Json *e = pf->find("enable")) out.profiler_enable = e->boolean();
        if (const Json *o = pf->find("output")) out.profiler_output = o->str();
        if (const Json *a = pf->find("area")) {
            if (!a->is_array() || a->size() != 4) throw std::logic_error("profiler area must be [x0, y0, x1, y1]");
            for (int c = 0; c < 4; ++c) out.profiler_area[c] = uint32_t(int64_t((*a)[c].number()));
        }
        if (const Json *pp = pf->find("props"))
            if (const Json *d = pp->find("depth")) out.profiler_depth = uint32_t(d->number());
    }
    std::vector<std::string> names;
    bool imported = false;
    if (const Json *assets = scene.find("assets")) {
        for (const Json &a : assets->items()) {
            const std::string &name = a[0].str();
            Json empty = Json::object();
            const Json &props = a.size() > 1 ? a[1] : empty;
            if (name == "import") {
                if (imported) throw std::logic_error("importing multiple scenes is not allowed");
                imported = true;
                std::string path = props.at("path").str();
                if (!path.empty() && path[0] != '/') {
                    std::ifstream probe(path);
                    if (!probe) path = base_dir + "/" + path;
                }
                std::vector<std::string> mats;
                auto meshes = load_mesh_file(path, mats);
                // assimp's material list comes first; spheres append (scene.cu:52,141-154)
                names = mats;
                for (auto &m : meshes) {
                    uint32_t base_v = uint32_t(out.positions.size() / 3);
                    kr_prim_desc p{};
                    p.type = KR_PRIM_MESH;
                    p.shader = uint32_t(std::find(names.begin(), names.end(), m.material) - names.begin());
                    p.tri_begin = uint32_t(out.indices.size() / 3);
                    p.tri_count = uint32_t(m.indices.size() / 3);
                    out.positions.insert(out.positions.end(), m.positions.begin(), m.positions.end());
                    out.normals.insert(out.normals.end(), m.normals.begin(), m.normals.end());
                    for (uint32_t i : m.indices) out.indices.push_back(base_v + i);
                    out.prims.push_back(p);
                }
                KINFO("scene", "Imported", meshes.size(), "meshes");
            } else if (name == "sphere") {
                kr_prim_desc p{};
                p.type = KR_PRIM_SPHERE;
                f3(props.at("o"), p.center, "o");
                p.radius = float(props.at("r").number());
                const std::string &sn = props.at("shader").str();
                auto it = std::find(names.begin(), names.end(), sn);
                if (it == names.end()) {
                    names.push_back(sn);
                    it = names.end() - 1;
                }
                p.shader = uint32_t(it - names.begin());
                out.prims.push_back(p);
            } else if (name.find("light") != std::string::npos || name.find("Light") != std::string::npos) {
                if (name != "pointLight") throw std::logic_error("unknown light type: " + name);
                // PointLight ctor (src/core/light/pointLight.hpp:12-17)
                kr_point_light pl{};
                f3(props.at("position"), pl.position, "position");
                const float one[3] = {1, 1, 1};
                float col[3];
                get3(props, "color", one, col);
                float strength = getf(props, "strength", 2.f);
                for (int k = 0; k < 3; ++k) pl.emission[k] = col[k] * strength;
                out.point_lights.push_back(pl);
            } else {
                throw std::logic_error("unknown asset type: " + name);
            }
        }
    }
    if (const Json *cams = scene.find("camera"))
        for (const Json &c : cams->items()) out.cameras.push_back(camera_from_json(c));
    // shaders by name (scene.cu:180-198); a missing config disables light
    // promotion (scene.cu:193-200)
    out.shader_names = names;
    out.shaders.resize(names.size());
    out.shader_defined.assign(names.size(), false);
    const Json *table = scene.find("shaders");
    for (size_t i = 0; i < names.size(); ++i) {
        KLOG("Looking for shader:", names[i]);
        const Json *cfg = table ? table->find(names[i]) : nullptr;
        if (cfg) {
            fill_shader(*cfg, out.shaders, i, 0);  // Mixed children go after the named shaders
            out.shader_defined[i] = true;
            KLOG(" =", i);
        } else {
            KLOG("Configuration for shader {", names[i], "} not found.");
            out.lights_enabled = false;
            out.shaders[i] = kr_shader_desc{};
            out.shaders[i].type = KR_SHADER_LAMBERT;
        }
    }
    for (auto &p : out.prims)
        if (!out.shader_defined[p.shader])
            throw std::logic_error("primitive uses shader {" + names[p.shader] + "}, which has no configuration");
    return out;
}

LoadedScene load_scene_file(const std::string &path) {
    Json doc = Json::parse(read_file(path));
    return load_scene_json(doc, dir_of(path));
}

kr_scene_desc LoadedScene::desc(uint32_t bvh_width, uint32_t leaf_size) const {
    kr_scene_desc d{};
    d.positions = positions.data();
    d.normals = normals.data();
    d.num_vertices = uint32_t(positions.size() / 3);
    d.indices = indices.data();
    d.num_triangles = uint32_t(indices.size() / 3);
    d.prims = prims.data();
    d.num_prims = uint32_t(prims.size());
    d.shaders = shaders.data();
    d.num_shaders = uint32_t(shaders.size());
    d.light_seed = light_seed;
    d.bvh_width = bvh_width;
    d.leaf_size = leaf_size;
    d.flags = lights_enabled ? 0u : KR_SCENE_NO_AREA_LIGHTS;
    d.point_lights = point_lights.empty() ? nullptr : point_lights.data();
    d.num_point_lights = uint32_t(point_lights.size());
    return d;
}

}  // namespace koishi
