// host/io.cpp - JSON, OBJ/MTL, PLY, PNG and PFM.
#include "host/io.h"
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>

namespace cwp {

// ------------------------------------------------------------------------------------------------------------- JSON
const Json* Json::get(const std::string& key) const {
    if (type != Object) return nullptr;
    for (const auto& kv : obj) if (kv.first == key) return &kv.second;
    return nullptr;
}
double Json::number(const std::string& key, double def) const { const Json* v = get(key); return v && v->type == Number ? v->num : def; }
std::string Json::string(const std::string& key, const std::string& def) const { const Json* v = get(key); return v && v->type == String ? v->str : def; }
static void flattenNumbers(const Json& v, std::vector<double>& out) {
    if (v.type == Json::Number) out.push_back(v.num);
    else if (v.type == Json::Array) for (const Json& e : v.arr) flattenNumbers(e, out);
}
std::vector<double> Json::numbers(const std::string& key) const {
    std::vector<double> r; const Json* v = get(key);
    if (v && v->type == Json::Array) flattenNumbers(*v, r);
    return r;
}

namespace {
struct Parser {
    const std::string& s; size_t i = 0; std::string err;
    explicit Parser(const std::string& t) : s(t) {}
    void ws() { while (i < s.size() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\n' || s[i] == '\r')) i++; }
    bool fail(const char* m) { if (err.empty()) err = std::string(m) + " at offset " + std::to_string(i); return false; }
    static void utf8(std::string& out, unsigned cp) {
        if (cp < 0x80) out += (char)cp;
        else if (cp < 0x800) { out += (char)(0xC0 | (cp >> 6)); out += (char)(0x80 | (cp & 63)); }
        else { out += (char)(0xE0 | (cp >> 12)); out += (char)(0x80 | ((cp >> 6) & 63)); out += (char)(0x80 | (cp & 63)); }
    }
    bool string(std::string& out) {
        if (i >= s.size() || s[i] != '"') return fail("expected string");
        i++;
        while (i < s.size() && s[i] != '"') {
            char c = s[i++];
            if (c != '\\') { out += c; continue; }
            if (i >= s.size()) break;
            char e = s[i++];
            switch (e) {
                case 'n': out += '\n'; break; case 't': out += '\t'; break; case 'r': out += '\r'; break;
                case 'b': out += '\b'; break; case 'f': out += '\f'; break;
                case 'u': if (i + 4 <= s.size()) { utf8(out, (unsigned)strtoul(s.substr(i, 4).c_str(), nullptr, 16)); i += 4; } break;
                default: out += e;
            }
        }
        if (i >= s.size()) return fail("unterminated string");
        i++;
        return true;
    }
    bool value(Json& v, int depth) {
        if (depth > 200) return fail("nesting too deep");
        ws();
        if (i >= s.size()) return fail("unexpected end");
        char c = s[i];
        if (c == '{') {
            v.type = Json::Object; i++; ws();
            if (i < s.size() && s[i] == '}') { i++; return true; }
            for (;;) {
                ws();
                std::pair<std::string, Json> kv;
                if (!string(kv.first)) return fail("expected key");
                ws();
                if (i >= s.size() || s[i] != ':') return fail("expected ':'");
                i++;
                if (!value(kv.second, depth + 1)) return false;
                v.obj.push_back(std::move(kv));
                ws();
                if (i < s.size() && s[i] == ',') { i++; continue; }
                if (i < s.size() && s[i] == '}') { i++; return true; }
                return fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            v.type = Json::Array; i++; ws();
            if (i < s.size() && s[i] == ']') { i++; return true; }
            for (;;) {
                Json e;
                if (!value(e, depth + 1)) return false;
                v.arr.push_back(std::move(e));
                ws();
                if (i < s.size() && s[i] == ',') { i++; continue; }
                if (i < s.size() && s[i] == ']') { i++; return true; }
                return fail("expected ',' or ']'");
            }
        }
        if (c == '"') { v.type = Json::String; return string(v.str); }
        if (s.compare(i, 4, "true") == 0) { v.type = Json::Bool; v.b = true; i += 4; return true; }
        if (s.compare(i, 5, "false") == 0) { v.type = Json::Bool; v.b = false; i += 5; return true; }
        if (s.compare(i, 4, "null") == 0) { v.type = Json::Null; i += 4; return true; }
        char* e; double x = strtod(s.c_str() + i, &e);
        if (e == s.c_str() + i) return fail("unexpected character");
        v.type = Json::Number; v.num = x; i = (size_t)(e - s.c_str());
        return true;
    }
};
}  // namespace

bool parseJson(const std::string& text, Json& out, std::string* err) {
    Parser p(text); out = Json();
    bool ok = p.value(out, 0);
    if (ok) { p.ws(); if (p.i != text.size()) ok = p.fail("trailing characters"); }
    if (!ok && err) *err = p.err;
    return ok;
}

std::string readTextFile(const std::string& path) {
    FILE* fp = fopen(path.c_str(), "rb");
    if (!fp) return "";
    std::string s; char buf[65536]; size_t n;
    while ((n = fread(buf, 1, sizeof buf, fp)) > 0) s.append(buf, n);
    fclose(fp);
    return s;
}

const Json* findKey(const Json& d, const std::string& key) {
    if (d.type == Json::Object)
        for (const auto& kv : d.obj) {
            if (kv.first == key) return &kv.second;
            if (const Json* r = findKey(kv.second, key)) return r;
        }
    if (d.type == Json::Array)
        for (const Json& e : d.arr) if (const Json* r = findKey(e, key)) return r;
    return nullptr;
}
std::vector<double> keyNumbers(const Json& d, const std::string& key) {
    std::vector<double> r; const Json* v = findKey(d, key);
    if (v) flattenNumbers(*v, r);
    return r;
}
static bool keyNumber(const Json& d, const std::string& key, double& out) {
    const Json* v = findKey(d, key);
    if (!v || v->type != Json::Number) return false;
    out = v->num; return true;
}

bool loadLightingFile(const std::string& path, Lighting& L, Mesh* m, std::string& err) {
    std::string text = readTextFile(path);
    if (text.empty()) { err = "lighting file " + path + " unreadable"; return false; }
    Json J; std::string perr;
    if (!parseJson(text, J, &perr)) { err = "lighting file " + path + ": " + perr; return false; }
    std::vector<double> v = keyNumbers(J, "sun_dir_to_light_world_zup");
    if (v.size() == 3) L.sunTo = v3((float)v[0], (float)v[1], (float)v[2]);
    double x;
    if (keyNumber(J, "sun_irradiance_W_m2", x)) L.sunE = v3((float)x, (float)x, (float)x);
    if (keyNumber(J, "sun_angle_rad", x)) L.sunAngleDeg = (float)(x * 180.0 / M_PI);
    v = keyNumbers(J, "sky_radiance_rgb_linear");
    if (v.size() == 3) L.sky = v3((float)v[0], (float)v[1], (float)v[2]);
    // environment map keys (reference: src/io/scene_files.cpp:22); a relative path is relative to the file
    std::string env = J.string("env_map", "");
    if (!env.empty()) {
        const size_t slash = path.find_last_of('/');
        L.envMap = env[0] == '/' || slash == std::string::npos ? env : path.substr(0, slash + 1) + env;
    }
    L.envStrength = (float)J.number("env_strength", L.envStrength);
    L.envRotDeg = (float)J.number("env_rotation_deg", L.envRotDeg);
    if (m) {
        v = keyNumbers(J, "albedo_rgb_linear_by_material_index");
        if (v.size() >= 3) { m->kd.clear(); for (size_t j = 0; j + 2 < v.size(); j += 3) m->kd.push_back(v3((float)v[j], (float)v[j + 1], (float)v[j + 2])); }
        int maxMat = 0; for (int t : m->fm) maxMat = t > maxMat ? t : maxMat;
        while ((int)m->kd.size() <= maxMat) m->kd.push_back(v3(0.8f, 0.8f, 0.8f));
    }
    return true;
}

static Vec3 vec3Or(const Json& o, const char* key, Vec3 def) {
    const std::vector<double> v = o.numbers(key);
    return v.size() == 3 ? v3((float)v[0], (float)v[1], (float)v[2]) : def;
}

bool loadMaterialsFile(const std::string& path, Mesh& m, std::string& err) {
    const std::string text = readTextFile(path);
    if (text.empty()) { err = "materials file " + path + " unreadable"; return false; }
    Json J; std::string perr;
    if (!parseJson(text, J, &perr)) { err = "materials file " + path + ": " + perr; return false; }
    const Json* arr = J.get("materials");
    if (!arr || arr->type != Json::Array) { err = "materials file " + path + ": no \"materials\" array"; return false; }
    std::vector<Material> mats;
    std::vector<bool> given;
    int next = 0;
    for (const Json& o : arr->arr) {
        Material x = defaultMaterial();
        x.baseColor = vec3Or(o, "base_color", x.baseColor);
        x.metallic = (float)o.number("metallic", x.metallic);
        x.roughness = (float)o.number("roughness", x.roughness);
        x.ior = (float)o.number("ior", x.ior);
        x.specularWeight = (float)o.number("specular_weight", x.specularWeight);
        x.specularColor = vec3Or(o, "specular_color", x.specularColor);
        x.transmission = (float)o.number("transmission", x.transmission);
        x.emission = vec3Or(o, "emission_color", v3(1.f, 1.f, 1.f)) * (float)o.number("emission_strength", 0.0);
        if (!(x.ior > 0.f) || x.metallic < 0.f || x.metallic > 1.f || x.transmission < 0.f || x.transmission > 1.f ||
            x.roughness < 0.f || x.roughness > 1.f) {
            err = "materials file " + path + ": entry " + std::to_string(next) + " out of range (ior > 0; metallic, transmission, roughness in [0, 1])";
            return false;
        }
        int id = next++;
        const std::string name = o.string("name", "");
        for (size_t j = 0; !name.empty() && j < m.matName.size(); j++)
            if (m.matName[j] == name) { id = (int)j; break; }
        if ((int)mats.size() <= id) { mats.resize(id + 1, lambertMaterial(v3(0.8f, 0.8f, 0.8f))); given.resize(id + 1, false); }
        mats[id] = x; given[id] = true;
    }
    for (size_t j = 0; j < mats.size(); j++)
        if (!given[j]) mats[j] = lambertMaterial(j < m.kd.size() ? m.kd[j] : v3(0.8f, 0.8f, 0.8f));
    m.mats = mats;
    return true;
}

bool loadCameraFile(const std::string& path, const std::string& key, int frame, int w, int h, float& fx, float& fy,
                    float& hfov, float c2w[16], float& apertureR, float& focusD, std::string& err) {
    std::string text = readTextFile(path);
    if (text.empty()) { err = "camera file " + path + " unreadable"; return false; }
    Json J; std::string perr;
    if (!parseJson(text, J, &perr)) { err = "camera file " + path + ": " + perr; return false; }
    double v;
    if (fx == 0 && keyNumber(J, "fx", v)) {
        double fw = 0, fh = 0; keyNumber(J, "width", fw); keyNumber(J, "height", fh);
        if ((int)fw == w && (int)fh == h) { fx = (float)v; if (keyNumber(J, "fy", v)) fy = (float)v; }
        else if (hfov == 0 && keyNumber(J, "hfov_deg", v)) hfov = (float)v;
    }
    std::vector<double> nums = keyNumbers(J, key);
    if (nums.size() < 16) { err = "no matrix " + key + " in camera file"; return false; }
    size_t off = nums.size() == 16 ? 0 : (size_t)frame * 16;
    if (off + 16 > nums.size()) { err = "camera frame " + std::to_string(frame) + " out of range"; return false; }
    for (int j = 0; j < 16; j++) c2w[j] = (float)nums[off + j];
    if (apertureR < 0.f && keyNumber(J, "aperture_radius", v)) apertureR = (float)v;
    if (focusD < 0.f && keyNumber(J, "focus_distance", v)) focusD = (float)v;
    return true;
}

// ------------------------------------------------------------------------------------------------------------- OBJ
static std::string dirOf(const std::string& p) {
    size_t k = p.find_last_of("/\\");
    return k == std::string::npos ? std::string(".") : p.substr(0, k);
}
static void loadMTL(const std::string& path, Mesh& m, std::map<std::string, int>& ids) {
    FILE* fp = fopen(path.c_str(), "rb");
    if (!fp) { fprintf(stderr, "warn: mtl %s not found\n", path.c_str()); return; }
    char line[1024]; int cur = -1;
    while (fgets(line, sizeof line, fp)) {
        if (strncmp(line, "newmtl", 6) == 0) {
            char name[512];
            if (sscanf(line + 6, "%511s", name) == 1) {
                cur = (int)m.kd.size(); ids[name] = cur; m.kd.push_back(v3(0.8f, 0.8f, 0.8f));
                m.matName.resize(m.kd.size()); m.matName[cur] = name;
            }
        } else if (cur >= 0 && strncmp(line, "Kd", 2) == 0) {
            float r, g, b;
            if (sscanf(line + 2, "%f %f %f", &r, &g, &b) == 3) m.kd[cur] = v3(r, g, b);
        }
    }
    fclose(fp);
}

// OBJ faces are fanned from their first corner; "v/t/n" normals are scattered onto the vertices; usemtl selects an
// MTL material by name (unknown names -> material 0) (reference: src/io/mesh_loader.cpp:14-62).
bool loadOBJ(const std::string& path, Mesh& m) {
    FILE* fp = fopen(path.c_str(), "rb");
    if (!fp) return false;
    std::vector<Vec3> vn; std::map<std::string, int> ids; int curMat = 0;
    if (m.kd.empty()) { m.kd.push_back(v3(0.8f, 0.8f, 0.8f)); m.matName.push_back(""); }
    std::vector<char> line(1 << 16);
    std::vector<int> vi, ni;
    while (fgets(line.data(), (int)line.size(), fp)) {
        char* s = line.data();
        if (s[0] == 'v' && s[1] == ' ') { float x, y, z; if (sscanf(s + 2, "%f %f %f", &x, &y, &z) == 3) m.v.push_back(v3(x, y, z)); }
        else if (s[0] == 'v' && s[1] == 'n') { float x, y, z; if (sscanf(s + 2, "%f %f %f", &x, &y, &z) == 3) vn.push_back(v3(x, y, z)); }
        else if (s[0] == 'f' && (s[1] == ' ' || s[1] == '\t')) {
            vi.clear(); ni.clear();
            char* p = s + 1;
            while (*p) {
                while (*p == ' ' || *p == '\t') p++;
                if (*p == '\n' || *p == '\r' || *p == 0) break;
                int a = atoi(p), b = 0; bool hasNrm = false;
                char* q = p; while (*q && *q != ' ' && *q != '\t' && *q != '\n' && *q != '\r') q++;
                char* sl = (char*)memchr(p, '/', q - p);
                if (sl) { char* sl2 = (char*)memchr(sl + 1, '/', q - sl - 1); if (sl2 && sl2 + 1 < q) { b = atoi(sl2 + 1); hasNrm = true; } }
                vi.push_back(a > 0 ? a - 1 : (int)m.v.size() + a);
                ni.push_back(hasNrm ? (b > 0 ? b - 1 : (int)vn.size() + b) : -1);
                p = q;
            }
            for (size_t k = 1; k + 1 < vi.size(); k++) { m.f.push_back({ vi[0], vi[k], vi[k + 1] }); m.fm.push_back(curMat); }
            for (size_t k = 0; k < vi.size(); k++)
                if (ni[k] >= 0 && ni[k] < (int)vn.size() && vi[k] >= 0 && vi[k] < (int)m.v.size()) {
                    if (m.n.size() != m.v.size()) m.n.resize(m.v.size(), v3(0, 0, 0));
                    m.n[vi[k]] = vn[ni[k]];
                }
        } else if (strncmp(s, "mtllib", 6) == 0) {
            char name[1024]; if (sscanf(s + 6, "%1023s", name) == 1) loadMTL(dirOf(path) + "/" + name, m, ids);
        } else if (strncmp(s, "usemtl", 6) == 0) {
            char name[1024]; curMat = 0;
            if (sscanf(s + 6, "%1023s", name) == 1) { auto it = ids.find(name); if (it != ids.end()) curMat = it->second; }
        }
    }
    fclose(fp);
    if (!m.n.empty()) {
        bool any = false; for (const Vec3& x : m.n) if (length(x) > 0) { any = true; break; }
        if (!any) m.n.clear();
    }
    return !m.f.empty();
}

// ------------------------------------------------------------------------------------------------------------- PLY
namespace {
struct PlyProp { std::string name, type, ltype, itype; bool list = false; };
struct PlyElem { std::string name; long count = 0; std::vector<PlyProp> props; };
int plySize(const std::string& t) {
    if (t == "char" || t == "uchar" || t == "int8" || t == "uint8") return 1;
    if (t == "short" || t == "ushort" || t == "int16" || t == "uint16") return 2;
    if (t == "double" || t == "float64") return 8;
    return 4;
}
double plyBin(FILE* fp, const std::string& t) {
    unsigned char b[8] = { 0 }; int n = plySize(t);
    if (fread(b, 1, n, fp) != (size_t)n) return 0;
    if (t == "char" || t == "int8") return (double)(int8_t)b[0];
    if (t == "uchar" || t == "uint8") return (double)b[0];
    if (t == "short" || t == "int16") { int16_t x; memcpy(&x, b, 2); return x; }
    if (t == "ushort" || t == "uint16") { uint16_t x; memcpy(&x, b, 2); return x; }
    if (t == "int" || t == "int32") { int32_t x; memcpy(&x, b, 4); return x; }
    if (t == "uint" || t == "uint32") { uint32_t x; memcpy(&x, b, 4); return x; }
    if (t == "float" || t == "float32") { float x; memcpy(&x, b, 4); return x; }
    if (t == "double" || t == "float64") { double x; memcpy(&x, b, 8); return x; }
    return 0;
}
}  // namespace

// ASCII or binary little-endian PLY: vertex x y z [nx ny nz], face vertex_indices (fanned) [material_index]
// (reference: src/io/mesh_loader.cpp:64-157).
bool loadPLY(const std::string& path, Mesh& m) {
    FILE* fp = fopen(path.c_str(), "rb");
    if (!fp) return false;
    char line[1024]; bool ascii = false, ok = false; std::vector<PlyElem> elems;
    if (!fgets(line, sizeof line, fp) || strncmp(line, "ply", 3) != 0) { fclose(fp); return false; }
    while (fgets(line, sizeof line, fp)) {
        if (strncmp(line, "format", 6) == 0) {
            ascii = strstr(line, "ascii") != nullptr;
            if (strstr(line, "big_endian")) { fprintf(stderr, "PLY big endian unsupported\n"); fclose(fp); return false; }
        } else if (strncmp(line, "element", 7) == 0) {
            PlyElem e; char name[256]; long c;
            if (sscanf(line + 7, "%255s %ld", name, &c) == 2) { e.name = name; e.count = c; elems.push_back(e); }
        } else if (strncmp(line, "property", 8) == 0 && !elems.empty()) {
            PlyProp p; char a[64], b[64], c[64];
            int got = sscanf(line + 8, "%63s %63s %63s", a, b, c);
            if (got >= 2) {
                if (strcmp(a, "list") == 0) { p.list = true; p.ltype = b; p.itype = got >= 3 ? c : ""; char d[64]; if (sscanf(line + 8, "%*s %*s %*s %63s", d) == 1) p.name = d; }
                else { p.type = a; p.name = b; }
                elems.back().props.push_back(p);
            }
        } else if (strncmp(line, "end_header", 10) == 0) { ok = true; break; }
    }
    if (!ok) { fclose(fp); return false; }
    if (m.kd.empty()) { m.kd.push_back(v3(0.8f, 0.8f, 0.8f)); m.matName.push_back(""); }
    auto rd = [&](const std::string& t) { double x = 0; if (ascii) { if (fscanf(fp, "%lf", &x) != 1) x = 0; } else x = plyBin(fp, t); return x; };
    for (auto& e : elems) {
        if (e.name == "vertex") {
            int ix = -1, iy = -1, iz = -1, inx = -1, iny = -1, inz = -1;
            for (size_t k = 0; k < e.props.size(); k++) {
                const std::string& nm = e.props[k].name;
                if (nm == "x") ix = (int)k; else if (nm == "y") iy = (int)k; else if (nm == "z") iz = (int)k;
                else if (nm == "nx") inx = (int)k; else if (nm == "ny") iny = (int)k; else if (nm == "nz") inz = (int)k;
            }
            if (ix < 0 || iy < 0 || iz < 0) { fclose(fp); return false; }
            bool hasNrm = inx >= 0 && iny >= 0 && inz >= 0;
            std::vector<double> vals(e.props.size());
            for (long i = 0; i < e.count; i++) {
                for (size_t k = 0; k < e.props.size(); k++) {
                    if (e.props[k].list) { int cnt = (int)rd(e.props[k].ltype); for (int j = 0; j < cnt; j++) rd(e.props[k].itype); vals[k] = 0; }
                    else vals[k] = rd(e.props[k].type);
                }
                m.v.push_back(v3((float)vals[ix], (float)vals[iy], (float)vals[iz]));
                if (hasNrm) m.n.push_back(v3((float)vals[inx], (float)vals[iny], (float)vals[inz]));
            }
        } else if (e.name == "face") {
            std::vector<int> idx; int maxMat = 0;
            for (long i = 0; i < e.count; i++) {
                int mat = 0;
                for (const PlyProp& p : e.props) {
                    if (p.list) {
                        int cnt = (int)rd(p.ltype);
                        bool isIdx = p.name == "vertex_indices" || p.name == "vertex_index";
                        if (isIdx) { idx.resize(cnt); for (int j = 0; j < cnt; j++) idx[j] = (int)rd(p.itype); }
                        else for (int j = 0; j < cnt; j++) rd(p.itype);
                    } else { double x = rd(p.type); if (p.name == "material_index") mat = (int)x; }
                }
                maxMat = mat > maxMat ? mat : maxMat;
                for (size_t k = 1; k + 1 < idx.size(); k++) { m.f.push_back({ idx[0], idx[k], idx[k + 1] }); m.fm.push_back(mat); }
            }
            while ((int)m.kd.size() <= maxMat) m.kd.push_back(v3(0.8f, 0.8f, 0.8f));
        } else {
            for (long i = 0; i < e.count; i++)
                for (const PlyProp& p : e.props) {
                    if (p.list) { int cnt = (int)rd(p.ltype); for (int j = 0; j < cnt; j++) rd(p.itype); } else rd(p.type);
                }
        }
    }
    fclose(fp);
    if (!m.n.empty() && m.n.size() != m.v.size()) m.n.clear();
    for (const Tri& t : m.f)
        if (t.a < 0 || t.b < 0 || t.c < 0 || t.a >= (int)m.v.size() || t.b >= (int)m.v.size() || t.c >= (int)m.v.size()) return false;
    return !m.f.empty();
}

bool loadMesh(const std::string& path, Mesh& m) {
    std::string ext = path.substr(path.find_last_of('.') + 1);
    for (auto& ch : ext) ch = (char)tolower((unsigned char)ch);
    bool ok = ext == "ply" ? loadPLY(path, m) : loadOBJ(path, m);
    if (!ok) return false;
    for (const Tri& t : m.f)
        if (t.a < 0 || t.b < 0 || t.c < 0 || t.a >= (int)m.v.size() || t.b >= (int)m.v.size() || t.c >= (int)m.v.size()) return false;
    return true;
}

// ------------------------------------------------------------------------------------------------------------- images
static uint32_t crcTab[256];
static void crcInit() {
    for (uint32_t n = 0; n < 256; n++) { uint32_t c = n; for (int k = 0; k < 8; k++) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1; crcTab[n] = c; }
}
static uint32_t crcOf(const uint8_t* b, size_t n) { uint32_t c = 0xFFFFFFFFu; for (size_t i = 0; i < n; i++) c = crcTab[(c ^ b[i]) & 0xFF] ^ (c >> 8); return c ^ 0xFFFFFFFFu; }
static void be32(std::vector<uint8_t>& v, uint32_t x) { v.push_back(x >> 24); v.push_back(x >> 16); v.push_back(x >> 8); v.push_back(x); }
static void pngChunk(FILE* fp, const char* type, const std::vector<uint8_t>& data) {
    std::vector<uint8_t> b; be32(b, (uint32_t)data.size());
    b.insert(b.end(), type, type + 4); b.insert(b.end(), data.begin(), data.end());
    uint32_t c = crcOf(b.data() + 4, b.size() - 4);
    be32(b, c);
    fwrite(b.data(), 1, b.size(), fp);
}
bool writePNG(const std::string& path, int w, int h, const uint8_t* rgb) {
    static bool init = false; if (!init) { crcInit(); init = true; }
    FILE* fp = fopen(path.c_str(), "wb");
    if (!fp) return false;
    const uint8_t sig[8] = { 137, 80, 78, 71, 13, 10, 26, 10 }; fwrite(sig, 1, 8, fp);
    std::vector<uint8_t> ihdr; be32(ihdr, w); be32(ihdr, h);
    ihdr.push_back(8); ihdr.push_back(2); ihdr.push_back(0); ihdr.push_back(0); ihdr.push_back(0);
    pngChunk(fp, "IHDR", ihdr);
    std::vector<uint8_t> raw; raw.reserve((size_t)h * (w * 3 + 1));
    for (int y = 0; y < h; y++) { raw.push_back(0); raw.insert(raw.end(), rgb + (size_t)y * w * 3, rgb + (size_t)(y + 1) * w * 3); }
    std::vector<uint8_t> z = { 0x78, 0x01 };
    uint32_t a = 1, b = 0;
    for (uint8_t x : raw) { a = (a + x) % 65521; b = (b + a) % 65521; }
    size_t pos = 0;
    do {
        size_t n = raw.size() - pos; if (n > 65535) n = 65535;
        bool last = pos + n == raw.size();
        z.push_back(last ? 1 : 0); z.push_back(n & 0xFF); z.push_back((n >> 8) & 0xFF); z.push_back((~n) & 0xFF); z.push_back(((~n) >> 8) & 0xFF);
        z.insert(z.end(), raw.begin() + pos, raw.begin() + pos + n); pos += n;
    } while (pos < raw.size());
    be32(z, (b << 16) | a);
    pngChunk(fp, "IDAT", z);
    pngChunk(fp, "IEND", std::vector<uint8_t>());
    fclose(fp);
    return true;
}
bool writePFM(const std::string& path, int w, int h, const float* rgb) {
    FILE* fp = fopen(path.c_str(), "wb");
    if (!fp) return false;
    fprintf(fp, "PF\n%d %d\n-1.0\n", w, h);
    for (int y = h - 1; y >= 0; y--) fwrite(rgb + (size_t)y * w * 3, sizeof(float), (size_t)w * 3, fp);
    fclose(fp);
    return true;
}
void tonemapSRGB8(const float* lin, size_t nPix, uint8_t* rgb) {
    for (size_t i = 0; i < nPix * 3; i++) {
        float x = fminf(fmaxf(lin[i], 0.f), 1.f);
        x = x <= 0.0031308f ? 12.92f * x : 1.055f * powf(x, 1.f / 2.4f) - 0.055f;
        rgb[i] = (uint8_t)(x * 255.f + 0.5f);
    }
}

}  // namespace cwp
