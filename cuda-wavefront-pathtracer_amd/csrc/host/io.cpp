// host/io.cpp - JSON, OBJ/MTL, PLY, PNG and PFM.
#include "host/io.h"
#include "host/texture.h"
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cctype>
#include <cstring>
#include <algorithm>
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

// JSON in two stages: a lexer that turns the text into tokens (strings unescaped, numbers converted), and an assembler
// that builds the tree with an explicit stack of the containers still open, so nesting depth costs no native stack.
struct Token {
    enum Kind { OpenObj, CloseObj, OpenArr, CloseArr, Colon, Comma, Str, Num, True, False, Null, End, Bad } kind = Bad;
    size_t at = 0;
    std::string text;
    double num = 0;
};

class JsonLexer {
public:
    explicit JsonLexer(const std::string& s) : s_(s) {}
    Token next() {
        while (pos_ < s_.size() && (s_[pos_] == ' ' || s_[pos_] == '\t' || s_[pos_] == '\n' || s_[pos_] == '\r')) pos_++;
        Token t;
        t.at = pos_;
        if (pos_ >= s_.size()) { t.kind = Token::End; return t; }
        const char c = s_[pos_];
        switch (c) {
            case '{': pos_++; t.kind = Token::OpenObj; return t;
            case '}': pos_++; t.kind = Token::CloseObj; return t;
            case '[': pos_++; t.kind = Token::OpenArr; return t;
            case ']': pos_++; t.kind = Token::CloseArr; return t;
            case ':': pos_++; t.kind = Token::Colon; return t;
            case ',': pos_++; t.kind = Token::Comma; return t;
            case '"': return stringToken(t);
            default: break;
        }
        static const struct { const char* word; Token::Kind kind; } words[] = { { "true", Token::True }, { "false", Token::False },
                                                                                { "null", Token::Null } };
        for (const auto& w : words) {
            const size_t n = strlen(w.word);
            if (s_.compare(pos_, n, w.word) == 0) { pos_ += n; t.kind = w.kind; return t; }
        }
        const char* b = s_.c_str() + pos_;
        char* e = nullptr;
        const double v = strtod(b, &e);
        if (e == b) { t.kind = Token::Bad; t.text = "unexpected character"; return t; }
        pos_ += (size_t)(e - b);
        t.kind = Token::Num; t.num = v;
        return t;
    }

private:
    static void appendUtf8(std::string& out, uint32_t cp) {
        if (cp < 0x80) { out += (char)cp; return; }
        if (cp < 0x800) { out += (char)(0xC0 | (cp >> 6)); out += (char)(0x80 | (cp & 0x3F)); return; }
        out += (char)(0xE0 | (cp >> 12)); out += (char)(0x80 | ((cp >> 6) & 0x3F)); out += (char)(0x80 | (cp & 0x3F));
    }
    Token stringToken(Token& t) {
        pos_++;   // opening quote
        for (;;) {
            if (pos_ >= s_.size()) { t.kind = Token::Bad; t.text = "unterminated string"; return t; }
            const char c = s_[pos_++];
            if (c == '"') break;
            if (c != '\\') { t.text += c; continue; }
            if (pos_ >= s_.size()) { t.kind = Token::Bad; t.text = "unterminated string"; return t; }
            const char e = s_[pos_++];
            const char* from = "nrtbf";
            const char* to = "\n\r\t\b\f";
            if (const char* k = strchr(from, e); k && e) { t.text += to[k - from]; continue; }
            if (e == 'u') {
                uint32_t cp = 0;
                int digits = 0;
                for (; digits < 4 && pos_ < s_.size() && isxdigit((unsigned char)s_[pos_]); digits++, pos_++)
                    cp = cp * 16 + (uint32_t)(isdigit((unsigned char)s_[pos_]) ? s_[pos_] - '0' : (tolower((unsigned char)s_[pos_]) - 'a' + 10));
                if (digits == 4) appendUtf8(t.text, cp);
                continue;
            }
            t.text += e;   // \" \\ \/ and anything else: the character itself
        }
        t.kind = Token::Str;
        return t;
    }
    const std::string& s_;
    size_t pos_ = 0;
};

bool jsonError(std::string* err, const char* what, size_t at) {
    if (err) *err = std::string(what) + " at offset " + std::to_string(at);
    return false;
}

}  // namespace

bool parseJson(const std::string& text, Json& out, std::string* err) {
    out = Json();
    JsonLexer lex(text);
    // each open container, and for an object the key whose value comes next
    struct Open { Json* v; std::string key; };
    std::vector<Open> open;
    bool done = false;                 // the top-level value is complete
    enum Expect { Value, KeyOrClose, Key, ColonNext, CommaOrClose, ValueOrClose } want = Value;
    for (;;) {
        Token t = lex.next();
        if (done) {
            if (t.kind != Token::End) return jsonError(err, "trailing characters", t.at);
            return true;
        }
        if (t.kind == Token::Bad) return jsonError(err, t.text.c_str(), t.at);
        if (t.kind == Token::End) return jsonError(err, "unexpected end", t.at);
        if (open.size() > 200) return jsonError(err, "nesting too deep", t.at);
        // the slot the next value goes to
        auto place = [&]() -> Json* {
            if (open.empty()) return &out;
            Open& o = open.back();
            if (o.v->type == Json::Array) { o.v->arr.emplace_back(); return &o.v->arr.back(); }
            o.v->obj.emplace_back(o.key, Json());
            return &o.v->obj.back().second;
        };
        auto closeOne = [&]() {
            open.pop_back();
            if (open.empty()) done = true;
            want = CommaOrClose;
        };
        switch (want) {
            case KeyOrClose:
            case Key:
                if (want == KeyOrClose && t.kind == Token::CloseObj) { closeOne(); continue; }
                if (t.kind != Token::Str) return jsonError(err, "expected key", t.at);
                open.back().key = t.text;
                want = ColonNext;
                continue;
            case ColonNext:
                if (t.kind != Token::Colon) return jsonError(err, "expected ':'", t.at);
                want = Value;
                continue;
            case CommaOrClose: {
                const bool inObj = open.back().v->type == Json::Object;
                if (t.kind == Token::Comma) { want = inObj ? Key : Value; continue; }
                if (t.kind == (inObj ? Token::CloseObj : Token::CloseArr)) { closeOne(); continue; }
                return jsonError(err, inObj ? "expected ',' or '}'" : "expected ',' or ']'", t.at);
            }
            case ValueOrClose:
                if (t.kind == Token::CloseArr) { closeOne(); continue; }
                [[fallthrough]];
            case Value: {
                Json* v = nullptr;
                switch (t.kind) {
                    case Token::OpenObj: v = place(); v->type = Json::Object; open.push_back({ v }); want = KeyOrClose; continue;
                    case Token::OpenArr: v = place(); v->type = Json::Array; open.push_back({ v }); want = ValueOrClose; continue;
                    case Token::Str: v = place(); v->type = Json::String; v->str = t.text; break;
                    case Token::Num: v = place(); v->type = Json::Number; v->num = t.num; break;
                    case Token::True: case Token::False: v = place(); v->type = Json::Bool; v->b = t.kind == Token::True; break;
                    case Token::Null: v = place(); v->type = Json::Null; break;
                    default: return jsonError(err, "unexpected character", t.at);
                }
                if (open.empty()) done = true; else want = CommaOrClose;
                continue;
            }
        }
    }
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

// ------------------------------------------------------------------------------------------------------ text meshes
namespace {

std::string dirOf(const std::string& p) {
    const size_t k = p.find_last_of("/\\");
    return k == std::string::npos ? std::string(".") : p.substr(0, k);
}

// A text file read line by line, each line split at blanks into words (no word longer than the line).
class WordReader {
public:
    explicit WordReader(const std::string& path) : fp_(fopen(path.c_str(), "rb")) {}
    ~WordReader() { if (fp_) fclose(fp_); }
    bool ok() const { return fp_ != nullptr; }
    // the words of the next line (false at end of file)
    bool line(std::vector<std::string>& words) {
        words.clear();
        std::string cur;
        int c;
        bool any = false;
        while ((c = fgetc(fp_)) != EOF) {
            any = true;
            if (c == '\n') break;
            if (c == ' ' || c == '\t' || c == '\r') { if (!cur.empty()) { words.push_back(cur); cur.clear(); } continue; }
            cur += (char)c;
        }
        if (!cur.empty()) words.push_back(cur);
        return any;
    }
private:
    FILE* fp_;
};

bool parseFloats(const std::vector<std::string>& w, size_t from, float* out, int n) {
    if (w.size() < from + (size_t)n) return false;
    for (int k = 0; k < n; k++) {
        char* e = nullptr;
        out[k] = strtof(w[from + k].c_str(), &e);
        if (e == w[from + k].c_str()) return false;
    }
    return true;
}

// A texture statement of a material: which map, the image file (relative to the MTL), normal strength, channel.
enum MapKind { kMapBase, kMapEmission, kMapNormal, kMapMetallic, kMapRoughness };
struct MapRef { int mat, kind; std::string file; float scale; int channel; };

// map statement "<key> [options] file": options -bm s (normal strength), -imfchan r|g|b|m|l|z (channel of a scalar
// map), and the other standard options, skipped with their arguments; the file name is the rest of the line
bool parseMapStatement(const std::vector<std::string>& w, std::string& file, float& scale, int& channel) {
    size_t k = 1;
    scale = 1.f; channel = -1;
    auto isNum = [](const std::string& s) { char* e = nullptr; strtod(s.c_str(), &e); return e != s.c_str() && *e == 0; };
    while (k < w.size() && w[k].size() > 1 && w[k][0] == '-' && !isNum(w[k])) {
        const std::string& o = w[k++];
        if (o == "-bm" && k < w.size()) scale = (float)atof(w[k++].c_str());
        else if (o == "-imfchan" && k < w.size()) {
            const char c = w[k++][0];
            channel = c == 'g' ? 1 : (c == 'b' ? 2 : (c == 'z' ? 3 : 0));
        } else if (o == "-o" || o == "-s" || o == "-t") { for (int j = 0; j < 3 && k < w.size() && isNum(w[k]); j++) k++; }
        else if (o == "-mm") k += 2;
        else k++;   // -blendu, -blendv, -clamp, -cc, -boost, -texres: one argument
    }
    if (k >= w.size()) return false;
    file = w[k];
    for (size_t j = k + 1; j < w.size(); j++) file += " " + w[j];
    return true;
}

// MTL: "newmtl name" opens a material (diffuse 0.8 until a "Kd r g b" line sets it; "Ke r g b" gives it an emission
// colour); texture maps map_Kd (base colour), map_Ke (emission), map_Bump / bump / norm / map_Kn (tangent-space normal
// map), map_Pr (roughness) and map_Pm (metalness)
void readMTL(const std::string& path, Mesh& m, std::map<std::string, int>& byName, std::vector<MapRef>& refs,
             std::vector<Vec3>& ke) {
    WordReader in(path);
    if (!in.ok()) { fprintf(stderr, "warn: mtl %s not found\n", path.c_str()); return; }
    std::vector<std::string> w;
    int cur = -1;
    while (in.line(w)) {
        if (w.empty()) continue;
        if (w[0] == "newmtl" && w.size() >= 2) {
            cur = (int)m.kd.size();
            byName[w[1]] = cur;
            m.kd.push_back(v3(0.8f, 0.8f, 0.8f));
            m.matName.resize(m.kd.size());
            m.matName[cur] = w[1];
        } else if (w[0] == "Kd" && cur >= 0) {
            float c[3];
            if (parseFloats(w, 1, c, 3)) m.kd[cur] = v3(c[0], c[1], c[2]);
        } else if (w[0] == "Ke" && cur >= 0) {
            float c[3];
            if (parseFloats(w, 1, c, 3)) { ke.resize(m.kd.size(), v3(0.f, 0.f, 0.f)); ke[cur] = v3(c[0], c[1], c[2]); }
        } else if (cur >= 0) {
            const std::string& k = w[0];
            const int kind = k == "map_Kd" ? kMapBase : k == "map_Ke" ? kMapEmission :
                             (k == "map_Bump" || k == "map_bump" || k == "bump" || k == "norm" || k == "map_Kn") ? kMapNormal :
                             k == "map_Pm" ? kMapMetallic : k == "map_Pr" ? kMapRoughness : -1;
            MapRef r;
            if (kind >= 0 && parseMapStatement(w, r.file, r.scale, r.channel)) {
                r.mat = cur; r.kind = kind;
                r.file = dirOf(path) + "/" + r.file;
                refs.push_back(r);
            }
        }
    }
}

// Loads the images the map statements name (one image per (file, decode); an unreadable image leaves its map off,
// with a warning, as the reference does) and sets the materials' bindings.
void bindMaps(Mesh& m, const std::vector<MapRef>& refs) {
    if (refs.empty()) return;
    m.maps.assign(m.kd.size(), noMaps());
    std::map<std::pair<std::string, bool>, int> loaded;
    for (const MapRef& r : refs) {
        const bool srgb = r.kind == kMapBase || r.kind == kMapEmission;
        const auto key = std::make_pair(r.file, srgb);
        auto it = loaded.find(key);
        if (it == loaded.end()) {
            TextureImage img;
            std::string err;
            int id = -1;
            if (readImage(r.file, img, err)) { img.srgb = srgb; id = (int)m.tex.size(); m.tex.push_back(std::move(img)); }
            else fprintf(stderr, "warn: texture %s\n", err.c_str());
            it = loaded.emplace(key, id).first;
        }
        MaterialMaps& x = m.maps[r.mat];
        const int id = it->second;
        switch (r.kind) {
            case kMapBase: x.baseColor = id; break;
            case kMapEmission: x.emission = id; break;
            case kMapNormal: x.normal = id; x.normalScale = r.scale; break;
            case kMapMetallic: x.metallic = id; if (r.channel >= 0) x.metallicChannel = r.channel; break;
            default: x.roughness = id; if (r.channel >= 0) x.roughnessChannel = r.channel; break;
        }
    }
    bool any = false;
    for (const MaterialMaps& x : m.maps) any = any || hasMaps(x);
    if (!any) { m.maps.clear(); m.tex.clear(); }
}

// one OBJ index (1-based, or negative: counted back from the end of the list so far); 0 or garbage gives an index
// outside the list, which the mesh check rejects
int objIndex(const std::string& s, int listSize) {
    const int k = atoi(s.c_str());
    return k > 0 ? k - 1 : listSize + k;
}

}  // namespace

// OBJ: "v x y z", "vt u v", "vn x y z", faces "f a b c ..." whose corners are "v", "v/t", "v//n" or "v/t/n" (fanned
// from the first corner into triangles), "mtllib file" (relative to the OBJ), "usemtl name" (unknown names: material
// 0). A corner's normal is stored on its vertex; a mesh whose normals are all zero has none (reference behaviour:
// src/io/mesh_loader.cpp:14-62). Texture coordinates are kept per triangle corner (corners without one get (0, 0));
// the MTL's texture maps are bound to their materials.
bool loadOBJ(const std::string& path, Mesh& m) {
    WordReader in(path);
    if (!in.ok()) return false;
    if (m.kd.empty()) { m.kd.push_back(v3(0.8f, 0.8f, 0.8f)); m.matName.push_back(""); }
    std::vector<Vec3> normals;
    std::vector<float> texcoords;   // 2 per vt
    std::map<std::string, int> byName;
    std::vector<MapRef> refs;
    std::vector<Vec3> ke;   // MTL emission colours per material id
    int material = 0;
    bool anyUV = false;
    std::vector<std::string> w;
    std::vector<int> cv, cn, ct;   // corners of the current face: vertex, normal, texture coordinate (-1: none)
    std::vector<float> fuv;        // corner uvs of every triangle so far
    while (in.line(w)) {
        if (w.empty()) continue;
        const std::string& k = w[0];
        float x[3];
        if (k == "v") { if (parseFloats(w, 1, x, 3)) m.v.push_back(v3(x[0], x[1], x[2])); }
        else if (k == "vn") { if (parseFloats(w, 1, x, 3)) normals.push_back(v3(x[0], x[1], x[2])); }
        else if (k == "vt") { if (parseFloats(w, 1, x, 2)) { texcoords.push_back(x[0]); texcoords.push_back(x[1]); } }
        else if (k == "f") {
            cv.clear(); cn.clear(); ct.clear();
            for (size_t j = 1; j < w.size(); j++) {
                const std::string& c = w[j];
                const size_t s1 = c.find('/'), s2 = s1 == std::string::npos ? s1 : c.find('/', s1 + 1);
                cv.push_back(objIndex(c, (int)m.v.size()));
                const bool hasN = s2 != std::string::npos && s2 + 1 < c.size();
                cn.push_back(hasN ? objIndex(c.substr(s2 + 1), (int)normals.size()) : -1);
                const bool hasT = s1 != std::string::npos && s1 + 1 < c.size() && c[s1 + 1] != '/';
                const int t = hasT ? objIndex(c.substr(s1 + 1), (int)texcoords.size() / 2) : -1;
                ct.push_back(t >= 0 && t < (int)texcoords.size() / 2 ? t : -1);
                anyUV = anyUV || ct.back() >= 0;
            }
            auto cornerUV = [&](size_t j) {
                fuv.push_back(ct[j] >= 0 ? texcoords[2 * ct[j]] : 0.f);
                fuv.push_back(ct[j] >= 0 ? texcoords[2 * ct[j] + 1] : 0.f);
            };
            for (size_t j = 2; j < cv.size(); j++) {
                m.f.push_back({ cv[0], cv[j - 1], cv[j] }); m.fm.push_back(material);
                cornerUV(0); cornerUV(j - 1); cornerUV(j);
            }
            for (size_t j = 0; j < cv.size(); j++) {
                if (cn[j] < 0 || cn[j] >= (int)normals.size() || cv[j] < 0 || cv[j] >= (int)m.v.size()) continue;
                if (m.n.size() != m.v.size()) m.n.resize(m.v.size(), v3(0.f, 0.f, 0.f));
                m.n[cv[j]] = normals[cn[j]];
            }
        } else if (k == "mtllib" && w.size() >= 2) readMTL(dirOf(path) + "/" + w[1], m, byName, refs, ke);
        else if (k == "usemtl") {
            const auto it = w.size() >= 2 ? byName.find(w[1]) : byName.end();
            material = it == byName.end() ? 0 : it->second;
        }
    }
    bool anyNormal = false;
    for (const Vec3& x : m.n) anyNormal = anyNormal || length(x) > 0.f;
    if (!anyNormal) m.n.clear();
    if (anyUV) m.uv.swap(fuv);
    bool emissive = false;
    for (const Vec3& e : ke) emissive = emissive || e.x > 0.f || e.y > 0.f || e.z > 0.f;
    if (emissive) {   // Lambertian materials with their emission colours
        m.mats.clear();
        for (size_t j = 0; j < m.kd.size(); j++) {
            Material x = lambertMaterial(m.kd[j]);
            if (j < ke.size()) x.emission = ke[j];
            m.mats.push_back(x);
        }
    }
    bindMaps(m, refs);
    return !m.f.empty();
}

// ------------------------------------------------------------------------------------------------------------- PLY
namespace {

enum class PlyType { I8, U8, I16, U16, I32, U32, F32, F64 };
bool plyType(const std::string& t, PlyType& out) {
    static const struct { const char* a; const char* b; PlyType t; } names[] = {
        { "char", "int8", PlyType::I8 }, { "uchar", "uint8", PlyType::U8 }, { "short", "int16", PlyType::I16 },
        { "ushort", "uint16", PlyType::U16 }, { "int", "int32", PlyType::I32 }, { "uint", "uint32", PlyType::U32 },
        { "float", "float32", PlyType::F32 }, { "double", "float64", PlyType::F64 } };
    for (const auto& n : names) if (t == n.a || t == n.b) { out = n.t; return true; }
    return false;
}

struct PlyProperty { std::string name; bool list = false; PlyType count = PlyType::U8, value = PlyType::F32; };
struct PlyElement { std::string name; long n = 0; std::vector<PlyProperty> props; };

// scalar values from the body of a PLY file, ASCII words or little-endian binary
class PlyBody {
public:
    PlyBody(FILE* fp, bool ascii) : fp_(fp), ascii_(ascii) {}
    double read(PlyType t) {
        if (ascii_) { double x = 0; return fscanf(fp_, "%lf", &x) == 1 ? x : 0.0; }
        unsigned char b[8] = { 0, 0, 0, 0, 0, 0, 0, 0 };
        static const int bytes[] = { 1, 1, 2, 2, 4, 4, 4, 8 };
        const int n = bytes[(int)t];
        if (fread(b, 1, (size_t)n, fp_) != (size_t)n) return 0.0;
        switch (t) {
            case PlyType::I8: return (double)(int8_t)b[0];
            case PlyType::U8: return (double)b[0];
            case PlyType::I16: { int16_t x; memcpy(&x, b, 2); return x; }
            case PlyType::U16: { uint16_t x; memcpy(&x, b, 2); return x; }
            case PlyType::I32: { int32_t x; memcpy(&x, b, 4); return x; }
            case PlyType::U32: { uint32_t x; memcpy(&x, b, 4); return x; }
            case PlyType::F32: { float x; memcpy(&x, b, 4); return x; }
            case PlyType::F64: { double x; memcpy(&x, b, 8); return x; }
        }
        return 0.0;
    }
    // a list property: its values (or nothing when `keep` is null)
    void list(const PlyProperty& p, std::vector<double>* keep) {
        const long n = (long)read(p.count);
        if (keep) keep->clear();
        for (long k = 0; k < n; k++) { const double x = read(p.value); if (keep) keep->push_back(x); }
    }
private:
    FILE* fp_;
    bool ascii_;
};

}  // namespace

// PLY, ASCII or binary little-endian: vertex x y z [nx ny nz]; face vertex_indices (or vertex_index), fanned from the
// first corner, with an optional material_index; other elements are skipped (reference behaviour:
// src/io/mesh_loader.cpp:64-157).
bool loadPLY(const std::string& path, Mesh& m) {
    FILE* fp = fopen(path.c_str(), "rb");
    if (!fp) return false;
    std::vector<PlyElement> elems;
    bool ascii = false, header = false;
    {
        char buf[1024];
        if (!fgets(buf, sizeof buf, fp) || strncmp(buf, "ply", 3) != 0) { fclose(fp); return false; }
        while (fgets(buf, sizeof buf, fp)) {
            std::vector<std::string> w;
            for (char* tok = strtok(buf, " \t\r\n"); tok; tok = strtok(nullptr, " \t\r\n")) w.push_back(tok);
            if (w.empty()) continue;
            if (w[0] == "end_header") { header = true; break; }
            if (w[0] == "format" && w.size() >= 2) {
                if (w[1] == "binary_big_endian") { fprintf(stderr, "PLY big endian unsupported\n"); fclose(fp); return false; }
                ascii = w[1] == "ascii";
            } else if (w[0] == "element" && w.size() >= 3) {
                PlyElement e; e.name = w[1]; e.n = atol(w[2].c_str());
                elems.push_back(e);
            } else if (w[0] == "property" && !elems.empty()) {
                PlyProperty p;
                if (w.size() >= 5 && w[1] == "list") {
                    p.list = true;
                    if (!plyType(w[2], p.count) || !plyType(w[3], p.value)) { fclose(fp); return false; }
                    p.name = w[4];
                } else if (w.size() >= 3) {
                    if (!plyType(w[1], p.value)) { fclose(fp); return false; }
                    p.name = w[2];
                } else continue;
                elems.back().props.push_back(p);
            }
        }
    }
    if (!header) { fclose(fp); return false; }
    if (m.kd.empty()) { m.kd.push_back(v3(0.8f, 0.8f, 0.8f)); m.matName.push_back(""); }
    PlyBody body(fp, ascii);
    std::vector<double> vals, idx;
    for (const PlyElement& e : elems) {
        auto column = [&](const char* name) {
            for (size_t k = 0; k < e.props.size(); k++) if (!e.props[k].list && e.props[k].name == name) return (int)k;
            return -1;
        };
        if (e.name == "vertex") {
            const int cx = column("x"), cy = column("y"), cz = column("z"), nx = column("nx"), ny = column("ny"), nz = column("nz");
            if (cx < 0 || cy < 0 || cz < 0) { fclose(fp); return false; }
            const bool withN = nx >= 0 && ny >= 0 && nz >= 0;
            vals.assign(e.props.size(), 0.0);
            for (long r = 0; r < e.n; r++) {
                for (size_t k = 0; k < e.props.size(); k++) {
                    if (e.props[k].list) { body.list(e.props[k], nullptr); vals[k] = 0.0; }
                    else vals[k] = body.read(e.props[k].value);
                }
                m.v.push_back(v3((float)vals[cx], (float)vals[cy], (float)vals[cz]));
                if (withN) m.n.push_back(v3((float)vals[nx], (float)vals[ny], (float)vals[nz]));
            }
        } else if (e.name == "face") {
            int topMaterial = 0;
            for (long r = 0; r < e.n; r++) {
                int material = 0;
                for (const PlyProperty& p : e.props) {
                    if (p.list) body.list(p, p.name == "vertex_indices" || p.name == "vertex_index" ? &idx : nullptr);
                    else { const double x = body.read(p.value); if (p.name == "material_index") material = (int)x; }
                }
                topMaterial = std::max(topMaterial, material);
                for (size_t j = 2; j < idx.size(); j++) { m.f.push_back({ (int)idx[0], (int)idx[j - 1], (int)idx[j] }); m.fm.push_back(material); }
            }
            while ((int)m.kd.size() <= topMaterial) m.kd.push_back(v3(0.8f, 0.8f, 0.8f));
        } else {
            for (long r = 0; r < e.n; r++)
                for (const PlyProperty& p : e.props) { if (p.list) body.list(p, nullptr); else body.read(p.value); }
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
uint8_t encodeSRGB8(float x) {
    x = fminf(fmaxf(x, 0.f), 1.f);
    x = x <= 0.0031308f ? 12.92f * x : 1.055f * powf(x, 1.f / 2.4f) - 0.055f;
    return (uint8_t)(x * 255.f + 0.5f);
}
void tonemapSRGB8(const float* lin, size_t nPix, uint8_t* rgb) {
    for (size_t i = 0; i < nPix * 3; i++) rgb[i] = encodeSRGB8(lin[i]);
}
void srgbSteps(float t[256]) {
    t[0] = -INFINITY;
    uint32_t lo = 0;                       // bits of 0.0f: the search range of each step starts at the previous one
    const uint32_t one = 0x3F800000u;      // bits of 1.0f, which encodes to 255
    for (int k = 1; k < 256; k++) {
        uint32_t a = lo, b = one;          // invariant: encode(bits b) >= k
        while (a < b) {
            const uint32_t m = a + (b - a) / 2;
            float x; memcpy(&x, &m, 4);
            if (encodeSRGB8(x) >= k) b = m; else a = m + 1;
        }
        memcpy(&t[k], &a, 4);
        lo = a;
    }
}

}  // namespace cwp
