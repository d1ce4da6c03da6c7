// host/instances.cpp - scene files with instances, per-frame transforms, the generated two-level scene.
#include "host/instances.h"
#include "host/io.h"
#include <cmath>

namespace cwp {

namespace {

bool readXform(const std::vector<double>& v, Xform& X) {
    if (v.size() != 12 && v.size() != 16) return false;
    for (int i = 0; i < 12; i++) X.m[i] = (float)v[i];
    return true;
}

Vec3 vec3Of(const Json& o, const char* key, Vec3 def, bool allowScalar = false) {
    const Json* v = o.get(key);
    if (!v) return def;
    if (allowScalar && v->type == Json::Number) return v3((float)v->num, (float)v->num, (float)v->num);
    const std::vector<double> a = o.numbers(key);
    return a.size() == 3 ? v3((float)a[0], (float)a[1], (float)a[2]) : def;
}

// rotation about a unit axis (Rodrigues) times a per-column scale, translation t; all in double, rounded once
Xform composeXform(const double t[3], const double axis[3], double ang, const double s[3]) {
    double ax = axis[0], ay = axis[1], az = axis[2];
    const double l = std::sqrt(ax * ax + ay * ay + az * az);
    if (l > 0) { ax /= l; ay /= l; az /= l; } else { ax = 0; ay = 0; az = 1; }
    const double c = std::cos(ang), sn = std::sin(ang), C = 1 - c;
    const double r[3][3] = { { c + ax * ax * C, ax * ay * C - az * sn, ax * az * C + ay * sn },
                             { ay * ax * C + az * sn, c + ay * ay * C, ay * az * C - ax * sn },
                             { az * ax * C - ay * sn, az * ay * C + ax * sn, c + az * az * C } };
    Xform X;
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) X.m[4 * i + j] = (float)(r[i][j] * s[j]);
        X.m[4 * i + 3] = (float)t[i];
    }
    return X;
}

}  // namespace

bool loadSceneFile(const std::string& path, SceneDesc& S, std::string& err) {
    const std::string js = readTextFile(path);
    if (js.empty()) { err = "scene file " + path + " unreadable"; return false; }
    Json J;
    if (!parseJson(js, J, &err)) { err = "scene file " + path + ": " + err; return false; }
    const size_t slash = path.find_last_of('/');
    const std::string dir = slash == std::string::npos ? std::string() : path.substr(0, slash + 1);
    auto rel = [&](const std::string& p) { return p.empty() || p[0] == '/' ? p : dir + p; };
    S = SceneDesc();
    S.frames = (int)J.number("frames", 1);
    if (S.frames < 1) S.frames = 1;
    const Json* ms = J.get("meshes");
    const Json* is = J.get("instances");
    if (!ms || ms->type != Json::Array || !is || is->type != Json::Array) {
        err = "scene file " + path + ": needs \"meshes\" and \"instances\" arrays"; return false;
    }
    for (const Json& m : ms->arr) {
        MeshDesc d; d.file = rel(m.string("file", "")); d.yup = (int)m.number("yup_to_zup", -1);
        const std::string mf = m.string("materials", "");
        if (!mf.empty()) d.materials = rel(mf);
        if (d.file.empty()) { err = "scene file: mesh without \"file\""; return false; }
        S.meshes.push_back(d);
    }
    for (const Json& o : is->arr) {
        InstanceDesc I;
        I.mesh = (int)o.number("mesh", 0);
        if (I.mesh < 0 || I.mesh >= (int)S.meshes.size()) { err = "scene file: instance mesh " + std::to_string(I.mesh) + " out of range"; return false; }
        if (o.get("transform")) {
            if (!readXform(o.numbers("transform"), I.matrix)) { err = "scene file: transform needs 12 or 16 numbers"; return false; }
            I.hasMatrix = true;
        }
        if (const Json* tf = o.get("transforms")) {
            for (const Json& t : tf->arr) {
                std::vector<double> v;
                for (const Json& x : t.arr) v.push_back(x.num);
                Xform X;
                if (!readXform(v, X)) { err = "scene file: transforms entries need 12 or 16 numbers"; return false; }
                I.perFrame.push_back(X);
            }
        }
        I.translate = vec3Of(o, "translate", I.translate); I.translateVel = vec3Of(o, "translate_per_frame", I.translateVel);
        I.axis = vec3Of(o, "axis", I.axis);
        I.angleDeg = (float)o.number("angle_deg", 0.0); I.angleVel = (float)o.number("angle_deg_per_frame", 0.0);
        I.scale = vec3Of(o, "scale", I.scale, true); I.scaleVel = vec3Of(o, "scale_per_frame", I.scaleVel, true);
        S.instances.push_back(I);
    }
    if (S.instances.empty()) { err = "scene file: no instances"; return false; }
    return true;
}

Xform instanceXform(const InstanceDesc& I, int f) {
    if (I.random) {
        SplitMix r(I.randomKey ^ (0x9E3779B97F4A7C15ull * (uint64_t)(f + 1)));
        const double z = 2.0 * r.next01() - 1.0, ph = 6.283185307179586 * r.next01(), ang = 6.283185307179586 * r.next01();
        const double rr = std::sqrt(std::fmax(0.0, 1.0 - z * z));
        const double axis[3] = { rr * std::cos(ph), rr * std::sin(ph), z };
        double t[3];
        for (int k = 0; k < 3; k++) t[k] = comp(I.spanLo, k) + (comp(I.spanHi, k) - comp(I.spanLo, k)) * (double)r.next01();
        const double s[3] = { 1, 1, 1 };
        return composeXform(t, axis, ang, s);
    }
    if (!I.perFrame.empty()) return I.perFrame[f < (int)I.perFrame.size() ? (f < 0 ? 0 : f) : I.perFrame.size() - 1];
    if (I.hasMatrix) return I.matrix;
    const double t[3] = { I.translate.x + (double)f * I.translateVel.x, I.translate.y + (double)f * I.translateVel.y,
                          I.translate.z + (double)f * I.translateVel.z };
    const double s[3] = { I.scale.x + (double)f * I.scaleVel.x, I.scale.y + (double)f * I.scaleVel.y, I.scale.z + (double)f * I.scaleVel.z };
    const double axis[3] = { I.axis.x, I.axis.y, I.axis.z };
    const double ang = (I.angleDeg + (double)f * I.angleVel) * 3.14159265358979323846 / 180.0;
    return composeXform(t, axis, ang, s);
}

void invertXform(const Xform& M, Xform& inv, float& detSign) {
    const float* m = M.m;
    const double a[3][3] = { { m[0], m[1], m[2] }, { m[4], m[5], m[6] }, { m[8], m[9], m[10] } };
    const double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                       a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
    detSign = det < 0 ? -1.f : 1.f;
    const double id = det != 0 ? 1.0 / det : 0.0;
    double b[3][3];
    b[0][0] = (a[1][1] * a[2][2] - a[1][2] * a[2][1]) * id; b[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) * id;
    b[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) * id; b[1][0] = (a[1][2] * a[2][0] - a[1][0] * a[2][2]) * id;
    b[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) * id; b[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) * id;
    b[2][0] = (a[1][0] * a[2][1] - a[1][1] * a[2][0]) * id; b[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) * id;
    b[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) * id;
    for (int i = 0; i < 3; i++) {
        const double t = -(b[i][0] * m[3] + b[i][1] * m[7] + b[i][2] * m[11]);
        for (int j = 0; j < 3; j++) inv.m[4 * i + j] = (float)b[i][j];
        inv.m[4 * i + 3] = (float)t;
    }
}

void makeInstanceScene(std::vector<Mesh>& meshes, SceneDesc& S, int nInst, uint64_t seed, bool small) {
    meshes.clear();
    S = SceneDesc();
    const Vec3 alb[5] = { v3(0.8f, 0.3f, 0.2f), v3(0.3f, 0.7f, 0.3f), v3(0.25f, 0.4f, 0.8f), v3(0.8f, 0.7f, 0.3f), v3(0.7f, 0.7f, 0.7f) };
    auto heightfield = [&](int grid, int k) {   // fBm heightfield over the unit square, centred on the origin
        Mesh m;
        makeTerrain(m, grid, 6, 0.2f, 0, seed + 17 * (uint64_t)k, 0.8f);
        for (Vec3& p : m.v) { p.x -= 0.5f; p.y -= 0.5f; }
        return m;
    };
    auto icosphere = [&](int level) { Mesh m; makeIcosphere(m, level, 0); return m; };
    meshes.push_back(icosphere(small ? 2 : 3));            // 1,280 triangles (small: 320)
    meshes.push_back(heightfield(small ? 20 : 72, 1));    // 10,082 (722)
    meshes.push_back(heightfield(small ? 40 : 159, 2));   // 49,928 (3,042)
    meshes.push_back(icosphere(small ? 3 : 6));            // 81,920 (1,280)
    meshes.push_back(heightfield(small ? 60 : 317, 4));   // 199,712 (6,962)
    for (int k = 0; k < 5; k++) { meshes[k].kd.assign(1, alb[k]); meshes[k].fm.assign(meshes[k].f.size(), 0); }
    Mesh ground;   // a 40 x 40 ground square at z = -1
    ground.v = { v3(-20.f, -20.f, -1.f), v3(20.f, -20.f, -1.f), v3(20.f, 20.f, -1.f), v3(-20.f, 20.f, -1.f) };
    ground.f = { { 0, 1, 2 }, { 0, 2, 3 } };
    ground.fm = { 0, 0 };
    ground.kd = { v3(0.5f, 0.5f, 0.5f) };
    meshes.push_back(ground);
    for (int k = 0; k < 6; k++) { MeshDesc d; d.file = "generated:" + std::to_string(k); S.meshes.push_back(d); }
    InstanceDesc g; g.mesh = 5; g.hasMatrix = true;
    const float id[12] = { 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0 };
    for (int i = 0; i < 12; i++) g.matrix.m[i] = id[i];
    S.instances.push_back(g);
    SplitMix r(seed * 0x2545F4914F6CDD1Dull + 7);
    for (int i = 0; i < nInst; i++) {
        InstanceDesc I;
        I.mesh = i % 5;
        I.random = true;
        I.randomKey = r.next();
        I.spanLo = v3(-12.f, -12.f, 0.f); I.spanHi = v3(12.f, 12.f, 2.f);
        S.instances.push_back(I);
    }
    S.frames = 60;
}

void instanceSceneCamera(float c2w[16]) { lookAtC2W(v3(0.f, -26.f, 14.f), v3(0.f, 0.f, 0.f), v3(0.f, 0.f, 1.f), c2w); }

void appendTransformed(Mesh& dst, const Mesh& src, int matOffset, const Xform& X) {
    Xform inv; float ds;
    invertXform(X, inv, ds);
    const int v0 = (int)dst.v.size();
    const bool wantN = !dst.n.empty() || src.hasN();
    if (wantN) dst.n.resize(v0, v3(0.f, 0.f, 0.f));
    for (size_t i = 0; i < src.v.size(); i++) {
        dst.v.push_back(xfPoint(X, src.v[i]));
        if (wantN) dst.n.push_back(src.hasN() ? xfNormal(inv, src.n[i]) : v3(0.f, 0.f, 0.f));
    }
    const bool wantUV = !dst.uv.empty() || src.hasUV();   // corner uvs stay with their triangles; (0, 0) where none
    if (wantUV) dst.uv.resize(6 * dst.f.size(), 0.f);
    for (size_t t = 0; t < src.f.size(); t++) {
        dst.f.push_back({ src.f[t].a + v0, src.f[t].b + v0, src.f[t].c + v0 });
        dst.fm.push_back(src.mat(t) + matOffset);
        if (wantUV) for (int c = 0; c < 6; c++) dst.uv.push_back(src.hasUV() ? src.uv[6 * t + c] : 0.f);
    }
}

}  // namespace cwp
