// app/viewer.cpp - the progressive viewer without a window. It keeps the reference viewer's render and display path
// and its automation options (src/app/viewer.cpp:1-19, 345-378): the batch renderer's scene setup, progressive
// accumulation (each viewer frame renders --passes passes of --batch samples at the next sample indices, so N frames
// equal the batch renderer's image at --spp N * passes * batch), the display transform on the GPU, and the outputs
// of the displayed image (PNG) and the linear mean (PFM). --test_updates 1 then changes the camera, the lens, the sun,
// the environment and material 0 through the runtime-update path and renders the same number of frames again,
// writing the outputs with the suffix _upd (src/app/viewer.cpp:385-393).
//
// The interactive window (GLFW, OpenGL interop, Dear ImGui, file dialogs, mouse and keyboard camera) is not built:
// the library targets headless GPU servers. Its options are rejected with a message.
//
// Options: every pt option, plus
//   --frames N           viewer frames (default: enough to reach --spp)   --passes P   passes per frame (default 1)
//   --screenshot f.png   the displayed image                              --screenshot_pfm f.pfm   linear sums / spp
//   --exposure EV        --tonemap clip|reinhard|aces                     --test_updates 1         --exit (accepted)
// One JSON line per viewer frame on stdout: frame, spp so far, render and display times.
#include "core/display.h"
#include "device/renderer.h"
#include "host/io.h"
#include "host/setup.h"
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

using namespace cwp;

namespace {

struct ViewOpts {
    int frames = 0, passes = 1, tonemap = kDisplayClip, testUpdates = 0;
    float exposure = 0.f;
    bool sppGiven = false;
    std::string screenshot, screenshotPfm;
};

bool splitArgs(int argc, char** argv, ViewOpts& V, std::vector<const char*>& rest, std::string& err) {
    rest.push_back(argv[0]);
    for (int i = 1; i < argc; i++) {
        const std::string k = argv[i];
        const bool hasV = i + 1 < argc;
        auto num = [&](int& dst, int lo) {
            if (!hasV) { err = k + " needs a value"; return false; }
            char* e = nullptr; const long v = strtol(argv[++i], &e, 10);
            if (!e || *e || v < lo) { err = k + " must be an integer >= " + std::to_string(lo); return false; }
            dst = (int)v; return true;
        };
        if (k == "--exit") continue;
        if (k == "--interop" || k == "--win_width" || k == "--win_height" || k == "--hide_ui" || k == "--screenshot_window") {
            err = k + " belongs to the interactive window, which this viewer does not have"; return false;
        }
        if (k == "--frames") { if (!num(V.frames, 1)) return false; }
        else if (k == "--passes") { if (!num(V.passes, 1)) return false; }
        else if (k == "--test_updates") { if (!num(V.testUpdates, 0)) return false; }
        else if (k == "--screenshot" || k == "--screenshot_pfm" || k == "--exposure" || k == "--tonemap") {
            if (!hasV) { err = k + " needs a value"; return false; }
            const std::string v = argv[++i];
            if (k == "--screenshot") V.screenshot = v;
            else if (k == "--screenshot_pfm") V.screenshotPfm = v;
            else if (k == "--exposure") {
                char* e = nullptr; V.exposure = strtof(v.c_str(), &e);
                if (!e || *e || !std::isfinite(V.exposure)) { err = "--exposure must be a finite number"; return false; }
            } else {
                if (v == "clip") V.tonemap = kDisplayClip;
                else if (v == "reinhard") V.tonemap = kDisplayReinhard;
                else if (v == "aces") V.tonemap = kDisplayACES;
                else { err = "--tonemap must be clip, reinhard or aces"; return false; }
            }
        } else {
            if (k == "--spp") V.sppGiven = true;
            rest.push_back(argv[i]);
        }
    }
    return true;
}

std::string withSuffix(const std::string& f, const char* suf) {
    if (f.empty()) return f;
    const size_t d = f.find_last_of('.');
    return d == std::string::npos ? f + suf : f.substr(0, d) + suf + f.substr(d);
}

// the viewer's camera state: yaw / pitch of the view direction, horizontal field of view, lens; the basis is
// rebuilt from them only when they change, so an untouched viewer renders the batch renderer's camera exactly
struct ViewCamera {
    float yaw, pitch, hfovDeg, fyRatio, lensR, focusD;
    void from(const SceneSetup& Z) {
        const Camera& c = Z.P.cam;
        const Vec3 f = -c.az;
        yaw = atan2f(f.y, f.x);
        pitch = asinf(std::max(-1.f, std::min(1.f, f.z)));
        hfovDeg = 2.f * atanf(0.5f * (float)c.w / c.fx) * 180.f / 3.14159265f;
        fyRatio = c.fy / c.fx;
        lensR = c.lensR; focusD = c.focusD;
        if (lensR <= 0.f) {   // a focus distance for a lens added later: the scene box centre along the view
            if (!Z.M.v.empty()) {
                Vec3 lo = Z.M.v[0], hi = lo;
                for (const Vec3& v : Z.M.v) { lo = vmin(lo, v); hi = vmax(hi, v); }
                const float d = dot((lo + hi) * 0.5f - c.pos, f);
                focusD = d > 0.f ? d : Z.extent;
            } else focusD = Z.extent;
        }
    }
    void apply(Camera& c) const {
        const Vec3 f = v3(cosf(pitch) * cosf(yaw), cosf(pitch) * sinf(yaw), sinf(pitch));
        const Vec3 r = normalize(cross(f, v3(0.f, 0.f, 1.f))), u = cross(r, f);
        c.ax = r; c.ay = u; c.az = -f;
        c.fx = focalFromHfov(c.w, hfovDeg); c.fy = c.fx * fyRatio;
        c.lensR = lensR; c.focusD = focusD;
    }
};

double msSince(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

int main(int argc, char** argv) {
    ViewOpts V;
    std::vector<const char*> rest;
    std::string err;
    if (!splitArgs(argc, argv, V, rest, err)) { fprintf(stderr, "%s\n", err.c_str()); return 1; }
    SceneSetup Z;
    if (!parseArgs((int)rest.size(), rest.data(), Z.A, err)) { fprintf(stderr, "%s\n", err.c_str()); return 1; }
    if (!setupScene(Z, err)) { fprintf(stderr, "%s\n", err.c_str()); return 1; }
    const Args& A = Z.A;
    const int W = A.w, H = A.h, N = W * H, K = std::max(1, A.batch);
    if (V.frames == 0) {   // without --frames: enough frames to reach --spp
        const int perFrame = V.passes * K;
        V.frames = std::max(1, (A.spp + perFrame - 1) / perFrame);
    }
    Renderer R;
    if (!R.init(Z.S, Z.P, Z.O, err)) { fprintf(stderr, "renderer: %s\n", err.c_str()); return 2; }
    fprintf(stderr, "device %s: %d CUs; %dx%d, %d frames x %d passes x %d spp\n", R.gpuName.c_str(), R.numCU, W, H,
            V.frames, V.passes, K);
    if (Z.inst) {   // the scene's first frame (the viewer's animation frame)
        std::vector<InstanceRecord> inst;
        double tlasMs = 0;
        if (!frameInstances(Z, A.frameStart, inst) || !R.setInstances(inst.data(), (int)inst.size(), tlasMs, err)) {
            fprintf(stderr, "instances: %s\n", err.c_str()); return 2;
        }
        if (Z.full) {
            std::vector<Vec4> emTris; std::vector<float> emCdf, mlp;
            frameEmitters(Z, A.frameStart, emTris, emCdf, mlp);
            if (!R.setEmitters(emTris.data(), (int)(emTris.size() / 3), emCdf.data(), mlp.data(), (int)mlp.size(), err)) {
                fprintf(stderr, "emitters: %s\n", err.c_str()); return 2;
            }
        }
    }
    uint32_t* dDisp = nullptr;
    if (!deviceMalloc((size_t)N * 4, (void**)&dDisp, err)) { fprintf(stderr, "%s\n", err.c_str()); return 2; }
    std::vector<uint32_t> disp((size_t)N);
    std::vector<float> acc((size_t)N * 4);
    int spp = 0;

    // V.frames viewer frames of V.passes passes each, the display after every frame, then the outputs
    auto run = [&](const char* suffix) -> bool {
        spp = 0;
        for (int f = 0; f < V.frames; f++) {
            const auto t0 = std::chrono::steady_clock::now();
            for (int p = 0; p < V.passes; p++) {
                FrameStats st;
                if (!R.render(K, spp, spp == 0, st, err)) return false;
                spp += K;
            }
            const double renderMs = msSince(t0);
            const auto t1 = std::chrono::steady_clock::now();
            if (!R.display(dDisp, spp, V.exposure, V.tonemap, err)) return false;
            const double displayMs = msSince(t1);
            printf("{\"frame\": %d, \"spp\": %d, \"render_ms\": %.3f, \"display_ms\": %.3f%s}\n", f, spp, renderMs,
                   displayMs, *suffix ? ", \"updated\": true" : "");
            fflush(stdout);
        }
        if (!V.screenshot.empty()) {
            if (!deviceCopy(disp.data(), dDisp, (size_t)N * 4, 2, err)) return false;
            std::vector<uint8_t> rgb((size_t)N * 3);
            for (int i = 0; i < N; i++)
                for (int c = 0; c < 3; c++) rgb[3 * (size_t)i + c] = (uint8_t)(disp[i] >> (8 * c));
            const std::string f = withSuffix(V.screenshot, suffix);
            if (!writePNG(f, W, H, rgb.data())) { err = "could not write " + f; return false; }
            fprintf(stderr, "wrote %s (%dx%d, %d spp, displayed image)\n", f.c_str(), W, H, spp);
        }
        if (!V.screenshotPfm.empty()) {
            if (!R.readAccum(acc.data(), (size_t)N, err)) return false;
            std::vector<float> lin((size_t)N * 3);
            for (int i = 0; i < N; i++)
                for (int c = 0; c < 3; c++) lin[3 * (size_t)i + c] = acc[4 * (size_t)i + c] / spp;
            const std::string f = withSuffix(V.screenshotPfm, suffix);
            if (!writePFM(f, W, H, lin.data())) { err = "could not write " + f; return false; }
            fprintf(stderr, "wrote %s\n", f.c_str());
        }
        fprintf(stderr, "automation: %d frames x %d passes x %d spp = %d spp\n", V.frames, V.passes, K, spp);
        return true;
    };
    bool ok = run("");
    if (ok && V.testUpdates == 1) {
        // one more round through every runtime-update path: camera yaw and field of view, a thin lens, the sun,
        // the environment's strength and rotation, and material 0
        FrameParams P = Z.P;
        ViewCamera vc; vc.from(Z);
        vc.yaw += 0.08f; vc.lensR = std::max(Z.extent, 1e-3f) * 0.01f; vc.hfovDeg *= 0.9f;
        vc.apply(P.cam);
        const Vec3 st0 = normalize(P.sunTo);
        const float sunAz = atan2f(st0.y, st0.x) * 180.f / 3.14159265f + 40.f;
        const float sunEl = std::max(asinf(std::max(-1.f, std::min(1.f, st0.z))) * 180.f / 3.14159265f, 30.f);
        const float a = sunAz * 3.14159265f / 180.f, e = sunEl * 3.14159265f / 180.f;
        P.sunTo = v3(cosf(e) * cosf(a), cosf(e) * sinf(a), sinf(e));
        ok = R.setParams(P, err);
        if (ok && Z.full && Z.S.env) {
            std::vector<Vec4> tex = Z.S.env->texels;
            for (Vec4& t : tex) { t.x *= 2.f; t.y *= 2.f; t.z *= 2.f; }
            ok = R.setEnvironment(tex.data(), tex.size(), Z.S.envRotDeg + 30.f, luminance(P.sunE) > 0.f, err);
        }
        if (ok) {
            std::vector<Material> mats;
            if (Z.full) mats = Z.mats;
            else for (const Vec4& c : Z.S.albedo) mats.push_back(lambertMaterial(v3(c.x, c.y, c.z)));
            if (!mats.empty()) {
                mats[0].baseColor = v3(0.9f, 0.2f, 0.1f); mats[0].roughness = 0.2f; mats[0].specularWeight = 1.f;
                ok = R.setMaterials(mats.data(), (int)mats.size(), err);
            }
        }
        if (ok) ok = run("_upd");
    }
    deviceFree(dDisp);
    if (!ok) { fprintf(stderr, "viewer: %s\n", err.c_str()); return 2; }
    return 0;
}
