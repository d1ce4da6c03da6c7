// app/pt.cpp - batch renderer CLI: set up the scene, render on the GPU, print one JSON line of statistics per timed
// frame, write <out>.png (sRGB) and <out>.pfm (linear radiance). Same flags and JSON keys as the reference's pt
// (src/app/pt.cpp:13-53): for each of the scene's frames (two-level scenes: instance transforms and emitters of that
// frame, the TLAS rebuilt on the GPU), an untimed warm-up render of `--warm` samples before the first, then `--reps`
// timed renders of `--spp`. The image files hold the last frame.
#include "device/renderer.h"
#include "host/io.h"
#include "host/setup.h"
#include <cstdio>
#include <vector>

int main(int argc, char** argv) {
    using namespace cwp;
    SceneSetup Z;
    std::string err;
    if (!parseArgs(argc, argv, Z.A, err)) { fprintf(stderr, "%s\n", err.c_str()); return 1; }
    if (!setupScene(Z, err)) { fprintf(stderr, "%s\n", err.c_str()); return 1; }
    Args& A = Z.A;
    fprintf(stderr, "mesh: %zu verts %zu tris, normals=%d, mats=%zu, extent=%.3f\n", Z.M.v.size(), Z.M.f.size(), (int)Z.M.hasN(), Z.M.kd.size(), Z.extent);
    fprintf(stderr, "BVH%d: %zu nodes (%.1f MB), depth %d, SAH %.3f, build %.2fs, tri=%s\n", Z.variant, Z.nNodes, Z.nodeBytes / 1e6, Z.maxDepth, Z.sah, Z.bvhS, A.tri.c_str());
    if (A.spp % (A.batch < 1 ? 1 : A.batch)) { fprintf(stderr, "spp must be a multiple of batch\n"); return 1; }
    Renderer R;
    if (!R.init(Z.S, Z.P, Z.O, err)) { fprintf(stderr, "renderer: %s\n", err.c_str()); return 2; }
    fprintf(stderr, "device %s: %d CUs, grid %d workgroups, %.0f MB\n", R.gpuName.c_str(), R.numCU, R.gridBlocks, R.vramMB);
    const int K = A.batch < 1 ? 1 : A.batch;
    std::vector<InstanceRecord> inst;
    std::vector<Vec4> emTris;
    std::vector<float> emCdf, matLightPdf;
    for (int fi = 0; fi < Z.frames; fi++) {
        const int f = A.frameStart + fi;
        if (Z.inst) {
            double tlasMs = 0;
            if (!frameInstances(Z, f, inst) || !R.setInstances(inst.data(), (int)inst.size(), tlasMs, err)) {
                fprintf(stderr, "frame %d: %s\n", f, err.c_str()); return 2;
            }
            if (Z.full) {
                frameEmitters(Z, f, emTris, emCdf, matLightPdf);
                if (!R.setEmitters(emTris.data(), (int)(emTris.size() / 3), emCdf.data(), matLightPdf.data(),
                                   (int)matLightPdf.size(), err)) { fprintf(stderr, "frame %d: %s\n", f, err.c_str()); return 2; }
            }
        }
        if (fi == 0 && A.warm > 0) {
            FrameStats W;
            if (!R.render(((A.warm + K - 1) / K) * K, 0, true, W, err)) { fprintf(stderr, "render: %s\n", err.c_str()); return 2; }
            fprintf(stderr, "warmup: %.1f ms, %.1f Mrays/s\n", W.frameMs, (W.rays + W.shadowRays) / (W.frameMs * 1e3));
        }
        for (int r = 0; r < A.reps; r++) {
            FrameStats st;
            if (!R.render(A.spp, 0, true, st, err)) { fprintf(stderr, "render: %s\n", err.c_str()); return 2; }
            printf("%s\n", frameJson(Z, R, st, r, f).c_str());
            fflush(stdout);
        }
    }
    const int N = A.w * A.h;
    std::vector<float> acc((size_t)N * 4), lin((size_t)N * 3);
    std::vector<uint8_t> rgb((size_t)N * 3);
    if (!R.readAccum(acc.data(), (size_t)N, err)) { fprintf(stderr, "%s\n", err.c_str()); return 2; }
    for (int i = 0; i < N; i++) for (int c = 0; c < 3; c++) lin[3 * i + c] = acc[4 * i + c] / A.spp;
    tonemapSRGB8(lin.data(), (size_t)N, rgb.data());
    writePNG(A.out + ".png", A.w, A.h, rgb.data());
    writePFM(A.out + ".pfm", A.w, A.h, lin.data());
    fprintf(stderr, "wrote %s.png / .pfm (tonemap: clip[0,1] + sRGB OETF, exposure 0)\n", A.out.c_str());
    return 0;
}
