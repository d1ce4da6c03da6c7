// host/setup.h - command-line options and scene setup of the batch renderer, shared by the pt CLI and the C ABI.
// Mirrors the reference's Args / parseArgs / setupScene (src/app/scene_setup.h:12-50, src/app/scene_setup.cpp:27-271)
// for the legacy (Lambert + sun + sky) integrator: load a mesh or generate a procedural scene, read lighting and camera
// files, build the binned-SAH BVH2 and its GPU layout, pack triangles, and fill the frame parameters.
#pragma once
#include "device/renderer.h"
#include "host/bvh.h"
#include "host/camera.h"
#include "host/instances.h"
#include "host/mesh.h"
#include <memory>
#include <string>
#include <vector>

namespace cwp {

struct Args {
    std::string mesh, camera, out = "out", variant = "bvh8", c2wKey = "c2w_blender_zup", tri = "woop", c2w, lighting;
    std::string gen;                 // procedural scene: "boxes" (1,024-box cloud), "terrain" (fBm heightfield + boxes) or
                                     // "instances" (5 meshes, 1,000 instances with per-frame rigid transforms)
    std::string scene;               // two-level scene file (meshes + instances)
    int genInstances = -1, genSmall = 0, flatten = 0, frames = 0, frameStart = 0;
    std::string integrator = "auto", builder = "sah", collapse = "sah";
    int genBoxes = -1, genGrid = 242; uint64_t genSeed = 1;
    int fuse = -1, leaf = 4, bins = 16, batch = 64, synthetic = 0, frame = 0, w = 512, h = 512, spp = 256, bounces = 8;
    int persistent = 0, shadowRefill = 0, filterBH = 1, smooth = 1, warm = 8, reps = 1, stack = 8, blocksPerCU = 0, timeKernels = 1;
    int bounceWaves = 8;
    int sort = 0;                    // continuations grouped by direction bucket between bounces (same image)
    int tail = -1;                   // straggler hand-off threshold in lanes (-1 auto, 0 off; same image)
    int tailLevels = 2;              // hand-off levels (1 or 2)
    // full material integrator (reference: scene_setup.h:20-22)
    std::string materials, env, lightSampling = "mis", envSampling = "importance";
    float envStrength = -1.f, envRot = -1e30f;   // unset: the lighting file's (or 1, 0)
    int rr = 0, rrDepth = 3, gdEnergy = 1;
    std::string genEnv;              // generated environment "W,H,SIGMA[,SEED]": log-normal luminance (mu 0)
    int genMaterials = -1;           // >= 0: seed of the generated OpenPBR materials of a --gen scene's boxes
    int graph = -1;                  // frames replayed from a captured HIP graph (-1: on for the full integrator)
    std::string bvhCache;            // directory of built-scene cache files (empty: no cache)
    int gpuCollapse = 1;             // --builder hploc with bvh8: the CWBVH collapse on the GPU (bit-identical to the host's)
    int mortonBits = 0;              // H-PLOC keys (--builder hploc BLAS builds, two-level scenes' TLAS): 0 adaptive, 30
    std::string tlas = "bvh8";       // two-level scenes with BVH8 BLASes: per-frame TLAS as bvh2 or compressed bvh8
    std::string envOut;              // write the environment map in use (e.g. a generated one) as a Radiance HDR file
    // textures (full integrator): the kinds to use ("none": no maps), the filter, terminator shadowing, generated textures
    std::string textures = "base,emission,normal,metallic,roughness", texFilter = "linear";
    int bumpShadow = 0;
    int genTextures = 0;             // > 0: --gen terrain / boxes get texture coordinates and generated maps of this size
    float fx = 0, fy = 0, hfov = 0, aperture = -1.f, focusDist = -1.f, epsScale = 1e-4f, nearT = 0.f;
    unsigned seed = 1000;
    Vec3 albedo = v3(-1, 0, 0);
    Lighting light;
};

// Parses the options; false with a message in err on an unknown option, a missing value or an invalid value.
bool parseArgs(int argc, const char* const* argv, Args& A, std::string& err);

struct SceneSetup {
    Args A;
    Mesh M;
    BinaryBVH tree;                  // binary SAH tree (before the layout)
    SceneUpload S;                   // GPU-ready scene
    FrameParams P;
    RenderOptions O;
    float c2w[16];
    float extent = 0.f;
    int variant = 8, maxDepth = 0, reportLeaf = 0;
    double bvhS = 0, treeMs = 0, postMs = 0, sah = 0;
    size_t nNodes = 0, nodeBytes = 0;
    std::string sceneName;
    // full material integrator: the material table and the environment map (shared by every layout of the scene)
    bool full = false;
    bool textured = false;           // texture maps bound (full integrator)
    std::vector<Material> mats;
    std::shared_ptr<const EnvMap> env;
    int genBoxFirst = -1, genBoxCount = 0;   // generated scenes: triangle range of the boxes (12 each)
    int bvhCacheHit = -1;            // -1: no cache directory, 0: built (and saved), 1: loaded from the cache
    // two-level scenes: the meshes (one BLAS each, material ids into M.kd) and the instances
    bool inst = false;
    std::vector<Mesh> meshes;
    SceneDesc SD;
    int frames = 1;
};

// Instance records of frame f of a two-level scene (false if the scene is not instanced).
bool frameInstances(const SceneSetup& Z, int f, std::vector<InstanceRecord>& out);
// World-space emissive triangles of frame f of a two-level full-integrator scene, 3 float4 each: (v0, leaf position
// bits), (e1, material bits), (e2, 0), in instance order then mesh order; their power CDF; and per material the
// light pdf per unit area luminance(emission) / total power (reference: frameData, src/app/scene_setup.cpp:273-299).
// False if the scene has no instanced emitters.
bool frameEmitters(const SceneSetup& Z, int f, std::vector<Vec4>& emTris, std::vector<float>& emCdf,
                   std::vector<float>& matLightPdf);

// Runs the setup from Z.A; false with a message in err.
bool setupScene(SceneSetup& Z, std::string& err);

// The GPU-ready scene of a single-level setup in another BVH layout (2, 4 or 8), built from its mesh with its builder
// options.
bool relayoutScene(const SceneSetup& Z, int variant, SceneUpload& S, std::string& err);

// One JSON object (no newline) with the reference's per-frame statistics keys (src/app/pt.cpp:35-41).
std::string frameJson(const SceneSetup& Z, const Renderer& R, const FrameStats& st, int rep, int frame);

}  // namespace cwp
