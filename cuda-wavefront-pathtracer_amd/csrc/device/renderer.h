// device/renderer.h - host interface of the gfx950 wavefront renderer: the legacy integrator (Lambert + sun + sky) and
// the full material integrator (OpenPBR-subset materials, emissive triangles, environment map, MIS).
//
// Plain C++ types only (no HIP types), so the g++-compiled host code and the C ABI can include it. Mirrors the
// reference's WavefrontRenderer (src/render/renderer.h:47-93): init uploads the scene and allocates path state for
// w*h*batch paths, render() runs spp samples (a multiple of batch) with sample indices [sBase, sBase + spp) into the
// accumulation buffer, readAccum() returns per-pixel radiance sums. Errors are reported through the return value and
// an error string, never by exiting.
#pragma once
#include "core/material.h"
#include "core/texture.h"
#include "device/traverse_types.h"
#include "host/camera.h"
#include "host/envmap.h"
#include <memory>
#include <string>
#include <vector>

namespace cwp {

enum LightMode { kLightMIS = 0, kLightNEE = 1, kLightBSDF = 2 };

constexpr int kTailDefault = 12;   // hand-off threshold in lanes of the automatic setting (measured best of 6-16)
struct RenderOptions {
    int variant = 8;        // BVH layout: 2 or 8 (compressed)
    int tri = kTriWoop;     // triangle test: kTriMT, kTriWoop, kTriWatertight
    int batch = 64;         // samples per pixel in flight per wavefront pass
    int bounces = 8;        // maximum bounce depth (camera hit = depth 0)
    int stackLds = 16;      // LDS traversal-stack entries per lane (8, 16 or 32); deeper entries spill to HBM
    int blocksPerCU = 0;    // persistent grid: workgroups per CU (0 = occupancy limit)
    int bounceWaves = 8;    // trace+shade kernel: register budget for at least this many waves per SIMD (1, 6 or 8)
    int timeKernels = 1;    // per-launch hipEvent timing of the trace/shadow kernels
    int fuse = 1;           // 1: trace and shade in one kernel; 0: a closest-hit trace kernel, then a shade kernel
    int persistent = 0;     // 1: split kernels whose waves refill idle lanes with new rays mid-traversal (implies fuse 0)
    int shadowRefill = 0;   // shadow kernel: 1 = refill idle lanes mid-traversal, 0 = one ray per lane per round
    int graph = 0;          // 1: a frame is captured once into a HIP graph and replayed (per-kernel timing off)
    int sort = 0;           // 1: between bounces, group each partition's continuations by direction bucket
    int mortonBits = 0;     // two-level scenes: Morton keys of the per-frame TLAS build (0 adaptive, 30: 10 bits per axis)
    int tlasWidth = 8;      // two-level scenes with BVH8 BLASes: the per-frame TLAS as a BVH2 (2) or a compressed BVH8 (8)
    int tailLevels = 2;     // hand-off levels: 1 (the tail kernels finish every handed-off traversal) or 2 (they hand
                            // their own stragglers on to a second tail kernel)
    int tail = -1;          // straggler hand-off (BVH8, single-level, fused kernel): a trace+shade or shadow wave hands
                            // its tracing lanes to a tail kernel once fewer than this many remain (0: off; 2..64;
                            // -1: kTailDefault where supported, else off)
};

struct FrameStats {
    double rays = 0, shadowRays = 0;     // closest-hit rays (camera + continuation) and shadow rays traced
    double traceMs = 0, frameMs = 0;     // sum of the trace + shadow kernel times; whole frame (one event pair)
    double bounceMs = 0, shadowMs = 0;   // trace+shade kernel time and shadow kernel time
    int bounceLaunches = 0, shadowLaunches = 0;
    double hitRays = 0, contRays = 0, missRays = 0, unoccluded = 0;   // event counters (bytes-per-ray accounting)
    double cameraMiss = 0;                                             // camera rays that miss
    double shadeMs = 0;                  // split kernels: the shade kernels' part of bounceMs
    double sortMs = 0;                   // continuation sort kernels (per-kernel timing on)
    double tailMs = 0;                   // straggler hand-off: the tail kernels' part of bounceMs and shadowMs
};

// Scene data ready for upload, in BVH leaf order.
struct SceneUpload {
    int variant = 8, root = 0;
    std::vector<Vec4> nodes;      // BVH2 (4 float4/node) or CWBVH (5 float4/node)
    std::vector<Vec4> woop;       // Woop rows (3 float4/tri)
    std::vector<Vec4> verts;      // corners (3 float4/tri): (v0, orig id), (v1, material), (v2, 0)
    std::vector<Vec4> normals;    // per-corner normals (3 float4/tri) or empty
    std::vector<Vec4> albedo;     // per material (rgb, 0)
    std::vector<int> order;       // leaf position -> original triangle id
    int stackNeed = 0;            // traversal-stack entries the tree can require (BLASes only for two-level scenes)
    // two-level scenes: one BLAS per mesh in `nodes` (absolute indices); instances are set per frame
    int inst = 0;
    std::vector<int> meshRoot;    // BLAS root node per mesh
    std::vector<Vec4> meshBox;    // object-space box per mesh: lo, hi
    // full material integrator (full = 0: the legacy integrator, which uses `albedo` only)
    int full = 0;
    std::vector<Material> mats;   // per material id
    std::vector<float> triLightPdf;   // per leaf position: light-selection probability / area (empty: no emitters)
    std::vector<int> emTri;       // leaf positions of the emissive triangles, in power-CDF order
    std::vector<float> emCdf;     // emTri.size() + 1 entries
    std::shared_ptr<const EnvMap> env;   // null: no environment map (the sky constant is used)
    float envRotDeg = 0.f;
    int sunOn = 1;                // the sun is light-sampled
    int lightMode = kLightMIS;    // emitters / environment: MIS of light and BSDF sampling, light only, BSDF only
    int rr = 0, rrDepth = 3;      // Russian roulette from this depth on
    int gdEnergy = 1;             // glossy-diffuse energy compensation
    int envUniform = 0;           // environment light samples uniform on the sphere (instead of by luminance)
    int numEmInst = 0;            // two-level scenes: emissive triangles over all instances (set per frame)
    std::vector<int> meshTris;    // two-level scenes: first leaf position and triangle count per mesh
    // textures (full integrator; empty maps: none)
    std::vector<MaterialMaps> maps;   // per material id
    std::vector<Vec4> uvs;        // 2 float4 per leaf position: (u0, v0, u1, v1), (u2, v2, 0, 0)
    std::vector<Vec4> tans;       // 3 float4 per leaf position (tangent, handedness) or empty (no normal maps)
    std::vector<TexImage> texImages;
    std::vector<uint32_t> texels; // every image's RGBA8 texels, rows top first
    int texNearest = 0;           // 1: nearest-texel lookups; 0: bilinear
    int bumpShadow = 0;           // 1: terminator shadowing of normal-mapped diffuse shading
};

class Renderer {
public:
    Renderer();
    ~Renderer();
    Renderer(const Renderer&) = delete;
    Renderer& operator=(const Renderer&) = delete;

    bool init(const SceneUpload& S, const FrameParams& P, const RenderOptions& O, std::string& err);
    bool render(int spp, int sBase, bool clear, FrameStats& st, std::string& err);
    bool readAccum(float* rgba, size_t nPix, std::string& err);
    // the same sums on the device (w * h float4), valid for the renderer's lifetime; read them after a render has
    // finished (render() returns once its frame is done)
    const Vec4* accumDevice() const;
    // The viewer's display transform of the accumulation (device/display.h) into out (device, w * h packed RGBA8),
    // ordered after the queued frames; returns when it is done.
    bool display(uint32_t* out, int spp, float exposureEV, int mode, std::string& err);
    bool setParams(const FrameParams& P, std::string& err);
    // Closest (anyHit = false) or any hit of n rays; rays = 2 float4 per ray (o, tmin), (d, tmax); hits = float4
    // (t, u, v, leaf position as int bits, -1 on a miss). Both are device pointers; synchronous.
    // Two-level scenes also write the hit instance (-1 on a miss) to inst (a device array here) when given.
    bool trace(const void* rays, int n, bool anyHit, void* hits, std::string& err, int* inst = nullptr);
    // The same with host arrays, staged through device buffers owned by the renderer.
    bool traceHost(const float* rays, int n, bool anyHit, float* hits, std::string& err, int* inst = nullptr);
    // Two-level scenes: the instances of the next frames (records validated against the meshes' BLAS roots); builds
    // the TLAS on the GPU (H-PLOC over the instance boxes) and reports its device time.
    bool setInstances(const InstanceRecord* rec, int n, double& tlasMs, std::string& err);
    // Two-level full-integrator scenes: the frame's world-space emitters (3 float4 each), their power CDF (nEm + 1)
    // and the per-material light pdf per unit area; nEm must be the scene's instanced emitter count.
    bool setEmitters(const Vec4* emTris, int nEm, const float* emCdf, const float* matLightPdf, int nMats, std::string& err);
    // runtime updates (reference: setMaterials / setEnvironment, src/render/renderer.h:59-63): same material count
    // (legacy: the albedo is the base colour; full: emission unchanged); environment texels of the same size or null,
    // rotation in degrees, sun light-sampled or not (light-type probabilities recomputed)
    bool setMaterials(const Material* m, int n, std::string& err);
    bool setEnvironment(const Vec4* texels, size_t count, float envRotDeg, bool sunOn, std::string& err);

    double vramMB = 0, tlasMs = 0;   // tlasMs: device time of the last TLAS build
    int tlasWide = 0;                // two-level scenes: 1 when the TLAS is a compressed BVH8
    int gridBlocks = 0, numCU = 0;
    std::string gpuName;
    const RenderOptions& options() const;

private:
    struct Impl;
    Impl* p_;
};

// Number of compute units and device name of device 0 (false if no HIP device)
bool deviceInfo(int& numCU, std::string& name, std::string& err);

// Plain device memory for callers without a GPU array library: allocate, free, copy (kind 1 host->device,
// 2 device->host, 3 device->device; synchronous).
bool deviceMalloc(size_t bytes, void** ptr, std::string& err);
void deviceFree(void* ptr);
bool deviceCopy(void* dst, const void* src, size_t bytes, int kind, std::string& err);

}  // namespace cwp
