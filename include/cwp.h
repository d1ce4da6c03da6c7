/* cwp.h - C ABI of the MI355X wavefront path tracer (libcwp.so).
 *
 * Plain pointers and sizes only. Every call returns CWP_OK (0) or an error code; cwp_last_error() gives the message of
 * the calling thread's last failure. Nothing in the library exits the process. Calls on one context are not
 * thread-safe; distinct contexts are independent. Inputs are copied; outputs go to caller buffers.
 *
 * The entry points replace the reference's statically linked C++ interfaces (paths relative to the reference repo):
 *   scene setup      src/app/scene_setup.h:12-50    parseArgs / setupScene / SceneSetup
 *   renderer         src/render/renderer.h:47-93    WavefrontRenderer::{init, render, readAccum, setParams}
 *   scene buffers    src/render/renderer.h:11-43    RenderOptions, SceneBuffers (legacy-integrator fields)
 *   BVH build        src/accel/bvh_builder.h:9-16, src/accel/bvh_layouts.h:14-40, src/accel/bvh_tools.h:23-29
 *   traversal        src/accel/traverse.cuh:155-160 trav<V, ANY, TRI> (as a kernel over explicit rays)
 *   frame statistics src/render/renderer.h:45 FrameStats, src/app/pt.cpp:35-41 (JSON line)
 */
#ifndef CWP_H
#define CWP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    CWP_OK = 0,
    CWP_ERR_ARG = 1,      /* invalid argument or option */
    CWP_ERR_IO = 2,       /* unreadable / malformed input file */
    CWP_ERR_DEVICE = 3,   /* HIP runtime failure or no device */
    CWP_ERR_STATE = 4     /* call not valid in the object's state */
};

typedef struct cwp_setup cwp_setup; /* parsed options + scene + BVH + packed triangles (host memory) */
typedef struct cwp_ctx cwp_ctx;     /* renderer: device buffers, stream, kernels */
typedef struct cwp_bvh cwp_bvh;     /* BVH built from caller arrays (host memory) */

/* Frame parameters of the legacy integrator (camera, sun, sky, bounce limit, seed, ray offset, filter table).
 * Replaces the reference's Params block (src/render/params.h:8-16). */
typedef struct cwp_params {
    float pos[3], ax[3], ay[3], az[3]; /* camera position and camera-to-world axes (looks down -az) */
    float fx, fy, px, py;              /* focal lengths and principal point in pixels */
    int w, h;                          /* image size */
    float lens_r, focus_d;             /* thin lens (lens_r = 0: pinhole) */
    float sun_to[3], sun_e[3], sky[3]; /* unit direction toward the sun, sun irradiance, sky radiance */
    float sun_cos;                     /* cos of the sun's angular radius */
    int max_bounce;                    /* paths stop after this bounce (camera hit = 0) */
    uint32_t seed;                     /* RNG seed */
    float eps, near_t;                 /* ray-origin offset (world units), camera near clip */
    int has_normals, filter_bh;        /* interpolate vertex normals; Blackman-Harris 1.5 px filter (else 1 px box) */
    float ftab[65];                    /* filter inverse CDF */
} cwp_params;

/* Per-frame counters and times. Replaces FrameStats (src/render/renderer.h:45) plus the event counters used for
 * the bytes-per-ray accounting. */
typedef struct cwp_frame_stats {
    double rays_closest, rays_shadow;  /* rays traced: camera + continuation, shadow */
    double t_trace_ms, t_frame_ms;     /* trace+shadow kernel time, whole frame (hipEvent pair) */
    double t_bounce_ms, t_shadow_ms;   /* fused trace+shade kernels, shadow kernels */
    int bounce_launches, shadow_launches;
    double rays_hit, rays_miss, rays_cont, shadow_unoccluded;
    double camera_miss;               /* camera rays that miss (rays_miss includes them) */
    double t_shade_ms;                /* split kernels: shade-kernel part of t_bounce_ms (0 when fused) */
    double t_sort_ms;                 /* continuation-sort kernels (options.sort; 0 when off or untimed) */
    double t_tail_ms;                 /* straggler hand-off: the tail kernels' part of t_bounce_ms and t_shadow_ms */
} cwp_frame_stats;

/* A scene in the renderer's upload layout: host arrays in BVH leaf order, copied by cwp_create_scene. Replaces
 * SceneBuffers (src/render/renderer.h:31-43); the layouts are those of cwp_bvh_array. A zero-initialised tail (mats
 * NULL) is a legacy-integrator scene; with mats the scene renders with the full material integrator. */
typedef struct cwp_scene_desc {
    int variant;              /* 2: BVH2, 16 floats per node; 8: compressed BVH8, 20 floats (80 B) per node */
    const float* nodes;
    size_t num_nodes;
    const float* corners;     /* 12 floats per triangle: (v0, original id bits), (v1, material id bits), (v2, 0) */
    size_t num_tris;
    const float* woop;        /* 12 floats per triangle (Woop rows), or NULL; required by tri_test 1 */
    const float* normals;     /* 12 floats per triangle (per-corner normals), or NULL */
    const float* albedo;      /* 4 floats per material (rgb, unused) */
    size_t num_materials;
    int tri_form;             /* CWP_TRIS_CORNERS, or CWP_TRIS_EDGES: the reference's (v0, id), (e1, material), (e2, 0)
                                 (src/accel/tri_pack.cpp:6-16), turned into corners v0 + e1, v0 + e2 on upload */
    int root;                 /* root node (0 for the layouts of cwp_bvh_build; n - 2 for cwp_hploc_to_bvh2) */
    /* full material integrator (renderer.h:35-38; the arrays of CWP_ARR_MATERIALS, CWP_ARR_TRI_LIGHT_PDF,
     * CWP_ARR_EM_TRI, CWP_ARR_EM_CDF and CWP_ARR_ENV_TEX) */
    const float* mats;          /* 16 floats per material (the 64-byte record), num_materials of them, or NULL */
    const float* tri_light_pdf; /* per triangle (leaf order): light-selection probability / area; NULL: no emitters */
    const int* em_tri;          /* emissive triangles (leaf positions) in power-CDF order */
    const float* em_cdf;        /* num_emitters + 1 entries, 0 .. 1 */
    size_t num_emitters;
    const float* env_texels;    /* 4 floats per texel (radiance), env_w x env_h, row 0 = +z; NULL: constant sky */
    int env_w, env_h;
    float env_rot_deg;          /* environment rotation about +z */
    int sun_on;                 /* 1: the sun is light-sampled */
    int light_mode;             /* CWP_LIGHT_MIS, CWP_LIGHT_NEE or CWP_LIGHT_BSDF */
    int rr, rr_depth;           /* Russian roulette from depth rr_depth on */
    int gd_energy;              /* glossy-diffuse energy compensation */
    int env_uniform;            /* environment light samples uniform on the sphere instead of by luminance */
    /* texture maps (full integrator; the arrays of CWP_ARR_MAPS, CWP_ARR_UVS, CWP_ARR_TANGENTS, CWP_ARR_TEX_IMAGES and
     * CWP_ARR_TEXELS; reference: RenderOptions / SceneBuffers texture fields, src/render/renderer.h:24, params.h:41).
     * maps NULL: no textures. */
    const int* maps;            /* 8 per material: base colour, emission, normal, metallic, roughness image ids (-1:
                                   none), metallic channel, roughness channel (0..3 = r..a), normal strength (float bits) */
    const float* uvs;           /* 8 floats per triangle (leaf order): (u0, v0, u1, v1), (u2, v2, 0, 0) */
    const float* tangents;      /* 12 floats per triangle: per corner (tangent xyz, handedness +-1); needed by normal maps */
    const int* tex_images;      /* 4 per image: first texel, width, height, 1 = sRGB colour data / 0 = linear */
    size_t num_images;
    const uint32_t* texels;     /* RGBA8 texels (r in the low byte), each image's rows top first */
    size_t num_texels;
    int tex_nearest;            /* 1: nearest-texel lookups; 0: bilinear */
    int bump_shadow;            /* 1: terminator shadowing of normal-mapped diffuse shading (Chiang et al. 2019) */
} cwp_scene_desc;
enum { CWP_LIGHT_MIS = 0, CWP_LIGHT_NEE = 1, CWP_LIGHT_BSDF = 2 };
enum { CWP_TRIS_CORNERS = 0, CWP_TRIS_EDGES = 1 };

/* Kernel and schedule options. Replaces RenderOptions (src/render/renderer.h:11-28) for the legacy integrator. */
typedef struct cwp_options {
    int tri_test;        /* 0 Moller-Trumbore, 1 Woop, 2 watertight */
    int batch;           /* samples per pixel in flight per wavefront pass */
    int bounces;         /* maximum bounce depth (camera hit = depth 0), 0..64 */
    int fuse;            /* 1: trace and shade in one kernel; 0: split trace and shade kernels */
    int persistent;      /* waves refill idle lanes with new rays: 1 split kernels (needs fuse 0), 2 fused kernel */
    int shadow_refill;   /* 1: the shadow kernel refills idle lanes mid-traversal */
    int stack_lds;       /* LDS traversal-stack entries per lane: 8, 16 or 32 (deeper entries spill to HBM) */
    int blocks_per_cu;   /* workgroups per CU of the persistent grids (0 = occupancy limit) */
    int bounce_waves;    /* fused kernel register budget: at least this many waves per SIMD (1, 6 or 8) */
    int time_kernels;    /* per-launch event timing (t_bounce_ms, t_shade_ms, t_shadow_ms) */
    int graph;           /* 1: each frame replays a captured HIP graph (per-launch timing off) */
    int sort;            /* continuations reordered before they are traced (reference option --sort,
                            src/render/wavefront.cu:717; different sorts, same results): 1 every bounce's, grouped by
                            direction bucket within each XCD partition; 2 the first bounce's, grouped by elevation
                            bucket within each pixel's samples (fused kernel only) */
    int tail_lanes;      /* straggler hand-off (compressed BVH8, single-level scenes, fused kernel): once fewer than
                            this many lanes of a trace+shade wave are still tracing, their traversals are handed to a
                            tail kernel that resumes them in full waves (same results); 0 off, 2..64 */
} cwp_options;

/* Scene and BVH summary of a setup. */
typedef struct cwp_scene_info {
    int variant;        /* 2 (BVH2) or 8 (compressed BVH8) */
    int tri_test;       /* 0 Moller-Trumbore, 1 Woop, 2 watertight */
    int num_tris, num_verts, num_materials;
    int num_nodes, node_floats4;  /* nodes and float4 per node */
    int bvh_depth;      /* binary SAH tree depth */
    int stack_need;     /* traversal stack entries the layout can require */
    double sah, bvh_build_s;
    int width, height, spp, batch, bounces;
    int num_meshes, num_instances, frames, frame_start;   /* two-level scenes (num_meshes = 0 otherwise) */
    /* full material integrator (full = 0: legacy Lambert + sun + sky) */
    int full, num_emitters, env_w, env_h, sun_on;
    int light_mode;     /* 0 MIS, 1 light sampling only, 2 BSDF sampling only */
    int rr, rr_depth, gd_energy, env_uniform;
    float env_rot_deg;
    int bvh_cache;      /* -1 no cache directory (--bvh_cache), 0 built and saved, 1 loaded from the cache */
    int textured, num_images, tex_nearest, bump_shadow;   /* texture maps bound (full integrator), images, filter */
} cwp_scene_info;

/* Instance record of a two-level scene (112 bytes): world-to-object rows, object-to-world rows (3 x 4, row-major), BLAS
 * root node of the mesh, mesh id, sign of det(fwd). Replaces DevInstance (src/accel/instance.h:7). */
typedef struct cwp_instance {
    float inv[12], fwd[12];
    int root, mesh;
    float det_sign;
    int pad;
} cwp_instance;

/* Array selectors for cwp_setup_array / cwp_bvh_array (element type, count unit). */
enum {
    CWP_ARR_VERTICES = 0, /* float[3] per mesh vertex */
    CWP_ARR_TRIANGLES = 1,/* int[3] per triangle (mesh order) */
    CWP_ARR_MATERIAL = 2, /* int per triangle: material id */
    CWP_ARR_ALBEDO = 3,   /* float[4] per material */
    CWP_ARR_NODES = 4,    /* float[4] x node_floats4 per node (GPU layout) */
    CWP_ARR_ORDER = 5,    /* int per leaf position: original triangle id */
    CWP_ARR_WOOP = 6,     /* float[4] x 3 per leaf position: Woop rows */
    CWP_ARR_CORNERS = 7,  /* float[4] x 3 per leaf position: (v0, id), (v1, material), (v2, 0) */
    CWP_ARR_NORMALS = 8,  /* float[3] per mesh vertex (empty without normals) */
    CWP_ARR_TREE = 9      /* int[4] per binary SAH node: left, right, first, count; then float[6] box via CWP_ARR_TREE_BOX */,
    CWP_ARR_TREE_BOX = 10,/* float[6] per binary SAH node: lo.xyz, hi.xyz */
    CWP_ARR_PRIM = 11,    /* int per binary-tree primitive slot: triangle id */
    CWP_ARR_MESH_ROOTS = 12, /* int per mesh of a two-level scene: BLAS root node */
    CWP_ARR_MESH_BOXES = 13, /* float[8] per mesh of a two-level scene: object-space box lo (xyz, 0), hi (xyz, 0) */
    /* full material integrator */
    CWP_ARR_MATERIALS = 14,  /* float[16] per material: base rgb, metallic, specular rgb, roughness, emission rgb, ior,
                                specular weight, transmission, 0, 0 */
    CWP_ARR_EM_TRI = 15,     /* int per emitter: leaf position, in power-CDF order */
    CWP_ARR_EM_CDF = 16,     /* float per emitter + 1: power CDF */
    CWP_ARR_TRI_LIGHT_PDF = 17, /* float per leaf position: light-selection probability / area */
    CWP_ARR_ENV_TEX = 18,    /* float[4] per environment texel (radiance x strength), row 0 = +z */
    CWP_ARR_ENV_MARG = 19,   /* float per environment row + 1: marginal CDF */
    CWP_ARR_ENV_COND = 20,   /* float per (row, column + 1): conditional CDFs */
    CWP_ARR_MESH_TRIS = 21,  /* int[2] per mesh of a two-level scene: first leaf position, triangle count */
    CWP_ARR_UVS = 22,        /* float[8] per leaf position: corner texture coordinates (u0, v0, u1, v1), (u2, v2, 0, 0) */
    CWP_ARR_TANGENTS = 23,   /* float[12] per leaf position: per corner (tangent xyz, handedness); empty without normal maps */
    CWP_ARR_MAPS = 24,       /* int[8] per material: texture bindings (see cwp_scene_desc.maps) */
    CWP_ARR_TEX_IMAGES = 25, /* int[4] per image: first texel, width, height, sRGB */
    CWP_ARR_TEXELS = 26,     /* uint32 per texel: RGBA8, r in the low byte, rows top first */
    CWP_ARR_SRGB_TABLE = 27, /* float[256]: the sRGB decode of each byte value */
    CWP_ARR_MESH_UVS = 28    /* float[6] per mesh triangle (single-level scenes): corner texture coordinates; empty without */
};

const char* cwp_last_error(void);
int cwp_version(void);

/* Number of compute units and the name of HIP device 0. */
int cwp_device_info(int* num_cu, char* name, size_t name_cap);

/* Scene setup from command-line style options (the reference's pt flags, e.g. "--mesh", "--camera", "--lighting",
 * "--width", "--spp", "--batch", "--bounces", "--variant", "--tri", plus "--gen boxes|terrain", "--bins", "--stack").
 * argv[0] is ignored. Replaces parseArgs + setupScene (src/app/scene_setup.cpp:27-57, 85-271). Host only. */
int cwp_setup_create(int argc, const char* const* argv, cwp_setup** out);
void cwp_setup_destroy(cwp_setup* s);
int cwp_setup_info(const cwp_setup* s, cwp_scene_info* info);
int cwp_setup_params(const cwp_setup* s, cwp_params* params);
/* Borrowed view of one of the setup's arrays (valid until cwp_setup_destroy); *count = number of elements. */
int cwp_setup_array(const cwp_setup* s, int which, const void** data, size_t* count);

/* Renderer: uploads the setup's scene and allocates path state for width * height * batch paths.
 * Replaces WavefrontRenderer::init (src/render/wavefront.cu:469-502). */
int cwp_create(const cwp_setup* s, cwp_ctx** out);
void cwp_destroy(cwp_ctx* c);
/* Renders spp samples per pixel (a multiple of batch), sample indices [sample_base, sample_base + spp), into the
 * accumulation buffer (cleared first if clear != 0). Replaces WavefrontRenderer::render (src/render/wavefront.cu:684). */
int cwp_render(cwp_ctx* c, int spp, int sample_base, int clear, cwp_frame_stats* stats);
/* Per-pixel radiance sums (divide by spp), float[4] per pixel, rows top to bottom. Replaces readAccum. */
int cwp_read_accum(cwp_ctx* c, float* rgba, size_t num_pixels);
/* Device pointer to the same sums (w * h float4), valid for the context's lifetime (NULL for a null context). Frames
 * render on the context's own stream and cwp_render returns when its frame is done. Replaces
 * WavefrontRenderer::accumDevice (src/render/renderer.h:59). */
const float* cwp_accum_device(cwp_ctx* c);

/* ---- progressive viewer: the display transform (reference: src/app/display.h:5-8, display.cu:9-33) ----
 * Each channel: sum / spp (spp 0 is read as 1), times 2^exposure_ev when exposure_ev != 0, then for modes 1 and 2 a
 * tone curve on max(x, 0) (Reinhard x / (1 + x); the ACES fit of Narkowicz 2015 after a 0.6 scale), then the PNG
 * writer's encode: clip to [0, 1], sRGB OETF, x * 255 + 0.5 truncated. Output: packed RGBA8 per pixel (r in the lowest
 * address byte, a = 255). Mode CWP_DISPLAY_CLIP at exposure 0 gives the bytes of the PNG file that pt writes. */
enum { CWP_DISPLAY_CLIP = 0, CWP_DISPLAY_REINHARD = 1, CWP_DISPLAY_ACES = 2 };
/* The context's accumulation into rgba8 (device pointer, w * h * 4 bytes), after the queued frames; synchronous.
 * Replaces launchDisplay(R.accumDevice(), out, N, spp, exposure, mode) as the viewer calls it
 * (src/app/viewer.cpp:176,182). */
int cwp_display(cwp_ctx* c, void* rgba8, int spp, float exposure_ev, int mode);
/* The same for any n device float4 sums (accum) into n device RGBA8 pixels, on the default stream; synchronous.
 * Replaces launchDisplay (src/app/display.cu:30-33). */
int cwp_display_buffer(const float* accum, void* rgba8, int n, int spp, float exposure_ev, int mode);
/* The host side of the same encode, for checks: the 8-bit sRGB encode of one value (cwp_srgb_encode8), and the step
 * positions the kernel searches (steps[k], k = 1..255: the smallest float that encodes to k or more; steps[0] = -inf). */
int cwp_srgb_encode8(float x);
void cwp_srgb_steps(float steps[256]);
/* Replaces setParams (src/render/wavefront.cu:744); the image size must not change. */
int cwp_set_params(cwp_ctx* c, const cwp_params* params);
/* Runtime updates, effective from the next render (the context waits for queued frames). Replace
 * WavefrontRenderer::setMaterials / setEnvironment (src/render/renderer.h:59-63). Materials: 16 floats each, the
 * scene's count; the legacy integrator takes their base colours as albedo; the full integrator needs the emission of
 * every material unchanged (the light tables depend on it). Environment (full integrator): texels of the scene's size
 * (or NULL to keep them; the sampling distribution is rebuilt from them), the rotation about +z in degrees, and
 * whether the sun is light-sampled (the light-type probabilities are recomputed). */
int cwp_set_materials(cwp_ctx* c, const float* mats, int num_materials);
int cwp_set_environment(cwp_ctx* c, const float* texels, size_t num_texels, float env_rot_deg, int sun_on);
/* Closest (any_hit = 0) or any hit of n rays with the context's BVH and triangle test. rays: device pointer to
 * 8 floats per ray (o.xyz, tmin, d.xyz, tmax); hits: device pointer to 4 floats per ray (t, u, v, leaf position as
 * int bits, -1 = miss). Synchronous. Kernel form of trav<V, ANY, TRI> (src/accel/traverse.cuh:155). */
int cwp_trace(cwp_ctx* c, const void* rays, int n, int any_hit, void* hits);
/* Same as cwp_trace with host arrays: the library stages the rays and hits through device buffers it owns. */
int cwp_trace_host(cwp_ctx* c, const float* rays, int n, int any_hit, float* hits);
/* One JSON object (no newline) with the reference's per-frame statistics keys (src/app/pt.cpp:35-41). */
int cwp_frame_json(const cwp_setup* s, const cwp_ctx* c, const cwp_frame_stats* st, int rep, int frame, char* buf, size_t cap);
/* The defaults of cwp_options: Woop test, fused trace+shade kernel at 8 waves per SIMD, one-ray-per-lane shadow
 * kernel, 8 LDS stack entries, batch 64, 8 bounces. */
void cwp_options_default(cwp_options* o);
/* Checks a scene description without a GPU: array sizes, every child-node and leaf-triangle index in range, BVH8
 * internal masks consistent with the slot metadata, material
 * ids in range, no node with two parents and none for the root (so every traversal ends). *stack_need receives the
 * traversal-stack entries the tree can need. cwp_create_scene runs the same
 * check; the kernels trust every index that passes it. */
int cwp_scene_check(const cwp_scene_desc* d, int* stack_need);
/* Renderer over caller buffers. Replaces WavefrontRenderer::init(Params, SceneBuffers, RenderOptions)
 * (src/render/wavefront.cu:469-502). cwp_frame_json needs a setup and is not available for such a context. */
int cwp_create_scene(const cwp_scene_desc* d, const cwp_params* params, const cwp_options* opts, cwp_ctx** out);
/* Two-level scenes (--scene file or --gen instances): the instance records of animation frame `frame` (transforms
 * evaluated in double, inverted in double; reference SceneSetup::frameData, src/app/scene_setup.cpp:274-299); *count
 * receives the instance count, at most `cap` records are written. Host only. */
int cwp_setup_instances(const cwp_setup* s, int frame, cwp_instance* out, int cap, int* count);
/* Sets the instances of the next frames and rebuilds the TLAS on the GPU (H-PLOC over the instance boxes); *tlas_ms
 * receives its device time. Every record must name a mesh and that mesh's BLAS root. Replaces setInstances
 * (src/render/renderer.h:52). cwp_create does this for the setup's first frame. */
int cwp_set_instances(cwp_ctx* c, const cwp_instance* inst, int n, double* tlas_ms);
/* cwp_setup_instances + cwp_set_instances (and, for the full integrator, cwp_setup_emitters + cwp_set_emitters) for a
 * frame of the context's setup. */
int cwp_set_frame(cwp_ctx* c, int frame, double* tlas_ms);
/* Two-level full-integrator scenes: the frame's world-space emissive triangles (3 float[4] each: (v0, leaf position
 * bits), (e1, material bits), (e2, 0)), their power CDF (num_emitters + 1) and per material the light pdf per unit
 * area; num_emitters must equal the scene's instanced emitter count (reference: the emTris / emCdf / matLightPdf
 * arguments of setInstances, src/render/renderer.h:52). */
int cwp_set_emitters(cwp_ctx* c, const float* em_tris, int num_emitters, const float* em_cdf, const float* mat_light_pdf,
                     int num_materials);
/* the setup's emitters of a frame (em_tris = NULL: count only; em_tris: cap emitters, em_cdf: num + 1 floats,
 * mat_light_pdf: mat_cap floats or NULL) (reference: frameData, src/app/scene_setup.cpp:273-299) */
int cwp_setup_emitters(const cwp_setup* s, int frame, float* em_tris, int cap, int* num_emitters, float* em_cdf,
                       float* mat_light_pdf, int mat_cap);
/* cwp_trace_host that also reports the hit instance of two-level scenes (-1 on a miss). */
int cwp_trace_host_inst(cwp_ctx* c, const float* rays, int n, int any_hit, float* hits, int* inst);
/* Grid size (workgroups) and VRAM used by the context. */
int cwp_ctx_info(const cwp_ctx* c, int* grid_blocks, double* vram_mb);

/* Binned-SAH BVH2 over caller triangles and its GPU layout: variant 2 (BVH2, leaves <= leaf_max) or 8 (CWBVH collapsed
 * from a 1-primitive-leaf tree). Replaces BVHBuilder::build + BVH2Flat / CWBVH::build (src/accel/bvh_builder.cpp:5,
 * src/accel/bvh_layouts.cpp:31, src/accel/bvh_tools.cpp:75). Host only. */
int cwp_bvh_build(const float* vertices, int num_vertices, const int* triangles, int num_tris, int variant, int bins,
                  int leaf_max, cwp_bvh** out);
void cwp_bvh_destroy(cwp_bvh* b);
/* CWP_ARR_NODES, CWP_ARR_ORDER, CWP_ARR_WOOP, CWP_ARR_CORNERS, CWP_ARR_TREE, CWP_ARR_TREE_BOX, CWP_ARR_PRIM */
int cwp_bvh_array(const cwp_bvh* b, int which, const void** data, size_t* count);
/* depth of the binary tree, traversal stack need, SAH cost of the layout */
int cwp_bvh_stats(const cwp_bvh* b, int* depth, int* stack_need, double* sah);
/* The device encoding of compressed BVH8 nodes (what cwp_create / cwp_create_scene upload): 32 floats (128 bytes) per
 * 20-float node of CWP_ARR_NODES: the node's 20 floats, except that the child base (float 4) carries the leaf-slot
 * mask in its top byte, then 12 floats of padding. out holds 32 * num_nodes floats. No reference counterpart (the
 * reference uploads its 80-byte nodes); for inspection and tests. Host only. */
int cwp_bvh8_device_nodes(const float* nodes, size_t num_nodes, float* out);
/* The same outputs from H-PLOC nodes over the mesh's triangle boxes (8 floats per node, 2 nt - 1 nodes, children
 * before parents): SAH leaf collapse (leaves <= leaf_max, 1 for variant 8), then the layout. Replaces treeFromHploc
 * + BVH2Flat / CWBVH::build (src/accel/bvh_tools.cpp:12-38, src/app/scene_setup.cpp:76-83). Host only. */
int cwp_bvh_from_hploc(const float* vertices, int num_vertices, const int* triangles, int num_tris, const float* hnodes,
                       int root, int variant, int leaf_max, cwp_bvh** out);

/* ---- H-PLOC GPU BVH2 builder (src/accel/hploc.h:21-40) ----
 * Nodes: 8 floats per node, (lo.xyz, left id bits), (hi.xyz, right id bits), 2n - 1 nodes; [0, n) are the leaves in
 * Morton order (left = primitive id, right = -1), internal node ids come from a counter (children before parents),
 * the root is 2n - 2 (0 when n == 1). Internal node numbering depends on scheduling; the tree does not. */
typedef struct cwp_hploc cwp_hploc;
/* Morton code of the keys: CWP_MORTON_ADAPTIVE = min(21, (64 - ceil(log2 n)) / 3) bits per axis, the reference's keys
 * (src/accel/hploc.cu:181-183); CWP_MORTON_30 = 10 bits per axis (30-bit codes). The primitive index below the code
 * keeps every key unique either way. */
enum { CWP_MORTON_ADAPTIVE = 0, CWP_MORTON_30 = 30 };
/* Host arrays: lo / hi 4 floats per primitive box; builds twice (the second timed with events, *gpu_ms) and copies
 * the nodes out. Replaces hplocBuildHost (adaptive keys); the _mode variant takes the Morton mode. */
int cwp_hploc_build_host(const float* lo, const float* hi, int n, float* nodes, int* root, double* gpu_ms);
int cwp_hploc_build_host_mode(const float* lo, const float* hi, int n, int morton_bits, float* nodes, int* root,
                              double* gpu_ms);
/* Stream-ordered builder over device arrays; scratch sized for `capacity` primitives (grown on demand). */
int cwp_hploc_create(int capacity, cwp_hploc** out);
void cwp_hploc_destroy(cwp_hploc* b);
/* lo / hi: device pointers, 4 floats per primitive; stream: a hipStream_t or NULL (default stream). Asynchronous: the
 * nodes (cwp_hploc_nodes, a device pointer owned by the builder) are ready when the stream gets there. */
int cwp_hploc_build(cwp_hploc* b, const void* lo, const void* hi, int n, void* stream, int* root);
/* as cwp_hploc_build, with the Morton mode (CWP_MORTON_ADAPTIVE or CWP_MORTON_30) */
int cwp_hploc_build_mode(cwp_hploc* b, const void* lo, const void* hi, int n, int morton_bits, void* stream, int* root);
/* No counterpart in the reference. An H-PLOC build that allocates nothing, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid allocation): the same kernels as
 * cwp_hploc_build, with their scratch in a fixed static device array (16 MiB per GPU, held until the process ends) for
 * at most CWP_HPLOC_STATIC_MAX primitives (larger n is rejected with CWP_ERR_ARG). lo / hi: device arrays (4 floats
 * per box); nodes: the caller's device buffer of 2n - 1 nodes (8 floats each); asynchronous on `stream` like
 * cwp_hploc_build. Every call on a GPU shares that GPU's scratch: calls that overlap on one GPU (other streams or
 * threads) race. */
enum { CWP_HPLOC_STATIC_MAX = 65536 };
int cwp_hploc_build_static(const void* lo, const void* hi, int n, int morton_bits, void* nodes, void* stream, int* root);
int cwp_hploc_nodes(const cwp_hploc* b, const void** nodes);
/* 64 B BVH2 layout (16 floats per node, max(n - 1, 1) nodes, root n - 2 or 0) of device H-PLOC nodes into `out`
 * (device). Replaces hplocToBVH2 (src/accel/hploc.h:36). */
int cwp_hploc_to_bvh2(const void* hnodes, int n, void* out, void* stream);
/* Compressed BVH8 (80-byte nodes, CWP_ARR_NODES layout) from H-PLOC nodes of n primitives, on the device: the
 * collapse of cwp_bvh_from_hploc with variant 8 (one primitive per binary leaf, SAH-optimal 8-wide collapse, octant
 * slot order, quantised child boxes), bit-identical to it. hnodes, nodes (room for 20 floats per primitive) and prims
 * (n ints: the leaf primitive order, i.e. CWP_ARR_ORDER) are device pointers; the call synchronises with `stream`
 * once per level of the wide tree. Replaces treeFromHploc + CWBVH::build on the host (src/accel/bvh_tools.cpp:12-168,
 * src/app/scene_setup.cpp:76-83). */
int cwp_cwbvh_from_hploc(const void* hnodes, int n, void* nodes, int* num_nodes, void* prims, int* depth, void* stream);
/* Host convenience for the same: uploads, collapses (timed with events: gpu_ms), downloads. */
int cwp_cwbvh_from_hploc_host(const float* hnodes, int n, float* nodes, int* num_nodes, int* prims, int* depth, double* gpu_ms);

/* ---- render-job server: batched multi-job passes, continuous batching, resident scenes with LRU eviction ----
 * Jobs (scene, resolution, samples, bounce limit, seed, arrival time) are admitted in arrival order once arrived,
 * their scene resident (within the scene budget, evicting least recently used scenes no admitted job needs) and their
 * paths fit the free path-state slots; each step is one pass over every admitted job (one launch per bounce for all of
 * them); finished jobs leave with their image and newly arrived ones are admitted. Plain mode: the same kernels, one
 * job at a time. A job's image does not depend on the mode or on the other jobs. */
typedef struct cwp_server cwp_server;
typedef struct cwp_server_options {
    int variant, tri_test, stack_lds;  /* every scene's BVH layout, the triangle test, LDS stack entries */
    int bounces;                       /* deepest bounce limit of any job */
    size_t max_paths;                  /* path-state slots (paths in flight over all admitted jobs) */
    int max_jobs;                      /* jobs per pass (<= 256) */
    double scene_budget_mb;            /* device memory for resident scenes */
    int batched;                       /* 1: continuous batching; 0: plain, one job at a time */
    int rays_per_lane;                 /* 1, 2 or 4: persistent grid = resident waves / rays_per_lane */
    int use_graph;                     /* 1: each pass replays a captured HIP graph */
    int full;                          /* 1: full material integrator (legacy setups render their Lambert materials) */
    int rolling;                       /* batched mode: 1 = admission at every step (one trace+shade and one shadow
                                          launch over all live paths of all live jobs, whatever their depth, plus the
                                          camera rays of the jobs admitted at that step); 0 = between whole passes */
    int compact;                       /* rolling steps' queue compaction: 0 per wave (one ballot and one atomic per
                                          wave), 1 per lane (one atomic per appended ray) */
    int tail_lanes;                    /* straggler hand-off (BVH8 only) of the trace+shade launches, and with the
                                          full integrator of the shadow launches: a wave whose lanes still tracing
                                          drop below this many hands their traversals to a tail launch that resumes
                                          them in full waves (images unchanged); -1 automatic
                                          (12 with BVH8, else off), 0 off, 2..64; cwp_server_get_options reports
                                          the resolved value. The autotuner keeps it for BVH8 trials and turns it
                                          off for the others. */
    int tail_levels;                   /* 2 (default): the rolling steps' trace+shade tail launches hand their own
                                          stragglers on to a second tail launch; 1: they resume every record to the
                                          end (as the pass mode's always do: a second level made it slower) */
} cwp_server_options;
typedef struct cwp_tune_trial {
    int variant, stack_lds, rays_per_lane, compact;
    double rays_per_s, wall_s;
} cwp_tune_trial;
typedef struct cwp_job_info {
    int scene, width, height, spp, bounces, samples_per_pass, passes;
    double arrival_s, admitted_s, finished_s, latency_s;   /* seconds after the start of cwp_server_run */
} cwp_job_info;
typedef struct cwp_serve_stats {
    int jobs, passes, scenes_loaded, scenes_evicted;   /* passes: rounds of K samples admitted */
    double wall_s, rays, rays_per_s, latency_p50_s, latency_p99_s, latency_mean_s, max_paths_in_flight;
    long long steps;                                   /* rolling steps queued */
} cwp_serve_stats;
void cwp_server_options_default(cwp_server_options* o);
int cwp_server_create(const cwp_server_options* o, cwp_server** out);
void cwp_server_destroy(cwp_server* s);
/* copies the setup's scene, camera and lighting (a single-level scene in any BVH layout: the server builds its own
 * layout from the mesh when they differ); scene_bytes (optional) receives the device bytes the scene takes when
 * resident in the server's layout, counted against the scene budget */
int cwp_server_add_scene(cwp_server* s, const cwp_setup* setup, int* scene_id, size_t* scene_bytes);
int cwp_server_submit(cwp_server* s, int scene_id, int width, int height, int spp, int bounces, uint32_t seed,
                      double arrival_s, int* job_id);
/* processes every submitted unfinished job; batched overrides the option for this run */
int cwp_server_run(cwp_server* s, int batched);
int cwp_server_job(const cwp_server* s, int job_id, cwp_job_info* info);
/* per-pixel radiance sums of a finished job (divide by spp), float[4] per pixel */
int cwp_server_image(const cwp_server* s, int job_id, float* rgba, size_t num_pixels);
int cwp_server_stats(const cwp_server* s, cwp_serve_stats* stats);
int cwp_server_clear_jobs(cwp_server* s);
/* marks every job unfinished again (images dropped), e.g. to run the same stream in the other mode */
int cwp_server_reset_jobs(cwp_server* s);
/* Autotuner: runs the submitted jobs, batched, under every combination of BVH width (variants: 2, 4, 8), LDS stack
 * entries (8, 16, 32), rays per lane (1, 2, 4) and queue-compaction granularity (compacts: 0 per wave, 1 per lane;
 * null = {0}), keeps the fastest and leaves the jobs unfinished. trials receives up to max_trials results
 * (num_trials = the number run); chosen (optional) receives the server's new options. */
int cwp_server_autotune(cwp_server* s, const int* variants, int num_variants, const int* stacks, int num_stacks,
                        const int* rays_per_lane, int num_rays_per_lane, const int* compacts, int num_compacts,
                        cwp_tune_trial* trials, int max_trials, int* num_trials, cwp_server_options* chosen);
int cwp_server_get_options(const cwp_server* s, cwp_server_options* o);
/* Live serving. No counterpart in the reference, which renders one job per process. cwp_server_start runs the rolling
 * loop on a worker thread of the library; from then on cwp_job_submit (the arrival is the time of the call),
 * cwp_job_poll and cwp_job_result may be called from any thread while other jobs are on the GPU; cwp_server_stop
 * waits until every submitted job is done and joins the worker. cwp_server_run / _autotune / _preload and the
 * per-job reads cwp_server_job / _image are for a stopped server. */
enum { CWP_JOB_QUEUED = 0, CWP_JOB_RUNNING = 1, CWP_JOB_DONE = 2 };
int cwp_server_start(cwp_server* s);
int cwp_job_submit(cwp_server* s, int scene_id, int width, int height, int spp, int bounces, uint32_t seed, int* job_id);
int cwp_job_poll(const cwp_server* s, int job_id, int* state);
/* per-pixel radiance sums of a done job (divide by spp), float[4] per pixel; CWP_ERR_STATE while it is not done */
int cwp_job_result(const cwp_server* s, int job_id, float* rgba, size_t num_pixels);
int cwp_server_stop(cwp_server* s);
/* uploads the scenes of the unfinished jobs, in arrival order, while they fit the scene budget (so that a run's
 * timing excludes the first uploads) */
int cwp_server_preload(cwp_server* s);

/* ---- material response on the host (the code the full-integrator kernels run, compiled for the CPU) ----
 * mat: the 64-byte material record (16 floats: baseColor, metallic, specularColor, roughness, emission, ior,
 * specularWeight, transmission, 2 unused). Local frame z = shading normal, wo.z > 0. eta = index ratio of the
 * dielectric base for this side; gd_energy / ux1 / ux2 = glossy-diffuse energy compensation and its random numbers.
 * Replace the reference's bsdfEval / bsdfSample (src/render/bsdf.h:196-220) for host-side checks. */
int cwp_bsdf_eval(const float* mat, const float wo[3], const float wi[3], float eta, int gd_energy, float ux1, float ux2,
                  float f[3], float* pdf);
/* returns CWP_OK and sets *ok (1: a direction was sampled; pdf 0 marks a delta sample) */
int cwp_bsdf_sample(const float* mat, const float wo[3], float eta, int gd_energy, float ux1, float ux2, float u, float u1,
                    float u2, float wi[3], float weight[3], float* pdf, int* ok);

/* ---- textures on the host (the readers and the lookup the renderer uses) ----
 * cwp_image_read: a PNG (every colour type and bit depth, Adam7), TGA (true colour / grey, RLE) or PPM / PGM file as
 * RGBA8 texels, r in the low byte, rows top first; with rgba NULL only w and h are set. Replaces the reference's
 * stb_image loads (src/io/assimp_loader.cpp:42-66). */
int cwp_image_read(const char* path, int* w, int* h, uint32_t* rgba, size_t capacity);
/* zlib stream -> bytes (at most capacity; CWP_ERR_ARG on malformed input, a checksum mismatch or more output) */
int cwp_inflate(const uint8_t* in, size_t n, uint8_t* out, size_t capacity, size_t* out_len);
/* RGBA at texture coordinates (u, v) of an image (first texel, width, height, sRGB) in `texels`: the lookup of the
 * full-integrator kernels (wrap, nearest or bilinear, sRGB decode). Replaces the reference's tex2D lookups
 * (src/render/wavefront.cu:203) for host-side checks. */
int cwp_texture_sample(const int image[4], const uint32_t* texels, size_t num_texels, float u, float v, int nearest,
                       float rgba[4]);

/* ---- device memory for callers without a GPU array library ---- */
enum { CWP_COPY_H2D = 1, CWP_COPY_D2H = 2, CWP_COPY_D2D = 3 };
int cwp_device_malloc(size_t bytes, void** ptr);
void cwp_device_free(void* ptr);
/* synchronous copy; kind: CWP_COPY_H2D, CWP_COPY_D2H or CWP_COPY_D2D */
int cwp_memcpy(void* dst, const void* src, size_t bytes, int kind);

#ifdef __cplusplus
}
#endif

#endif /* CWP_H */
