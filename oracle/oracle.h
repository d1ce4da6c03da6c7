/* oracle.h - CPU restatement of the reference's algorithms for the wavefront hot path (plain C11).
 *
 * TEST INFRASTRUCTURE ONLY. This library is the parity checker: the tests, the smoke check and the CPU-baseline leg
 * of bench.py call it to compute expected results. The product (libcwp.so, the pt CLI) never links or calls it.
 *
 * Every function restates one reference function (paths relative to the reference repo); the pins that check each
 * restatement independently of the product are listed in the header of oracle.c.
 */
#ifndef CWP_ORACLE_H
#define CWP_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* same layout as cwp_params in include/cwp.h */
typedef struct orc_params {
    float pos[3], ax[3], ay[3], az[3];
    float fx, fy, px, py;
    int w, h;
    float lens_r, focus_d;
    float sun_to[3], sun_e[3], sky[3];
    float sun_cos;
    int max_bounce;
    uint32_t seed;
    float eps, near_t;
    int has_normals, filter_bh;
    float ftab[65];
} orc_params;

typedef struct orc_scene {
    int variant;            /* 2: BVH2 64 B nodes, 4: BVH4 128 B nodes, 8: compressed BVH8 80 B nodes */
    int tri_test;           /* 0 Moller-Trumbore, 1 Woop rows, 2 watertight */
    const float* nodes;     /* float4 array */
    int root;
    const float* woop;      /* 3 float4 per leaf position */
    const float* corners;   /* 3 float4 per leaf position: (v0, id), (v1, material), (v2, 0) */
    const float* normals;   /* 3 float4 per leaf position or NULL */
    const float* albedo;    /* float4 per material */
    int num_tris;
    /* two-level scenes (tlas != NULL): BVH2 TLAS over instance ids, its root, instance records (28 floats each:
     * world-to-object rows, object-to-world rows, BLAS root bits, mesh bits, det sign, pad); nodes then hold every
     * BLAS with absolute indices */
    const float* tlas;
    int tlas_root;
    const float* inst;
    int num_inst;
    /* full material integrator (full != 0): materials (16 floats each: base rgb, metallic, specular rgb, roughness,
     * emission rgb, ior, specular weight, transmission, 0, 0), emitter tables (orc_emitter_tables), environment map
     * (orc_env_build; env_w = 0: none, the sky constant is used), light-selection probabilities and options */
    int full;
    const float* mats;
    int num_mats;
    const float* tri_light_pdf;
    const int* em_tri;
    const float* em_cdf;
    int num_em;
    const float* env_tex;
    const float* env_marg;
    const float* env_cond;
    int env_w, env_h;
    float env_rot;
    float p_sun, p_tri, p_env;
    int light_mode, rr, rr_depth, gd_energy, env_uniform;
    /* two-level scenes with the full integrator: the frame's world emitters (orc_inst_emitters) and per-material
     * light pdf per unit area */
    const float* em_tris;
    const float* mat_light_pdf;
    /* textures (maps != NULL): 8 ints per material (base, emission, normal, metallic, roughness image ids, metallic /
     * roughness channel, normal strength float bits), corner uvs (8 floats per leaf position), tangents (12 floats per
     * leaf position, or NULL), images (first texel, w, h, srgb), RGBA8 texels, the 256-entry sRGB decode table,
     * nearest-texel filter, terminator shadowing of normal-mapped diffuse shading */
    const int* maps;
    int num_maps;
    const float* uvs;
    const float* tans;
    const int* tex_images;
    const uint32_t* texels;
    const float* srgb;
    int tex_nearest, bump_shadow;
} orc_scene;

/* binary SAH tree: node = int[4] (left, right, first, count) + float[6] box (lo, hi); prim = triangle order */
typedef struct orc_tree {
    int num_nodes, num_prims, max_depth;
    int* node;
    float* box;
    int* prim;
} orc_tree;

typedef struct orc_layout {
    int num_float4;         /* node array length in float4 */
    float* nodes;
    int num_prims;
    int* order;             /* leaf position -> triangle id */
    int depth;              /* wide-tree levels (CWBVH) or internal-node levels (BVH2) */
    int max_leaf;
    double sah;
} orc_layout;

/* ---- RNG and warps: src/core/rng.h:6-12, src/core/sampling.h:13-58 ---- */
uint32_t orc_hash32(uint32_t x);
float orc_rng4(uint32_t pix, uint32_t smp, uint32_t depth, uint32_t dim, uint32_t seed);
void orc_blackman_harris_table(float* tab, int n, double width);
void orc_cosine_hemisphere(const float n[3], float u0, float u1, float out[3]);
void orc_uniform_cone(const float axis[3], float cos_max, float u0, float u1, float out[3]);

/* ---- BVH construction: src/accel/bvh_builder.cpp:5-68, bvh_layouts.cpp:6-30, bvh_tools.cpp:75-168, tri_pack.cpp ---- */
int orc_build_sah(const float* v, int nv, const int* tri, int nt, int bins, int leaf_max, orc_tree* out);
void orc_tree_free(orc_tree* t);
int orc_layout_bvh2(const orc_tree* t, orc_layout* out);
int orc_collapse_cwbvh(const orc_tree* t, orc_layout* out);
int orc_layout_bvh4(const orc_tree* t, orc_layout* out);
void orc_layout_free(orc_layout* l);
void orc_pack_woop(const float* v, const int* tri, const int* order, int n, float* rows);
void orc_pack_corners(const float* v, const int* tri, const int* mat, const int* order, int n, float* corners);

/* ---- H-PLOC GPU builder and its SAH leaf collapse: src/accel/hploc.cu:11-160, src/accel/bvh_tools.cpp:12-38 ---- */
/* lo / hi: float4 per primitive; nodes: 8 floats per node, (lo.xyz, left id bits), (hi.xyz, right id bits); leaves
 * [0, n) in Morton order with left = primitive id, right = -1; internal nodes [n, 2n - 1) */
void orc_hp_centroid_box(const float* lo, const float* hi, int n, float box[6]);
uint64_t orc_expand_bits21(uint32_t v);
void orc_hp_key_bits(int n, int* mb, int* ib);
/* morton: 0 = adaptive bits per axis (the reference), 30 = 10 bits per axis */
void orc_hp_key_bits_mode(int n, int morton, int* mb, int* ib);
void orc_hp_keys(const float* lo, const float* hi, int n, const float box[6], int mb, int ib, uint64_t* keys);
/* full build (keys, sort, leaves, H-PLOC); sorted_keys (n) may be NULL; returns the root id (2n - 2, or 0 for n = 1) */
int orc_hp_build(const float* lo, const float* hi, int n, float* nodes, uint64_t* sorted_keys);
int orc_hp_build_mode(const float* lo, const float* hi, int n, int morton, float* nodes, uint64_t* sorted_keys);
/* BVH2 layout (16 floats per node, max(n - 1, 1) nodes, root n - 2 or 0) */
void orc_hp_to_bvh2(const float* hn, int n, float* out);
/* binary tree with SAH-collapsed leaves (<= leaf_max primitives), pre-order; prim = primitive ids */
int orc_tree_from_hploc(const float* hn, int n, int root, int leaf_max, orc_tree* out);

/* ---- traversal: src/accel/traverse.cuh ---- */
/* rays: 8 floats (o, tmin, d, tmax); hits: 4 floats (t, u, v, leaf position bits, -1 = miss); returns 0 on success */
int orc_trace(const orc_scene* s, const float* rays, int n, int any_hit, float* hits, int threads);
/* the same with the hit instance of two-level scenes (-1 on a miss) */
int orc_trace_inst(const orc_scene* s, const float* rays, int n, int any_hit, float* hits, int* inst, int threads);
/* world boxes of instances (TLAS input): the 8 corners of the mesh box (mesh_box: lo, hi float4 per mesh) through the
 * object-to-world rows, min / max (reference: k_inst_boxes, src/render/wavefront.cu:445-451) */
void orc_inst_boxes(const float* inst, int n, const float* mesh_box, float* lo, float* hi);
/* closest hit by testing every triangle (same triangle test), for pinning the traversal */
int orc_trace_brute(const orc_scene* s, const float* rays, int n, float* hits, int threads);
/* node visits and triangle tests of a closest-hit traversal, summed over the rays (cache-side traffic model) */
int orc_trace_work(const orc_scene* s, const float* rays, int n, double* node_visits, double* tri_tests);

/* ---- legacy integrator: src/render/wavefront.cu:42-58 (raygen), 94-136 (shadeOne), 67-72 (shadow) ---- */
/* Renders spp samples (passes of `batch` samples; per-path radiance summed per pass in sample order) into accum
 * (float4 per pixel, += unless clear). stats[0..3] += closest rays, shadow rays, misses, unoccluded shadow rays.
 * Pixels rows are split over `threads` threads; each pixel is owned by one thread. */
/* ---- full material integrator: src/render/bsdf.h, src/render/wavefront.cu:177-356, src/scene/environment.cpp,
 * src/app/scene_setup.cpp:225-247 ---- */
/* f (3) and pdf of a material (16 floats) for local directions wo, wi (z = shading normal) */
/* the texture lookup of the full integrator: image (first texel, w, h, srgb), RGBA8 texels, sRGB decode table */
void orc_texture_sample(const int img[4], const uint32_t* texels, const float* srgb_table, float u, float v, int nearest,
                        float out[4]);
void orc_bsdf_eval(const float* mat, const float wo[3], const float wi[3], float eta, int gd_energy, float ux1, float ux2,
                   float f[3], float* pdf);
/* returns 1 and wi, weight = f|cos|/pdf, pdf (0: delta) on success */
int orc_bsdf_sample(const float* mat, const float wo[3], float eta, int gd_energy, float ux1, float ux2, float u, float u1,
                    float u2, float wi[3], float weight[3], float* pdf);
/* n samples with a seeded splitmix64 stream (5 numbers per sample: u, u1, u2, ux1, ux2): wi, weight, pdf, ok */
void orc_bsdf_sample_batch(const float* mat, const float wo[3], float eta, int gd_energy, int n, uint64_t seed, float* wi,
                           float* weight, float* pdf, int* ok);
/* n evaluations (ux1 = ux2 = 0.5): pdf only */
void orc_bsdf_pdf_batch(const float* mat, const float wo[3], float eta, int gd_energy, const float* wi, int n, float* pdf);
float orc_fresnel_dielectric(float cos_i, float eta);
void orc_fresnel_f82(const float f0[3], const float tint[3], float mu, float out[3]);
float orc_ggx_d(float cos_m, float a2);
float orc_ggx_g2(float cos_o, float cos_i, float a2);
/* deterministic directional albedo by quadrature over microfacet normals: kind 0 conductor F = 1, 1 dielectric (R + T),
 * 2 the glossy-diffuse specular lobe alone (F = Fdiel) */
double orc_quad_albedo(int kind, float rough, float eta, const float wo[3]);
/* environment distribution of w x h rgb texels (row 0 = +z): texels rgba * strength, marginal (h + 1) and conditional
 * (h rows of w + 1) CDFs over luminance * sin(theta) */
void orc_env_build(const float* rgb, int w, int h, float strength, float* tex, float* marg, float* cond);
/* solid-angle pdf of the environment's luminance sampling in direction d (s->env_* set) */
float orc_env_pdf(const orc_scene* s, const float d[3]);
/* n luminance samples of the environment (u pairs from a seeded stream): directions and their pdf */
void orc_env_sample_batch(const orc_scene* s, int n, uint64_t seed, float* dirs, float* pdf);
/* emissive triangles by power (mesh order = order[] ascending); returns the emitter count; em_tri (n), em_cdf (n + 1),
 * tri_light_pdf (num_tris) */
int orc_emitter_tables(const float* corners, const int* order, int num_tris, const float* mats, int* em_tri, float* em_cdf,
                       float* tri_light_pdf);

/* two-level scenes: world-space emissive triangles of the given instance records (28 floats each), in instance order
 * then mesh triangle order (mesh_tris: first leaf and count per mesh; order: leaf -> first leaf of its mesh + local
 * triangle), 12 floats each (v0, leaf bits, e1, material bits, e2, 0); power CDF (n + 1); mat_light_pdf (num_mats).
 * Call with em_tris = NULL to get the count. */
int orc_inst_emitters(const float* inst, int n_inst, const int* mesh_tris, const float* corners, const int* order,
                      const float* mats, int num_mats, float* em_tris, float* em_cdf, float* mat_light_pdf);

int orc_render(const orc_scene* s, const orc_params* p, int spp, int sample_base, int batch, int bounces, int threads,
               int clear, float* accum, double* stats);
/* the camera ray of (pixel, sample): origin and direction */
void orc_camera_ray(const orc_params* p, int pix, uint32_t smp, float o[3], float d[3]);

/* the viewer's display transform (src/app/display.cu:9-33): n rgba sums of spp samples (spp 0 read as 1) -> n RGBA8
 * pixels (4 bytes each, a = 255); mode 0 clip, 1 Reinhard, 2 ACES; exposure in stops (scale 2^ev unless ev == 0) */
void orc_display(const float* acc, int n, int spp, float exposure_ev, int mode, uint8_t* out);
/* the 8-bit encode of one value: clip to [0, 1], sRGB OETF, x * 255 + 0.5 truncated (src/app/display.cu:9-13) */
int orc_srgb_encode8(float x);
/* every float in [lo, hi] (as bit patterns, lo <= hi, both in [0, 1]) checked against 255 step positions (steps[k]:
 * the smallest float that encodes to k or more): returns the number of floats whose encode differs from the search */
uint64_t orc_srgb_steps_check(const float* steps, float lo, float hi, int threads);

#ifdef __cplusplus
}
#endif
#endif
