// capi.cpp - the extern "C" entry points of libcwp.so (declared in include/cwp.h).
#include "cwp.h"
#include "core/bsdf.h"
#include "device/renderer.h"
#include "host/bvh.h"
#include "host/envmap.h"
#include "host/io.h"
#include "host/scene_check.h"
#include "device/collapse.h"
#include "device/display.h"
#include "device/hploc.h"
#include "host/serve.h"
#include "host/setup.h"
#include "host/texture.h"
#include <cstddef>
#include <cstring>
#include <new>
#include <string>

using namespace cwp;

struct cwp_setup { SceneSetup Z; std::vector<int> triFlat; std::vector<int> tree; std::vector<float> treeBox; };
struct cwp_ctx { Renderer R; const cwp_setup* setup = nullptr; };
struct cwp_hploc { HplocBuilder B; };
struct cwp_server { Server S; };
struct cwp_bvh {
    BinaryBVH tree; std::vector<Vec4> nodes, woop, corners, nrm; std::vector<int> order, treeFlat; std::vector<float> treeBox;
    int variant = 2, stackNeed = 0; double sah = 0;
};

static_assert(sizeof(cwp_params) == sizeof(FrameParams), "cwp_params must mirror FrameParams");
static_assert(CWP_LIGHT_MIS == kLightMIS && CWP_LIGHT_NEE == kLightNEE && CWP_LIGHT_BSDF == kLightBSDF, "light modes");
static_assert(CWP_MORTON_ADAPTIVE == kMortonAdaptive && CWP_MORTON_30 == kMorton30, "Morton modes");
static_assert(CWP_HPLOC_STATIC_MAX == kHplocStaticMax, "allocation-free build limit");
static_assert(sizeof(cwp_instance) == sizeof(InstanceRecord), "cwp_instance must mirror InstanceRecord");
static_assert(offsetof(cwp_params, sun_to) == offsetof(FrameParams, sunTo), "cwp_params layout");
static_assert(offsetof(cwp_params, seed) == offsetof(FrameParams, seed), "cwp_params layout");
static_assert(offsetof(cwp_params, ftab) == offsetof(FrameParams, ftab), "cwp_params layout");

static thread_local std::string g_err;
static int fail(int code, const std::string& msg) { g_err = msg; return code; }

extern "C" {

const char* cwp_last_error(void) { return g_err.c_str(); }
int cwp_version(void) { return 1; }

int cwp_device_info(int* num_cu, char* name, size_t cap) {
    int n = 0; std::string nm, err;
    if (!deviceInfo(n, nm, err)) return fail(CWP_ERR_DEVICE, err);
    if (num_cu) *num_cu = n;
    if (name && cap) { strncpy(name, nm.c_str(), cap - 1); name[cap - 1] = 0; }
    return CWP_OK;
}

int cwp_setup_create(int argc, const char* const* argv, cwp_setup** out) {
    if (!out || argc < 0 || (argc > 0 && !argv)) return fail(CWP_ERR_ARG, "cwp_setup_create: null argument");
    *out = nullptr;
    cwp_setup* s = new (std::nothrow) cwp_setup();
    if (!s) return fail(CWP_ERR_ARG, "out of host memory");
    std::string err;
    if (!parseArgs(argc, argv, s->Z.A, err)) { delete s; return fail(CWP_ERR_ARG, err); }
    if (!setupScene(s->Z, err)) { delete s; return fail(CWP_ERR_IO, err); }
    const Mesh& M = s->Z.M;
    s->triFlat.resize(M.f.size() * 3);
    for (size_t t = 0; t < M.f.size(); t++) { s->triFlat[3 * t] = M.f[t].a; s->triFlat[3 * t + 1] = M.f[t].b; s->triFlat[3 * t + 2] = M.f[t].c; }
    if (s->Z.M.fm.empty()) s->Z.M.fm.assign(M.f.size(), 0);
    for (const BinNode& n : s->Z.tree.nodes) {
        s->tree.insert(s->tree.end(), { n.left, n.right, n.first, n.count });
        s->treeBox.insert(s->treeBox.end(), { n.box.lo.x, n.box.lo.y, n.box.lo.z, n.box.hi.x, n.box.hi.y, n.box.hi.z });
    }
    *out = s;
    return CWP_OK;
}

void cwp_setup_destroy(cwp_setup* s) { delete s; }

int cwp_setup_info(const cwp_setup* s, cwp_scene_info* info) {
    if (!s || !info) return fail(CWP_ERR_ARG, "cwp_setup_info: null argument");
    const SceneSetup& Z = s->Z;
    memset(info, 0, sizeof *info);
    info->variant = Z.variant; info->tri_test = Z.O.tri;
    info->num_tris = (int)Z.M.f.size(); info->num_verts = (int)Z.M.v.size(); info->num_materials = (int)Z.M.kd.size();
    info->node_floats4 = nodeFloat4s(Z.variant); info->num_nodes = (int)Z.nNodes;
    info->bvh_depth = Z.maxDepth; info->stack_need = Z.S.stackNeed; info->sah = Z.sah; info->bvh_build_s = Z.bvhS;
    info->width = Z.A.w; info->height = Z.A.h; info->spp = Z.A.spp; info->batch = Z.A.batch; info->bounces = Z.A.bounces;
    info->num_meshes = Z.inst ? (int)Z.meshes.size() : 0; info->num_instances = Z.inst ? (int)Z.SD.instances.size() : 0;
    info->frames = Z.frames; info->frame_start = Z.A.frameStart;
    const SceneUpload& S = Z.S;
    info->full = S.full; info->num_emitters = S.inst ? S.numEmInst : (int)S.emTri.size();
    info->env_w = S.env ? S.env->w : 0; info->env_h = S.env ? S.env->h : 0; info->sun_on = S.sunOn;
    info->light_mode = S.lightMode; info->rr = S.rr; info->rr_depth = S.rrDepth; info->gd_energy = S.gdEnergy;
    info->env_uniform = S.envUniform; info->env_rot_deg = S.envRotDeg; info->bvh_cache = Z.bvhCacheHit;
    info->textured = S.maps.empty() ? 0 : 1; info->num_images = (int)S.texImages.size();
    info->tex_nearest = S.texNearest; info->bump_shadow = S.bumpShadow;
    return CWP_OK;
}

int cwp_setup_params(const cwp_setup* s, cwp_params* p) {
    if (!s || !p) return fail(CWP_ERR_ARG, "cwp_setup_params: null argument");
    memcpy(p, &s->Z.P, sizeof *p);
    return CWP_OK;
}

int cwp_setup_array(const cwp_setup* s, int which, const void** data, size_t* count) {
    if (!s || !data || !count) return fail(CWP_ERR_ARG, "cwp_setup_array: null argument");
    const SceneSetup& Z = s->Z;
    switch (which) {
        case CWP_ARR_VERTICES: *data = Z.M.v.data(); *count = Z.M.v.size(); break;
        case CWP_ARR_TRIANGLES: *data = s->triFlat.data(); *count = Z.M.f.size(); break;
        case CWP_ARR_MATERIAL: *data = Z.M.fm.data(); *count = Z.M.fm.size(); break;
        case CWP_ARR_ALBEDO: *data = Z.S.albedo.data(); *count = Z.S.albedo.size(); break;
        case CWP_ARR_NODES: *data = Z.S.nodes.data(); *count = Z.nNodes; break;
        case CWP_ARR_ORDER: *data = Z.S.order.data(); *count = Z.S.order.size(); break;
        case CWP_ARR_WOOP: *data = Z.S.woop.data(); *count = Z.S.woop.size() / 3; break;
        case CWP_ARR_CORNERS: *data = Z.S.verts.data(); *count = Z.S.verts.size() / 3; break;
        case CWP_ARR_NORMALS: *data = Z.M.n.data(); *count = Z.M.n.size(); break;
        case CWP_ARR_TREE: *data = s->tree.data(); *count = s->tree.size() / 4; break;
        case CWP_ARR_TREE_BOX: *data = s->treeBox.data(); *count = s->treeBox.size() / 6; break;
        case CWP_ARR_PRIM: *data = Z.tree.prim.data(); *count = Z.tree.prim.size(); break;
        case CWP_ARR_MESH_ROOTS: *data = Z.S.meshRoot.data(); *count = Z.S.meshRoot.size(); break;
        case CWP_ARR_MESH_BOXES: *data = Z.S.meshBox.data(); *count = Z.S.meshBox.size() / 2; break;
        case CWP_ARR_MATERIALS: *data = Z.S.mats.data(); *count = Z.S.mats.size(); break;
        case CWP_ARR_EM_TRI: *data = Z.S.emTri.data(); *count = Z.S.emTri.size(); break;
        case CWP_ARR_EM_CDF: *data = Z.S.emCdf.data(); *count = Z.S.emCdf.size(); break;
        case CWP_ARR_TRI_LIGHT_PDF: *data = Z.S.triLightPdf.data(); *count = Z.S.triLightPdf.size(); break;
        case CWP_ARR_ENV_TEX: *data = Z.S.env ? (const void*)Z.S.env->texels.data() : nullptr; *count = Z.S.env ? Z.S.env->texels.size() : 0; break;
        case CWP_ARR_ENV_MARG: *data = Z.S.env ? (const void*)Z.S.env->marg.data() : nullptr; *count = Z.S.env ? Z.S.env->marg.size() : 0; break;
        case CWP_ARR_MESH_TRIS: *data = Z.S.meshTris.data(); *count = Z.S.meshTris.size() / 2; break;
        case CWP_ARR_ENV_COND: *data = Z.S.env ? (const void*)Z.S.env->cond.data() : nullptr; *count = Z.S.env ? Z.S.env->cond.size() : 0; break;
        case CWP_ARR_UVS: *data = Z.S.uvs.data(); *count = Z.S.uvs.size() / 2; break;
        case CWP_ARR_TANGENTS: *data = Z.S.tans.data(); *count = Z.S.tans.size() / 3; break;
        case CWP_ARR_MAPS: *data = Z.S.maps.data(); *count = Z.S.maps.size(); break;
        case CWP_ARR_TEX_IMAGES: *data = Z.S.texImages.data(); *count = Z.S.texImages.size(); break;
        case CWP_ARR_TEXELS: *data = Z.S.texels.data(); *count = Z.S.texels.size(); break;
        case CWP_ARR_SRGB_TABLE: *data = srgbDecodeTable(); *count = 256; break;
        case CWP_ARR_MESH_UVS: *data = Z.M.uv.data(); *count = Z.M.hasUV() ? Z.M.f.size() : 0; break;
        default: return fail(CWP_ERR_ARG, "cwp_setup_array: unknown array");
    }
    return CWP_OK;
}

int cwp_create(const cwp_setup* s, cwp_ctx** out) {
    if (!s || !out) return fail(CWP_ERR_ARG, "cwp_create: null argument");
    *out = nullptr;
    cwp_ctx* c = new (std::nothrow) cwp_ctx();
    if (!c) return fail(CWP_ERR_ARG, "out of host memory");
    std::string err;
    if (!c->R.init(s->Z.S, s->Z.P, s->Z.O, err)) { delete c; return fail(CWP_ERR_DEVICE, err); }
    c->setup = s;
    if (s->Z.inst) {
        std::vector<InstanceRecord> rec;
        frameInstances(s->Z, s->Z.A.frameStart, rec);
        double ms = 0;
        if (!c->R.setInstances(rec.data(), (int)rec.size(), ms, err)) { delete c; return fail(CWP_ERR_DEVICE, err); }
        if (s->Z.full) {
            std::vector<Vec4> em;
            std::vector<float> cdf, mlp;
            frameEmitters(s->Z, s->Z.A.frameStart, em, cdf, mlp);
            if (!c->R.setEmitters(em.data(), (int)(em.size() / 3), cdf.data(), mlp.data(), (int)mlp.size(), err)) {
                delete c; return fail(CWP_ERR_DEVICE, err);
            }
        }
    }
    *out = c;
    return CWP_OK;
}

int cwp_setup_instances(const cwp_setup* s, int frame, cwp_instance* out, int cap, int* count) {
    if (!s || !count || (cap > 0 && !out) || frame < 0) return fail(CWP_ERR_ARG, "cwp_setup_instances: bad arguments");
    std::vector<InstanceRecord> rec;
    if (!frameInstances(s->Z, frame, rec)) return fail(CWP_ERR_STATE, "cwp_setup_instances: the scene is not two-level");
    *count = (int)rec.size();
    if (cap > 0) memcpy(out, rec.data(), std::min<size_t>((size_t)cap, rec.size()) * sizeof(InstanceRecord));
    return CWP_OK;
}

int cwp_set_instances(cwp_ctx* c, const cwp_instance* inst, int n, double* tlas_ms) {
    if (!c || !inst || n < 1) return fail(CWP_ERR_ARG, "cwp_set_instances: bad arguments");
    double ms = 0; std::string err;
    if (!c->R.setInstances((const InstanceRecord*)inst, n, ms, err))
        return fail(err.find("hip") != std::string::npos ? CWP_ERR_DEVICE : CWP_ERR_ARG, err);
    if (tlas_ms) *tlas_ms = ms;
    return CWP_OK;
}

int cwp_set_frame(cwp_ctx* c, int frame, double* tlas_ms) {
    if (!c || !c->setup || frame < 0) return fail(CWP_ERR_ARG, "cwp_set_frame: needs a context made from a setup");
    std::vector<InstanceRecord> rec;
    if (!frameInstances(c->setup->Z, frame, rec)) return fail(CWP_ERR_STATE, "cwp_set_frame: the scene is not two-level");
    const int rc = cwp_set_instances(c, (const cwp_instance*)rec.data(), (int)rec.size(), tlas_ms);
    if (rc != CWP_OK || !c->setup->Z.full) return rc;
    std::vector<Vec4> em;
    std::vector<float> cdf, mlp;
    frameEmitters(c->setup->Z, frame, em, cdf, mlp);
    return cwp_set_emitters(c, (const float*)em.data(), (int)(em.size() / 3), cdf.data(), mlp.data(), (int)mlp.size());
}

int cwp_set_emitters(cwp_ctx* c, const float* em_tris, int num_emitters, const float* em_cdf, const float* mat_light_pdf,
                     int num_materials) {
    if (!c) return fail(CWP_ERR_ARG, "cwp_set_emitters: null context");
    std::string err;
    if (!c->R.setEmitters((const Vec4*)em_tris, num_emitters, em_cdf, mat_light_pdf, num_materials, err)) return fail(CWP_ERR_ARG, err);
    return CWP_OK;
}

int cwp_setup_emitters(const cwp_setup* s, int frame, float* em_tris, int cap, int* num_emitters, float* em_cdf,
                       float* mat_light_pdf, int mat_cap) {
    if (!s || !num_emitters || frame < 0) return fail(CWP_ERR_ARG, "cwp_setup_emitters: bad argument");
    std::vector<Vec4> em;
    std::vector<float> cdf, mlp;
    frameEmitters(s->Z, frame, em, cdf, mlp);
    *num_emitters = (int)(em.size() / 3);
    if (!em_tris) return CWP_OK;   // count only
    if (cap < *num_emitters || !em_cdf || (mat_light_pdf && mat_cap < (int)mlp.size()))
        return fail(CWP_ERR_ARG, "cwp_setup_emitters: buffers too small");
    memcpy(em_tris, em.data(), em.size() * sizeof(Vec4));
    memcpy(em_cdf, cdf.data(), cdf.size() * sizeof(float));
    if (mat_light_pdf) memcpy(mat_light_pdf, mlp.data(), mlp.size() * sizeof(float));
    return CWP_OK;
}

int cwp_trace_host_inst(cwp_ctx* c, const float* rays, int n, int any_hit, float* hits, int* inst) {
    if (!c || (n > 0 && (!rays || !hits))) return fail(CWP_ERR_ARG, "cwp_trace_host_inst: null argument");
    std::string err;
    if (!c->R.traceHost(rays, n, any_hit != 0, hits, err, inst)) return fail(CWP_ERR_DEVICE, err);
    return CWP_OK;
}

void cwp_destroy(cwp_ctx* c) { delete c; }

int cwp_render(cwp_ctx* c, int spp, int sample_base, int clear, cwp_frame_stats* stats) {
    if (!c) return fail(CWP_ERR_ARG, "cwp_render: null context");
    FrameStats st; std::string err;
    if (!c->R.render(spp, sample_base, clear != 0, st, err)) return fail(err.find("hip") != std::string::npos ? CWP_ERR_DEVICE : CWP_ERR_ARG, err);
    if (stats) {
        stats->rays_closest = st.rays; stats->rays_shadow = st.shadowRays; stats->t_trace_ms = st.traceMs; stats->t_frame_ms = st.frameMs;
        stats->t_bounce_ms = st.bounceMs; stats->t_shadow_ms = st.shadowMs;
        stats->bounce_launches = st.bounceLaunches; stats->shadow_launches = st.shadowLaunches;
        stats->rays_hit = st.hitRays; stats->rays_miss = st.missRays; stats->rays_cont = st.contRays; stats->shadow_unoccluded = st.unoccluded;
        stats->camera_miss = st.cameraMiss; stats->t_shade_ms = st.shadeMs; stats->t_sort_ms = st.sortMs; stats->t_tail_ms = st.tailMs;
    }
    return CWP_OK;
}

int cwp_read_accum(cwp_ctx* c, float* rgba, size_t npix) {
    if (!c || !rgba) return fail(CWP_ERR_ARG, "cwp_read_accum: null argument");
    std::string err;
    if (!c->R.readAccum(rgba, npix, err)) return fail(CWP_ERR_ARG, err);
    return CWP_OK;
}

const float* cwp_accum_device(cwp_ctx* c) { return c ? (const float*)c->R.accumDevice() : nullptr; }

int cwp_display(cwp_ctx* c, void* rgba8, int spp, float exposure_ev, int mode) {
    if (!c || !rgba8) return fail(CWP_ERR_ARG, "cwp_display: null argument");
    std::string err;
    if (!c->R.display((uint32_t*)rgba8, spp, exposure_ev, mode, err)) return fail(CWP_ERR_ARG, err);
    return CWP_OK;
}

int cwp_display_buffer(const float* accum, void* rgba8, int n, int spp, float exposure_ev, int mode) {
    std::string err;
    if (!displaySync((const Vec4*)accum, (uint32_t*)rgba8, n, spp, exposure_ev, mode, err)) return fail(CWP_ERR_ARG, err);
    return CWP_OK;
}

int cwp_srgb_encode8(float x) { return (int)encodeSRGB8(x); }
void cwp_srgb_steps(float steps[256]) { if (steps) srgbSteps(steps); }

int cwp_set_params(cwp_ctx* c, const cwp_params* p) {
    if (!c || !p) return fail(CWP_ERR_ARG, "cwp_set_params: null argument");
    FrameParams F; memcpy(&F, p, sizeof F);
    std::string err;
    if (!c->R.setParams(F, err)) return fail(CWP_ERR_ARG, err);
    return CWP_OK;
}

int cwp_trace(cwp_ctx* c, const void* rays, int n, int any_hit, void* hits) {
    if (!c || (n > 0 && (!rays || !hits))) return fail(CWP_ERR_ARG, "cwp_trace: null argument");
    std::string err;
    if (!c->R.trace(rays, n, any_hit != 0, hits, err)) return fail(CWP_ERR_DEVICE, err);
    return CWP_OK;
}

int cwp_trace_host(cwp_ctx* c, const float* rays, int n, int any_hit, float* hits) {
    if (!c || (n > 0 && (!rays || !hits))) return fail(CWP_ERR_ARG, "cwp_trace_host: null argument");
    std::string err;
    if (!c->R.traceHost(rays, n, any_hit != 0, hits, err)) return fail(CWP_ERR_DEVICE, err);
    return CWP_OK;
}

int cwp_frame_json(const cwp_setup* s, const cwp_ctx* c, const cwp_frame_stats* st, int rep, int frame, char* buf, size_t cap) {
    if (!s || !c || !st || !buf || !cap) return fail(CWP_ERR_ARG, "cwp_frame_json: null argument");
    FrameStats F;
    F.rays = st->rays_closest; F.shadowRays = st->rays_shadow; F.traceMs = st->t_trace_ms; F.frameMs = st->t_frame_ms;
    F.bounceMs = st->t_bounce_ms; F.shadowMs = st->t_shadow_ms; F.hitRays = st->rays_hit; F.missRays = st->rays_miss;
    F.contRays = st->rays_cont; F.unoccluded = st->shadow_unoccluded; F.cameraMiss = st->camera_miss; F.shadeMs = st->t_shade_ms;
    F.sortMs = st->t_sort_ms; F.tailMs = st->t_tail_ms;
    std::string j = frameJson(s->Z, c->R, F, rep, frame);
    if (j.size() + 1 > cap) return fail(CWP_ERR_ARG, "cwp_frame_json: buffer too small");
    memcpy(buf, j.c_str(), j.size() + 1);
    return CWP_OK;
}

void cwp_options_default(cwp_options* o) {
    if (!o) return;
    const RenderOptions d;
    o->tri_test = d.tri; o->batch = d.batch; o->bounces = d.bounces; o->fuse = d.fuse; o->persistent = d.persistent;
    o->shadow_refill = d.shadowRefill; o->stack_lds = 8; o->blocks_per_cu = d.blocksPerCU; o->bounce_waves = d.bounceWaves;
    o->time_kernels = d.timeKernels; o->graph = d.graph; o->sort = d.sort; o->tail_lanes = d.tail;
}

static bool descToUpload(const cwp_scene_desc* d, SceneUpload& S, std::string& err) {
    if (!nodeFloat4s(d->variant)) { err = "variant must be 2, 4 or 8"; return false; }
    if (!d->nodes || !d->num_nodes || !d->corners || !d->num_tris || !d->albedo || !d->num_materials) {
        err = "nodes, corners and albedo are required"; return false;
    }
    if (d->num_nodes > 0x7fffffff || d->num_tris > 0x7fffffff || d->num_materials > 0x7fffffff) { err = "arrays too large"; return false; }
    const size_t per = nodeFloat4s(d->variant);
    auto take = [](const float* p, size_t n4, std::vector<Vec4>& out) {
        out.resize(n4);
        if (n4) memcpy(out.data(), p, n4 * sizeof(Vec4));
    };
    S.variant = d->variant; S.root = d->root;
    take(d->nodes, d->num_nodes * per, S.nodes);
    take(d->corners, d->num_tris * 3, S.verts);
    if (d->tri_form != CWP_TRIS_CORNERS && d->tri_form != CWP_TRIS_EDGES) { err = "tri_form must be corners or edges"; return false; }
    if (d->tri_form == CWP_TRIS_EDGES)
        for (size_t t = 0; t < d->num_tris; t++) {
            const Vec3 v0 = xyz(S.verts[3 * t]);
            for (int k = 1; k < 3; k++) { Vec4& e = S.verts[3 * t + k]; e = v4(v0 + xyz(e), e.w); }
        }
    if (d->woop) take(d->woop, d->num_tris * 3, S.woop); else S.woop.clear();
    if (d->normals) take(d->normals, d->num_tris * 3, S.normals); else S.normals.clear();
    take(d->albedo, d->num_materials, S.albedo);
    if (d->mats) {   // the full material integrator
        S.full = 1;
        S.mats.resize(d->num_materials);
        memcpy(S.mats.data(), d->mats, d->num_materials * sizeof(Material));
        if (d->tri_light_pdf) {
            if (!d->em_tri || !d->em_cdf || !d->num_emitters) { err = "emitters: tri_light_pdf needs em_tri, em_cdf and num_emitters"; return false; }
            S.triLightPdf.assign(d->tri_light_pdf, d->tri_light_pdf + d->num_tris);
            S.emTri.assign(d->em_tri, d->em_tri + d->num_emitters);
            S.emCdf.assign(d->em_cdf, d->em_cdf + d->num_emitters + 1);
        }
        if (d->env_texels) {
            if (d->env_w < 1 || d->env_h < 1) { err = "environment: env_w and env_h must be positive"; return false; }
            std::vector<float> rgb((size_t)d->env_w * d->env_h * 3);
            for (size_t t = 0; t < (size_t)d->env_w * d->env_h; t++)
                for (int k = 0; k < 3; k++) rgb[3 * t + k] = d->env_texels[4 * t + k];
            auto E = std::make_shared<EnvMap>();
            buildEnvMap(rgb, d->env_w, d->env_h, 1.f, *E);
            S.env = E;
        }
        S.envRotDeg = d->env_rot_deg; S.sunOn = d->sun_on; S.lightMode = d->light_mode;
        S.rr = d->rr; S.rrDepth = d->rr_depth; S.gdEnergy = d->gd_energy; S.envUniform = d->env_uniform;
        if (S.lightMode < CWP_LIGHT_MIS || S.lightMode > CWP_LIGHT_BSDF) { err = "light_mode must be MIS, NEE or BSDF"; return false; }
        if (d->maps) {   // textures (checkScene validates ids, sizes and ranges)
            if (!d->uvs || !d->tex_images || !d->texels) { err = "textures: maps need uvs, tex_images and texels"; return false; }
            S.maps.resize(d->num_materials);
            memcpy(S.maps.data(), d->maps, d->num_materials * sizeof(MaterialMaps));
            take(d->uvs, d->num_tris * 2, S.uvs);
            if (d->tangents) take(d->tangents, d->num_tris * 3, S.tans); else S.tans.clear();
            S.texImages.resize(d->num_images);
            if (d->num_images) memcpy(S.texImages.data(), d->tex_images, d->num_images * sizeof(TexImage));
            S.texels.assign(d->texels, d->texels + d->num_texels);
            S.texNearest = d->tex_nearest; S.bumpShadow = d->bump_shadow;
        }
    } else if (d->maps) {
        err = "textures need the full integrator (mats)"; return false;
    }
    return checkScene(S, S.stackNeed, err);
}

int cwp_set_materials(cwp_ctx* c, const float* mats, int num_materials) {
    if (!c || !mats || num_materials < 1) return fail(CWP_ERR_ARG, "cwp_set_materials: bad arguments");
    std::string err;
    if (!c->R.setMaterials((const Material*)mats, num_materials, err)) return fail(CWP_ERR_ARG, err);
    return CWP_OK;
}

int cwp_set_environment(cwp_ctx* c, const float* texels, size_t num_texels, float env_rot_deg, int sun_on) {
    if (!c) return fail(CWP_ERR_ARG, "cwp_set_environment: null context");
    std::string err;
    if (!c->R.setEnvironment((const Vec4*)texels, num_texels, env_rot_deg, sun_on != 0, err)) return fail(CWP_ERR_ARG, err);
    return CWP_OK;
}

int cwp_scene_check(const cwp_scene_desc* d, int* stack_need) {
    if (!d) return fail(CWP_ERR_ARG, "cwp_scene_check: null argument");
    SceneUpload S; std::string err;
    if (!descToUpload(d, S, err)) return fail(CWP_ERR_ARG, err);
    if (stack_need) *stack_need = S.stackNeed;
    return CWP_OK;
}

int cwp_create_scene(const cwp_scene_desc* d, const cwp_params* p, const cwp_options* o, cwp_ctx** out) {
    if (!d || !p || !o || !out) return fail(CWP_ERR_ARG, "cwp_create_scene: null argument");
    *out = nullptr;
    SceneUpload S; std::string err;
    if (!descToUpload(d, S, err)) return fail(CWP_ERR_ARG, err);
    FrameParams F; memcpy(&F, p, sizeof F);
    RenderOptions O;
    O.variant = d->variant; O.tri = o->tri_test; O.batch = o->batch; O.bounces = o->bounces; O.fuse = o->fuse;
    O.persistent = o->persistent; O.shadowRefill = o->shadow_refill; O.stackLds = o->stack_lds;
    O.blocksPerCU = o->blocks_per_cu; O.bounceWaves = o->bounce_waves; O.timeKernels = o->time_kernels; O.graph = o->graph;
    O.sort = o->sort;
    O.tail = o->tail_lanes;
    if (O.sort < 0 || O.sort > 2) return fail(CWP_ERR_ARG, "sort must be 0, 1 or 2");
    if (O.batch < 1) return fail(CWP_ERR_ARG, "batch must be >= 1");
    if (O.fuse != 0 && O.fuse != 1) return fail(CWP_ERR_ARG, "fuse must be 0 or 1");
    if (O.persistent && O.fuse) return fail(CWP_ERR_ARG, "persistent needs fuse 0");
    cwp_ctx* c = new (std::nothrow) cwp_ctx();
    if (!c) return fail(CWP_ERR_ARG, "out of host memory");
    if (!c->R.init(S, F, O, err)) { delete c; return fail(CWP_ERR_DEVICE, err); }
    *out = c;
    return CWP_OK;
}

int cwp_ctx_info(const cwp_ctx* c, int* grid_blocks, double* vram_mb) {
    if (!c) return fail(CWP_ERR_ARG, "cwp_ctx_info: null context");
    if (grid_blocks) *grid_blocks = c->R.gridBlocks;
    if (vram_mb) *vram_mb = c->R.vramMB;
    return CWP_OK;
}

static bool meshFromArrays(const float* vertices, int nv, const int* tris, int nt, Mesh& m, std::string& err) {
    if ((nv > 0 && !vertices) || (nt > 0 && !tris) || nv < 0 || nt < 1) { err = "bad mesh arguments"; return false; }
    m.v.resize(nv);
    for (int i = 0; i < nv; i++) m.v[i] = v3(vertices[3 * i], vertices[3 * i + 1], vertices[3 * i + 2]);
    m.f.resize(nt);
    for (int t = 0; t < nt; t++) {
        Tri q{ tris[3 * t], tris[3 * t + 1], tris[3 * t + 2] };
        if (q.a < 0 || q.b < 0 || q.c < 0 || q.a >= nv || q.b >= nv || q.c >= nv) { err = "vertex index out of range"; return false; }
        m.f[t] = q;
    }
    return true;
}

// layout, triangle packs and flat tree arrays of b->tree
static void finishBVH(const Mesh& m, int variant, cwp_bvh* b) {
    b->variant = variant;
    if (variant == 2 || variant == 4) {
        if (variant == 2) layoutBVH2(b->tree, b->nodes); else layoutBVH4(b->tree, b->nodes);
        b->order = b->tree.prim; b->sah = sahCost(b->tree, 1, 1);
    } else {
        WideBVH W; collapseCWBVH(b->tree, W);
        b->nodes.swap(W.nodes); b->order.swap(W.prim); b->sah = W.sah;
    }
    packWoop(m, b->order, b->woop);
    packVerts(m, b->order, b->corners, b->nrm);
    SceneUpload S;   // the traversal-stack need of the layout, as the renderer will compute it
    S.variant = variant; S.nodes = b->nodes; S.verts = b->corners;
    for (Vec4& c : S.verts) c.w = 0.f;   // corners carry ids in w; only the material slot is checked
    S.albedo.assign(1, v4(1.f, 1.f, 1.f, 0.f));
    std::string err;
    if (!checkScene(S, b->stackNeed, err)) b->stackNeed = -1;
    for (const BinNode& n : b->tree.nodes) {
        b->treeFlat.insert(b->treeFlat.end(), { n.left, n.right, n.first, n.count });
        b->treeBox.insert(b->treeBox.end(), { n.box.lo.x, n.box.lo.y, n.box.lo.z, n.box.hi.x, n.box.hi.y, n.box.hi.z });
    }
}

int cwp_bvh_build(const float* vertices, int nv, const int* tris, int nt, int variant, int bins, int leafMax, cwp_bvh** out) {
    if (!out) return fail(CWP_ERR_ARG, "cwp_bvh_build: null argument");
    if (!nodeFloat4s(variant)) return fail(CWP_ERR_ARG, "cwp_bvh_build: variant must be 2, 4 or 8");
    if (bins < 2 || bins > 256 || leafMax < 1 || leafMax > 32) return fail(CWP_ERR_ARG, "cwp_bvh_build: bins in [2,256], leaf_max in [1,32]");
    *out = nullptr;
    Mesh m; std::string err;
    if (!meshFromArrays(vertices, nv, tris, nt, m, err)) return fail(CWP_ERR_ARG, "cwp_bvh_build: " + err);
    cwp_bvh* b = new (std::nothrow) cwp_bvh();
    if (!b) return fail(CWP_ERR_ARG, "out of host memory");
    buildSAH(m, bins, variant == 8 ? 1 : leafMax, b->tree);
    finishBVH(m, variant, b);
    *out = b;
    return CWP_OK;
}

int cwp_bvh_from_hploc(const float* vertices, int nv, const int* tris, int nt, const float* hnodes, int root, int variant,
                       int leafMax, cwp_bvh** out) {
    if (!out || !hnodes) return fail(CWP_ERR_ARG, "cwp_bvh_from_hploc: null argument");
    if (!nodeFloat4s(variant)) return fail(CWP_ERR_ARG, "cwp_bvh_from_hploc: variant must be 2, 4 or 8");
    if (leafMax < 1 || leafMax > 32) return fail(CWP_ERR_ARG, "cwp_bvh_from_hploc: leaf_max in [1,32]");
    *out = nullptr;
    Mesh m; std::string err;
    if (!meshFromArrays(vertices, nv, tris, nt, m, err)) return fail(CWP_ERR_ARG, "cwp_bvh_from_hploc: " + err);
    const size_t tot = 2 * (size_t)nt - 1;
    std::vector<Vec4> hn(2 * tot);
    memcpy(hn.data(), hnodes, hn.size() * sizeof(Vec4));
    // the conversion trusts the ids: check them (children before parents, leaves name triangles, root in range)
    if (root < 0 || (size_t)root >= tot) return fail(CWP_ERR_ARG, "cwp_bvh_from_hploc: root out of range");
    for (size_t h = 0; h < tot; h++) {
        const int l = asInt(hn[2 * h].w), r = asInt(hn[2 * h + 1].w);
        if (h < (size_t)nt) {
            if (l < 0 || l >= nt || r != -1) return fail(CWP_ERR_ARG, "cwp_bvh_from_hploc: bad leaf node");
        } else if (l < 0 || r < 0 || (size_t)l >= h || (size_t)r >= h) {
            return fail(CWP_ERR_ARG, "cwp_bvh_from_hploc: child ids must be smaller than the parent's");
        }
    }
    cwp_bvh* b = new (std::nothrow) cwp_bvh();
    if (!b) return fail(CWP_ERR_ARG, "out of host memory");
    treeFromHploc(hn, nt, root, variant == 8 ? 1 : leafMax, b->tree);
    if ((int)b->tree.prim.size() != nt) { delete b; return fail(CWP_ERR_ARG, "cwp_bvh_from_hploc: the tree does not hold every triangle once"); }
    finishBVH(m, variant, b);
    *out = b;
    return CWP_OK;
}

int cwp_hploc_build_host_mode(const float* lo, const float* hi, int n, int morton_bits, float* nodes, int* root, double* gpu_ms) {
    if (!lo || !hi || !nodes || n < 1) return fail(CWP_ERR_ARG, "cwp_hploc_build_host: bad arguments");
    if (!validMortonMode(morton_bits)) return fail(CWP_ERR_ARG, "cwp_hploc_build_host: morton_bits must be 0 (adaptive) or 30");
    int r = 0; double ms = 0; std::string err;
    if (!hplocBuildHost((const Vec4*)lo, (const Vec4*)hi, n, (Vec4*)nodes, r, ms, err, morton_bits)) return fail(CWP_ERR_DEVICE, err);
    if (root) *root = r;
    if (gpu_ms) *gpu_ms = ms;
    return CWP_OK;
}
int cwp_hploc_build_host(const float* lo, const float* hi, int n, float* nodes, int* root, double* gpu_ms) {
    return cwp_hploc_build_host_mode(lo, hi, n, CWP_MORTON_ADAPTIVE, nodes, root, gpu_ms);
}

int cwp_hploc_create(int capacity, cwp_hploc** out) {
    if (!out || capacity < 1) return fail(CWP_ERR_ARG, "cwp_hploc_create: bad arguments");
    *out = nullptr;
    cwp_hploc* b = new (std::nothrow) cwp_hploc();
    if (!b) return fail(CWP_ERR_ARG, "out of host memory");
    std::string err;
    if (!b->B.reserve(capacity, err)) { delete b; return fail(CWP_ERR_DEVICE, err); }
    *out = b;
    return CWP_OK;
}

void cwp_hploc_destroy(cwp_hploc* b) { delete b; }

int cwp_hploc_build_mode(cwp_hploc* b, const void* lo, const void* hi, int n, int morton_bits, void* stream, int* root) {
    if (!b || !lo || !hi || n < 1) return fail(CWP_ERR_ARG, "cwp_hploc_build: bad arguments");
    if (!validMortonMode(morton_bits)) return fail(CWP_ERR_ARG, "cwp_hploc_build: morton_bits must be 0 (adaptive) or 30");
    int r = 0; std::string err;
    if (!b->B.build((const Vec4*)lo, (const Vec4*)hi, n, stream, r, err, morton_bits)) return fail(CWP_ERR_DEVICE, err);
    if (root) *root = r;
    return CWP_OK;
}
int cwp_hploc_build(cwp_hploc* b, const void* lo, const void* hi, int n, void* stream, int* root) {
    return cwp_hploc_build_mode(b, lo, hi, n, CWP_MORTON_ADAPTIVE, stream, root);
}

int cwp_hploc_build_static(const void* lo, const void* hi, int n, int morton_bits, void* nodes, void* stream, int* root) {
    if (!lo || !hi || !nodes) return fail(CWP_ERR_ARG, "cwp_hploc_build_static: null argument");
    if (n < 1 || n > CWP_HPLOC_STATIC_MAX)
        return fail(CWP_ERR_ARG, "cwp_hploc_build_static: " + std::to_string(n) + " primitives; the fixed scratch holds 1 to " +
                                     std::to_string(CWP_HPLOC_STATIC_MAX));
    if (!validMortonMode(morton_bits)) return fail(CWP_ERR_ARG, "cwp_hploc_build_static: morton_bits must be 0 (adaptive) or 30");
    int r = 0; std::string err;
    if (!hplocBuildStatic((const Vec4*)lo, (const Vec4*)hi, n, morton_bits, (Vec4*)nodes, stream, r, err)) return fail(CWP_ERR_DEVICE, err);
    if (root) *root = r;
    return CWP_OK;
}

int cwp_hploc_nodes(const cwp_hploc* b, const void** nodes) {
    if (!b || !nodes) return fail(CWP_ERR_ARG, "cwp_hploc_nodes: null argument");
    *nodes = b->B.nodes();
    return CWP_OK;
}

int cwp_hploc_to_bvh2(const void* hnodes, int n, void* out, void* stream) {
    std::string err;
    if (!hplocToBVH2((const Vec4*)hnodes, n, (Vec4*)out, stream, err)) return fail(CWP_ERR_ARG, err);
    return CWP_OK;
}

int cwp_cwbvh_from_hploc(const void* hnodes, int n, void* nodes, int* num_nodes, void* prims, int* depth, void* stream) {
    if (!hnodes || !nodes || !prims || !num_nodes || n < 1) return fail(CWP_ERR_ARG, "cwp_cwbvh_from_hploc: bad arguments");
    // scratch kept between calls of a thread; never destroyed (the device memory goes with the process, and a
    // destructor at exit could run after the HIP runtime has shut down)
    static thread_local CwbvhCollapser* collapser = new CwbvhCollapser();
    std::string err;
    int d = 0;
    if (!collapser->run((const Vec4*)hnodes, n, (Vec4*)nodes, (int*)prims, stream, *num_nodes, d, err)) return fail(CWP_ERR_DEVICE, err);
    if (depth) *depth = d;
    return CWP_OK;
}

int cwp_cwbvh_from_hploc_host(const float* hnodes, int n, float* nodes, int* num_nodes, int* prims, int* depth, double* gpu_ms) {
    if (!hnodes || !nodes || !prims || !num_nodes || n < 1) return fail(CWP_ERR_ARG, "cwp_cwbvh_from_hploc_host: bad arguments");
    std::vector<Vec4> out;
    std::vector<int> pr;
    int d = 0;
    double ms = 0;
    std::string err;
    float sah = 0.f;
    if (!cwbvhFromHplocHost((const Vec4*)hnodes, n, out, pr, d, ms, sah, err)) return fail(CWP_ERR_DEVICE, err);
    memcpy(nodes, out.data(), out.size() * sizeof(Vec4));
    memcpy(prims, pr.data(), pr.size() * sizeof(int));
    *num_nodes = (int)(out.size() / 5);
    if (depth) *depth = d;
    if (gpu_ms) *gpu_ms = ms;
    return CWP_OK;
}

int cwp_device_malloc(size_t bytes, void** ptr) {
    std::string err;
    if (!ptr) return fail(CWP_ERR_ARG, "cwp_device_malloc: null argument");
    if (!deviceMalloc(bytes, ptr, err)) return fail(CWP_ERR_DEVICE, err);
    return CWP_OK;
}

void cwp_device_free(void* ptr) { deviceFree(ptr); }

int cwp_memcpy(void* dst, const void* src, size_t bytes, int kind) {
    std::string err;
    if ((!dst || !src) && bytes) return fail(CWP_ERR_ARG, "cwp_memcpy: null pointer");
    if (kind < CWP_COPY_H2D || kind > CWP_COPY_D2D) return fail(CWP_ERR_ARG, "cwp_memcpy: unknown kind");
    if (!deviceCopy(dst, src, bytes, kind, err)) return fail(CWP_ERR_DEVICE, err);
    return CWP_OK;
}

void cwp_bvh_destroy(cwp_bvh* b) { delete b; }

int cwp_bvh_array(const cwp_bvh* b, int which, const void** data, size_t* count) {
    if (!b || !data || !count) return fail(CWP_ERR_ARG, "cwp_bvh_array: null argument");
    switch (which) {
        case CWP_ARR_NODES: *data = b->nodes.data(); *count = b->nodes.size() / nodeFloat4s(b->variant); break;
        case CWP_ARR_ORDER: *data = b->order.data(); *count = b->order.size(); break;
        case CWP_ARR_WOOP: *data = b->woop.data(); *count = b->woop.size() / 3; break;
        case CWP_ARR_CORNERS: *data = b->corners.data(); *count = b->corners.size() / 3; break;
        case CWP_ARR_TREE: *data = b->treeFlat.data(); *count = b->treeFlat.size() / 4; break;
        case CWP_ARR_TREE_BOX: *data = b->treeBox.data(); *count = b->treeBox.size() / 6; break;
        case CWP_ARR_PRIM: *data = b->tree.prim.data(); *count = b->tree.prim.size(); break;
        default: return fail(CWP_ERR_ARG, "cwp_bvh_array: unknown array");
    }
    return CWP_OK;
}

int cwp_bvh_stats(const cwp_bvh* b, int* depth, int* stackNeed, double* sah) {
    if (!b) return fail(CWP_ERR_ARG, "cwp_bvh_stats: null argument");
    if (depth) *depth = b->tree.maxDepth;
    if (stackNeed) *stackNeed = b->stackNeed;
    if (sah) *sah = b->sah;
    return CWP_OK;
}


int cwp_bvh8_device_nodes(const float* nodes, size_t num_nodes, float* out) {
    if ((!nodes || !out) && num_nodes) return fail(CWP_ERR_ARG, "cwp_bvh8_device_nodes: null argument");
    std::vector<Vec4> n80(num_nodes * 5), dev;
    if (num_nodes) memcpy(n80.data(), nodes, n80.size() * sizeof(Vec4));
    std::string err;
    if (!deviceBVH8(n80, dev, err)) return fail(CWP_ERR_ARG, err);
    if (!dev.empty()) memcpy(out, dev.data(), dev.size() * sizeof(Vec4));
    return CWP_OK;
}

static Material materialOf(const float* m) {
    Material M;
    static_assert(sizeof(Material) == 16 * sizeof(float), "material record");
    memcpy(&M, m, sizeof(Material));
    return M;
}
int cwp_bsdf_eval(const float* mat, const float wo[3], const float wi[3], float eta, int gd_energy, float ux1, float ux2,
                  float f[3], float* pdf) {
    if (!mat || !wo || !wi || !f || !pdf) return fail(CWP_ERR_ARG, "cwp_bsdf_eval: null argument");
    const Vec3 o = v3(wo[0], wo[1], wo[2]);
    const BsdfValue e = evalResponse(surfaceResponse(materialOf(mat), o, eta, gd_energy, ux1, ux2), o, v3(wi[0], wi[1], wi[2]));
    f[0] = e.f.x; f[1] = e.f.y; f[2] = e.f.z; *pdf = e.pdf;
    return CWP_OK;
}
int cwp_bsdf_sample(const float* mat, const float wo[3], float eta, int gd_energy, float ux1, float ux2, float u, float u1,
                    float u2, float wi[3], float weight[3], float* pdf, int* ok) {
    if (!mat || !wo || !wi || !weight || !pdf || !ok) return fail(CWP_ERR_ARG, "cwp_bsdf_sample: null argument");
    const Vec3 o = v3(wo[0], wo[1], wo[2]);
    Vec3 a = v3(0.f, 0.f, 0.f), w = v3(0.f, 0.f, 0.f);
    float p = 0.f;
    *ok = sampleResponse(surfaceResponse(materialOf(mat), o, eta, gd_energy, ux1, ux2), o, u, u1, u2, a, w, p) ? 1 : 0;
    wi[0] = a.x; wi[1] = a.y; wi[2] = a.z; weight[0] = w.x; weight[1] = w.y; weight[2] = w.z; *pdf = p;
    return CWP_OK;
}

static void serverOptionsOut(const ServeOptions& d, cwp_server_options* o) {
    o->variant = d.batch.variant; o->tri_test = d.batch.tri; o->stack_lds = d.batch.stackLds; o->bounces = d.batch.bounces;
    o->max_paths = d.batch.maxPaths; o->max_jobs = d.batch.maxJobs; o->scene_budget_mb = d.sceneBudgetMB;
    o->batched = d.batched ? 1 : 0; o->rays_per_lane = d.batch.raysPerLane; o->use_graph = d.batch.useGraph;
    o->full = d.batch.full; o->rolling = d.rolling ? 1 : 0; o->compact = d.batch.compact; o->tail_lanes = d.batch.tail;
    o->tail_levels = d.batch.tailLevels;
}

void cwp_server_options_default(cwp_server_options* o) {
    if (o) serverOptionsOut(ServeOptions(), o);
}

int cwp_server_get_options(const cwp_server* s, cwp_server_options* o) {
    if (!s || !o) return fail(CWP_ERR_ARG, "cwp_server_get_options: null argument");
    serverOptionsOut(s->S.options(), o);
    return CWP_OK;
}

int cwp_server_create(const cwp_server_options* o, cwp_server** out) {
    if (!o || !out) return fail(CWP_ERR_ARG, "cwp_server_create: null argument");
    *out = nullptr;
    ServeOptions so;
    so.batch.variant = o->variant; so.batch.tri = o->tri_test; so.batch.stackLds = o->stack_lds; so.batch.bounces = o->bounces;
    so.batch.maxPaths = o->max_paths; so.batch.maxJobs = o->max_jobs; so.sceneBudgetMB = o->scene_budget_mb;
    so.batched = o->batched != 0; so.batch.raysPerLane = o->rays_per_lane; so.batch.useGraph = o->use_graph;
    so.batch.full = o->full; so.rolling = o->rolling != 0; so.batch.compact = o->compact; so.batch.tail = o->tail_lanes;
    so.batch.tailLevels = o->tail_levels;
    cwp_server* s = new (std::nothrow) cwp_server();
    if (!s) return fail(CWP_ERR_ARG, "out of host memory");
    std::string err;
    if (!s->S.init(so, err)) { delete s; return fail(err.find("hip") != std::string::npos || err.find("device") != std::string::npos ? CWP_ERR_DEVICE : CWP_ERR_ARG, err); }
    *out = s;
    return CWP_OK;
}

void cwp_server_destroy(cwp_server* s) { delete s; }

int cwp_server_add_scene(cwp_server* s, const cwp_setup* setup, int* scene_id, size_t* scene_bytes) {
    if (!s || !setup || !scene_id) return fail(CWP_ERR_ARG, "cwp_server_add_scene: null argument");
    std::string err;
    const int id = s->S.addScene(setup->Z, err);
    if (id < 0) return fail(CWP_ERR_ARG, err);
    *scene_id = id;
    if (scene_bytes) *scene_bytes = s->S.sceneBytes(id);
    return CWP_OK;
}

int cwp_server_submit(cwp_server* s, int scene_id, int width, int height, int spp, int bounces, uint32_t seed,
                      double arrival_s, int* job_id) {
    if (!s || !job_id) return fail(CWP_ERR_ARG, "cwp_server_submit: null argument");
    ServeJob J;
    J.scene = scene_id; J.w = width; J.h = height; J.spp = spp; J.bounces = bounces; J.seed = seed; J.arrival = arrival_s;
    std::string err;
    const int id = s->S.submit(J, err);
    if (id < 0) return fail(CWP_ERR_ARG, err);
    *job_id = id;
    return CWP_OK;
}

int cwp_server_run(cwp_server* s, int batched) {
    if (!s) return fail(CWP_ERR_ARG, "cwp_server_run: null server");
    s->S.setBatched(batched != 0);
    std::string err;
    if (!s->S.run(err)) return fail(CWP_ERR_DEVICE, err);
    return CWP_OK;
}

int cwp_server_job(const cwp_server* s, int job_id, cwp_job_info* info) {
    if (!s || !info || job_id < 0 || job_id >= s->S.numJobs()) return fail(CWP_ERR_ARG, "cwp_server_job: bad job");
    const ServeJob& J = s->S.job(job_id);
    info->scene = J.scene; info->width = J.w; info->height = J.h; info->spp = J.spp; info->bounces = J.bounces;
    info->samples_per_pass = J.K; info->passes = J.passes;
    info->arrival_s = J.arrival; info->admitted_s = J.admitted; info->finished_s = J.finished;
    info->latency_s = J.finished >= 0 ? J.finished - J.arrival : -1;
    return CWP_OK;
}

int cwp_server_image(const cwp_server* s, int job_id, float* rgba, size_t num_pixels) {
    if (!s || !rgba || job_id < 0 || job_id >= s->S.numJobs()) return fail(CWP_ERR_ARG, "cwp_server_image: bad job");
    const ServeJob& J = s->S.job(job_id);
    if (J.finished < 0) return fail(CWP_ERR_STATE, "cwp_server_image: the job has not finished");
    if (num_pixels < (size_t)J.w * J.h) return fail(CWP_ERR_ARG, "cwp_server_image: buffer too small");
    memcpy(rgba, J.image.data(), J.image.size() * sizeof(float));
    return CWP_OK;
}

int cwp_server_stats(const cwp_server* s, cwp_serve_stats* st) {
    if (!s || !st) return fail(CWP_ERR_ARG, "cwp_server_stats: null argument");
    const ServeStats& x = s->S.stats();
    st->jobs = x.jobs; st->passes = x.passes; st->scenes_loaded = x.scenesLoaded; st->scenes_evicted = x.scenesEvicted;
    st->wall_s = x.wallS; st->rays = x.rays; st->rays_per_s = x.raysPerS; st->latency_p50_s = x.latP50;
    st->latency_p99_s = x.latP99; st->latency_mean_s = x.latMean; st->max_paths_in_flight = x.maxPathsInFlight;
    st->steps = x.steps;
    return CWP_OK;
}

int cwp_server_start(cwp_server* s) {
    if (!s) return fail(CWP_ERR_ARG, "cwp_server_start: null server");
    std::string err;
    if (!s->S.start(err)) return fail(CWP_ERR_STATE, err);
    return CWP_OK;
}

int cwp_job_submit(cwp_server* s, int scene_id, int width, int height, int spp, int bounces, uint32_t seed, int* job_id) {
    if (!s || !job_id) return fail(CWP_ERR_ARG, "cwp_job_submit: null argument");
    ServeJob J;
    J.scene = scene_id; J.w = width; J.h = height; J.spp = spp; J.bounces = bounces; J.seed = seed;
    std::string err;
    const int id = s->S.submitLive(J, err);
    if (id < 0) return fail(s->S.running() ? CWP_ERR_ARG : CWP_ERR_STATE, err);
    *job_id = id;
    return CWP_OK;
}

int cwp_job_poll(const cwp_server* s, int job_id, int* state) {
    if (!s || !state) return fail(CWP_ERR_ARG, "cwp_job_poll: null argument");
    *state = (int)s->S.poll(job_id);
    return CWP_OK;
}

int cwp_job_result(const cwp_server* s, int job_id, float* rgba, size_t num_pixels) {
    if (!s || !rgba) return fail(CWP_ERR_ARG, "cwp_job_result: null argument");
    std::vector<float> img;
    std::string err;
    if (!s->S.result(job_id, img, err)) return fail(CWP_ERR_STATE, err);
    if (num_pixels * 4 < img.size()) return fail(CWP_ERR_ARG, "cwp_job_result: buffer too small");
    memcpy(rgba, img.data(), img.size() * sizeof(float));
    return CWP_OK;
}

int cwp_server_stop(cwp_server* s) {
    if (!s) return fail(CWP_ERR_ARG, "cwp_server_stop: null server");
    std::string err;
    if (!s->S.stop(err)) return fail(err.find("not started") != std::string::npos ? CWP_ERR_STATE : CWP_ERR_DEVICE, err);
    return CWP_OK;
}

int cwp_server_clear_jobs(cwp_server* s) {
    if (!s) return fail(CWP_ERR_ARG, "cwp_server_clear_jobs: null server");
    s->S.clearJobs();
    return CWP_OK;
}

int cwp_server_preload(cwp_server* s) {
    if (!s) return fail(CWP_ERR_ARG, "cwp_server_preload: null server");
    std::string err;
    if (!s->S.preload(err)) return fail(CWP_ERR_DEVICE, err);
    return CWP_OK;
}

int cwp_server_reset_jobs(cwp_server* s) {
    if (!s) return fail(CWP_ERR_ARG, "cwp_server_reset_jobs: null server");
    s->S.resetJobs();
    return CWP_OK;
}

int cwp_server_autotune(cwp_server* s, const int* variants, int nv, const int* stacks, int ns, const int* rpl, int nr,
                        const int* compacts, int nc, cwp_tune_trial* trials, int max_trials, int* num_trials,
                        cwp_server_options* chosen) {
    if (!s || !variants || !stacks || !rpl || nv < 1 || ns < 1 || nr < 1 || (compacts && nc < 1) ||
        (max_trials > 0 && !trials))
        return fail(CWP_ERR_ARG, "cwp_server_autotune: bad arguments");
    std::vector<TuneTrial> T;
    std::string err;
    const std::vector<int> cm = compacts ? std::vector<int>(compacts, compacts + nc) : std::vector<int>{0};
    const bool ok = s->S.autotune(std::vector<int>(variants, variants + nv), std::vector<int>(stacks, stacks + ns),
                                  std::vector<int>(rpl, rpl + nr), cm, T, err);
    if (num_trials) *num_trials = (int)T.size();
    for (int k = 0; k < (int)T.size() && k < max_trials; k++) {
        trials[k].variant = T[k].variant; trials[k].stack_lds = T[k].stackLds; trials[k].rays_per_lane = T[k].raysPerLane;
        trials[k].compact = T[k].compact;
        trials[k].rays_per_s = T[k].raysPerS; trials[k].wall_s = T[k].wallS;
    }
    if (!ok) return fail(CWP_ERR_DEVICE, err);
    if (chosen) serverOptionsOut(s->S.options(), chosen);
    return CWP_OK;
}

int cwp_image_read(const char* path, int* w, int* h, uint32_t* rgba, size_t capacity) {
    if (!path || !w || !h) return fail(CWP_ERR_ARG, "cwp_image_read: null argument");
    TextureImage img;
    std::string err;
    if (!readImage(path, img, err)) return fail(CWP_ERR_ARG, err);
    *w = img.w; *h = img.h;
    if (!rgba) return CWP_OK;
    if (capacity < img.rgba.size()) return fail(CWP_ERR_ARG, "cwp_image_read: buffer too small");
    memcpy(rgba, img.rgba.data(), img.rgba.size() * 4);
    return CWP_OK;
}

int cwp_inflate(const uint8_t* in, size_t n, uint8_t* out, size_t capacity, size_t* out_len) {
    if (!in || !out_len || (capacity && !out)) return fail(CWP_ERR_ARG, "cwp_inflate: null argument");
    std::vector<uint8_t> o;
    std::string err;
    if (!inflateZlib(in, n, o, capacity, err)) return fail(CWP_ERR_ARG, err);
    if (!o.empty()) memcpy(out, o.data(), o.size());
    *out_len = o.size();
    return CWP_OK;
}

int cwp_texture_sample(const int image[4], const uint32_t* texels, size_t num_texels, float u, float v, int nearest,
                       float rgba[4]) {
    if (!image || !texels || !rgba) return fail(CWP_ERR_ARG, "cwp_texture_sample: null argument");
    TexImage T;
    T.first = (uint32_t)image[0]; T.w = image[1]; T.h = image[2]; T.srgb = image[3];
    if (T.w < 1 || T.h < 1 || (uint64_t)T.first + (uint64_t)T.w * T.h > num_texels || (T.srgb != 0 && T.srgb != 1))
        return fail(CWP_ERR_ARG, "cwp_texture_sample: image outside the texel array");
    const Vec4 c = sampleTexture(T, texels, srgbDecodeTable(), u, v, nearest ? 1 : 0);
    rgba[0] = c.x; rgba[1] = c.y; rgba[2] = c.z; rgba[3] = c.w;
    return CWP_OK;
}

}  // extern "C"
