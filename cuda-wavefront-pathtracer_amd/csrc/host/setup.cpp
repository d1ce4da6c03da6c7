// host/setup.cpp - option parsing and scene setup (see host/setup.h).
#include "host/setup.h"
#include "device/collapse.h"
#include "device/hploc.h"
#include "host/bvhcache.h"
#include "host/scene_check.h"
#include "host/instances.h"
#include "host/io.h"
#include "host/texture.h"
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace cwp {

bool parseArgs(int argc, const char* const* argv, Args& A, std::string& err) {
    for (int i = 1; i < argc; i++) {
        const std::string k = argv[i];
        auto need = [&](int n) {
            if (i + n >= argc) { err = "option " + k + " needs " + std::to_string(n) + " value(s)"; return false; }
            return true;
        };
        auto str = [&](std::string& dst) { if (!need(1)) return false; dst = argv[++i]; return true; };
        auto num = [&](float& dst) { if (!need(1)) return false; dst = (float)atof(argv[++i]); return true; };
        auto inum = [&](int& dst) { if (!need(1)) return false; dst = atoi(argv[++i]); return true; };
        auto vec = [&](Vec3& dst) {
            if (!need(3)) return false;
            dst = v3((float)atof(argv[i + 1]), (float)atof(argv[i + 2]), (float)atof(argv[i + 3])); i += 3; return true;
        };
        bool ok = true;
        if (k == "--mesh") ok = str(A.mesh);
        else if (k == "--camera") ok = str(A.camera);
        else if (k == "--out") ok = str(A.out);
        else if (k == "--variant") ok = str(A.variant);
        else if (k == "--c2w_key") ok = str(A.c2wKey);
        else if (k == "--c2w") ok = str(A.c2w);
        else if (k == "--tri") ok = str(A.tri);
        else if (k == "--lighting") ok = str(A.lighting);
        else if (k == "--integrator") ok = str(A.integrator);
        else if (k == "--builder") ok = str(A.builder);
        else if (k == "--collapse") ok = str(A.collapse);
        else if (k == "--gen") ok = str(A.gen);
        else if (k == "--gen_boxes") ok = inum(A.genBoxes);
        else if (k == "--gen_grid") ok = inum(A.genGrid);
        else if (k == "--gen_seed") { int s = 0; ok = inum(s); A.genSeed = (uint64_t)(unsigned)s; }
        else if (k == "--synthetic") ok = inum(A.synthetic);
        else if (k == "--frame") ok = inum(A.frame);
        else if (k == "--width") ok = inum(A.w);
        else if (k == "--height") ok = inum(A.h);
        else if (k == "--spp") ok = inum(A.spp);
        else if (k == "--bounces") ok = inum(A.bounces);
        else if (k == "--batch") ok = inum(A.batch);
        else if (k == "--leaf") ok = inum(A.leaf);
        else if (k == "--bins") ok = inum(A.bins);
        else if (k == "--fuse") ok = inum(A.fuse);
        else if (k == "--persistent") ok = inum(A.persistent);
        else if (k == "--scene") ok = str(A.scene);
        else if (k == "--flatten") ok = inum(A.flatten);
        else if (k == "--frames") ok = inum(A.frames);
        else if (k == "--frame_start") ok = inum(A.frameStart);
        else if (k == "--gen_instances") ok = inum(A.genInstances);
        else if (k == "--gen_small") ok = inum(A.genSmall);
        else if (k == "--shadow_refill") ok = inum(A.shadowRefill);
        else if (k == "--smooth") ok = inum(A.smooth);
        else if (k == "--warm") ok = inum(A.warm);
        else if (k == "--reps") ok = inum(A.reps);
        else if (k == "--stack") ok = inum(A.stack);
        else if (k == "--bounce_waves") ok = inum(A.bounceWaves);
        else if (k == "--blocks_per_cu") ok = inum(A.blocksPerCU);
        else if (k == "--time_kernels") ok = inum(A.timeKernels);
        else if (k == "--filter") { std::string f; ok = str(f); A.filterBH = f == "bh"; if (ok && f != "bh" && f != "box") { err = "--filter must be bh or box"; return false; } }
        else if (k == "--fx") ok = num(A.fx);
        else if (k == "--fy") ok = num(A.fy);
        else if (k == "--hfov") ok = num(A.hfov);
        else if (k == "--aperture") ok = num(A.aperture);
        else if (k == "--focus_dist") ok = num(A.focusDist);
        else if (k == "--eps") ok = num(A.epsScale);
        else if (k == "--near") ok = num(A.nearT);
        else if (k == "--sun_angle") ok = num(A.light.sunAngleDeg);
        else if (k == "--seed") { int s = 0; ok = inum(s); A.seed = (unsigned)s; }
        else if (k == "--sun_to") ok = vec(A.light.sunTo);
        else if (k == "--sun_rgb") ok = vec(A.light.sunE);
        else if (k == "--sky") ok = vec(A.light.sky);
        else if (k == "--albedo") ok = vec(A.albedo);
        else if (k == "--materials") ok = str(A.materials);
        else if (k == "--env") ok = str(A.env);
        else if (k == "--env_strength") ok = num(A.envStrength);
        else if (k == "--env_rot") ok = num(A.envRot);
        else if (k == "--light_sampling") ok = str(A.lightSampling);
        else if (k == "--env_sampling") ok = str(A.envSampling);
        else if (k == "--rr") ok = inum(A.rr);
        else if (k == "--rr_depth") ok = inum(A.rrDepth);
        else if (k == "--gd_energy") ok = inum(A.gdEnergy);
        else if (k == "--gen_env") ok = str(A.genEnv);
        else if (k == "--gen_materials") ok = inum(A.genMaterials);
        else if (k == "--graph") ok = inum(A.graph);
        else if (k == "--bvh_cache") ok = str(A.bvhCache);
        else if (k == "--gpu_collapse") ok = inum(A.gpuCollapse);
        else if (k == "--morton_bits") ok = inum(A.mortonBits);
        else if (k == "--tlas") ok = str(A.tlas);
        else if (k == "--textures") ok = str(A.textures);
        else if (k == "--tex_filter") ok = str(A.texFilter);
        else if (k == "--bump_shadow") ok = inum(A.bumpShadow);
        else if (k == "--gen_textures") ok = inum(A.genTextures);
        else if (k == "--env_out") ok = str(A.envOut);
        else if (k == "--sort") ok = inum(A.sort);
        else if (k == "--tail") ok = inum(A.tail);
        else if (k == "--tail_levels") ok = inum(A.tailLevels);
        else if (k == "--devloop") {
            int v = 0; ok = inum(v);
            if (ok && v) { err = k + " selects an experimental variant of the reference that this renderer does not provide"; return false; }
        } else { err = "unknown arg " + k; return false; }
        if (!ok) return false;
    }
    if (A.variant != "bvh2" && A.variant != "bvh4" && A.variant != "bvh8") { err = "--variant must be bvh2, bvh4 or bvh8"; return false; }
    if (A.tri != "mt" && A.tri != "woop" && A.tri != "watertight") { err = "--tri must be mt, woop or watertight"; return false; }
    if (A.integrator != "auto" && A.integrator != "legacy" && A.integrator != "full") { err = "--integrator must be auto, legacy or full"; return false; }
    if (A.lightSampling != "mis" && A.lightSampling != "nee" && A.lightSampling != "bsdf") { err = "--light_sampling must be mis, nee or bsdf"; return false; }
    if (A.envSampling != "importance" && A.envSampling != "uniform") { err = "--env_sampling must be importance or uniform"; return false; }
    if (A.rrDepth < 0) { err = "--rr_depth must be >= 0"; return false; }
    if (A.builder != "sah" && A.builder != "hploc") { err = "--builder must be sah or hploc"; return false; }
    if (A.mortonBits != 0 && A.mortonBits != 30) { err = "--morton_bits must be 0 (adaptive) or 30"; return false; }
    if (A.tlas != "bvh2" && A.tlas != "bvh8") { err = "--tlas must be bvh2 or bvh8"; return false; }
    if (A.texFilter != "linear" && A.texFilter != "nearest") { err = "--tex_filter must be linear or nearest"; return false; }
    {
        std::string t = A.textures + ",", cur;
        for (char c : t) {
            if (c != ',') { cur += c; continue; }
            if (!cur.empty() && cur != "base" && cur != "emission" && cur != "normal" && cur != "metallic" && cur != "roughness" &&
                cur != "none") {
                err = "--textures: unknown kind " + cur + " (base, emission, normal, metallic, roughness, or none)"; return false;
            }
            cur.clear();
        }
    }
    if (A.genTextures < 0 || A.genTextures > 8192) { err = "--gen_textures must be 0 (off) or an image size up to 8192"; return false; }
    if (A.collapse != "sah") { err = "--collapse must be sah"; return false; }
    if (A.fuse < -1 || A.fuse > 1) { err = "--fuse must be 0 (split trace and shade kernels) or 1 (fused)"; return false; }
    if (A.persistent < 0 || A.persistent > 2) {
        err = "--persistent must be 0, 1 (split kernels with refilled lanes) or 2 (fused kernel with refilled lanes)"; return false;
    }
    if (A.shadowRefill != 0 && A.shadowRefill != 1) { err = "--shadow_refill must be 0 or 1"; return false; }
    if (A.sort < 0 || A.sort > 2) { err = "--sort must be 0, 1 or 2"; return false; }
    if (A.tail < -1 || A.tail == 1 || A.tail > 64) { err = "--tail must be -1 (auto), 0 (off) or 2..64 lanes"; return false; }
    if (A.tailLevels != 1 && A.tailLevels != 2) { err = "--tail_levels must be 1 or 2"; return false; }
    if (A.fuse < 0) A.fuse = 1;
    if (A.persistent == 1) A.fuse = 0;   // persistent traversal 1 runs in the split trace kernel
    if (A.persistent == 2) A.fuse = 1;   // 2: the fused trace+shade kernel with refilled lanes
    if (A.bins < 2 || A.bins > 256) { err = "--bins must be in [2, 256]"; return false; }
    if (A.leaf < 1 || A.leaf > 32) { err = "--leaf must be in [1, 32]"; return false; }
    if (A.w < 1 || A.h < 1) { err = "--width/--height must be positive"; return false; }
    if (A.batch < 1) A.batch = 1;
    if (A.bounces < 0) { err = "--bounces must be >= 0"; return false; }
    if (A.stack != 8 && A.stack != 16 && A.stack != 32) { err = "--stack must be 8, 16 or 32"; return false; }
    if (A.bounceWaves != 1 && A.bounceWaves != 6 && A.bounceWaves != 8) { err = "--bounce_waves must be 1, 6 or 8"; return false; }
    if (!A.gen.empty() && A.gen != "boxes" && A.gen != "terrain" && A.gen != "instances") {
        err = "--gen must be boxes, terrain or instances"; return false;
    }
    if (A.frames < 0 || A.frameStart < 0) { err = "--frames and --frame_start must be >= 0"; return false; }
    return true;
}

static double msSince(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// binary tree of one mesh: binned SAH on the host, or H-PLOC on the GPU followed by the host SAH leaf collapse
// (reference: buildTree, src/app/scene_setup.cpp:76-83); treeMs is the build time (device time for H-PLOC), postMs
// the host conversion
static bool buildTree(const Mesh& M, const Args& A, int leafMax, BinaryBVH& tree, double& treeMs, double& postMs,
                      std::string& err, WideBVH* wide = nullptr) {
    const auto t0 = std::chrono::steady_clock::now();
    if (A.builder != "hploc") {
        buildSAH(M, A.bins, leafMax, tree);
        treeMs = msSince(t0); postMs = 0;
        return true;
    }
    const int n = (int)M.f.size();
    std::vector<Vec4> lo(n), hi(n), hn((size_t)2 * (2 * (size_t)n - 1));
    for (int t = 0; t < n; t++) {
        AABB b; b.grow(M.v[M.f[t].a]); b.grow(M.v[M.f[t].b]); b.grow(M.v[M.f[t].c]);
        lo[t] = v4(b.lo, 0.f); hi[t] = v4(b.hi, 0.f);
    }
    int root = 0; double gpuMs = 0;
    if (!hplocBuildHost(lo.data(), hi.data(), n, hn.data(), root, gpuMs, err, A.mortonBits)) return false;
    const auto t1 = std::chrono::steady_clock::now();
    treeFromHploc(hn, n, root, leafMax, tree);
    if (wide) {   // the compressed BVH8 collapse on the device (bit-identical to collapseCWBVH of the tree above)
        int depth = 0; double collapseMs = 0; float sah = 0.f;
        if (!cwbvhFromHplocHost(hn.data(), n, wide->nodes, wide->prim, depth, collapseMs, sah, err)) return false;
        wide->sah = sah; wide->depth = depth;
    }
    postMs = msSince(t1); treeMs = gpuMs;
    return true;
}

bool frameEmitters(const SceneSetup& Z, int f, std::vector<Vec4>& emTris, std::vector<float>& emCdf,
                   std::vector<float>& matLightPdf) {
    emTris.clear(); emCdf.clear(); matLightPdf.assign(Z.S.mats.size(), 0.f);
    if (!Z.inst || !Z.full || Z.S.numEmInst == 0) return false;
    // leaf position of every mesh triangle: S.order maps leaf -> mesh triangle offset + local triangle
    std::vector<int> posOf(Z.S.order.size(), 0);
    for (size_t kk = 0; kk < Z.S.order.size(); kk++) posOf[Z.S.order[kk]] = (int)kk;
    std::vector<double> pw;
    double tot = 0;
    for (const InstanceDesc& I : Z.SD.instances) {
        const Xform X = instanceXform(I, f);
        const Mesh& m = Z.meshes[I.mesh];
        const int off = Z.S.meshTris[2 * I.mesh];
        for (size_t t = 0; t < m.f.size(); t++) {
            const int mi = m.mat(t);
            if (!isEmissive(Z.S.mats[mi])) continue;
            const Vec3 w0 = xfPoint(X, m.v[m.f[t].a]), w1 = xfPoint(X, m.v[m.f[t].b]), w2 = xfPoint(X, m.v[m.f[t].c]);
            const Vec3 e1 = w1 - w0, e2 = w2 - w0;
            const double area = 0.5 * (double)length(cross(e1, e2));
            emTris.push_back(v4(w0, asFloat(posOf[off + (int)t]))); emTris.push_back(v4(e1, asFloat(mi))); emTris.push_back(v4(e2, 0.f));
            double p = (double)luminance(Z.S.mats[mi].emission) * area;
            p = p > 0 ? p : 0;
            pw.push_back(p); tot += p;
        }
    }
    emCdf.assign(pw.size() + 1, 0.f);
    double c = 0;
    for (size_t e = 0; e < pw.size(); e++) { c += pw[e]; emCdf[e + 1] = (float)(c / tot); }
    emCdf.back() = 1.f;
    for (size_t j = 0; j < Z.S.mats.size(); j++)
        if (isEmissive(Z.S.mats[j])) matLightPdf[j] = (float)(luminance(Z.S.mats[j].emission) / tot);
    return true;
}

bool frameInstances(const SceneSetup& Z, int f, std::vector<InstanceRecord>& out) {
    out.clear();
    if (!Z.inst) return false;
    for (const InstanceDesc& I : Z.SD.instances) {
        const Xform X = instanceXform(I, f);
        Xform Xi; float ds;
        invertXform(X, Xi, ds);
        InstanceRecord R;
        for (int r = 0; r < 3; r++) {
            R.inv[r] = v4(Xi.m[4 * r], Xi.m[4 * r + 1], Xi.m[4 * r + 2], Xi.m[4 * r + 3]);
            R.fwd[r] = v4(X.m[4 * r], X.m[4 * r + 1], X.m[4 * r + 2], X.m[4 * r + 3]);
        }
        R.root = Z.S.meshRoot[I.mesh]; R.mesh = I.mesh; R.detSign = ds; R.pad = 0;
        out.push_back(R);
    }
    return true;
}

// Emissive triangles as lights, sampled in proportion to power = luminance(emission) * area: the CDF runs over the
// triangles in mesh order (independent of the BVH), and triLightPdf[leaf position] = selection probability / area
// (reference: src/app/scene_setup.cpp:225-247).
static void emitterTables(const Mesh& M, const std::vector<Material>& mats, SceneUpload& S) {
    const std::vector<int>& order = S.order;
    std::vector<std::pair<int, int>> em;   // (mesh triangle, leaf position)
    for (size_t kk = 0; kk < order.size(); kk++)
        if (isEmissive(mats[M.mat(order[kk])])) em.push_back({ order[kk], (int)kk });
    std::sort(em.begin(), em.end());
    std::vector<double> pw, area;
    double tot = 0;
    S.emTri.clear(); S.emCdf.clear(); S.triLightPdf.clear();
    for (const auto& e : em) {
        const Tri& f = M.f[e.first];
        const Vec3 e1 = M.v[f.b] - M.v[f.a], e2 = M.v[f.c] - M.v[f.a];
        const double a = 0.5 * (double)length(cross(e1, e2));
        const double p = (double)luminance(mats[M.mat(e.first)].emission) * a;
        if (!(p > 0)) continue;
        S.emTri.push_back(e.second); pw.push_back(p); area.push_back(a); tot += p;
    }
    if (S.emTri.empty()) return;
    S.triLightPdf.assign(order.size(), 0.f);
    S.emCdf.assign(S.emTri.size() + 1, 0.f);
    double c = 0;
    for (size_t e = 0; e < S.emTri.size(); e++) {
        S.triLightPdf[S.emTri[e]] = (float)(pw[e] / tot / area[e]);
        c += pw[e];
        S.emCdf[e + 1] = (float)(c / tot);
    }
    S.emCdf.back() = 1.f;
}

// BVH of the setup's mesh (or one BLAS per mesh of a two-level scene) in the layout Z.A.variant names, the packed
// triangles and the GPU-ready scene Z.S; uses Z.M / Z.meshes, Z.inst, Z.P.hasN and the builder options of Z.A
static bool buildSceneBVH(SceneSetup& Z, std::string& err) {
    const Args& A = Z.A;
    const Mesh& M = Z.M;
    auto t0 = std::chrono::steady_clock::now();
    const int V = Z.variant = A.variant == "bvh2" ? 2 : (A.variant == "bvh4" ? 4 : 8);
    const int leafMax = V == 8 ? 1 : A.leaf;   // the CWBVH collapse forms its own leaves from a 1-primitive-leaf tree
    Z.reportLeaf = V == 8 ? 3 : leafMax;
    SceneUpload& S = Z.S;
    S = SceneUpload();
    S.variant = V; S.root = 0; S.inst = Z.inst ? 1 : 0;
    // one BLAS per mesh for two-level scenes (material ids already offset into the combined table), else the mesh
    std::vector<const Mesh*> parts;
    if (Z.inst) for (const Mesh& m : Z.meshes) parts.push_back(&m);
    else parts.push_back(&M);
    Z.treeMs = 0; Z.postMs = 0; Z.maxDepth = 0; Z.sah = 0;
    // a cached build of the same meshes with the same layout and builder options replaces the build
    Z.bvhCacheHit = -1;
    uint64_t cacheKey = 0;
    std::string cachePath;
    if (!A.bvhCache.empty()) {
        // the Morton mode changes H-PLOC trees: it is part of the builder's identity
        const std::string builderId = A.builder == "hploc" && A.mortonBits ? A.builder + std::to_string(A.mortonBits) : A.builder;
        cacheKey = bvhCacheKey(parts, V, builderId, leafMax, A.bins, Z.P.hasN);
        cachePath = bvhCachePath(A.bvhCache, cacheKey);
        BvhCacheEntry E;
        Z.bvhCacheHit = loadBvhCache(cachePath, cacheKey, S, E) && S.variant == V && S.inst == (Z.inst ? 1 : 0) ? 1 : 0;
        if (Z.bvhCacheHit) { Z.tree = std::move(E.tree); Z.maxDepth = E.maxDepth; Z.sah = E.sah; }
        else { S = SceneUpload(); S.variant = V; S.root = 0; S.inst = Z.inst ? 1 : 0; }
    }
    for (size_t k = 0; Z.bvhCacheHit != 1 && k < parts.size(); k++) {
        const Mesh& mk = *parts[k];
        BinaryBVH tree;
        double treeMs = 0, postMs = 0, sah = 0;
        const bool gpuWide = V == 8 && A.builder == "hploc" && A.gpuCollapse;
        WideBVH W;
        if (!buildTree(mk, A, leafMax, tree, treeMs, postMs, err, gpuWide ? &W : nullptr)) return false;
        std::vector<Vec4> nodes;
        std::vector<int> order;
        if (V == 4 || V == 2) {
            if (tree.maxDepth > 63) { err = "BVH depth exceeds the traversal stack"; return false; }
            if (V == 4) layoutBVH4(tree, nodes); else layoutBVH2(tree, nodes);
            order = tree.prim; sah = sahCost(tree, 1, 1);
        } else {
            if (!gpuWide) collapseCWBVH(tree, W);
            nodes.swap(W.nodes); order.swap(W.prim); sah = W.sah;
        }
        const int triBase = (int)S.verts.size() / 3;
        const int root = V == 2 ? appendBVH2(S.nodes, nodes, triBase) : (V == 4 ? appendBVH4(S.nodes, nodes, triBase)
                                                                               : appendCWBVH(S.nodes, nodes, triBase));
        std::vector<Vec4> verts, woop, normals;
        packVerts(mk, order, verts, normals);
        packWoop(mk, order, woop);
        if (normals.empty() && Z.P.hasN) normals.assign(verts.size(), v4(0.f, 0.f, 0.f, 0.f));   // geometric normal
        S.verts.insert(S.verts.end(), verts.begin(), verts.end());
        S.woop.insert(S.woop.end(), woop.begin(), woop.end());
        S.normals.insert(S.normals.end(), normals.begin(), normals.end());
        for (int o : order) S.order.push_back(o + triBase);
        if (Z.inst) {
            const AABB mb = meshBox(mk);
            S.meshRoot.push_back(root); S.meshBox.push_back(v4(mb.lo, 0.f)); S.meshBox.push_back(v4(mb.hi, 0.f));
            S.meshTris.push_back(triBase); S.meshTris.push_back((int)mk.f.size());
        }
        if (k == 0) Z.tree = tree;
        Z.treeMs += treeMs; Z.postMs += postMs; Z.maxDepth = std::max(Z.maxDepth, tree.maxDepth);
        Z.sah += sah / (double)parts.size();
    }
    if (Z.bvhCacheHit == 0) {
        BvhCacheEntry E;
        E.tree = Z.tree; E.maxDepth = Z.maxDepth; E.sah = Z.sah;
        std::string werr;
        if (!saveBvhCache(cachePath, cacheKey, S, E, werr)) fprintf(stderr, "warning: %s\n", werr.c_str());
    }
    S.maps.clear(); S.uvs.clear(); S.tans.clear(); S.texImages.clear(); S.texels.clear();
    if (Z.full && Z.textured) {   // corner uvs and tangents in leaf order (S.order numbers the parts' triangles in turn)
        bool anyNormalMap = false;
        for (const MaterialMaps& x : M.maps) anyNormalMap = anyNormalMap || x.normal >= 0;
        std::vector<float> uvAll;
        std::vector<Vec4> tanAll;
        for (const Mesh* p : parts) {
            const std::vector<float> uv = p->hasUV() ? p->uv : std::vector<float>(6 * p->f.size(), 0.f);
            if (anyNormalMap) { std::vector<Vec4> t; cornerTangents(*p, uv, t); tanAll.insert(tanAll.end(), t.begin(), t.end()); }
            uvAll.insert(uvAll.end(), uv.begin(), uv.end());
        }
        const size_t n = S.order.size();
        S.uvs.resize(2 * n);
        if (anyNormalMap) S.tans.resize(3 * n);
        for (size_t kk = 0; kk < n; kk++) {
            const size_t t = (size_t)S.order[kk];
            const float* q = &uvAll[6 * t];
            S.uvs[2 * kk] = v4(q[0], q[1], q[2], q[3]); S.uvs[2 * kk + 1] = v4(q[4], q[5], 0.f, 0.f);
            if (anyNormalMap) for (int c = 0; c < 3; c++) S.tans[3 * kk + c] = tanAll[3 * t + c];
        }
        S.maps = M.maps;
        for (const TextureImage& T : M.tex) {
            if (S.texels.size() + T.rgba.size() > 0xFFFFFFFFull) { err = "textures: more than 2^32 texels"; return false; }
            TexImage r;
            r.first = (uint32_t)S.texels.size(); r.w = T.w; r.h = T.h; r.srgb = T.srgb ? 1 : 0;
            S.texImages.push_back(r);
            S.texels.insert(S.texels.end(), T.rgba.begin(), T.rgba.end());
        }
        S.texNearest = A.texFilter == "nearest" ? 1 : 0;
        S.bumpShadow = A.bumpShadow ? 1 : 0;
    }
    S.albedo.clear();
    for (const Vec3& k : M.kd) S.albedo.push_back(v4(k, 0.f));
    if (Z.full) {
        S.full = 1; S.mats = Z.mats; S.env = Z.env; S.envRotDeg = A.light.envRotDeg;
        S.sunOn = luminance(A.light.sunE) > 0.f ? 1 : 0;
        S.lightMode = A.lightSampling == "nee" ? kLightNEE : (A.lightSampling == "bsdf" ? kLightBSDF : kLightMIS);
        S.rr = A.rr; S.rrDepth = A.rrDepth; S.gdEnergy = A.gdEnergy; S.envUniform = A.envSampling == "uniform" ? 1 : 0;
        if (Z.inst) {   // emitters are per frame in world space; here only their count
            S.numEmInst = 0;
            for (const InstanceDesc& I : Z.SD.instances) {
                const Mesh& m = Z.meshes[I.mesh];
                for (size_t t = 0; t < m.f.size(); t++) if (isEmissive(S.mats[m.mat(t)])) S.numEmInst++;
            }
        } else {
            emitterTables(M, S.mats, S);
        }
    }
    Z.bvhS = msSince(t0) / 1000.0;
    Z.frames = A.frames > 0 ? A.frames : (Z.inst ? Z.SD.frames : 1);
    if (!Z.inst && Z.frames > 1) { err = "--frames > 1 needs a two-level scene (--scene or --gen instances, no --flatten)"; return false; }
    Z.nNodes = S.nodes.size() / nodeFloat4s(V);
    if (!checkScene(S, S.stackNeed, err)) return false;
    Z.nodeBytes = S.nodes.size() * 16;
    return true;
}

bool relayoutScene(const SceneSetup& Z, int variant, SceneUpload& S, std::string& err) {
    if (Z.inst) { err = "relayout: two-level scenes keep their layout"; return false; }
    if (!nodeFloat4s(variant)) { err = "relayout: variant must be 2, 4 or 8"; return false; }
    SceneSetup T;
    T.A = Z.A;
    T.A.variant = variant == 2 ? "bvh2" : (variant == 4 ? "bvh4" : "bvh8");
    T.M = Z.M;
    T.P.hasN = Z.P.hasN;
    T.full = Z.full; T.mats = Z.mats; T.env = Z.env; T.textured = Z.textured;
    if (!buildSceneBVH(T, err)) return false;
    S = std::move(T.S);
    return true;
}

bool setupScene(SceneSetup& Z, std::string& err) {
    Args& A = Z.A;
    Mesh& M = Z.M;
    float* c2w = Z.c2w;
    identityC2W(c2w);
    bool haveC2w = false;
    // ---- scene ----
    Z.inst = false; Z.meshes.clear(); Z.SD = SceneDesc();
    if (A.gen == "instances" || !A.scene.empty()) {
        std::vector<Mesh> meshes;
        if (A.gen == "instances") {
            makeInstanceScene(meshes, Z.SD, A.genInstances >= 0 ? A.genInstances : 1000, A.genSeed, A.genSmall != 0);
            instanceSceneCamera(c2w); haveC2w = true; Z.sceneName = "gen:instances";
            if (A.genMaterials >= 0) {   // one OpenPBR material per mesh: glossy-diffuse, conductor, dielectric, emitter
                SplitMix r((uint64_t)A.genMaterials * 0x9E3779B97F4A7C15ull + 11);
                for (size_t k = 0; k < meshes.size(); k++) {
                    Material x = defaultMaterial();
                    x.baseColor = meshes[k].kd.empty() ? v3(0.8f, 0.8f, 0.8f) : meshes[k].kd[0];
                    x.roughness = r.uniform(0.05f, 0.6f);
                    const int kind = k + 1 == meshes.size() ? 0 : (int)(k % 4);
                    if (kind == 1) { x.metallic = 1.f; x.specularColor = v3(0.95f, 0.86f, 0.55f); }
                    else if (kind == 2) { x.transmission = 1.f; x.baseColor = v3(0.95f, 0.95f, 1.f); }
                    else if (kind == 3) { x = lambertMaterial(v3(0.5f, 0.5f, 0.5f)); x.emission = x.baseColor * 6.f; }
                    meshes[k].mats.assign(1, x);
                }
            }
        } else {
            if (!loadSceneFile(A.scene, Z.SD, err)) return false;
            for (const MeshDesc& d : Z.SD.meshes) {
                Mesh m;
                if (!loadMesh(d.file, m)) { err = "failed to load mesh " + d.file; return false; }
                if (!d.materials.empty() && !loadMaterialsFile(d.materials, m, err)) return false;
                if (d.yup == 1) for (Vec3& p : m.v) p = v3(p.x, -p.z, p.y);   // y-up to z-up
                if (d.yup == 1) for (Vec3& q : m.n) q = v3(q.x, -q.z, q.y);
                meshes.push_back(m);
            }
            Z.sceneName = A.scene;
        }
        // one material table: each mesh's materials appended, its material ids offset
        M = Mesh();
        std::vector<int> matOff;
        bool anyMats = false;
        for (const Mesh& m : meshes) anyMats = anyMats || !m.mats.empty();
        bool anyMaps = false;
        for (const Mesh& m : meshes) anyMaps = anyMaps || !m.maps.empty();
        for (Mesh& m : meshes) {
            if (m.kd.empty()) m.kd.push_back(v3(0.8f, 0.8f, 0.8f));
            int maxMat = 0; for (int t : m.fm) maxMat = std::max(maxMat, t);
            while ((int)m.kd.size() <= maxMat) m.kd.push_back(v3(0.8f, 0.8f, 0.8f));
            while (m.kd.size() < m.mats.size()) m.kd.push_back(m.mats[m.kd.size()].baseColor);
            matOff.push_back((int)M.kd.size());
            const int texOff = (int)M.tex.size();
            for (size_t j = 0; j < m.kd.size(); j++) {
                M.kd.push_back(m.kd[j]); M.matName.push_back(j < m.matName.size() ? m.matName[j] : "");
                if (anyMats) M.mats.push_back(j < m.mats.size() ? m.mats[j] : lambertMaterial(m.kd[j]));
                if (anyMaps) {   // the mesh's bindings, their image ids moved past the images of the meshes before
                    MaterialMaps x = j < m.maps.size() ? m.maps[j] : noMaps();
                    for (int* id : { &x.baseColor, &x.emission, &x.normal, &x.metallic, &x.roughness }) if (*id >= 0) *id += texOff;
                    M.maps.push_back(x);
                }
            }
            M.tex.insert(M.tex.end(), m.tex.begin(), m.tex.end());
            m.tex.clear(); m.maps.clear();
        }
        if (A.flatten) {   // every instance transformed into one mesh at the first frame (reference --flatten 1)
            for (const InstanceDesc& I : Z.SD.instances) appendTransformed(M, meshes[I.mesh], matOff[I.mesh], instanceXform(I, A.frameStart));
        } else {
            Z.inst = true;
            for (size_t k = 0; k < meshes.size(); k++) {
                Mesh& m = meshes[k];
                if (m.fm.empty()) m.fm.assign(m.f.size(), 0);
                for (int& t : m.fm) t += matOff[k];
                m.kd.clear();
            }
            Z.meshes = meshes;
            // the combined mesh only carries the material table and the world box of the first frame (ray offset eps)
            for (const InstanceDesc& I : Z.SD.instances) {
                const Xform X = instanceXform(I, A.frameStart);
                const AABB b = meshBox(meshes[I.mesh]);
                for (int c = 0; c < 8; c++)
                    M.v.push_back(xfPoint(X, v3(c & 1 ? b.hi.x : b.lo.x, c & 2 ? b.hi.y : b.lo.y, c & 4 ? b.hi.z : b.lo.z)));
            }
            bool anyN = false;
            for (const Mesh& m : meshes) anyN = anyN || m.hasN();
            if (anyN) M.n.assign(M.v.size(), v3(0.f, 0.f, 1.f));
        }
    } else if (A.gen == "boxes") {
        Z.genBoxCount = A.genBoxes >= 0 ? A.genBoxes : 1024;
        makeBoxCloud(M, Z.genBoxCount, A.genSeed, 0.7f);
        Z.genBoxFirst = 0;
        boxCloudCamera(c2w); haveC2w = true; Z.sceneName = "gen:boxes";
    } else if (A.gen == "terrain") {
        Z.genBoxCount = A.genBoxes >= 0 ? A.genBoxes : 200;
        makeTerrain(M, A.genGrid, 6, 0.2f, Z.genBoxCount, A.genSeed, 0.8f);
        Z.genBoxFirst = 2 * (A.genGrid - 1) * (A.genGrid - 1);
        terrainCamera(c2w); haveC2w = true; Z.sceneName = "gen:terrain";
    } else if (A.synthetic) {
        makeSphereGrid(M, A.synthetic, 200, 100, 200);
        sphereGridCamera(c2w); haveC2w = true; Z.sceneName = "synthetic";
    } else {
        if (A.mesh.empty()) { err = "no scene: give --mesh, --gen or --synthetic"; return false; }
        if (!loadMesh(A.mesh, M)) { err = "failed to load mesh " + A.mesh; return false; }
        Z.sceneName = A.mesh;
    }
    if (A.synthetic && A.hfov == 0 && A.fx == 0) A.hfov = 50.f;
    if (!A.smooth) { M.n.clear(); for (Mesh& m : Z.meshes) m.n.clear(); }
    if (!A.lighting.empty() && !loadLightingFile(A.lighting, A.light, &M, err)) return false;
    if (A.albedo.x >= 0) for (auto& kd : M.kd) kd = A.albedo;
    if (M.kd.empty()) M.kd.push_back(v3(0.8f, 0.8f, 0.8f));
    if (!A.materials.empty() && !loadMaterialsFile(A.materials, M, err)) return false;
    if (A.genMaterials >= 0 && A.gen != "instances") {
        if (Z.genBoxFirst < 0) { err = "--gen_materials needs a --gen boxes, terrain or instances scene"; return false; }
        assignGeneratedMaterials(M, Z.genBoxFirst, Z.genBoxCount, (uint64_t)A.genMaterials);
    }
    { int maxMat = 0; for (int t : M.fm) maxMat = t > maxMat ? t : maxMat;
      while ((int)M.kd.size() <= maxMat) M.kd.push_back(v3(0.8f, 0.8f, 0.8f)); }
    // the material table: materials-file / generated entries, else Lambertian kd; --albedo overrides base colours
    Z.mats.clear();
    for (size_t j = 0; j < std::max(M.kd.size(), M.mats.size()); j++) {
        Material x = j < M.mats.size() ? M.mats[j] : lambertMaterial(M.kd[j]);
        if (A.albedo.x >= 0) x.baseColor = A.albedo;
        Z.mats.push_back(x);
    }
    while (M.kd.size() < Z.mats.size()) M.kd.push_back(Z.mats[M.kd.size()].baseColor);
    // textures: generated maps, the kinds asked for (--textures), --albedo replacing textured base colours too
    if (A.genTextures > 0) {
        if (A.gen != "terrain" && A.gen != "boxes") { err = "--gen_textures needs a --gen terrain or boxes scene"; return false; }
        addGeneratedTextures(M, A.genTextures, A.genSeed);
    }
    if (!M.maps.empty()) {
        M.maps.resize(Z.mats.size(), noMaps());
        const std::string kinds = "," + A.textures + ",";
        auto on = [&](const char* k) { return kinds.find(std::string(",") + k + ",") != std::string::npos; };
        for (MaterialMaps& x : M.maps) {
            if (!on("base") || A.albedo.x >= 0) x.baseColor = -1;
            if (!on("emission")) x.emission = -1;
            if (!on("normal")) x.normal = -1;
            if (!on("metallic")) x.metallic = -1;
            if (!on("roughness")) x.roughness = -1;
        }
    }
    Z.textured = false;
    for (const MaterialMaps& x : M.maps) Z.textured = Z.textured || hasMaps(x);
    if (Z.textured) {
        bool anyUV = Z.inst ? false : M.hasUV();
        for (const Mesh& m : Z.meshes) anyUV = anyUV || m.hasUV();
        if (!anyUV) { fprintf(stderr, "warning: texture maps without texture coordinates are ignored\n"); Z.textured = false; }
    }
    if (!Z.textured) { M.maps.clear(); M.tex.clear(); }
    if (!A.env.empty()) A.light.envMap = A.env;
    if (A.envStrength >= 0.f) A.light.envStrength = A.envStrength;
    if (A.envRot > -1e29f) A.light.envRotDeg = A.envRot;
    bool needFull = A.rr != 0 || !A.light.envMap.empty() || !A.genEnv.empty() || Z.textured;
    for (const Material& x : Z.mats) needFull = needFull || !isLambert(x) || isEmissive(x);
    Z.full = A.integrator == "full" || (A.integrator == "auto" && needFull);
    Z.env.reset();
    if (Z.full && (!A.light.envMap.empty() || !A.genEnv.empty())) {
        std::vector<float> rgb;
        int ew = 0, eh = 0;
        if (!A.genEnv.empty()) {
            double g[4] = { 2048, 1024, 2, 1 };
            int ng = 0;
            std::string s = A.genEnv; for (auto& ch : s) if (ch == ',') ch = ' ';
            const char* q = s.c_str(); char* e;
            while (ng < 4) { const double x = strtod(q, &e); if (e == q) break; g[ng++] = x; q = e; }
            if (ng < 3 || g[0] < 2 || g[1] < 1 || g[0] > 16384 || g[1] > 16384 || g[2] < 0) { err = "--gen_env needs W,H,SIGMA[,SEED]"; return false; }
            ew = (int)g[0]; eh = (int)g[1];
            makeLogNormalEnv(ew, eh, 0.f, (float)g[2], (uint64_t)g[3], rgb);
        } else if (!loadHDR(A.light.envMap, ew, eh, rgb, err)) return false;
        if (!A.envOut.empty() && !writeHDR(A.envOut, ew, eh, rgb, err)) return false;   // radiance before the strength
        std::shared_ptr<EnvMap> E = std::make_shared<EnvMap>();
        buildEnvMap(rgb, ew, eh, A.light.envStrength, *E);
        Z.env = E;
    }
    if (!A.envOut.empty() && !Z.env) { err = "--env_out needs an environment map (--env or --gen_env) of the full integrator"; return false; }
    AABB sb = meshBox(M);
    Z.extent = maxc(sb.extent());
    // ---- camera ----
    if (!A.camera.empty()) {
        if (!loadCameraFile(A.camera, A.c2wKey, A.frame, A.w, A.h, A.fx, A.fy, A.hfov, c2w, A.aperture, A.focusDist, err)) return false;
        haveC2w = true;
    }
    if (!A.c2w.empty()) {
        std::string s = A.c2w; for (auto& ch : s) if (ch == ',') ch = ' ';
        const char* q = s.c_str(); char* e;
        for (int j = 0; j < 16; j++) { c2w[j] = (float)strtod(q, &e); if (e == q) { err = "--c2w needs 16 numbers"; return false; } q = e; }
        haveC2w = true;
    }
    if (!haveC2w) { err = "no camera"; return false; }
    if (A.fx == 0) { if (A.hfov == 0) A.hfov = 50.f; A.fx = focalFromHfov(A.w, A.hfov); }
    if (A.fy == 0) A.fy = A.fx;
    FrameParams& H = Z.P;
    memset(&H, 0, sizeof H);
    H.cam = makeCamera(c2w, A.fx, A.fy, A.w, A.h);
    if (A.aperture > 0.f) {
        if (!(A.focusDist > 0.f)) { err = "thin lens: --aperture needs --focus_dist"; return false; }
        H.cam.lensR = A.aperture; H.cam.focusD = A.focusDist;
    }
    H.sunTo = normalize(A.light.sunTo); H.sunE = A.light.sunE; H.sky = A.light.sky;
    H.sunCos = cosf(0.5f * A.light.sunAngleDeg * 3.14159265f / 180.f);
    H.maxB = A.bounces; H.seed = A.seed; H.eps = A.epsScale * Z.extent; H.nearT = A.nearT;
    H.hasN = M.hasN() ? 1 : 0; H.filterBH = A.filterBH;
    blackmanHarrisTable(H.ftab, kFilterTable, 1.5);
    if (!buildSceneBVH(Z, err)) return false;
    // ---- renderer options ----
    RenderOptions& O = Z.O;
    O.variant = Z.variant; O.tri = A.tri == "mt" ? kTriMT : (A.tri == "watertight" ? kTriWatertight : kTriWoop);
    O.batch = A.batch; O.bounces = A.bounces; O.stackLds = A.stack; O.blocksPerCU = A.blocksPerCU; O.timeKernels = A.timeKernels;
    O.bounceWaves = A.bounceWaves; O.fuse = A.fuse; O.persistent = A.persistent;
    O.shadowRefill = A.shadowRefill; O.sort = A.sort; O.tail = A.tail; O.tailLevels = A.tailLevels; O.mortonBits = A.mortonBits;
    O.tlasWidth = A.tlas == "bvh2" ? 2 : 8;
    O.graph = A.graph < 0 ? (Z.full ? 1 : 0) : A.graph;   // the reference's default: graphs for the full integrator
    return true;
}

std::string frameJson(const SceneSetup& Z, const Renderer& R, const FrameStats& st, int rep, int frame) {
    const Args& A = Z.A;
    const double tot = st.rays + st.shadowRays;
    const int N = A.w * A.h;
    char buf[4096];
    snprintf(buf, sizeof buf,
             "{\"rep\":%d,\"variant\":\"bvh%d\",\"tri\":\"%s\",\"persistent\":%d,\"devloop\":0,\"sort\":%d,\"fuse\":%d,\"leaf\":%d,"
             "\"batch\":%d,\"minb\":0,\"scene\":\"%s\",\"tris\":%zu,\"bvh_nodes\":%zu,\"bvh_bytes\":%zu,\"bvh_depth\":%d,"
             "\"bvh_build_s\":%.2f,\"w\":%d,\"h\":%d,\"spp\":%d,\"bounces\":%d,\"rays_closest\":%.0f,\"rays_shadow\":%.0f,"
             "\"rays_total\":%.0f,\"t_trace_ms\":%.2f,\"t_frame_ms\":%.2f,\"mrays_s_trace\":%.1f,\"mrays_s_e2e\":%.1f,"
             "\"avg_path_len\":%.3f,\"vram_MB\":%.0f,\"pt_blocks\":%d,\"gpu\":\"%s\",\"filter\":\"%s\",\"smooth\":%d,"
             "\"seed\":%u,\"integrator\":\"%s\",\"lens_r\":%.4f,\"focus_d\":%.4f,\"env_sampling\":\"%s\",\"textures\":0,"
             "\"builder\":\"%s\",\"collapse\":\"%s\",\"bvh_tree_ms\":%.2f,\"bvh_post_ms\":%.2f,\"bvh_sah\":%.4f,\"graph\":%d,"
             "\"instances\":%d,\"flatten\":%d,\"frame\":%d,\"tlas_ms\":%.3f,\"tlas\":\"%s\",\"bins\":%d,\"stack_lds\":%d,"
             "\"t_bounce_ms\":%.3f,\"t_shade_ms\":%.3f,\"t_shadow_ms\":%.3f,\"rays_hit\":%.0f,\"rays_miss\":%.0f,\"rays_cont\":%.0f,\"shadow_unoccluded\":%.0f,"
             "\"bvh_cache\":\"%s\",\"t_sort_ms\":%.3f,\"tail\":%d}",
             rep, Z.variant, A.tri.c_str(), A.persistent, A.sort, A.fuse, Z.reportLeaf, A.batch, Z.sceneName.c_str(), Z.M.f.size(), Z.nNodes,
             Z.nodeBytes, Z.maxDepth, Z.bvhS, A.w, A.h, A.spp, A.bounces, st.rays, st.shadowRays, tot, st.traceMs, st.frameMs,
             st.traceMs > 0 ? tot / (st.traceMs * 1e3) : 0.0, st.frameMs > 0 ? tot / (st.frameMs * 1e3) : 0.0,
             st.rays / ((double)N * A.spp), R.vramMB, R.gridBlocks, R.gpuName.c_str(), A.filterBH ? "blackman-harris1.5" : "box",
             Z.P.hasN, A.seed, Z.full ? "full" : "legacy", Z.P.cam.lensR, Z.P.cam.focusD, A.envSampling.c_str(), A.builder.c_str(), Z.variant == 8 ? A.collapse.c_str() : "-",
             Z.treeMs, Z.postMs, Z.sah, R.options().graph, Z.inst || !Z.A.scene.empty() || A.gen == "instances" ? (int)Z.SD.instances.size() : 0,
             A.flatten, frame, Z.inst ? R.tlasMs : 0.0, Z.inst ? (R.tlasWide ? "bvh8" : "bvh2") : "-", A.bins, R.options().stackLds, st.bounceMs, st.shadeMs, st.shadowMs, st.hitRays, st.missRays, st.contRays,
             st.unoccluded, Z.bvhCacheHit < 0 ? "off" : (Z.bvhCacheHit ? "hit" : "miss"), st.sortMs, R.options().tail);
    return buf;
}

}  // namespace cwp
