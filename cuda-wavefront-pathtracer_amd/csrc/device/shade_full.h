// device/shade_full.h - shading of the full material integrator (included by device/kernels.h inside namespace
// cwp::dev, after the path-state and shading-result types).
//
// Restates the reference's shadeFull (src/render/wavefront.cu:223-356) for single-level scenes: OpenPBR-subset
// materials (core/bsdf.h); emissive triangles as area lights sampled by power; an equirectangular environment map
// sampled by luminance * sin(theta) (or uniformly); the sun. One light sample per vertex (light type chosen with
// probabilities pSun / pTri / pEnv), combined with BSDF sampling by the power heuristic; optional Russian roulette.
// Random dimensions of vertex depth d (key depth d + 1): 0, 1 BSDF direction, 2, 3 light position, 4 lobe choice,
// 5 light choice, 6 roulette, 7 (and 7 under seed ^ 0x68E31DA4) the glossy-diffuse albedo estimate.
//
// Differences in mechanism (not in the estimator): radiance goes into the path's own L[i] (no atomics), the shadow
// ray is returned to the caller's wave-level append, and the pdf of the last BSDF sample travels in D.w of the path
// state (the reference keeps a separate pAux array). Textures (core/texture.h) are looked up by this build's own
// filtering code instead of a hardware texture unit, so the host oracle evaluates the same arithmetic.
#pragma once

// y with cdf[y] <= u < cdf[y + 1] (cdf has n + 1 ascending entries, cdf[0] = 0)
__device__ __forceinline__ int cdfFind(const float* cdf, int n, float u) {
    int lo = 0, hi = n;
    while (hi - lo > 1) { const int mid = (lo + hi) >> 1; if (cdf[mid] <= u) lo = mid; else hi = mid; }
    return lo;
}

// equirectangular mapping: u = 0.5 - (atan2(y, x) - rot) / 2 pi (wrapped), v = acos(z) / pi
__device__ __forceinline__ void envUV(const DevLights& E, Vec3 d, float& u, float& v) {
    u = 0.5f - (atan2f(d.y, d.x) - E.envRot) * (1.f / kTwoPi);
    u -= floorf(u);
    v = acosf(fminf(fmaxf(d.z, -1.f), 1.f)) * kInvPi;
}
__device__ __forceinline__ Vec3 envDir(const DevLights& E, float u, float v) {
    const float ph = (0.5f - u) * kTwoPi + E.envRot, th = v * kPi;
    const float st = sinf(th);
    return v3(st * cosf(ph), st * sinf(ph), cosf(th));
}
__device__ __forceinline__ void envTexel(const DevLights& E, Vec3 d, int& x, int& y, float& v) {
    float u;
    envUV(E, d, u, v);
    x = min((int)(u * (float)E.envW), E.envW - 1);
    y = min((int)(v * (float)E.envH), E.envH - 1);
}
__device__ __forceinline__ Vec3 envRadiance(const DevLights& E, Vec3 d) {
    int x, y; float v;
    envTexel(E, d, x, y, v);
    return xyz(E.envTex[y * E.envW + x]);
}
// solid-angle density of the luminance sampling at texel (x, y) whose row coordinate is v
__device__ __forceinline__ float envPdfTexel(const DevLights& E, int x, int y, float v) {
    const float* cr = E.envCond + (size_t)y * (E.envW + 1);
    const float puv = (E.envMarg[y + 1] - E.envMarg[y]) * (cr[x + 1] - cr[x]) * (float)E.envW * (float)E.envH;
    const float st = sinf(v * kPi);
    return st > 0.f ? puv / (2.f * kPi * kPi * st) : 0.f;
}
__device__ __forceinline__ float envPdf(const DevLights& E, Vec3 d) {
    if (E.envUniform) return 0.25f * kInvPi;
    int x, y; float v;
    envTexel(E, d, x, y, v);
    return envPdfTexel(E, x, y, v);
}
__device__ __forceinline__ Material loadMaterial(const Material* mats, int id) {
    const Vec4* p = (const Vec4*)(mats + id);
    const Vec4 a = p[0], b = p[1], c = p[2], e = p[3];
    Material m;
    m.baseColor = xyz(a); m.metallic = a.w; m.specularColor = xyz(b); m.roughness = b.w;
    m.emission = xyz(c); m.ior = c.w; m.specularWeight = e.x; m.transmission = e.y; m.pad0 = 0.f; m.pad1 = 0.f;
    return m;
}
__device__ __forceinline__ MaterialMaps loadMaps(const MaterialMaps* maps, int id) {
    const Vec4* p = (const Vec4*)(maps + id);
    const Vec4 a = p[0], b = p[1];
    MaterialMaps m;
    m.baseColor = asInt(a.x); m.emission = asInt(a.y); m.normal = asInt(a.z); m.metallic = asInt(a.w);
    m.roughness = asInt(b.x); m.metallicChannel = asInt(b.y); m.roughnessChannel = asInt(b.z); m.normalScale = b.w;
    return m;
}
__device__ __forceinline__ Vec4 texLookup(const DevLights& E, int id, float u, float v) {
    const Vec4 r = *(const Vec4*)(E.texImages + id);
    TexImage T;
    T.first = asUint(r.x); T.w = asInt(r.y); T.h = asInt(r.z); T.srgb = asInt(r.w);
    return sampleTexture(T, E.texels, E.srgbTable, u, v, E.texNearest);
}
// texture coordinates of a point of leaf triangle t with barycentric weights b1, b2 of corners 1 and 2
__device__ __forceinline__ void triUV(const DevLights& E, int t, float b1, float b2, float& u, float& v) {
    const Vec4 q0 = E.uvs[2 * t], q1 = E.uvs[2 * t + 1];
    const float b0 = 1.f - b1 - b2;
    u = q0.x * b0 + q0.z * b1 + q1.x * b2;
    v = q0.y * b0 + q0.w * b1 + q1.y * b2;
}
__device__ __forceinline__ void addRadiance(const DevPaths& P, int i, int depth, Vec3 c) {
    if (depth == 0) { P.L[i] = v4(c, 0.f); return; }
    Vec4 l = P.L[i]; l.x += c.x; l.y += c.y; l.z += c.z; P.L[i] = l;
}

// Shades the closest hit (or miss) of path i; lastPdf = pdf of the BSDF sample that made the ray (0: camera ray or a
// delta lobe, whose emission hits count fully). INST: the hit lies in instance h.inst: the triangle is in object
// space, normals go to world space through inv^T and keep the geometric winding by sign(det); emitters are the
// frame's world-space list and an emitter's area pdf is its material's.
template <bool INST>
__device__ __forceinline__ ShadeOut shadeFull(const DevScene& S, const DevPaths& P, const FrameParams& F, int depth,
                                              int i, int pix, uint32_t smp, Vec3 o, Vec3 d, Vec3 T, float lastPdf,
                                              const Hit& h) {
    const DevLights& E = S.Lt;
    ShadeOut out;
    out.cont = false; out.shadow = false; out.miss = false;
    out.po = v3(0.f, 0.f, 0.f); out.nd = out.po; out.Tn = out.po; out.sdir = out.po; out.scon = out.po; out.so = out.po;
    out.stmax = 1e30f; out.pdf = 0.f;
    if (h.tri < 0) {   // environment (MIS-weighted against its light sampling) or the constant sky
        if (h.tri == kHandedOff) return out;
        out.miss = true;
        Vec3 Lm;
        if (E.hasEnv) {
            Lm = envRadiance(E, d);
            if (lastPdf > 0.f) {
                if (E.lightMode == kLightNEE) Lm = v3(0.f, 0.f, 0.f);
                else if (E.lightMode == kLightMIS) Lm = Lm * misPower(lastPdf, E.pEnv * envPdf(E, d));
            }
        } else {
            Lm = F.sky;
        }
        addRadiance(P, i, depth, T * Lm);
        return out;
    }
    const Vec3 p = o + d * h.t;
    const Vec4 a = S.verts[3 * h.tri], b = S.verts[3 * h.tri + 1], c = S.verts[3 * h.tri + 2];
    const Vec3 e1 = xyz(b) - xyz(a), e2 = xyz(c) - xyz(a);
    Vec4 xr[3];
    float dsg = 1.f;
    if constexpr (INST) {
        const InstanceRecord* I = S.inst + h.inst;
        xr[0] = I->inv[0]; xr[1] = I->inv[1]; xr[2] = I->inv[2]; dsg = I->detSign;
    }
    const Vec3 ng0 = INST ? normalize(xfRowsN(xr, cross(e1, e2))) * dsg : normalize(cross(e1, e2));
    Vec3 ns = ng0;
    if (F.hasN) {
        const Vec3 n0 = xyz(S.nrm[3 * h.tri]), n1 = xyz(S.nrm[3 * h.tri + 1]), n2 = xyz(S.nrm[3 * h.tri + 2]);
        Vec3 nn = n0 * (1.f - h.u - h.v) + n1 * h.u + n2 * h.v;
        if constexpr (INST) nn = xfRowsN(xr, nn);
        const float l = length(nn);
        if (l > 0.f) ns = nn * (1.f / l);
        if (dot(ns, ng0) < 0.f) ns = -ns;
    }
    const int mid = asInt(b.w);
    Material M = loadMaterial(E.mats, mid);
    Vec3 nsU = ns;   // the shading normal before normal mapping (terminator shadowing)
    bool mapped = false;
    if (E.maps && mid < E.nMaps) {   // maps multiply the material's factors; a normal map bends ns
        const MaterialMaps X = loadMaps(E.maps, mid);
        if (hasMaps(X)) {
            float tu, tv;
            triUV(E, h.tri, h.u, h.v, tu, tv);
            if (X.baseColor >= 0) M.baseColor = M.baseColor * xyz(texLookup(E, X.baseColor, tu, tv));
            if (X.emission >= 0) M.emission = M.emission * xyz(texLookup(E, X.emission, tu, tv));
            if (X.metallic >= 0) M.metallic *= channelOf(texLookup(E, X.metallic, tu, tv), X.metallicChannel);
            if (X.roughness >= 0) M.roughness *= channelOf(texLookup(E, X.roughness, tu, tv), X.roughnessChannel);
            if (X.normal >= 0 && E.tans) {
                const float w0 = 1.f - h.u - h.v;
                const Vec4 t0 = E.tans[3 * h.tri], t1 = E.tans[3 * h.tri + 1], t2 = E.tans[3 * h.tri + 2];
                Vec3 tg = xyz(t0) * w0 + xyz(t1) * h.u + xyz(t2) * h.v;
                float sg = (t0.w * w0 + t1.w * h.u + t2.w * h.v) < 0.f ? -1.f : 1.f;
                if constexpr (INST) {   // tangents are directions: the forward rows
                    const InstanceRecord* I = S.inst + h.inst;
                    const Vec4 xf[3] = { I->fwd[0], I->fwd[1], I->fwd[2] };
                    tg = xfRowsV(xf, tg);
                    sg *= dsg;
                }
                tg = tg - ns * dot(ns, tg);
                const float tl = length(tg);
                if (tl > 0.f) {
                    tg = tg * (1.f / tl);
                    const Vec3 bt = cross(ns, tg) * sg;
                    const Vec4 c = texLookup(E, X.normal, tu, tv);
                    const Vec3 nm = tg * ((2.f * c.x - 1.f) * X.normalScale) + bt * ((2.f * c.y - 1.f) * X.normalScale) +
                                    ns * (2.f * c.z - 1.f);
                    const float nl = length(nm);
                    if (nl > 0.f) { ns = nm * (1.f / nl); mapped = true; }
                }
            }
        }
    }
    const bool front = dot(ng0, d) < 0.f;
    const Vec3 ng = front ? ng0 : -ng0;
    if (!front) { ns = -ns; nsU = -nsU; }
    // emission (two-sided), MIS-weighted against light sampling of this triangle
    Vec3 Lv = v3(0.f, 0.f, 0.f);
    if (isEmissive(M)) {
        float w = 1.f;
        if (lastPdf > 0.f) {
            if (E.lightMode == kLightNEE) w = 0.f;
            else if (E.lightMode == kLightMIS) {
                const float cl = fabsf(dot(ng0, d));
                const float pa = INST ? E.matLightPdf[mid] : E.triLightPdf[h.tri];
                const float pl = cl > 0.f ? E.pTri * pa * h.t * h.t / cl : 0.f;
                w = misPower(lastPdf, pl);
            }
        }
        Lv = T * M.emission * w;
    }
    if (depth == 0 || Lv.x != 0.f || Lv.y != 0.f || Lv.z != 0.f) addRadiance(P, i, depth, Lv);
    if (depth > F.maxB) return out;
    const uint32_t dd = (uint32_t)depth + 1u;
    if (E.rr && depth >= E.rrDepth) {
        const float qs = fminf(fmaxf(maxc(T), 0.05f), 1.f);
        if (rnd((uint32_t)pix, smp, dd, 6, F.seed) >= qs) return out;
        T = T * (1.f / qs);
    }
    const Vec3 wo = -d;
    if (dot(ns, wo) <= 1e-6f) ns = ng;
    Vec3 tx, ty;
    basis(ns, tx, ty);
    const Vec3 woL = v3(dot(wo, tx), dot(wo, ty), dot(wo, ns));
    Vec3 nrL = v3(0.f, 0.f, 1.f);   // the unperturbed normal in the local frame (terminator shadowing off: +z)
    if (E.bumpShadow && mapped) {
        nrL = v3(dot(nsU, tx), dot(nsU, ty), dot(nsU, ns));
        if (!(nrL.z > 0.f)) nrL = v3(0.f, 0.f, 1.f);
    }
    const float eta = front ? M.ior : 1.f / M.ior;
    const Vec3 poUp = p + ng * F.eps, poDn = p - ng * F.eps;
    const float ux1 = rnd((uint32_t)pix, smp, dd, 7, F.seed), ux2 = rnd((uint32_t)pix, smp, dd, 7, F.seed ^ 0x68E31DA4u);
    // ---- next-event estimation: one light sample
    // the material's response is rebuilt where it is used (light sample, BSDF sample), so it holds no registers across
    // the light-selection code
    if (hasDensity(surfaceResponse(M, woL, eta, E.gdEnergy, ux1, ux2, nrL))) {
        const float us = rnd((uint32_t)pix, smp, dd, 5, F.seed);
        const float u2 = rnd((uint32_t)pix, smp, dd, 2, F.seed), u3 = rnd((uint32_t)pix, smp, dd, 3, F.seed);
        Vec3 l = v3(0.f, 0.f, 1.f), Le = v3(0.f, 0.f, 0.f);
        float pl = 0.f, tmax = 1e30f;
        bool isSun = false, ok = false;
        int orgSide = 0;   // +1 / -1: the shadow-ray origin side fixed by a triangle sample
        if (us < E.pSun) {
            l = uniformCone(F.sunTo, F.sunCos, u2, u3); Le = F.sunE; pl = E.pSun; isSun = true; ok = true;
        } else if (us < E.pSun + E.pTri) {
            const float ut = fminf((us - E.pSun) / E.pTri, kOneMinusEps);
            const int em = cdfFind(E.emCdf, E.nEm, ut);
            Vec4 a0, a1;
            Vec3 f1, f2;
            int kt;
            if constexpr (INST) {
                a0 = E.emTris[3 * em]; a1 = E.emTris[3 * em + 1];
                f1 = xyz(a1); f2 = xyz(E.emTris[3 * em + 2]); kt = asInt(a0.w);
            } else {
                kt = E.emTri[em];
                a0 = S.verts[3 * kt]; a1 = S.verts[3 * kt + 1];
                f1 = xyz(a1) - xyz(a0); f2 = xyz(S.verts[3 * kt + 2]) - xyz(a0);
            }
            const float su = sqrtf(u2);
            const Vec3 pL = xyz(a0) + f1 * (su * (1.f - u3)) + f2 * (su * u3);
            const Vec3 nL = normalize(cross(f1, f2));
            orgSide = dot(ng, pL - p) > 0.f ? 1 : -1;   // connect from the shadow ray's actual (offset) origin
            const Vec3 org = orgSide > 0 ? poUp : poDn;
            const Vec3 dv = pL - org;
            const float dist2 = dot(dv, dv), dist = sqrtf(dist2);
            if (dist > 0.f) {
                l = dv * (1.f / dist);
                const float cl = fabsf(dot(nL, l));
                if (cl > 0.f) {
                    const int lm = asInt(a1.w);
                    pl = E.pTri * (INST ? E.matLightPdf[lm] : E.triLightPdf[kt]) * dist2 / cl;
                    Le = loadMaterial(E.mats, lm).emission;
                    if (E.maps && lm < E.nMaps) {   // an emission map at the sampled point
                        const int emap = loadMaps(E.maps, lm).emission;
                        if (emap >= 0) {
                            float tu, tv;
                            triUV(E, kt, su * (1.f - u3), su * u3, tu, tv);
                            Le = Le * xyz(texLookup(E, emap, tu, tv));
                        }
                    }
                    tmax = dist * (1.f - 1e-3f);
                    ok = true;
                }
            }
        } else if (E.pEnv > 0.f && E.envUniform) {
            l = uniformSphere(u2, u3); pl = E.pEnv * (0.25f * kInvPi); Le = envRadiance(E, l); ok = true;
        } else if (E.pEnv > 0.f) {
            const float ue = fminf((us - E.pSun - E.pTri) / E.pEnv, kOneMinusEps);
            const int y = cdfFind(E.envMarg, E.envH, ue);
            const float* cr = E.envCond + (size_t)y * (E.envW + 1);
            const int x = cdfFind(cr, E.envW, u2);
            const float v = ((float)y + u3) / (float)E.envH;
            float u = ((float)x + 0.5f) / (float)E.envW;
            const float w0 = cr[x + 1] - cr[x];   // continuous position inside the texel from the rest of u2
            if (w0 > 0.f) u = ((float)x + fminf((u2 - cr[x]) / w0, kOneMinusEps)) / (float)E.envW;
            l = envDir(E, u, v);
            const float pe = envPdfTexel(E, x, y, v);
            if (pe > 0.f) { pl = E.pEnv * pe; Le = xyz(E.envTex[y * E.envW + x]); ok = true; }
        }
        if (ok) {
            const Vec3 lL = v3(dot(l, tx), dot(l, ty), dot(l, ns));
            const float gl = dot(ng, l);
            if (lL.z != 0.f && (gl > 0.f) == (lL.z > 0.f)) {
                const BsdfValue e = evalResponse(surfaceResponse(M, woL, eta, E.gdEnergy, ux1, ux2, nrL), woL, lL);
                const float w = (isSun || E.lightMode == kLightNEE) ? 1.f : misPower(pl, e.pdf);
                const Vec3 con = T * e.f * Le * (fabsf(lL.z) * w / pl);
                if (con.x > 0.f || con.y > 0.f || con.z > 0.f) {
                    out.shadow = true;
                    out.so = (orgSide ? orgSide > 0 : gl > 0.f) ? poUp : poDn;
                    out.sdir = l; out.stmax = tmax; out.scon = con;
                }
            }
        }
    }
    // ---- BSDF sample: the continuation
    Vec3 wiL, wgt;
    float pdf;
    if (!sampleResponse(surfaceResponse(M, woL, eta, E.gdEnergy, ux1, ux2, nrL), woL, rnd((uint32_t)pix, smp, dd, 4, F.seed),
                        rnd((uint32_t)pix, smp, dd, 0, F.seed), rnd((uint32_t)pix, smp, dd, 1, F.seed), wiL, wgt, pdf))
        return out;
    const Vec3 wi = normalize(tx * wiL.x + ty * wiL.y + ns * wiL.z);
    const float gi = dot(ng, wi);
    if (gi == 0.f || (gi > 0.f) != (wiL.z > 0.f)) return out;
    const Vec3 Tn = T * wgt;
    if (!(maxc(Tn) > 0.f)) return out;
    out.cont = true; out.po = gi > 0.f ? poUp : poDn; out.nd = wi; out.Tn = Tn; out.pdf = pdf;
    return out;
}
