// core/bsdf.h - BSDFs of the full integrator's materials (host and device).
//
// Local shading frame: z = shading normal, wo (toward the viewer) has wo.z > 0. The lobes are the ones the reference
// uses (src/render/bsdf.h:1-220), written here from the papers it cites:
//   GGX microfacet distribution, height-correlated Smith masking-shadowing (Walter et al. 2007; Heitz 2014);
//   visible-normal sampling by spherical caps (Dupuy & Benyoub 2023);
//   rough dielectric reflection + refraction with exact unpolarised Fresnel (Walter et al. 2007), transmission in the
//   energy (symmetric) convention: a lossless smooth interface transmits 1 - F;
//   conductor Fresnel: Schlick with the F82 tint term (Kutz et al. 2021, as OpenPBR specifies it);
//   glossy-diffuse: dielectric GGX specular over a Lambertian lobe weighted by 1 - X, with X = F(wo.z) (plain mix) or,
//   with energy compensation, X = an unbiased one-sample estimate of the specular lobe's directional albedo.
// A material blends conductor (weight metallic), dielectric ((1 - metallic) transmission) and glossy-diffuse
// ((1 - metallic)(1 - transmission)); sampling picks one base by its weight and returns the whole material's
// f |cos| / pdf (one-sample MIS over the bases). GGX alpha below kDeltaAlpha is a perfectly specular (delta) lobe,
// reported with pdf = 0.
#pragma once
#include "core/material.h"
#include "core/sampling.h"

namespace cwp {

constexpr float kPi = 3.14159265f;
constexpr float kDeltaAlpha = 1e-3f;
constexpr float kOneMinusEps = 0.99999994f;   // largest float below 1

struct BsdfValue { Vec3 f; float pdf; };

// ------------------------------------------------------------------------------------------------------------ GGX
CWP_HD float ggxD(float cosM, float a2) {
    if (cosM <= 0.f) return 0.f;
    const float c2 = cosM * cosM;
    const float d = c2 * (a2 - 1.f) + 1.f;
    return a2 / (kPi * d * d);
}
CWP_HD float ggxLambda(float cosW, float a2) {
    const float c2 = cosW * cosW;
    if (c2 < 1e-12f) return 1e12f;
    return 0.5f * (sqrtf(1.f + a2 * (1.f - c2) / c2) - 1.f);
}
CWP_HD float ggxG1(float cosW, float a2) { return 1.f / (1.f + ggxLambda(cosW, a2)); }
CWP_HD float ggxG2(float cosO, float cosI, float a2) { return 1.f / (1.f + ggxLambda(cosO, a2) + ggxLambda(cosI, a2)); }
// density of the normals visible from wo: G1(wo) max(0, wo.m) D(m) / wo.z
CWP_HD float ggxVisiblePdf(Vec3 wo, Vec3 m, float a2) {
    const float om = dot(wo, m);
    if (om <= 0.f || m.z <= 0.f) return 0.f;
    return ggxG1(wo.z, a2) * om * ggxD(m.z, a2) / wo.z;
}
// a visible normal for wo: stretch wo to the unit-roughness configuration, sample the spherical cap of directions
// around it, unstretch (Dupuy & Benyoub 2023, isotropic alpha)
CWP_HD Vec3 ggxSampleVisible(Vec3 wo, float a, float u1, float u2) {
    const Vec3 ws = normalize(v3(a * wo.x, a * wo.y, wo.z));
    const float ph = kTwoPi * u1;
    const float z = (1.f - u2) * (1.f + ws.z) - ws.z;
    const float sn = sqrtf(fminf(fmaxf(1.f - z * z, 0.f), 1.f));
    const Vec3 h = v3(sn * cosf(ph), sn * sinf(ph), z) + ws;
    return normalize(v3(a * h.x, a * h.y, fmaxf(h.z, 0.f)));
}

// -------------------------------------------------------------------------------------------------------- Fresnel
// unpolarised dielectric Fresnel reflectance; cosI > 0 on the incident side, eta = eta_transmitted / eta_incident
CWP_HD float fresnelDielectric(float cosI, float eta) {
    const float s2t = (1.f - cosI * cosI) / (eta * eta);
    if (s2t >= 1.f) return 1.f;
    const float cosT = sqrtf(1.f - s2t);
    const float rs = (cosI - eta * cosT) / (cosI + eta * cosT);
    const float rp = (eta * cosI - cosT) / (eta * cosI + cosT);
    return 0.5f * (rs * rs + rp * rp);
}
// Schlick with the F82 tint: F(mu) = Fs(mu) - b mu (1 - mu)^6, b such that F(1/7) = tint * Fs(1/7); clamped to [0, 1]
CWP_HD Vec3 fresnelF82(Vec3 F0, Vec3 tint, float mu) {
    mu = fminf(fmaxf(mu, 0.f), 1.f);
    const float m = 1.f - mu, m2 = m * m, m5 = m2 * m2 * m, m6 = m5 * m;
    const float mb = 1.f / 7.f;
    const float mb5 = 0.46266436f;   // (6/7)^5
    const float mb6 = 0.39656945f;   // (6/7)^6
    const Vec3 one = v3(1.f, 1.f, 1.f);
    const Vec3 Fs = F0 + (one - F0) * m5;
    const Vec3 Fsb = F0 + (one - F0) * mb5;
    const Vec3 b = Fsb * (one - tint) * (1.f / (mb * mb6));
    const Vec3 F = Fs - b * (mu * m6);
    return v3(fminf(fmaxf(F.x, 0.f), 1.f), fminf(fmaxf(F.y, 0.f), 1.f), fminf(fmaxf(F.z, 0.f), 1.f));
}

// ------------------------------------------------------------------------------------------------ glossy-diffuse
// Bump-terminator shadowing of the diffuse lobe (Chiang, Li, Burley 2019): nr = the unperturbed normal in the local
// frame; nr = (0, 0, 1) gives 1 (off).
CWP_HD float bumpShadowing(Vec3 nr, Vec3 wi) {
    if (nr.z >= 1.f) return 1.f;
    const float g = dot(nr, wi) / (wi.z * nr.z);
    if (g >= 1.f) return 1.f;
    if (g < 0.f) return 0.f;
    return g + g * g - g * g * g;
}

struct GlossyDiffuse {
    Vec3 base; float ws, eta, a, a2; bool delta; float Fo, pSpec, kd; Vec3 nr;
    // energy 1: kd = 1 - X with X = ws F(wo.m) G2 / G1(wo) for one visible normal drawn with (ux1, ux2)
    CWP_HD GlossyDiffuse(const Material& m, Vec3 wo, int energy, float ux1, float ux2, Vec3 nrm) {
        base = m.baseColor; ws = m.specularWeight; eta = m.ior; a = m.roughness * m.roughness; a2 = a * a;
        delta = a < kDeltaAlpha; nr = nrm;
        Fo = ws > 0.f ? ws * fresnelDielectric(wo.z, eta) : 0.f;
        const float pd = (1.f - Fo) * luminance(base);
        pSpec = Fo + pd > 0.f ? Fo / (Fo + pd) : 0.f;
        kd = 1.f - Fo;
        if (energy && ws > 0.f && !delta) {
            const Vec3 mm = ggxSampleVisible(wo, a, ux1, ux2);
            const float om = dot(wo, mm);
            const Vec3 wr = mm * (2.f * om) - wo;
            const float X = (om > 0.f && wr.z > 0.f) ? ws * fresnelDielectric(om, eta) * ggxG2(wo.z, wr.z, a2) / ggxG1(wo.z, a2) : 0.f;
            kd = 1.f - X;
        }
        kd = fmaxf(kd, 0.f);
    }
    CWP_HD BsdfValue eval(Vec3 wo, Vec3 wi) const {
        BsdfValue r; r.f = v3(0.f, 0.f, 0.f); r.pdf = 0.f;
        if (wi.z <= 0.f) return r;
        r.f = base * (kd * kInvPi * bumpShadowing(nr, wi));
        r.pdf = (1.f - pSpec) * wi.z * kInvPi;
        if (!delta && Fo > 0.f) {
            const Vec3 m = normalize(wo + wi);
            const float om = dot(wo, m);
            if (om > 0.f) {
                const float F = ws * fresnelDielectric(om, eta);
                const float s = F * ggxD(m.z, a2) * ggxG2(wo.z, wi.z, a2) / (4.f * wo.z * wi.z);
                r.f = r.f + v3(s, s, s);
                r.pdf += pSpec * ggxVisiblePdf(wo, m, a2) / (4.f * om);
            }
        }
        return r;
    }
    CWP_HD bool sample(Vec3 wo, float u, float u1, float u2, Vec3& wi, Vec3& w, float& pdf) const {
        if (u < pSpec) {
            if (delta) {
                wi = v3(-wo.x, -wo.y, wo.z);
                const float s = Fo / pSpec;
                w = v3(s, s, s); pdf = 0.f;
                return true;
            }
            const Vec3 m = ggxSampleVisible(wo, a, u1, u2);
            wi = m * (2.f * dot(wo, m)) - wo;
        } else {
            wi = cosineHemisphere(v3(0.f, 0.f, 1.f), u1, u2);
        }
        if (wi.z <= 0.f) return false;
        const BsdfValue e = eval(wo, wi);
        if (e.pdf <= 0.f) return false;
        w = e.f * (wi.z / e.pdf); pdf = e.pdf;
        return true;
    }
};

// ------------------------------------------------------------------------------------------------------ conductor
struct Conductor {
    Vec3 F0, tint; float ws, a, a2; bool delta;
    CWP_HD explicit Conductor(const Material& m) {
        F0 = m.baseColor; tint = m.specularColor; ws = m.specularWeight; a = m.roughness * m.roughness; a2 = a * a;
        delta = a < kDeltaAlpha;
    }
    CWP_HD BsdfValue eval(Vec3 wo, Vec3 wi) const {
        BsdfValue r; r.f = v3(0.f, 0.f, 0.f); r.pdf = 0.f;
        if (delta || wi.z <= 0.f) return r;
        const Vec3 m = normalize(wo + wi);
        const float om = dot(wo, m);
        if (om <= 0.f) return r;
        r.f = fresnelF82(F0, tint, om) * (ws * ggxD(m.z, a2) * ggxG2(wo.z, wi.z, a2) / (4.f * wo.z * wi.z));
        r.pdf = ggxVisiblePdf(wo, m, a2) / (4.f * om);
        return r;
    }
    CWP_HD bool sample(Vec3 wo, float u1, float u2, Vec3& wi, Vec3& w, float& pdf) const {
        if (delta) {
            wi = v3(-wo.x, -wo.y, wo.z); w = fresnelF82(F0, tint, wo.z) * ws; pdf = 0.f;
            return true;
        }
        const Vec3 m = ggxSampleVisible(wo, a, u1, u2);
        const float om = dot(wo, m);
        wi = m * (2.f * om) - wo;
        if (wi.z <= 0.f || om <= 0.f) return false;
        w = fresnelF82(F0, tint, om) * (ws * ggxG2(wo.z, wi.z, a2) / ggxG1(wo.z, a2));   // F G2 / G1(wo)
        pdf = ggxVisiblePdf(wo, m, a2) / (4.f * om);
        return true;
    }
};

// ----------------------------------------------------------------------------------------------------- dielectric
// eta = index of the far side / index of the wo side; refraction tinted by baseColor
struct Dielectric {
    Vec3 tint; float eta, a, a2; bool delta;
    CWP_HD Dielectric(const Material& m, float etaRel) {
        tint = m.baseColor; eta = etaRel; a = m.roughness * m.roughness; a2 = a * a; delta = a < kDeltaAlpha;
    }
    CWP_HD BsdfValue eval(Vec3 wo, Vec3 wi) const {
        BsdfValue r; r.f = v3(0.f, 0.f, 0.f); r.pdf = 0.f;
        if (delta) return r;
        if (wi.z > 0.f) {
            const Vec3 m = normalize(wo + wi);
            const float om = dot(wo, m);
            if (om <= 0.f) return r;
            const float F = fresnelDielectric(om, eta);
            const float s = F * ggxD(m.z, a2) * ggxG2(wo.z, wi.z, a2) / (4.f * wo.z * wi.z);
            r.f = v3(s, s, s); r.pdf = F * ggxVisiblePdf(wo, m, a2) / (4.f * om);
        } else if (wi.z < 0.f) {
            Vec3 m = normalize(wo + wi * eta);
            if (m.z < 0.f) m = -m;
            const float om = dot(wo, m), im = dot(wi, m);
            if (om <= 0.f || im >= 0.f) return r;
            const float F = fresnelDielectric(om, eta);
            if (F >= 1.f) return r;
            const float den = om + eta * im;
            const float jac = eta * eta * (-im) / (den * den);   // |d m / d wi|
            const float s = (1.f - F) * ggxD(m.z, a2) * ggxG2(wo.z, -wi.z, a2) * om * jac / (wo.z * -wi.z);
            r.f = tint * s; r.pdf = (1.f - F) * ggxVisiblePdf(wo, m, a2) * jac;
        }
        return r;
    }
    CWP_HD bool sample(Vec3 wo, float u, float u1, float u2, Vec3& wi, Vec3& w, float& pdf) const {
        const Vec3 m = delta ? v3(0.f, 0.f, 1.f) : ggxSampleVisible(wo, a, u1, u2);
        const float om = dot(wo, m);
        if (om <= 0.f) return false;
        const float F = fresnelDielectric(om, eta);
        if (u < F) {
            wi = m * (2.f * om) - wo;
            if (wi.z <= 0.f) return false;
            if (delta) { w = v3(1.f, 1.f, 1.f); pdf = 0.f; return true; }
        } else {
            const float cosT = sqrtf(fmaxf(1.f - (1.f - om * om) / (eta * eta), 0.f));
            wi = wo * (-1.f / eta) + m * (om / eta - cosT);
            if (wi.z >= 0.f) return false;
            wi = normalize(wi);
            if (delta) { w = tint; pdf = 0.f; return true; }
        }
        const BsdfValue e = eval(wo, wi);
        if (e.pdf <= 0.f) return false;
        w = e.f * (fabsf(wi.z) / e.pdf); pdf = e.pdf;
        return true;
    }
};

// ------------------------------------------------------------------------------------------------ whole material
CWP_HD void baseWeights(const Material& m, float& wM, float& wT, float& wD) {
    wM = m.metallic; wT = (1.f - m.metallic) * m.transmission; wD = (1.f - m.metallic) * (1.f - m.transmission);
}
// true if some lobe has a density (light sampling can be combined with it)
CWP_HD bool bsdfHasNonDelta(const Material& m) {
    float wM, wT, wD; baseWeights(m, wM, wT, wD);
    const bool rough = m.roughness * m.roughness >= kDeltaAlpha;
    return (wD > 0.f && (luminance(m.baseColor) > 0.f || rough)) || ((wM > 0.f || wT > 0.f) && rough);
}
// f and solid-angle pdf of the non-delta lobes; etaRel for the dielectric base; (gdEnergy, ux1, ux2) the
// glossy-diffuse energy compensation and its random numbers; nr the unperturbed normal (bump terminator)
CWP_HD BsdfValue bsdfEval(const Material& m, Vec3 wo, Vec3 wi, float etaRel, int gdEnergy, float ux1, float ux2,
                          Vec3 nr = v3(0.f, 0.f, 1.f)) {
    float wM, wT, wD; baseWeights(m, wM, wT, wD);
    BsdfValue r; r.f = v3(0.f, 0.f, 0.f); r.pdf = 0.f;
    if (wM > 0.f) { const BsdfValue e = Conductor(m).eval(wo, wi); r.f = r.f + e.f * wM; r.pdf += wM * e.pdf; }
    if (wT > 0.f) { const BsdfValue e = Dielectric(m, etaRel).eval(wo, wi); r.f = r.f + e.f * wT; r.pdf += wT * e.pdf; }
    if (wD > 0.f) {
        const BsdfValue e = GlossyDiffuse(m, wo, gdEnergy, ux1, ux2, nr).eval(wo, wi);
        r.f = r.f + e.f * wD; r.pdf += wD * e.pdf;
    }
    return r;
}
// samples wi; weight = f |cos| / pdf of the whole material; pdf = 0 for a delta sample
CWP_HD bool bsdfSample(const Material& m, Vec3 wo, float etaRel, int gdEnergy, float ux1, float ux2, float u, float u1,
                       float u2, Vec3& wi, Vec3& weight, float& pdf, Vec3 nr = v3(0.f, 0.f, 1.f)) {
    float wM, wT, wD; baseWeights(m, wM, wT, wD);
    bool ok;
    if (u < wM) {
        ok = Conductor(m).sample(wo, u1, u2, wi, weight, pdf);
    } else if (u < wM + wT) {
        const float uu = (u - wM) / wT;
        ok = Dielectric(m, etaRel).sample(wo, fminf(uu, kOneMinusEps), u1, u2, wi, weight, pdf);
    } else {
        const float uu = wD > 0.f ? (u - wM - wT) / wD : 0.f;
        ok = GlossyDiffuse(m, wo, gdEnergy, ux1, ux2, nr).sample(wo, fminf(uu, kOneMinusEps), u1, u2, wi, weight, pdf);
    }
    if (!ok) return false;
    if (pdf <= 0.f) return true;                             // delta: the other bases have no density there
    if (wM == 1.f || wT == 1.f || wD == 1.f) return true;    // one base: its weight is already f |cos| / pdf
    const BsdfValue e = bsdfEval(m, wo, wi, etaRel, gdEnergy, ux1, ux2, nr);
    if (e.pdf <= 0.f) return false;
    weight = e.f * (fabsf(wi.z) / e.pdf); pdf = e.pdf;
    return true;
}

// power heuristic (beta = 2) weight of the strategy with density a against b (Veach 1997)
CWP_HD float misPower(float a, float b) { const float a2 = a * a; return a2 / (a2 + b * b); }

}  // namespace cwp
