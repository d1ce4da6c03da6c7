// core/bsdf.h - the full integrator's material response (host and device).
//
// A material of the OpenPBR subset blends three bases by fixed weights: a conductor (weight metallic), a dielectric
// interface ((1 - metallic) transmission) and a glossy-diffuse layer ((1 - metallic)(1 - transmission)). All three
// share one GGX roughness, so in this design the response is built once per shading point (SurfaceResponse: the base
// weights, the microfacet terms that depend only on wo, the layer's Fresnel-derived lobe probability and diffuse
// scale) and every evaluation computes the reflection half-vector, its distribution value and the height-correlated
// masking term ONCE, then adds each base's Fresnel factor on top. The reference evaluates each base as its own object,
// recomputing the microfacet terms per base (src/render/bsdf.h:60-220); the estimator is the same:
//   GGX distribution, height-correlated Smith masking-shadowing (Walter et al. 2007; Heitz 2014);
//   visible-normal sampling by spherical caps (Dupuy & Benyoub 2023, isotropic);
//   exact unpolarised dielectric Fresnel; conductor Fresnel = Schlick with the F82 tint term (Kutz et al. 2021, as
//   OpenPBR specifies it);
//   transmission in the energy (symmetric) convention: a lossless smooth interface transmits 1 - F;
//   glossy-diffuse = dielectric GGX coat over a Lambertian lobe scaled by 1 - X, X = F(wo.z) (plain) or, with energy
//   compensation, a one-sample estimate ws F(wo.m) G2 / G1(wo) of the coat's directional albedo;
//   optional bump-terminator shadowing of the diffuse lobe (Chiang, Li & Burley 2019).
// Local frame: z = shading normal, wo.z > 0. Alpha = roughness^2; alpha below kDeltaAlpha makes every microfacet lobe
// a delta lobe, which evaluation ignores and sampling reports with pdf 0. Sampling picks a base by its weight; with
// more than one base present the returned weight and pdf are those of the whole material (one-sample MIS over bases).
#pragma once
#include "core/material.h"
#include "core/sampling.h"

namespace cwp {

constexpr float kPi = 3.14159265f;
constexpr float kDeltaAlpha = 1e-3f;
constexpr float kOneMinusEps = 0.99999994f;   // largest float below 1

struct BsdfValue { Vec3 f; float pdf; };

// --------------------------------------------------------------------------------------- microfacet building blocks
// GGX normal distribution D(m) for cos(theta_m) = c and alpha^2 = a2
CWP_HD float ggxDistribution(float c, float a2) {
    if (c <= 0.f) return 0.f;
    const float t = c * c * (a2 - 1.f) + 1.f;
    return a2 / (kPi * t * t);
}
// Smith Lambda of a direction with cos(theta) = c (large for grazing directions)
CWP_HD float smithLambda(float c, float a2) {
    const float c2 = c * c;
    if (c2 < 1e-12f) return 1e12f;
    return 0.5f * (sqrtf(1.f + a2 * (1.f - c2) / c2) - 1.f);
}
// a visible normal for wo: stretch to unit roughness, sample the spherical cap around the stretched direction, unstretch
CWP_HD Vec3 sampleVisibleNormal(Vec3 wo, float a, float u1, float u2) {
    const Vec3 s = normalize(v3(a * wo.x, a * wo.y, wo.z));
    const float phi = kTwoPi * u1;
    const float z = (1.f - u2) * (1.f + s.z) - s.z;
    const float r = sqrtf(fminf(fmaxf(1.f - z * z, 0.f), 1.f));
    const Vec3 h = s + v3(r * cosf(phi), r * sinf(phi), z);
    return normalize(v3(a * h.x, a * h.y, fmaxf(h.z, 0.f)));
}
CWP_HD Vec3 reflectAbout(Vec3 w, Vec3 m) { return m * (2.f * dot(w, m)) - w; }

// unpolarised Fresnel reflectance of a dielectric interface; c = cosine on the incident side (> 0), eta = eta_t / eta_i
CWP_HD float dielectricFresnel(float c, float eta) {
    const float sin2t = (1.f - c * c) / (eta * eta);
    if (sin2t >= 1.f) return 1.f;   // total internal reflection
    const float ct = sqrtf(1.f - sin2t);
    const float perp = (c - eta * ct) / (c + eta * ct);
    const float para = (eta * c - ct) / (eta * c + ct);
    return 0.5f * (perp * perp + para * para);
}
// conductor reflectance: Schlick's curve lowered by b mu (1 - mu)^6, b fixed so that F(1/7) = tint * Schlick(1/7);
// clamped to [0, 1]
CWP_HD Vec3 conductorFresnel(Vec3 f0, Vec3 tint, float mu) {
    mu = fminf(fmaxf(mu, 0.f), 1.f);
    const float q = 1.f - mu, q2 = q * q, q5 = q2 * q2 * q, q6 = q5 * q;
    const float muB = 1.f / 7.f, qB5 = 0.46266436f /* (6/7)^5 */, qB6 = 0.39656945f /* (6/7)^6 */;
    const Vec3 one = v3(1.f, 1.f, 1.f);
    const Vec3 schlick = f0 + (one - f0) * q5;
    const Vec3 b = (f0 + (one - f0) * qB5) * (one - tint) * (1.f / (muB * qB6));
    const Vec3 F = schlick - b * (mu * q6);
    return v3(fminf(fmaxf(F.x, 0.f), 1.f), fminf(fmaxf(F.y, 0.f), 1.f), fminf(fmaxf(F.z, 0.f), 1.f));
}
// bump-terminator shadowing of the diffuse lobe: nr = the unperturbed normal in the local frame ((0, 0, 1): off)
CWP_HD float terminatorShadowing(Vec3 nr, Vec3 wi) {
    if (nr.z >= 1.f) return 1.f;
    const float g = dot(nr, wi) / (wi.z * nr.z);
    if (g >= 1.f) return 1.f;
    if (g < 0.f) return 0.f;
    return g + g * g - g * g * g;
}

// ------------------------------------------------------------------------------------------------ surface response
enum : int { kBaseMetal = 1, kBaseGlass = 2, kBaseLayer = 4, kBaseAll = 7 };

struct SurfaceResponse {
    float wMetal, wGlass, wLayer;   // base weights (sum 1)
    float a, a2;                    // GGX alpha and alpha^2
    bool smooth;                    // alpha below kDeltaAlpha: delta lobes
    float cosO, lambdaO, g1O;       // wo.z, Smith Lambda(wo), G1(wo)
    Vec3 f0, tint;                  // conductor: normal-incidence colour and F82 tint
    float specW;                    // specular weight (conductor and coat)
    float eta;                      // dielectric base: far-side / wo-side index
    Vec3 glassTint;                 // refraction tint (baseColor)
    float coatEta, coatF, coatP;    // coat index, ws F(wo.z), probability of sampling the coat within the layer
    float diffuseK;                 // Lambertian scale (1 - X, clamped at 0)
    Vec3 diffuse, nr;               // Lambertian colour, unperturbed normal for terminator shadowing
};

// the response of material m seen from wo; etaRel = index ratio of the dielectric base for this side; (gdEnergy, ux1,
// ux2) = the layer's energy compensation and the random numbers of its one-sample estimate
CWP_HD SurfaceResponse surfaceResponse(const Material& m, Vec3 wo, float etaRel, int gdEnergy, float ux1, float ux2,
                                       Vec3 nr = v3(0.f, 0.f, 1.f)) {
    SurfaceResponse R;
    R.wMetal = m.metallic;
    R.wGlass = (1.f - m.metallic) * m.transmission;
    R.wLayer = (1.f - m.metallic) * (1.f - m.transmission);
    R.a = m.roughness * m.roughness; R.a2 = R.a * R.a;
    R.smooth = R.a < kDeltaAlpha;
    R.cosO = wo.z;
    R.lambdaO = smithLambda(wo.z, R.a2);
    R.g1O = 1.f / (1.f + R.lambdaO);
    R.f0 = m.baseColor; R.tint = m.specularColor; R.specW = m.specularWeight;
    R.eta = etaRel; R.glassTint = m.baseColor;
    R.coatEta = m.ior; R.diffuse = m.baseColor; R.nr = nr;
    R.coatF = R.specW > 0.f ? R.specW * dielectricFresnel(wo.z, R.coatEta) : 0.f;
    const float diffuseLum = (1.f - R.coatF) * luminance(m.baseColor);
    R.coatP = R.coatF + diffuseLum > 0.f ? R.coatF / (R.coatF + diffuseLum) : 0.f;
    float X = R.coatF;
    if (gdEnergy && R.specW > 0.f && !R.smooth) {
        const Vec3 mm = sampleVisibleNormal(wo, R.a, ux1, ux2);
        const float om = dot(wo, mm);
        const Vec3 wr = mm * (2.f * om) - wo;
        X = 0.f;
        if (om > 0.f && wr.z > 0.f)
            X = R.specW * dielectricFresnel(om, R.coatEta) * (1.f + R.lambdaO) / (1.f + R.lambdaO + smithLambda(wr.z, R.a2));
    }
    R.diffuseK = fmaxf(1.f - X, 0.f);
    return R;
}

// true if some lobe of m has a density (light samples can be weighed against it)
CWP_HD bool hasDensity(const SurfaceResponse& R) {
    return (R.wLayer > 0.f && (luminance(R.diffuse) > 0.f || !R.smooth)) || ((R.wMetal > 0.f || R.wGlass > 0.f) && !R.smooth);
}

// f and solid-angle pdf of the bases in `bases`, each scaled by its weight; delta lobes contribute nothing
CWP_HD BsdfValue evalResponse(const SurfaceResponse& R, Vec3 wo, Vec3 wi, int bases = kBaseAll) {
    BsdfValue r; r.f = v3(0.f, 0.f, 0.f); r.pdf = 0.f;
    const float wM = (bases & kBaseMetal) ? R.wMetal : 0.f, wG = (bases & kBaseGlass) ? R.wGlass : 0.f,
                wL = (bases & kBaseLayer) ? R.wLayer : 0.f;
    if (wi.z > 0.f) {
        if (wL > 0.f) {   // Lambertian part of the layer
            r.f = R.diffuse * (wL * (R.diffuseK * kInvPi * terminatorShadowing(R.nr, wi)));
            r.pdf = wL * ((1.f - R.coatP) * wi.z * kInvPi);
        }
        if (R.smooth) return r;
        const Vec3 h = normalize(wo + wi);
        const float oh = dot(wo, h);
        if (oh <= 0.f) return r;
        // shared microfacet reflection terms: D G2 / (4 cos_o cos_i), and the visible-normal density mapped to wi
        const float D = ggxDistribution(h.z, R.a2);
        const float spec = D / (1.f + R.lambdaO + smithLambda(wi.z, R.a2)) / (4.f * R.cosO * wi.z);
        const float dens = h.z > 0.f ? R.g1O * oh * D / R.cosO / (4.f * oh) : 0.f;
        if (wM > 0.f) {
            r.f = r.f + conductorFresnel(R.f0, R.tint, oh) * (wM * R.specW * spec);
            r.pdf += wM * dens;
        }
        if (wG > 0.f) {
            const float F = dielectricFresnel(oh, R.eta);
            r.f = r.f + v3(1.f, 1.f, 1.f) * (wG * F * spec);
            r.pdf += wG * F * dens;
        }
        if (wL > 0.f && R.coatF > 0.f) {
            const float F = R.specW * dielectricFresnel(oh, R.coatEta);
            r.f = r.f + v3(1.f, 1.f, 1.f) * (wL * F * spec);
            r.pdf += wL * R.coatP * dens;
        }
    } else if (wi.z < 0.f && wG > 0.f && !R.smooth) {
        // refraction through the dielectric base: generalised half-vector on the wo side
        Vec3 h = normalize(wo + wi * R.eta);
        if (h.z < 0.f) h = -h;
        const float oh = dot(wo, h), ih = dot(wi, h);
        if (oh <= 0.f || ih >= 0.f) return r;
        const float F = dielectricFresnel(oh, R.eta);
        if (F >= 1.f) return r;
        const float den = oh + R.eta * ih;
        const float jac = R.eta * R.eta * (-ih) / (den * den);   // |d h / d wi|
        const float D = ggxDistribution(h.z, R.a2);
        const float g2 = 1.f / (1.f + R.lambdaO + smithLambda(wi.z, R.a2));
        r.f = R.glassTint * (wG * (1.f - F) * D * g2 * oh * jac / (R.cosO * -wi.z));
        r.pdf = h.z > 0.f ? wG * (1.f - F) * (R.g1O * oh * D / R.cosO) * jac : 0.f;
    }
    return r;
}

// Samples wi from base `base` (one of kBaseMetal / kBaseGlass / kBaseLayer) with its lobe choice u in [0, 1).
// Returns false when the sample is rejected; delta samples return pdf 0 and their weight.
CWP_HD bool sampleBase(const SurfaceResponse& R, int base, Vec3 wo, float u, float u1, float u2, Vec3& wi, Vec3& weight,
                       float& pdf) {
    if (base == kBaseMetal) {
        if (R.smooth) {
            wi = v3(-wo.x, -wo.y, wo.z); weight = conductorFresnel(R.f0, R.tint, wo.z) * R.specW; pdf = 0.f;
            return true;
        }
        const Vec3 h = sampleVisibleNormal(wo, R.a, u1, u2);
        const float oh = dot(wo, h);
        wi = reflectAbout(wo, h);
        if (wi.z <= 0.f || oh <= 0.f) return false;
        // F G2 / G1(wo) with one visible normal
        weight = conductorFresnel(R.f0, R.tint, oh) * (R.specW * (1.f + R.lambdaO) / (1.f + R.lambdaO + smithLambda(wi.z, R.a2)));
        pdf = h.z > 0.f ? R.g1O * oh * ggxDistribution(h.z, R.a2) / R.cosO / (4.f * oh) : 0.f;
        return true;
    }
    if (base == kBaseGlass) {
        const Vec3 h = R.smooth ? v3(0.f, 0.f, 1.f) : sampleVisibleNormal(wo, R.a, u1, u2);
        const float oh = dot(wo, h);
        if (oh <= 0.f) return false;
        const float F = dielectricFresnel(oh, R.eta);
        if (u < F) {
            wi = reflectAbout(wo, h);
            if (wi.z <= 0.f) return false;
            if (R.smooth) { weight = v3(1.f, 1.f, 1.f); pdf = 0.f; return true; }
        } else {
            const float ct = sqrtf(fmaxf(1.f - (1.f - oh * oh) / (R.eta * R.eta), 0.f));
            wi = wo * (-1.f / R.eta) + h * (oh / R.eta - ct);
            if (wi.z >= 0.f) return false;
            wi = normalize(wi);
            if (R.smooth) { weight = R.glassTint; pdf = 0.f; return true; }
        }
    } else {
        if (u < R.coatP) {
            if (R.smooth) {
                wi = v3(-wo.x, -wo.y, wo.z);
                const float s = R.coatF / R.coatP;
                weight = v3(s, s, s); pdf = 0.f;
                return true;
            }
            wi = reflectAbout(wo, sampleVisibleNormal(wo, R.a, u1, u2));
        } else {
            wi = cosineHemisphere(v3(0.f, 0.f, 1.f), u1, u2);
        }
        if (wi.z <= 0.f) return false;
    }
    const BsdfValue e = evalResponse(R, wo, wi, base);
    if (e.pdf <= 0.f) return false;
    weight = e.f * (fabsf(wi.z) / e.pdf); pdf = e.pdf;
    return true;
}

// Samples the whole material: a base chosen by u against the weights (the rest of u is the base's lobe choice), then,
// unless the sample is a delta or the material has one base, weight and pdf of the whole material at wi.
CWP_HD bool sampleResponse(const SurfaceResponse& R, Vec3 wo, float u, float u1, float u2, Vec3& wi, Vec3& weight, float& pdf) {
    int base;
    float ub;
    if (u < R.wMetal) { base = kBaseMetal; ub = 0.f; }
    else if (u < R.wMetal + R.wGlass) { base = kBaseGlass; ub = fminf((u - R.wMetal) / R.wGlass, kOneMinusEps); }
    else { base = kBaseLayer; ub = R.wLayer > 0.f ? fminf((u - R.wMetal - R.wGlass) / R.wLayer, kOneMinusEps) : 0.f; }
    if (!sampleBase(R, base, wo, ub, u1, u2, wi, weight, pdf)) return false;
    if (pdf <= 0.f) return true;                                              // delta: no other base has density there
    if (R.wMetal == 1.f || R.wGlass == 1.f || R.wLayer == 1.f) return true;   // one base: its weight is the material's
    const BsdfValue e = evalResponse(R, wo, wi);
    if (e.pdf <= 0.f) return false;
    weight = e.f * (fabsf(wi.z) / e.pdf); pdf = e.pdf;
    return true;
}

// power heuristic (beta = 2) weight of the strategy with density a against b (Veach 1997)
CWP_HD float misPower(float a, float b) { const float a2 = a * a; return a2 / (a2 + b * b); }

}  // namespace cwp
