// core/material.h - the material of the full integrator: an OpenPBR Surface subset (base, specular, transmission and
// emission layers), 64 bytes, shared by host and device. Same fields and meaning as the reference's Material
// (src/scene/material.h:12-17), so that materials files and generated materials mean the same thing:
//   metallic 1                 rough conductor, F0 = baseColor, F82 tint = specularColor
//   metallic 0, transmission 1 rough dielectric, refraction tinted by baseColor, index of refraction ior
//   metallic 0, transmission 0 glossy-diffuse: GGX dielectric specular (weight specularWeight, ior) over a Lambertian
//                              baseColor lobe weighted by 1 - F
//   emission                   radiance (colour * strength), two-sided, added on top
// metallic and transmission blend the three bases linearly; roughness r gives GGX alpha = r^2.
#pragma once
#include "core/vecmath.h"

namespace cwp {

struct Material {
    Vec3 baseColor; float metallic;
    Vec3 specularColor; float roughness;
    Vec3 emission; float ior;
    float specularWeight, transmission, pad0, pad1;
};
static_assert(sizeof(Material) == 64, "Material is 64 bytes (four float4 loads)");

// pure Lambertian material (what OBJ Kd, generated albedos and lighting albedos give)
CWP_HD Material lambertMaterial(Vec3 kd) {
    Material m;
    m.baseColor = kd; m.metallic = 0.f; m.specularColor = v3(1.f, 1.f, 1.f); m.roughness = 1.f;
    m.emission = v3(0.f, 0.f, 0.f); m.ior = 1.5f; m.specularWeight = 0.f; m.transmission = 0.f; m.pad0 = 0.f; m.pad1 = 0.f;
    return m;
}
// OpenPBR's defaults (specular on, roughness 0.5): what a materials-file entry starts from
CWP_HD Material defaultMaterial() {
    Material m = lambertMaterial(v3(0.8f, 0.8f, 0.8f));
    m.roughness = 0.5f; m.specularWeight = 1.f;
    return m;
}
CWP_HD bool isLambert(const Material& m) { return m.metallic == 0.f && m.transmission == 0.f && m.specularWeight == 0.f; }
CWP_HD bool isEmissive(const Material& m) { return m.emission.x > 0.f || m.emission.y > 0.f || m.emission.z > 0.f; }
CWP_HD float luminance(Vec3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }

}  // namespace cwp
