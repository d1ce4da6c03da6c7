// host/io.h - the reference's file formats: camera / lighting JSON, OBJ (+MTL) and PLY meshes, PNG / PFM images.
// Formats and key semantics follow README.md "Input formats" and src/io/{json,scene_files,mesh_loader,image}.cpp.
#pragma once
#include "host/camera.h"
#include "host/mesh.h"
#include <string>
#include <vector>

namespace cwp {

// Recursive-descent JSON value (objects keep key order).
struct Json {
    enum Type { Null, Bool, Number, String, Array, Object } type = Null;
    bool b = false; double num = 0; std::string str;
    std::vector<Json> arr;
    std::vector<std::pair<std::string, Json>> obj;
    const Json* get(const std::string& key) const;
    double number(const std::string& key, double def) const;
    std::string string(const std::string& key, const std::string& def) const;
    std::vector<double> numbers(const std::string& key) const;   // flattened numbers of an array member
};
bool parseJson(const std::string& text, Json& out, std::string* err);
std::string readTextFile(const std::string& path);   // "" if unreadable

// First value stored under `key` anywhere in the document (depth-first, document order).
const Json* findKey(const Json& doc, const std::string& key);
// All numbers inside the value of `key`, nested arrays flattened.
std::vector<double> keyNumbers(const Json& doc, const std::string& key);

// Lighting file: sun_dir_to_light_world_zup, sun_irradiance_W_m2, sun_angle_rad, sky_radiance_rgb_linear and
// albedo_rgb_linear_by_material_index (replaces the mesh albedo table, then padded to every used id).
bool loadLightingFile(const std::string& path, Lighting& L, Mesh* m, std::string& err);
// Materials file {"materials": [{name?, base_color, metallic, roughness, ior, specular_weight, specular_color,
// transmission, emission_color, emission_strength}, ...]}: entries start from OpenPBR's defaults; an entry whose name
// matches a mesh material name replaces that id, the others take ids in file order; ids the file does not give stay
// Lambertian with the mesh's kd (reference: loadMaterialsFile, src/io/scene_files.cpp:28-55).
bool loadMaterialsFile(const std::string& path, Mesh& m, std::string& err);
// Camera file: fx/fy used only when still 0 and width/height match; else hfov_deg when still 0; c2w from key[frame]
// (the key holds one matrix or a list of per-frame matrices); optional aperture_radius / focus_distance.
bool loadCameraFile(const std::string& path, const std::string& key, int frame, int w, int h, float& fx, float& fy,
                    float& hfov, float c2w[16], float& apertureR, float& focusD, std::string& err);

bool loadOBJ(const std::string& path, Mesh& m);
bool loadPLY(const std::string& path, Mesh& m);
bool loadMesh(const std::string& path, Mesh& m);   // by extension

// PNG (8-bit RGB, stored deflate blocks), PFM (linear float RGB, bottom-to-top rows), clip + sRGB OETF tone map
bool writePNG(const std::string& path, int w, int h, const uint8_t* rgb);
bool writePFM(const std::string& path, int w, int h, const float* rgb);
void tonemapSRGB8(const float* lin, size_t nPix, uint8_t* rgb);
// the 8-bit encode of tonemapSRGB8 for one value: clip to [0, 1], sRGB OETF, x * 255 + 0.5 truncated
uint8_t encodeSRGB8(float x);
// step positions of encodeSRGB8 for the device display transform: t[k] (k = 1..255) is the smallest float in [0, 1]
// that encodes to k or more, t[0] = -inf (the encode is non-decreasing in x)
void srgbSteps(float t[256]);

}  // namespace cwp
