// host/texture.h - texture images on the host: readers for PNG, TGA and PPM / PGM files, the image set of a scene
// (one RGBA8 image per (file, decode) pair), the sRGB decode table and per-corner tangent frames for normal maps.
//
// The reference reads textures with stb_image through its Assimp loader (src/io/assimp_loader.cpp:42-66); those
// libraries are fetched from the network and absent here, so the readers below are this build's own:
//   PNG  zlib / deflate decoder (RFC 1950 / 1951: stored, fixed and dynamic Huffman blocks, Adler-32 and chunk CRCs
//        checked), every colour type (grey, RGB, palette, grey + alpha, RGBA) at every bit depth (1-16), the five
//        scanline filters and Adam7 interlacing, tRNS transparency;
//   TGA  uncompressed and run-length encoded true-colour (24 / 32 bit) and grey (8 bit) images, either row origin;
//   PPM / PGM  binary (P6 / P5) and ASCII (P3 / P2), maxval up to 65535.
// 16-bit samples keep their high byte. Every image becomes RGBA8 rows top first.
#pragma once
#include "core/texture.h"
#include "host/mesh.h"
#include <string>
#include <vector>

namespace cwp {

bool readImage(const std::string& path, TextureImage& img, std::string& err);
// zlib stream (RFC 1950) -> bytes; false with a reason on malformed input, a bad checksum or more than maxOut bytes
bool inflateZlib(const uint8_t* in, size_t n, std::vector<uint8_t>& out, size_t maxOut, std::string& err);

// sRGB electro-optical transfer of the 256 byte values (computed in double, rounded to float)
const float* srgbDecodeTable();

// Tangent frames per triangle corner (3 per triangle, (tangent, handedness w = +-1), bitangent = w cross(n, t)) from
// texture-coordinate derivatives (Lengyel 2001): per triangle T = dP/du and B = dP/dv, summed at its vertices, made
// orthogonal to the vertex normal (the mesh's, or the area-weighted sum of its faces' normals); a vertex without a
// usable derivative gets any unit tangent orthogonal to its normal. uv: 6 floats per triangle (corner u, v).
void cornerTangents(const Mesh& m, const std::vector<float>& uv, std::vector<Vec4>& tan);

}  // namespace cwp
