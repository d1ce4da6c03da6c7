// host/texture.cpp - image readers (PNG with its own zlib / deflate decoder, TGA, PPM / PGM), the sRGB decode table
// and per-corner tangent frames.
#include "host/texture.h"
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>

namespace cwp {

namespace {

constexpr size_t kMaxTexels = (size_t)1 << 28;   // 1 GiB of RGBA8 per image

bool readFile(const std::string& path, std::vector<uint8_t>& data) {
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) return false;
    data.clear();
    uint8_t buf[1 << 16];
    size_t k;
    while ((k = fread(buf, 1, sizeof buf, f)) > 0) data.insert(data.end(), buf, buf + k);
    const bool ok = !ferror(f);
    fclose(f);
    return ok;
}

uint32_t be32(const uint8_t* p) { return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3]; }
uint32_t rgba8(uint32_t r, uint32_t g, uint32_t b, uint32_t a) { return r | (g << 8) | (b << 16) | (a << 24); }

// ------------------------------------------------------------------------------------------------ deflate
// Bits are consumed least significant first from a 64-bit window refilled byte by byte; reading past the end yields
// zero bits and marks the stream as truncated.
struct BitReader {
    const uint8_t* p; size_t n, pos = 0;
    uint64_t acc = 0; int cnt = 0;
    bool truncated = false;
    BitReader(const uint8_t* d, size_t len) : p(d), n(len) {}
    void refill() {
        while (cnt <= 56 && pos < n) { acc |= (uint64_t)p[pos++] << cnt; cnt += 8; }
    }
    uint32_t peek(int k) { refill(); return (uint32_t)(acc & ((1ull << k) - 1)); }
    void drop(int k) {
        if (k > cnt) { truncated = true; acc = 0; cnt = 0; return; }
        acc >>= k; cnt -= k;
    }
    uint32_t get(int k) { const uint32_t v = peek(k); drop(k); return v; }
    void toByte() { drop(cnt & 7); }
};

// Canonical Huffman code as a direct table over maxLen bits (codes are read bit-reversed, so entry r holds the symbol
// whose reversed code is the low bits of r): symbol << 4 | length, 0 for bit patterns no code starts with.
struct HuffTable {
    std::vector<uint16_t> t;
    int maxLen = 0;
    bool build(const uint8_t* lens, int n) {
        int count[16] = { 0 };
        for (int s = 0; s < n; s++) count[lens[s]]++;
        count[0] = 0;
        int left = 1;
        for (int l = 1; l <= 15; l++) { left = 2 * left - count[l]; if (left < 0) return false; }   // over-subscribed
        maxLen = 0;
        for (int l = 1; l <= 15; l++) if (count[l]) maxLen = l;
        if (maxLen == 0) { t.assign(2, 0); maxLen = 1; return true; }   // no codes: every lookup fails
        int next[16] = { 0 };
        for (int l = 1, code = 0; l <= 15; l++) { code = (code + count[l - 1]) << 1; next[l] = code; }
        t.assign((size_t)1 << maxLen, 0);
        for (int s = 0; s < n; s++) {
            const int l = lens[s];
            if (!l) continue;
            const int code = next[l]++;
            int rev = 0;
            for (int b = 0; b < l; b++) rev |= ((code >> b) & 1) << (l - 1 - b);
            for (int r = rev; r < (1 << maxLen); r += 1 << l) t[r] = (uint16_t)((s << 4) | l);
        }
        return true;
    }
    int decode(BitReader& br) const {
        const uint16_t e = t[br.peek(maxLen)];
        if (!(e & 15)) return -1;
        br.drop(e & 15);
        return br.truncated ? -1 : e >> 4;
    }
};

struct DeflateTables {
    int lenBase[29], lenExtra[29], distBase[30], distExtra[30];
    HuffTable fixedLit, fixedDist;
    DeflateTables() {
        for (int i = 0, base = 3; i < 28; i++) { lenBase[i] = base; lenExtra[i] = i < 8 ? 0 : (i - 4) >> 2; base += 1 << lenExtra[i]; }
        lenBase[28] = 258; lenExtra[28] = 0;
        for (int d = 0, base = 1; d < 30; d++) { distBase[d] = base; distExtra[d] = d < 4 ? 0 : (d - 2) >> 1; base += 1 << distExtra[d]; }
        uint8_t l[288];
        for (int s = 0; s < 288; s++) l[s] = s < 144 ? 8 : (s < 256 ? 9 : (s < 280 ? 7 : 8));
        fixedLit.build(l, 288);
        for (int s = 0; s < 30; s++) l[s] = 5;
        fixedDist.build(l, 30);
    }
};
const DeflateTables& tables() { static const DeflateTables T; return T; }

bool inflateBlocks(BitReader& br, std::vector<uint8_t>& out, size_t maxOut, std::string& err) {
    const DeflateTables& T = tables();
    for (;;) {
        const bool last = br.get(1) != 0;
        const uint32_t type = br.get(2);
        if (br.truncated) { err = "deflate: truncated stream"; return false; }
        if (type == 0) {   // stored
            br.toByte();
            const uint32_t len = br.get(16), nlen = br.get(16);
            if (br.truncated || (len ^ 0xFFFFu) != nlen) { err = "deflate: bad stored block length"; return false; }
            if (out.size() + len > maxOut) { err = "deflate: more data than expected"; return false; }
            for (uint32_t k = 0; k < len; k++) out.push_back((uint8_t)br.get(8));
            if (br.truncated) { err = "deflate: truncated stored block"; return false; }
        } else if (type == 1 || type == 2) {
            HuffTable dynLit, dynDist;
            const HuffTable* lit = &T.fixedLit;
            const HuffTable* dist = &T.fixedDist;
            if (type == 2) {
                static const int order[19] = { 16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15 };
                const int hlit = (int)br.get(5) + 257, hdist = (int)br.get(5) + 1, hclen = (int)br.get(4) + 4;
                if (hlit > 286 || hdist > 30) { err = "deflate: bad code counts"; return false; }
                uint8_t cl[19] = { 0 };
                for (int k = 0; k < hclen; k++) cl[order[k]] = (uint8_t)br.get(3);
                HuffTable clt;
                if (!clt.build(cl, 19)) { err = "deflate: bad code-length code"; return false; }
                uint8_t lens[286 + 30] = { 0 };
                for (int k = 0; k < hlit + hdist;) {
                    const int s = clt.decode(br);
                    if (s < 0) { err = "deflate: bad code-length symbol"; return false; }
                    if (s < 16) { lens[k++] = (uint8_t)s; continue; }
                    int rep, val = 0;
                    if (s == 16) {
                        if (k == 0) { err = "deflate: repeat with no previous length"; return false; }
                        val = lens[k - 1]; rep = 3 + (int)br.get(2);
                    } else if (s == 17) rep = 3 + (int)br.get(3);
                    else rep = 11 + (int)br.get(7);
                    if (k + rep > hlit + hdist) { err = "deflate: code lengths overflow"; return false; }
                    while (rep--) lens[k++] = (uint8_t)val;
                }
                if (lens[256] == 0) { err = "deflate: no end-of-block code"; return false; }
                if (!dynLit.build(lens, hlit) || !dynDist.build(lens + hlit, hdist)) { err = "deflate: bad Huffman code"; return false; }
                lit = &dynLit; dist = &dynDist;
            }
            for (;;) {
                const int s = lit->decode(br);
                if (s < 0 || s > 285) { err = "deflate: bad literal/length code"; return false; }
                if (s < 256) {
                    if (out.size() >= maxOut) { err = "deflate: more data than expected"; return false; }
                    out.push_back((uint8_t)s);
                    continue;
                }
                if (s == 256) break;
                const int li = s - 257;
                const size_t len = (size_t)T.lenBase[li] + br.get(T.lenExtra[li]);
                const int ds = dist->decode(br);
                if (ds < 0 || ds >= 30) { err = "deflate: bad distance code"; return false; }
                const size_t d = (size_t)T.distBase[ds] + br.get(T.distExtra[ds]);
                if (br.truncated) { err = "deflate: truncated stream"; return false; }
                if (d > out.size()) { err = "deflate: distance beyond the output"; return false; }
                if (out.size() + len > maxOut) { err = "deflate: more data than expected"; return false; }
                const size_t from = out.size() - d;
                for (size_t k = 0; k < len; k++) out.push_back(out[from + k]);   // overlapping copies repeat
            }
        } else {
            err = "deflate: reserved block type";
            return false;
        }
        if (last) return true;
    }
}

// ------------------------------------------------------------------------------------------------ PNG
uint32_t crc32(const uint8_t* p, size_t n) {
    static uint32_t table[256];
    static bool init = false;
    if (!init) {
        for (uint32_t i = 0; i < 256; i++) {
            uint32_t c = i;
            for (int k = 0; k < 8; k++) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            table[i] = c;
        }
        init = true;
    }
    uint32_t c = 0xFFFFFFFFu;
    for (size_t i = 0; i < n; i++) c = table[(c ^ p[i]) & 255] ^ (c >> 8);
    return c ^ 0xFFFFFFFFu;
}

uint8_t paeth(int a, int b, int c) {
    const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    return (uint8_t)(pa <= pb && pa <= pc ? a : (pb <= pc ? b : c));
}

bool unfilterRows(uint8_t* data, int rows, size_t rowBytes, int bpp, std::string& err) {
    std::vector<uint8_t> prev(rowBytes, 0);
    for (int y = 0; y < rows; y++) {
        uint8_t* line = data + (size_t)y * (rowBytes + 1);
        const int f = line[0];
        uint8_t* x = line + 1;
        for (size_t i = 0; i < rowBytes; i++) {
            const int a = i >= (size_t)bpp ? x[i - bpp] : 0, b = prev[i], c = i >= (size_t)bpp ? prev[i - bpp] : 0;
            switch (f) {
                case 0: break;
                case 1: x[i] = (uint8_t)(x[i] + a); break;
                case 2: x[i] = (uint8_t)(x[i] + b); break;
                case 3: x[i] = (uint8_t)(x[i] + ((a + b) >> 1)); break;
                case 4: x[i] = (uint8_t)(x[i] + paeth(a, b, c)); break;
                default: err = "PNG: unknown scanline filter " + std::to_string(f); return false;
            }
        }
        memcpy(prev.data(), x, rowBytes);
    }
    return true;
}

bool readPNG(const std::vector<uint8_t>& d, TextureImage& img, std::string& err) {
    size_t pos = 8;
    int w = 0, h = 0, depth = 0, ctype = -1, interlace = 0;
    std::vector<uint8_t> idat, pal, trns;
    bool seenIHDR = false, seenIEND = false;
    while (pos + 12 <= d.size()) {
        const uint32_t len = be32(&d[pos]);
        if (len > d.size() - pos - 12) { err = "PNG: truncated chunk"; return false; }
        const uint8_t* type = &d[pos + 4];
        const uint8_t* data = &d[pos + 8];
        if (crc32(type, (size_t)len + 4) != be32(data + len)) { err = "PNG: chunk CRC mismatch"; return false; }
        const std::string t((const char*)type, 4);
        if (t == "IHDR") {
            if (len != 13) { err = "PNG: bad IHDR"; return false; }
            w = (int)be32(data); h = (int)be32(data + 4); depth = data[8]; ctype = data[9]; interlace = data[12];
            if (data[10] != 0 || data[11] != 0 || interlace > 1) { err = "PNG: unsupported compression, filter or interlace method"; return false; }
            seenIHDR = true;
        } else if (!seenIHDR) {
            err = "PNG: IHDR is not the first chunk"; return false;
        } else if (t == "PLTE") {
            if (len % 3 || len > 768 || len == 0) { err = "PNG: bad palette"; return false; }
            pal.assign(data, data + len);
        } else if (t == "tRNS") {
            trns.assign(data, data + len);
        } else if (t == "IDAT") {
            idat.insert(idat.end(), data, data + len);
        } else if (t == "IEND") {
            seenIEND = true;
            break;
        } else if (!(type[0] & 0x20)) {
            err = "PNG: unknown critical chunk " + t; return false;
        }
        pos += (size_t)len + 12;
    }
    if (!seenIHDR || !seenIEND) { err = "PNG: missing IHDR or IEND"; return false; }
    int ch = 0;
    switch (ctype) {
        case 0: ch = 1; if (depth != 1 && depth != 2 && depth != 4 && depth != 8 && depth != 16) ch = 0; break;
        case 2: ch = 3; if (depth != 8 && depth != 16) ch = 0; break;
        case 3: ch = 1; if (depth != 1 && depth != 2 && depth != 4 && depth != 8) ch = 0; break;
        case 4: ch = 2; if (depth != 8 && depth != 16) ch = 0; break;
        case 6: ch = 4; if (depth != 8 && depth != 16) ch = 0; break;
        default: ch = 0;
    }
    if (!ch) { err = "PNG: unsupported colour type / bit depth " + std::to_string(ctype) + "/" + std::to_string(depth); return false; }
    if (w <= 0 || h <= 0 || (size_t)w * (size_t)h > kMaxTexels) { err = "PNG: bad image size"; return false; }
    if (ctype == 3 && pal.empty()) { err = "PNG: palette image without PLTE"; return false; }
    const int bitsPP = ch * depth, bpp = std::max(1, bitsPP / 8);
    // Adam7 passes (x start, y start, x step, y step); a non-interlaced image is one pass
    static const int adam7[7][4] = { { 0, 0, 8, 8 }, { 4, 0, 8, 8 }, { 0, 4, 4, 8 }, { 2, 0, 4, 4 }, { 0, 2, 2, 4 }, { 1, 0, 2, 2 }, { 0, 1, 1, 2 } };
    static const int single[1][4] = { { 0, 0, 1, 1 } };
    const int (*passes)[4] = interlace ? adam7 : single;
    const int np = interlace ? 7 : 1;
    size_t expect = 0;
    for (int k = 0; k < np; k++) {
        const int pw = (w - passes[k][0] + passes[k][2] - 1) / passes[k][2], ph = (h - passes[k][1] + passes[k][3] - 1) / passes[k][3];
        if (pw > 0 && ph > 0) expect += (size_t)ph * (1 + ((size_t)pw * bitsPP + 7) / 8);
    }
    std::vector<uint8_t> raw;
    raw.reserve(expect);
    if (!inflateZlib(idat.data(), idat.size(), raw, expect, err)) return false;
    if (raw.size() != expect) { err = "PNG: image data has the wrong size"; return false; }
    img.w = w; img.h = h;
    img.rgba.assign((size_t)w * h, 0);
    const uint32_t maxv = (1u << depth) - 1;
    auto sample = [&](const uint8_t* row, int x, int c) -> uint32_t {   // raw sample value
        const size_t bit = ((size_t)x * ch + c) * depth;
        if (depth == 16) return ((uint32_t)row[bit / 8] << 8) | row[bit / 8 + 1];
        if (depth == 8) return row[bit / 8];
        return (row[bit / 8] >> (8 - depth - (int)(bit % 8))) & maxv;
    };
    auto to8 = [&](uint32_t v) -> uint32_t { return depth == 16 ? v >> 8 : (depth == 8 ? v : v * 255u / maxv); };
    size_t off = 0;
    for (int k = 0; k < np; k++) {
        const int xs = passes[k][0], ys = passes[k][1], dx = passes[k][2], dy = passes[k][3];
        const int pw = (w - xs + dx - 1) / dx, ph = (h - ys + dy - 1) / dy;
        if (pw <= 0 || ph <= 0) continue;
        const size_t rb = ((size_t)pw * bitsPP + 7) / 8;
        if (!unfilterRows(raw.data() + off, ph, rb, bpp, err)) return false;
        for (int y = 0; y < ph; y++) {
            const uint8_t* row = raw.data() + off + (size_t)y * (rb + 1) + 1;
            for (int x = 0; x < pw; x++) {
                uint32_t r, g, b, a = 255;
                if (ctype == 0 || ctype == 4) {
                    const uint32_t v = sample(row, x, 0);
                    r = g = b = to8(v);
                    if (ctype == 4) a = to8(sample(row, x, 1));
                    else if (trns.size() >= 2 && v == (((uint32_t)trns[0] << 8) | trns[1])) a = 0;
                } else if (ctype == 2 || ctype == 6) {
                    const uint32_t vr = sample(row, x, 0), vg = sample(row, x, 1), vb = sample(row, x, 2);
                    r = to8(vr); g = to8(vg); b = to8(vb);
                    if (ctype == 6) a = to8(sample(row, x, 3));
                    else if (trns.size() >= 6 && vr == (((uint32_t)trns[0] << 8) | trns[1]) &&
                             vg == (((uint32_t)trns[2] << 8) | trns[3]) && vb == (((uint32_t)trns[4] << 8) | trns[5])) a = 0;
                } else {
                    const uint32_t idx = sample(row, x, 0);
                    if (idx * 3 >= pal.size()) { err = "PNG: palette index out of range"; return false; }
                    r = pal[3 * idx]; g = pal[3 * idx + 1]; b = pal[3 * idx + 2];
                    if (idx < trns.size()) a = trns[idx];
                }
                img.rgba[(size_t)(ys + y * dy) * w + (xs + x * dx)] = rgba8(r, g, b, a);
            }
        }
        off += (size_t)ph * (rb + 1);
    }
    return true;
}

// ------------------------------------------------------------------------------------------------ TGA
bool readTGA(const std::vector<uint8_t>& d, TextureImage& img, std::string& err) {
    if (d.size() < 18) { err = "TGA: truncated header"; return false; }
    const int idLen = d[0], cmapType = d[1], type = d[2];
    const int cmapLen = d[5] | (d[6] << 8), cmapDepth = d[7];
    const int w = d[12] | (d[13] << 8), h = d[14] | (d[15] << 8), depth = d[16], desc = d[17];
    if (type != 2 && type != 3 && type != 10 && type != 11) { err = "TGA: unsupported image type " + std::to_string(type); return false; }
    const bool grey = type == 3 || type == 11, rle = type >= 10;
    if (grey ? depth != 8 : (depth != 24 && depth != 32)) { err = "TGA: unsupported pixel depth " + std::to_string(depth); return false; }
    if (w <= 0 || h <= 0) { err = "TGA: bad image size"; return false; }
    size_t pos = 18 + (size_t)idLen + (cmapType ? (size_t)cmapLen * ((cmapDepth + 7) / 8) : 0);
    const int bpp = depth / 8;
    const size_t n = (size_t)w * h;
    std::vector<uint32_t> px;
    px.reserve(n);
    auto pixel = [&](const uint8_t* p) {
        return grey ? rgba8(p[0], p[0], p[0], 255) : rgba8(p[2], p[1], p[0], bpp == 4 ? p[3] : 255);
    };
    while (px.size() < n) {
        if (!rle) {
            if (pos + bpp > d.size()) { err = "TGA: truncated pixel data"; return false; }
            px.push_back(pixel(&d[pos])); pos += bpp;
            continue;
        }
        if (pos >= d.size()) { err = "TGA: truncated pixel data"; return false; }
        const int hdr = d[pos++], count = (hdr & 127) + 1;
        if (px.size() + count > n) { err = "TGA: run crosses the end of the image"; return false; }
        if (hdr & 128) {
            if (pos + bpp > d.size()) { err = "TGA: truncated pixel data"; return false; }
            const uint32_t v = pixel(&d[pos]); pos += bpp;
            for (int k = 0; k < count; k++) px.push_back(v);
        } else {
            if (pos + (size_t)count * bpp > d.size()) { err = "TGA: truncated pixel data"; return false; }
            for (int k = 0; k < count; k++) { px.push_back(pixel(&d[pos])); pos += bpp; }
        }
    }
    img.w = w; img.h = h;
    img.rgba.resize(n);
    const bool top = (desc & 0x20) != 0, right = (desc & 0x10) != 0;   // origin at the top / at the right
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++)
            img.rgba[(size_t)(top ? y : h - 1 - y) * w + (right ? w - 1 - x : x)] = px[(size_t)y * w + x];
    return true;
}

// ------------------------------------------------------------------------------------------------ PPM / PGM
bool readPNM(const std::vector<uint8_t>& d, TextureImage& img, std::string& err) {
    const char kind = (char)d[1];
    const bool grey = kind == '2' || kind == '5', ascii = kind == '2' || kind == '3';
    size_t pos = 2;
    auto skipSpace = [&]() {
        for (;;) {
            while (pos < d.size() && isspace(d[pos])) pos++;
            if (pos < d.size() && d[pos] == '#') { while (pos < d.size() && d[pos] != '\n') pos++; continue; }
            return;
        }
    };
    auto number = [&](long& v) {
        skipSpace();
        if (pos >= d.size() || !isdigit(d[pos])) return false;
        v = 0;
        while (pos < d.size() && isdigit(d[pos])) { v = v * 10 + (d[pos++] - '0'); if (v > (1L << 30)) return false; }
        return true;
    };
    long w, h, maxv;
    if (!number(w) || !number(h) || !number(maxv) || w <= 0 || h <= 0 || maxv <= 0 || maxv > 65535 ||
        (size_t)w * (size_t)h > kMaxTexels) { err = "PNM: bad header"; return false; }
    const int ch = grey ? 1 : 3, bytes = maxv > 255 ? 2 : 1;
    const size_t n = (size_t)w * h;
    if (!ascii) {
        if (pos >= d.size() || !isspace(d[pos])) { err = "PNM: bad header"; return false; }
        pos++;   // exactly one whitespace byte before binary data
        if (d.size() - pos < n * ch * bytes) { err = "PNM: truncated pixel data"; return false; }
    }
    img.w = (int)w; img.h = (int)h;
    img.rgba.resize(n);
    auto scale = [&](long v) -> uint32_t { return (uint32_t)((v * 255 + maxv / 2) / maxv); };
    for (size_t i = 0; i < n; i++) {
        uint32_t c[3];
        for (int k = 0; k < ch; k++) {
            long v;
            if (ascii) { if (!number(v)) { err = "PNM: truncated pixel data"; return false; } }
            else if (bytes == 2) { v = (d[pos] << 8) | d[pos + 1]; pos += 2; }
            else v = d[pos++];
            if (v > maxv) { err = "PNM: sample above maxval"; return false; }
            c[k] = scale(v);
        }
        img.rgba[i] = grey ? rgba8(c[0], c[0], c[0], 255) : rgba8(c[0], c[1], c[2], 255);
    }
    return true;
}

}  // namespace

bool inflateZlib(const uint8_t* in, size_t n, std::vector<uint8_t>& out, size_t maxOut, std::string& err) {
    out.clear();
    if (n < 6) { err = "zlib: truncated stream"; return false; }
    if ((in[0] & 15) != 8 || (in[0] >> 4) > 7 || ((in[0] << 8) | in[1]) % 31 || (in[1] & 0x20)) { err = "zlib: bad header"; return false; }
    BitReader br(in + 2, n - 2);
    if (!inflateBlocks(br, out, maxOut, err)) return false;
    br.toByte();
    uint32_t want = 0;
    for (int k = 0; k < 4; k++) want = (want << 8) | br.get(8);
    uint32_t a = 1, b = 0;
    for (uint8_t c : out) { a = (a + c) % 65521u; b = (b + a) % 65521u; }
    if (br.truncated || want != ((b << 16) | a)) { err = "zlib: checksum mismatch"; return false; }
    return true;
}

bool readImage(const std::string& path, TextureImage& img, std::string& err) {
    std::vector<uint8_t> d;
    if (!readFile(path, d)) { err = "cannot read image " + path; return false; }
    img = TextureImage();
    img.path = path;
    static const uint8_t pngSig[8] = { 0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A };
    bool ok;
    if (d.size() >= 8 && memcmp(d.data(), pngSig, 8) == 0) ok = readPNG(d, img, err);
    else if (d.size() >= 2 && d[0] == 'P' && d[1] >= '2' && d[1] <= '6' && d[1] != '4') ok = readPNM(d, img, err);
    else {
        const std::string ext = path.size() >= 4 ? path.substr(path.size() - 4) : "";
        if (ext == ".tga" || ext == ".TGA") ok = readTGA(d, img, err);
        else { err = "unsupported image format (PNG, TGA, PPM / PGM)"; ok = false; }
    }
    if (!ok) err = path + ": " + err;
    return ok;
}

const float* srgbDecodeTable() {
    static float table[256];
    static bool init = false;
    if (!init) {
        for (int b = 0; b < 256; b++) {
            const double c = b / 255.0;
            table[b] = (float)(c <= 0.04045 ? c / 12.92 : std::pow((c + 0.055) / 1.055, 2.4));
        }
        init = true;
    }
    return table;
}

void cornerTangents(const Mesh& m, const std::vector<float>& uv, std::vector<Vec4>& tan) {
    const size_t nv = m.v.size(), nt = m.f.size();
    std::vector<Vec3> su(nv, v3(0.f, 0.f, 0.f)), sv(nv, v3(0.f, 0.f, 0.f)), fn;
    const bool useN = m.hasN();
    if (!useN) fn.assign(nv, v3(0.f, 0.f, 0.f));
    for (size_t t = 0; t < nt; t++) {
        const int id[3] = { m.f[t].a, m.f[t].b, m.f[t].c };
        const Vec3 e1 = m.v[id[1]] - m.v[id[0]], e2 = m.v[id[2]] - m.v[id[0]];
        const float* q = &uv[6 * t];
        const float du1 = q[2] - q[0], dv1 = q[3] - q[1], du2 = q[4] - q[0], dv2 = q[5] - q[1];
        if (!useN) { const Vec3 c = cross(e1, e2); for (int k = 0; k < 3; k++) fn[id[k]] = fn[id[k]] + c; }
        const float det = du1 * dv2 - du2 * dv1;
        if (!(std::fabs(det) >= 1e-20f)) continue;
        const float r = 1.f / det;
        const Vec3 dPdu = (e1 * dv2 - e2 * dv1) * r, dPdv = (e2 * du1 - e1 * du2) * r;
        for (int k = 0; k < 3; k++) { su[id[k]] = su[id[k]] + dPdu; sv[id[k]] = sv[id[k]] + dPdv; }
    }
    std::vector<Vec4> perVertex(nv);
    for (size_t i = 0; i < nv; i++) {
        const Vec3 n = normalize(useN ? m.n[i] : fn[i]);
        Vec3 t = su[i] - n * dot(n, su[i]);
        if (dot(t, t) < 1e-24f) {   // no usable derivative: any unit vector orthogonal to n
            const Vec3 a = std::fabs(n.x) < 0.9f ? v3(1.f, 0.f, 0.f) : v3(0.f, 1.f, 0.f);
            t = cross(a, n);
        }
        t = normalize(t);
        perVertex[i] = v4(t, dot(cross(n, t), sv[i]) < 0.f ? -1.f : 1.f);
    }
    tan.resize(3 * nt);
    for (size_t t = 0; t < nt; t++) {
        tan[3 * t] = perVertex[m.f[t].a]; tan[3 * t + 1] = perVertex[m.f[t].b]; tan[3 * t + 2] = perVertex[m.f[t].c];
    }
}

}  // namespace cwp
