// host/bvhcache.cpp - see bvhcache.h.
#include "host/bvhcache.h"
#include <cstdio>
#include <cstring>
#include <type_traits>

namespace cwp {

namespace {

constexpr uint64_t kFormat = 1;
const char kMagic[8] = { 'C', 'W', 'P', 'B', 'V', 'H', '1', 0 };

static_assert(std::is_trivially_copyable<BinNode>::value, "BinNode is stored as raw bytes");

struct Fnv {
    uint64_t h = 0xcbf29ce484222325ull;
    void bytes(const void* p, size_t n) {
        const unsigned char* b = (const unsigned char*)p;
        for (size_t i = 0; i < n; i++) { h ^= b[i]; h *= 0x100000001b3ull; }
    }
    template <class T> void pod(const T& v) { bytes(&v, sizeof v); }
    template <class T> void vec(const std::vector<T>& v) {
        const uint64_t n = v.size(); pod(n);
        if (n) bytes(v.data(), n * sizeof(T));
    }
};

struct Writer {
    FILE* f; Fnv sum; bool ok = true;
    void bytes(const void* p, size_t n) { if (n && fwrite(p, 1, n, f) != n) ok = false; sum.bytes(p, n); }
    template <class T> void pod(const T& v) { bytes(&v, sizeof v); }
    template <class T> void vec(const std::vector<T>& v) {
        const uint64_t n = v.size(); pod(n);
        if (n) bytes(v.data(), n * sizeof(T));
    }
};

// reads with a running checksum; a vector longer than the file is refused before it is allocated
struct Reader {
    FILE* f; uint64_t limit; Fnv sum; bool ok = true;
    void bytes(void* p, size_t n) {
        if (!ok || !n) return;
        if (fread(p, 1, n, f) != n) { ok = false; return; }
        sum.bytes(p, n);
    }
    template <class T> void pod(T& v) { bytes(&v, sizeof v); }
    template <class T> void vec(std::vector<T>& v) {
        uint64_t n = 0; pod(n);
        if (!ok || n > limit / sizeof(T)) { ok = false; return; }
        v.resize((size_t)n);
        bytes(v.data(), (size_t)n * sizeof(T));
    }
};

template <class IO, class U, class E> void fields(IO& io, U& S, E& B) {
    io.pod(S.variant); io.pod(S.root); io.pod(S.inst);
    io.pod(B.maxDepth); io.pod(B.sah); io.pod(B.tree.maxDepth); io.pod(B.tree.leafMax);
    io.vec(S.nodes); io.vec(S.woop); io.vec(S.verts); io.vec(S.normals); io.vec(S.order);
    io.vec(S.meshRoot); io.vec(S.meshBox); io.vec(S.meshTris);
    io.vec(B.tree.nodes); io.vec(B.tree.prim);
}

}  // namespace

uint64_t bvhCacheKey(const std::vector<const Mesh*>& parts, int variant, const std::string& builder, int leaf, int bins,
                     bool withNormals) {
    Fnv k;
    k.pod(kFormat); k.pod(variant); k.pod(leaf); k.pod(bins);
    const uint64_t bl = builder.size(); k.pod(bl); k.bytes(builder.data(), builder.size());
    const int wn = withNormals ? 1 : 0; k.pod(wn);
    const uint64_t np = parts.size(); k.pod(np);
    for (const Mesh* m : parts) { k.vec(m->v); k.vec(m->f); k.vec(m->fm); k.vec(m->n); }
    return k.h;
}

std::string bvhCachePath(const std::string& dir, uint64_t key) {
    char name[32];
    snprintf(name, sizeof name, "%016llx.cwpbvh", (unsigned long long)key);
    return dir.empty() || dir.back() == '/' ? dir + name : dir + "/" + name;
}

bool saveBvhCache(const std::string& path, uint64_t key, const SceneUpload& S, const BvhCacheEntry& E, std::string& err) {
    const std::string tmp = path + ".tmp";
    FILE* f = fopen(tmp.c_str(), "wb");
    if (!f) { err = "bvh cache: cannot write " + tmp; return false; }
    Writer w{ f };
    w.bytes(kMagic, 8); w.pod(key);
    fields(w, S, E);
    const uint64_t sum = w.sum.h;
    w.pod(sum);
    const bool closed = fclose(f) == 0;
    if (!w.ok || !closed || rename(tmp.c_str(), path.c_str()) != 0) {
        remove(tmp.c_str());
        err = "bvh cache: write failed: " + path;
        return false;
    }
    return true;
}

bool loadBvhCache(const std::string& path, uint64_t key, SceneUpload& S, BvhCacheEntry& E) {
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) return false;
    long size = -1;
    if (fseek(f, 0, SEEK_END) == 0) size = ftell(f);
    if (size <= 0 || fseek(f, 0, SEEK_SET) != 0) { fclose(f); return false; }
    Reader r{ f, (uint64_t)size };
    char magic[8] = {};
    uint64_t k = 0;
    r.bytes(magic, 8); r.pod(k);
    if (!r.ok || memcmp(magic, kMagic, 8) != 0 || k != key) { fclose(f); return false; }
    SceneUpload T;
    BvhCacheEntry B;
    fields(r, T, B);
    const uint64_t expect = r.sum.h;
    uint64_t sum = 0;
    if (r.ok && fread(&sum, sizeof sum, 1, f) != 1) r.ok = false;
    const bool atEnd = r.ok && fgetc(f) == EOF;
    fclose(f);
    if (!r.ok || !atEnd || sum != expect) return false;
    S.variant = T.variant; S.root = T.root; S.inst = T.inst;
    S.nodes.swap(T.nodes); S.woop.swap(T.woop); S.verts.swap(T.verts); S.normals.swap(T.normals);
    S.order.swap(T.order); S.meshRoot.swap(T.meshRoot); S.meshBox.swap(T.meshBox); S.meshTris.swap(T.meshTris);
    E = std::move(B);
    return true;
}

}  // namespace cwp
