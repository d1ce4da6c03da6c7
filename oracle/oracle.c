/* oracle.c - CPU restatement of the reference's wavefront hot path (plain C11, IEEE, -ffp-contract=off).
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for libcwp.so. Only the tests, the smoke check and bench.py's
 * CPU-baseline leg use it; the product never links it.
 *
 * What pins each part (the reference itself cannot be built here: its headers need the CUDA toolkit's
 * cuda_runtime.h for the float3/float4 types, and it has no golden vectors; see tests/test_oracle.py):
 *  - orc_hash32: Wellons' lowbias32 (hash-prospector), undone by its published inverse on 69,635 inputs.
 *  - orc_rng4: uniformity and independence statistics only (parity unpinned beyond them: no vectors exist).
 *  - orc_blackman_harris_table: monotone, antisymmetric, endpoints +-width/2, and the tabulated CDF matches the
 *    closed-form integral of the Blackman-Harris window.
 *  - warps: unit length, pdf moments (cosine hemisphere: E[cos] = 2/3; uniform cone: E[cos] = (1 + cos_max)/2).
 *  - orc_build_sah: every triangle exactly once, node boxes equal the union of their triangles' boxes, leaf sizes,
 *    depth cap, and a hand-computable split on two separated clusters; bin count 16 vs 32 both checked.
 *  - orc_collapse_cwbvh / orc_layout_bvh2: dequantised child boxes contain the exact child boxes, every triangle
 *    referenced once, <= 3 triangles per leaf, and traversal equals brute force.
 *  - orc_trace: closest hits equal orc_trace_brute (every triangle tested) on random and scene rays; any-hit agrees
 *    with closest-hit occlusion; the watertight test leaves no gaps on shared edges (rays aimed at shared vertices).
 *  - orc_render: closed forms (infinite Lambert plane under sky: a*L; plus sun: a*(L + E cos / pi) on average),
 *    sky-only miss image = L, progressive accumulation = single call.
 * Functions with no check beyond these are marked "parity unpinned" in the test file.
 */
#include "oracle.h"
#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef struct { float x, y, z; } f3;
static inline f3 mk(float x, float y, float z) { f3 r = { x, y, z }; return r; }
static inline f3 add(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline f3 sub(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline f3 mul(f3 a, f3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline f3 scl(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline f3 neg(f3 a) { return mk(-a.x, -a.y, -a.z); }
static inline float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline f3 crs(f3 a, f3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static inline f3 nrm3(f3 a) { float l = sqrtf(dot3(a, a)); return l > 0.f ? scl(a, 1.0f / l) : a; }
static inline f3 ld3(const float* p) { return mk(p[0], p[1], p[2]); }
static inline float cmp3(f3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
static inline int f2i(float f) { int i; memcpy(&i, &f, 4); return i; }
static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float i2f(int i) { float f; memcpy(&f, &i, 4); return f; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* ------------------------------------------------------------------------------------------ RNG: src/core/rng.h:6-12 */
uint32_t orc_hash32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU; x ^= x >> 16;
    return x;
}
float orc_rng4(uint32_t pix, uint32_t smp, uint32_t depth, uint32_t dim, uint32_t seed) {
    uint32_t a = orc_hash32((depth * 8u + dim) * 0xC2B2AE3DU ^ seed);
    uint32_t b = orc_hash32(smp * 0x85EBCA77U ^ a);
    uint32_t h = orc_hash32(pix * 0x9E3779B1U ^ b);
    return (float)(h >> 8) * (1.0f / 16777216.0f);
}

/* -------------------------------------------------------------------------------- warps: src/core/sampling.h:13-58 */
static void onb(f3 n, f3* t, f3* b) {
    float s = n.z >= 0.f ? 1.f : -1.f;
    float a = -1.f / (s + n.z);
    float c = n.x * n.y * a;
    *t = mk(1.f + s * n.x * n.x * a, s * c, -s * n.x);
    *b = mk(c, s + n.y * n.y * a, -n.y);
}
static f3 cos_hemi(f3 n, float u0, float u1) {
    float r = sqrtf(u0), ph = 6.2831853f * u1;
    f3 t, b; onb(n, &t, &b);
    f3 x = scl(t, r * cosf(ph)), y = scl(b, r * sinf(ph)), z = scl(n, sqrtf(fmaxf(0.f, 1.f - u0)));
    return nrm3(add(add(x, y), z));
}
static f3 cone(f3 axis, float cmax, float u0, float u1) {
    float cz = 1.f - u0 * (1.f - cmax);
    float sz = sqrtf(fmaxf(0.f, 1.f - cz * cz));
    float ph = 6.2831853f * u1;
    f3 t, b; onb(axis, &t, &b);
    f3 x = scl(axis, cz), y = scl(t, sz * cosf(ph)), z = scl(b, sz * sinf(ph));
    return nrm3(add(add(x, y), z));
}
void orc_cosine_hemisphere(const float n[3], float u0, float u1, float out[3]) {
    f3 r = cos_hemi(ld3(n), u0, u1); out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
void orc_uniform_cone(const float axis[3], float cmax, float u0, float u1, float out[3]) {
    f3 r = cone(ld3(axis), cmax, u0, u1); out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
static float table_lookup(const float* tab, int n, float u) {
    float f = u * (float)n;
    int k = (int)f; if (k > n - 1) k = n - 1;
    float t = f - (float)k;
    return tab[k] * (1.f - t) + tab[k + 1] * t;
}
static void concentric(float u0, float u1, float* x, float* y) {
    float a = 2.f * u0 - 1.f, b = 2.f * u1 - 1.f;
    if (a == 0.f && b == 0.f) { *x = 0.f; *y = 0.f; return; }
    float r, ph;
    if (fabsf(a) > fabsf(b)) { r = a; ph = 0.78539816f * (b / a); } else { r = b; ph = 1.57079633f - 0.78539816f * (a / b); }
    *x = r * cosf(ph); *y = r * sinf(ph);
}
void orc_blackman_harris_table(float* tab, int n, double width) {
    enum { CELLS = 4096 };
    double* c = (double*)malloc(sizeof(double) * (CELLS + 1));
    c[0] = 0.0;
    for (int i = 0; i < CELLS; i++) {
        double t = (i + 0.5) / CELLS;
        double f = 0.35875 - 0.48829 * cos(2 * M_PI * t) + 0.14128 * cos(4 * M_PI * t) - 0.01168 * cos(6 * M_PI * t);
        c[i + 1] = c[i] + fmax(f, 0.0);
    }
    for (int i = 0; i <= CELLS; i++) c[i] /= c[CELLS];
    for (int k = 0; k <= n; k++) {
        double u = (double)k / n;
        int j = 0;
        while (j < CELLS - 1 && c[j + 1] < u) j++;
        double t = c[j + 1] > c[j] ? (u - c[j]) / (c[j + 1] - c[j]) : 0.0;
        tab[k] = (float)(((j + t) / CELLS - 0.5) * width);
    }
    free(c);
}

/* ------------------------------------------------------------------------ boxes (src/accel/aabb.h:6-13) */
typedef struct { f3 lo, hi; } box3;
static box3 box_empty(void) { box3 b = { { FLT_MAX, FLT_MAX, FLT_MAX }, { -FLT_MAX, -FLT_MAX, -FLT_MAX } }; return b; }
static void grow_p(box3* b, f3 p) {
    b->lo = mk(fminf(b->lo.x, p.x), fminf(b->lo.y, p.y), fminf(b->lo.z, p.z));
    b->hi = mk(fmaxf(b->hi.x, p.x), fmaxf(b->hi.y, p.y), fmaxf(b->hi.z, p.z));
}
static void grow_b(box3* b, box3 o) {
    b->lo = mk(fminf(b->lo.x, o.lo.x), fminf(b->lo.y, o.lo.y), fminf(b->lo.z, o.lo.z));
    b->hi = mk(fmaxf(b->hi.x, o.hi.x), fmaxf(b->hi.y, o.hi.y), fmaxf(b->hi.z, o.hi.z));
}
static float box_area(box3 b) {
    if (b.hi.x < b.lo.x) return 0.f;
    f3 d = sub(b.hi, b.lo);
    return 2.f * (d.x * d.y + d.y * d.z + d.z * d.x);
}

/* growable int / float arrays */
typedef struct { int* p; int n, cap; } ivec;
typedef struct { float* p; int n, cap; } fvec;
static void ipush(ivec* v, int x) { if (v->n == v->cap) { v->cap = v->cap ? 2 * v->cap : 64; v->p = (int*)realloc(v->p, sizeof(int) * v->cap); } v->p[v->n++] = x; }
static void fgrow(fvec* v, int n) { if (n > v->cap) { int c = v->cap ? v->cap : 64; while (c < n) c *= 2; v->p = (float*)realloc(v->p, sizeof(float) * c); v->cap = c; } if (n > v->n) memset(v->p + v->n, 0, sizeof(float) * (n - v->n)); v->n = n; }

/* ------------------------------------------------------------ binned SAH BVH2: src/accel/bvh_builder.cpp:5-68 */
typedef struct {
    const box3* tb; const f3* cent; int* idx;
    ivec node;   /* 4 ints per node */
    fvec box;    /* 6 floats per node */
    int bins, leaf_max, max_depth;
    box3* bb; int* bc; float* rA; int* rC;
} sah_ctx;

static int bin_of(float c, float lo, float scale, int bins) {
    int b = (int)((c - lo) * scale);
    return b < 0 ? 0 : (b > bins - 1 ? bins - 1 : b);
}
static int new_node(sah_ctx* S) {
    int id = S->node.n / 4;
    for (int k = 0; k < 4; k++) ipush(&S->node, k < 2 ? -1 : 0);
    fgrow(&S->box, S->box.n + 6);
    return id;
}
static int sah_rec(sah_ctx* S, int begin, int end, int depth) {
    int id = new_node(S);
    if (depth > S->max_depth) S->max_depth = depth;
    box3 box = box_empty(), cb = box_empty();
    int n = end - begin;
    for (int i = begin; i < end; i++) { grow_b(&box, S->tb[S->idx[i]]); grow_p(&cb, S->cent[S->idx[i]]); }
    float* bx = S->box.p + 6 * id;
    bx[0] = box.lo.x; bx[1] = box.lo.y; bx[2] = box.lo.z; bx[3] = box.hi.x; bx[4] = box.hi.y; bx[5] = box.hi.z;
    if (n <= S->leaf_max || depth >= 62) { S->node.p[4 * id + 2] = begin; S->node.p[4 * id + 3] = n; return id; }
    float best = FLT_MAX; int bestAxis = -1, bestBin = -1;
    f3 ext = sub(cb.hi, cb.lo);
    const int B = S->bins;
    for (int axis = 0; axis < 3; axis++) {
        float lo = cmp3(cb.lo, axis), e = cmp3(ext, axis);
        if (e <= 1e-12f) continue;
        float scale = (float)B / e;
        for (int b = 0; b < B; b++) { S->bb[b] = box_empty(); S->bc[b] = 0; }
        for (int i = begin; i < end; i++) {
            int b = bin_of(cmp3(S->cent[S->idx[i]], axis), lo, scale, B);
            grow_b(&S->bb[b], S->tb[S->idx[i]]); S->bc[b]++;
        }
        box3 acc = box_empty(); int cnt = 0;
        for (int b = B - 1; b > 0; b--) { grow_b(&acc, S->bb[b]); cnt += S->bc[b]; S->rA[b] = box_area(acc); S->rC[b] = cnt; }
        box3 lacc = box_empty(); int lcnt = 0;
        for (int b = 0; b < B - 1; b++) {
            grow_b(&lacc, S->bb[b]); lcnt += S->bc[b];
            if (lcnt == 0 || S->rC[b + 1] == 0) continue;
            float cost = box_area(lacc) * (float)lcnt + S->rA[b + 1] * (float)S->rC[b + 1];
            if (cost < best) { best = cost; bestAxis = axis; bestBin = b; }
        }
    }
    int mid;
    if (bestAxis < 0) mid = (begin + end) / 2;
    else {
        float lo = cmp3(cb.lo, bestAxis), e = cmp3(ext, bestAxis), scale = (float)B / e;
        /* two-sided in-place partition (left scan for a failing element, right scan for a passing one, swap) */
        int* first = S->idx + begin; int* last = S->idx + end;
        for (;;) {
            while (first != last && bin_of(cmp3(S->cent[*first], bestAxis), lo, scale, B) <= bestBin) first++;
            if (first == last) break;
            last--;
            while (first != last && !(bin_of(cmp3(S->cent[*last], bestAxis), lo, scale, B) <= bestBin)) last--;
            if (first == last) break;
            int t = *first; *first = *last; *last = t;
            first++;
        }
        mid = (int)(first - S->idx);
        if (mid == begin || mid == end) mid = (begin + end) / 2;
    }
    int l = sah_rec(S, begin, mid, depth + 1);
    int r = sah_rec(S, mid, end, depth + 1);
    S->node.p[4 * id] = l; S->node.p[4 * id + 1] = r;
    return id;
}

int orc_build_sah(const float* v, int nv, const int* tri, int nt, int bins, int leaf_max, orc_tree* out) {
    if (nt < 1 || bins < 2 || leaf_max < 1) return 1;
    for (int t = 0; t < 3 * nt; t++) if (tri[t] < 0 || tri[t] >= nv) return 1;
    sah_ctx S; memset(&S, 0, sizeof S);
    box3* tb = (box3*)malloc(sizeof(box3) * nt);
    f3* cent = (f3*)malloc(sizeof(f3) * nt);
    int* idx = (int*)malloc(sizeof(int) * nt);
    for (int t = 0; t < nt; t++) {
        box3 b = box_empty();
        for (int k = 0; k < 3; k++) grow_p(&b, ld3(v + 3 * tri[3 * t + k]));
        tb[t] = b; cent[t] = scl(add(b.lo, b.hi), 0.5f); idx[t] = t;
    }
    S.tb = tb; S.cent = cent; S.idx = idx; S.bins = bins; S.leaf_max = leaf_max;
    S.bb = (box3*)malloc(sizeof(box3) * bins); S.bc = (int*)malloc(sizeof(int) * bins);
    S.rA = (float*)malloc(sizeof(float) * bins); S.rC = (int*)malloc(sizeof(int) * bins);
    sah_rec(&S, 0, nt, 0);
    out->num_nodes = S.node.n / 4; out->num_prims = nt; out->max_depth = S.max_depth;
    out->node = S.node.p; out->box = S.box.p; out->prim = idx;
    free(tb); free(cent); free(S.bb); free(S.bc); free(S.rA); free(S.rC);
    return 0;
}
void orc_tree_free(orc_tree* t) { free(t->node); free(t->box); free(t->prim); memset(t, 0, sizeof *t); }

static box3 tree_box(const orc_tree* t, int n) {
    const float* b = t->box + 6 * n; box3 r = { { b[0], b[1], b[2] }, { b[3], b[4], b[5] } }; return r;
}

/* ------------------------------------------------------------------- BVH2 layout: src/accel/bvh_layouts.cpp:6-30 */
typedef struct { const int* node; const float* box; int wrapped_root; } b2src;
static int b2_left(const orc_tree* t, int n) { return n == t->num_nodes ? 0 : t->node[4 * n]; }
static int b2_right(const orc_tree* t, int n) { return n == t->num_nodes ? -1 : t->node[4 * n + 1]; }
static int b2_leaf(const orc_tree* t, int n) { return n < t->num_nodes && t->node[4 * n + 3] > 0; }
static int b2_emit(const orc_tree* t, int n, fvec* out, int level, int* depth) {
    int id = out->n / 16;
    fgrow(out, out->n + 16);
    if (level > *depth) *depth = level;
    int ch[2] = { b2_left(t, n), b2_right(t, n) }, c[2], k[2];
    box3 bx[2];
    for (int j = 0; j < 2; j++) {
        if (ch[j] < 0) { c[j] = -1; k[j] = 0; bx[j].lo = bx[j].hi = mk(3e37f, 3e37f, 3e37f); continue; }
        bx[j] = tree_box(t, ch[j]);
        if (b2_leaf(t, ch[j])) { c[j] = t->node[4 * ch[j] + 2]; k[j] = t->node[4 * ch[j] + 3]; }
        else { k[j] = 0; c[j] = b2_emit(t, ch[j], out, level + 1, depth); }
    }
    float* o = out->p + 16 * id;
    o[0] = bx[0].lo.x; o[1] = bx[0].hi.x; o[2] = bx[0].lo.y; o[3] = bx[0].hi.y;
    o[4] = bx[1].lo.x; o[5] = bx[1].hi.x; o[6] = bx[1].lo.y; o[7] = bx[1].hi.y;
    o[8] = bx[0].lo.z; o[9] = bx[0].hi.z; o[10] = bx[1].lo.z; o[11] = bx[1].hi.z;
    o[12] = i2f(c[0]); o[13] = i2f(c[1]); o[14] = i2f(k[0]); o[15] = i2f(k[1]);
    return id;
}
int orc_layout_bvh2(const orc_tree* t, orc_layout* out) {
    fvec v = { 0, 0, 0 };
    int depth = 0;
    /* a leaf root is wrapped in an internal node whose second child is empty (index num_nodes = the wrapper) */
    int root = t->node[3] > 0 ? t->num_nodes : 0;
    b2_emit(t, root, &v, 1, &depth);
    out->num_float4 = v.n / 4; out->nodes = v.p; out->num_prims = t->num_prims;
    out->order = (int*)malloc(sizeof(int) * t->num_prims); memcpy(out->order, t->prim, sizeof(int) * t->num_prims);
    out->depth = depth; out->max_leaf = 0; out->sah = 0;
    double a0 = box_area(tree_box(t, 0)), s = 0;
    if (a0 > 0) {
        for (int n = 0; n < t->num_nodes; n++) {
            double a = box_area(tree_box(t, n));
            s += t->node[4 * n + 3] > 0 ? a * t->node[4 * n + 3] : a;
        }
        out->sah = s / a0;
    }
    return 0;
}
void orc_layout_free(orc_layout* l) { free(l->nodes); free(l->order); memset(l, 0, sizeof *l); }

/* ------------------------------------------------------------------- BVH4 layout: src/accel/bvh_layouts.cpp:32-63 */
/* children: start from the two children, then repeatedly open the internal child of largest surface area (first of
 * equal areas) until there are 4 or only leaves remain; pre-order node ids */
static int b4_emit(const orc_tree* t, int n, fvec* out, int level, int* depth) {
    int id = out->n / 32;
    fgrow(out, out->n + 32);
    if (level > *depth) *depth = level;
    int ch[4], nc = 2;
    ch[0] = b2_left(t, n); ch[1] = b2_right(t, n);
    while (nc < 4) {
        int best = -1; float bestA = -1.f;
        for (int j = 0; j < nc; j++)
            if (ch[j] >= 0 && !b2_leaf(t, ch[j])) {
                float a = box_area(tree_box(t, ch[j]));
                if (a > bestA) { bestA = a; best = j; }
            }
        if (best < 0) break;
        int nb = ch[best];
        ch[best] = t->node[4 * nb]; ch[nc++] = t->node[4 * nb + 1];
    }
    int c[4], k[4];
    box3 bx[4];
    for (int j = 0; j < 4; j++) {
        if (j >= nc || ch[j] < 0) { c[j] = -1; k[j] = 0; bx[j].lo = bx[j].hi = mk(3e37f, 3e37f, 3e37f); continue; }
        bx[j] = tree_box(t, ch[j]);
        if (b2_leaf(t, ch[j])) { c[j] = t->node[4 * ch[j] + 2]; k[j] = t->node[4 * ch[j] + 3]; }
        else { k[j] = 0; c[j] = b4_emit(t, ch[j], out, level + 1, depth); }
    }
    float* o = out->p + 32 * id;
    for (int j = 0; j < 4; j++) {
        o[j] = bx[j].lo.x; o[4 + j] = bx[j].hi.x; o[8 + j] = bx[j].lo.y; o[12 + j] = bx[j].hi.y;
        o[16 + j] = bx[j].lo.z; o[20 + j] = bx[j].hi.z; o[24 + j] = i2f(c[j]); o[28 + j] = i2f(k[j]);
    }
    return id;
}
int orc_layout_bvh4(const orc_tree* t, orc_layout* out) {
    fvec v = { 0, 0, 0 };
    int depth = 0;
    int root = t->node[3] > 0 ? t->num_nodes : 0;
    b4_emit(t, root, &v, 1, &depth);
    out->num_float4 = v.n / 4; out->nodes = v.p; out->num_prims = t->num_prims;
    out->order = (int*)malloc(sizeof(int) * t->num_prims); memcpy(out->order, t->prim, sizeof(int) * t->num_prims);
    out->depth = depth; out->max_leaf = 0; out->sah = 0;
    return 0;
}

/* ------------------------------------------------------- CWBVH collapse: src/accel/bvh_tools.cpp:75-168 (Ylitie 2017) */
enum { D_LEAF = 0, D_NODE = 1, D_PREV = 2, D_DIST = 10 };
typedef struct {
    box3* box; int *L, *R, *first, *cnt; int n, cap;
    float* C; unsigned char* D; unsigned char* K8;
    const int* prim;
} cw_t;
static void cw_push(cw_t* w, box3 b, int l, int r, int f, int c) {
    if (w->n == w->cap) {
        w->cap = w->cap ? 2 * w->cap : 64;
        w->box = (box3*)realloc(w->box, sizeof(box3) * w->cap);
        w->L = (int*)realloc(w->L, sizeof(int) * w->cap); w->R = (int*)realloc(w->R, sizeof(int) * w->cap);
        w->first = (int*)realloc(w->first, sizeof(int) * w->cap); w->cnt = (int*)realloc(w->cnt, sizeof(int) * w->cap);
    }
    w->box[w->n] = b; w->L[w->n] = l; w->R[w->n] = r; w->first[w->n] = f; w->cnt[w->n] = c; w->n++;
}
static void cw_collect(const cw_t* w, int n, int i, ivec* out) {
    for (;;) {
        if (i == 1) { ipush(out, n); return; }
        unsigned char d = w->D[(size_t)n * 8 + i];
        if (d == D_PREV) { i--; continue; }
        int k = d - D_DIST;
        cw_collect(w, w->L[n], k, out);
        cw_collect(w, w->R[n], i - k, out);
        return;
    }
}
static void cw_prims(const cw_t* w, int n, ivec* out) {
    ivec st = { 0, 0, 0 }; ipush(&st, n);
    while (st.n) {
        int h = st.p[--st.n];
        if (w->L[h] < 0) { for (int t = 0; t < w->cnt[h]; t++) ipush(out, w->prim[w->first[h] + t]); }
        else { ipush(&st, w->R[h]); ipush(&st, w->L[h]); }
    }
    free(st.p);
}
static int qexp(float ext, float lo, float hi) {
    if (!(ext > 0.f)) return -100;
    int e = (int)ceilf(log2f(ext / 255.f));
    if (e < -100) e = -100;
    while (lo + 255.f * ldexpf(1.f, e) < hi) e++;
    return e;
}
static int quant(float v, float o, float sc, int up) {
    float f = (v - o) / sc;
    int i = up ? (int)ceilf(f) : (int)floorf(f);
    return i < 0 ? 0 : (i > 255 ? 255 : i);
}
static uint32_t pk4(const int* q, int b) { return (uint32_t)q[b] | ((uint32_t)q[b + 1] << 8) | ((uint32_t)q[b + 2] << 16) | ((uint32_t)q[b + 3] << 24); }

int orc_collapse_cwbvh(const orc_tree* t, orc_layout* out) {
    const int pMax = 3; const float cNode = 1.f, cPrim = 0.3f, INF = 3e38f;
    cw_t w; memset(&w, 0, sizeof w); w.prim = t->prim;
    for (int i = 0; i < t->num_nodes; i++) {
        const int* nd = t->node + 4 * i; int leaf = nd[3] > 0;
        cw_push(&w, tree_box(t, i), leaf ? -1 : nd[0], leaf ? -1 : nd[1], nd[2], leaf ? nd[3] : 0);
    }
    for (int i = 0; i < w.n; i++) {   /* oversize leaves -> binary subtrees with the same box */
        if (w.L[i] >= 0 || w.cnt[i] <= pMax) continue;
        int h = w.cnt[i] / 2, a = w.n, b = a + 1, f0 = w.first[i], c0 = w.cnt[i]; box3 bx = w.box[i];
        cw_push(&w, bx, -1, -1, f0, h); cw_push(&w, bx, -1, -1, f0 + h, c0 - h);
        w.L[i] = a; w.R[i] = b; w.cnt[i] = 0;
    }
    const int N = w.n;
    w.C = (float*)calloc((size_t)N * 8, sizeof(float)); w.D = (unsigned char*)calloc((size_t)N * 8, 1);
    w.K8 = (unsigned char*)malloc(N); memset(w.K8, 1, N);
    int* p = (int*)calloc(N, sizeof(int));
    for (int n = N - 1; n >= 0; n--) {
        float A = box_area(w.box[n]);
        float* c = w.C + (size_t)n * 8; unsigned char* d = w.D + (size_t)n * 8;
        if (w.L[n] < 0) {
            p[n] = w.cnt[n];
            float cl = A * (float)p[n] * cPrim;
            for (int i = 1; i < 8; i++) { c[i] = cl; d[i] = i == 1 ? D_LEAF : D_PREV; }
            continue;
        }
        int l = w.L[n], r = w.R[n];
        p[n] = p[l] + p[r];
        const float *cl_ = w.C + (size_t)l * 8, *cr_ = w.C + (size_t)r * 8;
        float dist[9]; int dk[9];
        for (int j = 2; j <= 8; j++) {
            dist[j] = INF; dk[j] = 1;
            for (int k = 1; k < j; k++) { float x = cl_[k] + cr_[j - k]; if (x < dist[j]) { dist[j] = x; dk[j] = k; } }
        }
        float cLeaf = p[n] <= pMax ? A * (float)p[n] * cPrim : INF;
        float cInt = dist[8] + A * cNode;
        if (cLeaf <= cInt) { c[1] = cLeaf; d[1] = D_LEAF; } else { c[1] = cInt; d[1] = D_NODE; }
        w.K8[n] = (unsigned char)dk[8];
        for (int i = 2; i < 8; i++) {
            if (dist[i] < c[i - 1]) { c[i] = dist[i]; d[i] = (unsigned char)(D_DIST + dk[i]); }
            else { c[i] = c[i - 1]; d[i] = D_PREV; }
        }
    }
    const float Ar = box_area(w.box[0]);
    int rootLeaf = w.L[0] < 0 || w.D[1] == D_LEAF;
    if (rootLeaf) out->sah = Ar > 0 ? (Ar * (float)p[0] * cPrim + Ar * cNode) / Ar : 0.f;
    else {
        float d8 = INF;
        for (int k = 1; k < 8; k++) { float x = w.C[(size_t)w.L[0] * 8 + k] + w.C[(size_t)w.R[0] * 8 + 8 - k]; d8 = x < d8 ? x : d8; }
        out->sah = Ar > 0 ? (d8 + Ar * cNode) / Ar : 0.f;
    }
    fvec nodes = { 0, 0, 0 }; ivec order = { 0, 0, 0 };
    fgrow(&nodes, 20);
    ivec jb = { 0, 0, 0 };   /* job stack: (binary node, output node, level, asLeaf) */
    ipush(&jb, 0); ipush(&jb, 0); ipush(&jb, 1); ipush(&jb, rootLeaf);
    ivec ch = { 0, 0, 0 }, pr = { 0, 0, 0 };
    int depth = 0, maxLeaf = 0;
    while (jb.n) {
        int asLeaf = jb.p[jb.n - 1], level = jb.p[jb.n - 2], outId = jb.p[jb.n - 3], n = jb.p[jb.n - 4];
        jb.n -= 4;
        if (level > depth) depth = level;
        ch.n = 0;
        if (asLeaf) ipush(&ch, n); else { cw_collect(&w, w.L[n], w.K8[n], &ch); cw_collect(&w, w.R[n], 8 - w.K8[n], &ch); }
        box3 box = w.box[n];
        f3 pc = scl(add(box.lo, box.hi), 0.5f);
        int slot[8]; float cost[8][8]; int used[8] = { 0 };
        for (int s = 0; s < 8; s++) slot[s] = -1;
        for (int c = 0; c < ch.n; c++) {
            f3 cc = sub(scl(add(w.box[ch.p[c]].lo, w.box[ch.p[c]].hi), 0.5f), pc);
            for (int s = 0; s < 8; s++)
                cost[c][s] = -(cc.x * (s & 1 ? -1.f : 1.f) + cc.y * (s & 2 ? -1.f : 1.f) + cc.z * (s & 4 ? -1.f : 1.f));
        }
        for (int it = 0; it < ch.n; it++) {
            float bc = INF; int bi = -1, bs = -1;
            for (int c = 0; c < ch.n; c++)
                if (!used[c]) for (int s = 0; s < 8; s++) if (slot[s] < 0 && cost[c][s] < bc) { bc = cost[c][s]; bi = c; bs = s; }
            for (int c = 0; c < ch.n && bi < 0; c++)
                if (!used[c]) for (int s = 0; s < 8; s++) if (slot[s] < 0) { bi = c; bs = s; break; }
            used[bi] = 1; slot[bs] = ch.p[bi];
        }
        f3 ext = sub(box.hi, box.lo);
        int ex = qexp(ext.x, box.lo.x, box.hi.x), ey = qexp(ext.y, box.lo.y, box.hi.y), ez = qexp(ext.z, box.lo.z, box.hi.z);
        float sx = ldexpf(1.f, ex), sy = ldexpf(1.f, ey), sz = ldexpf(1.f, ez);
        int qlx[8], qly[8], qlz[8], qhx[8], qhy[8], qhz[8], meta[8], childLeaf[8];
        int triBase = order.n, nInt = 0; uint32_t imask = 0;
        for (int s = 0; s < 8; s++) {
            qlx[s] = qly[s] = qlz[s] = qhx[s] = qhy[s] = qhz[s] = meta[s] = childLeaf[s] = 0;
            int cn = slot[s];
            if (cn < 0) continue;
            box3 cb = w.box[cn];
            qlx[s] = quant(cb.lo.x, box.lo.x, sx, 0); qhx[s] = quant(cb.hi.x, box.lo.x, sx, 1);
            qly[s] = quant(cb.lo.y, box.lo.y, sy, 0); qhy[s] = quant(cb.hi.y, box.lo.y, sy, 1);
            qlz[s] = quant(cb.lo.z, box.lo.z, sz, 0); qhz[s] = quant(cb.hi.z, box.lo.z, sz, 1);
            if (asLeaf || w.L[cn] < 0 || w.D[(size_t)cn * 8 + 1] == D_LEAF) {
                pr.n = 0; cw_prims(&w, cn, &pr);
                int off = order.n - triBase;
                for (int k = 0; k < pr.n; k++) ipush(&order, pr.p[k]);
                meta[s] = 0x80 | ((pr.n - 1) << 5) | off;
                if (pr.n > maxLeaf) maxLeaf = pr.n;
                childLeaf[s] = 1;
            } else { meta[s] = 0x40; imask |= 1u << s; nInt++; }
        }
        int childBase = nodes.n / 20;
        fgrow(&nodes, nodes.n + 20 * nInt);
        float* o = nodes.p + (size_t)outId * 20;
        uint32_t ew = (uint32_t)(ex + 127) | ((uint32_t)(ey + 127) << 8) | ((uint32_t)(ez + 127) << 16) | (imask << 24);
        o[0] = box.lo.x; o[1] = box.lo.y; o[2] = box.lo.z; o[3] = u2f(ew);
        o[4] = i2f(childBase); o[5] = i2f(triBase); o[6] = u2f(pk4(meta, 0)); o[7] = u2f(pk4(meta, 4));
        uint32_t u[12] = { pk4(qlx, 0), pk4(qlx, 4), pk4(qly, 0), pk4(qly, 4), pk4(qlz, 0), pk4(qlz, 4),
                           pk4(qhx, 0), pk4(qhx, 4), pk4(qhy, 0), pk4(qhy, 4), pk4(qhz, 0), pk4(qhz, 4) };
        for (int j = 0; j < 12; j++) o[8 + j] = u2f(u[j]);
        int ord = 0;
        for (int s = 0; s < 8; s++)
            if (slot[s] >= 0 && !childLeaf[s]) { ipush(&jb, slot[s]); ipush(&jb, childBase + ord); ipush(&jb, level + 1); ipush(&jb, 0); ord++; }
    }
    out->num_float4 = nodes.n / 4; out->nodes = nodes.p; out->num_prims = order.n; out->order = order.p;
    out->depth = depth; out->max_leaf = maxLeaf;
    free(jb.p); free(ch.p); free(pr.p); free(p);
    free(w.box); free(w.L); free(w.R); free(w.first); free(w.cnt); free(w.C); free(w.D); free(w.K8);
    return 0;
}

/* ------------------------------------------------------------------------ triangle packing: src/accel/tri_pack.cpp */
void orc_pack_woop(const float* v, const int* tri, const int* order, int n, float* rows) {
    for (int k = 0; k < n; k++) {
        const int* f = tri + 3 * order[k];
        f3 v0 = ld3(v + 3 * f[0]), e1 = sub(ld3(v + 3 * f[1]), v0), e2 = sub(ld3(v + 3 * f[2]), v0), nn = crs(e1, e2);
        double m[3][3] = { { e1.x, e2.x, nn.x }, { e1.y, e2.y, nn.y }, { e1.z, e2.z, nn.z } }, iv[3][3];
        double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                     m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
        memset(iv, 0, sizeof iv);
        if (!(fabs(det) < 1e-300)) {
            double id = 1.0 / det;
            iv[0][0] = (m[1][1] * m[2][2] - m[1][2] * m[2][1]) * id; iv[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) * id;
            iv[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * id; iv[1][0] = (m[1][2] * m[2][0] - m[1][0] * m[2][2]) * id;
            iv[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * id; iv[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) * id;
            iv[2][0] = (m[1][0] * m[2][1] - m[1][1] * m[2][0]) * id; iv[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) * id;
            iv[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * id;
        }
        for (int r = 0; r < 3; r++) {
            double tr = -(iv[r][0] * v0.x + iv[r][1] * v0.y + iv[r][2] * v0.z);
            float* o = rows + 12 * k + 4 * r;
            o[0] = (float)iv[r][0]; o[1] = (float)iv[r][1]; o[2] = (float)iv[r][2]; o[3] = (float)tr;
        }
    }
}
void orc_pack_corners(const float* v, const int* tri, const int* mat, const int* order, int n, float* c) {
    for (int k = 0; k < n; k++) {
        int t = order[k]; const int* f = tri + 3 * t;
        for (int j = 0; j < 3; j++) { memcpy(c + 12 * k + 4 * j, v + 3 * f[j], 12); }
        c[12 * k + 3] = i2f(t); c[12 * k + 7] = i2f(mat ? mat[t] : 0); c[12 * k + 11] = 0.f;
    }
}

/* ---------------------------------------------------------------------- triangle tests: src/accel/traverse.cuh:10-42 */
typedef struct {
    f3 o, d; float tmin; float ix, iy, iz; uint32_t oct; int kx, ky, kz; float Sx, Sy, Sz;
} ray_t;
typedef struct { float t, u, v; int tri, inst; } hit_t;
static float safe_rcp(float a) { const float e = 1e-20f; return 1.f / (fabsf(a) > e ? a : copysignf(e, a)); }
static ray_t mkray(f3 o, f3 d, float tmin, int tri_test) {
    ray_t r; memset(&r, 0, sizeof r);
    r.o = o; r.d = d; r.tmin = tmin;
    r.ix = safe_rcp(d.x); r.iy = safe_rcp(d.y); r.iz = safe_rcp(d.z);
    r.oct = (d.x > 0.f ? 1u : 0u) | (d.y > 0.f ? 2u : 0u) | (d.z > 0.f ? 4u : 0u);
    if (tri_test == 2) {   /* watertight: dominant axis kz, kx/ky keep winding; shear constants */
        float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
        int kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
        int kx = kz == 2 ? 0 : kz + 1, ky = kx == 2 ? 0 : kx + 1;
        if (cmp3(d, kz) < 0.f) { int t = kx; kx = ky; ky = t; }
        r.kx = kx; r.ky = ky; r.kz = kz;
        float dz = cmp3(d, kz);
        r.Sx = cmp3(d, kx) / dz; r.Sy = cmp3(d, ky) / dz; r.Sz = 1.f / dz;
    }
    return r;
}
static int tri_test(const orc_scene* s, int t, const ray_t* r, hit_t* h) {
    if (s->tri_test == 1) {   /* Woop: t from row 2 first, then u, v */
        const float* R = s->woop + 12 * t;
        float Oz = fmaf(r->o.z, R[10], fmaf(r->o.y, R[9], fmaf(r->o.x, R[8], R[11])));
        float Dz = fmaf(r->d.z, R[10], fmaf(r->d.y, R[9], r->d.x * R[8]));
        float tt = -Oz / Dz;
        if (!(tt > r->tmin && tt < h->t)) return 0;
        float Ox = fmaf(r->o.z, R[2], fmaf(r->o.y, R[1], fmaf(r->o.x, R[0], R[3])));
        float Dx = fmaf(r->d.z, R[2], fmaf(r->d.y, R[1], r->d.x * R[0]));
        float u = fmaf(tt, Dx, Ox);
        if (!(u >= 0.f && u <= 1.f)) return 0;
        float Oy = fmaf(r->o.z, R[6], fmaf(r->o.y, R[5], fmaf(r->o.x, R[4], R[7])));
        float Dy = fmaf(r->d.z, R[6], fmaf(r->d.y, R[5], r->d.x * R[4]));
        float v = fmaf(tt, Dy, Oy);
        if (!(v >= 0.f && u + v <= 1.f)) return 0;
        h->t = tt; h->u = u; h->v = v; h->tri = t;
        return 1;
    }
    const float* P = s->corners + 12 * t;
    f3 p0 = ld3(P), p1 = ld3(P + 4), p2 = ld3(P + 8);
    if (s->tri_test == 0) {   /* Moller-Trumbore on (v0, e1 = v1 - v0, e2 = v2 - v0) */
        f3 e1 = sub(p1, p0), e2 = sub(p2, p0);
        f3 pv = crs(r->d, e2);
        float det = dot3(e1, pv);
        if (fabsf(det) < 1e-12f) return 0;
        float inv = 1.f / det;
        f3 tv = sub(r->o, p0);
        float u = dot3(tv, pv) * inv;
        if (u < 0.f || u > 1.f) return 0;
        f3 qv = crs(tv, e1);
        float v = dot3(r->d, qv) * inv;
        if (v < 0.f || u + v > 1.f) return 0;
        float tt = dot3(e2, qv) * inv;
        if (!(tt > r->tmin && tt < h->t)) return 0;
        h->t = tt; h->u = u; h->v = v; h->tri = t;
        return 1;
    }
    /* watertight (Woop, Benthin, Wald 2013) */
    f3 A = sub(p0, r->o), B = sub(p1, r->o), C = sub(p2, r->o);
    float Akz = cmp3(A, r->kz), Bkz = cmp3(B, r->kz), Ckz = cmp3(C, r->kz);
    float Ax = cmp3(A, r->kx) - r->Sx * Akz, Ay = cmp3(A, r->ky) - r->Sy * Akz;
    float Bx = cmp3(B, r->kx) - r->Sx * Bkz, By = cmp3(B, r->ky) - r->Sy * Bkz;
    float Cx = cmp3(C, r->kx) - r->Sx * Ckz, Cy = cmp3(C, r->ky) - r->Sy * Ckz;
    float U = Cx * By - Cy * Bx, V = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
    if (U == 0.f || V == 0.f || W == 0.f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.f || V < 0.f || W < 0.f) && (U > 0.f || V > 0.f || W > 0.f)) return 0;
    float det = U + V + W;
    if (det == 0.f) return 0;
    float Az = r->Sz * Akz, Bz = r->Sz * Bkz, Cz = r->Sz * Ckz;
    float T = U * Az + V * Bz + W * Cz;
    float rcp = 1.f / det, tt = T * rcp;
    if (!(tt > r->tmin && tt < h->t)) return 0;
    h->t = tt; h->u = V * rcp; h->v = W * rcp; h->tri = t;
    return 1;
}

/* ------------------------------------------------------------------ BVH2 traversal: src/accel/traverse.cuh:45-70 */
static int trav2(const orc_scene* s, int root, const ray_t* r, hit_t* h, int any, double* nv, double* nt) {
    int stack[128], sp = 0, node = root;
    const float odx = r->o.x * r->ix, ody = r->o.y * r->iy, odz = r->o.z * r->iz;
    for (;;) {
        const float* n = s->nodes + 16 * (size_t)node;
        if (nv) *nv += 1;
        int c0 = f2i(n[12]), c1 = f2i(n[13]), k0 = f2i(n[14]), k1 = f2i(n[15]);
        float a0 = fmaf(n[0], r->ix, -odx), a1 = fmaf(n[1], r->ix, -odx), a2 = fmaf(n[2], r->iy, -ody), a3 = fmaf(n[3], r->iy, -ody);
        float a4 = fmaf(n[8], r->iz, -odz), a5 = fmaf(n[9], r->iz, -odz);
        float b0 = fmaf(n[4], r->ix, -odx), b1 = fmaf(n[5], r->ix, -odx), b2 = fmaf(n[6], r->iy, -ody), b3 = fmaf(n[7], r->iy, -ody);
        float b4 = fmaf(n[10], r->iz, -odz), b5 = fmaf(n[11], r->iz, -odz);
        float t0n = fmaxf(fmaxf(fminf(a0, a1), fminf(a2, a3)), fmaxf(fminf(a4, a5), r->tmin));
        float t0f = fminf(fminf(fmaxf(a0, a1), fmaxf(a2, a3)), fminf(fmaxf(a4, a5), h->t));
        float t1n = fmaxf(fmaxf(fminf(b0, b1), fminf(b2, b3)), fmaxf(fminf(b4, b5), r->tmin));
        float t1f = fminf(fminf(fmaxf(b0, b1), fmaxf(b2, b3)), fminf(fmaxf(b4, b5), h->t));
        int h0 = t0n <= t0f, h1 = t1n <= t1f;
        if (h0 && k0 > 0) {
            for (int t = c0; t < c0 + k0; t++) { if (nt) *nt += 1; if (tri_test(s, t, r, h) && any) return 1; }
            h0 = 0;
        }
        if (h1 && k1 > 0) {
            for (int t = c1; t < c1 + k1; t++) { if (nt) *nt += 1; if (tri_test(s, t, r, h) && any) return 1; }
            h1 = 0;
        }
        if (!h0 && !h1) { if (sp == 0) break; node = stack[--sp]; }
        else if (h0 && h1) { if (t1n < t0n) { int x = c0; c0 = c1; c1 = x; } stack[sp++] = c1; node = c0; }
        else node = h0 ? c0 : c1;
    }
    return h->tri >= 0;
}

/* ---------------------------------------------------------------- BVH8 traversal: src/accel/traverse.cuh:112-152 */
static uint32_t byte_of(uint32_t w, int s) { return (w >> (8 * (s & 3))) & 255u; }
static int trav8(const orc_scene* s, int root, const ray_t* r, hit_t* h, int any, double* nv, double* nt) {
    uint32_t stB[128], stM[128]; int sp = 0;
    uint32_t gB = 0, gM = 0, gH = 0;
    int node = root;
    for (;;) {
        const float* n = s->nodes + 20 * (size_t)node;
        if (nv) *nv += 1;
        uint32_t ew = f2u(n[3]);
        float sx = u2f((ew & 255u) << 23) * r->ix, sy = u2f(((ew >> 8) & 255u) << 23) * r->iy, sz = u2f(((ew >> 16) & 255u) << 23) * r->iz;
        float ox = (n[0] - r->o.x) * r->ix, oy = (n[1] - r->o.y) * r->iy, oz = (n[2] - r->o.z) * r->iz;
        uint32_t imask = ew >> 24, meta[2] = { f2u(n[6]), f2u(n[7]) };
        int childBase = f2i(n[4]), triBase = f2i(n[5]);
        uint32_t qlx[2] = { f2u(n[8]), f2u(n[9]) }, qly[2] = { f2u(n[10]), f2u(n[11]) }, qlz[2] = { f2u(n[12]), f2u(n[13]) };
        uint32_t qhx[2] = { f2u(n[14]), f2u(n[15]) }, qhy[2] = { f2u(n[16]), f2u(n[17]) }, qhz[2] = { f2u(n[18]), f2u(n[19]) };
        /* all 8 child slabs against the hit distance at node entry, then the hit leaves in slot order (the reference
         * interleaves them; the closest hit is the same, and the order of triangle tests within a node is kept) */
        const float tEntry = h->t;
        uint32_t hitSlots = 0, leafSlots = 0;
        for (int sl = 0; sl < 8; sl++) {
            int w = sl >> 2;
            uint32_t m = byte_of(meta[w], sl);
            if (m == 0) continue;
            float x0 = fmaf((float)byte_of(qlx[w], sl), sx, ox), x1 = fmaf((float)byte_of(qhx[w], sl), sx, ox);
            float y0 = fmaf((float)byte_of(qly[w], sl), sy, oy), y1 = fmaf((float)byte_of(qhy[w], sl), sy, oy);
            float z0 = fmaf((float)byte_of(qlz[w], sl), sz, oz), z1 = fmaf((float)byte_of(qhz[w], sl), sz, oz);
            float tn = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), r->tmin));
            float tf = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), tEntry));
            if (tn <= tf) hitSlots |= 1u << sl;
            if (m & 0x80u) leafSlots |= 1u << sl;
        }
        uint32_t hits = 0;
        for (int sl = 0; sl < 8; sl++) {
            if (!(hitSlots >> sl & 1u)) continue;
            if (leafSlots >> sl & 1u) {
                uint32_t m = byte_of(meta[sl >> 2], sl);
                int first = triBase + (int)(m & 31u), cnt = (int)((m >> 5) & 3u) + 1;
                for (int t = first; t < first + cnt; t++) { if (nt) *nt += 1; if (tri_test(s, t, r, h) && any) return 1; }
            } else hits |= 1u << (sl ^ r->oct);
        }
        if (hits) {
            if (gH) { stB[sp] = gB; stM[sp] = (gM << 8) | gH; sp++; }
            gB = (uint32_t)childBase; gM = imask; gH = hits;
        }
        while (gH == 0) {
            if (sp == 0) return h->tri >= 0;
            sp--; gB = stB[sp]; gM = stM[sp] >> 8; gH = stM[sp] & 255u;
        }
        uint32_t k = (uint32_t)__builtin_ctz(gH);
        gH &= gH - 1;
        uint32_t slot = k ^ r->oct;
        node = (int)gB + __builtin_popcount(gM & ((1u << slot) - 1u));
    }
}

/* ---------------------------------------------------------------- BVH4 traversal: src/accel/traverse.cuh:72-107 */
/* the 4 child slabs against the hit distance at node entry, the hit leaves in slot order, then the internal hits that
 * can still hold a closer hit (entry distance <= current hit distance), nearest first: sorted by entry distance with
 * the reference's 5-comparator network (ties keep the network's order), the far ones pushed farthest first */
static int trav4(const orc_scene* s, int root, const ray_t* r, hit_t* h, int any, double* nv, double* nt) {
    int stack[256], sp = 0, node = root;
    const float odx = r->o.x * r->ix, ody = r->o.y * r->iy, odz = r->o.z * r->iz;
    const float INF = 3.4e38f;
    for (;;) {
        const float* n = s->nodes + 32 * (size_t)node;
        if (nv) *nv += 1;
        const float tEntry = h->t;
        float tn[4];
        int hit[4], c[4], k[4];
        for (int j = 0; j < 4; j++) {
            float x0 = fmaf(n[j], r->ix, -odx), x1 = fmaf(n[4 + j], r->ix, -odx);
            float y0 = fmaf(n[8 + j], r->iy, -ody), y1 = fmaf(n[12 + j], r->iy, -ody);
            float z0 = fmaf(n[16 + j], r->iz, -odz), z1 = fmaf(n[20 + j], r->iz, -odz);
            tn[j] = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), r->tmin));
            float tf = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), tEntry));
            hit[j] = tn[j] <= tf;
            c[j] = f2i(n[24 + j]); k[j] = f2i(n[28 + j]);
        }
        for (int j = 0; j < 4; j++)
            if (hit[j] && k[j] > 0)
                for (int t = c[j]; t < c[j] + k[j]; t++) { if (nt) *nt += 1; if (tri_test(s, t, r, h) && any) return 1; }
        float ht[4];
        int hc[4];
        for (int j = 0; j < 4; j++) { ht[j] = hit[j] && k[j] == 0 && tn[j] <= h->t ? tn[j] : INF; hc[j] = c[j]; }
        static const int net[5][2] = { { 0, 1 }, { 2, 3 }, { 0, 2 }, { 1, 3 }, { 1, 2 } };
        for (int q = 0; q < 5; q++) {
            int a = net[q][0], b = net[q][1];
            if (ht[b] < ht[a]) { float tf = ht[a]; ht[a] = ht[b]; ht[b] = tf; int ti = hc[a]; hc[a] = hc[b]; hc[b] = ti; }
        }
        if (ht[0] == INF) { if (sp == 0) break; node = stack[--sp]; continue; }
        if (ht[3] < INF) stack[sp++] = hc[3];
        if (ht[2] < INF) stack[sp++] = hc[2];
        if (ht[1] < INF) stack[sp++] = hc[1];
        node = hc[0];
    }
    return h->tri >= 0;
}

static int trace_blas(const orc_scene* s, int root, const ray_t* r, hit_t* h, int any, double* nv, double* nt) {
    if (s->variant == 2) return trav2(s, root, r, h, any, nv, nt);
    if (s->variant == 4) return trav4(s, root, r, h, any, nv, nt);
    return trav8(s, root, r, h, any, nv, nt);
}

/* ------------------------------------------------ two-level traversal: src/accel/traverse_inst.cuh:10-41 */
/* rows of an instance record (28 floats: inv rows 0-11, fwd rows 12-23, root, mesh, det sign, pad) applied to a point
 * or a vector, summed left to right */
static f3 xf_p(const float* r, f3 p) {
    return mk(r[0] * p.x + r[1] * p.y + r[2] * p.z + r[3], r[4] * p.x + r[5] * p.y + r[6] * p.z + r[7],
              r[8] * p.x + r[9] * p.y + r[10] * p.z + r[11]);
}
static f3 xf_v(const float* r, f3 p) {
    return mk(r[0] * p.x + r[1] * p.y + r[2] * p.z, r[4] * p.x + r[5] * p.y + r[6] * p.z, r[8] * p.x + r[9] * p.y + r[10] * p.z);
}
static f3 xf_n(const float* r, f3 n) {   /* inv^T n */
    return mk(r[0] * n.x + r[4] * n.y + r[8] * n.z, r[1] * n.x + r[5] * n.y + r[9] * n.z, r[2] * n.x + r[6] * n.y + r[10] * n.z);
}
/* TLAS: BVH2 over instance ids (leaf count = instances); an instance is entered with its object-space ray
 * (o' = inv o, d' = inv d, not normalised: hit distances are shared); an instance that lowers the hit distance becomes
 * the hit instance */
static int trav_inst(const orc_scene* s, const ray_t* r, hit_t* h, int any, double* nv, double* nt) {
    int stack[128], sp = 0, node = s->tlas_root;
    const float odx = r->o.x * r->ix, ody = r->o.y * r->iy, odz = r->o.z * r->iz;
    for (;;) {
        const float* n = s->tlas + 16 * (size_t)node;
        int c[2] = { f2i(n[12]), f2i(n[13]) }, k[2] = { f2i(n[14]), f2i(n[15]) };
        float a0 = fmaf(n[0], r->ix, -odx), a1 = fmaf(n[1], r->ix, -odx), a2 = fmaf(n[2], r->iy, -ody), a3 = fmaf(n[3], r->iy, -ody);
        float a4 = fmaf(n[8], r->iz, -odz), a5 = fmaf(n[9], r->iz, -odz);
        float b0 = fmaf(n[4], r->ix, -odx), b1 = fmaf(n[5], r->ix, -odx), b2 = fmaf(n[6], r->iy, -ody), b3 = fmaf(n[7], r->iy, -ody);
        float b4 = fmaf(n[10], r->iz, -odz), b5 = fmaf(n[11], r->iz, -odz);
        float tn[2], tf[2];
        tn[0] = fmaxf(fmaxf(fminf(a0, a1), fminf(a2, a3)), fmaxf(fminf(a4, a5), r->tmin));
        tf[0] = fminf(fminf(fmaxf(a0, a1), fmaxf(a2, a3)), fminf(fmaxf(a4, a5), h->t));
        tn[1] = fmaxf(fmaxf(fminf(b0, b1), fminf(b2, b3)), fmaxf(fminf(b4, b5), r->tmin));
        tf[1] = fminf(fminf(fmaxf(b0, b1), fmaxf(b2, b3)), fminf(fmaxf(b4, b5), h->t));
        int hit[2] = { tn[0] <= tf[0], tn[1] <= tf[1] };
        for (int q = 0; q < 2; q++) {
            if (!(hit[q] && k[q] > 0)) continue;
            for (int t = c[q]; t < c[q] + k[q]; t++) {
                const float* I = s->inst + 28 * (size_t)t;
                ray_t rb = mkray(xf_p(I, r->o), xf_v(I, r->d), r->tmin, s->tri_test);
                const float before = h->t;
                int got = trace_blas(s, f2i(I[24]), &rb, h, any, nv, nt);
                if (h->t < before) h->inst = t;
                if (any && got) return 1;
            }
            hit[q] = 0;
        }
        if (!hit[0] && !hit[1]) { if (sp == 0) break; node = stack[--sp]; }
        else if (hit[0] && hit[1]) { if (tn[1] < tn[0]) { int x = c[0]; c[0] = c[1]; c[1] = x; } stack[sp++] = c[1]; node = c[0]; }
        else node = hit[0] ? c[0] : c[1];
    }
    return h->tri >= 0;
}

static int trace_one(const orc_scene* s, const ray_t* r, hit_t* h, int any, double* nv, double* nt) {
    if (s->tlas) return trav_inst(s, r, h, any, nv, nt);
    return trace_blas(s, s->root, r, h, any, nv, nt);
}

void orc_inst_boxes(const float* inst, int n, const float* mesh_box, float* lo, float* hi) {
    for (int i = 0; i < n; i++) {
        const float* I = inst + 28 * (size_t)i;
        const int m = f2i(I[25]);
        const float* bl = mesh_box + 8 * (size_t)m;
        const float* bh = bl + 4;
        f3 mn = mk(3.4e38f, 3.4e38f, 3.4e38f), mx = mk(-3.4e38f, -3.4e38f, -3.4e38f);
        for (int c = 0; c < 8; c++) {
            f3 w = xf_p(I + 12, mk(c & 1 ? bh[0] : bl[0], c & 2 ? bh[1] : bl[1], c & 4 ? bh[2] : bl[2]));
            mn = mk(w.x < mn.x ? w.x : mn.x, w.y < mn.y ? w.y : mn.y, w.z < mn.z ? w.z : mn.z);
            mx = mk(w.x > mx.x ? w.x : mx.x, w.y > mx.y ? w.y : mx.y, w.z > mx.z ? w.z : mx.z);
        }
        lo[4 * i] = mn.x; lo[4 * i + 1] = mn.y; lo[4 * i + 2] = mn.z; lo[4 * i + 3] = 0.f;
        hi[4 * i] = mx.x; hi[4 * i + 1] = mx.y; hi[4 * i + 2] = mx.z; hi[4 * i + 3] = 0.f;
    }
}

/* --------------------------------------------------------------------------------------------- threading helper */
typedef struct { void (*fn)(void*, int, int); void* arg; int tid, nthreads; } job_t;
static void* job_main(void* a) { job_t* j = (job_t*)a; j->fn(j->arg, j->tid, j->nthreads); return NULL; }
static void run_threads(void (*fn)(void*, int, int), void* arg, int threads) {
    if (threads < 1) threads = 1;
    if (threads == 1) { fn(arg, 0, 1); return; }
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * threads);
    job_t* jobs = (job_t*)malloc(sizeof(job_t) * threads);
    for (int t = 0; t < threads; t++) { jobs[t].fn = fn; jobs[t].arg = arg; jobs[t].tid = t; jobs[t].nthreads = threads; pthread_create(&th[t], NULL, job_main, &jobs[t]); }
    for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
    free(th); free(jobs);
}

typedef struct { const orc_scene* s; const float* rays; int n, any, brute; float* hits; int* inst; } trace_job;
static void trace_worker(void* a, int tid, int nth) {
    trace_job* J = (trace_job*)a;
    for (int k = tid; k < J->n; k += nth) {
        const float* R = J->rays + 8 * (size_t)k;
        ray_t r = mkray(ld3(R), ld3(R + 4), R[3], J->s->tri_test);
        hit_t h = { R[7], 0.f, 0.f, -1, -1 };
        if (J->brute) { for (int t = 0; t < J->s->num_tris; t++) tri_test(J->s, t, &r, &h); }
        else trace_one(J->s, &r, &h, J->any, NULL, NULL);
        float* o = J->hits + 4 * (size_t)k;
        o[0] = h.t; o[1] = h.u; o[2] = h.v; o[3] = i2f(h.tri);
        if (J->inst) J->inst[k] = h.tri >= 0 ? h.inst : -1;
    }
}
int orc_trace(const orc_scene* s, const float* rays, int n, int any, float* hits, int threads) {
    trace_job J = { s, rays, n, any, 0, hits, NULL };
    run_threads(trace_worker, &J, threads);
    return 0;
}
int orc_trace_inst(const orc_scene* s, const float* rays, int n, int any, float* hits, int* inst, int threads) {
    trace_job J = { s, rays, n, any, 0, hits, inst };
    run_threads(trace_worker, &J, threads);
    return 0;
}
int orc_trace_brute(const orc_scene* s, const float* rays, int n, float* hits, int threads) {
    trace_job J = { s, rays, n, 0, 1, hits, NULL };
    run_threads(trace_worker, &J, threads);
    return 0;
}
int orc_trace_work(const orc_scene* s, const float* rays, int n, double* node_visits, double* tri_tests) {
    *node_visits = 0; *tri_tests = 0;
    for (int k = 0; k < n; k++) {
        const float* R = rays + 8 * (size_t)k;
        ray_t r = mkray(ld3(R), ld3(R + 4), R[3], s->tri_test);
        hit_t h = { R[7], 0.f, 0.f, -1, -1 };
        trace_one(s, &r, &h, 0, node_visits, tri_tests);
    }
    return 0;
}

/* ----------------------------------------------------------- camera rays: src/render/wavefront.cu:43-57, camera.cpp */
void orc_camera_ray(const orc_params* p, int pix, uint32_t smp, float o[3], float d[3]) {
    const int x = pix % p->w, y = pix / p->w;
    float u0 = orc_rng4((uint32_t)pix, smp, 0, 0, p->seed), u1 = orc_rng4((uint32_t)pix, smp, 0, 1, p->seed);
    float jx, jy;
    if (p->filter_bh) { jx = table_lookup(p->ftab, 64, u0); jy = table_lookup(p->ftab, 64, u1); }
    else { jx = u0 - 0.5f; jy = u1 - 0.5f; }
    float dx = ((float)x + 0.5f + jx - p->px) / p->fx, dy = -((float)y + 0.5f + jy - p->py) / p->fy;
    f3 ax = ld3(p->ax), ay = ld3(p->ay), az = ld3(p->az), pos = ld3(p->pos);
    f3 dir = sub(add(scl(ax, dx), scl(ay, dy)), az);
    f3 rd = nrm3(dir), ro = pos;
    if (p->lens_r > 0.f) {
        float lx, ly;
        concentric(orc_rng4((uint32_t)pix, smp, 0, 2, p->seed), orc_rng4((uint32_t)pix, smp, 0, 3, p->seed), &lx, &ly);
        f3 pf = add(pos, scl(dir, p->focus_d));
        ro = add(add(pos, scl(ax, lx * p->lens_r)), scl(ay, ly * p->lens_r));
        rd = nrm3(sub(pf, ro));
    }
    o[0] = ro.x; o[1] = ro.y; o[2] = ro.z; d[0] = rd.x; d[1] = rd.y; d[2] = rd.z;
}

/* --------------------------------------------------------- viewer display transform: src/app/display.cu:9-33 */
int orc_srgb_encode8(float x) {
    x = fminf(fmaxf(x, 0.f), 1.f);
    x = x <= 0.0031308f ? 12.92f * x : 1.055f * powf(x, 1.f / 2.4f) - 0.055f;
    return (int)(unsigned char)(x * 255.f + 0.5f);
}
static float tone_curve(float x, int mode) {
    if (mode == 1) return x / (1.f + x);
    if (mode == 2) { x *= 0.6f; return (x * (2.51f * x + 0.03f)) / (x * (2.43f * x + 0.59f) + 0.14f); }
    return x;
}
void orc_display(const float* acc, int n, int spp, float exposure_ev, int mode, uint8_t* out) {
    const float scale = exposure_ev == 0.f ? 1.f : exp2f(exposure_ev), fs = (float)(spp > 0 ? spp : 1);
    for (int i = 0; i < n; i++) {
        for (int c = 0; c < 3; c++) {
            float x = acc[4 * (size_t)i + c] / fs;
            if (scale != 1.f) x = x * scale;
            if (mode != 0) x = tone_curve(fmaxf(x, 0.f), mode);
            out[4 * (size_t)i + c] = (uint8_t)orc_srgb_encode8(x);
        }
        out[4 * (size_t)i + 3] = 255;
    }
}
typedef struct { const float* steps; uint32_t lo, hi; uint64_t* bad; } steps_job;
static void steps_worker(void* a, int tid, int nth) {
    steps_job* J = (steps_job*)a;
    uint64_t bad = 0;
    for (uint64_t b = (uint64_t)J->lo + (uint64_t)tid; b <= J->hi; b += (uint64_t)nth) {
        const float x = u2f((uint32_t)b);
        int k = 0;
        for (int s = 128; s > 0; s >>= 1) if (x >= J->steps[k + s]) k += s;
        if (k != orc_srgb_encode8(x)) bad++;
    }
    J->bad[tid] = bad;
}
uint64_t orc_srgb_steps_check(const float* steps, float lo, float hi, int threads) {
    if (threads < 1) threads = 1;
    uint64_t* bad = (uint64_t*)calloc((size_t)threads, sizeof(uint64_t));
    steps_job J = { steps, f2u(lo), f2u(hi), bad };
    run_threads(steps_worker, &J, threads);
    uint64_t t = 0;
    for (int k = 0; k < threads; k++) t += bad[k];
    free(bad);
    return t;
}

/* ------------------------------------------------- legacy path of one (pixel, sample): src/render/wavefront.cu:94-136 */
static f3 trace_path(const orc_scene* s, const orc_params* p, int pix, uint32_t smp, int bounces, double* st) {
    float o3[3], d3[3];
    orc_camera_ray(p, pix, smp, o3, d3);
    f3 o = ld3(o3), d = ld3(d3), T = mk(1.f, 1.f, 1.f), L = mk(0.f, 0.f, 0.f);
    const f3 sky = ld3(p->sky), sunE = ld3(p->sun_e), sunTo = ld3(p->sun_to);
    const int depths = bounces + 2;
    for (int depth = 0; depth < depths; depth++) {
        ray_t r = mkray(o, d, depth == 0 ? p->near_t : 0.f, s->tri_test);
        hit_t h = { 1e30f, 0.f, 0.f, -1, -1 };
        trace_one(s, &r, &h, 0, NULL, NULL);
        st[0] += 1;
        if (h.tri < 0) { st[2] += 1; L = add(L, mul(T, sky)); break; }
        if (depth > p->max_bounce) break;
        f3 ph = add(o, scl(d, h.t));
        const float* P = s->corners + 12 * h.tri;
        const float* XI = s->tlas ? s->inst + 28 * (size_t)h.inst : NULL;   /* two-level: normals through inv^T */
        f3 e1 = sub(ld3(P + 4), ld3(P)), e2 = sub(ld3(P + 8), ld3(P));
        f3 ng = XI ? nrm3(xf_n(XI, crs(e1, e2))) : nrm3(crs(e1, e2)), ns = ng;
        if (p->has_normals && s->normals) {
            const float* N = s->normals + 12 * h.tri;
            f3 nn = add(add(scl(ld3(N), 1.f - h.u - h.v), scl(ld3(N + 4), h.u)), scl(ld3(N + 8), h.v));
            if (XI) nn = xf_n(XI, nn);
            float l = sqrtf(dot3(nn, nn));
            if (l > 0.f) ns = scl(nn, 1.f / l);
            if (dot3(ns, ng) < 0.f) ns = neg(ns);
        }
        if (dot3(ng, d) > 0.f) { ng = neg(ng); ns = neg(ns); }
        int mat = f2i(P[7]);
        f3 alb = ld3(s->albedo + 4 * mat);
        f3 po = add(ph, scl(ng, p->eps));
        uint32_t dd = (uint32_t)depth + 1u;
        f3 l = cone(sunTo, p->sun_cos, orc_rng4((uint32_t)pix, smp, dd, 2, p->seed), orc_rng4((uint32_t)pix, smp, dd, 3, p->seed));
        float cl = dot3(ns, l);
        if (cl > 0.f && dot3(ng, l) > 0.f) {
            f3 contrib = scl(mul(mul(T, alb), sunE), cl * 0.31830989f);
            ray_t sr = mkray(po, l, 0.f, s->tri_test);
            hit_t sh = { 1e30f, 0.f, 0.f, -1, -1 };
            st[1] += 1;
            if (!trace_one(s, &sr, &sh, 1, NULL, NULL)) { st[3] += 1; L = add(L, contrib); }
        }
        f3 nd = cos_hemi(ns, orc_rng4((uint32_t)pix, smp, dd, 0, p->seed), orc_rng4((uint32_t)pix, smp, dd, 1, p->seed));
        if (dot3(nd, ng) <= 0.f) break;
        T = mul(T, alb);
        o = po; d = nd;
    }
    return L;
}

/* ============================================================================================================
 * Full material integrator: the reference's shadeFull (src/render/wavefront.cu:223-356) with its BSDFs
 * (src/render/bsdf.h:1-220) and light tables (src/scene/environment.cpp:7-37, src/app/scene_setup.cpp:225-247).
 * The path's radiance is summed in the order the product adds it: emission of each hit, then that vertex's
 * unoccluded light sample.
 * ============================================================================================================ */
#define ORC_PI 3.14159265f
#define ORC_TWO_PI 6.2831853f
#define ORC_INV_PI 0.31830989f
#define ORC_DELTA_ALPHA 1e-3f
#define ORC_ONE_MINUS 0.99999994f

typedef struct { f3 base; float metallic; f3 spec; float rough; f3 emis; float ior; float ws, transm; } orc_mat;
static orc_mat ld_mat(const float* m) {
    orc_mat r;
    r.base = ld3(m); r.metallic = m[3]; r.spec = ld3(m + 4); r.rough = m[7]; r.emis = ld3(m + 8); r.ior = m[11];
    r.ws = m[12]; r.transm = m[13];
    return r;
}
static float lum3(f3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }
static int emissive(const orc_mat* m) { return m->emis.x > 0.f || m->emis.y > 0.f || m->emis.z > 0.f; }

/* ---- GGX, Smith height-correlated G2, visible normals (Walter 2007, Heitz 2014, Dupuy & Benyoub 2023) */
static float g_d(float c, float a2) { if (c <= 0.f) return 0.f; float c2 = c * c; float d = c2 * (a2 - 1.f) + 1.f; return a2 / (ORC_PI * d * d); }
static float g_lambda(float c, float a2) { float c2 = c * c; if (c2 < 1e-12f) return 1e12f; return 0.5f * (sqrtf(1.f + a2 * (1.f - c2) / c2) - 1.f); }
static float g_g1(float c, float a2) { return 1.f / (1.f + g_lambda(c, a2)); }
static float g_g2(float co, float ci, float a2) { return 1.f / (1.f + g_lambda(co, a2) + g_lambda(ci, a2)); }
static float g_vpdf(f3 wo, f3 m, float a2) { float om = dot3(wo, m); if (om <= 0.f || m.z <= 0.f) return 0.f; return g_g1(wo.z, a2) * om * g_d(m.z, a2) / wo.z; }
static f3 g_vsample(f3 wo, float a, float u1, float u2) {
    f3 ws = nrm3(mk(a * wo.x, a * wo.y, wo.z));
    float ph = ORC_TWO_PI * u1;
    float z = (1.f - u2) * (1.f + ws.z) - ws.z;
    float sn = sqrtf(fminf(fmaxf(1.f - z * z, 0.f), 1.f));
    f3 h = add(mk(sn * cosf(ph), sn * sinf(ph), z), ws);
    return nrm3(mk(a * h.x, a * h.y, fmaxf(h.z, 0.f)));
}
static float fr_diel(float ci, float eta) {
    float s2t = (1.f - ci * ci) / (eta * eta);
    if (s2t >= 1.f) return 1.f;
    float ct = sqrtf(1.f - s2t);
    float rs = (ci - eta * ct) / (ci + eta * ct), rp = (eta * ci - ct) / (eta * ci + ct);
    return 0.5f * (rs * rs + rp * rp);
}
static f3 fr_f82(f3 F0, f3 tint, float mu) {
    mu = fminf(fmaxf(mu, 0.f), 1.f);
    float m = 1.f - mu, m2 = m * m, m5 = m2 * m2 * m, m6 = m5 * m;
    const float mb = 1.f / 7.f, mb5 = 0.46266436f, mb6 = 0.39656945f;
    f3 one = mk(1.f, 1.f, 1.f);
    f3 Fs = add(F0, scl(sub(one, F0), m5));
    f3 Fsb = add(F0, scl(sub(one, F0), mb5));
    f3 b = scl(mul(Fsb, sub(one, tint)), 1.f / (mb * mb6));
    f3 F = sub(Fs, scl(b, mu * m6));
    return mk(fminf(fmaxf(F.x, 0.f), 1.f), fminf(fmaxf(F.y, 0.f), 1.f), fminf(fmaxf(F.z, 0.f), 1.f));
}

/* ---- lobes; local frame, wo.z > 0 */
typedef struct { f3 f; float pdf; } bval;
/* bump-terminator shadowing of the diffuse lobe (Chiang, Li and Burley 2019): nr = the unperturbed normal in the local
 * frame of the mapped shading normal, g = (nr . wi) / (wi.z nr.z), G = clamp(-g^3 + g^2 + g); nr = +z: off
 * (src/render/bsdf.h:71-75) */
static float term_shadow(f3 nr, f3 wi) {
    if (nr.z >= 1.f) return 1.f;
    float g = dot3(nr, wi) / (wi.z * nr.z);
    if (g >= 1.f) return 1.f;
    if (g < 0.f) return 0.f;
    return g + g * g - g * g * g;
}
typedef struct { f3 base; float ws, eta, a, a2; int delta; float Fo, pSpec, kd; f3 nr; } gd_t;
static gd_t gd_make(const orc_mat* m, f3 wo, int energy, float ux1, float ux2, f3 nr) {
    gd_t g;
    g.nr = nr;
    g.base = m->base; g.ws = m->ws; g.eta = m->ior; g.a = m->rough * m->rough; g.a2 = g.a * g.a; g.delta = g.a < ORC_DELTA_ALPHA;
    g.Fo = g.ws > 0.f ? g.ws * fr_diel(wo.z, g.eta) : 0.f;
    float pd = (1.f - g.Fo) * lum3(g.base);
    g.pSpec = g.Fo + pd > 0.f ? g.Fo / (g.Fo + pd) : 0.f;
    g.kd = 1.f - g.Fo;
    if (energy && g.ws > 0.f && !g.delta) {
        f3 mm = g_vsample(wo, g.a, ux1, ux2);
        float om = dot3(wo, mm);
        f3 wr = sub(scl(mm, 2.f * om), wo);
        float X = (om > 0.f && wr.z > 0.f) ? g.ws * fr_diel(om, g.eta) * g_g2(wo.z, wr.z, g.a2) / g_g1(wo.z, g.a2) : 0.f;
        g.kd = 1.f - X;
    }
    g.kd = fmaxf(g.kd, 0.f);
    return g;
}
static bval gd_eval(const gd_t* g, f3 wo, f3 wi) {
    bval r = { { 0.f, 0.f, 0.f }, 0.f };
    if (wi.z <= 0.f) return r;
    r.f = scl(g->base, g->kd * ORC_INV_PI * term_shadow(g->nr, wi));
    r.pdf = (1.f - g->pSpec) * wi.z * ORC_INV_PI;
    if (!g->delta && g->Fo > 0.f) {
        f3 m = nrm3(add(wo, wi));
        float om = dot3(wo, m);
        if (om > 0.f) {
            float F = g->ws * fr_diel(om, g->eta);
            float sp = F * g_d(m.z, g->a2) * g_g2(wo.z, wi.z, g->a2) / (4.f * wo.z * wi.z);
            r.f = add(r.f, mk(sp, sp, sp));
            r.pdf += g->pSpec * g_vpdf(wo, m, g->a2) / (4.f * om);
        }
    }
    return r;
}
static int gd_sample(const gd_t* g, f3 wo, float u, float u1, float u2, f3* wi, f3* w, float* pdf) {
    if (u < g->pSpec) {
        if (g->delta) { *wi = mk(-wo.x, -wo.y, wo.z); float s = g->Fo / g->pSpec; *w = mk(s, s, s); *pdf = 0.f; return 1; }
        f3 m = g_vsample(wo, g->a, u1, u2);
        *wi = sub(scl(m, 2.f * dot3(wo, m)), wo);
    } else {
        *wi = cos_hemi(mk(0.f, 0.f, 1.f), u1, u2);
    }
    if (wi->z <= 0.f) return 0;
    bval e = gd_eval(g, wo, *wi);
    if (e.pdf <= 0.f) return 0;
    *w = scl(e.f, wi->z / e.pdf); *pdf = e.pdf;
    return 1;
}
static bval cd_eval(const orc_mat* mt, f3 wo, f3 wi) {
    bval r = { { 0.f, 0.f, 0.f }, 0.f };
    float a = mt->rough * mt->rough, a2 = a * a;
    if (a < ORC_DELTA_ALPHA || wi.z <= 0.f) return r;
    f3 m = nrm3(add(wo, wi));
    float om = dot3(wo, m);
    if (om <= 0.f) return r;
    r.f = scl(fr_f82(mt->base, mt->spec, om), mt->ws * g_d(m.z, a2) * g_g2(wo.z, wi.z, a2) / (4.f * wo.z * wi.z));
    r.pdf = g_vpdf(wo, m, a2) / (4.f * om);
    return r;
}
static int cd_sample(const orc_mat* mt, f3 wo, float u1, float u2, f3* wi, f3* w, float* pdf) {
    float a = mt->rough * mt->rough, a2 = a * a;
    if (a < ORC_DELTA_ALPHA) { *wi = mk(-wo.x, -wo.y, wo.z); *w = scl(fr_f82(mt->base, mt->spec, wo.z), mt->ws); *pdf = 0.f; return 1; }
    f3 m = g_vsample(wo, a, u1, u2);
    float om = dot3(wo, m);
    *wi = sub(scl(m, 2.f * om), wo);
    if (wi->z <= 0.f || om <= 0.f) return 0;
    *w = scl(fr_f82(mt->base, mt->spec, om), mt->ws * g_g2(wo.z, wi->z, a2) / g_g1(wo.z, a2));
    *pdf = g_vpdf(wo, m, a2) / (4.f * om);
    return 1;
}
static bval dl_eval(const orc_mat* mt, float eta, f3 wo, f3 wi) {
    bval r = { { 0.f, 0.f, 0.f }, 0.f };
    float a = mt->rough * mt->rough, a2 = a * a;
    if (a < ORC_DELTA_ALPHA) return r;
    if (wi.z > 0.f) {
        f3 m = nrm3(add(wo, wi));
        float om = dot3(wo, m);
        if (om <= 0.f) return r;
        float F = fr_diel(om, eta);
        float sp = F * g_d(m.z, a2) * g_g2(wo.z, wi.z, a2) / (4.f * wo.z * wi.z);
        r.f = mk(sp, sp, sp); r.pdf = F * g_vpdf(wo, m, a2) / (4.f * om);
    } else if (wi.z < 0.f) {
        f3 m = nrm3(add(wo, scl(wi, eta)));
        if (m.z < 0.f) m = neg(m);
        float om = dot3(wo, m), im = dot3(wi, m);
        if (om <= 0.f || im >= 0.f) return r;
        float F = fr_diel(om, eta);
        if (F >= 1.f) return r;
        float den = om + eta * im, jac = eta * eta * (-im) / (den * den);
        float sp = (1.f - F) * g_d(m.z, a2) * g_g2(wo.z, -wi.z, a2) * om * jac / (wo.z * -wi.z);
        r.f = scl(mt->base, sp); r.pdf = (1.f - F) * g_vpdf(wo, m, a2) * jac;
    }
    return r;
}
static int dl_sample(const orc_mat* mt, float eta, f3 wo, float u, float u1, float u2, f3* wi, f3* w, float* pdf) {
    float a = mt->rough * mt->rough;
    int delta = a < ORC_DELTA_ALPHA;
    f3 m = delta ? mk(0.f, 0.f, 1.f) : g_vsample(wo, a, u1, u2);
    float om = dot3(wo, m);
    if (om <= 0.f) return 0;
    float F = fr_diel(om, eta);
    if (u < F) {
        *wi = sub(scl(m, 2.f * om), wo);
        if (wi->z <= 0.f) return 0;
        if (delta) { *w = mk(1.f, 1.f, 1.f); *pdf = 0.f; return 1; }
    } else {
        float ct = sqrtf(fmaxf(1.f - (1.f - om * om) / (eta * eta), 0.f));
        *wi = add(scl(wo, -1.f / eta), scl(m, om / eta - ct));
        if (wi->z >= 0.f) return 0;
        *wi = nrm3(*wi);
        if (delta) { *w = mt->base; *pdf = 0.f; return 1; }
    }
    bval e = dl_eval(mt, eta, wo, *wi);
    if (e.pdf <= 0.f) return 0;
    *w = scl(e.f, fabsf(wi->z) / e.pdf); *pdf = e.pdf;
    return 1;
}
static void base_w(const orc_mat* m, float* wM, float* wT, float* wD) {
    *wM = m->metallic; *wT = (1.f - m->metallic) * m->transm; *wD = (1.f - m->metallic) * (1.f - m->transm);
}
static int has_non_delta(const orc_mat* m) {
    float wM, wT, wD; base_w(m, &wM, &wT, &wD);
    int rough = m->rough * m->rough >= ORC_DELTA_ALPHA;
    return (wD > 0.f && (lum3(m->base) > 0.f || rough)) || ((wM > 0.f || wT > 0.f) && rough);
}
static bval mat_eval(const orc_mat* m, f3 wo, f3 wi, float eta, int gde, float ux1, float ux2, f3 nr) {
    float wM, wT, wD; base_w(m, &wM, &wT, &wD);
    bval r = { { 0.f, 0.f, 0.f }, 0.f };
    if (wM > 0.f) { bval e = cd_eval(m, wo, wi); r.f = add(r.f, scl(e.f, wM)); r.pdf += wM * e.pdf; }
    if (wT > 0.f) { bval e = dl_eval(m, eta, wo, wi); r.f = add(r.f, scl(e.f, wT)); r.pdf += wT * e.pdf; }
    if (wD > 0.f) { gd_t g = gd_make(m, wo, gde, ux1, ux2, nr); bval e = gd_eval(&g, wo, wi); r.f = add(r.f, scl(e.f, wD)); r.pdf += wD * e.pdf; }
    return r;
}
static int mat_sample(const orc_mat* m, f3 wo, float eta, int gde, float ux1, float ux2, float u, float u1, float u2, f3* wi, f3* w, float* pdf, f3 nr) {
    float wM, wT, wD; base_w(m, &wM, &wT, &wD);
    int ok;
    if (u < wM) ok = cd_sample(m, wo, u1, u2, wi, w, pdf);
    else if (u < wM + wT) { float uu = (u - wM) / wT; ok = dl_sample(m, eta, wo, fminf(uu, ORC_ONE_MINUS), u1, u2, wi, w, pdf); }
    else { float uu = wD > 0.f ? (u - wM - wT) / wD : 0.f; gd_t g = gd_make(m, wo, gde, ux1, ux2, nr); ok = gd_sample(&g, wo, fminf(uu, ORC_ONE_MINUS), u1, u2, wi, w, pdf); }
    if (!ok) return 0;
    if (*pdf <= 0.f) return 1;
    if (wM == 1.f || wT == 1.f || wD == 1.f) return 1;
    bval e = mat_eval(m, wo, *wi, eta, gde, ux1, ux2, nr);
    if (e.pdf <= 0.f) return 0;
    *w = scl(e.f, fabsf(wi->z) / e.pdf); *pdf = e.pdf;
    return 1;
}
static float mis2(float a, float b) { float a2 = a * a; return a2 / (a2 + b * b); }

void orc_bsdf_eval(const float* mat, const float wo[3], const float wi[3], float eta, int gde, float ux1, float ux2, float f[3], float* pdf) {
    orc_mat m = ld_mat(mat);
    bval e = mat_eval(&m, ld3(wo), ld3(wi), eta, gde, ux1, ux2, mk(0.f, 0.f, 1.f));
    f[0] = e.f.x; f[1] = e.f.y; f[2] = e.f.z; *pdf = e.pdf;
}
int orc_bsdf_sample(const float* mat, const float wo[3], float eta, int gde, float ux1, float ux2, float u, float u1, float u2,
                    float wi[3], float weight[3], float* pdf) {
    orc_mat m = ld_mat(mat);
    f3 a, w; float p = 0.f;
    int ok = mat_sample(&m, ld3(wo), eta, gde, ux1, ux2, u, u1, u2, &a, &w, &p, mk(0.f, 0.f, 1.f));
    wi[0] = a.x; wi[1] = a.y; wi[2] = a.z; weight[0] = w.x; weight[1] = w.y; weight[2] = w.z; *pdf = p;
    return ok;
}
static uint64_t sm_next(uint64_t* s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull; z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static float sm_u(uint64_t* s) { float f = (float)(sm_next(s) >> 40) * (1.0f / 16777216.0f); return f < 1.f ? f : ORC_ONE_MINUS; }
void orc_bsdf_sample_batch(const float* mat, const float wo[3], float eta, int gde, int n, uint64_t seed, float* wi, float* weight,
                           float* pdf, int* ok) {
    orc_mat m = ld_mat(mat);
    uint64_t st = seed;
    for (int k = 0; k < n; k++) {
        float u = sm_u(&st), u1 = sm_u(&st), u2 = sm_u(&st), x1 = sm_u(&st), x2 = sm_u(&st);
        f3 a = { 0, 0, 0 }, w = { 0, 0, 0 }; float p = 0.f;
        ok[k] = mat_sample(&m, ld3(wo), eta, gde, x1, x2, u, u1, u2, &a, &w, &p, mk(0.f, 0.f, 1.f));
        wi[3 * k] = a.x; wi[3 * k + 1] = a.y; wi[3 * k + 2] = a.z;
        weight[3 * k] = w.x; weight[3 * k + 1] = w.y; weight[3 * k + 2] = w.z; pdf[k] = p;
    }
}
void orc_bsdf_pdf_batch(const float* mat, const float wo[3], float eta, int gde, const float* wi, int n, float* pdf) {
    orc_mat m = ld_mat(mat);
    for (int k = 0; k < n; k++) pdf[k] = mat_eval(&m, ld3(wo), ld3(wi + 3 * k), eta, gde, 0.5f, 0.5f, mk(0.f, 0.f, 1.f)).pdf;
}
float orc_fresnel_dielectric(float ci, float eta) { return fr_diel(ci, eta); }
void orc_fresnel_f82(const float f0[3], const float tint[3], float mu, float out[3]) {
    f3 r = fr_f82(ld3(f0), ld3(tint), mu); out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
float orc_ggx_d(float c, float a2) { return g_d(c, a2); }
float orc_ggx_g2(float co, float ci, float a2) { return g_g2(co, ci, a2); }
double orc_quad_albedo(int kind, float rough, float eta, const float wo3[3]) {
    const double a = (double)rough * rough, a2 = a * a;
    const int NU = 1024, NV = 256;
    f3 wo = ld3(wo3);
    double sum = 0;
    for (int i = 0; i < NU; i++)
        for (int j = 0; j < NV; j++) {
            /* m ~ D(m) cos(theta_m) by the GGX inverse CDF; u1 = 1 - s^2 removes the endpoint singularity */
            double sq = (i + 0.5) / NU, u1 = 1.0 - sq * sq, u2 = (j + 0.5) / NV, jac = 2.0 * sq;
            double t2 = a2 * u1 / (1.0 - u1), ct = 1.0 / sqrt(1.0 + t2), st = sqrt(fmax(0.0, 1 - ct * ct)), ph = 2 * M_PI * u2;
            f3 m = mk((float)(st * cos(ph)), (float)(st * sin(ph)), (float)ct);
            double om = dot3(wo, m);
            if (om <= 0) continue;
            double g = om / (wo.z * ct);
            f3 wr = sub(scl(m, (float)(2 * om)), wo);
            double val = 0;
            if (kind == 0) { if (wr.z > 0) val = g_g2(wo.z, wr.z, (float)a2); }
            else if (kind == 1) {
                double F = fr_diel((float)om, eta);
                if (wr.z > 0) val += F * g_g2(wo.z, wr.z, (float)a2);
                if (F < 1) {
                    double cT = sqrt(fmax(0.0, 1 - (1 - om * om) / ((double)eta * eta)));
                    f3 wt = add(scl(wo, -1.f / eta), scl(m, (float)(om / eta - cT)));
                    if (wt.z < 0) val += (1 - F) * g_g2(wo.z, -nrm3(wt).z, (float)a2);
                }
            } else { if (wr.z > 0) val = fr_diel((float)om, eta) * g_g2(wo.z, wr.z, (float)a2); }
            sum += g * val * jac;
        }
    return sum / ((double)NU * NV);
}

/* ---- environment: equirectangular, row 0 = +z, u = 0.5 - atan2(y, x) / 2pi (src/scene/environment.h:1-5) */
void orc_env_build(const float* rgb, int w, int h, float strength, float* tex, float* marg, float* cond) {
    double* row = (double*)malloc(sizeof(double) * (w + 1));
    double* rs = (double*)malloc(sizeof(double) * h);
    double* mc = (double*)malloc(sizeof(double) * (h + 1));
    mc[0] = 0.0;
    for (int y = 0; y < h; y++) {
        double st = sin(M_PI * (y + 0.5) / h);
        row[0] = 0.0;
        for (int x = 0; x < w; x++) {
            const float* c = rgb + 3 * ((size_t)y * w + x);
            f3 L = scl(mk(c[0], c[1], c[2]), strength);
            float* t = tex + 4 * ((size_t)y * w + x);
            t[0] = L.x; t[1] = L.y; t[2] = L.z; t[3] = 0.f;
            double lum = 0.2126 * L.x + 0.7152 * L.y + 0.0722 * L.z;
            row[x + 1] = row[x] + (lum > 0 ? lum : 0.0) * st;
        }
        rs[y] = row[w];
        float* cr = cond + (size_t)y * (w + 1);
        for (int x = 0; x <= w; x++) cr[x] = rs[y] > 0 ? (float)(row[x] / rs[y]) : (float)x / (float)w;
        cr[w] = 1.f;
        mc[y + 1] = mc[y] + rs[y];
    }
    for (int y = 0; y <= h; y++) marg[y] = mc[h] > 0 ? (float)(mc[y] / mc[h]) : (float)y / (float)h;
    marg[h] = 1.f;
    free(row); free(rs); free(mc);
}
static int cdf_find(const float* cdf, int n, float u) {
    int lo = 0, hi = n;
    while (hi - lo > 1) { int mid = (lo + hi) >> 1; if (cdf[mid] <= u) lo = mid; else hi = mid; }
    return lo;
}
static void env_uv(const orc_scene* s, f3 d, float* u, float* v) {
    *u = 0.5f - (atan2f(d.y, d.x) - s->env_rot) * (1.f / ORC_TWO_PI);
    *u -= floorf(*u);
    *v = acosf(fminf(fmaxf(d.z, -1.f), 1.f)) * ORC_INV_PI;
}
static f3 env_dir(const orc_scene* s, float u, float v) {
    float ph = (0.5f - u) * ORC_TWO_PI + s->env_rot, th = v * ORC_PI, st = sinf(th);
    return mk(st * cosf(ph), st * sinf(ph), cosf(th));
}
static void env_texel(const orc_scene* s, f3 d, int* x, int* y, float* v) {
    float u; env_uv(s, d, &u, v);
    *x = (int)(u * (float)s->env_w); if (*x > s->env_w - 1) *x = s->env_w - 1;
    *y = (int)(*v * (float)s->env_h); if (*y > s->env_h - 1) *y = s->env_h - 1;
}
static f3 env_rad(const orc_scene* s, f3 d) { int x, y; float v; env_texel(s, d, &x, &y, &v); return ld3(s->env_tex + 4 * ((size_t)y * s->env_w + x)); }
static float env_pdf_texel(const orc_scene* s, int x, int y, float v) {
    const float* cr = s->env_cond + (size_t)y * (s->env_w + 1);
    float puv = (s->env_marg[y + 1] - s->env_marg[y]) * (cr[x + 1] - cr[x]) * (float)s->env_w * (float)s->env_h;
    float st = sinf(v * ORC_PI);
    return st > 0.f ? puv / (2.f * ORC_PI * ORC_PI * st) : 0.f;
}
static float env_pdf_dir(const orc_scene* s, f3 d) {
    if (s->env_uniform) return 0.25f * ORC_INV_PI;
    int x, y; float v; env_texel(s, d, &x, &y, &v); return env_pdf_texel(s, x, y, v);
}
float orc_env_pdf(const orc_scene* s, const float d[3]) { return env_pdf_dir(s, ld3(d)); }
/* luminance sample from (ue, u2, u3): direction, texel and pdf (0 if none) */
static float env_sample(const orc_scene* s, float ue, float u2, float u3, f3* l, int* tx, int* ty) {
    int y = cdf_find(s->env_marg, s->env_h, ue);
    const float* cr = s->env_cond + (size_t)y * (s->env_w + 1);
    int x = cdf_find(cr, s->env_w, u2);
    float v = ((float)y + u3) / (float)s->env_h, u = ((float)x + 0.5f) / (float)s->env_w;
    float w0 = cr[x + 1] - cr[x];
    if (w0 > 0.f) u = ((float)x + fminf((u2 - cr[x]) / w0, ORC_ONE_MINUS)) / (float)s->env_w;
    *l = env_dir(s, u, v); *tx = x; *ty = y;
    return env_pdf_texel(s, x, y, v);
}
void orc_env_sample_batch(const orc_scene* s, int n, uint64_t seed, float* dirs, float* pdf) {
    uint64_t st = seed;
    for (int k = 0; k < n; k++) {
        float ue = sm_u(&st), u2 = sm_u(&st), u3 = sm_u(&st);
        f3 l; int x, y;
        pdf[k] = env_sample(s, ue, u2, u3, &l, &x, &y);
        dirs[3 * k] = l.x; dirs[3 * k + 1] = l.y; dirs[3 * k + 2] = l.z;
    }
}

/* ---- emissive triangles by power, in mesh order (src/app/scene_setup.cpp:225-247) */
typedef struct { int tri, pos; } em_pair;
static int em_cmp(const void* a, const void* b) { int x = ((const em_pair*)a)->tri, y = ((const em_pair*)b)->tri; return (x > y) - (x < y); }
int orc_emitter_tables(const float* corners, const int* order, int nt, const float* mats, int* em_tri, float* em_cdf, float* tri_light_pdf) {
    em_pair* em = (em_pair*)malloc(sizeof(em_pair) * (nt > 0 ? nt : 1));
    int ne = 0;
    for (int k = 0; k < nt; k++) {
        orc_mat m = ld_mat(mats + 16 * f2i(corners[12 * (size_t)k + 7]));
        if (emissive(&m)) { em[ne].tri = order[k]; em[ne].pos = k; ne++; }
    }
    qsort(em, ne, sizeof(em_pair), em_cmp);
    double* pw = (double*)malloc(sizeof(double) * (ne + 1));
    double* ar = (double*)malloc(sizeof(double) * (ne + 1));
    double tot = 0;
    int n = 0;
    for (int e = 0; e < ne; e++) {
        const float* C = corners + 12 * (size_t)em[e].pos;
        f3 e1 = sub(ld3(C + 4), ld3(C)), e2 = sub(ld3(C + 8), ld3(C));
        f3 cr = crs(e1, e2);
        double a = 0.5 * (double)sqrtf(dot3(cr, cr));
        orc_mat m = ld_mat(mats + 16 * f2i(C[7]));
        double p = (double)lum3(m.emis) * a;
        if (!(p > 0)) continue;
        em_tri[n] = em[e].pos; pw[n] = p; ar[n] = a; tot += p; n++;
    }
    for (int k = 0; k < nt; k++) tri_light_pdf[k] = 0.f;
    em_cdf[0] = 0.f;
    double c = 0;
    for (int e = 0; e < n; e++) {
        tri_light_pdf[em_tri[e]] = (float)(pw[e] / tot / ar[e]);
        c += pw[e];
        em_cdf[e + 1] = (float)(c / tot);
    }
    if (n) em_cdf[n] = 1.f;
    free(em); free(pw); free(ar);
    return n;
}

/* ---- two-level scenes: world emitters per frame (src/app/scene_setup.cpp:273-299) */
int orc_inst_emitters(const float* inst, int n_inst, const int* mesh_tris, const float* corners, const int* order,
                      const float* mats, int num_mats, float* em_tris, float* em_cdf, float* mlp) {
    int n = 0, nl = 0;
    double tot = 0;
    double* pw = NULL;
    for (int i = 0; i < n_inst; i++) {   /* leaves of all meshes: the largest first + count */
        const int mesh = f2i(inst[28 * (size_t)i + 25]);
        if (mesh_tris[2 * mesh] + mesh_tris[2 * mesh + 1] > nl) nl = mesh_tris[2 * mesh] + mesh_tris[2 * mesh + 1];
    }
    int* pos_of = (int*)malloc(sizeof(int) * (nl > 0 ? nl : 1));
    for (int k = 0; k < nl; k++) pos_of[order[k]] = k;
    for (int pass = 0; pass < 2; pass++) {
        n = 0; tot = 0;
        for (int i = 0; i < n_inst; i++) {
            const float* R = inst + 28 * (size_t)i;
            const int mesh = f2i(R[25]);
            const int first = mesh_tris[2 * mesh], cnt = mesh_tris[2 * mesh + 1];
            /* the mesh's triangles in mesh order: leaf of local triangle t = the leaf whose order entry is first + t */
            for (int t = 0; t < cnt; t++) {
                const int leaf = pos_of[first + t];
                const float* Cn = corners + 12 * (size_t)leaf;
                const int mi = f2i(Cn[7]);
                orc_mat m = ld_mat(mats + 16 * mi);
                if (!emissive(&m)) continue;
                if (pass == 1) {
                    f3 w0 = xf_p(R + 12, ld3(Cn)), w1 = xf_p(R + 12, ld3(Cn + 4)), w2 = xf_p(R + 12, ld3(Cn + 8));
                    f3 e1 = sub(w1, w0), e2 = sub(w2, w0);
                    f3 cr = crs(e1, e2);
                    double area = 0.5 * (double)sqrtf(dot3(cr, cr));
                    float* E = em_tris + 12 * (size_t)n;
                    E[0] = w0.x; E[1] = w0.y; E[2] = w0.z; E[3] = i2f(leaf);
                    E[4] = e1.x; E[5] = e1.y; E[6] = e1.z; E[7] = i2f(mi);
                    E[8] = e2.x; E[9] = e2.y; E[10] = e2.z; E[11] = 0.f;
                    double p = (double)lum3(m.emis) * area;
                    pw[n] = p > 0 ? p : 0; tot += pw[n];
                }
                n++;
            }
        }
        if (pass == 0) {
            if (!em_tris) { free(pos_of); return n; }
            pw = (double*)malloc(sizeof(double) * (n + 1));
        }
    }
    em_cdf[0] = 0.f;
    double c = 0;
    for (int e = 0; e < n; e++) { c += pw[e]; em_cdf[e + 1] = (float)(c / tot); }
    if (n) em_cdf[n] = 1.f;
    for (int j = 0; j < num_mats; j++) {
        orc_mat m = ld_mat(mats + 16 * j);
        mlp[j] = emissive(&m) && n ? (float)(lum3(m.emis) / tot) : 0.f;
    }
    free(pw);
    free(pos_of);
    return n;
}

/* ---- textures: the product's lookup (its core/texture.h): coordinates wrap (t - floor t, NaN -> 0), nearest takes
 * texel floor(t size), bilinear blends the 2 x 2 texels around t size - 0.5 (rows top first, so v is flipped), each
 * texel decoded first (colour channels of sRGB images through the 256-entry table, otherwise byte / 255); maps
 * multiply the material's factors; a normal map bends the shading normal in its tangent frame
 * (src/render/wavefront.cu:199-204, 252-271) */
typedef struct { float x, y, z, w; } f4;
static float tex_chan(uint32_t px, int c, int srgb, const float* tab) {
    uint32_t b = (px >> (8 * c)) & 255u;
    return (srgb && c < 3) ? tab[b] : (float)b * (1.f / 255.f);
}
static f4 tex_texel(uint32_t px, int srgb, const float* tab) {
    f4 r = { tex_chan(px, 0, srgb, tab), tex_chan(px, 1, srgb, tab), tex_chan(px, 2, srgb, tab), tex_chan(px, 3, srgb, tab) };
    return r;
}
static float lerp1(float a, float b, float t) { return fmaf(t, b - a, a); }
static f4 lerp4(f4 a, f4 b, float t) { f4 r = { lerp1(a.x, b.x, t), lerp1(a.y, b.y, t), lerp1(a.z, b.z, t), lerp1(a.w, b.w, t) }; return r; }
static f4 orc_tex_sample_img(const int* img, const uint32_t* texels, const float* tab, float u, float v, int nearest) {
    const uint32_t first = (uint32_t)img[0];
    const int w = img[1], h = img[2], srgb = img[3];
    const uint32_t* px = texels + first;
    float vv = 1.f - v, tu = u - floorf(u), tv = vv - floorf(vv);
    if (!(tu >= 0.f && tu <= 1.f)) tu = 0.f;
    if (!(tv >= 0.f && tv <= 1.f)) tv = 0.f;
    if (nearest) {
        int xi = (int)(tu * (float)w), yi = (int)(tv * (float)h);
        if (xi > w - 1) xi = w - 1;
        if (yi > h - 1) yi = h - 1;
        return tex_texel(px[(size_t)yi * w + xi], srgb, tab);
    }
    float x = tu * (float)w - 0.5f, y = tv * (float)h - 0.5f;
    float fx = floorf(x), fy = floorf(y);
    float ax = x - fx, ay = y - fy;
    int x0 = (int)fx, y0 = (int)fy;
    if (x0 < 0) x0 += w;
    if (y0 < 0) y0 += h;
    int x1 = x0 + 1 < w ? x0 + 1 : 0, y1 = y0 + 1 < h ? y0 + 1 : 0;
    f4 top = lerp4(tex_texel(px[(size_t)y0 * w + x0], srgb, tab), tex_texel(px[(size_t)y0 * w + x1], srgb, tab), ax);
    f4 bot = lerp4(tex_texel(px[(size_t)y1 * w + x0], srgb, tab), tex_texel(px[(size_t)y1 * w + x1], srgb, tab), ax);
    return lerp4(top, bot, ay);
}
void orc_texture_sample(const int img[4], const uint32_t* texels, const float* srgb_table, float u, float v, int nearest,
                        float out[4]) {
    f4 c = orc_tex_sample_img(img, texels, srgb_table, u, v, nearest);
    out[0] = c.x; out[1] = c.y; out[2] = c.z; out[3] = c.w;
}
static f4 tex_sample(const orc_scene* s, int id, float u, float v) {
    return orc_tex_sample_img(s->tex_images + 4 * id, s->texels, s->srgb, u, v, s->tex_nearest);
}
static float chan4(f4 c, int k) { return k == 0 ? c.x : (k == 1 ? c.y : (k == 2 ? c.z : c.w)); }
/* texture coordinates of a point of leaf triangle t with barycentric weights b1, b2 of corners 1 and 2 */
static void tri_uv(const orc_scene* s, int t, float b1, float b2, float* u, float* v) {
    const float* q = s->uvs + 8 * (size_t)t;
    float b0 = 1.f - b1 - b2;
    *u = q[0] * b0 + q[2] * b1 + q[4] * b2;
    *v = q[1] * b0 + q[3] * b1 + q[5] * b2;
}

/* ---- one path of the full integrator */
static f3 trace_path_full(const orc_scene* s, const orc_params* p, int pix, uint32_t smp, int bounces, double* st) {
    float o3[3], d3[3];
    orc_camera_ray(p, pix, smp, o3, d3);
    f3 o = ld3(o3), d = ld3(d3), T = mk(1.f, 1.f, 1.f), L = mk(0.f, 0.f, 0.f);
    const f3 sky = ld3(p->sky), sunE = ld3(p->sun_e), sunTo = ld3(p->sun_to);
    const uint32_t seed = p->seed;
    float lastPdf = 0.f;
    const int depths = bounces + 2;
    for (int depth = 0; depth < depths; depth++) {
        ray_t r = mkray(o, d, depth == 0 ? p->near_t : 0.f, s->tri_test);
        hit_t h = { 1e30f, 0.f, 0.f, -1, -1 };
        trace_one(s, &r, &h, 0, NULL, NULL);
        st[0] += 1;
        if (h.tri < 0) {
            st[2] += 1;
            f3 Lm;
            if (s->env_w > 0) {
                Lm = env_rad(s, d);
                if (lastPdf > 0.f) {
                    if (s->light_mode == 1) Lm = mk(0.f, 0.f, 0.f);
                    else if (s->light_mode == 0) Lm = scl(Lm, mis2(lastPdf, s->p_env * env_pdf_dir(s, d)));
                }
            } else Lm = sky;
            L = add(L, mul(T, Lm));
            break;
        }
        f3 ph = add(o, scl(d, h.t));
        const float* P = s->corners + 12 * h.tri;
        const float* XI = s->tlas ? s->inst + 28 * (size_t)h.inst : NULL;   /* two-level: normals through inv^T */
        f3 e1 = sub(ld3(P + 4), ld3(P)), e2 = sub(ld3(P + 8), ld3(P));
        f3 ng0 = XI ? scl(nrm3(xf_n(XI, crs(e1, e2))), XI[26]) : nrm3(crs(e1, e2)), ns = ng0;
        if (p->has_normals && s->normals) {
            const float* N = s->normals + 12 * h.tri;
            f3 nn = add(add(scl(ld3(N), 1.f - h.u - h.v), scl(ld3(N + 4), h.u)), scl(ld3(N + 8), h.v));
            if (XI) nn = xf_n(XI, nn);
            float l = sqrtf(dot3(nn, nn));
            if (l > 0.f) ns = scl(nn, 1.f / l);
            if (dot3(ns, ng0) < 0.f) ns = neg(ns);
        }
        const int mid = f2i(P[7]);
        orc_mat M = ld_mat(s->mats + 16 * mid);
        f3 nsU = ns;   /* the shading normal before normal mapping */
        int mapped = 0;
        if (s->maps && mid < s->num_maps) {
            const int* X = s->maps + 8 * mid;
            if (X[0] >= 0 || X[1] >= 0 || X[2] >= 0 || X[3] >= 0 || X[4] >= 0) {
                float tu, tv;
                tri_uv(s, h.tri, h.u, h.v, &tu, &tv);
                if (X[0] >= 0) { f4 c = tex_sample(s, X[0], tu, tv); M.base = mul(M.base, mk(c.x, c.y, c.z)); }
                if (X[1] >= 0) { f4 c = tex_sample(s, X[1], tu, tv); M.emis = mul(M.emis, mk(c.x, c.y, c.z)); }
                if (X[3] >= 0) M.metallic *= chan4(tex_sample(s, X[3], tu, tv), X[5]);
                if (X[4] >= 0) M.rough *= chan4(tex_sample(s, X[4], tu, tv), X[6]);
                if (X[2] >= 0 && s->tans) {
                    const float* T0 = s->tans + 12 * (size_t)h.tri;
                    float w0 = 1.f - h.u - h.v;
                    f3 tg = add(add(scl(ld3(T0), w0), scl(ld3(T0 + 4), h.u)), scl(ld3(T0 + 8), h.v));
                    float sg = (T0[3] * w0 + T0[7] * h.u + T0[11] * h.v) < 0.f ? -1.f : 1.f;
                    if (XI) { tg = xf_v(XI + 12, tg); sg *= XI[26]; }   /* tangents: forward rows */
                    tg = sub(tg, scl(ns, dot3(ns, tg)));
                    float tl = sqrtf(dot3(tg, tg));
                    if (tl > 0.f) {
                        tg = scl(tg, 1.f / tl);
                        f3 bt = scl(crs(ns, tg), sg);
                        f4 c = tex_sample(s, X[2], tu, tv);
                        float sc = i2f(X[7]);
                        f3 nm = add(add(scl(tg, (2.f * c.x - 1.f) * sc), scl(bt, (2.f * c.y - 1.f) * sc)), scl(ns, 2.f * c.z - 1.f));
                        float nl = sqrtf(dot3(nm, nm));
                        if (nl > 0.f) { ns = scl(nm, 1.f / nl); mapped = 1; }
                    }
                }
            }
        }
        int front = dot3(ng0, d) < 0.f;
        f3 ng = front ? ng0 : neg(ng0);
        if (!front) { ns = neg(ns); nsU = neg(nsU); }
        if (emissive(&M)) {
            float w = 1.f;
            if (lastPdf > 0.f) {
                if (s->light_mode == 1) w = 0.f;
                else if (s->light_mode == 0) {
                    float cl = fabsf(dot3(ng0, d));
                    float pa = XI ? s->mat_light_pdf[f2i(P[7])] : s->tri_light_pdf[h.tri];
                    float pl = cl > 0.f ? s->p_tri * pa * h.t * h.t / cl : 0.f;
                    w = mis2(lastPdf, pl);
                }
            }
            L = add(L, scl(mul(T, M.emis), w));
        }
        if (depth > p->max_bounce) break;
        uint32_t dd = (uint32_t)depth + 1u;
        if (s->rr && depth >= s->rr_depth) {
            float mx = fmaxf(T.x, fmaxf(T.y, T.z));
            float qs = fminf(fmaxf(mx, 0.05f), 1.f);
            if (orc_rng4((uint32_t)pix, smp, dd, 6, seed) >= qs) break;
            T = scl(T, 1.f / qs);
        }
        f3 wo = neg(d);
        if (dot3(ns, wo) <= 1e-6f) ns = ng;
        f3 tx, ty; onb(ns, &tx, &ty);
        f3 woL = mk(dot3(wo, tx), dot3(wo, ty), dot3(wo, ns));
        f3 nrL = mk(0.f, 0.f, 1.f);   /* unperturbed normal in the local frame (+z: terminator shadowing off) */
        if (s->bump_shadow && mapped) {
            nrL = mk(dot3(nsU, tx), dot3(nsU, ty), dot3(nsU, ns));
            if (!(nrL.z > 0.f)) nrL = mk(0.f, 0.f, 1.f);
        }
        float eta = front ? M.ior : 1.f / M.ior;
        f3 poUp = add(ph, scl(ng, p->eps)), poDn = sub(ph, scl(ng, p->eps));
        float ux1 = orc_rng4((uint32_t)pix, smp, dd, 7, seed), ux2 = orc_rng4((uint32_t)pix, smp, dd, 7, seed ^ 0x68E31DA4u);
        if (has_non_delta(&M)) {
            float us = orc_rng4((uint32_t)pix, smp, dd, 5, seed);
            float u2 = orc_rng4((uint32_t)pix, smp, dd, 2, seed), u3 = orc_rng4((uint32_t)pix, smp, dd, 3, seed);
            f3 l = mk(0.f, 0.f, 1.f), Le = mk(0.f, 0.f, 0.f);
            float pl = 0.f, tmax = 1e30f;
            int isSun = 0, ok = 0, side = 0;
            if (us < s->p_sun) { l = cone(sunTo, p->sun_cos, u2, u3); Le = sunE; pl = s->p_sun; isSun = 1; ok = 1; }
            else if (us < s->p_sun + s->p_tri) {
                float ut = fminf((us - s->p_sun) / s->p_tri, ORC_ONE_MINUS);
                int em = cdf_find(s->em_cdf, s->num_em, ut);
                const float* A;
                f3 f1, f2;
                int kt = -1;
                if (s->tlas) { A = s->em_tris + 12 * (size_t)em; f1 = ld3(A + 4); f2 = ld3(A + 8); }
                else { kt = s->em_tri[em]; A = s->corners + 12 * (size_t)kt; f1 = sub(ld3(A + 4), ld3(A)); f2 = sub(ld3(A + 8), ld3(A)); }
                float su = sqrtf(u2);
                f3 pL = add(add(ld3(A), scl(f1, su * (1.f - u3))), scl(f2, su * u3));
                f3 nL = nrm3(crs(f1, f2));
                side = dot3(ng, sub(pL, ph)) > 0.f ? 1 : -1;
                f3 org = side > 0 ? poUp : poDn;
                f3 dv = sub(pL, org);
                float dist2 = dot3(dv, dv), dist = sqrtf(dist2);
                if (dist > 0.f) {
                    l = scl(dv, 1.f / dist);
                    float cl = fabsf(dot3(nL, l));
                    if (cl > 0.f) {
                        pl = s->p_tri * (s->tlas ? s->mat_light_pdf[f2i(A[7])] : s->tri_light_pdf[kt]) * dist2 / cl;
                        const int lm = f2i(A[7]);
                        Le = ld_mat(s->mats + 16 * lm).emis;
                        if (s->maps && lm < s->num_maps && s->maps[8 * lm + 1] >= 0) {   /* emission map at the sample */
                            float tu, tv;
                            tri_uv(s, s->tlas ? f2i(A[3]) : kt, su * (1.f - u3), su * u3, &tu, &tv);
                            f4 c = tex_sample(s, s->maps[8 * lm + 1], tu, tv);
                            Le = mul(Le, mk(c.x, c.y, c.z));
                        }
                        tmax = dist * (1.f - 1e-3f);
                        ok = 1;
                    }
                }
            } else if (s->p_env > 0.f && s->env_uniform) {
                float z = 1.f - 2.f * u2, rr = sqrtf(fmaxf(0.f, 1.f - z * z)), phi = ORC_TWO_PI * u3;
                l = mk(rr * cosf(phi), rr * sinf(phi), z); pl = s->p_env * (0.25f * ORC_INV_PI); Le = env_rad(s, l); ok = 1;
            } else if (s->p_env > 0.f) {
                float ue = fminf((us - s->p_sun - s->p_tri) / s->p_env, ORC_ONE_MINUS);
                int x, y;
                float pe = env_sample(s, ue, u2, u3, &l, &x, &y);
                if (pe > 0.f) { pl = s->p_env * pe; Le = ld3(s->env_tex + 4 * ((size_t)y * s->env_w + x)); ok = 1; }
            }
            if (ok) {
                f3 lL = mk(dot3(l, tx), dot3(l, ty), dot3(l, ns));
                float gl = dot3(ng, l);
                if (lL.z != 0.f && (gl > 0.f) == (lL.z > 0.f)) {
                    bval e = mat_eval(&M, woL, lL, eta, s->gd_energy, ux1, ux2, nrL);
                    float w = (isSun || s->light_mode == 1) ? 1.f : mis2(pl, e.pdf);
                    f3 con = scl(mul(mul(T, e.f), Le), fabsf(lL.z) * w / pl);
                    if (con.x > 0.f || con.y > 0.f || con.z > 0.f) {
                        f3 so = (side ? side > 0 : gl > 0.f) ? poUp : poDn;
                        ray_t sr = mkray(so, l, 0.f, s->tri_test);
                        hit_t sh = { tmax, 0.f, 0.f, -1, -1 };
                        st[1] += 1;
                        if (!trace_one(s, &sr, &sh, 1, NULL, NULL)) { st[3] += 1; L = add(L, con); }
                    }
                }
            }
        }
        f3 wiL, wgt; float pdf = 0.f;
        if (!mat_sample(&M, woL, eta, s->gd_energy, ux1, ux2, orc_rng4((uint32_t)pix, smp, dd, 4, seed),
                        orc_rng4((uint32_t)pix, smp, dd, 0, seed), orc_rng4((uint32_t)pix, smp, dd, 1, seed), &wiL, &wgt, &pdf, nrL))
            break;
        f3 wi = nrm3(add(add(scl(tx, wiL.x), scl(ty, wiL.y)), scl(ns, wiL.z)));
        float gi = dot3(ng, wi);
        if (gi == 0.f || (gi > 0.f) != (wiL.z > 0.f)) break;
        T = mul(T, wgt);
        if (!(fmaxf(T.x, fmaxf(T.y, T.z)) > 0.f)) break;
        o = gi > 0.f ? poUp : poDn; d = wi; lastPdf = pdf;
    }
    return L;
}

typedef struct { const orc_scene* s; const orc_params* p; int spp, sbase, batch, bounces; float* accum; double st[64][4]; } render_job;
static void render_worker(void* a, int tid, int nth) {
    render_job* J = (render_job*)a;
    const int N = J->p->w * J->p->h;
    double* st = J->st[tid];
    for (int pix = tid; pix < N; pix += nth) {
        float* acc = J->accum + 4 * (size_t)pix;
        for (int s0 = J->sbase; s0 < J->sbase + J->spp; s0 += J->batch) {
            float r = 0.f, g = 0.f, b = 0.f;
            for (int s = s0; s < s0 + J->batch; s++) {
                f3 L = J->s->full ? trace_path_full(J->s, J->p, pix, (uint32_t)s, J->bounces, st)
                                  : trace_path(J->s, J->p, pix, (uint32_t)s, J->bounces, st);
                r += L.x; g += L.y; b += L.z;
            }
            acc[0] += r; acc[1] += g; acc[2] += b;
        }
    }
}
int orc_render(const orc_scene* s, const orc_params* p, int spp, int sample_base, int batch, int bounces, int threads,
               int clear, float* accum, double* stats) {
    if (batch < 1 || spp % batch || threads > 64) return 1;
    if (threads < 1) threads = 1;
    if (clear) memset(accum, 0, sizeof(float) * 4 * (size_t)p->w * p->h);
    render_job* J = (render_job*)calloc(1, sizeof(render_job));
    J->s = s; J->p = p; J->spp = spp; J->sbase = sample_base; J->batch = batch; J->bounces = bounces; J->accum = accum;
    run_threads(render_worker, J, threads);
    for (int t = 0; t < threads; t++) for (int k = 0; k < 4; k++) stats[k] += J->st[t][k];
    free(J);
    return 0;
}
