/* hploc.c - CPU restatement of the reference's GPU BVH builder (H-PLOC, Benthin et al. 2024) and of the conversion
 * of its output into a SAH binary tree (plain C11, IEEE, -ffp-contract=off).
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the product's HIP builder. Only tests/ use it.
 *
 * Restated from (paths relative to the reference repo):
 *   orc_hp_centroid_box  src/accel/hploc.cu:11-26   centroid box (exact min / max of (lo + hi) * 0.5)
 *   orc_expand_bits21,   src/accel/hploc.cu:27-42   Morton code, mb bits per axis, x << 2 | y << 1 | z, then << ib |
 *   orc_hp_keys                                     primitive index; ib = ceil(log2 n), mb = min(21, (64 - ib) / 3)
 *   (sort)               src/accel/hploc.cu:168,184 CUB radix sort; the keys are unique, so any sort gives the same order
 *   orc_hp_build         src/accel/hploc.cu:43-136  leaves, bottom-up LBVH by key deltas with first-arrival hand-over,
 *                                                   PLOC merges with threshold T = 16 and radius R = 8 in the lane
 *                                                   order of the reference's 32-lane merge
 *   orc_hp_to_bvh2       src/accel/hploc.cu:138-160 64 B BVH2 layout of the H-PLOC nodes
 *   orc_tree_from_hploc  src/accel/bvh_tools.cpp:12-38 SAH leaf collapse (C = min(A p, A + C_l + C_r)), pre-order tree
 *
 * The merge order of parallel threads decides only the ids of internal nodes (an atomic counter), not the tree: the
 * clusters of an LBVH node, their order and every merge decision are functions of the sorted keys and boxes. This
 * restatement numbers nodes in its own sequential order; tests compare trees in a canonical pre-order form.
 *
 * Pins (tests/test_oracle.py): expandBits21 closed form; keys against a direct bit-by-bit interleave; sorted keys
 * strictly increasing; the tree holds 2n - 1 nodes, every primitive once, parent box = union of child boxes; for
 * n <= 32 (one PLOC window at the root) every merge joins a mutually nearest pair in the window; the SAH collapse
 * against a brute-force recursion on small trees.
 */
#include "oracle.h"
#include <stdlib.h>
#include <string.h>

#define HP_T 16
#define HP_R 8

static float hp_f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t hp_u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static float hp_fi(int i) { float f; memcpy(&f, &i, 4); return f; }
static int hp_if(float f) { int i; memcpy(&i, &f, 4); return i; }

void orc_hp_centroid_box(const float* lo, const float* hi, int n, float box[6]) {
    for (int a = 0; a < 3; a++) { box[a] = 3.4e38f; box[3 + a] = -3.4e38f; }
    for (int i = 0; i < n; i++)
        for (int a = 0; a < 3; a++) {
            const float c = (lo[4 * i + a] + hi[4 * i + a]) * 0.5f;
            if (c < box[a]) box[a] = c;
            if (c > box[3 + a]) box[3 + a] = c;
        }
}

uint64_t orc_expand_bits21(uint32_t v) {
    uint64_t x = v & 0x1fffffu, r = 0;
    for (int b = 0; b < 21; b++) r |= ((x >> b) & 1u) << (3 * b);
    return r;
}

void orc_hp_key_bits_mode(int n, int morton, int* mb, int* ib) {
    int b = 1;
    while ((1ll << b) < n) b++;
    *ib = b;
    /* adaptive (the reference, src/accel/hploc.cu:181-183): as many bits per axis as fit beside the index, at most 21;
       morton == 30: the classic 30-bit code, 10 bits per axis */
    *mb = morton == 30 ? 10 : ((64 - b) / 3 < 21 ? (64 - b) / 3 : 21);
}
void orc_hp_key_bits(int n, int* mb, int* ib) { orc_hp_key_bits_mode(n, 0, mb, ib); }

void orc_hp_keys(const float* lo, const float* hi, int n, const float box[6], int mb, int ib, uint64_t* keys) {
    const float s = (float)(1u << mb), smax = s - 1.f;
    for (int i = 0; i < n; i++) {
        uint32_t q[3];
        for (int a = 0; a < 3; a++) {
            const float c = (lo[4 * i + a] + hi[4 * i + a]) * 0.5f, e = box[3 + a] - box[a];
            const float t = e > 0.f ? (c - box[a]) / e : 0.f;
            float v = t * s;
            v = v > 0.f ? v : 0.f;
            v = v < smax ? v : smax;
            q[a] = (uint32_t)v;
        }
        const uint64_t m = (orc_expand_bits21(q[0]) << 2) | (orc_expand_bits21(q[1]) << 1) | orc_expand_bits21(q[2]);
        keys[i] = (m << ib) | (uint64_t)i;
    }
}

static int cmp_u64(const void* a, const void* b) {
    const uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

static float half_area(const float* lo, const float* hi) {
    const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    return dx * dy + dy * dz + dz * dx;
}

/* PLOC on the clusters of LBVH node [L, R] split at s, in the lane order of a 32-lane merge */
static void hp_merge(int L, int R, int s, int root, float* nodes, int* C, int* counter) {
    int nL = s - L + 1 < HP_T ? s - L + 1 : HP_T, nR = R - s < HP_T ? R - s : HP_T;
    int slot[32], id[32], num = 0;
    float lo[32][3], hi[32][3];
    for (int lane = 0; lane < 32; lane++) {
        slot[lane] = lane < nL ? L + lane : (lane < nL + nR ? s + 1 + (lane - nL) : -1);
        const int c = slot[lane] >= 0 ? C[slot[lane]] : -1;
        if (c >= 0) id[num++] = c;
    }
    for (int k = 0; k < num; k++)
        for (int a = 0; a < 3; a++) { lo[k][a] = nodes[8 * id[k] + a]; hi[k][a] = nodes[8 * id[k] + 4 + a]; }
    const int target = root ? 1 : HP_T;
    while (num > target) {
        uint64_t best[32];
        for (int lane = 0; lane < num; lane++) best[lane] = ~0ull;
        for (int lane = 0; lane < num; lane++)
            for (int r = 1; r <= HP_R && lane + r < num; r++) {
                float ml[3], mh[3];
                for (int a = 0; a < 3; a++) {
                    ml[a] = lo[lane][a] < lo[lane + r][a] ? lo[lane][a] : lo[lane + r][a];
                    mh[a] = hi[lane][a] > hi[lane + r][a] ? hi[lane][a] : hi[lane + r][a];
                }
                const uint64_t key = ((uint64_t)hp_u(half_area(ml, mh)) << 32) | ((uint64_t)lane << 5) | (uint64_t)(lane + r);
                if (key < best[lane]) best[lane] = key;
                if (key < best[lane + r]) best[lane + r] = key;
            }
        int nn[32];
        for (int lane = 0; lane < num; lane++) {
            const int pa = (int)((best[lane] >> 5) & 31u), pb = (int)(best[lane] & 31u);
            nn[lane] = pa == lane ? pb : pa;
        }
        int nid = *counter;
        int kid[32], kn = 0;
        float klo[32][3], khi[32][3];
        for (int lane = 0; lane < num; lane++) {
            const int p = nn[lane];
            const int mutual = p != lane && nn[p] == lane;
            if (mutual && lane < p) {   /* the lower lane leads: new node (lower, upper) */
                float* nd = nodes + 8 * (size_t)nid;
                for (int a = 0; a < 3; a++) {
                    lo[lane][a] = lo[lane][a] < lo[p][a] ? lo[lane][a] : lo[p][a];
                    hi[lane][a] = hi[lane][a] > hi[p][a] ? hi[lane][a] : hi[p][a];
                    nd[a] = lo[lane][a]; nd[4 + a] = hi[lane][a];
                }
                nd[3] = hp_fi(id[lane]); nd[7] = hp_fi(id[p]);
                id[lane] = nid++;
            }
        }
        /* merged boxes above were written in lane order, before any partner reads them: partners of later leads are
         * always upper lanes that are not leads themselves, so their boxes are unchanged */
        for (int lane = 0; lane < num; lane++) {
            const int p = nn[lane];
            const int mutual = p != lane && nn[p] == lane;
            if (mutual && lane > p) continue;
            kid[kn] = id[lane];
            for (int a = 0; a < 3; a++) { klo[kn][a] = lo[lane][a]; khi[kn][a] = hi[lane][a]; }
            kn++;
        }
        *counter = nid;
        num = kn;
        for (int k = 0; k < num; k++) {
            id[k] = kid[k];
            for (int a = 0; a < 3; a++) { lo[k][a] = klo[k][a]; hi[k][a] = khi[k][a]; }
        }
    }
    for (int lane = 0; lane < 32; lane++)
        if (slot[lane] >= 0) C[slot[lane]] = -1;
    for (int k = 0; k < num; k++) C[L + k] = id[k];
}

int orc_hp_build(const float* plo, const float* phi, int n, float* nodes, uint64_t* sorted_keys) {
    return orc_hp_build_mode(plo, phi, n, 0, nodes, sorted_keys);
}
int orc_hp_build_mode(const float* plo, const float* phi, int n, int morton, float* nodes, uint64_t* sorted_keys) {
    if (n < 1) return -1;
    float box[6];
    int mb, ib;
    orc_hp_centroid_box(plo, phi, n, box);
    orc_hp_key_bits_mode(n, morton, &mb, &ib);
    uint64_t* k = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)n);
    orc_hp_keys(plo, phi, n, box, mb, ib, k);
    qsort(k, (size_t)n, sizeof(uint64_t), cmp_u64);
    if (sorted_keys) memcpy(sorted_keys, k, sizeof(uint64_t) * (size_t)n);
    int* C = (int*)malloc(sizeof(int) * (size_t)n);
    int* bounds = (int*)malloc(sizeof(int) * (size_t)n);
    const uint64_t imask = (1ull << ib) - 1ull;
    for (int i = 0; i < n; i++) {   /* leaves in Morton order */
        const int p = (int)(k[i] & imask);
        float* nd = nodes + 8 * (size_t)i;
        for (int a = 0; a < 3; a++) { nd[a] = plo[4 * p + a]; nd[4 + a] = phi[4 * p + a]; }
        nd[3] = hp_fi(p); nd[7] = hp_fi(-1);
        C[i] = i; bounds[i] = -1;
    }
    int counter = n;
    /* each primitive's thread climbs until it is the first to reach a node; the second continues with the union range */
    for (int i = 0; i < n && n > 1; i++) {
        int L = i, R = i;
        for (;;) {
            int s, prev;
            if (L == 0 || (R != n - 1 && (k[R] ^ k[R + 1]) < (k[L - 1] ^ k[L]))) {
                s = R; prev = bounds[R]; bounds[R] = L;
                if (prev != -1) R = prev;
            } else {
                s = L - 1; prev = bounds[L - 1]; bounds[L - 1] = R;
                if (prev != -1) L = prev;
            }
            if (prev == -1) break;
            const int root = L == 0 && R == n - 1;
            if (R - L + 1 > HP_T || root) hp_merge(L, R, s, root, nodes, C, &counter);
            if (root) break;
        }
    }
    free(k); free(C); free(bounds);
    (void)hp_if;
    (void)hp_f;
    return n > 1 ? 2 * n - 2 : 0;
}

void orc_hp_to_bvh2(const float* hn, int n, float* out) {
    const float E = 3e37f;
    if (n == 1) {
        const float* a = hn; const float* b = hn + 4;
        float* o = out;
        o[0] = a[0]; o[1] = b[0]; o[2] = a[1]; o[3] = b[1];
        o[4] = E; o[5] = E; o[6] = E; o[7] = E;
        o[8] = a[2]; o[9] = b[2]; o[10] = E; o[11] = E;
        o[12] = a[3]; o[13] = hp_fi(-1); o[14] = hp_fi(1); o[15] = hp_fi(0);
        return;
    }
    for (int j = 0; j < n - 1; j++) {
        const float* nd = hn + 8 * (size_t)(n + j);
        const int ch[2] = { hp_if(nd[3]), hp_if(nd[7]) };
        const float *cl[2], *chh[2];
        int c[2], cnt[2];
        for (int q = 0; q < 2; q++) {
            const int x = ch[q];
            cl[q] = hn + 8 * (size_t)x; chh[q] = cl[q] + 4;
            if (x < n) { c[q] = hp_if(cl[q][3]); cnt[q] = 1; } else { c[q] = x - n; cnt[q] = 0; }
        }
        float* o = out + 16 * (size_t)j;
        o[0] = cl[0][0]; o[1] = chh[0][0]; o[2] = cl[0][1]; o[3] = chh[0][1];
        o[4] = cl[1][0]; o[5] = chh[1][0]; o[6] = cl[1][1]; o[7] = chh[1][1];
        o[8] = cl[0][2]; o[9] = chh[0][2]; o[10] = cl[1][2]; o[11] = chh[1][2];
        o[12] = hp_fi(c[0]); o[13] = hp_fi(c[1]); o[14] = hp_fi(cnt[0]); o[15] = hp_fi(cnt[1]);
    }
}

static double box_area6(const float* lo, const float* hi) {
    const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    return 2.0 * (double)(dx * dy + dy * dz + dz * dx);
}

int orc_tree_from_hploc(const float* hn, int n, int root, int leaf_max, orc_tree* out) {
    const int tot = 2 * n - 1;
    int *lc = (int*)malloc(sizeof(int) * tot), *rc = (int*)malloc(sizeof(int) * tot), *cnt = (int*)malloc(sizeof(int) * tot);
    double* cost = (double*)malloc(sizeof(double) * tot);
    char* leaf = (char*)malloc((size_t)tot);
    for (int h = 0; h < tot; h++) {   /* children have smaller ids than their parent */
        const float* nd = hn + 8 * (size_t)h;
        lc[h] = hp_if(nd[3]); rc[h] = hp_if(nd[7]);
        const double A = box_area6(nd, nd + 4);
        if (h < n) { cnt[h] = 1; cost[h] = A; leaf[h] = 1; continue; }
        cnt[h] = cnt[lc[h]] + cnt[rc[h]];
        const double ci = A + cost[lc[h]] + cost[rc[h]], cl = cnt[h] <= leaf_max ? A * cnt[h] : 1e300;
        leaf[h] = cl <= ci;
        cost[h] = cl < ci ? cl : ci;
    }
    memset(out, 0, sizeof *out);
    out->node = (int*)malloc(sizeof(int) * 4 * (size_t)tot);
    out->box = (float*)malloc(sizeof(float) * 6 * (size_t)tot);
    out->prim = (int*)malloc(sizeof(int) * (size_t)n);
    int* st = (int*)malloc(sizeof(int) * 3 * (size_t)(tot + 1));   /* (h, parent slot, depth) */
    int* gs = (int*)malloc(sizeof(int) * (size_t)(tot + 1));
    int sp = 0, nn = 0, np = 0;
    st[0] = root; st[1] = -1; st[2] = 0; sp = 1;
    while (sp) {
        sp--;
        const int h = st[3 * sp], ps = st[3 * sp + 1], d = st[3 * sp + 2];
        const int id = nn++;
        const float* nd = hn + 8 * (size_t)h;
        for (int a = 0; a < 3; a++) { out->box[6 * id + a] = nd[a]; out->box[6 * id + 3 + a] = nd[4 + a]; }
        out->node[4 * id] = -1; out->node[4 * id + 1] = -1; out->node[4 * id + 2] = 0; out->node[4 * id + 3] = 0;
        if (ps >= 0) out->node[4 * (ps >> 1) + (ps & 1)] = id;
        if (d > out->max_depth) out->max_depth = d;
        if (leaf[h]) {
            const int first = np;
            int gp = 0;
            gs[gp++] = h;
            while (gp) {
                const int x = gs[--gp];
                if (x < n) out->prim[np++] = lc[x];
                else { gs[gp++] = rc[x]; gs[gp++] = lc[x]; }
            }
            out->node[4 * id + 2] = first; out->node[4 * id + 3] = np - first;
        } else {
            st[3 * sp] = rc[h]; st[3 * sp + 1] = 2 * id + 1; st[3 * sp + 2] = d + 1; sp++;
            st[3 * sp] = lc[h]; st[3 * sp + 1] = 2 * id; st[3 * sp + 2] = d + 1; sp++;
        }
    }
    out->num_nodes = nn; out->num_prims = np;
    free(lc); free(rc); free(cnt); free(cost); free(leaf); free(st); free(gs);
    return 0;
}
