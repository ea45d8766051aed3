/*
 * bh_oracle.c — CPU restatement of the reference's Barnes-Hut step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never the product.
 *
 * Arithmetic contract: every FP32 expression keeps the GLSL operation order,
 * uses IEEE division and square root, and contracts no multiply-add
 * (compiled with -ffp-contract=off). Against the reference GLSL executed by
 * Mesa llvmpipe (the .npz fixtures in tests/golden) this reproduces bounds, Morton codes,
 * the permutation, the tree, the fill and the integrated bodies bit for bit.
 *
 * How each function is pinned (tests/test_oracle_*.py):
 *  - bho_expand_bits21 / bho_morton3d: brute force over all 2^21 inputs
 *    against a per-bit interleave loop; llvmpipe fixtures.
 *  - bho_morton_encode, bho_aabb: llvmpipe fixtures (bit-exact); AABB min/max
 *    also against numpy min/max (order-free, exact).
 *  - bho_dead_partition: numpy stable partition; llvmpipe fixtures with dead bodies.
 *  - bho_sort: numpy stable argsort; llvmpipe fixtures.
 *  - bho_safe_lcp / bho_build: Karras-2012 invariants (one root, every node two
 *    children, parent/child agreement, contiguous ranges, split = highest
 *    differing bit found by brute force on small N); llvmpipe fixtures.
 *  - bho_fill: closed form per node (mass sum, mass-weighted COM in float64,
 *    AABB union, height); llvmpipe fixtures bit-exact.
 *  - bho_force: theta = 0 reduces to a direct sum in walk order (float64
 *    comparison); llvmpipe acceleration fixtures; opened-node counter.
 *  - bho_merge / bho_merge_pair: closed form (mass, momentum, volume
 *    conservation); llvmpipe fixtures for isolated pairs.
 *  - bho_star_color: llvmpipe fixtures within a tolerance (Mesa's pow/log
 *    are approximations, so colours are compared to 1e-3).
 *  - bho_units: parity unpinned at the bit level (Mesa's pow is an
 *    approximation); tests feed llvmpipe's G bits where bits matter.
 */
#include "bh_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define EMPTY_IDX 0xFFFFFFFFu

static int finite3(const float *v) {
    return !(isnan(v[0]) || isnan(v[1]) || isnan(v[2]) || isinf(v[0]) || isinf(v[1]) || isinf(v[2]));
}
/* GLSL min/max as llvmpipe executes them: IEEE minNum/maxNum (a NaN operand
 * yields the other operand), pinned by the NaN fixture */
static float gmin(float a, float b) { return fminf(a, b); }
static float gmax(float a, float b) { return fmaxf(a, b); }

/* common/common.glsl:346-352 */
void bho_units(float mass, float density, float len, float time, float *G, float *blc) {
    const float GC = 6.67430e-11f;
    *G = GC * powf(len, -3.0f) * powf(time, 2.0f) * powf(mass, 1.0f);
    *blc = 0.6203504909f * powf(mass / density, 1.0f / 3.0f) / len;
}

/* ------------------------------------------------------------ dead */
/* shaders/compute/dead/bh_dead.comp:13-112 */
uint32_t bho_dead_partition(float *src, float *dst, const uint32_t *in, uint32_t num, uint32_t *out) {
    uint32_t dead = 0;
    for (uint32_t g = 0; g < num; g++) if (src[(size_t)in[g] * 8 + 3] == 0.0f) dead++;
    uint32_t a = 0, d = num - dead;
    for (uint32_t g = 0; g < num; g++) {
        uint32_t b = in[g];
        if (src[(size_t)b * 8 + 3] == 0.0f) {
            memset(src + (size_t)b * 8, 0, 32);
            if (dst) memset(dst + (size_t)b * 8, 0, 32);
            out[d++] = b;
        } else {
            out[a++] = b;
        }
    }
    return dead;
}

/* ------------------------------------------------------------ AABB */
/* bh_morton.comp:50-74: point box of a finite inlier, or nothing */
static int contributor(const float *pos, const float *prev, float *lo_hi_out) {
    if (!finite3(pos)) return 0;
    if (prev[0] > prev[3]) {                       /* emptyAABB(prev) */
        memcpy(lo_hi_out, pos, 12); memcpy(lo_hi_out + 3, pos, 12);
        return 1;
    }
    for (int k = 0; k < 3; k++) {
        float center = 0.5f * (prev[k] + prev[3 + k]);
        float half = gmax(0.5f * (prev[3 + k] - prev[k]) * 1.05f, 1e-9f);
        float lo = center - half, hi = center + half;
        if (pos[k] < lo || pos[k] > hi) return 0;
    }
    memcpy(lo_hi_out, pos, 12); memcpy(lo_hi_out + 3, pos, 12);
    return 1;
}

/* bh_morton.comp:114-231: repopulate + collapse + growth clamp */
void bho_aabb(const float *bodies, const uint32_t *index, uint32_t num, float *bounds) {
    float prev[6];
    memcpy(prev, bounds, sizeof prev);
    float res[6] = {1e38f, 1e38f, 1e38f, -1e38f, -1e38f, -1e38f};
    for (uint32_t g = 0; g < num; g++) {
        const float *b = bodies + (size_t)index[g] * 8;
        if (b[3] == 0.0f) continue;
        float box[6];
        if (!contributor(b, prev, box)) continue;
        for (int k = 0; k < 3; k++) { res[k] = gmin(res[k], box[k]); res[3 + k] = gmax(res[3 + k], box[3 + k]); }
    }
    if (res[0] > res[3]) {                         /* no finite inliers */
        if (prev[0] > prev[3]) memcpy(bounds, res, sizeof res);
        return;
    }
    if (prev[0] > prev[3]) { memcpy(bounds, res, sizeof res); return; }
    for (int k = 0; k < 3; k++) {                  /* clampAABBExtentGrowth, :77-88 */
        float prevExt = gmax(prev[3 + k] - prev[k], 1e-9f);
        float maxExt = prevExt * 1.1f;
        float c = 0.5f * (res[k] + res[3 + k]);
        float h = 0.5f * (res[3 + k] - res[k]);
        float mh = 0.5f * maxExt;
        h = gmin(h, mh);
        bounds[k] = c - h;
        bounds[3 + k] = c + h;
    }
}

/* ------------------------------------------------------------ Morton */
/* bh_morton.comp:6-15 */
uint64_t bho_expand_bits21(uint32_t v) {
    uint64_t x = (uint64_t)v & 0x1FFFFFull;
    x = (x | (x << 32)) & 0x1F00000000FFFFull;
    x = (x | (x << 16)) & 0x1F0000FF0000FFull;
    x = (x | (x << 8)) & 0x100F00F00F00F00Full;
    x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}
/* bh_morton.comp:18-21 */
uint64_t bho_morton3d(uint32_t x, uint32_t y, uint32_t z) {
    return (bho_expand_bits21(x) << 2) | (bho_expand_bits21(y) << 1) | bho_expand_bits21(z);
}

/* bh_morton.comp:24-34,238-262 */
void bho_morton_encode(const float *bodies, const uint32_t *index, uint32_t num, const float *bounds, uint64_t *codes) {
    const float MAXV = 2097151.0f;
    int empty = bounds[0] > bounds[3];
    for (uint32_t g = 0; g < num; g++) {
        const float *b = bodies + (size_t)index[g] * 8;
        float p[3], q[3];
        if (!finite3(b) || empty) {
            for (int k = 0; k < 3; k++) p[k] = 0.5f * (bounds[k] + bounds[3 + k]);
        } else {
            for (int k = 0; k < 3; k++) p[k] = gmin(gmax(b[k], bounds[k]), bounds[3 + k]);
        }
        uint32_t qi[3];
        for (int k = 0; k < 3; k++) {
            float ext = gmax(bounds[3 + k] - bounds[k], 1e-9f);
            float pn = (p[k] - bounds[k]) / ext;
            q[k] = gmin(gmax(floorf(pn * MAXV), 0.0f), MAXV);
            qi[k] = (uint32_t)floorf(q[k]);
        }
        codes[g] = bho_morton3d(qi[0], qi[1], qi[2]);
    }
}

/* ------------------------------------------------------------ sort */
/* stable LSD radix sort, 16 bits per pass (same order as 16 passes of 4 bits) */
void bho_sort(uint64_t *keys, uint32_t *index, uint32_t num) {
    if (num < 2) return;
    uint64_t *k2 = malloc((size_t)num * 8);
    uint32_t *i2 = malloc((size_t)num * 4);
    size_t *cnt = malloc(65536 * sizeof(size_t));
    uint64_t *ks = keys, *kd = k2;
    uint32_t *is = index, *id = i2;
    for (int pass = 0; pass < 4; pass++) {
        int sh = pass * 16;
        memset(cnt, 0, 65536 * sizeof(size_t));
        for (uint32_t g = 0; g < num; g++) cnt[(ks[g] >> sh) & 0xFFFF]++;
        size_t s = 0;
        for (int d = 0; d < 65536; d++) { size_t c = cnt[d]; cnt[d] = s; s += c; }
        for (uint32_t g = 0; g < num; g++) {
            size_t dpos = cnt[(ks[g] >> sh) & 0xFFFF]++;
            kd[dpos] = ks[g]; id[dpos] = is[g];
        }
        uint64_t *tk = ks; ks = kd; kd = tk;
        uint32_t *ti = is; is = id; id = ti;
    }
    /* four passes: the result is back in keys/index */
    free(k2); free(i2); free(cnt);
}

/* ------------------------------------------------------------ build */
static int clz64(uint64_t x) { return x ? __builtin_clzll(x) : 64; }
static int clz32(uint32_t x) { return x ? __builtin_clz(x) : 32; }

/* bh_tree.comp:6-37 */
int bho_safe_lcp(const uint64_t *keys, uint32_t num, int i, int j) {
    if (i < 0 || j < 0 || i >= (int)num || j >= (int)num) return -1;
    uint64_t a = keys[i], b = keys[j];
    if (a == b) {
        if (i == j) return 64;
        return 64 + clz32((uint32_t)i ^ (uint32_t)j);
    }
    return clz64(a ^ b);
}

static void default_aabb(float *a) {
    a[0] = a[1] = a[2] = 1e38f;
    a[3] = a[4] = a[5] = -1e38f;
}

/* bh_tree.comp:43-133 */
void bho_build(const uint64_t *keys, uint32_t num, uint32_t n_init, bho_node *leaf, bho_node *internal) {
    if (num <= 1) {
        if (num == 1 || n_init >= 2) {
            bho_node *r = &internal[0];
            r->parentId = EMPTY_IDX; r->childA = 0; r->childB = EMPTY_IDX; r->readyChildren = 1;
            memset(r->comMass, 0, 16); default_aabb(r->aabb); r->nodeDepth = 0; r->bodiesContained = 0;
            memset(r->avgColor, 0, 16);
            leaf[0].parentId = n_init;
        }
        return;
    }
    for (int i = 0; i < (int)num - 1; i++) {
        int lcpR = bho_safe_lcp(keys, num, i, i + 1);
        int lcpL = bho_safe_lcp(keys, num, i, i - 1);
        int dir = (lcpL > lcpR) ? -1 : 1;
        int dmin = bho_safe_lcp(keys, num, i, i - dir);
        int lmax = 2;
        while (bho_safe_lcp(keys, num, i, i + dir * lmax) > dmin) lmax *= 2;
        int l = 0, t = lmax / 2;
        while (t > 0) {
            if (bho_safe_lcp(keys, num, i, i + dir * (l + t)) > dmin) l += t;
            t /= 2;
        }
        int j = i + l * dir;
        int dnode = bho_safe_lcp(keys, num, i, j);
        int s = 0;
        t = l;
        while (t > 1) {
            t = (t + 1) / 2;
            if (bho_safe_lcp(keys, num, i, i + (s + t) * dir) > dnode) s += t;
        }
        int gamma = i + s * dir + (dir < 0 ? dir : 0);
        int lo = i < j ? i : j, hi = i < j ? j : i;
        uint32_t left = (lo == gamma) ? (uint32_t)gamma : (uint32_t)gamma + n_init;
        uint32_t right = (hi == gamma + 1) ? (uint32_t)(gamma + 1) : (uint32_t)(gamma + 1) + n_init;
        bho_node *nd = &internal[i];
        uint32_t gidx = (uint32_t)i + n_init;
        nd->childA = left; nd->childB = right; nd->readyChildren = 0;
        memset(nd->comMass, 0, 16); default_aabb(nd->aabb); memset(nd->avgColor, 0, 16);
        if (left < n_init) leaf[left].parentId = gidx; else internal[left - n_init].parentId = gidx;
        if (right < n_init) leaf[right].parentId = gidx; else internal[right - n_init].parentId = gidx;
        nd->nodeDepth = (uint32_t)lo;
        nd->bodiesContained = (uint32_t)(hi - lo + 1);
        if (i == 0) nd->parentId = EMPTY_IDX;
    }
}

/* ------------------------------------------------------------ colour */
/* common/common.glsl:235-276 */
static float star_temp(float mass, float density) {
    const float c = 5.95589e-19f;
    return c * powf(mass, 0.875f - 1.0f / 6.0f) * powf(density, 1.0f / 6.0f);
}
static float clampf(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }
void bho_star_color(float mass, float density, float rgb[3]) {
    float temp = star_temp(mass, density) / 100.0f;
    float r, g, b;
    if (temp <= 66.0f) r = 255.0f;
    else { r = temp - 60.0f; r = 329.698727446f * powf(r, -0.1332047592f); }
    if (temp <= 66.0f) g = 99.4708025861f * logf(gmax(temp, 1e-3f)) - 161.1195681661f;
    else { g = temp; g = 288.1221695283f * powf(g, -0.0755148492f); }
    if (temp >= 66.0f) b = 255.0f;
    else { b = temp - 10.0f; b = 138.5177312231f * logf(gmax(b, 1e-3f)) - 305.0447927307f; }
    rgb[0] = clampf(r, 0.0f, 255.0f) / 255.0f;
    rgb[1] = clampf(g, 0.0f, 255.0f) / 255.0f;
    rgb[2] = clampf(b, 0.0f, 255.0f) / 255.0f;
}

/* ------------------------------------------------------------ fill */
static const bho_node *get_node(const bho_node *leaf, const bho_node *internal, uint32_t n_init, uint32_t idx) {
    return idx < n_init ? &leaf[idx] : &internal[idx - n_init];
}

/* bh_tree.comp:183-259, one internal node from its two children */
static void combine(bho_node *nd, const bho_node *L, const bho_node *R) {
    float mL = L->comMass[3], mR = R->comMass[3];
    float M = mL + mR;
    float com[3], col[3];
    if (M > 0.0f) {
        for (int k = 0; k < 3; k++) {
            float a = mL * L->comMass[k];
            float b = mR * R->comMass[k];
            com[k] = (a + b) / M;
            float ca = mL * L->avgColor[k];
            float cb = mR * R->avgColor[k];
            col[k] = (ca + cb) / M;
        }
    } else {
        for (int k = 0; k < 3; k++) {
            com[k] = (L->comMass[k] + R->comMass[k]) * 0.5f;
            col[k] = (L->avgColor[k] + R->avgColor[k]) * 0.5f;
        }
    }
    nd->comMass[0] = com[0]; nd->comMass[1] = com[1]; nd->comMass[2] = com[2]; nd->comMass[3] = M;
    for (int k = 0; k < 3; k++) {
        nd->aabb[k] = gmin(L->aabb[k], R->aabb[k]);
        nd->aabb[3 + k] = gmax(L->aabb[3 + k], R->aabb[3 + k]);
    }
    nd->avgColor[0] = col[0]; nd->avgColor[1] = col[1]; nd->avgColor[2] = col[2]; nd->avgColor[3] = 0.0f;
    nd->readyChildren = EMPTY_IDX;
    uint32_t dl = L->nodeDepth, dr = R->nodeDepth;
    nd->nodeDepth = 1u + (dl > dr ? dl : dr);
}

/* bh_tree.comp:140-175 (leaves) and :183-259 (levels); returns the height */
uint32_t bho_fill(const float *bodies, const uint32_t *index, uint32_t num, uint32_t n_init,
                  float unit_mass, float unit_density, int max_passes, int with_colors,
                  bho_node *leaf, bho_node *internal) {
    for (uint32_t g = 0; g < num; g++) {
        const float *b = bodies + (size_t)index[g] * 8;
        bho_node *l = &leaf[g];
        l->comMass[0] = b[0]; l->comMass[1] = b[1]; l->comMass[2] = b[2]; l->comMass[3] = b[3];
        l->aabb[0] = b[0]; l->aabb[1] = b[1]; l->aabb[2] = b[2];
        l->aabb[3] = b[0]; l->aabb[4] = b[1]; l->aabb[5] = b[2];
        l->childA = l->childB = EMPTY_IDX;
        l->readyChildren = EMPTY_IDX;
        l->nodeDepth = 0;
        if (with_colors && b[3] > 0.0f && b[7] > 0.0f) {
            float rgb[3];
            bho_star_color(b[3] * unit_mass, b[7] * unit_density, rgb);
            l->avgColor[0] = rgb[0]; l->avgColor[1] = rgb[1]; l->avgColor[2] = rgb[2]; l->avgColor[3] = 0.0f;
        } else {
            memset(l->avgColor, 0, 16);
        }
    }
    if (num < 2) return 0;
    /* heights by an explicit post-order walk from the root (internal 0) */
    uint32_t nint = num - 1;
    uint32_t *height = calloc(nint, 4);
    uint32_t *stack = malloc((size_t)2 * nint * 4 + 16);
    uint8_t *state = calloc(nint, 1);
    size_t sp = 0;
    stack[sp++] = 0;
    while (sp) {
        uint32_t k = stack[sp - 1];
        const bho_node *nd = &internal[k];
        if (!state[k]) {
            state[k] = 1;
            if (nd->childA >= n_init) stack[sp++] = nd->childA - n_init;
            if (nd->childB >= n_init) stack[sp++] = nd->childB - n_init;
            continue;
        }
        sp--;
        uint32_t ha = nd->childA >= n_init ? height[nd->childA - n_init] : 0;
        uint32_t hb = nd->childB >= n_init ? height[nd->childB - n_init] : 0;
        height[k] = 1 + (ha > hb ? ha : hb);
    }
    uint32_t tree_h = height[0];
    /* fill level by level (pass p handles height p+1), as the queue does */
    uint32_t maxh = tree_h;
    if (max_passes > 0 && (uint32_t)max_passes < maxh) maxh = (uint32_t)max_passes;
    uint32_t **lev = calloc(maxh + 1, sizeof(uint32_t *));
    uint32_t *levn = calloc(maxh + 1, 4);
    for (uint32_t k = 0; k < nint; k++) if (height[k] <= maxh) levn[height[k]]++;
    for (uint32_t h = 1; h <= maxh; h++) { lev[h] = malloc((size_t)levn[h] * 4 + 4); levn[h] = 0; }
    for (uint32_t k = 0; k < nint; k++) if (height[k] <= maxh) lev[height[k]][levn[height[k]]++] = k;
    for (uint32_t h = 1; h <= maxh; h++) {
        for (uint32_t q = 0; q < levn[h]; q++) {
            bho_node *nd = &internal[lev[h][q]];
            combine(nd, get_node(leaf, internal, n_init, nd->childA), get_node(leaf, internal, n_init, nd->childB));
        }
    }
    /* nodes above the cap: readyChildren counts the children that completed */
    for (uint32_t k = 0; k < nint; k++) {
        if (height[k] <= maxh) continue;
        const bho_node *nd = &internal[k];
        uint32_t c = 0;
        c += (nd->childA < n_init || height[nd->childA - n_init] <= maxh);
        c += (nd->childB < n_init || height[nd->childB - n_init] <= maxh);
        internal[k].readyChildren = c;
    }
    for (uint32_t h = 1; h <= maxh; h++) free(lev[h]);
    free(lev); free(levn); free(height); free(stack); free(state);
    return tree_h;
}

/* ------------------------------------------------------------ walk */
/* GLSL dot(a, b) as llvmpipe evaluates it: a.x*b.x + (a.y*b.y + a.z*b.z),
 * no contraction (pinned bit-exactly by the llvmpipe acceleration fixtures) */
static float dot3(const float *a, const float *b) {
    float yz = a[1] * b[1] + a[2] * b[2];
    return a[0] * b[0] + yz;
}

static float body_radius(const float *b, float blc) {   /* common/common.glsl:228-230 */
    return blc * powf(b[3] / b[7], 1.0f / 3.0f);
}

/* bh_force.comp:40-142 for the body at sorted position g */
static void walk_one(uint32_t g, const float *src, float *dst, const uint32_t *index, uint32_t n_init,
                     const bho_node *leaf, const bho_node *internal, const bho_params *p,
                     float G, float blc, const float *bounds, uint32_t *stack, int stack_cap,
                     float *accel_out, uint32_t *inter_out, uint32_t *tasks, uint32_t task_cap,
                     uint32_t *n_tasks, bho_walk_stats *st) {
    uint32_t bodyIdx = index[g];
    float body[8];
    memcpy(body, src + (size_t)bodyIdx * 8, 32);
    float acc[3] = {0.0f, 0.0f, 0.0f};
    uint32_t sp = 0;
    uint32_t opened = 0, inter = 0, tests = 0, maxsp = 1;
    stack[sp++] = n_init;
    while (sp > 0) {
        uint32_t nodeIdx = stack[--sp];
        const bho_node *nd = get_node(leaf, internal, n_init, nodeIdx);
        float r[3] = {nd->comMass[0] - body[0], nd->comMass[1] - body[1], nd->comMass[2] - body[2]};
        float d2 = dot3(r, r) + p->softening;
        float inv = 1.0f / sqrtf(d2);
        if (nd->childA != EMPTY_IDX) {
            tests++;
            float ex = nd->aabb[3] - nd->aabb[0], ey = nd->aabb[4] - nd->aabb[1], ez = nd->aabb[5] - nd->aabb[2];
            float longest = gmax(ex, gmax(ey, ez)) / 2.0f;
            if (longest * inv < p->theta) {
                float gm = G * nd->comMass[3];
                for (int k = 0; k < 3; k++) acc[k] = acc[k] + gm * r[k] * inv * inv * inv;
                inter++;
            } else {
                if ((int)sp + 2 > stack_cap) { st->max_stack = 0xFFFFFFFFu; break; }
                stack[sp++] = nd->childA;
                stack[sp++] = nd->childB;
                if (sp > maxsp) maxsp = sp;
                opened++;
            }
        } else {
            uint32_t other = index[nodeIdx];
            if (other == bodyIdx) continue;
            float gm = G * nd->comMass[3];
            for (int k = 0; k < 3; k++) acc[k] = acc[k] + gm * r[k] * inv * inv * inv;
            inter++;
            if (p->merging == 0) continue;
            const float *ob = src + (size_t)other * 8;
            float ri = body_radius(body, blc), rj = body_radius(ob, blc);
            float dist = sqrtf(dot3(r, r));
            float pen = ri + rj - dist;
            if (p->merging == 2 && pen > 0.0f) {
                float dv[3] = {ob[4] - body[4], ob[5] - body[5], ob[6] - body[6]};
                float il = 1.0f / sqrtf(dot3(r, r));
                float nrm[3] = {r[0] * il, r[1] * il, r[2] * il};
                float vimp = dot3(dv, nrm);
                if (vimp < 0.0f) {
                    float meff = 1.0f / (1.0f / body[3] + 1.0f / ob[3]);
                    float imp = (1.0f + p->elasticity) * meff * vimp;
                    for (int k = 0; k < 3; k++) body[4 + k] += nrm[k] * imp / body[3];
                }
                float cs = (pen / (body[3] + ob[3])) * p->restitution;
                for (int k = 0; k < 3; k++) body[k] -= cs * nrm[k];
            }
            if (p->merging == 1 && pen > 0.0f && bodyIdx < other) {
                uint32_t slot;
#pragma omp atomic capture
                slot = (*n_tasks)++;
                if (slot < task_cap) { tasks[2 * slot] = bodyIdx; tasks[2 * slot + 1] = other; }
            }
        }
    }
    if (accel_out) { accel_out[3 * g] = acc[0]; accel_out[3 * g + 1] = acc[1]; accel_out[3 * g + 2] = acc[2]; }
    if (inter_out) inter_out[g] = inter;
#pragma omp atomic
    st->opened += opened;
#pragma omp atomic
    st->interactions += inter;
#pragma omp atomic
    st->mac_tests += tests;
    if (st->max_stack != 0xFFFFFFFFu && maxsp > st->max_stack) {
#pragma omp critical
        { if (st->max_stack != 0xFFFFFFFFu && maxsp > st->max_stack) st->max_stack = maxsp; }
    }
    if (!dst) return;
    float nv[3], np[3];
    for (int k = 0; k < 3; k++) { nv[k] = body[4 + k] + acc[k] * p->dt; np[k] = body[k] + nv[k] * p->dt; }
    if (p->dynamic == 0) {
        if (p->wrap_around) {
            for (int k = 0; k < 3; k++) {
                float x = np[k] - bounds[k], y = bounds[3 + k] - bounds[k];
                float m = x - y * floorf(x / y);
                np[k] = m + bounds[k];
            }
        } else if (body[0] < bounds[0] || body[0] > bounds[3] || body[1] < bounds[1] || body[1] > bounds[4] ||
                   body[2] < bounds[2] || body[2] > bounds[5]) {
            memset(dst + (size_t)bodyIdx * 8, 0, 32);
#pragma omp atomic
            st->out_of_bounds++;
            return;
        }
    }
    float *o = dst + (size_t)bodyIdx * 8;
    o[4] = nv[0]; o[5] = nv[1]; o[6] = nv[2]; o[7] = body[7];
    o[0] = np[0]; o[1] = np[1]; o[2] = np[2]; o[3] = body[3];
}

static int cmp_pair(const void *a, const void *b) {
    const uint32_t *x = a, *y = b;
    if (x[0] != y[0]) return x[0] < y[0] ? -1 : 1;
    if (x[1] != y[1]) return x[1] < y[1] ? -1 : 1;
    return 0;
}

int bho_force(const float *src, float *dst, const uint32_t *index, uint32_t num, uint32_t n_init,
              const bho_node *leaf, const bho_node *internal, const bho_params *p,
              float G, float blc, const float *bounds, int stack_cap,
              float *accel, uint32_t *interactions, uint32_t *tasks, uint32_t task_cap,
              bho_walk_stats *st, int threads) {
    return bho_force_range(src, dst, index, num, n_init, leaf, internal, p, G, blc, bounds, stack_cap, accel,
                           interactions, tasks, task_cap, st, threads, 0, num);
}

static int force_impl(const float *src, float *dst, const uint32_t *index, uint32_t num, uint32_t n_init,
                      const bho_node *leaf, const bho_node *internal, const bho_params *p,
                      float G, float blc, const float *bounds, int stack_cap,
                      float *accel, uint32_t *interactions, uint32_t *tasks, uint32_t task_cap,
                      bho_walk_stats *st, int threads, uint32_t g_begin, uint32_t g_end) {
    memset(st, 0, sizeof *st);
    if (g_end > num) g_end = num;
    if (stack_cap < 2) stack_cap = 2;
    uint32_t nt = 0;
    if (num <= 1) {
        /* a single body feels no force (the reference reads past its one-node tree) */
        for (uint32_t g = 0; g < num; g++) {
            uint32_t b = index[g];
            const float *s = src + (size_t)b * 8;
            if (accel) accel[0] = accel[1] = accel[2] = 0.0f;
            if (dst) {
                float *o = dst + (size_t)b * 8;
                for (int k = 0; k < 3; k++) { o[4 + k] = s[4 + k]; o[k] = s[k] + o[4 + k] * p->dt; }
                o[3] = s[3]; o[7] = s[7];
            }
        }
        return 0;
    }
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
#pragma omp parallel
    {
        uint32_t *stack = malloc((size_t)stack_cap * 4);
#pragma omp for schedule(dynamic, 256)
        for (uint32_t g = g_begin; g < g_end; g++)
            walk_one(g, src, dst, index, n_init, leaf, internal, p, G, blc, bounds, stack, stack_cap,
                     accel, interactions, tasks, task_cap, &nt, st);
        free(stack);
    }
    st->n_tasks = nt < task_cap ? nt : task_cap;
    st->task_overflow = nt > task_cap ? nt - task_cap : 0;
    if (tasks && st->n_tasks > 1) qsort(tasks, st->n_tasks, 8, cmp_pair);
    return st->max_stack == 0xFFFFFFFFu ? -1 : 0;
}

int bho_force_range(const float *src, float *dst, const uint32_t *index, uint32_t num, uint32_t n_init,
                    const bho_node *leaf, const bho_node *internal, const bho_params *p,
                    float G, float blc, const float *bounds, int stack_cap,
                    float *accel, uint32_t *interactions, uint32_t *tasks, uint32_t task_cap,
                    bho_walk_stats *st, int threads, uint32_t g_begin, uint32_t g_end) {
    return force_impl(src, dst, index, num, n_init, leaf, internal, p, G, blc, bounds, stack_cap, accel,
                      interactions, tasks, task_cap, st, threads, g_begin, g_end);
}

/* The corrected walk (no reference counterpart) for the bodies at sorted
 * positions [0, num), in groups of `group` consecutive positions (64: one GPU
 * wave, the GPU kernel's rule; 1: every body on its own):
 * a group walks the tree in the reference's order (childB before childA) and
 * takes an internal node when every body of the group accepts it, opening it
 * for all of them otherwise; a leaf is taken by every body but its own. A
 * body accepts node k, in FP32 with r = com - position, s^2 = fma(x, x, fma(y,
 * y, fma(z, z, eps^2))) and b^2 = fma(dx, dx, fma(dy, dy, dz dz)), d = max(com
 * - min, max - com) (b: distance from the centre of mass to the farthest box
 * corner), when
 *   geometric (amag == NULL): fl(s^2 theta^2) > b^2;
 *   relative: s^2 > b^2 and fl(s^2 A) > K, A = (float)(alpha amag[g])^0.4 and
 *   K = (float)((G M)^0.4 (b^2)^0.6) in FP64 (G M the FP32 product);
 * the GPU kernel's exact sequence, so both take the same nodes. Every node
 * taken adds, in FP64 from those FP32 inputs,
 *     G M r / s^3 + (7.5 (r.GT.r) / s^2 - 1.5 G tr T) r / s^5 - 3 GT r / s^5
 * with quad[8k .. 8k+6] = G T (xx xy xz yy yz zz), -1.5 G tr T for internal
 * node k (a leaf: its monopole). Outputs per sorted position: accel[3g..]
 * the FP64 sum, absum[g] the sum over the nodes taken of |G M r / s^3| +
 * |quadrupole term| (the scale of the FP32 rounding a GPU sum may carry),
 * inter[g] the nodes taken. The stats get the opened nodes (per body), the
 * quadrupole terms (in mac_tests) and the deepest stack. Pinned by
 * tests/test_oracle_quadrupole.py (exact field of clusters, the error falling
 * one order faster than the monopole's, the taken set against an independent
 * top-down restatement). */
int bho_force_corrected(const bho_node *leaf, const bho_node *internal, uint32_t num, uint32_t n_init,
                        const float *quad, float theta, float softening, float G, const float *amag, float alpha,
                        uint32_t group, int stack_cap, double *accel, double *absum, uint32_t *inter,
                        bho_walk_stats *st, int threads) {
    memset(st, 0, sizeof *st);
    if (group < 1 || group > 64) return -2;
    if (stack_cap < 2) stack_cap = 2;
    for (uint32_t g = 0; g < num; g++) {
        accel[3 * g] = accel[3 * g + 1] = accel[3 * g + 2] = 0.0;
        absum[g] = 0.0;
        inter[g] = 0;
    }
    if (num < 2) return 0;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    const float th2 = theta * theta;
    const uint32_t groups = (num + group - 1) / group;
    int overflow = 0;
    uint64_t opened_all = 0, quads_all = 0;
    uint32_t maxsp_all = 0;
#pragma omp parallel reduction(+ : opened_all, quads_all) reduction(max : maxsp_all) reduction(| : overflow)
    {
        uint32_t *stack = malloc((size_t)stack_cap * 4);
        float afac[64];
        int ok[64];
#pragma omp for schedule(dynamic, 16)
        for (uint32_t gi = 0; gi < groups; gi++) {
            const uint32_t g0 = group * gi, cnt = num - g0 < group ? num - g0 : group;
            for (uint32_t j = 0; j < cnt; j++)
                afac[j] = amag ? (float)pow((double)alpha * (double)amag[g0 + j], 0.4) : th2;
            uint32_t sp = 0, maxsp = 1;
            stack[sp++] = n_init;
            while (sp > 0) {
                const uint32_t id = stack[--sp];
                const int internal_node = id >= n_init;
                const bho_node *nd = get_node(leaf, internal, n_init, id);
                float b2 = -1.0f, K = -1.0f;
                if (internal_node) {
                    const float *c = nd->comMass, *bb = nd->aabb;
                    const float dx = gmax(c[0] - bb[0], bb[3] - c[0]);
                    const float dy = gmax(c[1] - bb[1], bb[4] - c[1]);
                    const float dz = gmax(c[2] - bb[2], bb[5] - c[2]);
                    b2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                    const float gm = G * nd->comMass[3];
                    K = (float)(pow((double)gm, 0.4) * pow((double)b2, 0.6));
                    int all = 1;
                    for (uint32_t j = 0; j < cnt; j++) {
                        const float *p = leaf[g0 + j].comMass;
                        const float rx = c[0] - p[0], ry = c[1] - p[1], rz = c[2] - p[2];
                        const float d2f = fmaf(rx, rx, fmaf(ry, ry, fmaf(rz, rz, softening)));
                        ok[j] = amag ? (d2f > b2 && d2f * afac[j] > K) : (d2f * afac[j] > b2);
                        all &= ok[j];
                    }
                    if (!all) {
                        if ((int)sp + 2 > stack_cap) { overflow = 1; break; }
                        stack[sp++] = nd->childA;
                        stack[sp++] = nd->childB;
                        if (sp > maxsp) maxsp = sp;
                        opened_all += cnt;
                        continue;
                    }
                }
                for (uint32_t j = 0; j < cnt; j++) {
                    const uint32_t g = g0 + j;
                    if (!internal_node && id == g) continue;
                    const float *p = leaf[g].comMass;
                    const double r[3] = {(float)(nd->comMass[0] - p[0]), (float)(nd->comMass[1] - p[1]),
                                         (float)(nd->comMass[2] - p[2])};
                    const double s2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2] + (double)softening;
                    const double inv = 1.0 / sqrt(s2);
                    const double i2 = inv * inv, i3 = i2 * inv, i5 = i3 * i2;
                    const double gm = (double)(G * nd->comMass[3]);
                    double term[3], mono = 0.0, qabs = 0.0;
                    for (int k = 0; k < 3; k++) { term[k] = gm * i3 * r[k]; mono += term[k] * term[k]; }
                    if (internal_node) {
                        const float *Q = quad + 8 * (size_t)(id - n_init);
                        const double t[3] = {Q[0] * r[0] + Q[1] * r[1] + Q[2] * r[2],
                                             Q[1] * r[0] + Q[3] * r[1] + Q[4] * r[2],
                                             Q[2] * r[0] + Q[4] * r[1] + Q[5] * r[2]};
                        const double rtr = r[0] * t[0] + r[1] * t[1] + r[2] * t[2];
                        const double cq = (7.5 * rtr * i2 + (double)Q[6]) * i5;
                        for (int k = 0; k < 3; k++) {
                            const double qk = cq * r[k] - 3.0 * i5 * t[k];
                            term[k] += qk;
                            qabs += qk * qk;
                        }
                        quads_all++;
                    }
                    for (int k = 0; k < 3; k++) accel[3 * g + k] += term[k];
                    absum[g] += sqrt(mono) + sqrt(qabs);
                    inter[g]++;
                }
            }
            if (maxsp > maxsp_all) maxsp_all = maxsp;
        }
        free(stack);
    }
    st->opened = opened_all;
    st->mac_tests = quads_all;
    st->max_stack = maxsp_all;
    st->interactions = 0;
    for (uint32_t g = 0; g < num; g++) st->interactions += inter[g];
    return overflow ? -1 : 0;
}

/* Second moment of the bodies of each internal node about its centre of
 * mass, FP64, by direct summation over the node's leaves (sorted positions
 * [lo, lo + bodiesContained), lo its leftmost leaf): m2[6k..6k+5] = xx, xy,
 * xz, yy, yz, zz. No reference counterpart; pinned by closed forms (two and
 * three bodies) in tests/test_oracle_quadrupole.py. */
void bho_node_moments(const bho_node *leaf, const bho_node *internal, uint32_t num, uint32_t n_init, double *m2,
                      int threads) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    if (num < 2) return;
#pragma omp parallel for schedule(dynamic, 1024)
    for (uint32_t k = 0; k < num - 1; k++) {
        uint32_t lo = n_init + k;
        while (lo >= n_init) lo = internal[lo - n_init].childA;
        const bho_node *nd = &internal[k];
        double c[3] = {nd->comMass[0], nd->comMass[1], nd->comMass[2]}, acc[6] = {0, 0, 0, 0, 0, 0};
        for (uint32_t j = lo; j < lo + nd->bodiesContained; j++) {
            const float *b = leaf[j].comMass;
            double d[3] = {b[0] - c[0], b[1] - c[1], b[2] - c[2]}, m = b[3];
            acc[0] += m * d[0] * d[0]; acc[1] += m * d[0] * d[1]; acc[2] += m * d[0] * d[2];
            acc[3] += m * d[1] * d[1]; acc[4] += m * d[1] * d[2]; acc[5] += m * d[2] * d[2];
        }
        for (int q = 0; q < 6; q++) m2[6 * (size_t)k + q] = acc[q];
    }
}

/* ------------------------------------------------------------ merge */
/* bh_merge.comp:5-23 */
void bho_merge_pair(const float *b1, const float *b2, float *o) {
    float m = b1[3] + b2[3];
    if (m == 0.0f) { memset(o, 0, 32); return; }
    if (b1[3] == 0.0f) { memcpy(o, b2, 32); return; }
    if (b2[3] == 0.0f) { memcpy(o, b1, 32); return; }
    for (int k = 0; k < 3; k++) {
        float a = b1[k] * b1[3];
        float b = b2[k] * b2[3];
        o[k] = (a + b) / m;
    }
    o[3] = m;
    for (int k = 0; k < 3; k++) {
        float a = b1[4 + k] * b1[3];
        float b = b2[4 + k] * b2[3];
        o[4 + k] = (a + b) / m;
    }
    o[7] = (b1[3] + b2[3]) / (b1[3] / b1[7] + b2[3] / b2[7]);
}

/* bh_merge.comp:30-71 with the tasks resolved in ascending (a, b) order */
uint32_t bho_merge(float *dst, uint32_t *tasks, uint32_t n) {
    if (n > 1) qsort(tasks, n, 8, cmp_pair);
    uint32_t applied = 0;
    for (uint32_t t = 0; t < n; t++) {
        uint32_t a = tasks[2 * t], b = tasks[2 * t + 1];
        if (a == b) continue;
        float *A = dst + (size_t)a * 8, *B = dst + (size_t)b * 8;
        if (A[3] != 0.0f && B[3] != 0.0f) {
            float o[8];
            bho_merge_pair(A, B, o);
            memcpy(A, o, 32);
            memset(B, 0, 32);
            applied++;
        }
    }
    return applied;
}

/* ------------------------------------------------------------ step */
int bho_state_init(bho_state *s, uint32_t n, const float *bodies, const float *bounds,
                   float um, float ud, float ul, float ut) {
    memset(s, 0, sizeof *s);
    if (n < 1) return -1;
    s->n_init = n;
    for (int k = 0; k < 2; k++) {
        s->bodies[k] = calloc((size_t)n, 32);
        s->index[k] = calloc((size_t)n, 4);
    }
    s->morton = calloc((size_t)n, 8);
    s->leaf = calloc((size_t)n, sizeof(bho_node));
    s->internal = calloc((size_t)(n > 1 ? n - 1 : 1), sizeof(bho_node));
    s->task_cap = n;
    s->tasks = calloc((size_t)n, 8);
    memcpy(s->bodies[0], bodies, (size_t)n * 32);
    for (uint32_t i = 0; i < n; i++) s->index[0][i] = i;      /* bh_init.comp:20 */
    s->sim.numBodies = n; s->sim.initialNumBodies = n;
    memcpy(s->sim.minCorner, bounds, 12); memcpy(s->sim.maxCorner, bounds + 3, 12);
    s->sim.unitMass = um; s->sim.unitDensity = ud; s->sim.unitLen = ul; s->sim.unitTime = ut;
    bho_units(um, ud, ul, ut, &s->sim.gravitationalConstant, &s->sim.bodyLengthConstant);
    s->cur = 0;
    s->fill_cap = 64;
    s->stack_cap = 64;
    s->threads = 0;
    return 0;
}

void bho_state_free(bho_state *s) {
    for (int k = 0; k < 2; k++) { free(s->bodies[k]); free(s->index[k]); }
    free(s->morton); free(s->leaf); free(s->internal); free(s->tasks);
    memset(s, 0, sizeof *s);
}

/* simulation/BarnesHut.java:99-164 */
int bho_step(bho_state *s, const bho_params *p) {
    float *src = s->bodies[s->cur], *dst = s->bodies[1 - s->cur];
    uint32_t num = s->sim.numBodies;
    /* partitionDeadBodies + swap + decrementDeadBodies */
    uint32_t died = bho_dead_partition(src, dst, s->index[0], num, s->index[1]);
    uint32_t *t = s->index[0]; s->index[0] = s->index[1]; s->index[1] = t;
    num -= died;
    s->sim.numBodies = num;
    s->sim.justDied = 0;
    s->sim.uintDebug[0] = 0;
    float bounds[6];
    memcpy(bounds, s->sim.minCorner, 12); memcpy(bounds + 3, s->sim.maxCorner, 12);
    if (p->dynamic == 1) {
        bho_aabb(src, s->index[0], num, bounds);
        memcpy(s->sim.minCorner, bounds, 12); memcpy(s->sim.maxCorner, bounds + 3, 12);
    }
    bho_morton_encode(src, s->index[0], num, bounds, s->morton);
    bho_sort(s->morton, s->index[0], num);
    bho_build(s->morton, num, s->n_init, s->leaf, s->internal);
    s->last_height = bho_fill(src, s->index[0], num, s->n_init, s->sim.unitMass, s->sim.unitDensity,
                              s->fill_cap, 1, s->leaf, s->internal);
    bho_walk_stats st;
    int rc = bho_force(src, dst, s->index[0], num, s->n_init, s->leaf, s->internal, p,
                       s->sim.gravitationalConstant, s->sim.bodyLengthConstant, bounds, s->stack_cap,
                       NULL, NULL, s->tasks, s->task_cap, &st, s->threads);
    s->sim.uintDebug[0] = (uint32_t)st.opened;
    s->sim.outOfBounds += st.out_of_bounds;
    if (p->merging == 1) bho_merge(dst, s->tasks, st.n_tasks);
    s->last = st;
    s->cur = 1 - s->cur;
    return rc;
}

/* -------------------------------------------------------- FP64 direct sum */
void bho_direct_sum_f64(const float *bodies, uint32_t n, const uint32_t *sample, uint32_t n_sample,
                        double G, double soft, double *out, int threads) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
#pragma omp parallel for schedule(dynamic, 16)
    for (uint32_t s = 0; s < n_sample; s++) {
        uint32_t i = sample[s];
        double xi = bodies[(size_t)i * 8], yi = bodies[(size_t)i * 8 + 1], zi = bodies[(size_t)i * 8 + 2];
        double ax = 0, ay = 0, az = 0;
        for (uint32_t j = 0; j < n; j++) {
            if (j == i) continue;
            const float *b = bodies + (size_t)j * 8;
            if (b[3] == 0.0f) continue;
            double rx = b[0] - xi, ry = b[1] - yi, rz = b[2] - zi;
            double d2 = rx * rx + ry * ry + rz * rz + soft;
            double inv = 1.0 / sqrt(d2);
            double f = G * (double)b[3] * inv * inv * inv;
            ax += f * rx; ay += f * ry; az += f * rz;
        }
        out[3 * s] = ax; out[3 * s + 1] = ay; out[3 * s + 2] = az;
    }
}
