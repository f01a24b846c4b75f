/*
 * sbgi_oracle.c — TEST INFRASTRUCTURE ONLY: the CPU oracle of the DDGI probe update.
 *
 * Built by oracle/Makefile into oracle/_build/liboracle.so. Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it, and only as the
 * checker; the product library (speedball_amd/libsbgi.so) never links or calls it.
 *
 *   or64_*  FP64 oracle: the reference algorithm evaluated in double precision.
 *   or32_*  FP32 restatement: the same operation order as the reference's TSL graphs
 *           in single precision (built with -ffp-contract=off, no fast-math).
 *   or_refit_*  bounds refits (exact min/max, precision-independent).
 *
 * Restated reference code: js/gi_oct.js:22-85, js/gi_probes.js:1692-1726, 1876-1885,
 * 1941-2214, 2317-2532, 2737-2846, js/spectral_traverse.js:116-229, 354-569,
 * js/gi_refit.js:34-430. Each function below cites its lines.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/sbgi.h"

#define GOLDEN_ANGLE 2.39996322972865332 /* pi (3 - sqrt 5), js/gi_probes.js:106 */
#define TWO_PI 6.283185307179586
#define PI_D 3.141592653589793 /* spectral_traverse.js:25 */
#define RAY_EPS 1e-3           /* spectral_traverse.js:35 */
#define T_MAX 1e30             /* spectral_traverse.js:34 */

/* IEEE binary16 round-to-nearest-even of a float, returned as float: the RGBA16F
 * textureStore of js/gi_probes.js:2838-2844. */
float half_round(float f) {
    uint32_t x;
    memcpy(&x, &f, 4);
    uint32_t sign = x & 0x80000000u, ax = x & 0x7FFFFFFFu;
    float o;
    if (ax >= 0x7F800000u) return f;
    if (ax >= 0x477FF000u) { /* >= 65520 rounds to infinity */
        uint32_t r = sign | 0x7F800000u;
        memcpy(&o, &r, 4);
        return o;
    }
    if (ax < 0x33000000u) { /* < 2^-25 rounds to zero */
        memcpy(&o, &sign, 4);
        return o;
    }
    int e = (int)(ax >> 23) - 127;
    int shift = e < -14 ? 13 + (-14 - e) : 13;
    uint32_t m = (ax & 0x7FFFFFu) | 0x800000u;
    uint32_t q = 1u << shift, rem = m & (q - 1), base = m & ~(q - 1), hq = q >> 1;
    if (rem > hq || (rem == hq && (base & q))) base += q;
    o = (float)ldexp((double)base, e - 23);
    return sign ? -o : o;
}

/* ---------------- FP64 oracle ---------------- */
/* Direction mode of the FP64 oracle: with f32 directions on, rayDir returns the FP32
 * restatement's direction (phi = k GA evaluated in f32, as the reference's WGSL does),
 * widened to double, and everything downstream stays FP64. Comparing the FP64 oracle in
 * both modes splits the error of an FP32 path into the part the f32 direction definition
 * causes and the part its arithmetic causes. */
void or32_ray_dir(uint32_t k, uint32_t rays, float jitter, float out[3]);
static int g_f32_dirs = 0;
void or64_set_f32_dirs(int on) { g_f32_dirs = on; }
static int ray_dir_f32_hook(uint32_t k, uint32_t rays, float jitter, double out[3]) {
    if (!g_f32_dirs) return 0;
    float f[3];
    or32_ray_dir(k, rays, jitter, f);
    out[0] = f[0];
    out[1] = f[1];
    out[2] = f[2];
    return 1;
}
#define RAY_DIR_HOOK ray_dir_f32_hook
#define REAL double
#define P(n) or64_##n
#define RABS fabs
#define RSQRT sqrt
#define RFLOOR floor
#define RSIN sin
#define RCOS cos
#define RSINCR sin
#define RCOSCR cos
#define RPOW pow
#define REXP exp
#define RFMIN fmin
#define RFMAX fmax
#define RGAUSS(jx, jy) exp(-(double)((jx) * (jx) + (jy) * (jy)) * 0.5)
#include "oracle_impl.h"
#undef REAL
#undef P
#undef RABS
#undef RSQRT
#undef RFLOOR
#undef RSIN
#undef RCOS
#undef RSINCR
#undef RCOSCR
#undef RPOW
#undef REXP
#undef RFMIN
#undef RFMAX
#undef RGAUSS
#undef VSTRIDE

#undef RAY_DIR_HOOK
#undef ROBUST_FAR

/* ---------------- FP32 restatement ---------------- */
#define REAL float
#define P(n) or32_##n
#define RABS fabsf
#define RSQRT sqrtf
#define RFLOOR floorf
#define RSIN sinf
#define RCOS cosf
#define RSINCR(x) ((float)sin((double)(x)))
#define RCOSCR(x) ((float)cos((double)(x)))
#define RPOW powf
#define REXP expf
#define RFMIN fminf
#define RFMAX fmaxf
#define RGAUSS(jx, jy) (float)exp(-(double)((jx) * (jx) + (jy) * (jy)) * 0.5)
#include "oracle_impl.h"

/* ---------------- refits (exact; precision-independent) ---------------- */

/* refitFlatBlasRange / createFlatBlasRefitStepper js/gi_refit.js:34-165: reverse
 * pre-order sweep over [root, end). Returns 1, or 0 when the layout is invalid (the
 * caller treats that as "rebuild required"). nodes is u32[8*N] with f32 bounds. */
int or_refit_blas(uint32_t *nodes, uint32_t node_count, const uint32_t *tri_index, uint32_t tri_count,
                  const float *vdata, uint32_t vert_count, uint32_t root, uint32_t end) {
    if (!(end > root) || end > node_count) return 0;
    float *nf = (float *)nodes;
    for (int64_t i = (int64_t)end - 1; i >= (int64_t)root; i--) {
        uint32_t *nd = nodes + (size_t)i * 8;
        float *ndf = nf + (size_t)i * 8;
        uint32_t info = nd[7];
        if (info != 0xFFFFFFFFu) {
            uint32_t off = info & 0x00FFFFFFu, cnt = info >> 24;
            if (cnt == 0) continue; /* empty leaf keeps its stale bounds (gi_refit.js:77-85) */
            float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
            for (uint32_t t = off; t < off + cnt; t++) {
                if ((uint64_t)t * 3 + 2 >= (uint64_t)tri_count * 3) return 0;
                for (int k = 0; k < 3; k++) {
                    uint32_t v = tri_index[(size_t)t * 3 + k];
                    if (v >= vert_count) return 0;
                    const float *p = vdata + (size_t)v * 8;
                    for (int a = 0; a < 3; a++) {
                        if (p[a] < mn[a]) mn[a] = p[a];
                        if (p[a] > mx[a]) mx[a] = p[a];
                    }
                }
            }
            for (int a = 0; a < 3; a++) {
                ndf[a] = mn[a];
                ndf[3 + a] = mx[a];
            }
        } else {
            uint64_t l = (uint64_t)i + 1;
            if (l >= end) return 0;
            uint32_t r = nodes[l * 8 + 6];
            if (!(r > l && r < end)) return 0;
            const float *lb = nf + l * 8, *rb = nf + (size_t)r * 8;
            for (int a = 0; a < 3; a++) {
                ndf[a] = fminf(lb[a], rb[a]);
                ndf[3 + a] = fmaxf(lb[3 + a], rb[3 + a]);
            }
        }
    }
    return 1;
}

/* refitFlatTlasRange js/gi_refit.js:245-292 over 12-float TLAS nodes and 6-float
 * instance world bounds per TLAS slot. Returns 1, or 0 on an invalid layout. */
int or_refit_tlas(float *tlas, uint32_t node_count, const float *inst_bounds, uint32_t inst_count) {
    for (int64_t i = (int64_t)node_count - 1; i >= 0; i--) {
        float *nd = tlas + (size_t)i * 12;
        float cnt = nd[8];
        float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
        if (cnt > 0) {
            float off = nd[7];
            if (off != floorf(off) || cnt != floorf(cnt) || off < 0 || cnt < 1 || off + cnt > (float)inst_count)
                return 0;
            for (uint32_t s = (uint32_t)off; s < (uint32_t)(off + cnt); s++) {
                const float *b = inst_bounds + (size_t)s * 6;
                for (int a = 0; a < 3; a++) {
                    if (b[a] < mn[a]) mn[a] = b[a];
                    if (b[3 + a] > mx[a]) mx[a] = b[3 + a];
                }
            }
        } else {
            uint64_t l = (uint64_t)i + 1;
            if (l >= node_count) return 0;
            float rf = tlas[l * 12 + 6];
            if (rf != floorf(rf) || !(rf > (float)l && rf < (float)node_count)) return 0;
            uint32_t r = (uint32_t)rf;
            const float *lb = tlas + l * 12, *rb = tlas + (size_t)r * 12;
            for (int a = 0; a < 3; a++) {
                mn[a] = fminf(lb[a], rb[a]);
                mx[a] = fmaxf(lb[3 + a], rb[3 + a]);
            }
        }
        for (int a = 0; a < 3; a++) {
            nd[a] = mn[a];
            nd[3 + a] = mx[a];
        }
    }
    return 1;
}

/* ---------------- per-frame TLAS (exact; precision-independent) ---------------- */

/* World AABB of every TLAS slot: Box3.applyMatrix4 of the instance's BLAS root box in
 * double, each corner divided by w (js/spectral_scene.js:1266-1285, three.js
 * Vector3.applyMatrix4); an inactive slot is the point at its translation
 * (js/spectral_scene.js:1279-1284). world is column-major 4x4 per slot. padded gets the
 * box inflated per axis by 1e-5 * (|extent| + |min| + |max|) + 1e-6 in float, the input
 * of the SAH build below. */
void or_inst_bounds(const float *mats, uint32_t inst_base, uint32_t n, const uint32_t *nodes, const double *world,
                    float *bounds, float *padded) {
    for (uint32_t i = 0; i < n; i++) {
        const float *rec = mats + inst_base + (size_t)i * 28;
        const double *e = world + (size_t)i * 16;
        float out[6];
        if (!(rec[15] > 0.5f)) {
            for (int a = 0; a < 3; a++) out[a] = out[3 + a] = (float)e[12 + a];
        } else {
            const float *rb = (const float *)(nodes + (size_t)(uint32_t)rec[12] * 8);
            double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
            for (int c = 0; c < 8; c++) {
                const double x = (c & 1) ? rb[3] : rb[0], y = (c & 2) ? rb[4] : rb[1], z = (c & 4) ? rb[5] : rb[2];
                const double w = 1 / (e[3] * x + e[7] * y + e[11] * z + e[15]);
                const double p[3] = {(e[0] * x + e[4] * y + e[8] * z + e[12]) * w,
                                     (e[1] * x + e[5] * y + e[9] * z + e[13]) * w,
                                     (e[2] * x + e[6] * y + e[10] * z + e[14]) * w};
                for (int a = 0; a < 3; a++) {
                    if (p[a] < mn[a]) mn[a] = p[a];
                    if (mx[a] < p[a]) mx[a] = p[a];
                }
            }
            for (int a = 0; a < 3; a++) {
                out[a] = (float)mn[a];
                out[3 + a] = (float)mx[a];
            }
        }
        for (int a = 0; a < 6; a++) bounds[(size_t)i * 6 + a] = out[a];
        for (int a = 0; a < 3; a++) {
            const float pad = 1e-5f * (fabsf(out[3 + a] - out[a]) + fabsf(out[a]) + fabsf(out[3 + a])) + 1e-6f;
            padded[(size_t)i * 6 + a] = out[a] - pad;
            padded[(size_t)i * 6 + 3 + a] = out[3 + a] + pad;
        }
    }
}

/* The SAH TLAS that the library builds on the device every frame (no reference
 * counterpart: the reference TLAS is a frozen median split, js/spectral_scene.js:973-1003).
 * Restated here top-down and recursively so the device's level-parallel build can be
 * compared word for word: one slot per leaf, 32 centroid bins per axis with extent > 0,
 * SAH cost area(L)·|L| + area(R)·|R| in double with the first minimum over (axis, bin),
 * a stable partition, the middle split when no axis has extent; depth-first node order,
 * miss = first node after the subtree. out is u32[8 (2n-1)] in the walk format
 * {min.xyz, miss}, {max.xyz, payload = 1 << 24 | slot, or 0xFFFFFFFF for interiors}. */
typedef struct {
    float mn[3], mx[3];
} or_box;

static void or_box_empty(or_box *b) {
    for (int a = 0; a < 3; a++) {
        b->mn[a] = INFINITY;
        b->mx[a] = -INFINITY;
    }
}
static void or_box_grow(or_box *b, const float *mn, const float *mx) {
    for (int a = 0; a < 3; a++) {
        if (mn[a] < b->mn[a]) b->mn[a] = mn[a];
        if (b->mx[a] < mx[a]) b->mx[a] = mx[a];
    }
}
static double or_box_area(const or_box *b) {
    double d[3];
    for (int a = 0; a < 3; a++) {
        d[a] = (double)b->mx[a] - (double)b->mn[a];
        if (!(d[a] >= 0)) return 0.0;
    }
    return 2.0 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0]);
}
static int or_bin(float c, float cmn, float ext) {
    int k = (int)((c - cmn) / ext * 32);
    return k < 0 ? 0 : (k > 31 ? 31 : k);
}

static uint32_t or_sah_emit(const float *bnd, uint32_t *ids, uint32_t *tmp, uint32_t lo, uint32_t hi, uint32_t pos,
                            uint32_t *out) {
    or_box b, cb;
    or_box_empty(&b);
    or_box_empty(&cb);
    for (uint32_t i = lo; i < hi; i++) {
        const float *bb = bnd + (size_t)ids[i] * 6;
        float c[3];
        for (int a = 0; a < 3; a++) c[a] = 0.5f * (bb[a] + bb[3 + a]);
        or_box_grow(&b, bb, bb + 3);
        or_box_grow(&cb, c, c);
    }
    uint32_t *node = out + (size_t)pos * 8;
    memcpy(node, b.mn, 12);
    memcpy(node + 4, b.mx, 12);
    const uint32_t m = hi - lo;
    if (m == 1) {
        node[3] = pos + 1;
        node[7] = (1u << 24) | (ids[lo] & 0x00FFFFFFu);
        return pos + 1;
    }
    double best = INFINITY;
    int best_axis = -1, best_bin = -1;
    for (int axis = 0; axis < 3; axis++) {
        const float ext = cb.mx[axis] - cb.mn[axis];
        if (!(ext > 0)) continue;
        or_box bins[32];
        uint32_t cnt[32] = {0}, rc[32];
        double ra[32];
        for (int k = 0; k < 32; k++) or_box_empty(&bins[k]);
        for (uint32_t i = lo; i < hi; i++) {
            const float *bb = bnd + (size_t)ids[i] * 6;
            const int k = or_bin(0.5f * (bb[axis] + bb[3 + axis]), cb.mn[axis], ext);
            cnt[k]++;
            or_box_grow(&bins[k], bb, bb + 3);
        }
        or_box acc;
        or_box_empty(&acc);
        uint32_t c = 0;
        for (int k = 31; k > 0; k--) {
            or_box_grow(&acc, bins[k].mn, bins[k].mx);
            c += cnt[k];
            ra[k] = or_box_area(&acc);
            rc[k] = c;
        }
        or_box lacc;
        or_box_empty(&lacc);
        uint32_t lc = 0;
        for (int k = 0; k < 31; k++) {
            or_box_grow(&lacc, bins[k].mn, bins[k].mx);
            lc += cnt[k];
            if (!lc || !rc[k + 1]) continue;
            const double cost = or_box_area(&lacc) * lc + ra[k + 1] * rc[k + 1];
            if (cost < best) {
                best = cost;
                best_axis = axis;
                best_bin = k;
            }
        }
    }
    uint32_t mid = lo + m / 2;
    if (best_axis >= 0) {
        const float ext = cb.mx[best_axis] - cb.mn[best_axis];
        uint32_t nl = 0, nr = 0;
        for (uint32_t i = lo; i < hi; i++) {
            const float *bb = bnd + (size_t)ids[i] * 6;
            if (or_bin(0.5f * (bb[best_axis] + bb[3 + best_axis]), cb.mn[best_axis], ext) <= best_bin)
                ids[lo + nl++] = ids[i];
            else
                tmp[nr++] = ids[i];
        }
        memcpy(ids + lo + nl, tmp, nr * sizeof(uint32_t));
        if (nl > 0 && nl < m) mid = lo + nl;
    }
    node[7] = 0xFFFFFFFFu;
    uint32_t next = or_sah_emit(bnd, ids, tmp, lo, mid, pos + 1, out);
    next = or_sah_emit(bnd, ids, tmp, mid, hi, next, out);
    node[3] = next;
    return next;
}

int or_build_sah_tlas(const float *padded, uint32_t n, uint32_t *out) {
    if (n == 0) return 0;
    uint32_t *ids = (uint32_t *)malloc((size_t)n * 4), *tmp = (uint32_t *)malloc((size_t)n * 4);
    if (!ids || !tmp) {
        free(ids);
        free(tmp);
        return 0;
    }
    for (uint32_t i = 0; i < n; i++) ids[i] = i;
    const uint32_t end = or_sah_emit(padded, ids, tmp, 0, n, 0, out);
    free(ids);
    free(tmp);
    return end == 2 * n - 1;
}

/* Linear-blend skinning of a vertex set (textbook LBS; the reference receives skinned
 * positions from its host — CPU skinning — and re-gathers them in updateDeforms,
 * js/spectral_scene.js:1831-1927). bones are 3x4 rows per bone; for each vertex
 * p' = sum_q w_q (B_q p), n' = normalize(sum_q w_q (B_q[3x3] n)), evaluated in float in
 * the order written (q = 0..3; row dot products left to right, translation last). rest
 * is 6 floats (position, normal) per vertex; out is 6 floats per vertex. */
void or_skin(const float *rest, const uint16_t *joints, const float *weights, uint32_t n, const float *bones,
             float *out) {
    for (uint32_t v = 0; v < n; v++) {
        const float *p = rest + (size_t)v * 6, *nr = p + 3;
        float acc[3] = {0.f, 0.f, 0.f}, nacc[3] = {0.f, 0.f, 0.f};
        for (int q = 0; q < 4; q++) {
            const float w = weights[(size_t)v * 4 + q];
            const float *B = bones + (size_t)joints[(size_t)v * 4 + q] * 12;
            for (int r = 0; r < 3; r++) {
                const float *row = B + r * 4;
                const float t = row[0] * p[0] + row[1] * p[1] + row[2] * p[2] + row[3];
                const float tn = row[0] * nr[0] + row[1] * nr[1] + row[2] * nr[2];
                acc[r] = acc[r] + w * t;
                nacc[r] = nacc[r] + w * tn;
            }
        }
        const float len = sqrtf(nacc[0] * nacc[0] + nacc[1] * nacc[1] + nacc[2] * nacc[2]);
        for (int r = 0; r < 3; r++) {
            out[(size_t)v * 6 + r] = acc[r];
            out[(size_t)v * 6 + 3 + r] = len > 0.f ? nacc[r] / len : nacc[r];
        }
    }
}
