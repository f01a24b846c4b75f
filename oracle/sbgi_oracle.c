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
#define REAL double
#define P(n) or64_##n
#define RABS fabs
#define RSQRT sqrt
#define RFLOOR floor
#define RSIN sin
#define RCOS cos
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
#undef RPOW
#undef REXP
#undef RFMIN
#undef RFMAX
#undef RGAUSS
#undef VSTRIDE

/* ---------------- FP32 restatement ---------------- */
#define REAL float
#define P(n) or32_##n
#define RABS fabsf
#define RSQRT sqrtf
#define RFLOOR floorf
#define RSIN sinf
#define RCOS cosf
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
