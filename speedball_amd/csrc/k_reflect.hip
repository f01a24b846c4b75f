// k_reflect.hip — local reflection lobes from the probe rays: the rough power-8 lobe in the
// irradiance tiles and the glossy power-64 cache.
//
// Restates the rough lobe of blendKernel (js/gi_probes.js:2358-2426, 2504-2522) and
// glossyKernel (:2538-2706). Both are contractions of the same ray records the irradiance
// blend reads, with texel x ray weights that depend only on the two directions:
//     S[t, (p, c)] = sum_k W[t, k] X[k, (p, c)],  W = max(dir_t . dir_k, 0)^8 (rough)
//                                                   or ^64 (glossy)
// over the channels c = {r v, g v, b v, v}, v = 1 for a hit, skyValid for a miss and 0
// for an unshaded (metal / glass) hit. The weight tables and their normalisers (sum_k W)
// are built once per solve. Emitters add their virtual rays with the same lobe. The rough
// lobe's temporal blend uses the probe's depth moments as the "seeded" sentinel, so its
// kernel runs before the irradiance blend rewrites them. FP32 path: one lane per probe and
// texel group; MFMA path: v_mfma_f32_16x16x32_{bf16,f16}, one wave per 16 probes.
#pragma clang fp contract(off)

#include "kernels.h"
#include "mfma_util.h"
#include "shade.h"

namespace {

constexpr uint32_t kIrr = 36;
constexpr int kGroup = 12;

__device__ __forceinline__ uint32_t kb_count(uint32_t rays) { return (rays + 31) / 32; }
__device__ __forceinline__ uint32_t mt_count(uint32_t texels) { return (texels + 15) / 16; }

__device__ __forceinline__ float pow8(float cd) {
    const float c2 = cd * cd, c4 = c2 * c2;
    return c4 * c4;
}
__device__ __forceinline__ float pow64(float cd) {
    const float c8 = pow8(cd), c16 = c8 * c8, c32 = c16 * c16;
    return c32 * c32;
}

// canonical gutter source of tile texel (lx, ly) (js/gi_probes.js:2574-2592)
__device__ __forceinline__ void gutter_src(uint32_t lx, uint32_t ly, uint32_t tile, uint32_t &sx, uint32_t &sy) {
    const uint32_t edge = tile - 1, lo = 1, hi = tile - 2;
    const bool on_l = lx == 0, on_r = lx == edge, on_t = ly == 0, on_b = ly == edge;
    const bool on_col = on_l || on_r, on_row = on_t || on_b, on_corner = on_col && on_row;
    sx = on_corner ? (on_l ? hi : lo) : (on_row ? edge - lx : (on_col ? (on_l ? lo : hi) : lx));
    sy = on_corner ? (on_t ? hi : lo) : (on_row ? (on_t ? lo : hi) : (on_col ? edge - ly : ly));
}

// direction of rough texel t (6x6 interior) or glossy tile texel t (its mirrored source)
__device__ __forceinline__ float3 rough_dir(uint32_t t) {
    return oct_decode_uv(((float)(t % 6) + 0.5f) / 6.f, ((float)(t / 6) + 0.5f) / 6.f);
}
__device__ __forceinline__ float3 glossy_dir(uint32_t t, uint32_t G) {
    uint32_t sx, sy;
    gutter_src(t % G, t / G, G, sx, sy);
    const float oct = (float)(G - 2);
    return oct_decode_uv(((float)sx - 1.f + 0.5f) / oct, ((float)sy - 1.f + 0.5f) / oct);
}

// coverage of a ray record: 1 for a hit, skyValid for a true miss, 0 for an unshaded hit
__device__ __forceinline__ float ray_valid(float ht, float sky_valid) {
    return ht >= 0.f ? 1.f : (ht == -1.f ? sky_valid : 0.f);
}

// Weight tables: block b < 36 owns rough texel b, blocks 36.. the glossy tile texels.
__global__ void __launch_bounds__(256) k_weights_refl(uint32_t rays, const float4 *__restrict__ dirs, ReflTables t,
                                                      uint32_t precision) {
    __shared__ float w_s[256];
    const uint32_t b = blockIdx.x;
    const bool rough = b < kIrr;
    const uint32_t row = rough ? b : b - kIrr;
    const uint32_t T = rough ? kIrr : t.gtile * t.gtile;
    const float3 dir = rough ? rough_dir(row) : glossy_dir(row, t.gtile);
    float *wt = rough ? t.w_rough_t : t.w_gloss_t;
    uint16_t *fhi = rough ? t.frag_rough : t.frag_gloss;
    const size_t frag = (size_t)mt_count(T) * kb_count(rays) * 64 * 8;
    // k runs over the K padding of the last block too (rays..32 KB - 1): zero weights there,
    // so the MFMA contraction never depends on the table's initial contents
    for (uint32_t k = threadIdx.x; k < 32 * kb_count(rays); k += blockDim.x) {
        float w = 0.f;
        if (k < rays) {
            const float cd = fmaxf(dot3(dir, xyz(dirs[k])), 0.f);
            w = rough ? pow8(cd) : pow64(cd);
            w_s[k] = w;
            wt[(size_t)k * T + row] = w;
        }
        if (precision != SBGI_BLEND_FP32) {
            uint16_t hi, lo;
            if (precision == SBGI_BLEND_FP16) {
                hi = to_f16(w);
                lo = 0;
            } else {
                hi = to_bf16(w);
                lo = precision == SBGI_BLEND_BF16X3 ? to_bf16(w - from_bf16(hi)) : (uint16_t)0;
            }
            const size_t o = frag_index(row, k, kb_count(rays));
            fhi[o] = hi;
            fhi[frag + o] = lo;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float s = 0.f, sq = 0.f;
        for (uint32_t k = 0; k < rays; k++) {
            s += w_s[k];
            sq += quantized(w_s[k], precision);
        }
        (rough ? t.norm_rough : t.norm_gloss)[row] = s;
        (rough ? t.norm_rough_q : t.norm_gloss_q)[row] = sq;
    }
}

// emitter e's virtual rays in a lobe of the given power (injectEmitterVirtualRays)
template <int POW>
__device__ __forceinline__ void inject_lobe(const EmitterTap &e, float3 tdir, float &r, float &g, float &b, float &h,
                                            float &ws) {
    const float cd = fmaxf(dot3(tdir, e.d), 0.f);
    const float wk = (POW == 8 ? pow8(cd) : pow64(cd)) * e.kv;
    r += e.L.x * wk;
    g += e.L.y * wk;
    b += e.L.z * wk;
    h += wk;
    ws += wk;
}

__device__ __forceinline__ EmitterTap tap_of(const DevCascade &c, const EmitterInject &em, uint32_t probe, uint32_t e,
                                            float3 ro) {
    return emitter_tap(em.lights + (size_t)(em.base + e) * SBGI_LIGHT_STRIDE, ro, (float)c.rays,
                       em.vis[(size_t)probe * SBGI_MAX_EMITTERS + e]);
}

// the depth texel whose first write marks the probe as seeded, for rough texel t
__device__ __forceinline__ size_t sentinel_index(const DevCascade &c, uint32_t probe, uint32_t t) {
    const uint32_t td = c.depth_oct == 6 ? 8u : c.depth_oct + 2u;
    return (size_t)probe * td * td + (t / 6 + 1) * td + (t % 6 + 1);
}

// rough temporal blend of one texel (js/gi_probes.js:2504-2522): steady retention, reset
// while the probe's depth moments are unseeded
__device__ __forceinline__ void rough_policy(const DevCascade &c, const sbgi_tick_params &p, float *__restrict__ rough,
                                             const float *__restrict__ depth, uint32_t probe, uint32_t t, float r,
                                             float g, float b, float h, float ws) {
    const float den = fmaxf(ws, 1e-4f);
    const float cr = r / den, cg = g / den, cb = b / den, ch = fminf(fmaxf(h / den, 0.f), 1.f);
    const bool seeded = !(depth[sentinel_index(c, probe, t) * 2] < 1e-6f);
    const float raw = p.hysteresis + (1.f - p.hysteresis) * p.dbg_noise_h_boost;
    const float sh = seeded ? powf(fminf(fmaxf(raw, 1e-6f), 1.f), p.hysteresis_exponent) : 0.f;
    float4 *S = reinterpret_cast<float4 *>(rough + ((size_t)probe * 64 + (t / 6 + 1) * 8 + (t % 6 + 1)) * 4);
    const float4 prev = *S;
    *S = make_float4(cr * (1.f - sh) + prev.x * sh, cg * (1.f - sh) + prev.y * sh, cb * (1.f - sh) + prev.z * sh,
                     ch * (1.f - sh) + prev.w * sh);
}

__device__ __forceinline__ unsigned short f2h(float x) { return __half_as_ushort(__float2half_rn(x)); }

// glossy history of one tile texel and its atlas texel (js/gi_probes.js:2638-2675)
__device__ __forceinline__ void glossy_policy(const DevCascade &c, const sbgi_tick_params &p, const ReflState &s,
                                              uint32_t G, uint32_t interval, float inv_rays, uint32_t probe,
                                              uint32_t t, float r, float g, float b, float h, float ws) {
    const size_t gt = (size_t)probe * G * G + t;
    float4 *Nm = reinterpret_cast<float4 *>(s.gnum) + gt;
    const float4 pn = *Nm;
    const float pd = s.gden[gt];
    const float4 cur = make_float4(r * inv_rays, g * inv_rays, b * inv_rays, h * inv_rays);
    const float cd = ws * inv_rays;
    const float ref = p.hysteresis + (1.f - p.hysteresis) * p.dbg_noise_h_boost;
    const float gh = pd < 1e-6f ? 0.f : powf(fminf(fmaxf(ref, 1e-6f), 1.f), p.hysteresis_exponent * (float)interval);
    const float4 num = make_float4(cur.x * (1.f - gh) + pn.x * gh, cur.y * (1.f - gh) + pn.y * gh,
                                   cur.z * (1.f - gh) + pn.z * gh, cur.w * (1.f - gh) + pn.w * gh);
    const float den = fmaxf(cd * (1.f - gh) + pd * gh, 1e-6f);
    *Nm = num;
    s.gden[gt] = den;
    const uint32_t col = probe % s.gtiles_x, row = probe / s.gtiles_x;
    const size_t o = (size_t)(row * G + t / G) * s.gaw + col * G + t % G;
    s.gatlas[o] = make_ushort4(f2h(num.x / den), f2h(num.y / den), f2h(num.z / den),
                               f2h(fminf(fmaxf(num.w / den, 0.f), 1.f)));
}

// ---------------------------------------------------------------- rough lobe
__global__ void __launch_bounds__(64) k_rough_fp32(DevCascade c, sbgi_tick_params p, const float4 *__restrict__ rays,
                                                   ReflTables t, const float *__restrict__ depth, ReflState s,
                                                   EmitterInject em, float sky_valid) {
    const uint32_t slot = blockIdx.x * 64 + threadIdx.x;
    if (slot >= c.updated) return;
    const uint32_t probe = (c.offset + slot) % c.total;
    const uint32_t R = c.rays, U = c.updated, gy = blockIdx.y;
    float a[kGroup][4];
#pragma unroll
    for (int j = 0; j < kGroup; j++) a[j][0] = a[j][1] = a[j][2] = a[j][3] = 0.f;
    const float *wbase = t.w_rough_t + gy * kGroup;
    for (uint32_t k = 0; k < R; k++) {
        const float4 x = rays[(size_t)k * U + slot];
        const float v = ray_valid(x.w, sky_valid);
        const float *w = wbase + (size_t)k * kIrr;
#pragma unroll
        for (int j = 0; j < kGroup; j++) {
            const float sw = w[j];
            a[j][0] += (x.x * sw) * v;
            a[j][1] += (x.y * sw) * v;
            a[j][2] += (x.z * sw) * v;
            a[j][3] += sw * v;
        }
    }
    float ws[kGroup];
#pragma unroll
    for (int j = 0; j < kGroup; j++) ws[j] = t.norm_rough[gy * kGroup + j];
    if (em.count) {
        const float3 ro = probe_origin(c, em.classify, em.meta, probe);
        for (uint32_t e = 0; e < em.count; e++) {
            const EmitterTap tap = tap_of(c, em, probe, e, ro);
            if (!tap.valid) continue;
#pragma unroll
            for (int j = 0; j < kGroup; j++)
                inject_lobe<8>(tap, rough_dir(gy * kGroup + j), a[j][0], a[j][1], a[j][2], a[j][3], ws[j]);
        }
    }
#pragma unroll
    for (int j = 0; j < kGroup; j++)
        rough_policy(c, p, s.rough, depth, probe, gy * kGroup + j, a[j][0], a[j][1], a[j][2], a[j][3], ws[j]);
}

// B operand of the reflection contractions: 8 consecutive rays x one channel
template <uint32_t PREC>
__device__ __forceinline__ void refl_operand(const float4 *__restrict__ rays, uint32_t U, uint32_t R, uint32_t slot,
                                             bool valid, uint32_t k0, float sky_valid, u16x8 bh[4], u16x8 bl[4]) {
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const uint32_t k = k0 + j;
        float v[4] = {0.f, 0.f, 0.f, 0.f};
        if (valid && k < R) {
            const float4 x = rays[(size_t)k * U + slot];
            const float vv = ray_valid(x.w, sky_valid);
            v[0] = x.x * vv;
            v[1] = x.y * vv;
            v[2] = x.z * vv;
            v[3] = vv;
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            uint16_t hi, lo;
            split<PREC>(v[q], hi, lo);
            bh[q][j] = hi;
            bl[q][j] = lo;
        }
    }
}

// One wave per 16 probes; MT M-tiles of texels starting at tile mt0 of a table with
// MT_TOTAL tiles.
template <uint32_t PREC, int MT>
__device__ __forceinline__ void refl_contract(const float4 *__restrict__ rays, uint32_t U, uint32_t R, uint32_t slot,
                                              bool valid, float sky_valid, const uint16_t *__restrict__ frag,
                                              size_t frag_half, uint32_t mt0, uint32_t mt_total, f32x4 acc[MT][4]) {
    constexpr bool SPLIT = PREC == SBGI_BLEND_BF16X3;
    const uint32_t lane = threadIdx.x & 63u, g = lane >> 4, KB = (R + 31) / 32;
    const u16x8 *f_hi = reinterpret_cast<const u16x8 *>(frag);
    const u16x8 *f_lo = reinterpret_cast<const u16x8 *>(frag + frag_half);
#pragma unroll
    for (int m = 0; m < MT; m++)
#pragma unroll
        for (int q = 0; q < 4; q++) acc[m][q] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (uint32_t kb = 0; kb < KB; kb++) {
        u16x8 bh[4], bl[4];
        refl_operand<PREC>(rays, U, R, slot, valid, kb * 32 + g * 8, sky_valid, bh, bl);
#pragma unroll
        for (int m = 0; m < MT; m++) {
            if (mt0 + m >= mt_total) break;
            const size_t fo = ((size_t)(mt0 + m) * KB + kb) * 64 + lane;
            const u16x8 a_hi = f_hi[fo];
            u16x8 a_lo;
            if (SPLIT) a_lo = f_lo[fo];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                acc[m][q] = mfma<PREC>(a_hi, bh[q], acc[m][q]);
                if (SPLIT) {
                    acc[m][q] = mfma<PREC>(a_hi, bl[q], acc[m][q]);
                    acc[m][q] = mfma<PREC>(a_lo, bh[q], acc[m][q]);
                }
            }
        }
    }
}

template <uint32_t PREC>
__global__ void __launch_bounds__(64) k_rough_mfma(DevCascade c, sbgi_tick_params p, const float4 *__restrict__ rays,
                                                   ReflTables t, const float *__restrict__ depth, ReflState s,
                                                   EmitterInject em, float sky_valid) {
    constexpr int MT = 3;
    const uint32_t lane = threadIdx.x, g = lane >> 4, n = lane & 15;
    const uint32_t slot = blockIdx.x * 16 + n;
    const bool valid = slot < c.updated;
    const size_t half = (size_t)MT * kb_count(c.rays) * 64 * 8;
    f32x4 acc[MT][4];
    refl_contract<PREC, MT>(rays, c.updated, c.rays, slot, valid, sky_valid, t.frag_rough, half, 0, MT, acc);
    if (!valid) return;
    const uint32_t probe = (c.offset + slot) % c.total;
    float v[MT][4][5];
#pragma unroll
    for (int m = 0; m < MT; m++)
#pragma unroll
        for (int jj = 0; jj < 4; jj++) {
            const uint32_t tt = m * 16 + g * 4 + jj;
#pragma unroll
            for (int q = 0; q < 4; q++) v[m][jj][q] = acc[m][q][jj];
            v[m][jj][4] = tt < kIrr ? t.norm_rough_q[tt] : 0.f;
        }
    if (em.count) {
        const float3 ro = probe_origin(c, em.classify, em.meta, probe);
        for (uint32_t e = 0; e < em.count; e++) {
            const EmitterTap tap = tap_of(c, em, probe, e, ro);
            if (!tap.valid) continue;
#pragma unroll
            for (int m = 0; m < MT; m++)
#pragma unroll
                for (int jj = 0; jj < 4; jj++) {
                    const uint32_t tt = m * 16 + g * 4 + jj;
                    if (tt < kIrr)
                        inject_lobe<8>(tap, rough_dir(tt), v[m][jj][0], v[m][jj][1], v[m][jj][2], v[m][jj][3],
                                       v[m][jj][4]);
                }
        }
    }
#pragma unroll
    for (int m = 0; m < MT; m++)
#pragma unroll
        for (int jj = 0; jj < 4; jj++) {
            const uint32_t tt = m * 16 + g * 4 + jj;
            if (tt < kIrr)
                rough_policy(c, p, s.rough, depth, probe, tt, v[m][jj][0], v[m][jj][1], v[m][jj][2], v[m][jj][3],
                             v[m][jj][4]);
        }
}

// ---------------------------------------------------------------- glossy cache
// FP32: one lane per (probe, tile texel); blockIdx.x = updated slot.
__global__ void __launch_bounds__(64) k_glossy_fp32(DevCascade c, sbgi_tick_params p, const float4 *__restrict__ rays,
                                                    ReflTables t, ReflState s, EmitterInject em, float sky_valid,
                                                    uint32_t interval, float inv_rays) {
    const uint32_t G = t.gtile, T = G * G;
    const uint32_t slot = blockIdx.x, tt = blockIdx.y * 64 + threadIdx.x;
    if (tt >= T || (interval > 1 && tt % interval != p.glossy_phase)) return;
    const uint32_t probe = (c.offset + slot) % c.total;
    const uint32_t R = c.rays, U = c.updated;
    float r = 0.f, g = 0.f, b = 0.f, h = 0.f;
    for (uint32_t k = 0; k < R; k++) {
        const float4 x = rays[(size_t)k * U + slot];
        const float v = ray_valid(x.w, sky_valid);
        const float gw = t.w_gloss_t[(size_t)k * T + tt];
        r += (x.x * gw) * v;
        g += (x.y * gw) * v;
        b += (x.z * gw) * v;
        h += gw * v;
    }
    float ws = t.norm_gloss[tt];
    if (em.count) {
        const float3 ro = probe_origin(c, em.classify, em.meta, probe);
        const float3 dir = glossy_dir(tt, G);
        for (uint32_t e = 0; e < em.count; e++) {
            const EmitterTap tap = tap_of(c, em, probe, e, ro);
            if (tap.valid) inject_lobe<64>(tap, dir, r, g, b, h, ws);
        }
    }
    glossy_policy(c, p, s, G, interval, inv_rays, probe, tt, r, g, b, h, ws);
}

// MFMA: one wave per 16 probes and 4 M-tiles (64 tile texels): blockIdx.y = texel chunk.
template <uint32_t PREC>
__global__ void __launch_bounds__(64) k_glossy_mfma(DevCascade c, sbgi_tick_params p, const float4 *__restrict__ rays,
                                                    ReflTables t, ReflState s, EmitterInject em, float sky_valid,
                                                    uint32_t interval, float inv_rays) {
    constexpr int MT = 4;
    const uint32_t G = t.gtile, T = G * G, MTT = mt_count(T);
    const uint32_t lane = threadIdx.x, g = lane >> 4, n = lane & 15;
    const uint32_t slot = blockIdx.x * 16 + n;
    const bool valid = slot < c.updated;
    const uint32_t mt0 = blockIdx.y * MT;
    const size_t half = (size_t)MTT * kb_count(c.rays) * 64 * 8;
    f32x4 acc[MT][4];
    refl_contract<PREC, MT>(rays, c.updated, c.rays, slot, valid, sky_valid, t.frag_gloss, half, mt0, MTT, acc);
    if (!valid) return;
    const uint32_t probe = (c.offset + slot) % c.total;
    float3 ro = f3(0.f, 0.f, 0.f);
    if (em.count) ro = probe_origin(c, em.classify, em.meta, probe);
#pragma unroll
    for (int m = 0; m < MT; m++)
#pragma unroll
        for (int jj = 0; jj < 4; jj++) {
            const uint32_t tt = (mt0 + m) * 16 + g * 4 + jj;
            if (tt >= T || (interval > 1 && tt % interval != p.glossy_phase)) continue;
            float r = acc[m][0][jj], gg = acc[m][1][jj], b = acc[m][2][jj], h = acc[m][3][jj];
            float ws = t.norm_gloss_q[tt];
            if (em.count) {
                const float3 dir = glossy_dir(tt, G);
                for (uint32_t e = 0; e < em.count; e++) {
                    const EmitterTap tap = tap_of(c, em, probe, e, ro);
                    if (tap.valid) inject_lobe<64>(tap, dir, r, gg, b, h, ws);
                }
            }
            glossy_policy(c, p, s, G, interval, inv_rays, probe, tt, r, gg, b, h, ws);
        }
}

__global__ void k_fill_u16x4(ushort4 *__restrict__ p, size_t n, ushort4 v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

float sky_valid_of(const sbgi_tick_params &p) {
    return p.reflection_sky_fallback * p.sky_configured * (p.sky_intensity > 0.f ? 1.f : 0.f);
}

}  // namespace

size_t refl_tables_bytes(uint32_t rays, uint32_t gtile) {
    const size_t KB = (rays + 31) / 32, Tg = (size_t)gtile * gtile;
    const size_t f = 4 * ((size_t)rays * (kIrr + Tg) + 2 * (kIrr + Tg));
    const size_t frag = 2 * 2 * 64 * 8 * KB * (3 + (Tg + 15) / 16);
    return f + frag + 1024;
}

void refl_tables_carve(void *base, uint32_t rays, uint32_t gtile, ReflTables &t) {
    char *p = static_cast<char *>(base);
    auto take = [&](size_t bytes) {
        char *r = p;
        p += (bytes + 63) & ~(size_t)63;
        return r;
    };
    const size_t KB = (rays + 31) / 32, Tg = (size_t)gtile * gtile;
    t.gtile = gtile;
    t.w_rough_t = reinterpret_cast<float *>(take(4ull * rays * kIrr));
    t.norm_rough = reinterpret_cast<float *>(take(4ull * kIrr));
    t.norm_rough_q = reinterpret_cast<float *>(take(4ull * kIrr));
    t.frag_rough = reinterpret_cast<uint16_t *>(take(2ull * 2 * 64 * 8 * KB * 3));
    t.w_gloss_t = reinterpret_cast<float *>(take(4ull * rays * Tg));
    t.norm_gloss = reinterpret_cast<float *>(take(4ull * Tg));
    t.norm_gloss_q = reinterpret_cast<float *>(take(4ull * Tg));
    t.frag_gloss = reinterpret_cast<uint16_t *>(take(2ull * 2 * 64 * 8 * KB * ((Tg + 15) / 16)));
}

hipError_t launch_weights_refl(uint32_t rays, const float4 *dirs, const ReflTables &t, uint32_t precision,
                               hipStream_t st) {
    const uint32_t blocks = kIrr + t.gtile * t.gtile;
    hipLaunchKernelGGL(k_weights_refl, dim3(blocks), dim3(256), 0, st, rays, dirs, t, precision);
    return hipGetLastError();
}

hipError_t launch_rough(const DevCascade &c, const sbgi_tick_params &p, const float4 *rays, const ReflTables &t,
                        uint32_t precision, const float *depth, const ReflState &s, const EmitterInject &em,
                        hipStream_t st) {
    if (c.updated == 0) return hipSuccess;
    const float sv = sky_valid_of(p);
    switch (precision) {
        case SBGI_BLEND_FP32:
            hipLaunchKernelGGL(k_rough_fp32, dim3((c.updated + 63) / 64, kIrr / kGroup), dim3(64), 0, st, c, p, rays,
                               t, depth, s, em, sv);
            break;
        case SBGI_BLEND_BF16:
            hipLaunchKernelGGL(k_rough_mfma<SBGI_BLEND_BF16>, dim3((c.updated + 15) / 16), dim3(64), 0, st, c, p,
                               rays, t, depth, s, em, sv);
            break;
        case SBGI_BLEND_BF16X3:
            hipLaunchKernelGGL(k_rough_mfma<SBGI_BLEND_BF16X3>, dim3((c.updated + 15) / 16), dim3(64), 0, st, c, p,
                               rays, t, depth, s, em, sv);
            break;
        case SBGI_BLEND_FP16:
            hipLaunchKernelGGL(k_rough_mfma<SBGI_BLEND_FP16>, dim3((c.updated + 15) / 16), dim3(64), 0, st, c, p,
                               rays, t, depth, s, em, sv);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_glossy(const DevCascade &c, const sbgi_tick_params &p, const float4 *rays, const ReflTables &t,
                         uint32_t precision, uint32_t interval, const ReflState &s, const EmitterInject &em,
                         hipStream_t st) {
    if (c.updated == 0 || t.gtile == 0) return hipSuccess;
    const float sv = sky_valid_of(p);
    const float inv = (float)(1.0 / (double)c.rays);
    const uint32_t T = t.gtile * t.gtile;
    if (precision == SBGI_BLEND_FP32) {
        hipLaunchKernelGGL(k_glossy_fp32, dim3(c.updated, (T + 63) / 64), dim3(64), 0, st, c, p, rays, t, s, em, sv,
                           interval, inv);
        return hipGetLastError();
    }
    const dim3 grid((c.updated + 15) / 16, (((T + 15) / 16) + 3) / 4);
    switch (precision) {
        case SBGI_BLEND_BF16:
            hipLaunchKernelGGL(k_glossy_mfma<SBGI_BLEND_BF16>, grid, dim3(64), 0, st, c, p, rays, t, s, em, sv,
                               interval, inv);
            break;
        case SBGI_BLEND_BF16X3:
            hipLaunchKernelGGL(k_glossy_mfma<SBGI_BLEND_BF16X3>, grid, dim3(64), 0, st, c, p, rays, t, s, em, sv,
                               interval, inv);
            break;
        case SBGI_BLEND_FP16:
            hipLaunchKernelGGL(k_glossy_mfma<SBGI_BLEND_FP16>, grid, dim3(64), 0, st, c, p, rays, t, s, em, sv,
                               interval, inv);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_fill_atlas(ushort4 *atlas, size_t texels, ushort4 v, hipStream_t st) {
    if (texels == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u16x4, dim3((unsigned)((texels + 255) / 256)), dim3(256), 0, st, atlas, texels, v);
    return hipGetLastError();
}
