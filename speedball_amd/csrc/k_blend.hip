// k_blend.hip — the ray->texel blend as a batched contraction, plus the temporal policy.
//
// Restates blendKernel (js/gi_probes.js:2317-2532). The reference recomputes the cosine
// weight cw = pow(max(dir_t . dir_k, 0), cosinePower) and the depth weight
// dw = pow(cw, depthSharpness) per (texel, ray) inside every probe's workgroup. Both
// depend only on the texel and the ray index — rayDir() ignores the probe
// (js/gi_probes.js:1692-1704) — so the gather of ALL probes is one contraction with a
// shared weight matrix:
//     S[t, (p, c)] = sum_k W[t, k] * X[k, (p, c)]
// with c = {r, g, b, L, L^2} against W_irr and c = {d, d^2} against W_d. The weights
// and their normalisers (sum_k W) are built once per solve by k_weights.
//
//   * FP32 path: one lane per probe, 12 texels per lane, weights broadcast through
//     scalar loads (they are wave-uniform), X read once per lane and ray.
//   * MFMA path: v_mfma_f32_16x16x32_{bf16,f16}; a wave owns 16 probes, A = W tiles
//     (16 texels x 32 rays, pre-swizzled into per-lane fragments), B = X tiles
//     (32 rays x 16 probes, one channel per tile) built in registers from the k-major
//     ray buffer, FP32 accumulation. BF16X3 splits both operands into hi + lo BF16
//     parts and accumulates hi*hi + hi*lo + lo*hi.
// Normalisation, the adaptive hysteresis policy and the depth blend stay FP32.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include "kernels.h"
#include "mfma_util.h"
#include "shade.h"

namespace {

constexpr uint32_t kIrrTexels = 36;

__device__ __forceinline__ uint32_t kb_count(uint32_t rays) { return (rays + 31) / 32; }
__device__ __forceinline__ uint32_t mt_count(uint32_t texels) { return (texels + 15) / 16; }

// Weight tables. Block b < 36 owns irradiance texel b (and, with the shared 8x8 depth
// tile, its depth weights); blocks 36.. own the 14x14 depth texels.
__global__ void __launch_bounds__(256) k_weights(DevCascade c, sbgi_tick_params p, const float4 *__restrict__ dirs,
                                                 BlendTables t, uint32_t precision) {
    __shared__ float w_s[256];
    __shared__ float d_s[256];
    __shared__ float qw_s[256];  // the weights as the MFMA operands round them
    __shared__ float qd_s[256];
    const uint32_t b = blockIdx.x;
    const bool parity = c.depth_oct == 6;
    const uint32_t td_count = parity ? kIrrTexels : c.depth_oct * c.depth_oct;
    const bool is_irr = b < kIrrTexels;
    uint32_t n, tx, ty, drow;
    if (is_irr) {
        n = 6;
        tx = b % 6;
        ty = b / 6;
        drow = b;
    } else {
        n = c.depth_oct;
        drow = b - kIrrTexels;
        tx = drow % n;
        ty = drow / n;
    }
    const float3 dir = oct_decode_uv(((float)tx + 0.5f) / (float)n, ((float)ty + 0.5f) / (float)n);
    const float cpow = fmaxf(p.dbg_cosine_power, 1e-4f);
    const uint32_t R = c.rays, KB = kb_count(R);
    const bool want_d = parity || !is_irr;
    for (uint32_t k = threadIdx.x; k < R; k += blockDim.x) {
        float cd = fmaxf(dot3(dir, xyz(dirs[k])), 0.f);
        float cw = powf(cd, cpow);
        float dw = powf(cw, p.depth_sharpness);
        w_s[k] = cw;
        d_s[k] = dw;
        qw_s[k] = quantized(cw, precision);
        qd_s[k] = quantized(dw, precision);
        if (is_irr) t.w_irr_t[(size_t)k * kIrrTexels + b] = cw;
        if (want_d) t.w_d_t[(size_t)k * td_count + drow] = dw;
    }
    __syncthreads();
    // MFMA fragments: element j of lane l of tile (mt, kb) is W[16 mt + l%16][32 kb + 8 (l/16) + j]
    const uint32_t mt_irr = mt_count(kIrrTexels);
    const size_t tile_elems = 64 * 8;
    const size_t irr_frag = (size_t)mt_irr * KB * tile_elems;
    const size_t d_frag = (size_t)mt_count(td_count) * KB * tile_elems;
    uint16_t *f_irr_hi = t.frag, *f_irr_lo = t.frag + irr_frag;
    uint16_t *f_d_hi = t.frag + 2 * irr_frag, *f_d_lo = f_d_hi + d_frag;
    if (precision != SBGI_BLEND_FP32) {
        // the K padding of the last block (rays R..32 KB - 1, when R is not a multiple of 32)
        // is written as zeros here, so the contraction never depends on the table's
        // initial contents
        for (uint32_t k = threadIdx.x; k < 32 * KB; k += blockDim.x) {
            const uint32_t kb = k / 32, g = (k % 32) / 8, j = k % 8;
            for (int which = 0; which < 2; which++) {
                if (which == 0 && !is_irr) continue;
                if (which == 1 && !want_d) continue;
                const uint32_t row = which == 0 ? b : drow;
                const float w = k < R ? (which == 0 ? w_s[k] : d_s[k]) : 0.f;
                const uint32_t mt = row / 16, lane = g * 16 + row % 16;
                const size_t o = (((size_t)mt * KB + kb) * 64 + lane) * 8 + j;
                uint16_t hi, lo;
                if (precision == SBGI_BLEND_FP16) {
                    hi = to_f16(w);
                    lo = 0;
                } else {
                    hi = to_bf16(w);
                    lo = precision == SBGI_BLEND_BF16X3 ? to_bf16(w - from_bf16(hi)) : (uint16_t)0;
                }
                if (which == 0) {
                    f_irr_hi[o] = hi;
                    f_irr_lo[o] = lo;
                } else {
                    f_d_hi[o] = hi;
                    f_d_lo[o] = lo;
                }
            }
        }
    }
    if (threadIdx.x == 0) {
        // normalisers, summed in ray order as the reference loop does (js/gi_probes.js:2368-2377)
        // (the rounded weights come from every thread above; one lane only adds)
        float s = 0.f, sq = 0.f, ds = 0.f, dsq = 0.f;
        for (uint32_t k = 0; k < R; k++) {
            s += w_s[k];
            ds += d_s[k];
            sq += qw_s[k];
            dsq += qd_s[k];
        }
        if (is_irr) {
            t.norm_irr[b] = s;
            t.norm_irr_q[b] = sq;
        }
        if (want_d) {
            t.norm_d[drow] = ds;
            t.norm_d_q[drow] = dsq;
        }
    }
}

// Adaptive temporal policy for one irradiance texel (js/gi_probes.js:2420-2493).
__device__ __forceinline__ void irr_policy(const sbgi_tick_params &p, float *__restrict__ irr, uint32_t probe,
                                           uint32_t local, float s_r, float s_g, float s_b, float s_l, float s_l2,
                                           float wsum) {
    const float ws = fmaxf(wsum, 1e-4f);
    const float mr = s_r / ws, mg = s_g / ws, mb = s_b / ws;
    const float cur_l = s_l / ws, cur_m2 = s_l2 / ws;
    float4 *I = reinterpret_cast<float4 *>(irr + ((size_t)probe * 64 + local) * 4);
    const float4 prev = *I;
    const bool was_black = prev.x + prev.y + prev.z < 1e-6f;
    const float prev_l = prev.x * 0.2126f + prev.y * 0.7152f + prev.z * 0.0722f;
    const float prev_m2 = prev.w;
    const float prev_var = fmaxf(prev_m2 - prev_l * prev_l, 0.f);
    const float cur_var = fmaxf(cur_m2 - cur_l * cur_l, 0.f);
    const float lref = fmaxf(fmaxf(cur_l, prev_l), 0.f);
    const float var_floor = p.dbg_var_eps + lref * lref * p.dbg_var_rel;
    const float sigma = sqrtf(prev_var + cur_var + var_floor);
    const float ad = fabsf(cur_l - prev_l);
    const float s0 = sigma * p.dbg_change_sigma0;
    const float s1 = sigma * fmaxf(p.temp_change_sigma1, p.dbg_change_sigma0 + 0.01f);
    const float cw = fminf(fmaxf((ad - s0) / fmaxf(s1 - s0, 1e-6f), 0.f), 1.f);
    const float hy = p.hysteresis;
    const float noise_h = hy + (1.f - hy) * p.dbg_noise_h_boost;
    const float change_h = fmaxf(hy - p.temp_change_h_drop, p.dbg_min_change_h);
    const float raw = fminf(fmaxf(noise_h * (1.f - cw) + change_h * cw, 1e-6f), 1.f);
    // pow(x, 1) is x exactly: the default exponent skips the general pow
    const float hexp = p.hysteresis_exponent;
    const float h = was_black ? 0.f : (hexp == 1.f ? raw : powf(raw, hexp));
    const float band = sqrtf(prev_var + var_floor) * p.temp_clamp_sigma;
    const float cs = fminf(1.f, band / fmaxf(ad, 1e-6f));
    const float cr = was_black ? mr : prev.x + (mr - prev.x) * cs;
    const float cg = was_black ? mg : prev.y + (mg - prev.y) * cs;
    const float cb = was_black ? mb : prev.z + (mb - prev.z) * cs;
    *I = make_float4(cr * (1.f - h) + prev.x * h, cg * (1.f - h) + prev.y * h, cb * (1.f - h) + prev.z * h,
                     cur_m2 * (1.f - h) + prev_m2 * h);
}

// Depth-moment blend for one depth texel (js/gi_probes.js:2500-2530).
__device__ __forceinline__ void depth_policy(const sbgi_tick_params &p, float *__restrict__ depth, size_t idx,
                                             float s_d, float s_d2, float dwsum) {
    const float dws = fmaxf(dwsum, 1e-4f);
    const float mr = s_d / dws, mr2 = s_d2 / dws;
    float2 *D = reinterpret_cast<float2 *>(depth + idx * 2);
    const float2 prev = *D;
    const bool dzero = prev.x < 1e-6f;
    const float rdh = fminf(fmaxf(p.hysteresis * p.dbg_depth_history, 0.f), 0.999f);
    const float dhb = fmaxf(rdh, 1e-6f);
    const float dh = dzero ? 0.f : (p.hysteresis_exponent == 1.f ? dhb : powf(dhb, p.hysteresis_exponent));
    *D = make_float2(mr * (1.f - dh) + prev.x * dh, mr2 * (1.f - dh) + prev.y * dh);
}

__device__ __forceinline__ size_t depth_index(const DevCascade &c, uint32_t probe, uint32_t t) {
    if (c.depth_oct == 6) return (size_t)probe * 64 + (t / 6 + 1) * 8 + (t % 6 + 1);
    const uint32_t n = c.depth_oct, td = n + 2;
    return (size_t)probe * td * td + (t / n + 1) * td + (t % n + 1);
}

constexpr int kGroup = 12;

// direction of interior irradiance texel tt of the 6x6 octahedral interior
__device__ __forceinline__ float3 irr_texel_dir(uint32_t tt) {
    return oct_decode_uv(((float)(tt % 6) + 0.5f) / 6.f, ((float)(tt / 6) + 0.5f) / 6.f);
}

// Emitter e's virtual rays added to one texel's sums after the ray loop, as the reference
// does (js/gi_probes.js:2262-2292, 2407-2426): w = pow(max(dir_t . d, 0), cosinePower),
// wk = w * kVirtual; acc += L wk, wsum += wk, and the luminance moments.
__device__ __forceinline__ void inject_tap(const EmitterTap &e, float3 tdir, float cpow, float &r, float &g, float &b,
                                           float &l, float &l2, float &ws) {
#pragma clang fp contract(off)
    const float cd = fmaxf(dot3(tdir, e.d), 0.f);
    // pow(x, 1) is x exactly: the default cosine power skips the general pow
    const float wk = (cpow == 1.f ? cd : powf(cd, cpow)) * e.kv;
    r += e.L.x * wk;
    g += e.L.y * wk;
    b += e.L.z * wk;
    ws += wk;
    const float lum = e.L.x * 0.2126f + e.L.y * 0.7152f + e.L.z * 0.0722f;
    l += lum * wk;
    l2 += (lum * lum) * wk;
}

__device__ __forceinline__ EmitterTap emitter_of(const DevCascade &c, const EmitterInject &em, uint32_t probe, uint32_t e,
                                                 float3 ro) {
    return emitter_tap(em.lights + (size_t)(em.base + e) * SBGI_LIGHT_STRIDE, ro, (float)c.rays,
                       em.vis[(size_t)probe * SBGI_MAX_EMITTERS + e]);
}

// FP32 blend. blockIdx.y < 3: irradiance texel group; else depth texel group. EMIT: the
// variant with emitter injection (it holds more registers, so it runs only for scenes
// with emitters).
template <bool EMIT>
__global__ void __launch_bounds__(64) k_blend_fp32(DevCascade c, sbgi_tick_params p, const float4 *__restrict__ rays,
                                                   BlendTables t, float *__restrict__ irr, float *__restrict__ depth,
                                                   EmitterInject em) {
    const uint32_t slot = blockIdx.x * 64 + threadIdx.x;
    if (slot >= c.updated) return;
    const uint32_t probe = (c.offset + slot) % c.total;
    const uint32_t R = c.rays, U = c.updated;
    const uint32_t gy = blockIdx.y;
    if (gy < kIrrTexels / kGroup) {
        float a[kGroup][5];
#pragma unroll
        for (int j = 0; j < kGroup; j++)
#pragma unroll
            for (int q = 0; q < 5; q++) a[j][q] = 0.f;
        const float *wbase = t.w_irr_t + gy * kGroup;
        for (uint32_t k = 0; k < R; k++) {
            const float4 x = rays[(size_t)k * U + slot];
            const float L = x.x * 0.2126f + x.y * 0.7152f + x.z * 0.0722f;
            const float L2 = L * L;
            const float *w = wbase + (size_t)k * kIrrTexels;
#pragma unroll
            for (int j = 0; j < kGroup; j++) {
                const float cw = w[j];
                a[j][0] += x.x * cw;
                a[j][1] += x.y * cw;
                a[j][2] += x.z * cw;
                a[j][3] += L * cw;
                a[j][4] += L2 * cw;
            }
        }
        float ws[kGroup];
        if (EMIT) {
#pragma unroll
            for (int j = 0; j < kGroup; j++) ws[j] = t.norm_irr[gy * kGroup + j];
            const float3 ro = probe_origin(c, em.classify, em.meta, probe);
            const float cpow = fmaxf(p.dbg_cosine_power, 1e-4f);
            for (uint32_t e = 0; e < em.count; e++) {
                const EmitterTap tap = emitter_of(c, em, probe, e, ro);
                if (!tap.valid) continue;
#pragma unroll
                for (int j = 0; j < kGroup; j++)
                    inject_tap(tap, irr_texel_dir(gy * kGroup + j), cpow, a[j][0], a[j][1], a[j][2], a[j][3], a[j][4],
                               ws[j]);
            }
        }
#pragma unroll
        for (int j = 0; j < kGroup; j++) {
            const uint32_t tt = gy * kGroup + j;
            const uint32_t local = (tt / 6 + 1) * 8 + (tt % 6 + 1);
            irr_policy(p, irr, probe, local, a[j][0], a[j][1], a[j][2], a[j][3], a[j][4], EMIT ? ws[j] : t.norm_irr[tt]);
        }
    } else {
        const uint32_t g = gy - kIrrTexels / kGroup;
        const uint32_t tdn = c.depth_oct == 6 ? kIrrTexels : c.depth_oct * c.depth_oct;
        const uint32_t t0 = g * kGroup;
        float a[kGroup][2];
#pragma unroll
        for (int j = 0; j < kGroup; j++) a[j][0] = a[j][1] = 0.f;
        const bool full = t0 + kGroup <= tdn;
        for (uint32_t k = 0; k < R; k++) {
            const float4 x = rays[(size_t)k * U + slot];
            const float d = ray_dist(x.w, c.max_dist);
            const float d2 = d * d;
            const float *w = t.w_d_t + (size_t)k * tdn + t0;
            if (full) {
#pragma unroll
                for (int j = 0; j < kGroup; j++) {
                    a[j][0] += d * w[j];
                    a[j][1] += d2 * w[j];
                }
            } else {
#pragma unroll
                for (int j = 0; j < kGroup; j++) {
                    if (t0 + j < tdn) {
                        a[j][0] += d * w[j];
                        a[j][1] += d2 * w[j];
                    }
                }
            }
        }
#pragma unroll
        for (int j = 0; j < kGroup; j++) {
            const uint32_t tt = t0 + j;
            if (tt < tdn) depth_policy(p, depth, depth_index(c, probe, tt), a[j][0], a[j][1], t.norm_d[tt]);
        }
    }
}

// MFMA blend: one wave per 16 probes and part. Part 0 (blockIdx.y) contracts the
// irradiance texels (MTI M-tiles, five channels) and runs their epilogue, with the emitter
// injection when EMIT; parts 1.. contract a range of the depth M-tiles (MTD in all: 3 for
// 36 texels, 13 for 196), two channels each. Splitting the texels over waves gives the
// grid three waves per 16 probes instead of one (12,288 probes: 2,304 waves for 1,024
// SIMDs), and each wave holds only its own accumulators.
constexpr int kBlendIrrTiles = 3;

template <uint32_t PREC, int NCH>
__device__ __forceinline__ void blend_b_frags(const float4 *__restrict__ rays, const DevCascade &c, bool valid,
                                              uint32_t slot, uint32_t kb, uint32_t g, u16x8 (&bh)[NCH],
                                              u16x8 (&bl)[NCH]) {
    // channels: NCH == 5: r, g, b, L, L^2; NCH == 2: d, d^2
    const uint32_t R = c.rays, U = c.updated;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const uint32_t k = kb * 32 + g * 8 + j;
        float v[NCH];
#pragma unroll
        for (int q = 0; q < NCH; q++) v[q] = 0.f;
        if (valid && k < R) {
            const float4 x = rays[(size_t)k * U + slot];
            if (NCH == 5) {
                const float L = x.x * 0.2126f + x.y * 0.7152f + x.z * 0.0722f;
                v[0] = x.x;
                v[1] = x.y;
                v[2] = x.z;
                v[3] = L;
                v[NCH - 1] = L * L;
            } else {
                const float d = ray_dist(x.w, c.max_dist);
                v[0] = d;
                v[NCH - 1] = d * d;
            }
        }
#pragma unroll
        for (int q = 0; q < NCH; q++) {
            uint16_t hi, lo;
            split<PREC>(v[q], hi, lo);
            bh[q][j] = hi;
            bl[q][j] = lo;
        }
    }
}

// Depth M-tiles [M0, M1) of one wave's 16 probes, contraction and epilogue.
template <uint32_t PREC, int MTD, int M0, int M1>
__device__ __forceinline__ void blend_depth_part(const DevCascade &c, const sbgi_tick_params &p,
                                                 const float4 *__restrict__ rays, const BlendTables &t,
                                                 float *__restrict__ depth) {
    constexpr bool SPLIT = PREC == SBGI_BLEND_BF16X3;
    const uint32_t lane = threadIdx.x;
    const uint32_t g = lane >> 4, n = lane & 15;
    const uint32_t slot = blockIdx.x * 16 + n;
    const bool valid = slot < c.updated;
    const uint32_t KB = (c.rays + 31) / 32;
    const uint32_t tdn = c.depth_oct == 6 ? kIrrTexels : c.depth_oct * c.depth_oct;
    const size_t tile_elems = 64 * 8;
    const size_t irr_frag = (size_t)kBlendIrrTiles * KB * tile_elems;
    const size_t d_frag = (size_t)MTD * KB * tile_elems;
    const u16x8 *f_d_hi = reinterpret_cast<const u16x8 *>(t.frag + 2 * irr_frag);
    const u16x8 *f_d_lo = reinterpret_cast<const u16x8 *>(t.frag + 2 * irr_frag + d_frag);
    f32x4 ad[M1 - M0][2];
#pragma unroll
    for (int m = 0; m < M1 - M0; m++) ad[m][0] = ad[m][1] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (uint32_t kb = 0; kb < KB; kb++) {
        u16x8 bh[2], bl[2];
        blend_b_frags<PREC, 2>(rays, c, valid, slot, kb, g, bh, bl);
#pragma unroll
        for (int m = 0; m < M1 - M0; m++) {
            const size_t fo = ((size_t)(M0 + m) * KB + kb) * 64 + lane;
            const u16x8 a_hi = f_d_hi[fo];
            u16x8 a_lo;
            if (SPLIT) a_lo = f_d_lo[fo];
#pragma unroll
            for (int q = 0; q < 2; q++) {
                ad[m][q] = mfma<PREC>(a_hi, bh[q], ad[m][q]);
                if (SPLIT) {
                    ad[m][q] = mfma<PREC>(a_hi, bl[q], ad[m][q]);
                    ad[m][q] = mfma<PREC>(a_lo, bh[q], ad[m][q]);
                }
            }
        }
    }
    if (!valid) return;
    const uint32_t probe = (c.offset + slot) % c.total;
    // D layout: lane holds rows 4*(lane/16) + jj of column lane%16
#pragma unroll
    for (int m = 0; m < M1 - M0; m++) {
#pragma unroll
        for (int jj = 0; jj < 4; jj++) {
            const uint32_t tt = (M0 + m) * 16 + g * 4 + jj;
            if (tt < tdn) depth_policy(p, depth, depth_index(c, probe, tt), ad[m][0][jj], ad[m][1][jj], t.norm_d_q[tt]);
        }
    }
}

// Irradiance M-tiles of one wave's 16 probes: contraction, emitter injection, epilogue.
template <uint32_t PREC, int MTD, bool EMIT>
__device__ __forceinline__ void blend_irr_part(const DevCascade &c, const sbgi_tick_params &p,
                                               const float4 *__restrict__ rays, const BlendTables &t,
                                               float *__restrict__ irr, const EmitterInject &em) {
    constexpr int MTI = kBlendIrrTiles;
    constexpr bool SPLIT = PREC == SBGI_BLEND_BF16X3;
    const uint32_t lane = threadIdx.x;
    const uint32_t g = lane >> 4, n = lane & 15;
    const uint32_t slot = blockIdx.x * 16 + n;
    const bool valid = slot < c.updated;
    const uint32_t KB = (c.rays + 31) / 32;
    const size_t tile_elems = 64 * 8;
    const size_t irr_frag = (size_t)MTI * KB * tile_elems;
    const u16x8 *f_irr_hi = reinterpret_cast<const u16x8 *>(t.frag);
    const u16x8 *f_irr_lo = reinterpret_cast<const u16x8 *>(t.frag + irr_frag);
    f32x4 ai[MTI][5];
#pragma unroll
    for (int m = 0; m < MTI; m++)
#pragma unroll
        for (int q = 0; q < 5; q++) ai[m][q] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (uint32_t kb = 0; kb < KB; kb++) {
        u16x8 bh[5], bl[5];
        blend_b_frags<PREC, 5>(rays, c, valid, slot, kb, g, bh, bl);
#pragma unroll
        for (int m = 0; m < MTI; m++) {
            const size_t fo = ((size_t)m * KB + kb) * 64 + lane;
            const u16x8 a_hi = f_irr_hi[fo];
            u16x8 a_lo;
            if (SPLIT) a_lo = f_irr_lo[fo];
#pragma unroll
            for (int q = 0; q < 5; q++) {
                ai[m][q] = mfma<PREC>(a_hi, bh[q], ai[m][q]);
                if (SPLIT) {
                    ai[m][q] = mfma<PREC>(a_hi, bl[q], ai[m][q]);
                    ai[m][q] = mfma<PREC>(a_lo, bh[q], ai[m][q]);
                }
            }
        }
    }
    if (!valid) return;
    const uint32_t probe = (c.offset + slot) % c.total;
    float ws[MTI][4];
    if (EMIT) {
#pragma unroll
        for (int m = 0; m < MTI; m++)
#pragma unroll
            for (int jj = 0; jj < 4; jj++) {
                const uint32_t tt = m * 16 + g * 4 + jj;
                ws[m][jj] = tt < kIrrTexels ? t.norm_irr_q[tt] : 0.f;
            }
        const float3 ro = probe_origin(c, em.classify, em.meta, probe);
        const float cpow = fmaxf(p.dbg_cosine_power, 1e-4f);
        for (uint32_t e = 0; e < em.count; e++) {
            const EmitterTap tap = emitter_of(c, em, probe, e, ro);
            if (!tap.valid) continue;
#pragma unroll
            for (int m = 0; m < MTI; m++)
#pragma unroll
                for (int jj = 0; jj < 4; jj++) {
                    const uint32_t tt = m * 16 + g * 4 + jj;
                    if (tt < kIrrTexels) {
                        float r = ai[m][0][jj], gg = ai[m][1][jj], b = ai[m][2][jj], l = ai[m][3][jj], l2 = ai[m][4][jj];
                        inject_tap(tap, irr_texel_dir(tt), cpow, r, gg, b, l, l2, ws[m][jj]);
                        ai[m][0][jj] = r;
                        ai[m][1][jj] = gg;
                        ai[m][2][jj] = b;
                        ai[m][3][jj] = l;
                        ai[m][4][jj] = l2;
                    }
                }
        }
    }
#pragma unroll
    for (int m = 0; m < MTI; m++) {
#pragma unroll
        for (int jj = 0; jj < 4; jj++) {
            const uint32_t tt = m * 16 + g * 4 + jj;
            if (tt < kIrrTexels) {
                const uint32_t local = (tt / 6 + 1) * 8 + (tt % 6 + 1);
                irr_policy(p, irr, probe, local, ai[m][0][jj], ai[m][1][jj], ai[m][2][jj], ai[m][3][jj],
                           ai[m][4][jj], EMIT ? ws[m][jj] : t.norm_irr_q[tt]);
            }
        }
    }
}

// Parts per 16 probes: irradiance, then the depth tiles in one part (3 tiles) or two halves.
template <int MTD>
constexpr uint32_t blend_parts() { return MTD > 3 ? 3u : 2u; }

template <uint32_t PREC, int MTD, bool EMIT>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) k_blend_mfma(DevCascade c, sbgi_tick_params p,
                                                   const float4 *__restrict__ rays, BlendTables t,
                                                   float *__restrict__ irr, float *__restrict__ depth,
                                                   EmitterInject em) {
    constexpr int H = (MTD + 1) / 2;
    if (blockIdx.y == 0) blend_irr_part<PREC, MTD, EMIT>(c, p, rays, t, irr, em);
    else if (MTD <= 3) blend_depth_part<PREC, MTD, 0, MTD>(c, p, rays, t, depth);
    else if (blockIdx.y == 1) blend_depth_part<PREC, MTD, 0, H>(c, p, rays, t, depth);
    else blend_depth_part<PREC, MTD, H, MTD>(c, p, rays, t, depth);
}

}  // namespace

size_t blend_tables_bytes(uint32_t rays, uint32_t depth_texels) {
    const size_t KB = (rays + 31) / 32;
    const size_t f = 4 * ((size_t)rays * kIrrTexels + (size_t)rays * depth_texels + 2 * kIrrTexels + 2 * depth_texels);
    const size_t frag = 2 * 2 * 64 * 8 * KB * (3 + (depth_texels + 15) / 16);
    return f + frag + 1024;  // + alignment slack of blend_tables_carve
}

void blend_tables_carve(void *base, uint32_t rays, uint32_t depth_texels, BlendTables &t) {
    char *p = static_cast<char *>(base);
    auto take = [&](size_t bytes) {
        char *r = p;
        p += (bytes + 63) & ~(size_t)63;
        return r;
    };
    t.w_irr_t = reinterpret_cast<float *>(take(4ull * rays * kIrrTexels));
    t.w_d_t = reinterpret_cast<float *>(take(4ull * rays * depth_texels));
    t.norm_irr = reinterpret_cast<float *>(take(4ull * kIrrTexels));
    t.norm_d = reinterpret_cast<float *>(take(4ull * depth_texels));
    t.norm_irr_q = reinterpret_cast<float *>(take(4ull * kIrrTexels));
    t.norm_d_q = reinterpret_cast<float *>(take(4ull * depth_texels));
    const size_t KB = (rays + 31) / 32;
    t.frag = reinterpret_cast<uint16_t *>(take(2ull * 2 * 64 * 8 * KB * (3 + (depth_texels + 15) / 16)));
}

hipError_t launch_weights(const DevCascade &c, const sbgi_tick_params &p, const float4 *dirs, const BlendTables &t,
                          uint32_t precision, hipStream_t st) {
    const uint32_t blocks = kIrrTexels + (c.depth_oct == 6 ? 0 : c.depth_oct * c.depth_oct);
    hipLaunchKernelGGL(k_weights, dim3(blocks), dim3(256), 0, st, c, p, dirs, t, precision);
    return hipGetLastError();
}

hipError_t launch_blend(const DevCascade &c, const sbgi_tick_params &p, const float4 *rays, const BlendTables &t,
                        uint32_t precision, float *irr, float *depth, const EmitterInject &em, hipStream_t st) {
    if (c.updated == 0) return hipSuccess;
    const uint32_t tdn = c.depth_oct == 6 ? kIrrTexels : c.depth_oct * c.depth_oct;
    if (precision == SBGI_BLEND_FP32) {
        dim3 grid((c.updated + 63) / 64, kIrrTexels / kGroup + (tdn + kGroup - 1) / kGroup);
        if (em.count)
            hipLaunchKernelGGL(k_blend_fp32<true>, grid, dim3(64), 0, st, c, p, rays, t, irr, depth, em);
        else
            hipLaunchKernelGGL(k_blend_fp32<false>, grid, dim3(64), 0, st, c, p, rays, t, irr, depth, em);
        return hipGetLastError();
    }
    const bool parity = c.depth_oct == 6;
    dim3 grid((c.updated + 15) / 16, parity ? blend_parts<3>() : blend_parts<13>());
    if (!parity && c.depth_oct != 14) return hipErrorInvalidValue;
#define SBGI_LAUNCH(PREC)                                                                                        \
    if (parity && em.count)                                                                                      \
        hipLaunchKernelGGL((k_blend_mfma<PREC, 3, true>), grid, dim3(64), 0, st, c, p, rays, t, irr, depth, em);  \
    else if (parity)                                                                                             \
        hipLaunchKernelGGL((k_blend_mfma<PREC, 3, false>), grid, dim3(64), 0, st, c, p, rays, t, irr, depth, em); \
    else if (em.count)                                                                                           \
        hipLaunchKernelGGL((k_blend_mfma<PREC, 13, true>), grid, dim3(64), 0, st, c, p, rays, t, irr, depth, em); \
    else                                                                                                         \
        hipLaunchKernelGGL((k_blend_mfma<PREC, 13, false>), grid, dim3(64), 0, st, c, p, rays, t, irr, depth, em);
    switch (precision) {
        case SBGI_BLEND_BF16:
            SBGI_LAUNCH(SBGI_BLEND_BF16);
            break;
        case SBGI_BLEND_BF16X3:
            SBGI_LAUNCH(SBGI_BLEND_BF16X3);
            break;
        case SBGI_BLEND_FP16:
            SBGI_LAUNCH(SBGI_BLEND_FP16);
            break;
        default:
            return hipErrorInvalidValue;
    }
#undef SBGI_LAUNCH
    return hipGetLastError();
}
