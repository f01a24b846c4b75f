// k_upload.hip — bilateral filter, octahedral gutters, RGBA16F atlas store.
//
// Restates uploadKernel (js/gi_probes.js:2737-2846) and clearAtlasKernel (:2690-2707).
// One 64-lane wave per probe: the 36 interior lanes filter once into LDS, then all 64
// lanes store the 8x8 tile with the mirrored gutter. The 16x16 depth tile of the
// 14-texel depth mode is copied with the same gutter rule.
#include <hip/hip_fp16.h>

#include "kernels.h"

namespace {

__device__ __forceinline__ unsigned short f2h(float x) { return __half_as_ushort(__float2half_rn(x)); }
__device__ __forceinline__ ushort4 pack4(float a, float b, float c, float d) {
    return make_ushort4(f2h(a), f2h(b), f2h(c), f2h(d));
}

// canonical gutter source (js/gi_probes.js:2810-2837)
__device__ __forceinline__ void gutter_src(uint32_t lx, uint32_t ly, uint32_t tile, uint32_t &sx, uint32_t &sy) {
    const uint32_t edge = tile - 1, lo = 1, hi = tile - 2;
    const bool on_l = lx == 0, on_r = lx == edge, on_t = ly == 0, on_b = ly == edge;
    const bool on_col = on_l || on_r, on_row = on_t || on_b, on_corner = on_col && on_row;
    sx = on_corner ? (on_l ? hi : lo) : (on_row ? edge - lx : (on_col ? (on_l ? lo : hi) : lx));
    sy = on_corner ? (on_t ? hi : lo) : (on_row ? (on_t ? lo : hi) : (on_col ? edge - ly : ly));
}

__global__ void __launch_bounds__(64) k_upload(DevCascade c, sbgi_tick_params p, const float *__restrict__ irr,
                                               const float *__restrict__ depth, ushort4 *__restrict__ atlas,
                                               ushort4 *__restrict__ datlas, float4 *__restrict__ atlas_f32,
                                               const float *__restrict__ rough, ushort4 *__restrict__ ratlas) {
    __shared__ float4 cache[36];
    __shared__ float2 dcache[36];
    const uint32_t slot = blockIdx.x;
    const uint32_t lane = threadIdx.x;
    const uint32_t probe = (c.offset + slot) % c.total;
    const float *I = irr + (size_t)probe * 64 * 4;
    const bool parity = c.depth_oct == 6;
    if (lane < 36) {
        const int sx = (int)(lane % 6) + 1, sy = (int)(lane / 6) + 1;
        const float4 e = *reinterpret_cast<const float4 *>(I + (sy * 8 + sx) * 4);
        const float lc = e.x * 0.2126f + e.y * 0.7152f + e.z * 0.0722f;
        const float vc = fmaxf(e.w - lc * lc, 0.f);
        const float sm_w = 1.f + p.filter_smooth * 6.f;
        const float k_eff = 8.f * sm_w * p.dbg_filter_k;
        const float rel_eff = 0.0225f * sm_w * p.dbg_filter_rel;
        const float eps = 0.001f * p.dbg_filter_eps;
        float fx = 0.f, fy = 0.f, fz = 0.f, fw = 0.f;
#pragma unroll
        for (int jy = -1; jy <= 1; jy++) {
#pragma unroll
            for (int jx = -1; jx <= 1; jx++) {
                const int d2 = jx * jx + jy * jy;
                const float gw = d2 == 0 ? 1.0f : (d2 == 1 ? 0.6065306597126334f : 0.36787944117144233f);
                const int nx = min(max(sx + jx, 1), 6), ny = min(max(sy + jy, 1), 6);
                const float4 en = *reinterpret_cast<const float4 *>(I + (ny * 8 + nx) * 4);
                const float dl = (en.x * 0.2126f + en.y * 0.7152f + en.z * 0.0722f) - lc;
                const float den = fmaxf(vc * k_eff + fmaxf(eps, lc * lc * rel_eff), 1e-8f);
                const float w = gw * expf(dl * dl / den * -1.f);
                fx += en.x * w;
                fy += en.y * w;
                fz += en.z * w;
                fw += w;
            }
        }
        const float fd = fmaxf(fw, 1e-4f);
        const float fs = p.filter_strength;
        cache[lane] = make_float4(e.x * (1.f - fs) + (fx / fd) * fs, e.y * (1.f - fs) + (fy / fd) * fs,
                                  e.z * (1.f - fs) + (fz / fd) * fs, 1.f);
        if (parity) dcache[lane] = *reinterpret_cast<const float2 *>(depth + ((size_t)probe * 64 + sy * 8 + sx) * 2);
    }
    __syncthreads();
    const uint32_t col = probe % c.res[0];
    const uint32_t row = (probe / (c.res[0] * c.res[1])) * c.res[1] + (probe / c.res[0]) % c.res[1];
    {
        const uint32_t lx = lane % 8, ly = lane / 8;
        uint32_t sx, sy;
        gutter_src(lx, ly, 8, sx, sy);
        const uint32_t src = (sy - 1) * 6 + (sx - 1);
        const size_t o = (size_t)(row * 8 + ly) * c.aw + col * 8 + lx;
        const float4 v = cache[src];
        atlas[o] = pack4(v.x, v.y, v.z, v.w);
        if (atlas_f32) atlas_f32[o] = v;
        if (parity) {
            const float2 d = dcache[src];
            datlas[o] = pack4(d.x, d.y, 0.f, 1.f);
        }
        // the rough lobe: unfiltered, same mirrored gutter (js/gi_probes.js:2792-2799, 2841-2843)
        if (ratlas) {
            const float4 r = *reinterpret_cast<const float4 *>(rough + ((size_t)probe * 64 + sy * 8 + sx) * 4);
            ratlas[o] = pack4(r.x, r.y, r.z, r.w);
        }
    }
    if (!parity) {
        const uint32_t td = c.depth_oct + 2;
        for (uint32_t local = lane; local < td * td; local += 64) {
            const uint32_t lx = local % td, ly = local / td;
            uint32_t sx, sy;
            gutter_src(lx, ly, td, sx, sy);
            const float2 d = *reinterpret_cast<const float2 *>(depth + ((size_t)probe * td * td + sy * td + sx) * 2);
            datlas[(size_t)(row * td + ly) * c.daw + col * td + lx] = pack4(d.x, d.y, 0.f, 1.f);
        }
    }
}

__global__ void k_clear_atlas(ushort4 *__restrict__ atlas, size_t texels) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < texels) atlas[i] = pack4(0.f, 0.f, 0.f, 1.f);
}

}  // namespace

hipError_t launch_upload(const DevCascade &c, const sbgi_tick_params &p, const float *irr, const float *depth,
                         ushort4 *atlas, ushort4 *datlas, float4 *atlas_f32, hipStream_t st, const float *rough,
                         ushort4 *ratlas) {
    if (c.updated == 0) return hipSuccess;
    hipLaunchKernelGGL(k_upload, dim3(c.updated), dim3(64), 0, st, c, p, irr, depth, atlas, datlas, atlas_f32, rough,
                       ratlas);
    return hipGetLastError();
}

hipError_t launch_clear_atlas(ushort4 *atlas, size_t texels, hipStream_t st) {
    if (texels == 0) return hipSuccess;
    hipLaunchKernelGGL(k_clear_atlas, dim3((unsigned)((texels + 255) / 256)), dim3(256), 0, st, atlas, texels);
    return hipGetLastError();
}
