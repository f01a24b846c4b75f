// k_sample.hip — the probe field sampled at receiver points, as a batch kernel.
//
// Restates the receiver's irradiance gather (GiProbeNode.sampleIrradiance, _sampleCascade
// and _tapEW, js/gi_probes.js:480-614, 825-847) for N points at once, one thread per point:
//   P = X + Nvis * normalBias                     (the visibility-biased point, :884-885)
//   per cascade: the 8 probes of P's grid cell, trilinear weights + 1e-4, each tap
//     weighted by the Chebyshev test of the probe's depth moments toward P, the wrap
//     weight of the probe's side of the surface and the probe's state (both gated by
//     chebyStrength / classifyStrength), and fetched from the irradiance atlas in the
//     direction Ndir; E_c = sum e w / max(sum w, 1e-4)
//   cascades: E = E_0, then E = mix(E, E_c, wFine_c) for each finer cascade c, wFine_c
//     ramping from 0 at the cascade box's faces to 1 at borderBand inside it.
// Atlas fetches are the software bilinear of the bounce lookup (texel centres, clamp to
// edge) over the RGBA16F atlases the update writes; the probe state and relocation come
// from the per-probe meta buffer.
#pragma clang fp contract(off)

#include <cstring>

#include "kernels.h"
#include "shade.h"

namespace {

// one cascade's atlases and grid
struct SampleCascade {
    DevCascade c;
    const ushort4 *atlas, *datlas;
    const float *meta;
    float cheby_bias;
};

struct SampleArgs {
    SampleCascade casc[SBGI_MAX_CASCADES];
    uint32_t n_casc;
    float normal_bias, cheby_strength, classify_strength, border_band;
};

__device__ __forceinline__ float mixf(float a, float b, float t) { return a * (1.f - t) + b * t; }

// _tileUV: probe tile (col, row) of a tile x tile atlas with an oct x oct interior
__device__ __forceinline__ float3 fetch_tile(const ushort4 *__restrict__ atlas, uint32_t w, uint32_t h, uint32_t tile,
                                             uint32_t oct, float col, float row, float2 uv) {
    const float ox = col * (float)tile + 1.f + uv.x * (float)oct;
    const float oy = row * (float)tile + 1.f + uv.y * (float)oct;
    return bilinear(atlas, w, h, ox / (float)w, oy / (float)h);
}

// _sampleCascade + _tapEW for one cascade
__device__ float3 sample_cascade(const SampleCascade &s, const SampleArgs &a, float3 P, float3 Nn, float2 octN) {
    const DevCascade &c = s.c;
    float gf[3], base[3], fr[3];
    const float pp[3] = {P.x, P.y, P.z};
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const float cell = c.gsize[k] / fmaxf((float)c.res[k] - 1.f, 1.f);
        gf[k] = (pp[k] - c.gmin[k]) / fmaxf(cell, 1e-6f);
        base[k] = fminf(fmaxf(floorf(gf[k]), 0.f), fmaxf((float)c.res[k] - 2.f, 0.f));
        fr[k] = fminf(fmaxf(gf[k] - base[k], 0.f), 1.f);
    }
    const uint32_t tile_d = c.depth_oct == 6 ? 8u : c.depth_oct + 2u, oct_d = c.depth_oct;
    float3 acc = f3(0.f, 0.f, 0.f);
    float wsum = 0.f;
    for (int i = 0; i < 8; i++) {
        const int dx = i & 1, dy = (i >> 1) & 1, dz = (i >> 2) & 1;
        // a 1-probe axis has no second layer: its tap reads the same probe
        const float px = fminf(base[0] + (float)dx, (float)c.res[0] - 1.f);
        const float py = fminf(base[1] + (float)dy, (float)c.res[1] - 1.f);
        const float pz = fminf(base[2] + (float)dz, (float)c.res[2] - 1.f);
        const float wx = dx ? fr[0] : 1.f - fr[0];
        const float wy = dy ? fr[1] : 1.f - fr[1];
        const float wz = dz ? fr[2] : 1.f - fr[2];
        const float wtri = wx * wy * wz + 1e-4f;
        const float col = px, row = pz * (float)c.res[1] + py;
        const uint32_t probe = (uint32_t)px + (uint32_t)py * c.res[0] + (uint32_t)pz * c.res[0] * c.res[1];
        const float4 m = reinterpret_cast<const float4 *>(s.meta)[probe];
        const float3 reloc = f3(m.y, m.z, m.w) * a.classify_strength;
        const float3 ppos = f3(c.gmin[0] + (px / fmaxf((float)c.res[0] - 1.f, 1.f)) * c.gsize[0],
                               c.gmin[1] + (py / fmaxf((float)c.res[1] - 1.f, 1.f)) * c.gsize[1],
                               c.gmin[2] + (pz / fmaxf((float)c.res[2] - 1.f, 1.f)) * c.gsize[2]) +
                            reloc;
        const float3 to_surf = P - ppos;
        const float dist = sqrtf(dot3(to_surf, to_surf));
        const float dd = fmaxf(dist, 1e-6f);
        const float2 oct_d_uv = oct_encode(f3(to_surf.x / dd, to_surf.y / dd, to_surf.z / dd));
        const float3 mom = fetch_tile(s.datlas, c.daw, c.dah, tile_d, oct_d, col, row, oct_d_uv);
        const float m1 = mom.x, m2 = mom.y;
        const float variance = fabsf(m2 - m1 * m1);
        const float dm = fmaxf((dist - m1) - s.cheby_bias, 0.f);
        const float cheby_raw = variance / fmaxf(variance + dm * dm, 1e-6f);
        const float cheby = dist <= m1 + s.cheby_bias ? 1.f : cheby_raw;
        const float vis_w = mixf(1.f, fmaxf((cheby * cheby) * cheby, 0.05f), a.cheby_strength);
        const float3 to_probe = f3((ppos.x - P.x) / dd, (ppos.y - P.y) / dd, (ppos.z - P.z) / dd);
        const float wrap = dot3(to_probe, Nn) * 0.5f + 0.5f;
        const float wrap_w = mixf(1.f, fmaxf(wrap * wrap, 0.05f), a.cheby_strength);
        const float state = mixf(1.f, m.x, a.classify_strength);
        const float w = ((wtri * wrap_w) * vis_w) * state;
        const float3 e = fetch_tile(s.atlas, c.aw, c.ah, 8u, 6u, col, row, octN);
        acc = acc + e * w;
        wsum += w;
    }
    const float ws = fmaxf(wsum, 1e-4f);
    return f3(acc.x / ws, acc.y / ws, acc.z / ws);
}

// 6 waves per SIMD: 78 VGPRs, none spilled (99 and 4 waves without the hint).
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) k_sample(SampleArgs a, const float *__restrict__ pos,
                                                const float *__restrict__ nvis, const float *__restrict__ ndir,
                                                uint32_t n, float *__restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float3 X = f3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    const float3 Nv = f3(nvis[3 * i], nvis[3 * i + 1], nvis[3 * i + 2]);
    const float3 Nd = ndir ? f3(ndir[3 * i], ndir[3 * i + 1], ndir[3 * i + 2]) : Nv;
    const float3 Nn = normalize3(Nv);
    const float3 P = X + Nn * a.normal_bias;
    const float2 octN = oct_encode(normalize3(Nd));
    // fine-cascade weights (sampleIrradiance, js/gi_probes.js:829-846)
    float wf[SBGI_MAX_CASCADES] = {0.f, 0.f, 0.f, 0.f};
    for (uint32_t c = 1; c < a.n_casc; c++) {
        const DevCascade &g = a.casc[c].c;
        const float fx = (P.x - g.gmin[0]) / fmaxf(g.gsize[0], 1e-6f);
        const float fy = (P.y - g.gmin[1]) / fmaxf(g.gsize[1], 1e-6f);
        const float fz = (P.z - g.gmin[2]) / fmaxf(g.gsize[2], 1e-6f);
        const float lo = fminf(fminf(fx, fy), fz);
        const float hi = fminf(fminf(1.f - fx, 1.f - fy), 1.f - fz);
        wf[c] = fminf(fmaxf(fminf(lo, hi) / fmaxf(a.border_band, 1e-4f), 0.f), 1.f);
    }
    float3 E = f3(0.f, 0.f, 0.f);
    if (a.n_casc < 2 || wf[1] < 1.f) E = sample_cascade(a.casc[0], a, P, Nn, octN);
    for (uint32_t c = 1; c < a.n_casc; c++) {
        if (!(wf[c] > 0.f)) continue;
        const float3 Ec = sample_cascade(a.casc[c], a, P, Nn, octN);
        E = f3(mixf(E.x, Ec.x, wf[c]), mixf(E.y, Ec.y, wf[c]), mixf(E.z, Ec.z, wf[c]));
    }
    out[3 * i] = E.x;
    out[3 * i + 1] = E.y;
    out[3 * i + 2] = E.z;
}

}  // namespace

hipError_t launch_sample(const SampleCascadeIn *casc, uint32_t n_casc, const sbgi_sample_params &p, const float *pos,
                         const float *nvis, const float *ndir, uint32_t n, float *out, hipStream_t st) {
    if (!n) return hipSuccess;
    SampleArgs a;
    memset(&a, 0, sizeof(a));
    a.n_casc = n_casc;
    for (uint32_t c = 0; c < n_casc; c++) {
        a.casc[c].c = casc[c].c;
        a.casc[c].atlas = casc[c].atlas;
        a.casc[c].datlas = casc[c].datlas;
        a.casc[c].meta = casc[c].meta;
        a.casc[c].cheby_bias = p.cheby_bias[c];
    }
    a.normal_bias = p.normal_bias;
    a.cheby_strength = p.cheby_strength;
    a.classify_strength = p.classify_strength;
    a.border_band = p.border_band;
    hipLaunchKernelGGL(k_sample, dim3((n + 255) / 256), dim3(256), 0, st, a, pos, nvis, ndir, n, out);
    return hipGetLastError();
}
