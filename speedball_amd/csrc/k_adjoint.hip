// k_adjoint.hip — hand-written adjoint of the differentiable two-update solve.
//
// The differentiable solve is two first-updates from cleared state (no history, so each
// update's irradiance is the cosine-weighted mean of its rays), the second one bouncing
// off the first's unrounded, unfiltered FP32 atlas A1. Its output is the second update's
// interior irradiance E2. Parameters: the per-instance-triangle albedo (it sets kd) and
// one scale per light (on its colour x intensity). No reference counterpart: the
// reference is not differentiable.
//
// The hit sets are frozen from the forward pass: both updates trace the same rays
// through the same geometry, so the second update's closest hits and light-visibility
// masks serve both, and every shading term is recomputed from them rather than stored.
//
//   gL  = W^T gE2                      per ray (k_adj_rays)
//   bounce of update 2 -> g kd, gB     per ray (k_adj_bounce), gB kept per ray
//   gA1 = S^T gB                       gathered per probe (k_adj_gather): hits are binned
//                                      by the trilinear cell they sample, and each probe
//                                      accumulates the cells around it in LDS, folding the
//                                      gutter copies into their interior source texels
//   gL += W^T gE1                      (k_adj_rays, accumulate)
//   direct light of both updates       per ray (k_adj_direct) -> g albedo, g light scale
// where W is the blend's cosine weight matrix (divided by its normaliser) and S the
// bounce lookup's trilinear x bilinear weights.
//
// Atomics: the albedo gradient uses float atomics (each triangle is hit by few rays),
// the light-scale gradient is reduced per block in LDS and then across blocks without
// atomics, and the atlas gradient uses LDS atomics only: the gather replaces the ~100
// global scatter atomics per hit that a direct S^T would need.
#pragma clang fp contract(off)

#include "kernels.h"
#include "shade.h"

namespace {

constexpr uint32_t kAdjBlock = 256;
constexpr float kLuma[3] = {0.2126f, 0.7152f, 0.0722f};

// gL[gid] (+)= sum_t W[k][t] / max(norm[t], 1e-4) * gE[probe][t]   (ray gid = k * P + probe)
__global__ void __launch_bounds__(kAdjBlock) k_adj_rays(DevCascade c, BlendTables t, const float *__restrict__ gE,
                                                        float4 *__restrict__ gL, int accumulate) {
    const uint32_t gid = blockIdx.x * kAdjBlock + threadIdx.x;
    const uint32_t n = c.updated * c.rays;
    if (gid >= n) return;
    const uint32_t k = gid / c.updated, probe = gid - k * c.updated;
    const float *w = t.w_irr_t + (size_t)k * 36;
    const float *g = gE + (size_t)probe * 108;
    float r = 0.f, gg = 0.f, b = 0.f;
    for (uint32_t tt = 0; tt < 36; tt++) {
        const float f = w[tt] / fmaxf(t.norm_irr[tt], 1e-4f);
        r += f * g[tt * 3 + 0];
        gg += f * g[tt * 3 + 1];
        b += f * g[tt * 3 + 2];
    }
    float4 o = make_float4(r, gg, b, 0.f);
    if (accumulate) {
        const float4 prev = gL[gid];
        o = make_float4(prev.x + r, prev.y + gg, prev.z + b, 0.f);
    }
    gL[gid] = o;
}

// d kd / d albedo (1 inside the clamp of kd = clamp(albedo * fac, 0, 0.95), else 0) x fac
__device__ __forceinline__ float3 kd_slope(const DevScene &s, const sbgi_tick_params &p, uint32_t tri, uint32_t inst,
                                           float3 &base) {
    const float4 B = s.tri[(size_t)tri * 3 + 1];
    const float *mat = s.mats + (size_t)__float_as_uint(B.w) * SBGI_MAT_STRIDE;
    const float fac = p.dbg_albedo * (1.f - mat[4]) * (1.f - mat[5]);
    const float *a = s.albedo + (size_t)albedo_index(s, inst, tri) * 3;
    base = f3(a[0], a[1], a[2]);
    const float vx = base.x * fac, vy = base.y * fac, vz = base.z * fac;
    return f3((vx > 0.f && vx < 0.95f) ? fac : 0.f, (vy > 0.f && vy < 0.95f) ? fac : 0.f,
              (vz > 0.f && vz < 0.95f) ? fac : 0.f);
}

// trilinear cell of a bounce lookup at P (the `base` corner of sample_atlas)
__device__ __forceinline__ uint32_t bounce_cell(const DevCascade &c, float3 P) {
    const float pp[3] = {P.x, P.y, P.z};
    uint32_t b[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const float cell = c.gsize[a] / fmaxf(1.f, (float)c.res[a] - 1.f);
        const float gf = (pp[a] - c.gmin[a]) / fmaxf(cell, 1e-6f);
        const float hi = fmaxf((float)c.res[a] - 2.f, 0.f);
        b[a] = (uint32_t)fminf(fmaxf(floorf(gf), 0.f), hi);
    }
    const uint32_t cx = max(c.res[0], 2u) - 1u, cy = max(c.res[1], 2u) - 1u;
    return b[0] + cx * (b[1] + cy * b[2]);
}

// Bounce term of update 2 (js/gi_probes.js:2181-2188 with the luminance roll-off):
// u = B * kd / pi, out = u * roll(u) * dbg_bounce. Writes g albedo (atomics), the lookup
// gradient gB and the hit's lookup point, and counts the hit into its cell.
// 6 waves per SIMD: 79 VGPRs, none spilled (115 and 4 waves without the hint).
__global__ void __launch_bounds__(kAdjBlock) __attribute__((amdgpu_waves_per_eu(6))) k_adj_bounce(DevScene s, DevCascade c, sbgi_tick_params p,
                                                          const float4 *__restrict__ dirs, const float4 *__restrict__ hits,
                                                          const float4 *__restrict__ atlas1, const float4 *__restrict__ gL,
                                                          float *__restrict__ g_albedo, float4 *__restrict__ gB,
                                                          float4 *__restrict__ gpos, uint32_t *__restrict__ cell_count) {
    const uint32_t gid = blockIdx.x * kAdjBlock + threadIdx.x;
    const uint32_t n = c.updated * c.rays;
    if (gid >= n) return;
    const float4 hr = hits[gid];
    const int tri = __float_as_int(hr.y), inst = __float_as_int(hr.z);
    if (tri < 0) {
        gB[gid] = make_float4(0.f, 0.f, 0.f, __uint_as_float(0xFFFFFFFFu));
        return;
    }
    float3 ro, rd;
    make_ray(c, 0.f, dirs, nullptr, gid, ro, rd);
    const HitFrame h = hit_frame(s, c, p, ro, rd, hr.x, (uint32_t)tri, (uint32_t)inst);
    const float3 B = sample_atlas(c, atlas1, h.hit, h.ng);
    const float3 u = B * h.kd * (1.0f / SBGI_PI);
    const float lum = u.x * kLuma[0] + u.y * kLuma[1] + u.z * kLuma[2];
    const float rc = p.radiance_clamp;
    const float roll = rc / fmaxf(fmaxf(rc, lum), 1e-6f);
    const float4 g = gL[gid];
    const float3 gob = f3(g.x, g.y, g.z) * p.dbg_bounce;
    const float gdot = gob.x * u.x + gob.y * u.y + gob.z * u.z;
    const bool clamped = lum > rc && lum > 1e-6f;
    const float k2 = clamped ? rc / (lum * lum) * gdot : 0.f;
    const float3 gu = f3(roll * gob.x - k2 * kLuma[0], roll * gob.y - k2 * kLuma[1], roll * gob.z - k2 * kLuma[2]);
    float3 base;
    const float3 slope = kd_slope(s, p, (uint32_t)tri, (uint32_t)inst, base);
    float *ga = g_albedo + (size_t)albedo_index(s, (uint32_t)inst, (uint32_t)tri) * 3;
    atomicAdd(ga + 0, gu.x * B.x * (1.0f / SBGI_PI) * slope.x);
    atomicAdd(ga + 1, gu.y * B.y * (1.0f / SBGI_PI) * slope.y);
    atomicAdd(ga + 2, gu.z * B.z * (1.0f / SBGI_PI) * slope.z);
    const float3 gb = gu * h.kd * (1.0f / SBGI_PI);
    const uint32_t cell = bounce_cell(c, h.hit);
    const float2 uv = oct_encode(normalize3(h.ng));
    gB[gid] = make_float4(gb.x, gb.y, gb.z, __uint_as_float(cell));
    // the lookup point for the gather: position and octahedral uv (gpos[n + gid].x = v)
    gpos[gid] = make_float4(h.hit.x, h.hit.y, h.hit.z, uv.x);
    gpos[n + gid] = make_float4(uv.y, 0.f, 0.f, 0.f);
    atomicAdd(cell_count + cell, 1u);
}

// exclusive scan of the cell counts (one block), resetting the fill cursors
__global__ void __launch_bounds__(1024) k_adj_scan(const uint32_t *__restrict__ count, uint32_t *__restrict__ off,
                                                   uint32_t *__restrict__ cursor, uint32_t cells) {
    __shared__ uint32_t part[1024];
    const uint32_t per = (cells + 1023) / 1024;
    const uint32_t lo = threadIdx.x * per, hi = min(lo + per, cells);
    uint32_t s = 0;
    for (uint32_t i = lo; i < hi; i++) s += count[i];
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (uint32_t i = 0; i < 1024; i++) {
            const uint32_t v = part[i];
            part[i] = acc;
            acc += v;
        }
    }
    __syncthreads();
    uint32_t acc = part[threadIdx.x];
    for (uint32_t i = lo; i < hi; i++) {
        off[i] = acc;
        cursor[i] = 0u;
        acc += count[i];
    }
}

__global__ void __launch_bounds__(kAdjBlock) k_adj_fill(DevCascade c, const float4 *__restrict__ gB,
                                                        const uint32_t *__restrict__ off, uint32_t *__restrict__ cursor,
                                                        uint32_t *__restrict__ list) {
    const uint32_t gid = blockIdx.x * kAdjBlock + threadIdx.x;
    if (gid >= c.updated * c.rays) return;
    const uint32_t cell = __float_as_uint(gB[gid].w);
    if (cell == 0xFFFFFFFFu) return;
    list[off[cell] + atomicAdd(cursor + cell, 1u)] = gid;
}

// One 64-lane block per probe: accumulate S^T gB over the hits of the (up to) 8 cells
// that have this probe as a corner, in LDS over the 8x8 tile, then fold the gutter texels
// into their interior source: gE1[probe][t] (3 floats per texel).
__global__ void __launch_bounds__(64) k_adj_gather(DevCascade c, const float4 *__restrict__ gB,
                                                   const float4 *__restrict__ gpos, const uint32_t *__restrict__ off,
                                                   const uint32_t *__restrict__ count, const uint32_t *__restrict__ list,
                                                   float *__restrict__ gE1) {
    __shared__ float acc[64 * 3];
    __shared__ float fold[36 * 3];
    const uint32_t probe = blockIdx.x, lane = threadIdx.x;
    const uint32_t n = c.updated * c.rays;
    for (uint32_t i = lane; i < 64 * 3; i += 64) acc[i] = 0.f;
    for (uint32_t i = lane; i < 36 * 3; i += 64) fold[i] = 0.f;
    __syncthreads();
    const uint32_t px = probe % c.res[0], py = (probe / c.res[0]) % c.res[1], pz = probe / (c.res[0] * c.res[1]);
    const uint32_t cx = max(c.res[0], 2u) - 1u, cy = max(c.res[1], 2u) - 1u, cz = max(c.res[2], 2u) - 1u;
    const float col = (float)px, row = (float)pz * (float)c.res[1] + (float)py;
    for (int corner = 0; corner < 8; corner++) {
        const int dx = corner & 1, dy = (corner >> 1) & 1, dz = (corner >> 2) & 1;
        const int bx = (int)px - dx, by = (int)py - dy, bz = (int)pz - dz;
        if (bx < 0 || by < 0 || bz < 0 || bx >= (int)cx || by >= (int)cy || bz >= (int)cz) continue;
        // the cell's base corner must map back to this probe (res-1 axes clamp to one cell)
        const uint32_t cell = (uint32_t)bx + cx * ((uint32_t)by + cy * (uint32_t)bz);
        const uint32_t o = off[cell], m = count[cell];
        for (uint32_t j = lane; j < m; j += 64) {
            const uint32_t gid = list[o + j];
            const float4 g = gB[gid];
            const float4 ps = gpos[gid];
            const float uvx = ps.w, uvy = gpos[n + gid].x;
            // trilinear weights exactly as sample_atlas forms them
            const float pp[3] = {ps.x, ps.y, ps.z};
            float fr[3];
#pragma unroll
            for (int a = 0; a < 3; a++) {
                const float cellw = c.gsize[a] / fmaxf(1.f, (float)c.res[a] - 1.f);
                const float gf = (pp[a] - c.gmin[a]) / fmaxf(cellw, 1e-6f);
                const float hi = fmaxf((float)c.res[a] - 2.f, 0.f);
                const float base = fminf(fmaxf(floorf(gf), 0.f), hi);
                fr[a] = fminf(fmaxf(gf - base, 0.f), 1.f);
            }
            float wsum = 0.f, wme = 0.f;
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const int ex = i & 1, ey = (i >> 1) & 1, ez = (i >> 2) & 1;
                const float wx = ex ? fr[0] : 1.f - fr[0];
                const float wy = ey ? fr[1] : 1.f - fr[1];
                const float wz = ez ? fr[2] : 1.f - fr[2];
                const float w = wx * wy * wz + 1e-4f;
                wsum += w;
                if (i == corner) wme = w;
            }
            const float pw = wme * (1.0f / fmaxf(wsum, 1e-4f));
            // bilinear footprint in this probe's tile (sample_atlas -> bilinear)
            const float ox = col * 8.f + 1.f + uvx * 6.f, oy = row * 8.f + 1.f + uvy * 6.f;
            const float x = (ox / (float)c.aw) * (float)c.aw - 0.5f, y = (oy / (float)c.ah) * (float)c.ah - 0.5f;
            const float x0f = floorf(x), y0f = floorf(y);
            const float fx = x - x0f, fy = y - y0f;
            int x0 = (int)x0f, y0 = (int)y0f, x1 = x0 + 1, y1 = y0 + 1;
            x0 = min(max(x0, 0), (int)c.aw - 1);
            x1 = min(max(x1, 0), (int)c.aw - 1);
            y0 = min(max(y0, 0), (int)c.ah - 1);
            y1 = min(max(y1, 0), (int)c.ah - 1);
            const int tx0 = x0 - (int)px * 8, tx1 = x1 - (int)px * 8, ty0 = y0 - (int)row * 8, ty1 = y1 - (int)row * 8;
            const int tx[4] = {tx0, tx1, tx0, tx1}, ty[4] = {ty0, ty0, ty1, ty1};
            const float bw[4] = {(1.f - fx) * (1.f - fy), fx * (1.f - fy), (1.f - fx) * fy, fx * fy};
#pragma unroll
            for (int q = 0; q < 4; q++) {
                if (tx[q] < 0 || tx[q] > 7 || ty[q] < 0 || ty[q] > 7) continue;  // never: footprint is in-tile
                const float wq = pw * bw[q];
                float *a = acc + (ty[q] * 8 + tx[q]) * 3;
                atomicAdd(a + 0, wq * g.x);
                atomicAdd(a + 1, wq * g.y);
                atomicAdd(a + 2, wq * g.z);
            }
        }
    }
    __syncthreads();
    {
        const uint32_t lx = lane % 8, ly = lane / 8;
        const uint32_t edge = 7, lo = 1, hi = 6;
        const bool on_l = lx == 0, on_r = lx == edge, on_t = ly == 0, on_b = ly == edge;
        const bool on_col = on_l || on_r, on_row = on_t || on_b, on_corner = on_col && on_row;
        const uint32_t sx = on_corner ? (on_l ? hi : lo) : (on_row ? edge - lx : (on_col ? (on_l ? lo : hi) : lx));
        const uint32_t sy = on_corner ? (on_t ? hi : lo) : (on_row ? (on_t ? lo : hi) : (on_col ? edge - ly : ly));
        const uint32_t t = (sy - 1) * 6 + (sx - 1);
        atomicAdd(fold + t * 3 + 0, acc[lane * 3 + 0]);
        atomicAdd(fold + t * 3 + 1, acc[lane * 3 + 1]);
        atomicAdd(fold + t * 3 + 2, acc[lane * 3 + 2]);
    }
    __syncthreads();
    for (uint32_t i = lane; i < 36 * 3; i += 64) gE1[(size_t)probe * 108 + i] = fold[i];
}

// Direct light of both updates with g = gL1 + gL2: contribution = kd * A * scale for each
// visible light, A = ndl * lcol0 * att * dbg_direct / pi. Albedo gradient by atomics;
// light-scale gradient reduced per block into part[block][light].
__global__ void __launch_bounds__(kAdjBlock) k_adj_direct(DevScene s, DevCascade c, sbgi_tick_params p,
                                                          const float4 *__restrict__ dirs, const float4 *__restrict__ hits,
                                                          const unsigned long long *__restrict__ vis,
                                                          const float4 *__restrict__ gL, float *__restrict__ g_albedo,
                                                          float *__restrict__ part) {
    __shared__ float gl[SBGI_MAX_LIGHTS];
    for (uint32_t i = threadIdx.x; i < SBGI_MAX_LIGHTS; i += kAdjBlock) gl[i] = 0.f;
    __syncthreads();
    const uint32_t gid = blockIdx.x * kAdjBlock + threadIdx.x;
    const uint32_t n = c.updated * c.rays;
    const float4 hr = gid < n ? hits[gid] : make_float4(0.f, __int_as_float(-1), 0.f, 0.f);
    const int tri = __float_as_int(hr.y), inst = __float_as_int(hr.z);
    if (gid < n && tri >= 0) {
        float3 ro, rd;
        make_ray(c, 0.f, dirs, nullptr, gid, ro, rd);
        const HitFrame h = hit_frame(s, c, p, ro, rd, hr.x, (uint32_t)tri, (uint32_t)inst);
        const float4 g4 = gL[gid];
        const float3 g = f3(g4.x, g4.y, g4.z);
        const unsigned long long vm = vis[gid];
        float3 gkd = f3(0.f, 0.f, 0.f);
        for (uint32_t li = 0; li < s.light_count; li++) {
            if (!((vm >> li) & 1ull)) continue;
            const float *L = s.lights + li * SBGI_LIGHT_STRIDE;
            const float sc = light_scale(s, li);
            const LightSample ls = light_gate(L, h, p, sc);
            if (!ls.pass) continue;
            // light_contrib with kd = 1 and the scale divided out of lcol
            HitFrame unit = h;
            unit.kd = f3(1.f, 1.f, 1.f);
            LightSample l0 = ls;
            l0.lcol = ls.lcol * (sc != 0.f ? 1.0f / sc : 0.f);
            const float3 A = light_contrib(L, unit, l0, p);
            gkd = gkd + g * A * sc;
            atomicAdd(&gl[li], g.x * h.kd.x * A.x + g.y * h.kd.y * A.y + g.z * h.kd.z * A.z);
        }
        float3 base;
        const float3 slope = kd_slope(s, p, (uint32_t)tri, (uint32_t)inst, base);
        float *ga = g_albedo + (size_t)albedo_index(s, (uint32_t)inst, (uint32_t)tri) * 3;
        atomicAdd(ga + 0, gkd.x * slope.x);
        atomicAdd(ga + 1, gkd.y * slope.y);
        atomicAdd(ga + 2, gkd.z * slope.z);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < s.light_count; i += kAdjBlock) part[(size_t)blockIdx.x * s.light_count + i] = gl[i];
}

__global__ void k_adj_reduce(const float *__restrict__ part, uint32_t blocks, uint32_t lights, float *__restrict__ out) {
    const uint32_t li = blockIdx.x * blockDim.x + threadIdx.x;
    if (li >= lights) return;
    float s = 0.f;
    for (uint32_t b = 0; b < blocks; b++) s += part[(size_t)b * lights + li];
    out[li] = s;
}

}  // namespace

size_t adjoint_scratch_bytes(uint32_t rays_total, uint32_t cells, uint32_t lights) {
    const size_t blocks = (rays_total + kAdjBlock - 1) / kAdjBlock;
    return (size_t)rays_total * 16 * 4          // gL, gB, gpos (2 x)
           + (size_t)cells * 4 * 3              // count, off, cursor
           + (size_t)rays_total * 4             // list
           + blocks * (lights ? lights : 1) * 4 + 4096;  // + alignment of the 8 carved arrays
}

hipError_t launch_adjoint(const DevScene &s, const DevCascade &c, const sbgi_tick_params &p, const float4 *dirs,
                          const BlendTables &t, const TraceScratch &w, const float4 *atlas1, const float *gE2,
                          float *gE1, float *g_albedo, float *g_light, void *scratch, hipStream_t st) {
    const uint32_t n = c.updated * c.rays;
    if (n == 0) return hipSuccess;
    const uint32_t cells = (max(c.res[0], 2u) - 1u) * (max(c.res[1], 2u) - 1u) * (max(c.res[2], 2u) - 1u);
    const uint32_t blocks = (n + kAdjBlock - 1) / kAdjBlock;
    char *q = static_cast<char *>(scratch);
    auto take = [&](size_t bytes) {
        void *r = q;
        q += (bytes + 255) & ~size_t(255);
        return r;
    };
    float4 *gL = static_cast<float4 *>(take((size_t)n * 16));
    float4 *gB = static_cast<float4 *>(take((size_t)n * 16));
    float4 *gpos = static_cast<float4 *>(take((size_t)n * 32));
    uint32_t *count = static_cast<uint32_t *>(take((size_t)cells * 4));
    uint32_t *off = static_cast<uint32_t *>(take((size_t)cells * 4));
    uint32_t *cursor = static_cast<uint32_t *>(take((size_t)cells * 4));
    uint32_t *list = static_cast<uint32_t *>(take((size_t)n * 4));
    float *part = static_cast<float *>(take((size_t)blocks * (s.light_count ? s.light_count : 1) * 4));
    hipError_t e;
    if ((e = hipMemsetAsync(count, 0, (size_t)cells * 4, st)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_adj_rays, dim3(blocks), dim3(kAdjBlock), 0, st, c, t, gE2, gL, 0);
    hipLaunchKernelGGL(k_adj_bounce, dim3(blocks), dim3(kAdjBlock), 0, st, s, c, p, dirs, w.hits, atlas1, gL, g_albedo,
                       gB, gpos, count);
    hipLaunchKernelGGL(k_adj_scan, dim3(1), dim3(1024), 0, st, count, off, cursor, cells);
    hipLaunchKernelGGL(k_adj_fill, dim3(blocks), dim3(kAdjBlock), 0, st, c, gB, off, cursor, list);
    hipLaunchKernelGGL(k_adj_gather, dim3(c.updated), dim3(64), 0, st, c, gB, gpos, off, count, list, gE1);
    hipLaunchKernelGGL(k_adj_rays, dim3(blocks), dim3(kAdjBlock), 0, st, c, t, gE1, gL, 1);
    hipLaunchKernelGGL(k_adj_direct, dim3(blocks), dim3(kAdjBlock), 0, st, s, c, p, dirs, w.hits, w.vis, gL, g_albedo,
                       part);
    if (s.light_count)
        hipLaunchKernelGGL(k_adj_reduce, dim3(1), dim3(64), 0, st, part, blocks, s.light_count, g_light);
    return hipGetLastError();
}
