// shade.h — hit shading shared by the forward trace and the adjoint: ray origin and
// direction, the geometric shading frame, the NEE light gate and contribution, and the
// previous-atlas bounce lookup (js/gi_probes.js:1714-1726, 1876-1885, 1941-2188).
//
// Include it after `#pragma clang fp contract(off)`: the adjoint recomputes the forward's
// values and must round them the same way. (The probe-origin, light-list and emitter
// helpers turn contraction off themselves: the blend, which contracts, calls them too.)
#pragma once
#include <hip/hip_fp16.h>

#include "sbgi_dev.h"

// scale of light li's colour x intensity (differentiable mode), 1 otherwise
__device__ __forceinline__ float light_scale(const DevScene &s, uint32_t li) {
    return s.light_scale ? s.light_scale[li] : 1.f;
}

__device__ __forceinline__ float h2f(unsigned short h) { return __half2float(__ushort_as_half(h)); }

// software bilinear of the previous RGBA16F atlas: texel-centre convention, clamp to edge
__device__ __forceinline__ float3 bilinear(const ushort4 *__restrict__ atlas, uint32_t w, uint32_t h, float u, float v) {
    float x = u * (float)w - 0.5f, y = v * (float)h - 0.5f;
    float x0f = floorf(x), y0f = floorf(y);
    float fx = x - x0f, fy = y - y0f;
    int x0 = (int)x0f, y0 = (int)y0f, x1 = x0 + 1, y1 = y0 + 1;
    x0 = min(max(x0, 0), (int)w - 1);
    x1 = min(max(x1, 0), (int)w - 1);
    y0 = min(max(y0, 0), (int)h - 1);
    y1 = min(max(y1, 0), (int)h - 1);
    ushort4 t00 = atlas[(size_t)y0 * w + x0], t10 = atlas[(size_t)y0 * w + x1];
    ushort4 t01 = atlas[(size_t)y1 * w + x0], t11 = atlas[(size_t)y1 * w + x1];
    float w00 = (1.f - fx) * (1.f - fy), w10 = fx * (1.f - fy), w01 = (1.f - fx) * fy, w11 = fx * fy;
    return f3(h2f(t00.x) * w00 + h2f(t10.x) * w10 + h2f(t01.x) * w01 + h2f(t11.x) * w11,
              h2f(t00.y) * w00 + h2f(t10.y) * w10 + h2f(t01.y) * w01 + h2f(t11.y) * w11,
              h2f(t00.z) * w00 + h2f(t10.z) * w10 + h2f(t01.z) * w01 + h2f(t11.z) * w11);
}

__device__ __forceinline__ float2 oct_encode(float3 n) {
    float s = fabsf(n.x) + fabsf(n.y) + fabsf(n.z);
    float inv = 1.0f / fmaxf(s, 1e-20f);
    float ox = n.x * inv, oy = n.y * inv;
    if (n.z < 0.f) {
        float fx = (1.f - fabsf(oy)) * (ox >= 0.f ? 1.f : -1.f);
        float fy = (1.f - fabsf(ox)) * (oy >= 0.f ? 1.f : -1.f);
        ox = fx;
        oy = fy;
    }
    return make_float2(ox * 0.5f + 0.5f, oy * 0.5f + 0.5f);
}

// software bilinear of an FP32 atlas (the differentiable mode's unrounded bounce source)
__device__ __forceinline__ float3 bilinear(const float4 *__restrict__ atlas, uint32_t w, uint32_t h, float u, float v) {
    float x = u * (float)w - 0.5f, y = v * (float)h - 0.5f;
    float x0f = floorf(x), y0f = floorf(y);
    float fx = x - x0f, fy = y - y0f;
    int x0 = (int)x0f, y0 = (int)y0f, x1 = x0 + 1, y1 = y0 + 1;
    x0 = min(max(x0, 0), (int)w - 1);
    x1 = min(max(x1, 0), (int)w - 1);
    y0 = min(max(y0, 0), (int)h - 1);
    y1 = min(max(y1, 0), (int)h - 1);
    const float4 t00 = atlas[(size_t)y0 * w + x0], t10 = atlas[(size_t)y0 * w + x1];
    const float4 t01 = atlas[(size_t)y1 * w + x0], t11 = atlas[(size_t)y1 * w + x1];
    float w00 = (1.f - fx) * (1.f - fy), w10 = fx * (1.f - fy), w01 = (1.f - fx) * fy, w11 = fx * fy;
    return f3(t00.x * w00 + t10.x * w10 + t01.x * w01 + t11.x * w11, t00.y * w00 + t10.y * w10 + t01.y * w01 + t11.y * w11,
              t00.z * w00 + t10.z * w10 + t01.z * w01 + t11.z * w11);
}

// sampleAtlas (js/gi_probes.js:1941-1974); ATLAS is the RGBA16F atlas (ushort4) or the
// differentiable mode's FP32 atlas (float4)
template <class ATLAS>
__device__ float3 sample_atlas(const DevCascade &c, const ATLAS *__restrict__ atlas, float3 P, float3 N) {
    float cell[3], gf[3], base[3], fr[3];
    float pp[3] = {P.x, P.y, P.z};
#pragma unroll
    for (int a = 0; a < 3; a++) {
        cell[a] = c.gsize[a] / fmaxf(1.f, (float)c.res[a] - 1.f);
        gf[a] = (pp[a] - c.gmin[a]) / fmaxf(cell[a], 1e-6f);
        float hi = fmaxf((float)c.res[a] - 2.f, 0.f);
        base[a] = fminf(fmaxf(floorf(gf[a]), 0.f), hi);
        fr[a] = fminf(fmaxf(gf[a] - base[a], 0.f), 1.f);
    }
    float2 uv = oct_encode(normalize3(N));
    float3 acc = f3(0.f, 0.f, 0.f);
    float wsum = 0.f;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        int dx = i & 1, dy = (i >> 1) & 1, dz = (i >> 2) & 1;
        float px = base[0] + dx, py = base[1] + dy, pz = base[2] + dz;
        float wx = dx ? fr[0] : 1.f - fr[0];
        float wy = dy ? fr[1] : 1.f - fr[1];
        float wz = dz ? fr[2] : 1.f - fr[2];
        float w = wx * wy * wz + 1e-4f;
        float col = px, row = pz * (float)c.res[1] + py;
        float ox = col * 8.f + 1.f + uv.x * 6.f;
        float oy = row * 8.f + 1.f + uv.y * 6.f;
        float3 t = bilinear(atlas, c.aw, c.ah, ox / (float)c.aw, oy / (float)c.ah);
        acc = acc + t * w;
        wsum += w;
    }
    return acc * (1.0f / fmaxf(wsum, 1e-4f));
}

// probeTraceOrigin of a probe: its grid position plus the relocation offset scaled by
// classifyStrength (js/gi_probes.js:1714-1726, 1876-1885)
__device__ __forceinline__ float3 probe_origin(const DevCascade &c, float classify_strength,
                                               const float *__restrict__ meta, uint32_t probe) {
#pragma clang fp contract(off)
    const uint32_t ix = probe % c.res[0], iy = (probe / c.res[0]) % c.res[1], iz = probe / (c.res[0] * c.res[1]);
    float3 ro = f3(c.gmin[0] + ((float)ix / fmaxf(1.f, (float)c.res[0] - 1.f)) * c.gsize[0],
                   c.gmin[1] + ((float)iy / fmaxf(1.f, (float)c.res[1] - 1.f)) * c.gsize[1],
                   c.gmin[2] + ((float)iz / fmaxf(1.f, (float)c.res[2] - 1.f)) * c.gsize[2]);
    if (classify_strength != 0.f) {
        const float *m = meta + (size_t)probe * 4;
        ro = ro + f3(m[1], m[2], m[3]) * classify_strength;
    }
    return ro;
}

// ray origin and direction of ray gid (index k * updated + slot)
__device__ __forceinline__ void make_ray(const DevCascade &c, float classify_strength, const float4 *__restrict__ dirs,
                                         const float *__restrict__ meta, uint32_t gid, float3 &ro, float3 &rd) {
    const uint32_t k = gid / c.updated;
    const uint32_t slot = gid - k * c.updated;
    ro = probe_origin(c, classify_strength, meta, (c.offset + slot) % c.total);
    rd = xyz(dirs[k]);
}

// Lights the NEE loop visits for a hit: every light, or with clustered lighting the list
// of the cascade-0 cell that holds the hit (js/gi_probes.js:2112-2127): all lights when
// the hit lies outside the grid or the cell's list overflowed. Bit li = light li.
__device__ __forceinline__ uint64_t light_members(const DevScene &s, float3 hit) {
#pragma clang fp contract(off)
    const uint64_t all = s.light_count >= 64 ? ~0ull : ((1ull << s.light_count) - 1ull);
    if (!s.light_grid) return all;
    const float pp[3] = {hit.x, hit.y, hit.z};
    float cf[3];
    bool outside = false;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const float cells = (float)(s.lg_cells[a] + 1u) - 1.f;
        const float size = s.lg_size[a] / cells;
        cf[a] = floorf((pp[a] - s.lg_min[a]) / size);
        outside = outside || cf[a] < 0.f || cf[a] >= cells;
        cf[a] = fminf(fmaxf(cf[a], 0.f), cells - 1.f);
    }
    const uint32_t cell = (uint32_t)cf[2] * (s.lg_cells[0] * s.lg_cells[1]) + (uint32_t)cf[1] * s.lg_cells[0] +
                          (uint32_t)cf[0];
    const uint32_t *list = s.light_grid + (size_t)cell * (SBGI_LIGHTS_PER_CELL + 1);
    const uint32_t n = list[0];
    if (outside || n > SBGI_LIGHTS_PER_CELL) return all;
    uint64_t m = 0ull;
    for (uint32_t i = 0; i < n; i++) m |= 1ull << list[1 + i];
    return m;
}

// The per-probe part of an emitter's virtual rays (injectEmitterVirtualRays,
// js/gi_probes.js:2262-2292): direction d from the probe's trace origin to the emitter
// centre, the expected ray count kVirtual = min(R, R * Omega / 4 pi) with the Cauchy solid
// angle Omega = min(2 pi, (A/4) / dist^2), and the radiance times the retained visibility.
// valid is false when the probe lies inside the emitter's bounding sphere.
struct EmitterTap {
    float3 d, L;
    float kv;
    bool valid;
};

__device__ __forceinline__ EmitterTap emitter_tap(const float *__restrict__ L, float3 ro, float rays, float vis) {
#pragma clang fp contract(off)
    EmitterTap t;
    const float3 axis = f3(L[1], L[2], L[3]) - ro;
    const float dist = sqrtf(dot3(axis, axis));
    const float radius = fmaxf(L[14], 0.f);
    t.valid = dist > fmaxf(radius, 1e-3f);
    t.d = f3(axis.x / dist, axis.y / dist, axis.z / dist);
    const float proj = fmaxf(L[15], 0.f);
    const float omega = fminf(6.2831853071795862f, proj / (dist * dist));
    t.kv = fminf(rays, (rays * omega) / 12.566370614359172f);
    t.L = f3(L[7], L[8], L[9]) * vis;
    return t;
}

// sin / cos rounded once from double: the correctly rounded f32 values, so that the
// emitter sample hash (which multiplies sin by ~4e4) is reproducible
__device__ __forceinline__ float sin_cr(float x) { return (float)sin((double)x); }
__device__ __forceinline__ float cos_cr(float x) { return (float)cos((double)x); }

// The visibility ray of emitterVisKernel (js/gi_probes.js:2216-2260): from the probe's
// trace origin toward a hashed point of the emitter's disc facing the probe, ending
// `bias` short of the emitter's bounding sphere. False when the probe lies inside it.
__device__ __forceinline__ bool emitter_ray(const float *__restrict__ L, float3 ro, uint32_t probe, uint32_t e,
                                            float seed, float bias, float3 &rd, float &tmax) {
#pragma clang fp contract(off)
    const float3 center = f3(L[1], L[2], L[3]);
    const float radius = fmaxf(L[14], 0.f);
    const float3 axis = center - ro;
    const float dist = sqrtf(dot3(axis, axis));
    if (!(dist > fmaxf(radius, 1e-3f))) return false;
    const float3 d = f3(axis.x / dist, axis.y / dist, axis.z / dist);
    const float3 up = fabsf(d.y) < 0.999f ? f3(0.f, 1.f, 0.f) : f3(1.f, 0.f, 0.f);
    const float3 tg = normalize3(cross3(up, d));
    const float3 bt = cross3(d, tg);
    const float es = ((float)probe * 12.9898f + (float)e * 78.233f) + seed * 37.719f;
    const float h0 = sin_cr(es) * 43758.5453f;
    const float h1 = sin_cr(es + 19.19f) * 24634.6345f;
    const float u = h0 - floorf(h0), v = h1 - floorf(h1);
    const float rad = sqrtf(u) * radius;
    const float ang = v * (SBGI_PI * 2.f);
    const float3 target = center + (tg * (cos_cr(ang) * rad) + bt * (sin_cr(ang) * rad));
    rd = normalize3(target - ro);
    tmax = fmaxf((dist - radius) - bias, 1e-4f);
    return true;
}

// Unit vector and distance from pos to light record L (a directional light: -dir at 1e4).
// Evaluated without FMA contraction, so the shadow walk re-derives a queued ray's direction
// and extent from its origin word for word as the emission computed them. The direction
// is tl times the hardware reciprocal square root (about 1 ulp; three IEEE divisions cost
// ~30 instructions here, on every light of every hit); the distance keeps the correctly
// rounded square root.
__device__ __forceinline__ void light_vector(const float *__restrict__ L, float3 pos, float3 &wi, float &dist) {
#pragma clang fp contract(off)
    const bool is_dir = L[0] < 0.5f;
    const float3 tl = is_dir ? f3(L[4], L[5], L[6]) * -1.f : f3(L[1], L[2], L[3]) - pos;
    const float dd = dot3(tl, tl);
    dist = is_dir ? 1e4f : fmaxf(sqrtf(dd), 1e-4f);
    wi = tl * __builtin_amdgcn_rsqf(dd);
}

// Geometric shading frame of a hit (js/gi_probes.js:1993-2054).
struct HitFrame {
    float3 hit, ng, kd, emis;
    float bias;
};

__device__ __forceinline__ HitFrame hit_frame(const DevScene &s, const DevCascade &c, const sbgi_tick_params &p,
                                              float3 ro, float3 rd, float t, uint32_t tri, uint32_t inst) {
    HitFrame h;
    const float4 *tp = s.tri + (size_t)tri * 3;
    const float4 B = tp[1], C = tp[2];
    const uint32_t m = __float_as_uint(B.w);
    const float4 *ip = s.inst + (size_t)inst * 4;
    const float4 r0 = ip[0], r1 = ip[1], r2 = ip[2];
    float3 nl = normalize3(cross3(xyz(B), xyz(C)));
    float3 ng = normalize3(xyz(r0) * nl.x + xyz(r1) * nl.y + xyz(r2) * nl.z);
    h.ng = ng * (dot3(ng, rd) < 0.f ? 1.f : -1.f);
    h.bias = hit_bias(c.cell_min, p.dbg_trace_bias, p.dbg_ray_eps);
    h.hit = (ro + rd * t) + h.ng * h.bias;
    const float *mat = s.mats + (size_t)m * SBGI_MAT_STRIDE;
    const float metal = mat[4], transm = mat[5];
    h.emis = f3(mat[7], mat[8], mat[9]) * p.dbg_emissive;
    float3 base = f3(mat[0], mat[1], mat[2]);
    if (s.albedo) {
        const float *a = s.albedo + (size_t)albedo_index(s, inst, tri) * 3;
        base = f3(a[0], a[1], a[2]);
    }
    h.kd.x = fminf(fmaxf(base.x * p.dbg_albedo * (1.f - metal) * (1.f - transm), 0.f), 0.95f);
    h.kd.y = fminf(fmaxf(base.y * p.dbg_albedo * (1.f - metal) * (1.f - transm), 0.f), 0.95f);
    h.kd.z = fminf(fmaxf(base.z * p.dbg_albedo * (1.f - metal) * (1.f - transm), 0.f), 0.95f);
    return h;
}

// One light of the NEE loop (js/gi_probes.js:2129-2176): the gate before the shadow ray,
// and the contribution when it is not blocked.
struct LightSample {
    float3 wi, lcol;
    float dist, ndl;
    bool pass;
};

__device__ __forceinline__ LightSample light_gate(const float *__restrict__ L, const HitFrame &h,
                                                  const sbgi_tick_params &p, float scale = 1.f) {
    LightSample ls;
    const float ltype = L[0];
    ls.pass = false;
    if (!(ltype < 2.5f)) return ls;
    const bool is_ir = fabsf(L[16] - 4.0f) < 0.25f;
    const float gate = (is_ir ? p.nir_gate * p.nir_gain : 1.f) * scale;
    ls.lcol = f3(L[7], L[8], L[9]) * gate;
    const float lrange = L[10];
    const bool is_dir = ltype < 0.5f;
    light_vector(L, h.hit, ls.wi, ls.dist);
    ls.ndl = fmaxf(dot3(h.ng, ls.wi), 0.f);
    const float lpeak = fmaxf(fmaxf(ls.lcol.x, ls.lcol.y), ls.lcol.z);
    ls.pass = ls.ndl * lpeak > 0.f && (is_dir || lrange <= 0.f || ls.dist < lrange);
    return ls;
}

__device__ __forceinline__ float3 light_contrib(const float *__restrict__ L, const HitFrame &h, const LightSample &ls,
                                                const sbgi_tick_params &p) {
    const float ltype = L[0];
    const float lrange = L[10], ldecay = L[11], lca = L[12], lcp = L[13];
    const bool is_dir = ltype < 0.5f;
    const bool is_spot = fabsf(ltype - 2.0f) < 0.5f;
    const float3 ldir = f3(L[4], L[5], L[6]);
    // pow(d, 2) is d * d exactly (one rounding either way), so the reference's default
    // decay skips the general pow
    const float pw = ldecay == 2.f ? ls.dist * ls.dist : powf(ls.dist, ldecay);
    const float falloff = 1.0f / fmaxf(pw, 0.01f);
    const float rr = ls.dist / fmaxf(lrange, 1e-4f);
    const float rr2 = rr * rr;
    const float win = fminf(fmaxf(1.f - rr2 * rr2, 0.f), 1.f);
    const float ranged = falloff * (win * win);
    const float pos_att = lrange > 0.f ? ranged : falloff;
    const float dist_att = is_dir ? 1.f : pos_att;
    const float acos_ = dot3(ls.wi, ldir) * -1.f;
    const float spot = fminf(fmaxf((acos_ - lca) / fmaxf(lcp - lca, 1e-4f), 0.f), 1.f);
    const float att = dist_att * (is_spot ? spot : 1.f);
    const float3 diffuse = h.kd * (1.0f / SBGI_PI);
    return ((diffuse * ls.ndl) * ls.lcol) * (att * p.dbg_direct);
}

