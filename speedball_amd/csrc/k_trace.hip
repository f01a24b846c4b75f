// k_trace.hip — the probe trace: ray generation, closest hit, NEE shadow rays, bounce.
//
// Restates traceKernel (js/gi_probes.js:1977-2214) for the configuration the reference
// ships by default: non-clustered NEE over all lights, no texture maps, reflections off.
//
// MI355X layout of the work, four launches per solve instead of one monolithic kernel:
//   1. k_walk<CLOSEST>  persistent waves pull probe rays from a fetch counter (index
//                 k * updated + slot, slot fastest-varying, so neighbouring lanes start on
//                 one direction k from neighbouring probes — rayDir ignores the probe,
//                 js/gi_probes.js:1692-1704) and walk them with the wave-scheduled
//                 two-level walk (traverse2.h): finished lanes are refilled, the expensive
//                 triangle and instance phases run only with enough lanes waiting, and the
//                 TLAS and instance table sit in LDS. Result: hits[i] = {t, tri, inst}.
//   2. k_emit     per hit and per light that passes the reference's gate (n.l > 0, in
//                 range, not band-gated) a shadow ray is appended to a dense queue; appends
//                 are aggregated per wave with a ballot + one atomic, so each 64-ray chunk
//                 holds one light's rays from neighbouring probes.
//   3. k_walk<ANY>  the same persistent walk over the queue, any-hit; an unblocked ray
//                 sets bit `light` of its ray's visibility mask.
//   4. k_resolve  per ray: emissive + sum over lights, in light order, of the visible NEE
//                 terms + the bounce from the previous atlas — the reference's summation
//                 order, independent of the queue order, so the result is deterministic.
// Output is stored k-major (rays[k * updated + slot]): stores coalesce and the blend reads
// it as the K x N operand of its contraction.
#include <hip/hip_fp16.h>

// No FMA contraction anywhere in this translation unit (including the inlined helpers of
// the headers below): ray origins, instance transforms and the triangle test are
// evaluated in the operation order of the reference's graph, as the FP32 restatement
// does, so hits and t agree bit for bit with it.
#pragma clang fp contract(off)

#include "kernels.h"
#include "traverse2.h"

namespace {

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

// sampleAtlas (js/gi_probes.js:1941-1974)
__device__ float3 sample_atlas(const DevCascade &c, const ushort4 *__restrict__ atlas, float3 P, float3 N) {
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

// ray origin and direction of ray gid (probeTraceOrigin js/gi_probes.js:1714-1726, 1876-1885)
__device__ __forceinline__ void make_ray(const DevCascade &c, float classify_strength, const float4 *__restrict__ dirs,
                                         const float *__restrict__ meta, uint32_t gid, float3 &ro, float3 &rd) {
    const uint32_t k = gid / c.updated;
    const uint32_t slot = gid - k * c.updated;
    const uint32_t probe = (c.offset + slot) % c.total;
    const uint32_t ix = probe % c.res[0], iy = (probe / c.res[0]) % c.res[1], iz = probe / (c.res[0] * c.res[1]);
    ro = f3(c.gmin[0] + ((float)ix / fmaxf(1.f, (float)c.res[0] - 1.f)) * c.gsize[0],
            c.gmin[1] + ((float)iy / fmaxf(1.f, (float)c.res[1] - 1.f)) * c.gsize[1],
            c.gmin[2] + ((float)iz / fmaxf(1.f, (float)c.res[2] - 1.f)) * c.gsize[2]);
    if (classify_strength != 0.f) {
        const float *m = meta + (size_t)probe * 4;
        ro = ro + f3(m[1], m[2], m[3]) * classify_strength;
    }
    rd = xyz(dirs[k]);
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
    h.bias = fmaxf(c.cell_min * 0.005f * p.dbg_trace_bias, SBGI_RAY_EPS * p.dbg_ray_eps);
    h.hit = (ro + rd * t) + h.ng * h.bias;
    const float *mat = s.mats + (size_t)m * SBGI_MAT_STRIDE;
    const float metal = mat[4], transm = mat[5];
    h.emis = f3(mat[7], mat[8], mat[9]) * p.dbg_emissive;
    h.kd.x = fminf(fmaxf(mat[0] * p.dbg_albedo * (1.f - metal) * (1.f - transm), 0.f), 0.95f);
    h.kd.y = fminf(fmaxf(mat[1] * p.dbg_albedo * (1.f - metal) * (1.f - transm), 0.f), 0.95f);
    h.kd.z = fminf(fmaxf(mat[2] * p.dbg_albedo * (1.f - metal) * (1.f - transm), 0.f), 0.95f);
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
                                                  const sbgi_tick_params &p) {
    LightSample ls;
    const float ltype = L[0];
    ls.pass = false;
    if (!(ltype < 2.5f)) return ls;
    const bool is_ir = fabsf(L[16] - 4.0f) < 0.25f;
    const float gate = is_ir ? p.nir_gate * p.nir_gain : 1.f;
    ls.lcol = f3(L[7], L[8], L[9]) * gate;
    const float lrange = L[10];
    const bool is_dir = ltype < 0.5f;
    const float3 ldir = f3(L[4], L[5], L[6]);
    const float3 tl = is_dir ? ldir * -1.f : f3(L[1], L[2], L[3]) - h.hit;
    ls.dist = is_dir ? 1e4f : fmaxf(sqrtf(dot3(tl, tl)), 1e-4f);
    ls.wi = normalize3(tl);
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
    const float falloff = 1.0f / fmaxf(powf(ls.dist, ldecay), 0.01f);
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

constexpr uint32_t kBlock = 256;
// Walk blocks of 8 waves share one LDS copy of the TLAS, so the copy does not limit how
// many waves a CU holds.
constexpr uint32_t kWalkBlock = 512;

// Ray sources of the walk kernels (see walk_wave).

// Probe rays of this solve, index = k * updated + slot; result hits[i] = {t, tri, inst, 0}.
struct PrimarySrc {
    DevCascade c;
    float classify;
    const float4 *dirs;
    const float *meta;
    float4 *hits;
    __device__ __forceinline__ void load(const DevScene &s, uint32_t i, LaneRay &r) const {
        float3 ro, rd;
        make_ray(c, classify, dirs, meta, i, ro, rd);
        lane_begin(s, r, ro, rd, SBGI_T_MAX);
    }
    __device__ __forceinline__ void store(uint32_t i, const LaneRay &r) const {
        hits[i] = make_float4(r.tmax, __int_as_float(r.hit_tri), __int_as_float(r.hit_inst), 0.f);
    }
};

// Shadow rays of the dense queue; an unblocked ray sets bit `light` of its ray's mask.
struct ShadowSrc {
    const float4 *rec;
    unsigned long long *vis;
    __device__ __forceinline__ void load(const DevScene &s, uint32_t i, LaneRay &r) const {
        const float4 a = rec[(size_t)i * 2], b = rec[(size_t)i * 2 + 1];
        lane_begin(s, r, xyz(a), xyz(b), a.w);
        r.tag = __float_as_uint(b.w);
    }
    __device__ __forceinline__ void store(uint32_t, const LaneRay &r) const {
        if (!r.blocked) atomicOr(&vis[r.tag >> 6], 1ull << (r.tag & 63u));
    }
};

// Ray queries: in[i] = {ro.xyz, tmax}, {rd.xyz, mode}; mode 0 closest hit, 1 any hit
// within tmax, 2 closest hit reported as blocked when before tmax.
// out[i] = {t, tri, inst, blocked}.
struct QuerySrc {
    const float4 *in;
    float4 *out;
    __device__ __forceinline__ void load(const DevScene &s, uint32_t i, LaneRay &r) const {
        const float4 a = in[(size_t)i * 2], b = in[(size_t)i * 2 + 1];
        const bool any = b.w >= 0.5f && b.w < 1.5f;
        lane_begin(s, r, xyz(a), xyz(b), any ? a.w : SBGI_T_MAX);
        r.any = any;
        r.tag = b.w >= 1.5f ? __float_as_uint(a.w) : 0x7FC00000u;  // quiet NaN: never "before"
    }
    __device__ __forceinline__ void store(uint32_t i, const LaneRay &r) const {
        if (r.any) {
            out[i] = make_float4(0.f, 0.f, 0.f, r.blocked ? 1.f : 0.f);
        } else {
            const float blocked = (r.hit_tri >= 0 && r.tmax < __uint_as_float(r.tag)) ? 1.f : 0.f;
            out[i] = make_float4(r.tmax, __int_as_float(r.hit_tri), __int_as_float(r.hit_inst), blocked);
        }
    }
};

// Persistent walk: every wave pulls rays from the source until it is drained. The number
// of rays is n_host, or min(*n_dev, n_host) when the count is only known on the device.
template <int MODE, bool COUNT, bool LDS, class Src>
__global__ void __launch_bounds__(kWalkBlock) k_walk(DevScene s, Src src, const uint32_t *__restrict__ n_dev,
                                                      uint32_t n_host, uint32_t *heads,
                                                      unsigned long long *__restrict__ counters) {
    if (LDS) stage_tlas_lds(s);
    const uint32_t n = n_dev ? min(*n_dev, n_host) : n_host;
    TravCounters<COUNT> cnt;
    walk_wave<MODE, COUNT, LDS>(s, src, n, heads, blockIdx.x * (kWalkBlock / 64u) + (threadIdx.x >> 6), cnt);
    if (COUNT) {
        atomicAdd(&counters[0], (unsigned long long)cnt.blas);
        atomicAdd(&counters[1], (unsigned long long)cnt.tlas);
        atomicAdd(&counters[2], (unsigned long long)cnt.tris);
        atomicAdd(&counters[3], (unsigned long long)cnt.insts);
    }
}

// Shadow-ray emission: for every primary hit and every light that passes the reference's
// gate (n.l > 0, in range, not band-gated), one record is appended to the queue. A block
// counts its records (per wave, per light, by ballot), reserves them with ONE atomic, and
// writes them wave by wave and light by light, so each run of the queue holds one light's
// rays from neighbouring probes and the queue costs one atomic per kEmitBlock rays.
constexpr uint32_t kEmitBlock = 512;
__global__ void __launch_bounds__(kEmitBlock) k_emit(DevScene s, DevCascade c, sbgi_tick_params p,
                                                      const float4 *__restrict__ dirs, const float *__restrict__ meta,
                                                      const float4 *__restrict__ hits, ShadowQueue q,
                                                      Counters *__restrict__ counters) {
    __shared__ uint32_t wave_base[kEmitBlock / 64];
    const uint32_t n = c.updated * c.rays;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint64_t lt_mask = lane ? ((~0ull) >> (64 - lane)) : 0ull;
    const uint32_t gid = blockIdx.x * kEmitBlock + threadIdx.x;
    const bool active = gid < n;
    const float4 hr = active ? hits[gid] : make_float4(0.f, __int_as_float(-1), __int_as_float(-1), 0.f);
    const int tri = __float_as_int(hr.y), inst = __float_as_int(hr.z);
    const bool hit = active && tri >= 0;
    const bool wave_hit = __ballot(hit) != 0ull;
    HitFrame h;
    if (hit) {
        float3 ro, rd;
        make_ray(c, p.classify_strength, dirs, meta, gid, ro, rd);
        h = hit_frame(s, c, p, ro, rd, hr.x, (uint32_t)tri, (uint32_t)inst);
    }
    // pass 1: which lights pass, and how many records this wave appends
    uint64_t pass_bits = 0;
    uint32_t count = 0;
    if (wave_hit) {
        for (uint32_t li = 0; li < s.light_count; li++) {
            bool pass = false;
            if (hit) pass = light_gate(s.lights + li * SBGI_LIGHT_STRIDE, h, p).pass;
            if (pass) pass_bits |= 1ull << li;
            count += (uint32_t)__popcll(__ballot(pass));
        }
    }
    if (lane == 0) wave_base[wave] = count;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t total = 0;
        for (uint32_t w = 0; w < kEmitBlock / 64; w++) {
            const uint32_t c_w = wave_base[w];
            wave_base[w] = total;
            total += c_w;
        }
        const uint32_t base = total ? atomicAdd(q.count, total) : 0u;
        for (uint32_t w = 0; w < kEmitBlock / 64; w++) wave_base[w] += base;
        if (counters && total) atomicAdd(&counters->shadow_rays, (unsigned long long)total);
    }
    __syncthreads();
    if (!wave_hit) return;
    // pass 2: write the records
    uint32_t pos = wave_base[wave];
    for (uint32_t li = 0; li < s.light_count; li++) {
        const bool pass = (pass_bits >> li) & 1ull;
        const uint64_t m = __ballot(pass);
        if (m == 0ull) continue;
        if (pass) {
            const LightSample ls = light_gate(s.lights + li * SBGI_LIGHT_STRIDE, h, p);
            const uint32_t idx = pos + (uint32_t)__popcll(m & lt_mask);
            if (idx < q.capacity) {
                q.rec[(size_t)idx * 2] = make_float4(h.hit.x, h.hit.y, h.hit.z, ls.dist - h.bias);
                q.rec[(size_t)idx * 2 + 1] = make_float4(ls.wi.x, ls.wi.y, ls.wi.z, __uint_as_float((gid << 6) | li));
            }
        }
        pos += (uint32_t)__popcll(m);
    }
}

// Stage 3: shading in the reference's summation order.
__global__ void __launch_bounds__(kBlock) k_resolve(DevScene s, DevCascade c, sbgi_tick_params p,
                                                     const float4 *__restrict__ dirs, const float *__restrict__ meta,
                                                     const ushort4 *__restrict__ prev_atlas,
                                                     const float4 *__restrict__ hits,
                                                     const unsigned long long *__restrict__ vis,
                                                     float4 *__restrict__ rays, int2 *__restrict__ hit_ids) {
    const uint32_t gid = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t n = c.updated * c.rays;
    if (gid >= n) return;
    float3 ro, rd;
    make_ray(c, p.classify_strength, dirs, meta, gid, ro, rd);
    const float4 hr = hits[gid];
    const int tri = __float_as_int(hr.y), inst = __float_as_int(hr.z);
    float3 out;
    float hit_t = -1.f;
    if (tri >= 0) {
        hit_t = hr.x;
        const HitFrame h = hit_frame(s, c, p, ro, rd, hr.x, (uint32_t)tri, (uint32_t)inst);
        float3 rad = h.emis;
        const unsigned long long vm = vis[gid];
        for (uint32_t li = 0; li < s.light_count; li++) {
            if (!((vm >> li) & 1ull)) continue;
            const float *L = s.lights + li * SBGI_LIGHT_STRIDE;
            const LightSample ls = light_gate(L, h, p);
            if (ls.pass) rad = rad + light_contrib(L, h, ls, p);
        }
        // multibounce from the previous filtered atlas (js/gi_probes.js:2181-2188)
        const float3 bounce = sample_atlas(c, prev_atlas, h.hit, h.ng) * h.kd * (1.0f / SBGI_PI);
        const float bl = bounce.x * 0.2126f + bounce.y * 0.7152f + bounce.z * 0.0722f;
        const float roll = p.radiance_clamp / fmaxf(fmaxf(p.radiance_clamp, bl), 1e-6f);
        out = rad + bounce * (roll * p.dbg_bounce);
    } else {
        // miss -> SH-9 sky (js/gi_probes.js:2189-2207)
        const float x = rd.x, y = rd.y, z = rd.z;
        const float *sh = p.sky_sh;
        float o[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            float v = sh[0 + a] * 0.282095f + sh[3 + a] * (y * 0.488603f) + sh[6 + a] * (z * 0.488603f) +
                      sh[9 + a] * (x * 0.488603f) + sh[12 + a] * (x * y * 1.092548f) +
                      sh[15 + a] * (y * z * 1.092548f) + sh[18 + a] * ((z * z * 3.f - 1.f) * 0.315392f) +
                      sh[21 + a] * (x * z * 1.092548f) + sh[24 + a] * ((x * x - y * y) * 0.546274f);
            o[a] = fmaxf(v, 0.f) * p.sky_intensity;
        }
        out = f3(o[0], o[1], o[2]);
    }
    rays[gid] = make_float4(out.x, out.y, out.z, hit_t);
    if (hit_ids) hit_ids[gid] = make_int2(tri, inst);
}

// Clears the visibility masks, the queue length and the fetch counters of the primary and
// shadow walks (the solve counters of kernels.h).
__global__ void k_clear_u64(unsigned long long *__restrict__ p, uint32_t n, uint32_t *__restrict__ counters) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
    if (i < 1u + 2u * kFetchHeads) counters[(size_t)i * kHeadStride] = 0u;
}

// Direction table: normalize(rayDir(k, jitter)) (js/gi_probes.js:1692-1704, 1984).
__global__ void k_raydirs(float4 *__restrict__ dirs, uint32_t rays, float jitter) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= rays) return;
    float sk = ((float)k + 0.5f) / (float)rays;
    float u = sk + jitter;
    float uw = u - floorf(u);
    float z = 1.f - uw * 2.f;
    float r = sqrtf(fmaxf(0.f, 1.f - z * z));
    float phi = (float)k * SBGI_GOLDEN_ANGLE + jitter * SBGI_TWO_PI;
    // cos/sin of the f32 phi, rounded once from double: the correctly rounded f32 values
    // (R evaluations per solve, so the double path costs nothing)
    const float cphi = (float)cos((double)phi), sphi = (float)sin((double)phi);
    float3 d = normalize3(f3(r * cphi, r * sphi, z));
    dirs[k] = make_float4(d.x, d.y, d.z, 0.f);
}

int device_cus() {
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
    }
    return cus;
}

// Persistent grid: as many blocks as can be resident at once (occupancy API), capped by
// the work. Every block then strides over the work, so no block waits for a slot.
template <typename K>
uint32_t persistent_blocks(K kernel, uint32_t block, size_t lds, uint32_t work_items) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, (int)block, lds) != hipSuccess || per_cu <= 0)
        per_cu = 1;
    const uint32_t cap = (uint32_t)device_cus() * (uint32_t)per_cu;
    const uint32_t want = (work_items + block - 1) / block;
    return want < cap ? (want ? want : 1u) : cap;
}

}  // namespace

size_t trace_lds_bytes(uint32_t tlas_count, uint32_t inst_count) {
    return ((size_t)tlas_count * 2 + (size_t)inst_count * 4) * sizeof(uint4);
}

hipError_t launch_query(const DevScene &s, const float4 *in, float4 *out, uint32_t n, uint32_t *head, hipStream_t st) {
    if (!n) return hipSuccess;
    hipError_t e = hipMemsetAsync(head, 0, (size_t)kFetchHeads * kHeadStride * sizeof(uint32_t), st);
    if (e != hipSuccess) return e;
    const QuerySrc src{in, out};
    auto kern = k_walk<WALK_MIXED, false, false, QuerySrc>;
    hipLaunchKernelGGL(kern, dim3(persistent_blocks(kern, kWalkBlock, 0, n)), dim3(kWalkBlock), 0, st, s, src, nullptr,
                       n, head, nullptr);
    return hipGetLastError();
}

hipError_t launch_raydirs(float4 *dirs, uint32_t rays, float jitter, hipStream_t st) {
    hipLaunchKernelGGL(k_raydirs, dim3((rays + 63) / 64), dim3(64), 0, st, dirs, rays, jitter);
    return hipGetLastError();
}

// One persistent walk launch: counting and LDS variants picked at run time.
template <int MODE, class Src>
void launch_walk(const DevScene &s, const Src &src, const uint32_t *n_dev, uint32_t n_host, uint32_t *heads,
                 unsigned long long *cnt, size_t lds, hipStream_t st) {
    const bool use_lds = lds <= kTraceLdsBudget;
    auto go = [&](auto kern, size_t l) {
        hipLaunchKernelGGL(kern, dim3(persistent_blocks(kern, kWalkBlock, l, n_host)), dim3(kWalkBlock), l, st, s, src,
                           n_dev, n_host, heads, cnt);
    };
    if (cnt) {
        if (use_lds) go(k_walk<MODE, true, true, Src>, lds);
        else go(k_walk<MODE, true, false, Src>, 0);
    } else {
        if (use_lds) go(k_walk<MODE, false, true, Src>, lds);
        else go(k_walk<MODE, false, false, Src>, 0);
    }
}

hipError_t launch_trace(const DevScene &s, const DevCascade &c, const sbgi_tick_params &p, const float4 *dirs,
                        const float *meta, const ushort4 *prev_atlas, const TraceScratch &w, float4 *rays,
                        int2 *hit_ids, Counters *counters, const hipEvent_t *ev, hipStream_t st) {
    const uint32_t n = c.updated * c.rays;
    if (n == 0) return hipSuccess;
    auto mark = [&](int i) { return ev ? hipEventRecord(ev[i], st) : hipSuccess; };
    uint32_t *ctr = w.queue.count;
    const uint32_t nclear = std::max(n, 1u + 2u * kFetchHeads);
    hipLaunchKernelGGL(k_clear_u64, dim3((nclear + 255) / 256), dim3(256), 0, st, w.vis, n, ctr);
    const size_t lds = trace_lds_bytes(s.tlas_count, s.inst_count);
    hipError_t e = mark(0);
    if (e != hipSuccess) return e;
    launch_walk<WALK_CLOSEST>(s, PrimarySrc{c, p.classify_strength, dirs, meta, w.hits}, nullptr, n,
                              ctr + trace_heads(0), counters ? counters->walk : nullptr, lds, st);
    if ((e = mark(1)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_emit, dim3((n + kEmitBlock - 1) / kEmitBlock), dim3(kEmitBlock), 0, st, s, c, p, dirs, meta,
                       w.hits, w.queue, counters);
    if ((e = mark(2)) != hipSuccess) return e;
    // the queue length is only known on the device: the kernel reads it, bounded by the capacity
    launch_walk<WALK_ANY>(s, ShadowSrc{w.queue.rec, w.vis}, w.queue.count, w.queue.capacity, ctr + trace_heads(1),
                          counters ? counters->shadow_walk : nullptr, lds, st);
    if ((e = mark(3)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_resolve, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, s, c, p, dirs, meta,
                       prev_atlas, w.hits, w.vis, rays, hit_ids);
    return hipGetLastError();
}
