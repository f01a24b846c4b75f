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
#include "shade.h"
#include "traverse2.h"

namespace {

constexpr uint32_t kBlock = 256;
// Walk blocks of 8 waves share one LDS copy of the TLAS, so the copy does not limit how
// many waves a CU holds.
constexpr uint32_t kWalkBlock = 512;
constexpr uint32_t kWalkBlockSmall = 256;

// Ray sources of the walk kernels (see walk_wave).

// Probe rays of this solve, index = k * updated + slot; result hits[i] = {t, tri, inst, 0}.
struct PrimarySrc {
    static constexpr bool kList = false;
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

// Shadow rays of the dense queue, 16 bytes each: {origin.xyz, (ray << 6) | light}. The
// direction and extent are re-derived from the light record exactly as the emission did
// (light_vector, and the cascade's hit bias). An unblocked ray sets bit `light` of its
// ray's mask.
struct ShadowSrc {
    static constexpr bool kList = false;
    const float4 *rec;
    unsigned long long *vis;
    float bias;
    __device__ __forceinline__ void load(const DevScene &s, uint32_t i, LaneRay &r) const {
        const float4 a = rec[i];
        const uint32_t tag = __float_as_uint(a.w);
        float3 wi;
        float dist;
        light_vector(s.lights + (size_t)(tag & 63u) * SBGI_LIGHT_STRIDE, xyz(a), wi, dist);
        lane_begin<true>(s, r, xyz(a), wi, dist - bias);
        r.tag = tag;
    }
    __device__ __forceinline__ void store(uint32_t, const LaneRay &r) const {
        if (!r.blocked) atomicOr(&vis[r.tag >> 6], 1ull << (r.tag & 63u));
    }
};

// The same shadow rays walked over their light's candidate list (k_shadow.hip) instead of
// the TLAS: the load finds the cell of the ray's origin for its light and walks that cell's
// list (never empty: a cell that no box reaches lists one box that no ray hits).
struct ShadowListSrc {
    static constexpr bool kList = true;
    const float4 *rec;
    unsigned long long *vis;
    float bias;
    __device__ __forceinline__ void load(const DevScene &s, uint32_t i, LaneRay &r) const {
        const float4 a = rec[i];
        const uint32_t tag = __float_as_uint(a.w), l = tag & 63u;
        float3 wi;
        float dist;
        light_vector(s.lights + (size_t)l * SBGI_LIGHT_STRIDE, xyz(a), wi, dist);
        lane_begin<true>(s, r, xyz(a), wi, dist - bias);
        r.tag = tag;
        const uint2 ce = s.sl_cells[(size_t)l * kShadowRecs + shadow_cell(s.sl_proj + (size_t)l * kShadowProj, xyz(a))];
        r.cur = ce.x;
        r.end = ce.y;
        r.lend = ce.y;
        r.lquad = reinterpret_cast<const uint2 *>(s.sl_list)[ce.x >> 2];
    }
    __device__ __forceinline__ void store(uint32_t, const LaneRay &r) const {
        if (!r.blocked) atomicOr(&vis[r.tag >> 6], 1ull << (r.tag & 63u));
    }
};

// Emitter visibility (emitterVisKernel, js/gi_probes.js:2216-2260): ray i = slot * E + e
// tests the probe's view of a hashed point on emitter e; the store blends the 0/1 sample
// into the retained visibility, vis = mix(sample, vis, 0.8). The rays come from a table that
// k_emitter_rays writes just before the walk, {origin, extent}, {direction, tag} per ray:
// the hash's double-precision sines and the sample frame cost the walk's refill ~300
// instructions and 40 VGPRs per ray if evaluated there. A probe inside the emitter's
// bounding sphere walks nothing (its extent is NaN, so no box test passes) and samples 0.
struct EmitterSrc {
    static constexpr bool kList = false;
    const float4 *rays;  // 2 per ray
    float *vis;          // [probe * SBGI_MAX_EMITTERS + e]
    __device__ __forceinline__ void load(const DevScene &s, uint32_t i, LaneRay &r) const {
        const float4 a = rays[(size_t)i * 2], b = rays[(size_t)i * 2 + 1];
        lane_begin<true>(s, r, xyz(a), xyz(b), a.w);
        r.tag = __float_as_uint(b.w);
    }
    __device__ __forceinline__ void store(uint32_t, const LaneRay &r) const {
        const float sample = (!(r.tag >> 31) && !r.blocked) ? 1.f : 0.f;
        float *v = vis + (r.tag & 0x7FFFFFFFu);
        *v = sample * (1.f - SBGI_EMITTER_RETENTION) + *v * SBGI_EMITTER_RETENTION;
    }
};

// Classification rays (classifyKernel, js/gi_probes.js:2850-2899): ray i = probe * 32 + k of
// every probe of the cascade, from the probe's grid position (no relocation) along fixed
// direction k (classifyRayDir, :1707-1712: rayDir with 32 rays and no jitter); result
// hits[i] = {t, tri, inst, 0}.
struct ClassifySrc {
    static constexpr bool kList = false;
    DevCascade c;
    const float4 *dirs;
    float4 *hits;
    __device__ __forceinline__ void load(const DevScene &s, uint32_t i, LaneRay &r) const {
        const uint32_t probe = i / kClassifyRays, k = i - probe * kClassifyRays;
        lane_begin(s, r, probe_origin(c, 0.f, nullptr, probe), xyz(dirs[k]), SBGI_T_MAX);
    }
    __device__ __forceinline__ void store(uint32_t i, const LaneRay &r) const {
        hits[i] = make_float4(r.tmax, __int_as_float(r.hit_tri), __int_as_float(r.hit_inst), 0.f);
    }
};

// Ray queries: in[i] = {ro.xyz, tmax}, {rd.xyz, mode}; mode 0 closest hit, 1 any hit
// within tmax, 2 closest hit reported as blocked when before tmax.
// out[i] = {t, tri, inst, blocked}.
struct QuerySrc {
    static constexpr bool kList = false;
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
template <int MODE, bool COUNT, bool LDS, class Src, bool WT = false, int BLOCK = kWalkBlock>
__global__ void __launch_bounds__(BLOCK) k_walk(DevScene s, Src src, const uint32_t *__restrict__ n_dev,
                                                 uint32_t n_host, uint32_t *heads,
                                                 unsigned long long *__restrict__ counters) {
    if (LDS) {
        if constexpr (Src::kList) stage_list_lds(s);
        else stage_tlas_lds(s);
    }
    const uint32_t n = n_dev ? min(*n_dev, n_host) : n_host;
    TravCounters<COUNT> cnt;
    walk_wave<MODE, COUNT, LDS, WT>(s, src, n, heads, blockIdx.x * (BLOCK / 64u) + (threadIdx.x >> 6), cnt);
    if (COUNT) {
        atomicAdd(&counters[0], (unsigned long long)cnt.blas);
        atomicAdd(&counters[1], (unsigned long long)cnt.tlas);
        atomicAdd(&counters[2], (unsigned long long)cnt.tris);
        atomicAdd(&counters[3], (unsigned long long)cnt.insts);
        if ((threadIdx.x & 63u) == 0) {
            atomicAdd(&counters[4], (unsigned long long)cnt.node_w);
            atomicAdd(&counters[5], (unsigned long long)cnt.tri_w);
            atomicAdd(&counters[6], (unsigned long long)cnt.inst_w);
            atomicAdd(&counters[7], (unsigned long long)cnt.refill_w);
        }
    }
}

// Batched walk: one segment per (field, cascade) solve. A wave starts on the segment that
// holds its share of the frame's probe rays (prefix), walks it with walk_wave until the
// segment's rays are all taken, marks it done and moves on to the next unfinished segment,
// so waves that finish early take work from segments that are still running. The scene
// changes only between segments, when every lane of the wave is idle.
template <class Src>
struct WalkSeg {
    DevScene s;
    Src src;
    const uint32_t *n_dev;  // device ray count (shadow queue) or nullptr
    uint32_t n_host;
    uint32_t *heads;
};

// LDS: every segment walks the same scene (the cascades of one field), `shared`, whose TLAS
// and instance records each block stages in its LDS as the single-field walk does.
template <int MODE, int BLOCK, class Src, bool WT = false, bool LDS = false>
__global__ void __launch_bounds__(BLOCK) k_walk_batch(const WalkSeg<Src> *__restrict__ segs, uint32_t n_seg,
                                                      const uint32_t *__restrict__ prefix, uint32_t total,
                                                      uint32_t *__restrict__ done, DevScene shared) {
    if (LDS) stage_tlas_lds(shared);
    // wave-uniform by construction; readfirstlane tells the compiler, so the segment table
    // is read with scalar loads into SGPRs instead of per lane on every step
    const uint32_t waves = gridDim.x * (BLOCK / 64u);
    const uint32_t wave_id = __builtin_amdgcn_readfirstlane(blockIdx.x * (BLOCK / 64u) + (threadIdx.x >> 6));
    const uint32_t target = (uint32_t)(((uint64_t)total * wave_id) / waves);
    uint32_t lo = 0, hi = n_seg;  // last segment whose prefix <= target
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (prefix[mid] <= target) lo = mid;
        else hi = mid;
    }
    TravCounters<false> cnt;
    for (uint32_t k = 0; k < n_seg; k++) {
        const uint32_t j = lo + k < n_seg ? lo + k : lo + k - n_seg;
        if (__hip_atomic_load(&done[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) continue;
        const WalkSeg<Src> &g = segs[j];
        const uint32_t n = g.n_dev ? min(*g.n_dev, g.n_host) : g.n_host;
        // copies read once per segment: the walk's stores cannot then force a reload of the
        // scene and source fields on every step
        const DevScene sc = g.s;
        const Src src = g.src;
        walk_wave<MODE, false, LDS, WT>(sc, src, n, g.heads, wave_id, cnt);
        if ((threadIdx.x & 63u) == 0) __hip_atomic_store(&done[j], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
    uint64_t members = 0ull;
    if (hit) {
        float3 ro, rd;
        make_ray(c, p.classify_strength, dirs, meta, gid, ro, rd);
        h = hit_frame(s, c, p, ro, rd, hr.x, (uint32_t)tri, (uint32_t)inst);
        members = light_members(s, h.hit);
    }
    // pass 1: which lights pass, and how many records this wave appends. A light outside
    // every lane's cell list is skipped by the whole wave.
    uint64_t pass_bits = 0;
    uint32_t count = 0;
    if (wave_hit) {
        for (uint32_t li = 0; li < s.light_count; li++) {
            const bool member = (members >> li) & 1ull;
            if (!__ballot(member)) continue;
            bool pass = false;
            if (member) pass = light_gate(s.lights + li * SBGI_LIGHT_STRIDE, h, p, light_scale(s, li)).pass;
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
            const uint32_t idx = pos + (uint32_t)__popcll(m & lt_mask);
            if (idx < q.capacity) {
                q.rec[idx] = make_float4(h.hit.x, h.hit.y, h.hit.z, __uint_as_float((gid << 6) | li));
            }
        }
        pos += (uint32_t)__popcll(m);
    }
}

// Stage 3: shading in the reference's summation order. Asked for at least 5 waves per
// SIMD, the compiler fits it in 80 VGPRs (6 waves) without spilling; left alone it takes
// 101 (4 waves).
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(5))) k_resolve(DevScene s, DevCascade c, sbgi_tick_params p,
                                                     const float4 *__restrict__ dirs, const float *__restrict__ meta,
                                                     const ushort4 *__restrict__ prev_atlas,
                                                     const float4 *__restrict__ prev_atlas_f32,
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
            const LightSample ls = light_gate(L, h, p, light_scale(s, li));
            if (ls.pass) rad = rad + light_contrib(L, h, ls, p);
        }
        // with reflection lobes, a hit without Lambert or emissive radiance (pure metal or
        // glass) keeps occluding depth but is marked as carrying no reflection radiance
        // (js/gi_probes.js:2055-2062)
        if (c.rough) {
            const float4 *tp = s.tri + (size_t)tri * 3;
            const float *mat = s.mats + (size_t)__float_as_uint(tp[1].w) * SBGI_MAT_STRIDE;
            const float energy = (1.f - mat[4]) * (1.f - mat[5]);
            const float peak = fmaxf(fmaxf(mat[7], mat[8]), mat[9]);
            if (!(energy > 0.02f || peak > 1e-6f)) hit_t = -(hr.x + 2.f);
        }
        // multibounce from the previous filtered atlas (js/gi_probes.js:2181-2188)
        const float3 smp = prev_atlas_f32 ? sample_atlas(c, prev_atlas_f32, h.hit, h.ng)
                                          : sample_atlas(c, prev_atlas, h.hit, h.ng);
        const float3 bounce = smp * h.kd * (1.0f / SBGI_PI);
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

// Clears the visibility masks, the queue length and the fetch counters of every walk of
// the solve (the solve counters of kernels.h).
static_assert(2 * SBGI_MAX_EMITTERS <= SBGI_RAYS_MIN, "the emitter ray table must fit the hit records");
// The emitter-visibility rays of a solve (EmitterSrc): ray i = slot * E + e, from the
// probe's traced position toward a hashed point of emitter e (emitter_ray), as
// {origin, extent (NaN when the probe is inside the emitter)}, {direction, tag}.
__global__ void __launch_bounds__(kBlock) k_emitter_rays(DevScene s, DevCascade c, float classify, float seed,
                                                          float bias, const float *__restrict__ meta,
                                                          float4 *__restrict__ out) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= c.updated * s.emitter_count) return;
    const uint32_t slot = i / s.emitter_count, e = i - slot * s.emitter_count;
    const uint32_t probe = (c.offset + slot) % c.total;
    const float3 ro = probe_origin(c, classify, meta, probe);
    float3 rd = f3(0.f, 0.f, 1.f);
    float tmax = 0.f;
    const bool valid = emitter_ray(s.lights + (size_t)(s.emitter_base + e) * SBGI_LIGHT_STRIDE, ro, probe, e, seed,
                                   bias, rd, tmax);
    const uint32_t tag = (probe * SBGI_MAX_EMITTERS + e) | (valid ? 0u : 0x80000000u);
    out[(size_t)i * 2] = make_float4(ro.x, ro.y, ro.z, valid ? tmax : __builtin_nanf(""));
    out[(size_t)i * 2 + 1] = make_float4(rd.x, rd.y, rd.z, __uint_as_float(tag));
}

__global__ void k_clear_u64(unsigned long long *__restrict__ p, uint32_t n, uint32_t *__restrict__ counters) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
    if (i < 1u + 4u * kFetchHeads) counters[(size_t)i * kHeadStride] = 0u;
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

// classifyKernel's reduction per probe (js/gi_probes.js:2866-2898): a hit whose world
// geometric normal faces along the ray (dot(rd, ng) > 0) sees a back face; above
// BACKFACE_FRACTION (0.25) of the hits the probe is buried (state 0, else 1). A probe whose
// nearest hit is a back face moves along that ray past the surface by the back distance plus
// half of min(front distance, cellMin), the move clamped to relocClamp. meta[p] = {state,
// offset.xyz}. The geometric normal is the one the shading uses (hit_frame), before any
// face-forwarding.
__global__ void k_classify(DevScene s, DevCascade c, const float4 *__restrict__ dirs, const float4 *__restrict__ hits,
                           float reloc_clamp, float *__restrict__ meta) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= c.total) return;
    float back = 0.f, nhit = 0.f, close_back = 1e30f, close_front = 1e30f;
    float3 back_dir = f3(0.f, 0.f, 0.f);
    for (uint32_t k = 0; k < kClassifyRays; k++) {
        const float4 h = hits[(size_t)p * kClassifyRays + k];
        const int tri = __float_as_int(h.y);
        if (tri < 0) continue;
        nhit += 1.f;
        const float3 rd = xyz(dirs[k]);
        const float4 *tp = s.tri + (size_t)tri * 3;
        const float4 *ip = s.inst + (size_t)__float_as_uint(h.z) * 4;
        const float3 nl = normalize3(cross3(xyz(tp[1]), xyz(tp[2])));
        const float3 ng = normalize3(xyz(ip[0]) * nl.x + xyz(ip[1]) * nl.y + xyz(ip[2]) * nl.z);
        if (dot3(rd, ng) > 0.f) {
            back += 1.f;
            if (h.x < close_back) {
                close_back = h.x;
                back_dir = rd;
            }
        } else if (h.x < close_front) {
            close_front = h.x;
        }
    }
    const float state = back / fmaxf(nhit, 1.f) > kBackfaceFraction ? 0.f : 1.f;
    float3 off = f3(0.f, 0.f, 0.f);
    if (back > 0.5f && close_back < close_front) {
        const float step = close_back + 0.5f * fminf(close_front, c.cell_min);
        const float3 raw = back_dir * step;
        const float len = sqrtf(dot3(raw, raw));
        off = raw * (fminf(len, reloc_clamp) / fmaxf(len, 1e-6f));
    }
    float *m = meta + (size_t)p * 4;
    m[0] = state;
    m[1] = off.x;
    m[2] = off.y;
    m[3] = off.z;
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
// waves_cap > 0 limits the resident waves per CU (at least one block).
template <typename K>
uint32_t persistent_blocks(K kernel, uint32_t block, size_t lds, uint32_t work_items, uint32_t waves_cap = 0) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, (int)block, lds) != hipSuccess || per_cu <= 0)
        per_cu = 1;
    if (waves_cap) per_cu = min(per_cu, max(1, (int)(waves_cap / (block / 64u))));
    const uint32_t cap = (uint32_t)device_cus() * (uint32_t)per_cu;
    const uint32_t want = (uint32_t)(((uint64_t)work_items + block - 1) / block);
    return want < cap ? (want ? want : 1u) : cap;
}

}  // namespace

size_t walk_tables_bytes(uint32_t n_seg) {
    return (size_t)n_seg * (sizeof(WalkSeg<PrimarySrc>) + sizeof(WalkSeg<ShadowSrc>));
}

void walk_tables(const TraceSeg *segs, uint32_t n_seg, std::vector<unsigned char> &blob) {
    blob.assign(walk_tables_bytes(n_seg), 0);
    auto *pr = reinterpret_cast<WalkSeg<PrimarySrc> *>(blob.data());
    auto *sh = reinterpret_cast<WalkSeg<ShadowSrc> *>(blob.data() + (size_t)n_seg * sizeof(WalkSeg<PrimarySrc>));
    for (uint32_t j = 0; j < n_seg; j++) {
        const TraceSeg &g = segs[j];
        pr[j] = WalkSeg<PrimarySrc>{g.s, PrimarySrc{g.c, g.classify, g.dirs, g.meta, g.w.hits}, nullptr,
                                    g.c.updated * g.c.rays, g.heads};
        DevScene gs = g.s;
        gs.blas_octants = gs.shadow_octants;  // the shadow walk's BLAS copies
        sh[j] = WalkSeg<ShadowSrc>{gs, ShadowSrc{g.w.queue.rec, g.w.vis, g.bias}, g.w.queue.count, g.w.queue.capacity,
                                   g.heads + (size_t)kFetchHeads * kHeadStride};
    }
}

hipError_t launch_walk_batch(int mode, const void *tables, uint32_t n_seg, const uint32_t *prefix, uint32_t total,
                             uint32_t *done, uint32_t block, uint32_t waves, bool wt, const DevScene *shared,
                             hipStream_t st) {
    if (!n_seg || !total) return hipSuccess;
    const auto *pr = static_cast<const WalkSeg<PrimarySrc> *>(tables);
    const auto *sh = reinterpret_cast<const WalkSeg<ShadowSrc> *>(static_cast<const unsigned char *>(tables) +
                                                                 (size_t)n_seg * sizeof(WalkSeg<PrimarySrc>));
    DevScene one{};
    const size_t lds = shared ? trace_lds_bytes(shared->tlas_count, shared->inst_count) : 0;
    const bool use_lds = shared && !wt && lds <= kTraceLdsBudget;
    if (use_lds) one = *shared;
    auto go = [&](auto kern, uint32_t bs, auto table) {
        const size_t l = use_lds ? lds : 0;
        hipLaunchKernelGGL(kern, dim3(persistent_blocks(kern, bs, l, total, waves)), dim3(bs), l, st, table, n_seg,
                           prefix, total, done, one);
    };
    if (use_lds) {
        if (mode == 0) {
            if (block == 256) go(k_walk_batch<WALK_CLOSEST, 256, PrimarySrc, false, true>, 256u, pr);
            else go(k_walk_batch<WALK_CLOSEST, 512, PrimarySrc, false, true>, 512u, pr);
        } else {
            if (block == 256) go(k_walk_batch<WALK_ANY, 256, ShadowSrc, false, true>, 256u, sh);
            else go(k_walk_batch<WALK_ANY, 512, ShadowSrc, false, true>, 512u, sh);
        }
        return hipGetLastError();
    }
    // the watertight variants only when a segment's scene asks for it: the watertight test
    // costs registers that the Moller-Trumbore walks do not pay for
    if (mode == 0 && !wt) {
        if (block == 256) go(k_walk_batch<WALK_CLOSEST, 256, PrimarySrc>, 256u, pr);
        else go(k_walk_batch<WALK_CLOSEST, 512, PrimarySrc>, 512u, pr);
    } else if (mode == 0) {
        if (block == 256) go(k_walk_batch<WALK_CLOSEST, 256, PrimarySrc, true>, 256u, pr);
        else go(k_walk_batch<WALK_CLOSEST, 512, PrimarySrc, true>, 512u, pr);
    } else if (!wt) {
        if (block == 256) go(k_walk_batch<WALK_ANY, 256, ShadowSrc>, 256u, sh);
        else go(k_walk_batch<WALK_ANY, 512, ShadowSrc>, 512u, sh);
    } else {
        if (block == 256) go(k_walk_batch<WALK_ANY, 256, ShadowSrc, true>, 256u, sh);
        else go(k_walk_batch<WALK_ANY, 512, ShadowSrc, true>, 512u, sh);
    }
    return hipGetLastError();
}

hipError_t launch_emit(const DevScene &s, const DevCascade &c, const sbgi_tick_params &p, const float4 *dirs,
                       const float *meta, const float4 *hits, const ShadowQueue &q, hipStream_t st) {
    const uint32_t n = c.updated * c.rays;
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_emit, dim3((n + kEmitBlock - 1) / kEmitBlock), dim3(kEmitBlock), 0, st, s, c, p, dirs, meta,
                       hits, q, nullptr);
    return hipGetLastError();
}

hipError_t launch_resolve(const DevScene &s, const DevCascade &c, const sbgi_tick_params &p, const float4 *dirs,
                          const float *meta, const ushort4 *prev_atlas, const float4 *hits,
                          const unsigned long long *vis, float4 *rays, hipStream_t st) {
    const uint32_t n = c.updated * c.rays;
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_resolve, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, s, c, p, dirs, meta,
                       prev_atlas, nullptr, hits, vis, rays, nullptr);
    return hipGetLastError();
}

size_t trace_lds_bytes(uint32_t tlas_count, uint32_t inst_count) {
    return ((size_t)tlas_count * 2 + (size_t)inst_count * 4) * sizeof(uint4);
}

hipError_t launch_query(const DevScene &s, const float4 *in, float4 *out, uint32_t n, uint32_t *head, hipStream_t st) {
    if (!n) return hipSuccess;
    hipError_t e = hipMemsetAsync(head, 0, (size_t)kFetchHeads * kHeadStride * sizeof(uint32_t), st);
    if (e != hipSuccess) return e;
    const QuerySrc src{in, out};
    DevScene sq = s;
    sq.tlas_oct = nullptr;
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(persistent_blocks(kern, kWalkBlock, 0, n)), dim3(kWalkBlock), 0, st, sq, src,
                           nullptr, n, head, nullptr);
    };
    if (s.watertight) go(k_walk<WALK_MIXED, false, false, QuerySrc, true>);
    else go(k_walk<WALK_MIXED, false, false, QuerySrc>);
    return hipGetLastError();
}

hipError_t launch_raydirs(float4 *dirs, uint32_t rays, float jitter, hipStream_t st) {
    hipLaunchKernelGGL(k_raydirs, dim3((rays + 63) / 64), dim3(64), 0, st, dirs, rays, jitter);
    return hipGetLastError();
}

// One persistent walk launch: counting and LDS variants picked at run time.
template <int MODE, class Src>
void launch_walk(const DevScene &scene, const Src &src, const uint32_t *n_dev, uint32_t n_host, uint32_t *heads,
                 unsigned long long *cnt, size_t lds, hipStream_t st) {
    // only the closest-hit walk takes the near-first TLAS copies; it reads them from global
    // memory and stages only the instance records, in blocks of 4 waves (its VGPRs allow 7
    // waves per SIMD, which blocks of 8 waves would round down to 6)
    DevScene s = scene;
    if (MODE != WALK_CLOSEST) s.tlas_oct = nullptr;
    if constexpr (Src::kList) {  // LDS: instance records, then the boxes the list steps test
        s.lds_inst = 0;
        s.lds_box = s.inst_count * 4;
        lds = ((size_t)s.inst_count * 6 + 2) * sizeof(uint4);
    }
    const bool use_lds = lds <= kTraceLdsBudget;
    if (MODE == WALK_CLOSEST && s.tlas_oct && use_lds && !s.watertight) {
        s.lds_inst = 0;
        const size_t l = (size_t)s.inst_count * 64;
        auto k = cnt ? k_walk<MODE, true, true, Src, false, kWalkBlockSmall>
                     : k_walk<MODE, false, true, Src, false, kWalkBlockSmall>;
        hipLaunchKernelGGL(k, dim3(persistent_blocks(k, kWalkBlockSmall, l, n_host)), dim3(kWalkBlockSmall), l, st, s,
                           src, n_dev, n_host, heads, cnt);
        return;
    }
    auto go = [&](auto kern, size_t l) {
        hipLaunchKernelGGL(kern, dim3(persistent_blocks(kern, kWalkBlock, l, n_host)), dim3(kWalkBlock), l, st, s, src,
                           n_dev, n_host, heads, cnt);
    };
    if (s.watertight) {
        if (cnt) {
            if (use_lds) go(k_walk<MODE, true, true, Src, true>, lds);
            else go(k_walk<MODE, true, false, Src, true>, 0);
        } else {
            if (use_lds) go(k_walk<MODE, false, true, Src, true>, lds);
            else go(k_walk<MODE, false, false, Src, true>, 0);
        }
    } else if (cnt) {
        if (use_lds) go(k_walk<MODE, true, true, Src>, lds);
        else go(k_walk<MODE, true, false, Src>, 0);
    } else {
        if (use_lds) go(k_walk<MODE, false, true, Src>, lds);
        else go(k_walk<MODE, false, false, Src>, 0);
    }
}

hipError_t launch_trace(const DevScene &s, const DevCascade &c, const sbgi_tick_params &p, const float4 *dirs,
                        const float *meta, const ushort4 *prev_atlas, const float4 *prev_atlas_f32,
                        const TraceScratch &w, float4 *rays, int2 *hit_ids, Counters *counters, const hipEvent_t *ev,
                        hipStream_t st, const StageMarker &marker) {
    const uint32_t n = c.updated * c.rays;
    if (n == 0) return hipSuccess;
    auto mark = [&](int i) { return ev ? marker(ev[i], st) : hipSuccess; };
    uint32_t *ctr = w.queue.count;
    const uint32_t nclear = std::max(n, 1u + 4u * kFetchHeads);
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
    const float bias = hit_bias(c.cell_min, p.dbg_trace_bias, p.dbg_ray_eps);
    DevScene ss = s;
    ss.blas_octants = s.shadow_octants;  // the shadow walk's BLAS copies
    if (s.sl_cells)
        launch_walk<WALK_ANY>(ss, ShadowListSrc{w.queue.rec, w.vis, bias}, w.queue.count, w.queue.capacity,
                              ctr + trace_heads(1), counters ? counters->shadow_walk : nullptr, lds, st);
    else
        launch_walk<WALK_ANY>(ss, ShadowSrc{w.queue.rec, w.vis, bias}, w.queue.count, w.queue.capacity,
                              ctr + trace_heads(1), counters ? counters->shadow_walk : nullptr, lds, st);
    if ((e = mark(3)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_resolve, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, s, c, p, dirs, meta,
                       prev_atlas, prev_atlas_f32, w.hits, w.vis, rays, hit_ids);
    return hipGetLastError();
}

hipError_t launch_classify(const DevScene &s, const DevCascade &c, float reloc_clamp, float4 *dirs, float4 *hits,
                           float *meta, uint32_t *heads, hipStream_t st) {
    if (!c.total) return hipSuccess;
    hipError_t e = launch_raydirs(dirs, kClassifyRays, 0.f, st);
    if (e != hipSuccess) return e;
    if ((e = hipMemsetAsync(heads, 0, (size_t)kFetchHeads * kHeadStride * sizeof(uint32_t), st)) != hipSuccess) return e;
    launch_walk<WALK_CLOSEST>(s, ClassifySrc{c, dirs, hits}, nullptr, c.total * kClassifyRays, heads, nullptr,
                              trace_lds_bytes(s.tlas_count, s.inst_count), st);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_classify, dim3((c.total + 255) / 256), dim3(256), 0, st, s, c, dirs, hits, reloc_clamp, meta);
    return hipGetLastError();
}

hipError_t launch_emitter_vis(const DevScene &s, const DevCascade &c, const sbgi_tick_params &p, const float *meta,
                              float *vis, float4 *scratch, uint32_t *heads, hipStream_t st) {
    const uint32_t n = c.updated * s.emitter_count;
    if (!n) return hipSuccess;
    const float bias = hit_bias(c.cell_min, p.dbg_trace_bias, p.dbg_ray_eps);
    hipLaunchKernelGGL(k_emitter_rays, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, s, c, p.classify_strength,
                       p.emitter_vis_seed, bias, meta, scratch);
    launch_walk<WALK_ANY>(s, EmitterSrc{scratch, vis}, nullptr, n, heads, nullptr,
                          trace_lds_bytes(s.tlas_count, s.inst_count), st);
    return hipGetLastError();
}
