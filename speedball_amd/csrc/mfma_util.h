// mfma_util.h — helpers shared by the blend and reflection contractions: the octahedral
// texel direction, BF16/FP16 operand conversion, the 16x16x32 MFMA step and the hi/lo
// operand split, and the decoded ray distance of a ray record.
#pragma once
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include "sbgi_dev.h"

__device__ __forceinline__ float3 oct_decode_uv(float u, float v) {
    float fx = u * 2.f - 1.f, fy = v * 2.f - 1.f;
    float nz = 1.f - fabsf(fx) - fabsf(fy);
    float t = fmaxf(-nz, 0.f);
    float nx = fx + (fx >= 0.f ? -t : t);
    float ny = fy + (fy >= 0.f ? -t : t);
    return normalize3(f3(nx, ny, nz));
}

__device__ __forceinline__ uint16_t to_bf16(float x) { return __bfloat16_as_ushort(__float2bfloat16(x)); }
__device__ __forceinline__ float from_bf16(uint16_t h) { return __bfloat162float(__ushort_as_bfloat16(h)); }
__device__ __forceinline__ uint16_t to_f16(float x) { return __half_as_ushort(__float2half_rn(x)); }
__device__ __forceinline__ float from_f16(uint16_t h) { return __half2float(__ushort_as_half(h)); }

__device__ __forceinline__ float ray_dist(float ht, float max_dist) {
    const float dec = ht < -1.5f ? -ht - 2.f : ht;
    return ht == -1.f ? max_dist : dec;
}

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

template <uint32_t PREC>
__device__ __forceinline__ f32x4 mfma(u16x8 a, u16x8 b, f32x4 acc) {
    if constexpr (PREC == SBGI_BLEND_FP16) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a), __builtin_bit_cast(f16x8, b), acc,
                                                      0, 0, 0);
    } else {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b),
                                                       acc, 0, 0, 0);
    }
}

template <uint32_t PREC>
__device__ __forceinline__ void split(float v, uint16_t &hi, uint16_t &lo) {
    if constexpr (PREC == SBGI_BLEND_FP16) {
        hi = to_f16(v);
        lo = 0;
    } else {
        hi = to_bf16(v);
        lo = PREC == SBGI_BLEND_BF16X3 ? to_bf16(v - from_bf16(hi)) : (uint16_t)0;
    }
}


// the value an MFMA A operand of blend precision `prec` carries for weight w (hi + lo for
// BF16x3): the normalisers sum these so that a uniform input is reproduced exactly
__device__ __forceinline__ float quantized(float w, uint32_t prec) {
    if (prec == SBGI_BLEND_FP16) return from_f16(to_f16(w));
    if (prec == SBGI_BLEND_BF16) return from_bf16(to_bf16(w));
    if (prec == SBGI_BLEND_BF16X3) {
        const uint16_t h = to_bf16(w);
        return from_bf16(h) + from_bf16(to_bf16(w - from_bf16(h)));
    }
    return w;
}

// A-fragment slot of weight (row, k) in the tables: element j of lane l of tile (mt, kb) is
// W[16 mt + l % 16][32 kb + 8 (l / 16) + j]
__device__ __forceinline__ size_t frag_index(uint32_t row, uint32_t k, uint32_t KB) {
    const uint32_t kb = k / 32, g = (k % 32) / 8, j = k % 8;
    const uint32_t mt = row / 16, lane = g * 16 + row % 16;
    return (((size_t)mt * KB + kb) * 64 + lane) * 8 + j;
}
