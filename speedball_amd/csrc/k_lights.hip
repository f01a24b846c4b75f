// k_lights.hip — the clustered light grid of the NEE loop.
//
// Restates lightGridKernel (js/gi_probes.js:1888-1938): one list per cell of cascade 0's
// probe grid, holding the lights whose range can reach the cell, in light order. A light
// belongs to a cell when it is directional, unranged, or its range (padded by
// max(1e-4, 1e-5 range) so that membership stays conservative) reaches the cell's box.
// Emitters (type 3) are never members: the NEE loop skips them. A cell with more than 16
// members stores the overflow count 17, and the NEE loop then visits every light.
#pragma clang fp contract(off)

#include "kernels.h"

namespace {

constexpr uint32_t kStride = SBGI_LIGHTS_PER_CELL + 1;

__global__ void __launch_bounds__(256) k_light_grid(const float *__restrict__ lights, uint32_t light_count, float3 gmin,
                                                    float3 gsize, uint3 cells, uint32_t *__restrict__ grid) {
    const uint32_t cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= cells.x * cells.y * cells.z) return;
    const uint32_t ix = cell % cells.x, iy = (cell / cells.x) % cells.y, iz = cell / (cells.x * cells.y);
    const float3 size = f3(gsize.x / (float)cells.x, gsize.y / (float)cells.y, gsize.z / (float)cells.z);
    const float3 cmin = gmin + f3((float)ix, (float)iy, (float)iz) * size;
    const float3 cmax = cmin + size;
    uint32_t *list = grid + (size_t)cell * kStride;
    uint32_t count = 0;
    for (uint32_t li = 0; li < light_count; li++) {
        const float *L = lights + (size_t)li * SBGI_LIGHT_STRIDE;
        const float ltype = L[0], range = L[10];
        const float dx = fmaxf(fmaxf(cmin.x - L[1], 0.f), L[1] - cmax.x);
        const float dy = fmaxf(fmaxf(cmin.y - L[2], 0.f), L[2] - cmax.y);
        const float dz = fmaxf(fmaxf(cmin.z - L[3], 0.f), L[3] - cmax.z);
        const float padded = range + fmaxf(1e-4f, range * 1e-5f);
        const bool reaches = (dx * dx + dy * dy) + dz * dz <= padded * padded;
        const bool member = ltype < 2.5f && (ltype < 0.5f || range <= 0.f || reaches);
        if (member) {
            if (count < SBGI_LIGHTS_PER_CELL) list[1 + count] = li;
            count++;
        }
    }
    // unused slots are cleared so that the lists compare word for word
    for (uint32_t i = count; i < SBGI_LIGHTS_PER_CELL; i++) list[1 + i] = 0u;
    list[0] = count > SBGI_LIGHTS_PER_CELL ? SBGI_LIGHTS_PER_CELL + 1 : count;
}

}  // namespace

hipError_t launch_light_grid(const float *lights, uint32_t light_count, const float *gmin, const float *gsize,
                             const uint32_t *cells, uint32_t *grid, hipStream_t st) {
    const uint32_t n = cells[0] * cells[1] * cells[2];
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_light_grid, dim3((n + 255) / 256), dim3(256), 0, st, lights, light_count,
                       make_float3(gmin[0], gmin[1], gmin[2]), make_float3(gsize[0], gsize[1], gsize[2]),
                       make_uint3(cells[0], cells[1], cells[2]), grid);
    return hipGetLastError();
}
