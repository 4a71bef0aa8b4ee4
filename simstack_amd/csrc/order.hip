// Queue order for repeated advances of one zone block: the cells longest
// first, by the step attempts each took in the previous advance of the block.
// Lanes take cells in queue order, so the costliest cells start in the first
// resident round instead of landing in the last one, where a few long cells
// would keep most lanes idle. Consecutive operator-split advances of a cell
// cost about the same, so the previous advance is the estimate. The order
// changes which lane runs a cell and when, never the cell's arithmetic.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "module_internal.h"

namespace {

__global__ void __launch_bounds__(256) simk_attempts_before_kernel(const f64 *steps, const f64 *rejects, f64 *before,
                                                                   u32 n) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) before[i] = steps[i] + rejects[i];
}

// key = the cell's attempts in the advance just run (the counter rows
// accumulate whole numbers, so the difference is exact); value = the cell
__global__ void __launch_bounds__(256) simk_attempts_key_kernel(const f64 *steps, const f64 *rejects,
                                                                const f64 *before, u32 *keys, u32 *cells, u32 n) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const f64 d = steps[i] + rejects[i] - before[i];
    keys[i] = d > 0.0 ? (d < 4.0e9 ? (u32)d : 4000000000u) : 0u;
    cells[i] = i;
}

__global__ void __launch_bounds__(256) simk_order_fill_kernel(u32 *keys, u32 *cells, u32 n) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    keys[i] = 0u;
    cells[i] = i;
}

}  // namespace

void simk_cost_order_free(SimkCostOrder &c) {
    if (c.before) (void)hipFree(c.before);
    if (c.keys) (void)hipFree(c.keys);
    if (c.temp) (void)hipFree(c.temp);
    c = SimkCostOrder();
}

hipError_t simk_cost_order_reserve(SimkCostOrder &c, hipStream_t st, u32 lanes) {
    if (lanes > c.cap) {
        // one allocation for the four u32 arrays (keys, sorted keys, cells, order)
        if (c.before) (void)hipFree(c.before);
        if (c.keys) (void)hipFree(c.keys);
        c.before = nullptr;
        c.keys = nullptr;
        c.cap = 0;
        c.valid = false;
        hipError_t e = hipMalloc(&c.before, (u64)lanes * sizeof(f64));
        if (e == hipSuccess) e = hipMalloc(&c.keys, 4ull * lanes * sizeof(u32));
        if (e != hipSuccess) return e;
        c.keys_out = c.keys + lanes;
        c.cells = c.keys_out + lanes;
        c.order = c.cells + lanes;
        c.cap = lanes;
    }
    size_t need = 0;
    hipError_t e = hipcub::DeviceRadixSort::SortPairsDescending(nullptr, need, c.keys, c.keys_out, c.cells, c.order,
                                                                (int)lanes, 0, 32, st);
    if (e != hipSuccess) return e;
    if (need > c.temp_bytes) {
        if (c.temp) (void)hipFree(c.temp);
        c.temp = nullptr;
        c.temp_bytes = 0;
        e = hipMalloc(&c.temp, need);
        if (e != hipSuccess) return e;
        c.temp_bytes = need;
    }
    if (!c.warmed) {
        hipLaunchKernelGGL(simk_order_fill_kernel, dim3((lanes + 255) / 256), dim3(256), 0, st, c.keys, c.cells, lanes);
        e = hipGetLastError();
        size_t bytes = c.temp_bytes;
        if (e == hipSuccess)
            e = hipcub::DeviceRadixSort::SortPairsDescending(c.temp, bytes, c.keys, c.keys_out, c.cells, c.order,
                                                             (int)lanes, 0, 32, st);
        if (e != hipSuccess) return e;
        c.warmed = true;
    }
    return hipSuccess;
}

hipError_t simk_cost_order_begin(SimkCostOrder &c, hipStream_t st, const f64 *steps, const f64 *rejects, u32 lanes) {
    hipError_t e = simk_cost_order_reserve(c, st, lanes);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(simk_attempts_before_kernel, dim3((lanes + 255) / 256), dim3(256), 0, st, steps, rejects,
                       c.before, lanes);
    return hipGetLastError();
}

hipError_t simk_cost_order_end(SimkCostOrder &c, hipStream_t st, const f64 *steps, const f64 *rejects, u32 lanes) {
    if (lanes > c.cap) return hipErrorInvalidValue;
    hipLaunchKernelGGL(simk_attempts_key_kernel, dim3((lanes + 255) / 256), dim3(256), 0, st, steps, rejects,
                       (const f64 *)c.before, c.keys, c.cells, lanes);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    size_t bytes = c.temp_bytes;
    return hipcub::DeviceRadixSort::SortPairsDescending(c.temp, bytes, c.keys, c.keys_out, c.cells, c.order,
                                                        (int)lanes, 0, 32, st);
}
