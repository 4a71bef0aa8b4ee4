// Operator-split upwind film advection on the device (the transport row
// update of nft_transport, src/sim/nft.cpp:1055-1081, and the plug-flow
// relaxation of dbd_flow_exchange, src/sim/dbd.cpp:468-483).
//
// Per substep every node of a channel takes x_j + c_j (x_{j-1} - x_j) from the
// OLD value of its upstream neighbour (node 0 reads the inflow): a Jacobi
// update, so all nodes of a substep are independent. One workgroup owns one
// (field, channel) row of `nodes` values: it reads the row once, keeps it in
// registers (chunk of consecutive nodes per thread) for every substep,
// exchanges only chunk boundaries through LDS, and writes the row once — HBM
// traffic per launch is one read and one write of the field, whatever the
// substep count. Rows are [field][padded] of the zone block with cell =
// channel * nodes + node.
#include <hip/hip_runtime.h>

#include "simk_gpu_api.h"
#include "module_internal.h"

namespace {

constexpr unsigned kThreads = 256;
constexpr unsigned kMaxChunk = 32;   // nodes per thread: up to 8192 nodes per channel

template <class T>
__device__ __forceinline__ T fma_t(T a, T b, T c);
template <>
__device__ __forceinline__ double fma_t<double>(double a, double b, double c) {
    return __builtin_fma(a, b, c);
}
template <>
__device__ __forceinline__ float fma_t<float>(float a, float b, float c) {
    return __builtin_fmaf(a, b, c);
}

// One (field, channel) row of `nodes` values at xr, advanced `substeps`
// substeps by the whole workgroup (kThreads threads, CHUNK nodes each).
// FUSED: x + c (up - x) as one fused multiply-add (nft.cpp:1072); otherwise
// the plug-flow form x + (c * (up - x)) with the product rounded on its own
// (dbd.cpp:480). cr null: every node's coefficient is c_all.
template <class T, unsigned CHUNK, bool FUSED = true>
__device__ __forceinline__ void upwind_row(T *xr, const T *cr, unsigned nodes, unsigned substeps, T in,
                                           T *outlet, T c_all = T(0)) {
    const unsigned t = threadIdx.x;
    __shared__ T stage[kThreads * CHUNK];
    __shared__ T last[kThreads];

    // Coalesced load of the row through LDS, then each thread takes its chunk.
    for (unsigned j = t; j < kThreads * CHUNK; j += kThreads) stage[j] = j < nodes ? xr[j] : T(0);
    __syncthreads();
    T v[CHUNK], c[CHUNK];
#pragma unroll
    for (unsigned i = 0; i < CHUNK; ++i) v[i] = stage[t * CHUNK + i];
    __syncthreads();
    if (cr) {
        for (unsigned j = t; j < kThreads * CHUNK; j += kThreads) stage[j] = j < nodes ? cr[j] : T(0);
        __syncthreads();
#pragma unroll
        for (unsigned i = 0; i < CHUNK; ++i) c[i] = stage[t * CHUNK + i];
    } else {
#pragma unroll
        for (unsigned i = 0; i < CHUNK; ++i) c[i] = c_all;
    }

    const unsigned first = t * CHUNK;
    const unsigned last_thread = (nodes - 1) / CHUNK;
    T out = (outlet && t == last_thread) ? *outlet : T(0);
    for (unsigned s = 0; s < substeps; ++s) {
        if (outlet && t == last_thread) out = out + v[(nodes - 1) - first];
        last[t] = v[CHUNK - 1];
        __syncthreads();
        const T left = t ? last[t - 1] : in;
        // Right to left inside the chunk so each update reads the old neighbour.
#pragma unroll
        for (unsigned i = CHUNK; i-- > 0;) {
            const T up = i ? v[i - 1] : left;
            if (first + i < nodes) v[i] = FUSED ? fma_t<T>(c[i], up - v[i], v[i]) : v[i] + c[i] * (up - v[i]);
        }
        __syncthreads();
    }
    if (outlet && t == last_thread) *outlet = out;

#pragma unroll
    for (unsigned i = 0; i < CHUNK; ++i) stage[t * CHUNK + i] = v[i];
    __syncthreads();
    for (unsigned j = t; j < nodes; j += kThreads) xr[j] = stage[j];
}

template <class T, unsigned CHUNK>
__global__ void __launch_bounds__(kThreads) simk_upwind_kernel(T *x, unsigned long long padded, unsigned channels,
                                                               unsigned nodes, const T *courant, unsigned substeps,
                                                               const T *inflow, T *outlet_sum) {
    const unsigned row = blockIdx.x;   // field * channels + channel
    const unsigned field = row / channels, ch = row % channels;
    upwind_row<T, CHUNK>(x + field * padded + (unsigned long long)ch * nodes, courant + (unsigned long long)ch * nodes,
                         nodes, substeps, inflow ? inflow[field] : T(0), outlet_sum ? outlet_sum + row : nullptr);
}

// Plug flow along a chain of zones (dbd_flow_exchange, dbd.cpp:465-483): one
// workgroup per species row; zone 0 relaxes toward the feed.
template <unsigned CHUNK>
__global__ void __launch_bounds__(kThreads) simk_plugflow_kernel(double *y, unsigned long long padded, unsigned zones,
                                                                 double a, const double *feed) {
    const unsigned s = blockIdx.x;
    upwind_row<double, CHUNK, false>(y + s * padded, nullptr, zones, 1, feed[s], nullptr, a);
}

// Many requests' rows in one launch: block b belongs to the segment whose
// [row_begin, row_begin + fields * channels) holds b (segments sorted by
// row_begin, the first at 0).
template <unsigned CHUNK>
__global__ void __launch_bounds__(kThreads) simk_upwind_batch_kernel(const SimkUpwindSeg *segs, unsigned count,
                                                                     unsigned substeps) {
    const unsigned b = blockIdx.x;
    unsigned lo = 0, hi = count;   // last segment with row_begin <= b
    while (hi - lo > 1) {
        const unsigned mid = (lo + hi) / 2;
        if (segs[mid].row_begin <= b) lo = mid;
        else hi = mid;
    }
    const SimkUpwindSeg &S = segs[lo];
    const unsigned row = b - S.row_begin;
    const unsigned field = row / S.channels, ch = row % S.channels;
    upwind_row<double, CHUNK>(S.x + field * S.padded + (unsigned long long)ch * S.nodes,
                              S.courant + (unsigned long long)ch * S.nodes, S.nodes, substeps,
                              S.inflow ? S.inflow[field] : 0.0, nullptr);
}

template <class T>
hipError_t launch_upwind(hipStream_t stream, T *x, u64 padded, u32 fields, u32 channels, u32 nodes,
                         const T *courant, u32 substeps, const T *inflow, T *outlet_sum) {
    if (!x || !courant || fields == 0 || channels == 0 || nodes == 0) return hipErrorInvalidValue;
    if ((u64)channels * nodes > padded) return hipErrorInvalidValue;
    const u32 chunk_needed = (nodes + kThreads - 1) / kThreads;
    const dim3 grid(fields * channels), block(kThreads);
    if (chunk_needed <= 4)
        hipLaunchKernelGGL((simk_upwind_kernel<T, 4>), grid, block, 0, stream, x, (unsigned long long)padded, channels,
                           nodes, courant, substeps, inflow, outlet_sum);
    else if (chunk_needed <= 16)
        hipLaunchKernelGGL((simk_upwind_kernel<T, 16>), grid, block, 0, stream, x, (unsigned long long)padded,
                           channels, nodes, courant, substeps, inflow, outlet_sum);
    else if (chunk_needed <= kMaxChunk)
        hipLaunchKernelGGL((simk_upwind_kernel<T, kMaxChunk>), grid, block, 0, stream, x, (unsigned long long)padded,
                           channels, nodes, courant, substeps, inflow, outlet_sum);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace

hipError_t simk_enqueue_upwind_batch(hipStream_t stream, const SimkUpwindSeg *segs, u32 count, u32 total_rows,
                                     u32 max_nodes, u32 substeps) {
    if (!segs || count == 0 || total_rows == 0 || max_nodes == 0) return hipErrorInvalidValue;
    const u32 chunk_needed = (max_nodes + kThreads - 1) / kThreads;
    const dim3 grid(total_rows), block(kThreads);
    if (chunk_needed <= 4)
        hipLaunchKernelGGL((simk_upwind_batch_kernel<4>), grid, block, 0, stream, segs, count, substeps);
    else if (chunk_needed <= 16)
        hipLaunchKernelGGL((simk_upwind_batch_kernel<16>), grid, block, 0, stream, segs, count, substeps);
    else if (chunk_needed <= kMaxChunk)
        hipLaunchKernelGGL((simk_upwind_batch_kernel<kMaxChunk>), grid, block, 0, stream, segs, count, substeps);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

extern "C" {

GPU_API s32 simk_upwind_f64(f64 *y, u64 padded, u32 fields, u32 channels, u32 nodes, const f64 *courant,
                            u32 substeps, const f64 *inflow) {
    return (s32)launch_upwind<f64>(simk_module_stream(), y, padded, fields, channels, nodes, courant, substeps,
                                   inflow, nullptr);
}

GPU_API s32 simk_plugflow_f64(f64 *y, u64 padded, u32 species, u32 zones, f64 a, const f64 *feed) {
    if (!y || !feed || species == 0 || zones == 0 || zones > padded) return (s32)hipErrorInvalidValue;
    if (!(a >= 0.0 && a <= 1.0)) return (s32)hipErrorInvalidValue;   // the reference asserts a <= 1 (dbd.cpp:473)
    const u32 chunk_needed = (zones + kThreads - 1) / kThreads;
    const dim3 grid(species), block(kThreads);
    hipStream_t st = simk_module_stream();
    if (chunk_needed <= 4)
        hipLaunchKernelGGL((simk_plugflow_kernel<4>), grid, block, 0, st, y, (unsigned long long)padded, zones, a, feed);
    else if (chunk_needed <= 16)
        hipLaunchKernelGGL((simk_plugflow_kernel<16>), grid, block, 0, st, y, (unsigned long long)padded, zones, a,
                           feed);
    else if (chunk_needed <= kMaxChunk)
        hipLaunchKernelGGL((simk_plugflow_kernel<kMaxChunk>), grid, block, 0, st, y, (unsigned long long)padded, zones,
                           a, feed);
    else
        return (s32)hipErrorInvalidValue;
    return (s32)hipGetLastError();
}

GPU_API s32 simk_upwind_f32(f32 *x, u64 padded, u32 fields, u32 channels, u32 nodes, const f32 *courant,
                            u32 substeps, f32 *outlet_sum) {
    return (s32)launch_upwind<f32>(simk_module_stream(), x, padded, fields, channels, nodes, courant, substeps,
                                   nullptr, outlet_sum);
}

}  // extern "C"
