// Internal interface between the translation units of libsimk_gpu.so (the
// solver module, the advection kernels and the serving loop). Not exported.
#ifndef SIMK_MODULE_INTERNAL_H
#define SIMK_MODULE_INTERNAL_H

#include <hip/hip_runtime.h>

#include "simk_gpu_api.h"

// The module's one stream (non-blocking, default priority): every launch of
// the library is ordered on it.
hipStream_t simk_module_stream();

// simk_plugflow_f64 on a given stream (the serving loop's DBD requests).
hipError_t simk_enqueue_plugflow(hipStream_t stream, f64 *y, u64 padded, u32 species, u32 zones, f64 a,
                                 const f64 *feed);

// Queue order by the previous advance's per-cell attempts (order.hip): the
// key (block rows, lane count) it was measured on, and device buffers.
struct SimkCostOrder {
    const f64 *y = nullptr;   // block the order belongs to
    u32 padded = 0, lanes = 0;
    bool valid = false;
    u32 cap = 0;              // cells the buffers hold
    f64 *before = nullptr;    // steps + rejects before the advance
    u32 *keys = nullptr, *keys_out = nullptr, *cells = nullptr, *order = nullptr;
    void *temp = nullptr;     // radix-sort scratch
    size_t temp_bytes = 0;
    bool warmed = false;      // the sort has run once on these buffers
};
// Sizes the buffers for `lanes` cells (allocates when they grow); the first
// time, also runs the sort once, so that its one-time set-up (kernel loading)
// is not counted in the first ordered advance.
hipError_t simk_cost_order_reserve(SimkCostOrder &c, hipStream_t stream, u32 lanes);
// Before an advance of `lanes` cells: sizes the buffers, records steps + rejects.
hipError_t simk_cost_order_begin(SimkCostOrder &c, hipStream_t stream, const f64 *steps, const f64 *rejects, u32 lanes);
// After it (same stream): c.order = the cells by descending attempts in that advance.
hipError_t simk_cost_order_end(SimkCostOrder &c, hipStream_t stream, const f64 *steps, const f64 *rejects, u32 lanes);
void simk_cost_order_free(SimkCostOrder &c);

// One request's advection rows for the batched upwind kernel: fields x
// channels rows of `nodes` values at x + field * padded + channel * nodes.
struct SimkUpwindSeg {
    double *x;
    unsigned long long padded;
    const double *courant;   // [channels * nodes]
    const double *inflow;    // [fields] or null (zero inflow)
    unsigned fields, channels, nodes;
    unsigned row_begin;      // first grid row of this segment (prefix sum of fields * channels)
};

// Enqueues the batched upwind kernel on `stream` (segs: device array sorted by
// row_begin, the first at 0; max_nodes <= 8192).
hipError_t simk_enqueue_upwind_batch(hipStream_t stream, const SimkUpwindSeg *segs, u32 count, u32 total_rows,
                                     u32 max_nodes, u32 substeps);

// The mechanisms of the last successful simk_kin_set_mechs.
struct SimkMultiShape {
    u32 count;               // 0: none set
    u32 slots;               // workspace f64 slots per lane of the multi kernel
    u32 resident_waves;      // waves the device holds at once for it
    u32 n[4], R[4];
};
SimkMultiShape simk_multi_shape();

// Time slicing of a multi-mechanism launch (see SimkKinArgs::slice).
struct SimkSlice {
    u64 ticks;               // 100 MHz real-time counter ticks per wave (0: off)
    double *resume[4];       // per mechanism: [5][padded] resume rows
    u32 *carry[4];           // per mechanism: stopped cells, waves * 64 entries
};

// Serving: one f32 delta-lane advance of L->duration for queue positions
// [0, lanes) of `order` (lanes of mechanism `index` of the simk_kin_set_mechs
// list), run to completion with that mechanism's generated f32 delta kernel
// (compiled on first use), on the module stream.
hipError_t simk_enqueue_d32_ext(u32 index, const GpuKinLaunch *L, u32 lanes, const u32 *order);

// Enqueues the queue reset and one multi-mechanism persistent launch on the
// module stream. Ls[i]: mechanism i's block (duration, padded, pointers);
// lanes[i]: queue positions of mechanism i (may be 0); orders[i]: device
// position -> lane map (null: identity). ws: waves * slots * 64 f64; queue:
// 8 u32 (queue heads [0, 4), carry counts [4, 8)), both owned by the caller.
// After the launch queue[i] (clamped to lanes[i]) is the number of queue
// positions taken and queue[4 + i] the number of carried cells.
hipError_t simk_enqueue_multi_ext(u32 count, const GpuKinLaunch *Ls, const u32 *lanes, const u32 *const *orders,
                                  u32 repeats, double *ws, u32 *queue, u32 waves, const SimkSlice *slice);

#endif  // SIMK_MODULE_INTERNAL_H
