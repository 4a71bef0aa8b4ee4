/* simk_gpu_api.h — C ABI of libsimk_gpu.so, the MI355X GPU module.
 *
 * The first block is the reference's GPU module interface, same names, same
 * signatures, same meaning (src/gpu/gpu_api.h:47-115), so the reference's
 * loader (src/platform/linux_gpu.cpp:7-31) can dlopen this library in place of
 * its CUDA module. Return values are hipError_t cast to s32 (0 = success),
 * decoded by gpu_error_string. The module owns one device bump arena, one
 * non-blocking HIP stream and its events; calls are not re-entrant (one
 * thread at a time), as in the reference.
 *
 * The second block extends it for what the reference lacks: mechanisms larger
 * than its compile-time capacities, the kernel choice, mixed precision, the
 * operator-split advection stencil and per-call statistics.
 */
#ifndef SIMK_GPU_API_H
#define SIMK_GPU_API_H

#include "simk_types.h"

#ifdef __cplusplus
extern "C" {
#endif

#define GPU_API __attribute__((visibility("default")))

/* ---- reference interface (src/gpu/gpu_api.h) ---------------------------- */

/* gpu_api.h:47 — selects the device, reserves arena_bytes of device memory
 * for gpu_alloc, creates the stream and events, fills *info. */
GPU_API s32 gpu_init(u32 device_index, u64 arena_bytes, GpuInfo *info);
/* gpu_api.h:50 */
GPU_API const char *gpu_error_string(s32 error);
/* gpu_api.h:54 — bump allocation from the arena, 256-byte aligned; 0 when full. */
GPU_API void *gpu_alloc(u64 size);
/* gpu_api.h:57 */
GPU_API u64 gpu_arena_used(void);
/* gpu_api.h:62 — waits for the device, then rolls the arena back to empty. */
GPU_API void gpu_arena_reset(void);
/* gpu_api.h:65 — copy on the module stream, then wait for it. */
GPU_API s32 gpu_copy_to_device(void *dst_device, const void *src_host, u64 size);
/* gpu_api.h:68 */
GPU_API s32 gpu_copy_to_host(void *dst_host, const void *src_device, u64 size);
/* gpu_api.h:71 — waits for the module stream. */
GPU_API s32 gpu_sync(void);
/* gpu_api.h:75 — the reference's point-particle demo step (kept so the
 * reference loader finds every export). */
GPU_API s32 gpu_step_points(f32 *x, f32 *y, f32 *z, f32 *vx, f32 *vy, f32 *vz, u32 count, f32 dt);
/* gpu_api.h:80 — mechanism (KinMech at the default capacities), method and
 * control for the next advances. Generates and JIT-compiles the mechanism's
 * straight-line kernel (cached on disk) and uploads the table-walk tables. */
GPU_API s32 gpu_kin_set_mech(const KinMech *mech, const KinMethod *method, const KinControl *control);
/* gpu_api.h:85 — advances lanes [0, lane_count or padded) of the block over
 * [0, duration] and waits; *kernel_ms = kernel time from stream events.
 * Results are bit-identical to the reference's FP64 CPU path (precision 0). */
GPU_API s32 gpu_kin_advance(const GpuKinLaunch *launch, f32 *kernel_ms);
/* gpu_api.h:89 — records the range; the platform's pinned-memory path is
 * not used (copies from pageable memory are staged by the runtime). */
GPU_API s32 gpu_host_register(void *host, u64 size);
/* gpu_api.h:91 */
GPU_API s32 gpu_host_unregister(void *host);
/* gpu_api.h:99 — uploads lanes [0, lane_count) of every row of the
 * [rows][padded] block, launches, downloads the first mutable_rows rows, all
 * on the module stream; returns at once. One advance may be outstanding. */
GPU_API s32 gpu_kin_advance_async(const GpuKinLaunch *launch, void *host_block, void *device_block, u64 rows,
                                  u64 mutable_rows);
/* gpu_api.h:105 — waits for the async advance; kernel and total milliseconds. */
GPU_API s32 gpu_kin_wait(f32 *kernel_ms, f32 *total_ms);
/* gpu_api.h:109 — 0: one cell per lane (default); 1: a team of lanes per cell. */
GPU_API s32 gpu_kin_set_mode(u32 mode);

/* ---- extensions (no counterpart in the reference) ----------------------- */

/* gpu_kin_set_mech for a KinMech compiled with capacities caps (mechanisms
 * beyond the default 32 species / 96 reactions / 512 non-zeros). */
GPU_API s32 simk_kin_set_mech_caps(const void *mech, const SimkCaps *caps, const KinMethod *method,
                                   const KinControl *control);

/* `repeats` consecutive advances of the launch's duration in ONE persistent
 * launch, carrying h — the result of `repeats` gpu_kin_advance calls with the
 * block's inputs unchanged between them (for operator-split steps with no
 * host work in between). */
GPU_API s32 simk_kin_advance_n(const GpuKinLaunch *launch, u32 repeats, f32 *kernel_ms);

/* simk_kin_advance_n without any device allocation. It has no counterpart in
 * the reference. It is for callers that cannot allocate device memory: code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc. The persistent kernels' workspace and queue counter come from a
 * fixed 64 MiB array in the library's code object (one copy per GPU, held
 * until the process ends, its address resolved at gpu_init). The grid is
 * lowered to the waves whose workspace fits it, which changes no result.
 * Errors: hipErrorInvalidValue when one wave's workspace alone exceeds
 * 64 MiB - 256 B (a mechanism of about 130,000 workspace slots per lane);
 * hipErrorNotReady when the launch would first have to load a kernel (the
 * mixed-precision and f32 delta kernels load on their first ordinary
 * advance; call gpu_kin_advance once with that precision beforehand). Calls
 * share the one array per GPU, so two calls that overlap on a GPU race. */
GPU_API s32 simk_kin_advance_static(const GpuKinLaunch *launch, u32 repeats, f32 *kernel_ms);

/* Lanes per cell in cooperative mode (gpu_kin_set_mode(1)): 64 (one wave per
 * cell, the default), 32, 16 or 8 (that many cells share a wave, each team
 * with its own LDS region and control flow). Results do not depend on it. */
GPU_API s32 simk_kin_set_team(u32 lanes_per_cell);

/* Control of mechanism `index` of the simk_kin_set_mechs list (ragged
 * batches and serving; simk_kin_set_mechs gives every mechanism the same
 * one): for a DBD mechanism among synthetic ones, say. */
GPU_API s32 simk_kin_set_mech_control(u32 index, const KinControl *control);

/* Persistent grid size in waves (0: as many as are resident on the device). */
GPU_API s32 simk_kin_set_grid(u32 waves);

/* Queue order (device array of lane_count cell indices, a permutation) in
 * which lanes take cells; null restores the natural order. Results per cell
 * do not depend on it. The array must stay valid while advances run. */
GPU_API s32 simk_kin_set_order(const u32 *order_device);

/* Queue order of single-mechanism advances when no order is set:
 * 1 (default): an advance of the same block as the previous advance (the
 * same rows, row stride and cell count) of at least 512 cells takes its
 * cells longest first, by the step attempts each took in that previous
 * advance; lanes take cells in queue order, so the costliest cells start
 * early instead of stretching the last round, and cells of like cost share
 * a wave. Consecutive operator-split advances of a cell cost about the
 * same. 0: natural order. Results per cell are the same either way; the
 * first advance of a block, the allocation-free advance, cooperative mode
 * and the multi-mechanism kernel keep the natural order. Setting the policy
 * forgets the recorded costs. */
GPU_API s32 simk_kin_set_queue_policy(u32 policy);

/* Kernel used by the FP64 and f32 delta advances: 0 auto (generated when it
 * compiled, else table), 1 generated straight-line, 2 table walk. */
GPU_API s32 simk_kin_set_kernel(u32 kind);

/* Generator knobs (speed only; results unchanged): see SimkGenOptions. */
typedef struct SimkGenOptions {
    u32 fence_rows;      /* 1: factored rows re-read from the workspace; 0: kept in registers */
    u32 waves_per_eu;    /* amdgpu_waves_per_eu hint (0: none) */
    u32 unroll_stages;   /* 1: unroll the Rosenbrock stage loop */
    u32 reg_window;      /* rows back a finished row stays in registers (0: 16 in one-mechanism FP64 kernels, 8 in
                            mixed precision and the multi-mechanism kernel; ~0u: all) */
    u32 batch;           /* loads issued this many statements ahead (0: 24; 16 in f32 delta lanes and refined mixed
                            precision) */
    u32 precision;       /* simk_jit_precompile only: 0 FP64 kernel, 1 f32 delta kernel, 2 mixed-precision kernel,
                            3 FP32 factor-storage kernel */
    u32 sweeps;          /* mixed precision: FP64-residual refinement sweeps per stage (0: 1; ~0u: none) */
    u32 reuse;           /* a rate constant loaded this many statements back is reused (0: 384 in one-mechanism FP64
                            kernels, 1000 in unrefined mixed precision, else 96; ~0u: never) */
    u32 lds_slots;       /* LU entries per lane held in LDS instead of the workspace (0: default: 160 f64 entries
                            = 80 KiB per wave; 200 f32 = 50 KiB in unrefined mixed precision; 160 f32 =
                            40 KiB in f32 delta lanes; none for refined mixed kernels; ~0u: none; at most
                            160 KiB). FP32 factor storage: the lane's LDS budget in f32 words, the FP64
                            pivots (two words each, always in LDS) first, off-diagonal entries in the rest
                            (0: 304 words = 76 KiB per wave) */
    u32 fun_in_jac;      /* 0: default (on); 1: the -J pass also evaluates f(y) for stage 0 (one rate-constant
                            pass fewer per attempt); ~0u: stage 0 evaluates f(y) itself */
    u32 pair_slots;      /* FP64 kernels: 1: workspace slots in adjacent pairs per lane ([pair][64 lanes][2]), the
                            LU entries in the triangular solves' read order, so that two slots move in one
                            16-byte access; ~0u: one 8-byte slot per access; 0: default (paired) */
    u32 k_pairs;         /* paired layouts: 1: a second copy of the cell's rate constants in paired slots, which the
                            f passes of stages read two per 16-byte access (the -J pass keeps the single copy);
                            ~0u: none; 0: default (on for one-mechanism kernels) */
    u32 stage_carry;     /* 1: a stage whose combination ends with the previous stage's K (the point of an f
                            evaluation, or the saved f plus the c-combination) takes that term from the
                            registers that still hold it at the end of the previous stage: one stage-vector
                            load fewer per such stage, the same bits; ~0u: every term is loaded; 0: default
                            (off; the autotuners try it) */
} SimkGenOptions;
GPU_API s32 simk_kin_set_gen_options(const SimkGenOptions *options);

/* Persistent launches keep the workspace of all resident waves within this
 * many bytes (default 208 MiB, below the MI355X's 256 MiB Infinity Cache; the
 * environment variable SIMK_CACHE_BUDGET_MB sets it at gpu_init; 0 = launch
 * as many waves as fit the register file). simk_kin_set_grid overrides it. */
GPU_API s32 simk_kin_set_cache_budget(u64 bytes);

/* Autotuner: times one advance of `duration` seconds (0: L's duration) of
 * cells [0, sample_cells) of L (all cells when 0; on a device copy, so L's
 * block is untouched) under candidate settings and
 * keeps the fastest for later FP64 advances of the current mechanism:
 *   1. persistent waves (the cache-aware default, 1, 1.5, 2, 2.5, 3 or 4 per CU);
 *   2. at the best grid, the generator's load-pipeline depth (8, 16), register
 *      window (8, 32), unrolled stages, rate-constant reuse window (none, 48, 96),
 *      a two-waves-per-SIMD register budget, single instead of paired workspace
 *      slots, the paired rate-constant copy off, the stage carry on, and LU
 *      entries in LDS (none, 80, 128, 160), one knob at a time;
 *   3. teams of 64, 32, 16 or 8 lanes per cell (the cooperative form) instead
 *      of one lane per cell.
 * Every candidate computes the same bits; only time differs. The tuned
 * settings last until the next gpu_kin_set_mech / simk_kin_set_mechs, which
 * restores the caller's own.
 * With L->precision 2 it tunes the mixed-precision kernel instead, for later
 * mixed advances only (FP64 settings untouched): LU entries in LDS (the
 * default, 320, 240, 160), each at the cache-aware grid and at as many waves
 * per CU as that LDS share allows, then at the best pair the load-pipeline
 * depth (8, 12, 20, 24, 28, 32), register window (4, 16), rate-constant reuse
 * (none, 48, 96, 384), unrolled stages and the stage carry. With L->precision 3 the same for the
 * FP32 factor-storage kernel, its LDS budgets being the default (304), 320,
 * 288, 256 and 208 f32 words per lane, pivots included. simk_kin_set_gen_options
 * drops these settings. Precision 1 is refused (hipErrorInvalidValue). */
typedef struct SimkTuneResult {
    u32 mode;            /* 0 one lane per cell, 1 one wave per cell */
    u32 lanes_per_cell;  /* 1 or 64 */
    u32 waves;           /* persistent grid (0: cache-aware default) */
    u32 batch, reg_window, unroll_stages, reuse;
    u32 lds_slots;       /* LU entries in LDS of the chosen kernel */
    u32 candidates, sample_cells;
    f32 best_ms;         /* kernel time of the chosen settings on the sample */
    f32 default_ms;      /* the same under the defaults */
} SimkTuneResult;
GPU_API s32 simk_kin_autotune(const GpuKinLaunch *L, u32 sample_cells, f64 duration, SimkTuneResult *out);
/* The same for the multi-mechanism kernel of simk_kin_set_mechs (ragged
 * batches): a sample of sample_cells cells of every launches[i] (all when 0),
 * over the persistent grid and, at the best grid, the load-pipeline depth,
 * register window, unrolled stages, LU entries in LDS, single workspace
 * slots, the paired rate-constant copy and the stage carry. One lane per cell
 * is this kernel's only lane layout. Destroys every captured graph. The tuned
 * settings last until the next simk_kin_set_mechs / gpu_kin_set_mech. */
GPU_API s32 simk_kin_autotune_multi(u32 count, const GpuKinLaunch *launches, u32 sample_cells, f64 duration,
                                    SimkTuneResult *out);

/* Statistics of the current mechanism and the last advance. */
typedef struct SimkKinInfo {
    u32 kernel;          /* last launch: 1 generated FP64, 2 table FP64, 3 mixed precision, 4 multi-mechanism,
                            5 f32 delta (table walk), 6 / 7 one wave per cell (FP64 / f32 delta), 8 generated f32 delta,
                            9 FP32 factor storage, 10 multi-mechanism FP32 factor storage */
    u32 ws_slots;        /* workspace f64 slots per lane */
    u32 waves;           /* persistent waves of the last launch */
    u32 ordered;         /* 1: the last advance took its cells longest first (simk_kin_set_queue_policy) */
    u32 species, reactions, nnz, lu_div, lu_sub;
    u32 jit_cached;      /* 1: the generated kernel came from the disk cache */
    f64 jit_ms;          /* generate + compile (or cache load) time */
    f64 last_kernel_ms;
    u64 signature;       /* FNV-1a of the table content */
    char kernel_name[64];
} SimkKinInfo;
GPU_API s32 simk_kin_info(SimkKinInfo *out);

/* Generates and compiles the straight-line kernel for a mechanism into the
 * on-disk cache without touching a device (build step). 0 on success; the
 * compile log is written to log (may be null). */
GPU_API s32 simk_jit_precompile(const void *mech, const SimkCaps *caps, const KinMethod *method,
                                const SimkGenOptions *options, char *log, u32 log_size);

/* simk_jit_precompile for the multi-mechanism kernel of simk_kin_set_mechs
 * (FP64; same cache). */
GPU_API s32 simk_jit_precompile_multi(u32 count, const void *const *mechs, const SimkCaps *caps,
                                      const KinMethod *method, const SimkGenOptions *options, char *log,
                                      u32 log_size);

/* Ragged batches: up to four mechanisms (KinMech at capacities caps[i])
 * with one method and control, compiled into one persistent kernel. Destroys
 * every captured graph (they launch the previous kernel). */
GPU_API s32 simk_kin_set_mechs(u32 count, const void *const *mechs, const SimkCaps *caps, const KinMethod *method,
                               const KinControl *control);
/* One launch advancing every mechanism's block (launches[i] for mechanism i,
 * FP64): waves start on a mechanism in proportion to its estimated work and
 * move to the others as queues empty; each cell's result is that of
 * gpu_kin_advance on its own block. `repeats` as in simk_kin_advance_n. */
GPU_API s32 simk_kin_advance_multi(u32 count, const GpuKinLaunch *launches, u32 repeats, f32 *kernel_ms);
/* Captures `steps` consecutive simk_kin_advance_multi sequences (queue reset
 * kernel + persistent launch each, no memset nodes) into a HIP graph with its
 * own workspace and queue counters (allocated here, freed by
 * simk_graph_destroy); simk_graph_replay launches it on the module stream and
 * waits. The graph keeps the launches' device pointers: the blocks must stay
 * allocated while it is replayed. */
GPU_API s32 simk_graph_capture_multi(u32 count, const GpuKinLaunch *launches, u32 repeats, u32 steps,
                                     u32 *graph_id);
GPU_API s32 simk_graph_replay(u32 graph_id, f32 *ms);
GPU_API s32 simk_graph_destroy(u32 graph_id);

/* Operator-split upwind advection along channels (the transport row update
 * of nft_transport, src/sim/nft.cpp:1055-1081): fields are rows of a
 * [field][padded] array, cell = channel * nodes + node (nodes <= 8192). Each
 * of `substeps` substeps sets x_j <- x_j + c_j (x_{j-1} - x_j) from the old
 * upstream value, node 0 reading inflow[field] (0 when inflow is null), as
 * one fused multiply-add. courant: [channels * nodes]. Enqueued on the module
 * stream (ordered with the solver launches); gpu_sync waits for it. */
GPU_API s32 simk_upwind_f64(f64 *y, u64 padded, u32 fields, u32 channels, u32 nodes, const f64 *courant,
                            u32 substeps, const f64 *inflow);
/* The same in f32 with zero inflow; outlet_sum[field * channels + channel]
 * (may be null) accumulates the outlet node's value before every substep. */
GPU_API s32 simk_upwind_f32(f32 *x, u64 padded, u32 fields, u32 channels, u32 nodes, const f32 *courant,
                            u32 substeps, f32 *outlet_sum);

/* ---- serving loop: device-resident requests, continuous batching -------
 *
 * Solve requests over the mechanisms of simk_kin_set_mechs. Each request is
 * a zone block of `cells` cells (the reference's SoA layout, row stride
 * `padded`) laid out as `channels` channels of `nodes` nodes
 * (cell = channel * nodes + node). It runs `steps` operator-split steps: the
 * upwind advection of every species row along its channels (simk_upwind_f64
 * with the request's courant and inflow), then the chemistry advance of `dt`
 * seconds (h carried). Its block lives in HBM from admission to completion.
 *
 * simk_serve_submit queues a request. Each simk_serve_step is one iteration
 * of continuous batching: admit queued requests in arrival order while their
 * blocks fit the HBM budget (and a free lane range of their mechanism's
 * slab); advect, in one batched launch, every request whose cells all
 * finished their previous chemistry step; run one time-sliced
 * multi-mechanism persistent launch over every cell with chemistry pending
 * (smallest requests first; a cell still integrating when the slice ends
 * keeps its exact state and continues in the next iteration); copy finished
 * requests' solver rows (y, h, steps, rejects, pulses, failed) into their
 * result buffers. A cell's result does not depend on which other requests
 * share its iterations or where slices fall. Input blocks, courant and inflow arrays must stay
 * valid and unchanged from submit until the request is admitted, result
 * buffers until it is reported done. Not re-entrant, like the rest of the
 * module. */
typedef struct SimkServeConfig {
    u64 hbm_budget;      /* bytes of admitted zone blocks at once (0: 8 GiB) */
    u32 max_requests;    /* queued + admitted at once (0: 4096) */
    u32 steps;           /* operator-split steps per request (0: 20) */
    u32 substeps;        /* advection substeps per step (0: 4) */
    u32 slice_us;        /* chemistry time slice per iteration, microseconds (0: 20000; ~0u: no slicing); four
                            times as long while no submitted request waits for admission */
    f64 dt;              /* chemistry seconds per step (0: 1e-3) */
} SimkServeConfig;

struct SimkServeDbd;     /* simk_dbd.h */

enum SimkServeKind {
    SimkServe_Transport = 0,   /* per step: upwind advection along the channels, then chemistry */
    SimkServe_Dbd = 1,         /* per step: discharge source, chemistry, pulses, plug flow (simk_dbd.h) */
};

typedef struct SimkServeRequest {
    u32 mech;            /* index into the simk_kin_set_mechs list */
    u32 cells;           /* transport: channels * nodes; DBD: zones (<= 8192 with plug flow) */
    u32 channels, nodes; /* transport: nodes <= 8192; DBD: unused */
    u32 padded;          /* row stride of block (>= cells) */
    u32 kind;            /* SimkServeKind */
    const f64 *block;    /* host zone block (all rows, reference order) holding the request's inputs */
    const f64 *courant;  /* transport: [channels * nodes], each in [0, 1]; DBD: unused */
    const f64 *inflow;   /* [species] concentrations entering node 0 (DBD: the feed gas of the chain), or null */
    f64 *result;         /* host [species + 5][cells]: y, h, steps, rejects, pulses, failed after the last step */
    u32 steps;           /* operator-split steps of this request (0: the server's) */
    u32 precision;       /* 0: FP64 lanes (time-sliced with every other request); 1: f32 delta lanes (the
                            reference's default chemistry precision), each step run whole in a launch of the
                            mechanism's f32 delta kernel beside the FP64 one */
    const struct SimkServeDbd *dbd;   /* DBD requests: the zones' discharge description (read by simk_serve_submit only) */
} SimkServeRequest;

typedef struct SimkServeStats {
    u64 submitted, admitted, completed;
    u32 queued, active;          /* now */
    u64 iterations;
    u64 admitted_bytes;          /* now */
    u64 peak_admitted_bytes;
    f64 advect_ms, chem_ms;      /* summed kernel time (module stream events) */
    f64 copy_ms;                 /* summed admission + completion copy time */
    u64 carried;                 /* cells stopped by a slice end and continued later */
} SimkServeStats;

/* Allocates the slabs (one per mechanism: the zone block rows plus a courant
 * row and five resume rows per lane, sized so each could hold the whole
 * budget, capped where a mechanism's row offsets would pass 2^31 bytes),
 * workspace and staging. Needs simk_kin_set_mechs first; a second open
 * closes the first. */
GPU_API s32 simk_serve_open(const SimkServeConfig *config);
/* Queues a request; *ticket identifies it in simk_serve_step's done list.
 * hipErrorInvalidValue for a malformed request or one larger than the
 * budget; hipErrorOutOfMemory when max_requests are in flight. */
GPU_API s32 simk_serve_submit(const SimkServeRequest *request, u32 *ticket);
/* One iteration (waits for it). done[0 .. *n_done) = tickets finished in it
 * (at most max_done are reported; the rest in later calls). */
GPU_API s32 simk_serve_step(u32 *done, u32 max_done, u32 *n_done);
GPU_API s32 simk_serve_stats(SimkServeStats *out);
GPU_API s32 simk_serve_close(void);

/* Plug flow along the zone chain (dbd_flow_exchange, src/sim/dbd.cpp:465-483):
 * zone i of the block relaxes toward zone i - 1's old state and zone 0
 * toward feed[species], y_i += a (up - y_i) with the product rounded on its
 * own, for zones [0, zones) of every species row (row stride padded,
 * zones <= 8192). a = dt * zones / residence must be in [0, 1]. Enqueued on
 * the module stream. */
GPU_API s32 simk_plugflow_f64(f64 *y, u64 padded, u32 species, u32 zones, f64 a, const f64 *feed);

/* Measurement aid: device-to-device copy of `bytes` (16-byte multiple and
 * alignment) by a plain streaming kernel, timed with stream events. HBM
 * byte counters of this kernel calibrate the solver kernels' counters. */
GPU_API s32 simk_copy_device(void *dst, const void *src, u64 bytes, f32 *ms);
/* The same copy with one 8-byte access per lane, the access width of the
 * solver and advection kernels: counter calibration for them. */
GPU_API s32 simk_copy_device_b64(void *dst, const void *src, u64 bytes, f32 *ms);

#ifdef __cplusplus
}
#endif

#endif /* SIMK_GPU_API_H */
