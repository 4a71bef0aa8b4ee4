/* ORACLE — test infrastructure only. A plain-C, scalar FP64 restatement of the
 * reference's batched Rosenbrock path, used by tests/, smoke() and bench.py's
 * cpu_baseline leg as the checker. Nothing in the product links or calls it.
 *
 * Every function names the reference file:line it follows. Mechanisms are the
 * reference KinMech byte layout at capacities `caps` (simk_mech_layout).
 *
 * oracle_factor_store32 / oracle_advance_store32 restate this build's FP32
 * factor-storage form (GpuKinLaunch precision 3), which has no reference
 * counterpart: the reference's FP64 advance with each finished factor row's
 * off-diagonal entries rounded to f32.
 */
#ifndef SIMK_ORACLE_H
#define SIMK_ORACLE_H

#include "../include/simk_types.h"

#define ORACLE_API __attribute__((visibility("default")))

/* The SoA block of one advance call (DbdLaneArrays, src/sim/dbd_lanes.h:13-19). */
typedef struct OracleLaunch {
    f64 *y, *h, *steps, *rejects, *pulses, *failed;
    const f64 *k, *src, *pulse, *period, *pulse_first, *pulse_count;
    u32 padded;
    u32 pad_;
    f64 duration;
} OracleLaunch;

ORACLE_API s32 oracle_finalize(u32 n, u32 R, const u8 *reactants, const u8 *products, const SimkCaps *caps, void *out);
ORACLE_API void oracle_method(u32 id, KinMethod *out);
ORACLE_API void oracle_fun(const void *mech, const SimkCaps *caps, const f64 *k, const f64 *src, const f64 *y, f64 *f);
ORACLE_API void oracle_jac(const void *mech, const SimkCaps *caps, const f64 *k, const f64 *y, f64 *jac);
ORACLE_API void oracle_neg_jac(const void *mech, const SimkCaps *caps, const f64 *k, const f64 *y, f64 *a);
ORACLE_API void oracle_factor(const void *mech, const SimkCaps *caps, f64 ghinv, f64 *a);
ORACLE_API void oracle_factor_store32(const void *mech, const SimkCaps *caps, f64 ghinv, f64 *a);
ORACLE_API void oracle_solve(const void *mech, const SimkCaps *caps, const f64 *a, f64 *x);
ORACLE_API f64 oracle_cbrt_clamped(f64 x);
ORACLE_API f64 oracle_step_factor(f64 err, u32 order);
ORACLE_API void oracle_integrate(const void *mech, const SimkCaps *caps, const KinMethod *method, const KinControl *ctl,
                                 const f64 *k, const f64 *src, f64 *y, f64 *t, f64 t_end, f64 *h, f64 *stats);
ORACLE_API s32 oracle_advance(const void *mech, const SimkCaps *caps, const KinMethod *method, const KinControl *ctl,
                              const OracleLaunch *L, const u32 *lanes, u32 lane_count, u32 threads);
ORACLE_API s32 oracle_advance_store32(const void *mech, const SimkCaps *caps, const KinMethod *method,
                                      const KinControl *ctl, const OracleLaunch *L, const u32 *lanes, u32 lane_count,
                                      u32 threads);
ORACLE_API s32 oracle_advance_delta32(const void *mech, const SimkCaps *caps, const KinMethod *method,
                                      const KinControl *ctl, const OracleLaunch *L, const u32 *lanes, u32 lane_count,
                                      u32 threads);
ORACLE_API f32 oracle_cbrt_clamped_f32(f32 x);
ORACLE_API f32 oracle_step_factor_f32(f32 err, u32 order);
ORACLE_API void oracle_plugflow_f64(f64 *y, u32 stride, u32 species, u32 zones, f64 a, const f64 *feed);
ORACLE_API void oracle_upwind(f32 *x, const f32 *c, u32 nodes, u32 fields, u32 stride, u32 substeps, f32 *outlet_sum);
ORACLE_API void oracle_upwind_f64(f64 *x, const f64 *c, u32 nodes, u32 fields, u32 stride, u32 substeps,
                                  const f64 *inflow);

#endif
