/* simk_dbd.h — the DBD reactor workload: SimStack's humid-air mechanism, its
 * feed gas and discharge source, and the per-step preparation of a zone block
 * (rate constants, pulse schedule and pulse increments) that the device
 * advance consumes. Host side, in libsimk_host.so; every function names the
 * reference function whose behaviour it has. Units are the reference's:
 * number densities in cm^-3, energies in J/cm^3, T in K, p in Pa. */
#ifndef SIMK_DBD_H
#define SIMK_DBD_H

#include "simk_host.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Species indices of the mechanism (AirSpecies, src/sim/kinetics/air_mech.cpp:23-26). */
typedef struct SimkAirSpecies {
    u32 O, O3, N, NO, NO2, NO3, N2O5, N2O, HNO2, HNO3, OH, HO2, H2O2, H, N2, O2, H2O;
} SimkAirSpecies;

/* Discharge source parameters (AirSourceModel, air_mech.cpp:133-148). */
typedef struct SimkAirSource {
    f64 g_O2_diss;       /* O2 dissociations per 100 eV at x_O2 = 0.21 */
    f64 f_O1D;           /* share of them in the O + O(1D) channel */
    f64 g_N2A;           /* N2(A) per 100 eV at x_N2 = 0.78 */
    f64 g_N2_diss;       /* N2 dissociations per 100 eV at x_N2 = 0.78 */
    f64 g_OH_per_100eV;  /* H2O dissociations at x_H2O = 0.02 */
    b32 boltzmann;       /* 1: G-values from the reference's LoKI-B table (not provided here: rejected) */
    u32 pad_;
    f64 reduced_field_Td;
    f64 filament_J_cm3;  /* micro-discharge energy density; 0: one pulse per half-cycle */
} SimkAirSource;

/* air_mech_build (air_mech.cpp:30-102): the 17-species / 44-reaction humid-air
 * mechanism as a new finalized model (free with simk_model_free). */
SIMK_HOST_API SimkModel *simk_air_model(SimkAirSpecies *species);
/* air_source_defaults (air_mech.cpp:165-176). */
SIMK_HOST_API void simk_air_source_defaults(SimkAirSource *src);
/* air_feed (air_mech.cpp:280-288): humid air at T, p; returns the total density. */
SIMK_HOST_API f64 simk_air_feed(const SimkAirSpecies *sp, u32 n, f64 T_K, f64 p_Pa, f64 x_H2O, f64 *y);
/* air_source_delta (air_mech.cpp:205-276): species change per cm^3 for
 * energy_J_cm3 deposited into gas y at T. 0 on success; 1 when src asks for
 * the table source. */
SIMK_HOST_API s32 simk_air_source_delta(const SimkAirSpecies *sp, const SimkAirSource *src, const f64 *y, u32 n,
                                        f64 energy_J_cm3, f64 T_K, f64 *delta);

/* A DBD zone block: the solver block (simk_zones_layout order) plus each
 * zone's operating point and pulse clock (DbdZones, src/sim/dbd.cpp:40-83). */
typedef struct SimkDbdZones {
    u32 n_species, n_reactions, count, padded;
    f64 *block;                                 /* simk_zones_layout(n, R, padded).total doubles */
    f64 *T_K, *p_Pa, *x_H2O, *power_W_cm3, *freq_Hz;   /* [padded] */
    f64 *pulse_index;                           /* [padded]: pulses fired so far */
} SimkDbdZones;

enum SimkDbdDischarge { SimkDbd_Pulsed = 0, SimkDbd_Averaged = 1 };

/* dbd_zones_fill_padding (dbd.cpp:97-103): padding lanes copy zone 0's operating point. */
SIMK_HOST_API void simk_dbd_fill_padding(SimkDbdZones *z);
/* dbd_zones_reset (dbd.cpp:86-95): feed gas in every lane, counters and clocks cleared. */
SIMK_HOST_API void simk_dbd_reset(SimkDbdZones *z, const SimkAirSpecies *sp);
/* dbd_zones_prepare_rates (dbd.cpp:107-116): k per lane from T and M = p / (k_B T). */
SIMK_HOST_API void simk_dbd_prepare_rates(SimkDbdZones *z, const SimkModel *model);
/* dbd_zones_prepare_source (dbd.cpp:123-154): this master step's source terms
 * and pulse schedule over (t0, t1]. 0 on success. */
SIMK_HOST_API s32 simk_dbd_prepare_source(SimkDbdZones *z, const SimkAirSpecies *sp, const SimkAirSource *src,
                                          u32 mode, f64 t0, f64 t1);
/* dbd_zones_commit_pulses (dbd.cpp:156-158). */
SIMK_HOST_API void simk_dbd_commit_pulses(SimkDbdZones *z);
/* dbd_default_control (dbd.cpp:486-496). */
SIMK_HOST_API void simk_dbd_default_control(KinControl *out);

/* A DBD reactor as a request of the GPU module's serving loop
 * (SimkServeRequest.kind = SimkServe_Dbd): `cells` zones advanced per master
 * step as the reference's scene does (dbd_scene_step, dbd_scene.cpp:226-298):
 * the discharge source and pulse schedule over (t, t + dt] prepared on the
 * device from each zone's composition, the chemistry advance, the pulses
 * committed, then (residence_s > 0) plug flow along the chain of zones toward
 * the request's inflow (its feed gas), a = dt * cells / residence_s. The
 * request's clock starts at 0. */
typedef struct SimkServeDbd {
    SimkAirSpecies species;       /* the mechanism's species indices (simk_air_model) */
    u32 mode;                     /* SimkDbd_Pulsed or SimkDbd_Averaged */
    SimkAirSource source;         /* boltzmann must be 0 */
    f64 residence_s;              /* 0: independent zones */
    const f64 *T_K, *power_W_cm3, *freq_Hz, *pulse_index;   /* [cells]: operating point, pulses fired so far */
} SimkServeDbd;

/* Split of a zone block between the GPU module and the caller's CPU path, for
 * a caller that advances both at once as dbd_advance_hybrid does
 * (src/sim/dbd.cpp:296-463): lanes [0, split) through gpu_kin_advance_async
 * (GpuKinLaunch.lane_count = split), [split, padded) on its CPU path, then
 * gpu_kin_wait. Every lane gives the same bits on either side, so the split
 * changes only the wall time. The controller fits t_cpu = c (padded - n) and
 * t_gpu = o + g n to the reported timings and picks the n that ends both
 * together (multiples of 8 lanes; 0 when the device saves under 3 %),
 * dithering the split while the fit needs a second point and every 16th
 * advance after. Its own design; the reference chooses by bucketed timings
 * and local probes. */
typedef struct SimkHybrid {
    u32 padded;          /* lanes in the block */
    u32 split;           /* lanes for the device at the next advance */
    u32 best;            /* the model's optimum (split differs from it on dither steps) */
    u32 steps, cpu_samples, gpu_samples, dither, pad_;
    f64 cpu_lane_ms;     /* c: CPU milliseconds per lane */
    f64 gpu_fixed_ms;    /* o: device milliseconds of an advance independent of its lanes (launch, copies) */
    f64 gpu_lane_ms;     /* g: device milliseconds per lane */
    f64 w, sn, snn, st, snt;   /* decaying least-squares sums of the device samples */
} SimkHybrid;

/* A new controller; its first split is about half the block. */
SIMK_HOST_API void simk_hybrid_init(SimkHybrid *h, u32 padded);
/* Reports one advance: the split used, the device's time for its share
 * (gpu_kin_wait's total_ms, copies included) and the CPU path's time for the
 * rest; sets h->split for the next advance. */
SIMK_HOST_API void simk_hybrid_record(SimkHybrid *h, u32 split, f64 gpu_ms, f64 cpu_ms);

#ifdef __cplusplus
}
#endif

#endif /* SIMK_DBD_H */
