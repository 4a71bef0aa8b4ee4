/* simk_host.h — C ABI of libsimk_host.so, the host side of the solver:
 * mechanism construction, rate constants, the symbolic sparse LU, the SoA
 * zone-block layout and the seeded synthetic-problem generator.
 *
 * The reference keeps these in its sim module, compiled into the caller; here
 * they are a small C++ library with a C ABI so C++ hosts and Python tests
 * share one implementation. Each entry point names the reference function
 * whose behaviour it has.
 */
#ifndef SIMK_HOST_H
#define SIMK_HOST_H

#include "simk_types.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SIMK_HOST_API __attribute__((visibility("default")))

/* ---- mechanism model (src/sim/kinetics/kin_build.cpp) ------------------- */

/* Rate law kinds (kin_build.cpp:16-21). */
enum SimkRateKind { SimkRate_Arrhenius = 0, SimkRate_ThirdBody = 1, SimkRate_Troe = 2, SimkRate_EqReverse = 3 };

typedef struct SimkRateLaw { /* kin_build.cpp:23-29 */
    u32 kind;
    f64 A, n, E;
    f64 A_inf, n_inf, fc;
    u32 ref;
    f64 keq_A, keq_B;
} SimkRateLaw;

typedef struct SimkModel SimkModel;

SIMK_HOST_API SimkModel *simk_model_new(void);
SIMK_HOST_API void simk_model_free(SimkModel *model);
/* kin_add_species (kin_build.cpp:78-89): abstract = 1 skips element bookkeeping. */
SIMK_HOST_API u32 simk_model_add_species(SimkModel *model, const char *name, b32 abstract_species);
/* kin_add_reaction (kin_build.cpp:100-148): "A + B -> C + D"; M is the third body. */
SIMK_HOST_API u32 simk_model_add_reaction(SimkModel *model, const char *equation, const SimkRateLaw *law);
/* kin_mech_finalize (kin_build.cpp:152-223). */
SIMK_HOST_API void simk_model_finalize(SimkModel *model);
/* 0 when no error so far, else 1 and the first error text (KinModel::error_text). */
SIMK_HOST_API s32 simk_model_error(const SimkModel *model, char *text, u32 size);
SIMK_HOST_API u32 simk_model_species_count(const SimkModel *model);
SIMK_HOST_API u32 simk_model_reaction_count(const SimkModel *model);
/* The finalized tables in the reference KinMech byte layout at capacities caps
 * (out: simk_mech_layout(caps).size bytes). 0 on success. */
SIMK_HOST_API s32 simk_model_mech_bytes(const SimkModel *model, const SimkCaps *caps, void *out);
/* kin_eval_rate_constants (kin_build.cpp:226-244): k[r] at temperature T [K], density M. */
SIMK_HOST_API void simk_model_rate_constants(const SimkModel *model, f64 T, f64 M, f64 *k);
/* kin_count_elements (kin_build.cpp:247-251): atoms[3] = N, O, H totals. */
SIMK_HOST_API void simk_model_count_elements(const SimkModel *model, const f64 *y, f64 *atoms);

/* Builds abstract species S0..S(n-1) and reactions from index lists
 * (reactants/products: [R][3], 0xFF = none), finalizes, and writes the KinMech
 * bytes at capacities caps. 0 on success, else 1 with the error text. */
SIMK_HOST_API s32 simk_mech_build(u32 n_species, u32 n_reactions, const u8 *reactants, const u8 *products,
                                  const SimkCaps *caps, void *out, char *error, u32 error_size);

/* Smallest-fitting statistics of KinMech bytes at caps: out[0..5] = species,
 * reactions, nnz, lu_div, lu_sub, signature low 32 bits. 0 on success. */
SIMK_HOST_API s32 simk_mech_stats(const void *mech, const SimkCaps *caps, u64 *out);

/* Rosenbrock coefficients (kin_solver.h:352-395). 0 on success. */
SIMK_HOST_API s32 simk_method(u32 id, KinMethod *out);

/* FNV-1a 64 (base_types.h:120-128). */
SIMK_HOST_API u64 simk_fnv1a64(const void *data, u64 size, u64 seed);

/* ---- SoA zone block (src/sim/dbd.cpp:40-83) ----------------------------- */

/* Row offsets (in doubles) of a zone block for n species, R reactions and
 * `padded` lanes, in the reference's order: y, h, steps, rejects, pulses,
 * failed | k, src, pulse, period, pulse_first, pulse_count. out[12] = total,
 * out[13] = mutable prefix (everything the solver writes). */
SIMK_HOST_API void simk_zones_layout(u32 n_species, u32 n_reactions, u32 padded, u64 *out);

/* ---- seeded synthetic problems ------------------------------------------ */

typedef struct SimkSynthSpec {
    u64 seed;
    u32 n_species, n_reactions;
    u32 band;                    /* half-width of the species window a reaction draws from; 0 = all */
    u32 max_attempts;            /* mechanism draws before giving up */
    f64 long_range;              /* probability one species is drawn from the whole range */
    f64 density_lo, density_hi;  /* accept a draw whose post-fill nnz / n^2 is inside; 0,0 = first draw */
    f64 log10_k_lo, log10_k_hi;  /* k at 298 K, log-uniform */
    f64 EaR_hi;                  /* Ea/R uniform [0, EaR_hi] K */
    f64 T_lo, T_hi;              /* per-cell temperature, uniform */
    f64 log10_y_lo, log10_y_hi;  /* initial concentrations, log-uniform */
    f64 pH_lo, pH_hi;            /* per-cell pH, uniform; both 0 = no pH dependence */
    f64 acid_frac, base_frac;    /* fraction of reactions acid / base catalysed */
    f64 water_frac;              /* fraction of unimolecular reactions with water (55.5 M) folded into k */
    f64 oxidant_log10_spread;    /* per-cell factor 10^U[0, spread] on the first oxidant_species y0 */
    u32 oxidant_species;
    u32 pad_;
    f64 load_log10_spread;       /* per-cell factor 10^-U[0, spread] on every initial concentration (dissolved load) */
} SimkSynthSpec;

/* Draws a mass-conserving mechanism. reactants/products: [R][3] (0xFF none);
 * params: [R][4] = log10 k_ref, Ea/R, pH class (+1 acid, -1 base, 0), water (0/1).
 * Returns the number of draws used (>= 1) or 0 when none was accepted. */
SIMK_HOST_API u32 simk_synth_mechanism(const SimkSynthSpec *spec, u8 *reactants, u8 *products, f64 *params);

/* Per-cell inputs of cells [first, first + count): k [R][padded] and y [n][padded]
 * at lane offset (cell - first), plus T and pH per cell (may be null). Each
 * cell draws from its own stream, so cell c is the same in any batch. */
SIMK_HOST_API void simk_synth_cells(const SimkSynthSpec *spec, const f64 *params, u64 first, u32 count, u32 padded,
                                    f64 *k, f64 *y, f64 *T, f64 *pH);

#ifdef __cplusplus
}
#endif

#endif /* SIMK_HOST_H */
