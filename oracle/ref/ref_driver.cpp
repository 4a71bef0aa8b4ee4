// Test-only driver that compiles the reference's own kinetics sources (read in
// place under the reference tree; nothing is copied) into oracle/_ref/libref.so
// so the oracle's C restatement and the golden fixtures can be checked against
// the real thing. This file is part of the checker, never of the product.
//
// Compiled by oracle/ref/Makefile with -I<reference>/src. It uses, unmodified:
//   src/sim/kinetics/kin_build.cpp   kin_add_species / kin_add_reaction / kin_mech_finalize
//   src/sim/kinetics/kin_solver.h    kin_fun / kin_jac / kin_factor / kin_solve / kin_integrate / kin_method
//   src/sim/dbd_lanes.h              dbd_advance_group<V> over the [row][padded] SoA block
//   src/base/base_simd.h             f64x4 (AVX2) lane type
//   src/sim/kinetics/air_mech.cpp    the humid-air mechanism, its feed and discharge source
// and re-states only the thread chunking of dbd_advance_threads
// (src/sim/dbd.cpp:228-260) over pthreads instead of the platform work queue.
//
// Caps are raised the way tools/ref/n2vib/n2vib_ref.cpp:17-21 raises them, so
// the 48- and 60-species synthetic mechanisms fit.

#define KIN_MAX_SPECIES   64
#define KIN_MAX_REACTIONS 256
#define KIN_MAX_NNZ       4096
#define KIN_MAX_LU_DIV    4096
#define KIN_MAX_LU_SUB    32768
#define SIM_SLOW 0

#include <immintrin.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include "base/base_types.h"
#include "base/base_simd.h"
#include "sim/kinetics/kin_build.cpp"
#include "sim/dbd_lanes.h"
#include "sim/kinetics/air_mech.cpp"

#define REF_API extern "C" __attribute__((visibility("default")))

REF_API void ref_caps(u32 *out) {
    out[0] = KIN_MAX_SPECIES; out[1] = KIN_MAX_REACTIONS; out[2] = KIN_MAX_NNZ;
    out[3] = KIN_MAX_LU_DIV; out[4] = KIN_MAX_LU_SUB;
}

REF_API u64 ref_sizeof(u32 what) {
    switch (what) {
        case 0: return sizeof(KinMech);
        case 1: return sizeof(KinMethod);
        case 2: return sizeof(KinControl);
        case 3: return sizeof(KinReaction);
        default: return 0;
    }
}

// Builds a mechanism of abstract species "S0".."S<n-1>" from reactant/product
// index lists ([r][3], 0xFF = none) through the reference's own parser and
// symbolic LU. Returns 0 and fills *out (sizeof(KinMech) bytes), or 1 with the
// builder's error text.
REF_API s32 ref_build(u32 n_species, u32 n_reactions, const u8 *reactants, const u8 *products, void *out,
                      char *error, u32 error_size) {
    static KinModel model;
    memset(&model, 0, sizeof(model));
    char name[16];
    for (u32 s = 0; s < n_species; ++s) {
        snprintf(name, sizeof(name), "S%u", s);
        kin_add_species(&model, name, 1);
    }
    for (u32 r = 0; r < n_reactions; ++r) {
        char eq[64];
        u32 len = 0;
        for (u32 p = 0; p < 3; ++p) {
            u8 s = reactants[r * 3 + p];
            if (s == 0xFF) continue;
            len += (u32)snprintf(eq + len, sizeof(eq) - len, "%sS%u", len ? " + " : "", s);
        }
        len += (u32)snprintf(eq + len, sizeof(eq) - len, " ->");
        u32 np = 0;
        for (u32 p = 0; p < 3; ++p) {
            u8 s = products[r * 3 + p];
            if (s == 0xFF) continue;
            len += (u32)snprintf(eq + len, sizeof(eq) - len, "%s S%u", np ? " +" : "", s);
            ++np;
        }
        kin_add_reaction(&model, eq, kin_arr(1.0));
    }
    kin_mech_finalize(&model);
    if (model.error) {
        snprintf(error, error_size, "%s", model.error_text);
        return 1;
    }
    memcpy(out, &model.mech, sizeof(KinMech));
    return 0;
}

REF_API void ref_method(u32 id, void *out) {
    KinMethod m = kin_method((KinMethodId)id);
    memcpy(out, &m, sizeof(KinMethod));
}

REF_API void ref_fun(const void *mech, const f64 *k, const f64 *src, const f64 *y, f64 *f) {
    kin_fun<f64>((const KinMech *)mech, k, src, y, f);
}
REF_API void ref_jac(const void *mech, const f64 *k, const f64 *y, f64 *jac) {
    kin_jac<f64>((const KinMech *)mech, k, y, jac);
}
REF_API void ref_neg_jac(const void *mech, const f64 *k, const f64 *y, f64 *a) {
    kin_neg_jac<f64>((const KinMech *)mech, k, y, a);
}
REF_API void ref_factor(const void *mech, f64 ghinv, f64 *a) {
    kin_factor<f64>((const KinMech *)mech, ghinv, a);
}
REF_API void ref_solve(const void *mech, const f64 *a, f64 *x) {
    kin_solve<f64>((const KinMech *)mech, a, x);
}
REF_API f64 ref_step_factor(f64 err, u32 order) { return kin_step_factor<f64>(err, order); }
REF_API f64 ref_cbrt_clamped(f64 x) { return lane_cbrt_clamped<f64>(x); }

// One scalar lane: kin_integrate<f64> from t to t_end. stats = {steps, rejects, pulses, failed}.
REF_API void ref_integrate(const void *mech, const void *method, const void *ctl, const f64 *k, const f64 *src, f64 *y,
                           f64 *t, f64 t_end, f64 *h, f64 *stats) {
    KinLaneStats<f64> st;
    st.steps = stats[0]; st.rejects = stats[1]; st.pulses = stats[2]; st.failed = stats[3];
    kin_integrate<f64>((const KinMech *)mech, (const KinMethod *)method, (const KinControl *)ctl, k, src, y, *t, t_end,
                       *h, st);
    stats[0] = st.steps; stats[1] = st.rejects; stats[2] = st.pulses; stats[3] = st.failed;
}

// The SoA block of one advance call, as in DbdLaneArrays (src/sim/dbd_lanes.h:13-19).
struct RefLaunch {
    f64 *y, *h, *steps, *rejects, *pulses, *failed;
    const f64 *k, *src, *pulse, *period, *pulse_first, *pulse_count;
    u32 padded;
    u32 lane_width;   // 1: scalar f64 lanes, 4: f64x4 (AVX2); f32 delta lanes: 32 scalar, 8 f32x8 (AVX2)
    f64 duration;
};

struct RefWork {
    const KinMech *m;
    const KinMethod *method;
    const KinControl *ctl;
    const RefLaunch *L;
    u32 first, end;
};

static void ref_run(const RefWork *w) {
    DbdLaneArrays a;
    const RefLaunch *L = w->L;
    a.y = L->y; a.h = L->h; a.steps = L->steps; a.rejects = L->rejects; a.pulses = L->pulses; a.failed = L->failed;
    a.k = L->k; a.src = L->src; a.pulse = L->pulse; a.period = L->period;
    a.pulse_first = L->pulse_first; a.pulse_count = L->pulse_count;
    a.padded = L->padded;
    if (L->lane_width == 4) {
        for (u32 i0 = w->first; i0 < w->end; i0 += 4)
            dbd_advance_group<f64x4>(w->m, w->method, w->ctl, a, i0, L->duration);
    } else if (L->lane_width == 8) {
        for (u32 i0 = w->first; i0 < w->end; i0 += 8)
            dbd_advance_group<f32x8, KinOpsDefault, true>(w->m, w->method, w->ctl, a, i0, L->duration);
    } else if (L->lane_width == 32) {
        for (u32 i0 = w->first; i0 < w->end; ++i0)
            dbd_advance_group<f32, KinOpsDefault, true>(w->m, w->method, w->ctl, a, i0, L->duration);
    } else {
        for (u32 i0 = w->first; i0 < w->end; ++i0) dbd_advance_group<f64>(w->m, w->method, w->ctl, a, i0, L->duration);
    }
}

struct RefPool {
    RefWork *items;
    u32 count;
    u32 next;   // atomically incremented
};

static void *ref_worker(void *arg) {
    RefPool *pool = (RefPool *)arg;
    for (;;) {
        u32 i = __atomic_fetch_add(&pool->next, 1u, __ATOMIC_RELAXED);
        if (i >= pool->count) break;
        ref_run(pool->items + i);
    }
    return 0;
}

// Advances lanes [0, lanes) (lanes a multiple of lane_width) over [0, duration].
// threads > 1 splits the lane groups into min(groups, 4 * threads) chunks handed
// to a pool, as dbd_advance_threads does; results do not depend on threads.
REF_API s32 ref_advance(const void *mech, const void *method, const void *ctl, const RefLaunch *L, u32 lanes,
                        u32 threads) {
    const u32 w = L->lane_width == 4 ? 4 : (L->lane_width == 8 ? 8 : 1);
    if (lanes % w) return 1;
    u32 groups = lanes / w;
    if (groups == 0) return 0;
    if (threads <= 1 || groups <= 1) {
        RefWork all = { (const KinMech *)mech, (const KinMethod *)method, (const KinControl *)ctl, L, 0, lanes };
        ref_run(&all);
        return 0;
    }
    u32 chunks = Minimum(groups, 4 * threads);
    u32 per = (groups + chunks - 1) / chunks;
    RefWork *items = (RefWork *)calloc(chunks, sizeof(RefWork));
    u32 used = 0;
    for (u32 g = 0; g < groups; g += per) {
        RefWork *it = items + used++;
        it->m = (const KinMech *)mech; it->method = (const KinMethod *)method; it->ctl = (const KinControl *)ctl;
        it->L = L;
        it->first = g * w;
        it->end = Minimum(groups, g + per) * w;
    }
    RefPool pool = { items, used, 0 };
    pthread_t *tids = (pthread_t *)calloc(threads, sizeof(pthread_t));
    for (u32 t = 0; t < threads; ++t) pthread_create(&tids[t], 0, ref_worker, &pool);
    for (u32 t = 0; t < threads; ++t) pthread_join(tids[t], 0);
    free(tids);
    free(items);
    return 0;
}

// ---- the humid-air DBD mechanism (src/sim/kinetics/air_mech.cpp) ------------
static KinModel g_air;
static AirSpecies g_air_sp;
static b32 g_air_built;

static void ref_air_ensure() {
    if (g_air_built) return;
    air_mech_build(&g_air, &g_air_sp);
    g_air_built = 1;
}

// KinMech bytes (sizeof(KinMech)) and the 17 species indices in AirSpecies
// order; returns 1 when the builder reported an error.
REF_API s32 ref_air_build(void *mech_out, u32 *species_out) {
    ref_air_ensure();
    memcpy(mech_out, &g_air.mech, sizeof(KinMech));
    memcpy(species_out, &g_air_sp, sizeof(AirSpecies));
    return g_air.error ? 1 : 0;
}

REF_API void ref_air_rate_constants(f64 T, f64 M, f64 *k) {
    ref_air_ensure();
    kin_eval_rate_constants(&g_air, T, M, k);
}

REF_API f64 ref_air_feed(f64 T, f64 p, f64 x_H2O, f64 *y) {
    ref_air_ensure();
    return air_feed(&g_air_sp, g_air.mech.species_count, T, p, x_H2O, y);
}

// air_source_delta with air_source_defaults(), optionally the table source.
REF_API void ref_air_source_delta(const f64 *y, f64 energy_J_cm3, f64 T, u32 boltzmann, f64 field_Td, f64 *delta) {
    ref_air_ensure();
    AirSourceModel src = air_source_defaults();
    src.boltzmann = boltzmann;
    if (boltzmann) src.reduced_field_Td = field_Td;
    air_source_delta(&g_air_sp, &src, y, g_air.mech.species_count, energy_J_cm3, delta, T);
}

REF_API void ref_air_count_elements(const f64 *y, f64 *atoms) {
    ref_air_ensure();
    kin_count_elements(&g_air, y, atoms);
}
