// The DBD reactor workload on the host (include/simk_dbd.h): SimStack's
// humid-air mechanism, its feed gas and discharge source, and the per-step
// preparation of a zone block for the device advance.
//
// The discharge source is data-driven: a table of primary electron-impact
// yields, then one table per metastable (N2(A), N(2D), O(1D)) of quenching
// channels and of the stoichiometric updates their shares make. The
// evaluator applies the updates in table order with a fixed grouping of
// every sum and product, so the source increments handed to the solver are
// the reference's bits (src/sim/kinetics/air_mech.cpp:185-288, checked
// against the reference's own build). Zone preparation follows
// src/sim/dbd.cpp:84-158: feed gas, rate constants from T and the third-body
// density, and the absolute pulse clock of each zone.
#include <math.h>
#include <string.h>

#include "air_source.h"
#include "simk_dbd.h"

namespace {

constexpr f64 kBoltzmann = 1.380649e-23;     // J/K (air_mech.cpp:278)

// One reaction of the mechanism: equation and rate law (kin_arr / kin_m /
// kin_troe / kin_eq_reverse, kin_build.cpp:42-53); forward: the reaction index
// an equilibrium reverse refers to.
struct AirReaction {
    const char *equation;
    u32 kind;
    f64 A, n, E, A_inf, n_inf, fc;
    u32 forward;
    f64 keq_A, keq_B;
};

constexpr u32 kArr = SimkRate_Arrhenius, kM = SimkRate_ThirdBody, kTroe = SimkRate_Troe, kRev = SimkRate_EqReverse;

// The 44 reactions and their JPL 19-5 / IUPAC rate data, in the reference's
// order (air_mech.cpp:48-96; the order fixes the reaction indices).
const AirReaction kAirReactions[] = {
    // Ox
    {"O + O2 + M -> O3 + M", kM, 6.0e-34, -2.4, 0, 0, 0, 0, 0, 0, 0},
    {"O + O3 -> O2 + O2", kArr, 8.0e-12, 0.0, 2060.0, 0, 0, 0, 0, 0, 0},
    {"O + O + M -> O2 + M", kM, 5.2e-35, 0.0, -900.0, 0, 0, 0, 0, 0, 0},
    // NOx
    {"O + NO + M -> NO2 + M", kTroe, 9.0e-32, -1.5, 0, 3.0e-11, 0.0, 0.6, 0, 0, 0},
    {"O + NO2 -> NO + O2", kArr, 5.1e-12, 0.0, -210.0, 0, 0, 0, 0, 0, 0},
    {"O + NO2 + M -> NO3 + M", kTroe, 2.5e-31, -1.8, 0, 2.2e-11, -0.7, 0.6, 0, 0, 0},
    {"O3 + NO -> NO2 + O2", kArr, 3.0e-12, 0.0, 1500.0, 0, 0, 0, 0, 0, 0},
    {"O3 + NO2 -> NO3 + O2", kArr, 1.2e-13, 0.0, 2450.0, 0, 0, 0, 0, 0, 0},
    {"NO3 + NO -> NO2 + NO2", kArr, 1.5e-11, 0.0, -170.0, 0, 0, 0, 0, 0, 0},
    {"NO2 + NO3 + M -> N2O5 + M", kTroe, 2.0e-30, -4.4, 0, 1.4e-12, -0.7, 0.6, 0, 0, 0},
    {"N2O5 + M -> NO2 + NO3 + M", kRev, 0, 0, 0, 0, 0, 0, 9, 2.7e-27, 11000.0},
    {"NO + NO + O2 -> NO2 + NO2", kArr, 3.3e-39, 0.0, -530.0, 0, 0, 0, 0, 0, 0},
    {"N2O5 + H2O -> HNO3 + HNO3", kArr, 2.5e-22, 0.0, 0.0, 0, 0, 0, 0, 0, 0},
    // N atoms
    {"N + NO -> N2 + O", kArr, 2.1e-11, 0.0, -100.0, 0, 0, 0, 0, 0, 0},
    {"N + O2 -> NO + O", kArr, 1.5e-11, 0.0, 3600.0, 0, 0, 0, 0, 0, 0},
    // N + NO2: total rate split by the product branching of the reference
    {"N + NO2 -> N2O + O", kArr, 0.434 * 5.8e-12, 0.0, -220.0, 0, 0, 0, 0, 0, 0},
    {"N + NO2 -> NO + NO", kArr, 0.333 * 5.8e-12, 0.0, -220.0, 0, 0, 0, 0, 0, 0},
    {"N + NO2 -> N2 + O + O", kArr, 0.132 * 5.8e-12, 0.0, -220.0, 0, 0, 0, 0, 0, 0},
    {"N + NO2 -> N2 + O2", kArr, 0.101 * 5.8e-12, 0.0, -220.0, 0, 0, 0, 0, 0, 0},
    {"N + O + M -> NO + M", kM, 5.5e-33, 0.0, -155.0, 0, 0, 0, 0, 0, 0},
    {"N + N + M -> N2 + M", kM, 8.3e-34, 0.0, -500.0, 0, 0, 0, 0, 0, 0},
    {"N + OH -> NO + H", kArr, 4.7e-11, 0.0, 0.0, 0, 0, 0, 0, 0, 0},
    // HOx
    {"OH + OH + M -> H2O2 + M", kTroe, 6.9e-31, -1.0, 0, 2.6e-11, 0.0, 0.6, 0, 0, 0},
    {"OH + OH -> H2O + O", kArr, 1.8e-12, 0.0, 0.0, 0, 0, 0, 0, 0, 0},
    {"OH + O -> H + O2", kArr, 1.8e-11, 0.0, -180.0, 0, 0, 0, 0, 0, 0},
    {"OH + O3 -> HO2 + O2", kArr, 1.7e-12, 0.0, 940.0, 0, 0, 0, 0, 0, 0},
    {"OH + HO2 -> H2O + O2", kArr, 4.8e-11, 0.0, -250.0, 0, 0, 0, 0, 0, 0},
    {"OH + H2O2 -> HO2 + H2O", kArr, 1.8e-12, 0.0, 0.0, 0, 0, 0, 0, 0, 0},
    {"HO2 + HO2 -> H2O2 + O2", kArr, 3.0e-13, 0.0, -460.0, 0, 0, 0, 0, 0, 0},
    {"HO2 + O -> OH + O2", kArr, 3.0e-11, 0.0, -200.0, 0, 0, 0, 0, 0, 0},
    {"HO2 + O3 -> OH + O2 + O2", kArr, 1.0e-14, 0.0, 490.0, 0, 0, 0, 0, 0, 0},
    {"H + O2 + M -> HO2 + M", kTroe, 4.4e-32, -1.3, 0, 7.5e-11, 0.2, 0.6, 0, 0, 0},
    {"H + O3 -> OH + O2", kArr, 1.4e-10, 0.0, 470.0, 0, 0, 0, 0, 0, 0},
    {"H + HO2 -> OH + OH", kArr, 7.2e-11, 0.0, 0.0, 0, 0, 0, 0, 0, 0},
    {"O + H2O2 -> OH + HO2", kArr, 1.4e-12, 0.0, 2000.0, 0, 0, 0, 0, 0, 0},
    // HOx-NOx coupling, nitrous and nitric acid
    {"OH + NO + M -> HNO2 + M", kTroe, 7.0e-31, -2.6, 0, 3.6e-11, -0.1, 0.6, 0, 0, 0},
    {"OH + NO2 + M -> HNO3 + M", kTroe, 1.8e-30, -3.0, 0, 2.8e-11, 0.0, 0.6, 0, 0, 0},
    {"OH + HNO2 -> NO2 + H2O", kArr, 1.8e-11, 0.0, 390.0, 0, 0, 0, 0, 0, 0},
    {"OH + HNO3 -> NO3 + H2O", kArr, 1.5e-13, 0.0, 0.0, 0, 0, 0, 0, 0, 0},
    {"HO2 + NO -> OH + NO2", kArr, 3.3e-12, 0.0, -270.0, 0, 0, 0, 0, 0, 0},
    {"H + NO2 -> OH + NO", kArr, 4.0e-10, 0.0, 340.0, 0, 0, 0, 0, 0, 0},
    {"OH + NO3 -> HO2 + NO2", kArr, 2.2e-11, 0.0, 0.0, 0, 0, 0, 0, 0, 0},
    {"HO2 + NO3 -> OH + NO2 + O2", kArr, 3.5e-12, 0.0, 0.0, 0, 0, 0, 0, 0, 0},
    // thermal ozone decomposition
    {"O3 + M -> O2 + O + M", kM, 7.2e-10, 0.0, 11400.0, 0, 0, 0, 0, 0, 0},
};

}  // namespace

extern "C" {

SIMK_HOST_API SimkModel *simk_air_model(SimkAirSpecies *sp) {
    SimkModel *m = simk_model_new();
    if (!m) return nullptr;
    const char *names[17] = {"O", "O3", "N", "NO", "NO2", "NO3", "N2O5", "N2O", "HNO2",
                             "HNO3", "OH", "HO2", "H2O2", "H", "N2", "O2", "H2O"};
    u32 idx[17];
    for (u32 i = 0; i < 17; ++i) idx[i] = simk_model_add_species(m, names[i], 0);
    for (const AirReaction &r : kAirReactions) {
        SimkRateLaw law;
        memset(&law, 0, sizeof(law));
        law.kind = r.kind;
        law.A = r.A;
        law.n = r.n;
        law.E = r.E;
        law.A_inf = r.A_inf;
        law.n_inf = r.n_inf;
        law.fc = r.fc;
        law.ref = r.forward;
        law.keq_A = r.keq_A;
        law.keq_B = r.keq_B;
        simk_model_add_reaction(m, r.equation, &law);
    }
    simk_model_finalize(m);
    if (sp) memcpy(sp, idx, sizeof(idx));
    return m;
}

SIMK_HOST_API void simk_air_source_defaults(SimkAirSource *s) {
    memset(s, 0, sizeof(*s));
    s->g_O2_diss = 1.5;
    s->f_O1D = 0.5;
    s->g_N2A = 1.5;
    s->g_N2_diss = 0.4;
    s->g_OH_per_100eV = 0.3;
    s->boltzmann = 0;
    s->reduced_field_Td = 180.0;
    s->filament_J_cm3 = 0.0;
}

// Feed gas: the ideal-gas number density at T and p (in cm^-3), split into
// water at x_H2O and dry air of 79 % N2 (argon counted as N2) and 21 % O2.
SIMK_HOST_API f64 simk_air_feed(const SimkAirSpecies *sp, u32 n, f64 T_K, f64 p_Pa, f64 x_H2O, f64 *y) {
    const f64 density = p_Pa / (kBoltzmann * T_K) * 1e-6;
    memset(y, 0, sizeof(f64) * n);
    const f64 dry_fraction = 1.0 - x_H2O;
    y[sp->N2] = density * dry_fraction * 0.79;
    y[sp->O2] = density * dry_fraction * 0.21;
    y[sp->H2O] = density * x_H2O;
    return density;
}

// The species change per cm^3 caused by depositing energy_J_cm3 into gas of
// composition y at T: the primary electron-impact yields, then the quenching
// of N2(A), N(2D) and O(1D), in that order (N2(A) feeds N(2D)).
SIMK_HOST_API s32 simk_air_source_delta(const SimkAirSpecies *sp, const SimkAirSource *src, const f64 *y, u32 n,
                                        f64 energy_J_cm3, f64 T_K, f64 *delta) {
    if (src->boltzmann) {   // the tabulated Boltzmann G-values are not provided
        memset(delta, 0, sizeof(f64) * n);
        return 1;
    }
    u32 species_of[17];
    memcpy(species_of, sp, sizeof(species_of));
    f64 rates[simk_air::kQuenchRates];
    simk_air::air_quench_rates(T_K, rates);
    const f64 g[4] = {src->g_O2_diss, src->g_N2A, src->g_N2_diss, src->g_OH_per_100eV};
    simk_air::air_source_delta_rates(species_of, g, src->f_O1D, y, n, energy_J_cm3, rates, delta);
    return 0;
}

// ---- zone blocks ------------------------------------------------------------
// The row offsets of a zone block, by name.
struct Rows {
    u64 p;
    f64 *y, *h, *steps, *rejects, *pulses, *failed, *k, *src, *pulse, *period, *first, *count;
    explicit Rows(const SimkDbdZones *z) {
        u64 off[14];
        simk_zones_layout(z->n_species, z->n_reactions, z->padded, off);
        p = z->padded;
        f64 *b = z->block;
        y = b + off[0];
        h = b + off[1];
        steps = b + off[2];
        rejects = b + off[3];
        pulses = b + off[4];
        failed = b + off[5];
        k = b + off[6];
        src = b + off[7];
        pulse = b + off[8];
        period = b + off[9];
        first = b + off[10];
        count = b + off[11];
    }
};

SIMK_HOST_API void simk_dbd_fill_padding(SimkDbdZones *z) {
    f64 *per_zone[5] = {z->T_K, z->p_Pa, z->x_H2O, z->power_W_cm3, z->freq_Hz};
    for (u32 i = z->count; i < z->padded; ++i)
        for (f64 *a : per_zone) a[i] = a[0];
}

SIMK_HOST_API void simk_dbd_reset(SimkDbdZones *z, const SimkAirSpecies *sp) {
    const Rows r(z);
    f64 gas[64];
    for (u32 i = 0; i < z->padded; ++i) {
        simk_air_feed(sp, z->n_species, z->T_K[i], z->p_Pa[i], z->x_H2O[i], gas);
        for (u32 s = 0; s < z->n_species; ++s) r.y[s * r.p + i] = gas[s];
        r.h[i] = r.steps[i] = r.rejects[i] = r.pulses[i] = r.failed[i] = 0.0;
        z->pulse_index[i] = 0.0;
    }
}

SIMK_HOST_API void simk_dbd_prepare_rates(SimkDbdZones *z, const SimkModel *model) {
    const Rows r(z);
    f64 k[256];
    for (u32 i = 0; i < z->padded; ++i) {
        const f64 third_body = z->p_Pa[i] / (kBoltzmann * z->T_K[i]) * 1e-6;   // cm^-3
        simk_model_rate_constants(model, z->T_K[i], third_body, k);
        for (u32 q = 0; q < z->n_reactions; ++q) r.k[q * r.p + i] = k[q];
    }
}

namespace {
// The discharge of one zone during a master step. Pulsed: a burst every
// `period` seconds of absolute time (bursts numbered from 1; the zone's
// pulse_index counts those already applied), each depositing power * period;
// the bursts with absolute time in (t0, t1] are handed to the solver as a
// count and the first one's local time. A filament energy density sets the
// burst rate to power / filament energy; otherwise there are two bursts per
// voltage cycle. Averaged: the same power as a steady source per second.
struct Discharge {
    bool on;
    f64 bursts_per_s;
};
Discharge discharge_of(const SimkDbdZones *z, u32 i, const SimkAirSource *src) {
    Discharge d;
    d.on = z->power_W_cm3[i] > 0.0 && z->freq_Hz[i] > 0.0;
    d.bursts_per_s = (d.on && src->filament_J_cm3 > 0.0) ? z->power_W_cm3[i] / src->filament_J_cm3
                                                          : 2.0 * z->freq_Hz[i];
    return d;
}
}  // namespace

SIMK_HOST_API s32 simk_dbd_prepare_source(SimkDbdZones *z, const SimkAirSpecies *sp, const SimkAirSource *src,
                                          u32 mode, f64 t0, f64 t1) {
    if (src->boltzmann || z->n_species > 64) return 1;
    const Rows r(z);
    const u32 n = z->n_species;
    f64 gas[64], change[64];
    for (u32 i = 0; i < z->padded; ++i) {
        for (u32 s = 0; s < n; ++s) gas[s] = r.y[s * r.p + i];
        const Discharge d = discharge_of(z, i, src);
        const bool pulsed = mode == SimkDbd_Pulsed && d.on;
        // energy of one burst, or the power itself as a rate
        const f64 energy = pulsed ? z->power_W_cm3[i] / d.bursts_per_s : (d.on ? z->power_W_cm3[i] : 0.0);
        simk_air_source_delta(sp, src, gas, n, energy, z->T_K[i], change);
        f64 *to = pulsed ? r.pulse : r.src, *zeroed = pulsed ? r.src : r.pulse;
        for (u32 s = 0; s < n; ++s) {
            to[s * r.p + i] = change[s];
            zeroed[s * r.p + i] = 0.0;
        }
        if (!pulsed) {
            r.period[i] = r.first[i] = r.count[i] = 0.0;
            continue;
        }
        const f64 period = 1.0 / d.bursts_per_s;
        const f64 first_burst = z->pulse_index[i] + 1.0;
        f64 due = 0.0;   // bursts first_burst, first_burst + 1, ... with absolute time <= t1
        for (;;) {
            if (!((first_burst + due) * period <= t1)) break;
            due += 1.0;
        }
        r.period[i] = period;
        r.first[i] = first_burst * period - t0;
        r.count[i] = due;
    }
    return 0;
}

SIMK_HOST_API void simk_dbd_commit_pulses(SimkDbdZones *z) {
    const Rows r(z);
    for (u32 i = 0; i < z->padded; ++i) z->pulse_index[i] = z->pulse_index[i] + r.count[i];
}

SIMK_HOST_API void simk_dbd_default_control(KinControl *c) {
    memset(c, 0, sizeof(*c));
    c->rtol = 1e-3;
    c->atol = 1e8;
    c->h_min = 1e-14;
    c->h_max = 0.0;
    c->h_start = 1e-9;
    c->max_steps = 20000;
    c->fixed_step = 0;
    c->clamp_negative = 1;
}

}  // extern "C"
