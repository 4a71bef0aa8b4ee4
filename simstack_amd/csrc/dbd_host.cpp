// The DBD reactor workload on the host (include/simk_dbd.h): SimStack's
// humid-air mechanism, feed gas, discharge source and the per-step zone-block
// preparation, restated from src/sim/kinetics/air_mech.cpp and
// src/sim/dbd.cpp. The arithmetic follows the reference expression by
// expression (host libm), so the inputs handed to the device are the
// reference's bits.
#include <math.h>
#include <string.h>

#include "simk_dbd.h"

namespace {

constexpr f64 kBoltzmann = 1.380649e-23;     // J/K (air_mech.cpp:278)
constexpr f64 kEvPerJoule = 6.241509074e18;  // air_mech.cpp:180

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

// Quasi-steady distribution of `amount` of a metastable over its channels
// (air_branch, air_mech.cpp:185-201): share_c = k_c [partner_c] / sum, no
// channel consuming more than half its partner (the excess goes to the
// dominant channel).
struct Channel {
    f64 k;
    u32 partner;
};

void branch(const Channel *ch, u32 count, const f64 *y, f64 amount, f64 *share) {
    f64 sum = 0.0;
    u32 big = 0;
    for (u32 c = 0; c < count; ++c) {
        share[c] = ch[c].k * y[ch[c].partner];
        if (share[c] > share[big]) big = c;
        sum += share[c];
    }
    if (sum <= 0.0) {
        for (u32 c = 0; c < count; ++c) share[c] = c == big ? 1.0 : 0.0;
        return;
    }
    f64 moved = 0.0;
    for (u32 c = 0; c < count; ++c) {
        share[c] /= sum;
        f64 cap = amount > 0.0 ? 0.5 * y[ch[c].partner] / amount : 1.0;
        if (c != big && share[c] > cap) {
            moved += share[c] - cap;
            share[c] = cap;
        }
    }
    share[big] += moved;
}

void zone_rows(const SimkDbdZones *z, u64 off[14]) { simk_zones_layout(z->n_species, z->n_reactions, z->padded, off); }

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

SIMK_HOST_API f64 simk_air_feed(const SimkAirSpecies *sp, u32 n, f64 T_K, f64 p_Pa, f64 x_H2O, f64 *y) {
    f64 total = p_Pa / (kBoltzmann * T_K) * 1e-6;   // m^-3 -> cm^-3
    for (u32 s = 0; s < n; ++s) y[s] = 0.0;
    f64 dry = 1.0 - x_H2O;
    y[sp->N2] = total * dry * 0.79;   // argon lumped into N2
    y[sp->O2] = total * dry * 0.21;
    y[sp->H2O] = total * x_H2O;
    return total;
}

SIMK_HOST_API s32 simk_air_source_delta(const SimkAirSpecies *sp, const SimkAirSource *src, const f64 *y, u32 n,
                                        f64 energy_J_cm3, f64 T_K, f64 *delta) {
    for (u32 s = 0; s < n; ++s) delta[s] = 0.0;
    if (src->boltzmann) return 1;
    f64 total = 0.0;
    for (u32 s = 0; s < n; ++s) total += y[s];
    if (total <= 0.0 || energy_J_cm3 <= 0.0) return 0;
    const f64 hundreds_eV = energy_J_cm3 * kEvPerJoule * 0.01;
    const f64 g_o2 = src->g_O2_diss, f_o1d = src->f_O1D, g_n2a = src->g_N2A, g_n2d = src->g_N2_diss;
    // electron-impact channels, scaled by the parents' mole fractions
    const f64 dO2 = g_o2 * (y[sp->O2] / total / 0.21) * hundreds_eV;
    const f64 dA = g_n2a * (y[sp->N2] / total / 0.78) * hundreds_eV;
    const f64 dN2 = g_n2d * (y[sp->N2] / total / 0.78) * hundreds_eV;
    const f64 dOH = src->g_OH_per_100eV * (y[sp->H2O] / total / 0.02) * hundreds_eV;
    const f64 d1D = f_o1d * dO2;
    delta[sp->O2] -= dO2;
    delta[sp->O] += 2.0 * dO2 - d1D;
    delta[sp->N2] -= dA + dN2;
    delta[sp->N] += dN2;
    delta[sp->OH] += dOH;
    delta[sp->H] += dOH;
    delta[sp->H2O] -= dOH;
    const f64 t3 = T_K / 300.0;
    const f64 k_o2 = pow(t3, 0.55);
    f64 f[10];
    // N2(A): every channel returns the N2 (Kossyi et al. 1992)
    const Channel a_ch[] = {
        {2.5e-12 * k_o2, sp->O2},   // -> N2 + O + O
        {1.3e-12 * k_o2, sp->O2},   // -> N2 + O2
        {7.8e-14 * k_o2, sp->O2},   // -> N2O + O
        {7.0e-12, sp->O},           // -> NO + N(2D)
        {7.0e-11, sp->NO},          // -> N2 + NO
        {1.0e-12, sp->NO2},         // -> N2 + NO + O
        {8.0e-11, sp->N2O},         // -> N2 + N + NO
        {5.0e-14, sp->H2O},         // -> N2 + H + OH
        {3.0e-16, sp->N2},          // -> N2 + N2
    };
    branch(a_ch, 9, y, dA, f);
    delta[sp->N2] += dA * (f[0] + f[1] + f[4] + f[5] + f[6] + f[7] + f[8]);
    delta[sp->O2] -= dA * (f[0] + f[2]);
    delta[sp->O] += dA * (2.0 * f[0] + f[2] + f[5]);
    delta[sp->N2O] += dA * (f[2] - f[6]);
    delta[sp->O] -= dA * f[3];
    delta[sp->NO] += dA * (f[3] + f[5] + f[6]);
    const f64 n2d = dN2 + dA * f[3];
    delta[sp->NO2] -= dA * f[5];
    delta[sp->N] += dA * f[6];
    delta[sp->H2O] -= dA * f[7];
    delta[sp->H] += dA * f[7];
    delta[sp->OH] += dA * f[7];
    // N(2D) (Herron 1999)
    const Channel n_ch[] = {
        {5.2e-12, sp->O2},    // -> NO + O
        {1.7e-14, sp->N2},    // -> N + N2
        {6.9e-13, sp->O},     // -> N + O
        {6.0e-11, sp->NO},    // -> N2 + O
        {2.2e-12, sp->N2O},   // -> NO + N2
    };
    branch(n_ch, 5, y, n2d, f);
    delta[sp->O2] -= n2d * f[0];
    delta[sp->NO] += n2d * (f[0] - f[3] + f[4]);
    delta[sp->O] += n2d * (f[0] + f[3]);
    delta[sp->N] += n2d * (f[1] + f[2]);
    delta[sp->N2] += n2d * (f[3] + f[4]);
    delta[sp->N2O] -= n2d * f[4];
    // O(1D) (JPL 19-5)
    const Channel o_ch[] = {
        {2.15e-11 * exp(110.0 / T_K), sp->N2},    // -> O + N2
        {3.3e-11 * exp(55.0 / T_K), sp->O2},      // -> O + O2
        {1.63e-10 * exp(60.0 / T_K), sp->H2O},    // -> OH + OH
        {7.26e-11 * exp(20.0 / T_K), sp->N2O},    // -> NO + NO
        {4.63e-11 * exp(20.0 / T_K), sp->N2O},    // -> N2 + O2
        {2.4e-10, sp->O3},                        // -> O2 + O2 or O2 + O + O
    };
    branch(o_ch, 6, y, d1D, f);
    delta[sp->O] += d1D * (f[0] + f[1]);
    delta[sp->O3] -= d1D * f[5];
    delta[sp->O2] += d1D * f[5] * 1.5;
    delta[sp->O] += d1D * f[5];
    delta[sp->H2O] -= d1D * f[2];
    delta[sp->OH] += d1D * f[2] * 2.0;
    delta[sp->N2O] -= d1D * (f[3] + f[4]);
    delta[sp->NO] += d1D * f[3] * 2.0;
    delta[sp->N2] += d1D * f[4];
    delta[sp->O2] += d1D * f[4];
    return 0;
}

SIMK_HOST_API void simk_dbd_fill_padding(SimkDbdZones *z) {
    f64 *arrays[5] = {z->T_K, z->p_Pa, z->x_H2O, z->power_W_cm3, z->freq_Hz};
    for (f64 *a : arrays)
        for (u32 i = z->count; i < z->padded; ++i) a[i] = a[0];
}

SIMK_HOST_API void simk_dbd_reset(SimkDbdZones *z, const SimkAirSpecies *sp) {
    u64 off[14];
    zone_rows(z, off);
    const u64 p = z->padded;
    f64 *b = z->block;
    f64 feed[64];
    for (u32 i = 0; i < z->padded; ++i) {
        simk_air_feed(sp, z->n_species, z->T_K[i], z->p_Pa[i], z->x_H2O[i], feed);
        for (u32 s = 0; s < z->n_species; ++s) b[off[0] + s * p + i] = feed[s];
        b[off[1] + i] = 0.0;   // h
        z->pulse_index[i] = 0.0;
        b[off[2] + i] = b[off[3] + i] = b[off[4] + i] = b[off[5] + i] = 0.0;   // steps, rejects, pulses, failed
    }
}

SIMK_HOST_API void simk_dbd_prepare_rates(SimkDbdZones *z, const SimkModel *model) {
    u64 off[14];
    zone_rows(z, off);
    const u64 p = z->padded;
    f64 k[256];
    for (u32 i = 0; i < z->padded; ++i) {
        const f64 M = z->p_Pa[i] / (kBoltzmann * z->T_K[i]) * 1e-6;
        simk_model_rate_constants(model, z->T_K[i], M, k);
        for (u32 r = 0; r < z->n_reactions; ++r) z->block[off[6] + r * p + i] = k[r];
    }
}

SIMK_HOST_API s32 simk_dbd_prepare_source(SimkDbdZones *z, const SimkAirSpecies *sp, const SimkAirSource *src,
                                          u32 mode, f64 t0, f64 t1) {
    if (src->boltzmann || z->n_species > 64) return 1;
    u64 off[14];
    zone_rows(z, off);
    const u64 p = z->padded;
    const u32 n = z->n_species;
    f64 *b = z->block;
    f64 y[64], delta[64];
    for (u32 i = 0; i < z->padded; ++i) {
        for (u32 s = 0; s < n; ++s) y[s] = b[off[0] + s * p + i];
        const f64 freq = z->freq_Hz[i];
        const bool on = z->power_W_cm3[i] > 0.0 && freq > 0.0;
        // one micro-discharge burst per half-cycle, or a filament of
        // filament_J_cm3 at rate power / filament_J_cm3
        const f64 pulses_per_s =
            src->filament_J_cm3 > 0.0 && on ? z->power_W_cm3[i] / src->filament_J_cm3 : 2.0 * freq;
        if (mode == SimkDbd_Pulsed && on) {
            simk_air_source_delta(sp, src, y, n, z->power_W_cm3[i] / pulses_per_s, z->T_K[i], delta);
            for (u32 s = 0; s < n; ++s) {
                b[off[8] + s * p + i] = delta[s];   // pulse
                b[off[7] + s * p + i] = 0.0;        // src
            }
            const f64 period = 1.0 / pulses_per_s;
            b[off[9] + i] = period;
            f64 next = z->pulse_index[i] + 1.0, count = 0.0;
            while ((next + count) * period <= t1) count += 1.0;
            b[off[10] + i] = next * period - t0;   // pulse_first, local time
            b[off[11] + i] = count;                // pulse_count
        } else {
            simk_air_source_delta(sp, src, y, n, on ? z->power_W_cm3[i] : 0.0, z->T_K[i], delta);   // per second
            for (u32 s = 0; s < n; ++s) {
                b[off[7] + s * p + i] = delta[s];
                b[off[8] + s * p + i] = 0.0;
            }
            b[off[9] + i] = 0.0;
            b[off[10] + i] = 0.0;
            b[off[11] + i] = 0.0;
        }
    }
    return 0;
}

SIMK_HOST_API void simk_dbd_commit_pulses(SimkDbdZones *z) {
    u64 off[14];
    zone_rows(z, off);
    for (u32 i = 0; i < z->padded; ++i) z->pulse_index[i] += z->block[off[11] + i];
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
