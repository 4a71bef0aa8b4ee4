// The discharge source of the DBD reactor workload, shared by the host
// library (zone preparation, simk_air_source_delta) and the GPU module (the
// serving loop prepares DBD requests' sources on the device).
//
// A data-driven evaluator: a table of primary electron-impact yields, then
// one table per metastable (N2(A), N(2D), O(1D)) of quenching channels and of
// the stoichiometric updates their shares make. The updates are applied in
// table order with a fixed grouping of every sum and product, so the
// increments are the reference's bits (src/sim/kinetics/air_mech.cpp:185-276).
// The only host-library arithmetic is the channels' rate coefficients at T
// (pow, exp): they are computed on the host (air_quench_rates) and the rest
// runs from them, identically on host and device.
#pragma once

#include <math.h>

#include "simk_types.h"

#ifdef __HIPCC__
#define SIMK_AIR_HD __host__ __device__
#else
#define SIMK_AIR_HD
#endif

namespace simk_air {

// Metastable bookkeeping.
//
// A discharge pulse leaves electronically excited species (N2(A), N(2D),
// O(1D)) that are not species of the mechanism: each is spent at once on its
// quenching channels. A channel's weight is its rate coefficient times its
// collision partner's density; the metastable's amount is split in
// proportion to the weights. A channel may use at most half of its partner's
// density; whatever a capped channel cannot take is handed to the channel of
// largest weight (the first such channel on ties). With no weight at all,
// everything goes to channel 0.
enum RateForm : u32 {
    kConstant = 0,   // k0
    kPowerT300 = 1,  // k0 (T / 300)^q
    kExpOverT = 2,   // k0 exp(q / T)
};
struct Quench {
    f64 k0;
    RateForm form;
    f64 q;
    u32 partner;   // SimkAirSpecies field index (species_of below)
};

// A channel's rate coefficient at T (host libm: pow, exp).
inline f64 quench_rate(const Quench &c, f64 T_K) {
    switch (c.form) {
    case kPowerT300: return c.k0 * pow(T_K / 300.0, c.q);
    case kExpOverT: return c.k0 * exp(c.q / T_K);
    default: return c.k0;
    }
}

// Shares of `amount` over `count` channels with rate coefficients rate[c]
// and partner species partner[c], written to share[0..count).
SIMK_AIR_HD inline void split_metastable(const f64 *rate, u32 count, const u32 *partner, const f64 *y, f64 amount,
                                         f64 *share) {
    u32 lead = 0;
    f64 weight_sum = 0.0;
    for (u32 c = 0; c < count; ++c) {
        share[c] = rate[c] * y[partner[c]];
        weight_sum += share[c];
        if (share[c] > share[lead]) lead = c;
    }
    if (weight_sum <= 0.0) {
        for (u32 c = 0; c < count; ++c) share[c] = c == lead ? 1.0 : 0.0;
        return;
    }
    f64 surplus = 0.0;
    for (u32 c = 0; c < count; ++c) {
        share[c] /= weight_sum;
        if (c == lead) continue;
        const f64 limit = amount > 0.0 ? 0.5 * y[partner[c]] / amount : 1.0;
        if (share[c] > limit) {
            surplus += share[c] - limit;
            share[c] = limit;
        }
    }
    share[lead] += surplus;
}

// One stoichiometric update of the source vector:
//   delta[species] += sign * ((scale * (c_0 v_0 + c_1 v_1 + ...)) * post)
// where the v are shares of the current stage (or, in the primary stage,
// the primary amounts), the sum runs left to right, scale is the stage's
// amount (1: none) and post a trailing factor (1: none). The updates of a
// stage are applied in table order; that order, and the grouping above,
// fix every rounding, so they are part of the data.
struct Term {
    f64 c;
    u32 v;
};
struct Update {
    u32 species;
    s32 sign;
    u32 nterms;
    Term t[7];
    f64 post;
};

SIMK_AIR_HD inline f64 update_value(const Update &u, const f64 *v, f64 scale, bool scaled) {
    f64 sum = u.t[0].c == 1.0 ? v[u.t[0].v] : u.t[0].c * v[u.t[0].v];
    for (u32 i = 1; i < u.nterms; ++i) sum = sum + (u.t[i].c == 1.0 ? v[u.t[i].v] : u.t[i].c * v[u.t[i].v]);
    f64 x = scaled ? scale * sum : sum;
    if (u.post != 1.0) x = x * u.post;
    return x;
}

SIMK_AIR_HD inline void apply(const Update *ups, u32 count, const u32 *species_of, const f64 *v, f64 scale, bool scaled,
                              f64 *delta) {
    for (u32 i = 0; i < count; ++i) {
        const f64 x = update_value(ups[i], v, scale, scaled);
        f64 &d = delta[species_of[ups[i].species]];
        d = ups[i].sign > 0 ? d + x : d - x;
    }
}

// Species by SimkAirSpecies field position.
enum AirField : u32 { O, O3, N, NO, NO2, NO3, N2O5, N2O, HNO2, HNO3, OH, HO2, H2O2, H, N2, O2, H2O };

// Primary electron-impact yields, per 100 eV deposited, scaled by the
// parent's mole fraction over the reference gas's (x / x_ref) and consumed
// from the parent. amounts: 0 O2 dissociations, 1 N2(A), 2 N2
// dissociations, 3 H2O dissociations, 4 O(1D) = f_O1D x amount 0.
struct Primary {
    u32 parent;
    f64 x_ref;
};
constexpr Primary kPrimary[4] = {{O2, 0.21}, {N2, 0.78}, {N2, 0.78}, {H2O, 0.02}};
enum { kDissO2 = 0, kN2A = 1, kDissN2 = 2, kDissH2O = 3, kO1D = 4 };
constexpr Update kPrimaryUpdates[] = {
    {O2, -1, 1, {{1.0, kDissO2}}, 1.0},                     // O2 -> O + O, O + O(1D)
    {O, +1, 2, {{2.0, kDissO2}, {-1.0, kO1D}}, 1.0},        // the ground-state O atoms
    {N2, -1, 2, {{1.0, kN2A}, {1.0, kDissN2}}, 1.0},        // N2 -> N2(A); N2 -> N + N(2D)
    {N, +1, 1, {{1.0, kDissN2}}, 1.0},
    {OH, +1, 1, {{1.0, kDissH2O}}, 1.0},                    // H2O -> OH + H
    {H, +1, 1, {{1.0, kDissH2O}}, 1.0},
    {H2O, -1, 1, {{1.0, kDissH2O}}, 1.0},
};

// N2(A) quenching (Kossyi et al. 1992). Every channel gives the N2 back.
constexpr Quench kN2AQ[9] = {
    {2.5e-12, kPowerT300, 0.55, O2},   // 0: N2 + O + O
    {1.3e-12, kPowerT300, 0.55, O2},   // 1: N2 + O2
    {7.8e-14, kPowerT300, 0.55, O2},   // 2: N2O + O
    {7.0e-12, kConstant, 0.0, O},      // 3: NO + N(2D)
    {7.0e-11, kConstant, 0.0, NO},     // 4: N2 + NO
    {1.0e-12, kConstant, 0.0, NO2},    // 5: N2 + NO + O
    {8.0e-11, kConstant, 0.0, N2O},    // 6: N2 + N + NO
    {5.0e-14, kConstant, 0.0, H2O},    // 7: N2 + H + OH
    {3.0e-16, kConstant, 0.0, N2},     // 8: N2 + N2
};
constexpr Update kN2AUpdates[] = {
    {N2, +1, 7, {{1, 0}, {1, 1}, {1, 4}, {1, 5}, {1, 6}, {1, 7}, {1, 8}}, 1.0},
    {O2, -1, 2, {{1, 0}, {1, 2}}, 1.0},
    {O, +1, 3, {{2, 0}, {1, 2}, {1, 5}}, 1.0},
    {N2O, +1, 2, {{1, 2}, {-1, 6}}, 1.0},
    {O, -1, 1, {{1, 3}}, 1.0},
    {NO, +1, 3, {{1, 3}, {1, 5}, {1, 6}}, 1.0},
    {NO2, -1, 1, {{1, 5}}, 1.0},
    {N, +1, 1, {{1, 6}}, 1.0},
    {H2O, -1, 1, {{1, 7}}, 1.0},
    {H, +1, 1, {{1, 7}}, 1.0},
    {OH, +1, 1, {{1, 7}}, 1.0},
};
// N(2D) quenching (Herron 1999). Its amount is the primary N2 dissociations
// plus N2(A)'s channel 3.
constexpr Quench kN2DQ[5] = {
    {5.2e-12, kConstant, 0.0, O2},    // 0: NO + O
    {1.7e-14, kConstant, 0.0, N2},    // 1: N + N2
    {6.9e-13, kConstant, 0.0, O},     // 2: N + O
    {6.0e-11, kConstant, 0.0, NO},    // 3: N2 + O
    {2.2e-12, kConstant, 0.0, N2O},   // 4: NO + N2
};
constexpr Update kN2DUpdates[] = {
    {O2, -1, 1, {{1, 0}}, 1.0},
    {NO, +1, 3, {{1, 0}, {-1, 3}, {1, 4}}, 1.0},
    {O, +1, 2, {{1, 0}, {1, 3}}, 1.0},
    {N, +1, 2, {{1, 1}, {1, 2}}, 1.0},
    {N2, +1, 2, {{1, 3}, {1, 4}}, 1.0},
    {N2O, -1, 1, {{1, 4}}, 1.0},
};
// O(1D) quenching (JPL 19-5). Channel 5 (O3) splits evenly between O2 + O2
// and O2 + O + O: per O(1D), -1 O3, +1.5 O2, +1 O.
constexpr Quench kO1DQ[6] = {
    {2.15e-11, kExpOverT, 110.0, N2},   // 0: O + N2
    {3.3e-11, kExpOverT, 55.0, O2},     // 1: O + O2
    {1.63e-10, kExpOverT, 60.0, H2O},   // 2: OH + OH
    {7.26e-11, kExpOverT, 20.0, N2O},   // 3: NO + NO
    {4.63e-11, kExpOverT, 20.0, N2O},   // 4: N2 + O2
    {2.4e-10, kConstant, 0.0, O3},      // 5: see above
};
constexpr Update kO1DUpdates[] = {
    {O, +1, 2, {{1, 0}, {1, 1}}, 1.0},
    {O3, -1, 1, {{1, 5}}, 1.0},
    {O2, +1, 1, {{1, 5}}, 1.5},
    {O, +1, 1, {{1, 5}}, 1.0},
    {H2O, -1, 1, {{1, 2}}, 1.0},
    {OH, +1, 1, {{1, 2}}, 2.0},
    {N2O, -1, 2, {{1, 3}, {1, 4}}, 1.0},
    {NO, +1, 1, {{1, 3}}, 2.0},
    {N2, +1, 1, {{1, 4}}, 1.0},
    {O2, +1, 1, {{1, 4}}, 1.0},
};


// Rate-coefficient slots of the three metastables' channels.
constexpr u32 kRatesN2A = 0, kRatesN2D = 9, kRatesO1D = 14, kQuenchRates = 20;

// Every channel's rate coefficient at T, in slot order.
inline void air_quench_rates(f64 T_K, f64 *rates) {
    for (u32 c = 0; c < 9; ++c) rates[kRatesN2A + c] = quench_rate(kN2AQ[c], T_K);
    for (u32 c = 0; c < 5; ++c) rates[kRatesN2D + c] = quench_rate(kN2DQ[c], T_K);
    for (u32 c = 0; c < 6; ++c) rates[kRatesO1D + c] = quench_rate(kO1DQ[c], T_K);
}

// One metastable's quenching: shares over its channels, then its updates.
template <u32 C>
SIMK_AIR_HD inline void quench_stage(const Quench (&ch)[C], const f64 *rate, const Update *ups, u32 nups,
                                     const u32 *species_of, const f64 *y, f64 amount, f64 *share, f64 *delta) {
    u32 partner[C];
    for (u32 c = 0; c < C; ++c) partner[c] = species_of[ch[c].partner];
    split_metastable(rate, C, partner, y, amount, share);
    apply(ups, nups, species_of, share, amount, true, delta);
}

// The species change per cm^3 of depositing `energy` J/cm^3 into gas of
// composition y (n species; species_of: the 17 species indices in
// SimkAirSpecies field order), from the channels' rate coefficients at the
// gas temperature (air_quench_rates) and the G-values g = {O2 dissociations,
// N2(A), N2 dissociations, H2O dissociations} per 100 eV with f_O1D the O(1D)
// share of O2 dissociations. delta must hold n values.
SIMK_AIR_HD inline void air_source_delta_rates(const u32 *species_of, const f64 *g, f64 f_O1D, const f64 *y, u32 n,
                                               f64 energy_J_cm3, const f64 *rates, f64 *delta) {
    for (u32 s = 0; s < n; ++s) delta[s] = 0.0;
    f64 density = 0.0;
    for (u32 s = 0; s < n; ++s) density += y[s];
    if (density <= 0.0 || energy_J_cm3 <= 0.0) return;
    const f64 per_100eV = energy_J_cm3 * 6.241509074e18 * 0.01;   // eV per joule (air_mech.cpp:180), per 100 eV
    f64 amount[5];
    for (u32 k = 0; k < 4; ++k)
        amount[k] = g[k] * (y[species_of[kPrimary[k].parent]] / density / kPrimary[k].x_ref) * per_100eV;
    amount[kO1D] = f_O1D * amount[kDissO2];
    apply(kPrimaryUpdates, sizeof(kPrimaryUpdates) / sizeof(Update), species_of, amount, 1.0, false, delta);
    f64 share[9];
    quench_stage(kN2AQ, rates + kRatesN2A, kN2AUpdates, sizeof(kN2AUpdates) / sizeof(Update), species_of, y,
                 amount[kN2A], share, delta);
    const f64 n2d = amount[kDissN2] + amount[kN2A] * share[3];
    quench_stage(kN2DQ, rates + kRatesN2D, kN2DUpdates, sizeof(kN2DUpdates) / sizeof(Update), species_of, y, n2d,
                 share, delta);
    quench_stage(kO1DQ, rates + kRatesO1D, kO1DUpdates, sizeof(kO1DUpdates) / sizeof(Update), species_of, y,
                 amount[kO1D], share, delta);
}

}  // namespace simk_air
