// libsimk_host.so: mechanism model, rate constants, zone layout and the seeded
// synthetic-problem generator, behind the C ABI declared in simk_host.h.
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "mech.hpp"
#include "simk_host.h"

using namespace simk;

// ---------------------------------------------------------------------------
// Model: species, reactions from equations, element balance, rate laws.
// Behaviour of kin_build.cpp:13-251 (elements N, O, H; "M" = third body).

static const char kElementLetters[3] = {'N', 'O', 'H'};

struct SimkModel {
    Mech mech;
    std::vector<SimkRateLaw> laws;
    std::vector<std::string> names;
    std::vector<std::array<u8, 3>> elements;
    std::vector<std::string> equations;
    bool error = false;
    std::string error_text;

    void fail(const char *what, const char *detail) {
        if (error) return;
        error = true;
        error_text = std::string(what) + ": " + detail;
    }
};

// Atoms of each element in a species name: a run of element letters, each
// optionally followed by a count ("HNO3", "N2O5"); anything from '(' on names
// an excited state and carries no atoms ("N2(A)", "O(1D)"). False when a
// letter is not an element.
static bool parse_formula(const char *name, u8 counts[3]) {
    counts[0] = counts[1] = counts[2] = 0;
    for (const char *c = name; *c && *c != '(';) {
        const char *hit = strchr(kElementLetters, *c);
        if (!hit || !*c) return false;
        const int element = (int)(hit - kElementLetters);
        ++c;
        u32 atoms = 0;
        for (; *c >= '0' && *c <= '9'; ++c) atoms = atoms * 10 + (u32)(*c - '0');
        counts[element] = (u8)(counts[element] + (atoms == 0 ? 1 : atoms));
    }
    return true;
}

// The species names of one side of an equation: '+'-separated, surrounding
// blanks ignored, names of at most 15 characters (a longer run is cut into
// 15-character pieces). The third body "M" is dropped: its density is part of
// the rate constant.
static std::vector<std::string> side_species(const std::string &side) {
    std::vector<std::string> out;
    std::string name;
    auto flush = [&]() {
        if (!name.empty() && name != "M") out.push_back(name);
        name.clear();
    };
    for (char ch : side) {
        if (ch == ' ' || ch == '+') {
            flush();
            continue;
        }
        if (name.size() == 15) flush();
        name += ch;
    }
    flush();
    return out;
}

extern "C" {

SIMK_HOST_API SimkModel *simk_model_new(void) { return new SimkModel(); }
SIMK_HOST_API void simk_model_free(SimkModel *model) { delete model; }

SIMK_HOST_API u32 simk_model_add_species(SimkModel *m, const char *name, b32 abstract_species) {
    u32 s = m->mech.n;
    if (s >= 255) {
        m->fail("too many species", name);
        return s;
    }
    m->mech.n++;
    m->names.emplace_back(name);
    std::array<u8, 3> el{0, 0, 0};
    if (!abstract_species && !parse_formula(name, el.data())) m->fail("species formula must use N, O, H", name);
    if (abstract_species) el = {0, 0, 0};
    m->elements.push_back(el);
    return s;
}

static int find_species(const SimkModel *m, const std::string &name) {
    for (size_t s = 0; s < m->names.size(); ++s)
        if (m->names[s] == name) return (int)s;
    return -1;
}

SIMK_HOST_API u32 simk_model_add_reaction(SimkModel *m, const char *equation, const SimkRateLaw *law) {
    const u32 r = m->mech.R++;
    m->mech.rx.emplace_back();
    Reaction &rx = m->mech.rx.back();
    m->laws.push_back(law ? *law : SimkRateLaw{});
    m->equations.emplace_back(equation);
    const SimkRateLaw &l = m->laws.back();

    // "reactants -> products"; a later "->" only separates products
    const std::string eq(equation);
    const size_t arrow = eq.find("->");
    if (arrow == std::string::npos) {
        m->fail("no '->' between reactants and products", equation);
        return r;
    }
    std::string right = eq.substr(arrow + 2);
    for (size_t at; (at = right.find("->")) != std::string::npos;) right.replace(at, 2, " ");
    const std::vector<std::string> lhs = side_species(eq.substr(0, arrow)), rhs = side_species(right);

    // stoichiometry: -1 per reactant occurrence, +1 per product occurrence;
    // atoms must balance element by element
    std::vector<int> net(m->mech.n, 0);
    int atoms[3] = {0, 0, 0};
    for (int side = 0; side < 2; ++side) {
        const int sign = side == 0 ? -1 : 1;
        for (const std::string &name : side == 0 ? lhs : rhs) {
            const int s = find_species(m, name);
            if (s < 0) {
                m->fail("unknown species in reaction", equation);
                return r;
            }
            if (side == 0) {
                if (rx.reactant_count == SIMK_MAX_REACTANTS) {
                    m->fail("more reactants than a reaction may have", equation);
                    return r;
                }
                rx.reactant[rx.reactant_count++] = (u8)s;
            }
            net[s] += sign;
            for (int e = 0; e < 3; ++e) atoms[e] += sign * m->elements[s][e];
        }
    }
    if (rx.reactant_count == 0) m->fail("reaction without reactants", equation);
    if (atoms[0] || atoms[1] || atoms[2]) m->fail("element balance violated", equation);
    if (l.kind == SimkRate_EqReverse && l.ref >= r)
        m->fail("an equilibrium reverse must come after the reaction it reverses", equation);
    // species whose count changes, in species order
    for (u32 s = 0; s < m->mech.n; ++s) {
        if (!net[s]) continue;
        if (rx.net_count == SIMK_MAX_NET) {
            m->fail("more changed species than a reaction may have", equation);
            return r;
        }
        rx.net_species[rx.net_count] = (u8)s;
        rx.net_coef[rx.net_count] = (s8)net[s];
        ++rx.net_count;
    }
    return r;
}

SIMK_HOST_API void simk_model_finalize(SimkModel *m) {
    std::string err = finalize(m->mech);
    if (!err.empty()) m->fail("mechanism", err.c_str());
}

SIMK_HOST_API s32 simk_model_error(const SimkModel *m, char *text, u32 size) {
    if (text && size) snprintf(text, size, "%s", m->error_text.c_str());
    return m->error ? 1 : 0;
}

SIMK_HOST_API u32 simk_model_species_count(const SimkModel *m) { return m->mech.n; }
SIMK_HOST_API u32 simk_model_reaction_count(const SimkModel *m) { return m->mech.R; }

SIMK_HOST_API s32 simk_model_mech_bytes(const SimkModel *m, const SimkCaps *caps, void *out) {
    if (m->error || m->mech.nnz == 0) return 1;
    SimkMechLayout L = simk_mech_layout(caps);
    memset(out, 0, L.size);
    return to_bytes(m->mech, *caps, out).empty() ? 0 : 1;
}

SIMK_HOST_API void simk_model_rate_constants(const SimkModel *m, f64 T, f64 M, f64 *k) {
    const f64 t300 = T / 300.0;
    for (u32 r = 0; r < m->mech.R; ++r) {
        const SimkRateLaw &l = m->laws[r];
        switch (l.kind) {
            case SimkRate_Arrhenius: k[r] = l.A * pow(t300, l.n) * exp(-l.E / T); break;
            case SimkRate_ThirdBody: k[r] = l.A * pow(t300, l.n) * exp(-l.E / T) * M; break;
            case SimkRate_Troe: {
                f64 k0m = l.A * pow(t300, l.n) * M;
                f64 kinf = l.A_inf * pow(t300, l.n_inf);
                f64 ratio = k0m / kinf;
                f64 lg = log10(ratio);
                k[r] = (k0m / (1.0 + ratio)) * pow(l.fc, 1.0 / (1.0 + lg * lg));
            } break;
            case SimkRate_EqReverse: k[r] = k[l.ref] / (l.keq_A * exp(l.keq_B / T)); break;
            default: k[r] = 0.0; break;
        }
    }
}

SIMK_HOST_API void simk_model_count_elements(const SimkModel *m, const f64 *y, f64 *atoms) {
    for (int e = 0; e < 3; ++e) atoms[e] = 0.0;
    for (u32 s = 0; s < m->mech.n; ++s)
        for (int e = 0; e < 3; ++e) atoms[e] += (f64)m->elements[s][e] * y[s];
}

// Net stoichiometry in ascending species order, as the equation parser stores it.
static void set_net(Reaction &rx, const u8 *reac, const u8 *prod, u32 n, std::vector<int> &net) {
    net.assign(n, 0);
    rx = Reaction();
    for (u32 p = 0; p < 3; ++p)
        if (reac[p] != SIMK_NONE) {
            rx.reactant[rx.reactant_count++] = reac[p];
            net[reac[p]] -= 1;
        }
    for (u32 p = 0; p < 3; ++p)
        if (prod[p] != SIMK_NONE) net[prod[p]] += 1;
    for (u32 s = 0; s < n; ++s)
        if (net[s] != 0 && rx.net_count < SIMK_MAX_NET) {
            rx.net_species[rx.net_count] = (u8)s;
            rx.net_coef[rx.net_count] = (s8)net[s];
            ++rx.net_count;
        }
}

SIMK_HOST_API s32 simk_mech_build(u32 n_species, u32 n_reactions, const u8 *reactants, const u8 *products,
                                  const SimkCaps *caps, void *out, char *error, u32 error_size) {
    Mech m;
    m.n = n_species;
    m.R = n_reactions;
    m.rx.resize(n_reactions);
    std::vector<int> net;
    for (u32 r = 0; r < n_reactions; ++r) {
        for (u32 p = 0; p < 3; ++p) {
            u8 a = reactants[r * 3 + p], b = products[r * 3 + p];
            if ((a != SIMK_NONE && a >= n_species) || (b != SIMK_NONE && b >= n_species)) {
                snprintf(error, error_size, "species index out of range in reaction %u", r);
                return 1;
            }
        }
        set_net(m.rx[r], reactants + r * 3, products + r * 3, n_species, net);
        if (m.rx[r].reactant_count == 0) {
            snprintf(error, error_size, "reaction %u has no reactants", r);
            return 1;
        }
    }
    std::string err = finalize(m);
    if (err.empty()) {
        SimkMechLayout L = simk_mech_layout(caps);
        memset(out, 0, L.size);
        err = to_bytes(m, *caps, out);
    }
    if (!err.empty()) {
        snprintf(error, error_size, "%s", err.c_str());
        return 1;
    }
    return 0;
}

SIMK_HOST_API s32 simk_mech_stats(const void *bytes, const SimkCaps *caps, u64 *out) {
    Mech m;
    if (!from_bytes(bytes, *caps, m).empty()) return 1;
    out[0] = m.n;
    out[1] = m.R;
    out[2] = m.nnz;
    out[3] = m.lu_div.size();
    out[4] = m.lu_sub.size();
    out[5] = signature(m) & 0xFFFFFFFFu;
    return 0;
}

SIMK_HOST_API s32 simk_method(u32 id, KinMethod *out) { return method(id, *out) ? 0 : 1; }

SIMK_HOST_API u64 simk_fnv1a64(const void *data, u64 size, u64 seed) { return fnv1a64(data, size, seed); }

SIMK_HOST_API void simk_zones_layout(u32 n, u32 R, u32 padded, u64 *out) {
    const u64 p = padded;
    const u64 rows[12] = {n * p, p, p, p, p, p, R * p, n * p, n * p, p, p, p};
    u64 off = 0;
    for (int i = 0; i < 12; ++i) {
        out[i] = off;
        off += rows[i];
        if (i == 5) out[13] = off;
    }
    out[12] = off;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Seeded synthetic problems. RNG: SplitMix64 seeding xoshiro256**.

namespace {

struct Rng {
    u64 s[4];
    static u64 splitmix(u64 &x) {
        u64 z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    Rng(u64 seed, u64 stream) {
        u64 x = seed ^ (stream * 0xD1B54A32D192ED03ull);
        for (int i = 0; i < 4; ++i) s[i] = splitmix(x);
    }
    static u64 rotl(u64 x, int k) { return (x << k) | (x >> (64 - k)); }
    u64 next() {
        u64 result = rotl(s[1] * 5, 7) * 9;
        u64 t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return result;
    }
    f64 uniform() { return (f64)(next() >> 11) * (1.0 / 9007199254740992.0); }   // [0, 1)
    u32 below(u32 n) { return (u32)(uniform() * n); }
};

constexpr u64 kStreamMechanism = 1;
constexpr u64 kStreamRates = 2;
constexpr u64 kStreamCells = 1ull << 32;

}  // namespace

extern "C" SIMK_HOST_API u32 simk_synth_mechanism(const SimkSynthSpec *sp, u8 *reactants, u8 *products, f64 *params) {
    const u32 n = sp->n_species, R = sp->n_reactions;
    if (n < 2 || n > 255 || R == 0) return 0;
    const u32 attempts = sp->max_attempts ? sp->max_attempts : 1000;
    std::vector<int> net;
    for (u32 attempt = 0; attempt < attempts; ++attempt) {
        Rng rng(sp->seed, kStreamMechanism + 16ull * attempt);
        // Species masses 1..3: every reaction conserves total mass, so the
        // trajectories stay bounded like a real element-balanced mechanism.
        std::vector<u32> mass(n);
        for (u32 s = 0; s < n; ++s) mass[s] = 1 + rng.below(3);
        Mech m;
        m.n = n;
        m.R = R;
        m.rx.resize(R);
        for (u32 r = 0; r < R; ++r) {
            for (u32 tries = 0;; ++tries) {
                u32 c = rng.below(n);
                u32 lo = 0, hi = n - 1;
                if (sp->band) {
                    lo = c > sp->band ? c - sp->band : 0;
                    hi = c + sp->band < n ? c + sp->band : n - 1;
                }
                auto pick = [&]() -> u32 {
                    if (rng.uniform() < sp->long_range) return rng.below(n);
                    return lo + rng.below(hi - lo + 1);
                };
                u32 nr = 1 + (rng.uniform() < 0.5 ? 1u : 0u);
                u32 np = 1 + (rng.uniform() < 0.5 ? 1u : 0u);
                u8 reac[3] = {SIMK_NONE, SIMK_NONE, SIMK_NONE}, prod[3] = {SIMK_NONE, SIMK_NONE, SIMK_NONE};
                u32 total = 0;
                for (u32 p = 0; p < nr; ++p) {
                    reac[p] = (u8)pick();
                    total += mass[reac[p]];
                }
                // Products with the same total mass, from the same window.
                bool ok = false;
                for (u32 t = 0; t < 32 && !ok; ++t) {
                    if (np == 1) {
                        u32 a = pick();
                        if (mass[a] == total) { prod[0] = (u8)a; ok = true; }
                    } else {
                        u32 a = pick(), b = pick();
                        if (mass[a] + mass[b] == total) { prod[0] = (u8)a; prod[1] = (u8)b; ok = true; }
                    }
                }
                if (!ok) continue;
                Reaction rx;
                set_net(rx, reac, prod, n, net);
                if (rx.net_count == 0) continue;   // A -> A, A + B -> B + A: no net change
                m.rx[r] = rx;
                for (u32 p = 0; p < 3; ++p) {
                    reactants[r * 3 + p] = reac[p];
                    products[r * 3 + p] = prod[p];
                }
                break;
            }
        }
        if (!finalize(m).empty()) continue;
        f64 density = (f64)m.nnz / ((f64)n * n);
        if ((sp->density_lo > 0.0 || sp->density_hi > 0.0) && (density < sp->density_lo || density > sp->density_hi))
            continue;
        Rng rr(sp->seed, kStreamRates);
        for (u32 r = 0; r < R; ++r) {
            f64 *pr = params + 4ull * r;
            pr[0] = sp->log10_k_lo + (sp->log10_k_hi - sp->log10_k_lo) * rr.uniform();
            pr[1] = sp->EaR_hi * rr.uniform();
            f64 u = rr.uniform();
            pr[2] = u < sp->acid_frac ? 1.0 : (u < sp->acid_frac + sp->base_frac ? -1.0 : 0.0);
            bool unimolecular = m.rx[r].reactant_count == 1;
            pr[3] = (unimolecular && rr.uniform() < sp->water_frac) ? 1.0 : 0.0;
        }
        return attempt + 1;
    }
    return 0;
}

extern "C" SIMK_HOST_API void simk_synth_cells(const SimkSynthSpec *sp, const f64 *params, u64 first, u32 count,
                                               u32 padded, f64 *k, f64 *y, f64 *T_out, f64 *pH_out) {
    const u32 n = sp->n_species, R = sp->n_reactions;
    const bool use_pH = sp->pH_lo != 0.0 || sp->pH_hi != 0.0;
    for (u32 i = 0; i < count; ++i) {
        Rng rng(sp->seed, kStreamCells + first + i);
        f64 T = sp->T_lo + (sp->T_hi - sp->T_lo) * rng.uniform();
        f64 pH = use_pH ? sp->pH_lo + (sp->pH_hi - sp->pH_lo) * rng.uniform() : 7.0;
        f64 ox = sp->oxidant_log10_spread > 0.0 ? pow(10.0, sp->oxidant_log10_spread * rng.uniform()) : 1.0;
        f64 load = sp->load_log10_spread > 0.0 ? pow(10.0, -sp->load_log10_spread * rng.uniform()) : 1.0;
        for (u32 s = 0; s < n; ++s) {
            f64 ly = sp->log10_y_lo + (sp->log10_y_hi - sp->log10_y_lo) * rng.uniform();
            f64 v = pow(10.0, ly);
            if (s < sp->oxidant_species) v *= ox;
            v *= load;
            if (y) y[(u64)s * padded + i] = v;
        }
        if (k) {
            const f64 inv_dT = 1.0 / T - 1.0 / 298.0;
            for (u32 r = 0; r < R; ++r) {
                const f64 *pr = params + 4ull * r;
                f64 kr = pow(10.0, pr[0]) * exp(-pr[1] * inv_dT);
                if (use_pH && pr[2] > 0.0) kr *= pow(10.0, 7.0 - pH);
                if (use_pH && pr[2] < 0.0) kr *= pow(10.0, pH - 7.0);
                if (pr[3] > 0.0) kr *= 55.5;
                k[(u64)r * padded + i] = kr;
            }
        }
        if (T_out) T_out[i] = T;
        if (pH_out) pH_out[i] = pH;
    }
}
