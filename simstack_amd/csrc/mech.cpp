// Symbolic sparse LU and mechanism table (de)serialisation.
#include "mech.hpp"

#include <cstring>

namespace simk {

std::string finalize(Mech &m) {
    const u32 n = m.n;
    if (n == 0) return "mechanism has no species";
    if (n > 255) return "more than 255 species (u8 columns)";
    if (m.rx.size() != m.R) return "reaction count mismatch";
    // Dense boolean pattern: J's structure (net species x reactant) and the diagonal.
    std::vector<u8> pat((size_t)n * n, 0);
    auto P = [&](u32 i, u32 j) -> u8 & { return pat[(size_t)i * n + j]; };
    for (u32 i = 0; i < n; ++i) P(i, i) = 1;
    for (const Reaction &r : m.rx) {
        for (u32 p = 0; p < r.reactant_count; ++p) {
            if (r.reactant[p] >= n) return "reactant index out of range";
            for (u32 e = 0; e < r.net_count; ++e) {
                if (r.net_species[e] >= n) return "net species index out of range";
                P(r.net_species[e], r.reactant[p]) = 1;
            }
        }
    }
    // Fill closure of elimination without pivoting, rows top to bottom: row i
    // picks up row k's upper pattern for every k < i it has an entry in. Row k
    // is complete before row i reads it, so one pass suffices.
    for (u32 i = 1; i < n; ++i)
        for (u32 k = 0; k < i; ++k) {
            if (!P(i, k)) continue;
            for (u32 j = k + 1; j < n; ++j)
                if (P(k, j)) P(i, j) = 1;
        }
    // Compact row-major numbering of the pattern.
    std::vector<u32> idx((size_t)n * n, 0xFFFFFFFFu);
    u32 count = 0;
    for (u32 i = 0; i < n; ++i)
        for (u32 j = 0; j < n; ++j)
            if (P(i, j)) idx[(size_t)i * n + j] = count++;
    if (count > 65535) return "more than 65535 non-zeros (u16 indices)";
    m.nnz = count;
    auto I = [&](u32 i, u32 j) { return (u16)idx[(size_t)i * n + j]; };
    m.diag.assign(n, 0);
    for (u32 i = 0; i < n; ++i) m.diag[i] = I(i, i);

    m.div_begin.assign(n + 1, 0);
    m.lu_div.clear();
    m.lu_sub.clear();
    for (u32 i = 0; i < n; ++i) {
        m.div_begin[i] = (u16)m.lu_div.size();
        for (u32 k = 0; k < i; ++k) {
            if (!P(i, k)) continue;
            KinLuDiv d;
            d.ik = I(i, k);
            d.kk = I(k, k);
            d.sub_begin = (u16)m.lu_sub.size();
            for (u32 j = k + 1; j < n; ++j) {
                if (!P(k, j)) continue;
                KinLuSub s;
                s.ij = I(i, j);
                s.kj = I(k, j);
                m.lu_sub.push_back(s);
                if (m.lu_sub.size() > 65535) return "more than 65535 LU updates (u16 indices)";
            }
            d.sub_end = (u16)m.lu_sub.size();
            m.lu_div.push_back(d);
            if (m.lu_div.size() > 65535) return "more than 65535 LU divisions (u16 indices)";
        }
    }
    m.div_begin[n] = (u16)m.lu_div.size();

    m.lower_begin.assign(n + 1, 0);
    m.upper_begin.assign(n + 1, 0);
    m.lower.clear();
    m.upper.clear();
    for (u32 i = 0; i < n; ++i) {
        m.lower_begin[i] = (u16)m.lower.size();
        m.upper_begin[i] = (u16)m.upper.size();
        for (u32 j = 0; j < i; ++j)
            if (P(i, j)) m.lower.push_back(KinRowEntry{I(i, j), (u8)j, 0});
        // Upper entries in descending column order, the order a column-oriented
        // back substitution applies them in (kin_build.cpp:209-213).
        for (u32 j = n; j-- > i + 1;)
            if (P(i, j)) m.upper.push_back(KinRowEntry{I(i, j), (u8)j, 0});
    }
    m.lower_begin[n] = (u16)m.lower.size();
    m.upper_begin[n] = (u16)m.upper.size();

    for (Reaction &r : m.rx) {
        for (u32 p = 0; p < SIMK_MAX_REACTANTS; ++p)
            for (u32 e = 0; e < SIMK_MAX_NET; ++e) r.jac_index[p][e] = 0;
        for (u32 p = 0; p < r.reactant_count; ++p)
            for (u32 e = 0; e < r.net_count; ++e) r.jac_index[p][e] = I(r.net_species[e], r.reactant[p]);
    }
    return "";
}

std::string validate(const Mech &m) {
    const u32 n = m.n, nnz = m.nnz;
    if (n == 0 || n > 255) return "species count out of range";
    if (nnz < n || nnz > 65535) return "nnz out of range";
    if (m.rx.size() != m.R) return "reaction count mismatch";
    if (m.diag.size() != n || m.div_begin.size() != n + 1 || m.lower_begin.size() != n + 1 ||
        m.upper_begin.size() != n + 1)
        return "row table sizes";
    for (u32 r = 0; r < m.R; ++r) {
        const Reaction &x = m.rx[r];
        if (x.reactant_count == 0 || x.reactant_count > SIMK_MAX_REACTANTS) return "reactant count out of range";
        if (x.net_count > SIMK_MAX_NET) return "net count out of range";
        for (u32 p = 0; p < x.reactant_count; ++p) {
            if (x.reactant[p] >= n) return "reactant index out of range";
            for (u32 e = 0; e < x.net_count; ++e)
                if (x.jac_index[p][e] >= nnz) return "jacobian index out of range";
        }
        for (u32 e = 0; e < x.net_count; ++e)
            if (x.net_species[e] >= n) return "net species out of range";
    }
    for (u32 i = 0; i < n; ++i)
        if (m.diag[i] >= nnz) return "diag index out of range";
    if (m.div_begin[0] != 0 || m.div_begin[n] != m.lu_div.size()) return "div_begin bounds";
    for (u32 i = 0; i < n; ++i)
        if (m.div_begin[i] > m.div_begin[i + 1]) return "div_begin not monotone";
    for (const KinLuDiv &d : m.lu_div) {
        if (d.ik >= nnz || d.kk >= nnz) return "lu_div index out of range";
        if (d.sub_begin > d.sub_end || d.sub_end > m.lu_sub.size()) return "lu_div sub range";
    }
    for (const KinLuSub &s : m.lu_sub)
        if (s.ij >= nnz || s.kj >= nnz) return "lu_sub index out of range";
    if (m.lower_begin[0] != 0 || m.lower_begin[n] != m.lower.size()) return "lower_begin bounds";
    if (m.upper_begin[0] != 0 || m.upper_begin[n] != m.upper.size()) return "upper_begin bounds";
    for (u32 i = 0; i < n; ++i) {
        if (m.lower_begin[i] > m.lower_begin[i + 1]) return "lower_begin not monotone";
        if (m.upper_begin[i] > m.upper_begin[i + 1]) return "upper_begin not monotone";
    }
    for (const KinRowEntry &e : m.lower)
        if (e.index >= nnz || e.col >= n) return "lower entry out of range";
    for (const KinRowEntry &e : m.upper)
        if (e.index >= nnz || e.col >= n) return "upper entry out of range";
    return "";
}

template <class T>
static T rd(const u8 *b, u64 off) {
    T v;
    memcpy(&v, b + off, sizeof(T));
    return v;
}
template <class T>
static void wr(u8 *b, u64 off, const T &v) {
    memcpy(b + off, &v, sizeof(T));
}

std::string from_bytes(const void *bytes, const SimkCaps &caps, Mech &out) {
    const u8 *b = (const u8 *)bytes;
    SimkMechLayout L = simk_mech_layout(&caps);
    Mech m;
    m.n = rd<u32>(b, L.species_count);
    m.R = rd<u32>(b, L.reaction_count);
    m.nnz = rd<u32>(b, L.nnz);
    u32 ndiv = rd<u32>(b, L.lu_div_count), nsub = rd<u32>(b, L.lu_sub_count);
    if (m.n == 0 || m.n > caps.species) return "species_count exceeds capacity";
    if (m.R > caps.reactions) return "reaction_count exceeds capacity";
    if (m.nnz > caps.nnz) return "nnz exceeds capacity";
    if (ndiv > caps.lu_div || nsub > caps.lu_sub) return "LU op counts exceed capacity";
    m.rx.resize(m.R);
    for (u32 r = 0; r < m.R; ++r) memcpy(&m.rx[r], b + L.reactions + (u64)r * sizeof(KinReaction), sizeof(KinReaction));
    m.diag.resize(m.n);
    memcpy(m.diag.data(), b + L.diag, 2ull * m.n);
    m.div_begin.resize(m.n + 1);
    memcpy(m.div_begin.data(), b + L.div_begin, 2ull * (m.n + 1));
    m.lu_div.resize(ndiv);
    memcpy(m.lu_div.data(), b + L.lu_div, sizeof(KinLuDiv) * ndiv);
    m.lu_sub.resize(nsub);
    memcpy(m.lu_sub.data(), b + L.lu_sub, sizeof(KinLuSub) * nsub);
    m.lower_begin.resize(m.n + 1);
    memcpy(m.lower_begin.data(), b + L.lower_begin, 2ull * (m.n + 1));
    m.upper_begin.resize(m.n + 1);
    memcpy(m.upper_begin.data(), b + L.upper_begin, 2ull * (m.n + 1));
    u32 nlo = m.lower_begin[m.n], nup = m.upper_begin[m.n];
    if (nlo > caps.nnz || nup > caps.nnz) return "row lists exceed capacity";
    m.lower.resize(nlo);
    memcpy(m.lower.data(), b + L.lower, sizeof(KinRowEntry) * nlo);
    m.upper.resize(nup);
    memcpy(m.upper.data(), b + L.upper, sizeof(KinRowEntry) * nup);
    std::string err = validate(m);
    if (!err.empty()) return err;
    out = std::move(m);
    return "";
}

std::string to_bytes(const Mech &m, const SimkCaps &caps, void *bytes) {
    if (m.n > caps.species || m.R > caps.reactions || m.nnz > caps.nnz || m.lu_div.size() > caps.lu_div ||
        m.lu_sub.size() > caps.lu_sub || m.lower.size() > caps.nnz || m.upper.size() > caps.nnz)
        return "mechanism exceeds the given capacities";
    u8 *b = (u8 *)bytes;
    SimkMechLayout L = simk_mech_layout(&caps);
    wr<u32>(b, L.species_count, m.n);
    wr<u32>(b, L.reaction_count, m.R);
    for (u32 r = 0; r < m.R; ++r) memcpy(b + L.reactions + (u64)r * sizeof(KinReaction), &m.rx[r], sizeof(KinReaction));
    wr<u32>(b, L.nnz, m.nnz);
    memcpy(b + L.diag, m.diag.data(), 2ull * m.n);
    wr<u32>(b, L.lu_div_count, (u32)m.lu_div.size());
    wr<u32>(b, L.lu_sub_count, (u32)m.lu_sub.size());
    memcpy(b + L.div_begin, m.div_begin.data(), 2ull * (m.n + 1));
    memcpy(b + L.lu_div, m.lu_div.data(), sizeof(KinLuDiv) * m.lu_div.size());
    memcpy(b + L.lu_sub, m.lu_sub.data(), sizeof(KinLuSub) * m.lu_sub.size());
    memcpy(b + L.lower_begin, m.lower_begin.data(), 2ull * (m.n + 1));
    memcpy(b + L.upper_begin, m.upper_begin.data(), 2ull * (m.n + 1));
    memcpy(b + L.lower, m.lower.data(), sizeof(KinRowEntry) * m.lower.size());
    memcpy(b + L.upper, m.upper.data(), sizeof(KinRowEntry) * m.upper.size());
    return "";
}

u64 fnv1a64(const void *data, u64 size, u64 seed) {
    const u8 *p = (const u8 *)data;
    u64 h = seed;
    for (u64 i = 0; i < size; ++i) {
        h ^= p[i];
        h *= 1099511628211ull;
    }
    return h;
}

u64 signature(const Mech &m) {
    u64 h = fnv1a64(&m.n, 4);
    h = fnv1a64(&m.R, 4, h);
    h = fnv1a64(&m.nnz, 4, h);
    h = fnv1a64(m.rx.data(), sizeof(Reaction) * m.rx.size(), h);
    h = fnv1a64(m.diag.data(), 2 * m.diag.size(), h);
    h = fnv1a64(m.div_begin.data(), 2 * m.div_begin.size(), h);
    h = fnv1a64(m.lu_div.data(), sizeof(KinLuDiv) * m.lu_div.size(), h);
    h = fnv1a64(m.lu_sub.data(), sizeof(KinLuSub) * m.lu_sub.size(), h);
    h = fnv1a64(m.lower_begin.data(), 2 * m.lower_begin.size(), h);
    h = fnv1a64(m.upper_begin.data(), 2 * m.upper_begin.size(), h);
    h = fnv1a64(m.lower.data(), sizeof(KinRowEntry) * m.lower.size(), h);
    h = fnv1a64(m.upper.data(), sizeof(KinRowEntry) * m.upper.size(), h);
    return h;
}

bool method(u32 id, KinMethod &r) {
    memset(&r, 0, sizeof(r));
    switch (id) {
        case KinMethod_Ros2: {
            // Two stages, order 2(1), L-stable; gamma = 1 + 1/sqrt(2).
            const f64 g = 1.0 + 1.0 / 1.4142135623730950488;
            r.name = "ros2";
            r.stages = 2;
            r.order = 2;
            r.a[0] = 1.0 / g;
            r.c[0] = -2.0 / g;
            r.m[0] = 3.0 / (2.0 * g);
            r.m[1] = 1.0 / (2.0 * g);
            r.e[0] = 1.0 / (2.0 * g);
            r.e[1] = 1.0 / (2.0 * g);
            r.gamma0 = g;
            r.new_f[0] = r.new_f[1] = 1;
            return true;
        }
        case KinMethod_Ros3: {
            // Three stages, order 3(2), L-stable (Sandu et al. 1997, table 3).
            r.name = "ros3";
            r.stages = 3;
            r.order = 3;
            r.a[0] = 1.0;
            r.a[1] = 1.0;
            r.a[2] = 0.0;
            r.c[0] = -0.10156171083877702091975600115545e+01;
            r.c[1] = 0.40759956452537699824805835358067e+01;
            r.c[2] = 0.92076794298330791242156818474003e+01;
            r.m[0] = 0.1e+01;
            r.m[1] = 0.61697947043828245592553615689730e+01;
            r.m[2] = -0.42772256543218573326238373806514e+00;
            r.e[0] = 0.5e+00;
            r.e[1] = -0.29079558716805469821718236208017e+01;
            r.e[2] = 0.22354069897811569627360909276199e+00;
            r.gamma0 = 0.43586652150845899941601945119356;
            r.new_f[0] = 1;
            r.new_f[1] = 1;
            r.new_f[2] = 0;
            return true;
        }
        case KinMethod_Rodas3: {
            // Four stages, order 3(2), stiffly accurate; gamma = 1/2.
            r.name = "rodas3";
            r.stages = 4;
            r.order = 3;
            const f64 a[6] = {0.0, 2.0, 0.0, 2.0, 0.0, 1.0};
            const f64 c[6] = {4.0, 1.0, -1.0, 1.0, -1.0, -8.0 / 3.0};
            const f64 mm[4] = {2.0, 0.0, 1.0, 1.0};
            const f64 e[4] = {0.0, 0.0, 0.0, 1.0};
            for (int i = 0; i < 6; ++i) { r.a[i] = a[i]; r.c[i] = c[i]; }
            for (int i = 0; i < 4; ++i) { r.m[i] = mm[i]; r.e[i] = e[i]; r.new_f[i] = 1; }
            r.gamma0 = 0.5;
            return true;
        }
        default:
            return false;
    }
}

}  // namespace simk
