/* ORACLE — test infrastructure only (see simk_oracle.h). Plain C, scalar f64,
 * built with -ffp-contract=off so every fused multiply-add is an explicit fma()
 * call, exactly where the reference writes lane_fmadd / lane_fnmadd
 * (src/base/base_lane.h:13-21, 51-56). min/max are the select forms
 * "a < b ? a : b" / "a > b ? a : b"; |a| keeps -0 (base_lane.h:53-55).
 */
#include "simk_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ---- lane vocabulary (src/base/base_lane.h:47-80) ---------------------- */
static inline f64 o_min(f64 a, f64 b) { return a < b ? a : b; }
static inline f64 o_max(f64 a, f64 b) { return a > b ? a : b; }
static inline f64 o_abs(f64 a) { return a < 0.0 ? -a : a; }
static inline f64 o_fmadd(f64 a, f64 b, f64 c) { return fma(a, b, c); }
static inline f64 o_fnmadd(f64 a, f64 b, f64 c) { return fma(-a, b, c); }
static inline int o_is_finite(f64 a) {
    volatile f64 d = a - a; /* x - x is 0 for finite x, NaN otherwise (base_lane.h:71-72) */
    return d == d;
}

/* ---- table view over the reference KinMech bytes at any capacities ------- */
typedef struct OMech {
    u32 n, R, nnz;
    const KinReaction *rx;
    const u16 *diag, *div_begin, *lower_begin, *upper_begin;
    const KinLuDiv *lu_div;
    const KinLuSub *lu_sub;
    const KinRowEntry *lower, *upper;
} OMech;

static OMech o_view(const void *bytes, const SimkCaps *caps) {
    const u8 *b = (const u8 *)bytes;
    SimkMechLayout L = simk_mech_layout(caps);
    OMech m;
    memcpy(&m.n, b + L.species_count, 4);
    memcpy(&m.R, b + L.reaction_count, 4);
    memcpy(&m.nnz, b + L.nnz, 4);
    m.rx = (const KinReaction *)(b + L.reactions);
    m.diag = (const u16 *)(b + L.diag);
    m.div_begin = (const u16 *)(b + L.div_begin);
    m.lu_div = (const KinLuDiv *)(b + L.lu_div);
    m.lu_sub = (const KinLuSub *)(b + L.lu_sub);
    m.lower_begin = (const u16 *)(b + L.lower_begin);
    m.upper_begin = (const u16 *)(b + L.upper_begin);
    m.lower = (const KinRowEntry *)(b + L.lower);
    m.upper = (const KinRowEntry *)(b + L.upper);
    return m;
}

/* ---- symbolic phase: src/sim/kinetics/kin_build.cpp:100-148 (net
 * stoichiometry, ascending species) and :152-223 (pattern, fill, index, op
 * lists, lower ascending / upper descending, jac_index). ------------------- */
ORACLE_API s32 oracle_finalize(u32 n, u32 R, const u8 *reactants, const u8 *products, const SimkCaps *caps, void *out) {
    if (n == 0 || n > caps->species || R > caps->reactions) return 1;
    SimkMechLayout L = simk_mech_layout(caps);
    u8 *b = (u8 *)out;
    memset(b, 0, L.size);
    memcpy(b + L.species_count, &n, 4);
    memcpy(b + L.reaction_count, &R, 4);
    KinReaction *rx = (KinReaction *)(b + L.reactions);
    int *net = (int *)calloc(n, sizeof(int));
    for (u32 r = 0; r < R; ++r) {
        KinReaction *x = rx + r;
        memset(x, 0, sizeof(*x));
        for (u32 p = 0; p < 3; ++p) x->reactant[p] = 0xFF;
        memset(net, 0, n * sizeof(int));
        for (u32 p = 0; p < 3; ++p) {
            u8 s = reactants[r * 3 + p];
            if (s == 0xFF) continue;
            x->reactant[x->reactant_count++] = s;
            net[s] -= 1;
        }
        for (u32 p = 0; p < 3; ++p) {
            u8 s = products[r * 3 + p];
            if (s != 0xFF) net[s] += 1;
        }
        for (u32 s = 0; s < n; ++s) {
            if (!net[s]) continue;
            if (x->net_count >= 6) { free(net); return 2; }
            x->net_species[x->net_count] = (u8)s;
            x->net_coef[x->net_count] = (s8)net[s];
            x->net_count++;
        }
    }
    free(net);

    u8 *pat = (u8 *)calloc((size_t)n * n, 1);
#define PAT(i, j) pat[(size_t)(i) * n + (j)]
    for (u32 i = 0; i < n; ++i) PAT(i, i) = 1;
    for (u32 r = 0; r < R; ++r)
        for (u32 p = 0; p < rx[r].reactant_count; ++p)
            for (u32 e = 0; e < rx[r].net_count; ++e) PAT(rx[r].net_species[e], rx[r].reactant[p]) = 1;
    for (u32 i = 1; i < n; ++i)
        for (u32 k = 0; k < i; ++k)
            if (PAT(i, k))
                for (u32 j = k + 1; j < n; ++j)
                    if (PAT(k, j)) PAT(i, j) = 1;
    u32 *index = (u32 *)malloc((size_t)n * n * sizeof(u32));
    u32 nnz = 0;
    for (u32 i = 0; i < n; ++i)
        for (u32 j = 0; j < n; ++j) index[(size_t)i * n + j] = PAT(i, j) ? nnz++ : 0xFFFFFFFFu;
#define IDX(i, j) ((u16)index[(size_t)(i) * n + (j)])
    s32 status = 0;
    if (nnz > caps->nnz) { status = 3; goto done; }
    memcpy(b + L.nnz, &nnz, 4);
    u16 *diag = (u16 *)(b + L.diag);
    for (u32 i = 0; i < n; ++i) diag[i] = IDX(i, i);
    {
        u16 *div_begin = (u16 *)(b + L.div_begin);
        KinLuDiv *lu_div = (KinLuDiv *)(b + L.lu_div);
        KinLuSub *lu_sub = (KinLuSub *)(b + L.lu_sub);
        u32 nd = 0, ns = 0;
        for (u32 i = 0; i < n; ++i) {
            div_begin[i] = (u16)nd;
            for (u32 k = 0; k < i; ++k) {
                if (!PAT(i, k)) continue;
                if (nd >= caps->lu_div) { status = 4; goto done; }
                KinLuDiv *d = lu_div + nd++;
                d->ik = IDX(i, k);
                d->kk = IDX(k, k);
                d->sub_begin = (u16)ns;
                for (u32 j = k + 1; j < n; ++j) {
                    if (!PAT(k, j)) continue;
                    if (ns >= caps->lu_sub) { status = 5; goto done; }
                    lu_sub[ns].ij = IDX(i, j);
                    lu_sub[ns].kj = IDX(k, j);
                    ++ns;
                }
                d->sub_end = (u16)ns;
            }
        }
        div_begin[n] = (u16)nd;
        memcpy(b + L.lu_div_count, &nd, 4);
        memcpy(b + L.lu_sub_count, &ns, 4);
    }
    {
        u16 *lower_begin = (u16 *)(b + L.lower_begin), *upper_begin = (u16 *)(b + L.upper_begin);
        KinRowEntry *lower = (KinRowEntry *)(b + L.lower), *upper = (KinRowEntry *)(b + L.upper);
        u32 lo = 0, up = 0;
        for (u32 i = 0; i < n; ++i) {
            lower_begin[i] = (u16)lo;
            upper_begin[i] = (u16)up;
            for (u32 j = 0; j < i; ++j)
                if (PAT(i, j)) { lower[lo].index = IDX(i, j); lower[lo].col = (u8)j; ++lo; }
            for (u32 j = n; j-- > i + 1;)
                if (PAT(i, j)) { upper[up].index = IDX(i, j); upper[up].col = (u8)j; ++up; }
        }
        lower_begin[n] = (u16)lo;
        upper_begin[n] = (u16)up;
    }
    for (u32 r = 0; r < R; ++r)
        for (u32 p = 0; p < rx[r].reactant_count; ++p)
            for (u32 e = 0; e < rx[r].net_count; ++e) rx[r].jac_index[p][e] = IDX(rx[r].net_species[e], rx[r].reactant[p]);
done:
    free(index);
    free(pat);
#undef PAT
#undef IDX
    return status;
}

/* ---- method coefficients: src/sim/kinetics/kin_solver.h:352-395 --------- */
ORACLE_API void oracle_method(u32 id, KinMethod *r) {
    memset(r, 0, sizeof(*r));
    if (id == 0) {
        f64 g = 1.0 + 1.0 / 1.4142135623730950488;
        r->stages = 2; r->order = 2;
        r->a[0] = 1.0 / g;
        r->c[0] = -2.0 / g;
        r->m[0] = 3.0 / (2.0 * g); r->m[1] = 1.0 / (2.0 * g);
        r->e[0] = 1.0 / (2.0 * g); r->e[1] = 1.0 / (2.0 * g);
        r->gamma0 = g;
        r->new_f[0] = r->new_f[1] = 1;
    } else if (id == 1) {
        r->stages = 3; r->order = 3;
        r->a[0] = 1.0; r->a[1] = 1.0; r->a[2] = 0.0;
        r->c[0] = -0.10156171083877702091975600115545e+01;
        r->c[1] = 0.40759956452537699824805835358067e+01;
        r->c[2] = 0.92076794298330791242156818474003e+01;
        r->m[0] = 0.1e+01;
        r->m[1] = 0.61697947043828245592553615689730e+01;
        r->m[2] = -0.42772256543218573326238373806514e+00;
        r->e[0] = 0.5e+00;
        r->e[1] = -0.29079558716805469821718236208017e+01;
        r->e[2] = 0.22354069897811569627360909276199e+00;
        r->gamma0 = 0.43586652150845899941601945119356;
        r->new_f[0] = 1; r->new_f[1] = 1; r->new_f[2] = 0;
    } else {
        r->stages = 4; r->order = 3;
        r->a[0] = 0.0; r->a[1] = 2.0; r->a[2] = 0.0; r->a[3] = 2.0; r->a[4] = 0.0; r->a[5] = 1.0;
        r->c[0] = 4.0; r->c[1] = 1.0; r->c[2] = -1.0; r->c[3] = 1.0; r->c[4] = -1.0; r->c[5] = -8.0 / 3.0;
        r->m[0] = 2.0; r->m[1] = 0.0; r->m[2] = 1.0; r->m[3] = 1.0;
        r->e[0] = 0.0; r->e[1] = 0.0; r->e[2] = 0.0; r->e[3] = 1.0;
        r->gamma0 = 0.5;
        r->new_f[0] = r->new_f[1] = r->new_f[2] = r->new_f[3] = 1;
    }
}

/* ---- the four kernels ---------------------------------------------------- */

/* kin_fun, src/sim/kinetics/kin_solver.h:39-53 */
static void o_fun(const OMech *m, const f64 *k, const f64 *src, const f64 *y, f64 *f) {
    for (u32 s = 0; s < m->n; ++s) f[s] = src[s];
    for (u32 r = 0; r < m->R; ++r) {
        const KinReaction *rx = m->rx + r;
        f64 rate = k[r];
        for (u32 p = 0; p < rx->reactant_count; ++p) rate = rate * y[rx->reactant[p]];
        for (u32 e = 0; e < rx->net_count; ++e) {
            u32 s = rx->net_species[e];
            f[s] = o_fmadd((f64)rx->net_coef[e], rate, f[s]);
        }
    }
}

/* kin_jac (coef) and kin_neg_jac (-coef), kin_solver.h:55-93 */
static void o_jac_signed(const OMech *m, const f64 *k, const f64 *y, f64 *a, f64 sign) {
    for (u32 i = 0; i < m->nnz; ++i) a[i] = 0.0;
    for (u32 r = 0; r < m->R; ++r) {
        const KinReaction *rx = m->rx + r;
        for (u32 p = 0; p < rx->reactant_count; ++p) {
            f64 d = k[r];
            for (u32 q = 0; q < rx->reactant_count; ++q)
                if (q != p) d = d * y[rx->reactant[q]];
            for (u32 e = 0; e < rx->net_count; ++e) {
                u32 idx = rx->jac_index[p][e];
                a[idx] = o_fmadd(sign * (f64)rx->net_coef[e], d, a[idx]);
            }
        }
    }
}

/* kin_factor, kin_solver.h:95-113: row i's eliminations, then the diagonal
 * holds the reciprocal pivot 1 / (a_ii + ghinv).
 *
 * store32 (not the reference: the FP32 factor-storage form, GpuKinLaunch
 * precision 3): once row i is final, each of its off-diagonal entries (L and U)
 * is rounded to the nearest f32 and kept as that value, so later rows eliminate
 * against the rounded U entries and the solves read rounded L and U entries.
 * The reciprocal pivots and every operation stay FP64. With store32 0 this is
 * the reference's factor. */
static void o_factor_rows(const OMech *m, f64 ghinv, f64 *a, int store32) {
    for (u32 i = 0; i < m->n; ++i) {
        for (u32 d = m->div_begin[i]; d < m->div_begin[i + 1]; ++d) {
            KinLuDiv div = m->lu_div[d];
            f64 l = a[div.ik] * a[div.kk];
            a[div.ik] = l;
            for (u32 s = div.sub_begin; s < div.sub_end; ++s) {
                KinLuSub sub = m->lu_sub[s];
                a[sub.ij] = o_fnmadd(l, a[sub.kj], a[sub.ij]);
            }
        }
        u32 ii = m->diag[i];
        a[ii] = 1.0 / (a[ii] + ghinv);
        if (store32) {
            for (u32 e = m->lower_begin[i]; e < m->lower_begin[i + 1]; ++e)
                a[m->lower[e].index] = (f64)(f32)a[m->lower[e].index];
            for (u32 e = m->upper_begin[i]; e < m->upper_begin[i + 1]; ++e)
                a[m->upper[e].index] = (f64)(f32)a[m->upper[e].index];
        }
    }
}

static void o_factor(const OMech *m, f64 ghinv, f64 *a) { o_factor_rows(m, ghinv, a, 0); }

/* kin_solve, kin_solver.h:122-139 */
static void o_solve(const OMech *m, const f64 *a, f64 *x) {
    const u32 n = m->n;
    for (u32 i = 0; i < n; ++i)
        for (u32 e = m->lower_begin[i]; e < m->lower_begin[i + 1]; ++e)
            x[i] = o_fnmadd(a[m->lower[e].index], x[m->lower[e].col], x[i]);
    for (u32 ii = n; ii-- > 0;) {
        for (u32 e = m->upper_begin[ii]; e < m->upper_begin[ii + 1]; ++e)
            x[ii] = o_fnmadd(a[m->upper[e].index], x[m->upper[e].col], x[ii]);
        x[ii] = x[ii] * a[m->diag[ii]];
    }
}

ORACLE_API void oracle_fun(const void *mech, const SimkCaps *caps, const f64 *k, const f64 *src, const f64 *y, f64 *f) {
    OMech m = o_view(mech, caps);
    o_fun(&m, k, src, y, f);
}
ORACLE_API void oracle_jac(const void *mech, const SimkCaps *caps, const f64 *k, const f64 *y, f64 *jac) {
    OMech m = o_view(mech, caps);
    o_jac_signed(&m, k, y, jac, 1.0);
}
ORACLE_API void oracle_neg_jac(const void *mech, const SimkCaps *caps, const f64 *k, const f64 *y, f64 *a) {
    OMech m = o_view(mech, caps);
    o_jac_signed(&m, k, y, a, -1.0);
}
ORACLE_API void oracle_factor(const void *mech, const SimkCaps *caps, f64 ghinv, f64 *a) {
    OMech m = o_view(mech, caps);
    o_factor(&m, ghinv, a);
}
ORACLE_API void oracle_factor_store32(const void *mech, const SimkCaps *caps, f64 ghinv, f64 *a) {
    OMech m = o_view(mech, caps);
    o_factor_rows(&m, ghinv, a, 1);
}
ORACLE_API void oracle_solve(const void *mech, const SimkCaps *caps, const f64 *a, f64 *x) {
    OMech m = o_view(mech, caps);
    o_solve(&m, a, x);
}

/* lane_cbrt_clamped, src/base/base_lane.h:123-135: x^(21/64) from six square
 * roots, then two Newton steps for the cube root. */
ORACLE_API f64 oracle_cbrt_clamped(f64 x) {
    x = o_min(o_max(x, 1e-4), 1e4);
    f64 r4 = sqrt(sqrt(x));
    f64 r16 = sqrt(sqrt(r4));
    f64 r64 = sqrt(sqrt(r16));
    f64 g = r4 * r16 * r64;
    const f64 third = 1.0 / 3.0;
    g = (2.0 * g + x / (g * g)) * third;
    g = (2.0 * g + x / (g * g)) * third;
    return g;
}

/* kin_step_factor, kin_solver.h:169-180 */
ORACLE_API f64 oracle_step_factor(f64 err, u32 order) {
    f64 e = o_min(o_max(err, 1e-4), 1e4);
    f64 root;
    if (order == 2) root = sqrt(e);
    else if (order == 4) root = sqrt(sqrt(e));
    else root = oracle_cbrt_clamped(e);
    f64 fac = 0.9 / root;
    return o_min(o_max(fac, 0.2), 6.0);
}

/* ---- the integrator --------------------------------------------------------- */

typedef struct OWork {
    f64 *lu, *kstage, *fcn, *ytmp; /* nnz, 4*n, n, n */
} OWork;

/* kin_integrate<f64>, kin_solver.h:189-313, one lane. stats = steps, rejects,
 * pulses, failed. */
static void o_integrate(const OMech *m, const KinMethod *method, const KinControl *ctl, const f64 *k, const f64 *src,
                        f64 *y, f64 *t_io, f64 t_end, f64 *h_io, f64 *stats, OWork *w, int store32) {
    const u32 n = m->n;
    const f64 h_min = ctl->h_min, rtol = ctl->rtol, atol = ctl->atol;
    const f64 inv_n = 1.0 / (f64)n;
    f64 t = *t_io, h = *h_io;
    f64 *lu = w->lu, *fcn = w->fcn, *ytmp = w->ytmp;
#define KST(st, s) w->kstage[(st) * n + (s)]

    int failed = stats[3] > 0.0;
    int active = (t < t_end) && !failed;
    int reject_last = 0, reject_more = 0;
    h = h > 0.0 ? h : ctl->h_start;

    u32 iterations = 0;
    while (active) {
        if (++iterations > ctl->max_steps) {
            stats[3] = 1.0;
            break;
        }
        f64 remaining = t_end - t;
        f64 h_cap = remaining;
        if (ctl->h_max > 0.0) h_cap = o_min(h_cap, ctl->h_max);
        f64 h_wanted = o_max(h, h_min);
        int last = h_wanted >= remaining;
        f64 hs = last ? remaining : o_min(h_wanted, h_cap);

        o_fun(m, k, src, y, fcn);
        o_jac_signed(m, k, y, lu, -1.0);
        f64 ghinv = 1.0 / (hs * method->gamma0);
        o_factor_rows(m, ghinv, lu, store32);
        f64 hinv = 1.0 / hs;

        for (u32 st = 0; st < method->stages; ++st) {
            u32 packed = st * (st - 1) / 2;
            if (st > 0 && method->new_f[st]) {
                for (u32 s = 0; s < n; ++s) ytmp[s] = y[s];
                for (u32 j = 0; j < st; ++j) {
                    f64 a = method->a[packed + j];
                    if (a == 0.0) continue;
                    for (u32 s = 0; s < n; ++s) ytmp[s] = o_fmadd(a, KST(j, s), ytmp[s]);
                }
                o_fun(m, k, src, ytmp, fcn);
            }
            for (u32 s = 0; s < n; ++s) KST(st, s) = fcn[s];
            for (u32 j = 0; j < st; ++j) {
                f64 c = method->c[packed + j];
                if (c == 0.0) continue;
                f64 hc = c * hinv;
                for (u32 s = 0; s < n; ++s) KST(st, s) = o_fmadd(hc, KST(j, s), KST(st, s));
            }
            o_solve(m, lu, &KST(st, 0));
        }

        f64 err = 0.0;
        for (u32 s = 0; s < n; ++s) {
            f64 ynew = y[s], yerr = 0.0;
            for (u32 st = 0; st < method->stages; ++st) {
                if (method->m[st] != 0.0) ynew = o_fmadd(method->m[st], KST(st, s), ynew);
                if (method->e[st] != 0.0) yerr = o_fmadd(method->e[st], KST(st, s), yerr);
            }
            f64 scale = o_fmadd(rtol, o_max(o_abs(y[s]), o_abs(ynew)), atol);
            f64 q = yerr / scale;
            err = o_fmadd(q, q, err);
            ytmp[s] = ynew;
        }
        err = sqrt(err * inv_n);
        if (!o_is_finite(err)) err = 1e10;

        f64 fac = oracle_step_factor(err, method->order);
        int ok = (err <= 1.0) || (hs <= h_min);
        if (ctl->fixed_step) {
            ok = 1;
            fac = 1.0;
        }
        int accept = ok, reject = !ok;

        if (accept) {
            for (u32 s = 0; s < n; ++s) {
                f64 ynew = ytmp[s];
                if (ctl->clamp_negative) ynew = o_max(ynew, 0.0);
                y[s] = ynew;
            }
            t = last ? t_end : t + hs;
        }
        f64 h_acc = hs * fac;
        if (reject_last) h_acc = o_min(h_acc, hs);
        if (last && fac >= 1.0) h_acc = o_max(h_acc, h_wanted);
        f64 h_rej = reject_more ? hs * 0.1 : o_min(hs * fac, hs);
        if (ctl->fixed_step) {
            h_acc = h_wanted;
            h_rej = h_wanted;
        }
        h = accept ? h_acc : h_rej;
        reject_more = reject && reject_last;
        reject_last = reject;
        stats[0] = stats[0] + (accept ? 1.0 : 0.0);
        stats[1] = stats[1] + (reject ? 1.0 : 0.0);
        active = t < t_end;
    }
#undef KST
    *t_io = t;
    *h_io = h;
}

static void o_alloc(OWork *w, const OMech *m) {
    w->lu = (f64 *)malloc(sizeof(f64) * (m->nnz + 7 * (size_t)m->n));
    w->kstage = w->lu + m->nnz;
    w->fcn = w->kstage + 4 * (size_t)m->n;
    w->ytmp = w->fcn + m->n;
}

ORACLE_API void oracle_integrate(const void *mech, const SimkCaps *caps, const KinMethod *method, const KinControl *ctl,
                                 const f64 *k, const f64 *src, f64 *y, f64 *t, f64 t_end, f64 *h, f64 *stats) {
    OMech m = o_view(mech, caps);
    OWork w;
    o_alloc(&w, &m);
    o_integrate(&m, method, ctl, k, src, y, t, t_end, h, stats, &w, 0);
    free(w.lu);
}

/* One lane of dbd_advance_group<f64> (src/sim/dbd_lanes.h:28-78) with
 * kin_advance's pulse segments (kin_solver.h:322-347): local time [0, duration],
 * restart at h_start after each pulse, counters accumulated into the block. */
static void o_advance_lane(const OMech *m, const KinMethod *method, const KinControl *ctl, const OracleLaunch *L,
                           u32 i, OWork *w, f64 *y, f64 *k, f64 *src, f64 *pulse, int store32) {
    const u32 n = m->n, p = L->padded;
    for (u32 s = 0; s < n; ++s) {
        y[s] = L->y[(size_t)s * p + i];
        src[s] = L->src[(size_t)s * p + i];
        pulse[s] = L->pulse[(size_t)s * p + i];
    }
    for (u32 r = 0; r < m->R; ++r) k[r] = L->k[(size_t)r * p + i];
    f64 h = L->h[i];
    const f64 period = L->period[i], first = L->pulse_first[i], count = L->pulse_count[i];
    f64 stats[4] = {0.0, 0.0, 0.0, L->failed[i]};
    const f64 duration = L->duration;

    f64 t = 0.0, applied = 0.0;
    for (;;) {
        int pending = applied < count;
        f64 next_pulse = o_fmadd(applied, period, first);
        f64 seg_end = pending ? o_min(next_pulse, duration) : duration;
        o_integrate(m, method, ctl, k, src, y, &t, seg_end, &h, stats, w, store32);
        int alive = !(stats[3] > 0.0);
        int hit = pending && alive && (t >= seg_end);
        if (hit) {
            for (u32 s = 0; s < n; ++s) y[s] = y[s] + pulse[s];
            applied = applied + 1.0;
            stats[2] = stats[2] + 1.0;
            h = ctl->h_start;
        }
        int done = ((t >= duration) && (applied >= count)) || !alive;
        if (done) break;
    }
    for (u32 s = 0; s < n; ++s) L->y[(size_t)s * p + i] = y[s];
    L->h[i] = h;
    L->failed[i] = stats[3];
    L->steps[i] += stats[0];
    L->rejects[i] += stats[1];
    L->pulses[i] += stats[2];
}

typedef struct OPool {
    const OMech *m;
    const KinMethod *method;
    const KinControl *ctl;
    const OracleLaunch *L;
    const u32 *lanes;
    u32 count;
    u32 next;
    int store32;
} OPool;

static void *o_worker(void *arg) {
    OPool *pool = (OPool *)arg;
    OWork w;
    o_alloc(&w, pool->m);
    f64 *buf = (f64 *)malloc(sizeof(f64) * (3 * (size_t)pool->m->n + pool->m->R + 1));
    f64 *y = buf, *src = y + pool->m->n, *pulse = src + pool->m->n, *k = pulse + pool->m->n;
    for (;;) {
        u32 i = __atomic_fetch_add(&pool->next, 1u, __ATOMIC_RELAXED);
        if (i >= pool->count) break;
        u32 lane = pool->lanes ? pool->lanes[i] : i;
        o_advance_lane(pool->m, pool->method, pool->ctl, pool->L, lane, &w, y, k, src, pulse, pool->store32);
    }
    free(buf);
    free(w.lu);
    return 0;
}

static s32 o_run_pool(OPool *pool, u32 threads) {
    if (threads <= 1) {
        o_worker(pool);
        return 0;
    }
    pthread_t *tid = (pthread_t *)calloc(threads, sizeof(pthread_t));
    for (u32 t = 0; t < threads; ++t) pthread_create(&tid[t], 0, o_worker, pool);
    for (u32 t = 0; t < threads; ++t) pthread_join(tid[t], 0);
    free(tid);
    return 0;
}

/* Advances the listed lanes (all of [0, lane_count) when lanes is null). Lanes
 * are independent, so the result does not depend on `threads`. */
ORACLE_API s32 oracle_advance(const void *mech, const SimkCaps *caps, const KinMethod *method, const KinControl *ctl,
                              const OracleLaunch *L, const u32 *lanes, u32 lane_count, u32 threads) {
    OMech m = o_view(mech, caps);
    OPool pool = {&m, method, ctl, L, lanes, lane_count, 0, 0};
    return o_run_pool(&pool, threads);
}

/* The same advance in the FP32 factor-storage form (GpuKinLaunch precision 3;
 * not a reference path): the factor of every attempt is o_factor_rows with
 * store32, everything else is the FP64 advance above. */
ORACLE_API s32 oracle_advance_store32(const void *mech, const SimkCaps *caps, const KinMethod *method,
                                      const KinControl *ctl, const OracleLaunch *L, const u32 *lanes, u32 lane_count,
                                      u32 threads) {
    OMech m = o_view(mech, caps);
    OPool pool = {&m, method, ctl, L, lanes, lane_count, 0, 1};
    return o_run_pool(&pool, threads);
}

/* ---- f32 delta lanes (GpuKinLaunch precision 1) ----------------------------
 * dbd_advance_group<f32, KinOpsDefault, Delta = true> (src/sim/dbd_lanes.h:28-78)
 * around kin_advance / kin_integrate<f32, Delta> (kin_solver.h:189-347), one
 * scalar lane:
 *  - every input is rounded to f32 on load (lane_load_f64, base_lane.h:101);
 *    method and control constants are rounded by lane_splat (base_lane.h:98);
 *  - y holds the increment over base = f32(y_f64), starting at 0, and every
 *    rate and Jacobian is evaluated at base + y (kin_solver.h:221-229, 245-250);
 *  - the error scale uses |base + y| and |base + y_new| (:268-270) and the
 *    clamp keeps base + y >= 0 (:288-291);
 *  - the increment is added to the f64 storage, which is then clamped at 0
 *    (dbd_lanes.h:58-67); h and failed are stored widened, the per-call
 *    counters added. */
static inline f32 of_min(f32 a, f32 b) { return a < b ? a : b; }
static inline f32 of_max(f32 a, f32 b) { return a > b ? a : b; }
static inline f32 of_abs(f32 a) { return a < 0.0f ? -a : a; }
static inline int of_is_finite(f32 a) { f32 d = a - a; return d == d; }

static void of_fun(const OMech *m, const f32 *k, const f32 *src, const f32 *y, f32 *f) {
    for (u32 s = 0; s < m->n; ++s) f[s] = src[s];
    for (u32 r = 0; r < m->R; ++r) {
        const KinReaction *rx = m->rx + r;
        f32 rate = k[r];
        for (u32 p = 0; p < rx->reactant_count; ++p) rate = rate * y[rx->reactant[p]];
        for (u32 e = 0; e < rx->net_count; ++e) {
            u32 s = rx->net_species[e];
            f[s] = fmaf((f32)(f64)rx->net_coef[e], rate, f[s]);
        }
    }
}

static void of_neg_jac(const OMech *m, const f32 *k, const f32 *y, f32 *a) {
    for (u32 i = 0; i < m->nnz; ++i) a[i] = 0.0f;
    for (u32 r = 0; r < m->R; ++r) {
        const KinReaction *rx = m->rx + r;
        for (u32 p = 0; p < rx->reactant_count; ++p) {
            f32 d = k[r];
            for (u32 q = 0; q < rx->reactant_count; ++q)
                if (q != p) d = d * y[rx->reactant[q]];
            for (u32 e = 0; e < rx->net_count; ++e) {
                u32 idx = rx->jac_index[p][e];
                a[idx] = fmaf((f32)(-(f64)rx->net_coef[e]), d, a[idx]);
            }
        }
    }
}

static void of_factor(const OMech *m, f32 ghinv, f32 *a) {
    for (u32 i = 0; i < m->n; ++i) {
        for (u32 d = m->div_begin[i]; d < m->div_begin[i + 1]; ++d) {
            KinLuDiv div = m->lu_div[d];
            f32 l = a[div.ik] * a[div.kk];
            a[div.ik] = l;
            for (u32 s = div.sub_begin; s < div.sub_end; ++s) {
                KinLuSub sub = m->lu_sub[s];
                a[sub.ij] = fmaf(-l, a[sub.kj], a[sub.ij]);
            }
        }
        u32 ii = m->diag[i];
        a[ii] = 1.0f / (a[ii] + ghinv);
    }
}

static void of_solve(const OMech *m, const f32 *a, f32 *x) {
    const u32 n = m->n;
    for (u32 i = 0; i < n; ++i)
        for (u32 e = m->lower_begin[i]; e < m->lower_begin[i + 1]; ++e)
            x[i] = fmaf(-a[m->lower[e].index], x[m->lower[e].col], x[i]);
    for (u32 ii = n; ii-- > 0;) {
        for (u32 e = m->upper_begin[ii]; e < m->upper_begin[ii + 1]; ++e)
            x[ii] = fmaf(-a[m->upper[e].index], x[m->upper[e].col], x[ii]);
        x[ii] = x[ii] * a[m->diag[ii]];
    }
}

/* lane_cbrt_clamped<f32> and kin_step_factor<f32>: the f64 forms with every
 * constant rounded to f32 (base_lane.h:123-135, kin_solver.h:169-180). */
ORACLE_API f32 oracle_cbrt_clamped_f32(f32 x) {
    x = of_min(of_max(x, (f32)1e-4), (f32)1e4);
    f32 r4 = sqrtf(sqrtf(x));
    f32 r16 = sqrtf(sqrtf(r4));
    f32 r64 = sqrtf(sqrtf(r16));
    f32 g = r4 * r16 * r64;
    const f32 two = 2.0f, third = (f32)(1.0 / 3.0);
    g = (two * g + x / (g * g)) * third;
    g = (two * g + x / (g * g)) * third;
    return g;
}

ORACLE_API f32 oracle_step_factor_f32(f32 err, u32 order) {
    f32 e = of_min(of_max(err, (f32)1e-4), (f32)1e4);
    f32 root;
    if (order == 2) root = sqrtf(e);
    else if (order == 4) root = sqrtf(sqrtf(e));
    else root = oracle_cbrt_clamped_f32(e);
    f32 fac = (f32)0.9 / root;
    return of_min(of_max(fac, (f32)0.2), (f32)6.0);
}

typedef struct OWork32 {
    f32 *lu, *kstage, *fcn, *ytmp, *yeval; /* nnz, 4*n, n, n, n */
} OWork32;

static void of_integrate(const OMech *m, const KinMethod *method, const KinControl *ctl, const f32 *k, const f32 *src,
                         const f32 *base, f32 *y, f32 *t_io, f32 t_end, f32 *h_io, f32 *stats, OWork32 *w) {
    const u32 n = m->n;
    const f32 one = 1.0f, zero = 0.0f;
    const f32 h_min = (f32)ctl->h_min, rtol = (f32)ctl->rtol, atol = (f32)ctl->atol;
    const f32 inv_n = (f32)(1.0 / (f64)n);
    f32 t = *t_io, h = *h_io;
    f32 *lu = w->lu, *fcn = w->fcn, *ytmp = w->ytmp, *yeval = w->yeval;
#define KST(st, s) w->kstage[(st) * n + (s)]

    int failed = stats[3] > zero;
    int active = (t < t_end) && !failed;
    int reject_last = 0, reject_more = 0;
    h = h > zero ? h : (f32)ctl->h_start;

    u32 iterations = 0;
    while (active) {
        if (++iterations > ctl->max_steps) {
            stats[3] = one;
            break;
        }
        f32 remaining = t_end - t;
        f32 h_cap = remaining;
        if (ctl->h_max > 0.0) h_cap = of_min(h_cap, (f32)ctl->h_max);
        f32 h_wanted = of_max(h, h_min);
        int last = h_wanted >= remaining;
        f32 hs = last ? remaining : of_min(h_wanted, h_cap);

        for (u32 s = 0; s < n; ++s) yeval[s] = base[s] + y[s];
        of_fun(m, k, src, yeval, fcn);
        of_neg_jac(m, k, yeval, lu);
        f32 ghinv = one / (hs * (f32)method->gamma0);
        of_factor(m, ghinv, lu);
        f32 hinv = one / hs;

        for (u32 st = 0; st < method->stages; ++st) {
            u32 packed = st * (st - 1) / 2;
            if (st > 0 && method->new_f[st]) {
                for (u32 s = 0; s < n; ++s) ytmp[s] = y[s];
                for (u32 j = 0; j < st; ++j) {
                    f64 a = method->a[packed + j];
                    if (a == 0.0) continue;
                    for (u32 s = 0; s < n; ++s) ytmp[s] = fmaf((f32)a, KST(j, s), ytmp[s]);
                }
                for (u32 s = 0; s < n; ++s) yeval[s] = base[s] + ytmp[s];
                of_fun(m, k, src, yeval, fcn);
            }
            for (u32 s = 0; s < n; ++s) KST(st, s) = fcn[s];
            for (u32 j = 0; j < st; ++j) {
                f64 c = method->c[packed + j];
                if (c == 0.0) continue;
                f32 hc = (f32)c * hinv;
                for (u32 s = 0; s < n; ++s) KST(st, s) = fmaf(hc, KST(j, s), KST(st, s));
            }
            of_solve(m, lu, &KST(st, 0));
        }

        f32 err = zero;
        for (u32 s = 0; s < n; ++s) {
            f32 ynew = y[s], yerr = zero;
            for (u32 st = 0; st < method->stages; ++st) {
                if (method->m[st] != 0.0) ynew = fmaf((f32)method->m[st], KST(st, s), ynew);
                if (method->e[st] != 0.0) yerr = fmaf((f32)method->e[st], KST(st, s), yerr);
            }
            f32 scale = fmaf(rtol, of_max(of_abs(base[s] + y[s]), of_abs(base[s] + ynew)), atol);
            f32 q = yerr / scale;
            err = fmaf(q, q, err);
            ytmp[s] = ynew;
        }
        err = sqrtf(err * inv_n);
        if (!of_is_finite(err)) err = (f32)1e10;

        f32 fac = oracle_step_factor_f32(err, method->order);
        int ok = (err <= one) || (hs <= h_min);
        if (ctl->fixed_step) {
            ok = 1;
            fac = one;
        }
        int accept = ok, reject = !ok;

        if (accept) {
            for (u32 s = 0; s < n; ++s) {
                f32 ynew = ytmp[s];
                if (ctl->clamp_negative) ynew = of_max(ynew, -base[s]);
                y[s] = ynew;
            }
            t = last ? t_end : t + hs;
        }
        f32 h_acc = hs * fac;
        if (reject_last) h_acc = of_min(h_acc, hs);
        if (last && fac >= one) h_acc = of_max(h_acc, h_wanted);
        f32 h_rej = reject_more ? hs * (f32)0.1 : of_min(hs * fac, hs);
        if (ctl->fixed_step) {
            h_acc = h_wanted;
            h_rej = h_wanted;
        }
        h = accept ? h_acc : h_rej;
        reject_more = reject && reject_last;
        reject_last = reject;
        stats[0] = stats[0] + (accept ? one : zero);
        stats[1] = stats[1] + (reject ? one : zero);
        active = t < t_end;
    }
#undef KST
    *t_io = t;
    *h_io = h;
}

static void of_advance_lane(const OMech *m, const KinMethod *method, const KinControl *ctl, const OracleLaunch *L,
                            u32 i, OWork32 *w, f32 *buf) {
    const u32 n = m->n, p = L->padded;
    f32 *base = buf, *y = base + n, *src = y + n, *pulse = src + n, *k = pulse + n;
    for (u32 s = 0; s < n; ++s) {
        base[s] = (f32)L->y[(size_t)s * p + i];
        y[s] = 0.0f;
        src[s] = (f32)L->src[(size_t)s * p + i];
        pulse[s] = (f32)L->pulse[(size_t)s * p + i];
    }
    for (u32 r = 0; r < m->R; ++r) k[r] = (f32)L->k[(size_t)r * p + i];
    f32 h = (f32)L->h[i];
    const f32 period = (f32)L->period[i], first = (f32)L->pulse_first[i], count = (f32)L->pulse_count[i];
    f32 stats[4] = {0.0f, 0.0f, 0.0f, (f32)L->failed[i]};
    const f32 duration = (f32)L->duration;

    f32 t = 0.0f, applied = 0.0f;
    for (;;) {
        int pending = applied < count;
        f32 next_pulse = fmaf(applied, period, first);
        f32 seg_end = pending ? of_min(next_pulse, duration) : duration;
        of_integrate(m, method, ctl, k, src, base, y, &t, seg_end, &h, stats, w);
        int alive = !(stats[3] > 0.0f);
        int hit = pending && alive && (t >= seg_end);
        if (hit) {
            for (u32 s = 0; s < n; ++s) y[s] = y[s] + pulse[s];
            applied = applied + 1.0f;
            stats[2] = stats[2] + 1.0f;
            h = (f32)ctl->h_start;
        }
        int done = ((t >= duration) && (applied >= count)) || !alive;
        if (done) break;
    }
    for (u32 s = 0; s < n; ++s) {
        f64 *dst = L->y + (size_t)s * p + i;
        *dst += (f64)y[s];
        if (ctl->clamp_negative) *dst = *dst > 0.0 ? *dst : 0.0;
    }
    L->h[i] = (f64)h;
    L->failed[i] = (f64)stats[3];
    L->steps[i] += (f64)stats[0];
    L->rejects[i] += (f64)stats[1];
    L->pulses[i] += (f64)stats[2];
}

typedef struct OPool32 {
    const OMech *m;
    const KinMethod *method;
    const KinControl *ctl;
    const OracleLaunch *L;
    const u32 *lanes;
    u32 count;
    u32 next;
} OPool32;

static void *of_worker(void *arg) {
    OPool32 *pool = (OPool32 *)arg;
    const OMech *m = pool->m;
    OWork32 w;
    w.lu = (f32 *)malloc(sizeof(f32) * (m->nnz + 8 * (size_t)m->n));
    w.kstage = w.lu + m->nnz;
    w.fcn = w.kstage + 4 * (size_t)m->n;
    w.ytmp = w.fcn + m->n;
    w.yeval = w.ytmp + m->n;
    f32 *buf = (f32 *)malloc(sizeof(f32) * (4 * (size_t)m->n + m->R + 1));
    for (;;) {
        u32 i = __atomic_fetch_add(&pool->next, 1u, __ATOMIC_RELAXED);
        if (i >= pool->count) break;
        of_advance_lane(m, pool->method, pool->ctl, pool->L, pool->lanes ? pool->lanes[i] : i, &w, buf);
    }
    free(buf);
    free(w.lu);
    return 0;
}

ORACLE_API s32 oracle_advance_delta32(const void *mech, const SimkCaps *caps, const KinMethod *method,
                                      const KinControl *ctl, const OracleLaunch *L, const u32 *lanes, u32 lane_count,
                                      u32 threads) {
    OMech m = o_view(mech, caps);
    OPool32 pool = {&m, method, ctl, L, lanes, lane_count, 0};
    if (threads <= 1) {
        of_worker(&pool);
        return 0;
    }
    pthread_t *tid = (pthread_t *)calloc(threads, sizeof(pthread_t));
    for (u32 t = 0; t < threads; ++t) pthread_create(&tid[t], 0, of_worker, &pool);
    for (u32 t = 0; t < threads; ++t) pthread_join(tid[t], 0);
    free(tid);
    return 0;
}

/* ---- operator-split upwind advection -------------------------------------
 * nft_transport's flux-form row update (src/sim/nft.cpp:1055-1081): per
 * substep, node j takes x_j + c_j (x_{j-1} - x_j) from the OLD neighbour
 * value (the sweep runs outlet -> inlet, nft.cpp:1058-1060), node 0 takes
 * x_0 - c_0 x_0 (inflow delta 0). Before each substep the outlet row is added
 * to outlet_sum (nft.cpp:1056). x: [field][stride] f32, c: [nodes]. */
ORACLE_API void oracle_upwind(f32 *x, const f32 *c, u32 nodes, u32 fields, u32 stride, u32 substeps, f32 *outlet_sum) {
    for (u32 s = 0; s < substeps; ++s) {
        for (u32 f = 0; f < fields; ++f) {
            f32 *row = x + (size_t)f * stride;
            if (outlet_sum) outlet_sum[f] = outlet_sum[f] + row[nodes - 1];
            for (u32 j = nodes; j-- > 0;) {
                f32 v = row[j];
                if (j > 0) row[j] = fmaf(c[j], row[j - 1] - v, v);
                else row[j] = fmaf(-c[j], v, v);
            }
        }
    }
}

/* The same fused update in f64 on the chemistry state, with an inflow value
 * per field (node 0: x_0 + c_0 (inflow - x_0); inflow 0 gives the reference's
 * x_0 - c_0 x_0 bit for bit). */
ORACLE_API void oracle_upwind_f64(f64 *x, const f64 *c, u32 nodes, u32 fields, u32 stride, u32 substeps,
                                  const f64 *inflow) {
    for (u32 s = 0; s < substeps; ++s)
        for (u32 f = 0; f < fields; ++f) {
            f64 *row = x + (size_t)f * stride;
            for (u32 j = nodes; j-- > 0;) {
                f64 v = row[j];
                f64 up = j > 0 ? row[j - 1] : (inflow ? inflow[f] : 0.0);
                row[j] = o_fmadd(c[j], up - v, v);
            }
        }
}

/* dbd_flow_exchange (src/sim/dbd.cpp:465-483): zones processed back to front,
 * zone i relaxes toward zone i-1's old state (zone 0 toward the feed) with
 * y += a * (up - y), the product and the sum rounded separately. y: [species]
 * rows of stride `stride`. */
ORACLE_API void oracle_plugflow_f64(f64 *y, u32 stride, u32 species, u32 zones, f64 a, const f64 *feed) {
    for (u32 i = zones; i-- > 0;)
        for (u32 s = 0; s < species; ++s) {
            f64 *v = y + (size_t)s * stride + i;
            f64 up = i == 0 ? feed[s] : y[(size_t)s * stride + i - 1];
            *v += a * (up - *v);
        }
}
