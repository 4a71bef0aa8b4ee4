// f32 delta lanes (GpuKinLaunch precision 1): the reference's f32 lane
// formulation, dbd_advance_group<f32, KinOpsDefault, Delta = true>
// (src/sim/dbd_lanes.h:28-78) around kin_advance / kin_integrate<f32, Delta>
// (src/sim/kinetics/kin_solver.h:189-347), one cell per lane.
//
// Every input is rounded to f32 on load (lane_load_f64, base_lane.h:101) and
// every method/control constant by lane_splat (base_lane.h:98). The lane's y
// is the increment over base = f32(y_f64), starting at 0; rates and the
// Jacobian are evaluated at base + y; the error scale takes |base + y| and
// |base + y_new|; the clamp keeps base + y >= 0. At the end the increment is
// added to the f64 storage, which is clamped at 0; h and failed are stored
// widened and the per-call counters added. Arithmetic is IEEE f32 with the
// fused multiply-adds written out (__builtin_fmaf), correctly rounded
// division and square root, and the reference's select-form min/max.
//
// Table walk over the device tables: every per-lane vector lives in f32
// workspace slots ([wave][slot][64] f32, one coalesced 256-byte line per
// slot access).
#pragma once

#include "device_kin.h"

// Method and control in the lane type T (lane_splat's rounding for f32),
// plus the f64 tests the reference makes on the unrounded values (a == 0,
// c == 0, h_max > 0, ...).
template <class T>
struct SimkTConsts {
    T a[6], c[6], m[4], e[4], gamma0;
    simk_u32 a_nz, c_nz, m_nz, e_nz;   // bit i: coefficient i is non-zero
    simk_u32 stages, order, new_f;     // new_f: bit st
    T rtol, atol, h_min, h_max, h_start, inv_n;
    simk_u32 h_max_pos, max_steps, fixed_step, clamp_negative;
};
typedef SimkTConsts<float> SimkD32Consts;

SIMK_DEV float simk_fminf_(float a, float b) { return a < b ? a : b; }
SIMK_DEV float simk_fmaxf_(float a, float b) { return a > b ? a : b; }
SIMK_DEV float simk_fabsf_(float a) { return a < 0.0f ? -a : a; }
SIMK_DEV bool simk_finitef_(float a) {
    float d = a - a;
    return d == d;
}

// lane_cbrt_clamped<f32> (base_lane.h:123-135)
SIMK_DEV float simk_cbrt_clamped_f32(float x) {
    x = simk_fminf_(simk_fmaxf_(x, (float)1e-4), (float)1e4);
    float r4 = __builtin_sqrtf(__builtin_sqrtf(x));
    float r16 = __builtin_sqrtf(__builtin_sqrtf(r4));
    float r64 = __builtin_sqrtf(__builtin_sqrtf(r16));
    float g = r4 * r16 * r64;
    const float two = 2.0f, third = (float)(1.0 / 3.0);
    g = (two * g + x / (g * g)) * third;
    g = (two * g + x / (g * g)) * third;
    return g;
}

// kin_step_factor<f32> (kin_solver.h:169-180)
SIMK_DEV float simk_step_factor_f32(float err, simk_u32 order) {
    float e = simk_fminf_(simk_fmaxf_(err, (float)1e-4), (float)1e4);
    float root;
    if (order == 2) root = __builtin_sqrtf(e);
    else if (order == 4) root = __builtin_sqrtf(__builtin_sqrtf(e));
    else root = simk_cbrt_clamped_f32(e);
    float fac = (float)0.9 / root;
    return simk_fminf_(simk_fmaxf_(fac, (float)0.2), (float)6.0);
}

// Workspace slots (f32) of one lane.
struct SimkD32Ws {
    SimkRows32 w;
    simk_u32 n, R, nnz;
    // [0, nnz) LU | base | y | yeval | fcn | ytmp | K0..K3 | src | pulse | k
    SIMK_DEV simk_u32 base() const { return nnz; }
    SIMK_DEV simk_u32 y() const { return nnz + n; }
    SIMK_DEV simk_u32 yeval() const { return nnz + 2 * n; }
    SIMK_DEV simk_u32 fcn() const { return nnz + 3 * n; }
    SIMK_DEV simk_u32 ytmp() const { return nnz + 4 * n; }
    SIMK_DEV simk_u32 kst(simk_u32 st) const { return nnz + (5 + st) * n; }
    SIMK_DEV simk_u32 src() const { return nnz + 9 * n; }
    SIMK_DEV simk_u32 pulse() const { return nnz + 10 * n; }
    SIMK_DEV simk_u32 k() const { return nnz + 11 * n; }
    SIMK_DEV float ld(simk_u32 i) const { return w.ld(i); }
    SIMK_DEV void st(simk_u32 i, float v) const { w.st(i, v); }
};

SIMK_DEV simk_u32 simk_d32_slots(simk_u32 n, simk_u32 R, simk_u32 nnz) { return nnz + 11 * n + R; }

// kin_fun<f32> at the point in slots [pt, pt + n) into slots [f, f + n).
SIMK_DEV void simk_d32_fun(const SimkDevTables *T, const SimkD32Ws &W, simk_u32 pt, simk_u32 f) {
    for (simk_u32 s = 0; s < W.n; ++s) W.st(f + s, W.ld(W.src() + s));
    for (simk_u32 r = 0; r < W.R; ++r) {
        float rate = W.ld(W.k() + r);
        const simk_u32 rc = T->reactant_count[r];
        for (simk_u32 p = 0; p < rc; ++p) rate = rate * W.ld(pt + T->reactant[r * 3 + p]);
        const simk_u32 nc = T->net_count[r];
        for (simk_u32 e = 0; e < nc; ++e) {
            const simk_u32 s = T->net_species[r * 6 + e];
            W.st(f + s, __builtin_fmaf((float)T->net_coef[r * 6 + e], rate, W.ld(f + s)));
        }
    }
}

// kin_neg_jac<f32> then kin_factor<f32> in the LU slots.
SIMK_DEV void simk_d32_jac_factor(const SimkDevTables *T, const SimkD32Ws &W, simk_u32 pt, float ghinv) {
    for (simk_u32 i = 0; i < W.nnz; ++i) W.st(i, 0.0f);
    for (simk_u32 r = 0; r < W.R; ++r) {
        const simk_u32 rc = T->reactant_count[r], nc = T->net_count[r];
        for (simk_u32 p = 0; p < rc; ++p) {
            float d = W.ld(W.k() + r);
            for (simk_u32 q = 0; q < rc; ++q)
                if (q != p) d = d * W.ld(pt + T->reactant[r * 3 + q]);
            for (simk_u32 e = 0; e < nc; ++e) {
                const simk_u32 idx = T->jac_index[(r * 3 + p) * 6 + e];
                W.st(idx, __builtin_fmaf((float)(-T->net_coef[r * 6 + e]), d, W.ld(idx)));
            }
        }
    }
    for (simk_u32 i = 0; i < W.n; ++i) {
        for (simk_u32 d = T->div_begin[i]; d < T->div_begin[i + 1]; ++d) {
            const unsigned short *dv = T->lu_div + 4 * d;
            const float l = W.ld(dv[0]) * W.ld(dv[1]);
            W.st(dv[0], l);
            for (simk_u32 s = dv[2]; s < dv[3]; ++s) {
                const simk_u32 ij = T->lu_sub[2 * s], kj = T->lu_sub[2 * s + 1];
                W.st(ij, __builtin_fmaf(-l, W.ld(kj), W.ld(ij)));
            }
        }
        const simk_u32 ii = T->diag[i];
        W.st(ii, 1.0f / (W.ld(ii) + ghinv));
    }
}

// kin_solve<f32> on slots [x, x + n).
SIMK_DEV void simk_d32_solve(const SimkDevTables *T, const SimkD32Ws &W, simk_u32 x) {
    for (simk_u32 i = 0; i < W.n; ++i)
        for (simk_u32 e = T->lower_begin[i]; e < T->lower_begin[i + 1]; ++e)
            W.st(x + i, __builtin_fmaf(-W.ld(T->lower[2 * e]), W.ld(x + T->lower[2 * e + 1]), W.ld(x + i)));
    for (simk_u32 ii = W.n; ii-- > 0;) {
        for (simk_u32 e = T->upper_begin[ii]; e < T->upper_begin[ii + 1]; ++e)
            W.st(x + ii, __builtin_fmaf(-W.ld(T->upper[2 * e]), W.ld(x + T->upper[2 * e + 1]), W.ld(x + ii)));
        W.st(x + ii, W.ld(x + ii) * W.ld(T->diag[ii]));
    }
}

// One step attempt of kin_integrate<f32, Delta>'s loop body (kin_solver.h:216-311).
SIMK_DEV void simk_d32_attempt(const SimkDevTables *T, const SimkD32Ws &W, const SimkD32Consts &C, float &t,
                               const float t_end, float &h, float &steps, float &rejects, bool &reject_last,
                               bool &reject_more) {
    const simk_u32 n = W.n;
    const float one = 1.0f;
    const float remaining = t_end - t;
    float h_cap = remaining;
    if (C.h_max_pos) h_cap = simk_fminf_(h_cap, C.h_max);
    const float h_wanted = simk_fmaxf_(h, C.h_min);
    const bool last = h_wanted >= remaining;
    const float hs = last ? remaining : simk_fminf_(h_wanted, h_cap);

    for (simk_u32 s = 0; s < n; ++s) W.st(W.yeval() + s, W.ld(W.base() + s) + W.ld(W.y() + s));
    simk_d32_fun(T, W, W.yeval(), W.fcn());
    const float ghinv = one / (hs * C.gamma0);
    simk_d32_jac_factor(T, W, W.yeval(), ghinv);
    const float hinv = one / hs;

    for (simk_u32 st = 0; st < C.stages; ++st) {
        const simk_u32 packed = st * (st - 1) / 2;
        if (st > 0 && ((C.new_f >> st) & 1u)) {
            for (simk_u32 s = 0; s < n; ++s) W.st(W.ytmp() + s, W.ld(W.y() + s));
            for (simk_u32 j = 0; j < st; ++j) {
                if (!((C.a_nz >> (packed + j)) & 1u)) continue;
                const float a = C.a[packed + j];
                for (simk_u32 s = 0; s < n; ++s)
                    W.st(W.ytmp() + s, __builtin_fmaf(a, W.ld(W.kst(j) + s), W.ld(W.ytmp() + s)));
            }
            for (simk_u32 s = 0; s < n; ++s) W.st(W.yeval() + s, W.ld(W.base() + s) + W.ld(W.ytmp() + s));
            simk_d32_fun(T, W, W.yeval(), W.fcn());
        }
        for (simk_u32 s = 0; s < n; ++s) W.st(W.kst(st) + s, W.ld(W.fcn() + s));
        for (simk_u32 j = 0; j < st; ++j) {
            if (!((C.c_nz >> (packed + j)) & 1u)) continue;
            const float hc = C.c[packed + j] * hinv;
            for (simk_u32 s = 0; s < n; ++s)
                W.st(W.kst(st) + s, __builtin_fmaf(hc, W.ld(W.kst(j) + s), W.ld(W.kst(st) + s)));
        }
        simk_d32_solve(T, W, W.kst(st));
    }

    float err = 0.0f;
    for (simk_u32 s = 0; s < n; ++s) {
        float ynew = W.ld(W.y() + s), yerr = 0.0f;
        for (simk_u32 st = 0; st < C.stages; ++st) {
            const float kv = W.ld(W.kst(st) + s);
            if ((C.m_nz >> st) & 1u) ynew = __builtin_fmaf(C.m[st], kv, ynew);
            if ((C.e_nz >> st) & 1u) yerr = __builtin_fmaf(C.e[st], kv, yerr);
        }
        const float b = W.ld(W.base() + s);
        const float scale = __builtin_fmaf(
            C.rtol, simk_fmaxf_(simk_fabsf_(b + W.ld(W.y() + s)), simk_fabsf_(b + ynew)), C.atol);
        const float q = yerr / scale;
        err = __builtin_fmaf(q, q, err);
        W.st(W.ytmp() + s, ynew);
    }
    err = __builtin_sqrtf(err * C.inv_n);
    if (!simk_finitef_(err)) err = (float)1e10;

    float fac = simk_step_factor_f32(err, C.order);
    bool ok = (err <= one) || (hs <= C.h_min);
    if (C.fixed_step) {
        ok = true;
        fac = one;
    }
    if (ok) {
        for (simk_u32 s = 0; s < n; ++s) {
            float v = W.ld(W.ytmp() + s);
            if (C.clamp_negative) v = simk_fmaxf_(v, -W.ld(W.base() + s));
            W.st(W.y() + s, v);
        }
        t = last ? t_end : t + hs;
    }
    float h_acc = hs * fac;
    if (reject_last) h_acc = simk_fminf_(h_acc, hs);
    if (last && fac >= one) h_acc = simk_fmaxf_(h_acc, h_wanted);
    float h_rej = reject_more ? hs * (float)0.1 : simk_fminf_(hs * fac, hs);
    if (C.fixed_step) {
        h_acc = h_wanted;
        h_rej = h_wanted;
    }
    h = ok ? h_acc : h_rej;
    reject_more = !ok && reject_last;
    reject_last = !ok;
    steps = steps + (ok ? one : 0.0f);
    rejects = rejects + (ok ? 0.0f : one);
}

// Persistent lane-refilling advance in f32 delta lanes: the same phase
// machine as simk_persistent_body (one wave iteration = one attempt per lane
// that has one), in f32, with the delta load and store of dbd_lanes.h.
SIMK_DEV void simk_d32_body(const SimkKinArgs &A, const SimkD32Consts &C) {
    const SimkDevTables *T = A.tables;
    const simk_u32 wave = blockIdx.x, lane = threadIdx.x;
    const simk_u32 n = T->n, R = T->R;
    const simk_u32 row_bytes = A.padded * 8u;
    SimkD32Ws W;
    W.n = n;
    W.R = R;
    W.nnz = T->nnz;
    const simk_u32 slots = simk_d32_slots(n, R, T->nnz);
    W.w.init((const float *)A.ws + (simk_u64)wave * slots * SIMK_WAVE, lane * 4u, SIMK_WAVE * 4u,
             slots * SIMK_WAVE * 4u);
    SimkRows ys, ks, ss, ps;
    ys.init(A.y, 0, row_bytes, n * row_bytes);
    ks.init(A.k, 0, row_bytes, R * row_bytes);
    ss.init(A.src, 0, row_bytes, n * row_bytes);
    ps.init(A.pulse, 0, row_bytes, n * row_bytes);

    simk_u32 pos = 0, cell = 0, phase = SIMK_PH_CELL, it = 0, rep = 0;
    float h = 0.0f, t = 0.0f, applied = 0.0f, seg_end = 0.0f, period = 0.0f, first = 0.0f, count = 0.0f;
    float steps = 0.0f, rejects = 0.0f, pulses = 0.0f, failed = 0.0f;
    const float duration = (float)A.duration;
    bool reject_last = false, reject_more = false, fresh = !A.queue_first;
    pos = wave * SIMK_WAVE + lane;

    // dbd_advance_group's load (dbd_lanes.h:31-54) for the current cell.
    auto load = [&](bool first_call) {
        const simk_u32 voff = cell * 8u;
        ys.voff = ks.voff = ss.voff = ps.voff = voff;
        for (simk_u32 s = 0; s < n; ++s) {
            W.st(W.base() + s, (float)ys.ld(s));
            W.st(W.y() + s, 0.0f);
            if (first_call) {
                W.st(W.src() + s, (float)ss.ld(s));
                W.st(W.pulse() + s, (float)ps.ld(s));
            }
        }
        if (first_call)
            for (simk_u32 r = 0; r < R; ++r) W.st(W.k() + r, (float)ks.ld(r));
        SimkRows one;
        one.init(A.h, voff, 0, row_bytes);
        h = (float)one.ld(0);
        one.init(A.failed, voff, 0, row_bytes);
        failed = (float)one.ld(0);
        one.init(A.period, voff, 0, row_bytes);
        period = (float)one.ld(0);
        one.init(A.pulse_first, voff, 0, row_bytes);
        first = (float)one.ld(0);
        one.init(A.pulse_count, voff, 0, row_bytes);
        count = (float)one.ld(0);
        steps = rejects = pulses = 0.0f;
        t = 0.0f;
        applied = 0.0f;
    };
    // dbd_advance_group's store (dbd_lanes.h:56-77).
    auto store = [&]() {
        const simk_u32 voff = cell * 8u;
        ys.voff = voff;
        for (simk_u32 s = 0; s < n; ++s) {
            double v = ys.ld(s) + (double)W.ld(W.y() + s);
            if (C.clamp_negative) v = v > 0.0 ? v : 0.0;
            ys.st(s, v);
        }
        SimkRows one;
        one.init(A.h, voff, 0, row_bytes);
        one.st(0, (double)h);
        one.init(A.failed, voff, 0, row_bytes);
        one.st(0, (double)failed);
        one.init(A.steps, voff, 0, row_bytes);
        one.st(0, one.ld(0) + (double)steps);
        one.init(A.rejects, voff, 0, row_bytes);
        one.st(0, one.ld(0) + (double)rejects);
        one.init(A.pulses, voff, 0, row_bytes);
        one.st(0, one.ld(0) + (double)pulses);
    };

    for (;;) {
        while (phase != SIMK_PH_ATTEMPT && phase != SIMK_PH_EXIT) {
            if (phase == SIMK_PH_CELL) {
                if (!fresh) {
                    const simk_u64 ask = __builtin_amdgcn_read_exec();
                    const simk_u32 rank = __builtin_amdgcn_mbcnt_hi((simk_u32)(ask >> 32),
                                                                    __builtin_amdgcn_mbcnt_lo((simk_u32)ask, 0u));
                    simk_u32 base = 0;
                    if (rank == 0) base = atomicAdd(A.queue, (simk_u32)__builtin_popcountll(ask));
                    const simk_u32 leader = (simk_u32)__builtin_ctzll(ask);
                    base = (simk_u32)__builtin_amdgcn_ds_bpermute((int)(leader << 2), (int)base);
                    pos = base + rank;
                }
                fresh = false;
                if (pos >= A.lanes) {
                    phase = SIMK_PH_EXIT;
                    break;
                }
                cell = A.order ? A.order[pos] : pos;
                load(true);
                rep = 0;
                phase = SIMK_PH_SEGMENT;
            } else if (phase == SIMK_PH_SEGMENT) {
                // kin_advance's segment (kin_solver.h:327-330) and kin_integrate's entry (:204-208)
                const bool pending = applied < count;
                const float next_pulse = __builtin_fmaf(applied, period, first);
                seg_end = pending ? simk_fminf_(next_pulse, duration) : duration;
                h = h > 0.0f ? h : C.h_start;
                reject_last = reject_more = false;
                it = 0;
                phase = ((t < seg_end) && !(failed > 0.0f)) ? SIMK_PH_CHECK : SIMK_PH_END;
            } else if (phase == SIMK_PH_CHECK) {
                if (++it > C.max_steps) {
                    failed = 1.0f;
                    phase = SIMK_PH_END;
                } else {
                    phase = SIMK_PH_ATTEMPT;
                }
            } else {   // SIMK_PH_END (kin_solver.h:332-345)
                // The segment's pending flag, recomputed: applied is unchanged
                // since SEGMENT. A bool carried from SEGMENT to here across
                // the wave's divergent loop exits was not kept per lane by the
                // compiler (the lane-mask register was rewritten for lanes
                // waiting outside the loop), so no lane mask is carried.
                const bool pending = applied < count;
                const bool alive = !(failed > 0.0f);
                const bool hit = pending && alive && (t >= seg_end);
                if (hit) {
                    for (simk_u32 s = 0; s < n; ++s) W.st(W.y() + s, W.ld(W.y() + s) + W.ld(W.pulse() + s));
                    applied = applied + 1.0f;
                    pulses = pulses + 1.0f;
                    h = C.h_start;
                }
                const bool done = ((t >= duration) && (applied >= count || !pending)) || !alive;
                if (!done) {
                    phase = SIMK_PH_SEGMENT;
                } else {
                    store();
                    if (++rep < A.repeats) {
                        load(false);   // the next operator-split call starts from the stored state
                        phase = SIMK_PH_SEGMENT;
                    } else {
                        phase = SIMK_PH_CELL;
                    }
                }
            }
        }
        if (__builtin_amdgcn_ballot_w64(phase == SIMK_PH_ATTEMPT) == 0) break;
        if (phase == SIMK_PH_ATTEMPT) {
            simk_d32_attempt(T, W, C, t, seg_end, h, steps, rejects, reject_last, reject_more);
            phase = (t < seg_end) ? SIMK_PH_CHECK : SIMK_PH_END;
        }
    }
}

// ---- generated f32 delta kernels -------------------------------------------
//
// The same lane formulation with a generated policy P (straight-line f32 fun,
// -J/LU and solves for one mechanism): base and the increment y live in
// registers, the stage vectors K in f32 workspace slots, the LU in LDS and f32
// slots. Every operation is the one simk_d32_attempt performs, in the same
// order, so both kernels (and the oracle) give the same bits. P provides
// N, fun(pt, f), jac_factor(pt, ghinv), solve(x), kstore(st, x),
// kaxpy(j, c, v) (v = fma(c, K_j, v) per element), pulse(s), init and set_cell.

// The argument block of a generated f32 delta kernel.
struct SimkD32Args {
    SimkKinArgs A;
    SimkD32Consts C;
};

// One step attempt (kin_solver.h:216-311 in f32 delta lanes).
template <class P>
SIMK_DEV void simk_d32g_attempt(const P &p, const SimkD32Consts &C, const float (&base)[P::N], float (&y)[P::N],
                                float &t, const float t_end, float &h, float &steps, float &rejects,
                                bool &reject_last, bool &reject_more) {
    constexpr simk_u32 N = P::N;
    const float one = 1.0f;
    const float remaining = t_end - t;
    float h_cap = remaining;
    if (C.h_max_pos) h_cap = simk_fminf_(h_cap, C.h_max);
    const float h_wanted = simk_fmaxf_(h, C.h_min);
    const bool last = h_wanted >= remaining;
    const float hs = last ? remaining : simk_fminf_(h_wanted, h_cap);

    float f[N], x[N];
    {
        float ev[N];
#pragma unroll
        for (simk_u32 s = 0; s < N; ++s) ev[s] = base[s] + y[s];
        p.fun(ev, f);
        const float ghinv = one / (hs * C.gamma0);
        p.jac_factor(ev, ghinv);
    }
    const float hinv = one / hs;

#pragma unroll 1
    for (simk_u32 st = 0; st < C.stages; ++st) {
        const simk_u32 packed = st * (st - 1) / 2;
        if (st > 0 && ((C.new_f >> st) & 1u)) {
            float ev[N];
#pragma unroll
            for (simk_u32 s = 0; s < N; ++s) ev[s] = y[s];
            for (simk_u32 j = 0; j < st; ++j)
                if ((C.a_nz >> (packed + j)) & 1u) p.kaxpy(j, C.a[packed + j], ev);
#pragma unroll
            for (simk_u32 s = 0; s < N; ++s) ev[s] = base[s] + ev[s];
            p.fun(ev, f);
        }
#pragma unroll
        for (simk_u32 s = 0; s < N; ++s) x[s] = f[s];
        for (simk_u32 j = 0; j < st; ++j)
            if ((C.c_nz >> (packed + j)) & 1u) p.kaxpy(j, C.c[packed + j] * hinv, x);
        p.solve(x);
        if (st + 1 < C.stages) p.kstore(st, x);
    }

    // y_new and the error estimate: stages in ascending order per element; the
    // last stage's K is still in x.
    const simk_u32 S = C.stages;
    float ynew[N], yerr[N];
#pragma unroll
    for (simk_u32 s = 0; s < N; ++s) {
        ynew[s] = y[s];
        yerr[s] = 0.0f;
    }
    for (simk_u32 st = 0; st + 1 < S; ++st) {
        if ((C.m_nz >> st) & 1u) p.kaxpy(st, C.m[st], ynew);
        if ((C.e_nz >> st) & 1u) p.kaxpy(st, C.e[st], yerr);
    }
    const bool m_last = (C.m_nz >> (S - 1)) & 1u, e_last = (C.e_nz >> (S - 1)) & 1u;
    float err = 0.0f;
#pragma unroll
    for (simk_u32 s = 0; s < N; ++s) {
        float yn = ynew[s], ye = yerr[s];
        if (m_last) yn = __builtin_fmaf(C.m[S - 1], x[s], yn);
        if (e_last) ye = __builtin_fmaf(C.e[S - 1], x[s], ye);
        const float b = base[s];
        const float scale = __builtin_fmaf(C.rtol, simk_fmaxf_(simk_fabsf_(b + y[s]), simk_fabsf_(b + yn)), C.atol);
        const float q = ye / scale;
        err = __builtin_fmaf(q, q, err);
        ynew[s] = yn;
    }
    err = __builtin_sqrtf(err * C.inv_n);
    if (!simk_finitef_(err)) err = (float)1e10;

    float fac = simk_step_factor_f32(err, C.order);
    bool ok = (err <= one) || (hs <= C.h_min);
    if (C.fixed_step) {
        ok = true;
        fac = one;
    }
    if (ok) {
#pragma unroll
        for (simk_u32 s = 0; s < N; ++s) {
            float v = ynew[s];
            if (C.clamp_negative) v = simk_fmaxf_(v, -base[s]);
            y[s] = v;
        }
        t = last ? t_end : t + hs;
    }
    float h_acc = hs * fac;
    if (reject_last) h_acc = simk_fminf_(h_acc, hs);
    if (last && fac >= one) h_acc = simk_fmaxf_(h_acc, h_wanted);
    float h_rej = reject_more ? hs * (float)0.1 : simk_fminf_(hs * fac, hs);
    if (C.fixed_step) {
        h_acc = h_wanted;
        h_rej = h_wanted;
    }
    h = ok ? h_acc : h_rej;
    reject_more = !ok && reject_last;
    reject_last = !ok;
    steps = steps + (ok ? one : 0.0f);
    rejects = rejects + (ok ? 0.0f : one);
}

// Persistent lane-refilling advance with a generated policy: simk_d32_body's
// phase machine, load and store (dbd_lanes.h:31-77).
template <class P>
SIMK_DEV void simk_d32g_body(const SimkKinArgs &A, const SimkD32Consts &C) {
    constexpr simk_u32 N = P::N;
    const simk_u32 wave = blockIdx.x, lane = threadIdx.x;
    const simk_u32 row_bytes = A.padded * 8u;
    P p;
    p.init(A, wave, lane);
    SimkRows ys;
    ys.init(A.y, 0, row_bytes, N * row_bytes);

    simk_u32 pos = 0, cell = 0, phase = SIMK_PH_CELL, it = 0, rep = 0;
    float h = 0.0f, t = 0.0f, applied = 0.0f, seg_end = 0.0f, period = 0.0f, first = 0.0f, count = 0.0f;
    float steps = 0.0f, rejects = 0.0f, pulses = 0.0f, failed = 0.0f;
    float base[N], y[N];
    const float duration = (float)A.duration;
    bool reject_last = false, reject_more = false, fresh = !A.queue_first;
    pos = wave * SIMK_WAVE + lane;

    auto load = [&](bool first_call) {
        const simk_u32 voff = cell * 8u;
        ys.voff = voff;
#pragma unroll
        for (simk_u32 s = 0; s < N; ++s) {
            base[s] = (float)ys.ld(s);
            y[s] = 0.0f;
        }
        if (first_call) p.set_cell(A, cell);
        SimkRows one;
        one.init(A.h, voff, 0, row_bytes);
        h = (float)one.ld(0);
        one.init(A.failed, voff, 0, row_bytes);
        failed = (float)one.ld(0);
        one.init(A.period, voff, 0, row_bytes);
        period = (float)one.ld(0);
        one.init(A.pulse_first, voff, 0, row_bytes);
        first = (float)one.ld(0);
        one.init(A.pulse_count, voff, 0, row_bytes);
        count = (float)one.ld(0);
        steps = rejects = pulses = 0.0f;
        t = 0.0f;
        applied = 0.0f;
    };
    auto store = [&]() {
        const simk_u32 voff = cell * 8u;
        ys.voff = voff;
#pragma unroll
        for (simk_u32 s = 0; s < N; ++s) {
            double v = ys.ld(s) + (double)y[s];
            if (C.clamp_negative) v = v > 0.0 ? v : 0.0;
            ys.st(s, v);
        }
        SimkRows one;
        one.init(A.h, voff, 0, row_bytes);
        one.st(0, (double)h);
        one.init(A.failed, voff, 0, row_bytes);
        one.st(0, (double)failed);
        one.init(A.steps, voff, 0, row_bytes);
        one.st(0, one.ld(0) + (double)steps);
        one.init(A.rejects, voff, 0, row_bytes);
        one.st(0, one.ld(0) + (double)rejects);
        one.init(A.pulses, voff, 0, row_bytes);
        one.st(0, one.ld(0) + (double)pulses);
    };

    for (;;) {
        while (phase != SIMK_PH_ATTEMPT && phase != SIMK_PH_EXIT) {
            if (phase == SIMK_PH_CELL) {
                if (!fresh) {
                    const simk_u64 ask = __builtin_amdgcn_read_exec();
                    const simk_u32 rank = __builtin_amdgcn_mbcnt_hi((simk_u32)(ask >> 32),
                                                                    __builtin_amdgcn_mbcnt_lo((simk_u32)ask, 0u));
                    simk_u32 qbase = 0;
                    if (rank == 0) qbase = atomicAdd(A.queue, (simk_u32)__builtin_popcountll(ask));
                    const simk_u32 leader = (simk_u32)__builtin_ctzll(ask);
                    qbase = (simk_u32)__builtin_amdgcn_ds_bpermute((int)(leader << 2), (int)qbase);
                    pos = qbase + rank;
                }
                fresh = false;
                if (pos >= A.lanes) {
                    phase = SIMK_PH_EXIT;
                    break;
                }
                cell = A.order ? A.order[pos] : pos;
                load(true);
                rep = 0;
                phase = SIMK_PH_SEGMENT;
            } else if (phase == SIMK_PH_SEGMENT) {
                const bool pending = applied < count;
                const float next_pulse = __builtin_fmaf(applied, period, first);
                seg_end = pending ? simk_fminf_(next_pulse, duration) : duration;
                h = h > 0.0f ? h : C.h_start;
                reject_last = reject_more = false;
                it = 0;
                phase = ((t < seg_end) && !(failed > 0.0f)) ? SIMK_PH_CHECK : SIMK_PH_END;
            } else if (phase == SIMK_PH_CHECK) {
                if (++it > C.max_steps) {
                    failed = 1.0f;
                    phase = SIMK_PH_END;
                } else {
                    phase = SIMK_PH_ATTEMPT;
                }
            } else {   // SIMK_PH_END
                // The segment's pending flag, recomputed: applied is unchanged
                // since SEGMENT. A bool carried from SEGMENT to here across
                // the wave's divergent loop exits was not kept per lane by the
                // compiler (the lane-mask register was rewritten for lanes
                // waiting outside the loop), so no lane mask is carried.
                const bool pending = applied < count;
                const bool alive = !(failed > 0.0f);
                const bool hit = pending && alive && (t >= seg_end);
                if (hit) {
#pragma unroll
                    for (simk_u32 s = 0; s < N; ++s) y[s] = y[s] + p.pulse(s);
                    applied = applied + 1.0f;
                    pulses = pulses + 1.0f;
                    h = C.h_start;
                }
                const bool done = ((t >= duration) && (applied >= count || !pending)) || !alive;
                if (!done) {
                    phase = SIMK_PH_SEGMENT;
                } else {
                    store();
                    if (++rep < A.repeats) {
                        load(false);
                        phase = SIMK_PH_SEGMENT;
                    } else {
                        phase = SIMK_PH_CELL;
                    }
                }
            }
        }
        if (__builtin_amdgcn_ballot_w64(phase == SIMK_PH_ATTEMPT) == 0) break;
        if (phase == SIMK_PH_ATTEMPT) {
            simk_d32g_attempt(p, C, base, y, t, seg_end, h, steps, rejects, reject_last, reject_more);
            phase = (t < seg_end) ? SIMK_PH_CHECK : SIMK_PH_END;
        }
    }
}
