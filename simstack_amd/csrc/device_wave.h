// Cooperative mode (gpu_kin_set_mode(1)): one 64-lane wave per cell, the
// cell's state in LDS, and each phase of the step spread over the wave. This is
// the decomposition of the reference's warp kernel (KinWarpOps,
// src/gpu/kernels.cu:51-225). Results are bit-identical to one lane per cell,
// because every value is produced by the same operation sequence; only the
// independent elements are spread across lanes:
//   f       rates k_r * prod y in parallel over reactions; then each species
//           sums its contributions in reaction order (its own list)
//   -J      reactant derivatives in parallel over (reaction, reactant); then
//           each LU entry sums its contributions in (r, p, e) order
//   LU      rows and eliminations in order; the updates of one elimination
//           touch distinct entries, so they run in parallel
//   solve   column-oriented: forward by ascending column, backward by
//           descending column. Each x_i receives its updates in the row-wise
//           order of kin_solve
//   vectors stage combinations, y_new and the error terms in parallel over
//           species; the error norm is summed in species order by one lane
// Control (step size, accept/reject, pulses) is uniform over the lanes of a
// cell and computed identically by each of them. T = double (FP64 lanes) or
// float with DELTA (the reference's f32 delta lanes, dbd_lanes.h:21-72).
//
// Teams: L lanes per cell (8, 16, 32 or 64), 64 / L cells per wave, each team
// with its own LDS region, queue position and control flow. Teams of one wave
// diverge (different step counts), so teams narrower than the wave order
// their LDS traffic with wave-scope fences instead of workgroup barriers: the
// lanes of a wave issue LDS operations in order, and the fences keep the
// compiler from moving a team's loads and stores across a phase boundary.
#pragma once

#include "device_delta32.h"

// Contribution lists in the reference's accumulation order.
struct SimkWaveTables {
    simk_u32 n, R, nnz, pad_;
    const unsigned short *fun_begin;    // [n + 1]: species -> its (reaction, coef) entries
    const unsigned short *fun_r;
    const double *fun_coef;
    const unsigned short *jac_begin;    // [nnz + 1]: LU entry -> its (r * 3 + p, -coef) entries
    const unsigned short *jac_rp;
    const double *jac_coef;
    const unsigned short *lcol_begin;   // [n + 1]: column -> (row, LU index) of L below the diagonal
    const unsigned short *lcol_row;
    const unsigned short *lcol_idx;
    const unsigned short *ucol_begin;   // [n + 1]: column -> (row, LU index) of U above the diagonal
    const unsigned short *ucol_row;
    const unsigned short *ucol_idx;
};

SIMK_DEV double simk_t_fma(double a, double b, double c) { return __builtin_fma(a, b, c); }
SIMK_DEV float simk_t_fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
SIMK_DEV double simk_t_sqrt(double a) { return __builtin_sqrt(a); }
SIMK_DEV float simk_t_sqrt(float a) { return __builtin_sqrtf(a); }
SIMK_DEV double simk_t_step_factor(double err, simk_u32 order) { return simk_step_factor(err, order); }
SIMK_DEV float simk_t_step_factor(float err, simk_u32 order) { return simk_step_factor_f32(err, order); }
template <class T>
SIMK_DEV T simk_t_min(T a, T b) { return a < b ? a : b; }
template <class T>
SIMK_DEV T simk_t_max(T a, T b) { return a > b ? a : b; }
template <class T>
SIMK_DEV T simk_t_abs(T a) { return a < T(0) ? -a : a; }
template <class T>
SIMK_DEV bool simk_t_finite(T a) {
    T d = a - a;
    return d == d;
}

// LDS layout of one cell (units of T).
SIMK_DEV simk_u32 simk_wave_lds_elems(simk_u32 n, simk_u32 R, simk_u32 nnz) { return 13 * n + 5 * R + nnz + 2; }

template <class T>
struct SimkWaveState {
    T *y, *base, *ytmp, *yeval, *fcn, *K, *src, *pulse, *q, *k, *rate, *deriv, *lu, *scal;
    simk_u32 n, R, nnz;
    SIMK_DEV void init(T *lds, simk_u32 n_, simk_u32 R_, simk_u32 nnz_) {
        n = n_;
        R = R_;
        nnz = nnz_;
        y = lds;
        base = y + n;
        ytmp = base + n;
        yeval = ytmp + n;
        fcn = yeval + n;
        K = fcn + n;         // 4 n
        src = K + 4 * n;
        pulse = src + n;
        q = pulse + n;
        k = q + n;           // R
        rate = k + R;        // R
        deriv = rate + R;    // 3 R
        lu = deriv + 3 * R;  // nnz
        scal = lu + nnz;     // 2
    }
};

// LDS writes by one lane visible to the other lanes of its team: a workgroup
// barrier when the team is the whole wave (uniform control), else wave-scope
// fences around a wave barrier (a compiler-level ordering point; the wave
// executes its LDS operations in order).
template <simk_u32 L>
SIMK_DEV void simk_team_sync() {
    if constexpr (L == SIMK_WAVE) {
        __syncthreads();
    } else {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// f = src + sum_r net_r k_r prod y at pt (kin_fun, kin_solver.h:39-53).
template <class T, simk_u32 L>
SIMK_DEV void simk_wave_fun(const SimkDevTables *D, const SimkWaveTables *W, SimkWaveState<T> &S, const T *pt,
                            T *f) {
    const simk_u32 lane = threadIdx.x % L;
    for (simk_u32 r = lane; r < S.R; r += L) {
        T rate = S.k[r];
        const simk_u32 rc = D->reactant_count[r];
        for (simk_u32 p = 0; p < rc; ++p) rate = rate * pt[D->reactant[r * 3 + p]];
        S.rate[r] = rate;
    }
    simk_team_sync<L>();
    for (simk_u32 s = lane; s < S.n; s += L) {
        T v = S.src[s];
        for (simk_u32 e = W->fun_begin[s]; e < W->fun_begin[s + 1]; ++e)
            v = simk_t_fma((T)W->fun_coef[e], S.rate[W->fun_r[e]], v);
        f[s] = v;
    }
    simk_team_sync<L>();
}

// a = -J at pt, then the in-place LU of I/(h gamma) - J (kin_neg_jac,
// kin_factor, kin_solver.h:75-113).
template <class T, simk_u32 L>
SIMK_DEV void simk_wave_jac_factor(const SimkDevTables *D, const SimkWaveTables *W, SimkWaveState<T> &S,
                                   const T *pt, const T ghinv) {
    const simk_u32 lane = threadIdx.x % L;
    for (simk_u32 item = lane; item < 3 * S.R; item += L) {
        const simk_u32 r = item / 3, p = item % 3;
        const simk_u32 rc = D->reactant_count[r];
        if (p >= rc) continue;
        T d = S.k[r];
        for (simk_u32 q = 0; q < rc; ++q)
            if (q != p) d = d * pt[D->reactant[r * 3 + q]];
        S.deriv[item] = d;
    }
    simk_team_sync<L>();
    for (simk_u32 idx = lane; idx < S.nnz; idx += L) {
        T v = T(0);
        for (simk_u32 e = W->jac_begin[idx]; e < W->jac_begin[idx + 1]; ++e)
            v = simk_t_fma((T)W->jac_coef[e], S.deriv[W->jac_rp[e]], v);
        S.lu[idx] = v;
    }
    simk_team_sync<L>();
    for (simk_u32 i = 0; i < S.n; ++i) {
        for (simk_u32 d = D->div_begin[i]; d < D->div_begin[i + 1]; ++d) {
            const unsigned short *dv = D->lu_div + 4 * d;
            const simk_u32 ik = dv[0], kk = dv[1], s0 = dv[2], s1 = dv[3];
            const T l = S.lu[ik] * S.lu[kk];
            simk_team_sync<L>();   // every lane has read lu[ik] before it is overwritten
            if (lane == 0) S.lu[ik] = l;
            for (simk_u32 s = s0 + lane; s < s1; s += L) {
                const simk_u32 ij = D->lu_sub[2 * s], kj = D->lu_sub[2 * s + 1];
                S.lu[ij] = simk_t_fma(-l, S.lu[kj], S.lu[ij]);
            }
            simk_team_sync<L>();
        }
        if (lane == 0) {
            const simk_u32 ii = D->diag[i];
            S.lu[ii] = T(1) / (S.lu[ii] + ghinv);
        }
        simk_team_sync<L>();
    }
}

// x = (LU)^-1 x by columns (kin_solve, kin_solver.h:122-139).
template <class T, simk_u32 L>
SIMK_DEV void simk_wave_solve(const SimkDevTables *D, const SimkWaveTables *W, SimkWaveState<T> &S, T *x) {
    const simk_u32 lane = threadIdx.x % L;
    for (simk_u32 c = 0; c < S.n; ++c) {
        const T xc = x[c];
        for (simk_u32 e = W->lcol_begin[c] + lane; e < W->lcol_begin[c + 1]; e += L) {
            const simk_u32 row = W->lcol_row[e];
            x[row] = simk_t_fma(-S.lu[W->lcol_idx[e]], xc, x[row]);
        }
        simk_team_sync<L>();
    }
    for (simk_u32 c = S.n; c-- > 0;) {
        if (lane == 0) x[c] = x[c] * S.lu[D->diag[c]];
        simk_team_sync<L>();
        const T xc = x[c];
        for (simk_u32 e = W->ucol_begin[c] + lane; e < W->ucol_begin[c + 1]; e += L) {
            const simk_u32 row = W->ucol_row[e];
            x[row] = simk_t_fma(-S.lu[W->ucol_idx[e]], xc, x[row]);
        }
        simk_team_sync<L>();
    }
}

// One step attempt (kin_integrate's loop body, kin_solver.h:216-311).
template <class T, bool DELTA, simk_u32 L>
SIMK_DEV void simk_wave_attempt(const SimkDevTables *D, const SimkWaveTables *W, SimkWaveState<T> &S,
                                const SimkTConsts<T> &C, T &t, const T t_end, T &h, T &steps, T &rejects,
                                bool &reject_last, bool &reject_more) {
    const simk_u32 lane = threadIdx.x % L, n = S.n;
    const T one = T(1);
    const T remaining = t_end - t;
    T h_cap = remaining;
    if (C.h_max_pos) h_cap = simk_t_min(h_cap, C.h_max);
    const T h_wanted = simk_t_max(h, C.h_min);
    const bool last = h_wanted >= remaining;
    const T hs = last ? remaining : simk_t_min(h_wanted, h_cap);

    const T *pt = S.y;
    if (DELTA) {
        for (simk_u32 s = lane; s < n; s += L) S.yeval[s] = S.base[s] + S.y[s];
        simk_team_sync<L>();
        pt = S.yeval;
    }
    simk_wave_fun<T, L>(D, W, S, pt, S.fcn);
    const T ghinv = one / (hs * C.gamma0);
    simk_wave_jac_factor<T, L>(D, W, S, pt, ghinv);
    const T hinv = one / hs;

    for (simk_u32 st = 0; st < C.stages; ++st) {
        const simk_u32 packed = st * (st - 1) / 2;
        T *Kst = S.K + st * n;
        if (st > 0 && ((C.new_f >> st) & 1u)) {
            for (simk_u32 s = lane; s < n; s += L) {
                T v = S.y[s];
                for (simk_u32 j = 0; j < st; ++j)
                    if ((C.a_nz >> (packed + j)) & 1u) v = simk_t_fma(C.a[packed + j], S.K[j * n + s], v);
                S.ytmp[s] = v;
                if (DELTA) S.yeval[s] = S.base[s] + v;
            }
            simk_team_sync<L>();
            simk_wave_fun<T, L>(D, W, S, DELTA ? S.yeval : S.ytmp, S.fcn);
        }
        for (simk_u32 s = lane; s < n; s += L) {
            T v = S.fcn[s];
            for (simk_u32 j = 0; j < st; ++j)
                if ((C.c_nz >> (packed + j)) & 1u) v = simk_t_fma(C.c[packed + j] * hinv, S.K[j * n + s], v);
            Kst[s] = v;
        }
        simk_team_sync<L>();
        simk_wave_solve<T, L>(D, W, S, Kst);
    }

    for (simk_u32 s = lane; s < n; s += L) {
        T ynew = S.y[s], yerr = T(0);
        for (simk_u32 st = 0; st < C.stages; ++st) {
            const T kv = S.K[st * n + s];
            if ((C.m_nz >> st) & 1u) ynew = simk_t_fma(C.m[st], kv, ynew);
            if ((C.e_nz >> st) & 1u) yerr = simk_t_fma(C.e[st], kv, yerr);
        }
        T scale;
        if (DELTA)
            scale = simk_t_fma(C.rtol, simk_t_max(simk_t_abs(S.base[s] + S.y[s]), simk_t_abs(S.base[s] + ynew)), C.atol);
        else
            scale = simk_t_fma(C.rtol, simk_t_max(simk_t_abs(S.y[s]), simk_t_abs(ynew)), C.atol);
        S.q[s] = yerr / scale;
        S.ytmp[s] = ynew;
    }
    simk_team_sync<L>();
    if (lane == 0) {
        T err = T(0);
        for (simk_u32 s = 0; s < n; ++s) err = simk_t_fma(S.q[s], S.q[s], err);
        S.scal[0] = err;
    }
    simk_team_sync<L>();
    T err = simk_t_sqrt(S.scal[0] * C.inv_n);
    if (!simk_t_finite(err)) err = T(1e10);

    T fac = simk_t_step_factor(err, C.order);
    bool ok = (err <= one) || (hs <= C.h_min);
    if (C.fixed_step) {
        ok = true;
        fac = one;
    }
    if (ok) {
        for (simk_u32 s = lane; s < n; s += L) {
            T v = S.ytmp[s];
            if (C.clamp_negative) v = DELTA ? simk_t_max(v, -S.base[s]) : simk_t_max(v, T(0));
            S.y[s] = v;
        }
        t = last ? t_end : t + hs;
    }
    T h_acc = hs * fac;
    if (reject_last) h_acc = simk_t_min(h_acc, hs);
    if (last && fac >= one) h_acc = simk_t_max(h_acc, h_wanted);
    T h_rej = reject_more ? hs * T(0.1) : simk_t_min(hs * fac, hs);
    if (C.fixed_step) {
        h_acc = h_wanted;
        h_rej = h_wanted;
    }
    h = ok ? h_acc : h_rej;
    reject_more = !ok && reject_last;
    reject_last = !ok;
    steps = steps + (ok ? one : T(0));
    rejects = rejects + (ok ? T(0) : one);
    simk_team_sync<L>();
}

// Bytes of LDS one cell's team uses: the cell's state (simk_wave_lds_elems
// values of T), then n doubles, the delta form's f64 storage between repeats.
template <class T>
SIMK_DEV simk_u32 simk_team_lds_bytes(simk_u32 n, simk_u32 R, simk_u32 nnz) {
    return ((simk_wave_lds_elems(n, R, nnz) * (simk_u32)sizeof(T) + 7u) & ~7u) + 8u * n;
}

// One cell advanced by a team: dbd_advance_group (load, kin_advance with
// pulse segments, store) with the cell's state in the team's LDS region
// `mine` (simk_team_lds_bytes of this mechanism).
template <class T, bool DELTA, simk_u32 L>
SIMK_DEV void simk_wave_cell(const SimkKinArgs &A, const SimkTConsts<T> &C, const SimkWaveTables *W,
                             unsigned char *mine, const simk_u64 cell) {
    const SimkDevTables *D = A.tables;
    const simk_u32 lane = threadIdx.x % L;
    double *y64 = (double *)(mine + ((simk_wave_lds_elems(D->n, D->R, D->nnz) * (simk_u32)sizeof(T) + 7u) & ~7u));
    SimkWaveState<T> S;
    S.init((T *)mine, D->n, D->R, D->nnz);
    const simk_u32 n = S.n, R = S.R;
    const simk_u64 p = A.padded;
    const T duration = (T)A.duration;
    // dbd_advance_group's load (dbd_lanes.h:31-54)
    for (simk_u32 s = lane; s < n; s += L) {
        const double v = A.y[s * p + cell];
        if (DELTA) {
            y64[s] = v;
            S.base[s] = (T)v;
            S.y[s] = T(0);
        } else {
            S.y[s] = (T)v;
        }
        S.src[s] = (T)A.src[s * p + cell];
        S.pulse[s] = (T)A.pulse[s * p + cell];
    }
    for (simk_u32 r = lane; r < R; r += L) S.k[r] = (T)A.k[r * p + cell];
    const T period = (T)A.period[cell], first = (T)A.pulse_first[cell], count = (T)A.pulse_count[cell];
    T h = (T)A.h[cell];
    T failed = (T)A.failed[cell];
    double acc_steps = 0.0, acc_rejects = 0.0, acc_pulses = 0.0;
    simk_team_sync<L>();
    for (simk_u32 rep = 0; rep < A.repeats; ++rep) {
        if (DELTA && rep > 0) {   // the next call restarts from the f64 storage
            for (simk_u32 s = lane; s < n; s += L) {
                S.base[s] = (T)y64[s];
                S.y[s] = T(0);
            }
            simk_team_sync<L>();
        }
        T steps = T(0), rejects = T(0), pulses = T(0), t = T(0), applied = T(0);
        // kin_advance (kin_solver.h:322-347)
        for (;;) {
            const bool pending = applied < count;
            const T next_pulse = simk_t_fma(applied, period, first);
            const T seg_end = pending ? simk_t_min(next_pulse, duration) : duration;
            // kin_integrate (kin_solver.h:189-313)
            h = h > T(0) ? h : C.h_start;
            bool reject_last = false, reject_more = false;
            bool active = (t < seg_end) && !(failed > T(0));
            simk_u32 it = 0;
            while (active) {
                if (++it > C.max_steps) {
                    failed = T(1);
                    break;
                }
                simk_wave_attempt<T, DELTA, L>(D, W, S, C, t, seg_end, h, steps, rejects, reject_last, reject_more);
                active = t < seg_end;
            }
            // pending recomputed (applied is unchanged since the segment
            // began): with teams narrower than the wave the loop above exits
            // divergently, and no lane-mask value is carried across it
            const bool pending_end = applied < count;
            const bool alive = !(failed > T(0));
            const bool hit = pending_end && alive && (t >= seg_end);
            if (hit) {
                for (simk_u32 s = lane; s < n; s += L) S.y[s] = S.y[s] + S.pulse[s];
                simk_team_sync<L>();
                applied = applied + T(1);
                pulses = pulses + T(1);
                h = C.h_start;
            }
            // "!pending" ends the loop for a NaN count as well; for every
            // finite count it is the reference's exit test
            if (((t >= duration) && (applied >= count || !pending_end)) || !alive) break;
        }
        if (DELTA) {   // dbd_lanes.h:58-67
            for (simk_u32 s = lane; s < n; s += L) {
                double v = y64[s] + (double)S.y[s];
                if (C.clamp_negative) v = v > 0.0 ? v : 0.0;
                y64[s] = v;
            }
            simk_team_sync<L>();
        }
        // per-call counters added to the f64 totals (integers: exact in any grouping)
        acc_steps = acc_steps + (double)steps;
        acc_rejects = acc_rejects + (double)rejects;
        acc_pulses = acc_pulses + (double)pulses;
    }
    // dbd_advance_group's store (dbd_lanes.h:56-77)
    for (simk_u32 s = lane; s < n; s += L) A.y[s * p + cell] = DELTA ? y64[s] : (double)S.y[s];
    if (lane == 0) {
        A.h[cell] = (double)h;
        A.failed[cell] = (double)failed;
        A.steps[cell] = A.steps[cell] + acc_steps;
        A.rejects[cell] = A.rejects[cell] + acc_rejects;
        A.pulses[cell] = A.pulses[cell] + acc_pulses;
    }
    simk_team_sync<L>();
}

// The next queue position of A for the calling team: its first lane takes
// it, the team shares it.
template <simk_u32 L>
SIMK_DEV simk_u32 simk_team_take(const SimkKinArgs &A) {
    const simk_u32 lane = threadIdx.x % L, team = threadIdx.x / L;
    simk_u32 got = 0;
    if (lane == 0) got = atomicAdd(A.queue, 1u);
    if constexpr (L == SIMK_WAVE)
        return (simk_u32)__builtin_amdgcn_readfirstlane((int)got);
    else
        return (simk_u32)__builtin_amdgcn_ds_bpermute((int)((team * L) << 2), (int)got);
}

// Persistent: each team takes cells from the queue one at a time and
// advances them. lds: the wave's LDS, 64 / L consecutive team regions.
template <class T, bool DELTA, simk_u32 L>
SIMK_DEV void simk_wave_body(const SimkKinArgs &A, const SimkTConsts<T> &C, const SimkWaveTables *W,
                             unsigned char *lds) {
    constexpr simk_u32 kTeams = SIMK_WAVE / L;
    const SimkDevTables *D = A.tables;
    const simk_u32 team = threadIdx.x / L;
    unsigned char *mine = lds + team * simk_team_lds_bytes<T>(D->n, D->R, D->nnz);
    simk_u32 pos = blockIdx.x * kTeams + team;   // the queue head starts at gridDim.x * kTeams
    bool first_cell = !A.queue_first;
    for (;;) {
        if (!first_cell) pos = simk_team_take<L>(A);
        first_cell = false;
        if (pos >= A.lanes) break;
        simk_wave_cell<T, DELTA, L>(A, C, W, mine, A.order ? A.order[pos] : pos);
    }
}

// Several mechanisms in one launch (ragged batches, FP64): per-mechanism
// arguments, constants and contribution lists; a team visits the
// mechanisms' queues in `seq` order (every queue position comes from its
// queue's counter, starting at 0) and moves on when one is empty.
struct SimkWaveMultiArgs {
    SimkKinArgs a[4];
    SimkTConsts<double> c[4];
    const SimkWaveTables *w[4];
    simk_u32 count;
    simk_u32 team_bytes;     // LDS bytes per team: the largest mechanism's
    simk_u32 seq[4];
};

template <simk_u32 L>
SIMK_DEV void simk_wave_multi_body(const SimkWaveMultiArgs &M, unsigned char *lds) {
    const simk_u32 team = threadIdx.x / L;
    unsigned char *mine = lds + team * M.team_bytes;
    for (simk_u32 k = 0; k < M.count;) {
        const simk_u32 i = M.seq[k];
        const SimkKinArgs &A = M.a[i];
        const simk_u32 pos = simk_team_take<L>(A);
        if (pos >= A.lanes) {
            ++k;
            continue;
        }
        simk_wave_cell<double, false, L>(A, M.c[i], M.w[i], mine, A.order ? A.order[pos] : pos);
    }
}
