// Device side of the batched Rosenbrock solver for gfx950: one cell per lane,
// 64-lane wavefronts, per-lane adaptive step control.
//
// This header is compiled twice: into libsimk_gpu.so (the table-walk kernel,
// any mechanism) and, as embedded source text, into every kernel the code
// generator emits for one mechanism through hipRTC. It includes no system or
// HIP headers so hipRTC compiles it as is.
//
// Arithmetic contract (the reference's lane rules, src/base/base_lane.h:13-21):
// every fused multiply-add is an explicit __builtin_fma (files are compiled
// with -ffp-contract=off); division and sqrt are IEEE correctly rounded f64;
// min/max are the select forms "a < b ? a : b"; |a| keeps -0; finiteness is
// tested as (x - x) == (x - x). So each lane performs exactly the operation
// sequence of the reference's scalar kin_advance on its cell.
//
// Memory: every [row][lanes] f64 array (the zone block rows and the per-lane
// workspace) is addressed through a buffer resource whose base is the wave's
// first lane: the lane offset is one VGPR, the row offset a scalar register,
// so no per-lane 64-bit address is ever kept live, and every access is one
// coalesced 512-byte wave transaction.
#pragma once

typedef unsigned int simk_u32;
typedef unsigned long long simk_u64;
typedef unsigned int simk_v2u __attribute__((ext_vector_type(2)));
typedef unsigned int simk_v4u __attribute__((ext_vector_type(4)));

#define SIMK_WAVE 64
#define SIMK_DEV __device__ __forceinline__

struct SimkDevCtl {
    double rtol, atol, h_min, h_max, h_start;
    double inv_n;            // 1.0 / n, rounded once on the host (kin_solver.h:197)
    simk_u32 max_steps, fixed_step, clamp_negative, pad_;
};

// Method coefficients with the host-derived stage plan.
struct SimkDevMethod {
    simk_u32 stages, order;
    double a[6], c[6], m[4], e[4], gamma0;
    simk_u32 new_f[4];
    simk_u32 eval[4];        // stage st evaluates f at its own point (else it reuses the saved f)
    simk_u32 save_f[4];      // stage st's f is reused by a later stage
};

// Mechanism tables of the table-walk kernel, in device memory (uniform reads).
struct SimkDevTables {
    simk_u32 n, R, nnz, pad_;
    const unsigned char *reactant;      // [R][3]
    const unsigned char *reactant_count;// [R]
    const unsigned char *net_count;     // [R]
    const unsigned char *net_species;   // [R][6]
    const double *net_coef;             // [R][6]
    const unsigned short *jac_index;    // [R][3][6]
    const unsigned short *diag;         // [n]
    const unsigned short *div_begin;    // [n + 1]
    const unsigned short *lu_div;       // [D][4]: ik, kk, sub_begin, sub_end
    const unsigned short *lu_sub;       // [S][2]: ij, kj
    const unsigned short *lower_begin;  // [n + 1]
    const unsigned short *upper_begin;  // [n + 1]
    const unsigned short *lower;        // [L][2]: index, col
    const unsigned short *upper;        // [U][2]
};

// Kernel argument: the zone block (GpuKinLaunch fields) plus workspace,
// control and method, passed by value.
struct SimkKinArgs {
    double *y, *h, *steps, *rejects, *pulses, *failed;
    const double *k, *src, *pulse, *period, *pulse_first, *pulse_count;
    simk_u32 padded;         // row stride of the block (lanes)
    simk_u32 lanes;          // cells [0, lanes) are integrated
    double duration;
    double *ws;              // workspace: [wave][slot][64] f64
    simk_u32 ws_slots;       // slots per lane
    simk_u32 n_species, n_reactions;
    simk_u32 repeats;        // consecutive advances of `duration` per cell (h carried)
    simk_u32 queue_first;    // 1: even the first queue positions come from the atomic (queue starts at 0)
    simk_u32 *queue;         // next unassigned queue position (starts at gridDim.x * 64)
    const simk_u32 *order;   // queue position -> cell (null: identity)
    const SimkDevTables *tables;
    SimkDevCtl ctl;
    SimkDevMethod method;
    // Time slicing (0: run every cell to the end of its advance). Otherwise a
    // wave takes no new cell once `slice` ticks of the 100 MHz real-time
    // counter have passed since it started, and a lane whose cell is between
    // two step attempts saves the cell's exact integration state (y, h,
    // failed and the counters to the block, the loop state to `resume`) and
    // appends the cell to `carry`; the next launch that takes the cell
    // continues it where it stopped.
    simk_u64 slice;
    double *resume;          // [5][padded]: t, pulses applied, repeat, attempt count, flags
    simk_u32 *carry;         // cells stopped mid-advance
    simk_u32 *carry_count;
};

// resume flags: SIMK_RESUME_VALID + reject_last + 2 * reject_more
#define SIMK_RESUME_VALID 4.0

// ---- lane vocabulary ------------------------------------------------------
SIMK_DEV double simk_fma(double a, double b, double c) { return __builtin_fma(a, b, c); }
SIMK_DEV double simk_fnma(double a, double b, double c) { return __builtin_fma(-a, b, c); }
SIMK_DEV double simk_min(double a, double b) { return a < b ? a : b; }
SIMK_DEV double simk_max(double a, double b) { return a > b ? a : b; }
SIMK_DEV double simk_abs(double a) { return a < 0.0 ? -a : a; }
SIMK_DEV double simk_sqrt(double a) { return __builtin_sqrt(a); }
SIMK_DEV bool simk_is_finite(double a) {
    double d = a - a;
    return d == d;
}

// x^(1/3) on [1e-4, 1e4] from arithmetic only (base_lane.h:123-135).
SIMK_DEV double simk_cbrt_clamped(double x) {
    x = simk_min(simk_max(x, 1e-4), 1e4);
    double r4 = simk_sqrt(simk_sqrt(x));
    double r16 = simk_sqrt(simk_sqrt(r4));
    double r64 = simk_sqrt(simk_sqrt(r16));
    double g = r4 * r16 * r64;
    const double third = 1.0 / 3.0;
    g = (2.0 * g + x / (g * g)) * third;
    g = (2.0 * g + x / (g * g)) * third;
    return g;
}

// Step-size factor (kin_solver.h:169-180).
SIMK_DEV double simk_step_factor(double err, simk_u32 order) {
    double e = simk_min(simk_max(err, 1e-4), 1e4);
    double root;
    if (order == 2) root = simk_sqrt(e);
    else if (order == 4) root = simk_sqrt(simk_sqrt(e));
    else root = simk_cbrt_clamped(e);
    double fac = 0.9 / root;
    return simk_min(simk_max(fac, 0.2), 6.0);
}

// Compiler-level memory fence: values stored to the workspace are re-read from
// memory after it instead of being kept live in registers (which would spill).
#define SIMK_MEMORY_FENCE() asm volatile("" ::: "memory")

// ---- [row][lanes] f64 arrays through a buffer resource ----------------------
// Raw buffer, 32-bit data format word (gfx9 family descriptor word 3).
#define SIMK_RSRC_FLAGS 0x00027000

struct SimkRows {
    __amdgpu_buffer_rsrc_t rs;
    simk_u32 voff;           // lane * 8
    simk_u32 stride;         // bytes between rows
    // base: first element seen by lane offset 0; voff_bytes: this lane's byte
    // offset; bytes: extent of the array from base (accesses beyond read 0).
    SIMK_DEV void init(const double *base, simk_u32 voff_bytes, simk_u32 row_stride_bytes, simk_u32 bytes) {
        rs = __builtin_amdgcn_make_buffer_rsrc((void *)base, 0, (int)bytes, SIMK_RSRC_FLAGS);
        voff = voff_bytes;
        stride = row_stride_bytes;
    }
    SIMK_DEV double ld(simk_u32 row) const {
        simk_v2u v = __builtin_amdgcn_raw_buffer_load_b64(rs, (int)voff, (int)(row * stride), 0);
        return __builtin_bit_cast(double, v);
    }
    SIMK_DEV void st(simk_u32 row, double x) const {
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(simk_v2u, x), rs, (int)voff, (int)(row * stride), 0);
    }
};

// Workspace slots held in adjacent pairs per lane, [pair][64 lanes][2] f64,
// so that two slots a lane reads or writes together move in one 16-byte
// access (a wave instruction then carries 1 KiB instead of 512 B). A single
// slot is still one 8-byte access.
struct SimkD2 {
    double x, y;
};

struct SimkRowPairs {
    __amdgpu_buffer_rsrc_t rs;
    simk_u32 voff;           // lane * 16
    static constexpr simk_u32 kPairBytes = SIMK_WAVE * 16u;
    SIMK_DEV void init(const double *base, simk_u32 lane, simk_u32 bytes) {
        rs = __builtin_amdgcn_make_buffer_rsrc((void *)base, 0, (int)bytes, SIMK_RSRC_FLAGS);
        voff = lane * 16u;
    }
    SIMK_DEV double ld(simk_u32 slot) const {
        simk_v2u v = __builtin_amdgcn_raw_buffer_load_b64(rs, (int)voff, (int)((slot >> 1) * kPairBytes + (slot & 1u) * 8u), 0);
        return __builtin_bit_cast(double, v);
    }
    SIMK_DEV void st(simk_u32 slot, double x) const {
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(simk_v2u, x), rs, (int)voff,
                                              (int)((slot >> 1) * kPairBytes + (slot & 1u) * 8u), 0);
    }
    SIMK_DEV SimkD2 ld2(simk_u32 pair) const {
        const simk_v4u v = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)voff, (int)(pair * kPairBytes), 0);
        SimkD2 r;
        r.x = __builtin_bit_cast(double, simk_v2u{v.x, v.y});
        r.y = __builtin_bit_cast(double, simk_v2u{v.z, v.w});
        return r;
    }
    SIMK_DEV void st2(simk_u32 pair, double a, double b) const {
        const simk_v2u lo = __builtin_bit_cast(simk_v2u, a), hi = __builtin_bit_cast(simk_v2u, b);
        __builtin_amdgcn_raw_buffer_store_b128(simk_v4u{lo.x, lo.y, hi.x, hi.y}, rs, (int)voff, (int)(pair * kPairBytes), 0);
    }
};

// f32 workspace slots in adjacent pairs per lane, [pair][64 lanes][2] f32:
// two slots in one 8-byte access (the f32 delta kernels).
struct SimkF2 {
    float x, y;
};

struct SimkRowPairs32 {
    __amdgpu_buffer_rsrc_t rs;
    simk_u32 voff;           // lane * 8
    static constexpr simk_u32 kPairBytes = SIMK_WAVE * 8u;
    SIMK_DEV void init(const void *base, simk_u32 lane, simk_u32 bytes) {
        rs = __builtin_amdgcn_make_buffer_rsrc((void *)base, 0, (int)bytes, SIMK_RSRC_FLAGS);
        voff = lane * 8u;
    }
    SIMK_DEV float ld(simk_u32 slot) const {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(
                                             rs, (int)voff, (int)((slot >> 1) * kPairBytes + (slot & 1u) * 4u), 0));
    }
    SIMK_DEV void st(simk_u32 slot, float x) const {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(simk_u32, x), rs, (int)voff,
                                              (int)((slot >> 1) * kPairBytes + (slot & 1u) * 4u), 0);
    }
    SIMK_DEV SimkF2 ld2(simk_u32 pair) const {
        // the whole 8-byte load reinterpreted at once: a bit cast applied to
        // one element of the vector (v.y) yielded element 0 with this compiler
        return __builtin_bit_cast(SimkF2, __builtin_amdgcn_raw_buffer_load_b64(rs, (int)voff, (int)(pair * kPairBytes), 0));
    }
    SIMK_DEV void st2(simk_u32 pair, float a, float b) const {
        __builtin_amdgcn_raw_buffer_store_b64(simk_v2u{__builtin_bit_cast(simk_u32, a), __builtin_bit_cast(simk_u32, b)},
                                              rs, (int)voff, (int)(pair * kPairBytes), 0);
    }
};

// f32 workspace slots in groups of four per lane, [quad][64 lanes][4] f32:
// four slots in one 16-byte access, two adjacent ones (2k, 2k + 1) in one
// 8-byte access (the FP32 factor-storage LU).
struct SimkF4 {
    float x, y, z, w;
};

struct SimkRowQuads32 {
    __amdgpu_buffer_rsrc_t rs;
    simk_u32 voff;           // lane * 16
    static constexpr simk_u32 kQuadBytes = SIMK_WAVE * 16u;
    SIMK_DEV void init(const void *base, simk_u32 lane, simk_u32 bytes) {
        rs = __builtin_amdgcn_make_buffer_rsrc((void *)base, 0, (int)bytes, SIMK_RSRC_FLAGS);
        voff = lane * 16u;
    }
    static SIMK_DEV int at(simk_u32 slot) { return (int)((slot >> 2) * kQuadBytes + (slot & 3u) * 4u); }
    SIMK_DEV float ld(simk_u32 slot) const {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, (int)voff, at(slot), 0));
    }
    SIMK_DEV void st(simk_u32 slot, float x) const {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(simk_u32, x), rs, (int)voff, at(slot), 0);
    }
    // slots 2 pair and 2 pair + 1
    SIMK_DEV SimkF2 ld2(simk_u32 pair) const {
        return __builtin_bit_cast(SimkF2, __builtin_amdgcn_raw_buffer_load_b64(rs, (int)voff, at(2u * pair), 0));
    }
    SIMK_DEV void st2(simk_u32 pair, float a, float b) const {
        __builtin_amdgcn_raw_buffer_store_b64(simk_v2u{__builtin_bit_cast(simk_u32, a), __builtin_bit_cast(simk_u32, b)},
                                              rs, (int)voff, at(2u * pair), 0);
    }
    // slots 4 quad .. 4 quad + 3
    SIMK_DEV SimkF4 ld4(simk_u32 quad) const {
        return __builtin_bit_cast(SimkF4, __builtin_amdgcn_raw_buffer_load_b128(rs, (int)voff, (int)(quad * kQuadBytes), 0));
    }
    SIMK_DEV void st4(simk_u32 quad, float a, float b, float c, float d) const {
        __builtin_amdgcn_raw_buffer_store_b128(simk_v4u{__builtin_bit_cast(simk_u32, a), __builtin_bit_cast(simk_u32, b),
                                                        __builtin_bit_cast(simk_u32, c), __builtin_bit_cast(simk_u32, d)},
                                               rs, (int)voff, (int)(quad * kQuadBytes), 0);
    }
};

// The same for f32 arrays (lane stride 4 bytes): the mixed-precision LU.
struct SimkRows32 {
    __amdgpu_buffer_rsrc_t rs;
    simk_u32 voff;
    simk_u32 stride;
    SIMK_DEV void init(const void *base, simk_u32 voff_bytes, simk_u32 row_stride_bytes, simk_u32 bytes) {
        rs = __builtin_amdgcn_make_buffer_rsrc((void *)base, 0, (int)bytes, SIMK_RSRC_FLAGS);
        voff = voff_bytes;
        stride = row_stride_bytes;
    }
    SIMK_DEV float ld(simk_u32 row) const {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, (int)voff, (int)(row * stride), 0));
    }
    SIMK_DEV void st(simk_u32 row, float x) const {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(simk_u32, x), rs, (int)voff, (int)(row * stride), 0);
    }
};

// A species vector in registers (generated kernels: N is a compile-time constant).
template <simk_u32 N>
struct SimkRegVec {
    double v[N];
    SIMK_DEV double get(simk_u32 s) const { return v[s]; }
    SIMK_DEV void set(simk_u32 s, double x) { v[s] = x; }
};

// A species vector in workspace slots [base, base + n) (table-walk kernel).
struct SimkWsVec {
    SimkRows w;
    simk_u32 base;
    SIMK_DEV double get(simk_u32 s) const { return w.ld(base + s); }
    SIMK_DEV void set(simk_u32 s, double x) const { w.st(base + s, x); }
};

// ---------------------------------------------------------------------------
// The integrator, written once over an Ops policy P:
//   P::Vec                 per-lane species vector (get/set)
//   p.n()                  species count (constexpr in generated code)
//   p.vec(i)               storage for vector i (0 y, 2 stage point / y_new, 3 stage solution, 4 error)
//   p.fun(y, f)            f = src + sum_r net_r k_r prod y         (kin_solver.h:39-53)
//   p.jac_factor(y, g)     LU of (I g - J(y)) into the workspace    (:75-113)
//   p.solve(x, y, g)       x = (LU)^-1 x (:122-139); the mixed-precision policy
//                          also refines against (I g - J(y)) in FP64
//   p.kstore(st, x)        keeps stage vector K_st (the last stage stays in x)
//   p.kaxpy(j, c, v)       v += c K_j, elementwise fma
//   p.fsave(x), p.fload(x) the f a later stage reuses
//   p.pulse(s)             the cell's pulse increment
// and a method policy M (compile-time constants in generated code).
//
// Order of work inside one step attempt differs from the reference only where
// results cannot depend on it: -J/LU is formed before f(y) (the two read y
// only), a stage whose point equals the previous evaluation point reuses
// the saved f instead of recomputing the same bits (Rodas3 stage 2), and a
// stage's combination whose last term is the previous stage's K is finished
// at the end of that stage, from the K still in registers.

// One step attempt of kin_integrate's loop body (kin_solver.h:216-311) for the
// active lanes: step size, -J/LU, the stages, error norm, accept or reject,
// next h. The caller has already counted the attempt against max_steps.
template <class P, class M>
SIMK_DEV void simk_attempt(P &p, const M &me, const SimkDevCtl &ctl, typename P::Vec &y, double &t,
                           const double t_end, double &h, double &steps, double &rejects, bool &reject_last,
                           bool &reject_more) {
    const simk_u32 n = p.n();
    const double remaining = t_end - t;
    double h_cap = remaining;
    if (ctl.h_max > 0.0) h_cap = simk_min(h_cap, ctl.h_max);
    const double h_wanted = simk_max(h, ctl.h_min);
    const bool last = h_wanted >= remaining;
    const double hs = last ? remaining : simk_min(h_wanted, h_cap);

    const double ghinv = 1.0 / (hs * me.gamma0());
    p.jac_factor(y, ghinv);
    const double hinv = 1.0 / hs;

    typename P::Vec x = p.vec(3);
    const simk_u32 S = me.stages();
    // Stage st >= 1 can be started at the end of stage st - 1, whose K is
    // still in x, when that K is the last term of st's combination (terms run
    // over j ascending, so the K_{st-1} term is applied last): its point
    // (a stage that evaluates f) or its right-hand side, the saved f plus the
    // c-combination (a stage that reuses f). Either is left in x for the next
    // stage. Every element sees the same fma sequence; the carried stage
    // skips one load of K_{st-1}.
    auto carry_point = [&](const simk_u32 st) {
        return P::kStageCarry && st >= 1 && st < S && me.eval(st) && me.a(st * (st - 1) / 2 + st - 1) != 0.0;
    };
    auto carry_rhs = [&](const simk_u32 st) {
        return P::kStageCarry && st >= 1 && st < S && !me.eval(st) && me.c(st * (st - 1) / 2 + st - 1) != 0.0;
    };
    // One stage: its point, f there (or the saved f), the c-combination of
    // earlier stages, the solve. Rolled by default so the solve and f are
    // emitted once; the generator may unroll it. Every vector update runs
    // over stages j in ascending order for each element, as in the
    // reference's loops (kin_solver.h:238-259); only the loop nesting differs.
    auto stage = [&](const simk_u32 st) {
        const simk_u32 packed = st * (st - 1) / 2;
        const bool carried_rhs = carry_rhs(st);
        bool have_f = false;
        if constexpr (P::kFunInJac) {
            // jac_factor left f(y) in the f slots, where a later stage that
            // reuses stage 0's f also finds it
            if (st == 0) {
                p.fjac(x);
                have_f = true;
            }
        }
        if (have_f) {
        } else if (st == 0 || me.eval(st)) {
            typename P::Vec pt = p.vec(2);
            const bool carried_point = carry_point(st);
#pragma unroll
            for (simk_u32 s = 0; s < n; ++s) pt.set(s, y.get(s));
            for (simk_u32 j = 0; j < st - (carried_point ? 1u : 0u); ++j) {
                const double a = me.a(packed + j);
                if (a != 0.0) p.kaxpy(j, a, pt);
            }
            if (carried_point) {
                const double a = me.a(packed + st - 1);
#pragma unroll
                for (simk_u32 s = 0; s < n; ++s) pt.set(s, simk_fma(a, x.get(s), pt.get(s)));
            }
            p.fun(pt, x);
            if (me.save_f(st)) p.fsave(x);
        } else if (!carried_rhs) {
            p.fload(x);
        }
        if (!carried_rhs) {
            for (simk_u32 j = 0; j < st; ++j) {
                const double c = me.c(packed + j);
                if (c != 0.0) p.kaxpy(j, c * hinv, x);
            }
        }
        p.solve(x, y, ghinv);
        p.kstore(st, x);
        const simk_u32 nx = st + 1, npacked = nx * st / 2;
        if (carry_rhs(nx)) {
            typename P::Vec acc = p.vec(2);
            p.fload(acc);
            for (simk_u32 j = 0; j < st; ++j) {
                const double c = me.c(npacked + j);
                if (c != 0.0) p.kaxpy(j, c * hinv, acc);
            }
            const double c = me.c(npacked + st) * hinv;
#pragma unroll
            for (simk_u32 s = 0; s < n; ++s) x.set(s, simk_fma(c, x.get(s), acc.get(s)));
        }
    };
    if constexpr (M::kUnrollStages) {
#pragma unroll
        for (simk_u32 st = 0; st < 4; ++st)
            if (st < me.stages()) stage(st);
    } else {
#pragma unroll 1
        for (simk_u32 st = 0; st < me.stages(); ++st) stage(st);
    }

    // y_new = y + sum m_st K_st and the error estimate sum e_st K_st (stages in
    // ascending order per element; kin_solver.h:264-277). The last stage's K
    // is still in x. The error vector is only materialised when an earlier
    // stage contributes to it.
    typename P::Vec ynew = p.vec(2);
#pragma unroll
    for (simk_u32 s = 0; s < n; ++s) ynew.set(s, y.get(s));
    bool early_e = false;
    for (simk_u32 st = 0; st + 1 < S; ++st) early_e = early_e || me.e(st) != 0.0;
    typename P::Vec yerr = p.vec(4);
    if (early_e) {
#pragma unroll
        for (simk_u32 s = 0; s < n; ++s) yerr.set(s, 0.0);
    }
    for (simk_u32 st = 0; st + 1 < S; ++st) {
        if (me.m(st) != 0.0) p.kaxpy(st, me.m(st), ynew);
        if (me.e(st) != 0.0) p.kaxpy(st, me.e(st), yerr);
    }
    const double m_last = me.m(S - 1), e_last = me.e(S - 1);
    double err = 0.0;
#pragma unroll
    for (simk_u32 s = 0; s < n; ++s) {
        double yn = ynew.get(s);
        if (m_last != 0.0) yn = simk_fma(m_last, x.get(s), yn);
        double ye = early_e ? yerr.get(s) : 0.0;
        if (e_last != 0.0) ye = simk_fma(e_last, x.get(s), ye);
        const double scale = simk_fma(ctl.rtol, simk_max(simk_abs(y.get(s)), simk_abs(yn)), ctl.atol);
        const double q = ye / scale;
        err = simk_fma(q, q, err);
        ynew.set(s, yn);
    }
    err = simk_sqrt(err * ctl.inv_n);
    if (!simk_is_finite(err)) err = 1e10;

    double fac = simk_step_factor(err, me.order());
    bool ok = (err <= 1.0) || (hs <= ctl.h_min);
    if (ctl.fixed_step) {
        ok = true;
        fac = 1.0;
    }
    if (ok) {
#pragma unroll
        for (simk_u32 s = 0; s < n; ++s) {
            double v = ynew.get(s);
            if (ctl.clamp_negative) v = simk_max(v, 0.0);
            y.set(s, v);
        }
        t = last ? t_end : t + hs;
    }
    const double hfac = hs * fac;
    double h_acc = hfac;
    if (reject_last) h_acc = simk_min(h_acc, hs);
    if (last && fac >= 1.0) h_acc = simk_max(h_acc, h_wanted);
    double h_rej = reject_more ? hs * 0.1 : simk_min(hfac, hs);
    if (ctl.fixed_step) {
        h_acc = h_wanted;
        h_rej = h_wanted;
    }
    h = ok ? h_acc : h_rej;
    reject_more = !ok && reject_last;
    reject_last = !ok;
    steps = steps + (ok ? 1.0 : 0.0);
    rejects = rejects + (ok ? 0.0 : 1.0);
}

// Lane phases of the persistent kernel.
#define SIMK_PH_CELL 0      // needs a cell from the queue
#define SIMK_PH_SEGMENT 1   // starts an integration segment (kin_advance loop head)
#define SIMK_PH_CHECK 2     // counts the next attempt against max_steps
#define SIMK_PH_ATTEMPT 3   // makes a step attempt this iteration
#define SIMK_PH_END 4       // segment ended: pulse, done test
#define SIMK_PH_EXIT 5      // queue exhausted

// Persistent, lane-refilling advance. Every lane runs the reference's per-cell
// control flow as a state machine — dbd_advance_group (dbd_lanes.h:28-78)
// around kin_advance (kin_solver.h:322-347) around kin_integrate's loop
// (:204-312) — and one wave-wide iteration makes one step attempt for every
// lane that has one pending. A lane whose cell is finished stores it and takes
// the next queue position, so lanes stay busy while the queue lasts. Each
// cell's arithmetic is exactly the reference's, whatever lane or order runs
// it. `repeats` runs that many consecutive advances of `duration` per cell
// (the result of that many calls with unchanged inputs), carrying h.
// SLICED: compile the time-slice paths (A.slice may then be non-zero).
template <class P, class M, bool SLICED = false>
SIMK_DEV void simk_persistent_body(const SimkKinArgs &A, const M &me) {
    const simk_u32 wave = blockIdx.x;
    const simk_u32 lane = threadIdx.x;
    const SimkDevCtl &ctl = A.ctl;
    const simk_u32 stride = A.padded * 8u;
    const simk_u32 row_bytes = A.padded * 8u;
    P p;
    p.init(A, wave, lane);
    typename P::Vec y = p.vec(0);
    const simk_u32 n = p.n();

    SimkRows ys, hs_, fs;
    ys.init(A.y, 0, stride, A.n_species * row_bytes);
    hs_.init(A.h, 0, 0, row_bytes);
    fs.init(A.failed, 0, 0, row_bytes);

    simk_u32 pos = wave * SIMK_WAVE + lane;   // first queue position without an atomic, unless queue_first
    simk_u32 cell = 0, phase = SIMK_PH_CELL, it = 0, rep = 0;
    double h = 0.0, t = 0.0, applied = 0.0, seg_end = 0.0, period = 0.0, first = 0.0, count = 0.0;
    double steps = 0.0, rejects = 0.0, pulses = 0.0, failed = 0.0;
    bool reject_last = false, reject_more = false, fresh = !A.queue_first;
    const simk_u64 t_begin = (SLICED && A.slice) ? __builtin_amdgcn_s_memrealtime() : 0;
    bool expired = false;

    for (;;) {
        if (SLICED && A.slice && !expired) expired = (__builtin_amdgcn_s_memrealtime() - t_begin) >= A.slice;
        // Resolve each lane's transitions until it has an attempt to make or
        // the queue is exhausted (bounded: every pass consumes a segment, a
        // pulse, an advance or a cell).
        while (phase != SIMK_PH_ATTEMPT && phase != SIMK_PH_EXIT) {
            if (phase == SIMK_PH_CELL) {
                if (SLICED && expired) {
                    phase = SIMK_PH_EXIT;
                    break;
                }
                if (!fresh) {
                    // Wave-aggregated refill: one atomic for every lane asking now.
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
                const simk_u32 voff = cell * 8u;
                ys.voff = hs_.voff = fs.voff = voff;
                p.set_cell(A, cell);
#pragma unroll
                for (simk_u32 s = 0; s < n; ++s) y.set(s, ys.ld(s));
                h = hs_.ld(0);
                failed = fs.ld(0);
                SimkRows one;
                one.init(A.period, voff, 0, row_bytes);
                period = one.ld(0);
                one.init(A.pulse_first, voff, 0, row_bytes);
                first = one.ld(0);
                one.init(A.pulse_count, voff, 0, row_bytes);
                count = one.ld(0);
                steps = rejects = pulses = 0.0;
                t = 0.0;
                applied = 0.0;
                rep = 0;
                phase = SIMK_PH_SEGMENT;
                if (SLICED && A.resume) {
                    SimkRows rr;
                    rr.init(A.resume, voff, row_bytes, 5u * row_bytes);
                    const double flags = rr.ld(4);
                    if (flags >= SIMK_RESUME_VALID) {
                        // Continue a stopped advance: the loop state as it was
                        // before the attempt that did not start (the segment's
                        // h, attempt count and reject flags are not reset).
                        t = rr.ld(0);
                        applied = rr.ld(1);
                        rep = (simk_u32)rr.ld(2);
                        it = (simk_u32)rr.ld(3);
                        const simk_u32 f = (simk_u32)(flags - SIMK_RESUME_VALID);
                        reject_last = (f & 1u) != 0;
                        reject_more = (f & 2u) != 0;
                        rr.st(4, 0.0);
                        const bool pending = applied < count;
                        const double next_pulse = simk_fma(applied, period, first);
                        seg_end = pending ? simk_min(next_pulse, A.duration) : A.duration;
                        phase = SIMK_PH_CHECK;
                    }
                }
            } else if (phase == SIMK_PH_SEGMENT) {
                const bool pending = applied < count;
                const double next_pulse = simk_fma(applied, period, first);
                seg_end = pending ? simk_min(next_pulse, A.duration) : A.duration;
                h = h > 0.0 ? h : ctl.h_start;
                reject_last = reject_more = false;
                it = 0;
                phase = ((t < seg_end) && !(failed > 0.0)) ? SIMK_PH_CHECK : SIMK_PH_END;
            } else if (phase == SIMK_PH_CHECK) {
                if (SLICED && expired) {
#pragma unroll
                    for (simk_u32 s = 0; s < n; ++s) ys.st(s, y.get(s));
                    hs_.st(0, h);
                    fs.st(0, failed);
                    SimkRows acc;
                    acc.init(A.steps, cell * 8u, 0, row_bytes);
                    acc.st(0, acc.ld(0) + steps);
                    acc.init(A.rejects, cell * 8u, 0, row_bytes);
                    acc.st(0, acc.ld(0) + rejects);
                    acc.init(A.pulses, cell * 8u, 0, row_bytes);
                    acc.st(0, acc.ld(0) + pulses);
                    SimkRows rr;
                    rr.init(A.resume, cell * 8u, row_bytes, 5u * row_bytes);
                    rr.st(0, t);
                    rr.st(1, applied);
                    rr.st(2, (double)rep);
                    rr.st(3, (double)it);
                    rr.st(4, SIMK_RESUME_VALID + (reject_last ? 1.0 : 0.0) + (reject_more ? 2.0 : 0.0));
                    // Wave-aggregated append to the carry list.
                    const simk_u64 who = __builtin_amdgcn_read_exec();
                    const simk_u32 rank = __builtin_amdgcn_mbcnt_hi((simk_u32)(who >> 32),
                                                                    __builtin_amdgcn_mbcnt_lo((simk_u32)who, 0u));
                    simk_u32 base = 0;
                    if (rank == 0) base = atomicAdd(A.carry_count, (simk_u32)__builtin_popcountll(who));
                    const simk_u32 leader = (simk_u32)__builtin_ctzll(who);
                    base = (simk_u32)__builtin_amdgcn_ds_bpermute((int)(leader << 2), (int)base);
                    A.carry[base + rank] = cell;
                    phase = SIMK_PH_EXIT;
                } else if (++it > ctl.max_steps) {
                    failed = 1.0;
                    phase = SIMK_PH_END;
                } else {
                    phase = SIMK_PH_ATTEMPT;
                }
            } else {   // SIMK_PH_END
                // The segment's pending flag, recomputed (applied is unchanged
                // since SEGMENT) rather than carried as a lane mask across the
                // wave's divergent loop exits.
                const bool pending = applied < count;
                const bool alive = !(failed > 0.0);
                const bool hit = pending && alive && (t >= seg_end);
                if (hit) {
#pragma unroll
                    for (simk_u32 s = 0; s < n; ++s) y.set(s, y.get(s) + p.pulse(s));
                    applied = applied + 1.0;
                    pulses = pulses + 1.0;
                    h = ctl.h_start;
                }
                // The reference's exit test; "!pending" is the same for every
                // finite count and also ends the loop for a NaN count.
                const bool done = ((t >= A.duration) && (applied >= count || !pending)) || !alive;
                if (!done) {
                    phase = SIMK_PH_SEGMENT;
                } else if (++rep < A.repeats) {
                    t = 0.0;
                    applied = 0.0;
                    phase = SIMK_PH_SEGMENT;
                } else {
#pragma unroll
                    for (simk_u32 s = 0; s < n; ++s) ys.st(s, y.get(s));
                    hs_.st(0, h);
                    fs.st(0, failed);
                    SimkRows acc;
                    acc.init(A.steps, cell * 8u, 0, row_bytes);
                    acc.st(0, acc.ld(0) + steps);
                    acc.init(A.rejects, cell * 8u, 0, row_bytes);
                    acc.st(0, acc.ld(0) + rejects);
                    acc.init(A.pulses, cell * 8u, 0, row_bytes);
                    acc.st(0, acc.ld(0) + pulses);
                    phase = SIMK_PH_CELL;
                }
            }
        }
        if (__builtin_amdgcn_ballot_w64(phase == SIMK_PH_ATTEMPT) == 0) break;
        if (phase == SIMK_PH_ATTEMPT) {
            simk_attempt(p, me, ctl, y, t, seg_end, h, steps, rejects, reject_last, reject_more);
            phase = (t < seg_end) ? SIMK_PH_CHECK : SIMK_PH_END;
        }
    }
}

// Several mechanisms in one launch (ragged batches): per-mechanism arguments
// and the first wave of each mechanism's starting range.
struct SimkMultiArgs {
    SimkKinArgs a[4];
    simk_u32 count;
    simk_u32 start[4];       // first wave that starts at position k of seq
    simk_u32 seq[4];         // the order in which a wave visits the mechanisms' queues
};
