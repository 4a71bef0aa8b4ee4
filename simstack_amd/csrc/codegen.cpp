// Straight-line kernel generator (see codegen.hpp).
//
// Emitted per mechanism:
//   fun         f = src; per reaction: rate = k_r * y_a [* y_b]; f_s = fma(coef, rate, f_s)
//               in reaction order (kin_codegen.cpp:58-72 order)
//   jac_factor  row by row: the row's -J entries accumulated in (reaction,
//               reactant, net entry) order starting from +0.0 (kin_solver.h:75-93),
//               then the row's eliminations against finished rows and its
//               reciprocal pivot (kin_solver.h:95-113), then the row is stored.
//               Row i of -J depends only on y, and row i of the factor reads only
//               rows k < i, so fusing the two per row gives the table walk's bits.
//   solve       forward over lower (ascending col), backward over upper
//               (descending col) then the reciprocal pivot (kin_solver.h:122-139)
// y, f and the stage vector being solved live in registers; the LU and earlier
// stage vectors live in the lane's workspace slots ([wave][slot][64] f64, so
// each access is one coalesced 512-byte wave transaction). Optionally the
// first LU entries live in LDS instead ([entry][64] f64 per wave; a lane only
// ever touches its own column, so no barrier is needed).
#include "codegen.hpp"

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <map>
#include <vector>

namespace simk {

namespace {

struct Out {
    std::string s;
    void operator()(const char *fmt, ...) __attribute__((format(printf, 2, 3))) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        s += buf;
    }
};

std::string lit(f64 v) {
    char buf[64];
    snprintf(buf, sizeof(buf), "%a", v);   // exact hexadecimal literal
    return buf;
}

// One generated statement that may consume a loaded value. The load is
// emitted ahead of the statement by the pipeline below so that many loads are
// in flight before the first use (a wave otherwise waits for each load).
struct Op {
    std::string load;      // "" when the statement needs no load
    std::string stmt;      // "%s" stands for the loaded value
    bool barrier = false;  // loads never move across it (fences, stores they depend on)
    bool immutable = false;   // the load reads a slot no statement writes (a rate constant)
    int slot = -1;         // paired workspace: the slot this op's load reads (W(slot)), else -1
    int from = -1;         // the value is half of op `from`'s pair load (this op loads nothing)
    std::string part;      // ".x" / ".y": the half of a pair load this op uses
    std::string type;      // the loaded value's type when not the segment's default
};

// Paired workspace: an op loading slot s and one of the next `lookahead` ops
// loading its partner s ^ 1 (no barrier between) share one 16-byte load.
void pair_loads(std::vector<Op> &ops, size_t lookahead = 4, const char *pair_type = "SimkD2",
                const char *accessor = "W2") {
    for (size_t i = 0; i < ops.size(); ++i) {
        Op &a = ops[i];
        if (a.slot < 0 || a.from >= 0 || !a.type.empty() || a.barrier) continue;
        for (size_t j = i + 1; j < ops.size() && j <= i + lookahead; ++j) {
            Op &b = ops[j];
            if (b.barrier) break;
            if (b.slot != (a.slot ^ 1) || b.from >= 0 || !b.type.empty()) continue;
            char buf[32];
            snprintf(buf, sizeof(buf), "%s(%d)", accessor, a.slot >> 1);
            a.load = buf;
            a.type = pair_type;
            a.part = (a.slot & 1) ? ".y" : ".x";
            a.immutable = false;   // a pair load is never shared by the reuse window
            b.load.clear();
            b.from = (int)i;
            b.part = (b.slot & 1) ? ".y" : ".x";
            break;
        }
    }
}

// f32 slots in groups of four ([quad][64 lanes][4]): an op loading slot 4q
// and ops loading 4q + 1 .. 4q + 3 within the next `lookahead` ops (no
// barrier between) share one 16-byte load.
void quad_loads(std::vector<Op> &ops, size_t lookahead, const char *quad_type, const char *accessor) {
    static const char *kPart[4] = {".x", ".y", ".z", ".w"};
    for (size_t i = 0; i < ops.size(); ++i) {
        Op &a = ops[i];
        if (a.slot < 0 || (a.slot & 3) != 0 || a.from >= 0 || !a.type.empty() || a.barrier) continue;
        size_t at[4] = {i, 0, 0, 0};
        int found = 1;
        for (size_t j = i + 1; j < ops.size() && j <= i + lookahead && found < 4; ++j) {
            Op &b = ops[j];
            if (b.barrier) break;
            const int k = b.slot - a.slot;
            if (k < 1 || k > 3 || b.from >= 0 || !b.type.empty() || at[k]) continue;
            at[k] = j;
            ++found;
        }
        if (found < 4) continue;
        char buf[32];
        snprintf(buf, sizeof(buf), "%s(%d)", accessor, a.slot >> 2);
        a.load = buf;
        a.type = quad_type;
        a.part = kPart[0];
        a.immutable = false;
        for (int k = 1; k < 4; ++k) {
            Op &b = ops[at[k]];
            b.load.clear();
            b.from = (int)i;
            b.part = kPart[k];
        }
    }
}

// Emits ops with their loads software-pipelined `batch` ops ahead: the loads
// of ops [b + batch, b + 2 batch) are issued before the statements of
// [b, b + batch) run. Every loaded value is its own SSA name, except that an
// immutable load issued at most `reuse` ops earlier (in any segment) is not
// repeated: the op uses the earlier value.
void emit_ops(Out &w, const std::vector<Op> &ops, u32 batch, const char *indent, u32 &uid, u32 reuse = 0,
              const char *type = "double") {
    if (batch == 0) batch = 1;
    std::map<std::string, std::pair<std::string, size_t>> recent;   // load -> (value, op index)
    size_t seg = 0;
    while (seg < ops.size()) {
        size_t end = seg;
        while (end < ops.size() && !ops[end].barrier) ++end;
        // ops [seg, end) form one pipelined segment
        std::vector<std::string> name(end - seg);
        size_t issued = seg;
        auto issue_until = [&](size_t limit) {
            for (; issued < limit && issued < end; ++issued) {
                const Op &op = ops[issued];
                if (op.load.empty()) continue;
                if (reuse && op.immutable) {
                    auto it = recent.find(op.load);
                    if (it != recent.end() && issued - it->second.second <= reuse) {
                        name[issued - seg] = it->second.first;
                        it->second.second = issued;   // the value stays live until this use
                        continue;
                    }
                }
                char buf[32];
                snprintf(buf, sizeof(buf), "q%u", uid++);
                name[issued - seg] = buf;
                w("%sconst %s %s = %s;\n", indent, op.type.empty() ? type : op.type.c_str(), buf, op.load.c_str());
                if (reuse && op.immutable) recent[op.load] = {buf, issued};
            }
        };
        issue_until(seg + 2 * batch);
        for (size_t i = seg; i < end; ++i) {
            if ((i - seg) % batch == 0) issue_until(i + 2 * batch);
            const Op &o = ops[i];
            std::string st = o.stmt;
            if (st.empty()) continue;
            size_t at = st.find("%s");
            if (at != std::string::npos) st.replace(at, 2, (o.from >= 0 ? name[o.from - seg] : name[i - seg]) + o.part);
            w("%s%s\n", indent, st.c_str());
        }
        if (end < ops.size()) w("%s%s\n", indent, ops[end].stmt.c_str());   // the barrier itself
        seg = end + 1;
    }
}

}  // namespace

u32 lds_entries(const Mech &m, const GenOptions &o) {
    if (o.precision == 3) {
        // FP32 factor storage: the n FP64 pivots always sit in LDS (two f32
        // words each). lds_slots is the lane's LDS budget in f32 words; the
        // off-diagonal f32 entries fill what the pivots leave, in compact
        // order, up to 160 KiB per wave. Returns the off-diagonal entries in LDS.
        const u32 off = m.nnz - m.n, piv = 2u * m.n;
        const u32 budget = o.lds_slots < 640u ? o.lds_slots : 640u;
        const u32 avail = budget > piv ? budget - piv : 0u;
        return off < avail ? off : avail;
    }
    // 160 KiB of LDS per workgroup (gfx950): 320 f64 or 640 f32 entries of 64 lanes.
    const u32 most = o.precision != 0 ? 640u : 320u;
    const u32 cap = m.nnz < most ? m.nnz : most;
    return o.lds_slots < cap ? o.lds_slots : cap;
}

// Per-lane workspace layout (f64 slots): FP64 [LU entries lds.. | K stages |
// f | k copy | src copy]; refined mixed [K stages | f | refinement rhs | FP32
// LU entries lds.. (two per slot) | k copy | src copy]; unrefined mixed [f |
// FP32 region: LU entries lds.., then the K stages | k copy | src copy]. The
// last stage's vector stays in registers. Without refinement every stage
// vector is the f64 image of an FP32 solve result, so storing it as FP32 keeps
// every bit.
struct Layout {
    u32 lds, koff, foff, rhsoff, lu32off, k32off, kcp, scp, slots;
    bool k32;
    // FP32 factor storage (precision 3): lds_rank[idx] is the LDS f32 row of
    // an off-diagonal entry held in LDS (~0u otherwise); lu_slot[idx] the f32
    // slot of one outside LDS; pivots are in LDS by row.
    bool s32 = false;
    std::vector<u32> lds_rank;
    // paired layouts: a second copy of the rate constants in the paired
    // region (kpoff, even; ~0u: none), which f passes read in pairs
    u32 kpoff = ~0u;
    // Paired FP64 workspace: [paired region: LU entries lds.. in the solves'
    // read order (forward: each row's lower entries; backward, rows n-1..0:
    // upper entries then the pivot), K stages and f at stride np (n rounded
    // up to even), then (k_pairs) the paired copy of the rate constants at
    // kpoff | unpaired region: k copy, src copy]; kcp and scp are offsets in
    // the unpaired region. Otherwise np = n and paired = 0.
    bool pairs = false;
    u32 np = 0, paired = 0;
    std::vector<u32> lu_slot;   // workspace slot of LU entry idx >= lds
};

// FP32 factor storage (precision 3): [paired region: K stages and f at
// stride np, then (k_pairs) the paired copy of the rate constants | unpaired
// region: k copy, src copy | f32 region: off-diagonal LU entries outside LDS
// in the solves' read order (forward: each row's lower entries; backward,
// rows n-1..0: upper entries), [quad][64 lanes][4]]. Without pairs: single
// slots [K stages | f | k copy | src copy | f32 entries, two per f64 slot].
static Layout layout_s32(const Mech &m, const KinMethod &me, const GenOptions &o) {
    Layout L;
    L.s32 = true;
    L.k32 = false;
    L.k32off = 0;
    const u32 stored_stages = me.stages > 0 ? me.stages - 1 : 0;
    L.lds = lds_entries(m, o);
    L.lds_rank.assign(m.nnz, ~0u);
    L.lu_slot.assign(m.nnz, 0);
    std::vector<bool> diag(m.nnz, false);
    for (u32 i = 0; i < m.n; ++i) diag[m.diag[i]] = true;
    u32 q = 0;
    for (u32 idx = 0; idx < m.nnz && q < L.lds; ++idx)
        if (!diag[idx]) L.lds_rank[idx] = q++;
    u32 t = 0;
    for (u32 i = 0; i < m.n; ++i)
        for (u32 e = m.lower_begin[i]; e < m.lower_begin[i + 1]; ++e)
            if (L.lds_rank[m.lower[e].index] == ~0u) L.lu_slot[m.lower[e].index] = t++;
    for (u32 ii = m.n; ii-- > 0;)
        for (u32 e = m.upper_begin[ii]; e < m.upper_begin[ii + 1]; ++e)
            if (L.lds_rank[m.upper[e].index] == ~0u) L.lu_slot[m.upper[e].index] = t++;
    L.pairs = o.pair_slots != 0;
    L.np = L.pairs ? m.n + (m.n & 1u) : m.n;
    L.koff = 0;
    L.foff = stored_stages * L.np;
    if (L.pairs) {
        L.paired = L.foff + L.np;
        if (o.k_pairs) {
            L.kpoff = L.paired;
            L.paired += m.R + (m.R & 1u);
        }
        L.kcp = 0;
        L.scp = m.R;
        L.lu32off = L.paired + m.R + m.n;
    } else {
        L.paired = 0;
        L.kcp = L.foff + m.n;
        L.scp = L.kcp + m.R;
        L.lu32off = L.scp + m.n;
    }
    L.rhsoff = L.lu32off;
    // f32 region: [quad][64 lanes][4] with pairs (16 bytes, two f64 slots, per quad), else [slot][64 lanes]
    L.slots = L.lu32off + (L.pairs ? 2 * ((t + 3) / 4) : (t + 1) / 2);
    return L;
}

static Layout layout(const Mech &m, const KinMethod &me, const GenOptions &o) {
    if (o.precision == 3) return layout_s32(m, me, o);
    Layout L;
    const bool mixed = o.precision == 2;
    const u32 sweeps = mixed ? o.sweeps : 0;
    const u32 stored_stages = me.stages > 0 ? me.stages - 1 : 0;
    L.lds = lds_entries(m, o);
    L.k32 = mixed && sweeps == 0;
    L.k32off = 0;
    u32 end;
    if (!mixed) {
        L.koff = m.nnz - L.lds;
        L.foff = L.koff + stored_stages * m.n;
        L.rhsoff = L.lu32off = L.foff + m.n;
        end = L.foff + m.n;
    } else if (!L.k32) {
        L.koff = 0;
        L.foff = stored_stages * m.n;
        L.rhsoff = L.foff + m.n;
        L.lu32off = L.rhsoff + m.n;
        end = L.lu32off + (m.nnz - L.lds + 1) / 2;
    } else {
        L.koff = 0;
        L.foff = 0;
        L.rhsoff = L.lu32off = m.n;
        L.k32off = m.nnz - L.lds;
        end = L.lu32off + (L.k32off + stored_stages * m.n + 1) / 2;
    }
    L.kcp = end;
    L.scp = L.kcp + m.R;
    L.slots = L.scp + m.n;
    L.np = m.n;
    L.lu_slot.assign(m.nnz, 0);
    for (u32 idx = L.lds; idx < m.nnz; ++idx) L.lu_slot[idx] = idx - L.lds;
    if (!mixed && o.pair_slots) {
        L.pairs = true;
        u32 t = 0;
        for (u32 i = 0; i < m.n; ++i)
            for (u32 e = m.lower_begin[i]; e < m.lower_begin[i + 1]; ++e)
                if (m.lower[e].index >= L.lds) L.lu_slot[m.lower[e].index] = t++;
        t += t & 1u;   // the backward sweep starts a pair
        for (u32 ii = m.n; ii-- > 0;) {
            for (u32 e = m.upper_begin[ii]; e < m.upper_begin[ii + 1]; ++e)
                if (m.upper[e].index >= L.lds) L.lu_slot[m.upper[e].index] = t++;
            if (m.diag[ii] >= L.lds) L.lu_slot[m.diag[ii]] = t++;
        }
        t += t & 1u;
        L.np = m.n + (m.n & 1u);
        L.koff = t;
        L.foff = L.koff + stored_stages * L.np;
        L.paired = L.foff + L.np;
        if (o.k_pairs) {
            L.kpoff = L.paired;
            L.paired += m.R + (m.R & 1u);
        }
        L.rhsoff = L.lu32off = L.paired;
        L.kcp = 0;
        L.scp = m.R;
        L.slots = L.paired + m.R + m.n;
    }
    return L;
}

// Measured on the MI355X. FP64 (48/150, 24/70, the DBD air mechanism and
// the three-mechanism kernel): two waves per CU with 80 KiB of LU each beat
// every grid with less LDS per wave by 20-100 %; more than 80 KiB leaves one
// wave per CU and loses. f32 delta lanes (fewer VGPRs): 40 KiB per wave and
// four waves per CU run the DBD air workload 1.9x faster than its whole LU
// (54 KiB, three waves per CU). Unrefined mixed precision (48/150, 262,144
// cells, 24-deep load pipelining, rate constants reused over the pass): 200
// f32 entries and three waves per CU run 8 % faster than 160 entries and
// four waves, and 1.5x faster than 320 entries and two waves. The refined
// mixed kernel is slower with LDS entries.
u32 auto_lds(const GenOptions &o) {
    if (o.precision == 1) return 160u;
    // FP32 factor storage, f32 words per lane with the pivots: 304 (76 KiB per
    // wave, 208 off-diagonal entries of 48/150 in LDS) runs 2-3 % faster than
    // 320 (80 KiB) on 48/150 at 262,144 cells, in repeated runs; 288 to 316
    // scatter between the two; above 320 one wave fits a CU and loses 35 %.
    if (o.precision == 3) return 304u;
    if (o.precision == 2) return o.sweeps ? 0u : 200u;
    return 160u;
}

u32 auto_batch(const GenOptions &o) { return o.precision == 1 || (o.precision == 2 && o.sweeps) ? 16u : 24u; }

// Unrefined mixed precision, 48/150 at 262,144 cells: reusing every earlier
// rate-constant load of a pass (the -J pass then loads 152 instead of 274)
// is 2 % faster than a 96-statement window, with no extra registers. FP64,
// one mechanism: 384 statements, 3 % faster than 96. The multi-mechanism
// kernel keeps 96 (larger windows spill there and run 25 % slower).
u32 auto_reuse(const GenOptions &o, bool multi) {
    if (o.precision == 2 && o.sweeps == 0) return 1000u;
    return (o.precision == 0 || o.precision == 3) && !multi ? 384u : 96u;
}

u32 auto_fun_in_jac(const GenOptions &) { return 1u; }

// FP64 kernels, paired against single workspace slots (MI355X, bit-identical
// results): 48/150 at 65,536 cells 529 -> 512 ms (+3 %), at 262,144 cells
// unchanged (1,674 / 1,671 ms); 60/200 at 65,536 cells +4.6 %; the DBD air
// mechanism at 65,536 zones 16.0 -> 14.2 ms (+13 %); the three-mechanism
// kernel at 221,000 cells 318 -> 286 ms (+11 %). Mixed-precision and f32
// delta kernels keep their own layouts.
u32 auto_pair(const GenOptions &o) { return o.precision == 0 || o.precision == 3 ? 1u : 0u; }

// Paired layouts, f passes reading a paired copy of the rate constants (bits
// unchanged): 48/150 at 262,144 cells, the FP64 kernel 4 % faster (1,672 ->
// 1,601 ms) and the FP32 factor-storage kernel 3 % faster (1,121 -> 1,087
// ms); the three-mechanism kernel at 221,000 cells 1.5 % faster. The extra R
// slots per lane fit the cache-aware grid's budget at the same waves.
u32 auto_k_pairs(const GenOptions &o, bool multi) {
    (void)multi;
    return o.precision == 0 || o.precision == 3 ? 1u : 0u;
}

// Off by default; the autotuners try it. The carried vector stays live across
// the stage loop. On 48/150 at 262,144 cells, 20 x 1 s (same bits either
// way): FP64 1,596 -> 1,584 ms with the carry, although the code object then
// spills 116 bytes per lane; FP32 factor storage 1,092-1,099 -> 1,113-1,123
// ms (508 registers, no scratch), fewer bytes but a longer critical path.
u32 auto_stage_carry(const GenOptions &o, bool multi) {
    (void)o;
    (void)multi;
    return 0u;
}

void stage_plan(const KinMethod &me, u32 eval[4], u32 save_f[4]) {
    // f holds f(point of stage `holder`); at_y: that point is y itself.
    u32 holder = 0;
    bool at_y = true;
    for (u32 st = 0; st < 4; ++st) eval[st] = save_f[st] = 0;
    for (u32 st = 1; st < me.stages && st < 4; ++st) {
        bool reuse = !me.new_f[st];
        if (!reuse) {
            const u32 packed = st * (st - 1) / 2;
            bool all_zero = true;
            for (u32 j = 0; j < st; ++j)
                if (me.a[packed + j] != 0.0) all_zero = false;
            // With every a_ij zero the stage point is y: if f(y) is the f held,
            // recomputing it would reproduce the same bits.
            reuse = all_zero && at_y;
            if (!reuse) {
                eval[st] = 1;
                holder = st;
                at_y = all_zero;
            }
        }
        if (reuse) save_f[holder] = 1;
    }
}

// Emits one mechanism's Ops policy struct `sname`; returns its workspace slots per lane.
static u32 emit_ops_struct(Out &w, const Mech &m, const KinMethod &me, const GenOptions &o, const char *sname) {
    const u32 n = m.n;
    const bool mixed = o.precision == 2;
    // FP32 factor storage: FP64 arithmetic throughout; a finished row's
    // off-diagonal entries are rounded to f32 and kept as that value (in
    // registers too), stored as f32 (LDS or the f32 workspace region); the
    // reciprocal pivots stay FP64 in LDS.
    const bool s32 = o.precision == 3;
    const u32 sweeps = mixed ? o.sweeps : 0;
    // LU entries below lds are in LDS. The cell's rate constants and sources
    // are copied into the lane's slots (kcp, scp) when the lane takes the
    // cell, so every later read is a coalesced wave transaction whatever cells
    // the lanes of a wave hold.
    const Layout Lo = layout(m, me, o);
    const u32 lds = Lo.lds, koff = Lo.koff, foff = Lo.foff, rhsoff = Lo.rhsoff, lu32off = Lo.lu32off;
    const u32 kcp = Lo.kcp, scp = Lo.scp, ws_slots = Lo.slots;
    const bool pairs = Lo.pairs;
    const bool lu_pairs = pairs && !s32;   // FP64 LU entries in paired slots
    // FP32 factor storage with pairs: the f32 region is [quad][64 lanes][4],
    // four entries per 16-byte lane access (WQ4 / WQS4), two adjacent ones
    // per 8-byte access (WQ2 / WQS2)
    const bool q_pairs = pairs && s32;
    w("struct %s {\n", sname);
    w("    static constexpr simk_u32 N = %u, R = %u;\n", n, m.R);
    w("    static constexpr simk_u32 KOFF = %u, FOFF = %u, RHSOFF = %u, STAGES = %u, K32OFF = %u;\n", koff, foff,
      rhsoff, me.stages, Lo.k32off);
    w("    static constexpr simk_u32 NP = %u, PAIRED = %u;\n", Lo.np, Lo.paired);
    // jac_factor also evaluates f(y) and leaves it in the f slots (fij).
    const bool fij = o.fun_in_jac != 0;
    w("    static constexpr bool kFunInJac = %s;\n", fij ? "true" : "false");
    w("    static constexpr bool kStageCarry = %s;\n", o.stage_carry ? "true" : "false");
    w("    using Vec = SimkRegVec<N>;\n");
    w("    SimkRows kr, sr, pr, ws;\n");
    if (pairs) w("    SimkRowPairs wp;\n");
    // srcnz: the lane's cell has a source row with a bit set. A wave whose
    // lanes all have all-zero (+0.0) rows starts f from the literal +0.0,
    // which is what loading those rows would give.
    w("    simk_u32 srcnz;\n");
    if (mixed || (s32 && !q_pairs)) w("    SimkRows32 w32;\n");
    if (q_pairs) w("    SimkRowQuads32 wq;\n");
    if (s32) {
        w("    simk_u32 ln;\n");
        w("    SIMK_DEV double P(simk_u32 i) const { return simk_lds[i * SIMK_WAVE + ln]; }\n");
        w("    SIMK_DEV void PS(simk_u32 i, double v) const { simk_lds[i * SIMK_WAVE + ln] = v; }\n");
        if (lds) {
            w("    SIMK_DEV float L32(simk_u32 i) const { return simk_lds32[i * SIMK_WAVE + ln]; }\n");
            w("    SIMK_DEV void LS32(simk_u32 i, float v) const { simk_lds32[i * SIMK_WAVE + ln] = v; }\n");
        }
    } else if (lds) {
        const char *T = mixed ? "float" : "double", *A = mixed ? "simk_lds32" : "simk_lds";
        w("    simk_u32 ln;\n");
        w("    SIMK_DEV %s L(simk_u32 i) const { return %s[i * SIMK_WAVE + ln]; }\n", T, A);
        w("    SIMK_DEV void LS(simk_u32 i, %s v) const { %s[i * SIMK_WAVE + ln] = v; }\n", T, A);
    }
    w("    SIMK_DEV void init(const SimkKinArgs &A, simk_u32 wave, simk_u32 lane) {\n");
    if (lds || s32) w("        ln = lane;\n");
    w("        const simk_u32 stride = A.padded * 8u;\n");
    w("        kr.init(A.k, 0, stride, R * stride);\n");
    w("        sr.init(A.src, 0, stride, N * stride);\n");
    w("        pr.init(A.pulse, 0, stride, N * stride);\n");
    if (pairs) {
        w("        wp.init(A.ws + (simk_u64)wave * A.ws_slots * SIMK_WAVE, lane, PAIRED * SIMK_WAVE * 8u);\n");
        w("        ws.init(A.ws + ((simk_u64)wave * A.ws_slots + PAIRED) * SIMK_WAVE, lane * 8u, SIMK_WAVE * 8u, (R + N) * SIMK_WAVE * 8u);\n");
    } else {
        w("        ws.init(A.ws + (simk_u64)wave * A.ws_slots * SIMK_WAVE, lane * 8u, SIMK_WAVE * 8u, A.ws_slots * SIMK_WAVE * 8u);\n");
    }
    if (mixed || (s32 && !q_pairs))
        w("        w32.init(A.ws + ((simk_u64)wave * A.ws_slots + %u) * SIMK_WAVE, lane * 4u, SIMK_WAVE * 4u, %u * SIMK_WAVE * 8u);\n",
          lu32off, ws_slots - lu32off);
    if (q_pairs)
        w("        wq.init(A.ws + ((simk_u64)wave * A.ws_slots + %u) * SIMK_WAVE, lane, %u * SIMK_WAVE * 8u);\n", lu32off,
          ws_slots - lu32off);
    w("    }\n");
    w("    SIMK_DEV void set_cell(const SimkKinArgs &, simk_u32 cell) {\n");
    w("        kr.voff = sr.voff = pr.voff = cell * 8u;\n");
    if (Lo.kpoff != ~0u) {
        // both copies: single slots for the -J pass, pairs for the f passes
        w("#pragma unroll 4\n        for (simk_u32 r = 0; r + 1 < R; r += 2) {\n"
          "            const double a = kr.ld(r), b = kr.ld(r + 1);\n"
          "            ws.st(%u + r, a); ws.st(%u + r + 1, b); WS2((%u + r) / 2, a, b);\n        }\n", kcp, kcp, Lo.kpoff);
        if (m.R & 1u)
            w("        { const double a = kr.ld(R - 1); ws.st(%u + R - 1, a); WS(%u + R - 1, a); }\n", kcp, Lo.kpoff);
    } else {
        w("#pragma unroll 8\n        for (simk_u32 r = 0; r < R; ++r) ws.st(%u + r, kr.ld(r));\n", kcp);
    }
    w("        srcnz = 0;\n");
    w("#pragma unroll 8\n        for (simk_u32 s = 0; s < N; ++s) { const double v = sr.ld(s); srcnz |= "
      "(simk_u32)(__builtin_bit_cast(simk_u64, v) != 0); ws.st(%u + s, v); }\n", scp);
    w("    }\n");
    w("    SIMK_DEV bool any_src() const { return __builtin_amdgcn_ballot_w64(srcnz != 0) != 0; }\n");
    w("    static SIMK_DEV constexpr simk_u32 n() { return N; }\n");
    w("    SIMK_DEV Vec vec(int) const { return Vec(); }\n");
    w("    SIMK_DEV double K(simk_u32 r) const { return ws.ld(%u + r); }\n", kcp);
    w("    SIMK_DEV double S(simk_u32 s) const { return ws.ld(%u + s); }\n", scp);
    w("    SIMK_DEV double pulse(simk_u32 s) const { return pr.ld(s); }\n");
    if (pairs) {
        w("    SIMK_DEV double W(simk_u32 i) const { return wp.ld(i); }\n");
        w("    SIMK_DEV void WS(simk_u32 i, double v) const { wp.st(i, v); }\n");
        w("    SIMK_DEV SimkD2 W2(simk_u32 pair) const { return wp.ld2(pair); }\n");
        w("    SIMK_DEV void WS2(simk_u32 pair, double a, double b) const { wp.st2(pair, a, b); }\n");
    } else {
        w("    SIMK_DEV double W(simk_u32 i) const { return ws.ld(i); }\n");
        w("    SIMK_DEV void WS(simk_u32 i, double v) const { ws.st(i, v); }\n");
    }
    if (q_pairs) {
        w("    SIMK_DEV float W32(simk_u32 i) const { return wq.ld(i); }\n");
        w("    SIMK_DEV void W32S(simk_u32 i, float v) const { wq.st(i, v); }\n");
        w("    SIMK_DEV SimkF2 WQ2(simk_u32 pair) const { return wq.ld2(pair); }\n");
        w("    SIMK_DEV void WQS2(simk_u32 pair, float a, float b) const { wq.st2(pair, a, b); }\n");
        w("    SIMK_DEV SimkF4 WQ4(simk_u32 quad) const { return wq.ld4(quad); }\n");
        w("    SIMK_DEV void WQS4(simk_u32 quad, float a, float b, float c, float d) const { wq.st4(quad, a, b, c, d); }\n");
    } else if (mixed || s32) {
        w("    SIMK_DEV float W32(simk_u32 i) const { return w32.ld(i); }\n");
        w("    SIMK_DEV void W32S(simk_u32 i, float v) const { w32.st(i, v); }\n");
    }
    if (Lo.k32) {
        w("    SIMK_DEV double kget(simk_u32 st, simk_u32 s) const { return (double)W32(K32OFF + st * N + s); }\n");
        w("    SIMK_DEV void kstore(simk_u32 st, const Vec &x) const {\n");
        w("        if (st + 1 >= STAGES) return;\n");
        w("#pragma unroll\n        for (simk_u32 s = 0; s < N; ++s) W32S(K32OFF + st * N + s, (float)x.v[s]);\n    }\n");
    } else if (pairs) {
        // stage vectors at even stride NP: slots s, s + 1 of a vector are a pair
        w("    SIMK_DEV double kget(simk_u32 st, simk_u32 s) const { return W(KOFF + st * NP + s); }\n");
        w("    SIMK_DEV void kstore(simk_u32 st, const Vec &x) const {\n");
        w("        if (st + 1 >= STAGES) return;\n");
        w("#pragma unroll\n        for (simk_u32 s = 0; s + 1 < N; s += 2) WS2((KOFF + st * NP + s) / 2, x.v[s], x.v[s + 1]);\n");
        w("        if (N & 1u) WS(KOFF + st * NP + N - 1, x.v[N - 1]);\n    }\n");
    } else {
        w("    SIMK_DEV double kget(simk_u32 st, simk_u32 s) const { return W(KOFF + st * N + s); }\n");
        w("    SIMK_DEV void kstore(simk_u32 st, const Vec &x) const {\n");
        w("        if (st + 1 >= STAGES) return;\n");
        w("#pragma unroll\n        for (simk_u32 s = 0; s < N; ++s) WS(KOFF + st * N + s, x.v[s]);\n    }\n");
    }
    {
        // v += c K_j, loads software-pipelined like the solve's.
        std::vector<Op> ops;
        char buf[96];
        for (u32 s2 = 0; s2 < n; ++s2) {
            Op op;
            if (pairs && s2 % 2 == 1 && s2 < n) {
                op.from = (int)ops.size() - 1;   // the odd slot of the pair loaded for s2 - 1
                op.part = ".y";
            } else if (pairs && s2 + 1 < n) {
                snprintf(buf, sizeof(buf), "W2((KOFF + j * NP + %u) / 2)", s2);
                op.load = buf;
                op.type = "SimkD2";
                op.part = ".x";
            } else {
                snprintf(buf, sizeof(buf), Lo.k32 ? "(double)W32(K32OFF + j * N + %u)"
                                                  : (pairs ? "W(KOFF + j * NP + %u)" : "W(KOFF + j * N + %u)"), s2);
                op.load = buf;
            }
            snprintf(buf, sizeof(buf), "v.v[%u] = simk_fma(c, %%s, v.v[%u]);", s2, s2);
            op.stmt = buf;
            ops.push_back(op);
        }
        w("    SIMK_DEV void kaxpy(simk_u32 j, double c, Vec &v) const {\n");
        u32 uid = 0;
        emit_ops(w, ops, o.batch, "        ", uid);
        w("    }\n");
    }
    // fjac: f(y) as jac_factor left it in the f slots. The saved f stays in
    // workspace slots: keeping it in registers was measured slower (the
    // register pressure costs more than the slot traffic it saves).
    if (pairs) {
        const char *ld = "#pragma unroll\n        for (simk_u32 s = 0; s + 1 < N; s += 2) { const SimkD2 t = W2((FOFF + s) / 2); "
                         "x.v[s] = t.x; x.v[s + 1] = t.y; }\n        if (N & 1u) x.v[N - 1] = W(FOFF + N - 1);\n";
        w("    SIMK_DEV void fjac(Vec &x) const {\n%s    }\n", ld);
        w("    SIMK_DEV void fsave(const Vec &x) const {\n");
        w("#pragma unroll\n        for (simk_u32 s = 0; s + 1 < N; s += 2) WS2((FOFF + s) / 2, x.v[s], x.v[s + 1]);\n");
        w("        if (N & 1u) WS(FOFF + N - 1, x.v[N - 1]);\n    }\n");
        w("    SIMK_DEV void fload(Vec &x) const {\n%s    }\n", ld);
    } else {
        w("    SIMK_DEV void fjac(Vec &x) const {\n");
        w("#pragma unroll\n        for (simk_u32 s = 0; s < N; ++s) x.v[s] = W(FOFF + s);\n    }\n");
        w("    SIMK_DEV void fsave(const Vec &x) const {\n");
        w("#pragma unroll\n        for (simk_u32 s = 0; s < N; ++s) WS(FOFF + s, x.v[s]);\n    }\n");
        w("    SIMK_DEV void fload(Vec &x) const {\n");
        w("#pragma unroll\n        for (simk_u32 s = 0; s < N; ++s) x.v[s] = W(FOFF + s);\n    }\n");
    }

    // fun: one op per reaction, rate = k * y_a [* y_b], then its net entries
    // in order (kin_codegen.cpp:58-72 order)
    auto fun_ops = [&](const char *yv, const char *fv) {
        std::vector<Op> ops;
        char buf[96];
        for (u32 r = 0; r < m.R; ++r) {
            const Reaction &x = m.rx[r];
            Op op;
            if (Lo.kpoff != ~0u) {
                snprintf(buf, sizeof(buf), "W(%u)", Lo.kpoff + r);
                op.slot = (int)(Lo.kpoff + r);
            } else {
                snprintf(buf, sizeof(buf), "K(%u)", r);
                op.immutable = true;
            }
            op.load = buf;
            std::string st = "{ const double rate = %s";
            for (u32 p = 0; p < x.reactant_count; ++p) {
                snprintf(buf, sizeof(buf), " * %s.v[%u]", yv, x.reactant[p]);
                st += buf;
            }
            st += ";";
            for (u32 e = 0; e < x.net_count; ++e) {
                snprintf(buf, sizeof(buf), " %s.v[%u] = simk_fma(%d.0, rate, %s.v[%u]);", fv, x.net_species[e],
                         (int)x.net_coef[e], fv, x.net_species[e]);
                st += buf;
            }
            st += " }";
            op.stmt = st;
            ops.push_back(op);
        }
        if (Lo.kpoff != ~0u) pair_loads(ops);
        return ops;
    };
    const char *src_init = "        if (any_src()) {\n#pragma unroll\n            for (simk_u32 s = 0; s < N; ++s) %s.v[s] = S(s);\n"
                           "        } else {\n#pragma unroll\n            for (simk_u32 s = 0; s < N; ++s) %s.v[s] = 0.0;\n        }\n";
    {
        w("    SIMK_DEV void fun(const Vec &y, Vec &f) const {\n");
        w(src_init, "f", "f");
        u32 uid = 0;
        emit_ops(w, fun_ops("y", "f"), o.batch, "        ", uid, o.reuse);
        w("    }\n");
    }

    // jac_factor: rows in order. Register mode (fence_rows = 0): every entry is
    // one SSA value, later rows read finished rows' values directly (no reload),
    // and each row is stored as soon as it is final. Fenced mode re-reads
    // finished rows from the workspace.
    const bool regs = !o.fence_rows;
    // Row of every compact index (to decide register use vs reload).
    std::vector<u32> row_of(m.nnz, 0);
    for (u32 i = 0; i < n; ++i) {
        row_of[m.diag[i]] = i;
        for (u32 e = m.lower_begin[i]; e < m.lower_begin[i + 1]; ++e) row_of[m.lower[e].index] = i;
        for (u32 e = m.upper_begin[i]; e < m.upper_begin[i + 1]; ++e) row_of[m.upper[e].index] = i;
    }
    // In register mode a finished entry is read from its register when its
    // row is at most reg_window rows back (0: always), else reloaded from the
    // workspace after a fence, which bounds live ranges on dense mechanisms.
    const u32 window = o.reg_window == GenOptions::kDefaultWindow ? (mixed ? 8u : 16u) : o.reg_window;
    auto near = [&](u32 idx, u32 i) { return window == 0 || i - row_of[idx] <= window; };
    // LU entry idx: LDS entry idx, or workspace entry idx - lds (W: f64
    // slots; W32: the FP32 LU region of mixed kernels).
    // FP32 factor storage: pivot of row r at P(r); an off-diagonal entry at
    // LDS f32 row lds_rank[idx] or f32 workspace slot lu_slot[idx], widened on load.
    std::vector<u32> diag_row(m.nnz, ~0u);
    for (u32 i = 0; i < n; ++i) diag_row[m.diag[i]] = i;
    auto lu_ld = [&](u32 idx) {
        char buf[40];
        if (s32) {
            if (diag_row[idx] != ~0u) snprintf(buf, sizeof(buf), "P(%u)", diag_row[idx]);
            else if (Lo.lds_rank[idx] != ~0u) snprintf(buf, sizeof(buf), "L32(%u)", Lo.lds_rank[idx]);
            else snprintf(buf, sizeof(buf), "W32(%u)", Lo.lu_slot[idx]);
            return std::string(buf);
        }
        snprintf(buf, sizeof(buf), idx < lds ? "L(%u)" : (mixed ? "W32(%u)" : "W(%u)"), idx < lds ? idx : Lo.lu_slot[idx]);
        return std::string(buf);
    };
    // An f32 entry is loaded as a float (and widened, in FP32 factor storage,
    // where the statement uses it): a conversion at the load would wait for
    // the load there, ahead of the pipelined use.
    auto lu_type = [&](u32 idx) {
        if (q_pairs && diag_row[idx] == ~0u && Lo.lds_rank[idx] == ~0u) return std::string();   // typed after pairing
        return std::string(mixed || (s32 && diag_row[idx] == ~0u) ? "float" : "");
    };
    // the f32 workspace slot an op's load reads, for pairing (-1: none)
    auto q_slot = [&](u32 idx) {
        return q_pairs && diag_row[idx] == ~0u && Lo.lds_rank[idx] == ~0u ? (int)Lo.lu_slot[idx] : -1;
    };
    auto q_pair_loads = [&](std::vector<Op> &ops) {
        quad_loads(ops, 12, "SimkF4", "WQ4");
        pair_loads(ops, 4, "SimkF2", "WQ2");
        for (Op &op : ops)
            if (op.slot >= 0 && op.type.empty() && op.from < 0 && !op.load.empty()) op.type = "float";
    };
    auto lu_st = [&](u32 idx, const char *v) {
        char buf[48];
        if (s32) {
            if (diag_row[idx] != ~0u) snprintf(buf, sizeof(buf), "PS(%u, %s%u);", diag_row[idx], v, idx);
            else if (Lo.lds_rank[idx] != ~0u) snprintf(buf, sizeof(buf), "LS32(%u, (float)%s%u);", Lo.lds_rank[idx], v, idx);
            else snprintf(buf, sizeof(buf), "W32S(%u, (float)%s%u);", Lo.lu_slot[idx], v, idx);
            return std::string(buf);
        }
        snprintf(buf, sizeof(buf), idx < lds ? "LS(%u, %s%u);" : (mixed ? "W32S(%u, %s%u);" : "WS(%u, %s%u);"),
                 idx < lds ? idx : Lo.lu_slot[idx], v, idx);
        return std::string(buf);
    };
    {
        std::vector<Op> ops;
        auto stmt = [&](const std::string &t) {
            Op op;
            op.stmt = t;
            ops.push_back(op);
        };
        auto barrier = [&]() {
            Op op;
            op.stmt = "SIMK_MEMORY_FENCE();";
            op.barrier = true;
            ops.push_back(op);
        };
        char buf[160];
        for (u32 i = 0; i < n; ++i) {
            std::vector<u32> row;   // compact indices of row i
            for (u32 e = m.lower_begin[i]; e < m.lower_begin[i + 1]; ++e) row.push_back(m.lower[e].index);
            row.push_back(m.diag[i]);
            for (u32 e = m.upper_begin[i]; e < m.upper_begin[i + 1]; ++e) row.push_back(m.upper[e].index);
            bool far = false;
            for (u32 d = m.div_begin[i]; d < m.div_begin[i + 1]; ++d) {
                if (!near(m.lu_div[d].kk, i)) far = true;
                for (u32 s = m.lu_div[d].sub_begin; s < m.lu_div[d].sub_end; ++s)
                    if (!near(m.lu_sub[s].kj, i)) far = true;
            }
            if (!regs || far) barrier();
            snprintf(buf, sizeof(buf), "// row %u", i);
            stmt(buf);
            for (u32 idx : row) {
                snprintf(buf, sizeof(buf), "double a%u = 0.0;", idx);
                stmt(buf);
            }
            // f_i of f(y): the source, then every reaction with a net entry
            // on species i in ascending order, exactly fun's sequence for
            // species i (so f_i has fun's bits). Those are the reactions this
            // row's -J visits, so they share its rate-constant loads.
            if (fij) {
                snprintf(buf, sizeof(buf), "double f%u = anysrc ? S(%u) : 0.0;", i, i);
                stmt(buf);
            }
            // -J contributions, per entry in (r, p, e) order: one op per
            // reaction touching row i, its reactants' terms in p order.
            for (u32 r = 0; r < m.R; ++r) {
                const Reaction &x = m.rx[r];
                bool any = false;
                for (u32 e = 0; e < x.net_count; ++e)
                    if (x.net_species[e] == i) any = true;
                if (!any || (!fij && x.reactant_count == 0)) continue;
                Op op;
                snprintf(buf, sizeof(buf), "K(%u)", r);
                op.load = buf;
                op.immutable = true;
                std::string st = "{ const double k = %s;";
                if (fij) {
                    st += " { const double rate = k";
                    for (u32 p = 0; p < x.reactant_count; ++p) {
                        snprintf(buf, sizeof(buf), " * y.v[%u]", x.reactant[p]);
                        st += buf;
                    }
                    st += ";";
                    for (u32 e = 0; e < x.net_count; ++e) {
                        if (x.net_species[e] != i) continue;
                        snprintf(buf, sizeof(buf), " f%u = simk_fma(%d.0, rate, f%u);", i, (int)x.net_coef[e], i);
                        st += buf;
                    }
                    st += " }";
                }
                for (u32 p = 0; p < x.reactant_count; ++p) {
                    st += " { const double d = k";
                    for (u32 q = 0; q < x.reactant_count; ++q)
                        if (q != p) {
                            snprintf(buf, sizeof(buf), " * y.v[%u]", x.reactant[q]);
                            st += buf;
                        }
                    st += ";";
                    for (u32 e = 0; e < x.net_count; ++e) {
                        if (x.net_species[e] != i) continue;
                        u32 idx = x.jac_index[p][e];
                        snprintf(buf, sizeof(buf), " a%u = simk_fma(%d.0, d, a%u);", idx, -(int)x.net_coef[e], idx);
                        st += buf;
                    }
                    st += " }";
                }
                st += " }";
                op.stmt = st;
                ops.push_back(op);
            }
            // Eliminations against finished rows k < i: register values, or
            // workspace reloads (fenced mode / beyond the register window).
            // Mixed precision: the row is rounded to FP32 here and factored in
            // FP32 (b = f32 entries, FP32 fma).
            const char *V = mixed ? "b" : "a";
            const char *T = mixed ? "float" : "double";
            if (mixed)
                for (u32 idx : row) {
                    snprintf(buf, sizeof(buf), "float b%u = (float)a%u;", idx, idx);
                    stmt(buf);
                }
            for (u32 d = m.div_begin[i]; d < m.div_begin[i + 1]; ++d) {
                const KinLuDiv &dv = m.lu_div[d];
                const bool kk_reg = regs && near(dv.kk, i);
                Op op;
                if (!kk_reg) {
                    op.load = lu_ld(dv.kk);
                    op.type = lu_type(dv.kk);
                    if (q_pairs) op.slot = q_slot(dv.kk);
                    if (lu_pairs && dv.kk >= lds) op.slot = (int)Lo.lu_slot[dv.kk];
                    snprintf(buf, sizeof(buf), "const %s l%u = %s%u * %%s; %s%u = l%u;", T, d, V, dv.ik, V, dv.ik, d);
                } else {
                    snprintf(buf, sizeof(buf), "const %s l%u = %s%u * %s%u; %s%u = l%u;", T, d, V, dv.ik, V, dv.kk, V,
                             dv.ik, d);
                }
                op.stmt = buf;
                ops.push_back(op);
                for (u32 su = dv.sub_begin; su < dv.sub_end; ++su) {
                    const KinLuSub &sb = m.lu_sub[su];
                    Op o2;
                    const char *fn = mixed ? "__builtin_fmaf(-l" : "simk_fnma(l";
                    if (regs && near(sb.kj, i)) {
                        snprintf(buf, sizeof(buf), "%s%u = %s%u, %s%u, %s%u);", V, sb.ij, fn, d, V, sb.kj, V, sb.ij);
                    } else {
                        o2.load = lu_ld(sb.kj);
                        o2.type = lu_type(sb.kj);
                        if (q_pairs) o2.slot = q_slot(sb.kj);
                        if (lu_pairs && sb.kj >= lds) o2.slot = (int)Lo.lu_slot[sb.kj];
                        snprintf(buf, sizeof(buf), "%s%u = %s%u, %%s, %s%u);", V, sb.ij, fn, d, V, sb.ij);
                    }
                    o2.stmt = buf;
                    ops.push_back(o2);
                }
            }
            if (mixed) {
                snprintf(buf, sizeof(buf), "b%u = 1.0f / (b%u + gf);", m.diag[i], m.diag[i]);
                stmt(buf);
                for (u32 idx : row) stmt(lu_st(idx, "b"));
            } else if (s32) {
                snprintf(buf, sizeof(buf), "a%u = 1.0 / (a%u + g);", m.diag[i], m.diag[i]);
                stmt(buf);
                // the row is final: its off-diagonal entries become their f32
                // values, which later rows and the solves then use
                for (u32 idx : row)
                    if (idx != m.diag[i]) {
                        snprintf(buf, sizeof(buf), "a%u = (double)(float)a%u;", idx, idx);
                        stmt(buf);
                    }
                if (!q_pairs) {
                    for (u32 idx : row) stmt(lu_st(idx, "a"));
                } else {
                    // workspace entries by slot, partners in one 8-byte store
                    std::vector<std::pair<u32, u32>> ws_row;   // (slot, compact index)
                    for (u32 idx : row) {
                        if (diag_row[idx] != ~0u || Lo.lds_rank[idx] != ~0u) stmt(lu_st(idx, "a"));
                        else ws_row.push_back({Lo.lu_slot[idx], idx});
                    }
                    std::sort(ws_row.begin(), ws_row.end());
                    for (size_t q = 0; q < ws_row.size(); ++q) {
                        const u32 sl = ws_row[q].first;
                        if (sl % 4 == 0 && q + 3 < ws_row.size() && ws_row[q + 3].first == sl + 3) {
                            snprintf(buf, sizeof(buf), "WQS4(%u, (float)a%u, (float)a%u, (float)a%u, (float)a%u);", sl / 4,
                                     ws_row[q].second, ws_row[q + 1].second, ws_row[q + 2].second, ws_row[q + 3].second);
                            q += 3;
                        } else if (sl % 2 == 0 && q + 1 < ws_row.size() && ws_row[q + 1].first == sl + 1) {
                            snprintf(buf, sizeof(buf), "WQS2(%u, (float)a%u, (float)a%u);", sl / 2, ws_row[q].second,
                                     ws_row[q + 1].second);
                            ++q;
                        } else {
                            snprintf(buf, sizeof(buf), "W32S(%u, (float)a%u);", sl, ws_row[q].second);
                        }
                        stmt(buf);
                    }
                }
            } else {
                snprintf(buf, sizeof(buf), "a%u = 1.0 / (a%u + g);", m.diag[i], m.diag[i]);
                stmt(buf);
                if (!pairs) {
                    for (u32 idx : row) stmt(lu_st(idx, "a"));
                } else {
                    // the row's workspace entries by slot, partners in one 16-byte store
                    std::vector<std::pair<u32, u32>> ws_row;   // (slot, compact index)
                    for (u32 idx : row) {
                        if (idx < lds) stmt(lu_st(idx, "a"));
                        else ws_row.push_back({Lo.lu_slot[idx], idx});
                    }
                    std::sort(ws_row.begin(), ws_row.end());
                    for (size_t q = 0; q < ws_row.size(); ++q) {
                        const u32 sl = ws_row[q].first;
                        if (sl % 2 == 0 && q + 1 < ws_row.size() && ws_row[q + 1].first == sl + 1) {
                            snprintf(buf, sizeof(buf), "WS2(%u, a%u, a%u);", sl / 2, ws_row[q].second,
                                     ws_row[q + 1].second);
                            ++q;
                        } else {
                            snprintf(buf, sizeof(buf), "WS(%u, a%u);", sl, ws_row[q].second);
                        }
                        stmt(buf);
                    }
                }
            }
            if (fij && pairs) {
                // f in pairs: rows 2k and 2k + 1 stored together
                if (i % 2 == 1) snprintf(buf, sizeof(buf), "WS2((FOFF + %u) / 2, f%u, f%u);", i - 1, i - 1, i);
                else if (i + 1 == n) snprintf(buf, sizeof(buf), "WS(FOFF + %u, f%u);", i, i);
                else buf[0] = 0;
                if (buf[0]) stmt(buf);
            } else if (fij) {
                snprintf(buf, sizeof(buf), "WS(FOFF + %u, f%u);", i, i);
                stmt(buf);
            }
        }
        barrier();
        if (q_pairs) q_pair_loads(ops);
        else if (pairs) pair_loads(ops);
        w("    SIMK_DEV void jac_factor(const Vec &y, const double g) const {\n");
        if (fij) w("        const bool anysrc = any_src();\n");
        if (mixed) w("        const float gf = (float)g;\n");
        u32 uid = 0;
        emit_ops(w, ops, o.batch, "        ", uid, o.reuse);
        w("    }\n");
    }

    // solve: forward over lower, backward over upper and the reciprocal pivot.
    auto solve_ops = [&](const char *vec, bool f32) {
        std::vector<Op> ops;
        char buf[160];
        const char *FN = f32 ? "__builtin_fmaf(-%s" : "simk_fnma(%s";
        for (u32 i = 0; i < n; ++i)
            for (u32 e = m.lower_begin[i]; e < m.lower_begin[i + 1]; ++e) {
                Op op;
                op.load = lu_ld(m.lower[e].index);
                op.type = lu_type(m.lower[e].index);
                if (q_pairs) op.slot = q_slot(m.lower[e].index);
                if (lu_pairs && m.lower[e].index >= lds) op.slot = (int)Lo.lu_slot[m.lower[e].index];
                std::string f = FN;
                snprintf(buf, sizeof(buf), "%s[%u] = ", vec, i);
                std::string st = buf + f + ", ";
                snprintf(buf, sizeof(buf), "%s[%u], %s[%u]);", vec, m.lower[e].col, vec, i);
                op.stmt = st + buf;
                ops.push_back(op);
            }
        for (u32 ii = n; ii-- > 0;) {
            for (u32 e = m.upper_begin[ii]; e < m.upper_begin[ii + 1]; ++e) {
                Op op;
                op.load = lu_ld(m.upper[e].index);
                op.type = lu_type(m.upper[e].index);
                if (q_pairs) op.slot = q_slot(m.upper[e].index);
                if (lu_pairs && m.upper[e].index >= lds) op.slot = (int)Lo.lu_slot[m.upper[e].index];
                std::string f = FN;
                snprintf(buf, sizeof(buf), "%s[%u] = ", vec, ii);
                std::string st = buf + f + ", ";
                snprintf(buf, sizeof(buf), "%s[%u], %s[%u]);", vec, m.upper[e].col, vec, ii);
                op.stmt = st + buf;
                ops.push_back(op);
            }
            Op op;
            op.load = lu_ld(m.diag[ii]);
            op.type = lu_type(m.diag[ii]);
            if (lu_pairs && m.diag[ii] >= lds) op.slot = (int)Lo.lu_slot[m.diag[ii]];
            snprintf(buf, sizeof(buf), "%s[%u] = %s[%u] * %%s;", vec, ii, vec, ii);
            op.stmt = buf;
            ops.push_back(op);
        }
        if (q_pairs) q_pair_loads(ops);
        else if (pairs) pair_loads(ops);
        return ops;
    };
    if (!mixed) {
        w("    SIMK_DEV void solve(Vec &x, const Vec &, double) const {\n        SIMK_MEMORY_FENCE();\n");
        u32 uid = 0;
        emit_ops(w, solve_ops("x.v", false), o.batch, "        ", uid);
        w("    }\n};\n\n");
    } else {
        // FP32 triangular solves on a float vector.
        w("    SIMK_DEV void solve32(float *z) const {\n        SIMK_MEMORY_FENCE();\n");
        u32 uid = 0;
        emit_ops(w, solve_ops("z", true), o.batch, "        ", uid);
        w("    }\n");
        // r += J(y) x in FP64, matrix-free from the reactions (d = k prod_{q!=p} y_q).
        std::vector<Op> jx;
        char buf[200];
        for (u32 r = 0; r < m.R; ++r) {
            const Reaction &x = m.rx[r];
            Op op;
            snprintf(buf, sizeof(buf), "K(%u)", r);
            op.load = buf;
            std::string st = "{ const double kr = %s;";
            for (u32 p = 0; p < x.reactant_count; ++p) {
                st += " { const double t = kr";
                for (u32 q = 0; q < x.reactant_count; ++q)
                    if (q != p) {
                        snprintf(buf, sizeof(buf), " * y.v[%u]", x.reactant[q]);
                        st += buf;
                    }
                snprintf(buf, sizeof(buf), " * x.v[%u];", x.reactant[p]);
                st += buf;
                for (u32 e = 0; e < x.net_count; ++e) {
                    snprintf(buf, sizeof(buf), " r[%u] = simk_fma(%d.0, t, r[%u]);", x.net_species[e],
                             (int)x.net_coef[e], x.net_species[e]);
                    st += buf;
                }
                st += " }";
            }
            st += " }";
            op.stmt = st;
            jx.push_back(op);
        }
        w("    SIMK_DEV void solve(Vec &x, const Vec &y, const double g) const {\n");
        if (sweeps) w("#pragma unroll\n        for (simk_u32 s = 0; s < N; ++s) WS(RHSOFF + s, x.v[s]);\n");
        w("        float z[N];\n");
        w("#pragma unroll\n        for (simk_u32 s = 0; s < N; ++s) z[s] = (float)x.v[s];\n");
        w("        solve32(z);\n");
        w("#pragma unroll\n        for (simk_u32 s = 0; s < N; ++s) x.v[s] = (double)z[s];\n");
        if (sweeps) {
            w("#pragma unroll 1\n        for (simk_u32 sweep = 0; sweep < %u; ++sweep) {\n", sweeps);
            w("            double r[N];\n");
            w("#pragma unroll\n            for (simk_u32 s = 0; s < N; ++s) r[s] = simk_fnma(g, x.v[s], W(RHSOFF + s));\n");
            u32 uid2 = 0;
            emit_ops(w, jx, o.batch, "            ", uid2);
            w("#pragma unroll\n            for (simk_u32 s = 0; s < N; ++s) z[s] = (float)r[s];\n");
            w("            solve32(z);\n");
            w("#pragma unroll\n            for (simk_u32 s = 0; s < N; ++s) x.v[s] = x.v[s] + (double)z[s];\n");
            w("        }\n");
        }
        w("    }\n};\n\n");
    }

    return ws_slots;
}

// Emits the f32 delta-lane policy `sname` (simk_d32g_body): the arithmetic of
// simk_d32_fun / simk_d32_jac_factor / simk_d32_solve as straight-line f32
// code. Workspace (f32 slots per lane): [LU entries lds.. | K stages (all
// but the last) | k | src | pulse], the first lds LU entries in LDS. Returns
// the workspace size in f64 slots.
static u32 emit_d32_struct(Out &w, const Mech &m, const KinMethod &me, const GenOptions &o, const char *sname) {
    const u32 n = m.n;
    const u32 lds = lds_entries(m, o);
    const u32 stored_stages = me.stages > 0 ? me.stages - 1 : 0;
    // Paired f32 slots ([pair][64 lanes][2], two per 8-byte access): LU
    // entries in the solves' read order as in the FP64 layout, then the stage
    // vectors at an even stride np, the k, src and pulse copies, each starting
    // on a pair.
    const bool pairs = o.pair_slots != 0;
    std::vector<u32> lu_slot(m.nnz, 0);
    for (u32 idx = lds; idx < m.nnz; ++idx) lu_slot[idx] = idx - lds;
    u32 np = n, koff = m.nnz - lds;
    if (pairs) {
        u32 t = 0;
        for (u32 i = 0; i < n; ++i)
            for (u32 e = m.lower_begin[i]; e < m.lower_begin[i + 1]; ++e)
                if (m.lower[e].index >= lds) lu_slot[m.lower[e].index] = t++;
        t += t & 1u;
        for (u32 ii = n; ii-- > 0;) {
            for (u32 e = m.upper_begin[ii]; e < m.upper_begin[ii + 1]; ++e)
                if (m.upper[e].index >= lds) lu_slot[m.upper[e].index] = t++;
            if (m.diag[ii] >= lds) lu_slot[m.diag[ii]] = t++;
        }
        koff = t + (t & 1u);
        np = n + (n & 1u);
    }
    const u32 kcp = koff + stored_stages * np, scp = kcp + (pairs ? m.R + (m.R & 1u) : m.R), pcp = scp + np;
    const u32 slots32 = pcp + np;
    w("struct %s {\n", sname);
    w("    static constexpr simk_u32 N = %u, R = %u, STAGES = %u, KOFF = %u, NP = %u;\n", n, m.R, me.stages, koff, np);
    w(pairs ? "    SimkRows kr, sr, pr;\n    SimkRowPairs32 wp;\n" : "    SimkRows kr, sr, pr;\n    SimkRows32 ws;\n");
    if (lds) {
        w("    simk_u32 ln;\n");
        w("    SIMK_DEV float L(simk_u32 i) const { return simk_lds32[i * SIMK_WAVE + ln]; }\n");
        w("    SIMK_DEV void LS(simk_u32 i, float v) const { simk_lds32[i * SIMK_WAVE + ln] = v; }\n");
    }
    if (pairs) {
        w("    SIMK_DEV float W32(simk_u32 i) const { return wp.ld(i); }\n");
        w("    SIMK_DEV void W32S(simk_u32 i, float v) const { wp.st(i, v); }\n");
        w("    SIMK_DEV SimkF2 W2(simk_u32 pair) const { return wp.ld2(pair); }\n");
        w("    SIMK_DEV void WS2(simk_u32 pair, float a, float b) const { wp.st2(pair, a, b); }\n");
    } else {
        w("    SIMK_DEV float W32(simk_u32 i) const { return ws.ld(i); }\n");
        w("    SIMK_DEV void W32S(simk_u32 i, float v) const { ws.st(i, v); }\n");
    }
    w("    SIMK_DEV void init(const SimkKinArgs &A, simk_u32 wave, simk_u32 lane) {\n");
    w("        const simk_u32 stride = A.padded * 8u;\n");
    w("        kr.init(A.k, 0, stride, R * stride);\n");
    w("        sr.init(A.src, 0, stride, N * stride);\n");
    w("        pr.init(A.pulse, 0, stride, N * stride);\n");
    if (pairs)
        w("        wp.init(A.ws + (simk_u64)wave * A.ws_slots * SIMK_WAVE, lane, A.ws_slots * SIMK_WAVE * 8u);\n");
    else
        w("        ws.init(A.ws + (simk_u64)wave * A.ws_slots * SIMK_WAVE, lane * 4u, SIMK_WAVE * 4u, A.ws_slots * SIMK_WAVE * 8u);\n");
    if (lds) w("        ln = lane;\n");
    w("    }\n");
    // lane_load_f64 (base_lane.h:101): inputs rounded to f32 once, when the lane takes the cell
    w("    SIMK_DEV void set_cell(const SimkKinArgs &, simk_u32 cell) {\n");
    w("        kr.voff = sr.voff = pr.voff = cell * 8u;\n");
    if (pairs) {
        // copies in pairs (every copy starts on a pair)
        w("#pragma unroll 8\n        for (simk_u32 r = 0; r + 1 < R; r += 2) WS2((%u + r) / 2, (float)kr.ld(r), (float)kr.ld(r + 1));\n", kcp);
        w("        if (R & 1u) W32S(%u + R - 1, (float)kr.ld(R - 1));\n", kcp);
        w("#pragma unroll 8\n        for (simk_u32 s = 0; s + 1 < N; s += 2) WS2((%u + s) / 2, (float)sr.ld(s), (float)sr.ld(s + 1));\n", scp);
        w("        if (N & 1u) W32S(%u + N - 1, (float)sr.ld(N - 1));\n", scp);
        w("#pragma unroll 8\n        for (simk_u32 s = 0; s + 1 < N; s += 2) WS2((%u + s) / 2, (float)pr.ld(s), (float)pr.ld(s + 1));\n", pcp);
        w("        if (N & 1u) W32S(%u + N - 1, (float)pr.ld(N - 1));\n", pcp);
    } else {
        w("#pragma unroll 8\n        for (simk_u32 r = 0; r < R; ++r) W32S(%u + r, (float)kr.ld(r));\n", kcp);
        w("#pragma unroll 8\n        for (simk_u32 s = 0; s < N; ++s) W32S(%u + s, (float)sr.ld(s));\n", scp);
        w("#pragma unroll 8\n        for (simk_u32 s = 0; s < N; ++s) W32S(%u + s, (float)pr.ld(s));\n", pcp);
    }
    w("    }\n");
    w("    SIMK_DEV float K(simk_u32 r) const { return W32(%u + r); }\n", kcp);
    w("    SIMK_DEV float S(simk_u32 s) const { return W32(%u + s); }\n", scp);
    w("    SIMK_DEV float pulse(simk_u32 s) const { return W32(%u + s); }\n", pcp);
    w("    SIMK_DEV void kstore(simk_u32 st, const float (&x)[N]) const {\n");
    if (pairs) {
        w("#pragma unroll\n        for (simk_u32 s = 0; s + 1 < N; s += 2) WS2((KOFF + st * NP + s) / 2, x[s], x[s + 1]);\n");
        w("        if (N & 1u) W32S(KOFF + st * NP + N - 1, x[N - 1]);\n    }\n");
    } else {
        w("#pragma unroll\n        for (simk_u32 s = 0; s < N; ++s) W32S(KOFF + st * N + s, x[s]);\n    }\n");
    }
    {
        std::vector<Op> ops;
        char buf[96];
        for (u32 s2 = 0; s2 < n; ++s2) {
            Op op;
            if (pairs && s2 % 2 == 1) {
                op.from = (int)ops.size() - 1;   // the odd slot of the pair loaded for s2 - 1
                op.part = ".y";
            } else if (pairs && s2 + 1 < n) {
                snprintf(buf, sizeof(buf), "W2((KOFF + j * NP + %u) / 2)", s2);
                op.load = buf;
                op.type = "SimkF2";
                op.part = ".x";
            } else {
                snprintf(buf, sizeof(buf), pairs ? "W32(KOFF + j * NP + %u)" : "W32(KOFF + j * N + %u)", s2);
                op.load = buf;
            }
            snprintf(buf, sizeof(buf), "v[%u] = __builtin_fmaf(c, %%s, v[%u]);", s2, s2);
            op.stmt = buf;
            ops.push_back(op);
        }
        w("    SIMK_DEV void kaxpy(simk_u32 j, float c, float (&v)[N]) const {\n");
        u32 uid = 0;
        emit_ops(w, ops, o.batch, "        ", uid, 0, "float");
        w("    }\n");
    }
    // fun: f = src, then per reaction rate = k * pt_a [* pt_b], f_s = fma(coef, rate, f_s)
    {
        std::vector<Op> ops;
        char buf[96];
        for (u32 s2 = 0; s2 < n; ++s2) {
            Op op;
            snprintf(buf, sizeof(buf), "S(%u)", s2);
            op.load = buf;
            if (pairs) op.slot = (int)(scp + s2);
            snprintf(buf, sizeof(buf), "f[%u] = %%s;", s2);
            op.stmt = buf;
            ops.push_back(op);
        }
        for (u32 r = 0; r < m.R; ++r) {
            const Reaction &x = m.rx[r];
            Op op;
            snprintf(buf, sizeof(buf), "K(%u)", r);
            op.load = buf;
            op.immutable = true;
            if (pairs) op.slot = (int)(kcp + r);   // fun reads the rate constants in order: pairs
            std::string st = "{ const float rate = %s";
            for (u32 p = 0; p < x.reactant_count; ++p) {
                snprintf(buf, sizeof(buf), " * pt[%u]", x.reactant[p]);
                st += buf;
            }
            st += ";";
            for (u32 e = 0; e < x.net_count; ++e) {
                snprintf(buf, sizeof(buf), " f[%u] = __builtin_fmaf(%d.0f, rate, f[%u]);", x.net_species[e],
                         (int)x.net_coef[e], x.net_species[e]);
                st += buf;
            }
            st += " }";
            op.stmt = st;
            ops.push_back(op);
        }
        if (pairs) pair_loads(ops, 4, "SimkF2");
        w("    SIMK_DEV void fun(const float (&pt)[N], float (&f)[N]) const {\n");
        u32 uid = 0;
        emit_ops(w, ops, o.batch, "        ", uid, o.reuse, "float");
        w("    }\n");
    }
    // jac_factor: per row, -J accumulated in (reaction, reactant, entry) order
    // from +0, the eliminations against finished rows (registers within the
    // window, else reloads), the reciprocal pivot, then the row is stored.
    std::vector<u32> row_of(m.nnz, 0);
    for (u32 i = 0; i < n; ++i) {
        row_of[m.diag[i]] = i;
        for (u32 e = m.lower_begin[i]; e < m.lower_begin[i + 1]; ++e) row_of[m.lower[e].index] = i;
        for (u32 e = m.upper_begin[i]; e < m.upper_begin[i + 1]; ++e) row_of[m.upper[e].index] = i;
    }
    const u32 window = o.reg_window == GenOptions::kDefaultWindow ? 8u : o.reg_window;
    auto near = [&](u32 idx, u32 i) { return window == 0 || i - row_of[idx] <= window; };
    auto lu_ld = [&](u32 idx) {
        char buf[24];
        snprintf(buf, sizeof(buf), idx < lds ? "L(%u)" : "W32(%u)", idx < lds ? idx : lu_slot[idx]);
        return std::string(buf);
    };
    {
        std::vector<Op> ops;
        auto stmt = [&](const std::string &t) {
            Op op;
            op.stmt = t;
            ops.push_back(op);
        };
        auto barrier = [&]() {
            Op op;
            op.stmt = "SIMK_MEMORY_FENCE();";
            op.barrier = true;
            ops.push_back(op);
        };
        char buf[160];
        for (u32 i = 0; i < n; ++i) {
            std::vector<u32> row;
            for (u32 e = m.lower_begin[i]; e < m.lower_begin[i + 1]; ++e) row.push_back(m.lower[e].index);
            row.push_back(m.diag[i]);
            for (u32 e = m.upper_begin[i]; e < m.upper_begin[i + 1]; ++e) row.push_back(m.upper[e].index);
            bool far = false;
            for (u32 d = m.div_begin[i]; d < m.div_begin[i + 1]; ++d) {
                if (!near(m.lu_div[d].kk, i)) far = true;
                for (u32 s2 = m.lu_div[d].sub_begin; s2 < m.lu_div[d].sub_end; ++s2)
                    if (!near(m.lu_sub[s2].kj, i)) far = true;
            }
            if (far) barrier();
            snprintf(buf, sizeof(buf), "// row %u", i);
            stmt(buf);
            for (u32 idx : row) {
                snprintf(buf, sizeof(buf), "float a%u = 0.0f;", idx);
                stmt(buf);
            }
            for (u32 r = 0; r < m.R; ++r) {
                const Reaction &x = m.rx[r];
                bool any = false;
                for (u32 e = 0; e < x.net_count; ++e)
                    if (x.net_species[e] == i) any = true;
                if (!any || x.reactant_count == 0) continue;
                Op op;
                snprintf(buf, sizeof(buf), "K(%u)", r);
                op.load = buf;
                op.immutable = true;
                std::string st = "{ const float k = %s;";
                for (u32 p = 0; p < x.reactant_count; ++p) {
                    st += " { const float d = k";
                    for (u32 q = 0; q < x.reactant_count; ++q)
                        if (q != p) {
                            snprintf(buf, sizeof(buf), " * pt[%u]", x.reactant[q]);
                            st += buf;
                        }
                    st += ";";
                    for (u32 e = 0; e < x.net_count; ++e) {
                        if (x.net_species[e] != i) continue;
                        const u32 idx = x.jac_index[p][e];
                        snprintf(buf, sizeof(buf), " a%u = __builtin_fmaf(%d.0f, d, a%u);", idx, -(int)x.net_coef[e],
                                 idx);
                        st += buf;
                    }
                    st += " }";
                }
                st += " }";
                op.stmt = st;
                ops.push_back(op);
            }
            for (u32 d = m.div_begin[i]; d < m.div_begin[i + 1]; ++d) {
                const KinLuDiv &dv = m.lu_div[d];
                Op op;
                if (!near(dv.kk, i)) {
                    op.load = lu_ld(dv.kk);
                    if (pairs && dv.kk >= lds) op.slot = (int)lu_slot[dv.kk];
                    snprintf(buf, sizeof(buf), "const float l%u = a%u * %%s; a%u = l%u;", d, dv.ik, dv.ik, d);
                } else {
                    snprintf(buf, sizeof(buf), "const float l%u = a%u * a%u; a%u = l%u;", d, dv.ik, dv.kk, dv.ik, d);
                }
                op.stmt = buf;
                ops.push_back(op);
                for (u32 su = dv.sub_begin; su < dv.sub_end; ++su) {
                    const KinLuSub &sb = m.lu_sub[su];
                    Op o2;
                    if (near(sb.kj, i)) {
                        snprintf(buf, sizeof(buf), "a%u = __builtin_fmaf(-l%u, a%u, a%u);", sb.ij, d, sb.kj, sb.ij);
                    } else {
                        o2.load = lu_ld(sb.kj);
                        if (pairs && sb.kj >= lds) o2.slot = (int)lu_slot[sb.kj];
                        snprintf(buf, sizeof(buf), "a%u = __builtin_fmaf(-l%u, %%s, a%u);", sb.ij, d, sb.ij);
                    }
                    o2.stmt = buf;
                    ops.push_back(o2);
                }
            }
            snprintf(buf, sizeof(buf), "a%u = 1.0f / (a%u + ghinv);", m.diag[i], m.diag[i]);
            stmt(buf);
            std::vector<std::pair<u32, u32>> ws_row;   // (slot, compact index)
            for (u32 idx : row) {
                if (idx < lds) {
                    snprintf(buf, sizeof(buf), "LS(%u, a%u);", idx, idx);
                    stmt(buf);
                } else {
                    ws_row.push_back({lu_slot[idx], idx});
                }
            }
            if (pairs) std::sort(ws_row.begin(), ws_row.end());
            for (size_t q = 0; q < ws_row.size(); ++q) {
                const u32 sl = ws_row[q].first;
                if (pairs && sl % 2 == 0 && q + 1 < ws_row.size() && ws_row[q + 1].first == sl + 1) {
                    snprintf(buf, sizeof(buf), "WS2(%u, a%u, a%u);", sl / 2, ws_row[q].second, ws_row[q + 1].second);
                    ++q;
                } else {
                    snprintf(buf, sizeof(buf), "W32S(%u, a%u);", sl, ws_row[q].second);
                }
                stmt(buf);
            }
        }
        barrier();
        if (pairs) pair_loads(ops, 4, "SimkF2");
        w("    SIMK_DEV void jac_factor(const float (&pt)[N], const float ghinv) const {\n");
        u32 uid = 0;
        emit_ops(w, ops, o.batch, "        ", uid, o.reuse, "float");
        w("    }\n");
    }
    // solve: forward over lower, backward over upper and the reciprocal pivot
    {
        std::vector<Op> ops;
        char buf[160];
        for (u32 i = 0; i < n; ++i)
            for (u32 e = m.lower_begin[i]; e < m.lower_begin[i + 1]; ++e) {
                Op op;
                op.load = lu_ld(m.lower[e].index);
                if (pairs && m.lower[e].index >= lds) op.slot = (int)lu_slot[m.lower[e].index];
                snprintf(buf, sizeof(buf), "x[%u] = __builtin_fmaf(-%%s, x[%u], x[%u]);", i, m.lower[e].col, i);
                op.stmt = buf;
                ops.push_back(op);
            }
        for (u32 ii = n; ii-- > 0;) {
            for (u32 e = m.upper_begin[ii]; e < m.upper_begin[ii + 1]; ++e) {
                Op op;
                op.load = lu_ld(m.upper[e].index);
                if (pairs && m.upper[e].index >= lds) op.slot = (int)lu_slot[m.upper[e].index];
                snprintf(buf, sizeof(buf), "x[%u] = __builtin_fmaf(-%%s, x[%u], x[%u]);", ii, m.upper[e].col, ii);
                op.stmt = buf;
                ops.push_back(op);
            }
            Op op;
            op.load = lu_ld(m.diag[ii]);
            if (pairs && m.diag[ii] >= lds) op.slot = (int)lu_slot[m.diag[ii]];
            snprintf(buf, sizeof(buf), "x[%u] = x[%u] * %%s;", ii, ii);
            op.stmt = buf;
            ops.push_back(op);
        }
        if (pairs) pair_loads(ops, 4, "SimkF2");
        w("    SIMK_DEV void solve(float (&x)[N]) const {\n        SIMK_MEMORY_FENCE();\n");
        u32 uid = 0;
        emit_ops(w, ops, o.batch, "        ", uid, 0, "float");
        w("    }\n};\n\n");
    }
    return (slots32 + 1) / 2;
}

// Emits the method policy with compile-time coefficients.
static void emit_method(Out &w, const KinMethod &me, const GenOptions &o) {
    u32 eval[4], save_f[4];
    stage_plan(me, eval, save_f);
    w("struct GenMethod {\n");
    w("    static constexpr bool kUnrollStages = %s;\n", o.unroll_stages ? "true" : "false");
    w("    SIMK_DEV static constexpr simk_u32 stages() { return %u; }\n", me.stages);
    w("    SIMK_DEV static constexpr simk_u32 order() { return %u; }\n", me.order);
    w("    SIMK_DEV static constexpr double gamma0() { return %s; }\n", lit(me.gamma0).c_str());
    auto table = [&](const char *fn, const f64 *v, int count) {
        w("    SIMK_DEV static constexpr double %s(simk_u32 i) {\n", fn);
        for (int q = 0; q < count; ++q) w("        if (i == %d) return %s;\n", q, lit(v[q]).c_str());
        w("        return 0.0;\n    }\n");
    };
    table("a", me.a, 6);
    table("c", me.c, 6);
    table("m", me.m, 4);
    table("e", me.e, 4);
    w("    SIMK_DEV static constexpr bool eval(simk_u32 i) {\n");
    for (int q = 0; q < 4; ++q) w("        if (i == %d) return %s;\n", q, eval[q] ? "true" : "false");
    w("        return false;\n    }\n");
    w("    SIMK_DEV static constexpr bool save_f(simk_u32 i) {\n");
    for (int q = 0; q < 4; ++q) w("        if (i == %d) return %s;\n", q, save_f[q] ? "true" : "false");
    w("        return false;\n    }\n};\n\n");

}

GenResult generate(const Mech &m, const KinMethod &me, const GenOptions &o, const char *device_header) {
    GenResult g;
    u64 sig = signature(m);
    char name[64];
    const bool mixed = o.precision == 2;
    if (mixed) snprintf(name, sizeof(name), "simk_mix%u_%016llx", o.sweeps, (unsigned long long)sig);
    else if (o.precision == 3) snprintf(name, sizeof(name), "simk_f32s_%016llx", (unsigned long long)sig);
    else if (o.precision == 1) snprintf(name, sizeof(name), "simk_d32g_%016llx", (unsigned long long)sig);
    else snprintf(name, sizeof(name), "simk_gen_%016llx", (unsigned long long)sig);
    g.kernel = name;
    Out w;
    w("// generated by simstack_amd codegen: %u species, %u reactions, %u non-zeros\n", m.n, m.R, m.nnz);
    w.s += device_header;
    w.s += "\n";
    if (o.precision == 3) {
        w("__shared__ double simk_lds[%u * SIMK_WAVE];\n", m.n);
        if (const u32 lds = lds_entries(m, o)) w("__shared__ float simk_lds32[%u * SIMK_WAVE];\n", lds);
        w("\n");
    } else if (const u32 lds = lds_entries(m, o))
        w("__shared__ %s[%u * SIMK_WAVE];\n\n", o.precision != 0 ? "float simk_lds32" : "double simk_lds", lds);
    if (o.precision == 1) {
        // f32 delta lanes: the method and control come in the argument block
        g.ws_slots = emit_d32_struct(w, m, me, o, "GenOps32");
        w("extern \"C\" __global__ void __launch_bounds__(SIMK_WAVE) %s(SimkD32Args DA) {\n", name);
        w("    simk_d32g_body<GenOps32>(DA.A, DA.C);\n}\n");
        g.source = std::move(w.s);
        return g;
    }
    g.ws_slots = emit_ops_struct(w, m, me, o, "GenOps");
    emit_method(w, me, o);
    if (o.waves_per_eu)
        w("extern \"C\" __global__ void __launch_bounds__(SIMK_WAVE) __attribute__((amdgpu_waves_per_eu(%u))) %s(SimkKinArgs A) {\n",
          o.waves_per_eu, name);
    else
        w("extern \"C\" __global__ void __launch_bounds__(SIMK_WAVE) %s(SimkKinArgs A) {\n", name);
    w("    GenMethod me;\n    simk_persistent_body<GenOps, GenMethod>(A, me);\n}\n");
    g.source = std::move(w.s);
    return g;
}

GenResult generate_multi(const std::vector<const Mech *> &ms, const KinMethod &me, const GenOptions &o,
                         const char *device_header) {
    GenResult g;
    u64 sig = 14695981039346656037ull;
    for (const Mech *m : ms) {
        u64 s1 = signature(*m);
        sig = fnv1a64(&s1, 8, sig);
    }
    char name[64];
    snprintf(name, sizeof(name), o.precision == 3 ? "simk_mf32s%u_%016llx" : "simk_multi%u_%016llx", (u32)ms.size(),
             (unsigned long long)sig);
    g.kernel = name;
    Out w;
    w("// generated by simstack_amd codegen: %u mechanisms in one persistent launch\n", (u32)ms.size());
    w.s += device_header;
    w.s += "\n";
    // One LDS array (FP32 factor storage: one for the pivots, one for the f32
    // entries) serves every mechanism: a wave runs one mechanism at a time.
    u32 piv = 0;
    for (const Mech *m : ms)
        if (m->n > piv) piv = m->n;
    // FP32 factor storage: every mechanism's f32 share is the LDS budget less
    // the shared pivot array of the largest mechanism
    auto mech_options = [&](const Mech &m) {
        GenOptions om = o;
        if (o.precision == 3) {
            const u32 extra = 2u * (piv - m.n);
            om.lds_slots = o.lds_slots > extra ? o.lds_slots - extra : 0u;
        }
        return om;
    };
    u32 lds = 0;
    for (const Mech *m : ms)
        if (lds_entries(*m, mech_options(*m)) > lds) lds = lds_entries(*m, mech_options(*m));
    if (o.precision == 3) {
        w("__shared__ double simk_lds[%u * SIMK_WAVE];\n", piv);
        if (lds) w("__shared__ float simk_lds32[%u * SIMK_WAVE];\n", lds);
        w("\n");
    } else if (lds) {
        w("__shared__ double simk_lds[%u * SIMK_WAVE];\n\n", lds);
    }
    for (size_t i = 0; i < ms.size(); ++i) {
        char sname[32];
        snprintf(sname, sizeof(sname), "GenOps%u", (u32)i);
        g.mech_slots.push_back(emit_ops_struct(w, *ms[i], me, mech_options(*ms[i]), sname));
        if (g.mech_slots.back() > g.ws_slots) g.ws_slots = g.mech_slots.back();
    }
    emit_method(w, me, o);
    // A wave starts at the position of MA.seq whose wave range holds its
    // index and visits the mechanisms' queues in seq order from there
    // (wrapping), moving on once a queue is empty, so no SIMD idles while any
    // mechanism still has cells.
    w("extern \"C\" __global__ void __launch_bounds__(SIMK_WAVE) %s(SimkMultiArgs MA) {\n", name);
    w("    GenMethod me;\n    simk_u32 first = 0;\n");
    w("    for (simk_u32 k = 1; k < MA.count; ++k)\n        if (blockIdx.x >= MA.start[k]) first = k;\n");
    w("    for (simk_u32 pass = 0; pass < MA.count; ++pass) {\n");
    w("        simk_u32 k = first + pass;\n        if (k >= MA.count) k -= MA.count;\n");
    w("        const simk_u32 m = MA.seq[k];\n");
    w("        switch (m) {\n");
    for (size_t i = 0; i < ms.size(); ++i)
        w("            case %u: simk_persistent_body<GenOps%u, GenMethod, true>(MA.a[%u], me); break;\n", (u32)i, (u32)i,
          (u32)i);
    w("            default: break;\n        }\n    }\n}\n");
    g.source = std::move(w.s);
    return g;
}

}  // namespace simk
