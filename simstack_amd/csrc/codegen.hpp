// Kernel generator: emits HIP source for one mechanism with the four Ops
// kernels unrolled into straight-line code from the symbolic pattern — the
// role of the reference's build-time kin_codegen (src/tools/kin_codegen.cpp:
// 32-138), here run when a mechanism is set and compiled for gfx950 by hipRTC.
#pragma once

#include <string>
#include <vector>

#include "mech.hpp"

namespace simk {

// Knobs the autotuner sweeps. Each changes speed only, never results.
struct GenOptions {
    u32 fence_rows = 0;      // 1: re-read finished rows from the workspace (fence per row); 0: keep them in registers
    u32 waves_per_eu = 0;    // amdgpu_waves_per_eu lower bound (0: compiler default)
    u32 unroll_stages = 0;   // unroll the stage loop (bigger code, constant stage offsets)
    u32 batch = kAutoBatch;  // loads issued this many statements ahead (software pipelining;
                             // kAutoBatch: resolved by auto_batch before generation)
    u32 precision = 0;       // 0: FP64 throughout; 1: f32 delta lanes; 2: FP32 Jacobian/LU + FP64 residual refinement
    u32 sweeps = 1;          // precision 2: refinement sweeps per stage solve
    u32 reg_window = kDefaultWindow;   // register mode: finished rows read from registers up to this many
                                       // rows back (0: all; kDefaultWindow: 16 in FP64, 8 in mixed precision)
    u32 reuse = kAutoReuse;  // a rate constant loaded within this many statements is reused, not reloaded (0: never;
                             // kAutoReuse: resolved by auto_reuse before generation)
    u32 lds_slots = kAutoLds;   // the first this many LU entries of each lane live in LDS, not the workspace
                                // (kAutoLds: resolved by auto_lds before generation)
    u32 fun_in_jac = kAutoFunInJac;   // 1: the -J pass also evaluates f(y) (one rate-constant pass fewer);
                                      // 0: stage 0 evaluates it (kAutoFunInJac: auto_fun_in_jac)
    u32 pair_slots = kAutoPair;   // FP64 kernels: 1: workspace slots in adjacent pairs per lane, LU entries
                                  // placed in the solves' read order, two slots per 16-byte access; 0: one
                                  // 8-byte slot per access (kAutoPair: auto_pair)
    static constexpr u32 kDefaultWindow = 0xFFFF;
    static constexpr u32 kAutoLds = 0xFFFF;
    static constexpr u32 kAutoBatch = 0xFFFF;
    static constexpr u32 kAutoReuse = 0xFFFF;
    static constexpr u32 kAutoFunInJac = 0xFFFF;
    u32 k_pairs = kAutoKPairs;   // paired layouts: 1: a second, paired copy of the rate constants that the f
                                 // passes read two per 16-byte access; 0: none (kAutoKPairs: auto_k_pairs)
    static constexpr u32 kAutoPair = 0xFFFF;
    static constexpr u32 kAutoKPairs = 0xFFFF;
    u32 stage_carry = kAutoCarry;   // 1: a stage whose combination ends with the previous stage's K takes
                                    // that term from the registers still holding it (one K load fewer per
                                    // such stage, same bits); 0: every K term is loaded (kAutoCarry:
                                    // auto_stage_carry)
    static constexpr u32 kAutoCarry = 0xFFFF;
};

struct GenResult {
    std::string source;   // complete translation unit (device header included)
    std::string kernel;   // extern "C" kernel name
    u32 ws_slots = 0;     // workspace slots per lane the kernel uses (the maximum over mechanisms)
    std::vector<u32> mech_slots;   // multi-mechanism kernels: slots of each mechanism
};

// LU entries per lane the kernel keeps in LDS (o.lds_slots clamped to the LU
// and to 160 KiB per wave).
u32 lds_entries(const Mech &m, const GenOptions &o);

// The default LU entries in LDS: 80 KiB per wave (160 f64), so two waves
// share a CU's 160 KiB; 40 KiB (160 f32) for f32 delta lanes (four waves per
// CU); 50 KiB (200 f32) for unrefined mixed-precision kernels (three waves
// per CU); none for mixed-precision kernels with refinement sweeps.
u32 auto_lds(const GenOptions &o);

// The default load-pipeline depth: 24 statements; 16 for f32 delta lanes and
// refined mixed-precision kernels.
u32 auto_batch(const GenOptions &o);

// The default rate-constant reuse window: 384 statements for one-mechanism
// FP64 kernels, 1000 (every earlier load of the pass) for unrefined
// mixed-precision kernels, 96 otherwise (multi: the multi-mechanism kernel).
u32 auto_reuse(const GenOptions &o, bool multi = false);

// The default for f(y) in the -J pass: on.
u32 auto_fun_in_jac(const GenOptions &o);
u32 auto_pair(const GenOptions &o);
u32 auto_k_pairs(const GenOptions &o, bool multi = false);
// The default stage carry: off (an autotuner knob).
u32 auto_stage_carry(const GenOptions &o, bool multi = false);

// Stage plan shared with the host launch code: which stages evaluate f.
void stage_plan(const KinMethod &me, u32 eval[4], u32 save_f[4]);

GenResult generate(const Mech &m, const KinMethod &me, const GenOptions &o, const char *device_header);

// One persistent kernel for up to four mechanisms (ragged batches).
GenResult generate_multi(const std::vector<const Mech *> &ms, const KinMethod &me, const GenOptions &o,
                         const char *device_header);

}  // namespace simk
