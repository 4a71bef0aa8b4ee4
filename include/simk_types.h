/* simk_types.h — plain-data types shared by the host library, the GPU module
 * and their callers. Every struct here is byte-compatible with the reference's
 * type of the same name so a caller built against the reference headers can
 * pass its own objects unchanged:
 *
 *   KinReaction, KinLuDiv, KinLuSub, KinRowEntry, KinMech  src/sim/kinetics/kin_mech.h:42-87
 *   KinMethod                                              src/sim/kinetics/kin_mech.h:94-102
 *   KinControl                                             src/sim/kinetics/kin_mech.h:112-119
 *   GpuKinLaunch, GpuInfo                                  src/gpu/gpu_api.h:30-44
 *
 * KinMech's size depends on the five capacity macros the caller compiled with.
 * The struct below is the layout at the reference's default capacities
 * (32 species, 96 reactions, 512 nnz, 512 LU divisions, 4096 LU updates:
 * 30,044 bytes). Callers that raise the capacities (as the reference's offline
 * tools do) describe them with SimkCaps and use the *_caps entry points; the
 * byte offsets for any capacities come from simk_mech_layout().
 */
#ifndef SIMK_TYPES_H
#define SIMK_TYPES_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int8_t s8;
typedef int32_t s32;
typedef uint8_t u8;
typedef uint16_t u16;
typedef uint32_t u32;
typedef uint64_t u64;
typedef double f64;
typedef float f32;
typedef int32_t b32;

#define SIMK_MAX_REACTANTS 3 /* kin_mech.h:28 */
#define SIMK_MAX_NET 6       /* kin_mech.h:29 */
#define SIMK_MAX_STAGES 4    /* kin_mech.h:93 */
#define SIMK_NONE 0xFF       /* kin_mech.h:40 */

/* Default capacities (kin_mech.h:22-38). */
#define SIMK_DEFAULT_SPECIES 32
#define SIMK_DEFAULT_REACTIONS 96
#define SIMK_DEFAULT_NNZ 512
#define SIMK_DEFAULT_LU_DIV 512
#define SIMK_DEFAULT_LU_SUB 4096

/* Capacities a KinMech was compiled with. */
typedef struct SimkCaps {
    u32 species, reactions, nnz, lu_div, lu_sub;
} SimkCaps;

/* One mass-action reaction (kin_mech.h:42-51): 54 bytes, 2-byte aligned. */
typedef struct KinReaction {
    u8 reactant[SIMK_MAX_REACTANTS];
    u8 reactant_count;
    u8 net_count;
    u8 net_species[SIMK_MAX_NET];
    s8 net_coef[SIMK_MAX_NET];
    u16 jac_index[SIMK_MAX_REACTANTS][SIMK_MAX_NET];
} KinReaction;

typedef struct KinLuDiv {
    u16 ik, kk;
    u16 sub_begin, sub_end;
} KinLuDiv;

typedef struct KinLuSub {
    u16 ij, kj;
} KinLuSub;

typedef struct KinRowEntry {
    u16 index;
    u8 col;
    u8 pad;
} KinRowEntry;

/* KinMech at the default capacities (kin_mech.h:72-87). */
typedef struct KinMech {
    u32 species_count;
    u32 reaction_count;
    KinReaction reactions[SIMK_DEFAULT_REACTIONS];
    u32 nnz;
    u16 diag[SIMK_DEFAULT_SPECIES];
    u32 lu_div_count, lu_sub_count;
    u16 div_begin[SIMK_DEFAULT_SPECIES + 1];
    KinLuDiv lu_div[SIMK_DEFAULT_LU_DIV];
    KinLuSub lu_sub[SIMK_DEFAULT_LU_SUB];
    u16 lower_begin[SIMK_DEFAULT_SPECIES + 1];
    u16 upper_begin[SIMK_DEFAULT_SPECIES + 1];
    KinRowEntry lower[SIMK_DEFAULT_NNZ];
    KinRowEntry upper[SIMK_DEFAULT_NNZ];
} KinMech;

/* Byte offsets of KinMech's fields for arbitrary capacities. */
typedef struct SimkMechLayout {
    u64 species_count, reaction_count, reactions, nnz, diag, lu_div_count, lu_sub_count, div_begin, lu_div, lu_sub,
        lower_begin, upper_begin, lower, upper, size;
} SimkMechLayout;

static inline u64 simk_align_up(u64 x, u64 a) { return (x + a - 1) & ~(a - 1); }

static inline SimkMechLayout simk_mech_layout(const SimkCaps *c) {
    SimkMechLayout L;
    u64 o = 0;
    L.species_count = o; o += 4;
    L.reaction_count = o; o += 4;
    L.reactions = o; o += (u64)sizeof(KinReaction) * c->reactions;
    o = simk_align_up(o, 4); L.nnz = o; o += 4;
    L.diag = o; o += 2ull * c->species;
    o = simk_align_up(o, 4); L.lu_div_count = o; o += 4;
    L.lu_sub_count = o; o += 4;
    L.div_begin = o; o += 2ull * (c->species + 1);
    L.lu_div = o; o += (u64)sizeof(KinLuDiv) * c->lu_div;
    L.lu_sub = o; o += (u64)sizeof(KinLuSub) * c->lu_sub;
    L.lower_begin = o; o += 2ull * (c->species + 1);
    L.upper_begin = o; o += 2ull * (c->species + 1);
    L.lower = o; o += (u64)sizeof(KinRowEntry) * c->nnz;
    L.upper = o; o += (u64)sizeof(KinRowEntry) * c->nnz;
    L.size = simk_align_up(o, 4);
    return L;
}

/* Rosenbrock coefficients in KPP form (kin_mech.h:89-102): 200 bytes. */
typedef struct KinMethod {
    u32 stages;
    u32 order;
    f64 a[6], c[6];
    f64 m[SIMK_MAX_STAGES], e[SIMK_MAX_STAGES];
    f64 gamma0;
    b32 new_f[SIMK_MAX_STAGES];
    const char *name;
} KinMethod;

enum KinMethodId { KinMethod_Ros2 = 0, KinMethod_Ros3 = 1, KinMethod_Rodas3 = 2, KinMethod_Count = 3 };

/* Tolerances and limits of one integration call (kin_mech.h:112-119): 56 bytes. */
typedef struct KinControl {
    f64 rtol, atol;
    f64 h_min, h_max;
    f64 h_start;
    u32 max_steps;
    b32 fixed_step;
    b32 clamp_negative;
} KinControl;

/* Device pointers of one batched advance (gpu_api.h:30-37): the [row][padded]
 * f64 SoA block, one lane per cell. 120 bytes. */
typedef struct GpuKinLaunch {
    f64 *y, *h, *steps, *rejects, *pulses, *failed;
    f64 *k, *src, *pulse, *period, *pulse_first, *pulse_count;
    u32 padded;
    u32 lane_count;
    u32 precision;   /* 0: FP64 (bit-identical to the CPU path); 1: the reference's f32 delta lanes;
                        2: mixed precision (FP32 Jacobian/LU, FP64 residual refinement and control);
                        3: FP32 factor storage (off-diagonal LU entries rounded to and stored as f32, FP64
                        reciprocal pivots, FP64 arithmetic and control) */
    f64 duration;
} GpuKinLaunch;

typedef struct GpuInfo {
    char name[128];
    u32 sm_major, sm_minor, multiprocessors;
    u64 total_bytes, free_bytes;
    u64 arena_bytes;
} GpuInfo;

#ifdef __cplusplus
}
#endif

#endif /* SIMK_TYPES_H */
