// Runtime-sized mechanism tables: the same content as the reference's KinMech
// (src/sim/kinetics/kin_mech.h:72-87) without compile-time capacities, plus
// conversion to and from the reference byte layout at any capacities.
//
// Shared by the host library (builder, generator) and the GPU module (which
// receives KinMech bytes at its C ABI and turns them into kernels).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "simk_types.h"

namespace simk {

struct Reaction {
    u8 reactant[SIMK_MAX_REACTANTS] = {SIMK_NONE, SIMK_NONE, SIMK_NONE};
    u8 reactant_count = 0;
    u8 net_count = 0;
    u8 net_species[SIMK_MAX_NET] = {};
    s8 net_coef[SIMK_MAX_NET] = {};
    u16 jac_index[SIMK_MAX_REACTANTS][SIMK_MAX_NET] = {};
};

struct Mech {
    u32 n = 0;     // species
    u32 R = 0;     // reactions
    u32 nnz = 0;   // J + diagonal + fill-in
    std::vector<Reaction> rx;
    std::vector<u16> diag;          // [n]
    std::vector<u16> div_begin;     // [n + 1]
    std::vector<KinLuDiv> lu_div;
    std::vector<KinLuSub> lu_sub;
    std::vector<u16> lower_begin;   // [n + 1], lower[] ascending column
    std::vector<u16> upper_begin;   // [n + 1], upper[] DESCENDING column
    std::vector<KinRowEntry> lower;
    std::vector<KinRowEntry> upper;
};

// Symbolic phase (behaviour of kin_build.cpp:150-223): pattern = J u diag,
// no-pivot IKJ fill closure in natural order, row-major compact index, the
// division/update op lists, lower (ascending) / upper (descending) row lists
// and each reaction's Jacobian scatter indices. Reactions must already hold
// reactants and net stoichiometry. Returns "" or an error.
std::string finalize(Mech &m);

// Checks every index a kernel would follow (so malformed tables can never
// send a device access out of bounds). Returns "" or the first problem.
std::string validate(const Mech &m);

// Reference byte layout at capacities `caps` (simk_mech_layout).
std::string from_bytes(const void *bytes, const SimkCaps &caps, Mech &out);
std::string to_bytes(const Mech &m, const SimkCaps &caps, void *bytes);   // bytes: layout.size, zero-filled first

// FNV-1a 64 (the reference's mechanism signature hash, base_types.h:120-128).
u64 fnv1a64(const void *data, u64 size, u64 seed = 14695981039346656037ull);
// Signature of the table content (independent of capacities).
u64 signature(const Mech &m);

// Rosenbrock coefficients (Sandu et al. 1997, KPP form; the reference's table is
// kin_solver.h:352-395). id: 0 Ros2, 1 Ros3, 2 Rodas3.
bool method(u32 id, KinMethod &out);

}  // namespace simk
