// Host-side check of the kernel generator under AddressSanitizer and
// UndefinedBehaviorSanitizer: decodes mechanism tables (KinMech bytes at the
// given capacities) and generates every kernel form for each of them: the
// four precisions, LDS shares from none to all, single and paired workspace
// slots, with and without the paired rate-constant copy, refinement sweeps,
// the stage carry on and off (alternating over the other options),
// and the multi-mechanism kernels in FP64 and FP32 factor storage. Built and
// run by tests/test_generator_sanitized.py (g++, host code only).
//
//   gen_harness method.bin mech0.bin caps0.bin [mech1.bin caps1.bin ...]
#include "codegen.hpp"
#include "mech.hpp"

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using namespace simk;

static std::vector<char> read_file(const char *path) {
    std::vector<char> v;
    FILE *f = fopen(path, "rb");
    if (!f) return v;
    fseek(f, 0, SEEK_END);
    const long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    v.resize((size_t)n);
    if (fread(v.data(), 1, v.size(), f) != v.size()) v.clear();
    fclose(f);
    return v;
}

int main(int argc, char **argv) {
    if (argc < 4) return 2;
    const std::vector<char> mb = read_file(argv[1]);
    if (mb.size() < sizeof(KinMethod)) return 2;
    KinMethod me;
    memcpy(&me, mb.data(), sizeof(me));
    me.name = nullptr;
    std::vector<Mech> ms;
    for (int i = 2; i + 1 < argc; i += 2) {
        const std::vector<char> b = read_file(argv[i]), c = read_file(argv[i + 1]);
        if (c.size() < sizeof(SimkCaps)) return 2;
        SimkCaps caps;
        memcpy(&caps, c.data(), sizeof(caps));
        Mech m;
        const std::string err = from_bytes(b.data(), caps, m);
        if (!err.empty()) {
            printf("decode: %s\n", err.c_str());
            return 1;
        }
        ms.push_back(m);
    }
    size_t bytes = 0, kernels = 0;
    for (const Mech &m : ms)
        for (u32 prec : {0u, 1u, 2u, 3u})
            for (u32 lds : {0u, 64u, 160u, 201u, 304u, 640u})
                for (u32 pair : {0u, 1u})
                    for (u32 kp : {0u, 1u})
                        for (u32 sweeps : {0u, 1u, 2u}) {
                            if (prec != 2 && sweeps) continue;
                            const u32 carry = (lds + kp + sweeps) & 1u;
                            GenOptions o;
                            o.precision = prec;
                            o.sweeps = sweeps;
                            o.lds_slots = lds;
                            o.pair_slots = pair;
                            o.k_pairs = kp;
                            o.stage_carry = carry;
                            o.batch = 24;
                            o.reuse = 384;
                            o.fun_in_jac = 1;
                            o.reg_window = 16;
                            bytes += generate(m, me, o, "").source.size();
                            ++kernels;
                        }
    std::vector<const Mech *> ptrs;
    for (const Mech &m : ms)
        if (ptrs.size() < 4) ptrs.push_back(&m);
    for (u32 prec : {0u, 3u})
        for (u32 kp : {0u, 1u}) {
            GenOptions o;
            o.precision = prec;
            o.lds_slots = prec == 3 ? 304u : 160u;
            o.pair_slots = 1;
            o.k_pairs = kp;
            o.stage_carry = kp;
            o.batch = 24;
            o.reuse = 96;
            o.fun_in_jac = 1;
            o.reg_window = 8;
            bytes += generate_multi(ptrs, me, o, "").source.size();
            ++kernels;
        }
    printf("generated %zu kernels, %zu bytes of source\n", kernels, bytes);
    return 0;
}
