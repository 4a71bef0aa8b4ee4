// libsimk_gpu.so — the GPU module: C ABI (simk_gpu_api.h), device arena,
// stream, the table-walk kernel, and the JIT pipeline for generated kernels.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <dlfcn.h>
#include <sys/stat.h>
#include <unistd.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "codegen.hpp"
#include "device_kin.h"
#include "mech.hpp"
#include "simk_gpu_api.h"
#include "module_internal.h"

#include "device_delta32.h"
#include "device_wave.h"
#include "device_kin_src.inc"   // kDeviceKinSource: device_kin.h + device_delta32.h as text, for hipRTC

using namespace simk;

// ---------------------------------------------------------------------------
// Table-walk policy: any mechanism, tables read uniformly from device memory
// (scalar loads), every per-lane vector in workspace slots. The behaviour of
// the reference's KinTableOps (src/sim/kinetics/kin_solver.h:152-161).

struct TableOps {
    static constexpr bool kFunInJac = false;
    static constexpr bool kStageCarry = false;
    using Vec = SimkWsVec;
    const SimkDevTables *T;
    simk_u32 n_, R_, nnz_;
    SimkRows kr, sr, pr, ws;

    SIMK_DEV void init(const SimkKinArgs &A, simk_u32 wave, simk_u32 lane) {
        T = A.tables;
        n_ = T->n;
        R_ = T->R;
        nnz_ = T->nnz;
        const simk_u32 stride = A.padded * 8u;
        kr.init(A.k, 0, stride, R_ * stride);
        sr.init(A.src, 0, stride, n_ * stride);
        pr.init(A.pulse, 0, stride, n_ * stride);
        ws.init(A.ws + (simk_u64)wave * A.ws_slots * SIMK_WAVE, lane * 8u, SIMK_WAVE * 8u,
                A.ws_slots * SIMK_WAVE * 8u);
    }
    SIMK_DEV void set_cell(const SimkKinArgs &, simk_u32 cell) { kr.voff = sr.voff = pr.voff = cell * 8u; }
    // Workspace slots: [0, nnz) LU; vectors i = 0..4 at nnz + i n; K_st at
    // nnz + (5 + st) n; the reused f at nnz + 9 n.
    SIMK_DEV simk_u32 n() const { return n_; }
    SIMK_DEV Vec vec(int i) const { return Vec{ws, nnz_ + (simk_u32)i * n_}; }
    SIMK_DEV double K(simk_u32 r) const { return kr.ld(r); }
    SIMK_DEV double S(simk_u32 s) const { return sr.ld(s); }
    SIMK_DEV double pulse(simk_u32 s) const { return pr.ld(s); }
    SIMK_DEV double W(simk_u32 i) const { return ws.ld(i); }
    SIMK_DEV void WS(simk_u32 i, double v) const { ws.st(i, v); }
    SIMK_DEV simk_u32 koff(simk_u32 st) const { return nnz_ + (5 + st) * n_; }
    SIMK_DEV double kget(simk_u32 st, simk_u32 s) const { return W(koff(st) + s); }
    SIMK_DEV void kstore(simk_u32 st, const Vec &x) const {
        for (simk_u32 s = 0; s < n_; ++s) WS(koff(st) + s, x.get(s));
    }
    SIMK_DEV void kaxpy(simk_u32 j, double c, const Vec &v) const {
        for (simk_u32 s = 0; s < n_; ++s) v.set(s, simk_fma(c, kget(j, s), v.get(s)));
    }
    SIMK_DEV void fsave(const Vec &x) const {
        for (simk_u32 s = 0; s < n_; ++s) WS(nnz_ + 9 * n_ + s, x.get(s));
    }
    SIMK_DEV void fload(const Vec &x) const {
        for (simk_u32 s = 0; s < n_; ++s) x.set(s, W(nnz_ + 9 * n_ + s));
    }

    // kin_fun (kin_solver.h:39-53)
    SIMK_DEV void fun(const Vec &y, const Vec &f) const {
        for (simk_u32 s = 0; s < n_; ++s) f.set(s, S(s));
        for (simk_u32 r = 0; r < R_; ++r) {
            double rate = K(r);
            const simk_u32 rc = T->reactant_count[r];
            for (simk_u32 p = 0; p < rc; ++p) rate = rate * y.get(T->reactant[r * 3 + p]);
            const simk_u32 nc = T->net_count[r];
            for (simk_u32 e = 0; e < nc; ++e) {
                const simk_u32 s = T->net_species[r * 6 + e];
                f.set(s, simk_fma(T->net_coef[r * 6 + e], rate, f.get(s)));
            }
        }
    }
    // kin_neg_jac then kin_factor (kin_solver.h:75-113), in place in the workspace.
    SIMK_DEV void jac_factor(const Vec &y, const double g) const {
        for (simk_u32 i = 0; i < nnz_; ++i) WS(i, 0.0);
        for (simk_u32 r = 0; r < R_; ++r) {
            const simk_u32 rc = T->reactant_count[r], nc = T->net_count[r];
            for (simk_u32 p = 0; p < rc; ++p) {
                double d = K(r);
                for (simk_u32 q = 0; q < rc; ++q)
                    if (q != p) d = d * y.get(T->reactant[r * 3 + q]);
                for (simk_u32 e = 0; e < nc; ++e) {
                    const simk_u32 idx = T->jac_index[(r * 3 + p) * 6 + e];
                    WS(idx, simk_fma(-T->net_coef[r * 6 + e], d, W(idx)));
                }
            }
        }
        for (simk_u32 i = 0; i < n_; ++i) {
            const simk_u32 d0 = T->div_begin[i], d1 = T->div_begin[i + 1];
            for (simk_u32 d = d0; d < d1; ++d) {
                const unsigned short *dv = T->lu_div + 4 * d;
                const double l = W(dv[0]) * W(dv[1]);
                WS(dv[0], l);
                for (simk_u32 s = dv[2]; s < dv[3]; ++s) {
                    const simk_u32 ij = T->lu_sub[2 * s], kj = T->lu_sub[2 * s + 1];
                    WS(ij, simk_fnma(l, W(kj), W(ij)));
                }
            }
            const simk_u32 ii = T->diag[i];
            WS(ii, 1.0 / (W(ii) + g));
        }
    }
    // kin_solve (kin_solver.h:122-139)
    SIMK_DEV void solve(const Vec &x, const Vec &, double) const {
        for (simk_u32 i = 0; i < n_; ++i)
            for (simk_u32 e = T->lower_begin[i]; e < T->lower_begin[i + 1]; ++e)
                x.set(i, simk_fnma(W(T->lower[2 * e]), x.get(T->lower[2 * e + 1]), x.get(i)));
        for (simk_u32 ii = n_; ii-- > 0;) {
            for (simk_u32 e = T->upper_begin[ii]; e < T->upper_begin[ii + 1]; ++e)
                x.set(ii, simk_fnma(W(T->upper[2 * e]), x.get(T->upper[2 * e + 1]), x.get(ii)));
            x.set(ii, x.get(ii) * W(T->diag[ii]));
        }
    }
};

struct ValueMethod {   // method copied by value from the kernel argument
    static constexpr bool kUnrollStages = false;
    SimkDevMethod mm;
    SIMK_DEV simk_u32 stages() const { return mm.stages; }
    SIMK_DEV simk_u32 order() const { return mm.order; }
    SIMK_DEV double gamma0() const { return mm.gamma0; }
    SIMK_DEV double a(simk_u32 i) const { return mm.a[i]; }
    SIMK_DEV double c(simk_u32 i) const { return mm.c[i]; }
    SIMK_DEV double m(simk_u32 i) const { return mm.m[i]; }
    SIMK_DEV double e(simk_u32 i) const { return mm.e[i]; }
    SIMK_DEV bool eval(simk_u32 i) const { return mm.eval[i] != 0; }
    SIMK_DEV bool save_f(simk_u32 i) const { return mm.save_f[i] != 0; }
};

extern "C" __global__ void __launch_bounds__(SIMK_WAVE) simk_table_kernel(SimkKinArgs A) {
    ValueMethod me{A.method};
    simk_persistent_body<TableOps, ValueMethod>(A, me);
}

// f32 delta lanes (GpuKinLaunch precision 1), table walk, any mechanism.
extern "C" __global__ void __launch_bounds__(SIMK_WAVE) simk_delta32_kernel(SimkKinArgs A, SimkD32Consts C) {
    simk_d32_body(A, C);
}

// Cooperative mode: one wave per cell, cell state in LDS.
template <class T, bool DELTA, simk_u32 L>
__global__ void __launch_bounds__(SIMK_WAVE) simk_wave_kernel(SimkKinArgs A, SimkTConsts<T> C,
                                                             const SimkWaveTables *W) {
    extern __shared__ __align__(16) unsigned char simk_wave_lds[];
    simk_wave_body<T, DELTA, L>(A, C, W, simk_wave_lds);
}

// The cooperative form of a ragged batch (FP64): teams of L lanes per cell
// over several mechanisms' queues.
template <simk_u32 L>
__global__ void __launch_bounds__(SIMK_WAVE) simk_wave_multi_kernel(SimkWaveMultiArgs M) {
    extern __shared__ __align__(16) unsigned char simk_wave_lds[];
    simk_wave_multi_body<L>(M, simk_wave_lds);
}

// The reference's M1 point demo (a spring/swirl particle step), kept so every
// reference export exists.
__global__ void simk_step_points_kernel(f32 *x, f32 *y, f32 *z, f32 *vx, f32 *vy, f32 *vz, u32 count, f32 dt) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const f32 spring = 4.0f, swirl = 0.6f, damp = 1.0f - 0.02f * dt;
    const f32 px = x[i], py = y[i], pz = z[i];
    const f32 ax = -spring * px + swirl * pz, ay = -spring * py, az = -spring * pz - swirl * px;
    const f32 nvx = __builtin_fmaf(ax, dt, vx[i]) * damp;
    const f32 nvy = __builtin_fmaf(ay, dt, vy[i]) * damp;
    const f32 nvz = __builtin_fmaf(az, dt, vz[i]) * damp;
    vx[i] = nvx;
    vy[i] = nvy;
    vz[i] = nvz;
    x[i] = __builtin_fmaf(nvx, dt, px);
    y[i] = __builtin_fmaf(nvy, dt, py);
    z[i] = __builtin_fmaf(nvz, dt, pz);
}

// Plain streaming copy (16 bytes per lane per iteration): the reference
// point for HBM byte counters (a copy of B bytes moves exactly 2 B).
__global__ void __launch_bounds__(256) simk_copy_kernel(double2 *dst, const double2 *src, u64 n) {
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

// The same with one 8-byte access per lane (512 bytes per wave access, the
// width of the solver and advection kernels' loads and stores).
__global__ void __launch_bounds__(256) simk_copy8_kernel(double *dst, const double *src, u64 n) {
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

// Sets the persistent kernels' queue head (a kernel, not a memset, so the
// sequence stays valid inside a captured graph).
__global__ void simk_queue_reset_kernel(u32 *queue, u32 value) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *queue = value;
}
__global__ void simk_queues_zero_kernel(u32 *queue, u32 count) {
    if (blockIdx.x == 0 && threadIdx.x < count) queue[threadIdx.x] = 0;
}

// Device memory of simk_kin_advance_static: the queue counter in the first 256
// bytes, the persistent kernels' workspace after it. Every GPU the library
// runs on holds its own copy from the first call until the process ends. Host
// code never passes this name to a kernel (on the host it stands for no device
// address): gpu_init resolves each device's copy with hipGetSymbolAddress and
// launches use that pointer.
#define SIMK_SCRATCH_BYTES (64ull << 20)
static __device__ __align__(256) double scratch[SIMK_SCRATCH_BYTES / sizeof(double)];

// ---------------------------------------------------------------------------
// Module state (one device, one stream; calls come from one thread at a time).

namespace {

struct Arena {
    u8 *base = nullptr;
    u64 size = 0, used = 0;
};

struct JitKernel {
    hipModule_t mod = nullptr;
    hipFunction_t fn = nullptr;
    u32 ws_slots = 0;
    std::vector<u32> mech_slots;   // multi-mechanism kernels: the slots each mechanism touches
    int blocks_per_cu = 1;
    std::string name;
    bool cached = false;
    f64 ms = 0.0;
};

struct State {
    bool ready = false;
    int device = 0;
    Arena arena;
    hipStream_t stream = nullptr;
    hipEvent_t kin_start = nullptr, kin_stop = nullptr, async_begin = nullptr, async_end = nullptr;
    u32 mode = 0;
    u32 kernel_kind = 0;   // 0 auto, 1 generated, 2 table
    GenOptions gen;
    // Settings the autotuner replaced, restored when the mechanism changes
    // (tuned settings belong to the mechanism they were measured on).
    bool tuned = false;
    GenOptions user_gen;
    u32 user_mode = 0, user_grid = 0;
    u32 team_lanes = SIMK_WAVE, user_team = SIMK_WAVE;   // lanes per cell in cooperative mode
    // current mechanism
    bool have_mech = false;
    Mech mech;
    u64 sig = 0;
    KinMethod method{};
    SimkDevCtl ctl{};
    SimkDevMethod dev_method{};
    void *d_tables = nullptr;   // SimkDevTables followed by its arrays
    void *d_wave = nullptr;     // SimkWaveTables followed by its lists (cooperative mode)
    u64 cache_budget = 208ull << 20;   // workspace of all resident waves kept within this (0: no limit)
    JitKernel jit;
    bool jit_ok = false;
    JitKernel jit_mixed;       // compiled on the first mixed-precision launch
    bool jit_mixed_ok = false;
    // Mixed-precision settings chosen by the autotuner (simk_kin_autotune with
    // precision 2); until then mixed launches use gen and grid_override.
    bool mix_tuned = false;
    GenOptions mix_gen;
    u32 mix_grid = 0;
    // The same for the FP32 factor-storage kernel (precision 3).
    bool s32_tuned = false;
    GenOptions s32_gen;
    u32 s32_grid = 0;
    JitKernel jit_s32;         // FP32 factor-storage kernel, compiled on the first precision-3 launch
    bool jit_s32_ok = false;
    JitKernel jit_d32;         // generated f32 delta kernel, compiled on the first precision-1 launch
    bool jit_d32_ok = false, jit_d32_tried = false;
    u32 sweeps = 1;
    std::string jit_log;
    // ragged batches: several mechanisms in one persistent launch
    std::vector<Mech> multi_mechs;
    std::vector<SimkDevCtl> multi_ctl;
    JitKernel multi_jit;
    bool multi_ok = false;
    JitKernel multi_jit_s32;   // FP32 factor storage form, compiled on the first precision-3 ragged launch
    // tables of each mechanism for the cooperative (team) form of the
    // multi-mechanism launch, and its team size (0: one lane per cell)
    void *multi_dtables[4] = {}, *multi_dwave[4] = {};
    // generated f32 delta kernels of the ragged list's mechanisms (serving
    // requests in f32 delta lanes), compiled on first use
    JitKernel multi_d32[4];
    bool multi_d32_tried[4] = {};
    u32 multi_team = 0, user_multi_team = 0;
    // captured step sequences
    // Captured multi-mechanism steps. Each graph owns its workspace and queue
    // counters, so later launches that grow the shared workspace never leave a
    // graph pointing at freed memory.
    struct Graph {
        hipGraphExec_t exec = nullptr;
        double *ws = nullptr;
        u32 *queue = nullptr;
    };
    std::vector<Graph> graphs;
    // workspace and queue
    double *ws = nullptr;
    u64 ws_bytes = 0;
    u32 *queue = nullptr;
    // scratch's address on each device (resolved at gpu_init), and whether the
    // launch being prepared must use it instead of allocating
    static constexpr int kMaxDevices = 64;
    u8 *scratch_at[kMaxDevices] = {};
    bool use_scratch = false;
    u32 cus = 0;
    u32 grid_override = 0;     // persistent waves per launch (0: resident waves)
    int blocks_per_cu_table = 0, blocks_per_cu_d32 = 0;
    const u32 *order = nullptr;
    // Queue order by the previous advance's per-cell attempts (policy 1, the
    // default; 0: natural order), and whether the last advance used it.
    u32 queue_policy = 1;
    SimkCostOrder cost_order;
    bool last_ordered = false;
    // last call
    u32 last_kernel = 0;
    u32 last_slots = 0;
    u32 last_waves = 0;
    std::string last_name;
    f64 last_ms = 0.0;
};

State g;

s32 err_code(hipError_t e) { return (s32)e; }

std::string module_dir() {
    Dl_info info;
    if (dladdr((void *)&gpu_init, &info) && info.dli_fname) {
        std::string p = info.dli_fname;
        size_t slash = p.rfind('/');
        if (slash != std::string::npos) return p.substr(0, slash);
    }
    return ".";
}

std::string cache_dir() {
    const char *env = getenv("SIMK_KCACHE");
    if (env && *env) return env;
    return module_dir() + "/kcache";
}

const char *kJitOptions[] = {"--offload-arch=gfx950", "-O3", "-ffp-contract=off", "-std=c++17",
                             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
const int kJitOptionCount = sizeof(kJitOptions) / sizeof(kJitOptions[0]);

std::string cache_path(const std::string &src) {
    u64 h = fnv1a64(src.data(), src.size());
    for (int i = 0; i < kJitOptionCount; ++i) h = fnv1a64(kJitOptions[i], strlen(kJitOptions[i]), h);
    int rtc_major = 0, rtc_minor = 0;
    hiprtcVersion(&rtc_major, &rtc_minor);
    h = fnv1a64(&rtc_major, sizeof(int), h);
    h = fnv1a64(&rtc_minor, sizeof(int), h);
    char name[64];
    snprintf(name, sizeof(name), "/%016llx.hsaco", (unsigned long long)h);
    return cache_dir() + name;
}

bool read_file(const std::string &path, std::vector<char> &out) {
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) return false;
    fseek(f, 0, SEEK_END);
    long size = ftell(f);
    fseek(f, 0, SEEK_SET);
    out.resize(size > 0 ? (size_t)size : 0);
    bool ok = size > 0 && fread(out.data(), 1, (size_t)size, f) == (size_t)size;
    fclose(f);
    return ok;
}

void write_file_atomic(const std::string &path, const std::vector<char> &data) {
    mkdir(cache_dir().c_str(), 0755);
    std::string tmp = path + ".tmp" + std::to_string((long)getpid());
    FILE *f = fopen(tmp.c_str(), "wb");
    if (!f) return;
    bool ok = fwrite(data.data(), 1, data.size(), f) == data.size();
    fclose(f);
    if (ok) rename(tmp.c_str(), path.c_str());
    else unlink(tmp.c_str());
}

// Compiles (or loads from the cache) the code object of a generated source.
bool jit_code(const GenResult &gr, std::vector<char> &code, bool &cached, std::string &log) {
    std::string path = cache_path(gr.source);
    if (const char *dump = getenv("SIMK_DUMP_SOURCE")) {   // tooling: keep the generated source
        FILE *f = fopen(dump, "wb");
        if (f) {
            fwrite(gr.source.data(), 1, gr.source.size(), f);
            fclose(f);
        }
    }
    if (read_file(path, code)) {
        cached = true;
        return true;
    }
    cached = false;
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, gr.source.c_str(), "simk_gen.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
        log = "hiprtcCreateProgram failed";
        return false;
    }
    hiprtcResult rc = hiprtcCompileProgram(prog, kJitOptionCount, kJitOptions);
    size_t log_size = 0;
    hiprtcGetProgramLogSize(prog, &log_size);
    if (log_size > 1) {
        log.resize(log_size);
        hiprtcGetProgramLog(prog, &log[0]);
    }
    if (rc != HIPRTC_SUCCESS) {
        hiprtcDestroyProgram(&prog);
        return false;
    }
    size_t size = 0;
    hiprtcGetCodeSize(prog, &size);
    code.resize(size);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    write_file_atomic(path, code);
    return true;
}

void unload_s32();

void free_mech_resources() {
    simk_cost_order_free(g.cost_order);   // recorded costs belong to the old mechanism's cells
    if (g.d_tables) (void)hipFree(g.d_tables);
    g.d_tables = nullptr;
    if (g.d_wave) (void)hipFree(g.d_wave);
    g.d_wave = nullptr;
    if (g.jit.mod) (void)hipModuleUnload(g.jit.mod);
    g.jit = JitKernel();
    g.jit_ok = false;
    if (g.jit_mixed.mod) (void)hipModuleUnload(g.jit_mixed.mod);
    g.jit_mixed = JitKernel();
    g.jit_mixed_ok = false;
    unload_s32();
    if (g.jit_d32.mod) (void)hipModuleUnload(g.jit_d32.mod);
    g.jit_d32 = JitKernel();
    g.jit_d32_ok = g.jit_d32_tried = false;
}

template <class T>
void put(std::vector<u8> &blob, u64 &off, const T *data, u64 count) {
    off = simk_align_up(off, 16);
    if (blob.size() < off + sizeof(T) * count) blob.resize(off + sizeof(T) * count);
    if (count) memcpy(blob.data() + off, data, sizeof(T) * count);
    off += sizeof(T) * count;
}

// Flattens the tables for the table-walk kernel into one device allocation.
hipError_t upload_tables(const Mech &m, void **out) {
    std::vector<u8> rc(m.R), nc(m.R), reac(m.R * 3), ns(m.R * 6);
    std::vector<f64> coef(m.R * 6, 0.0);
    std::vector<u16> jac(m.R * 18, 0), div(m.lu_div.size() * 4), sub(m.lu_sub.size() * 2), lo(m.lower.size() * 2),
        up(m.upper.size() * 2);
    for (u32 r = 0; r < m.R; ++r) {
        const Reaction &x = m.rx[r];
        rc[r] = x.reactant_count;
        nc[r] = x.net_count;
        for (int p = 0; p < 3; ++p) reac[r * 3 + p] = x.reactant[p];
        for (int e = 0; e < 6; ++e) {
            ns[r * 6 + e] = x.net_species[e];
            coef[r * 6 + e] = (f64)x.net_coef[e];
        }
        for (int p = 0; p < 3; ++p)
            for (int e = 0; e < 6; ++e) jac[(r * 3 + p) * 6 + e] = x.jac_index[p][e];
    }
    for (size_t d = 0; d < m.lu_div.size(); ++d) {
        div[4 * d] = m.lu_div[d].ik;
        div[4 * d + 1] = m.lu_div[d].kk;
        div[4 * d + 2] = m.lu_div[d].sub_begin;
        div[4 * d + 3] = m.lu_div[d].sub_end;
    }
    for (size_t s = 0; s < m.lu_sub.size(); ++s) {
        sub[2 * s] = m.lu_sub[s].ij;
        sub[2 * s + 1] = m.lu_sub[s].kj;
    }
    for (size_t e = 0; e < m.lower.size(); ++e) {
        lo[2 * e] = m.lower[e].index;
        lo[2 * e + 1] = m.lower[e].col;
    }
    for (size_t e = 0; e < m.upper.size(); ++e) {
        up[2 * e] = m.upper[e].index;
        up[2 * e + 1] = m.upper[e].col;
    }
    std::vector<u8> blob(sizeof(SimkDevTables), 0);
    u64 off = sizeof(SimkDevTables);
    u64 o_reac, o_rc, o_nc, o_ns, o_coef, o_jac, o_diag, o_db, o_div, o_sub, o_lb, o_ub, o_lo, o_up;
    put(blob, off, reac.data(), reac.size()); o_reac = off - reac.size();
    put(blob, off, rc.data(), rc.size()); o_rc = off - rc.size();
    put(blob, off, nc.data(), nc.size()); o_nc = off - nc.size();
    put(blob, off, ns.data(), ns.size()); o_ns = off - ns.size();
    put(blob, off, coef.data(), coef.size()); o_coef = off - 8 * coef.size();
    put(blob, off, jac.data(), jac.size()); o_jac = off - 2 * jac.size();
    put(blob, off, m.diag.data(), m.diag.size()); o_diag = off - 2 * m.diag.size();
    put(blob, off, m.div_begin.data(), m.div_begin.size()); o_db = off - 2 * m.div_begin.size();
    put(blob, off, div.data(), div.size()); o_div = off - 2 * div.size();
    put(blob, off, sub.data(), sub.size()); o_sub = off - 2 * sub.size();
    put(blob, off, m.lower_begin.data(), m.lower_begin.size()); o_lb = off - 2 * m.lower_begin.size();
    put(blob, off, m.upper_begin.data(), m.upper_begin.size()); o_ub = off - 2 * m.upper_begin.size();
    put(blob, off, lo.data(), lo.size()); o_lo = off - 2 * lo.size();
    put(blob, off, up.data(), up.size()); o_up = off - 2 * up.size();
    void *d = nullptr;
    hipError_t e = hipMalloc(&d, blob.size());
    if (e != hipSuccess) return e;
    u8 *base = (u8 *)d;
    SimkDevTables t;
    memset(&t, 0, sizeof(t));
    t.n = m.n;
    t.R = m.R;
    t.nnz = m.nnz;
    t.reactant = base + o_reac;
    t.reactant_count = base + o_rc;
    t.net_count = base + o_nc;
    t.net_species = base + o_ns;
    t.net_coef = (const double *)(base + o_coef);
    t.jac_index = (const unsigned short *)(base + o_jac);
    t.diag = (const unsigned short *)(base + o_diag);
    t.div_begin = (const unsigned short *)(base + o_db);
    t.lu_div = (const unsigned short *)(base + o_div);
    t.lu_sub = (const unsigned short *)(base + o_sub);
    t.lower_begin = (const unsigned short *)(base + o_lb);
    t.upper_begin = (const unsigned short *)(base + o_ub);
    t.lower = (const unsigned short *)(base + o_lo);
    t.upper = (const unsigned short *)(base + o_up);
    memcpy(blob.data(), &t, sizeof(t));
    e = hipMemcpy(d, blob.data(), blob.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(d);
        return e;
    }
    *out = d;
    return hipSuccess;
}

// Contribution lists of the cooperative kernel, in the reference's
// accumulation order (fun: reaction order per species; -J: (r, p, e) order per
// LU entry; solves: by column).
hipError_t upload_wave_tables(const Mech &m, void **out) {
    std::vector<u16> fb(m.n + 1), fr, jb(m.nnz + 1), jrp, lb(m.n + 1), lr, li, ub(m.n + 1), ur, ui;
    std::vector<f64> fc, jc;
    for (u32 sp = 0; sp < m.n; ++sp) {
        fb[sp] = (u16)fr.size();
        for (u32 r = 0; r < m.R; ++r)
            for (u32 e = 0; e < m.rx[r].net_count; ++e)
                if (m.rx[r].net_species[e] == sp) {
                    fr.push_back((u16)r);
                    fc.push_back((f64)m.rx[r].net_coef[e]);
                }
    }
    fb[m.n] = (u16)fr.size();
    std::vector<std::vector<std::pair<u16, f64>>> per(m.nnz);
    for (u32 r = 0; r < m.R; ++r)
        for (u32 p = 0; p < m.rx[r].reactant_count; ++p)
            for (u32 e = 0; e < m.rx[r].net_count; ++e)
                per[m.rx[r].jac_index[p][e]].push_back({(u16)(r * 3 + p), -(f64)m.rx[r].net_coef[e]});
    for (u32 idx = 0; idx < m.nnz; ++idx) {
        jb[idx] = (u16)jrp.size();
        for (const auto &c : per[idx]) {
            jrp.push_back(c.first);
            jc.push_back(c.second);
        }
    }
    jb[m.nnz] = (u16)jrp.size();
    std::vector<std::vector<std::pair<u16, u16>>> lcol(m.n), ucol(m.n);
    for (u32 i = 0; i < m.n; ++i) {
        for (u32 e = m.lower_begin[i]; e < m.lower_begin[i + 1]; ++e)
            lcol[m.lower[e].col].push_back({(u16)i, m.lower[e].index});
        for (u32 e = m.upper_begin[i]; e < m.upper_begin[i + 1]; ++e)
            ucol[m.upper[e].col].push_back({(u16)i, m.upper[e].index});
    }
    for (u32 c = 0; c < m.n; ++c) {
        lb[c] = (u16)lr.size();
        for (const auto &x : lcol[c]) {
            lr.push_back(x.first);
            li.push_back(x.second);
        }
        ub[c] = (u16)ur.size();
        for (const auto &x : ucol[c]) {
            ur.push_back(x.first);
            ui.push_back(x.second);
        }
    }
    lb[m.n] = (u16)lr.size();
    ub[m.n] = (u16)ur.size();
    if (fr.size() > 0xFFFF || jrp.size() > 0xFFFF) return hipErrorInvalidValue;
    std::vector<u8> blob(sizeof(SimkWaveTables), 0);
    u64 off = sizeof(SimkWaveTables);
    u64 o[12];
    put(blob, off, fb.data(), fb.size()); o[0] = off - 2 * fb.size();
    put(blob, off, fr.data(), fr.size()); o[1] = off - 2 * fr.size();
    put(blob, off, fc.data(), fc.size()); o[2] = off - 8 * fc.size();
    put(blob, off, jb.data(), jb.size()); o[3] = off - 2 * jb.size();
    put(blob, off, jrp.data(), jrp.size()); o[4] = off - 2 * jrp.size();
    put(blob, off, jc.data(), jc.size()); o[5] = off - 8 * jc.size();
    put(blob, off, lb.data(), lb.size()); o[6] = off - 2 * lb.size();
    put(blob, off, lr.data(), lr.size()); o[7] = off - 2 * lr.size();
    put(blob, off, li.data(), li.size()); o[8] = off - 2 * li.size();
    put(blob, off, ub.data(), ub.size()); o[9] = off - 2 * ub.size();
    put(blob, off, ur.data(), ur.size()); o[10] = off - 2 * ur.size();
    put(blob, off, ui.data(), ui.size()); o[11] = off - 2 * ui.size();
    void *d = nullptr;
    hipError_t e = hipMalloc(&d, blob.size());
    if (e != hipSuccess) return e;
    const u8 *base = (const u8 *)d;
    SimkWaveTables t;
    memset(&t, 0, sizeof(t));
    t.n = m.n;
    t.R = m.R;
    t.nnz = m.nnz;
    t.fun_begin = (const unsigned short *)(base + o[0]);
    t.fun_r = (const unsigned short *)(base + o[1]);
    t.fun_coef = (const double *)(base + o[2]);
    t.jac_begin = (const unsigned short *)(base + o[3]);
    t.jac_rp = (const unsigned short *)(base + o[4]);
    t.jac_coef = (const double *)(base + o[5]);
    t.lcol_begin = (const unsigned short *)(base + o[6]);
    t.lcol_row = (const unsigned short *)(base + o[7]);
    t.lcol_idx = (const unsigned short *)(base + o[8]);
    t.ucol_begin = (const unsigned short *)(base + o[9]);
    t.ucol_row = (const unsigned short *)(base + o[10]);
    t.ucol_idx = (const unsigned short *)(base + o[11]);
    memcpy(blob.data(), &t, sizeof(t));
    e = hipMemcpy(d, blob.data(), blob.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(d);
        return e;
    }
    *out = d;
    return hipSuccess;
}

void fill_dev_method(const KinMethod &me, SimkDevMethod &dm) {
    memset(&dm, 0, sizeof(dm));
    dm.stages = me.stages;
    dm.order = me.order;
    for (int i = 0; i < 6; ++i) {
        dm.a[i] = me.a[i];
        dm.c[i] = me.c[i];
    }
    for (int i = 0; i < 4; ++i) {
        dm.m[i] = me.m[i];
        dm.e[i] = me.e[i];
        dm.new_f[i] = me.new_f[i] ? 1 : 0;
    }
    dm.gamma0 = me.gamma0;
    stage_plan(me, dm.eval, dm.save_f);
}

// The workspace and queue counter for a persistent launch of `waves` waves
// with `wave_bytes` of workspace each. Ordinary launches grow the module's
// workspace as needed. Inside simk_kin_advance_static nothing is allocated:
// the device's scratch array serves both, `waves` is lowered to what it holds
// (results do not depend on the grid), and a wave whose workspace alone
// exceeds it is an error.
hipError_t ensure_ws(u64 bytes);
hipError_t launch_buffers(u64 wave_bytes, u32 &waves, double *&ws, u32 *&queue) {
    if (g.use_scratch) {
        u8 *base = (g.device >= 0 && g.device < State::kMaxDevices) ? g.scratch_at[g.device] : nullptr;
        if (!base) return hipErrorNotInitialized;
        const u64 room = SIMK_SCRATCH_BYTES - 256;
        if (wave_bytes > room) return hipErrorInvalidValue;
        if (wave_bytes) {
            const u64 fit = room / wave_bytes;
            if (fit < waves) waves = (u32)fit;
        }
        queue = (u32 *)base;
        ws = (double *)(base + 256);
        return hipSuccess;
    }
    hipError_t e = ensure_ws((u64)waves * wave_bytes);
    if (e != hipSuccess) return e;
    if (!g.queue) {
        e = hipMalloc(&g.queue, 256);
        if (e != hipSuccess) return e;
    }
    ws = g.ws;
    queue = g.queue;
    return hipSuccess;
}

hipError_t ensure_ws(u64 bytes) {
    if (bytes <= g.ws_bytes) return hipSuccess;
    if (g.ws) {
        hipError_t e = hipStreamSynchronize(g.stream);
        if (e != hipSuccess) return e;
        (void)hipFree(g.ws);
        g.ws = nullptr;
        g.ws_bytes = 0;
    }
    hipError_t e = hipMalloc(&g.ws, bytes);
    if (e != hipSuccess) {
        g.ws = nullptr;
        return e;
    }
    g.ws_bytes = bytes;
    return hipSuccess;
}

// Generates, compiles (or loads from the cache) and loads one kernel variant
// of the current mechanism.
// Options with the automatic LDS share resolved.
GenOptions resolved(GenOptions o);

// Options of the multi-mechanism kernel with the automatic values resolved.
GenOptions resolved_multi(GenOptions o) {
    if (o.reuse == GenOptions::kAutoReuse) o.reuse = auto_reuse(o, true);
    // Three mechanisms (60/200, 40/130, 24/70), 221,000 cells: finished rows
    // kept in registers 8 rows back instead of 16 run 7-10 % faster (the
    // multi-mechanism autotuner's pick; the kernel is at the register limit).
    if (o.reg_window == GenOptions::kDefaultWindow) o.reg_window = 8u;
    if (o.k_pairs == GenOptions::kAutoKPairs) o.k_pairs = auto_k_pairs(o, true);
    if (o.stage_carry == GenOptions::kAutoCarry) o.stage_carry = auto_stage_carry(o, true);
    return resolved(o);
}

GenOptions resolved(GenOptions o) {
    if (o.lds_slots == GenOptions::kAutoLds) o.lds_slots = auto_lds(o);
    if (o.batch == GenOptions::kAutoBatch) o.batch = auto_batch(o);
    if (o.reuse == GenOptions::kAutoReuse) o.reuse = auto_reuse(o);
    if (o.fun_in_jac == GenOptions::kAutoFunInJac) o.fun_in_jac = auto_fun_in_jac(o);
    if (o.pair_slots == GenOptions::kAutoPair) o.pair_slots = auto_pair(o);
    if (o.k_pairs == GenOptions::kAutoKPairs) o.k_pairs = auto_k_pairs(o);
    if (o.stage_carry == GenOptions::kAutoCarry) o.stage_carry = auto_stage_carry(o);
    return o;
}

bool load_generated_mech(const Mech &m, const GenOptions &opt, JitKernel &out) {
    auto t0 = std::chrono::steady_clock::now();
    GenResult gr = generate(m, g.method, resolved(opt), kDeviceKinSource);
    std::vector<char> code;
    bool cached = false, ok = false;
    if (jit_code(gr, code, cached, g.jit_log)) {
        hipModule_t mod;
        if (hipModuleLoadData(&mod, code.data()) == hipSuccess) {
            hipFunction_t fn;
            if (hipModuleGetFunction(&fn, mod, gr.kernel.c_str()) == hipSuccess) {
                int blocks = 0;
                if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, fn, SIMK_WAVE, 0) != hipSuccess)
                    blocks = 1;
                out.mod = mod;
                out.fn = fn;
                out.ws_slots = gr.ws_slots;
                out.name = gr.kernel;
                out.cached = cached;
                out.blocks_per_cu = blocks;
                ok = true;
            } else {
                (void)hipModuleUnload(mod);
            }
        }
    }
    out.ms = std::chrono::duration<f64, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return ok;
}

bool load_generated(const GenOptions &opt, JitKernel &out) { return load_generated_mech(g.mech, opt, out); }

void unload_s32() {
    if (g.jit_s32.mod) (void)hipModuleUnload(g.jit_s32.mod);
    g.jit_s32 = JitKernel();
    g.jit_s32_ok = false;
}

void unload_mixed() {
    if (g.jit_mixed.mod) (void)hipModuleUnload(g.jit_mixed.mod);
    g.jit_mixed = JitKernel();
    g.jit_mixed_ok = false;
}

void untune() {
    if (g.mix_tuned) {
        g.mix_tuned = false;
        unload_mixed();
    }
    if (g.s32_tuned) {
        g.s32_tuned = false;
        unload_s32();
    }
    if (!g.tuned) return;
    g.gen = g.user_gen;
    g.mode = g.user_mode;
    g.team_lanes = g.user_team;
    g.multi_team = g.user_multi_team;
    g.grid_override = g.user_grid;
    g.tuned = false;
}

s32 set_mech_impl(const Mech &m, const KinMethod *method, const KinControl *control) {
    if (!g.ready) return err_code(hipErrorNotInitialized);
    if (!method || !control) return err_code(hipErrorInvalidValue);
    if (method->stages < 1 || method->stages > 4) return err_code(hipErrorInvalidValue);
    std::string v = validate(m);
    if (!v.empty()) return err_code(hipErrorInvalidValue);
    hipError_t e = hipStreamSynchronize(g.stream);
    if (e != hipSuccess) return err_code(e);
    untune();
    free_mech_resources();
    g.mech = m;
    g.sig = signature(m);
    g.method = *method;
    g.method.name = nullptr;
    fill_dev_method(g.method, g.dev_method);
    g.ctl.rtol = control->rtol;
    g.ctl.atol = control->atol;
    g.ctl.h_min = control->h_min;
    g.ctl.h_max = control->h_max;
    g.ctl.h_start = control->h_start;
    g.ctl.inv_n = 1.0 / (f64)m.n;
    g.ctl.max_steps = control->max_steps;
    g.ctl.fixed_step = control->fixed_step ? 1 : 0;
    g.ctl.clamp_negative = control->clamp_negative ? 1 : 0;
    e = upload_tables(m, &g.d_tables);
    if (e == hipSuccess) e = upload_wave_tables(m, &g.d_wave);
    if (e != hipSuccess) return err_code(e);
    if (!g.blocks_per_cu_table) {
        int blocks = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, simk_table_kernel, SIMK_WAVE, 0) != hipSuccess)
            blocks = 1;
        g.blocks_per_cu_table = blocks;
    }
    g.have_mech = true;

    // Generated straight-line kernel (compiled or loaded from the disk cache).
    g.jit_ok = load_generated(g.gen, g.jit);
    if (!g.jit_ok && g.kernel_kind == 1) return err_code(hipErrorInvalidImage);
    return 0;
}

// Persistent waves to launch: the occupancy limit, lowered so the workspace
// of all resident waves stays within the last-level cache budget (the 256 MiB
// Infinity Cache of the MI355X). Every workspace access then stays on chip:
// for the 48/150 mechanism (449 KB of workspace per wave) 512 waves run 24 %
// faster than the 1,024 that fit the register file. At least one wave per CU.
u32 resident_waves(int per_cu, u64 ws_bytes_per_wave, u32 grid) {
    if (grid) return grid;
    const u32 cus = g.cus ? g.cus : 1;
    u64 r = (u64)(per_cu > 0 ? per_cu : 1) * cus;
    if (g.cache_budget && ws_bytes_per_wave) {
        u64 fit = g.cache_budget / ws_bytes_per_wave;
        if (fit < cus) fit = cus;
        if (fit < r) r = fit;
    }
    return (u32)r;
}

// Algorithmic flops of one Rodas3-like step attempt of a mechanism (FMA = 2):
// four f evaluations, -J, the LU factor, four solves and the vector work.
f64 attempt_flops(const Mech &m) {
    f64 fun = 0.0, jac = 0.0;
    for (const Reaction &x : m.rx) {
        const f64 r = x.reactant_count, e = x.net_count;
        fun += r + 2.0 * e;
        jac += r * (r - 1.0) + 2.0 * r * e;
    }
    const f64 lu = (f64)m.lu_div.size() + 2.0 * (f64)m.lu_sub.size() + 2.0 * m.n;
    const f64 solve = 2.0 * ((f64)m.nnz - m.n) + m.n;
    return 4.0 * fun + jac + lu + 4.0 * solve + 31.0 * m.n + 30.0;
}

// Longest processing time first across mechanisms: every wave starts at the
// mechanism whose cells cost the most per attempt and visits the rest in
// decreasing cost, so the costliest cells start at once on as many lanes as
// they need instead of queueing behind a small share of the waves and ending
// the launch on a tail of them. (Splitting the waves in proportion to an
// estimated work share put 6 of 512 waves on the 1,000 cells of the 60/200
// mechanism: 19 % slower on the 60/200 + 40/130 + 24/70 batch.)
void multi_schedule(SimkMultiArgs &MA, u32 count) {
    f64 cost[4];
    for (u32 i = 0; i < count; ++i) {
        cost[i] = attempt_flops(g.multi_mechs[i]);
        MA.seq[i] = i;
    }
    for (u32 i = 1; i < count; ++i)   // stable insertion sort, decreasing cost
        for (u32 j = i; j > 0 && cost[MA.seq[j]] > cost[MA.seq[j - 1]]; --j) {
            const u32 t = MA.seq[j];
            MA.seq[j] = MA.seq[j - 1];
            MA.seq[j - 1] = t;
        }
    MA.start[0] = 0;
    for (u32 k = 1; k < 4; ++k) MA.start[k] = 0xFFFFFFFFu;
}

// Workspace bytes per wave the multi-mechanism kernel actually touches: each
// wave's slots are sized for the largest mechanism, but a wave running a
// smaller one touches only that mechanism's slots. Weighted by the work
// share of each mechanism (weight[i]), this is what the cache budget must
// hold. (Sizing by the largest mechanism left 416 of 512 resident waves on
// the 60/200 + 40/130 + 24/70 batch, 84 % of whose work is 24/70: 24 % slower.)
u64 multi_wave_bytes(const f64 *weight, u32 count) {
    const JitKernel &k = g.multi_jit;
    f64 acc = 0.0, wsum = 0.0;
    for (u32 i = 0; i < count && i < k.mech_slots.size(); ++i) {
        acc += weight[i] * (f64)k.mech_slots[i];
        wsum += weight[i];
    }
    const f64 slots = wsum > 0.0 ? acc / wsum : (f64)k.ws_slots;
    return (u64)(slots * SIMK_WAVE * 8.0) + 1;
}

// Method and control in the lane type T (lane_splat's rounding for f32, base_lane.h:98).
template <class T>
SimkTConsts<T> t_consts_ctl(const SimkDevCtl &ctl);
template <class T>
SimkTConsts<T> t_consts() { return t_consts_ctl<T>(g.ctl); }
template <class T>
SimkTConsts<T> t_consts_ctl(const SimkDevCtl &ctl) {
    SimkTConsts<T> C;
    memset(&C, 0, sizeof(C));
    const KinMethod &me = g.method;
    for (int i = 0; i < 6; ++i) {
        C.a[i] = (T)me.a[i];
        C.c[i] = (T)me.c[i];
        if (me.a[i] != 0.0) C.a_nz |= 1u << i;
        if (me.c[i] != 0.0) C.c_nz |= 1u << i;
    }
    for (int i = 0; i < 4; ++i) {
        C.m[i] = (T)me.m[i];
        C.e[i] = (T)me.e[i];
        if (me.m[i] != 0.0) C.m_nz |= 1u << i;
        if (me.e[i] != 0.0) C.e_nz |= 1u << i;
        if (me.new_f[i]) C.new_f |= 1u << i;
    }
    C.gamma0 = (T)me.gamma0;
    C.stages = me.stages;
    C.order = me.order;
    C.rtol = (T)ctl.rtol;
    C.atol = (T)ctl.atol;
    C.h_min = (T)ctl.h_min;
    C.h_max = (T)ctl.h_max;
    C.h_start = (T)ctl.h_start;
    C.inv_n = (T)ctl.inv_n;   // (T)(1.0 / (f64)n), kin_solver.h:197
    C.h_max_pos = ctl.h_max > 0.0 ? 1u : 0u;
    C.max_steps = ctl.max_steps;
    C.fixed_step = ctl.fixed_step;
    C.clamp_negative = ctl.clamp_negative;
    return C;
}

// LDS bytes of one cooperative team's cell (simk_team_lds_bytes).
template <class T>
u64 team_cell_bytes(const Mech &m) {
    const u64 elems = 13ull * m.n + 5ull * m.R + m.nnz + 2;
    return ((elems * sizeof(T) + 7) & ~7ull) + 8ull * m.n;
}

// Cooperative mode (gpu_kin_set_mode(1)): teams of TL lanes per cell, 64 / TL
// cells per wave; FP64 or f32 delta.
template <class T, bool DELTA, u32 TL>
hipError_t launch_wave_t(const GpuKinLaunch *L, u32 lanes, u32 repeats) {
    const Mech &m = g.mech;
    constexpr u32 teams = SIMK_WAVE / TL;
    const u64 lds = teams * team_cell_bytes<T>(m);
    hipDeviceProp_t prop;
    hipError_t e = hipGetDeviceProperties(&prop, g.device);
    if (e != hipSuccess) return e;
    if (lds > prop.sharedMemPerBlock) return hipErrorInvalidValue;   // the wave's cells do not fit its LDS
    const auto kernel = simk_wave_kernel<T, DELTA, TL>;
    if (lds > 65536) {
        e = hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    int blocks = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, kernel, SIMK_WAVE, lds) != hipSuccess || blocks < 1)
        blocks = 1;
    u32 resident = (u32)blocks * (g.cus ? g.cus : 1);
    if (g.grid_override) resident = g.grid_override;
    const u32 needed = (lanes + teams - 1) / teams;
    u32 waves = needed < resident ? needed : resident;
    double *ws = nullptr;
    u32 *queue = nullptr;
    e = launch_buffers(0, waves, ws, queue);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(simk_queue_reset_kernel, dim3(1), dim3(64), 0, g.stream, queue, waves * teams);
    SimkKinArgs A;
    memset(&A, 0, sizeof(A));
    A.y = L->y; A.h = L->h; A.steps = L->steps; A.rejects = L->rejects; A.pulses = L->pulses; A.failed = L->failed;
    A.k = L->k; A.src = L->src; A.pulse = L->pulse; A.period = L->period; A.pulse_first = L->pulse_first;
    A.pulse_count = L->pulse_count;
    A.padded = L->padded;
    A.lanes = lanes;
    A.duration = L->duration;
    A.n_species = m.n;
    A.n_reactions = m.R;
    A.repeats = repeats;
    A.queue = queue;
    A.order = g.order;
    A.tables = (const SimkDevTables *)g.d_tables;
    A.ctl = g.ctl;
    A.method = g.dev_method;
    const SimkTConsts<T> C = t_consts<T>();
    g.last_kernel = DELTA ? 7 : 6;
    char name[64];
    snprintf(name, sizeof(name), "simk_wave_kernel<%s,%u lanes per cell>", DELTA ? "f32 delta" : "f64", TL);
    g.last_name = name;
    g.last_slots = 0;
    g.last_waves = waves;
    hipLaunchKernelGGL(kernel, dim3(waves), dim3(SIMK_WAVE), (u32)lds, g.stream, A, C,
                       (const SimkWaveTables *)g.d_wave);
    return hipGetLastError();
}

template <class T, bool DELTA>
hipError_t launch_wave_lanes(const GpuKinLaunch *L, u32 lanes, u32 repeats) {
    switch (g.team_lanes) {
    case 8: return launch_wave_t<T, DELTA, 8>(L, lanes, repeats);
    case 16: return launch_wave_t<T, DELTA, 16>(L, lanes, repeats);
    case 32: return launch_wave_t<T, DELTA, 32>(L, lanes, repeats);
    default: return launch_wave_t<T, DELTA, SIMK_WAVE>(L, lanes, repeats);
    }
}

hipError_t launch_wave(const GpuKinLaunch *L, u32 lanes, u32 repeats) {
    if (!g.d_wave) return hipErrorInvalidValue;
    if (L->precision == 0) return launch_wave_lanes<f64, false>(L, lanes, repeats);
    if (L->precision == 1) return launch_wave_lanes<f32, true>(L, lanes, repeats);
    return hipErrorNotSupported;   // mixed precision has no cooperative form
}

// Queue order of a single-mechanism advance: the caller's (simk_kin_set_order),
// else, under queue policy 1, the cells by descending attempts in the
// previous advance of the same block (rows, row stride and cell count). With
// more cells than resident lanes the costliest start first; with as many or
// fewer, cells of like cost share waves, so fewer lanes wait on a wave's
// slowest cell. Blocks under 8 waves of cells keep the natural order (the
// sort would cost more than it saves), as does the allocation-free entry
// point (ordering allocates its buffers on first use).
const u32 *order_for(const GpuKinLaunch *L, u32 lanes, u32 waves, bool &by_cost) {
    (void)waves;
    by_cost = !g.order && g.queue_policy == 1 && !g.use_scratch && lanes >= 8u * SIMK_WAVE;
    if (!by_cost) return g.order;
    SimkCostOrder &c = g.cost_order;
    const bool same = c.valid && c.y == L->y && c.padded == L->padded && c.lanes == lanes;
    if (simk_cost_order_begin(c, g.stream, L->steps, L->rejects, lanes) != hipSuccess) {
        (void)hipGetLastError();
        c.valid = false;
        by_cost = false;
        return nullptr;
    }
    g.last_ordered = same;
    return same ? c.order : nullptr;
}

// Before an advance is timed: the ordering buffers grow here, so that their
// allocation is not counted in the advance's kernel time.
void order_reserve(const GpuKinLaunch *L) {
    if (!L || g.order || g.queue_policy != 1 || g.use_scratch || g.mode != 0) return;
    const u32 lanes = L->lane_count ? L->lane_count : L->padded;
    if (lanes >= 8u * SIMK_WAVE && lanes <= L->padded)
        if (simk_cost_order_reserve(g.cost_order, g.stream, lanes) != hipSuccess) (void)hipGetLastError();
}

// After the advance is enqueued: this advance's per-cell attempts become the
// order of the next advance of the block.
hipError_t order_after(const GpuKinLaunch *L, u32 lanes, bool by_cost, hipError_t launched) {
    if (!by_cost) return launched;
    SimkCostOrder &c = g.cost_order;
    hipError_t e = launched;
    if (e == hipSuccess) e = simk_cost_order_end(c, g.stream, L->steps, L->rejects, lanes);
    c.valid = e == hipSuccess;
    c.y = L->y;
    c.padded = L->padded;
    c.lanes = lanes;
    return e;
}

// f32 delta lanes: the generated kernel (kernel kind 0 when it builds, or 1),
// else the table walk (kind 2, or 0 when generation fails).
hipError_t launch_delta32(const GpuKinLaunch *L, u32 lanes, u32 repeats) {
    const Mech &m = g.mech;
    if (g.kernel_kind != 2 && !g.jit_d32_tried) {
        if (g.use_scratch) return hipErrorNotReady;   // loading a kernel allocates
        g.jit_d32_tried = true;
        GenOptions o = g.gen;
        o.precision = 1;
        g.jit_d32_ok = load_generated(o, g.jit_d32);
    }
    if (g.kernel_kind == 1 && !g.jit_d32_ok) return hipErrorInvalidImage;
    const bool gen = g.kernel_kind != 2 && g.jit_d32_ok;
    // f32 slots per lane (simk_d32_slots), or the generated kernel's f64 slots
    const u64 slot_bytes = gen ? 8 : 4;
    const u32 slots = gen ? g.jit_d32.ws_slots : m.nnz + 11 * m.n + m.R;
    const u64 rows = m.R > m.n ? m.R : m.n;
    if (rows * L->padded * 8ull >= (1ull << 31) || (u64)slots * SIMK_WAVE * slot_bytes >= (1ull << 31))
        return hipErrorInvalidValue;
    if (!g.blocks_per_cu_d32) {
        int blocks = 1;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, simk_delta32_kernel, SIMK_WAVE, 0) != hipSuccess)
            blocks = 1;
        g.blocks_per_cu_d32 = blocks > 0 ? blocks : 1;
    }
    const u32 needed = (lanes + SIMK_WAVE - 1) / SIMK_WAVE;
    const u32 resident = resident_waves(gen ? g.jit_d32.blocks_per_cu : g.blocks_per_cu_d32,
                                        (u64)slots * SIMK_WAVE * slot_bytes, g.grid_override);
    u32 waves = needed < resident ? needed : resident;
    double *ws = nullptr;
    u32 *queue = nullptr;
    hipError_t e = launch_buffers((u64)slots * SIMK_WAVE * slot_bytes, waves, ws, queue);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(simk_queue_reset_kernel, dim3(1), dim3(64), 0, g.stream, queue, waves * SIMK_WAVE);
    SimkKinArgs A;
    memset(&A, 0, sizeof(A));
    A.y = L->y; A.h = L->h; A.steps = L->steps; A.rejects = L->rejects; A.pulses = L->pulses; A.failed = L->failed;
    A.k = L->k; A.src = L->src; A.pulse = L->pulse; A.period = L->period; A.pulse_first = L->pulse_first;
    A.pulse_count = L->pulse_count;
    A.padded = L->padded;
    A.lanes = lanes;
    A.duration = L->duration;
    A.ws = ws;
    A.ws_slots = slots;
    A.n_species = m.n;
    A.n_reactions = m.R;
    A.repeats = repeats;
    A.queue = queue;
    bool by_cost = false;
    A.order = order_for(L, lanes, waves, by_cost);
    A.tables = (const SimkDevTables *)g.d_tables;
    A.ctl = g.ctl;
    A.method = g.dev_method;
    const SimkD32Consts C = t_consts<f32>();
    g.last_kernel = gen ? 8 : 5;
    g.last_name = gen ? g.jit_d32.name : std::string("simk_delta32_kernel");
    g.last_slots = slots;
    g.last_waves = waves;
    if (gen) {
        SimkD32Args DA;
        DA.A = A;
        DA.C = C;
        size_t size = sizeof(DA);
        void *cfg[] = {HIP_LAUNCH_PARAM_BUFFER_POINTER, &DA, HIP_LAUNCH_PARAM_BUFFER_SIZE, &size, HIP_LAUNCH_PARAM_END};
        return order_after(L, lanes, by_cost,
                           hipModuleLaunchKernel(g.jit_d32.fn, waves, 1, 1, SIMK_WAVE, 1, 1, 0, g.stream, nullptr, cfg));
    }
    hipLaunchKernelGGL(simk_delta32_kernel, dim3(waves), dim3(SIMK_WAVE), 0, g.stream, A, C);
    return order_after(L, lanes, by_cost, hipGetLastError());
}

// Launches `repeats` consecutive advances of queue positions [0, lanes) on the
// module stream with the persistent kernel: grid = resident waves (or fewer
// when there are fewer cells), lanes refill from the queue.
hipError_t launch(const GpuKinLaunch *L, u32 repeats) {
    g.last_ordered = false;
    const u32 lanes = L->lane_count ? L->lane_count : L->padded;
    if (lanes == 0 || repeats == 0) return hipSuccess;
    if (lanes > L->padded) return hipErrorInvalidValue;
    if (!L->y || !L->h || !L->steps || !L->rejects || !L->pulses || !L->failed || !L->k || !L->src || !L->pulse ||
        !L->period || !L->pulse_first || !L->pulse_count)
        return hipErrorInvalidValue;
    if (L->precision > 3) return hipErrorNotSupported;
    if (g.mode == 1) return launch_wave(L, lanes, repeats);
    if (L->precision == 1) return launch_delta32(L, lanes, repeats);
    const JitKernel *jk = nullptr;
    if (L->precision == 2) {
        // Mixed precision exists only as a generated kernel.
        if (!g.jit_mixed_ok) {
            if (g.use_scratch) return hipErrorNotReady;   // loading a kernel allocates
            GenOptions mo = g.mix_tuned ? g.mix_gen : g.gen;
            mo.precision = 2;
            mo.sweeps = g.sweeps;
            g.jit_mixed_ok = load_generated(mo, g.jit_mixed);
            if (!g.jit_mixed_ok) return hipErrorInvalidImage;
        }
        jk = &g.jit_mixed;
    } else if (L->precision == 3) {
        // FP32 factor storage exists only as a generated kernel, built from
        // the current generator options.
        if (!g.jit_s32_ok) {
            if (g.use_scratch) return hipErrorNotReady;   // loading a kernel allocates
            GenOptions so = g.s32_tuned ? g.s32_gen : g.gen;
            so.precision = 3;
            g.jit_s32_ok = load_generated(so, g.jit_s32);
            if (!g.jit_s32_ok) return hipErrorInvalidImage;
        }
        jk = &g.jit_s32;
    } else {
        if (g.kernel_kind == 1 && !g.jit_ok) return hipErrorInvalidImage;
        if (g.jit_ok && g.kernel_kind != 2) jk = &g.jit;
    }
    const bool use_gen = jk != nullptr;
    const u32 slots = use_gen ? jk->ws_slots : g.mech.nnz + 10 * g.mech.n;
    // Buffer resources address each [row][padded] array with 32-bit offsets.
    const u64 rows = g.mech.R > g.mech.n ? g.mech.R : g.mech.n;
    if (rows * L->padded * 8ull >= (1ull << 31) || (u64)slots * SIMK_WAVE * 8ull >= (1ull << 31))
        return hipErrorInvalidValue;
    const u32 needed = (lanes + SIMK_WAVE - 1) / SIMK_WAVE;
    const int per_cu = use_gen ? jk->blocks_per_cu : g.blocks_per_cu_table;
    const u32 grid = L->precision == 2 && g.mix_tuned   ? g.mix_grid
                     : L->precision == 3 && g.s32_tuned ? g.s32_grid
                                                        : g.grid_override;
    const u32 resident = resident_waves(per_cu, (u64)slots * SIMK_WAVE * 8, grid);
    u32 waves = needed < resident ? needed : resident;
    double *ws = nullptr;
    u32 *queue = nullptr;
    hipError_t e = launch_buffers((u64)slots * SIMK_WAVE * sizeof(double), waves, ws, queue);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(simk_queue_reset_kernel, dim3(1), dim3(64), 0, g.stream, queue, waves * SIMK_WAVE);
    SimkKinArgs A;
    memset(&A, 0, sizeof(A));
    A.y = L->y; A.h = L->h; A.steps = L->steps; A.rejects = L->rejects; A.pulses = L->pulses; A.failed = L->failed;
    A.k = L->k; A.src = L->src; A.pulse = L->pulse; A.period = L->period; A.pulse_first = L->pulse_first;
    A.pulse_count = L->pulse_count;
    A.padded = L->padded;
    A.lanes = lanes;
    A.duration = L->duration;
    A.ws = ws;
    A.ws_slots = slots;
    A.n_species = g.mech.n;
    A.n_reactions = g.mech.R;
    A.repeats = repeats;
    A.queue = queue;
    bool by_cost = false;
    A.order = order_for(L, lanes, waves, by_cost);
    A.tables = (const SimkDevTables *)g.d_tables;
    A.ctl = g.ctl;
    A.method = g.dev_method;
    g.last_kernel = use_gen ? (L->precision == 2 ? 3 : (L->precision == 3 ? 9 : 1)) : 2;
    g.last_name = use_gen ? jk->name : std::string("simk_table_kernel");
    g.last_slots = slots;
    g.last_waves = waves;
    if (use_gen) {
        size_t size = sizeof(A);
        void *cfg[] = {HIP_LAUNCH_PARAM_BUFFER_POINTER, &A, HIP_LAUNCH_PARAM_BUFFER_SIZE, &size, HIP_LAUNCH_PARAM_END};
        return order_after(L, lanes, by_cost,
                           hipModuleLaunchKernel(jk->fn, waves, 1, 1, SIMK_WAVE, 1, 1, 0, g.stream, nullptr, cfg));
    }
    hipLaunchKernelGGL(simk_table_kernel, dim3(waves), dim3(SIMK_WAVE), 0, g.stream, A);
    return order_after(L, lanes, by_cost, hipGetLastError());
}

// Kernel arguments of one advance of a launch's block (shared by the single
// and the multi-mechanism launches).
SimkKinArgs make_args(const GpuKinLaunch *L, u32 lanes, const Mech &m, const SimkDevCtl &ctl, u32 slots,
                      u32 repeats, u32 *queue, u32 queue_first) {
    SimkKinArgs A;
    memset(&A, 0, sizeof(A));
    A.y = L->y; A.h = L->h; A.steps = L->steps; A.rejects = L->rejects; A.pulses = L->pulses; A.failed = L->failed;
    A.k = L->k; A.src = L->src; A.pulse = L->pulse; A.period = L->period; A.pulse_first = L->pulse_first;
    A.pulse_count = L->pulse_count;
    A.padded = L->padded;
    A.lanes = lanes;
    A.duration = L->duration;
    A.ws = g.ws;
    A.ws_slots = slots;
    A.n_species = m.n;
    A.n_reactions = m.R;
    A.repeats = repeats;
    A.queue_first = queue_first;
    A.queue = queue;
    A.order = nullptr;
    A.tables = nullptr;
    A.ctl = ctl;
    A.method = g.dev_method;
    return A;
}

bool launch_ok(const GpuKinLaunch *L) {
    return L && L->y && L->h && L->steps && L->rejects && L->pulses && L->failed && L->k && L->src && L->pulse &&
           L->period && L->pulse_first && L->pulse_count && L->padded > 0 &&
           (L->lane_count ? L->lane_count : L->padded) <= L->padded;
}

// SimkGenOptions (0 = default, ~0u = none where documented) to generator options.
GenOptions gen_options(const SimkGenOptions *options) {
    GenOptions o;
    if (!options) return o;
    o.fence_rows = options->fence_rows;
    o.waves_per_eu = options->waves_per_eu;
    o.unroll_stages = options->unroll_stages;
    o.reg_window = options->reg_window == 0 ? GenOptions::kDefaultWindow
                                            : (options->reg_window == 0xFFFFFFFFu ? 0u : options->reg_window);
    o.batch = options->batch == 0 ? GenOptions::kAutoBatch : options->batch;
    o.precision = options->precision;
    o.sweeps = options->sweeps == 0 ? 1u : (options->sweeps == 0xFFFFFFFFu ? 0u : options->sweeps);
    o.reuse = options->reuse == 0 ? GenOptions::kAutoReuse : (options->reuse == 0xFFFFFFFFu ? 0u : options->reuse);
    o.lds_slots = options->lds_slots == 0 ? GenOptions().lds_slots
                                          : (options->lds_slots == 0xFFFFFFFFu ? 0u : options->lds_slots);
    o.fun_in_jac = options->fun_in_jac == 0 ? GenOptions::kAutoFunInJac : (options->fun_in_jac == 0xFFFFFFFFu ? 0u : 1u);
    o.pair_slots = options->pair_slots == 0 ? GenOptions::kAutoPair : (options->pair_slots == 0xFFFFFFFFu ? 0u : 1u);
    o.k_pairs = options->k_pairs == 0 ? GenOptions::kAutoKPairs : (options->k_pairs == 0xFFFFFFFFu ? 0u : 1u);
    o.stage_carry = options->stage_carry == 0 ? GenOptions::kAutoCarry
                                              : (options->stage_carry == 0xFFFFFFFFu ? 0u : 1u);
    return o;
}

void free_graph(State::Graph &r) {
    if (r.exec) (void)hipGraphExecDestroy(r.exec);
    if (r.ws) (void)hipFree(r.ws);
    if (r.queue) (void)hipFree(r.queue);
    r = State::Graph();
}

// Caller has synchronized the stream.
void destroy_graphs() {
    for (State::Graph &r : g.graphs) free_graph(r);
    g.graphs.clear();
}

// One persistent launch over several mechanisms' blocks (ragged batch):
// prepare sizes everything (may allocate), enqueue only records stream work.
struct MultiPlan {
    SimkMultiArgs MA;
    u32 waves = 0, slots = 0;
    double *ws = nullptr;
    u32 *queue = nullptr;
    const JitKernel *jk = nullptr;   // the kernel launched (nullptr: the FP64 multi-mechanism kernel)
};

u32 multi_team_size();

}  // namespace

// Defined below, in the C-linkage block, outside the unnamed namespace.
extern "C" hipError_t load_multi(GenOptions o, JitKernel &out, u32 precision = 0);

namespace {

// own = true: allocate a workspace and queue counters for this plan alone (a
// captured graph keeps them); else use the module's shared ones.
hipError_t prepare_multi(u32 count, const GpuKinLaunch *Ls, u32 repeats, MultiPlan &P, bool own = false) {
    if (!g.multi_ok || count != g.multi_mechs.size() || repeats == 0 || !Ls) return hipErrorInvalidValue;
    // every block in FP64 (precision 0) or every block in FP32 factor storage (3)
    const u32 prec = Ls[0].precision;
    if (prec != 0 && prec != 3) return hipErrorInvalidValue;
    if (prec == 3) {
        if (multi_team_size()) return hipErrorNotSupported;   // no cooperative form
        if (!g.multi_jit_s32.mod) {
            const hipError_t le = load_multi(g.gen, g.multi_jit_s32, 3);
            if (le != hipSuccess) return le;
        }
        P.jk = &g.multi_jit_s32;
    } else {
        P.jk = &g.multi_jit;
    }
    memset(&P.MA, 0, sizeof(P.MA));
    P.MA.count = count;
    u64 total_waves = 0;
    f64 weight[4];   // work share of each mechanism, for the cache footprint
    for (u32 i = 0; i < count; ++i) {
        const GpuKinLaunch *L = Ls + i;
        if (!launch_ok(L) || L->precision != prec) return hipErrorInvalidValue;
        const Mech &m = g.multi_mechs[i];
        const u64 rows = m.R > m.n ? m.R : m.n;
        if (rows * L->padded * 8ull >= (1ull << 31)) return hipErrorInvalidValue;
        const u32 lanes = L->lane_count ? L->lane_count : L->padded;
        total_waves += (lanes + SIMK_WAVE - 1) / SIMK_WAVE;
        weight[i] = (f64)lanes * (f64)(m.nnz + 4 * m.n);   // per-cell cost ~ LU plus stage vectors
    }
    P.slots = P.jk->ws_slots;
    if ((u64)P.slots * SIMK_WAVE * 8ull >= (1ull << 31)) return hipErrorInvalidValue;
    const u32 resident = resident_waves(P.jk->blocks_per_cu, multi_wave_bytes(weight, count), g.grid_override);
    P.waves = (u32)(total_waves < resident ? total_waves : resident);
    if (P.waves == 0) P.waves = 1;
    const u64 ws_bytes = (u64)P.waves * P.slots * SIMK_WAVE * sizeof(double);
    hipError_t e;
    if (own) {
        e = hipMalloc(&P.ws, ws_bytes);
        if (e != hipSuccess) return e;
        e = hipMalloc(&P.queue, 256);
        if (e != hipSuccess) {
            (void)hipFree(P.ws);
            P.ws = nullptr;
            return e;
        }
    } else {
        e = ensure_ws(ws_bytes);
        if (e != hipSuccess) return e;
        if (!g.queue) {
            e = hipMalloc(&g.queue, 256);
            if (e != hipSuccess) return e;
        }
        P.ws = g.ws;
        P.queue = g.queue;
    }
    for (u32 i = 0; i < count; ++i) {
        const GpuKinLaunch *L = Ls + i;
        const u32 lanes = L->lane_count ? L->lane_count : L->padded;
        P.MA.a[i] = make_args(L, lanes, g.multi_mechs[i], g.multi_ctl[i], P.slots, repeats, P.queue + i, 1);
        P.MA.a[i].ws = P.ws;
    }
    multi_schedule(P.MA, count);
    return hipSuccess;
}

// The cooperative form of a planned multi-mechanism launch, teams of TL lanes.
template <u32 TL>
hipError_t enqueue_multi_team(const MultiPlan &P) {
    constexpr u32 teams = SIMK_WAVE / TL;
    SimkWaveMultiArgs M;
    memset(&M, 0, sizeof(M));
    M.count = P.MA.count;
    u64 team_bytes = 0, cells = 0;
    for (u32 i = 0; i < M.count; ++i) {
        M.a[i] = P.MA.a[i];
        M.a[i].tables = (const SimkDevTables *)g.multi_dtables[i];
        M.w[i] = (const SimkWaveTables *)g.multi_dwave[i];
        M.c[i] = t_consts_ctl<f64>(g.multi_ctl[i]);
        M.seq[i] = P.MA.seq[i];
        if (!M.a[i].tables || !M.w[i] || M.a[i].slice) return hipErrorInvalidValue;
        const u64 b = team_cell_bytes<f64>(g.multi_mechs[i]);
        if (b > team_bytes) team_bytes = b;
        cells += M.a[i].lanes;
    }
    M.team_bytes = (u32)team_bytes;
    const u64 lds = teams * team_bytes;
    hipDeviceProp_t prop;
    hipError_t e = hipGetDeviceProperties(&prop, g.device);
    if (e != hipSuccess) return e;
    if (lds > prop.sharedMemPerBlock) return hipErrorInvalidValue;
    const auto kernel = simk_wave_multi_kernel<TL>;
    if (lds > 65536) {
        e = hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    int blocks = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, kernel, SIMK_WAVE, lds) != hipSuccess || blocks < 1)
        blocks = 1;
    const u64 resident = (u64)blocks * (g.cus ? g.cus : 1);
    const u64 needed = (cells + teams - 1) / teams;
    const u32 waves = (u32)(needed < resident ? (needed ? needed : 1) : resident);
    hipLaunchKernelGGL(simk_queues_zero_kernel, dim3(1), dim3(64), 0, g.stream, P.queue, 8u);
    g.last_kernel = 6;
    char name[64];
    snprintf(name, sizeof(name), "simk_wave_multi_kernel<%u lanes per cell>", TL);
    g.last_name = name;
    g.last_slots = 0;
    g.last_waves = waves;
    hipLaunchKernelGGL(kernel, dim3(waves), dim3(SIMK_WAVE), (u32)lds, g.stream, M);
    return hipGetLastError();
}

// Team size of ragged launches: the multi-mechanism tuner's choice, else the
// caller's cooperative mode (gpu_kin_set_mode(1), simk_kin_set_team), else
// one lane per cell (0).
u32 multi_team_size() { return g.multi_team ? g.multi_team : (g.mode == 1 ? g.team_lanes : 0u); }

hipError_t enqueue_multi(const MultiPlan &P, bool allow_team = true) {
    const JitKernel &jk = P.jk ? *P.jk : g.multi_jit;
    const bool s32 = P.jk == &g.multi_jit_s32;
    if (allow_team && multi_team_size() && !s32) {
        switch (multi_team_size()) {
        case 8: return enqueue_multi_team<8>(P);
        case 16: return enqueue_multi_team<16>(P);
        case 32: return enqueue_multi_team<32>(P);
        default: return enqueue_multi_team<SIMK_WAVE>(P);
        }
    }
    hipLaunchKernelGGL(simk_queues_zero_kernel, dim3(1), dim3(64), 0, g.stream, P.queue, 8u);
    SimkMultiArgs MA = P.MA;
    size_t size = sizeof(MA);
    void *cfg[] = {HIP_LAUNCH_PARAM_BUFFER_POINTER, &MA, HIP_LAUNCH_PARAM_BUFFER_SIZE, &size, HIP_LAUNCH_PARAM_END};
    g.last_kernel = s32 ? 10 : 4;
    g.last_slots = P.slots;
    g.last_waves = P.waves;
    g.last_name = jk.name;
    return hipModuleLaunchKernel(jk.fn, P.waves, 1, 1, SIMK_WAVE, 1, 1, 0, g.stream, nullptr, cfg);
}

}  // namespace

hipStream_t simk_module_stream() { return g.stream; }

SimkMultiShape simk_multi_shape() {
    SimkMultiShape S;
    memset(&S, 0, sizeof(S));
    if (!g.multi_ok) return S;
    S.count = (u32)g.multi_mechs.size();
    S.slots = g.multi_jit.ws_slots;
    // serving: equal cells expected of every mechanism
    f64 weight[4] = {0.0, 0.0, 0.0, 0.0};
    for (u32 i = 0; i < S.count && i < 4; ++i) weight[i] = (f64)(g.multi_mechs[i].nnz + 4 * g.multi_mechs[i].n);
    S.resident_waves = resident_waves(g.multi_jit.blocks_per_cu, multi_wave_bytes(weight, S.count), g.grid_override);
    for (u32 i = 0; i < S.count; ++i) {
        S.n[i] = g.multi_mechs[i].n;
        S.R[i] = g.multi_mechs[i].R;
    }
    return S;
}

hipError_t simk_enqueue_multi_ext(u32 count, const GpuKinLaunch *Ls, const u32 *lanes, const u32 *const *orders,
                                  u32 repeats, double *ws, u32 *queue, u32 waves, const SimkSlice *slice) {
    if (!g.multi_ok || count != g.multi_mechs.size() || !Ls || !lanes || !ws || !queue || waves == 0 ||
        repeats == 0)
        return hipErrorInvalidValue;
    MultiPlan P;
    memset(&P.MA, 0, sizeof(P.MA));
    P.MA.count = count;
    P.slots = g.multi_jit.ws_slots;
    P.waves = waves;
    P.ws = ws;
    P.queue = queue;
    for (u32 i = 0; i < count; ++i) {
        const GpuKinLaunch *L = Ls + i;
        const Mech &m = g.multi_mechs[i];
        const u64 rows = m.R > m.n ? m.R : m.n;
        if (!launch_ok(L) || lanes[i] > L->padded || rows * L->padded * 8ull >= (1ull << 31))
            return hipErrorInvalidValue;
    }
    for (u32 i = 0; i < count; ++i) {
        P.MA.a[i] = make_args(Ls + i, lanes[i], g.multi_mechs[i], g.multi_ctl[i], P.slots, repeats, queue + i, 1);
        P.MA.a[i].ws = ws;
        P.MA.a[i].order = orders ? orders[i] : nullptr;
        if (slice && slice->ticks) {
            if (!slice->resume[i] || !slice->carry[i]) return hipErrorInvalidValue;
            P.MA.a[i].slice = slice->ticks;
            P.MA.a[i].resume = slice->resume[i];
            P.MA.a[i].carry = slice->carry[i];
            P.MA.a[i].carry_count = queue + 4 + i;
        }
    }
    multi_schedule(P.MA, count);
    return enqueue_multi(P, false);   // time slices and carry lists exist in the one-lane-per-cell kernel only
}

hipError_t simk_enqueue_d32_ext(u32 index, const GpuKinLaunch *L, u32 lanes, const u32 *order) {
    if (!g.multi_ok || index >= g.multi_mechs.size() || !launch_ok(L) || lanes == 0 || lanes > L->padded || !order)
        return hipErrorInvalidValue;
    const Mech &m = g.multi_mechs[index];
    if (!g.multi_d32_tried[index]) {
        g.multi_d32_tried[index] = true;
        GenOptions o = g.gen;
        o.precision = 1;
        if (!load_generated_mech(m, o, g.multi_d32[index])) g.multi_d32[index] = JitKernel();
    }
    const JitKernel &jk = g.multi_d32[index];
    if (!jk.mod) return hipErrorInvalidImage;
    const u64 rows = m.R > m.n ? m.R : m.n;
    const u64 wave_bytes = (u64)jk.ws_slots * SIMK_WAVE * 8;
    if (rows * L->padded * 8ull >= (1ull << 31) || wave_bytes >= (1ull << 31)) return hipErrorInvalidValue;
    const u32 needed = (lanes + SIMK_WAVE - 1) / SIMK_WAVE;
    const u32 resident = resident_waves(jk.blocks_per_cu, wave_bytes, 0);
    u32 waves = needed < resident ? needed : resident;
    double *ws = nullptr;
    u32 *queue = nullptr;
    hipError_t e = launch_buffers(wave_bytes, waves, ws, queue);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(simk_queue_reset_kernel, dim3(1), dim3(64), 0, g.stream, queue, waves * SIMK_WAVE);
    SimkD32Args DA;
    DA.A = make_args(L, lanes, m, g.multi_ctl[index], jk.ws_slots, 1, queue, 0);
    DA.A.ws = ws;
    DA.A.order = order;
    DA.A.tables = (const SimkDevTables *)g.multi_dtables[index];
    DA.C = t_consts_ctl<f32>(g.multi_ctl[index]);
    size_t size = sizeof(DA);
    void *cfg[] = {HIP_LAUNCH_PARAM_BUFFER_POINTER, &DA, HIP_LAUNCH_PARAM_BUFFER_SIZE, &size, HIP_LAUNCH_PARAM_END};
    return hipModuleLaunchKernel(jk.fn, waves, 1, 1, SIMK_WAVE, 1, 1, 0, g.stream, nullptr, cfg);
}

// ---------------------------------------------------------------------------
// Exports

extern "C" {

GPU_API s32 gpu_init(u32 device_index, u64 arena_bytes, GpuInfo *info) {
    if (info) memset(info, 0, sizeof(*info));
    hipError_t e = hipSetDevice((int)device_index);
    if (e != hipSuccess) return err_code(e);
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, (int)device_index);
    if (e != hipSuccess) return err_code(e);
    size_t free_b = 0, total_b = 0;
    (void)hipMemGetInfo(&free_b, &total_b);
    if (info) {
        snprintf(info->name, sizeof(info->name), "%s (%s)", prop.name, prop.gcnArchName);
        info->sm_major = (u32)prop.major;
        info->sm_minor = (u32)prop.minor;
        info->multiprocessors = (u32)prop.multiProcessorCount;
        info->free_bytes = free_b;
        info->total_bytes = total_b;
        info->arena_bytes = arena_bytes;
    }
    if (g.ready) {   // re-init: keep the stream and events, replace the arena
        (void)hipStreamSynchronize(g.stream);
        if (g.arena.base) (void)hipFree(g.arena.base);
        g.arena = Arena();
    } else {
        e = hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking);
        if (e != hipSuccess) return err_code(e);
        (void)hipEventCreate(&g.kin_start);
        (void)hipEventCreate(&g.kin_stop);
        (void)hipEventCreate(&g.async_begin);
        (void)hipEventCreate(&g.async_end);
    }
    void *base = nullptr;
    if (arena_bytes) {
        e = hipMalloc(&base, arena_bytes);
        if (e != hipSuccess) return err_code(e);
    }
    g.arena.base = (u8 *)base;
    g.arena.size = arena_bytes;
    g.arena.used = 0;
    g.device = (int)device_index;
    g.cus = (u32)prop.multiProcessorCount;
    // simk_kin_advance_static's scratch array: its address on this device,
    // resolved once per device with the device current.
    if (device_index < (u32)State::kMaxDevices && !g.scratch_at[device_index]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return err_code(e);
        g.scratch_at[device_index] = (u8 *)p;
    }
    if (const char *mb = getenv("SIMK_CACHE_BUDGET_MB")) g.cache_budget = (u64)strtoull(mb, nullptr, 10) << 20;
    g.ready = true;
    return 0;
}

GPU_API const char *gpu_error_string(s32 error) { return hipGetErrorString((hipError_t)error); }

GPU_API void *gpu_alloc(u64 size) {
    u64 aligned = simk_align_up(size, 256);
    if (!g.arena.base || g.arena.used + aligned > g.arena.size) return nullptr;
    void *p = g.arena.base + g.arena.used;
    g.arena.used += aligned;
    return p;
}

GPU_API u64 gpu_arena_used(void) { return g.arena.used; }

GPU_API void gpu_arena_reset(void) {
    if (g.ready) (void)hipDeviceSynchronize();
    g.arena.used = 0;
    g.cost_order.valid = false;   // blocks in the arena are gone
}

GPU_API s32 gpu_copy_to_device(void *dst_device, const void *src_host, u64 size) {
    hipError_t e = hipMemcpyAsync(dst_device, src_host, size, hipMemcpyHostToDevice, g.stream);
    if (e != hipSuccess) return err_code(e);
    return err_code(hipStreamSynchronize(g.stream));
}

GPU_API s32 gpu_copy_to_host(void *dst_host, const void *src_device, u64 size) {
    hipError_t e = hipMemcpyAsync(dst_host, src_device, size, hipMemcpyDeviceToHost, g.stream);
    if (e != hipSuccess) return err_code(e);
    return err_code(hipStreamSynchronize(g.stream));
}

GPU_API s32 gpu_sync(void) { return err_code(hipStreamSynchronize(g.stream)); }

GPU_API s32 gpu_step_points(f32 *x, f32 *y, f32 *z, f32 *vx, f32 *vy, f32 *vz, u32 count, f32 dt) {
    if (count == 0) return 0;
    const u32 block = 256, grid = (count + block - 1) / block;
    hipLaunchKernelGGL(simk_step_points_kernel, dim3(grid), dim3(block), 0, g.stream, x, y, z, vx, vy, vz, count, dt);
    return err_code(hipGetLastError());
}

GPU_API s32 simk_kin_set_mech_caps(const void *mech, const SimkCaps *caps, const KinMethod *method,
                                   const KinControl *control) {
    if (!mech || !caps) return err_code(hipErrorInvalidValue);
    Mech m;
    if (!from_bytes(mech, *caps, m).empty()) return err_code(hipErrorInvalidValue);
    return set_mech_impl(m, method, control);
}

GPU_API s32 gpu_kin_set_mech(const KinMech *mech, const KinMethod *method, const KinControl *control) {
    SimkCaps c = {SIMK_DEFAULT_SPECIES, SIMK_DEFAULT_REACTIONS, SIMK_DEFAULT_NNZ, SIMK_DEFAULT_LU_DIV,
                  SIMK_DEFAULT_LU_SUB};
    return simk_kin_set_mech_caps(mech, &c, method, control);
}

GPU_API s32 gpu_kin_advance(const GpuKinLaunch *L, f32 *kernel_ms) {
    if (!g.have_mech || !L) return err_code(hipErrorInvalidValue);
    order_reserve(L);
    (void)hipEventRecord(g.kin_start, g.stream);
    hipError_t e = launch(L, 1);
    if (e != hipSuccess) return err_code(e);
    (void)hipEventRecord(g.kin_stop, g.stream);
    e = hipEventSynchronize(g.kin_stop);
    if (e != hipSuccess) return err_code(e);
    f32 ms = 0.0f;
    (void)hipEventElapsedTime(&ms, g.kin_start, g.kin_stop);
    g.last_ms = ms;
    if (kernel_ms) *kernel_ms = ms;
    return err_code(hipGetLastError());
}

GPU_API s32 gpu_host_register(void *host, u64 size) { return (host && size) ? 0 : err_code(hipErrorInvalidValue); }

GPU_API s32 gpu_host_unregister(void *host) {
    (void)hipStreamSynchronize(g.stream);
    return host ? 0 : err_code(hipErrorInvalidValue);
}

GPU_API s32 gpu_kin_advance_async(const GpuKinLaunch *L, void *host_block, void *device_block, u64 rows,
                                  u64 mutable_rows) {
    if (!g.have_mech || !L || !host_block || !device_block || mutable_rows > rows) return err_code(hipErrorInvalidValue);
    const u64 pitch = (u64)L->padded * sizeof(f64);
    const u64 width = (u64)(L->lane_count ? L->lane_count : L->padded) * sizeof(f64);
    order_reserve(L);
    (void)hipEventRecord(g.async_begin, g.stream);
    hipError_t e = hipMemcpy2DAsync(device_block, pitch, host_block, pitch, width, rows, hipMemcpyHostToDevice, g.stream);
    if (e != hipSuccess) return err_code(e);
    (void)hipEventRecord(g.kin_start, g.stream);
    e = launch(L, 1);
    if (e != hipSuccess) return err_code(e);
    (void)hipEventRecord(g.kin_stop, g.stream);
    e = hipMemcpy2DAsync(host_block, pitch, device_block, pitch, width, mutable_rows, hipMemcpyDeviceToHost, g.stream);
    if (e != hipSuccess) return err_code(e);
    (void)hipEventRecord(g.async_end, g.stream);
    return err_code(hipGetLastError());
}

GPU_API s32 gpu_kin_wait(f32 *kernel_ms, f32 *total_ms) {
    hipError_t e = hipEventSynchronize(g.async_end);
    if (e != hipSuccess) return err_code(e);
    f32 k = 0.0f, t = 0.0f;
    (void)hipEventElapsedTime(&k, g.kin_start, g.kin_stop);
    (void)hipEventElapsedTime(&t, g.async_begin, g.async_end);
    g.last_ms = k;
    if (kernel_ms) *kernel_ms = k;
    if (total_ms) *total_ms = t;
    return err_code(hipGetLastError());
}

GPU_API s32 gpu_kin_set_mode(u32 mode) {
    if (mode > 1) return err_code(hipErrorInvalidValue);
    g.mode = mode;
    g.user_mode = mode;
    return 0;
}

GPU_API s32 simk_kin_set_team(u32 lanes_per_cell) {
    if (lanes_per_cell != 8 && lanes_per_cell != 16 && lanes_per_cell != 32 && lanes_per_cell != SIMK_WAVE)
        return err_code(hipErrorInvalidValue);
    g.team_lanes = lanes_per_cell;
    g.user_team = lanes_per_cell;
    return 0;
}

GPU_API s32 simk_kin_advance_n(const GpuKinLaunch *L, u32 repeats, f32 *kernel_ms) {
    if (!g.have_mech || !L) return err_code(hipErrorInvalidValue);
    order_reserve(L);
    (void)hipEventRecord(g.kin_start, g.stream);
    hipError_t e = launch(L, repeats);
    if (e != hipSuccess) return err_code(e);
    (void)hipEventRecord(g.kin_stop, g.stream);
    e = hipEventSynchronize(g.kin_stop);
    if (e != hipSuccess) return err_code(e);
    f32 ms = 0.0f;
    (void)hipEventElapsedTime(&ms, g.kin_start, g.kin_stop);
    g.last_ms = ms;
    if (kernel_ms) *kernel_ms = ms;
    return err_code(hipGetLastError());
}

GPU_API s32 simk_kin_advance_static(const GpuKinLaunch *L, u32 repeats, f32 *kernel_ms) {
    if (!g.have_mech || !L) return err_code(hipErrorInvalidValue);
    (void)hipEventRecord(g.kin_start, g.stream);
    g.use_scratch = true;
    hipError_t e = launch(L, repeats);
    g.use_scratch = false;
    if (e != hipSuccess) return err_code(e);
    (void)hipEventRecord(g.kin_stop, g.stream);
    e = hipEventSynchronize(g.kin_stop);
    if (e != hipSuccess) return err_code(e);
    f32 ms = 0.0f;
    (void)hipEventElapsedTime(&ms, g.kin_start, g.kin_stop);
    g.last_ms = ms;
    if (kernel_ms) *kernel_ms = ms;
    return err_code(hipGetLastError());
}

GPU_API s32 simk_kin_set_grid(u32 waves) {
    g.grid_override = waves;
    g.user_grid = waves;
    return 0;
}

GPU_API s32 simk_kin_set_queue_policy(u32 policy) {
    if (policy > 1) return err_code(hipErrorInvalidValue);
    g.queue_policy = policy;
    g.cost_order.valid = false;
    return 0;
}

GPU_API s32 simk_kin_set_order(const u32 *order_device) {
    g.order = order_device;
    return 0;
}

GPU_API s32 simk_kin_set_kernel(u32 kind) {
    if (kind > 2) return err_code(hipErrorInvalidValue);
    g.kernel_kind = kind;
    return 0;
}

GPU_API s32 simk_kin_set_gen_options(const SimkGenOptions *o) {
    if (!o) return err_code(hipErrorInvalidValue);
    const GenOptions n = gen_options(o);
    g.gen.fence_rows = n.fence_rows;
    g.gen.waves_per_eu = n.waves_per_eu;
    g.gen.unroll_stages = n.unroll_stages;
    g.gen.reg_window = n.reg_window;
    g.gen.reuse = n.reuse;
    g.gen.lds_slots = n.lds_slots;
    g.gen.fun_in_jac = n.fun_in_jac;
    g.gen.pair_slots = n.pair_slots;
    g.gen.k_pairs = n.k_pairs;
    g.gen.stage_carry = n.stage_carry;
    g.gen.batch = n.batch;
    unload_s32();   // the FP32 factor-storage kernels are built from these options
    if (g.multi_jit_s32.mod) (void)hipModuleUnload(g.multi_jit_s32.mod);
    g.multi_jit_s32 = JitKernel();
    // explicit options are what a later untune() restores
    if (g.tuned) g.user_gen = g.gen;
    const u32 sw = n.sweeps;
    // explicit options replace tuned mixed-precision settings
    if (sw != g.sweeps || g.mix_tuned) unload_mixed();
    g.mix_tuned = false;
    g.s32_tuned = false;
    g.sweeps = sw;
    return 0;
}

// ---- autotuner ------------------------------------------------------------
// A device copy of cells [0, cells) of a launch's block (compact, own row
// stride), so candidates run on identical inputs without touching the caller's.
constexpr f32 kTuneMargin = 0.98f;   // a candidate replaces the best only when 2 % faster

struct TuneSample {
    u8 *mem = nullptr;
    GpuKinLaunch orig{}, work{};
    u64 half = 0;   // bytes of one copy
};

hipError_t tune_sample(const Mech &m, const GpuKinLaunch *L, u32 cells, TuneSample &S) {
    const u64 p = ((u64)cells + 63) & ~63ull;
    const u32 rows[12] = {m.n, 1, 1, 1, 1, 1, m.R, m.n, m.n, 1, 1, 1};
    const f64 *src[12] = {L->y, L->h, L->steps, L->rejects, L->pulses, L->failed,
                          L->k, L->src, L->pulse, L->period, L->pulse_first, L->pulse_count};
    u64 total = 0;
    for (u32 r : rows) total += r;
    S.half = total * p * 8;
    hipError_t e = hipMalloc(&S.mem, 2 * S.half);
    if (e != hipSuccess) return e;
    f64 *dst[2][12];
    for (int c = 0; c < 2; ++c) {
        f64 *b = (f64 *)(S.mem + c * S.half);
        for (int i = 0; i < 12; ++i) {
            dst[c][i] = b;
            b += rows[i] * p;
        }
    }
    for (int i = 0; i < 12 && e == hipSuccess; ++i)
        e = hipMemcpy2DAsync(dst[0][i], p * 8, src[i], (u64)L->padded * 8, (u64)cells * 8, rows[i],
                             hipMemcpyDeviceToDevice, g.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g.stream);
    for (int c = 0; c < 2; ++c) {
        GpuKinLaunch &T = c ? S.work : S.orig;
        T = *L;
        T.y = dst[c][0]; T.h = dst[c][1]; T.steps = dst[c][2]; T.rejects = dst[c][3]; T.pulses = dst[c][4];
        T.failed = dst[c][5]; T.k = dst[c][6]; T.src = dst[c][7]; T.pulse = dst[c][8]; T.period = dst[c][9];
        T.pulse_first = dst[c][10]; T.pulse_count = dst[c][11];
        T.padded = (u32)p;
        T.lane_count = cells;
    }
    return e;
}

// Kernel milliseconds of one advance of the sample under the current settings
// (best of `reps`; inputs restored before each run).
hipError_t tune_time(TuneSample &S, u32 reps, f32 &best) {
    best = 1e30f;
    for (u32 r = 0; r < reps; ++r) {
        hipError_t e = hipMemcpyAsync(S.mem + S.half, S.mem, S.half, hipMemcpyDeviceToDevice, g.stream);
        if (e != hipSuccess) return e;
        (void)hipEventRecord(g.kin_start, g.stream);
        e = launch(&S.work, 1);
        if (e != hipSuccess) return e;
        (void)hipEventRecord(g.kin_stop, g.stream);
        e = hipEventSynchronize(g.kin_stop);
        if (e != hipSuccess) return e;
        f32 ms = 0.0f;
        (void)hipEventElapsedTime(&ms, g.kin_start, g.kin_stop);
        if (ms < best) best = ms;
    }
    return hipSuccess;
}

GPU_API s32 simk_kin_set_cache_budget(u64 bytes) {
    g.cache_budget = bytes;
    return 0;
}

bool same_gen(const GenOptions &a, const GenOptions &b) {
    return a.fence_rows == b.fence_rows && a.waves_per_eu == b.waves_per_eu && a.unroll_stages == b.unroll_stages &&
           a.batch == b.batch && a.precision == b.precision && a.sweeps == b.sweeps &&
           a.reg_window == b.reg_window && a.reuse == b.reuse && a.lds_slots == b.lds_slots &&
           a.fun_in_jac == b.fun_in_jac && a.pair_slots == b.pair_slots && a.k_pairs == b.k_pairs &&
           a.stage_carry == b.stage_carry;
}

// Mixed precision (precision 2) and FP32 factor storage (precision 3): the
// same protocol over that kernel's own knobs. The LU entries held in LDS set
// how many waves fit a CU, so each LDS share is timed at the cache-aware grid
// and at every per-CU wave count that share allows; then, at the best pair,
// load-pipeline depth, register window, rate-constant reuse and stage
// unrolling one at a time. The winner serves later advances of that precision
// only; FP64 settings are untouched.
s32 autotune_mixed(const GpuKinLaunch *L, u32 cells, f64 duration, SimkTuneResult *out, u32 prec) {
    hipError_t e = hipStreamSynchronize(g.stream);
    if (e != hipSuccess) return err_code(e);
    const bool p2 = prec == 2;
    bool &tuned = p2 ? g.mix_tuned : g.s32_tuned;
    GenOptions &tgen = p2 ? g.mix_gen : g.s32_gen;
    u32 &tgrid = p2 ? g.mix_grid : g.s32_grid;
    JitKernel &jk = p2 ? g.jit_mixed : g.jit_s32;
    bool &jk_ok = p2 ? g.jit_mixed_ok : g.jit_s32_ok;
    auto unload = [&]() {
        if (p2) unload_mixed();
        else unload_s32();
    };
    // the options a kernel of this precision is generated from
    auto kernel_options = [&](const GenOptions &opt) {
        GenOptions mo = opt;
        mo.precision = prec;
        if (p2) mo.sweeps = g.sweeps;
        return mo;
    };
    TuneSample S;
    e = tune_sample(g.mech, L, cells, S);
    if (duration > 0.0) S.orig.duration = S.work.duration = duration;
    S.work.precision = S.orig.precision = prec;
    const bool saved_tuned = tuned;
    const GenOptions saved_gen = tuned ? tgen : g.gen;
    const u32 saved_grid = tuned ? tgrid : g.grid_override;
    const u32 cus = g.cus ? g.cus : 1;
    f32 best = 1e30f, t = 0.0f;
    GenOptions best_gen = saved_gen, loaded = saved_gen;
    u32 best_grid = saved_grid;
    bool have_loaded = false;
    tuned = true;
    auto consider = [&](u32 grid, const GenOptions &opt) -> hipError_t {
        if (!have_loaded || !same_gen(opt, loaded)) {
            JitKernel k;
            if (!load_generated(kernel_options(opt), k)) return hipSuccess;   // variant does not build: skip it
            unload();
            jk = k;
            jk_ok = true;
            loaded = opt;
            have_loaded = true;
        }
        tgen = opt;
        tgrid = grid;
        hipError_t r = tune_time(S, 2, t);
        if (r != hipSuccess) return r;
        out->candidates++;
        if (grid == saved_grid && same_gen(opt, saved_gen)) out->default_ms = t;
        if (t < best * kTuneMargin) {
            best = t;
            best_gen = opt;
            best_grid = grid;
        }
        return hipSuccess;
    };
    // 1. LDS share x grid (precision 2: f32 LU entries; precision 3: the
    // lane's f32 words, the FP64 pivots included)
    e = e == hipSuccess ? consider(saved_grid, saved_gen) : e;
    const u32 ldss2[] = {GenOptions::kAutoLds, 320, 240, 160};
    const u32 ldss3[] = {GenOptions::kAutoLds, 320, 288, 256, 208};
    const u32 *ldss = p2 ? ldss2 : ldss3;
    const u32 nlds = p2 ? 4u : 5u;
    for (u32 li = 0; li < nlds; ++li) {
        GenOptions o = saved_gen;
        o.lds_slots = ldss[li];
        const GenOptions ro = resolved(kernel_options(o));
        const u32 entries = lds_entries(g.mech, ro);
        // LDS bytes per wave, and the waves per CU they allow (160 KiB per CU)
        const u32 bytes = (p2 ? entries * 4u : g.mech.n * 8u + entries * 4u) * SIMK_WAVE;
        const u32 per_cu = bytes ? (u32)((160u * 1024u) / bytes) : 4u;
        const u32 grids[] = {0, cus * (per_cu < 4 ? per_cu : 4)};
        for (u32 gr : grids)
            if (e == hipSuccess) e = consider(gr, o);
    }
    // 2. generator knobs at the best pair, one at a time
    const u32 batches[] = {8, 12, 20, 24, 28, 32}, windows[] = {4, 16}, reuses[] = {0u, 48, 96, 384};
    const GenOptions base = best_gen;
    const u32 base_grid = best_grid;
    for (u32 b : batches)
        if (e == hipSuccess) {
            GenOptions o = base;
            o.batch = b;
            e = consider(base_grid, o);
        }
    for (u32 w : windows)
        if (e == hipSuccess) {
            GenOptions o = best_gen;
            o.reg_window = w;
            e = consider(best_grid, o);
        }
    for (u32 r : reuses)
        if (e == hipSuccess) {
            GenOptions o = best_gen;
            o.reuse = r;
            e = consider(best_grid, o);
        }
    if (e == hipSuccess) {
        GenOptions o = best_gen;
        o.unroll_stages = 1;
        e = consider(best_grid, o);
    }
    if (e == hipSuccess) {
        GenOptions o = best_gen;
        o.stage_carry = resolved(kernel_options(best_gen)).stage_carry ? 0u : 1u;
        e = consider(best_grid, o);
    }
    if (S.mem) (void)hipFree(S.mem);
    if (e != hipSuccess) {
        tuned = saved_tuned;
        tgen = saved_gen;
        tgrid = saved_grid;
        unload();
        return err_code(e);
    }
    tgen = best_gen;
    tgrid = best_grid;
    if (!have_loaded || !same_gen(best_gen, loaded)) unload();   // reloaded on the next launch of that precision
    const GenOptions ro = resolved(kernel_options(best_gen));
    out->mode = 0;
    out->lanes_per_cell = 1;
    out->waves = best_grid;
    out->batch = ro.batch;
    out->reg_window = best_gen.reg_window;
    out->unroll_stages = best_gen.unroll_stages;
    out->reuse = ro.reuse;
    out->lds_slots = lds_entries(g.mech, ro);
    out->best_ms = best;
    out->sample_cells = cells;
    return 0;
}

static s32 autotune_impl(const GpuKinLaunch *L, u32 sample_cells, f64 duration, SimkTuneResult *out);

// The tuner times every candidate on the same copy of the same cells: a
// queue ordered by an earlier candidate's costs would favour the later ones,
// so candidates run in natural order.
GPU_API s32 simk_kin_autotune(const GpuKinLaunch *L, u32 sample_cells, f64 duration, SimkTuneResult *out) {
    const u32 saved = g.queue_policy;
    g.queue_policy = 0;
    const s32 r = autotune_impl(L, sample_cells, duration, out);
    g.queue_policy = saved;
    g.cost_order.valid = false;
    return r;
}

static s32 autotune_impl(const GpuKinLaunch *L, u32 sample_cells, f64 duration, SimkTuneResult *out) {
    if (!g.have_mech || !L || !out || !launch_ok(L) || L->precision == 1 || L->precision > 3)
        return err_code(hipErrorInvalidValue);
    if (L->precision >= 2) {
        memset(out, 0, sizeof(*out));
        const u32 lanes = L->lane_count ? L->lane_count : L->padded;
        const u32 cells = sample_cells && sample_cells < lanes ? sample_cells : lanes;
        if (cells == 0) return err_code(hipErrorInvalidValue);
        return autotune_mixed(L, cells, duration, out, L->precision);
    }
    memset(out, 0, sizeof(*out));
    const u32 lanes = L->lane_count ? L->lane_count : L->padded;
    const u32 cells = sample_cells && sample_cells < lanes ? sample_cells : lanes;
    if (cells == 0) return err_code(hipErrorInvalidValue);
    hipError_t e = hipStreamSynchronize(g.stream);
    if (e != hipSuccess) return err_code(e);
    TuneSample S;
    e = tune_sample(g.mech, L, cells, S);
    if (duration > 0.0) S.orig.duration = S.work.duration = duration;
    const u32 saved_mode = g.mode, saved_grid = g.grid_override, saved_kind = g.kernel_kind;
    const u32 saved_team = g.team_lanes;
    const GenOptions saved_gen = g.gen;
    const u32 cus = g.cus ? g.cus : 1;
    f32 best = 1e30f, t = 0.0f;
    u32 best_mode = 0, best_grid = 0, best_team = saved_team;
    GenOptions best_gen = g.gen;
    auto consider = [&](u32 mode, u32 grid, const GenOptions &opt, u32 sub = 0) -> hipError_t {
        g.mode = mode;
        g.grid_override = grid;
        // sub: time only the first `sub` cells and scale to the sample (a
        // persistent launch far beyond its resident cells scales linearly)
        const u32 use = sub && sub < cells ? sub : cells;
        S.work.lane_count = use;
        if (mode == 0 && (!same_gen(opt, g.gen) || !g.jit_ok)) {
            JitKernel k;
            if (!load_generated(opt, k)) return hipSuccess;   // variant does not build: skip it
            if (g.jit.mod) (void)hipModuleUnload(g.jit.mod);
            g.jit = k;
            g.jit_ok = true;
            g.gen = opt;
        }
        hipError_t r = tune_time(S, 2, t);
        S.work.lane_count = cells;
        if (r != hipSuccess) return r;
        t = t * ((f32)cells / (f32)use);
        out->candidates++;
        if (mode == 0 && grid == 0 && same_gen(opt, saved_gen)) out->default_ms = t;
        // a candidate must beat the best so far by 2 %: timing noise never
        // replaces the default with an equal setting
        if (t < best * kTuneMargin) {
            best = t;
            best_mode = mode;
            best_grid = grid;
            best_gen = opt;
            best_team = g.team_lanes;
        }
        return hipSuccess;
    };
    g.kernel_kind = 0;
    // 1. waves per CU (0: the cache-aware default), one cell per lane
    const u32 grids[] = {0, cus, cus * 3 / 2, cus * 2, cus * 5 / 2, cus * 3, cus * 4};
    for (u32 i = 0; i < sizeof(grids) / sizeof(grids[0]) && e == hipSuccess; ++i) e = consider(0, grids[i], g.gen);
    // 2. generator knobs at the best grid, one at a time
    const u32 batches[] = {8, 16}, windows[] = {8, 32}, unrolls[] = {1};
    for (u32 b : batches)
        if (e == hipSuccess) {
            GenOptions o = best_gen;
            o.batch = b;
            e = consider(0, best_grid, o);
        }
    for (u32 w : windows)
        if (e == hipSuccess) {
            GenOptions o = best_gen;
            o.reg_window = w;
            e = consider(0, best_grid, o);
        }
    for (u32 u : unrolls)
        if (e == hipSuccess) {
            GenOptions o = best_gen;
            o.unroll_stages = u;
            e = consider(0, best_grid, o);
        }
    const u32 reuses[] = {0, 48, 96};
    for (u32 r : reuses)
        if (e == hipSuccess) {
            GenOptions o = best_gen;
            o.reuse = r;
            e = consider(0, best_grid, o);
        }
    // register budget: at least two waves per SIMD (at most 256 VGPRs a lane)
    if (e == hipSuccess) {
        GenOptions o = best_gen;
        o.waves_per_eu = 2;
        e = consider(0, best_grid, o);
    }
    // workspace slots one per 8-byte access instead of paired (or paired, if off)
    if (e == hipSuccess) {
        GenOptions o = best_gen;
        o.pair_slots = resolved(best_gen).pair_slots ? 0u : 1u;
        e = consider(0, best_grid, o);
    }
    // the paired copy of the rate constants for the f passes, off (or on)
    if (e == hipSuccess && resolved(best_gen).pair_slots) {
        GenOptions o = best_gen;
        o.k_pairs = resolved(best_gen).k_pairs ? 0u : 1u;
        e = consider(0, best_grid, o);
    }
    // the stage carry, on (or off)
    if (e == hipSuccess) {
        GenOptions o = best_gen;
        o.stage_carry = resolved(best_gen).stage_carry ? 0u : 1u;
        e = consider(0, best_grid, o);
    }
    // LU entries in LDS: none, and shares for 4, 3 and 2 waves per CU
    const u32 lds_now = lds_entries(g.mech, resolved(best_gen));
    const u32 ldss[] = {0, 80, 128, 160};
    for (u32 l : ldss)
        if (e == hipSuccess && lds_entries(g.mech, [&] { GenOptions t = best_gen; t.lds_slots = l; return t; }()) != lds_now) {
            GenOptions o = best_gen;
            o.lds_slots = l;
            e = consider(0, best_grid, o);
        }
    // 3. teams of 64, 32, 16 and 8 lanes per cell (cooperative mode), each
    // timed on at most 16 resident rounds of its grid
    {
        hipDeviceProp_t prop;
        const bool have_prop = hipGetDeviceProperties(&prop, g.device) == hipSuccess;
        const u32 team_sizes[] = {SIMK_WAVE, 32, 16, 8};
        for (u32 tl : team_sizes) {
            if (e != hipSuccess) break;
            const u64 lds = (SIMK_WAVE / tl) * team_cell_bytes<f64>(g.mech);
            if (!have_prop || lds > prop.sharedMemPerBlock) continue;
            int blocks = 1;
            hipError_t oc = hipErrorInvalidValue;
            switch (tl) {
            case 8: oc = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, simk_wave_kernel<f64, false, 8>, SIMK_WAVE, lds); break;
            case 16: oc = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, simk_wave_kernel<f64, false, 16>, SIMK_WAVE, lds); break;
            case 32: oc = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, simk_wave_kernel<f64, false, 32>, SIMK_WAVE, lds); break;
            default: oc = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, simk_wave_kernel<f64, false, SIMK_WAVE>, SIMK_WAVE, lds); break;
            }
            if (oc != hipSuccess || blocks < 1) blocks = 1;
            g.team_lanes = tl;
            e = consider(1, 0, best_gen, 16u * (u32)blocks * cus * (SIMK_WAVE / tl));
        }
        g.team_lanes = saved_team;
    }
    if (S.mem) (void)hipFree(S.mem);
    g.kernel_kind = saved_kind;
    if (e != hipSuccess) {
        g.mode = saved_mode;
        g.grid_override = saved_grid;
        g.team_lanes = saved_team;
        return err_code(e);
    }
    // apply the winner
    if (!g.tuned) {
        g.tuned = true;
        g.user_gen = saved_gen;
        g.user_mode = saved_mode;
        g.user_team = saved_team;
        g.user_grid = saved_grid;
    }
    g.mode = best_mode;
    g.grid_override = best_grid;
    if (best_mode == 1) g.team_lanes = best_team;
    if (!same_gen(best_gen, g.gen)) {
        JitKernel k;
        if (load_generated(best_gen, k)) {
            if (g.jit.mod) (void)hipModuleUnload(g.jit.mod);
            g.jit = k;
            g.jit_ok = true;
            g.gen = best_gen;
        }
    }
    out->mode = best_mode;
    out->lanes_per_cell = best_mode ? g.team_lanes : 1;
    out->waves = best_grid;
    out->batch = resolved(best_gen).batch;
    out->reg_window = best_gen.reg_window;
    out->unroll_stages = best_gen.unroll_stages;
    out->reuse = resolved(best_gen).reuse;
    out->lds_slots = lds_entries(g.mech, resolved(best_gen));
    out->best_ms = best;
    out->sample_cells = cells;
    return 0;
}

GPU_API s32 simk_kin_info(SimkKinInfo *out) {
    if (!out) return err_code(hipErrorInvalidValue);
    memset(out, 0, sizeof(*out));
    out->kernel = g.last_kernel;
    out->ws_slots = g.last_slots;
    out->waves = g.last_waves;
    out->ordered = g.last_ordered ? 1u : 0u;
    out->species = g.mech.n;
    out->reactions = g.mech.R;
    out->nnz = g.mech.nnz;
    out->lu_div = (u32)g.mech.lu_div.size();
    out->lu_sub = (u32)g.mech.lu_sub.size();
    out->jit_cached = g.jit.cached ? 1 : 0;
    out->jit_ms = g.jit.ms;
    out->last_kernel_ms = g.last_ms;
    out->signature = g.sig;
    snprintf(out->kernel_name, sizeof(out->kernel_name), "%s",
             g.last_name.empty() ? g.jit.name.c_str() : g.last_name.c_str());
    return 0;
}

GPU_API s32 simk_jit_precompile(const void *mech, const SimkCaps *caps, const KinMethod *method,
                                const SimkGenOptions *options, char *log, u32 log_size) {
    if (!mech || !caps || !method) return err_code(hipErrorInvalidValue);
    Mech m;
    if (!from_bytes(mech, *caps, m).empty()) return err_code(hipErrorInvalidValue);
    GenOptions o = gen_options(options);
    KinMethod me = *method;
    me.name = nullptr;
    GenResult gr = generate(m, me, resolved(o), kDeviceKinSource);
    std::vector<char> code;
    bool cached = false;
    std::string text;
    bool ok = jit_code(gr, code, cached, text);
    if (log && log_size) snprintf(log, log_size, "%s", text.c_str());
    return ok ? 0 : err_code(hipErrorInvalidImage);
}

GPU_API s32 simk_jit_precompile_multi(u32 count, const void *const *mechs, const SimkCaps *caps,
                                      const KinMethod *method, const SimkGenOptions *options, char *log,
                                      u32 log_size) {
    if (count == 0 || count > 4 || !mechs || !caps || !method) return err_code(hipErrorInvalidValue);
    std::vector<Mech> ms(count);
    std::vector<const Mech *> ptrs;
    for (u32 i = 0; i < count; ++i) {
        if (!mechs[i] || !from_bytes(mechs[i], caps[i], ms[i]).empty()) return err_code(hipErrorInvalidValue);
        ptrs.push_back(&ms[i]);
    }
    GenOptions o = gen_options(options);
    if (o.precision != 3) o.precision = 0;   // the multi-mechanism kernel: FP64 or FP32 factor storage
    KinMethod me = *method;
    me.name = nullptr;
    GenResult gr = generate_multi(ptrs, me, resolved_multi(o), kDeviceKinSource);
    std::vector<char> code;
    bool cached = false;
    std::string text;
    bool ok = jit_code(gr, code, cached, text);
    if (log && log_size) snprintf(log, log_size, "%s", text.c_str());
    return ok ? 0 : err_code(hipErrorInvalidImage);
}

static s32 copy_device(void *dst, const void *src, u64 bytes, f32 *ms, u32 width) {
    if (!dst || !src || (bytes & 15) || ((u64)dst & 15) || ((u64)src & 15)) return err_code(hipErrorInvalidValue);
    const u64 n = bytes / width;
    const u32 block = 256;
    u64 grid = (n + block - 1) / block;
    if (grid > 65536) grid = 65536;
    if (grid == 0) return 0;
    (void)hipEventRecord(g.kin_start, g.stream);
    if (width == 16)
        hipLaunchKernelGGL(simk_copy_kernel, dim3((u32)grid), dim3(block), 0, g.stream, (double2 *)dst,
                           (const double2 *)src, n);
    else
        hipLaunchKernelGGL(simk_copy8_kernel, dim3((u32)grid), dim3(block), 0, g.stream, (double *)dst,
                           (const double *)src, n);
    (void)hipEventRecord(g.kin_stop, g.stream);
    hipError_t e = hipEventSynchronize(g.kin_stop);
    if (e != hipSuccess) return err_code(e);
    f32 t = 0.0f;
    (void)hipEventElapsedTime(&t, g.kin_start, g.kin_stop);
    if (ms) *ms = t;
    return err_code(hipGetLastError());
}

GPU_API s32 simk_copy_device(void *dst, const void *src, u64 bytes, f32 *ms) { return copy_device(dst, src, bytes, ms, 16); }

GPU_API s32 simk_copy_device_b64(void *dst, const void *src, u64 bytes, f32 *ms) {
    return copy_device(dst, src, bytes, ms, 8);
}

// Generates, compiles (or loads from the cache) and loads the multi-mechanism
// kernel of g.multi_mechs under options o.
hipError_t load_multi(GenOptions o, JitKernel &out, u32 precision) {
    auto t0 = std::chrono::steady_clock::now();
    std::vector<const Mech *> ptrs;
    for (const Mech &m : g.multi_mechs) ptrs.push_back(&m);
    o.precision = precision;
    GenResult gr = generate_multi(ptrs, g.method, resolved_multi(o), kDeviceKinSource);
    std::vector<char> code;
    bool cached = false;
    if (!jit_code(gr, code, cached, g.jit_log)) return hipErrorInvalidImage;
    hipModule_t mod;
    hipError_t e = hipModuleLoadData(&mod, code.data());
    if (e != hipSuccess) return e;
    hipFunction_t fn;
    e = hipModuleGetFunction(&fn, mod, gr.kernel.c_str());
    if (e != hipSuccess) {
        (void)hipModuleUnload(mod);
        return e;
    }
    int blocks = 1;
    if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, fn, SIMK_WAVE, 0) != hipSuccess) blocks = 1;
    out.mod = mod;
    out.fn = fn;
    out.ws_slots = gr.ws_slots;
    out.mech_slots = gr.mech_slots;
    out.name = gr.kernel;
    out.cached = cached;
    out.blocks_per_cu = blocks;
    out.ms = std::chrono::duration<f64, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return hipSuccess;
}

GPU_API s32 simk_kin_set_mechs(u32 count, const void *const *mechs, const SimkCaps *caps, const KinMethod *method,
                               const KinControl *control) {
    if (!g.ready) return err_code(hipErrorNotInitialized);
    if (count == 0 || count > 4 || !mechs || !caps || !method || !control) return err_code(hipErrorInvalidValue);
    if (method->stages < 1 || method->stages > 4) return err_code(hipErrorInvalidValue);
    std::vector<Mech> ms(count);
    for (u32 i = 0; i < count; ++i)
        if (!mechs[i] || !from_bytes(mechs[i], caps[i], ms[i]).empty()) return err_code(hipErrorInvalidValue);
    hipError_t e = hipStreamSynchronize(g.stream);
    if (e != hipSuccess) return err_code(e);
    untune();
    destroy_graphs();   // they launch the kernel of the module unloaded here
    if (g.multi_jit.mod) (void)hipModuleUnload(g.multi_jit.mod);
    g.multi_jit = JitKernel();
    if (g.multi_jit_s32.mod) (void)hipModuleUnload(g.multi_jit_s32.mod);
    g.multi_jit_s32 = JitKernel();
    g.multi_ok = false;
    for (u32 i = 0; i < 4; ++i) {
        if (g.multi_dtables[i]) (void)hipFree(g.multi_dtables[i]);
        if (g.multi_dwave[i]) (void)hipFree(g.multi_dwave[i]);
        g.multi_dtables[i] = g.multi_dwave[i] = nullptr;
        if (g.multi_d32[i].mod) (void)hipModuleUnload(g.multi_d32[i].mod);
        g.multi_d32[i] = JitKernel();
        g.multi_d32_tried[i] = false;
    }
    g.multi_mechs = ms;
    g.method = *method;
    g.method.name = nullptr;
    fill_dev_method(g.method, g.dev_method);
    g.multi_ctl.assign(count, SimkDevCtl());
    for (u32 i = 0; i < count; ++i) {
        SimkDevCtl &c = g.multi_ctl[i];
        c.rtol = control->rtol;
        c.atol = control->atol;
        c.h_min = control->h_min;
        c.h_max = control->h_max;
        c.h_start = control->h_start;
        c.inv_n = 1.0 / (f64)ms[i].n;
        c.max_steps = control->max_steps;
        c.fixed_step = control->fixed_step ? 1 : 0;
        c.clamp_negative = control->clamp_negative ? 1 : 0;
    }
    for (u32 i = 0; i < count && e == hipSuccess; ++i) {
        e = upload_tables(ms[i], &g.multi_dtables[i]);
        if (e == hipSuccess) e = upload_wave_tables(ms[i], &g.multi_dwave[i]);
    }
    if (e != hipSuccess) return err_code(e);
    e = load_multi(g.gen, g.multi_jit);
    if (e != hipSuccess) return err_code(e);
    g.multi_ok = true;
    return 0;
}

GPU_API s32 simk_kin_set_mech_control(u32 index, const KinControl *control) {
    if (!g.multi_ok || !control || index >= g.multi_ctl.size()) return err_code(hipErrorInvalidValue);
    hipError_t e = hipStreamSynchronize(g.stream);
    if (e != hipSuccess) return err_code(e);
    destroy_graphs();   // they hold the previous controls in their launch arguments
    SimkDevCtl &c = g.multi_ctl[index];
    c.rtol = control->rtol;
    c.atol = control->atol;
    c.h_min = control->h_min;
    c.h_max = control->h_max;
    c.h_start = control->h_start;
    c.inv_n = 1.0 / (f64)g.multi_mechs[index].n;
    c.max_steps = control->max_steps;
    c.fixed_step = control->fixed_step ? 1 : 0;
    c.clamp_negative = control->clamp_negative ? 1 : 0;
    return 0;
}

GPU_API s32 simk_kin_advance_multi(u32 count, const GpuKinLaunch *launches, u32 repeats, f32 *kernel_ms) {
    MultiPlan P;
    hipError_t e = prepare_multi(count, launches, repeats, P);
    if (e != hipSuccess) return err_code(e);
    (void)hipEventRecord(g.kin_start, g.stream);
    e = enqueue_multi(P);
    if (e != hipSuccess) return err_code(e);
    (void)hipEventRecord(g.kin_stop, g.stream);
    e = hipEventSynchronize(g.kin_stop);
    if (e != hipSuccess) return err_code(e);
    f32 ms = 0.0f;
    (void)hipEventElapsedTime(&ms, g.kin_start, g.kin_stop);
    g.last_ms = ms;
    if (kernel_ms) *kernel_ms = ms;
    return err_code(hipGetLastError());
}

// Autotuner of the multi-mechanism kernel (ragged batches): the same
// protocol as simk_kin_autotune on a sample of every mechanism's cells, over
// the persistent grid and the generator's knobs, the kernel rebuilt (or taken
// from the cache) for each candidate. One lane per cell is the only lane
// layout of this kernel.
GPU_API s32 simk_kin_autotune_multi(u32 count, const GpuKinLaunch *launches, u32 sample_cells, f64 duration,
                                    SimkTuneResult *out) {
    if (!g.multi_ok || !launches || !out || count != g.multi_mechs.size()) return err_code(hipErrorInvalidValue);
    for (u32 i = 0; i < count; ++i)
        if (!launch_ok(launches + i) || launches[i].precision != 0) return err_code(hipErrorInvalidValue);
    memset(out, 0, sizeof(*out));
    hipError_t e = hipStreamSynchronize(g.stream);
    if (e != hipSuccess) return err_code(e);
    destroy_graphs();   // candidates replace the kernel the graphs launch
    std::vector<TuneSample> S(count);
    std::vector<GpuKinLaunch> work(count);
    u32 sampled = 0;
    for (u32 i = 0; i < count && e == hipSuccess; ++i) {
        const u32 lanes = launches[i].lane_count ? launches[i].lane_count : launches[i].padded;
        const u32 cells = sample_cells && sample_cells < lanes ? sample_cells : lanes;
        e = tune_sample(g.multi_mechs[i], launches + i, cells, S[i]);
        if (duration > 0.0) S[i].orig.duration = S[i].work.duration = duration;
        work[i] = S[i].work;
        sampled += cells;
    }
    auto time_it = [&](f32 &best) -> hipError_t {
        best = 1e30f;
        for (u32 r = 0; r < 2; ++r) {
            for (u32 i = 0; i < count; ++i) {
                hipError_t c = hipMemcpyAsync(S[i].mem + S[i].half, S[i].mem, S[i].half, hipMemcpyDeviceToDevice,
                                              g.stream);
                if (c != hipSuccess) return c;
            }
            MultiPlan P;
            hipError_t c = prepare_multi(count, work.data(), 1, P);
            if (c != hipSuccess) return c;
            (void)hipEventRecord(g.kin_start, g.stream);
            c = enqueue_multi(P);
            if (c != hipSuccess) return c;
            (void)hipEventRecord(g.kin_stop, g.stream);
            c = hipEventSynchronize(g.kin_stop);
            if (c != hipSuccess) return c;
            f32 ms = 0.0f;
            (void)hipEventElapsedTime(&ms, g.kin_start, g.kin_stop);
            if (ms < best) best = ms;
        }
        return hipSuccess;
    };
    const u32 saved_grid = g.grid_override;
    const GenOptions saved_gen = g.gen;
    const u32 cus = g.cus ? g.cus : 1;
    f32 best = 1e30f, t = 0.0f;
    u32 best_grid = 0;
    GenOptions best_gen = g.gen;
    auto consider = [&](u32 grid, const GenOptions &opt) -> hipError_t {
        g.grid_override = grid;
        if (!same_gen(opt, g.gen)) {
            JitKernel k;
            if (load_multi(opt, k) != hipSuccess) return hipSuccess;   // variant does not build: skip it
            if (g.multi_jit.mod) (void)hipModuleUnload(g.multi_jit.mod);
            g.multi_jit = k;
            g.gen = opt;
        }
        hipError_t r = time_it(t);
        if (r != hipSuccess) return r;
        out->candidates++;
        if (grid == 0 && same_gen(opt, saved_gen)) out->default_ms = t;
        if (t < best * kTuneMargin) {
            best = t;
            best_grid = grid;
            best_gen = opt;
        }
        return hipSuccess;
    };
    const u32 grids[] = {0, cus, cus * 3 / 2, cus * 2, cus * 3, cus * 4};
    for (u32 i = 0; i < sizeof(grids) / sizeof(grids[0]) && e == hipSuccess; ++i) e = consider(grids[i], g.gen);
    auto knob = [&](auto set) {
        if (e != hipSuccess) return;
        GenOptions o = best_gen;
        set(o);
        if (!same_gen(o, best_gen)) e = consider(best_grid, o);
    };
    knob([](GenOptions &o) { o.batch = 8; });
    knob([](GenOptions &o) { o.batch = 16; });
    knob([](GenOptions &o) { o.reg_window = 16; });
    knob([](GenOptions &o) { o.reg_window = 32; });
    knob([](GenOptions &o) { o.unroll_stages = 1; });
    knob([](GenOptions &o) { o.lds_slots = 0; });
    knob([](GenOptions &o) { o.lds_slots = 80; });
    knob([](GenOptions &o) { o.lds_slots = 128; });
    knob([](GenOptions &o) { o.pair_slots = resolved(o).pair_slots ? 0u : 1u; });
    knob([](GenOptions &o) { o.k_pairs = resolved_multi(o).k_pairs ? 0u : 1u; });
    knob([](GenOptions &o) { o.stage_carry = resolved_multi(o).stage_carry ? 0u : 1u; });
    // teams of 64, 32, 16 and 8 lanes per cell (the cooperative form)
    const u32 saved_multi_team = g.multi_team;
    u32 best_team = 0;
    {
        const GenOptions keep = g.gen;
        // timed on 1/16 of every mechanism's sample (at least 64 cells) and
        // scaled: a team takes 15-80x longer per cell than a lane
        const std::vector<GpuKinLaunch> full = work;
        u64 all = 0, part = 0;
        for (GpuKinLaunch &w : work) {
            const u32 l = w.lane_count ? w.lane_count : w.padded;
            u32 sub = l / 16;
            if (sub < 64) sub = l < 64 ? l : 64;
            w.lane_count = sub;
            all += l;
            part += sub;
        }
        const u32 team_sizes[] = {SIMK_WAVE, 32, 16, 8};
        for (u32 tl : team_sizes) {
            if (e != hipSuccess) break;
            g.multi_team = tl;
            hipError_t r = time_it(t);
            if (r == hipErrorInvalidValue) continue;   // the team's cells do not fit the LDS
            e = r;
            if (e != hipSuccess) break;
            t = t * (f32)((f64)all / (f64)(part ? part : 1));
            out->candidates++;
            if (t < best * kTuneMargin) {
                best = t;
                best_team = tl;
            }
        }
        g.multi_team = saved_multi_team;
        g.gen = keep;
        work = full;
    }
    for (TuneSample &x : S)
        if (x.mem) (void)hipFree(x.mem);
    if (e == hipSuccess && !same_gen(best_gen, g.gen)) {
        JitKernel k;
        e = load_multi(best_gen, k);
        if (e == hipSuccess) {
            if (g.multi_jit.mod) (void)hipModuleUnload(g.multi_jit.mod);
            g.multi_jit = k;
            g.gen = best_gen;
        }
    }
    if (e != hipSuccess) {
        // leave a consistent kernel: the caller's own options
        g.grid_override = saved_grid;
        if (!same_gen(g.gen, saved_gen)) {
            JitKernel k;
            if (load_multi(saved_gen, k) == hipSuccess) {
                if (g.multi_jit.mod) (void)hipModuleUnload(g.multi_jit.mod);
                g.multi_jit = k;
            }
            g.gen = saved_gen;
        }
        return err_code(e);
    }
    if (!g.tuned) {
        g.tuned = true;
        g.user_gen = saved_gen;
        g.user_mode = g.mode;
        g.user_team = g.team_lanes;
        g.user_multi_team = saved_multi_team;
        g.user_grid = saved_grid;
    }
    g.grid_override = best_grid;
    g.multi_team = best_team;
    out->mode = best_team ? 1 : 0;
    out->lanes_per_cell = best_team ? best_team : 1;
    out->waves = best_grid;
    out->batch = resolved_multi(best_gen).batch;
    out->reg_window = resolved_multi(best_gen).reg_window;
    out->unroll_stages = best_gen.unroll_stages;
    out->reuse = resolved_multi(best_gen).reuse;
    u32 lds = 0;
    for (const Mech &m : g.multi_mechs) {
        const u32 l = lds_entries(m, resolved_multi(best_gen));
        if (l > lds) lds = l;
    }
    out->lds_slots = lds;
    out->best_ms = best;
    out->sample_cells = sampled;
    return 0;
}

GPU_API s32 simk_graph_capture_multi(u32 count, const GpuKinLaunch *launches, u32 repeats, u32 steps,
                                     u32 *graph_id) {
    if (!graph_id || steps == 0) return err_code(hipErrorInvalidValue);
    MultiPlan P;
    hipError_t e = prepare_multi(count, launches, repeats, P, true);   // every allocation happens here
    if (e != hipSuccess) return err_code(e);
    State::Graph rec;
    rec.ws = P.ws;
    rec.queue = P.queue;
    e = hipStreamSynchronize(g.stream);
    if (e == hipSuccess) e = hipStreamBeginCapture(g.stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        free_graph(rec);
        return err_code(e);
    }
    hipError_t le = hipSuccess;
    for (u32 s = 0; s < steps && le == hipSuccess; ++s) le = enqueue_multi(P);
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(g.stream, &graph);
    if (le == hipSuccess && e == hipSuccess) e = hipGraphInstantiate(&rec.exec, graph, nullptr, nullptr, 0);
    if (graph) (void)hipGraphDestroy(graph);
    if (le != hipSuccess || e != hipSuccess) {
        free_graph(rec);
        return err_code(le != hipSuccess ? le : e);
    }
    g.graphs.push_back(rec);
    *graph_id = (u32)(g.graphs.size() - 1);
    return 0;
}

GPU_API s32 simk_graph_replay(u32 graph_id, f32 *ms) {
    if (graph_id >= g.graphs.size() || !g.graphs[graph_id].exec) return err_code(hipErrorInvalidValue);
    (void)hipEventRecord(g.kin_start, g.stream);
    hipError_t e = hipGraphLaunch(g.graphs[graph_id].exec, g.stream);
    if (e != hipSuccess) return err_code(e);
    (void)hipEventRecord(g.kin_stop, g.stream);
    e = hipEventSynchronize(g.kin_stop);
    if (e != hipSuccess) return err_code(e);
    f32 t = 0.0f;
    (void)hipEventElapsedTime(&t, g.kin_start, g.kin_stop);
    if (ms) *ms = t;
    return err_code(hipGetLastError());
}

GPU_API s32 simk_graph_destroy(u32 graph_id) {
    if (graph_id >= g.graphs.size() || !g.graphs[graph_id].exec) return err_code(hipErrorInvalidValue);
    hipError_t e = hipStreamSynchronize(g.stream);
    free_graph(g.graphs[graph_id]);
    return err_code(e);
}

}  // extern "C"
