// Serving loop: device-resident solve requests with HBM-budget admission and
// continuous batching over the multi-mechanism persistent kernel.
//
// Memory: one slab per mechanism, a zone block of `cap` lanes in the
// reference's SoA row order (y, h, steps, rejects, pulses, failed, k, src,
// pulse, period, pulse_first, pulse_count), one courant row, the resume rows
// of time slices and the DBD rows of reactor requests. A request owns a
// contiguous lane range of its mechanism's slab from admission to completion,
// so its state never leaves HBM between steps. A per-slab order
// array lists the lanes of the admitted requests; the persistent kernel takes
// cells from it, so every request of every mechanism advances in one launch
// and a request joins or leaves at any iteration boundary.
//
// Two kinds of request share the launches. A transport request runs upwind
// advection along its channels before each chemistry step. A DBD reactor
// request runs the chained reactor's master step (dbd_scene_step,
// src/sim/dbd_scene.cpp:226-298): the discharge source and pulse schedule of
// (t, t + dt] prepared on the device from each zone's composition, the
// chemistry step, the pulses committed, then plug flow along the chain.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <deque>
#include <map>
#include <unordered_map>
#include <vector>

#include "air_source.h"
#include "module_internal.h"
#include "simk_dbd.h"
#include "simk_gpu_api.h"

namespace {

constexpr u32 kMaxSpecies = 64;
constexpr u32 kMaxNodes = 8192;
constexpr u32 kResumeRows = 5;
// DBD requests' per-zone rows: T, power, frequency, pulses fired so far, then
// the discharge channels' rate coefficients at T (computed on the host).
constexpr u32 kDbdT = 0, kDbdPower = 1, kDbdFreq = 2, kDbdPulseIndex = 3, kDbdRates = 4;
constexpr u32 kDbdRows = kDbdRates + simk_air::kQuenchRates;
constexpr u32 kDbdMaxSpecies = 32;

struct Slab {
    double *d = nullptr;     // zone block rows, courant row, resume rows, DBD rows; row stride cap
    u32 cap = 0;
    u32 n = 0, R = 0;
    u64 zone_rows = 0;       // 3n + R + 8
    std::map<u32, u32> free_ranges;   // offset -> length
    std::vector<u32> owner;  // lane -> ticket of the request holding it (0: none)
    u32 *d_order = nullptr;  // queue position -> lane
    u32 *h_order = nullptr;  // pinned staging
    u32 lanes = 0;           // current order length
    u32 *d_order32 = nullptr, *h_order32 = nullptr;   // the lanes of f32 delta requests
    u32 lanes32 = 0;
    u32 *d_carry = nullptr;  // cells stopped by the slice end
    u32 *h_carry = nullptr;
    double *courant() const { return d + zone_rows * cap; }
    double *resume() const { return d + (zone_rows + 1) * cap; }
    double *dbd() const { return d + (zone_rows + 1 + kResumeRows) * cap; }
};

enum { kAdvect = 0, kChem = 1 };

struct Request {
    SimkServeRequest r;
    u32 ticket = 0;
    u32 off = 0;
    u32 steps_done = 0;
    u32 steps = 0;               // operator-split steps of this request
    u32 phase = kAdvect;
    u64 bytes = 0;
    u32 slot = 0;                // inflow slot
    std::vector<u32> remaining;  // lanes whose chemistry step is pending: stopped cells first
    // DBD requests
    f64 t = 0.0;                 // the request's clock (accumulated step by step, as the scene's)
    u32 mode = 0;
    f64 residence = 0.0;
    u32 species[17] = {};
    f64 g[4] = {}, f_O1D = 0.0, filament = 0.0;
    std::vector<f64> dbd_rows;   // [kDbdRows][cells], uploaded at admission
};

// A DBD request's step-preparation work (dbd_zones_prepare_source, dbd.cpp:123-154).
struct DbdPrepSeg {
    f64 *y, *src, *pulse, *period, *first, *count;   // the request's zone-block rows in its slab
    const f64 *dbd;                                  // its DBD rows
    u64 stride;                                      // lanes between rows
    u32 n, zones, mode, pad_;
    u32 species[17], pad2_;
    f64 t0, t1, g[4], f_O1D, filament;
};

// Each zone of each segment: the discharge source (pulse increments or a
// steady rate) from its composition, and the pulses due in (t0, t1] on its
// absolute clock (bursts numbered from 1; pulse_index of them fired so far).
__global__ void __launch_bounds__(256) simk_dbd_prep_kernel(const DbdPrepSeg *segs) {
    const DbdPrepSeg &S = segs[blockIdx.y];
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S.zones) return;
    const u64 p = S.stride;
    f64 gas[kDbdMaxSpecies], change[kDbdMaxSpecies], rates[simk_air::kQuenchRates];
    for (u32 s = 0; s < S.n; ++s) gas[s] = S.y[s * p + i];
    for (u32 c = 0; c < simk_air::kQuenchRates; ++c) rates[c] = S.dbd[(kDbdRates + c) * p + i];
    const f64 power = S.dbd[kDbdPower * p + i], freq = S.dbd[kDbdFreq * p + i];
    const bool on = power > 0.0 && freq > 0.0;
    const f64 bursts_per_s = (on && S.filament > 0.0) ? power / S.filament : 2.0 * freq;
    const bool pulsed = S.mode == SimkDbd_Pulsed && on;
    const f64 energy = pulsed ? power / bursts_per_s : (on ? power : 0.0);
    simk_air::air_source_delta_rates(S.species, S.g, S.f_O1D, gas, S.n, energy, rates, change);
    f64 *to = pulsed ? S.pulse : S.src, *zeroed = pulsed ? S.src : S.pulse;
    for (u32 s = 0; s < S.n; ++s) {
        to[s * p + i] = change[s];
        zeroed[s * p + i] = 0.0;
    }
    if (!pulsed) {
        S.period[i] = S.first[i] = S.count[i] = 0.0;
        return;
    }
    const f64 period = 1.0 / bursts_per_s;
    const f64 first_burst = S.dbd[kDbdPulseIndex * p + i] + 1.0;
    f64 due = 0.0;
    for (;;) {
        if (!((first_burst + due) * period <= S.t1)) break;
        due += 1.0;
    }
    S.period[i] = period;
    S.first[i] = first_burst * period - S.t0;
    S.count[i] = due;
}

// After a DBD step's chemistry: the scheduled pulses are fired
// (dbd_zones_commit_pulses, dbd.cpp:156-158).
__global__ void __launch_bounds__(256) simk_dbd_commit_kernel(f64 *pulse_index, const f64 *count, u32 zones) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < zones) pulse_index[i] = pulse_index[i] + count[i];
}

struct Server {
    bool open = false;
    SimkServeConfig cfg;
    SimkMultiShape shape;
    std::vector<Slab> slabs;
    std::unordered_map<u32, Request> reqs;
    std::deque<u32> queued;          // tickets in arrival order
    std::vector<u32> active;         // admitted tickets
    std::vector<u32> free_slots;
    std::vector<u32> pending_done;   // finished, not yet reported
    u32 next_ticket = 1;
    double *d_ws = nullptr;
    u32 *d_queue = nullptr, *h_queue = nullptr;
    double *d_inflow = nullptr;      // [max_requests][kMaxSpecies]
    SimkUpwindSeg *d_segs = nullptr, *h_segs = nullptr;
    DbdPrepSeg *d_prep = nullptr, *h_prep = nullptr;   // [max_requests]
    hipEvent_t ev[5] = {};
    // Admission uploads run on a stream of their own (default priority), so
    // a request admitted while a chemistry launch runs is copied in during it;
    // the module stream waits for `admitted` before it touches those lanes.
    hipStream_t copy_stream = nullptr;
    hipEvent_t admitted = nullptr;
    SimkServeStats st;
};

Server S;

s32 code(hipError_t e) { return (s32)e; }

u64 zone_rows(u32 n, u32 R) { return 3ull * n + R + 8; }

// Row offsets (in lanes of the slab) of the zone block, reference order.
void zone_offsets(u32 n, u32 R, u64 p, u64 off[12]) {
    const u64 rows[12] = {n * p, p, p, p, p, p, R * p, n * p, n * p, p, p, p};
    u64 o = 0;
    for (int i = 0; i < 12; ++i) {
        off[i] = o;
        o += rows[i];
    }
}

bool take_range(Slab &sl, u32 len, u32 &off) {
    for (auto it = sl.free_ranges.begin(); it != sl.free_ranges.end(); ++it) {
        if (it->second < len) continue;
        off = it->first;
        const u32 rest = it->second - len;
        sl.free_ranges.erase(it);
        if (rest) sl.free_ranges[off + len] = rest;
        return true;
    }
    return false;
}

void give_range(Slab &sl, u32 off, u32 len) {
    auto it = sl.free_ranges.emplace(off, len).first;
    auto next = std::next(it);
    if (next != sl.free_ranges.end() && it->first + it->second == next->first) {
        it->second += next->second;
        sl.free_ranges.erase(next);
    }
    if (it != sl.free_ranges.begin()) {
        auto prev = std::prev(it);
        if (prev->first + prev->second == it->first) {
            prev->second += it->second;
            sl.free_ranges.erase(it);
        }
    }
}

void release_all() {
    for (Slab &sl : S.slabs) {
        if (sl.d) (void)hipFree(sl.d);
        if (sl.d_order) (void)hipFree(sl.d_order);
        if (sl.h_order) (void)hipHostFree(sl.h_order);
        if (sl.d_order32) (void)hipFree(sl.d_order32);
        if (sl.h_order32) (void)hipHostFree(sl.h_order32);
        if (sl.d_carry) (void)hipFree(sl.d_carry);
        if (sl.h_carry) (void)hipHostFree(sl.h_carry);
    }
    S.slabs.clear();
    if (S.d_ws) (void)hipFree(S.d_ws);
    if (S.d_queue) (void)hipFree(S.d_queue);
    if (S.h_queue) (void)hipHostFree(S.h_queue);
    if (S.d_inflow) (void)hipFree(S.d_inflow);
    if (S.d_segs) (void)hipFree(S.d_segs);
    if (S.h_segs) (void)hipHostFree(S.h_segs);
    if (S.d_prep) (void)hipFree(S.d_prep);
    if (S.h_prep) (void)hipHostFree(S.h_prep);
    for (hipEvent_t &e : S.ev)
        if (e) (void)hipEventDestroy(e);
    if (S.admitted) (void)hipEventDestroy(S.admitted);
    if (S.copy_stream) {
        (void)hipStreamSynchronize(S.copy_stream);
        (void)hipStreamDestroy(S.copy_stream);
    }
    S = Server();
}

bool same_shape(const SimkMultiShape &a, const SimkMultiShape &b) {
    if (a.count != b.count || a.slots != b.slots) return false;
    for (u32 i = 0; i < a.count; ++i)
        if (a.n[i] != b.n[i] || a.R[i] != b.R[i]) return false;
    return true;
}

// Admits queued requests in arrival order while the budget and a lane range
// allow; enqueues their uploads.
hipError_t admit(hipStream_t st) {
    while (!S.queued.empty()) {
        Request &q = S.reqs[S.queued.front()];
        if (S.st.admitted_bytes + q.bytes > S.cfg.hbm_budget) break;
        Slab &sl = S.slabs[q.r.mech];
        u32 off = 0;
        if (!take_range(sl, q.r.cells, off)) break;
        q.off = off;
        const u64 cap = sl.cap;
        hipError_t e = hipMemcpy2DAsync(sl.d + off, cap * 8, q.r.block, (u64)q.r.padded * 8, (u64)q.r.cells * 8,
                                        sl.zone_rows, hipMemcpyHostToDevice, st);
        if (e == hipSuccess && q.r.kind == SimkServe_Transport)
            e = hipMemcpyAsync(sl.courant() + off, q.r.courant, (u64)q.r.cells * 8, hipMemcpyHostToDevice, st);
        if (e == hipSuccess && q.r.inflow)
            e = hipMemcpyAsync(S.d_inflow + (u64)q.slot * kMaxSpecies, q.r.inflow, (u64)sl.n * 8,
                               hipMemcpyHostToDevice, st);
        if (e == hipSuccess && q.r.kind == SimkServe_Dbd)
            e = hipMemcpy2DAsync(sl.dbd() + off, cap * 8, q.dbd_rows.data(), (u64)q.r.cells * 8, (u64)q.r.cells * 8,
                                 kDbdRows, hipMemcpyHostToDevice, st);
        if (e != hipSuccess) return e;
        for (u32 c = 0; c < q.r.cells; ++c) sl.owner[off + c] = q.ticket;
        q.phase = kAdvect;
        S.st.admitted_bytes += q.bytes;
        if (S.st.admitted_bytes > S.st.peak_admitted_bytes) S.st.peak_admitted_bytes = S.st.admitted_bytes;
        S.st.admitted++;
        S.active.push_back(q.ticket);
        S.queued.pop_front();
    }
    return hipSuccess;
}

// Advects every request whose next step starts now; they enter the chemistry
// phase with all their lanes pending.
hipError_t advect(hipStream_t st) {
    u32 rows = 0, max_nodes = 0, count = 0, preps = 0, max_zones = 0;
    for (u32 t : S.active) {
        Request &q = S.reqs[t];
        if (q.phase != kAdvect) continue;
        Slab &sl = S.slabs[q.r.mech];
        if (q.r.kind == SimkServe_Dbd) {
            // the step's source and pulse schedule over (t, t + dt]
            DbdPrepSeg &g = S.h_prep[preps++];
            u64 off[12];
            zone_offsets(sl.n, sl.R, sl.cap, off);
            g.y = sl.d + off[0] + q.off;
            g.src = sl.d + off[7] + q.off;
            g.pulse = sl.d + off[8] + q.off;
            g.period = sl.d + off[9] + q.off;
            g.first = sl.d + off[10] + q.off;
            g.count = sl.d + off[11] + q.off;
            g.dbd = sl.dbd() + q.off;
            g.stride = sl.cap;
            g.n = sl.n;
            g.zones = q.r.cells;
            g.mode = q.mode;
            memcpy(g.species, q.species, sizeof(g.species));
            g.t0 = q.t;
            g.t1 = q.t + S.cfg.dt;
            memcpy(g.g, q.g, sizeof(g.g));
            g.f_O1D = q.f_O1D;
            g.filament = q.filament;
            if (q.r.cells > max_zones) max_zones = q.r.cells;
            q.phase = kChem;
            q.remaining.resize(q.r.cells);
            for (u32 c = 0; c < q.r.cells; ++c) q.remaining[c] = q.off + c;
            continue;
        }
        SimkUpwindSeg &g = S.h_segs[count++];
        g.x = sl.d + q.off;
        g.padded = sl.cap;
        g.courant = sl.courant() + q.off;
        g.inflow = q.r.inflow ? S.d_inflow + (u64)q.slot * kMaxSpecies : nullptr;
        g.fields = sl.n;
        g.channels = q.r.channels;
        g.nodes = q.r.nodes;
        g.row_begin = rows;
        rows += sl.n * q.r.channels;
        if (q.r.nodes > max_nodes) max_nodes = q.r.nodes;
        q.phase = kChem;
        q.remaining.resize(q.r.cells);
        for (u32 c = 0; c < q.r.cells; ++c) q.remaining[c] = q.off + c;
    }
    hipError_t e = hipSuccess;
    if (preps) {
        e = hipMemcpyAsync(S.d_prep, S.h_prep, preps * sizeof(DbdPrepSeg), hipMemcpyHostToDevice, st);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(simk_dbd_prep_kernel, dim3((max_zones + 255) / 256, preps), dim3(256), 0, st,
                               (const DbdPrepSeg *)S.d_prep);
            e = hipGetLastError();
        }
    }
    if (!count || e != hipSuccess) return e;
    e = hipMemcpyAsync(S.d_segs, S.h_segs, count * sizeof(SimkUpwindSeg), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = simk_enqueue_upwind_batch(st, S.d_segs, count, rows, max_nodes, S.cfg.substeps);
    return e;
}

// Queue order of every slab: requests with chemistry pending, smallest first
// (then oldest), each with its stopped cells before its untouched ones.
hipError_t build_orders(hipStream_t st, u64 &waves) {
    std::vector<const Request *> order;
    for (u32 t : S.active) {
        const Request &q = S.reqs[t];
        if (q.phase == kChem) order.push_back(&q);
    }
    std::sort(order.begin(), order.end(), [](const Request *a, const Request *b) {
        return a->r.cells != b->r.cells ? a->r.cells < b->r.cells : a->ticket < b->ticket;
    });
    for (Slab &sl : S.slabs) sl.lanes = sl.lanes32 = 0;
    for (const Request *q : order) {
        Slab &sl = S.slabs[q->r.mech];
        // f32 delta requests run in a launch of their own (whole steps, no slices)
        u32 *to = q->r.precision == 1 ? sl.h_order32 + sl.lanes32 : sl.h_order + sl.lanes;
        memcpy(to, q->remaining.data(), q->remaining.size() * 4);
        (q->r.precision == 1 ? sl.lanes32 : sl.lanes) += (u32)q->remaining.size();
    }
    waves = 0;
    hipError_t e = hipSuccess;
    for (Slab &sl : S.slabs) {
        waves += (sl.lanes + 63) / 64;
        if (e == hipSuccess && sl.lanes)
            e = hipMemcpyAsync(sl.d_order, sl.h_order, (u64)sl.lanes * 4, hipMemcpyHostToDevice, st);
        if (e == hipSuccess && sl.lanes32)
            e = hipMemcpyAsync(sl.d_order32, sl.h_order32, (u64)sl.lanes32 * 4, hipMemcpyHostToDevice, st);
    }
    return e;
}

// After a launch: the cells each request still has pending are the ones the
// slice stopped (carry lists) and the queue positions no lane took.
void collect(std::vector<u32> &step_done) {
    for (u32 t : S.active) {
        Request &q = S.reqs[t];
        if (q.phase == kChem) q.remaining.clear();
    }
    for (u32 i = 0; i < S.shape.count; ++i) {
        Slab &sl = S.slabs[i];
        const u32 head = S.h_queue[i] < sl.lanes ? S.h_queue[i] : sl.lanes;
        const u32 carried = S.h_queue[4 + i];
        S.st.carried += carried;
        for (u32 j = 0; j < carried; ++j) {
            const u32 cell = sl.h_carry[j];
            S.reqs[sl.owner[cell]].remaining.push_back(cell);
        }
        for (u32 pos = head; pos < sl.lanes; ++pos) {
            const u32 cell = sl.h_order[pos];
            S.reqs[sl.owner[cell]].remaining.push_back(cell);
        }
    }
    for (u32 t : S.active) {
        Request &q = S.reqs[t];
        if (q.phase == kChem && q.remaining.empty()) step_done.push_back(t);
    }
}

}  // namespace

extern "C" {

GPU_API s32 simk_serve_open(const SimkServeConfig *config) {
    if (S.open) {
        (void)hipStreamSynchronize(simk_module_stream());
        release_all();
    }
    const SimkMultiShape shape = simk_multi_shape();
    if (shape.count == 0 || !simk_module_stream()) return code(hipErrorNotInitialized);
    SimkServeConfig c;
    memset(&c, 0, sizeof(c));
    if (config) c = *config;
    if (!c.hbm_budget) c.hbm_budget = 8ull << 30;
    if (!c.max_requests) c.max_requests = 4096;
    if (!c.steps) c.steps = 20;
    if (!c.substeps) c.substeps = 4;
    if (!c.slice_us) c.slice_us = 20000;
    if (!(c.dt > 0.0)) c.dt = 1e-3;
    for (u32 i = 0; i < shape.count; ++i)
        if (shape.n[i] > kMaxSpecies) return code(hipErrorInvalidValue);
    S.cfg = c;
    S.shape = shape;
    memset(&S.st, 0, sizeof(S.st));
    hipStream_t st = simk_module_stream();
    const u64 carry_cap = (u64)shape.resident_waves * 64;
    hipError_t e = hipSuccess;
    S.slabs.resize(shape.count);
    for (u32 i = 0; i < shape.count && e == hipSuccess; ++i) {
        Slab &sl = S.slabs[i];
        sl.n = shape.n[i];
        sl.R = shape.R[i];
        sl.zone_rows = zone_rows(sl.n, sl.R);
        const u64 rows = sl.zone_rows + 1 + kResumeRows + kDbdRows;
        const u64 widest = sl.R > sl.n ? sl.R : sl.n;   // row group whose byte offsets must stay below 2^31
        // Rounded up, so any request within the budget fits an empty slab.
        u64 cap = (c.hbm_budget / (rows * 8) + 63) & ~63ull;
        const u64 lim = (((1ull << 31) - 1) / (widest * 8)) & ~63ull;
        if (cap > lim) cap = lim;
        if (cap < 64) {
            e = hipErrorInvalidValue;
            break;
        }
        sl.cap = (u32)cap;
        sl.free_ranges[0] = sl.cap;
        sl.owner.assign(cap, 0);
        e = hipMalloc(&sl.d, rows * cap * 8);
        // Resume flags start clear; a lane's flag is set only while its cell
        // is stopped mid-step, and cleared when the cell continues.
        if (e == hipSuccess) e = hipMemsetAsync(sl.resume(), 0, kResumeRows * cap * 8, st);
        if (e == hipSuccess) e = hipMalloc(&sl.d_order, cap * 4);
        if (e == hipSuccess) e = hipHostMalloc(&sl.h_order, cap * 4, 0);
        if (e == hipSuccess) e = hipMalloc(&sl.d_order32, cap * 4);
        if (e == hipSuccess) e = hipHostMalloc(&sl.h_order32, cap * 4, 0);
        if (e == hipSuccess) e = hipMalloc(&sl.d_carry, carry_cap * 4);
        if (e == hipSuccess) e = hipHostMalloc(&sl.h_carry, carry_cap * 4, 0);
    }
    if (e == hipSuccess) e = hipMalloc(&S.d_ws, (u64)shape.resident_waves * shape.slots * 64 * 8);
    if (e == hipSuccess) e = hipMalloc(&S.d_queue, 256);
    if (e == hipSuccess) e = hipHostMalloc(&S.h_queue, 256, 0);
    if (e == hipSuccess) e = hipMalloc(&S.d_inflow, (u64)c.max_requests * kMaxSpecies * 8);
    if (e == hipSuccess) e = hipMalloc(&S.d_segs, (u64)c.max_requests * sizeof(SimkUpwindSeg));
    if (e == hipSuccess) e = hipHostMalloc(&S.h_segs, (u64)c.max_requests * sizeof(SimkUpwindSeg), 0);
    if (e == hipSuccess) e = hipMalloc(&S.d_prep, (u64)c.max_requests * sizeof(DbdPrepSeg));
    if (e == hipSuccess) e = hipHostMalloc(&S.h_prep, (u64)c.max_requests * sizeof(DbdPrepSeg), 0);
    for (hipEvent_t &ev : S.ev)
        if (e == hipSuccess) e = hipEventCreate(&ev);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&S.admitted, hipEventDisableTiming);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&S.copy_stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        release_all();
        return code(e);
    }
    S.free_slots.reserve(c.max_requests);
    for (u32 i = c.max_requests; i-- > 0;) S.free_slots.push_back(i);
    S.open = true;
    return 0;
}

GPU_API s32 simk_serve_submit(const SimkServeRequest *request, u32 *ticket) {
    if (!S.open) return code(hipErrorNotInitialized);
    if (!request || !ticket) return code(hipErrorInvalidValue);
    const SimkServeRequest &r = *request;
    if (r.mech >= S.shape.count || r.cells == 0 || r.padded < r.cells || !r.block || !r.result ||
        r.kind > SimkServe_Dbd || r.precision > 1)
        return code(hipErrorInvalidValue);
    if (r.kind == SimkServe_Transport && (r.channels == 0 || r.nodes == 0 || r.nodes > kMaxNodes ||
                                          (u64)r.channels * r.nodes != r.cells || !r.courant))
        return code(hipErrorInvalidValue);
    const Slab &sl = S.slabs[r.mech];
    Request q;
    if (r.kind == SimkServe_Dbd) {
        const SimkServeDbd *d = r.dbd;
        if (!d || d->source.boltzmann || sl.n > kDbdMaxSpecies || !d->T_K || !d->power_W_cm3 || !d->freq_Hz ||
            !d->pulse_index || d->mode > SimkDbd_Averaged)
            return code(hipErrorInvalidValue);
        memcpy(q.species, &d->species, sizeof(q.species));
        for (u32 k = 0; k < 17; ++k)
            if (q.species[k] >= sl.n) return code(hipErrorInvalidValue);
        if (d->residence_s > 0.0) {
            // plug flow toward the feed: a = dt * zones / residence <= 1 (dbd.cpp:473)
            if (!r.inflow || r.cells > kMaxNodes || !(S.cfg.dt * (f64)r.cells / d->residence_s <= 1.0))
                return code(hipErrorInvalidValue);
        }
        q.mode = d->mode;
        q.residence = d->residence_s > 0.0 ? d->residence_s : 0.0;
        q.g[0] = d->source.g_O2_diss;
        q.g[1] = d->source.g_N2A;
        q.g[2] = d->source.g_N2_diss;
        q.g[3] = d->source.g_OH_per_100eV;
        q.f_O1D = d->source.f_O1D;
        q.filament = d->source.filament_J_cm3;
        q.dbd_rows.assign((u64)kDbdRows * r.cells, 0.0);
        f64 *rows = q.dbd_rows.data();
        for (u32 i = 0; i < r.cells; ++i) {
            rows[kDbdT * r.cells + i] = d->T_K[i];
            rows[kDbdPower * r.cells + i] = d->power_W_cm3[i];
            rows[kDbdFreq * r.cells + i] = d->freq_Hz[i];
            rows[kDbdPulseIndex * r.cells + i] = d->pulse_index[i];
            f64 rates[simk_air::kQuenchRates];
            simk_air::air_quench_rates(d->T_K[i], rates);
            for (u32 c = 0; c < simk_air::kQuenchRates; ++c) rows[(kDbdRates + c) * r.cells + i] = rates[c];
        }
    }
    const u64 bytes = (u64)r.cells * (sl.zone_rows + 1 + kResumeRows + kDbdRows) * 8;
    if (bytes > S.cfg.hbm_budget || r.cells > sl.cap) return code(hipErrorInvalidValue);
    if (S.free_slots.empty()) return code(hipErrorOutOfMemory);
    q.r = r;
    q.r.dbd = nullptr;   // the caller's description is staged; it need not outlive this call
    q.steps = r.steps ? r.steps : S.cfg.steps;
    q.ticket = S.next_ticket++;
    if (S.next_ticket == 0) S.next_ticket = 1;
    q.bytes = bytes;
    q.slot = S.free_slots.back();
    S.free_slots.pop_back();
    S.reqs[q.ticket] = q;
    S.queued.push_back(q.ticket);
    S.st.submitted++;
    *ticket = q.ticket;
    return 0;
}

GPU_API s32 simk_serve_step(u32 *done, u32 max_done, u32 *n_done) {
    if (!S.open) return code(hipErrorNotInitialized);
    if (n_done) *n_done = 0;
    if (!same_shape(S.shape, simk_multi_shape())) return code(hipErrorInvalidValue);   // mechanisms changed
    hipStream_t st = simk_module_stream();
    hipError_t e = hipEventRecord(S.ev[0], st);
    // requests submitted since the last launch: their uploads, then the
    // module stream waits for every upload enqueued so far
    if (e == hipSuccess) e = admit(S.copy_stream);
    if (e == hipSuccess) e = hipEventRecord(S.admitted, S.copy_stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(st, S.admitted, 0);
    if (e == hipSuccess) e = hipEventRecord(S.ev[1], st);
    if (e == hipSuccess) e = advect(st);
    if (e == hipSuccess) e = hipEventRecord(S.ev[2], st);
    if (e != hipSuccess) return code(e);
    u64 waves = 0;
    e = build_orders(st, waves);
    if (e != hipSuccess) return code(e);
    bool any32 = false;
    for (const Slab &sl : S.slabs) any32 = any32 || sl.lanes32 > 0;
    const bool ran = waves > 0 || any32;
    std::vector<u32> step_done, finished;
    if (ran) {
        GpuKinLaunch Ls[4];
        u32 lanes[4];
        const u32 *orders[4];
        SimkSlice slice;
        memset(&slice, 0, sizeof(slice));
        // 100 MHz counter. With nobody waiting for admission the slice is 4x
        // longer: fewer launch boundaries (each costs a host round trip and
        // the launch's ramp) while no arrival can use the lanes sooner.
        const u64 slice_us = S.queued.empty() ? 4ull * S.cfg.slice_us : (u64)S.cfg.slice_us;
        slice.ticks = S.cfg.slice_us == 0xFFFFFFFFu ? 0 : slice_us * 100;
        for (u32 i = 0; i < S.shape.count; ++i) {
            Slab &sl = S.slabs[i];
            u64 off[12];
            zone_offsets(sl.n, sl.R, sl.cap, off);
            f64 *rows[12];
            for (int j = 0; j < 12; ++j) rows[j] = sl.d + off[j];
            GpuKinLaunch &L = Ls[i];
            memset(&L, 0, sizeof(L));
            L.y = rows[0]; L.h = rows[1]; L.steps = rows[2]; L.rejects = rows[3]; L.pulses = rows[4];
            L.failed = rows[5]; L.k = rows[6]; L.src = rows[7]; L.pulse = rows[8]; L.period = rows[9];
            L.pulse_first = rows[10]; L.pulse_count = rows[11];
            L.padded = sl.cap;
            L.lane_count = sl.cap;
            L.precision = 0;
            L.duration = S.cfg.dt;
            lanes[i] = sl.lanes;
            orders[i] = sl.d_order;
            slice.resume[i] = sl.resume();
            slice.carry[i] = sl.d_carry;
        }
        if (waves > S.shape.resident_waves) waves = S.shape.resident_waves;
        if (waves)   // without FP64 lanes nothing is taken or carried (collect reads no position)
            e = simk_enqueue_multi_ext(S.shape.count, Ls, lanes, orders, 1, S.d_ws, S.d_queue, (u32)waves, &slice);
        // f32 delta requests: each slab's lanes advance one whole step
        for (u32 i = 0; i < S.shape.count && e == hipSuccess; ++i)
            if (S.slabs[i].lanes32) e = simk_enqueue_d32_ext(i, &Ls[i], S.slabs[i].lanes32, S.slabs[i].d_order32);
        if (e == hipSuccess) e = hipEventRecord(S.ev[3], st);
        // Uploads of requests that fit now overlap the launch: their lanes
        // are outside every order array of this launch, and the next
        // iteration's advection waits for them.
        if (e == hipSuccess) e = admit(S.copy_stream);
        if (e == hipSuccess) e = hipMemcpyAsync(S.h_queue, S.d_queue, 8 * 4, hipMemcpyDeviceToHost, st);
        for (u32 i = 0; i < S.shape.count && e == hipSuccess; ++i)
            if (slice.ticks)
                e = hipMemcpyAsync(S.slabs[i].h_carry, S.slabs[i].d_carry, (u64)waves * 64 * 4,
                                   hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) return code(e);
        if (!slice.ticks || !waves)
            for (u32 i = 0; i < S.shape.count; ++i) S.h_queue[4 + i] = 0;
        collect(step_done);
        // Requests whose last step just ended: copy their solver rows back.
        for (u32 t : step_done) {
            Request &q = S.reqs[t];
            if (q.r.kind == SimkServe_Dbd) {
                // the step's end: pulses committed, then plug flow along the chain (dbd_scene_step)
                Slab &sl = S.slabs[q.r.mech];
                u64 off[12];
                zone_offsets(sl.n, sl.R, sl.cap, off);
                hipLaunchKernelGGL(simk_dbd_commit_kernel, dim3((q.r.cells + 255) / 256), dim3(256), 0, st,
                                   sl.dbd() + (u64)kDbdPulseIndex * sl.cap + q.off,
                                   (const f64 *)(sl.d + off[11] + q.off), q.r.cells);
                e = hipGetLastError();
                if (e == hipSuccess && q.residence > 0.0)
                    e = simk_enqueue_plugflow(st, sl.d + off[0] + q.off, sl.cap, sl.n, q.r.cells,
                                              S.cfg.dt * (f64)q.r.cells / q.residence,
                                              S.d_inflow + (u64)q.slot * kMaxSpecies);
                if (e != hipSuccess) return code(e);
                q.t = q.t + S.cfg.dt;
            }
            if (++q.steps_done < q.steps) {
                q.phase = kAdvect;
                continue;
            }
            Slab &sl = S.slabs[q.r.mech];
            e = hipMemcpy2DAsync(q.r.result, (u64)q.r.cells * 8, sl.d + q.off, (u64)sl.cap * 8, (u64)q.r.cells * 8,
                                 (u64)sl.n + 5, hipMemcpyDeviceToHost, st);
            if (e != hipSuccess) return code(e);
            finished.push_back(t);
        }
    }
    if (e == hipSuccess) e = hipEventRecord(S.ev[4], st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return code(e);
    S.st.iterations++;
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, S.ev[0], S.ev[1]) == hipSuccess) S.st.copy_ms += ms;
    if (hipEventElapsedTime(&ms, S.ev[1], S.ev[2]) == hipSuccess) S.st.advect_ms += ms;
    if (ran) {
        if (hipEventElapsedTime(&ms, S.ev[2], S.ev[3]) == hipSuccess) S.st.chem_ms += ms;
        if (hipEventElapsedTime(&ms, S.ev[3], S.ev[4]) == hipSuccess) S.st.copy_ms += ms;
    }
    for (u32 t : finished) {
        Request &q = S.reqs[t];
        Slab &sl = S.slabs[q.r.mech];
        give_range(sl, q.off, q.r.cells);
        for (u32 c = 0; c < q.r.cells; ++c) sl.owner[q.off + c] = 0;
        S.st.admitted_bytes -= q.bytes;
        S.free_slots.push_back(q.slot);
        S.active.erase(std::find(S.active.begin(), S.active.end(), t));
        S.reqs.erase(t);
        S.st.completed++;
        S.pending_done.push_back(t);
    }
    u32 k = 0;
    while (k < max_done && k < S.pending_done.size() && done) {
        done[k] = S.pending_done[k];
        ++k;
    }
    S.pending_done.erase(S.pending_done.begin(), S.pending_done.begin() + k);
    if (n_done) *n_done = k;
    return 0;
}

GPU_API s32 simk_serve_stats(SimkServeStats *out) {
    if (!out) return code(hipErrorInvalidValue);
    *out = S.st;
    out->queued = (u32)S.queued.size();
    out->active = (u32)S.active.size();
    return S.open ? 0 : code(hipErrorNotInitialized);
}

GPU_API s32 simk_serve_close(void) {
    if (!S.open) return 0;
    hipError_t e = hipStreamSynchronize(simk_module_stream());
    release_all();
    return code(e);
}

}  // extern "C"
