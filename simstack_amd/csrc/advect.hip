// Operator-split upwind film advection on the device (the transport row
// update of nft_transport, src/sim/nft.cpp:1055-1081, and the plug-flow
// relaxation of dbd_flow_exchange, src/sim/dbd.cpp:468-483).
//
// Per substep every node of a channel takes x_j + c_j (x_{j-1} - x_j) from the
// OLD value of its upstream neighbour (node 0 reads the inflow): a Jacobi
// update, so all nodes of a substep are independent. A row is read once, kept
// in registers for every substep and written once, so HBM traffic per launch
// is one read and one write of the field, whatever the substep count. Rows are
// [field][padded] of the zone block with cell = channel * nodes + node.
//
// Upwind rows (up to 8,192 nodes): 1, 2 or 4 waves per (field, channel) row,
// each holding up to 2,048 nodes with node slot * 64 + lane in register slot
// `slot` of lane `lane`, so every global load and store is one coalesced
// 512-byte wave access straight to registers. The upstream neighbour is the
// same slot one lane down (DPP wave_shr:1); lane 0 takes lane 63 of the slot
// before (DPP wave_ror:1 of that slot), slots updated from the last down so
// every read sees the old value. Only a part's first node crosses waves
// (through LDS, one barrier per substep). f64 groups of fields with even row
// lengths hold two nodes per lane and slot (16-byte accesses). The plug-flow chain: one 256-thread
// workgroup per species row, a chunk of consecutive zones per thread, chunk
// boundaries through LDS.
#include <hip/hip_runtime.h>

#include "simk_gpu_api.h"
#include "module_internal.h"

namespace {

constexpr unsigned kThreads = 256;
constexpr unsigned kMaxChunk = 32;   // plug flow: zones per thread, up to 8192 zones
constexpr unsigned kRowsPerBlock = kThreads / 64u;   // waves per workgroup
constexpr unsigned kGroupFields = 8;   // fields of a channel advanced by one wave group

template <class T>
__device__ __forceinline__ T fma_t(T a, T b, T c);
template <>
__device__ __forceinline__ double fma_t<double>(double a, double b, double c) {
    return __builtin_fma(a, b, c);
}
template <>
__device__ __forceinline__ float fma_t<float>(float a, float b, float c) {
    return __builtin_fmaf(a, b, c);
}

// One (field, channel) row of `nodes` values at xr, advanced `substeps`
// substeps by the whole workgroup (kThreads threads, CHUNK nodes each).
// FUSED: x + c (up - x) as one fused multiply-add (nft.cpp:1072); otherwise
// the plug-flow form x + (c * (up - x)) with the product rounded on its own
// (dbd.cpp:480). cr null: every node's coefficient is c_all.
template <class T, unsigned CHUNK, bool FUSED = true>
__device__ __forceinline__ void upwind_row(T *xr, const T *cr, unsigned nodes, unsigned substeps, T in,
                                           T *outlet, T c_all = T(0)) {
    const unsigned t = threadIdx.x;
    __shared__ T stage[kThreads * CHUNK];
    __shared__ T last[kThreads];

    // Coalesced load of the row through LDS, then each thread takes its chunk.
    for (unsigned j = t; j < kThreads * CHUNK; j += kThreads) stage[j] = j < nodes ? xr[j] : T(0);
    __syncthreads();
    T v[CHUNK], c[CHUNK];
#pragma unroll
    for (unsigned i = 0; i < CHUNK; ++i) v[i] = stage[t * CHUNK + i];
    __syncthreads();
    if (cr) {
        for (unsigned j = t; j < kThreads * CHUNK; j += kThreads) stage[j] = j < nodes ? cr[j] : T(0);
        __syncthreads();
#pragma unroll
        for (unsigned i = 0; i < CHUNK; ++i) c[i] = stage[t * CHUNK + i];
    } else {
#pragma unroll
        for (unsigned i = 0; i < CHUNK; ++i) c[i] = c_all;
    }

    const unsigned first = t * CHUNK;
    const unsigned last_thread = (nodes - 1) / CHUNK;
    T out = (outlet && t == last_thread) ? *outlet : T(0);
    for (unsigned s = 0; s < substeps; ++s) {
        if (outlet && t == last_thread) out = out + v[(nodes - 1) - first];
        last[t] = v[CHUNK - 1];
        __syncthreads();
        const T left = t ? last[t - 1] : in;
        // Right to left inside the chunk so each update reads the old neighbour.
#pragma unroll
        for (unsigned i = CHUNK; i-- > 0;) {
            const T up = i ? v[i - 1] : left;
            if (first + i < nodes) v[i] = FUSED ? fma_t<T>(c[i], up - v[i], v[i]) : v[i] + c[i] * (up - v[i]);
        }
        __syncthreads();
    }
    if (outlet && t == last_thread) *outlet = out;

#pragma unroll
    for (unsigned i = 0; i < CHUNK; ++i) stage[t * CHUNK + i] = v[i];
    __syncthreads();
    for (unsigned j = t; j < nodes; j += kThreads) xr[j] = stage[j];
}

// DPP across the whole wave: wave_shr:1 (lane l takes lane l - 1; lane 0
// keeps `old`) and wave_ror:1 (lane l takes lane l - 1, lane 0 takes lane 63).
constexpr int kWaveShr1 = 0x138, kWaveRor1 = 0x13C;
template <int CTRL>
__device__ __forceinline__ float dpp(float old, float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, old),
                                                                 __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp(double old, double v) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned long long o = __builtin_bit_cast(unsigned long long, old);
    const int lo = __builtin_amdgcn_update_dpp((int)(unsigned)o, (int)(unsigned)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(unsigned)(o >> 32), (int)(unsigned)(b >> 32), CTRL, 0xF, 0xF,
                                               false);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
// The upstream values of slot k: lane l gets cur's lane l - 1, lane 0 gets
// prev's lane 63 (prev: the slot before, or `first` broadcast for slot 0).
template <class T>
__device__ __forceinline__ T shift_up(T cur, T prev_rotated) {
    return dpp<kWaveShr1>(prev_rotated, cur);
}
__device__ __forceinline__ float lane_value(float v, unsigned lane) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), (int)lane));
}
__device__ __forceinline__ double lane_value(double v, unsigned lane) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, (int)lane);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), (int)lane);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Raw buffer access (gfx9 family descriptor word 3: raw buffer, 32-bit data
// format). Loads at or beyond the resource's extent return 0 and stores there
// are dropped; the range check covers the lane offset and the instruction's
// immediate offset, and the scalar offset is always 0 here.
constexpr int kRawBufferFlags = 0x00027000;
template <class T>
__device__ __forceinline__ T buf_load(__amdgpu_buffer_rsrc_t r, unsigned off);
template <>
__device__ __forceinline__ double buf_load<double>(__amdgpu_buffer_rsrc_t r, unsigned off) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
}
template <>
__device__ __forceinline__ float buf_load<float>(__amdgpu_buffer_rsrc_t r, unsigned off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
}
template <class T>
__device__ __forceinline__ void buf_store(__amdgpu_buffer_rsrc_t r, unsigned off, T v);
template <>
__device__ __forceinline__ void buf_store<double>(__amdgpu_buffer_rsrc_t r, unsigned off, double v) {
    typedef unsigned int v2u __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(v2u, v), r, (int)off, 0, 0);
}
template <>
__device__ __forceinline__ void buf_store<float>(__amdgpu_buffer_rsrc_t r, unsigned off, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)off, 0, 0);
}

// One part of a row: nodes [part * SLOTS * 64, (part + 1) * SLOTS * 64) of a
// row of `nodes` values, advanced by one wave; a row of WPR parts is advanced
// by WPR consecutive waves of the workgroup. Each substep a part's first node
// takes its upstream value from the previous part's last node (old value)
// through LDS, one workgroup barrier per substep (double-buffered by
// substep parity); WPR = 1 needs neither. `active` false: a wave past the
// last row, which only joins the barriers. Nodes past the end hold 0 with
// coefficient 0: fma(0, up - 0, 0) leaves them unchanged for finite data,
// they only feed further nodes past the end, and they are never stored.
template <class T, unsigned SLOTS, unsigned WPR>
__device__ __forceinline__ void upwind_row_wave(T *xr, const T *cr, unsigned nodes, unsigned substeps, T in,
                                                T *outlet, bool active) {
    constexpr unsigned kPart = SLOTS * 64u;
    __shared__ T edge[2][kRowsPerBlock];
    const unsigned lane = threadIdx.x & 63u, wave = threadIdx.x / 64u, part = wave % WPR;
    const unsigned base = part * kPart;
    T v[SLOTS], c[SLOTS];
#pragma unroll
    for (unsigned k = 0; k < SLOTS; ++k) {
        const unsigned j = base + k * 64u + lane;
        const bool in_row = active && j < nodes;
        v[k] = in_row ? xr[j] : T(0);
        c[k] = in_row ? cr[j] : T(0);
    }
    // the outlet node (nodes - 1) lives in one part
    const bool owns_outlet = active && outlet && (nodes - 1) / kPart == part;
    const unsigned oslot = ((nodes - 1) % kPart) / 64u, olane = (nodes - 1) % 64u;
    T out = owns_outlet ? *outlet : T(0);
    for (unsigned s = 0; s < substeps; ++s) {
        if (owns_outlet) {
            T ov = T(0);
#pragma unroll
            for (unsigned k = 0; k < SLOTS; ++k)
                if (k == oslot) ov = v[k];
            out = out + lane_value(ov, olane);
        }
        T first = in;
        if constexpr (WPR > 1) {
            if (lane == 63) edge[s & 1u][wave] = v[SLOTS - 1];
            __syncthreads();
            if (part) first = edge[s & 1u][wave - 1];
        }
#pragma unroll
        for (unsigned k = SLOTS; k-- > 0;) {
            const T prev = k ? dpp<kWaveRor1>(T(0), v[k - 1]) : first;
            const T up = shift_up(v[k], prev);
            v[k] = fma_t<T>(c[k], up - v[k], v[k]);
        }
    }
    if (owns_outlet && lane == 0) *outlet = out;
    if (!active) return;
#pragma unroll
    for (unsigned k = 0; k < SLOTS; ++k) {
        const unsigned j = base + k * 64u + lane;
        if (j < nodes) xr[j] = v[k];
    }
}

// Rows of the field in row order (field * channels + channel), WPR waves per
// row, kRowsPerBlock / WPR rows per workgroup. (Measured: an XCD-aware order
// that keeps each XCD's coefficient rows in its L2 was no faster; the 256 MB
// Infinity Cache already serves their re-reads.)
template <class T, unsigned SLOTS, unsigned WPR>
__global__ void __launch_bounds__(kThreads) simk_upwind_wave_kernel(T *x, unsigned long long padded, unsigned rows,
                                                                    unsigned channels, unsigned nodes,
                                                                    const T *courant, unsigned substeps,
                                                                    const T *inflow, T *outlet_sum) {
    const unsigned row = blockIdx.x * (kRowsPerBlock / WPR) + threadIdx.x / 64u / WPR;
    const bool active = row < rows;
    const unsigned r = active ? row : 0;
    const unsigned field = r / channels, ch = r % channels;
    if (WPR == 1 && !active) return;   // whole waves, no barriers
    upwind_row_wave<T, SLOTS, WPR>(x + field * padded + (unsigned long long)ch * nodes,
                                   courant + (unsigned long long)ch * nodes, nodes, substeps,
                                   inflow ? inflow[field] : T(0), outlet_sum ? outlet_sum + r : nullptr, active);
}

// The same row update for G consecutive fields of one channel per wave group:
// the channel's coefficients are loaded once and kept in registers for the
// G rows, instead of once per row. Row groups are numbered field group major,
// channel minor (neighbouring groups touch neighbouring memory). Every wave
// runs G row iterations with the same barriers (a row past the last field is
// inactive: no loads, stores or outlet), and the edge buffers' parity follows
// one substep counter across the rows, so two writes of a buffer are always
// a barrier apart from the reads between them.
template <class T, unsigned SLOTS, unsigned WPR, unsigned G>
__global__ void __launch_bounds__(kThreads) simk_upwind_group_kernel(T *x, unsigned long long padded, unsigned fields,
                                                                     unsigned channels, unsigned nodes,
                                                                     const T *courant, unsigned substeps,
                                                                     const T *inflow, T *outlet_sum) {
    constexpr unsigned kPart = SLOTS * 64u;
    __shared__ T edge[2][kRowsPerBlock];
    // The wave index is uniform across the wave; read from lane 0 it is a
    // scalar, so every row base below is one too.
    const unsigned lane = threadIdx.x & 63u, wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64u);
    const unsigned part = wave % WPR;
    const unsigned groups = (fields + G - 1) / G;
    const unsigned gidx = blockIdx.x * (kRowsPerBlock / WPR) + wave / WPR;
    const bool gactive = gidx < groups * channels;
    if (WPR == 1 && !gactive) return;   // whole waves, no barriers
    const unsigned gi = gactive ? gidx : 0;
    const unsigned ch = gi % channels, fg = gi / channels;
    const unsigned base = part * kPart;
    // This wave's nodes of a row through a buffer resource whose extent ends
    // at the row's last node: a lane past the end (or every lane of an
    // inactive row, extent 0) reads 0 and its store is dropped, so no access
    // needs a per-slot guard.
    const unsigned part_bytes = base < nodes ? (nodes - base) * (unsigned)sizeof(T) : 0u;
    auto row_rsrc = [&](const T *row, const bool active) {
        return __builtin_amdgcn_make_buffer_rsrc((void *)(row + base), 0, active ? (int)part_bytes : 0,
                                                 kRawBufferFlags);
    };
    const unsigned loff = lane * (unsigned)sizeof(T);
    const __amdgpu_buffer_rsrc_t crs = row_rsrc(courant + (unsigned long long)ch * nodes, gactive);
    T c[SLOTS];
#pragma unroll
    for (unsigned k = 0; k < SLOTS; ++k) c[k] = buf_load<T>(crs, loff + k * 64u * (unsigned)sizeof(T));
    const bool outlet_part = outlet_sum && (nodes - 1) / kPart == part;
    const unsigned oslot = ((nodes - 1) % kPart) / 64u, olane = (nodes - 1) % 64u;
    auto field_rsrc = [&](const unsigned i) {
        const unsigned f = fg * G + i;
        const bool active = gactive && f < fields;
        return row_rsrc(x + (unsigned long long)(active ? f : 0u) * padded + (unsigned long long)ch * nodes, active);
    };
    // Field i + 1's part is loaded before field i is advanced, so its loads
    // are in flight during field i's substeps and stores.
    T v[SLOTS], next[SLOTS];
    {
        const __amdgpu_buffer_rsrc_t rs = field_rsrc(0);
#pragma unroll
        for (unsigned k = 0; k < SLOTS; ++k) next[k] = buf_load<T>(rs, loff + k * 64u * (unsigned)sizeof(T));
    }
    unsigned step = 0;
    for (unsigned i = 0; i < G; ++i) {
        const unsigned f = fg * G + i;
        const bool active = gactive && f < fields;
        const __amdgpu_buffer_rsrc_t rs = field_rsrc(i);
#pragma unroll
        for (unsigned k = 0; k < SLOTS; ++k) v[k] = next[k];
        if (i + 1 < G) {
            const __amdgpu_buffer_rsrc_t nrs = field_rsrc(i + 1);
#pragma unroll
            for (unsigned k = 0; k < SLOTS; ++k) next[k] = buf_load<T>(nrs, loff + k * 64u * (unsigned)sizeof(T));
        }
        const bool owns_outlet = active && outlet_part;
        const unsigned r = (active ? f : 0u) * channels + ch;
        T out = owns_outlet ? outlet_sum[r] : T(0);
        const T in = (active && inflow) ? inflow[f] : T(0);
        for (unsigned s = 0; s < substeps; ++s, ++step) {
            if (owns_outlet) {
                T ov = T(0);
#pragma unroll
                for (unsigned k = 0; k < SLOTS; ++k)
                    if (k == oslot) ov = v[k];
                out = out + lane_value(ov, olane);
            }
            T first = in;
            if constexpr (WPR > 1) {
                if (lane == 63) edge[step & 1u][wave] = v[SLOTS - 1];
                __syncthreads();
                if (part) first = edge[step & 1u][wave - 1];
            }
#pragma unroll
            for (unsigned k = SLOTS; k-- > 0;) {
                const T prev = k ? dpp<kWaveRor1>(T(0), v[k - 1]) : first;
                const T up = shift_up(v[k], prev);
                v[k] = fma_t<T>(c[k], up - v[k], v[k]);
            }
        }
        if (owns_outlet && lane == 0) outlet_sum[r] = out;
#pragma unroll
        for (unsigned k = 0; k < SLOTS; ++k) buf_store<T>(rs, loff + k * 64u * (unsigned)sizeof(T), v[k]);
    }
}

// The f64 group kernel with 16-byte lane accesses: lane l of slot k holds
// nodes base + 128 k + 2 l and 2 l + 1 of its row, so every load and store
// of a row moves 1 KiB per wave instruction. Node 2 l + 1 takes its upstream
// value from node 2 l of the same lane; node 2 l from node 2 l - 1, the .y of
// lane l - 1 (DPP wave_shr:1; lane 0 takes lane 63 of the slot before, DPP
// wave_ror:1). Both come from old values, and each node's update is the same
// fused multiply-add as in the kernels above. Used when every row base is
// 16-byte aligned and the row length is even, so a lane's pair is either
// wholly inside its row or wholly past the end.
struct SimkD2A {
    double x, y;
};
__device__ __forceinline__ SimkD2A buf_load2(__amdgpu_buffer_rsrc_t r, unsigned off) {
    typedef unsigned int v4u __attribute__((ext_vector_type(4)));
    const v4u q = __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0);
    typedef unsigned int v2u __attribute__((ext_vector_type(2)));
    return {__builtin_bit_cast(double, v2u{q.x, q.y}), __builtin_bit_cast(double, v2u{q.z, q.w})};
}
__device__ __forceinline__ void buf_store2(__amdgpu_buffer_rsrc_t r, unsigned off, SimkD2A v) {
    typedef unsigned int v4u __attribute__((ext_vector_type(4)));
    typedef unsigned int v2u __attribute__((ext_vector_type(2)));
    const v2u a = __builtin_bit_cast(v2u, v.x), b = __builtin_bit_cast(v2u, v.y);
    __builtin_amdgcn_raw_buffer_store_b128(v4u{a.x, a.y, b.x, b.y}, r, (int)off, 0, 0);
}
template <unsigned SLOTS, unsigned WPR, unsigned G>
__global__ void __launch_bounds__(kThreads) simk_upwind_group2_kernel(double *x, unsigned long long padded,
                                                                      unsigned fields, unsigned channels,
                                                                      unsigned nodes, const double *courant,
                                                                      unsigned substeps, const double *inflow) {
    constexpr unsigned kPart = SLOTS * 128u;
    __shared__ double edge[2][kRowsPerBlock];
    const unsigned lane = threadIdx.x & 63u, wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64u);
    const unsigned part = wave % WPR;
    const unsigned groups = (fields + G - 1) / G;
    const unsigned gidx = blockIdx.x * (kRowsPerBlock / WPR) + wave / WPR;
    const bool gactive = gidx < groups * channels;
    if (WPR == 1 && !gactive) return;   // whole waves, no barriers
    const unsigned gi = gactive ? gidx : 0;
    const unsigned ch = gi % channels, fg = gi / channels;
    const unsigned base = part * kPart;
    const unsigned part_bytes = base < nodes ? (nodes - base) * 8u : 0u;
    auto row_rsrc = [&](const double *row, const bool active) {
        return __builtin_amdgcn_make_buffer_rsrc((void *)(row + base), 0, active ? (int)part_bytes : 0,
                                                 kRawBufferFlags);
    };
    const unsigned loff = lane * 16u;
    const __amdgpu_buffer_rsrc_t crs = row_rsrc(courant + (unsigned long long)ch * nodes, gactive);
    SimkD2A c[SLOTS];
#pragma unroll
    for (unsigned k = 0; k < SLOTS; ++k) c[k] = buf_load2(crs, loff + k * 1024u);
    auto field_rsrc = [&](const unsigned i) {
        const unsigned f = fg * G + i;
        const bool active = gactive && f < fields;
        return row_rsrc(x + (unsigned long long)(active ? f : 0u) * padded + (unsigned long long)ch * nodes, active);
    };
    SimkD2A v[SLOTS], next[SLOTS];
    {
        const __amdgpu_buffer_rsrc_t rs = field_rsrc(0);
#pragma unroll
        for (unsigned k = 0; k < SLOTS; ++k) next[k] = buf_load2(rs, loff + k * 1024u);
    }
    unsigned step = 0;
    for (unsigned i = 0; i < G; ++i) {
        const unsigned f = fg * G + i;
        const bool active = gactive && f < fields;
        const __amdgpu_buffer_rsrc_t rs = field_rsrc(i);
#pragma unroll
        for (unsigned k = 0; k < SLOTS; ++k) v[k] = next[k];
        if (i + 1 < G) {
            const __amdgpu_buffer_rsrc_t nrs = field_rsrc(i + 1);
#pragma unroll
            for (unsigned k = 0; k < SLOTS; ++k) next[k] = buf_load2(nrs, loff + k * 1024u);
        }
        const double in = (active && inflow) ? inflow[f] : 0.0;
        for (unsigned s = 0; s < substeps; ++s, ++step) {
            double first = in;
            if constexpr (WPR > 1) {
                if (lane == 63) edge[step & 1u][wave] = v[SLOTS - 1].y;
                __syncthreads();
                if (part) first = edge[step & 1u][wave - 1];
            }
#pragma unroll
            for (unsigned k = SLOTS; k-- > 0;) {
                const double prev = k ? dpp<kWaveRor1>(0.0, v[k - 1].y) : first;
                const double up_x = shift_up(v[k].y, prev), up_y = v[k].x;
                v[k].y = fma_t<double>(c[k].y, up_y - v[k].y, v[k].y);
                v[k].x = fma_t<double>(c[k].x, up_x - v[k].x, v[k].x);
            }
        }
#pragma unroll
        for (unsigned k = 0; k < SLOTS; ++k) buf_store2(rs, loff + k * 1024u, v[k]);
    }
}

// Plug flow along a chain of zones (dbd_flow_exchange, dbd.cpp:465-483): one
// workgroup per species row; zone 0 relaxes toward the feed.
template <unsigned CHUNK>
__global__ void __launch_bounds__(kThreads) simk_plugflow_kernel(double *y, unsigned long long padded, unsigned zones,
                                                                 double a, const double *feed) {
    const unsigned s = blockIdx.x;
    upwind_row<double, CHUNK, false>(y + s * padded, nullptr, zones, 1, feed[s], nullptr, a);
}

// Many requests' rows in one launch: wave row b belongs to the segment whose
// [row_begin, row_begin + fields * channels) holds b (segments sorted by
// row_begin, the first at 0). SLOTS and WPR fit the longest row; shorter rows
// leave their last parts empty.
template <unsigned SLOTS, unsigned WPR>
__global__ void __launch_bounds__(kThreads) simk_upwind_batch_wave_kernel(const SimkUpwindSeg *segs, unsigned count,
                                                                          unsigned total_rows, unsigned substeps) {
    const unsigned b = blockIdx.x * (kRowsPerBlock / WPR) + threadIdx.x / 64u / WPR;
    const bool active = b < total_rows;
    if (WPR == 1 && !active) return;   // whole waves, no barriers
    const unsigned bb = active ? b : 0;
    unsigned lo = 0, hi = count;
    while (hi - lo > 1) {
        const unsigned mid = (lo + hi) / 2;
        if (segs[mid].row_begin <= bb) lo = mid;
        else hi = mid;
    }
    const SimkUpwindSeg &S = segs[lo];
    const unsigned row = bb - S.row_begin;
    const unsigned field = row / S.channels, ch = row % S.channels;
    upwind_row_wave<double, SLOTS, WPR>(S.x + field * S.padded + (unsigned long long)ch * S.nodes,
                                        S.courant + (unsigned long long)ch * S.nodes, S.nodes, substeps,
                                        S.inflow ? S.inflow[field] : 0.0, nullptr, active);
}

// Wave shape for rows of up to `nodes` values: 16 or 32 slots per wave, and
// 1, 2 or 4 waves per row (0: too long).
struct WaveShape {
    unsigned slots, wpr;
};
WaveShape wave_shape(u32 nodes) {
    if (nodes <= 1024) return {16, 1};
    if (nodes <= 2048) return {32, 1};
    if (nodes <= 4096) return {32, 2};
    if (nodes <= 8192) return {32, 4};
    return {0, 0};
}

template <class T>
hipError_t launch_upwind(hipStream_t stream, T *x, u64 padded, u32 fields, u32 channels, u32 nodes,
                         const T *courant, u32 substeps, const T *inflow, T *outlet_sum) {
    if (!x || !courant || fields == 0 || channels == 0 || nodes == 0) return hipErrorInvalidValue;
    if ((u64)channels * nodes > padded) return hipErrorInvalidValue;
    if ((u64)fields * channels >= (1ull << 32)) return hipErrorInvalidValue;
    const u32 rows = fields * channels;
    const WaveShape w = wave_shape(nodes);
    if (!w.wpr) return hipErrorInvalidValue;
    const u32 per_block = kRowsPerBlock / w.wpr;
    const dim3 block(kThreads);
    const unsigned long long pd = padded;
    if constexpr (sizeof(T) == 8) {
        // f64 rows with 16-byte lane accesses: every row base 16-byte aligned
        // and an even row length (a lane's pair never straddles a row's end)
        if (fields >= kGroupFields && !outlet_sum && nodes % 2 == 0 && padded % 2 == 0 &&
            (reinterpret_cast<uintptr_t>(x) & 15u) == 0 && (reinterpret_cast<uintptr_t>(courant) & 15u) == 0) {
            const u32 groups = (fields + kGroupFields - 1) / kGroupFields;
            const u64 units = (u64)groups * channels;
            const u32 wpr2 = nodes <= 2048 ? 1u : (nodes <= 4096 ? 2u : 4u);
            const u32 per2 = kRowsPerBlock / wpr2;
            const dim3 ggrid((u32)((units + per2 - 1) / per2));
            if (nodes <= 1024)
                hipLaunchKernelGGL((simk_upwind_group2_kernel<8, 1, kGroupFields>), ggrid, block, 0, stream, x, pd,
                                   fields, channels, nodes, courant, substeps, inflow);
            else if (wpr2 == 1)
                hipLaunchKernelGGL((simk_upwind_group2_kernel<16, 1, kGroupFields>), ggrid, block, 0, stream, x, pd,
                                   fields, channels, nodes, courant, substeps, inflow);
            else if (wpr2 == 2)
                hipLaunchKernelGGL((simk_upwind_group2_kernel<16, 2, kGroupFields>), ggrid, block, 0, stream, x, pd,
                                   fields, channels, nodes, courant, substeps, inflow);
            else
                hipLaunchKernelGGL((simk_upwind_group2_kernel<16, 4, kGroupFields>), ggrid, block, 0, stream, x, pd,
                                   fields, channels, nodes, courant, substeps, inflow);
            return hipGetLastError();
        }
    }
    if (fields >= kGroupFields) {
        // kGroupFields fields of a channel per wave group (coefficients loaded once)
        const u32 groups = (fields + kGroupFields - 1) / kGroupFields;
        const u64 units = (u64)groups * channels;
        const dim3 ggrid((u32)((units + per_block - 1) / per_block));
        if (w.slots == 16)
            hipLaunchKernelGGL((simk_upwind_group_kernel<T, 16, 1, kGroupFields>), ggrid, block, 0, stream, x, pd,
                               fields, channels, nodes, courant, substeps, inflow, outlet_sum);
        else if (w.wpr == 1)
            hipLaunchKernelGGL((simk_upwind_group_kernel<T, 32, 1, kGroupFields>), ggrid, block, 0, stream, x, pd,
                               fields, channels, nodes, courant, substeps, inflow, outlet_sum);
        else if (w.wpr == 2)
            hipLaunchKernelGGL((simk_upwind_group_kernel<T, 32, 2, kGroupFields>), ggrid, block, 0, stream, x, pd,
                               fields, channels, nodes, courant, substeps, inflow, outlet_sum);
        else
            hipLaunchKernelGGL((simk_upwind_group_kernel<T, 32, 4, kGroupFields>), ggrid, block, 0, stream, x, pd,
                               fields, channels, nodes, courant, substeps, inflow, outlet_sum);
        return hipGetLastError();
    }
    const dim3 grid((rows + per_block - 1) / per_block);
    if (w.slots == 16)
        hipLaunchKernelGGL((simk_upwind_wave_kernel<T, 16, 1>), grid, block, 0, stream, x, pd, rows, channels,
                           nodes, courant, substeps, inflow, outlet_sum);
    else if (w.wpr == 1)
        hipLaunchKernelGGL((simk_upwind_wave_kernel<T, 32, 1>), grid, block, 0, stream, x, pd, rows, channels,
                           nodes, courant, substeps, inflow, outlet_sum);
    else if (w.wpr == 2)
        hipLaunchKernelGGL((simk_upwind_wave_kernel<T, 32, 2>), grid, block, 0, stream, x, pd, rows, channels,
                           nodes, courant, substeps, inflow, outlet_sum);
    else
        hipLaunchKernelGGL((simk_upwind_wave_kernel<T, 32, 4>), grid, block, 0, stream, x, pd, rows, channels,
                           nodes, courant, substeps, inflow, outlet_sum);
    return hipGetLastError();
}

}  // namespace

hipError_t simk_enqueue_upwind_batch(hipStream_t stream, const SimkUpwindSeg *segs, u32 count, u32 total_rows,
                                     u32 max_nodes, u32 substeps) {
    if (!segs || count == 0 || total_rows == 0 || max_nodes == 0) return hipErrorInvalidValue;
    const WaveShape w = wave_shape(max_nodes);
    if (!w.wpr) return hipErrorInvalidValue;
    const u32 per_block = kRowsPerBlock / w.wpr;
    const dim3 grid((total_rows + per_block - 1) / per_block), block(kThreads);
    if (w.slots == 16)
        hipLaunchKernelGGL((simk_upwind_batch_wave_kernel<16, 1>), grid, block, 0, stream, segs, count, total_rows,
                           substeps);
    else if (w.wpr == 1)
        hipLaunchKernelGGL((simk_upwind_batch_wave_kernel<32, 1>), grid, block, 0, stream, segs, count, total_rows,
                           substeps);
    else if (w.wpr == 2)
        hipLaunchKernelGGL((simk_upwind_batch_wave_kernel<32, 2>), grid, block, 0, stream, segs, count, total_rows,
                           substeps);
    else
        hipLaunchKernelGGL((simk_upwind_batch_wave_kernel<32, 4>), grid, block, 0, stream, segs, count, total_rows,
                           substeps);
    return hipGetLastError();
}

extern "C" {

GPU_API s32 simk_upwind_f64(f64 *y, u64 padded, u32 fields, u32 channels, u32 nodes, const f64 *courant,
                            u32 substeps, const f64 *inflow) {
    return (s32)launch_upwind<f64>(simk_module_stream(), y, padded, fields, channels, nodes, courant, substeps,
                                   inflow, nullptr);
}

GPU_API s32 simk_plugflow_f64(f64 *y, u64 padded, u32 species, u32 zones, f64 a, const f64 *feed) {
    return (s32)simk_enqueue_plugflow(simk_module_stream(), y, padded, species, zones, a, feed);
}

}  // extern "C"

hipError_t simk_enqueue_plugflow(hipStream_t st, f64 *y, u64 padded, u32 species, u32 zones, f64 a, const f64 *feed) {
    if (!y || !feed || species == 0 || zones == 0 || zones > padded) return hipErrorInvalidValue;
    if (!(a >= 0.0 && a <= 1.0)) return hipErrorInvalidValue;   // the reference asserts a <= 1 (dbd.cpp:473)
    const u32 chunk_needed = (zones + kThreads - 1) / kThreads;
    const dim3 grid(species), block(kThreads);
    if (chunk_needed <= 4)
        hipLaunchKernelGGL((simk_plugflow_kernel<4>), grid, block, 0, st, y, (unsigned long long)padded, zones, a, feed);
    else if (chunk_needed <= 16)
        hipLaunchKernelGGL((simk_plugflow_kernel<16>), grid, block, 0, st, y, (unsigned long long)padded, zones, a,
                           feed);
    else if (chunk_needed <= kMaxChunk)
        hipLaunchKernelGGL((simk_plugflow_kernel<kMaxChunk>), grid, block, 0, st, y, (unsigned long long)padded, zones,
                           a, feed);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

extern "C" {

GPU_API s32 simk_upwind_f32(f32 *x, u64 padded, u32 fields, u32 channels, u32 nodes, const f32 *courant,
                            u32 substeps, f32 *outlet_sum) {
    return (s32)launch_upwind<f32>(simk_module_stream(), x, padded, fields, channels, nodes, courant, substeps,
                                   nullptr, outlet_sum);
}

}  // extern "C"
