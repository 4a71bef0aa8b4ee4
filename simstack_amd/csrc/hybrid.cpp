// Split of one zone block between the GPU module and the caller's CPU path
// (the reference caller dbd_advance_hybrid, src/sim/dbd.cpp:296-463, runs
// lanes [0, split) on the device asynchronously while its CPU work queue
// integrates [split, padded)). Every lane's result is the same on either
// side, so the split only changes the advance's wall time.
//
// This controller fits a two-part time model from the timings the caller
// reports after each advance:
//   CPU:    t_cpu(n) = c * (padded - n)      c from an average of ms per lane
//   device: t_gpu(n) = o + g * n             o, g by least squares over recent
//                                            (n, t) samples, older ones decaying
// and picks the n that ends both sides together, n = (c*padded - o)/(g + c),
// in steps of 8 lanes (the block's row padding), or no device at all when that
// saves under 3 %. The least-squares fit needs two distinct splits, so the
// split is dithered by one spacing around the optimum until the fit is
// determined, and once every 16 advances afterwards to follow a model that
// drifts as the chemistry stiffens or relaxes; with the device off, every
// 16th advance probes it on a quarter of the block.
#include <cmath>
#include <cstring>

#include "simk_dbd.h"

namespace {

constexpr f64 kForget = 0.85;   // weight of a device sample after each newer one
constexpr f64 kCpuKeep = 0.75;  // weight of the CPU average against the newest sample
constexpr u32 kTrackEvery = 16;
constexpr u32 kGroup = 8;

u32 round_group(f64 n, u32 padded) {
    if (!(n > 0.0)) return 0;
    if (n >= (f64)padded) return padded;
    u32 r = (u32)(n / kGroup + 0.5) * kGroup;
    return r > padded ? padded : r;
}

u32 spacing(u32 padded) {
    u32 s = (padded / 8) / kGroup * kGroup;
    return s < kGroup ? kGroup : s;
}

bool fit_determined(const SimkHybrid *h) {
    const f64 det = h->w * h->snn - h->sn * h->sn;
    return h->w > 0.0 && det > 1e-6 * h->w * h->snn;
}

}  // namespace

extern "C" {

SIMK_HOST_API void simk_hybrid_init(SimkHybrid *h, u32 padded) {
    memset(h, 0, sizeof(*h));
    h->padded = padded;
    h->split = padded >= 2 * kGroup ? round_group(padded / 2.0, padded) : padded;
}

SIMK_HOST_API void simk_hybrid_record(SimkHybrid *h, u32 split, f64 gpu_ms, f64 cpu_ms) {
    const u32 N = h->padded;
    if (split > N) split = N;
    h->steps++;
    if (N - split >= kGroup && cpu_ms > 0.0) {
        const f64 c = cpu_ms / (f64)(N - split);
        h->cpu_lane_ms = h->cpu_samples ? kCpuKeep * h->cpu_lane_ms + (1.0 - kCpuKeep) * c : c;
        h->cpu_samples++;
    }
    if (split > 0 && gpu_ms >= 0.0) {
        const f64 n = (f64)split;
        h->w = kForget * h->w + 1.0;
        h->sn = kForget * h->sn + n;
        h->snn = kForget * h->snn + n * n;
        h->st = kForget * h->st + gpu_ms;
        h->snt = kForget * h->snt + n * gpu_ms;
        h->gpu_samples++;
        if (fit_determined(h)) {
            const f64 det = h->w * h->snn - h->sn * h->sn;
            f64 g = (h->w * h->snt - h->sn * h->st) / det;
            f64 o = (h->st - g * h->sn) / h->w;
            if (o < 0.0 || g <= 0.0) {        // no fixed cost in the data: a line through 0
                o = 0.0;
                g = h->snt / h->snn;
            }
            h->gpu_fixed_ms = o;
            h->gpu_lane_ms = g > 1e-12 ? g : 1e-12;
        } else if (h->gpu_lane_ms > 0.0) {   // one split so far: keep g, refit o
            const f64 o = (h->st - h->gpu_lane_ms * h->sn) / h->w;
            h->gpu_fixed_ms = o > 0.0 ? o : 0.0;
        } else {
            h->gpu_fixed_ms = 0.0;
            h->gpu_lane_ms = gpu_ms / n > 1e-12 ? gpu_ms / n : 1e-12;
        }
    }

    const u32 d = spacing(N);
    if (!h->cpu_samples) {                    // never timed the CPU: give it a share
        h->split = N > d ? N - d : 0;
        return;
    }
    if (!h->gpu_samples) {
        h->split = round_group(N / 2.0, N);
        if (!h->split) h->split = N < kGroup ? N : kGroup;
        return;
    }
    const f64 c = h->cpu_lane_ms, o = h->gpu_fixed_ms, g = h->gpu_lane_ms;
    const f64 n_eq = ((f64)N * c - o) / (g + c);
    u32 best = round_group(n_eq, N);
    const f64 t_best = fmax(o + g * best, c * (f64)(N - best));
    if (best > 0 && !(t_best < 0.97 * c * (f64)N)) best = 0;
    h->best = best;

    const bool track = h->steps % kTrackEvery == 0;
    if (best == 0) {
        h->split = track ? round_group(N / 4.0, N) : 0;
        if (track && !h->split) h->split = N < kGroup ? N : kGroup;
        return;
    }
    if (!fit_determined(h) || track) {
        h->dither = !h->dither;
        if (h->dither || best < d) {
            h->split = best + d <= N ? best + d : (best >= d ? best - d : best);
        } else {
            h->split = best - d;
        }
        return;
    }
    h->split = best;
}

}  // extern "C"
