"""The serving workload and its driver over the simk_serve_* entry points.

Requests arrive as a Poisson process; each names one of the three ragged-batch
mechanisms (60/200, 40/130, 24/70, uniform) and a cell count log-uniform in
[64, 262144], laid out as channels of up to 4,096 nodes (counts above one
channel round up to whole channels). Every request runs the module's
operator-split loop: upwind advection along its channels, then the chemistry
advance, for the configured number of steps, device-resident from admission
to completion.

Request inputs are windows of one pre-generated pool of seeded cells per
mechanism (a request's block is a view into the pool with the pool's row
stride), so arrivals cost no host generation time.
"""
import math
import time

import numpy as np

from . import problem

MECHANISMS = ("c3a", "c3b", "c3c")
CHANNEL_NODES = 4096


def channel_layout(cells, nodes=CHANNEL_NODES):
    """(channels, nodes, cells) with cells rounded up to whole channels."""
    if cells <= nodes:
        return 1, cells, cells
    ch = -(-cells // nodes)
    return ch, nodes, ch * nodes


def courant_profile(channels, nodes):
    """Courant number per node: a film that speeds up along the channel
    (0.2 at the inlet to 0.8 at the outlet), the same in every channel."""
    c = 0.2 + 0.6 * np.arange(nodes) / max(nodes - 1, 1)
    return np.tile(c, channels)


class Arrivals:
    """Seeded Poisson arrivals: (time s, mechanism index, cells)."""

    def __init__(self, rate=200.0, window=1.0, lo=64, hi=262144, seed=5):
        rng = np.random.default_rng(seed)
        self.items = []
        t = 0.0
        while True:
            t += rng.exponential(1.0 / rate)
            if t >= window:
                break
            cells = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
            self.items.append((t, int(rng.integers(len(MECHANISMS))), cells))


class Pools:
    """Per mechanism: the mechanism and a seeded pool of `cells` cells."""

    def __init__(self, cells=262144 + CHANNEL_NODES, names=MECHANISMS):
        self.mechs, self.zones, self.specs = [], [], []
        for name in names:
            sp = problem.spec(name)
            m = problem.Mechanism.synthetic(sp)
            z = problem.Zones(m.n, m.R, cells).fill_synthetic(sp, m.params)
            self.mechs.append(m)
            self.zones.append(z)
            self.specs.append(sp)

    def window(self, mech, first, cells):
        """Host address of a request block of `cells` cells starting at pool lane
        `first` (row stride = the pool's padded)."""
        z = self.zones[mech]
        assert first + cells <= z.count
        return z.block.ctypes.data + first * 8, z.padded

    def feed(self, mech):
        """Inflow concentrations: the pool's first cell."""
        z = self.zones[mech]
        return np.ascontiguousarray(z.y[:, 0])


def run(gm, pools, arrivals, steps=20, substeps=4, dt=1e-3, hbm_budget=16 << 30, seed=9, realtime=True,
        on_done=None):
    """Drives the serving loop over `arrivals` in real time (a request is
    submitted once the clock passes its arrival time). Returns per-request
    latency (completion - arrival, s), totals and the module's statistics.
    on_done(index, result) is called with each finished request's result rows."""
    rng = np.random.default_rng(seed)
    meth, ctl = problem.method(2), problem.control()
    gm.set_mechs(pools.mechs, meth, ctl)
    gm.serve_open(hbm_budget=hbm_budget, max_requests=max(len(arrivals.items), 1), steps=steps,
                  substeps=substeps, dt=dt)
    feeds = [pools.feed(m) for m in range(len(pools.mechs))]
    flops_per_attempt = [m.flops_per_attempt() for m in pools.mechs]
    n = len(arrivals.items)
    lat = np.zeros(n)
    cells_total = 0
    ticket_of = {}
    results = {}
    i = done = failed = 0
    accepted = attempts = flops = 0.0
    t0 = time.perf_counter()
    try:
        while done < n:
            now = time.perf_counter() - t0
            while i < n and (arrivals.items[i][0] <= now or not realtime):
                _, mech, want = arrivals.items[i]
                channels, nodes, cells = channel_layout(want)
                first = int(rng.integers(0, pools.zones[mech].count - cells + 1))
                ptr, padded = pools.window(mech, first, cells)
                result = np.empty((pools.mechs[mech].n + 5, cells))
                t = gm.serve_submit(mech, ptr, padded, cells, channels, nodes, courant_profile(channels, nodes),
                                    result, feeds[mech])
                ticket_of[t] = i
                results[i] = result
                cells_total += cells
                i += 1
            st = gm.serve_stats()
            if st.queued == 0 and st.active == 0:
                if i < n:
                    time.sleep(max(0.0, min(arrivals.items[i][0] - (time.perf_counter() - t0), 0.01)))
                continue
            for t in gm.serve_step():
                k = ticket_of.pop(t)
                lat[k] = time.perf_counter() - t0 - arrivals.items[k][0]
                res = results.pop(k)
                ns = pools.mechs[arrivals.items[k][1]].n
                accepted += float(res[ns + 1].sum())
                att = float(res[ns + 1].sum() + res[ns + 2].sum())
                attempts += att
                flops += att * flops_per_attempt[arrivals.items[k][1]]
                failed += int(res[ns + 4].sum())
                if on_done:
                    on_done(k, res)
                done += 1
        wall = time.perf_counter() - t0
        st = gm.serve_stats()
    finally:
        gm.serve_close()
    return {"latency": lat, "wall": wall, "cells": cells_total, "requests": n, "stats": st, "accepted": accepted,
            "attempts": attempts, "flops": flops, "failed": failed}
