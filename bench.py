"""Benchmark: cell-steps/s of the batched Rosenbrock solver on one MI355X.

Headline workload (the metric's): the 48-species / 150-reaction synthetic
mechanism (~18 % fill after the symbolic LU), 262,144 cells, pH 2-7, rtol 1e-6
atol 1e-12, Rodas3, K operator-split advances of 1 s carrying h, fused in one
persistent launch (one bench "step" = one 1-s advance of every cell). Inputs
are resident in HBM before timing starts.

It runs the FP64 path (bit-identical to the FP64 oracle), the FP64 path
started one ulp higher (the sensitivity floor of any comparison), and the
mixed-precision variants: FP32 LU and solves with 0, 1 or 2 FP64 refinement
sweeps per stage (precision 2), and FP32 factor storage (precision 3: FP64
arithmetic, off-diagonal factor entries stored as f32). The accuracy criterion: the p99.9 over cells of each cell's max
relative error against the FP64 run, over species above atol/rtol, is at most
1e-6. value = accepted cell-steps / wall seconds (the reference's definition,
src/tests/test_kinetics.cpp:1037-1042) of the fastest variant that meets it.

The JSON line also carries each variant's speed, errors and roofline (FP64
and FP32 VALU flops per attempt from the tables over the kernel time from HIP
events on the module stream; counted traffic per attempt from the committed
rocprofv3 passes), the 65,536-cell FP64 workload, the ragged batch, serving,
the DBD reactor, advection, and the reference's CPU path on a bounded sample.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_VALU_TFLOPS = 78.6     # MI355X FP64 vector, dense
PEAK_FP32_VALU_TFLOPS = 157.3    # MI355X FP32 vector, dense
PEAK_HBM_GBS = 8000.0


def cpu_baseline(mech, meth, ctl, sp, cells, duration, repeats=1, threads=0, port=False):
    """The reference's own FP64 AVX2 path (oracle/_ref, f64x4 lanes, threads as
    dbd_advance_threads) on the first `cells` cells for `repeats` operator-split
    advances of `duration` from the seeded state (the timed GPU run's seconds);
    falls back to the oracle's scalar restatement if the reference build is
    absent, and uses it when `port` is set. threads 0: the host's CPU share
    (OMP_NUM_THREADS, else the CPU count, at most 64)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ffi
    from simstack_amd import problem
    if threads <= 0:
        threads = min(int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count() or 1, 64)
    z = problem.Zones(mech.n, mech.R, cells).fill_synthetic(sp, mech.params)
    t0 = time.perf_counter()
    if oracle_ffi.have_ref() and not port:
        mb = mech.rebuild(oracle_ffi.ref_caps()).bytes
        mm = oracle_ffi.ref_method_bytes(2)
        for _ in range(repeats):
            oracle_ffi.ref_advance(mb, mm, ctl, z, duration, z.padded, width=4, threads=threads)
        kind = "reference"
    else:
        for _ in range(repeats):
            oracle_ffi.oracle_advance(mech, meth, ctl, z, duration, threads=threads)
        kind = "port"
    wall = time.perf_counter() - t0
    steps = float(z.steps[:cells].sum())
    return {"value": steps / wall, "unit": "cell-steps/s", "cores": threads, "kind": kind,
            "sample": f"first {cells} cells of the workload, {repeats} x {duration:g}-s advances from the seeded "
                      f"state ({int(steps)} accepted steps, {wall:.1f} s wall, {wall * threads:.0f} core-s)"}


def run_workload(gpu, problem, mech, sp, meth, ctl, cells, steps, warmup, precision=0, sweeps=0, separate=False,
                 kernel=0, grid=0, reg_window=0, waves_per_eu=0, batch=0, tune=0.0, lds_slots=0, perturb=False,
                 reuse=0, order=0, stage_carry=0):
    """Advances `cells` cells of the workload `steps` x 1 s from the seeded initial
    state (after `warmup` untimed seconds on a copy). Returns timings and the block.
    perturb: every initial concentration moved up by one ulp. order: the
    module's queue policy; 0 (natural order) by default, because the timed
    run repeats the warm-up's cells from the same state, so costs recorded
    in the warm-up would be exact foreknowledge no real caller has."""
    z0 = problem.Zones(mech.n, mech.R, cells).fill_synthetic(sp, mech.params)
    if perturb:
        z0.y[:] = np.nextafter(z0.y, np.inf)
    gm = run_workload.gm
    if gm is None:
        gm = run_workload.gm = gpu.GpuModule(arena_bytes=1 << 30)
    gm.arena_reset()
    need = z0.block.nbytes + (64 << 20)
    if need > gm.info.arena_bytes:
        gm = run_workload.gm = gpu.GpuModule(arena_bytes=need)
    gm.set_kernel(kernel)
    gm.lib.gpu_kin_set_mode(0)
    gm.set_gen_options(reg_window=reg_window, waves_per_eu=waves_per_eu, batch=batch, sweeps=sweeps,
                       lds_slots=lds_slots, reuse=reuse, stage_carry=stage_carry)
    gm.set_grid(grid)
    gm.set_queue_policy(order)
    gm.set_mech(mech, meth, ctl)
    dz = gpu.DeviceZones(gm, z0.copy())
    dz.upload()
    tuned = None
    if tune > 0.0:
        # The whole workload: a sample of about one cell per resident lane
        # (65,536) is dominated by its slowest cells and picked settings that
        # are slower at 262,144 cells.
        t0 = time.perf_counter()
        tuned = dz.autotune(duration=tune, sample_cells=cells, precision=precision).as_dict()
        tuned["tuning_wall_s"] = time.perf_counter() - t0
    if warmup:
        dz.advance(1.0, repeats=warmup, precision=precision)
    dz.z.block[:] = z0.block
    dz.upload()
    kernel_ms = []
    gm.lib.gpu_sync()
    t0 = time.perf_counter()
    if separate:
        for _ in range(steps):
            kernel_ms.append(dz.advance(1.0, precision=precision))
    else:
        kernel_ms.append(dz.advance(1.0, repeats=steps, precision=precision))
    gm.lib.gpu_sync()
    wall = time.perf_counter() - t0
    dz.download()
    z = dz.z
    acc = float((z.steps[:cells] - z0.steps[:cells]).sum())
    att = acc + float((z.rejects[:cells] - z0.rejects[:cells]).sum())
    info = gm.kin_info()
    gm.lib.gpu_kin_set_mode(0)
    return {"wall": wall, "kernel_ms": kernel_ms, "accepted": acc, "attempts": att, "zones": z,
            "failed": int(z.failed[:cells].sum()), "kernel": info.kernel_name.decode(), "waves": info.waves,
            "tuned": tuned}


run_workload.gm = None


def rel_err_groups(got, ref, atol, rtol, same_path=None):
    """Max |got - ref| / max(|got|, |ref|) (test_kinetics.cpp:35-38) over entries
    above atol/rtol, and per concentration group of the reference. same_path:
    per-cell mask of cells whose accepted and rejected step counts equal the
    reference's; the errors are then also given over those cells only and
    their share reported."""
    d = np.abs(got - ref)
    s = np.maximum(np.abs(got), np.abs(ref))
    r = np.divide(d, s, out=np.zeros_like(d), where=s > 0)
    groups = (("above atol/rtol", ref > atol / rtol), ("1e-9..1e-6", (ref > 1e-9) & (ref <= 1e-6)),
              ("1e-11..1e-9", (ref > 1e-11) & (ref <= 1e-9)))
    out = {}
    for name, mask in groups:
        out[name] = float(r[mask].max()) if mask.any() else 0.0
    if same_path is not None:
        out["cells_same_step_counts"] = float(same_path.mean())
        for name, mask in groups:
            m2 = mask & same_path[None, :]
            out[name + ", same step counts"] = float(r[m2].max()) if m2.any() else 0.0
        worst = np.where(ref > atol / rtol, r, 0.0).max(axis=0)
        out["above atol/rtol, cell p99.9"] = float(np.percentile(worst, 99.9))
    return out


def same_counts(a, b, n):
    return (a.steps[:n] == b.steps[:n]) & (a.rejects[:n] == b.rejects[:n])


def ragged_block(gpu, problem, steps, cells=(1000, 20000, 200000), duration=1.0, lds_slots=0, tune=1.0):
    """Three mechanisms (60/200, 40/130, 24/70) with 1k / 20k / 200k cells in one
    persistent launch per operator-split step, the steps captured in one HIP
    graph; beside it the same steps as eager launches and as one mechanism at a
    time (three kernels, each over its own cells only)."""
    names = ("c3a", "c3b", "c3c")
    meth, ctl = problem.method(2), problem.control()
    mechs, z0s = [], []
    for name, n in zip(names, cells):
        sp = problem.spec(name)
        m = problem.Mechanism.synthetic(sp)
        mechs.append(m)
        z0s.append(problem.Zones(m.n, m.R, n).fill_synthetic(sp, m.params))
    gm = run_workload.gm
    need = sum(z.block.nbytes for z in z0s) + (64 << 20)
    if gm is None or need > gm.info.arena_bytes:
        gm = run_workload.gm = gpu.GpuModule(arena_bytes=need)
    gm.arena_reset()
    gm.set_grid(0)
    gm.set_gen_options(lds_slots=lds_slots)
    dzs = [gpu.DeviceZones(gm, z.copy()) for z in z0s]

    def reset():
        for d, z in zip(dzs, z0s):
            d.z.block[:] = z.block
            d.upload()

    fpa = [m.flops_per_attempt() for m in mechs]

    def totals():
        for d in dzs:
            d.download()
        acc = sum(float(d.z.steps[:d.z.count].sum()) for d in dzs)
        att = acc + sum(float(d.z.rejects[:d.z.count].sum()) for d in dzs)
        totals.flops = sum(f * float(d.z.steps[:d.z.count].sum() + d.z.rejects[:d.z.count].sum())
                           for f, d in zip(fpa, dzs))
        return acc, att, [d.z.steps[:d.z.count].copy() for d in dzs]

    out = {"workload": f"{names[0]} 60/200 x {cells[0]}, {names[1]} 40/130 x {cells[1]}, {names[2]} 24/70 x "
                       f"{cells[2]} cells; {steps} x {duration:g} s operator-split, Rodas3, rtol 1e-6 atol 1e-12",
           "steps": steps}
    # one mechanism at a time
    reset()
    gm.lib.gpu_sync()
    t0 = time.perf_counter()
    for m, d in zip(mechs, dzs):
        gm.set_mech(m, meth, ctl)
        for _ in range(steps):
            d.advance(duration)
    gm.lib.gpu_sync()
    sep_wall = time.perf_counter() - t0
    sep_acc, _, sep_steps = totals()
    # one persistent launch per step, eager
    gm.set_mechs(mechs, meth, ctl)
    reset()
    gm.lib.gpu_sync()
    t0 = time.perf_counter()
    kms = [gm.advance_multi(dzs, duration) for _ in range(steps)]
    gm.lib.gpu_sync()
    eager_wall = time.perf_counter() - t0
    eager_acc, eager_att, eager_steps = totals()
    # the same steps captured in one graph
    def graph_run():
        reset()
        gid = gm.capture_multi(dzs, duration, 1, steps)
        gm.lib.gpu_sync()
        t0 = time.perf_counter()
        gm.replay(gid)
        gm.lib.gpu_sync()
        wall = time.perf_counter() - t0
        gm.destroy_graph(gid)
        return (wall,) + totals()

    untuned_wall, untuned_acc, _, untuned_steps = graph_run()
    tuned = None
    if tune > 0.0:
        # simk_kin_autotune_multi on a copy of the whole workload, one advance
        # of `tune` seconds per candidate (a sample of 65,536 cells per
        # mechanism picked settings slower at full size: its slowest cells
        # dominate)
        reset()
        t0 = time.perf_counter()
        tuned = gm.autotune_multi(dzs, duration=tune, sample_cells=0).as_dict()
        tuned["tuning_wall_s"] = time.perf_counter() - t0
    graph_wall, graph_acc, graph_att, graph_steps = graph_run()
    graph_flops = totals.flops
    # the same steps in FP32 factor storage (precision 3), one launch per
    # step, judged by the accuracy criterion against the FP64 run just made
    ref = [d.z.copy() for d in dzs]
    reset()
    gm.lib.gpu_sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        gm.advance_multi(dzs, duration, precision=3)
    gm.lib.gpu_sync()
    s32_wall = time.perf_counter() - t0
    s32_acc, _, _ = totals()
    s32_name = gm.kin_info().kernel_name.decode()
    errs = np.concatenate([cell_errors(d.z, r, ctl.atol, ctl.rtol, d.z.count) for d, r in zip(dzs, ref)])
    s32 = {"value": s32_acc / s32_wall, "vs_fp64_eager": (s32_acc / s32_wall) / (eager_acc / eager_wall),
           "kernel": s32_name, "cell_p99_9": float(np.percentile(errs, 99.9)), "max_rel_err": float(errs.max()),
           "cells_above_1e-6": float(np.mean(errs > 1e-6)),
           "meets_criterion": bool(np.percentile(errs, 99.9) <= CRITERION_P999)}
    per_cell = np.concatenate(graph_steps)
    same = all(np.array_equal(a, b) and np.array_equal(a, c) and np.array_equal(a, d)
               for a, b, c, d in zip(graph_steps, eager_steps, sep_steps, untuned_steps))
    out.update({"value": graph_acc / graph_wall, "unit": "cell-steps/s", "graph_wall_s": graph_wall,
                "autotune": tuned, "untuned_graph_value": untuned_acc / untuned_wall,
                "eager_value": eager_acc / eager_wall, "separate_value": sep_acc / sep_wall,
                "graph_vs_separate": (graph_acc / graph_wall) / (sep_acc / sep_wall),
                "attempts_per_s": graph_att / graph_wall, "algorithmic_TFLOPs": graph_flops / graph_wall / 1e12,
                "kernel_ms_per_step": float(np.mean(kms)),
                "steps_per_cell_p5_p50_p95": [float(np.percentile(per_cell, q)) for q in (5, 50, 95)],
                "identical_step_counts_all_modes": bool(same), "fp32store": s32})
    return out


def team_sweep(gpu, problem, cells=16384, duration=0.2):
    """One advance of the 48/150 workload's first `cells` cells per lane
    layout: one cell per lane, then teams of 64, 32, 16 and 8 lanes per cell
    (gpu_kin_set_mode(1) with simk_kin_set_team), from the same state."""
    sp = problem.spec("c2")
    mech = problem.Mechanism.synthetic(sp)
    meth, ctl = problem.method(2), problem.control()
    z0 = problem.Zones(mech.n, mech.R, cells).fill_synthetic(sp, mech.params)
    gm = gpu.GpuModule(arena_bytes=z0.block.nbytes + (64 << 20))
    gm.set_kernel(0)
    out = []
    for team in (0, 64, 32, 16, 8):
        gm.arena_reset()
        gm.lib.gpu_kin_set_mode(1 if team else 0)
        if team:
            gm.lib.simk_kin_set_team(team)
        gm.set_mech(mech, meth, ctl)
        dz = gpu.DeviceZones(gm, z0.copy())
        dz.upload()
        ms = dz.advance(duration)
        dz.download()
        att = float(dz.z.steps[:cells].sum() + dz.z.rejects[:cells].sum())
        out.append({"lanes_per_cell": team or 1, "kernel_ms": ms, "attempts": att,
                    "kernel": gm.kin_info().kernel_name.decode()})
    gm.lib.gpu_kin_set_mode(0)
    return {"workload": f"48 species / 150 reactions, {cells} cells x {duration:g} s, FP64", "layouts": out}


def advect_block(species=48, channels=512, nodes=4096, substeps=4, reps=50):
    """The NFT upwind row update (simk_upwind_f64) on a [species][channels *
    nodes] f64 field of 805 MB, beyond the 256 MB Infinity Cache. One launch
    reads and writes each row once whatever the substep count (substeps run
    on registers and LDS), so its algorithmic bytes are 2 x field + courant
    + inflow. Wall time over `reps` back-to-back launches, synchronized."""
    gm = run_workload.gm
    gm.arena_reset()
    cells = channels * nodes
    field = species * cells * 8
    d_y, d_c, d_in = gm.alloc(field), gm.alloc(cells * 8), gm.alloc(species * 8)
    rng = np.random.default_rng(11)
    gm.to_device(d_y, rng.random(species * cells))
    gm.to_device(d_c, 0.2 + 0.6 * rng.random(cells))
    gm.to_device(d_in, rng.random(species))
    lib = gm.lib

    def launch():
        gm._check(lib.simk_upwind_f64(d_y, cells, species, channels, nodes, d_c, substeps, d_in), "simk_upwind_f64")

    for _ in range(3):
        launch()
    lib.gpu_sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        launch()
    lib.gpu_sync()
    per = (time.perf_counter() - t0) / reps
    algo = 2 * field + cells * 8 + species * 8
    # a plain device copy of half the field onto the other half moves the same
    # bytes (read + write): the bandwidth a streaming kernel reaches here
    half = field // 2
    cms = []
    for _ in range(10):
        ms = C.c_float()
        gm._check(lib.simk_copy_device(d_y + half, d_y, half, C.byref(ms)), "simk_copy_device")
        cms.append(ms.value)
    copy_gbs = 2 * half / (float(np.median(cms)) * 1e-3) / 1e9
    return {"workload": f"{species} species x {channels} channels x {nodes} nodes f64, {substeps} substeps per launch",
            "ms_per_launch": per * 1e3, "node_substeps_per_s": species * cells * substeps / per,
            "roofline": {"bound": "hbm", "achieved": algo / per / 1e9, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                         "frac": algo / per / 1e9 / PEAK_HBM_GBS, "algorithmic_bytes": algo},
            "copy_kernel_GBps": copy_gbs, "frac_of_copy": algo / per / 1e9 / copy_gbs}


def dbd_block(gpu, problem, zones=65536, steps=5, dt=1e-3):
    """The DBD reactor workload: SimStack's humid-air mechanism (17 species, 44
    reactions) through the reference ABI, the reference's zone operating points,
    pulsed discharge, `steps` master steps of dt with the host preparing each
    step's pulses (not timed). Kernel time per step for FP64 lanes and for the
    reference's default f32 delta lanes (generated kernel, and the table walk
    for comparison). Steps after the first take the zones longest first by
    the previous step's attempts (the module's default queue policy; a
    continuing simulation, so the previous step is a real estimate); the
    first step runs in natural order and is reported beside the mean."""
    from simstack_amd import _ffi, dbd
    air = dbd.AirMechanism(caps=_ffi.DEFAULT_CAPS)
    gm = run_workload.gm
    meth, ctl = problem.method(2), dbd.default_control()
    src = dbd.source_defaults()
    out = {"workload": f"DBD humid air 17/44, {zones} zones (T 300-340 K, 1 atm, 0-3 % H2O, 5-12.5 kHz), pulsed, "
                       f"{steps} x {dt * 1e3:g} ms, Rodas3, rtol {ctl.rtol:g} atol {ctl.atol:g}"}
    for name, precision, kind in (("fp64", 0, 1), ("f32_delta", 1, 1), ("f32_delta_table_walk", 1, 2)):
        z = dbd.reference_zones(air, zones)
        gm.arena_reset()
        gm.set_kernel(kind)
        gm.set_queue_policy(1)
        gm._check(gm.lib.gpu_kin_set_mech(air.bytes.ctypes.data, C.byref(meth), C.byref(ctl)), "gpu_kin_set_mech")
        dz = gpu.DeviceZones(gm, z)
        kms, ordered = [], []
        for step in range(steps):
            z.prepare_source(src, dbd.PULSED, step * dt, (step + 1) * dt)
            dz.upload()
            kms.append(dz.advance(dt, precision=precision))
            ordered.append(int(gm.kin_info().ordered))
            dz.download(mutable_only=True)
            z.commit_pulses()
        gm.set_kernel(0)
        acc = float(z.steps[:zones].sum())
        out[name] = {"kernel_ms_per_step": float(np.mean(kms)), "cell_steps_per_s": acc / (sum(kms) / 1e3),
                     "kernel_ms_first_step": float(kms[0]),
                     "kernel_ms_later_steps": float(np.mean(kms[1:])) if len(kms) > 1 else None,
                     "steps_cost_ordered": int(sum(ordered)),
                     "failed_zones": int(z.failed[:zones].sum()), "kernel": gm.kin_info().kernel_name.decode()}
    return out


def serving_run(gpu, problem, rate, window, steps=20, substeps=4, dt=1e-3, seed=5):
    """One serving measurement: Poisson arrivals at `rate` over `window`
    seconds through simk_serve_*, run until every request completed."""
    from simstack_amd import serving
    if serving_run.pools is None:
        serving_run.pools = serving.Pools()
    gm = run_workload.gm
    if gm is None:
        gm = run_workload.gm = gpu.GpuModule(arena_bytes=1 << 28)
    arr = serving.Arrivals(rate=rate, window=window, seed=seed)
    r = serving.run(gm, serving_run.pools, arr, steps=steps, substeps=substeps, dt=dt)
    lat, st = r["latency"], r["stats"]
    busy_s = (st.advect_ms + st.chem_ms + st.copy_ms) / 1e3
    return {"rate_req_per_s": rate, "window_s": window, "requests": r["requests"], "cells": r["cells"],
            "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
            "max_latency_s": float(lat.max()), "wall_s": r["wall"], "completed_req_per_s": r["requests"] / r["wall"],
            "cell_steps_per_s": r["accepted"] / r["wall"], "attempts_per_s": r["attempts"] / r["wall"],
            "algorithmic_TFLOPs": r["flops"] / r["wall"] / 1e12,
            "chem_algorithmic_TFLOPs": r["flops"] / (st.chem_ms * 1e-3) / 1e12 if st.chem_ms else None,
            "failed_cells": r["failed"], "iterations": int(st.iterations),
            "peak_admitted_GB": st.peak_admitted_bytes / 1e9, "gpu_busy_fraction": busy_s / r["wall"],
            "advect_ms": st.advect_ms, "chem_ms": st.chem_ms, "exposed_copy_ms": st.copy_ms}


serving_run.pools = None


def serving_block(gpu, problem, rate=200.0, window=0.25, sustained_rate=3.0, sustained_window=10.0):
    """The serving loop (the ragged batch's three mechanisms, Poisson arrivals,
    cells log-uniform 64-262,144 in 4,096-node channels, 20 operator-split
    steps of upwind advection then 1 ms of chemistry per request). `sustained`
    is a rate the GPU keeps up with over a 10-s window (the serving figure:
    latency is service time); the overloaded rate (200 req/s, far beyond
    capacity) shows queueing."""
    out = {"workload": "Poisson arrivals, cells log-uniform 64-262144 in 4096-node channels, mechanism uniform over "
                       "60/200, 40/130, 24/70; 20 x (upwind advection 4 substeps, then chemistry 0.001 s)"}
    if sustained_rate > 0.0:
        out["sustained"] = serving_run(gpu, problem, sustained_rate, sustained_window, seed=7)
    out["overload"] = serving_run(gpu, problem, rate, window)
    return out


CRITERION_P999 = 1e-6   # p99.9 over cells of the cell's max relative error vs FP64 (entries > atol/rtol)
NONE = 0xFFFFFFFF


def cell_errors(got, ref, atol, rtol, n):
    """Per-cell max relative error |got - ref| / max(|got|, |ref|) over the
    species whose FP64 value exceeds atol/rtol (test_kinetics.cpp:35-38)."""
    a, b = got.y[:, :n], ref.y[:, :n]
    d = np.abs(a - b)
    s = np.maximum(np.abs(a), np.abs(b))
    r = np.divide(d, s, out=np.zeros_like(d), where=s > 0)
    return np.where(b > atol / rtol, r, 0.0).max(axis=0)


def accuracy(got, ref, ctl, n):
    e = cell_errors(got, ref, ctl.atol, ctl.rtol, n)
    same = same_counts(got, ref, n)
    return {"max_rel_err": float(e.max()), "cell_p99_9": float(np.percentile(e, 99.9)),
            "cell_p99": float(np.percentile(e, 99)), "cells_above_1e-6": float(np.mean(e > 1e-6)),
            "cells_same_step_counts": float(same.mean()),
            "meets_criterion": bool(np.percentile(e, 99.9) <= CRITERION_P999),
            "max_rel_err_by_species_group": rel_err_groups(got.y[:, :n], ref.y[:, :n], ctl.atol, ctl.rtol, same)}


def host_cpu():
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count()


def traffic_for(name, attempts_per_launch):
    """Counted HBM traffic of one launch: bytes per attempt from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes (profiles/traffic_<name>.json,
    calibrated by the copy kernel) times this launch's attempts."""
    path = os.path.join(ROOT, "profiles", f"traffic_{name}.json")
    if not os.path.exists(path):
        return None, None
    t = json.load(open(path))
    per = t.get("bytes_per_attempt")
    if not per:
        return None, None
    return per * attempts_per_launch, per


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cells", type=int, default=262144)
    ap.add_argument("--fp64-cells", type=int, default=65536, help="the FP64 throughput workload's cells (extra block)")
    ap.add_argument("--cpu-cells", type=int, default=8192)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="FP64 only (no mixed-precision variants, no floor run)")
    ap.add_argument("--no-fp64-block", action="store_true", help="skip the 65,536-cell FP64 block")
    ap.add_argument("--kernel", type=int, default=0, help="0 auto, 1 generated, 2 table walk")
    ap.add_argument("--separate", action="store_true",
                    help="one gpu_kin_advance launch per step instead of one fused simk_kin_advance_n launch")
    ap.add_argument("--grid", type=int, default=0, help="persistent waves (0: resident waves)")
    ap.add_argument("--reg-window", type=int, default=0, help="generator register window (0: default)")
    ap.add_argument("--waves-per-eu", type=int, default=0)
    ap.add_argument("--batch", type=int, default=0, help="generator load-pipeline depth (0: default)")
    ap.add_argument("--lds-slots", type=int, default=0,
                    help="LU entries per lane held in LDS (0: default, -1: none)")
    ap.add_argument("--variants", default="norefine,1sweep,2sweeps,fp32store",
                    help="mixed-precision variants to run: norefine, 1sweep, 2sweeps (FP32 LU and solves), "
                         "fp32store (FP32 factor storage, FP64 pivots and arithmetic)")
    ap.add_argument("--tune", type=float, default=0.0,
                    help="autotune FP64 on the workload for this many simulated seconds before timing (0: defaults)")
    ap.add_argument("--no-ragged", action="store_true", help="skip the three-mechanism ragged batch")
    ap.add_argument("--ragged-steps", type=int, default=5)
    ap.add_argument("--no-serve", action="store_true", help="skip the serving loop")
    ap.add_argument("--no-dbd", action="store_true", help="skip the DBD reactor workload")
    ap.add_argument("--no-advect", action="store_true", help="skip the upwind advection bandwidth block")
    ap.add_argument("--serve-window", type=float, default=0.25, help="seconds of Poisson arrivals, overload rate")
    ap.add_argument("--serve-rate", type=float, default=200.0, help="the overload arrival rate")
    ap.add_argument("--serve-sustained", type=float, default=3.0,
                    help="the highest arrival rate the GPU sustains (req/s; profiles/serving_latency_curve.md)")
    ap.add_argument("--team-sweep", action="store_true",
                    help="only time one advance per lane layout (one lane, teams of 64/32/16/8 lanes per cell)")
    ap.add_argument("--serve-curve", default="",
                    help="comma-separated arrival rates: latency at each over a 10-s window (adds minutes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported (cells shard with no "
              "communication: each GPU would run the same kernel on its own share)", file=sys.stderr)
        sys.exit(2)

    from simstack_amd import gpu, problem
    if args.team_sweep:
        print(json.dumps({"team_sweep": team_sweep(gpu, problem)}), flush=True)
        return
    sp = problem.spec("c2")
    mech = problem.Mechanism.synthetic(sp)
    meth, ctl = problem.method(2), problem.control()
    knobs = dict(kernel=args.kernel, grid=args.grid, reg_window=args.reg_window, waves_per_eu=args.waves_per_eu,
                 batch=args.batch, lds_slots=args.lds_slots & 0xFFFFFFFF)
    n = args.cells
    rate = lambda r: r["accepted"] / r["wall"]

    def run(precision=0, sweeps=0, perturb=False, cells=n, tune=0.0):
        return run_workload(gpu, problem, mech, sp, meth, ctl, cells, args.steps, args.warmup, precision, sweeps,
                            args.separate, tune=tune if not args.grid else 0.0, perturb=perturb, **knobs)

    def roofline(r, precision, sweeps, traffic_name):
        kms = r["kernel_ms"]
        avg_ms = float(np.mean(kms))
        att = r["attempts"] / len(kms)
        if precision in (0, 3):
            f64, f32 = mech.flops_per_attempt(), 0
        else:
            f64, f32 = mech.flops_mixed(0 if sweeps == NONE else sweeps)
        flops = att * (f64 + f32)
        t_min = att * (f64 / (PEAK_FP64_VALU_TFLOPS * 1e12) + f32 / (PEAK_FP32_VALU_TFLOPS * 1e12))
        t = avg_ms * 1e-3
        traffic, per = traffic_for(traffic_name, att)
        out = {"bound": "valu", "achieved": flops / t / 1e12,
               "peak": PEAK_FP64_VALU_TFLOPS if precision == 0 else flops / t_min / 1e12, "unit": "TFLOP/s",
               "frac": t_min / t, "traffic": traffic, "kernel_ms_avg": avg_ms, "attempts_per_launch": att,
               "flops_per_attempt_fp64": f64, "flops_per_attempt_fp32": f32}
        if traffic:
            out.update({"traffic_bytes_per_attempt": per, "hbm_GBps": traffic / t / 1e9,
                        "hbm_frac": traffic / t / 1e9 / PEAK_HBM_GBS})
        return out

    # the workload the metric is quoted on: FP64 (bit-exact), then mixed precision
    f64r = run(0, 0, tune=args.tune)
    variants = {"fp64": {"value": rate(f64r), "wall_s": f64r["wall"], "vs_fp64": 1.0, "dtype": "f64",
                         "kernel": f64r["kernel"],
                         "waves": f64r["waves"], "failed_cells": f64r["failed"], "autotune": f64r["tuned"],
                         "accuracy": {"max_rel_err": 0.0, "cell_p99_9": 0.0, "cells_same_step_counts": 1.0,
                                      "meets_criterion": True, "note": "bit-identical to the FP64 oracle"},
                         "roofline": roofline(f64r, 0, 0, "c4_fp64")}}
    floor = None
    if not args.no_mixed:
        pr = run(0, 0, perturb=True)
        floor = accuracy(pr["zones"], f64r["zones"], ctl, n)
        floor.pop("meets_criterion")
        del pr
        sweeps_of = {"norefine": NONE, "1sweep": 1, "2sweeps": 2, "fp32store": 0}
        for name in [v for v in args.variants.split(",") if v]:
            sw = sweeps_of[name]
            prec = 3 if name == "fp32store" else 2
            mr = run(prec, sw)
            variants["mixed_" + name] = {
                "value": rate(mr), "wall_s": mr["wall"], "vs_fp64": rate(mr) / rate(f64r),
                "dtype": "f32 LU / f64 state" if prec == 2 else "f32 LU storage / f64 pivots and arithmetic",
                "kernel": mr["kernel"], "waves": mr["waves"], "failed_cells": mr["failed"],
                "steps_ratio": mr["accepted"] / f64r["accepted"],
                "accuracy": accuracy(mr["zones"], f64r["zones"], ctl, n),
                "roofline": roofline(mr, prec, sw, f"c4_mixed_{name}")}
            del mr
    best_name = max((k for k, v in variants.items() if v["accuracy"]["meets_criterion"]),
                    key=lambda k: variants[k]["value"])
    best = variants[best_name]
    ms_per_step = best["wall_s"] * 1e3 / args.steps
    out = {
        "metric": "cell-steps/s at rtol 1e-6 (mixed vs FP64, 1 GPU); p50/p99 solve latency",
        "value": best["value"],
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": best["dtype"],
        "data": "synthetic (seeded mechanism and cells)",
        "config": {"workload": f"48 species / 150 reactions, {n:,} cells, {args.steps} x 1 s operator-split, "
                               "Rodas3, rtol 1e-6 atol 1e-12; the fastest variant whose p99.9-cell max relative "
                               "error vs the FP64 oracle is <= 1e-6", "cells": n, "species": mech.n,
                   "reactions": mech.R, "nnz": mech.nnz, "lu_div": mech.lu_div, "lu_sub": mech.lu_sub,
                   "variant": best_name, "kernel": best["kernel"], "waves": best["waves"],
                   "parallelism": "1 GPU, one cell per lane, persistent lane-refilling kernel"},
        "accepted_steps_fp64": f64r["accepted"],
        "attempts_per_s_fp64": f64r["attempts"] / f64r["wall"],
        "failed_cells": best["failed_cells"],
        "roofline": best["roofline"],
        "criterion": {"statistic": "p99.9 over cells of the cell's max relative error |x - x64| / max(|x|, |x64|) "
                                   "over species above atol/rtol, vs the FP64 path (bit-identical to the oracle)",
                      "bound": CRITERION_P999,
                      "fp64_one_ulp_floor": floor},
        "variants": variants,
        "mixed_vs_fp64": {k: v["vs_fp64"] for k, v in variants.items() if k != "fp64"},
    }
    del f64r
    if not args.no_fp64_block:
        r2 = run(0, 0, cells=args.fp64_cells)
        out["fp64_65536"] = {"workload": f"48 species / 150 reactions, {args.fp64_cells:,} cells, {args.steps} x 1 s "
                                         "fused, FP64 (bit-identical to the oracle)",
                             "value": rate(r2), "attempts_per_s": r2["attempts"] / r2["wall"],
                             "kernel": r2["kernel"], "waves": r2["waves"], "failed_cells": r2["failed"],
                             "roofline": roofline(r2, 0, 0, "c2_fp64")}
        del r2
    if not args.no_ragged:
        out["ragged"] = ragged_block(gpu, problem, args.ragged_steps)
    if not args.no_serve:
        out["serving"] = serving_block(gpu, problem, rate=args.serve_rate, window=args.serve_window,
                                       sustained_rate=args.serve_sustained)
        head = out["serving"].get("sustained", out["serving"]["overload"])
        out["p50_latency_s"] = head["p50_latency_s"]
        out["p99_latency_s"] = head["p99_latency_s"]
        out["latency_at_req_per_s"] = head["rate_req_per_s"]
    if args.serve_curve:
        out["serving_curve"] = []
        for r in (x for x in args.serve_curve.split(",") if x):
            out["serving_curve"].append(serving_run(gpu, problem, float(r), 10.0, seed=7))
            c = out["serving_curve"][-1]
            print(f"[serving] {c['rate_req_per_s']:g} req/s: p50 {c['p50_latency_s']:.2f} s, p99 "
                  f"{c['p99_latency_s']:.2f} s", file=sys.stderr, flush=True)
    if not args.no_dbd:
        out["dbd"] = dbd_block(gpu, problem)
    if not args.no_advect:
        out["advect"] = advect_block()
    if not args.no_cpu:
        model, cpus = host_cpu()
        out["cpu_baseline"] = cpu_baseline(mech, meth, ctl, sp, args.cpu_cells, 1.0, repeats=args.steps)
        out["cpu_baseline"].update({"host_cpu_model": model, "host_logical_cpus": cpus})
        # the same path on one core, on a smaller sample
        out["cpu_baseline_1core"] = cpu_baseline(mech, meth, ctl, sp, max(args.cpu_cells // 16, 64), 1.0,
                                                 repeats=args.steps, threads=1)
        # the oracle's scalar FP64 restatement on one core, the same sample
        out["cpu_oracle_1core"] = cpu_baseline(mech, meth, ctl, sp, max(args.cpu_cells // 16, 64), 1.0,
                                               repeats=args.steps, threads=1, port=True)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
