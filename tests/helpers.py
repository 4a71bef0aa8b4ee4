"""Shared test plumbing: seeded problems, GPU and oracle runs, block comparison."""
import os
import time

import numpy as np

from simstack_amd import _ffi, problem
import oracle_ffi


def make_problem(name, cells, first_cell=0, caps=_ffi.RAISED_CAPS, **spec_over):
    sp = problem.spec(name, **spec_over)
    mech = problem.Mechanism.synthetic(sp, caps=caps)
    z = problem.Zones(mech.n, mech.R, cells).fill_synthetic(sp, mech.params, first_cell)
    return sp, mech, z


def add_pulses(z, seed=7, frac=0.5, max_count=3, duration=1.0):
    """Random pulse schedules (kin_advance segments) on a fraction of the cells."""
    rng = np.random.default_rng(seed)
    for i in range(z.count):
        if rng.random() < frac:
            z.period[i] = duration / (1 + rng.integers(1, 6))
            z.pulse_first[i] = rng.random() * z.period[i]
            z.pulse_count[i] = float(rng.integers(1, max_count + 1))
            z.pulse[:, i] = z.y[:, i] * rng.random() * 0.5
    z.period[z.count:] = z.period[0]
    z.pulse_first[z.count:] = z.pulse_first[0]
    z.pulse_count[z.count:] = z.pulse_count[0]
    z.pulse[:, z.count:] = z.pulse[:, :1]


def gpu_run(gm, mech, meth, ctl, z, durations, kernel=0, use_default_abi=False):
    """Uploads z, advances on the device once per duration, downloads; returns kernel ms list."""
    from simstack_amd import gpu
    import ctypes as C
    gm.arena_reset()
    gm.set_kernel(kernel)
    if use_default_abi:
        rc = gm.lib.gpu_kin_set_mech(mech.bytes.ctypes.data, C.byref(meth), C.byref(ctl))
        gm._check(rc, "gpu_kin_set_mech")
    else:
        gm.set_mech(mech, meth, ctl)
    dz = gpu.DeviceZones(gm, z)
    dz.upload()
    ms = []
    for d in durations:
        ms.append(dz.advance(d))
    dz.download(mutable_only=False)
    return ms


def oracle_run(mech, meth, ctl, z, durations, lanes=None, threads=16):
    for d in durations:
        oracle_ffi.oracle_advance(mech, meth, ctl, z, d, lanes=lanes, threads=threads)


SOLVER_ROWS = ("y", "h", "steps", "rejects", "pulses", "failed")


def compare_bits(a, b, lanes=None):
    """Per-row count of lanes whose f64 bits differ (0 everywhere = bit-identical)."""
    out = {}
    for name in SOLVER_ROWS:
        x, y = getattr(a, name), getattr(b, name)
        if lanes is not None:
            x, y = x[..., lanes], y[..., lanes]
        else:
            x, y = x[..., :a.count], y[..., :a.count]
        diff = x.view(np.uint64) != y.view(np.uint64)
        out[name] = int(diff.any(axis=0).sum()) if diff.ndim == 2 else int(diff.sum())
    return out


def max_rel_err(a, b, lanes=None, floor=0.0):
    x, y = a.y[:, :a.count], b.y[:, :b.count]
    if lanes is not None:
        x, y = a.y[:, lanes], b.y[:, lanes]
    d = np.abs(x - y)
    s = np.maximum(np.abs(x), np.abs(y))
    mask = s > floor
    return float((d[mask] / s[mask]).max()) if mask.any() else 0.0


def _threads():
    return max(1, min(int(os.environ.get("OMP_NUM_THREADS") or 0) or os.cpu_count() or 1, 32))


def checker(mech, meth_id, ctl, z, durations):
    """Advances z in place once per duration with the CPU checker; returns its name."""
    t0 = time.perf_counter()
    if oracle_ffi.have_ref():
        mb = mech.rebuild(oracle_ffi.ref_caps()).bytes
        meth_b = oracle_ffi.ref_method_bytes(meth_id)
        for d in durations:
            oracle_ffi.ref_advance(mb, meth_b, ctl, z, d, z.padded, width=4, threads=_threads())
        kind = "reference f64x4"
    else:
        for d in durations:
            oracle_ffi.oracle_advance(mech, problem.method(meth_id), ctl, z, d, threads=_threads())
        kind = "oracle"
    print(f"[checker] {kind}, {z.count} cells x {len(durations)} advances, {_threads()} threads, "
          f"{time.perf_counter() - t0:.1f} s", flush=True)
    return kind
