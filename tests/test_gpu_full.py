"""Whole-population parity at the workloads' own sizes: every cell of the
device run against the CPU checker, bit for bit (y, h, accepted and rejected
steps, pulses, failed), as the reference's cross-path test compares every
zone with memcmp (src/tests/test_kinetics.cpp:1106-1128).

The checker is the reference's own f64x4 AVX2 path (oracle/_ref, compiled in
place from the reference's sources) when it is present, else the oracle's
scalar restatement; the two are bit-identical (tests/test_oracle.py). Both run
threaded on the host's CPU share.

- 48 species / 150 reactions, 65,536 cells x 10 operator-split advances of
  1 s (the FP64 throughput workload), fused in one launch on the device;
- the same mechanism at 262,144 cells (the mixed-precision comparison's FP64
  run): the last 65,536 cells are checked;
- the ragged batch: 60/200, 40/130 and 24/70 at 20,000 cells each, one
  multi-mechanism launch of 1 s.
"""
import numpy as np
import pytest

from simstack_amd import gpu, problem
from helpers import checker, compare_bits, make_problem

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]

ZERO = {k: 0 for k in ("y", "h", "steps", "rejects", "pulses", "failed")}


def _device_fused(gm, mech, meth, ctl, z, duration, repeats):
    gm.arena_reset()
    gm.set_kernel(0)
    gm.set_grid(0)
    gm.set_mech(mech, meth, ctl)
    dz = gpu.DeviceZones(gm, z)
    dz.upload()
    ms = dz.advance(duration, repeats=repeats)
    dz.download(mutable_only=False)
    print(f"[device] {gm.kin_info().kernel_name.decode()}: {z.count} cells x {repeats} x {duration:g} s, "
          f"{ms:.1f} ms", flush=True)


def test_fp64_48x150_65536_cells_10s_every_cell(gm):
    print("[case] 48/150, 65,536 cells, 10 x 1 s fused, every cell", flush=True)
    sp, mech, z = make_problem("c2", 65536)
    meth, ctl = problem.method(2), problem.control()
    a, b = z.copy(), z.copy()
    _device_fused(gm, mech, meth, ctl, a, 1.0, 10)
    checker(mech, 2, ctl, b, [1.0] * 10)
    assert compare_bits(a, b) == ZERO
    assert b.failed[:65536].sum() == 0 and b.steps[:65536].min() > 100


def test_fp64_48x150_262144_cells_last_65536_checked(gm):
    print("[case] 48/150, 262,144 cells, 10 x 1 s fused, cells 196,608.. checked", flush=True)
    sp, mech, z = make_problem("c2", 262144)
    meth, ctl = problem.method(2), problem.control()
    a = z.copy()
    _device_fused(gm, mech, meth, ctl, a, 1.0, 10)
    del z
    first = 262144 - 65536
    _, _, b = make_problem("c2", 65536, first_cell=first)
    checker(mech, 2, ctl, b, [1.0] * 10)
    got = a.subset(np.arange(first, 262144))
    assert compare_bits(got, b) == ZERO


def test_ragged_three_mechanisms_20000_cells_each(gm):
    print("[case] ragged 60/200 + 40/130 + 24/70, 20,000 cells each, one launch of 1 s", flush=True)
    blocks = [make_problem(name, 20000) for name in ("c3a", "c3b", "c3c")]
    meth, ctl = problem.method(2), problem.control()
    gm.arena_reset()
    gm.set_grid(0)
    gm.set_mechs([m for _, m, _ in blocks], meth, ctl)
    work = [z.copy() for _, _, z in blocks]
    dzs = [gpu.DeviceZones(gm, z) for z in work]
    for d in dzs:
        d.upload()
    ms = gm.advance_multi(dzs, 1.0)
    for d in dzs:
        d.download(mutable_only=False)
    print(f"[device] multi-mechanism kernel {ms:.1f} ms", flush=True)
    for (_, mech, z), got in zip(blocks, work):
        want = z.copy()
        checker(mech, 2, ctl, want, [1.0])
        assert compare_bits(got, want) == ZERO
