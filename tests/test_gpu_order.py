"""Queue order by the previous advance's per-cell attempts
(simk_kin_set_queue_policy 1, the default): consecutive advances of one
block take its cells longest first. The order decides only which lane runs a
cell and when, so every bit equals the natural-order run (itself checked
against the oracle elsewhere). The grid is cut to 8 waves (512 lanes), so
the blocks below have more cells than resident lanes; the order applies to
blocks of at least 512 cells."""
import ctypes as C

import numpy as np
import pytest

from simstack_amd import _ffi, dbd, gpu, problem
from helpers import make_problem

pytestmark = pytest.mark.gpu


def _advances(gm, mech, meth, ctl, z, durations, policy, lanes=0, precision=0):
    gm.arena_reset()
    gm.set_grid(8)
    gm.set_queue_policy(policy)
    gm.set_mech(mech, meth, ctl)
    dz = gpu.DeviceZones(gm, z)
    dz.upload()
    ordered = []
    for d in durations:
        dz.advance(d, lane_count=lanes, precision=precision)
        ordered.append(gm.kin_info().ordered)
    dz.download(mutable_only=False)
    return ordered


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "f32-delta"])
def test_cost_order_keeps_bits(gm, precision):
    print(f"[case] 48/150, 3,000 cells, 4 advances, cost order vs natural, precision {precision}", flush=True)
    try:
        sp, mech, z = make_problem("c2", 3000)
        meth, ctl = problem.method(2), problem.control()
        a, b = z.copy(), z.copy()
        ordered = _advances(gm, mech, meth, ctl, a, [0.05] * 4, 1, precision=precision)
        natural = _advances(gm, mech, meth, ctl, b, [0.05] * 4, 0, precision=precision)
        assert ordered == [0, 1, 1, 1] and natural == [0, 0, 0, 0]
        assert a.block.tobytes() == b.block.tobytes()
    finally:
        gm.set_grid(0)
        gm.set_queue_policy(1)


def test_cost_order_forgets_other_blocks(gm):
    """A different cell count, an arena reset or a new mechanism starts again
    in natural order; a caller's own order (simk_kin_set_order) wins."""
    print("[case] cost order bookkeeping", flush=True)
    try:
        sp, mech, z = make_problem("c3c", 2000)
        meth, ctl = problem.method(2), problem.control()
        gm.arena_reset()
        gm.set_grid(8)
        gm.set_queue_policy(1)
        gm.set_mech(mech, meth, ctl)
        dz = gpu.DeviceZones(gm, z)
        dz.upload()
        seen = []
        for lanes in (2000, 2000, 1500, 1500, 2000):
            dz.advance(0.02, lane_count=lanes)
            seen.append(gm.kin_info().ordered)
        assert seen == [0, 1, 0, 1, 0]
        order = gm.alloc(4 * 2000)
        gm.to_device(order, np.arange(2000, dtype=np.uint32)[::-1].copy())
        gm._check(gm.lib.simk_kin_set_order(C.c_void_p(order)), "simk_kin_set_order")
        dz.advance(0.02)
        assert gm.kin_info().ordered == 0
        gm._check(gm.lib.simk_kin_set_order(None), "simk_kin_set_order")
        dz.advance(0.02)
        assert gm.kin_info().ordered == 1
        gm.set_mech(mech, meth, ctl)
        dz.advance(0.02)
        assert gm.kin_info().ordered == 0
        gm.arena_reset()
        dz2 = gpu.DeviceZones(gm, z)
        dz2.upload()
        dz2.advance(0.02)
        assert gm.kin_info().ordered == 0
        with pytest.raises(gpu.GpuError):
            gm.set_queue_policy(2)
    finally:
        gm.set_grid(0)
        gm.set_queue_policy(1)


def test_dbd_steps_cost_ordered_match_natural(gm):
    """Consecutive 1-ms DBD master steps (pulsed, zones at different
    operating points), cost-ordered from the second step, give the
    natural-order bits in f64 and f32 delta lanes."""
    print("[case] DBD 4,096 zones x 4 steps, cost order vs natural", flush=True)
    air = dbd.AirMechanism(caps=_ffi.DEFAULT_CAPS)
    meth, ctl = problem.method(2), dbd.default_control()
    src = dbd.source_defaults()
    try:
        for precision in (0, 1):
            blocks = []
            for policy in (1, 0):
                z = dbd.reference_zones(air, 4096)
                gm.arena_reset()
                gm.set_grid(8)
                gm.set_queue_policy(policy)
                gm._check(gm.lib.gpu_kin_set_mech(air.bytes.ctypes.data, C.byref(meth), C.byref(ctl)), "set_mech")
                dz = gpu.DeviceZones(gm, z)
                for step in range(4):
                    z.prepare_source(src, dbd.PULSED, step * 1e-3, (step + 1) * 1e-3)
                    dz.upload()
                    dz.advance(1e-3, precision=precision)
                    assert gm.kin_info().ordered == (1 if policy and step else 0)
                    dz.download(mutable_only=True)
                    z.commit_pulses()
                blocks.append(z.block.tobytes())
            assert blocks[0] == blocks[1], precision
    finally:
        gm.set_grid(0)
        gm.set_queue_policy(1)
