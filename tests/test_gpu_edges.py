"""Edge shapes of the device advance against the oracle, bit for bit: a single
cell and partial waves (FP64 generated and table-walk kernels, f32 delta
lanes, FP32 factor storage), a prefix of the block (GpuKinLaunch.lane_count below the cell count:
the cells past it keep every bit), and a zero-length advance with pulses due
at t = 0 (kin_advance applies them and takes no step, kin_solver.h:322-347)."""
import numpy as np
import pytest

import oracle_ffi
from simstack_amd import gpu, problem
from helpers import add_pulses, compare_bits, make_problem

pytestmark = pytest.mark.gpu

ZERO = {k: 0 for k in ("y", "h", "steps", "rejects", "pulses", "failed")}


def _device(gm, mech, meth, ctl, z, duration, kernel=1, precision=0, lane_count=0):
    gm.arena_reset()
    gm.set_kernel(kernel)
    gm.set_mech(mech, meth, ctl)
    dz = gpu.DeviceZones(gm, z)
    dz.upload()
    dz.advance(duration, lane_count=lane_count, precision=precision)
    dz.download(mutable_only=False)


def _oracle(mech, meth, ctl, z, duration, precision=0, lanes=None):
    adv = {1: oracle_ffi.oracle_advance_delta32, 3: oracle_ffi.oracle_advance_store32}.get(precision,
                                                                                         oracle_ffi.oracle_advance)
    adv(mech, meth, ctl, z, duration, lanes=lanes)


# Pulse schedules due after the end of the advance (the pulsed cells apply
# their pulses at the clipped segment end while the pulse-free cells of the
# same waves integrate to it) once left pulse-free f32 delta lanes with one
# recorded pulse and h = h_start: the segment's pending flag, carried as a
# lane mask across the wave's divergent loop exits, was rewritten for lanes
# waiting outside the loop. The flag is now recomputed where it is tested.
@pytest.mark.parametrize("cells", [1, 63, 65, 130])
@pytest.mark.parametrize("kernel,precision", [(1, 0), (2, 0), (1, 1), (1, 3)],
                         ids=["generated", "table", "f32delta", "fp32store"])
def test_single_cell_and_partial_waves(gm, cells, kernel, precision):
    sp, mech, z = make_problem("c3c", cells)
    add_pulses(z, seed=cells)
    meth, ctl = problem.method(2), problem.control(rtol=1e-5, atol=1e-10)
    a, b = z.copy(), z.copy()
    _device(gm, mech, meth, ctl, a, 0.05, kernel, precision)
    _oracle(mech, meth, ctl, b, 0.05, precision)
    assert compare_bits(a, b) == ZERO
    assert b.steps[:cells].min() > 0


@pytest.mark.parametrize("cells", [63, 65, 130])
def test_f32delta_partial_waves_pulses_inside_the_advance(gm, cells):
    sp, mech, z = make_problem("c3c", cells)
    add_pulses(z, seed=cells, duration=0.05)
    meth, ctl = problem.method(2), problem.control(rtol=1e-3, atol=1e-10, max_steps=1000)
    a, b = z.copy(), z.copy()
    _device(gm, mech, meth, ctl, a, 0.05, 1, 1)
    _oracle(mech, meth, ctl, b, 0.05, 1)
    assert compare_bits(a, b) == ZERO
    assert b.pulses[:cells].sum() > 0


@pytest.mark.parametrize("precision", [0, 1, 3], ids=["fp64", "f32delta", "fp32store"])
def test_prefix_lane_count_leaves_the_rest(gm, precision):
    sp, mech, z = make_problem("c3c", 200)
    meth, ctl = problem.method(2), problem.control()
    a, b = z.copy(), z.copy()
    _device(gm, mech, meth, ctl, a, 0.1, 1, precision, lane_count=77)
    _oracle(mech, meth, ctl, b, 0.1, precision, lanes=np.arange(77))
    assert compare_bits(a, b, lanes=np.arange(77)) == ZERO
    assert b.steps[:77].min() > 0
    rest = np.arange(77, 200)
    assert compare_bits(a, z, lanes=rest) == ZERO


@pytest.mark.parametrize("precision", [0, 3], ids=["fp64", "fp32store"])
def test_zero_duration_applies_due_pulses_only(gm, precision):
    sp, mech, z = make_problem("c3c", 96)
    add_pulses(z, seed=11, frac=1.0)
    z.pulse_first[:48] = 0.0
    meth, ctl = problem.method(2), problem.control()
    a, b = z.copy(), z.copy()
    _device(gm, mech, meth, ctl, a, 0.0, precision=precision)
    _oracle(mech, meth, ctl, b, 0.0, precision)
    assert compare_bits(a, b) == ZERO
    assert b.steps[:96].sum() == 0
