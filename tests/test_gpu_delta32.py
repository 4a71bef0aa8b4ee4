"""f32 delta lanes on the device (GpuKinLaunch precision 1) vs the oracle's
restatement of the reference's dbd_advance_group<f32, .., Delta = true>
(src/sim/dbd_lanes.h:28-78, src/sim/kinetics/kin_solver.h:189-347): bit for
bit on y (f64 storage after the increment is added and clamped), h, the
counters and the failure flag. This is the reference's default chemistry
precision (--precision f32). Every test runs the generated straight-line f32
kernel and the table walk."""
import ctypes as C

import numpy as np
import pytest

import oracle_ffi
from simstack_amd import _ffi, gpu, problem
from helpers import add_pulses, compare_bits, make_problem

pytestmark = pytest.mark.gpu

ZERO = {k: 0 for k in ("y", "h", "steps", "rejects", "pulses", "failed")}
KERNEL_CODE = {1: 8, 2: 5}   # kernel kind -> SimkKinInfo.kernel


@pytest.fixture(params=[1, 2], ids=["generated", "table"])
def kind(request, gm):
    gm.set_kernel(request.param)
    yield request.param
    gm.set_kernel(0)


def _gpu(gm, mech, meth, ctl, z, durations, repeats=1, default_abi=False):
    gm.arena_reset()
    if default_abi:
        gm._check(gm.lib.gpu_kin_set_mech(mech.bytes.ctypes.data, C.byref(meth), C.byref(ctl)), "set_mech")
    else:
        gm.set_mech(mech, meth, ctl)
    dz = gpu.DeviceZones(gm, z)
    dz.upload()
    for d in durations:
        dz.advance(d, precision=1, repeats=repeats)
    dz.download(mutable_only=False)
    return gm.kin_info()


def _oracle(mech, meth, ctl, z, durations, lanes=None):
    for d in durations:
        oracle_ffi.oracle_advance_delta32(mech, meth, ctl, z, d, lanes=lanes, threads=16)


@pytest.mark.parametrize("mid", [0, 1, 2], ids=["ros2", "ros3", "rodas3"])
def test_delta32_bit_exact_with_pulses(gm, kind, mid):
    """Every method, pulses on half the cells, two calls carrying h. max_steps
    bounds the cells f32 cannot converge (the reference's f32 lanes fail them
    too); the second-order method runs at rtol 1e-2."""
    sp, mech, z = make_problem("c1", 256, caps=_ffi.DEFAULT_CAPS)
    add_pulses(z, seed=mid + 3, duration=0.01)
    meth = problem.method(mid)
    ctl = problem.control(rtol=1e-2 if mid == 0 else 1e-3, atol=1e-10, max_steps=1000)
    a, b = z.copy(), z.copy()
    info = _gpu(gm, mech, meth, ctl, a, [0.01, 0.01], default_abi=True)
    _oracle(mech, meth, ctl, b, [0.01, 0.01])
    assert info.kernel == KERNEL_CODE[kind]
    assert compare_bits(a, b) == ZERO
    assert b.pulses[:256].sum() > 0 and b.steps[:256].min() > 2


def test_delta32_c2_sampled(gm, kind):
    """48/150 with pH spread: 1,024 cells, 128 checked by the oracle."""
    sp, mech, z = make_problem("c2", 1024)
    meth, ctl = problem.method(2), problem.control(rtol=1e-3, atol=1e-10, max_steps=1000)
    a = z.copy()
    _gpu(gm, mech, meth, ctl, a, [0.02])
    lanes = np.random.default_rng(4).choice(1024, 128, replace=False)
    b = z.subset(lanes)
    _oracle(mech, meth, ctl, b, [0.02])
    assert compare_bits(a.subset(lanes), b) == ZERO


def test_delta32_fused_repeats_equal_separate_calls(gm, kind):
    """simk_kin_advance_n(3) in delta lanes == three calls: each call restarts
    from the f64 storage (base re-rounded, counters added)."""
    sp, mech, z = make_problem("c3c", 500)
    meth, ctl = problem.method(2), problem.control(rtol=1e-3, atol=1e-10, max_steps=1000)
    a, b, c = z.copy(), z.copy(), z.copy()
    _gpu(gm, mech, meth, ctl, a, [0.005], repeats=3)
    _gpu(gm, mech, meth, ctl, b, [0.005, 0.005, 0.005])
    _oracle(mech, meth, ctl, c, [0.005, 0.005, 0.005])
    assert compare_bits(a, b) == ZERO
    assert compare_bits(a, c) == ZERO


def test_delta32_refill_with_few_waves(gm, kind):
    """Three resident waves for 400 cells of the 40/130 mechanism (lanes refill
    many times). max_steps bounds the cells f32 cannot converge: about 37,000
    attempts in all, at most 1,000 for one cell."""
    sp, mech, z = make_problem("c3b", 400)
    meth, ctl = problem.method(2), problem.control(rtol=1e-3, atol=1e-10, max_steps=1000)
    a, b = z.copy(), z.copy()
    try:
        gm.set_grid(3)
        _gpu(gm, mech, meth, ctl, a, [0.01])
    finally:
        gm.set_grid(0)
    _oracle(mech, meth, ctl, b, [0.01])
    assert compare_bits(a, b) == ZERO
    assert b.failed[:400].sum() > 0


@pytest.mark.parametrize("lds", [0xFFFFFFFF, 64, 0xFFFFFFFE], ids=["lds_none", "lds64", "lds_max"])
def test_delta32_generated_lds_shares_keep_bits(gm, lds):
    """The generated f32 delta kernel with its LU entirely in the workspace, 64
    entries in LDS, and as many as fit (all 407 of the 40/130 mechanism's LU,
    f32 entries of 64 lanes in 104 KB of LDS): the same bits as the oracle."""
    sp, mech, z = make_problem("c3b", 300)
    meth, ctl = problem.method(2), problem.control(rtol=1e-3, atol=1e-10, max_steps=1000)
    a, b = z.copy(), z.copy()
    try:
        gm.set_kernel(1)
        gm.set_gen_options(lds_slots=lds)
        info = _gpu(gm, mech, meth, ctl, a, [0.01])
    finally:
        gm.set_gen_options()
        gm.set_kernel(0)
    _oracle(mech, meth, ctl, b, [0.01])
    assert info.kernel == 8
    assert compare_bits(a, b) == ZERO


def test_delta32_repeats_after_a_previous_call(gm):
    """One advance, then simk_kin_advance_n(2) fused in one launch: the same
    bits as three single advances and as the oracle, at the tolerance where
    half the cells fail in f32 (rtol 1e-6). The fused launch stores each
    repeat's cell and re-reads the stored rows for the next one, starting
    here from the state a previous launch left."""
    sp, mech, z = make_problem("c2", 512)
    meth, ctl = problem.method(2), problem.control()
    a, b = z.copy(), z.copy()
    gm.arena_reset()
    gm.set_kernel(1)
    gm.set_mech(mech, meth, ctl)
    da, db = gpu.DeviceZones(gm, a), gpu.DeviceZones(gm, b)
    da.upload()
    db.upload()
    da.advance(0.2, precision=1)
    da.advance(0.2, precision=1, repeats=2)
    for _ in range(3):
        db.advance(0.2, precision=1)
    da.download(mutable_only=False)
    db.download(mutable_only=False)
    gm.set_kernel(0)
    assert a.block.tobytes() == b.block.tobytes()
    assert 0 < a.failed[:512].sum() < 512
    ok = np.nonzero(a.failed[:512] == 0)[0][:4]
    c = z.subset(ok)
    _oracle(mech, meth, ctl, c, [0.2, 0.2, 0.2])
    assert compare_bits(a.subset(ok), c) == ZERO


@pytest.mark.parametrize("name", ["c1", "c3c"])
def test_delta32_paired_slots_keep_bits(gm, name):
    """The generated f32 delta kernel with its workspace in adjacent f32
    pairs (pair_slots = 1: LU entries in the solves' read order, two slots
    per 8-byte access; an option, off by default): bit-identical to the
    oracle, with pulses, over two calls."""
    try:
        gm.set_kernel(1)
        gm.set_gen_options(pair_slots=1)
        sp, mech, z = make_problem(name, 256)
        add_pulses(z, seed=9, duration=0.01)
        meth, ctl = problem.method(2), problem.control(rtol=1e-3, atol=1e-10, max_steps=1000)
        a, b = z.copy(), z.copy()
        info = _gpu(gm, mech, meth, ctl, a, [0.01, 0.01])
        _oracle(mech, meth, ctl, b, [0.01, 0.01])
        assert info.kernel == KERNEL_CODE[1]
        assert compare_bits(a, b) == ZERO
    finally:
        gm.set_gen_options()
        gm.set_kernel(0)
