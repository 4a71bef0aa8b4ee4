"""Hybrid CPU + GPU advance (the reference caller dbd_advance_hybrid,
src/sim/dbd.cpp:296-463): lanes [0, split) on the device through
gpu_kin_advance_async / gpu_kin_wait while the CPU path integrates
[split, padded) at the same time. As the reference's test does
(src/tests/test_kinetics.cpp:1204-1243): 64 DBD zones, pulsed, three master
steps of 1 ms, splits 0/8/24/40/64, f64 and f32 delta lanes; every split gives
the CPU-only run's y bits. Beyond it: the split controller choosing a
different split from step to step (its dithering) changes no bit either.
The CPU path here is the oracle (the reference's own AVX2 path in a real
caller)."""
import ctypes as C

import pytest

import oracle_ffi
from simstack_amd import _ffi, dbd, gpu, problem

pytestmark = pytest.mark.gpu


def _cpu(air, meth, ctl, precision):
    adv = oracle_ffi.oracle_advance_delta32 if precision else oracle_ffi.oracle_advance

    def run(z, duration, lanes):
        adv(air, meth, ctl, z, duration, lanes=lanes, threads=4)
    return run


def _steps(sc, advance, count):
    for step in range(count):
        sc.z.prepare_source(sc.source, sc.mode, step * 1e-3, (step + 1) * 1e-3)
        advance(sc.z, 1e-3)
        sc.z.commit_pulses()


@pytest.mark.parametrize("precision", [0, 1], ids=["f64", "f32"])
def test_any_split_gives_the_cpu_bits(gm, precision):
    print(f"[case] hybrid splits, 64 DBD zones x 3 steps, precision {precision}", flush=True)
    air = dbd.AirMechanism(caps=_ffi.DEFAULT_CAPS)
    meth, ctl = problem.method(2), dbd.default_control()
    cpu = _cpu(air, meth, ctl, precision)
    ref = dbd.Scene(air)
    _steps(ref, lambda z, dt: cpu(z, dt, None), 3)
    want = ref.z.y[:, :64].tobytes()
    gm.arena_reset()
    gm.set_kernel(0)
    gm._check(gm.lib.gpu_kin_set_mech(air.bytes.ctypes.data, C.byref(meth), C.byref(ctl)), "set_mech")
    for split in (0, 8, 24, 40, 64):
        sc = dbd.Scene(air)
        hy = dbd.Hybrid(gm, gpu.DeviceZones(gm, sc.z), cpu, force_split=split)
        _steps(sc, lambda z, dt: hy.advance(dt, precision), 3)
        print(f"[hybrid] split {split}: {'same' if sc.z.y[:, :64].tobytes() == want else 'DIFFERENT'}", flush=True)
        assert sc.z.y[:, :64].tobytes() == want
        assert sc.z.steps[:64].tobytes() == ref.z.steps[:64].tobytes()


def test_controller_splits_change_no_bits(gm):
    print("[case] hybrid controller, 64 DBD zones x 12 steps, f64", flush=True)
    air = dbd.AirMechanism(caps=_ffi.DEFAULT_CAPS)
    meth, ctl = problem.method(2), dbd.default_control()
    cpu = _cpu(air, meth, ctl, 0)
    ref = dbd.Scene(air)
    _steps(ref, lambda z, dt: cpu(z, dt, None), 12)
    gm.arena_reset()
    gm.set_kernel(0)
    gm._check(gm.lib.gpu_kin_set_mech(air.bytes.ctypes.data, C.byref(meth), C.byref(ctl)), "set_mech")
    sc = dbd.Scene(air)
    hy = dbd.Hybrid(gm, gpu.DeviceZones(gm, sc.z), cpu)
    _steps(sc, lambda z, dt: hy.advance(dt), 12)
    splits = [s for s, _, _ in hy.history]
    print(f"[hybrid] splits {splits}; model c {hy.ctl.cpu_lane_ms:.3g} ms/lane, "
          f"o {hy.ctl.gpu_fixed_ms:.3g} ms, g {hy.ctl.gpu_lane_ms:.3g} ms/lane", flush=True)
    assert len(set(splits)) >= 2
    assert sc.z.y[:, :64].tobytes() == ref.z.y[:, :64].tobytes()
    assert sc.z.steps[:64].tobytes() == ref.z.steps[:64].tobytes()
