"""Cooperative mode (gpu_kin_set_mode(1), gpu_api.h:107-109): one wave per
cell, the cell in LDS, fun / -J / LU / solves spread over the wave by
contribution lists in the reference's accumulation order. The reference's own
GPU test (src/tests/test_kinetics.cpp:1176-1202) asks for warp-mode advances
bit-identical to thread mode in f64 and f32 lanes; here both are checked
against the oracle, which is bit-identical to the reference's CPU lanes."""
import numpy as np
import pytest

import oracle_ffi
from simstack_amd import _ffi, gpu, problem
from helpers import add_pulses, compare_bits, make_problem, oracle_run

pytestmark = pytest.mark.gpu

ZERO = {k: 0 for k in ("y", "h", "steps", "rejects", "pulses", "failed")}


def _run(gm, mech, meth, ctl, z, durations, precision=0, mode=1, repeats=1, team=64):
    gm.arena_reset()
    gm.set_mech(mech, meth, ctl)
    gm._check(gm.lib.gpu_kin_set_mode(mode), "gpu_kin_set_mode")
    gm._check(gm.lib.simk_kin_set_team(team), "simk_kin_set_team")
    try:
        dz = gpu.DeviceZones(gm, z)
        dz.upload()
        for d in durations:
            dz.advance(d, precision=precision, repeats=repeats)
        dz.download(mutable_only=False)
        return gm.kin_info()
    finally:
        gm.lib.gpu_kin_set_mode(0)
        gm.lib.simk_kin_set_team(64)


@pytest.mark.parametrize("mid", [0, 1, 2], ids=["ros2", "ros3", "rodas3"])
def test_wave_fp64_bit_exact_with_pulses(gm, mid):
    sp, mech, z = make_problem("c1", 192)
    add_pulses(z, seed=mid + 11, duration=0.05)
    meth, ctl = problem.method(mid), problem.control(rtol=1e-5, atol=1e-10)
    a, b = z.copy(), z.copy()
    info = _run(gm, mech, meth, ctl, a, [0.05, 0.05])
    oracle_run(mech, meth, ctl, b, [0.05, 0.05])
    assert info.kernel == 6
    assert compare_bits(a, b) == ZERO
    assert b.pulses[:192].sum() > 0


def test_wave_delta32_bit_exact_with_pulses(gm):
    """f32 delta lanes in cooperative mode (the reference's default precision)."""
    sp, mech, z = make_problem("c1", 192)
    add_pulses(z, seed=2, duration=0.01)
    meth, ctl = problem.method(2), problem.control(rtol=1e-3, atol=1e-10, max_steps=1000)
    a, b = z.copy(), z.copy()
    info = _run(gm, mech, meth, ctl, a, [0.01, 0.01], precision=1)
    for d in (0.01, 0.01):
        oracle_ffi.oracle_advance_delta32(mech, meth, ctl, b, d, threads=16)
    assert info.kernel == 7
    assert compare_bits(a, b) == ZERO


def test_wave_equals_thread_mode_48x150(gm):
    """48/150 mechanism: cooperative and one-lane-per-cell advances agree bit
    for bit (FP64), and fused repeats equal separate calls."""
    sp, mech, z = make_problem("c2", 256)
    meth, ctl = problem.method(2), problem.control()
    a, b, c = z.copy(), z.copy(), z.copy()
    _run(gm, mech, meth, ctl, a, [0.2, 0.2], mode=1)
    _run(gm, mech, meth, ctl, b, [0.2, 0.2], mode=0)
    _run(gm, mech, meth, ctl, c, [0.2], mode=1, repeats=2)
    assert compare_bits(a, b) == ZERO
    assert compare_bits(a, c) == ZERO


def test_wave_delta32_repeats_and_60x200(gm):
    """60/200 mechanism (683 LU entries in LDS), f32 delta lanes, three fused
    repeats == three calls == the oracle."""
    sp, mech, z = make_problem("c3a", 128)
    meth, ctl = problem.method(2), problem.control(rtol=1e-3, atol=1e-10, max_steps=1000)
    a, b, c = z.copy(), z.copy(), z.copy()
    _run(gm, mech, meth, ctl, a, [0.004], precision=1, repeats=3)
    _run(gm, mech, meth, ctl, b, [0.004, 0.004, 0.004], precision=1)
    for d in (0.004, 0.004, 0.004):
        oracle_ffi.oracle_advance_delta32(mech, meth, ctl, c, d, threads=16)
    assert compare_bits(a, b) == ZERO
    assert compare_bits(a, c) == ZERO


@pytest.mark.parametrize("precision", [2, 3], ids=["mixed", "fp32store"])
def test_wave_mode_rejects_mixed_precision(gm, precision):
    sp, mech, z = make_problem("c1", 64)
    meth, ctl = problem.method(2), problem.control()
    with pytest.raises(gpu.GpuError):
        _run(gm, mech, meth, ctl, z.copy(), [0.01], precision=precision)


@pytest.mark.parametrize("team", [8, 16, 32])
def test_teams_fp64_bit_exact_with_pulses(gm, team):
    """Teams of 8, 16 and 32 lanes per cell (64 / team cells per wave, each
    team with its own control flow): every cell bit-identical to the oracle,
    with pulse segments, two calls, a cell count that leaves the last wave's
    teams partly empty, and one fused pair of repeats."""
    print(f"[case] teams of {team} lanes, FP64, c1 x 203 cells", flush=True)
    sp, mech, z = make_problem("c1", 203)
    add_pulses(z, seed=team, duration=0.05)
    meth, ctl = problem.method(2), problem.control(rtol=1e-5, atol=1e-10)
    a, b, c = z.copy(), z.copy(), z.copy()
    info = _run(gm, mech, meth, ctl, a, [0.05, 0.05], team=team)
    _run(gm, mech, meth, ctl, c, [0.05], team=team, repeats=2)
    oracle_run(mech, meth, ctl, b, [0.05, 0.05])
    assert info.kernel == 6 and f"{team} lanes".encode() in info.kernel_name
    assert compare_bits(a, b) == ZERO
    assert compare_bits(c, b) == ZERO
    assert b.pulses[:203].sum() > 0


@pytest.mark.parametrize("team", [8, 16, 32])
def test_teams_delta32_and_48x150(gm, team):
    """f32 delta lanes in teams (the c1 mechanism, pulses) and the 48/150
    mechanism in FP64 teams: bit-identical to the oracle."""
    print(f"[case] teams of {team} lanes, f32 delta c1 and FP64 c2", flush=True)
    sp, mech, z = make_problem("c1", 150)
    add_pulses(z, seed=team + 1, duration=0.01)
    meth, ctl = problem.method(2), problem.control(rtol=1e-3, atol=1e-10, max_steps=1000)
    a, b = z.copy(), z.copy()
    _run(gm, mech, meth, ctl, a, [0.01, 0.01], precision=1, team=team)
    for d in (0.01, 0.01):
        oracle_ffi.oracle_advance_delta32(mech, meth, ctl, b, d, threads=16)
    assert compare_bits(a, b) == ZERO
    sp, mech, z = make_problem("c2", 100)
    meth, ctl = problem.method(2), problem.control()
    a, b = z.copy(), z.copy()
    _run(gm, mech, meth, ctl, a, [0.1], team=team)
    oracle_run(mech, meth, ctl, b, [0.1])
    assert compare_bits(a, b) == ZERO
