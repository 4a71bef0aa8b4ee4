"""The DBD reactor on the device, as the reference's own GPU test runs it
(src/tests/test_kinetics.cpp:1086-1130): the humid-air mechanism through the
reference ABI (gpu_kin_set_mech at the default capacities), 12 zones with
varied operating points, 5 master steps of 1 ms with the host preparing each
step's pulses or averaged source, pulsed and averaged discharges. Every step's
result is bit-identical to the oracle's: FP64 lanes (generated, table walk and
one wave per cell) and the reference's f32 delta lanes."""
import ctypes as C

import numpy as np
import pytest

import oracle_ffi
from simstack_amd import _ffi, dbd, gpu, problem
from helpers import compare_bits

pytestmark = pytest.mark.gpu

ZERO = {k: 0 for k in ("y", "h", "steps", "rejects", "pulses", "failed")}


@pytest.fixture(scope="module")
def air():
    return dbd.AirMechanism(caps=_ffi.DEFAULT_CAPS)


def _gpu_run(gm, air, z, mode, steps, precision=0, kernel=0, wave=False):
    meth, ctl = problem.method(2), dbd.default_control()
    gm.arena_reset()
    gm.set_kernel(kernel)
    gm._check(gm.lib.gpu_kin_set_mech(air.bytes.ctypes.data, C.byref(meth), C.byref(ctl)), "gpu_kin_set_mech")
    gm._check(gm.lib.gpu_kin_set_mode(1 if wave else 0), "gpu_kin_set_mode")
    src = dbd.source_defaults()
    dz = gpu.DeviceZones(gm, z)
    try:
        for step in range(steps):
            t0, t1 = step * 1e-3, (step + 1) * 1e-3
            z.prepare_source(src, mode, t0, t1)
            dz.upload()
            dz.advance(t1 - t0, precision=precision)
            dz.download(mutable_only=True)
            z.commit_pulses()
    finally:
        gm.lib.gpu_kin_set_mode(0)
        gm.set_kernel(0)


def _oracle_run(air, z, mode, steps, precision=0):
    meth, ctl = problem.method(2), dbd.default_control()
    src = dbd.source_defaults()
    adv = oracle_ffi.oracle_advance_delta32 if precision == 1 else oracle_ffi.oracle_advance
    for step in range(steps):
        t0, t1 = step * 1e-3, (step + 1) * 1e-3
        z.prepare_source(src, mode, t0, t1)
        adv(air, meth, ctl, z, t1 - t0, threads=8)
        z.commit_pulses()


@pytest.mark.parametrize("mode", [dbd.PULSED, dbd.AVERAGED], ids=["pulsed", "averaged"])
@pytest.mark.parametrize("variant", ["generated", "table", "wave"])
def test_dbd_fp64_bit_exact(gm, air, mode, variant):
    a = dbd.reference_zones(air, 12)
    b = a.copy()
    _gpu_run(gm, air, a, mode, 5, kernel={"generated": 1, "table": 2, "wave": 0}[variant], wave=variant == "wave")
    _oracle_run(air, b, mode, 5)
    assert compare_bits(a, b) == ZERO
    assert b.steps[:12].min() > 0
    if mode == dbd.PULSED:
        assert b.pulses[:12].min() >= 50


@pytest.mark.parametrize("mode", [dbd.PULSED, dbd.AVERAGED], ids=["pulsed", "averaged"])
@pytest.mark.parametrize("variant", ["generated", "table"])
def test_dbd_f32_delta_bit_exact(gm, air, mode, variant):
    """The reference scene's default precision (f32 delta lanes)."""
    a = dbd.reference_zones(air, 12)
    b = a.copy()
    _gpu_run(gm, air, a, mode, 5, precision=1, kernel={"generated": 1, "table": 2}[variant])
    _oracle_run(air, b, mode, 5, precision=1)
    assert gm.kin_info().kernel == {"generated": 8, "table": 5}[variant]
    assert compare_bits(a, b) == ZERO


def test_dbd_f32_delta_many_zones_generated(gm, air):
    """2,048 zones x 20 ms pulsed in f32 delta lanes (generated kernel): 64
    sampled zones equal the oracle's f32 delta lanes bit for bit."""
    z = dbd.reference_zones(air, 2048)
    lanes = np.random.default_rng(6).choice(2048, 64, replace=False)
    sample = z.subset(lanes)
    _gpu_run(gm, air, z, dbd.PULSED, 20, precision=1, kernel=1)
    _oracle_run(air, sample, dbd.PULSED, 20, precision=1)
    assert gm.kin_info().kernel == 8
    assert compare_bits(z.subset(lanes), sample) == ZERO
    assert z.failed[:2048].sum() == 0


def test_dbd_many_zones(gm, air):
    """2,048 zones x 20 ms pulsed on the device: 64 sampled zones equal the
    oracle bit for bit; N and O atoms (totals ~1e19 cm^-3) are kept to 1e-10
    and H atoms (totals ~1e17-1e18, so an atol-sized drift of ~1e8 is 1e-10
    of them) to 1e-9; no zone fails."""
    z = dbd.reference_zones(air, 2048)
    lanes = np.random.default_rng(2).choice(2048, 64, replace=False)
    sample = z.subset(lanes)
    before = np.array([air.count_elements(z.y[:, i]) for i in range(2048)])
    _gpu_run(gm, air, z, dbd.PULSED, 20)
    _oracle_run(air, sample, dbd.PULSED, 20)
    got = z.subset(lanes)
    assert compare_bits(got, sample) == ZERO
    after = np.array([air.count_elements(z.y[:, i]) for i in range(2048)])
    with np.errstate(invalid="ignore", divide="ignore"):
        rel = np.where(before > 0, np.abs(after - before) / np.maximum(np.abs(after), np.abs(before)), 0.0)
    assert rel[:, :2].max() < 1e-10 and rel[:, 2].max() < 1e-9
    assert z.failed[:2048].sum() == 0


def test_f32_delta_lanes_accuracy_bounds(gm, air):
    """The reference's f32-lane accuracy bounds (test_kinetics.cpp:278-341)
    on the device: 16 zones at the reference's operating points, 100 ms
    pulsed and 1,000 ms averaged in 1-ms master steps, f32 delta lanes
    against FP64 lanes. The worst relative error of the watched products
    (O3, NO, NO2, N2O5, HNO2, HNO3, H2O2, N2O, where the FP64 run holds at
    least 1e-3 ppm) stays below 1e-2, the N/O/H atom drift below 3e-7 per
    simulated second (each master step rounds the increments to f32), and no
    zone fails."""
    sp = air.sp
    watch = [sp.O3, sp.NO, sp.NO2, sp.N2O5, sp.HNO2, sp.HNO3, sp.H2O2, sp.N2O]
    for mode, ms in ((dbd.PULSED, 100), (dbd.AVERAGED, 1000)):
        print(f"[case] f32 delta vs FP64 lanes, 16 zones x {ms} ms, mode {mode}", flush=True)
        z32, z64 = dbd.reference_zones(air, 16), dbd.reference_zones(air, 16)
        before = np.array([air.count_elements(z32.y[:, i]) for i in range(16)])
        _gpu_run(gm, air, z32, mode, ms, precision=1)
        _gpu_run(gm, air, z64, mode, ms, precision=0)
        worst = 0.0
        for i in range(16):
            total = z64.y[:, i].sum()
            for s in watch:
                if z64.y[s, i] / total * 1e6 < 1e-3:
                    continue
                a, b = z32.y[s, i], z64.y[s, i]
                worst = max(worst, abs(a - b) / max(abs(a), abs(b)))
        after = np.array([air.count_elements(z32.y[:, i]) for i in range(16)])
        with np.errstate(invalid="ignore", divide="ignore"):
            drift = np.where(before > 0, np.abs(after - before) / np.maximum(np.abs(after), np.abs(before)), 0.0).max()
        print(f"[f32] worst product error {worst:.2e}, atom drift {drift / (ms * 1e-3):.2e} per s, "
              f"steps {z32.steps[:16].sum():.0f} vs {z64.steps[:16].sum():.0f}", flush=True)
        assert worst < 1e-2
        assert drift / (ms * 1e-3) < 3e-7
        assert z32.failed[:16].sum() == 0 and z64.failed[:16].sum() == 0
