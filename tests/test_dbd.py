"""The DBD reactor workload on the host: SimStack's humid-air mechanism, feed,
discharge source and per-step zone preparation (simk_dbd.h) against the
reference's own air_mech.cpp (compiled in place into oracle/_ref), and the
reference's physics known answers (src/tests/test_kinetics.cpp:343-398) run
through the oracle's solver."""
import ctypes as C

import numpy as np
import pytest

import oracle_ffi
from simstack_amd import _ffi, dbd, problem

REF = pytest.mark.skipif(not oracle_ffi.have_ref(), reason="oracle/_ref not built (reference tree absent)")


@pytest.fixture(scope="module")
def air():
    return dbd.AirMechanism()


def _ref_air():
    lib = oracle_ffi.ref()
    for name, res, args in (("ref_air_build", C.c_int32, [C.c_void_p, C.c_void_p]),
                            ("ref_air_rate_constants", None, [C.c_double, C.c_double, C.c_void_p]),
                            ("ref_air_feed", C.c_double, [C.c_double, C.c_double, C.c_double, C.c_void_p]),
                            ("ref_air_source_delta", None, [C.c_void_p, C.c_double, C.c_double, C.c_uint32,
                                                            C.c_double, C.c_void_p]),
                            ("ref_air_count_elements", None, [C.c_void_p, C.c_void_p])):
        f = getattr(lib, name)
        f.restype, f.argtypes = res, args
    return lib


@pytest.mark.ref
@REF
def test_air_mechanism_tables_equal_reference(air):
    lib = _ref_air()
    mb = np.zeros(oracle_ffi.ref().ref_sizeof(0), np.uint8)
    idx = (C.c_uint32 * 17)()
    assert lib.ref_air_build(mb.ctypes.data, idx) == 0
    assert [idx[i] for i in range(17)] == [getattr(air.sp, n) for n in _ffi.SimkAirSpecies.NAMES]
    out = np.zeros_like(mb)
    caps = _ffi.caps(oracle_ffi.ref_caps())
    assert _ffi.host().simk_model_mech_bytes(air.model, C.byref(caps), out.ctypes.data) == 0
    assert out.tobytes() == mb.tobytes()
    assert (air.n, air.R) == (17, 44)


@pytest.mark.ref
@REF
@pytest.mark.parametrize("T,p", [(300.0, 101325.0), (250.0, 5e4), (340.0, 2e5), (420.0, 101325.0)])
def test_air_rate_constants_equal_reference(air, T, p):
    """kin_eval_rate_constants (Arrhenius, third body, Troe fall-off,
    equilibrium reverse) bit for bit at operating points around 1 atm."""
    M = p / (1.380649e-23 * T) * 1e-6
    want = np.zeros(44)
    _ref_air().ref_air_rate_constants(T, M, want.ctypes.data)
    got = air.rate_constants(T, M)
    assert got.tobytes() == want.tobytes()
    assert np.all(got > 0)


@pytest.mark.ref
@REF
def test_air_feed_and_source_equal_reference(air):
    lib = _ref_air()
    src = dbd.source_defaults()
    rng = np.random.default_rng(8)
    for T, p, x in ((300.0, 101325.0, 0.0), (320.0, 101325.0, 0.02), (350.0, 9e4, 0.035)):
        y = np.zeros(17)
        want = np.zeros(17)
        tot = _ffi.host().simk_air_feed(C.byref(air.sp), 17, T, p, x, y.ctypes.data_as(_ffi.f64p))
        assert tot == lib.ref_air_feed(T, p, x, want.ctypes.data)
        assert y.tobytes() == want.tobytes()
        # discharge increments on the feed and on a state carrying radicals
        for scale in (0.0, 1.0):
            yy = y + scale * rng.random(17) * 1e13
            for energy in (1e-6, 2e-5, 3e-3):
                d = np.zeros(17)
                dw = np.zeros(17)
                assert _ffi.host().simk_air_source_delta(C.byref(air.sp), C.byref(src), yy.ctypes.data_as(_ffi.f64p),
                                                         17, energy, T, d.ctypes.data_as(_ffi.f64p)) == 0
                lib.ref_air_source_delta(yy.ctypes.data, energy, T, 0, 0.0, dw.ctypes.data)
                assert d.tobytes() == dw.tobytes()

    # trace partners under a large burst: channels hit the half-the-partner
    # cap and hand their surplus to the leading channel; a gas with no
    # partners at all sends everything to channel 0
    sp = air.sp
    for O2, NO, O, N2O, H2O in ((1e10, 1e14, 1e11, 1e9, 0.0), (0.0, 0.0, 0.0, 0.0, 0.0), (1e8, 3e12, 1e13, 5e12, 1e9)):
        yy = np.zeros(17)
        yy[sp.N2], yy[sp.O2], yy[sp.NO], yy[sp.O], yy[sp.N2O], yy[sp.H2O] = 2.4e19, O2, NO, O, N2O, H2O
        for energy in (2e-5, 3e-2):
            d, dw = np.zeros(17), np.zeros(17)
            _ffi.host().simk_air_source_delta(C.byref(air.sp), C.byref(src), yy.ctypes.data_as(_ffi.f64p), 17,
                                              energy, 330.0, d.ctypes.data_as(_ffi.f64p))
            lib.ref_air_source_delta(yy.ctypes.data, energy, 330.0, 0, 0.0, dw.ctypes.data)
            assert d.tobytes() == dw.tobytes()


def test_air_source_conserves_atoms(air):
    """Every channel of the discharge source conserves N, O and H atoms
    (air_mech.cpp:127-129)."""
    src = dbd.source_defaults()
    y = np.zeros(17)
    _ffi.host().simk_air_feed(C.byref(air.sp), 17, 310.0, 101325.0, 0.02, y.ctypes.data_as(_ffi.f64p))
    y[air.sp.O3] = 1e15
    y[air.sp.NO] = 3e14
    y[air.sp.N2O] = 2e13
    y[air.sp.O] = 1e12
    d = np.zeros(17)
    _ffi.host().simk_air_source_delta(C.byref(air.sp), C.byref(src), y.ctypes.data_as(_ffi.f64p), 17, 1e-4, 310.0,
                                      d.ctypes.data_as(_ffi.f64p))
    before, after = air.count_elements(y), air.count_elements(y + d)
    assert np.all(np.abs(after - before) <= 1e-12 * before)
    assert d[air.sp.O] > 0 and d[air.sp.O2] < 0 and d[air.sp.N] > 0


def test_pulse_schedule_counts_pulses_in_each_step(air):
    """dbd_zones_prepare_source's clock: pulse j of a zone fires at j * period
    (j >= 1); a step (t0, t1] gets every pulse in it, the first in local time."""
    z = dbd.DbdZones(air, 3, freq=7000.0).setup()
    src = dbd.source_defaults()
    fired = np.zeros(3)
    for step in range(10):
        t0, t1 = step * 1e-3, (step + 1) * 1e-3
        z.prepare_source(src, dbd.PULSED, t0, t1)
        period = z.period[0]
        assert abs(period - 1.0 / 14000.0) < 1e-18
        assert 0.0 < z.pulse_first[0] <= period * 1.000001 + 1e-12 or z.pulse_count[0] == 0
        fired += z.pulse_count[:3]
        z.commit_pulses()
    assert np.all(fired == 140)   # 14,000 half-cycles per second over 10 ms


def _run(z, src, mode, steps, dt=1e-3, method=2):
    meth, ctl = problem.method(method), dbd.default_control()
    for step in range(steps):
        t0, t1 = step * dt, (step + 1) * dt
        z.prepare_source(src, mode, t0, t1)
        oracle_ffi.oracle_advance(z.mech, meth, ctl, z, t1 - t0, threads=8)
        z.commit_pulses()


def test_ozone_yield_known_answer(air):
    """The reference's test_ozone_yield (test_kinetics.cpp:373-398): dry air at
    0.1 mJ/cm^3 turns nearly every O atom the discharge makes into O3, and 10
    kHz for 10 ms fires exactly 200 pulses."""
    z = dbd.DbdZones(air, 1, T=300.0, p=101325.0, x_H2O=0.0, power=0.01, freq=10000.0).setup()
    src = dbd.source_defaults()
    _run(z, src, dbd.PULSED, 10)
    energy = 0.01 * 10e-3
    per_n2a = (2.0 * 2.5e-12 + 7.8e-14) / (2.5e-12 + 1.3e-12 + 7.8e-14)
    o_made = (2.0 * src.g_O2_diss + per_n2a * src.g_N2A + 2.0 * src.g_N2_diss) * energy * 6.241509074e18 * 0.01
    ratio = z.y[air.sp.O3, 0] / o_made
    assert 0.8 < ratio < 1.0
    assert z.pulses[0] == 200.0


@pytest.mark.parametrize("mode", [dbd.PULSED, dbd.AVERAGED], ids=["pulsed", "averaged"])
def test_conservation_known_answer(air, mode):
    """The reference's test_conservation (test_kinetics.cpp:343-370): 8 zones
    x 50 ms keep every N, O and H atom to 1e-10, and no zone fails."""
    z = dbd.reference_zones(air, 8)
    before = [air.count_elements(z.y[:, i]) for i in range(8)]
    _run(z, dbd.source_defaults(), mode, 50)
    worst = 0.0
    for i in range(8):
        after = air.count_elements(z.y[:, i])
        for e in range(3):
            if before[i][e] > 0:
                worst = max(worst, abs(after[e] - before[i][e]) / max(abs(after[e]), abs(before[i][e])))
    assert worst < 1e-10
    assert z.failed[:8].sum() == 0 and z.steps[:8].min() > 50
