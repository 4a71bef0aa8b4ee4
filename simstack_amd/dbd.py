"""The DBD reactor workload: SimStack's humid-air mechanism (17 species, 44
reactions) in zones with per-zone operating points, prepared on the host each
master step (rate constants, pulse schedule, discharge increments) the way
src/sim/dbd.cpp does, then advanced by the solver (include/simk_dbd.h)."""
import ctypes as C

import numpy as np

from . import _ffi, problem


class AirMechanism(problem.Mechanism):
    """The humid-air mechanism as KinMech bytes (at `caps`), with its model
    handle (rate laws, element counts) and species indices."""

    def __init__(self, caps=_ffi.RAISED_CAPS):
        h = _ffi.host()
        self.sp = _ffi.SimkAirSpecies()
        self.model = h.simk_air_model(C.byref(self.sp))
        err = C.create_string_buffer(200)
        if h.simk_model_error(self.model, err, 200):
            raise ValueError(err.value.decode())
        self.n = int(h.simk_model_species_count(self.model))
        self.R = int(h.simk_model_reaction_count(self.model))
        self.caps_tuple = tuple(caps)
        self.caps = _ffi.caps(self.caps_tuple)
        self.bytes = np.zeros(_ffi.mech_size(self.caps), dtype=np.uint8)
        if h.simk_model_mech_bytes(self.model, C.byref(self.caps), self.bytes.ctypes.data) != 0:
            raise ValueError("air mechanism does not fit the capacities")
        st = (C.c_uint64 * 6)()
        h.simk_mech_stats(self.bytes.ctypes.data, C.byref(self.caps), st)
        self.nnz, self.lu_div, self.lu_sub = int(st[2]), int(st[3]), int(st[4])
        self.params = None

    def __del__(self):
        try:
            _ffi.host().simk_model_free(self.model)
        except Exception:
            pass

    def rate_constants(self, T, M):
        k = np.zeros(self.R)
        _ffi.host().simk_model_rate_constants(self.model, T, M, k.ctypes.data_as(_ffi.f64p))
        return k

    def count_elements(self, y):
        atoms = np.zeros(3)
        _ffi.host().simk_model_count_elements(self.model, np.ascontiguousarray(y).ctypes.data_as(_ffi.f64p),
                                              atoms.ctypes.data_as(_ffi.f64p))
        return atoms


def default_control():
    c = _ffi.KinControl()
    _ffi.host().simk_dbd_default_control(C.byref(c))
    return c


def source_defaults():
    s = _ffi.SimkAirSource()
    _ffi.host().simk_air_source_defaults(C.byref(s))
    return s


PULSED, AVERAGED = 0, 1


class DbdZones(problem.Zones):
    """A zone block of the air mechanism with each zone's operating point."""

    def __init__(self, mech, count, T=300.0, p=101325.0, x_H2O=0.0, power=0.05, freq=5000.0):
        super().__init__(mech.n, mech.R, count)
        self.mech = mech
        pad = self.padded
        self.T_K = np.full(pad, T, dtype=np.float64)
        self.p_Pa = np.full(pad, p, dtype=np.float64)
        self.x_H2O = np.full(pad, x_H2O, dtype=np.float64)
        self.power_W_cm3 = np.full(pad, power, dtype=np.float64)
        self.freq_Hz = np.full(pad, freq, dtype=np.float64)
        self.pulse_index = np.zeros(pad, dtype=np.float64)

    def desc(self):
        d = _ffi.SimkDbdZones(self.n, self.R, self.count, self.padded, self.block.ctypes.data, self.T_K.ctypes.data,
                              self.p_Pa.ctypes.data, self.x_H2O.ctypes.data, self.power_W_cm3.ctypes.data,
                              self.freq_Hz.ctypes.data, self.pulse_index.ctypes.data)
        return d

    def setup(self):
        """dbd_zones_fill_padding, dbd_zones_reset, dbd_zones_prepare_rates."""
        h = _ffi.host()
        d = self.desc()
        h.simk_dbd_fill_padding(C.byref(d))
        h.simk_dbd_reset(C.byref(d), C.byref(self.mech.sp))
        h.simk_dbd_prepare_rates(C.byref(d), self.mech.model)
        return self

    def prepare_source(self, src, mode, t0, t1):
        d = self.desc()
        if _ffi.host().simk_dbd_prepare_source(C.byref(d), C.byref(self.mech.sp), C.byref(src), mode, t0, t1) != 0:
            raise ValueError("source model not supported")

    def commit_pulses(self):
        d = self.desc()
        _ffi.host().simk_dbd_commit_pulses(C.byref(d))

    def subset(self, lanes):
        """A new DbdZones holding the given zones (block rows and operating points)."""
        lanes = np.asarray(lanes)
        base = problem.Zones.subset(self, lanes)
        z = DbdZones(self.mech, len(lanes))
        z.block[:] = base.block
        for a in ("T_K", "p_Pa", "x_H2O", "power_W_cm3", "freq_Hz", "pulse_index"):
            v = getattr(z, a)
            v[:len(lanes)] = getattr(self, a)[lanes]
            v[len(lanes):] = v[0]
        return z

    def copy(self):
        z = DbdZones(self.mech, self.count)
        z.block[:] = self.block
        for a in ("T_K", "p_Pa", "x_H2O", "power_W_cm3", "freq_Hz", "pulse_index"):
            getattr(z, a)[:] = getattr(self, a)
        return z


def reference_zones(mech, count):
    """The operating points of the reference's zone tests
    (src/tests/test_kinetics.cpp:214-227): T 300-340 K, 1 atm, 0-3 % water,
    0.05-0.25 W/cm^3, 5-12.5 kHz."""
    z = DbdZones(mech, count)
    for i in range(count):
        z.T_K[i] = 300.0 + 10.0 * (i % 5)
        z.p_Pa[i] = 101325.0
        z.x_H2O[i] = 0.005 * (i % 7)
        z.power_W_cm3[i] = 0.05 + 0.1 * (i % 3)
        z.freq_Hz[i] = 5000.0 + 2500.0 * (i % 4)
    return z.setup()


def fnv1a64(data, seed=0xCBF29CE484222325):
    """64-bit FNV-1a of a bytes-like object."""
    h = seed
    for b in bytes(data):
        h = ((h ^ b) * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return h


class Scene:
    """The chained DBD reactor as the reference's scene runs it
    (src/sim/dbd_scene.cpp:82-90 operating point, :226-298 master step):
    `count` zones (64, the command line's default, src/platform/
    platform_common.cpp:190) in a chain at T 310 K, 1 atm, 2 % water, 0.1 W/cm^3,
    10 kHz, fed with humid air and exchanging by plug flow over a 2-s
    residence. Each master step of dt prepares the zones' source and pulse
    schedule over (t, t + dt], advances the chemistry, commits the pulses and
    relaxes each zone toward its upstream neighbour (zone 0 toward the feed)
    with a = dt * count / residence.

    The chemistry advance and the plug flow are supplied by the caller:
    `advance(zones, dt)` integrates the block in place and
    `flow(zones, a, feed)` applies the plug-flow update to its y rows."""

    def __init__(self, mech, count=64, residence=2.0, mode=PULSED, source=None):
        if count % 8:
            raise ValueError("the scene's state hash covers whole groups of 8 zones (the reference's row padding)")
        self.mech, self.count, self.residence, self.mode = mech, count, residence, mode
        self.source = source if source is not None else source_defaults()
        self.z = DbdZones(mech, count, T=310.0, p=101325.0, x_H2O=0.02, power=0.1, freq=10000.0).setup()
        self.t = 0.0
        self.master_steps = 0

    def feed(self):
        y = np.zeros(self.mech.n)
        _ffi.host().simk_air_feed(C.byref(self.mech.sp), self.mech.n, float(self.z.T_K[0]), float(self.z.p_Pa[0]),
                                  float(self.z.x_H2O[0]), y.ctypes.data_as(_ffi.f64p))
        return y

    def step(self, advance, flow, dt=1e-3):
        t0, t1 = self.t, self.t + dt
        self.z.prepare_source(self.source, self.mode, t0, t1)
        advance(self.z, dt)
        self.z.commit_pulses()
        if self.residence > 0.0:
            flow(self.z, dt * self.count / self.residence, self.feed())
        self.t = t1
        self.master_steps += 1

    def state_hash(self):
        """FNV-1a 64 of every zone's species row (dbd_scene_hash,
        src/sim/dbd_scene.cpp:388-391: n rows of the zones, in row order)."""
        return fnv1a64(np.ascontiguousarray(self.z.y[:, :self.count]).tobytes())


class Hybrid:
    """One zone block advanced by the GPU module and a CPU path at once, as the
    reference caller dbd_advance_hybrid does (src/sim/dbd.cpp:296-463): lanes
    [0, split) go to the device through gpu_kin_advance_async (the host block
    is copied in and the lanes' mutable rows back, `split` lanes wide), while
    `cpu_advance(zones, duration, lanes)` integrates [split, padded) on the
    host; gpu_kin_wait joins them. Each lane gets the same bits on either
    side, so the split (`force_split`, or the simk_hybrid_* controller fed
    with each advance's timings) changes only the wall time.

    `dz` is a gpu.DeviceZones of the block; `cpu_advance` is the caller's CPU
    path (the tests pass the oracle)."""

    def __init__(self, gm, dz, cpu_advance, force_split=None):
        self.gm, self.dz, self.cpu_advance = gm, dz, cpu_advance
        self.force_split = force_split
        self.ctl = _ffi.SimkHybrid()
        _ffi.host().simk_hybrid_init(C.byref(self.ctl), dz.z.padded)
        self.history = []

    @property
    def split(self):
        s = self.ctl.split if self.force_split is None else self.force_split
        return min(s & ~7, self.dz.z.padded)

    def advance(self, duration, precision=0):
        import time
        z, gm = self.dz.z, self.gm
        split = self.split
        kms, tms = C.c_float(0.0), C.c_float(0.0)
        if split:
            L = self.dz.launch_desc(duration, lane_count=split, precision=precision)
            gm._check(gm.lib.gpu_kin_advance_async(C.byref(L), z.block.ctypes.data, self.dz.d_block,
                                                   z.total // z.padded, z.mutable // z.padded),
                      "gpu_kin_advance_async")
        t0 = time.perf_counter()
        if split < z.padded:
            self.cpu_advance(z, duration, np.arange(split, z.padded, dtype=np.uint32))
        cpu_ms = (time.perf_counter() - t0) * 1e3
        if split:
            gm._check(gm.lib.gpu_kin_wait(C.byref(kms), C.byref(tms)), "gpu_kin_wait")
        self.history.append((split, tms.value, cpu_ms))
        if self.force_split is None:
            _ffi.host().simk_hybrid_record(C.byref(self.ctl), split, tms.value, cpu_ms)
        return split
