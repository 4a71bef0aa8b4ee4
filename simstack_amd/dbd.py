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
