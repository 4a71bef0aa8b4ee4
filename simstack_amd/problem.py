"""Mechanisms, zone blocks and the seeded synthetic workloads.

The zone block is the reference's DbdZones SoA block (src/sim/dbd.cpp:40-83):
one contiguous f64 array, rows [species or reaction][padded lanes], solver-
written rows first so a download needs only the mutable prefix.
"""
import ctypes as C

import numpy as np

from . import _ffi

# Seeded synthetic workloads. Every field is a SimkSynthSpec member; cell c of a
# workload draws from its own stream, so any subset of cells is reproducible.
_COMMON = dict(max_attempts=4000, log10_k_lo=-3.0, log10_k_hi=9.0, EaR_hi=6000.0, T_lo=288.0, T_hi=303.0,
               log10_y_lo=-12.0, log10_y_hi=-2.0, water_frac=0.3)

PRESETS = {
    # 32 species / 96 reactions, 256 cells, 1 s
    "c1": dict(_COMMON, seed=1, n_species=32, n_reactions=96, band=3, long_range=0.02,
               density_lo=0.0, density_hi=0.40),
    # 48 / 150 with ~18 % fill after the symbolic LU, pH 2-7
    "c2": dict(_COMMON, seed=2, n_species=48, n_reactions=150, band=3, long_range=0.02,
               density_lo=0.16, density_hi=0.20, pH_lo=2.0, pH_hi=7.0, acid_frac=0.1, base_frac=0.1),
    # the three ragged-batch mechanisms: 60/200, 40/130, 24/70; per-cell T, oxidant and load
    # spread (attempts per 1 s vary about 10x between the 5th and 95th percentile cells)
    "c3a": dict(_COMMON, seed=31, n_species=60, n_reactions=200, band=3, long_range=0.02,
                density_lo=0.10, density_hi=0.22, T_lo=278.0, T_hi=318.0,
                oxidant_log10_spread=3.0, oxidant_species=8, load_log10_spread=5.0),
    "c3b": dict(_COMMON, seed=32, n_species=40, n_reactions=130, band=3, long_range=0.02,
                density_lo=0.12, density_hi=0.26, T_lo=278.0, T_hi=318.0,
                oxidant_log10_spread=3.0, oxidant_species=6, load_log10_spread=5.0),
    "c3c": dict(_COMMON, seed=33, n_species=24, n_reactions=70, band=3, long_range=0.02,
                density_lo=0.15, density_hi=0.40, T_lo=278.0, T_hi=318.0,
                oxidant_log10_spread=3.0, oxidant_species=4, load_log10_spread=5.0),
}


def spec(name, **over):
    d = dict(PRESETS[name])
    d.update(over)
    s = _ffi.SimkSynthSpec()
    for k, v in d.items():
        setattr(s, k, v)
    return s


def control(rtol=1e-6, atol=1e-12, h_min=0.0, h_max=0.0, h_start=1e-6, max_steps=100000, fixed_step=0,
            clamp_negative=1):
    """KinControl (kin_mech.h:112-119); defaults are the synthetic workloads'."""
    return _ffi.KinControl(rtol, atol, h_min, h_max, h_start, max_steps, fixed_step, clamp_negative)


def method(mid=2):
    m = _ffi.KinMethod()
    if _ffi.host().simk_method(mid, C.byref(m)) != 0:
        raise ValueError(f"unknown method {mid}")
    return m


class Mechanism:
    """Finalized tables in the reference KinMech byte layout (at `caps`)."""

    def __init__(self, n, R, reactants, products, params=None, caps=_ffi.RAISED_CAPS):
        self.n, self.R = int(n), int(R)
        self.reactants = np.ascontiguousarray(reactants, dtype=np.uint8).reshape(R, 3)
        self.products = np.ascontiguousarray(products, dtype=np.uint8).reshape(R, 3)
        self.params = None if params is None else np.ascontiguousarray(params, dtype=np.float64).reshape(R, 4)
        self.caps_tuple = tuple(caps)
        self.caps = _ffi.caps(self.caps_tuple)
        self.bytes = np.zeros(_ffi.mech_size(self.caps), dtype=np.uint8)
        err = C.create_string_buffer(256)
        rc = _ffi.host().simk_mech_build(self.n, self.R, self.reactants.ctypes.data, self.products.ctypes.data,
                                         C.byref(self.caps), self.bytes.ctypes.data, err, 256)
        if rc != 0:
            raise ValueError(f"mechanism build failed: {err.value.decode()}")
        st = (C.c_uint64 * 6)()
        _ffi.host().simk_mech_stats(self.bytes.ctypes.data, C.byref(self.caps), st)
        self.nnz, self.lu_div, self.lu_sub = int(st[2]), int(st[3]), int(st[4])

    @classmethod
    def synthetic(cls, sp, caps=_ffi.RAISED_CAPS):
        R = sp.n_reactions
        reac = np.zeros((R, 3), np.uint8)
        prod = np.zeros((R, 3), np.uint8)
        par = np.zeros((R, 4), np.float64)
        used = _ffi.host().simk_synth_mechanism(C.byref(sp), reac.ctypes.data, prod.ctypes.data,
                                                par.ctypes.data_as(_ffi.f64p))
        if used == 0:
            raise ValueError("no synthetic mechanism accepted within max_attempts")
        m = cls(sp.n_species, R, reac, prod, par, caps)
        m.spec = sp
        m.draws = used
        return m

    def rebuild(self, caps):
        m = Mechanism(self.n, self.R, self.reactants, self.products, self.params, caps)
        m.__dict__.setdefault("spec", getattr(self, "spec", None))
        return m

    def equation(self, r):
        a = " + ".join(f"S{s}" for s in self.reactants[r] if s != 0xFF)
        b = " + ".join(f"S{s}" for s in self.products[r] if s != 0xFF)
        return f"{a} -> {b}"

    @property
    def density(self):
        return self.nnz / float(self.n * self.n)

    def _flop_terms(self):
        nr = (self.reactants != 0xFF).sum(axis=1)
        net = np.zeros(self.R, np.int64)
        for r in range(self.R):
            d = {}
            for s in self.reactants[r]:
                if s != 0xFF:
                    d[s] = d.get(s, 0) - 1
            for s in self.products[r]:
                if s != 0xFF:
                    d[s] = d.get(s, 0) + 1
            net[r] = sum(1 for v in d.values() if v != 0)
        return nr, net

    def flops_mixed(self, sweeps=0, stages=4, fun_evals=4):
        """(FP64, FP32) algorithmic flops of one mixed-precision step attempt
        (FMA = 2). FP32: the LU factor and every triangular solve, (1 + sweeps)
        per stage. FP64: f, -J, the vector work, and per sweep and stage the
        residual b - g x + J x (J x matrix-free from the reactions) and the
        correction x += z."""
        nr, net = self._flop_terms()
        f_fun = int((nr + 2 * net).sum())
        f_jac = int((nr * (nr - 1) + 2 * nr * net).sum())
        f_jx = int((nr * (nr + 2 * net)).sum())
        f_lu = self.lu_div + 2 * self.lu_sub + 2 * self.n
        f_solve = 2 * (self.nnz - self.n) + self.n
        fp32 = f_lu + stages * (1 + sweeps) * f_solve
        fp64 = fun_evals * f_fun + f_jac + 31 * self.n + 30 + sweeps * stages * (f_jx + 2 * self.n + self.n)
        return fp64, fp32

    def flops_per_attempt(self, stages=4, fun_evals=4):
        """Algorithmic FP64 flops of one Rodas3 step attempt (FMA = 2), from the tables:
        fun_evals * F_fun + F_jac + F_lu + stages * F_solve + 31 n + 30."""
        net = np.zeros(self.R, np.int64)
        nr = (self.reactants != 0xFF).sum(axis=1)
        for r in range(self.R):
            d = {}
            for s in self.reactants[r]:
                if s != 0xFF:
                    d[s] = d.get(s, 0) - 1
            for s in self.products[r]:
                if s != 0xFF:
                    d[s] = d.get(s, 0) + 1
            net[r] = sum(1 for v in d.values() if v != 0)
        f_fun = int((nr + 2 * net).sum())
        f_jac = int((nr * (nr - 1) + 2 * nr * net).sum())
        f_lu = self.lu_div + 2 * self.lu_sub + 2 * self.n
        f_solve = 2 * (self.nnz - self.n) + self.n
        return fun_evals * f_fun + f_jac + f_lu + stages * f_solve + 31 * self.n + 30


def pad_lanes(count, multiple=64):
    return max(multiple, (count + multiple - 1) // multiple * multiple)


def _aligned_zeros(count, align=256):
    """f64 zeros whose first element is `align`-byte aligned (the reference
    aligns the zone block to 64 B, dbd.cpp:75; AVX2 loads need 32 B)."""
    raw = np.zeros(count + align // 8, np.float64)
    skip = (-raw.ctypes.data % align) // 8
    return raw[skip:skip + count]


ROWS = ("y", "h", "steps", "rejects", "pulses", "failed", "k", "src", "pulse", "period", "pulse_first", "pulse_count")


class Zones:
    """A DbdZones-style SoA block (src/sim/dbd.cpp:40-83) held in one numpy array."""

    def __init__(self, n, R, count, padded=None):
        self.n, self.R, self.count = n, R, count
        self.padded = padded or pad_lanes(count)
        lay = (C.c_uint64 * 14)()
        _ffi.host().simk_zones_layout(n, R, self.padded, lay)
        self.offsets = [int(lay[i]) for i in range(12)]
        self.total = int(lay[12])
        self.mutable = int(lay[13])
        self.block = _aligned_zeros(self.total)
        p = self.padded
        sizes = {"y": n * p, "k": R * p, "src": n * p, "pulse": n * p}
        for name, off in zip(ROWS, self.offsets):
            v = self.block[off:off + sizes.get(name, p)]
            if name in ("y", "src", "pulse"):
                v = v.reshape(n, p)
            elif name == "k":
                v = v.reshape(R, p)
            setattr(self, name, v)

    def fill_synthetic(self, sp, params, first_cell=0):
        """Seeded inputs for cells [first_cell, first_cell + count); padding lanes copy lane 0."""
        p = self.padded
        self.T = np.zeros(self.count)
        self.pH = np.zeros(self.count)
        k = np.zeros((self.R, p))
        y = np.zeros((self.n, p))
        _ffi.host().simk_synth_cells(C.byref(sp), np.ascontiguousarray(params).ctypes.data_as(_ffi.f64p),
                                     first_cell, self.count, p, k.ctypes.data, y.ctypes.data,
                                     self.T.ctypes.data, self.pH.ctypes.data)
        k[:, self.count:] = k[:, :1]
        y[:, self.count:] = y[:, :1]
        self.k[:] = k
        self.y[:] = y
        return self

    def copy(self):
        z = Zones(self.n, self.R, self.count, self.padded)
        z.block[:] = self.block
        for a in ("T", "pH"):
            if hasattr(self, a):
                setattr(z, a, getattr(self, a).copy())
        return z

    def subset(self, lanes):
        """A new block holding the given lanes (cell order kept)."""
        lanes = np.asarray(lanes)
        z = Zones(self.n, self.R, len(lanes))
        for name in ROWS:
            src = getattr(self, name)
            dst = getattr(z, name)
            if src.ndim == 2:
                dst[:, :len(lanes)] = src[:, lanes]
                dst[:, len(lanes):] = src[:, lanes[:1]]
            else:
                dst[:len(lanes)] = src[lanes]
                dst[len(lanes):] = src[lanes[:1]]
        return z
