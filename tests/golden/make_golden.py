"""Writes the golden fixtures under tests/golden/ by running the reference's own
kinetics sources (oracle/_ref/libref.so, compiled in place from the reference
tree by oracle/ref/Makefile). Run in a container that has the reference:

    python tests/golden/make_golden.py

Each fixture is data only: the inputs (mechanism reaction lists, seeds and
controls) and the reference's outputs (f64 bits of y, h and the counters).
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

from simstack_amd import problem  # noqa: E402
import oracle_ffi  # noqa: E402


def ref_robertson():
    """Robertson to t = 40 (test_kinetics.cpp:151-180) under all three methods."""
    reac = np.array([[0, 255, 255], [1, 1, 255], [1, 2, 255]], np.uint8)
    prod = np.array([[1, 255, 255], [2, 1, 255], [0, 2, 255]], np.uint8)
    mech = problem.Mechanism(3, 3, reac, prod, caps=oracle_ffi.ref_caps())
    mb = oracle_ffi.ref_mech(mech)
    k = np.array([0.04, 3.0e7, 1.0e4])
    out = {"reactants": reac, "products": prod, "k": k}
    for mid in range(3):
        mbytes = oracle_ffi.ref_method_bytes(mid)
        ctl = problem.control(rtol=1e-6, atol=1e-12, h_min=0.0, h_start=1e-6, max_steps=100000, clamp_negative=0)
        y = np.array([1.0, 0.0, 0.0])
        src = np.zeros(3)
        t, h = C.c_double(0.0), C.c_double(0.0)
        stats = np.zeros(4)
        oracle_ffi.ref().ref_integrate(mb.ctypes.data, mbytes.ctypes.data, C.addressof(ctl),
                                       k.ctypes.data_as(oracle_ffi.f64p), src.ctypes.data_as(oracle_ffi.f64p),
                                       y.ctypes.data_as(oracle_ffi.f64p), C.byref(t), 40.0, C.byref(h),
                                       stats.ctypes.data_as(oracle_ffi.f64p))
        out[f"y_{mid}"] = y
        out[f"t_{mid}"] = np.array([t.value])
        out[f"h_{mid}"] = np.array([h.value])
        out[f"stats_{mid}"] = stats
    return out


def ref_zones(name, cells, durations, mid=2, pulses=False, ctl_kw=None, seed_pulses=11, delta32=False):
    """A seeded workload advanced by the reference's dbd_advance_group<f64>, or
    with delta32 by its f32 delta lanes dbd_advance_group<f32, ..., true>."""
    from helpers import add_pulses
    sp = problem.spec(name)
    mech = problem.Mechanism.synthetic(sp)
    z = problem.Zones(mech.n, mech.R, cells).fill_synthetic(sp, mech.params)
    if pulses:
        add_pulses(z, seed=seed_pulses, duration=durations[0])
    ctl = problem.control(**(ctl_kw or {}))
    z_in = z.copy()
    mb = oracle_ffi.ref_mech(mech)
    meth_b = oracle_ffi.ref_method_bytes(mid)
    for d in durations:
        oracle_ffi.ref_advance(mb, meth_b, ctl, z, d, z.padded, width=32 if delta32 else 1)
    out = {"reactants": mech.reactants, "products": mech.products, "block_in": z_in.block,
           "block_out": z.block, "padded": np.array([z.padded]), "count": np.array([cells]),
           "n": np.array([mech.n]), "R": np.array([mech.R]), "durations": np.array(durations),
           "method": np.array([mid]),
           "ctl": np.array([ctl.rtol, ctl.atol, ctl.h_min, ctl.h_max, ctl.h_start, ctl.max_steps, ctl.fixed_step,
                            ctl.clamp_negative], np.float64),
           "mech_bytes_ref_caps": mb, "ref_caps": np.array(oracle_ffi.ref_caps(), np.uint32),
           "delta32": np.array([1 if delta32 else 0])}
    return out


def main():
    if not oracle_ffi.have_ref():
        sys.exit("oracle/_ref/libref.so is not built (needs the reference tree): run __graft_entry__.build()")
    np.savez_compressed(os.path.join(HERE, "robertson.npz"), **ref_robertson())
    np.savez_compressed(os.path.join(HERE, "c1_32cells_1s.npz"), **ref_zones("c1", 32, [1.0]))
    for mid in range(3):
        np.savez_compressed(os.path.join(HERE, f"c1_16cells_pulses_m{mid}.npz"),
                            **ref_zones("c1", 16, [0.05, 0.05], mid=mid, pulses=True,
                                        ctl_kw=dict(rtol=1e-5, atol=1e-10)))
    np.savez_compressed(os.path.join(HERE, "c3c_8cells_fixed.npz"),
                        **ref_zones("c3c", 8, [0.02], ctl_kw=dict(fixed_step=1, max_steps=1000, h_start=1e-3)))
    np.savez_compressed(os.path.join(HERE, "c2_8cells_failure.npz"),
                        **ref_zones("c2", 8, [1.0], ctl_kw=dict(max_steps=40)))
    # f32 delta lanes (the reference's default f32 precision)
    np.savez_compressed(os.path.join(HERE, "c1_d32_24cells_pulses.npz"),
                        **ref_zones("c1", 24, [0.05, 0.05], mid=2, pulses=True, delta32=True,
                                    ctl_kw=dict(rtol=1e-3, atol=1e-10)))
    for mid in (0, 1):
        np.savez_compressed(os.path.join(HERE, f"c3c_d32_16cells_m{mid}.npz"),
                            **ref_zones("c3c", 16, [0.1], mid=mid, delta32=True, ctl_kw=dict(rtol=1e-4, atol=1e-11)))
    np.savez_compressed(os.path.join(HERE, "c2_d32_8cells_failure.npz"),
                        **ref_zones("c2", 8, [1.0], delta32=True, ctl_kw=dict(rtol=1e-5, atol=1e-12, max_steps=60)))
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
