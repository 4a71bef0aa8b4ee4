"""The reference's DBD scene known answers on the device: the same 500 master
steps of 64 chained zones as tests/test_dbd_scene.py, with the chemistry
advanced by the GPU module through the reference ABI (gpu_kin_set_mech,
gpu_kin_advance) and the plug-flow chain run on the device
(simk_plugflow_f64); the host prepares each step's source and pulse clock.
State hashes 9beb634425faf2c2 (f64 lanes) and 32256818f6f35bc9 (f32 delta
lanes), the values docs/BUILD.md:125-128 gives for the reference's CPU, GPU
and hybrid runs."""
import ctypes as C

import pytest

from simstack_amd import _ffi, dbd, gpu, problem

pytestmark = pytest.mark.gpu

HASHES = {0: 0x9BEB634425FAF2C2, 1: 0x32256818F6F35BC9}


# The f32 table walk is left out: one launch per master step of a single
# wave, each vector update a dependent workspace round trip, takes minutes
# for the 500 steps (it is checked bit for bit on 5-step runs in
# test_gpu_dbd.py).
@pytest.mark.parametrize("precision,kernel", [(0, 1), (0, 2), (1, 1)], ids=["f64-generated", "f64-table", "f32-generated"])
def test_scene_hash_known_answer_device(gm, precision, kernel):
    print(f"[case] dbd scene, 64 zones x 500 steps, precision {precision}, kernel {kernel}", flush=True)
    air = dbd.AirMechanism(caps=_ffi.DEFAULT_CAPS)
    meth, ctl = problem.method(2), dbd.default_control()
    gm.arena_reset()
    gm.set_kernel(kernel)
    try:
        gm._check(gm.lib.gpu_kin_set_mech(air.bytes.ctypes.data, C.byref(meth), C.byref(ctl)), "gpu_kin_set_mech")
        sc = dbd.Scene(air)
        dz = gpu.DeviceZones(gm, sc.z)
        d_feed = gm.alloc(8 * air.n)
        gm.to_device(d_feed, sc.feed())

        def advance(z, dt):
            dz.upload()
            dz.advance(dt, precision=precision)

        def flow(z, a, feed):
            gm._check(gm.lib.simk_plugflow_f64(dz.d_block + z.offsets[0] * 8, z.padded, z.n, z.count, a, d_feed),
                      "simk_plugflow_f64")
            dz.download(mutable_only=True)

        for _ in range(500):
            sc.step(advance, flow)
    finally:
        gm.set_kernel(0)
    assert sc.state_hash() == HASHES[precision], f"{sc.state_hash():016x}"
    assert sc.z.failed[:64].sum() == 0
