"""simk_kin_advance_static (the advance that allocates no device memory: its
workspace and queue counter are the library's fixed scratch array) against
simk_kin_advance_n, the entry point it is a variant of, through the C ABI:
bit-identical solver rows on two calls in a row, the second with a smaller
block, and the grid lowered to what the 64 MiB array holds."""
import ctypes as C

import numpy as np
import pytest

from simstack_amd import gpu, problem
from helpers import compare_bits, make_problem

pytestmark = pytest.mark.gpu

ZERO = {k: 0 for k in ("y", "h", "steps", "rejects", "pulses", "failed")}
SCRATCH = 64 << 20


def _run(gm, mech, meth, ctl, z, duration, repeats, static, precision=0):
    dz = gpu.DeviceZones(gm, z)
    dz.upload()
    L = dz.launch_desc(duration, precision=precision)
    ms = C.c_float()
    fn = gm.lib.simk_kin_advance_static if static else gm.lib.simk_kin_advance_n
    gm._check(fn(C.byref(L), repeats, C.byref(ms)), "simk_kin_advance_static" if static else "simk_kin_advance_n")
    info = gm.kin_info()
    dz.download(mutable_only=False)
    return info


@pytest.mark.parametrize("precision", [0, 1, 2, 3], ids=["fp64", "f32delta", "mixed", "fp32store"])
def test_static_advance_equals_advance_n_twice(gm, precision):
    sp, mech, z1 = make_problem("c2", 9000)
    _, _, z2 = make_problem("c2", 700, first_cell=20000)
    meth = problem.method(2)
    ctl = problem.control(rtol=1e-3, atol=1e-10, max_steps=1000) if precision == 1 else problem.control()
    gm.arena_reset()
    gm.set_kernel(0)
    gm.set_grid(0)
    gm.set_mech(mech, meth, ctl)
    for z, repeats in ((z1, 2), (z2, 1)):
        print(f"[case] precision {precision}, {z.count} cells x {repeats} x 0.1 s", flush=True)
        a, b = z.copy(), z.copy()
        gm.arena_reset()
        ref = _run(gm, mech, meth, ctl, a, 0.1, repeats, static=False, precision=precision)
        got = _run(gm, mech, meth, ctl, b, 0.1, repeats, static=True, precision=precision)
        assert compare_bits(b, a) == ZERO
        assert got.kernel == ref.kernel
        # the static launch's grid holds its workspace in the scratch array
        assert got.waves >= 1 and got.waves * got.ws_slots * 64 * (4 if got.kernel == 5 else 8) <= SCRATCH - 256
        assert got.waves <= ref.waves
        assert a.steps[:z.count].min() > 0


def test_static_advance_before_the_kernel_is_loaded(gm):
    """A precision whose kernel has not been loaded yet is refused (loading
    allocates), and the rows are left untouched."""
    sp, mech, z = make_problem("c1", 256)
    meth, ctl = problem.method(2), problem.control()
    gm.arena_reset()
    gm.set_kernel(0)
    gm.set_mech(mech, meth, ctl)
    dz = gpu.DeviceZones(gm, z.copy())
    dz.upload()
    for precision in (2, 3):
        L = dz.launch_desc(0.1, precision=precision)
        ms = C.c_float()
        rc = gm.lib.simk_kin_advance_static(C.byref(L), 1, C.byref(ms))
        assert rc != 0 and gm.lib.gpu_error_string(rc)
        dz.download(mutable_only=False)
        assert dz.z.block.tobytes() == z.block.tobytes()
