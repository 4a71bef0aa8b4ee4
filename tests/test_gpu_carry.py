"""Stage carry (SimkGenOptions.stage_carry): a stage whose combination ends
with the previous stage's K takes that term from the registers that still
hold it. Every element sees the same fma sequence, so the bits must not move:
FP64 against oracle_advance (the reference's kin_advance,
src/sim/kinetics/kin_solver.h:189-347, stage loops :238-259) and FP32 factor
storage against oracle_advance_store32, with the carry on (the default,
off, is what every other parity test runs). Ros2
carries a stage point; Ros3 a point and a right-hand side with an earlier
term; Rodas3 a right-hand side and a point with an earlier term."""
import pytest

import oracle_ffi
from simstack_amd import gpu, problem
from helpers import add_pulses, compare_bits, make_problem

pytestmark = pytest.mark.gpu

ZERO = {k: 0 for k in ("y", "h", "steps", "rejects", "pulses", "failed")}


def _device(gm, mech, meth, ctl, z, durations, precision, **gen):
    gm.arena_reset()
    gm.set_kernel(1)
    gm.set_gen_options(**gen)
    try:
        gm.set_mech(mech, meth, ctl)
        dz = gpu.DeviceZones(gm, z)
        dz.upload()
        for d in durations:
            dz.advance(d, precision=precision)
        dz.download(mutable_only=False)
        info = gm.kin_info()
        assert info.kernel == (9 if precision == 3 else 1), info.kernel
        print(f"[device] {info.kernel_name.decode()}: {z.count} cells", flush=True)
    finally:
        gm.set_gen_options()


@pytest.mark.parametrize("precision,carry", [(0, 1), (3, 1)], ids=["fp64", "fp32store"])
@pytest.mark.parametrize("mid", [0, 1, 2], ids=["ros2", "ros3", "rodas3"])
def test_stage_carry_keeps_bits(gm, mid, precision, carry):
    print(f"[case] stage carry {carry:#x}, precision {precision}, c1 pulses, method {mid}", flush=True)
    sp, mech, z = make_problem("c1", 128)
    add_pulses(z, seed=10 + mid)
    meth, ctl = problem.method(mid), problem.control(rtol=1e-5, atol=1e-10)
    a, b = z.copy(), z.copy()
    _device(gm, mech, meth, ctl, a, [0.1, 0.1], precision, stage_carry=carry)
    adv = oracle_ffi.oracle_advance_store32 if precision == 3 else oracle_ffi.oracle_advance
    for d in (0.1, 0.1):
        adv(mech, meth, ctl, b, d, threads=16)
    assert compare_bits(a, b) == ZERO
    assert b.pulses[:128].sum() > 0


@pytest.mark.parametrize("precision,carry", [(0, 1), (3, 1)], ids=["fp64", "fp32store"])
def test_stage_carry_48x150_keeps_bits(gm, precision, carry):
    print(f"[case] stage carry {carry:#x}, precision {precision}, c2 1024 cells x 2 x 0.2 s", flush=True)
    sp, mech, z = make_problem("c2", 1024)
    meth, ctl = problem.method(2), problem.control()
    a, b = z.copy(), z.copy()
    _device(gm, mech, meth, ctl, a, [0.2, 0.2], precision, stage_carry=carry)
    adv = oracle_ffi.oracle_advance_store32 if precision == 3 else oracle_ffi.oracle_advance
    for d in (0.2, 0.2):
        adv(mech, meth, ctl, b, d, threads=16)
    assert compare_bits(a, b) == ZERO
