"""Generated kernels for seeded random mechanisms beyond the fixed workloads,
bit for bit against the oracle: odd species counts (the paired and quad
workspace layouts' odd tails), a mechanism at the 64-species capacity (the
FP32 factor-storage pivots take 32 KiB of LDS), and small ones whose whole LU
fits in LDS. FP64 against oracle_advance (the reference's kin_advance,
src/sim/kinetics/kin_solver.h:189-347); FP32 factor storage against
oracle_advance_store32. Each mechanism's kernels are compiled on first use."""
import pytest

import oracle_ffi
from simstack_amd import gpu, problem
from helpers import add_pulses, compare_bits, make_problem

pytestmark = pytest.mark.gpu

ZERO = {k: 0 for k in ("y", "h", "steps", "rejects", "pulses", "failed")}

# (seed, species, reactions, fill range)
MECHS = [(101, 9, 20, (0.2, 0.5)), (102, 17, 44, (0.1, 0.3)), (103, 33, 90, (0.1, 0.25)),
         (104, 47, 160, (0.12, 0.2)), (105, 64, 230, (0.08, 0.15))]


@pytest.mark.parametrize("precision", [0, 3], ids=["fp64", "fp32store"])
@pytest.mark.parametrize("seed,n,R,fill", MECHS, ids=[f"{n}x{R}" for _, n, R, _ in MECHS])
def test_random_mechanism_bit_exact(gm, seed, n, R, fill, precision):
    print(f"[case] random mechanism {n}/{R} seed {seed}, precision {precision}", flush=True)
    sp, mech, z = make_problem("c1", 200, seed=seed, n_species=n, n_reactions=R, density_lo=fill[0],
                               density_hi=fill[1])
    add_pulses(z, seed=seed, duration=0.1)
    meth, ctl = problem.method(2), problem.control(rtol=1e-5, atol=1e-10)
    a, b = z.copy(), z.copy()
    gm.arena_reset()
    gm.set_kernel(1)
    gm.set_mech(mech, meth, ctl)
    dz = gpu.DeviceZones(gm, a)
    dz.upload()
    for _ in range(2):
        dz.advance(0.05, precision=precision)
    dz.download(mutable_only=False)
    assert gm.kin_info().kernel == (9 if precision == 3 else 1)
    adv = oracle_ffi.oracle_advance_store32 if precision == 3 else oracle_ffi.oracle_advance
    for _ in range(2):
        adv(mech, meth, ctl, b, 0.05, threads=16)
    assert compare_bits(a, b) == ZERO
    assert b.steps[:200].min() > 0
