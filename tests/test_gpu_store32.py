"""FP32 factor storage (GpuKinLaunch precision 3) on the device, bit for bit
against the oracle's restatement of that form (oracle_advance_store32: the
reference's FP64 kin_integrate, src/sim/kinetics/kin_solver.h:189-313, with
each finished factor row's off-diagonal entries rounded to f32), and against
the FP64 path by the written accuracy criterion: the p99.9 over cells of each
cell's maximum relative error (species above atol/rtol) is at most 1e-6."""
import numpy as np
import pytest

from simstack_amd import gpu, problem
import oracle_ffi
from helpers import add_pulses, checker, compare_bits, make_problem

pytestmark = pytest.mark.gpu

ZERO = {k: 0 for k in ("y", "h", "steps", "rejects", "pulses", "failed")}
CRITERION_P999 = 1e-6


def _device(gm, mech, meth, ctl, z, durations, repeats=1, **gen):
    gm.arena_reset()
    gm.set_kernel(0)
    gm.set_gen_options(**gen)
    try:
        gm.set_mech(mech, meth, ctl)
        dz = gpu.DeviceZones(gm, z)
        dz.upload()
        for d in durations:
            dz.advance(d, precision=3, repeats=repeats)
        dz.download(mutable_only=False)
        info = gm.kin_info()
        assert info.kernel == 9, info.kernel
        print(f"[device] {info.kernel_name.decode()}: {z.count} cells, {info.ws_slots} slots per lane, "
              f"{info.waves} waves", flush=True)
        return info
    finally:
        gm.set_gen_options()


def _oracle(mech, meth, ctl, z, durations, repeats=1):
    for d in durations:
        for _ in range(repeats):
            oracle_ffi.oracle_advance_store32(mech, meth, ctl, z, d, threads=16)


def test_store32_c1_bit_exact(gm):
    print("[case] store32 c1 256 cells x 1 s", flush=True)
    sp, mech, z = make_problem("c1", 256)
    meth, ctl = problem.method(2), problem.control()
    a, b = z.copy(), z.copy()
    _device(gm, mech, meth, ctl, a, [1.0])
    _oracle(mech, meth, ctl, b, [1.0])
    assert compare_bits(a, b) == ZERO
    assert b.steps[:256].min() > 10 and b.failed[:256].sum() == 0


@pytest.mark.parametrize("mid", [0, 1, 2], ids=["ros2", "ros3", "rodas3"])
def test_store32_methods_with_pulses(gm, mid):
    print(f"[case] store32 c1 pulses method {mid}", flush=True)
    sp, mech, z = make_problem("c1", 128)
    add_pulses(z, seed=mid)
    meth, ctl = problem.method(mid), problem.control(rtol=1e-5, atol=1e-10)
    a, b = z.copy(), z.copy()
    _device(gm, mech, meth, ctl, a, [0.1, 0.1])
    _oracle(mech, meth, ctl, b, [0.1, 0.1])
    assert compare_bits(a, b) == ZERO
    assert b.pulses[:128].sum() > 0


@pytest.mark.parametrize("lds", [0xFFFFFFFF, 160, 0, 640], ids=["pivots_only", "160", "default", "all"])
def test_store32_lds_placements_keep_bits(gm, lds):
    """The LDS budget moves off-diagonal entries between LDS and the f32
    workspace region; the values and operations are unchanged."""
    print(f"[case] store32 c2 1024 cells x 2 x 0.2 s, lds budget {lds:#x}", flush=True)
    sp, mech, z = make_problem("c2", 1024)
    meth, ctl = problem.method(2), problem.control()
    a, b = z.copy(), z.copy()
    info = _device(gm, mech, meth, ctl, a, [0.2, 0.2], lds_slots=lds)
    off = mech.nnz - mech.n
    in_lds = {0xFFFFFFFF: 0, 160: 160 - 2 * mech.n, 0: 304 - 2 * mech.n, 640: off}[lds]
    # K stages and f, the paired copy of the rate constants, the single copy and the sources, the f32 quads
    assert info.ws_slots == 3 * 48 + 48 + 150 + 150 + 48 + 2 * ((off - in_lds + 3) // 4)
    _oracle(mech, meth, ctl, b, [0.2, 0.2])
    assert compare_bits(a, b) == ZERO


def test_store32_single_slots_keep_bits(gm):
    """Single 8-byte workspace slots (no pairs, no quads, one copy of the rate
    constants): another layout of the same values and operations."""
    print("[case] store32 c2 1024 cells x 2 x 0.2 s, single slots", flush=True)
    sp, mech, z = make_problem("c2", 1024)
    meth, ctl = problem.method(2), problem.control()
    a, b = z.copy(), z.copy()
    info = _device(gm, mech, meth, ctl, a, [0.2, 0.2], pair_slots=0xFFFFFFFF)
    off = mech.nnz - mech.n
    assert info.ws_slots == 3 * 48 + 48 + 150 + 48 + (off - (304 - 2 * mech.n) + 1) // 2
    _oracle(mech, meth, ctl, b, [0.2, 0.2])
    assert compare_bits(a, b) == ZERO


def test_store32_autotune_keeps_bits(gm):
    """simk_kin_autotune with precision 3 times the FP32 factor-storage
    kernel's LDS budgets and knobs on a copy of the block (left untouched);
    the tuned kernel's advance equals the oracle form bit for bit, and the
    FP64 settings are not touched."""
    print("[case] store32 autotune, 24/70, 3000 cells", flush=True)
    sp, mech, z = make_problem("c3c", 3000)
    meth, ctl = problem.method(2), problem.control()
    a, b = z.copy(), z.copy()
    gm.arena_reset()
    gm.set_kernel(0)
    gm.set_mech(mech, meth, ctl)
    dz = gpu.DeviceZones(gm, a)
    dz.upload()
    res = dz.autotune(duration=0.02, precision=3)
    assert res.candidates >= 10 and 0.0 < res.best_ms <= res.default_ms
    dz.download(mutable_only=False)
    assert a.block.tobytes() == z.block.tobytes()     # tuning ran on a copy
    dz.advance(0.2, precision=3)
    dz.download(mutable_only=False)
    assert gm.kin_info().kernel == 9
    _oracle(mech, meth, ctl, b, [0.2])
    assert compare_bits(a, b) == ZERO
    gm.set_gen_options()


def test_store32_ragged_three_mechanisms(gm):
    """The ragged batch (60/200, 40/130, 24/70) in one multi-mechanism launch
    in FP32 factor storage, two advances: every cell of every mechanism equals
    the oracle form."""
    print("[case] store32 ragged 60/200 + 40/130 + 24/70, 2 x 0.5 s", flush=True)
    blocks = [make_problem(name, cells) for name, cells in (("c3a", 300), ("c3b", 700), ("c3c", 1500))]
    meth, ctl = problem.method(2), problem.control()
    gm.arena_reset()
    gm.set_grid(0)
    gm.set_mechs([m for _, m, _ in blocks], meth, ctl)
    work = [z.copy() for _, _, z in blocks]
    dzs = [gpu.DeviceZones(gm, z) for z in work]
    for d in dzs:
        d.upload()
    for _ in range(2):
        gm.advance_multi(dzs, 0.5, precision=3)
    assert gm.kin_info().kernel == 10
    for d in dzs:
        d.download(mutable_only=False)
    for (_, mech, z), got in zip(blocks, work):
        want = z.copy()
        _oracle(mech, meth, ctl, want, [0.5, 0.5])
        assert compare_bits(got, want) == ZERO


def test_store32_c2_fused_repeats(gm):
    """4,096 cells of the 48/150 workload, 3 fused operator-split advances of 1 s."""
    print("[case] store32 c2 4096 cells x 3 x 1 s fused", flush=True)
    sp, mech, z = make_problem("c2", 4096)
    meth, ctl = problem.method(2), problem.control()
    a, b = z.copy(), z.copy()
    _device(gm, mech, meth, ctl, a, [1.0], repeats=3)
    _oracle(mech, meth, ctl, b, [1.0], repeats=3)
    assert compare_bits(a, b) == ZERO


@pytest.mark.timeout(900)
def test_store32_65536_cells_10s_every_cell(gm):
    """The FP64 throughput workload's size: 65,536 cells x 10 fused advances of 1 s, every cell."""
    print("[case] store32 c2 65,536 cells x 10 x 1 s fused, every cell", flush=True)
    sp, mech, z = make_problem("c2", 65536)
    meth, ctl = problem.method(2), problem.control()
    a, b = z.copy(), z.copy()
    _device(gm, mech, meth, ctl, a, [1.0], repeats=10)
    _oracle(mech, meth, ctl, b, [1.0], repeats=10)
    assert compare_bits(a, b) == ZERO
    assert b.failed[:65536].sum() == 0


def test_store32_accuracy_criterion(gm):
    """8,192 cells x 10 x 1 s against the FP64 checker: p99.9 cell error <= 1e-6."""
    print("[case] store32 accuracy, 8192 cells x 10 x 1 s vs FP64", flush=True)
    cells = 8192
    sp, mech, z = make_problem("c2", cells)
    meth, ctl = problem.method(2), problem.control()
    a, b = z.copy(), z.copy()
    _device(gm, mech, meth, ctl, a, [1.0], repeats=10)
    checker(mech, 2, ctl, b, [1.0] * 10)
    x, y = a.y[:, :cells], b.y[:, :cells]
    d = np.abs(x - y)
    s = np.maximum(np.abs(x), np.abs(y))
    r = np.divide(d, s, out=np.zeros_like(d), where=s > 0)
    e = np.where(y > ctl.atol / ctl.rtol, r, 0.0).max(axis=0)
    p999 = float(np.percentile(e, 99.9))
    same = float(np.mean((a.steps[:cells] == b.steps[:cells]) & (a.rejects[:cells] == b.rejects[:cells])))
    print(f"FP32 factor storage, {cells} cells x 10 s: cell p99.9 {p999:.3e}, max {e.max():.3e}, "
          f"cells above 1e-6 {np.mean(e > 1e-6):.4f}, same step counts {same:.3f}", flush=True)
    assert a.failed[:cells].sum() == 0
    assert p999 <= CRITERION_P999
