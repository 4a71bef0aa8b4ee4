"""Ragged batches on the device: three mechanisms (60/200, 40/130, 24/70) in
one persistent launch, and the same launches captured in a HIP graph. Every
cell must come out bit for bit as the oracle's scalar FP64 kin_advance
(src/sim/kinetics/kin_solver.h:322-347) gives it on its own, whichever
mechanism's waves ran it: the reference's partition-independence test
(src/tests/test_kinetics.cpp:1204-1243) asks the same of its lane groups."""
import numpy as np
import pytest

from simstack_amd import gpu, problem
from helpers import compare_bits, make_problem, oracle_run

pytestmark = pytest.mark.gpu

ZERO = {k: 0 for k in ("y", "h", "steps", "rejects", "pulses", "failed")}
NAMES = ("c3a", "c3b", "c3c")


def _blocks(counts, first=0):
    out = []
    for name, cells in zip(NAMES, counts):
        sp, mech, z = make_problem(name, cells, first_cell=first)
        out.append((mech, z))
    return out


def _set(gm, blocks):
    meth, ctl = problem.method(2), problem.control()
    gm.arena_reset()
    gm.set_mechs([m for m, _ in blocks], meth, ctl)
    return meth, ctl


@pytest.mark.parametrize("pair,carry", [(0, 0), (0xFFFFFFFF, 0), (0, 1)],
                         ids=["paired-slots", "single-slots", "stage-carry"])
def test_three_mechanisms_one_launch(gm, pair, carry):
    """Ragged cell counts (not multiples of 64), two operator-split calls;
    workspace slots paired (the default) or single; the stage carry off (the
    default) or on."""
    try:
        gm.set_gen_options(pair_slots=pair, stage_carry=carry)
        blocks = _blocks((200, 333, 130))
        meth, ctl = _set(gm, blocks)
        work = [z.copy() for _, z in blocks]
        dzs = [gpu.DeviceZones(gm, z) for z in work]
        for d in dzs:
            d.upload()
        gm.advance_multi(dzs, 0.25)
        gm.advance_multi(dzs, 0.25)
        for d in dzs:
            d.download(mutable_only=False)
        assert gm.kin_info().kernel == 4
        for (mech, z), got in zip(blocks, work):
            want = z.copy()
            oracle_run(mech, meth, ctl, want, [0.25, 0.25])
            assert compare_bits(got, want) == ZERO
            assert want.steps[:want.count].max() >= 5 * want.steps[:want.count].min()
    finally:
        gm.set_gen_options()


def test_multi_with_lds_lu_entries(gm):
    """The multi-mechanism kernel with LU entries in LDS (one array shared by
    the mechanisms a wave runs in turn): bit-identical to the oracle."""
    try:
        gm.set_gen_options(lds_slots=96)
        blocks = _blocks((150, 260, 190), first=500)
        meth, ctl = _set(gm, blocks)
        work = [z.copy() for _, z in blocks]
        dzs = [gpu.DeviceZones(gm, z) for z in work]
        for d in dzs:
            d.upload()
        gm.advance_multi(dzs, 0.25, repeats=2)
        for d in dzs:
            d.download(mutable_only=False)
        for (mech, z), got in zip(blocks, work):
            want = z.copy()
            oracle_run(mech, meth, ctl, want, [0.25, 0.25])
            assert compare_bits(got, want) == ZERO
    finally:
        gm.set_gen_options()


def test_multi_with_refill_repeats_and_small_grid(gm):
    """Three resident waves for 700 cells (every wave crosses mechanisms and
    refills many times) with two fused operator-split repeats."""
    blocks = _blocks((64, 500, 136), first=1000)
    meth, ctl = _set(gm, blocks)
    work = [z.copy() for _, z in blocks]
    dzs = [gpu.DeviceZones(gm, z) for z in work]
    for d in dzs:
        d.upload()
    try:
        gm.set_grid(3)
        gm.advance_multi(dzs, 0.2, repeats=2)
    finally:
        gm.set_grid(0)
    for d in dzs:
        d.download(mutable_only=False)
    for (mech, z), got in zip(blocks, work):
        want = z.copy()
        oracle_run(mech, meth, ctl, want, [0.2, 0.2])
        assert compare_bits(got, want) == ZERO


def test_graph_replays_match_eager(gm):
    """A graph of two captured steps, replayed three times with eager launches
    and host synchronisation between the replays; after every replay the
    graph's blocks equal the eager blocks bit for bit, and the first equals
    the oracle."""
    blocks = _blocks((96, 160, 64), first=5000)
    meth, ctl = _set(gm, blocks)
    g_work = [z.copy() for _, z in blocks]
    e_work = [z.copy() for _, z in blocks]
    g_dzs = [gpu.DeviceZones(gm, z) for z in g_work]
    e_dzs = [gpu.DeviceZones(gm, z) for z in e_work]
    for d in g_dzs + e_dzs:
        d.upload()
    gid = gm.capture_multi(g_dzs, 0.1, 1, 2)
    try:
        for rep in range(3):
            gm.replay(gid)
            gm.advance_multi(e_dzs, 0.1)
            gm.advance_multi(e_dzs, 0.1)
            for d in g_dzs + e_dzs:
                d.download(mutable_only=False)
            for a, b in zip(g_work, e_work):
                assert compare_bits(a, b) == ZERO, f"replay {rep}"
            if rep == 0:
                for (mech, z), got in zip(blocks, g_work):
                    want = z.copy()
                    oracle_run(mech, meth, ctl, want, [0.1, 0.1])
                    assert compare_bits(got, want) == ZERO
    finally:
        gm.destroy_graph(gid)


def test_multi_rejects_bad_launches(gm):
    """Count mismatch and a precision other than FP64 are refused with an
    error code, by the advance and by the autotuner."""
    blocks = _blocks((64, 64, 64))
    _set(gm, blocks)
    dzs = [gpu.DeviceZones(gm, z.copy()) for _, z in blocks]
    with pytest.raises(gpu.GpuError):
        gm.advance_multi(dzs[:2], 0.1)
    import ctypes as C
    from simstack_amd import _ffi
    Ls = (_ffi.GpuKinLaunch * 3)(*[d.launch_desc(0.1, precision=2) for d in dzs])
    assert gm.lib.simk_kin_advance_multi(3, Ls, 1, None) != 0
    assert np.all(np.isfinite(dzs[0].z.y))
    # the multi-mechanism autotuner refuses the same launches
    res = gpu.SimkTuneResult()
    assert gm.lib.simk_kin_autotune_multi(3, Ls, 64, 0.01, C.byref(res)) != 0
    with pytest.raises(gpu.GpuError):
        gm.autotune_multi(dzs[:2], duration=0.01, sample_cells=64)
    # misaligned or null copies are refused before any launch
    ms = C.c_float()
    assert gm.lib.simk_copy_device_b64(None, None, 1024, C.byref(ms)) != 0
    assert gm.lib.simk_copy_device_b64(dzs[0].d_block + 8, dzs[1].d_block, 1024, C.byref(ms)) != 0


def test_multi_autotune_keeps_bits_and_restores(gm):
    """simk_kin_autotune_multi times candidates on a copy of the blocks (the
    callers' rows are untouched), keeps the fastest, and the tuned kernel's
    results are the oracle's."""
    blocks = _blocks((120, 300, 260), first=2000)
    meth, ctl = _set(gm, blocks)
    work = [z.copy() for _, z in blocks]
    dzs = [gpu.DeviceZones(gm, z) for z in work]
    for d in dzs:
        d.upload()
    res = gm.autotune_multi(dzs, duration=0.05, sample_cells=128)
    assert res.candidates >= 6 and res.lanes_per_cell in (1, 8, 16, 32, 64) and res.best_ms > 0.0
    assert res.best_ms <= res.default_ms
    for d, (_, z) in zip(dzs, blocks):
        d.download(mutable_only=False)
        assert d.z.block.tobytes() == z.block.tobytes()   # tuning ran on copies
    gm.advance_multi(dzs, 0.25)
    for d in dzs:
        d.download(mutable_only=False)
    for (mech, z), got in zip(blocks, work):
        want = z.copy()
        oracle_run(mech, meth, ctl, want, [0.25])
        assert compare_bits(got, want) == ZERO
    _set(gm, blocks)   # restores the default settings for the tests after this one


@pytest.mark.parametrize("team", [64, 32, 16, 8])
def test_multi_teams_bit_exact(gm, team):
    """The cooperative form of a ragged launch (gpu_kin_set_mode(1) with
    simk_kin_set_team): teams of `team` lanes walk the three mechanisms'
    queues in one launch; every cell bit-identical to the oracle, with a
    fused pair of repeats."""
    print(f"[case] ragged launch in teams of {team} lanes", flush=True)
    blocks = _blocks((70, 150, 90), first=3000)
    meth, ctl = _set(gm, blocks)
    work = [z.copy() for _, z in blocks]
    dzs = [gpu.DeviceZones(gm, z) for z in work]
    for d in dzs:
        d.upload()
    gm._check(gm.lib.gpu_kin_set_mode(1), "gpu_kin_set_mode")
    gm._check(gm.lib.simk_kin_set_team(team), "simk_kin_set_team")
    try:
        gm.advance_multi(dzs, 0.05, repeats=2)
        name = gm.kin_info().kernel_name
    finally:
        gm.lib.gpu_kin_set_mode(0)
        gm.lib.simk_kin_set_team(64)
    for d in dzs:
        d.download(mutable_only=False)
    assert f"{team} lanes".encode() in name
    for (mech, z), got in zip(blocks, work):
        want = z.copy()
        oracle_run(mech, meth, ctl, want, [0.05, 0.05])
        assert compare_bits(got, want) == ZERO
