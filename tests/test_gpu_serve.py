"""The serving loop (simk_serve_*): device-resident requests, HBM-budget
admission and continuous batching over the three ragged mechanisms.

A request's result must not depend on the traffic around it: served alone,
served in a mix (other mechanisms and sizes admitted before and after it, so
its cells sit at other lanes of other waves), and computed by the oracle as
the operator split it is — nft_transport's upwind row update
(src/sim/nft.cpp:1055-1081) then kin_advance (src/sim/kinetics/kin_solver.h:
322-347) per step — the bits are the same."""
import numpy as np
import pytest

import oracle_ffi
from simstack_amd import gpu, problem, serving
from helpers import compare_bits

pytestmark = pytest.mark.gpu

STEPS, SUBSTEPS, DT = 5, 3, 1e-3


@pytest.fixture(scope="module")
def pools():
    return serving.Pools(cells=6000)


def _serve(gm, pools, reqs, budget=8 << 30, submit_after=None, slice_us=0):
    """reqs: list of (mech, first, channels, nodes). submit_after[i] = iterations
    to run before submitting request i (default 0). Returns the results."""
    meth, ctl = problem.method(2), problem.control()
    gm.set_mechs(pools.mechs, meth, ctl)
    gm.serve_open(hbm_budget=budget, max_requests=16, steps=STEPS, substeps=SUBSTEPS, dt=DT, slice_us=slice_us)
    submit_after = submit_after or [0] * len(reqs)
    results, ticket_of = {}, {}
    it = 0
    pending = list(range(len(reqs)))
    try:
        while pending or ticket_of:
            for i in [i for i in pending if submit_after[i] <= it]:
                mech, first, channels, nodes = reqs[i]
                cells = channels * nodes
                ptr, padded = pools.window(mech, first, cells)
                res = np.full((pools.mechs[mech].n + 5, cells), np.nan)
                t = gm.serve_submit(mech, ptr, padded, cells, channels, nodes,
                                    serving.courant_profile(channels, nodes), res, pools.feed(mech))
                ticket_of[t] = i
                results[i] = res
                pending.remove(i)
            for t in gm.serve_step():
                ticket_of.pop(t)
            it += 1
            assert it < 20000
        st = gm.serve_stats()
    finally:
        gm.serve_close()
    return results, st


def _oracle(pools, mech, first, channels, nodes):
    m, pool = pools.mechs[mech], pools.zones[mech]
    cells = channels * nodes
    z = problem.Zones(m.n, m.R, cells)
    for name in ("y", "k", "src", "pulse"):
        getattr(z, name)[:, :cells] = getattr(pool, name)[:, first:first + cells]
        getattr(z, name)[:, cells:] = getattr(z, name)[:, :1]
    c = serving.courant_profile(channels, nodes)
    feed = pools.feed(mech)
    meth, ctl = problem.method(2), problem.control()
    lib = oracle_ffi.oracle()
    for _ in range(STEPS):
        for ch in range(channels):
            sub = np.ascontiguousarray(z.y[:, ch * nodes:(ch + 1) * nodes])
            cs = np.ascontiguousarray(c[ch * nodes:(ch + 1) * nodes])
            lib.oracle_upwind_f64(sub.ctypes.data, cs.ctypes.data, nodes, m.n, nodes, SUBSTEPS, feed.ctypes.data)
            z.y[:, ch * nodes:(ch + 1) * nodes] = sub
        oracle_ffi.oracle_advance(m, meth, ctl, z, DT, lanes=np.arange(cells), threads=8)
    return z


def _as_zones(pools, mech, res):
    """Result rows (y, h, steps, rejects, pulses, failed) as a Zones block."""
    m = pools.mechs[mech]
    cells = res.shape[1]
    z = problem.Zones(m.n, m.R, cells)
    z.y[:, :cells] = res[:m.n]
    for j, name in enumerate(("h", "steps", "rejects", "pulses", "failed")):
        getattr(z, name)[:cells] = res[m.n + j]
    return z


def test_request_same_alone_in_mix_and_oracle(gm, pools):
    x = (2, 100, 1, 300)                       # 24/70 mechanism, one 300-node channel
    alone, _ = _serve(gm, pools, [x])
    mix_reqs = [(0, 0, 2, 500), (1, 50, 1, 640), x, (2, 1000, 3, 700), (0, 2000, 1, 64)]
    mixed, st = _serve(gm, pools, mix_reqs, submit_after=[0, 0, 2, 3, 1])
    assert st.completed == 5 and st.queued == 0 and st.active == 0
    assert alone[0].tobytes() == mixed[2].tobytes()
    want = _oracle(pools, *x)
    got = _as_zones(pools, 2, alone[0])
    assert compare_bits(got, want) == {k: 0 for k in ("y", "h", "steps", "rejects", "pulses", "failed")}
    assert np.all(want.steps[:300] >= STEPS)
    # the other requests of the mix against the oracle too (two channels, other mechanisms)
    for i in (0, 1):
        want = _oracle(pools, *mix_reqs[i])
        got = _as_zones(pools, mix_reqs[i][0], mixed[i])
        assert compare_bits(got, want)["y"] == 0


def test_time_slices_keep_bits(gm, pools):
    """Slices of 1 ms stop cells mid-step many times (their state saved to
    the block and resume rows, continued in later launches); the results equal
    the unsliced run's bit for bit."""
    reqs = [(0, 0, 1, 2000), (2, 100, 1, 300), (1, 0, 2, 400)]
    whole, st_whole = _serve(gm, pools, reqs, slice_us=0xFFFFFFFF)
    sliced, st = _serve(gm, pools, reqs, slice_us=1000)
    assert st_whole.carried == 0
    assert st.carried > 0 and st.iterations > st_whole.iterations
    for i in range(len(reqs)):
        assert whole[i].tobytes() == sliced[i].tobytes()


def test_budget_admission_serialises_and_keeps_results(gm, pools):
    reqs = [(1, 0, 2, 500), (1, 1000, 2, 500), (0, 3000, 1, 1000)]
    free, st_free = _serve(gm, pools, reqs)
    # zone rows (3n + R + 8) + courant + 5 resume rows + 24 DBD-request rows
    rows1 = 3 * pools.mechs[1].n + pools.mechs[1].R + 38
    rows0 = 3 * pools.mechs[0].n + pools.mechs[0].R + 38
    budget = max(1000 * rows1 * 8, 1000 * rows0 * 8)   # one request at a time
    tight, st = _serve(gm, pools, reqs, budget=budget)
    assert st.peak_admitted_bytes <= budget
    assert st_free.peak_admitted_bytes > budget
    assert st.iterations >= 3 * STEPS
    for i in range(3):
        assert free[i].tobytes() == tight[i].tobytes()


def test_serve_rejects_bad_requests(gm, pools):
    meth, ctl = problem.method(2), problem.control()
    gm.set_mechs(pools.mechs, meth, ctl)
    gm.serve_open(hbm_budget=1 << 30, max_requests=2, steps=1, substeps=1, dt=DT)
    try:
        ptr, padded = pools.window(0, 0, 100)
        res = np.zeros((pools.mechs[0].n + 5, 100))
        c = serving.courant_profile(1, 100)
        with pytest.raises(gpu.GpuError):          # channels * nodes != cells
            req = gpu.SimkServeRequest(0, 100, 2, 60, padded, 0, ptr, c.ctypes.data, None, res.ctypes.data)
            import ctypes as C
            t = C.c_uint32()
            gm._check(gm.lib.simk_serve_submit(C.byref(req), C.byref(t)), "submit")
        with pytest.raises(gpu.GpuError):          # unknown mechanism
            gm.serve_submit(7, ptr, padded, 100, 1, 100, c, res)
        with pytest.raises(gpu.GpuError):          # larger than the budget
            big = 600000
            gm.serve_submit(0, ptr, big, big, 100, 6000, np.zeros(big), np.zeros((pools.mechs[0].n + 5, big)))
        gm.serve_submit(0, ptr, padded, 100, 1, 100, c, res)
        gm.serve_submit(0, ptr, padded, 100, 1, 100, c, np.zeros((pools.mechs[0].n + 5, 100)))
        with pytest.raises(gpu.GpuError):          # max_requests in flight
            gm.serve_submit(0, ptr, padded, 100, 1, 100, c, np.zeros((pools.mechs[0].n + 5, 100)))
        done = []
        while len(done) < 2:
            done += gm.serve_step()
        assert np.all(res[pools.mechs[0].n + 1] >= 1)
    finally:
        gm.serve_close()
