"""DBD reactor requests in the serving loop (SimkServeRequest.kind =
SimkServe_Dbd): the reference's chained reactor scene (64 zones at the
scene's operating point, 500 master steps of 1 ms, pulsed discharge, plug
flow over a 2-s residence) served device-resident, with each step's
discharge source and pulse clock prepared on the device. Its state hash is
the reference's published `--scene dbd --headless 30` value
9beb634425faf2c2 (f64 lanes, docs/BUILD.md:125-128), alone and while
transport requests of another mechanism share the launches."""
import numpy as np
import pytest

from simstack_amd import _ffi, dbd, gpu, problem, serving
from helpers import make_problem

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(600)]

HASH_F64 = 0x9BEB634425FAF2C2
HASH_F32 = 0x32256818F6F35BC9


def _open(gm, air, c3c, slice_us):
    meth = problem.method(2)
    gm.set_mechs([c3c, air], meth, problem.control())
    gm.set_mech_control(1, dbd.default_control())
    gm.serve_open(hbm_budget=4 << 30, max_requests=64, steps=20, substeps=4, dt=1e-3, slice_us=slice_us)


def _drain(gm, tickets):
    left = set(tickets)
    while left:
        for t in gm.serve_step():
            left.discard(t)


def test_served_dbd_reactor_reproduces_the_scene_hash(gm):
    print("[case] served DBD reactor, 64 zones x 500 steps, alone then in a mix", flush=True)
    air = dbd.AirMechanism(caps=_ffi.DEFAULT_CAPS)
    sp, c3c, _ = make_problem("c3c", 64)
    results = []
    # alone; beside transport requests; beside them with 300-us time slices
    # (cells stopped mid-step and resumed in later launches)
    for mix, slice_us in ((False, 0), (True, 0), (True, 300)):
        sc = dbd.Scene(air)
        gm.arena_reset()
        _open(gm, air, c3c, slice_us)
        try:
            t, res = gm.serve_submit_dbd(1, sc.z, steps=500, residence=sc.residence, mode=dbd.PULSED,
                                         feed=sc.feed())
            tickets = [t]
            if mix:
                others = []
                for k, cells in enumerate((300, 4096, 1000)):
                    _, _, z = make_problem("c3c", cells, first_cell=100 * k)
                    tk, r = gm.serve_submit_zones(0, z, 1, cells, serving.courant_profile(1, cells),
                                                  inflow=z.y[:, 0].copy())
                    others.append((tk, r))
                    tickets.append(tk)
            _drain(gm, tickets)
            carried = gm.serve_stats().carried
        finally:
            gm.serve_close()
        y = np.ascontiguousarray(res[:air.n, :64])
        results.append(res.copy())
        h = dbd.fnv1a64(y.tobytes())
        print(f"[serve] mix={mix} slice_us={slice_us}: state hash {h:016x}, steps {int(res[air.n + 1].sum())}, "
              f"failed {int(res[air.n + 4].sum())}, cells carried {carried}", flush=True)
        assert h == HASH_F64
        assert res[air.n + 4].sum() == 0
    assert all(r.tobytes() == results[0].tobytes() for r in results[1:])


def _varied(air, mode, residence):
    """A 64-zone scene whose zones run at different discharge powers and
    frequencies (two of them off), so each zone keeps its own pulse clock."""
    rng = np.random.default_rng(11)
    power = rng.uniform(0.02, 0.3, 64)
    power[[5, 40]] = 0.0
    freq = rng.uniform(5e3, 1.5e4, 64)
    sc = dbd.Scene(air, residence=residence, mode=mode)
    sc.z.power_W_cm3[:64] = power
    sc.z.freq_Hz[:64] = freq
    return sc


@pytest.mark.parametrize("mode,residence", [(dbd.PULSED, 0.0), (dbd.PULSED, 0.5), (dbd.AVERAGED, 2.0)],
                         ids=["pulsed-independent", "pulsed-chain-0.5s", "averaged-chain-2s"])
def test_served_dbd_matches_the_oracle_scene(gm, mode, residence):
    """40 master steps served against the same steps driven on the host with
    the oracle's advance and plug flow: y, h, step and reject counts, pulses
    and failure flags equal bit for bit."""
    print(f"[case] served DBD vs oracle scene, mode {mode}, residence {residence}", flush=True)
    import oracle_ffi
    air = dbd.AirMechanism(caps=_ffi.DEFAULT_CAPS)
    _, c3c, _ = make_problem("c3c", 64)
    meth, ctl = problem.method(2), dbd.default_control()
    host = _varied(air, mode, residence)

    def advance(z, dt):
        oracle_ffi.oracle_advance(air, meth, ctl, z, dt, threads=8)

    def flow(z, a, feed):
        oracle_ffi.oracle().oracle_plugflow_f64(z.y.ctypes.data, z.padded, z.n, z.count, a, feed.ctypes.data)

    for _ in range(40):
        host.step(advance, flow)
    z = host.z
    want = np.vstack([z.y[:, :64], z.h[None, :64], z.steps[None, :64], z.rejects[None, :64], z.pulses[None, :64],
                      z.failed[None, :64]])
    served = _varied(air, mode, residence)
    gm.arena_reset()
    _open(gm, air, c3c, 0)
    try:
        t, res = gm.serve_submit_dbd(1, served.z, steps=40, residence=residence, mode=mode,
                                     feed=served.feed() if residence > 0 else None)
        _drain(gm, [t])
    finally:
        gm.serve_close()
    print(f"[serve] steps {int(res[air.n + 1].sum())} (oracle {int(want[air.n + 1].sum())}), "
          f"pulses {int(res[air.n + 3].sum())}", flush=True)
    assert res.tobytes() == np.ascontiguousarray(want).tobytes()
    if mode == dbd.PULSED:
        assert res[air.n + 3].sum() > 0 and res[air.n + 3][[5, 40]].sum() == 0


def test_served_dbd_rejects_bad_requests(gm):
    print("[case] served DBD request validation", flush=True)
    air = dbd.AirMechanism(caps=_ffi.DEFAULT_CAPS)
    _, c3c, _ = make_problem("c3c", 64)
    gm.arena_reset()
    _open(gm, air, c3c, 0)
    try:
        sc = dbd.Scene(air)
        feed = sc.feed()
        with pytest.raises(gpu.GpuError):      # plug flow without a feed
            gm.serve_submit_dbd(1, sc.z, steps=2, residence=2.0)
        with pytest.raises(gpu.GpuError):      # a = dt * zones / residence = 6.4 > 1
            gm.serve_submit_dbd(1, sc.z, steps=2, residence=0.01, feed=feed)
        with pytest.raises(gpu.GpuError):      # unknown source mode
            gm.serve_submit_dbd(1, sc.z, steps=2, feed=feed, mode=2)
        src = dbd.source_defaults()
        src.boltzmann = 1
        with pytest.raises(gpu.GpuError):      # the Boltzmann-solver source is not served
            gm.serve_submit_dbd(1, sc.z, steps=2, feed=feed, source=src)
        sp = _ffi.SimkAirSpecies.from_buffer_copy(air.sp)
        sp.NO = 30
        with pytest.raises(gpu.GpuError):      # a species index outside the mechanism
            gm.serve_submit_dbd(1, sc.z, steps=2, feed=feed, species=sp)
        assert gm.serve_stats().submitted == 0
    finally:
        gm.serve_close()


def test_served_dbd_reactor_f32_lanes_reproduces_the_f32_hash(gm):
    """The same reactor as a request in f32 delta lanes (each step run whole
    by the air mechanism's f32 delta kernel beside the FP64 launch): the
    reference's f32 hash, alone and beside time-sliced FP64 transport
    requests."""
    print("[case] served DBD reactor in f32 delta lanes, 64 zones x 500 steps", flush=True)
    air = dbd.AirMechanism(caps=_ffi.DEFAULT_CAPS)
    sp, c3c, _ = make_problem("c3c", 64)
    results = []
    for mix, slice_us in ((False, 0), (True, 300)):
        sc = dbd.Scene(air)
        gm.arena_reset()
        _open(gm, air, c3c, slice_us)
        try:
            t, res = gm.serve_submit_dbd(1, sc.z, steps=500, residence=sc.residence, mode=dbd.PULSED,
                                         feed=sc.feed(), precision=1)
            tickets = [t]
            if mix:
                for k, cells in enumerate((300, 4096, 1000)):
                    _, _, z = make_problem("c3c", cells, first_cell=100 * k)
                    tk, _ = gm.serve_submit_zones(0, z, 1, cells, serving.courant_profile(1, cells),
                                                  inflow=z.y[:, 0].copy())
                    tickets.append(tk)
            _drain(gm, tickets)
            carried = gm.serve_stats().carried
        finally:
            gm.serve_close()
        h = dbd.fnv1a64(np.ascontiguousarray(res[:air.n, :64]).tobytes())
        print(f"[serve f32] mix={mix} slice_us={slice_us}: state hash {h:016x}, cells carried {carried}", flush=True)
        assert h == HASH_F32
        assert res[air.n + 4].sum() == 0
        results.append(res.copy())
    assert results[0].tobytes() == results[1].tobytes()


def test_served_f32_transport_request_matches_the_oracle(gm):
    """A transport request in f32 delta lanes equals the oracle's operator
    split with its f32 delta advance: 20 x (upwind advection, 4 substeps,
    toward the inflow; then the chemistry of 1 ms), on all six solver rows."""
    print("[case] served transport request in f32 delta lanes vs the oracle", flush=True)
    import oracle_ffi
    sp, c3c, z = make_problem("c3c", 2048)
    air = dbd.AirMechanism(caps=_ffi.DEFAULT_CAPS)
    courant = serving.courant_profile(1, 2048)
    inflow = z.y[:, 0].copy()
    gm.arena_reset()
    _open(gm, air, c3c, 0)
    try:
        t, res = gm.serve_submit_zones(0, z.copy(), 1, 2048, courant, inflow=inflow, precision=1)
        _drain(gm, [t])
    finally:
        gm.serve_close()
    w = z.copy()
    lib = oracle_ffi.oracle()
    for _ in range(20):
        sub = np.ascontiguousarray(w.y[:, :2048])
        lib.oracle_upwind_f64(sub.ctypes.data, courant.ctypes.data, 2048, c3c.n, 2048, 4, inflow.ctypes.data)
        w.y[:, :2048] = sub
        oracle_ffi.oracle_advance_delta32(c3c, problem.method(2), problem.control(), w, 1e-3,
                                          lanes=np.arange(2048), threads=8)
    want = np.vstack([w.y[:, :2048], w.h[None, :2048], w.steps[None, :2048], w.rejects[None, :2048],
                      w.pulses[None, :2048], w.failed[None, :2048]])
    assert res.tobytes() == np.ascontiguousarray(want).tobytes()
