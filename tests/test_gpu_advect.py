"""Upwind film advection on the device vs the oracle's restatement of
nft_transport's row update (src/sim/nft.cpp:1055-1081): bit for bit, f32 with
the outlet sum and f64 with an inflow, at the 4,096-node channel size of the
serving workload and at ragged lengths."""
import ctypes as C

import numpy as np
import pytest

import oracle_ffi

pytestmark = pytest.mark.gpu


def _dev(gm, arr):
    p = gm.alloc(arr.nbytes)
    gm.to_device(p, arr)
    return p


@pytest.mark.parametrize("nodes,channels,fields,substeps", [(4096, 4, 48, 9), (1, 1, 3, 5), (255, 3, 2, 17),
                                                            (257, 2, 5, 3), (1000, 1, 9, 40), (8192, 1, 2, 4),
                                                            (2048, 2, 3, 4), (3001, 3, 5, 7), (5000, 1, 3, 6),
                                                            (6000, 2, 1, 3), (8192, 2, 10, 4), (5000, 3, 17, 6),
                                                            (2048, 5, 8, 3), (700, 3, 23, 5), (3001, 3, 9, 7),
                                                            (999, 2, 8, 5)])
def test_upwind_f64_matches_oracle(gm, nodes, channels, fields, substeps):
    """Row lengths over every wave shape: 16 or 32 slots per wave, 1, 2 or 4
    waves per row, ragged last parts, empty last parts (5,000 nodes in 4
    waves) and row counts that leave a workgroup's last row inactive; with 8
    or more fields, wave groups of 8 fields of a channel, including groups
    whose last fields are past the end: with even row lengths the groups use
    16-byte lane accesses (two nodes per lane and slot), with odd ones 8-byte
    accesses."""
    rng = np.random.default_rng(nodes + channels)
    padded = (channels * nodes + 63) // 64 * 64
    x = rng.random((fields, padded)) * 10.0 ** rng.uniform(-12, -2, (fields, 1))
    c = rng.uniform(0.05, 0.9, channels * nodes)
    inflow = rng.random(fields) * 1e-3
    gm.arena_reset()
    dx, dc, di = _dev(gm, x), _dev(gm, c), _dev(gm, inflow)
    gm._check(gm.lib.simk_upwind_f64(dx, padded, fields, channels, nodes, dc, substeps, di), "upwind")
    got = np.empty_like(x)
    gm.to_host(got, dx)
    want = x.copy()
    for ch in range(channels):
        sub = np.ascontiguousarray(want[:, ch * nodes:(ch + 1) * nodes])
        cs = np.ascontiguousarray(c[ch * nodes:(ch + 1) * nodes])
        oracle_ffi.oracle().oracle_upwind_f64(sub.ctypes.data, cs.ctypes.data, nodes, fields, nodes, substeps,
                                              inflow.ctypes.data)
        want[:, ch * nodes:(ch + 1) * nodes] = sub
    assert got.tobytes() == want.tobytes()


@pytest.mark.parametrize("nodes,substeps", [(4096, 12), (777, 30), (2500, 9), (8000, 5)])
def test_upwind_f32_matches_oracle_with_outlet(gm, nodes, substeps):
    rng = np.random.default_rng(nodes)
    fields = 9
    x = (rng.random((fields, nodes)) - 0.3).astype(np.float32)
    c = rng.uniform(0.05, 0.9, nodes).astype(np.float32)
    out = rng.random(fields).astype(np.float32)
    gm.arena_reset()
    dx, dc, do = _dev(gm, x), _dev(gm, c), _dev(gm, out)
    gm._check(gm.lib.simk_upwind_f32(dx, nodes, fields, 1, nodes, dc, substeps, do), "upwind32")
    got, got_out = np.empty_like(x), np.empty_like(out)
    gm.to_host(got, dx)
    gm.to_host(got_out, do)
    want, want_out = x.copy(), out.copy()
    oracle_ffi.oracle().oracle_upwind(want.ctypes.data, c.ctypes.data, nodes, fields, nodes, substeps,
                                      want_out.ctypes.data)
    assert got.tobytes() == want.tobytes()
    assert got_out.tobytes() == want_out.tobytes()


def test_upwind_conserves_mass_with_closed_outlet_balance(gm):
    """Flux form: sum over nodes changes only by inflow in minus outflow out
    (nft.cpp:1023-1027); checked in f64 against the per-substep outlet values."""
    nodes, fields, substeps = 4096, 3, 25
    rng = np.random.default_rng(0)
    x = rng.random((fields, nodes))
    c = np.full(nodes, 0.5)
    total0 = x.sum(axis=1)
    gm.arena_reset()
    dx, dc = _dev(gm, x), _dev(gm, c)
    gm._check(gm.lib.simk_upwind_f64(dx, nodes, fields, 1, nodes, dc, substeps, None), "upwind")
    got = np.empty_like(x)
    gm.to_host(got, dx)
    # outflow per substep = c * x_last(old); recompute with the oracle to get the outlet series
    series = np.zeros(fields)
    w = x.copy()
    for _ in range(substeps):
        series += 0.5 * w[:, -1]
        oracle_ffi.oracle().oracle_upwind_f64(w.ctypes.data, c.ctypes.data, nodes, fields, nodes, 1, None)
    assert np.allclose(got.sum(axis=1), total0 - series, rtol=1e-12)


@pytest.mark.parametrize("zones", [1, 300, 4096, 5000])
def test_plugflow_matches_oracle(gm, zones):
    """simk_plugflow_f64 (dbd_flow_exchange on the device): bit for bit, with
    the product rounded separately from the sum as the reference computes it."""
    rng = np.random.default_rng(zones)
    species = 17
    padded = (zones + 63) // 64 * 64
    y = rng.random((species, padded)) * 10.0 ** rng.uniform(8, 19, (species, 1))
    feed = rng.random(species) * 1e19
    a = 0.3
    gm.arena_reset()
    dy, df = _dev(gm, y), _dev(gm, feed)
    gm._check(gm.lib.simk_plugflow_f64(dy, padded, species, zones, a, df), "plugflow")
    got = np.empty_like(y)
    gm.to_host(got, dy)
    want = y.copy()
    oracle_ffi.oracle().oracle_plugflow_f64(want.ctypes.data, padded, species, zones, a, feed.ctypes.data)
    assert got.tobytes() == want.tobytes()
    with pytest.raises(Exception):
        gm._check(gm.lib.simk_plugflow_f64(dy, padded, species, zones, 1.5, df), "plugflow")
