"""The hybrid split controller (simk_hybrid_init / simk_hybrid_record) on
synthetic timings: device time o + g*n (+-2 % noise), CPU time c*(padded - n).
The split that ends both together is n* = (c*padded - o)/(g + c); the
controller must settle on it (to its 8-lane grid), follow it when the CPU
slows down, switch the device off when its fixed cost exceeds the whole CPU
advance, and keep every split a multiple of 8 inside the block."""
import ctypes as C

import numpy as np
import pytest

from simstack_amd import _ffi


def _run(N, o, g, c, steps, seed=1, change=None):
    h = _ffi.SimkHybrid()
    lib = _ffi.host()
    lib.simk_hybrid_init(C.byref(h), N)
    rng = np.random.default_rng(seed)
    splits, bests = [], []
    for s in range(steps):
        if change and s == change[0]:
            c = change[1]
        n = h.split
        assert n % 8 == 0 and 0 <= n <= N
        gpu = (o + g * n) * (1 + 0.02 * rng.standard_normal()) if n else 0.0
        cpu = c * (N - n) * (1 + 0.02 * rng.standard_normal())
        lib.simk_hybrid_record(C.byref(h), n, gpu, cpu)
        splits.append(n)
        bests.append(h.best)
    return np.array(splits), np.array(bests), h


def _n_star(N, o, g, c):
    return (N * c - o) / (g + c)


def test_settles_on_the_balanced_split():
    N, o, g, c = 4096, 0.4, 2e-4, 5e-3
    splits, bests, h = _run(N, o, g, c, 60)
    want = _n_star(N, o, g, c)
    assert np.all(np.abs(bests[8:] - want) <= 24), (bests, want)
    # apart from the dither steps (one in 16) the next split is the optimum
    assert np.mean(splits[9:] == bests[8:-1]) >= 0.9
    assert abs(h.gpu_fixed_ms - o) < 0.1 and abs(h.gpu_lane_ms / g - 1) < 0.1
    assert abs(h.cpu_lane_ms / c - 1) < 0.05


def test_follows_a_slower_cpu():
    N, o, g = 4096, 0.4, 2e-4
    splits, bests, _ = _run(N, o, g, 5e-3, 120, change=(40, 1.5e-2))
    want = _n_star(N, o, g, 1.5e-2)
    assert np.all(np.abs(bests[60:] - want) <= 24), (bests[60:], want)
    assert bests[39] < want - 100


def test_device_off_when_its_fixed_cost_exceeds_the_cpu_advance():
    N, o, g, c = 1024, 50.0, 1e-4, 1e-3     # the whole CPU advance takes ~1 ms
    splits, bests, _ = _run(N, o, g, c, 64)
    assert np.all(bests[4:] == 0)
    probes = np.flatnonzero(splits[4:])
    assert len(probes) <= 4 and np.all(splits[4:][probes] == 256)


def test_all_device_when_the_cpu_is_much_slower():
    N, o, g, c = 2048, 0.05, 1e-5, 1.0
    splits, bests, _ = _run(N, o, g, c, 48)
    assert np.all(bests[8:] == N)
    assert np.mean(splits[9:] == N) >= 0.9


@pytest.mark.parametrize("N", [8, 16, 24])
def test_tiny_blocks_stay_on_the_grid(N):
    for o, c in ((0.01, 1.0), (10.0, 1e-3), (0.5, 0.05)):
        splits, _, _ = _run(N, o, 1e-3, c, 40)
        assert np.all(splits % 8 == 0) and splits.max() <= N
