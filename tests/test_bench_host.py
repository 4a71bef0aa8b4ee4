"""Host-side parts of bench.py: the GPU-count refusal, the accuracy criterion
and the CPU baseline leg (no GPU needed)."""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle_ffi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
from simstack_amd import problem  # noqa: E402


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


class _Block:
    def __init__(self, y):
        self.y = y
        self.steps = np.full(y.shape[1], 7.0)
        self.rejects = np.zeros(y.shape[1])


def test_accuracy_criterion_counts_only_entries_above_atol_over_rtol():
    ctl = problem.control()                       # atol / rtol = 1e-6
    rng = np.random.default_rng(3)
    ref = rng.uniform(1e-5, 1e-2, (6, 2000))
    ref[5] = 1e-8                                 # below atol / rtol: never counted
    got = ref.copy()
    got[5] *= 1.5
    got[2, 17] *= 1 + 4e-5                        # one cell off by 4e-5
    got[1, 300:302] *= 1 + 2e-7                   # two cells off by 2e-7
    a = bench.accuracy(_Block(got), _Block(ref), ctl, 2000)
    assert a["max_rel_err"] == pytest.approx(4e-5 / (1 + 4e-5), rel=1e-9)
    assert a["cells_above_1e-6"] == pytest.approx(1 / 2000)
    # p99.9 over 2000 cells interpolates between the 2nd and 3rd worst cells
    assert 2e-7 / (1 + 2e-7) * 0.99 < a["cell_p99_9"] < 4e-5
    assert a["meets_criterion"] is (a["cell_p99_9"] <= bench.CRITERION_P999)
    assert a["cells_same_step_counts"] == 1.0
    g = a["max_rel_err_by_species_group"]
    assert g["above atol/rtol"] == pytest.approx(a["max_rel_err"])
    assert g["1e-9..1e-6"] == pytest.approx(0.5 / 1.5)


def test_accuracy_criterion_fails_when_many_cells_are_off():
    ctl = problem.control()
    ref = np.full((4, 1000), 1e-3)
    got = ref.copy()
    got[0, :10] *= 1 + 1e-5                       # 1 % of the cells
    a = bench.accuracy(_Block(got), _Block(ref), ctl, 1000)
    assert not a["meets_criterion"] and a["cell_p99_9"] > 1e-6


def test_cpu_baseline_reference_and_port_do_the_same_work():
    sp = problem.spec("c2")
    mech = problem.Mechanism.synthetic(sp)
    meth, ctl = problem.method(2), problem.control()
    port = bench.cpu_baseline(mech, meth, ctl, sp, 64, 0.05, repeats=2, threads=2, port=True)
    assert port["kind"] == "port" and port["cores"] == 2 and port["value"] > 0
    assert "2 x 0.05-s advances" in port["sample"]
    if not oracle_ffi.have_ref():
        pytest.skip("oracle/_ref not built")
    ref = bench.cpu_baseline(mech, meth, ctl, sp, 64, 0.05, repeats=2, threads=2)
    assert ref["kind"] == "reference"
    steps = lambda d: int(d["sample"].split("(")[1].split(" accepted")[0])  # noqa: E731
    assert steps(ref) == steps(port) > 0
