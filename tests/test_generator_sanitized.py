"""The kernel generator's host code under AddressSanitizer and
UndefinedBehaviorSanitizer (tests/sanitize/gen_harness.cpp, built here with
g++): every kernel form for the workload mechanisms and seeded random ones,
including an odd species count and the 64-species capacity. Any sanitizer
report fails the run."""
import os
import shutil
import subprocess

import pytest

from simstack_amd import problem
from helpers import make_problem

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "simstack_amd", "csrc")

MECHS = [("c1", {}), ("c2", {}), ("c3a", {}), ("c3b", {}), ("c3c", {}),
         ("c1", dict(seed=105, n_species=64, n_reactions=230, density_lo=0.08, density_hi=0.15)),
         ("c1", dict(seed=103, n_species=33, n_reactions=90, density_lo=0.1, density_hi=0.25)),
         ("c1", dict(seed=101, n_species=9, n_reactions=20, density_lo=0.2, density_hi=0.5))]


@pytest.mark.timeout(600)
def test_generator_under_sanitizers(tmp_path):
    if not shutil.which("g++"):
        pytest.skip("g++ is not available")
    exe = str(tmp_path / "gen_harness")
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
                    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
                    os.path.join(ROOT, "tests", "sanitize", "gen_harness.cpp"),
                    os.path.join(CSRC, "codegen.cpp"), os.path.join(CSRC, "mech.cpp"), "-o", exe], check=True)
    args = [exe]
    meth = tmp_path / "method.bin"
    meth.write_bytes(bytes(problem.method(2)))
    args.append(str(meth))
    for i, (name, over) in enumerate(MECHS):
        _, mech, _ = make_problem(name, 4, **over)
        (tmp_path / f"m{i}.bin").write_bytes(mech.bytes.tobytes())
        (tmp_path / f"c{i}.bin").write_bytes(bytes(mech.caps))
        args += [str(tmp_path / f"m{i}.bin"), str(tmp_path / f"c{i}.bin")]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    env.pop("LD_PRELOAD", None)
    r = subprocess.run(args, env=env, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "generated" in r.stdout and "runtime error" not in r.stderr, r.stderr[-4000:]
