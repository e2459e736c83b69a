"""The reference's own System and Sampler driving the engine: the reference
built from its sources with the binding a maintainer adds
(oracle/ref/binding: System::initializeEngine creates the engine from the
reference's configuration, System::runEngineCycles replaces
runMetropolisCycles with one vmc_run per iteration, Sampler::setTotals
replaces the MPI reduction of computeTotals, ref src/system.cpp:147-164,
src/sampler.cpp:111-187), everything else -- parser, initial weights,
optimisers, averages, terminal output -- the reference's code.

Run on two of the reference's shipped examples, compared with its logs as
the engine's own command line is (tests/test_gpu_cli.py): the last 20
iterations' mean energy within 3 x (the sum of both runs' spreads) of the
reference log's, the acceptance within 0.01.  The binary exists where the
reference sources were present at build time (`make -C oracle binding`).
"""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

from tests.test_gpu_cli import EX, EXPECTED, _compare
from vmachine_amd import host as H

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BOUND = os.path.join(ROOT, "oracle", "_ref", "binding", "vmachine_bound")


@pytest.mark.parametrize("name", ["small_fermionic_quantumdot_vmc",
                                  "large_fermionic_quantumdot_vmc"])
def test_reference_driving_the_engine(name, tmp_path):
    if not os.path.exists(BOUND):
        pytest.skip("the bound reference is built only where the reference sources are")
    shutil.copy(os.path.join(EX, name + ".in"), tmp_path / "input.in")
    mpiexec = shutil.which("mpiexec") or "/opt/conda/bin/mpiexec"
    cmd = ([mpiexec, "-n", "1"] if os.path.exists(mpiexec) else []) + [BOUND, "input.in"]
    print(f"\n[bound reference] {name}: {' '.join(cmd)}", flush=True)
    sys.stdout.flush()
    r = subprocess.run(cmd, cwd=tmp_path, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       text=True, timeout=240)
    assert r.returncode == 0, (r.stdout[-1500:], r.stderr[-1500:])
    rows, final = H.parse_output(r.stdout)
    _compare(name, np.array(rows), final)
