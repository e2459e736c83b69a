"""`vmachine input.in` on the GPU against the reference's shipped example runs
(examples/*/vmachine.log, stored in tests/golden/examples/expected.json).

The engine draws from Philox streams over many walkers, the reference from one
Mersenne-Twister chain, so the runs agree in distribution, not digit by digit:
 * the iteration table has the reference's shape (numIterations - 1 rows of 11
   columns, then the final block);
 * the non-interacting example is exact (E = 2, zero variance);
 * the mean energy of the last 20 iterations lies within 3 x (the sum of both
   runs' iteration-to-iteration spreads) of the reference's, and the acceptance
   ratio within 0.01: the importance-sampling acceptance depends on the
   reference's Green's-function ratio and Pade-Jastrow ratio semantics, so a
   change in either moves it well beyond that.
"""
import json
import os
import sys

import numpy as np
import pytest

from vmachine_amd import host as H

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
EX = os.path.join(HERE, "golden", "examples")
EXPECTED = json.load(open(os.path.join(EX, "expected.json")))


def _run(name, *extra, cwd=None):
    print(f"\n[vmachine] {name} {' '.join(map(str, extra))}", flush=True)
    sys.stdout.flush()
    r = H.run(list(extra) + [os.path.join(EX, name + ".in")], timeout=240, cwd=cwd)
    assert r.returncode == 0, r.stderr[-2000:]
    rows, final = H.parse_output(r.stdout)
    return np.array(rows), final, r.stdout


def _compare(name, rows, final):
    ref = EXPECTED[name]
    R = np.array(ref["rows"])
    assert rows.shape == R.shape, (rows.shape, R.shape)
    np.testing.assert_array_equal(rows[:, 0], np.arange(1, len(R) + 1))
    e_ours, e_ref = rows[-20:, 1], R[-20:, 1]
    tol = 3 * (e_ours.std() + e_ref.std())
    print(f"  last-20 mean {e_ours.mean():.6f} (ref {e_ref.mean():.6f}, tol {tol:.6f}); "
          f"acceptance {rows[-20:, 9].mean():.5f} (ref {R[-20:, 9].mean():.5f}); "
          f"final {final['energy']:.8f} (ref {ref['final']['energy']:.8f})")
    assert abs(e_ours.mean() - e_ref.mean()) < tol
    assert abs(rows[-20:, 9].mean() - R[-20:, 9].mean()) < 0.01
    for k in ("energy", "kinetic", "external", "interaction", "acceptance"):
        assert np.isfinite(final[k])


def test_no_interaction_example_is_exact():
    rows, final, _ = _run("no_interaction", "--seed", 11)
    assert rows.shape == (99, 11)
    np.testing.assert_allclose(rows[:, 1], 2.0, rtol=0, atol=1e-12)
    assert final["energy"] == pytest.approx(2.0, abs=1e-10)
    assert final["energy_std"] == pytest.approx(0.0, abs=1e-8)
    assert final["interaction"] == 0.0


@pytest.mark.parametrize("name", ["small_fermionic_quantumdot_vmc", "electron_density"])
def test_two_electron_dot_examples(name):
    rows, final, _ = _run(name, "--seed", 12)
    _compare(name, rows, final)


def test_twelve_electron_dot_example():
    rows, final, _ = _run("large_fermionic_quantumdot_vmc", "--seed", 13)
    _compare("large_fermionic_quantumdot_vmc", rows, final)


def test_rbm_example():
    # RBM-Gaussian x RBM-product, N=2, H=2, Xavier weights, 500 ADAM iterations
    rows, final, _ = _run("small_fermionic_quantumdot_rbm", "--seed", 17)
    _compare("small_fermionic_quantumdot_rbm", rows, final)


def test_resampling_example_blocks_the_last_iteration():
    rows, final, out = _run("resampling", "--seed", 14)
    _compare("resampling", rows, final)
    # the final standard error comes from blocking the recorded local energies:
    # correlated samples give an error above the naive sqrt(var/M) of the rows
    ref = EXPECTED["resampling"]["final"]
    assert 0.2 * ref["energy_std"] < final["energy_std"] < 5 * ref["energy_std"]


def test_fp32_walk_matches_fp64_statistically():
    r64, f64, _ = _run("small_fermionic_quantumdot_vmc", "--seed", 15)
    r32, f32, _ = _run("small_fermionic_quantumdot_vmc", "--seed", 15, "--precision", 32)
    tol = 3 * (r64[-20:, 1].std() + r32[-20:, 1].std())
    assert abs(r64[-20:, 1].mean() - r32[-20:, 1].mean()) < tol


def test_output_files(tmp_path):
    cfg = tmp_path / "input.in"
    text = open(os.path.join(EX, "small_fermionic_quantumdot_vmc.in")).read()
    text = text.replace("numIterations: 100", "numIterations: 12")
    cfg.write_text(text + f"printEnergyToFile: true\nprintParametersToFile: true\n"
                          f"path: {tmp_path}/\n")
    print("\n[vmachine] output files", flush=True)
    r = H.run(["--seed", 16, str(cfg)], timeout=120)
    assert r.returncode == 0, r.stderr
    E = np.loadtxt(tmp_path / "energy.dat")
    assert E.shape == (12,)
    rows, final = H.parse_output(r.stdout)
    np.testing.assert_allclose(E[:-1], np.array(rows)[:, 1], rtol=1e-5)
    assert (tmp_path / "weights_0.dat").exists()
    W = np.loadtxt(tmp_path / "weights_0.dat")
    assert W.shape == (2,)
