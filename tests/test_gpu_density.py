"""Electron densities counted by the walk (vmc_set_density / vmc_get_density)
against the oracle's restatement of the reference's three binning rules
(ref src/sampler.cpp:639-692), applied to the walkers' positions after every
sampled step.  The counts are integers and must agree exactly: the walk is
advanced one step per launch so that every intermediate configuration can be
read back, and a second engine on the same random stream runs the whole
segment (burn-in + sampled steps) in one launch and must give the same counts.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import OMEGA_FOR_N, engine_config, random_parameters

pytestmark = pytest.mark.gpu

KINDS = [O.SLATER, O.GAUSSIAN, O.PADE]
ALL = 1 | 2 | 4


def _engine(N, W, precision, dt=0.05):
    import vmachine_amd as V
    if V.device_count() < 1:
        pytest.fail("no GPU visible: GPU tests must run on the MI355X box")
    cfg = engine_config(N, OMEGA_FOR_N[N], KINDS, O.IMPORTANCE_EXACT, dt, quirks=0,
                        precision=precision, walkers=W, drift_limit=1.0)
    e = V.Engine(cfg)
    e.set_parameters(random_parameters(np.random.default_rng(N), KINDS))
    return e


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("N,W,bins,R", [(2, 70, 60, 3.0), (6, 37, 41, 4.0), (12, 29, 64, 3.5)])
def test_density_counts_match_the_oracle(N, W, bins, R, precision):
    seed, equil, steps = 300 + N, 3, 12
    print(f"density N={N} W={W} bins={bins} precision={precision}", flush=True)
    radial = np.zeros(bins, dtype=np.int64)
    spatial = np.zeros((bins, bins), dtype=np.int64)
    two = np.zeros((bins, bins), dtype=np.int64)
    with _engine(N, W, precision) as e:
        e.init_walkers(seed)
        e.run(equil)                      # burn-in, densities off
        e.set_density(ALL, bins, R)
        for t in range(steps):
            e.run(1)
            for x in e.positions():
                O.density_add(x, N, bins, R, radial, spatial, two)
        r, s, tb = e.density()
    np.testing.assert_array_equal(r, radial)
    np.testing.assert_array_equal(s, spatial)
    np.testing.assert_array_equal(tb, two)
    assert r.sum() > 0.5 * N * W * steps
    # one launch over burn-in + sampled steps on the same stream: the burn-in
    # steps are not counted, the sampled ones are the same configurations
    with _engine(N, W, precision) as e:
        e.init_walkers(seed)
        e.set_density(ALL, bins, R)
        e.run(steps, equil)
        r2, s2, t2 = e.density()
    if precision == 64:
        np.testing.assert_array_equal(r2, radial)
        np.testing.assert_array_equal(s2, spatial)
        np.testing.assert_array_equal(t2, two)
    else:
        # FP32: one launch keeps the Sherman-Morrison inverse where the
        # step-per-launch run rebuilds it, so chains agree to rounding only
        assert r2.sum() == pytest.approx(radial.sum(), rel=0.02)
        assert np.abs(r2 - radial).sum() <= 0.05 * radial.sum()


def test_density_off_and_cleared():
    import vmachine_amd as V
    with _engine(6, 33, 64) as e:
        e.init_walkers(5)
        with pytest.raises(V.VmcError, match="off"):
            e.density()
        e.set_density(ALL, 20, 3.0)
        e.run(4)
        r, _, _ = e.density()
        assert r.sum() > 0
        e.set_density(ALL, 20, 3.0)      # clears
        r, s, t = e.density()
        assert r.sum() == 0 and s.sum() == 0 and t.sum() == 0
        with pytest.raises(V.VmcError, match="numberOfBins"):
            e.set_density(ALL, 0, 3.0)
        e.set_density(0)
        with pytest.raises(V.VmcError, match="off"):
            e.density()


def test_instant_values_are_the_sums():
    """vmc_run_instant: the recorded E, T, V, I of every walker and step add up
    to the run's sums (FP64 walk, all walkers recorded)."""
    with _engine(6, 40, 64) as e:
        e.init_walkers(9)
        e.run(50)
        sums, _, _, out = e.run_instant(64, 5, 0, 1, record_walkers=40)
    assert out.shape == (4, 40, 64)
    for q, key in enumerate(["energy", "kinetic", "external", "interaction"]):
        assert out[q].astype(np.float64).sum() == pytest.approx(sums[key], rel=1e-6)
    np.testing.assert_allclose(out[0], out[1] + out[2] + out[3], rtol=1e-5, atol=1e-5)


def test_blocked_means_are_the_recorded_series():
    """vmc_run_blocked: the device block means equal the means of the per-step
    series vmc_run_recorded returns for the same chain (same positions, random
    state and walk), and the run's sums are the recorded series' sums."""
    with _engine(6, 40, 64) as e:
        e.init_walkers(9)
        e.run(50)
        pos = e.positions()
        seed, ctr = e.random_state()
        sums_r, _, _, en = e.run(96, 4, record_walkers=40)
        e.set_positions(pos)
        e.set_random_state(seed, ctr)
        sums_b, _, _, means = e.run_blocked(96, 8, 4, record_walkers=33)
    assert means.shape == (33, 12)
    ref = en[:33].astype(np.float64).reshape(33, 12, 8).mean(axis=2)
    np.testing.assert_allclose(means, ref, rtol=1e-12, atol=0)
    assert sums_b["energy"] == sums_r["energy"]
    import vmachine_amd as V
    with _engine(6, 8, 64) as e:
        e.init_walkers(3)
        with pytest.raises(V.VmcError, match="block"):
            e.run_blocked(10, 3)
