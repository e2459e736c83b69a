"""Parity of the HIP engine against the FP64 oracle (which is itself pinned to
the reference's golden vectors in test_oracle_golden.py).

Tolerances (errors scaled by the size of the terms a quantity is computed
from, see scaled_errors): FP64 walk 1e-9 (max over walkers), FP32 walk 2e-3
(max) and 2e-5 (median).  Configurations are drawn at random, so a few lie
close to a node of the Slater determinant, where FP32 cancellation is largest.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import (CLOSED_SHELLS, OMEGA_FOR_N, engine_config, oracle_config,
                           oracle_evaluate, random_parameters, zero_variance_energy)

pytestmark = pytest.mark.gpu

KINDS_SGP = [O.SLATER, O.GAUSSIAN, O.PADE]


def _engine(cfg):
    import vmachine_amd as V
    if V.device_count() < 1:
        pytest.fail("no GPU visible: GPU tests must run on the MI355X box")
    return V.Engine(cfg)


def _rel_err(a, b):
    scale = np.maximum(np.abs(b), 1.0)
    return np.max(np.abs(a - b) / scale)


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("N", CLOSED_SHELLS)
def test_evaluate_matches_oracle(N, precision):
    omega = OMEGA_FOR_N[N]
    rng = np.random.default_rng(1000 + N)
    W = 48
    ocfg = oracle_config(N, omega, KINDS_SGP, O.IMPORTANCE_COMPAT, 0.01)
    P = random_parameters(rng, KINDS_SGP)
    pos = rng.standard_normal((W, 2 * N)) / np.sqrt(omega)
    moves = rng.integers(0, 2 * N, W)
    newv = pos[np.arange(W), moves] + 0.2 / np.sqrt(omega) * rng.standard_normal(W)
    print(f"evaluate N={N} omega={omega} precision={precision}", flush=True)
    with _engine(engine_config(N, omega, KINDS_SGP, 1, 0.01, precision=precision,
                               walkers=W)) as eng:
        eng.set_parameters(P)
        g0, g1 = eng.evaluate(pos, moves, newv)
    r0, r1 = oracle_evaluate(ocfg, P, pos, moves, newv)
    errs = scaled_errors(g0, r0, g1, r1)
    print(f"N={N} precision={precision} " +
          " ".join(f"{k}: max {v.max():.2e} median {np.median(v):.2e}" for k, v in errs.items()))
    # 72 to 110 particles at independent random positions: the 36..55-row
    # Slater matrices of some walkers are conditioned to 1e8 and more, and the
    # drift (the gradient of ln det) carries that on both sides, so the bounds
    # are looser there; the FP64 chains at N = 72 and 110 (typical
    # configurations) still retrace the oracle to 1e-9
    tol_max, tol_med = (TOL if N <= 56 else TOL_LARGE)[precision]
    for k, v in errs.items():
        assert v.max() < tol_max, f"{k}: max scaled error {v.max():.3e} >= {tol_max}"
        assert np.median(v) < tol_med, f"{k}: median scaled error {np.median(v):.3e} >= {tol_med}"


# scaled-error bounds (max over walkers, median over walkers), see scaled_errors()
TOL = {64: (1e-9, 1e-11), 32: (2e-3, 2e-5)}
TOL_LARGE = {64: (2e-7, 5e-10), 32: (5e-2, 2e-4)}


def scaled_errors(g0, r0, g1, r1):
    """Per-walker errors of the engine (g) against the oracle (r), each scaled by
    the size of the terms it is computed from:
      energy:  |dE_L| / (|T| + |V_ext| + |V_int| + 1/2 sum_k G_k^2 + 1)  (the kinetic
               energy is a difference of the Laplacian and sum G^2)
      dlnpsi/dtheta:  |d g| / (|g| + 1)
      drift:   max_k |d F_k| / (max_k |F_k| + 1)
      ratios:  |d r| / max(|r|, 1) of the forced move's Slater, Pade and Gaussian
               ratios (the acceptance probability is min(1, r))."""
    out = {}
    for tag, g, r in [("before", g0, r0), ("after", g1, r1)]:
        drift = r[:, 9:]
        scale_e = (np.abs(r[:, 1]) + np.abs(r[:, 2]) + np.abs(r[:, 3]) +
                   0.125 * np.sum(drift ** 2, axis=1) + 1.0)
        out[f"E_{tag}"] = np.abs(g[:, 0] - r[:, 0]) / scale_e
        for c, nm in [(4, "Galpha"), (5, "Ggamma")]:
            out[f"{nm}_{tag}"] = np.abs(g[:, c] - r[:, c]) / (np.abs(r[:, c]) + 1.0)
        out[f"drift_{tag}"] = (np.max(np.abs(g[:, 9:] - drift), axis=1) /
                               (np.max(np.abs(drift), axis=1) + 1.0))
    for c, nm in [(6, "ratio_slater"), (7, "ratio_pade"), (8, "ratio_gauss")]:
        out[nm] = np.abs(g1[:, c] - r1[:, c]) / np.maximum(np.abs(r1[:, c]), 1.0)
    return out


@pytest.mark.parametrize("N,sampling,dt,kinds,drift", [
    (2, O.BRUTE_FORCE, 0.1, [O.GAUSSIAN, O.PADE], 0.0),
    (2, O.IMPORTANCE_COMPAT, 0.1, [O.GAUSSIAN, O.PADE], 0.0),
    (6, O.IMPORTANCE_COMPAT, 0.05, KINDS_SGP, 0.0),
    (12, O.IMPORTANCE_EXACT, 0.02, KINDS_SGP, 0.0),
    (12, O.IMPORTANCE_EXACT, 0.05, KINDS_SGP, 1.0),
    (20, O.IMPORTANCE_COMPAT, 0.01, KINDS_SGP, 0.0),
    (30, O.IMPORTANCE_EXACT, 0.01, KINDS_SGP, 1.0),
    (56, O.IMPORTANCE_EXACT, 0.005, KINDS_SGP, 1.0),
    (72, O.IMPORTANCE_EXACT, 0.005, KINDS_SGP, 1.0),
    (110, O.IMPORTANCE_COMPAT, 0.005, KINDS_SGP, 0.0),
])
def test_fp64_trajectories_match_oracle(N, sampling, dt, kinds, drift):
    """Same Philox streams, same positions: the FP64 walk must retrace the
    oracle's chains step for step (accept decisions, positions, sums).  The
    exact-Langevin Slater x Gaussian x Pade-Jastrow cases with Coulomb run the
    walk's instantiation specialised to that configuration."""
    omega = OMEGA_FOR_N[N]
    rng = np.random.default_rng(7 + N)
    W, steps, equil, seed = 64, 200, 20, 99 + N
    P = random_parameters(rng, kinds)
    ocfg = oracle_config(N, omega, kinds, sampling, dt, drift_limit=drift)
    pos0 = np.stack([O.initial_positions(seed, w, 2 * N, omega) for w in range(W)])
    # start from a gently spread configuration so that short chains are in a typical region
    pos0 = pos0 * omega / np.sqrt(omega)
    print(f"trajectory N={N} sampling={sampling}", flush=True)
    with _engine(engine_config(N, omega, kinds, sampling, dt, walkers=W,
                               drift_limit=drift)) as eng:
        eng.set_parameters(P)
        eng.set_positions(pos0)
        eng.set_random_state(seed, 0)
        sums, g, ge, _ = eng.run(steps, equil, 0, 1)
        gpos = eng.positions()
    opos = pos0.copy()
    counters = np.zeros(W, dtype=np.uint64)
    qf = None
    acc, og, oge = O.run_many(ocfg, P, opos, seed, counters, steps, equil, 0, 1, threads=8, qf=qf)
    assert sums["accepted"] == acc[8], (sums["accepted"], acc[8])
    assert abs(sums["energy"] - acc[0]) <= 1e-9 * abs(acc[0])
    assert abs(sums["energy2"] - acc[1]) <= 1e-9 * abs(acc[1])
    assert np.max(np.abs(gpos - opos)) < 1e-9 * max(1.0, np.max(np.abs(opos)))
    assert np.max(np.abs(g - og)) <= 1e-9 * max(1.0, np.max(np.abs(og)))
    assert np.max(np.abs(ge - oge)) <= 1e-9 * max(1.0, np.max(np.abs(oge)))


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("N", CLOSED_SHELLS)
def test_zero_variance_closed_shells(N, precision):
    """Non-interacting closed shells with alpha = 1: E_L is the exact energy at
    every step (ref examples/no_interaction/vmachine.log:180 for N=2)."""
    omega = OMEGA_FOR_N[N]
    kinds = [O.SLATER, O.GAUSSIAN]
    W = 256
    with _engine(engine_config(N, omega, kinds, 1, 0.01, interaction=False, precision=precision,
                               walkers=W, refresh=100 if precision == 32 else 0)) as eng:
        eng.set_parameters(np.array([[1.0], [1.0]]))
        eng.init_walkers(5)
        eng.run(200, 0)
        sums, _, _, en = eng.run(300, 0, record_walkers=W)
    E0 = zero_variance_energy(N, omega)
    mean = sums["energy"] / sums["samples"]
    dev = np.abs(en.astype(np.float64) - E0) / E0
    print(f"N={N} precision={precision}: <E> = {mean:.10f} (E0 = {E0}), max |E_L-E0|/E0 "
          f"{dev.max():.2e}, 99.9% {np.quantile(dev, 0.999):.2e}")
    # the recorded series is FP32, so its own rounding (6e-8) bounds the FP64 check
    if precision == 64:
        assert abs(mean - E0) < 1e-10 * E0, (mean, E0)
        assert dev.max() < 2e-7
    else:
        assert abs(mean - E0) < 1e-4 * E0, (mean, E0)
        assert np.quantile(dev, 0.999) < 1e-3
    assert sums["nonfinite"] == 0


def test_two_electron_dot_energy_is_three():
    """N=2, omega=1 (Taut's exactly solvable dot, E = 3): Gaussian x Pade with
    near-optimal parameters and exact MH Langevin sampling gives E within its
    blocking error of 3.0 (plus the small variational gap of this trial form)."""
    from vmachine_amd import stats
    kinds = [O.GAUSSIAN, O.PADE]
    W = 4096
    with _engine(engine_config(2, 1.0, kinds, O.IMPORTANCE_EXACT, 0.1, quirks=0,
                               walkers=W)) as eng:
        eng.set_parameters(np.array([[0.988], [0.40]]))
        eng.init_walkers(17)
        eng.run(2000, 0)
        sums, _, _, en = eng.run(4096, 0, record_walkers=256)
    E = sums["energy"] / sums["samples"]
    se = np.std([stats.blocking(s)[0] for s in en.astype(np.float64)]) / np.sqrt(256)
    print(f"N=2 E = {E:.6f}, spread-based error {se:.2e}, acceptance "
          f"{sums['accepted'] / sums['samples']:.4f}")
    assert abs(E - 3.0) < 2e-3, E
    assert sums["nonfinite"] == 0


@pytest.mark.parametrize("N", [6, 12, 30, 56])
def test_fp32_walk_agrees_with_fp64_walk(N):
    """The FP32 walk (FP64 inverse rebuild every 100 steps) samples the same
    distribution as the FP64 walk.  Both start from the same walkers and draw
    the same Philox streams, so walker w of one precision is paired with
    walker w of the other: the chains stay close for a while and then
    decorrelate, and the per-walker differences of the mean local energy are
    independent across walkers.  Their mean must be within 4 standard errors
    of zero (the standard error from the spread of the differences)."""
    omega = OMEGA_FOR_N[N]
    W = 8192
    res = {}
    for prec in (64, 32):
        with _engine(engine_config(N, omega, KINDS_SGP, O.IMPORTANCE_EXACT, 0.02, quirks=0,
                                   precision=prec, walkers=W, refresh=100,
                                   drift_limit=1.0)) as eng:
            eng.set_parameters(np.array([[0.0], [1.0], [0.5]]))
            eng.init_walkers(23)
            eng.run(3000, 0)
            sums, _, _, en = eng.run(2048, 0, record_walkers=W)
        means = en.astype(np.float64).mean(axis=1)
        res[prec] = (sums["energy"] / sums["samples"], means)
        print(f"N={N} precision={prec}: E = {res[prec][0]:.6f} "
              f"(E_L range {en.min():.2f} .. {en.max():.2f})")
    d = res[32][1] - res[64][1]
    se = d.std() / np.sqrt(W)
    print(f"  paired difference {d.mean():.3e} +- {se:.2e} "
          f"(independent-chain error {np.hypot(res[32][1].std(), res[64][1].std()) / np.sqrt(W):.2e})")
    assert abs(d.mean()) < 4.0 * se + 1e-7 * abs(res[64][0]), (d.mean(), se)
    assert abs(res[32][0] - res[64][0]) < 4.0 * se + 1e-6 * abs(res[64][0])


@pytest.mark.parametrize("N", [2, 12, 56])
def test_init_walkers_equal_oracle_initial_state(N):
    """vmc_init_walkers draws the reference's RandomNormal initial state
    x ~ N(0, 1/omega) (standard deviation 1/omega, ref
    src/InitialStates/randomnormal.cpp:17) from the same Philox words as the
    oracle's vmco_initial_positions, whose law is pinned on the CPU
    (test_oracle_pins.py).  The Philox words are integers and equal; the
    device's log, sqrt, cos and sin and the host C library's each round to
    within an ulp or two, so the positions agree to a few ulps (most bit for
    bit)."""
    omega = OMEGA_FOR_N[N]
    W, seed = 1024, 4242 + N
    with _engine(engine_config(N, omega, KINDS_SGP, O.IMPORTANCE_COMPAT, 0.01,
                               walkers=W)) as eng:
        eng.init_walkers(seed)
        g = eng.positions()
    o = np.stack([O.initial_positions(seed, w, 2 * N, omega) for w in range(W)])
    scale = np.sqrt(-2.0 * np.log(2.0 ** -24)) / omega   # largest |x| the draw can give
    ulps = np.abs(g - o) / np.spacing(np.maximum(np.abs(o), 1e-3 * scale))
    same = np.mean(g == o)
    print(f"N={N}: {same:.4f} of the coordinates bit for bit, max {ulps.max():.1f} ulp")
    assert ulps.max() <= 8.0
    assert same > 0.5
    assert abs(o.std() * omega - 1.0) < 0.05


@pytest.mark.parametrize("precision", [64, 32])
def test_random_draws_follow_the_protocol(precision):
    """The walk's draws equal the oracle's Philox protocol (the protocol the
    golden fixtures were generated with): identical coordinates and uniforms,
    normals to the precision's rounding; and the FP32 normals are N(0,1)."""
    from vmachine_amd import engine as E
    n, seed, walker = 4096, 777, 5
    k, z, um, ua = E.draws(precision, seed, walker, 1000, n, 112)
    for i in range(0, n, 97):
        ok, oz, oum, oua = O.step_draws(seed, walker, 1000 + i, 112)
        assert k[i] == ok and um[i] == oum and ua[i] == oua
        assert abs(z[i] - oz) <= (1e-14 if precision == 64 else 2e-6) * max(1.0, abs(oz))
    k, z, um, ua = E.draws(precision, seed, walker, 0, 1 << 20, 112)
    assert abs(z.mean()) < 5e-3 and abs(z.std() - 1.0) < 5e-3
    assert abs(np.mean(z ** 3)) < 1e-2 and abs(np.mean(z ** 4) - 3.0) < 3e-2
    assert k.min() == 0 and k.max() == 111
