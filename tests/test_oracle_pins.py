"""Pins of two oracle functions that no reference golden vector covers.

1. vmco_initial_positions (the RandomNormal initial state,
   ref src/InitialStates/randomnormal.cpp:10-25): every coordinate is drawn
   from N(0, sigma) with sigma = 1/omega, because the reference passes 1/omega
   to std::normal_distribution, whose second argument is the standard
   deviation (ref src/RNG/mersennetwister.cpp:75-77).  Checked by the law
   itself over 10^6 draws: mean, standard deviation, the fourth moment of a
   normal, and no correlation between the two coordinates a Box-Muller pair
   gives.  Taking 1/omega as the variance (sigma = 1/sqrt(omega)) or
   1/sqrt(omega) as the standard deviation fails the standard-deviation check
   at omega = 0.1 by a factor of three.

2. The exact Metropolis-Hastings Langevin step of the oracle (VMC_IMPORTANCE_EXACT:
   proposal x_k' = x_k + dt/2 F_k + sqrt(dt) z, acceptance
   |psi'/psi|^2 G(x|x')/G(x'|x)).  For a Gaussian trial function
   psi = exp(-alpha omega |x|^2 / 2) of N = 2 non-interacting particles in
   D = 2 (ref src/WaveFunctions/gaussian.cpp:8-78, harmonic oscillator
   ref src/Hamiltonians/harmonicoscillator.cpp:16-20) the sampled density
   |psi|^2 is a normal law of variance 1/(2 alpha omega) per coordinate, so
   <sum x^2> = N D / (2 alpha omega) and <E_L> = omega (alpha + 1/alpha)
   (N D = 4).  At dt = 0.5 the Langevin proposal alone is far from exact
   (its own stationary variance is dt / (1 - (1 - alpha omega dt)^2), 21 %
   above the target at alpha = 0.7, omega = 1), so a dropped or sign-flipped
   Green's-function ratio misses both closed forms by many standard errors.

3. The oracle's scaled basis (basis_scaled: the closed-shell orbitals as
   H_nx(sqrt(omega) x) H_ny(sqrt(omega) y) instead of the reference's
   H_nx(x) H_ny(y), ref src/Basis/hermite.cpp:33-71).  Both sets span the same
   polynomials of degree <= n per shell, so the Slater matrices differ by an
   invertible shell-triangular factor and every determinant ratio, local
   energy, drift and parameter gradient is the same.  Checked on random
   configurations at the closed shells up to N = 42, where the reference's
   basis is still well conditioned in FP64; the oracle uses the scaled basis
   past N = 56, where it is not.
"""
import numpy as np
import pytest

from oracle import oracle as O


@pytest.mark.parametrize("omega", [0.1, 1.0])
def test_initial_positions_follow_the_reference_law(omega):
    nd, walkers, seed = 200, 5000, 20261203
    x = np.stack([O.initial_positions(seed, w, nd, omega) for w in range(walkers)])
    n = x.size
    sigma = 1.0 / omega
    z = x.ravel() / sigma
    assert abs(z.mean()) < 5.0 / np.sqrt(n)
    assert abs(z.std() - 1.0) < 5.0 * np.sqrt(0.5 / n)
    assert abs(np.mean(z ** 4) - 3.0) < 5.0 * np.sqrt(96.0 / n)
    pairs = x.reshape(walkers, nd // 2, 2) / sigma
    assert abs(np.mean(pairs[..., 0] * pairs[..., 1])) < 5.0 / np.sqrt(n / 2)
    # a different walker or seed gives a different stream
    assert not np.array_equal(x[0], x[1])
    assert not np.array_equal(x[0], O.initial_positions(seed + 1, 0, nd, omega))
    # reproducible
    np.testing.assert_array_equal(x[3], O.initial_positions(seed, 3, nd, omega))


def _gaussian_chain_estimates(alpha, omega, dt, walkers, steps, equil, seed):
    cfg = O.make_config(2, 2, omega, 0, [O.GAUSSIAN], sampling=O.IMPORTANCE_EXACT,
                        step_length=dt, quirks=0)
    P = np.array([[alpha]])
    E, r2, acc = np.zeros(walkers), np.zeros(walkers), np.zeros(walkers)
    for w in range(walkers):
        walker = O.Walker(cfg, P, O.initial_positions(seed, w, 4, omega))
        _, a, _, _, _ = walker.run(seed, w, 0, steps, equil)
        E[w] = a[0] / steps
        r2[w] = 2.0 * a[3] / (omega * omega * steps)   # Vext = omega^2 |x|^2 / 2
        acc[w] = a[8] / steps
    return E, r2, acc


@pytest.mark.parametrize("alpha,omega,dt", [(0.7, 1.0, 0.5), (1.3, 0.5, 0.8)])
def test_exact_langevin_step_samples_psi_squared(alpha, omega, dt):
    E, r2, acc = _gaussian_chain_estimates(alpha, omega, dt, walkers=256, steps=8000, equil=500,
                                           seed=977)
    E_exact = omega * (alpha + 1.0 / alpha)
    r2_exact = 4.0 / (2.0 * alpha * omega)
    se_E = E.std() / np.sqrt(E.size)
    se_r2 = r2.std() / np.sqrt(r2.size)
    print(f"alpha={alpha} omega={omega} dt={dt}: E = {E.mean():.5f} +- {se_E:.1e} "
          f"(exact {E_exact:.5f}), <x^2> = {r2.mean():.5f} +- {se_r2:.1e} "
          f"(exact {r2_exact:.5f}), acceptance {acc.mean():.3f}")
    assert abs(E.mean() - E_exact) < 4.0 * se_E
    assert abs(r2.mean() - r2_exact) < 4.0 * se_r2
    # the test has power: the proposal's own bias is many standard errors
    q = 1.0 - alpha * omega * dt
    r2_proposal = 4.0 * dt / (1.0 - q * q)
    assert abs(r2_proposal - r2_exact) > 20.0 * se_r2
    assert 0.3 < acc.mean() < 0.99


@pytest.mark.parametrize("N,omega", [(2, 1.0), (6, 1.0), (12, 0.5), (20, 0.5), (30, 0.28),
                                     (42, 0.28)])
def test_scaled_basis_gives_the_reference_basis_values(N, omega):
    from tests.helpers import oracle_evaluate, random_parameters
    kinds = [O.SLATER, O.GAUSSIAN, O.PADE]
    rng = np.random.default_rng(500 + N)
    W = 12
    P = random_parameters(rng, kinds)
    pos = rng.standard_normal((W, 2 * N)) / np.sqrt(omega)
    moves = rng.integers(0, 2 * N, W)
    newv = pos[np.arange(W), moves] + 0.3 / np.sqrt(omega) * rng.standard_normal(W)
    res = []
    for scaled in (False, True):
        cfg = O.make_config(N, 2, omega, 1, kinds, None, None, False, O.IMPORTANCE_EXACT, 0.01,
                            quirks=O.Q2 | O.Q3, basis_scaled=scaled)
        res.append(oracle_evaluate(cfg, P, pos, moves, newv))
    for a, b in zip(res[0], res[1]):
        scale = np.maximum(np.abs(a), 1.0)
        assert np.max(np.abs(a - b) / scale) < 1e-8, np.max(np.abs(a - b) / scale)
    # not vacuous: the same comparison sees a 10 % change of the y coordinates
    cfg = O.make_config(N, 2, omega, 1, kinds, None, None, False, O.IMPORTANCE_EXACT, 0.01,
                        quirks=O.Q2 | O.Q3, basis_scaled=True)
    moved = oracle_evaluate(cfg, P, pos * np.array([1.0, 0.9] * N), moves, newv)
    assert np.max(np.abs(moved[0][:, 0] - res[0][0][:, 0])) > 1e-3
