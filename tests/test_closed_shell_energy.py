"""The mean local energy of non-interacting closed shells with a Gaussian
factor alpha != 1: a closed form that pins the samplers at N > 2, where the
determinant's drift enters the proposal.

psi = det_up det_down exp(-alpha omega |r|^2 / 2) over the closed-shell
orbitals H_nx(sqrt(omega) x) H_ny(sqrt(omega) y) (ref
src/WaveFunctions/slaterdeterminant.cpp, gaussian.cpp, src/Basis/hermite.cpp)
in the oscillator of frequency omega (ref src/Hamiltonians/harmonicoscillator.cpp).
Two derivations of <E_L> under |psi|^2:
  * Loewdin's rule for a determinant of non-orthogonal orbitals, 2 Tr(S^-1 H),
    with the overlap and one-body matrices from Gauss-Hermite quadrature
    (tests.helpers.closed_shell_mean_energy);
  * the virial theorem: the shells' Hermite polynomials of sqrt(omega) x and
    of sqrt(alpha omega) x span the same space, so psi is the ground state of
    the oscillator of frequency alpha omega, <T> = alpha E0 / 2 and
    <V> = E0 / (2 alpha), E0 = 2 sum_s (s+1)^2 omega.
They must agree to rounding; then the oracle's exact Langevin walk must sample
that mean within 4 standard errors at N = 6 and 12, where E(alpha) - E0 is
many standard errors, so a sampler off |psi|^2 misses it.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import closed_shell_mean_energy, zero_variance_energy


def virial(N, omega, alpha):
    return zero_variance_energy(N, omega) * (alpha + 1.0 / alpha) / 2.0


@pytest.mark.parametrize("N", [2, 6, 12, 20, 30, 42, 56, 72])
@pytest.mark.parametrize("omega,alpha", [(1.0, 0.7), (0.28, 1.3), (0.1, 0.9), (0.5, 1.0)])
def test_two_derivations_agree(N, omega, alpha):
    q = closed_shell_mean_energy(N, omega, alpha)
    assert q == pytest.approx(virial(N, omega, alpha), rel=1e-10)
    if N == 2:
        assert q == pytest.approx(omega * (alpha + 1.0 / alpha), rel=1e-12)
    if alpha == 1.0:
        assert q == pytest.approx(zero_variance_energy(N, omega), rel=1e-12)


@pytest.mark.parametrize("N,omega,alpha,dt", [(6, 1.0, 0.8, 0.1), (12, 0.5, 1.2, 0.2)])
def test_oracle_langevin_samples_the_closed_form(N, omega, alpha, dt):
    ocfg = O.make_config(N, 2, omega, 0, [O.SLATER, O.GAUSSIAN], None, None, False,
                         O.IMPORTANCE_EXACT, dt, quirks=0)
    P = np.array([[0.0], [alpha]])
    W, steps, equil, seed = 32, 6000, 1500, 41 + N
    means = []
    for w in range(W):
        wk = O.Walker(ocfg, P, O.initial_positions(seed, w, 2 * N, omega))
        _, _, _, _, en = wk.run(seed, w, 0, steps, equil, record=True)
        means.append(en.mean())
    means = np.array(means)
    E, se = means.mean(), means.std(ddof=1) / np.sqrt(W)
    exact = virial(N, omega, alpha)
    print(f"N={N} alpha={alpha}: E = {E:.5f} +- {se:.1e}, exact {exact:.5f}, "
          f"E0 {zero_variance_energy(N, omega):.5f}")
    assert abs(E - exact) < 4 * se
    assert abs(exact - zero_variance_energy(N, omega)) > 10 * se
