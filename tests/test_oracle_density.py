"""The oracle's electron densities (ref src/sampler.cpp:639-692).

Pinned two ways: against an independent vectorised numpy statement of the
three binning rules (truncation toward zero, the centre cell spanning 4 dr,
the two-body bins shifted by one), and, for the spatial grid, against the
mathematics: exact samples of the non-interacting 2-D oscillator ground state
(|psi|^2 = Gaussian of variance 1/(2 omega) per coordinate) must fill the
cells as the Gaussian's cell integrals say (chi-square).  The reference's own
example output (examples/electron_density/onebody2.dat, fixed under
tests/golden/examples) carries the same centre-row/column structure.
"""
import math
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "examples")


def numpy_density(X, bins, R):
    """X [S, N, 2]: counts over all S configurations."""
    dr = R / bins
    half = bins // 2
    r = np.sqrt(X[..., 0] ** 2 + X[..., 1] ** 2)
    radial = np.zeros(bins, dtype=np.int64)
    q = r / dr
    ok = (r < R) & (np.trunc(q) < bins)
    np.add.at(radial, np.trunc(q[ok]).astype(np.int64), 1)
    ix = np.trunc(X[..., 0] / (2 * dr)).astype(np.int64)
    iy = np.trunc(X[..., 1] / (2 * dr)).astype(np.int64)
    ok = (np.abs(ix) < half) & (np.abs(iy) < half)
    spatial = np.zeros((bins, bins), dtype=np.int64)
    np.add.at(spatial, (iy[ok] + half, ix[ok] + half), 1)
    b = np.trunc(q).astype(np.int64) + 1
    two = np.zeros((bins, bins), dtype=np.int64)
    N = X.shape[1]
    for i in range(N):
        for j in range(i + 1, N):
            ok = (b[:, i] < bins) & (b[:, j] < bins)
            np.add.at(two, (b[ok, i], b[ok, j]), 1)
    return radial, spatial, two


@pytest.mark.parametrize("N,bins,R", [(2, 100, 3.0), (6, 37, 4.5), (12, 64, 2.0)])
def test_density_against_numpy(N, bins, R):
    rng = np.random.default_rng(N)
    X = rng.standard_normal((200, N, 2)) * 1.3
    X[:5] = np.round(X[:5] / (R / bins)) * (R / bins)   # values on cell edges
    radial = np.zeros(bins, dtype=np.int64)
    spatial = np.zeros((bins, bins), dtype=np.int64)
    two = np.zeros((bins, bins), dtype=np.int64)
    for x in X:
        O.density_add(x.ravel(), N, bins, R, radial, spatial, two)
    r2, s2, t2 = numpy_density(X, bins, R)
    np.testing.assert_array_equal(radial, r2)
    np.testing.assert_array_equal(spatial, s2)
    np.testing.assert_array_equal(two, t2)
    assert two[0, :].sum() == 0 and two[:, 0].sum() == 0   # shifted bins never use 0


def test_spatial_density_of_the_oscillator_ground_state():
    """Exact samples of exp(-omega r^2): cell counts against the Gaussian's
    cell integrals, the centre row and column spanning (-2dr, 2dr)."""
    omega, bins, R, S = 1.0, 40, 3.0, 400000
    rng = np.random.default_rng(7)
    X = rng.standard_normal((S, 1, 2)) / math.sqrt(2 * omega)
    spatial = np.zeros((bins, bins), dtype=np.int64)
    # the oracle bins one configuration per call: pack 100 samples as one
    # "configuration" of 100 particles (the spatial rule is per particle)
    for k in range(0, S, 100):
        O.density_add(X[k:k + 100].ravel(), 100, bins, R, None, spatial, None)
    assert spatial.sum() <= S
    dr = R / bins
    half = bins // 2
    # cell c (column index) covers (c - half) * 2dr .. with truncation: index 0 covers (-2dr, 2dr)
    lo = np.array([(c - half) * 2 * dr if c > half else ((c - half - 1) * 2 * dr if c < half else -2 * dr)
                   for c in range(bins)])
    hi = np.array([(c - half + 1) * 2 * dr if c > half else ((c - half) * 2 * dr if c < half else 2 * dr)
                   for c in range(bins)])
    s = 1 / math.sqrt(2 * omega)
    cdf = lambda v: 0.5 * (1 + np.vectorize(math.erf)(v / (s * math.sqrt(2))))
    p1 = cdf(hi) - cdf(lo)
    expected = S * np.outer(p1, p1)
    mask = expected > 20
    chi2 = (((spatial - expected) ** 2) / expected)[mask].sum()
    dof = mask.sum()
    assert chi2 < dof + 6 * math.sqrt(2 * dof), (chi2, dof)
    # the centre row and column hold twice the width of their neighbours
    assert spatial[half, half] > 1.5 * spatial[half + 1, half + 1]


def test_reference_example_grid_structure():
    """examples/electron_density/onebody2.dat (N=2, 100 bins, maxRadius 3,
    100000 steps): 2 x 100000 counts, the doubled centre row and column."""
    g = np.loadtxt(os.path.join(GOLDEN, "electron_density", "onebody2.dat"), dtype=np.int64)
    assert g.shape == (100, 100) and g.sum() == 200000
    assert g[50, 50] == g.max()
    ring = g[48:53, 48:53].astype(float)
    assert ring[2, :].mean() > 1.5 * np.delete(ring, 2, axis=0).mean()
