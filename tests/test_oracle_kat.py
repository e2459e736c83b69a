"""Known-answer and closed-form pins of the oracle and of the product's
statistics module (no GPU)."""
import gzip
import os

import numpy as np
import pytest
from scipy.stats import chi2

from oracle import oracle as O
from tests.helpers import zero_variance_energy
from vmachine_amd import stats

FIX = os.path.join(os.path.dirname(__file__), "golden", "fixtures")


def _resampling_series():
    # the reference's own instant-energy file (examples/resampling/523861_0.dat,
    # 2^17 lines), read with strtod semantics as the reference does
    with gzip.open(os.path.join(FIX, "resampling_523861_0.dat.gz"), "rt") as f:
        return np.array([float(ln) for ln in f if ln.strip()])


def test_blocking_reproduces_reference_log():
    """ref examples/resampling/vmachine.log:182:
    Total energy : 3.0111013667 (with STD = 0.0030690608)"""
    x = _resampling_series()
    assert x.size == 2 ** 17
    m, se, nk = O.blocking(x)
    assert round(m, 10) == 3.0111013667
    assert round(se, 10) == 0.0030690608
    m2, se2, nk2 = stats.blocking(x)
    assert round(m2, 10) == 3.0111013667 and round(se2, 10) == 0.0030690608 and nk2 == nk


def test_blocking_rejects_non_power_of_two():
    with pytest.raises(ValueError):
        O.blocking(np.zeros(1000))
    with pytest.raises(ValueError):
        stats.blocking(np.zeros(1000))


def test_blocking_on_ar1_matches_closed_form():
    """AR(1) with coefficient phi: var(mean) -> sigma^2 (1+phi)/(1-phi) / n."""
    rng = np.random.default_rng(5)
    n, phi = 2 ** 18, 0.9
    e = rng.standard_normal(n)
    x = np.empty(n)
    x[0] = e[0]
    for i in range(1, n):
        x[i] = phi * x[i - 1] + np.sqrt(1 - phi * phi) * e[i]
    _, se, _ = stats.blocking(x)
    exact = np.sqrt((1 + phi) / (1 - phi) / n)
    assert abs(se / exact - 1) < 0.15
    tau = stats.iat_sokal(x)
    assert abs(tau - 0.5 * (1 + phi) / (1 - phi)) < 1.5


def test_chi2_quantiles():
    for k in [1, 2, 5, 17, 40, 64]:
        assert abs(O.chi2_quantile95(k) - chi2.ppf(0.95, k)) < 1e-9 * k


@pytest.mark.parametrize("ctr,key,expect", [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
])
def test_philox_known_answers(ctr, key, expect):
    """Random123 known-answer vectors for Philox4x32-10."""
    assert O.philox(ctr, key) == expect


def test_hermite_orbitals_against_recurrence_table():
    """The oracle's orbitals are c_n H_n(x) c_m H_m(y) with the textbook Hermite
    polynomials (H_0..H_3 written out) and c_n = 1/sqrt(2^n n!)."""
    cfg = O.make_config(6, kinds=[O.SLATER])
    w = O.Walker(cfg, np.zeros((1, 1)), np.linspace(-1, 1, 12))
    H = [lambda t: 1.0, lambda t: 2 * t, lambda t: 4 * t * t - 2, lambda t: 8 * t ** 3 - 12 * t]
    c = [1.0, 1 / np.sqrt(2), 1 / np.sqrt(8), 1 / np.sqrt(48)]
    r = np.array([0.3, -0.7])
    states = O.orbital_list(3)
    for j, (nx, ny) in enumerate(states):
        v = O.lib().vmco_basis_element(w.h, j, r.ctypes.data_as(O._DP))
        assert abs(v - c[nx] * H[nx](r[0]) * c[ny] * H[ny](r[1])) < 1e-14


@pytest.mark.parametrize("N", [2, 6, 12, 20, 30, 42, 56])
def test_oracle_zero_variance(N):
    """Non-interacting closed shells with alpha = 1: E_L = E0 exactly."""
    omega = 0.5
    cfg = O.make_config(N, 2, omega, 0, [O.SLATER, O.GAUSSIAN], step_length=0.01)
    rng = np.random.default_rng(N)
    w = O.Walker(cfg, np.array([[1.0], [1.0]]), rng.standard_normal(2 * N) / np.sqrt(omega))
    E0 = zero_variance_energy(N, omega)
    for _ in range(5):
        assert abs(w.local_energy() - E0) < 1e-8 * E0
        w.propose(int(rng.integers(2 * N)), float(rng.standard_normal()))
        w.accept()


@pytest.mark.parametrize("N,kinds", [
    (2, [O.GAUSSIAN, O.PADE]),
    (6, [O.SLATER, O.GAUSSIAN, O.PADE]),
    (12, [O.SLATER, O.GAUSSIAN, O.PADE]),
    (12, [O.SLATER, O.RBM_GAUSSIAN, O.RBM_PRODUCT, O.PADE]),
])
def test_oracle_derivatives_against_finite_differences(N, kinds):
    """Exact-mode elements: gradient and Laplacian of ln psi against central
    differences of ln|psi| built from the ratios of small moves."""
    cfg = O.make_config(N, 2, 1.0, 1, kinds, num_hidden=8, sorting=False, quirks=0)
    mp = O.max_parameters(cfg)
    rng = np.random.default_rng(3)
    P = 0.3 * rng.standard_normal((len(kinds), mp))
    for e, k in enumerate(kinds):
        if k in (O.GAUSSIAN, O.PADE):
            P[e, 0] = 0.9
    x0 = rng.standard_normal(2 * N)
    w = O.Walker(cfg, P, x0)
    grad = sum(w.gradient(e) for e in range(len(kinds)))
    lap = sum(w.laplacian(e) for e in range(len(kinds)))
    h = 1e-4
    fd_grad = np.zeros(2 * N)
    fd_lap = 0.0
    for kk in range(2 * N):
        lr = []
        for s in (+1, -1):
            w.propose(kk, x0[kk] + s * h)
            # ratios are |psi'/psi|^2 except the RBM-Gaussian (exact mode: also squared)
            lr.append(0.5 * np.log(w.ratio()))
            w.reject()
        fd_grad[kk] = (lr[0] - lr[1]) / (2 * h)
        fd_lap += (lr[0] + lr[1]) / (h * h)
    # lap ln psi = sum_k d2 ln psi, T uses lap + |grad|^2
    assert np.max(np.abs(fd_grad - grad)) < 1e-6 * max(1, np.max(np.abs(grad)))
    assert abs(fd_lap - lap) < 1e-4 * max(1, abs(lap))


def test_adam_one_step_closed_form():
    """First ADAM step from m = v = 1 (ref src/Optimization/adam.cpp:17-38)."""
    opt = O.Optimizer(0, 2, 0.1)
    g = 2 * (np.array([3.0, -1.0]) - 2.0 * np.array([1.0, 0.5]))
    upd = opt.update(2.0, np.array([1.0, 0.5]), np.array([3.0, -1.0]))
    m = 0.89 * 1 + 0.11 * g
    v = 0.999 * 1 + 0.001 * g * g
    exp = 0.1 * (m / 0.11) / (np.sqrt(v / 0.001) + 1e-8)
    assert np.allclose(upd, exp, rtol=1e-14)
