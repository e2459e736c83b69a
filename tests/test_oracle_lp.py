"""The oracle's learned proposal (features, network, MH step, ESJD gradient).

It has no counterpart in the reference, so it is pinned by mathematics:
 * dJ/dtheta against central finite differences of J (every parameter block);
 * the zero-head network is a symmetric Gaussian random walk: the proposal
   ratio is 1, so log A = min(0, log|psi'/psi|^2) and J = A sigma0^2 |xi|^2;
 * the features: own coordinates, then the 8 nearest electrons in order of
   distance (brute force in numpy), zero-padded below 9 electrons;
 * the MH step accepts exactly when log R + log q_r - log q_f >= log u.
"""
import numpy as np
import pytest

from oracle import oracle as O

SGP = [O.SLATER, O.GAUSSIAN, O.PADE]


def _walker(N, omega, seed):
    rng = np.random.default_rng(seed)
    cfg = O.make_config(N, 2, omega, 1, SGP, None, None, False, O.IMPORTANCE_EXACT, 0.01, quirks=0)
    P = np.array([[0.0], [1.0], [0.6]])
    x = rng.standard_normal(2 * N) / np.sqrt(omega)
    return cfg, P, O.Walker(cfg, P, x), rng


@pytest.mark.parametrize("N", [2, 6, 12])
def test_features_against_brute_force(N):
    cfg, _, w, rng = _walker(N, 0.5, 3)
    x = w.positions()
    s = np.sqrt(0.5)
    for i in range(N):
        xi, yi = x[2 * i] + 0.1, x[2 * i + 1] - 0.2
        f, nn = O.lp_features(cfg, x, i, xi, yi)
        others = [j for j in range(N) if j != i]
        d2 = [(x[2 * j] - xi) ** 2 + (x[2 * j + 1] - yi) ** 2 for j in others]
        order = [others[k] for k in np.argsort(d2, kind="stable")][:8]
        want = np.zeros(18)
        want[0:2] = s * xi, s * yi
        for m, j in enumerate(order):
            want[2 + 2 * m] = s * (x[2 * j] - xi)
            want[3 + 2 * m] = s * (x[2 * j + 1] - yi)
        assert nn == min(8, N - 1)
        np.testing.assert_allclose(f, want, rtol=0, atol=1e-15)


def test_forward_against_numpy():
    rng = np.random.default_rng(5)
    th = O.lp_init(rng, zero_head=False)
    f = rng.standard_normal(18)
    h1, h2, o = O.lp_forward(th, f)
    W1 = th[O.LP_W1:O.LP_B1].reshape(64, 18)
    W2 = th[O.LP_W2:O.LP_B2].reshape(64, 64)
    W3 = th[O.LP_W3:O.LP_B3].reshape(4, 64)
    a1 = np.tanh(W1 @ f + th[O.LP_B1:O.LP_W2])
    a2 = np.tanh(W2 @ a1 + th[O.LP_B2:O.LP_W3])
    np.testing.assert_allclose(h1, a1, rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(h2, a2, rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(o, W3 @ a2 + th[O.LP_B3:], rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("N", [6, 12])
def test_zero_head_is_a_symmetric_random_walk(N):
    cfg, P, w, rng = _walker(N, 1.0, 11)
    hp = O.lp_hyper(sigma0=0.3)
    th = O.lp_init(rng, zero_head=True)
    for _ in range(10):
        i = int(rng.integers(0, N))
        xi = rng.standard_normal(2)
        J, A, logR, xn, yn, _ = w.lp_sample_grad(th, hp, i, xi[0], xi[1])
        x = w.positions()
        np.testing.assert_allclose([xn - x[2 * i], yn - x[2 * i + 1]], 0.3 * xi, rtol=1e-13)
        assert A == pytest.approx(min(1.0, np.exp(logR)), rel=1e-12)
        assert J == pytest.approx(A * 0.09 * (xi @ xi), rel=1e-12)


@pytest.mark.parametrize("N,omega", [(6, 1.0), (12, 0.5), (20, 0.5)])
def test_gradient_against_finite_differences(N, omega):
    cfg, P, w, rng = _walker(N, omega, 21 + N)
    hp = O.lp_hyper()
    th = O.lp_init(rng, zero_head=False) * 0.7
    blocks = [(O.LP_W1, O.LP_B1), (O.LP_B1, O.LP_W2), (O.LP_W2, O.LP_B2), (O.LP_B2, O.LP_W3),
              (O.LP_W3, O.LP_B3), (O.LP_B3, O.LP_NPARAM)]
    checked = 0
    for trial in range(12):
        i = int(rng.integers(0, N))
        xi = rng.standard_normal(2)
        J, A, logR, _, _, g = w.lp_sample_grad(th, hp, i, xi[0], xi[1])
        if 0.999 < A < 1.0 + 1e-12 and A != 1.0:
            continue   # too close to the kink of min(1, .)
        for lo, hi in blocks:
            v = np.zeros(O.LP_NPARAM)
            v[lo:hi] = rng.standard_normal(hi - lo)
            eps = 1e-6
            Jp = w.lp_sample_grad(th + eps * v, hp, i, xi[0], xi[1])[0]
            Jm = w.lp_sample_grad(th - eps * v, hp, i, xi[0], xi[1])[0]
            fd = (Jp - Jm) / (2 * eps)
            an = g @ v
            assert abs(fd - an) <= 1e-6 * max(abs(fd), abs(an), 1e-3), (trial, lo, fd, an)
            checked += 1
    assert checked >= 30


def test_mh_step_accepts_by_the_logged_weight():
    N = 6
    cfg, P, w, rng = _walker(N, 1.0, 31)
    hp = O.lp_hyper()
    th = O.lp_init(rng, zero_head=False) * 0.5
    for _ in range(20):
        i = int(rng.integers(0, N))
        xi = rng.standard_normal(2)
        u = float(rng.random())
        x0 = w.positions()
        _, A, _, xn, yn, _ = w.lp_sample_grad(th, hp, i, xi[0], xi[1])
        acc = w.lp_step(th, hp, i, xi[0], xi[1], u)
        assert acc == (A >= u)
        x1 = w.positions()
        if acc:
            assert x1[2 * i] == xn and x1[2 * i + 1] == yn
        else:
            np.testing.assert_array_equal(x1, x0)
