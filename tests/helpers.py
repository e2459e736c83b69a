"""Shared helpers: matching oracle and engine configurations."""
import numpy as np

from oracle import oracle as O

# closed-shell particle counts in 2-D and the frequencies used for them (72,
# 90 and 110 are the shells past the configurations' largest, N = 56)
CLOSED_SHELLS = [2, 6, 12, 20, 30, 42, 56, 72, 90, 110]
OMEGA_FOR_N = {2: 1.0, 6: 1.0, 12: 1.0, 20: 0.5, 30: 0.5, 42: 0.28, 56: 0.1, 72: 0.1, 90: 0.1,
               110: 0.1}


def zero_variance_energy(N, omega):
    """Non-interacting closed shells: E0 = 2 sum_{s<S} (s+1)^2 omega."""
    E, s, filled = 0.0, 0, 0
    while filled < N:
        E += 2 * (s + 1) * (s + 1) * omega
        filled += 2 * (s + 1)
        s += 1
    assert filled == N
    return E


def closed_shell_mean_energy(N, omega, alpha):
    """<E_L> under |psi|^2 for N non-interacting electrons of a 2-D oscillator
    in closed shells, psi = det_up det_down exp(-alpha omega |r|^2 / 2) with
    the closed-shell orbitals H_nx(sqrt(omega) x) H_ny(sqrt(omega) y), nx + ny
    <= S (ref src/WaveFunctions/slaterdeterminant.cpp, gaussian.cpp,
    src/Hamiltonians/harmonicoscillator.cpp).  For alpha != 1 the orbitals
    phi_k = H_nx H_ny exp(-alpha omega r^2 / 2) are not orthogonal; a
    determinant is unchanged by an invertible mixing of its orbitals, so the
    mean energy is that of the orthogonalised ones, 2 Tr(S^-1 H) (both spins)
    with S and H the orbitals' overlap and one-body Hamiltonian matrices
    (H = -grad^2/2 + omega^2 r^2/2, the kinetic part as (1/2) int grad phi_k .
    grad phi_l).  Every matrix element is a product of 1-D integrals of a
    polynomial times exp(-alpha omega x^2), exact with Gauss-Hermite
    quadrature.  (The per-shell Hermite polynomials of sqrt(omega) x and of
    sqrt(alpha omega) x span the same space, so psi is the ground state of the
    oscillator of frequency alpha omega and the result is also
    E0 (alpha + 1/alpha) / 2 by the virial theorem; tests compare the two.)"""
    from numpy.polynomial import hermite as Hm
    NP, S = N // 2, 0
    while (S + 1) * (S + 2) // 2 < NP:
        S += 1
    assert (S + 1) * (S + 2) // 2 == NP, N
    orbs = [(nx, s - nx) for s in range(S + 1) for nx in range(s + 1)]
    u, wq = Hm.hermgauss(40)
    x = u / np.sqrt(alpha * omega)              # int f(x) e^{-alpha omega x^2} dx = sum wq f(x) / sqrt(alpha omega)
    jac = 1.0 / np.sqrt(alpha * omega)
    t = np.sqrt(omega) * x
    Hv = [Hm.hermval(t, [0] * n + [1]) for n in range(S + 1)]
    # d/dx [H_n(sqrt(omega) x) e^{-alpha omega x^2/2}] / e^{-alpha omega x^2/2}
    dv = [np.sqrt(omega) * (2 * n * Hv[n - 1] if n else 0.0) - alpha * omega * x * Hv[n]
          for n in range(S + 1)]
    m = S + 1
    s1 = np.array([[np.sum(wq * Hv[a] * Hv[b]) * jac for b in range(m)] for a in range(m)])
    v1 = np.array([[np.sum(wq * x * x * Hv[a] * Hv[b]) * jac for b in range(m)] for a in range(m)])
    t1 = np.array([[np.sum(wq * dv[a] * dv[b]) * jac for b in range(m)] for a in range(m)])
    Sm = np.array([[s1[a[0], b[0]] * s1[a[1], b[1]] for b in orbs] for a in orbs])
    Hmat = np.array([[0.5 * (t1[a[0], b[0]] * s1[a[1], b[1]] + s1[a[0], b[0]] * t1[a[1], b[1]]) +
                      0.5 * omega ** 2 * (v1[a[0], b[0]] * s1[a[1], b[1]] +
                                          s1[a[0], b[0]] * v1[a[1], b[1]])
                      for b in orbs] for a in orbs])
    return 2.0 * float(np.trace(np.linalg.solve(Sm, Hmat)))


def oracle_config(N, omega, kinds, sampling, dt, interaction=True, quirks=O.Q2 | O.Q3,
                  drift_limit=0.0, basis_scaled=None):
    """The oracle with the reference's basis H_n(x); past N = 56 (shells of
    degree 7 and more at omega = 0.1, where the unscaled matrix loses digits in
    FP64) with H_n(sqrt(omega) x), which spans the same polynomials, so every
    ratio, energy and derivative is the same (pinned in test_oracle_pins.py)."""
    if basis_scaled is None:
        basis_scaled = N > 56
    return O.make_config(N, 2, omega, int(interaction), kinds, None, None, False, sampling, dt,
                         quirks=quirks, basis_scaled=bool(basis_scaled), drift_limit=drift_limit)


def engine_config(N, omega, kinds, sampling, dt, interaction=True, quirks=0b1100,
                  precision=64, walkers=64, refresh=0, drift_limit=0.0):
    import vmachine_amd as V
    return V.make_config(N, omega=omega, interaction=interaction, kinds=kinds, sampling=sampling,
                         step_length=dt, quirks=quirks, precision=precision,
                         refresh_interval=refresh, num_walkers=walkers, drift_limit=drift_limit)


def random_parameters(rng, kinds):
    P = np.zeros((len(kinds), 1))
    for e, k in enumerate(kinds):
        if k == O.GAUSSIAN:
            P[e, 0] = 1.0 + 0.1 * rng.standard_normal()
        elif k == O.PADE:
            P[e, 0] = 0.3 + 0.7 * rng.random()
    return P


def oracle_evaluate(ocfg, P, positions, moves, new_values):
    """The oracle's version of vmc_evaluate: rows [E, T, Vext, Vint, Ga, Gg, rS, rP, rG, drift...]."""
    kinds = [ocfg.kinds[i] for i in range(ocfg.num_elements)]
    W, nd = positions.shape
    out0 = np.zeros((W, 9 + nd))
    out1 = np.zeros((W, 9 + nd))

    def fill(w, row):
        T, V, I = w.kinetic(), w.external(), w.interaction()
        row[0:4] = [T + V + I, T, V, I]
        for e, k in enumerate(kinds):
            if k == O.GAUSSIAN:
                row[4] = w.parameter_gradient(e)[0]
            if k == O.PADE:
                row[5] = w.parameter_gradient(e)[0]
        row[9:] = w.drift()

    for i in range(W):
        w = O.Walker(ocfg, P, positions[i])
        fill(w, out0[i])
        w.propose(int(moves[i]), float(new_values[i]))
        for e, k in enumerate(kinds):
            col = {O.SLATER: 6, O.PADE: 7, O.GAUSSIAN: 8}[k]
            out1[i, col] = w.ratio(e)
        w.accept()
        fill(w, out1[i])
        if O.SLATER not in kinds:
            out1[i, 6] = 1.0
        if O.PADE not in kinds:
            out1[i, 7] = 1.0
        if O.GAUSSIAN not in kinds:
            out1[i, 8] = 1.0
    return out0, out1


def oracle_evaluate_elements(ocfg, P, positions, moves, new_values):
    """The oracle's version of vmc_evaluate_elements: (o0, o1) as oracle_evaluate,
    per-element ratios of the forced move, and parameter-gradient matrices."""
    kinds = [ocfg.kinds[i] for i in range(ocfg.num_elements)]
    W, nd = positions.shape
    E = len(kinds)
    mp = O.max_parameters(ocfg)
    o0, o1 = np.zeros((W, 9 + nd)), np.zeros((W, 9 + nd))
    r = np.zeros((W, E))
    g0, g1 = np.zeros((W, E, mp)), np.zeros((W, E, mp))

    def fill(w, row, g):
        T, V, I = w.kinetic(), w.external(), w.interaction()
        row[0:4] = [T + V + I, T, V, I]
        for e, k in enumerate(kinds):
            pg = w.parameter_gradient(e)
            g[e, :] = pg
            if k == O.GAUSSIAN:
                row[4] = pg[0]
            if k == O.PADE:
                row[5] = pg[0]
        row[9:] = w.drift()

    for i in range(W):
        w = O.Walker(ocfg, P, positions[i])
        fill(w, o0[i], g0[i])
        w.propose(int(moves[i]), float(new_values[i]))
        for e in range(E):
            r[i, e] = w.ratio(e)
        w.accept()
        fill(w, o1[i], g1[i])
    return o0, o1, r, g0, g1


def rbm_parameters(rng, kinds, nd, H):
    """A parameter matrix with Xavier-like RBM weights (ref xavier.cpp:10-23)."""
    mp = 1
    for k in kinds:
        if k == O.RBM_GAUSSIAN:
            mp = max(mp, nd)
        if k == O.RBM_PRODUCT:
            mp = max(mp, H + nd * H)
    P = np.zeros((len(kinds), mp))
    a = 1.0 / np.sqrt(nd)
    for e, k in enumerate(kinds):
        if k == O.GAUSSIAN:
            P[e, 0] = 1.0 + 0.1 * rng.standard_normal()
        elif k == O.PADE:
            P[e, 0] = 0.3 + 0.7 * rng.random()
        elif k == O.RBM_GAUSSIAN:
            P[e, :nd] = rng.uniform(-a, a, nd)
        elif k == O.RBM_PRODUCT:
            P[e, :H + nd * H] = rng.uniform(-a, a, H + nd * H)
    return P
