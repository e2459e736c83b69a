"""The optimisation loop on the device (vmc_optimize in include/vmc.h): the
walk, the sums, the reference's energy gradient and optimiser step and the
parameter update, iteration after iteration without a return to the host.

Pinned against the host loop the reference runs (ref src/system.cpp:81-139):
vmc_run for one iteration's sums, the host optimiser (libvmc_host, whose ADAM,
SGD and GD are pinned to the reference's own iterations by
test_optimizer_replays_reference_iterations in test_host.py) for the update, vmc_set_parameters for the next
iteration.  FP64 walks on the same random streams: every iteration's sums and
the final parameter matrix agree to 1e-9 relative (the device's pow and the
host's may differ in the last bit, which the chains do not see).  A captured
iteration replayed as a hipGraph equals direct launches bit for bit.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import engine_config, rbm_parameters

pytestmark = pytest.mark.gpu

SUMS = ["energy", "energy2", "kinetic", "external", "interaction", "kinetic2", "external2",
        "interaction2", "accepted"]

CASES = [
    # N, omega, kinds, sampling, dt, optimiser, lr, gamma, monotonic, H
    (6, 1.0, [O.SLATER, O.GAUSSIAN, O.PADE], O.IMPORTANCE_COMPAT, 0.05, "adam", 0.1, 0.0, 0.0, 0),
    (2, 1.0, [O.GAUSSIAN, O.PADE], O.BRUTE_FORCE, 0.5, "sgd", 0.1, 0.5, 0.3, 0),
    (12, 1.0, [O.SLATER, O.RBM_GAUSSIAN, O.RBM_PRODUCT, O.PADE], O.IMPORTANCE_EXACT, 0.02,
     "adam", 0.01, 0.0, 0.0, 24),
]


def _engine(case, precision=64, walkers=96):
    import vmachine_amd as V
    if V.device_count() < 1:
        pytest.fail("no GPU visible: GPU tests must run on the MI355X box")
    N, omega, kinds, sampling, dt, kind, lr, gamma, mono, H = case
    cfg = engine_config(N, omega, kinds, sampling, dt, precision=precision, walkers=walkers,
                        refresh=40)
    if H:
        cfg.num_hidden = H
        cfg.sigma = 1.0
        cfg.sorting = 1
    return V.Engine(cfg)


def _start(case, rng):
    N, omega, kinds, sampling, dt, kind, lr, gamma, mono, H = case
    P = rbm_parameters(rng, kinds, 2 * N, H) if H else np.zeros((len(kinds), 1))
    for e, k in enumerate(kinds):
        if k in (O.GAUSSIAN, O.PADE):
            P[e, 0] = 0.9 if k == O.GAUSSIAN else 1.1
    return P


@pytest.mark.parametrize("case", range(len(CASES)))
def test_device_loop_equals_host_loop(case):
    from vmachine_amd import host as HST
    c = CASES[case]
    kind, lr, gamma, mono = c[5], c[6], c[7], c[8]
    iters, steps, equil, seed = 3, 120, 20, 31 + case
    P0 = _start(c, np.random.default_rng(700 + case))
    print(f"device optimise case={case} kind={kind} N={c[0]}", flush=True)
    with _engine(c) as e:
        e.init_walkers(seed)
        Pd, hist, _ = e.optimize(P0, iters, steps, equil, kind=kind, learning_rate=lr, gamma=gamma,
                                 monotonic=mono)
        pos_d = e.positions()
    with _engine(c) as e:
        opt = HST.Optimizer(kind, P0.size, lr, gamma, mono)
        B = opt.batches
        e.init_walkers(seed)
        P = P0.copy()
        for it in range(iters):
            e.set_parameters(P)
            sums, g, ge, _ = e.run(steps, equil, it, B)
            ref = np.array([sums[k] for k in SUMS])
            np.testing.assert_allclose(hist[it, :9], ref, rtol=1e-9, atol=1e-9)
            upd = opt.update(sums["energy"] / sums["samples"], g / sums["gradient_samples"],
                             ge / sums["gradient_samples"])
            P = P - upd.reshape(P.shape)
        pos_h = e.positions()
    print(f"  final params max |diff| {np.max(np.abs(Pd - P)):.2e}", flush=True)
    np.testing.assert_allclose(Pd, P, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(pos_d, pos_h, rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("precision", [64, 32])
def test_graph_replay_equals_direct_launches(precision):
    c = CASES[2] if precision == 32 else CASES[0]
    P0 = _start(c, np.random.default_rng(5))
    out = {}
    for g in (True, False):
        print(f"device optimise precision={precision} graph={g}", flush=True)
        with _engine(c, precision=precision) as e:
            e.init_walkers(9)
            P, hist, _ = e.optimize(P0, 4, 80, 10, kind="adam", learning_rate=0.05, use_graph=g)
            out[g] = (P, hist, e.positions())
    for a, b in zip(out[True], out[False]):
        np.testing.assert_array_equal(a, b)


def test_optimisation_lowers_the_energy():
    """N=2, omega=1 Gauss x Pade from alpha = gamma = 0.5: ADAM on the device
    moves toward the exact ground state E = 3 (ref examples/no_interaction
    and the C1 configuration)."""
    c = (2, 1.0, [O.GAUSSIAN, O.PADE], O.IMPORTANCE_EXACT, 0.05, "adam", 0.1, 0.0, 0.0, 0)
    with _engine(c, walkers=4096) as e:
        e.init_walkers(3)
        P0 = np.array([[0.5], [0.5]])
        P, hist, ms = e.optimize(P0, 60, 200, 50, kind="adam", learning_rate=0.1)
        E = hist[:, 0] / (4096 * 200)
        print(f"  E first {E[0]:.5f} last {E[-1]:.5f} alpha {P[0, 0]:.4f} gamma {P[1, 0]:.4f} "
              f"({ms:.1f} ms)", flush=True)
        assert E[-1] < E[0]
        assert abs(E[-1] - 3.0) < 0.02
        np.testing.assert_allclose(hist[-1, 14:16], [P[0, 0], P[1, 0]], rtol=0, atol=0)
