"""vmc_optimize_static, the optimisation loop that allocates no device memory,
against vmc_optimize (use_graph = 0), whose kernels it shares: the same
engine configuration, seed and starting parameters give the same final
parameter matrix, history and walker positions, bit for bit.

Two calls in a row on one GPU, the second with a smaller parameter matrix and
fewer iterations: the per-GPU array keeps the first call's contents, and the
second call must not see them.  Sizes beyond the array are rejected.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import engine_config, rbm_parameters

pytestmark = pytest.mark.gpu

CASES = [
    # N, omega, kinds, sampling, dt, precision, optimiser, lr, H, iterations
    (12, 1.0, [O.SLATER, O.RBM_GAUSSIAN, O.RBM_PRODUCT, O.PADE], O.IMPORTANCE_EXACT, 0.02, 32,
     "adam", 0.01, 24, 4),
    (6, 1.0, [O.SLATER, O.GAUSSIAN, O.PADE], O.IMPORTANCE_COMPAT, 0.05, 64, "sgd", 0.1, 0, 2),
]


def _engine(case):
    import vmachine_amd as V
    if V.device_count() < 1:
        pytest.fail("no GPU visible: GPU tests must run on the MI355X box")
    N, omega, kinds, sampling, dt, precision, kind, lr, H, iters = case
    cfg = engine_config(N, omega, kinds, sampling, dt, precision=precision, walkers=128,
                        refresh=40)
    if H:
        cfg.num_hidden = H
        cfg.sigma = 1.0
        cfg.sorting = 1
    return V.Engine(cfg)


def _start(case):
    N, omega, kinds, sampling, dt, precision, kind, lr, H, iters = case
    rng = np.random.default_rng(41 + N)
    P = rbm_parameters(rng, kinds, 2 * N, H) if H else np.zeros((len(kinds), 1))
    for e, k in enumerate(kinds):
        if k in (O.GAUSSIAN, O.PADE):
            P[e, 0] = 0.95 if k == O.GAUSSIAN else 1.05
    return P


def _run(case, static):
    N, omega, kinds, sampling, dt, precision, kind, lr, H, iters = case
    P0 = _start(case)
    with _engine(case) as e:
        e.init_walkers(123)
        if static:
            P, hist, _ = e.optimize_static(P0, iters, 60, 10, kind=kind, learning_rate=lr)
        else:
            P, hist, _ = e.optimize(P0, iters, 60, 10, kind=kind, learning_rate=lr,
                                    use_graph=False)
        return P, hist, e.positions()


def test_static_loop_equals_allocating_loop_twice():
    for i, case in enumerate(CASES):
        print(f"optimize_static call {i}: N={case[0]} params={_start(case).size} "
              f"iterations={case[-1]}", flush=True)
        ref = _run(case, static=False)
        got = _run(case, static=True)
        for a, b in zip(got, ref):
            np.testing.assert_array_equal(a, b)
        assert np.all(np.isfinite(got[1]))


def test_static_loop_rejects_too_many_iterations():
    import vmachine_amd as V
    case = CASES[1]
    with _engine(case) as e:
        e.init_walkers(5)
        with pytest.raises(V.VmcError, match="at most 8192 parameter-matrix entries and 1024"):
            e.optimize_static(_start(case), 1025, 10, 0, kind="sgd", learning_rate=0.1)
