"""Every GPU sampler samples |psi|^2 at large N: the mean local energy of
non-interacting closed shells with a Gaussian factor alpha != 1 against its
closed form E0 (alpha + 1/alpha) / 2 (two derivations, pinned against each
other and against the oracle's walk in test_closed_shell_energy.py).

The forms: exact Langevin in FP64 and FP32 (the 28-orbital and the 64-lane
large-shell kernels), and the learned whole-electron proposal in FP32 with an
untrained network of random weights, on each of its network paths (bf16 dot
products at N = 12, matrix-core tiles with four walkers per wave at N = 20 and
two at N = 56).  The proposal's moves do not matter to the target: a wrong
Metropolis-Hastings ratio (proposal densities, determinant or Gaussian ratio)
or a stale LDS exchange in the network's two passes would leave E(alpha),
which differs from E0 by many standard errors.  Walkers start from the
engine's N(0, 1/omega) draw and are burned in; the error is the spread of the
walkers' means (independent chains).
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import zero_variance_energy

pytestmark = pytest.mark.gpu

CASES = [
    # N, omega, alpha, sampler, precision, dt, burn-in, steps
    (12, 1.0, 0.85, "langevin", 64, 0.1, 4000, 4096),
    (30, 0.5, 1.15, "langevin", 32, 0.1, 8000, 4096),
    (72, 0.1, 1.1, "langevin", 32, 0.3, 30000, 8192),
    (12, 1.0, 1.1, "learned", 32, 0.01, 3000, 4096),
    (20, 0.5, 0.85, "learned", 32, 0.01, 10000, 4096),
    (56, 0.1, 0.9, "learned", 32, 0.01, 20000, 8192),
]


@pytest.mark.parametrize("N,omega,alpha,sampler,precision,dt,burn,steps", CASES)
def test_sampled_mean_energy_is_the_closed_form(N, omega, alpha, sampler, precision, dt, burn,
                                                steps):
    import vmachine_amd as V
    if V.device_count() < 1:
        pytest.fail("no GPU visible: GPU tests must run on the MI355X box")
    W = 2048
    samp = V.LEARNED if sampler == "learned" else V.IMPORTANCE_EXACT
    cfg = V.make_config(N, omega=omega, interaction=False, kinds=[V.SLATER, V.GAUSSIAN],
                        sampling=samp, step_length=dt, quirks=0, precision=precision,
                        refresh_interval=100 if precision == 32 else 0, num_walkers=W,
                        drift_limit=1.0)
    print(f"closed shell N={N} omega={omega} alpha={alpha} {sampler} FP{precision}", flush=True)
    with V.Engine(cfg) as e:
        e.set_parameters(np.array([[0.0], [alpha]]))
        if sampler == "learned":
            rng = np.random.default_rng(600 + N)
            e.lp_set((O.lp_init(rng, zero_head=False) * 0.3).astype(np.float32))
        e.init_walkers(900 + N)
        for _ in range(burn // 1000):
            e.run(1000, 0)
        sums, _, _, means = e.run_blocked(steps, 64)
    assert sums["nonfinite"] == 0
    walker_means = means.mean(axis=1)
    E = float(walker_means.mean())
    se = float(walker_means.std(ddof=1) / np.sqrt(W))
    E0 = zero_variance_energy(N, omega)
    exact = E0 * (alpha + 1.0 / alpha) / 2.0
    print(f"  E = {E:.5f} +- {se:.1e}, exact {exact:.5f} (E0 {E0:.5f}), "
          f"acceptance {sums['accepted'] / sums['samples']:.3f}")
    assert abs(E - exact) < 4 * se
    assert abs(exact - E0) > 10 * se
