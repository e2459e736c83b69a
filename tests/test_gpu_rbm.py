"""RBM-Gaussian and RBM-product elements in the HIP walk kernel against the
FP64 oracle (whose RBM functions are pinned to the reference's golden vectors
el_n2_rbm, el_n12_rbmpj_sorted/unsorted and it_n12_rbmpj_adam in
test_oracle_golden.py).

Per configuration the engine and the oracle evaluate the same positions and
parameters, force the same move through the incremental path and evaluate
again: local energy and parts, drift, per-element ratios and the full
parameter-gradient matrices (600 RBM-product parameters at N=12, H=24).  FP64
trajectories must retrace the oracle's chains step for step, including the
gradient-window sums of every RBM parameter.

The FP32 walk runs the RBM-product contractions (the proposal's pre-activation
W^T x', the gradient W n after an accepted move) three ways: FMA chains on the
walker's lanes, FP32 matrix-core tiles (V_MFMA_F32_16X16X4_F32) and BF16
matrix-core tiles with FP32 accumulation (V_MFMA_F32_16X16X32_BF16).

Tolerances: FP64 1e-9 scaled (as in test_gpu_parity); FP32 2e-3 max / 2e-5
median on energies and drifts, and for the ratio and parameter gradients
relative 1e-4 (single-precision pre-activations v = b + W^T x / sigma_p^2 over
24 terms), for both FP32 contraction modes.  BF16 inputs carry a relative
error of 2^-9 per product; the contractions enter v and d ln psi/dx scaled by
1/sigma_p^2 = 1/(100 sigma^2) (ref rbmproduct.cpp:15-17), so energies, drifts
and parameter gradients stay within the FP32 tolerances, while the RBM-product
ratio, a product over the H hidden units of factors that each carry the
error of one pre-activation, is allowed 5e-4 (measured 1.5e-4 at N=12, H=24).
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import oracle_evaluate_elements, rbm_parameters

pytestmark = pytest.mark.gpu

RBMPJ = [O.SLATER, O.RBM_GAUSSIAN, O.RBM_PRODUCT, O.PADE]
RBM = [O.RBM_GAUSSIAN, O.RBM_PRODUCT]
REF = O.Q2 | O.Q3 | O.Q6 | O.Q13

CASES = [
    # N, omega, kinds, H, sorting, quirks
    (2, 1.0, RBM, 2, True, REF),
    (2, 1.0, RBM, 2, False, REF),
    (2, 1.0, RBM, 12, True, REF),          # more hidden units than one lane holds
    (6, 1.0, RBM + [O.PADE], 8, True, REF),
    (12, 1.0, RBMPJ, 24, True, REF),
    (12, 1.0, RBMPJ, 24, False, REF),
    (12, 0.5, RBMPJ, 24, True, 0),          # the exact-form ratios and gradients
    (2, 1.0, [O.RBM_PRODUCT], 4, True, REF),
    (6, 1.0, [O.SLATER, O.RBM_GAUSSIAN], 0, True, REF),
]


def _engine(cfg):
    import vmachine_amd as V
    if V.device_count() < 1:
        pytest.fail("no GPU visible: GPU tests must run on the MI355X box")
    return V.Engine(cfg)


def _configs(N, omega, kinds, H, sorting, quirks, sampling, dt, precision, W, mode=0):
    import vmachine_amd as V
    ocfg = O.make_config(N, 2, omega, 1, kinds, H if H else None, None, sorting, sampling, dt,
                         quirks=quirks, basis_scaled=False)
    ecfg = V.make_config(N, omega=omega, interaction=True, kinds=kinds, sampling=sampling,
                         step_length=dt, quirks=quirks, precision=precision, num_walkers=W,
                         num_hidden=H if H else N, sorting=sorting, rbm_matrix=mode)
    return ocfg, ecfg


def _mf_applies(N, kinds, H):
    """Matrix-core contractions exist for an RBM-product element with >= 4
    lanes per walker (4 lanes hold 16 hidden units; N = 2 with H <= 4 has 1)."""
    if O.RBM_PRODUCT not in kinds:
        return False
    np_ = {2: 1, 6: 3, 12: 6}[N]
    K = {1: 1, 3: 4, 6: 8}[np_]
    while 4 * K < H:
        K *= 2
    return K >= 4


# (precision, rbm_matrix): FP64; FP32 lanes, FP32 matrix cores, BF16 matrix cores
MODES = [(64, 0), (32, 1), (32, 2), (32, 3)]


def _scaled(a, b):
    return np.abs(a - b) / (np.abs(b) + 1.0)


@pytest.mark.parametrize("mode", range(len(MODES)))
@pytest.mark.parametrize("case", range(len(CASES)))
def test_rbm_evaluate_matches_oracle(case, mode):
    N, omega, kinds, H, sorting, quirks = CASES[case]
    precision, rbm_matrix = MODES[mode]
    if rbm_matrix >= 2 and not _mf_applies(N, kinds, H):
        pytest.skip("no matrix-core contractions for this shape")
    rng = np.random.default_rng(300 + case)
    W, nd = 40, 2 * N
    ocfg, ecfg = _configs(N, omega, kinds, H, sorting, quirks, O.IMPORTANCE_COMPAT, 0.05,
                          precision, W, rbm_matrix)
    P = rbm_parameters(rng, kinds, nd, H)
    pos = rng.standard_normal((W, nd)) / np.sqrt(omega)
    moves = rng.integers(0, nd, W)
    newv = pos[np.arange(W), moves] + 0.3 * rng.standard_normal(W)
    print(f"rbm evaluate case={case} N={N} H={H} sorting={sorting} precision={precision} "
          f"rbm_matrix={rbm_matrix}", flush=True)
    with _engine(ecfg) as eng:
        eng.set_parameters(P)
        g0, g1, gr, gg0, gg1 = eng.evaluate_elements(pos, moves, newv)
    r0, r1, rr, rg0, rg1 = oracle_evaluate_elements(ocfg, P, pos, moves, newv)
    tol = 1e-9 if precision == 64 else 2e-3
    tol_g = 1e-9 if precision == 64 else 1e-4
    tol_r = 5e-4 if rbm_matrix == 3 else tol_g
    for tag, g, r in [("before", g0, r0), ("after", g1, r1)]:
        drift = r[:, 9:]
        scale_e = (np.abs(r[:, 1]) + np.abs(r[:, 2]) + np.abs(r[:, 3]) +
                   0.125 * np.sum(drift ** 2, axis=1) + 1.0)
        eE = np.abs(g[:, 0] - r[:, 0]) / scale_e
        eF = np.max(np.abs(g[:, 9:] - drift), axis=1) / (np.max(np.abs(drift), axis=1) + 1.0)
        print(f"  {tag}: E max {eE.max():.2e} drift max {eF.max():.2e}")
        assert eE.max() < tol and eF.max() < tol
        for c in (1, 2, 3):
            assert np.max(np.abs(g[:, c] - r[:, c]) / scale_e) < tol
    eR = np.abs(gr - rr) / np.maximum(np.abs(rr), 1.0)
    eG0, eG1 = _scaled(gg0, rg0), _scaled(gg1, rg1)
    print(f"  ratios max {eR.max():.2e}  dlnpsi/dtheta max {max(eG0.max(), eG1.max()):.2e}")
    assert eR.max() < tol_r
    assert eG0.max() < tol_g and eG1.max() < tol_g


TRAJ = [
    # N, omega, kinds, H, sorting, quirks, sampling, dt, batches, iteration
    (2, 1.0, RBM, 2, True, REF, O.IMPORTANCE_COMPAT, 0.1, 1, 0),
    (2, 1.0, RBM, 2, True, REF, O.BRUTE_FORCE, 0.5, 10, 3),
    (12, 1.0, RBMPJ, 24, True, REF, O.IMPORTANCE_COMPAT, 0.02, 1, 0),
    (12, 1.0, RBMPJ, 24, False, REF, O.IMPORTANCE_EXACT, 0.02, 10, 7),
    (6, 1.0, RBM + [O.PADE], 8, True, 0, O.IMPORTANCE_EXACT, 0.05, 1, 0),
]


@pytest.mark.parametrize("case", range(len(TRAJ)))
def test_rbm_fp64_trajectories_match_oracle(case):
    """Same Philox streams: accept decisions, positions, energy sums and the
    gradient-window sums of every parameter agree with the oracle's chains."""
    N, omega, kinds, H, sorting, quirks, sampling, dt, batches, it = TRAJ[case]
    rng = np.random.default_rng(500 + case)
    W, steps, equil, seed = 48, 200, 20, 77 + case
    nd = 2 * N
    ocfg, ecfg = _configs(N, omega, kinds, H, sorting, quirks, sampling, dt, 64, W)
    P = rbm_parameters(rng, kinds, nd, H)
    pos0 = np.stack([O.initial_positions(seed, w, nd, omega) for w in range(W)]) * np.sqrt(omega)
    print(f"rbm trajectory case={case} N={N} H={H} sampling={sampling}", flush=True)
    with _engine(ecfg) as eng:
        eng.set_parameters(P)
        eng.set_positions(pos0)
        eng.set_random_state(seed, 0)
        sums, g, ge, _ = eng.run(steps, equil, it, batches)
        gpos = eng.positions()
    opos = pos0.copy()
    counters = np.zeros(W, dtype=np.uint64)
    acc, og, oge = O.run_many(ocfg, P, opos, seed, counters, steps, equil, it, batches, threads=8)
    og = og.reshape(g.shape)
    oge = oge.reshape(ge.shape)
    print(f"  accepted {sums['accepted']:.0f} / {acc[8]:.0f}, E {sums['energy']:.10g} / {acc[0]:.10g}")
    assert sums["accepted"] == acc[8]
    assert abs(sums["energy"] - acc[0]) <= 1e-9 * abs(acc[0])
    assert abs(sums["energy2"] - acc[1]) <= 1e-9 * abs(acc[1])
    assert np.max(np.abs(gpos - opos)) < 1e-9 * max(1.0, np.max(np.abs(opos)))
    assert np.max(np.abs(g - og)) <= 1e-9 * max(1.0, np.max(np.abs(og)))
    assert np.max(np.abs(ge - oge)) <= 1e-9 * max(1.0, np.max(np.abs(oge)))
    assert np.any(g != 0)


@pytest.mark.parametrize("rbm_matrix", [1, 2, 3])
def test_rbm_fp32_walk_agrees_with_fp64(rbm_matrix):
    """RBMPJ N=12, H=24: FP32 and FP64 walks from the same walkers on the same
    Philox streams sample the same distribution, for each FP32 contraction
    mode.  Walker w of one precision is paired with walker w of the other;
    the mean of the per-walker differences of the mean local energy must be
    within 4 standard errors (from their spread) of zero, and the acceptance
    rates within 4 standard errors of their paired difference."""
    N, omega, kinds, H = 12, 1.0, RBMPJ, 24
    rng = np.random.default_rng(901)
    W, nd, steps = 512, 24, 2048
    P = rbm_parameters(rng, kinds, nd, H)
    res = {}
    for prec in (64, 32):
        _, ecfg = _configs(N, omega, kinds, H, True, REF, O.IMPORTANCE_COMPAT, 0.02, prec, W,
                           rbm_matrix if prec == 32 else 0)
        ecfg.refresh_interval = 100
        print(f"rbm fp{prec} walk rbm_matrix={rbm_matrix}", flush=True)
        with _engine(ecfg) as eng:
            eng.set_parameters(P)
            eng.init_walkers(5)
            eng.run(1000, 0)
            sums, g, ge, en = eng.run(steps, 0, 0, 1, record_walkers=W)
        res[prec] = (sums["energy"] / sums["samples"], en.astype(np.float64).mean(axis=1),
                     sums["accepted"] / sums["samples"])
    E64, m64, a64 = res[64]
    E32, m32, a32 = res[32]
    d = m32 - m64
    se = d.std() / np.sqrt(W)
    print(f"  E {E64:.6f} vs {E32:.6f}, paired difference {d.mean():.2e} +- {se:.1e}, "
          f"acceptance {a64:.5f} vs {a32:.5f}")
    assert np.all(np.isfinite(m32))
    assert abs(d.mean()) < 4.0 * se + 1e-6 * abs(E64)
    assert abs(a32 - a64) < 4.0 * np.sqrt(max(a64 * (1 - a64), 1e-4) * 2.0 / (W * steps)) + 2e-4


@pytest.mark.parametrize("sampling", [O.IMPORTANCE_COMPAT, O.IMPORTANCE_EXACT])
@pytest.mark.parametrize("rbm_matrix", [1, 2, 3])
def test_rbm_fp32_gradient_sums_retrace_fp64(rbm_matrix, sampling):
    """RBMPJ N=12, H=24: a short FP32 chain retraces the FP64 chain from the
    same walkers on the same streams (no accept decision differs, positions
    agree to FP32 rounding), so the two precisions' gradient-window sums of
    every one of the 624 RBM parameters -- b, W and the RBM-Gaussian a, the
    sums of dlnpsi/dtheta and of E_L dlnpsi/dtheta -- differ by rounding only.
    Tolerance, relative to the largest sum: 1e-4 for the FP32 lanes and FP32
    tiles; 1e-3 for BF16 tiles, whose inputs to the pre-activation
    v = b + W^T x / sigma_p^2 are rounded to 8 significant bits (2^-8
    relative per product, a few 1e-5 in v after the 1/sigma_p^2 = 1/100
    scale, so a few 1e-5 in n = 1/(1 + e^-v)) while the window sums of
    x n^T take FP32 products (V_MFMA_F32_16X16X4_F32 in both tile modes).
    Under the reference's importance sampling the FP32-tile modes run the
    instantiation specialised to Slater x RBM-Gaussian x RBM-product x
    Pade-Jastrow; under exact Langevin the generic one."""
    N, omega, kinds, H = 12, 1.0, RBMPJ, 24
    rng = np.random.default_rng(911)
    W, nd, steps = 256, 24, 24
    P = rbm_parameters(rng, kinds, nd, H)
    out = {}
    for prec in (64, 32):
        _, ecfg = _configs(N, omega, kinds, H, True, REF, sampling, 0.02, prec, W,
                           rbm_matrix if prec == 32 else 0)
        print(f"rbm gradient sums fp{prec} rbm_matrix={rbm_matrix} sampling={sampling}",
              flush=True)
        with _engine(ecfg) as eng:
            eng.set_parameters(P)
            eng.init_walkers(19)
            sums, g, ge, _ = eng.run(steps, 0, 0, 1)
            out[prec] = (sums, g, ge, eng.positions())
    s64, g64, ge64, x64 = out[64]
    s32, g32, ge32, x32 = out[32]
    dx = np.max(np.abs(x32 - x64))
    print(f"  accepted {s32['accepted']:.0f} / {s64['accepted']:.0f}, positions max |diff| {dx:.1e}")
    assert s32["accepted"] == s64["accepted"]
    assert dx < 1e-3
    assert abs(s32["energy"] - s64["energy"]) < 1e-4 * abs(s64["energy"])
    tol = 1e-3 if rbm_matrix == 3 else 1e-4
    for tag, a, b in (("dlnpsi/dtheta", g32, g64), ("E_L dlnpsi/dtheta", ge32, ge64)):
        scale = np.max(np.abs(b))
        err = np.max(np.abs(a - b)) / scale
        print(f"  {tag} sums: max |diff| / max {err:.2e}")
        assert scale > 0 and err < tol


@pytest.mark.parametrize("sampling", [O.IMPORTANCE_EXACT, O.IMPORTANCE_COMPAT])
def test_rbm_mfma_f32_retraces_lanes(sampling):
    """The FP32 matrix-core tiles add the products of each contraction in the
    order the lanes' FMA chains do (index 4e + q over the eight 16x16x4 steps
    is 0, 1, 2, ... ; the padding adds exact zeros), so the FP32 walk takes the
    same accept decisions in both modes and ends at the same positions.  The
    two modes are separate kernel instantiations, in which the compiler may
    fuse a multiply and an add differently, so energies and positions agree to
    FP32 rounding of the chains (energy sums 1e-6 relative, positions 1e-4)
    rather than bit for bit.  The gradient-window sums of dlnpsi/dW take FP32
    products and 4-walker sums on the tiles (FP64 products on the lanes)
    before the FP64 accumulation: they agree to 1e-6 relative.  Under the reference's importance sampling the
    FP32-tile walk of Slater x RBM-Gaussian x RBM-product x Pade-Jastrow runs
    the instantiation specialised to that element set."""
    N, omega, kinds, H = 12, 1.0, RBMPJ, 24
    rng = np.random.default_rng(907)
    W, nd = 256, 24
    P = rbm_parameters(rng, kinds, nd, H)
    out = {}
    for mode in (1, 2):
        _, ecfg = _configs(N, omega, kinds, H, True, REF, sampling, 0.02, 32, W, mode)
        ecfg.refresh_interval = 50
        print(f"rbm fp32 walk rbm_matrix={mode} sampling={sampling}", flush=True)
        with _engine(ecfg) as eng:
            eng.set_parameters(P)
            eng.init_walkers(11)
            sums, g, ge, _ = eng.run(300, 100, 0, 1)
            out[mode] = (np.array([sums[k] for k in ("energy", "energy2", "kinetic", "accepted")]),
                         g, ge, eng.positions())
    assert out[1][0][3] == out[2][0][3]                        # accepted moves
    np.testing.assert_allclose(out[2][0], out[1][0], rtol=1e-6)   # energy sums
    np.testing.assert_allclose(out[2][3], out[1][3], rtol=0, atol=1e-4)   # positions
    for k in (1, 2):                                      # gradient sums
        scale = np.max(np.abs(out[1][k]))
        print(f"  gradient sums max |diff| / max {np.max(np.abs(out[1][k] - out[2][k])) / scale:.2e}")
        np.testing.assert_allclose(out[2][k], out[1][k], rtol=1e-6, atol=1e-6 * scale)
