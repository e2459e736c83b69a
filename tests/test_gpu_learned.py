"""The learned Metropolis proposal on the GPU against the oracle, and what the
trained network gains over the method without it.

* training kernels (vmc_lp_gradient): per-sample J, A, log|psi'/psi|^2 and the
  proposed position, and the summed dJ/dtheta, against the oracle's
  vmco_lp_sample_grad (itself pinned by finite differences in
  test_oracle_lp.py).  The network part runs in FP32: tolerance 2e-4 relative
  per sample, 2e-3 on the gradient norm.  Samples within 1e-3 of the kink of
  min(1, e^L) are redrawn (the gradient is discontinuous there).  The proposed
  position carries the FP32 network's rounding (1e-7), log|psi'/psi|^2 inherits
  it (tolerance 2e-5).
* the learned walk in FP64 retraces the oracle's learned steps (same Philox
  draws, the same bfloat16-rounded weights) step for step.
* training raises the expected squared jump distance, and the trained
  proposal samples the same energy as exact Langevin with a shorter
  integrated autocorrelation time per unit of work.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

SGP = [O.SLATER, O.GAUSSIAN, O.PADE]


def _engine(cfg):
    import vmachine_amd as V
    if V.device_count() < 1:
        pytest.fail("no GPU visible: GPU tests must run on the MI355X box")
    return V.Engine(cfg)


def _bf16(x):
    """Round float32 values to the nearest-even bfloat16, returned as float32."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def _theta(seed, scale=0.5):
    rng = np.random.default_rng(seed)
    return (O.lp_init(rng, zero_head=False) * scale).astype(np.float32)


@pytest.mark.parametrize("N,omega", [(6, 1.0), (12, 0.5), (56, 0.1)])
def test_lp_gradient_matches_oracle(N, omega):
    import vmachine_amd as V
    rng = np.random.default_rng(40 + N)
    W, nd = 48, 2 * N
    th = _theta(7 + N)
    hp = O.lp_hyper()
    ocfg = O.make_config(N, 2, omega, 1, SGP, None, None, False, O.IMPORTANCE_EXACT, 0.01, quirks=0)
    P = np.array([[0.0], [1.0], [0.7]])
    pos = rng.standard_normal((W, nd)) / np.sqrt(omega)
    el = rng.integers(0, N, W)
    nrm = rng.standard_normal((W, 2))
    g_or = np.zeros(O.LP_NPARAM)
    per_or = np.zeros((W, 5))
    for b in range(W):
        w = O.Walker(ocfg, P, pos[b])
        while True:
            out = w.lp_sample_grad(th.astype(np.float64), hp, el[b], nrm[b, 0], nrm[b, 1],
                                   np.zeros(O.LP_NPARAM))
            if abs(w.last_L) > 1e-3:
                break
            nrm[b] = rng.standard_normal(2)
        g_or += out[5]
        per_or[b] = out[:5]
    cfg = V.make_config(N, omega=omega, kinds=SGP, sampling=V.LEARNED, step_length=0.01,
                        quirks=0, precision=64, num_walkers=W)
    print(f"lp gradient N={N}", flush=True)
    with _engine(cfg) as eng:
        eng.set_parameters(P)
        eng.lp_set(th)
        per, g = eng.lp_gradient(pos, el, nrm)
    errJ = np.abs(per[:, 0] - per_or[:, 0]) / (np.abs(per_or[:, 0]) + 1e-3)
    errA = np.abs(per[:, 1] - per_or[:, 1])
    errR = np.abs(per[:, 2] - per_or[:, 2]) / (np.abs(per_or[:, 2]) + 1.0)
    errX = np.max(np.abs(per[:, 3:5] - per_or[:, 3:5]), axis=1) / (1.0 + np.abs(per_or[:, 3:5]).max(axis=1))
    errG = np.linalg.norm(g - g_or) / np.linalg.norm(g_or)
    print(f"  J {errJ.max():.2e} A {errA.max():.2e} logR {errR.max():.2e} x' {errX.max():.2e} "
          f"grad {errG:.2e} (|grad| {np.linalg.norm(g_or):.3e})")
    # x' carries the FP32 network's rounding (1e-7 relative), and log R inherits it
    assert errJ.max() < 2e-4 and errA.max() < 2e-4 and errR.max() < 2e-5 and errX.max() < 1e-6
    assert errG < 2e-3


@pytest.mark.parametrize("N,omega", [(6, 1.0), (12, 0.5), (30, 0.5), (56, 0.1)])
def test_learned_walk_fp64_retraces_oracle(N, omega):
    import vmachine_amd as V
    W, steps, seed = 24, 60, 123 + N
    nd = 2 * N
    th = _bf16(_theta(3 + N))
    hp = O.lp_hyper()
    ocfg = O.make_config(N, 2, omega, 1, SGP, None, None, False, O.IMPORTANCE_EXACT, 0.01, quirks=0)
    P = np.array([[0.0], [1.0], [0.7]])
    pos0 = np.stack([O.initial_positions(seed, w, nd, omega) for w in range(W)]) * np.sqrt(omega)
    cfg = V.make_config(N, omega=omega, kinds=SGP, sampling=V.LEARNED, step_length=0.01,
                        quirks=0, precision=64, num_walkers=W)
    print(f"learned walk N={N}", flush=True)
    with _engine(cfg) as eng:
        eng.set_parameters(P)
        eng.lp_set(th)
        eng.set_positions(pos0)
        eng.set_random_state(seed, 0)
        sums, _, _, _ = eng.run(steps, 0)
        gpos = eng.positions()
    acc = 0
    esum = 0.0
    opos = np.zeros_like(pos0)
    for w in range(W):
        wk = O.Walker(ocfg, P, pos0[w])
        for t in range(steps):
            c = O.philox([t & 0xFFFFFFFF, t >> 32, w, 0], [seed & 0xFFFFFFFF, seed >> 32])
            i = ((int(c[0]) * nd) >> 32) >> 1
            U1 = ((int(c[1]) >> 9) + 0.5) / 8388608.0
            U2 = ((int(c[2]) >> 9) + 0.5) / 8388608.0
            ua = ((int(c[3]) >> 9) + 0.5) / 8388608.0
            r = np.sqrt(-2.0 * np.log(U1))
            acc += wk.lp_step(th.astype(np.float64), hp, i, r * np.cos(2 * np.pi * U2),
                              r * np.sin(2 * np.pi * U2), ua)
            esum += wk.local_energy()
        opos[w] = wk.positions()
    print(f"  accepted {sums['accepted']:.0f} / {acc}, E {sums['energy']:.10g} / {esum:.10g}")
    assert sums["accepted"] == acc
    assert np.max(np.abs(gpos - opos)) < 1e-9 * max(1.0, np.max(np.abs(opos)))
    assert abs(sums["energy"] - esum) <= 1e-9 * abs(esum)


def _series_stats(en):
    from vmachine_amd import stats
    tau = stats.iat_many(en.astype(np.float64), "blocking")
    m = en.mean(axis=1)
    return float(m.mean()), float(m.std() / np.sqrt(len(m))), tau


def _equilibrated(N, omega, P, W, seed, th, steps):
    """Walkers equilibrated by the learned proposal (it mixes in a few thousand
    steps where Langevin at small dt needs 10^5 or more from the initial
    N(0, 1/omega^2) positions); returned as positions for any method."""
    import vmachine_amd as V
    cfg = V.make_config(N, omega=omega, kinds=SGP, sampling=V.LEARNED, step_length=0.01, quirks=0,
                        precision=32, refresh_interval=100, num_walkers=W)
    with _engine(cfg) as eng:
        eng.set_parameters(P)
        eng.lp_set(th)
        eng.init_walkers(seed)
        for _ in range(steps // 1000):
            eng.run(1000, 0)
        return eng.positions()


def test_training_gain_over_untrained_and_langevin():
    """N=56, omega=0.1 (the headline system): from equilibrated walkers, train
    the network online, then record E_L with the trained proposal, the
    untrained one (a symmetric random walk of width sigma0) and exact Langevin
    at dt=0.005 (its series 4x longer: tau ~ 1200 steps).  The three energies
    agree within 4 blocking errors; training raises the jump distance; the
    trained proposal has the lowest autocorrelation time x kernel time per step."""
    import vmachine_amd as V
    N, omega, W = 56, 0.1, 8192
    P = np.array([[0.0], [1.0], [1.0]])
    rec_w, rec_steps = 256, 8192
    th0 = O.lp_init(np.random.default_rng(77), zero_head=True).astype(np.float32)
    print("learned gain: equilibrating", flush=True)
    pos = _equilibrated(N, omega, P, W, 5, th0, 40000)
    res = {}
    for label in ("untrained", "trained", "langevin"):
        samp = V.IMPORTANCE_EXACT if label == "langevin" else V.LEARNED
        cfg = V.make_config(N, omega=omega, kinds=SGP, sampling=samp, step_length=0.005, quirks=0,
                            precision=32, refresh_interval=100, num_walkers=W, drift_limit=1.0)
        print(f"learned gain: {label}", flush=True)
        with _engine(cfg) as eng:
            eng.set_parameters(P)
            if samp == V.LEARNED:
                eng.lp_set(th0)
            eng.set_positions(pos)
            eng.set_random_state(11, 0)
            if label == "trained":
                hist = eng.lp_train(150, 10, 3e-3)
                print(f"  ESJD {hist[:5, 0].mean():.4f} -> {hist[-5:, 0].mean():.4f}, "
                      f"A {hist[:5, 1].mean():.3f} -> {hist[-5:, 1].mean():.3f}", flush=True)
                res["hist"] = hist
            eng.run(1000, 0)
            # Langevin's tau (~1200 steps) needs a series many times longer
            steps = 4 * rec_steps if label == "langevin" else rec_steps
            s, _, _, en = eng.run(steps, 0, record_walkers=rec_w)
            E, Ee, tau = _series_stats(en)
            ns = s["kernel_ms"] * 1e6 / (W * steps)
            res[label] = (E, Ee, tau, ns, s["accepted"] / s["samples"])
            print(f"  E {E:.4f} +- {Ee:.4f}, tau {tau:.1f} steps, {ns:.3f} ns/walker-step, "
                  f"acceptance {res[label][4]:.3f}", flush=True)
    hist = res["hist"]
    assert hist[-10:, 0].mean() > 3.0 * hist[:10, 0].mean()
    El, el_err = res["langevin"][0], res["langevin"][1]
    for label in ("untrained", "trained"):
        E, Ee = res[label][0], res[label][1]
        assert abs(E - El) < 4 * np.hypot(Ee, el_err), (label, E, El)
    cost = {k: res[k][2] * res[k][3] for k in ("untrained", "trained", "langevin")}
    print("  tau x ns per step: " + ", ".join(f"{k} {v:.1f}" for k, v in cost.items()))
    assert cost["trained"] < cost["untrained"]
    assert cost["trained"] < cost["langevin"]


@pytest.mark.parametrize("N,omega", [(30, 0.5), (56, 0.1)])
def test_fp32_learned_walk_samples_psi_squared(N, omega):
    """Whole-electron moves at large N: from the same equilibrated walkers the
    FP32 learned walk and FP32 exact Langevin give the same energy within 4
    blocking errors."""
    import vmachine_amd as V
    W = 8192
    P = np.array([[0.0], [1.0], [0.8]])
    th = _theta(99, scale=0.3)
    print(f"fp32 learned vs Langevin N={N}: equilibrating", flush=True)
    pos = _equilibrated(N, omega, P, W, 9, th, 20000)
    res = {}
    for label in ("learned", "langevin"):
        samp = V.LEARNED if label == "learned" else V.IMPORTANCE_EXACT
        cfg = V.make_config(N, omega=omega, kinds=SGP, sampling=samp, step_length=0.05, quirks=0,
                            precision=32, refresh_interval=100, num_walkers=W, drift_limit=1.0)
        with _engine(cfg) as eng:
            eng.set_parameters(P)
            if samp == V.LEARNED:
                eng.lp_set(th)
            eng.set_positions(pos)
            eng.run(2000, 0)
            _, _, _, en = eng.run(8192, 0, record_walkers=512)
        res[label] = _series_stats(en)
        print(f"  {label}: E {res[label][0]:.4f} +- {res[label][1]:.4f} tau {res[label][2]:.1f}")
    (El, el, _), (Eg, eg, _) = res["learned"], res["langevin"]
    assert abs(El - Eg) < 4 * np.hypot(el, eg)


def test_learned_proposal_limited_to_56_particles():
    """The learned proposal's kernels exist for the closed shells up to N = 56;
    the engine refuses it past them at creation."""
    import vmachine_amd as V
    cfg = V.make_config(72, omega=0.1, kinds=SGP, sampling=V.LEARNED, step_length=0.01, quirks=0,
                        precision=32, num_walkers=8)
    with pytest.raises(V.VmcError, match="learned proposal"):
        V.Engine(cfg)
