"""The batch form: quantum dots of different electron counts walked side by
side (one kernel per system on streams of their own), an optimisation
iteration per hipGraph replay, the optimiser on the device (vmc_batch_* in
include/vmc.h).

Pinned three ways:
 * every system of a batch retraces the FP64 oracle's chains (Philox stream
   seed + s for system s) step for step, as a single-system engine does;
 * per system, the batch's sums and final positions equal those of a
   single-system engine seeded seed + s, in FP64 and FP32, also after the
   kernel shape (lanes per walker) is changed;
 * graph replay and direct launches give bit-identical results, and the
   device ADAM step reproduces the host optimiser (ref
   src/Optimization/adam.cpp:12-38) iteration by iteration.
Tolerances: FP64 1e-9 relative (summation order differs between the batch's
per-system reduction and the engine's), FP32 as stated in each test.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import OMEGA_FOR_N, engine_config, oracle_config

pytestmark = pytest.mark.gpu

KINDS_SGP = [O.SLATER, O.GAUSSIAN, O.PADE]
# (N, walkers, dt): deliberately not in cost order, ragged walker counts
SYSTEMS = [(12, 37, 0.02), (2, 100, 0.1), (6, 64, 0.05), (20, 45, 0.01)]
# the three largest closed shells, few walkers each (the oracle is the slow side)
SYSTEMS_LARGE = [(42, 19, 0.005), (30, 24, 0.01), (56, 13, 0.005)]
LANES = {1: 1, 3: 4, 6: 8, 10: 16, 15: 16, 21: 32, 28: 32}   # orbitals per spin -> default K
ORBITALS = {2: 1, 6: 3, 12: 6, 20: 10, 30: 15, 42: 21, 56: 28}


def _require_gpu():
    import vmachine_amd as V
    if V.device_count() < 1:
        pytest.fail("no GPU visible: GPU tests must run on the MI355X box")
    return V


def _systems(rng, systems=SYSTEMS):
    out = []
    for N, W, dt in systems:
        out.append(dict(num_particles=N, omega=OMEGA_FOR_N[N], num_walkers=W, step_length=dt,
                        alpha=1.0 + 0.1 * rng.standard_normal(), gamma=0.3 + 0.7 * rng.random()))
    return out


def _batch(V, systems, sampling, precision=64, quirks=0b1100, drift=0.0, refresh=0):
    return V.Batch(systems, kinds=KINDS_SGP, sampling=sampling, precision=precision, quirks=quirks,
                   drift_limit=drift, refresh_interval=refresh)


def _params(y):
    return np.array([[0.0], [y["alpha"]], [y["gamma"]]])


def _engine(V, y, sampling, precision=64, drift=0.0, refresh=0):
    cfg = engine_config(y["num_particles"], y["omega"], KINDS_SGP, sampling, y["step_length"],
                        precision=precision, walkers=y["num_walkers"], refresh=refresh,
                        drift_limit=drift)
    e = V.Engine(cfg)
    e.set_parameters(_params(y))
    return e


SUM_KEYS = ["energy", "energy2", "kinetic", "external", "interaction", "kinetic2", "external2",
            "interaction2", "accepted"]


def _check_sums(out_s, sums, g, ge, rtol):
    for k, key in enumerate(SUM_KEYS):
        assert abs(out_s[k] - sums[key]) <= rtol * max(1.0, abs(sums[key])), (key, out_s[k], sums[key])
    for k, ref in [(9, g[1, 0]), (10, g[2, 0]), (11, ge[1, 0]), (12, ge[2, 0])]:
        assert abs(out_s[k] - ref) <= rtol * max(1.0, abs(ref)), (k, out_s[k], ref)
    assert out_s[13] == sums["nonfinite"]


@pytest.mark.parametrize("large", [False, True])
@pytest.mark.parametrize("sampling,drift", [(O.IMPORTANCE_COMPAT, 0.0),
                                            (O.IMPORTANCE_EXACT, 1.0),
                                            (O.BRUTE_FORCE, 0.0)])
def test_batch_retraces_the_oracle(sampling, drift, large):
    V = _require_gpu()
    rng = np.random.default_rng(41 + sampling)
    systems = _systems(rng, SYSTEMS_LARGE if large else SYSTEMS)
    if sampling == O.BRUTE_FORCE:
        for y in systems:
            y["step_length"] = 0.5 / np.sqrt(y["omega"])
    steps, equil, seed = 120, 10, 500 + sampling
    print(f"batch vs oracle sampling={sampling} N={[y['num_particles'] for y in systems]}",
          flush=True)
    pos0 = []
    with _batch(V, systems, sampling, drift=drift) as b:
        b.init_walkers(seed)   # the random key: system s draws from seed + s
        for s, y in enumerate(systems):
            N, W = y["num_particles"], y["num_walkers"]
            p = np.stack([O.initial_positions(seed + s, w, 2 * N, y["omega"]) for w in range(W)])
            p = p * np.sqrt(y["omega"])   # a gently spread start, as in the single-engine test
            pos0.append(p)
            b.set_positions(s, p)
        out, _ = b.run(1, steps, equil)
        gpos = [b.get_positions(s) for s in range(len(systems))]
    for s, y in enumerate(systems):
        N, W = y["num_particles"], y["num_walkers"]
        ocfg = oracle_config(N, y["omega"], KINDS_SGP, sampling, y["step_length"],
                             drift_limit=drift)
        opos = pos0[s].copy()
        counters = np.zeros(W, dtype=np.uint64)
        acc, og, oge = O.run_many(ocfg, _params(y), opos, seed + s, counters, steps, equil, 0, 1,
                                  threads=8)
        o = out[0, s]
        assert o[8] == acc[8], (s, o[8], acc[8])
        for k in range(8):
            assert abs(o[k] - acc[k]) <= 1e-9 * max(1.0, abs(acc[k])), (s, k, o[k], acc[k])
        for k, ref in [(9, og[1, 0]), (10, og[2, 0]), (11, oge[1, 0]), (12, oge[2, 0])]:
            assert abs(o[k] - ref) <= 1e-9 * max(1.0, abs(ref)), (s, k, o[k], ref)
        assert np.max(np.abs(gpos[s] - opos)) < 1e-9 * max(1.0, np.max(np.abs(opos))), s


@pytest.mark.parametrize("precision", [64, 32])
def test_batch_matches_single_engines(precision):
    """Two iterations (gradient window of iteration 1 of 2 batches in the
    second) with a refresh every 50 steps; FP32 sums within 1e-5 relative."""
    V = _require_gpu()
    rng = np.random.default_rng(77)
    systems = _systems(rng)
    seed, steps, equil = 1234, 200, 20
    sampling = O.IMPORTANCE_EXACT
    print(f"batch vs engines precision={precision}", flush=True)
    with _batch(V, systems, sampling, precision=precision, refresh=50) as b:
        b.init_walkers(seed)
        out0, _ = b.run(1, steps, equil, batches=2, use_graph=False)
        out1, _ = b.run(1, steps, equil, batches=2, use_graph=True)
        bpos = [b.get_positions(s) for s in range(len(systems))]
    rtol = 1e-9 if precision == 64 else 1e-5
    for s, y in enumerate(systems):
        with _engine(V, y, sampling, precision=precision, refresh=50) as e:
            e.init_walkers(seed + s)
            r0 = e.run(steps, equil, 0, 2)
            r1 = e.run(steps, equil, 1, 2)
            epos = e.positions()
        _check_sums(out0[0, s], *r0[:3], rtol)
        _check_sums(out1[0, s], *r1[:3], rtol)
        if precision == 64:
            assert np.max(np.abs(bpos[s] - epos)) < 1e-9 * max(1.0, np.max(np.abs(epos)))
        else:
            # same code and draws: FP32 chains agree to rounding unless an
            # acceptance test sits on its threshold
            close = np.all(np.abs(bpos[s] - epos) < 1e-4, axis=1)
            assert close.mean() >= 0.95, (s, close.mean())


def test_graph_replay_equals_direct_launches():
    V = _require_gpu()
    rng = np.random.default_rng(5)
    systems = _systems(rng)
    res = []
    for use_graph in (True, False):
        with _batch(V, systems, O.IMPORTANCE_EXACT, precision=32, refresh=40) as b:
            b.init_walkers(99)
            out, ms = b.run(4, 100, 10, batches=2, learning_rate=0.05, use_graph=use_graph)
            assert ms > 0
            res.append((out, [b.get_positions(s) for s in range(len(systems))], b.parameters()))
    np.testing.assert_array_equal(res[0][0], res[1][0])
    for a, c in zip(res[0][1], res[1][1]):
        np.testing.assert_array_equal(a, c)
    np.testing.assert_array_equal(res[0][2], res[1][2])


# all seven closed shells, ragged walker counts that leave partial tiles
SYSTEMS_SEVEN = [(2, 83, 0.1), (6, 70, 0.05), (12, 37, 0.02), (20, 45, 0.01), (30, 29, 0.01),
                 (42, 19, 0.005), (56, 13, 0.005)]


@pytest.mark.parametrize("refresh", [40, 100])
def test_persistent_launch_equals_side_by_side(refresh):
    """One persistent launch per iteration (waves taking tiles of one system
    from a device counter, every refresh segment of a tile in one go) walks
    the same chains as the side-by-side kernels: the same walk code and draws
    per walker, compiled as a called function instead of a kernel body, so the
    FP32 results agree to rounding (sums 1e-5 relative, parameters after ADAM
    1e-6, >= 95 % of the walkers' positions within 1e-4, as in the FP32
    comparisons with single-system engines), and the persistent launch as a
    graph equals its direct launches bit for bit.  Configurations the
    persistent kernel does not carry are refused."""
    V = _require_gpu()
    rng = np.random.default_rng(21)
    systems = _systems(rng, SYSTEMS_SEVEN)
    res = []
    for launch, use_graph in ((V.BATCH_SIDE_BY_SIDE, False), (V.BATCH_PERSISTENT, False),
                              (V.BATCH_PERSISTENT, True)):
        print(f"persistent test: refresh={refresh} launch={launch} graph={use_graph}", flush=True)
        with _batch(V, systems, O.IMPORTANCE_EXACT, precision=32, quirks=0, drift=1.0,
                    refresh=refresh) as b:
            b.set_launch(launch)
            assert b.launch() == launch
            b.init_walkers(606)
            out, ms = b.run(3, 120, 10, batches=2, learning_rate=0.02, use_graph=use_graph)
            assert ms > 0
            res.append((out, [b.get_positions(s) for s in range(len(systems))], b.parameters()))
    assert np.all(np.isfinite(res[0][0])) and np.all(res[0][0][:, :, 13] == 0)
    side, pers, pers_graph = res
    np.testing.assert_array_equal(pers_graph[0], pers[0])
    for a, c in zip(pers_graph[1], pers[1]):
        np.testing.assert_array_equal(a, c)
    np.testing.assert_array_equal(pers_graph[2], pers[2])
    # sums of E, T, V_ext, V_int cancel between walkers far from equilibrium
    # (T < 0 near a node): their scale is sqrt(M * sum of squares), not the sum
    squares = {0: 1, 2: 5, 3: 6, 4: 7}
    for it in range(side[0].shape[0]):
        for s in range(len(systems)):
            o, r = pers[0][it, s], side[0][it, s]
            M = systems[s]["num_walkers"] * 130
            for k in range(13):
                scale = np.sqrt(M * r[squares[k]]) if k in squares else abs(r[k])
                assert abs(o[k] - r[k]) <= 1e-5 * max(1.0, scale), (it, s, k, o[k], r[k])
            assert o[13] == r[13] == 0
            np.testing.assert_allclose(o[14:16], r[14:16], rtol=1e-6)
    for s in range(len(systems)):
        close = np.all(np.abs(pers[1][s] - side[1][s]) < 1e-4, axis=1)
        assert close.mean() >= 0.95, (s, close.mean())
    with _batch(V, systems[:2], O.IMPORTANCE_EXACT, precision=64) as b:
        with pytest.raises(V.VmcError, match="persistent"):
            b.set_launch(V.BATCH_PERSISTENT)
        assert b.launch() == V.BATCH_SIDE_BY_SIDE


def test_device_adam_matches_host_optimiser():
    """Three iterations of lr 0.05: (alpha, gamma) after every device ADAM step
    against the host optimiser driven by single-engine sums."""
    V = _require_gpu()
    from vmachine_amd import host as H
    rng = np.random.default_rng(9)
    systems = _systems(rng)
    seed, steps, iters, lr, B = 4321, 200, 3, 0.05, 2
    with _batch(V, systems, O.IMPORTANCE_EXACT) as b:
        b.init_walkers(seed)
        out, _ = b.run(iters, steps, 0, batches=B, learning_rate=lr)
        final = b.parameters()
    for s, y in enumerate(systems):
        P = _params(y)
        opt = H.Optimizer("adam", P.size, lr)
        with _engine(V, y, O.IMPORTANCE_EXACT) as e:
            e.init_walkers(seed + s)
            for it in range(iters):
                sums, g, ge, _ = e.run(steps, 0, it, B)
                assert abs(out[it, s, 0] - sums["energy"]) <= 1e-9 * abs(sums["energy"])
                avgE = sums["energy"] / sums["samples"]
                upd = opt.update(avgE, g / sums["gradient_samples"], ge / sums["gradient_samples"])
                P = P - upd.reshape(P.shape)
                e.set_parameters(P)
                np.testing.assert_allclose(out[it, s, 14:16], [P[1, 0], P[2, 0]], rtol=1e-9,
                                           atol=1e-12)
        np.testing.assert_allclose(final[s], [P[1, 0], P[2, 0]], rtol=1e-9, atol=1e-12)


def test_shapes_and_autotune():
    """A changed shape (2K lanes per walker) walks the same chains; the
    autotuner picks valid shapes and leaves positions untouched."""
    V = _require_gpu()
    rng = np.random.default_rng(3)
    systems = _systems(rng)
    seed, steps = 77, 150
    with _batch(V, systems, O.IMPORTANCE_EXACT) as b:
        b.init_walkers(seed)
        ref, _ = b.run(1, steps)
        ref_pos = [b.get_positions(s) for s in range(len(systems))]
    with _batch(V, systems, O.IMPORTANCE_EXACT) as b:
        for s, y in enumerate(systems):
            K = LANES[ORBITALS[y["num_particles"]]]
            b.set_shape(s, 2 * K)
        b.init_walkers(seed)
        out, _ = b.run(1, steps)
        for s in range(len(systems)):
            np.testing.assert_allclose(out[0, s, :13], ref[0, s, :13], rtol=1e-9, atol=1e-9)
            assert np.max(np.abs(b.get_positions(s) - ref_pos[s])) < 1e-9
    with _batch(V, systems, O.IMPORTANCE_EXACT, precision=32) as b:
        b.init_walkers(seed)
        before = [b.get_positions(s) for s in range(len(systems))]
        lanes, ns = b.autotune(30)
        print("autotune", lanes, ns, flush=True)
        for s, y in enumerate(systems):
            K = LANES[ORBITALS[y["num_particles"]]]
            assert lanes[s] in (K, 2 * K) and ns[s] > 0
            np.testing.assert_array_equal(b.get_positions(s), before[s])
        assert b.launch() in (V.BATCH_SIDE_BY_SIDE, V.BATCH_PERSISTENT)
        out, _ = b.run(2, 100, 10)
        assert np.all(np.isfinite(out))
        for s, y in enumerate(systems):
            assert out[1, s, 8] > 0.5 * y["num_walkers"] * 100   # acceptance


@pytest.mark.parametrize("precision", [64, 32])
def test_uneven_dots_zero_variance(precision):
    """The seven closed shells N = 2..56 in one launch, non-interacting,
    Slater x Gaussian with alpha = 1: every local energy is the exact
    E0 = 2 sum (s+1)^2 omega, so every system's mean and variance are fixed."""
    V = _require_gpu()
    from tests.helpers import zero_variance_energy
    Ns = [2, 6, 12, 20, 30, 42, 56]
    omegas = [1.0, 0.5, 1.0, 0.28, 0.5, 1.0, 0.1]
    systems = [dict(num_particles=N, omega=w, num_walkers=96 + 7 * i, step_length=0.01,
                    alpha=1.0, gamma=0.0) for i, (N, w) in enumerate(zip(Ns, omegas))]
    with V.Batch(systems, kinds=[O.SLATER, O.GAUSSIAN], sampling=O.IMPORTANCE_EXACT,
                 precision=precision, interaction=False, quirks=0b1100, refresh_interval=100) as b:
        b.init_walkers(11)
        out, _ = b.run(2, 200, 50)
    tol = 1e-10 if precision == 64 else 1e-5
    for s, y in enumerate(systems):
        E0 = zero_variance_energy(y["num_particles"], y["omega"])
        M = y["num_walkers"] * 200
        E, E2 = out[1, s, 0] / M, out[1, s, 1] / M
        assert abs(E - E0) <= tol * E0, (y["num_particles"], E, E0)
        assert abs(E2 - E0 * E0) <= 4 * tol * E0 * E0
        assert out[1, s, 8] > 0.5 * M


# The mixed-size batch configuration: seven closed-shell dots, walkers in
# proportion to N^2 (131072 in total), FP32 walk with an FP64 inverse rebuild
# every 100 steps, exact Langevin dt 0.005 with drift limit 1, autotuned
# shapes, one optimisation iteration (device ADAM) per hipGraph replay.
C4_DOTS = [(2, 1.0), (6, 1.0), (12, 1.0), (20, 0.5), (30, 0.5), (42, 0.28), (56, 0.1)]
C4_TOTAL, C4_DT, C4_STEPS = 131072, 0.005, 200


def _c4_systems():
    w = np.array([N * N for N, _ in C4_DOTS], dtype=np.float64)
    Ws = [max(1, int(round(C4_TOTAL * x / w.sum()))) for x in w]
    return [dict(num_particles=N, omega=om, num_walkers=W, step_length=C4_DT, alpha=1.0,
                 gamma=1.0) for (N, om), W in zip(C4_DOTS, Ws)]


def _c4_batch(V, systems, precision):
    return V.Batch(systems, kinds=KINDS_SGP, sampling=O.IMPORTANCE_EXACT, precision=precision,
                   quirks=0, drift_limit=1.0, refresh_interval=100)


def test_c4_mixed_batch_at_full_size():
    """The C4 batch at full size.  (a) With the default shapes every system's
    sums equal those of a single-system engine seeded seed + s (FP32: 1e-5
    relative, >= 95 % of the walkers' positions within 1e-4).  (b) The FP64
    batch from the same walkers and streams gives per-system energies within
    4 standard errors of the FP32 batch's (the error of a mean over W walkers
    is at most sigma / sqrt(W) whatever the autocorrelation within each
    walker's 200 steps).  (c) After the autotuner (shapes and launch form),
    optimisation iterations with ADAM replayed as a hipGraph equal direct
    launches with the same shapes and launch form bit for bit.  No
    sampled local energy is non-finite."""
    V = _require_gpu()
    systems = _c4_systems()
    Ws = [y["num_walkers"] for y in systems]
    assert abs(sum(Ws) - C4_TOTAL) <= len(Ws)
    seed = 8080
    print(f"C4 batch: walkers {Ws}", flush=True)
    res = {}
    for prec in (32, 64):
        with _c4_batch(V, systems, prec) as b:
            b.init_walkers(seed)
            out, ms = b.run(2, C4_STEPS, 0, use_graph=False)
            res[prec] = (out, [b.get_positions(s) for s in range(len(systems))])
        print(f"  precision {prec}: {ms:.1f} ms for 2 iterations", flush=True)
    out32, pos32 = res[32]
    out64, _ = res[64]
    assert np.all(np.isfinite(out32)) and np.all(np.isfinite(out64))
    for s, y in enumerate(systems):
        N, W = y["num_particles"], y["num_walkers"]
        M = W * C4_STEPS
        assert out32[1, s, 13] == 0 and out64[1, s, 13] == 0
        # (a) single-system engine, same seed + s
        cfg = engine_config(N, y["omega"], KINDS_SGP, O.IMPORTANCE_EXACT, C4_DT, quirks=0,
                            precision=32, walkers=W, refresh=100, drift_limit=1.0)
        with V.Engine(cfg) as e:
            e.set_parameters(_params(y))
            e.init_walkers(seed + s)
            r0 = e.run(C4_STEPS, 0, 0, 1)
            r1 = e.run(C4_STEPS, 0, 1, 1)
            epos = e.positions()
        _check_sums(out32[0, s], *r0[:3], 1e-5)
        _check_sums(out32[1, s], *r1[:3], 1e-5)
        close = np.all(np.abs(pos32[s] - epos) < 1e-4, axis=1)
        assert close.mean() >= 0.95, (N, close.mean())
        # (b) FP32 against FP64
        E32, E64 = out32[1, s, 0] / M, out64[1, s, 0] / M
        v32 = out32[1, s, 1] / M - E32 * E32
        v64 = out64[1, s, 1] / M - E64 * E64
        se = np.sqrt((v32 + v64) / W)
        print(f"  N={N} W={W}: E32 {E32:.6f} E64 {E64:.6f} |diff| {abs(E32 - E64):.2e} "
              f"(4 se {4 * se:.2e}), acceptance {out32[1, s, 8] / M:.4f}", flush=True)
        assert abs(E32 - E64) <= 4.0 * se + 1e-9 * abs(E64)
        assert out32[1, s, 8] > 0.9 * M
    # (c) autotuned shapes, ADAM, graph against direct launches
    runs = []
    lanes = None
    for use_graph in (True, False):
        with _c4_batch(V, systems, 32) as b:
            b.init_walkers(seed)
            if lanes is None:
                lanes, ns = b.autotune(30)
                launch = b.launch()
                print(f"  autotune lanes {lanes.tolist()} launch {launch}", flush=True)
            else:
                for s in range(len(systems)):
                    b.set_shape(s, int(lanes[s]))
                b.set_launch(launch)
            out, ms = b.run(3, C4_STEPS, 0, learning_rate=0.01, use_graph=use_graph)
            runs.append((out, [b.get_positions(s) for s in range(len(systems))], b.parameters()))
            print(f"  graph={use_graph}: {ms:.1f} ms for 3 iterations", flush=True)
    np.testing.assert_array_equal(runs[0][0], runs[1][0])
    for a, c in zip(runs[0][1], runs[1][1]):
        np.testing.assert_array_equal(a, c)
    np.testing.assert_array_equal(runs[0][2], runs[1][2])
    assert np.all(np.isfinite(runs[0][0])) and np.all(runs[0][0][:, :, 13] == 0)
