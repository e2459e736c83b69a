"""Throughput benchmark of the VMC sampling hot path on one MI355X.

Workload (the headline configuration): 2-D quantum dot, N=56 electrons,
omega=0.1, Coulomb, Slater(Hermite, 28x28 per spin) x Gaussian x Pade-Jastrow
(alpha = gamma = 1, the reference's Automatize -> Constant(1.0) start),
Langevin importance sampling with the exact Metropolis-Hastings ratio,
dt=0.005, 32768 walkers.  A "step" of this benchmark is one sampling iteration:
every walker runs `--steps-per-iter` single-coordinate Metropolis steps, each
followed by the full local energy and dlnpsi/dtheta sample (the reference's
per-step semantics, ref src/system.cpp:147-164, src/sampler.cpp:52-90).

value = walker-steps/s of the FP32 walk (FP64 Slater-inverse rebuild every
`--refresh` steps, FP64 accumulation).  Beside it, all from the same
equilibrated walkers: the plain FP64 walk, the learned whole-electron proposal
(trained online on the engine's walkers), the integrated autocorrelation time
of E_L of each and the IAT-corrected samples/s, and the CPU baseline.

    python bench.py [--gpus 1] [--steps 10] [--warmup 2]
"""
import argparse
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# algorithmic cost of one walker-step at N=56 (single-coordinate move,
# acceptance a ~ 1, n = 28 per spin block, D = 2):
#   flops  F = 2n + a*4n^2 + 4D n^2 + 16 n(1+2D) + 40(N-1) + 6ND
#   bytes  B = b [(1+a) n^2 + 2D n^2 + (2D+1) n + 3ND]   (walker state a streaming
#              implementation would touch; b = 4 (FP32) or 8 (FP64))
def flops_per_step(N, n, D=2, a=1.0):
    return 2 * n + a * 4 * n * n + 4 * D * n * n + 16 * n * (1 + 2 * D) + 40 * (N - 1) + 6 * N * D


def bytes_per_step(N, n, b, D=2, a=1.0):
    return b * ((1 + a) * n * n + 2 * D * n * n + (2 * D + 1) * n + 3 * N * D)


# HBM bytes the FP32 N=56 walk kernel moves per walker per launch (positions in
# and out, accumulators), from rocprofv3 FETCH_SIZE / WRITE_SIZE passes scaled by
# a 1 GiB copy (profiles/walk_kernel_f32_N56_w0.1_hbm_traffic.md)
PMC_BYTES_PER_WALKER_LAUNCH = 2090.0

PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


def make_engine(V, args, precision, walkers):
    cfg = V.make_config(args.N, omega=args.omega, kinds=[V.SLATER, V.GAUSSIAN, V.PADE],
                        sampling=V.IMPORTANCE_EXACT, step_length=args.dt, quirks=0,
                        drift_limit=args.drift_limit,
                        precision=precision,
                        refresh_interval=args.refresh if precision == 32 else 0,
                        num_walkers=walkers)
    eng = V.Engine(cfg)
    eng.set_parameters(np.array([[1.0], [1.0], [1.0]]))
    eng.init_walkers(args.seed)
    return eng


def burn_in(eng, steps, label):
    done = 0
    t0 = time.time()
    while done < steps:
        chunk = min(steps - done, 20000)
        s, _, _, _ = eng.run(chunk, 0)
        done += chunk
        print(f"# {label}: burn-in {done}/{steps} steps/walker, <E> = "
              f"{s['energy'] / s['samples']:.4f}, acc = {s['accepted'] / s['samples']:.4f} "
              f"({time.time() - t0:.1f} s)", file=sys.stderr, flush=True)


def timed(eng, iters, warmup, spi):
    for _ in range(warmup):
        eng.run(spi, 0)
    kern = 0.0
    E = E2 = n = acc = 0.0
    t0 = time.perf_counter()
    for it in range(iters):
        s, _, _, _ = eng.run(spi, 0, it, 1)
        kern += s["kernel_ms"]
        E += s["energy"]
        E2 += s["energy2"]
        n += s["samples"]
        acc += s["accepted"]
    wall = time.perf_counter() - t0
    return wall, kern / 1000.0, E / n, acc / n


def measure_iat(eng, steps, block, chunk=1 << 17):
    """tau_int of E_L of the engine's walkers (all of them recorded) over
    `steps` further steps, run in chunks whose sampled energies the engine
    reduces to means of `block` steps on the device (vmc_run_blocked).  tau_int_sokal: Sokal's windowed sum over the
    autocorrelation about the mean pooled over all walkers (stats.iat_pooled),
    which needs series only a few windows long; tau_int_blocking: per-walker
    automated blocking, which reads low until each series is some hundred tau
    long.  The IAT-corrected rates use the larger of the two (tau_int)."""
    from vmachine_amd import stats
    W = eng.cfg.num_walkers
    Ys, s1, s2, cnt = [], 0.0, 0.0, 0
    chunk = min(chunk, steps)
    for _ in range(max(1, steps // chunk)):
        sm, _, _, y = eng.run_blocked(chunk, block, record_walkers=W)
        s1 += sm["energy"]
        s2 += sm["energy2"]
        cnt += sm["samples"]
        Ys.append(y)
    Y = np.concatenate(Ys, axis=1)
    mean = s1 / cnt
    var_step = s2 / cnt - mean * mean
    r = stats.iat_pooled(Y, block, var_step)
    tau = max(r["tau_int_sokal"], r["tau_int_blocking"])
    r.update({"tau_int": tau, "series_steps": int(Y.shape[1] * block), "series_walkers": W,
              "block_steps": block, "series_length_in_tau": Y.shape[1] * block / tau,
              "energy": float(mean),
              "energy_err": float(np.sqrt(2.0 * tau * var_step / cnt))})
    return r


def iat_engine(V, args, make, pos, stream):
    """A fresh engine of --iat-walkers walkers from the first of the
    equilibrated positions, with a random stream of its own."""
    eng = make(args.iat_walkers)
    eng.set_positions(pos[: args.iat_walkers])
    eng.set_random_state(args.seed + stream, 0)
    return eng


def cpu_description():
    """Host CPU model and core counts (the cores this process may use, and all)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return model, usable or 1, os.cpu_count() or 1


def cpu_baseline(args):
    """The reference binary (oracle/_ref/vmachine, built from the reference's own
    sources) on the host cores, one MPI rank per core, on a bounded sample of the
    same workload; falls back to the oracle's multithreaded restatement."""
    model, cores, total_cores = cpu_description()
    host = {"cpu_model": model, "cores_usable": cores, "cores_total": total_cores}
    ref = os.path.join(ROOT, "oracle", "_ref", "vmachine")
    mpiexec = shutil.which("mpiexec") or "/opt/conda/bin/mpiexec"
    if os.path.exists(ref) and os.path.exists(mpiexec) and not args.cpu_port:
        per_rank = args.cpu_steps_per_rank
        text = (f"numParticles: {args.N}\nnumDimensions: 2\nhamiltonian: harmonicOscillator\n"
                f"omega: {args.omega}\ninteractionStyle: coulomb\n"
                "waveFunctionElement: slaterDeterminant\nwaveFunctionElement: gaussian\n"
                "waveFunctionElement: padeJastrow\nbasis: hermite\n"
                f"numIterations: 3\nnumSteps: {per_rank * cores}\nlearningRate: 0.0\n"
                f"stepLength: {args.dt}\nsampling: importanceSampling\n")
        with tempfile.TemporaryDirectory() as td:
            with open(os.path.join(td, "input.in"), "w") as f:
                f.write(text)
            try:
                t0 = time.time()
                r = subprocess.run([mpiexec, "-n", str(cores), ref, "input.in"], cwd=td,
                                   stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                                   timeout=600)
                wall = time.time() - t0
                rows = [ln.split() for ln in r.stdout.splitlines()
                        if ln[:1].isdigit() and len(ln.split()) == 11]
                if r.returncode == 0 and len(rows) >= 2:
                    # last column: sampling time summed over ranks for one
                    # iteration.  The chains start from N(0, 1/omega) and keep
                    # their positions between iterations: the first iteration
                    # (the transient from that start) is discarded and the last
                    # one timed.
                    t_sum = float(rows[-1][-1])
                    steps = per_rank * cores + int(per_rank * cores * 0.001) * cores
                    rate = steps / (t_sum / cores)
                    out = {"value": rate, "unit": "walker-steps/s", "cores": cores,
                           "kind": "reference",
                           "sample": f"reference vmachine (FP64, importance sampling), "
                                     f"mpiexec -n {cores}, {per_rank} steps/rank per iteration, "
                                     f"3 iterations, the last timed by its own MPI_Wtime (wall "
                                     f"{wall:.1f} s incl. start-up and the discarded iterations)"}
                    out.update(host)
                    return out
            except (OSError, subprocess.SubprocessError) as ex:
                print(f"# reference baseline failed: {ex}", file=sys.stderr)
    from oracle import oracle as O
    cfg = O.make_config(args.N, 2, args.omega, 1, [O.SLATER, O.GAUSSIAN, O.PADE], None, None,
                        False, O.IMPORTANCE_EXACT, args.dt, quirks=0)
    W = cores * 4
    pos = np.stack([O.initial_positions(args.seed, w, 2 * args.N, args.omega) for w in range(W)])
    P = np.array([[1.0], [1.0], [1.0]])
    counters = np.zeros(W, dtype=np.uint64)
    steps = args.cpu_steps_per_rank // 4
    t0 = time.perf_counter()
    O.run_many(cfg, P, pos, args.seed, counters, steps, 0, threads=cores)
    dt = time.perf_counter() - t0
    out = {"value": W * steps / dt, "unit": "walker-steps/s", "cores": cores, "kind": "port",
           "sample": f"oracle restatement (FP64), {W} walkers x {steps} steps on {cores} threads"}
    out.update(host)
    return out


def rbmpj_workload(args, V):
    """RBM x Slater x Pade-Jastrow, N=12, H=24, sigma=1, Xavier weights from the
    seed (ref src/InitialWeights/xavier.cpp), reference importance sampling
    dt=0.01, 65536 walkers: walker-steps/s of the FP64 walk and of the FP32 walk
    with each RBM-product contraction mode (value: the default, matrix cores in
    FP32)."""
    N, H, nd = 12, 24, 24
    rng = np.random.default_rng(args.seed)
    a = 1.0 / np.sqrt(nd)
    P = np.zeros((4, H + nd * H))
    P[1, :nd] = rng.uniform(-a, a, nd)
    P[2, :] = rng.uniform(-a, a, H + nd * H)
    P[3, 0] = rng.uniform(0.0, 0.2)
    res = {}
    forms = [("fp32", 32, V.RBM_AUTO), ("fp32_lanes", 32, V.RBM_LANES),
             ("fp32_mfma_bf16", 32, V.RBM_MFMA_BF16), ("fp64", 64, V.RBM_AUTO)]
    keep = set(str(args.rbm_forms).split(","))
    for name, prec, mode in [f for f in forms if f[0] in keep]:
        cfg = V.make_config(N, omega=1.0, kinds=[V.SLATER, V.RBM_GAUSSIAN, V.RBM_PRODUCT, V.PADE],
                            sampling=V.IMPORTANCE, step_length=0.01, precision=prec,
                            refresh_interval=args.refresh if prec == 32 else 0,
                            num_walkers=args.walkers, num_hidden=H, sigma=1.0, sorting=True,
                            rbm_matrix=mode)
        eng = V.Engine(cfg)
        eng.set_parameters(P)
        eng.init_walkers(args.seed)
        burn_in(eng, args.burnin, f"rbmpj {name}")
        iters = args.steps if prec == 32 else args.fp64_iters
        wall, kern, E, acc = timed(eng, iters, args.warmup, args.steps_per_iter)
        ws = args.walkers * args.steps_per_iter * iters
        res[name] = {"walker_steps_per_s": ws / wall, "kernel_walker_steps_per_s": ws / kern,
                     "energy": E, "acceptance": acc, "ms_per_step": 1000.0 * wall / iters}
        print(f"rbmpj {name}: {ws / kern:.4g} walker-steps/s (kernel), E {E:.6f}",
              file=sys.stderr, flush=True)
        eng.close()
    if "opt" in keep:
        res["optimisation"] = rbmpj_optimisation(args, V, N, H, P)
    out = {"metric": "walker-steps/s, RBM x Slater x Pade-Jastrow N=12 H=24",
           "value": res.get("fp32", {}).get("walker_steps_per_s"), "unit": "walker-steps/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": res.get("fp32", {}).get("ms_per_step"),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32 walk, f64 inverse rebuild and accumulation",
           "data": "synthetic (walkers drawn from N(0, 1/omega^2), burned in; Xavier weights)",
           "config": {"workload": "2-D quantum dot N=12 omega=1, Coulomb, Slater x RBM-Gaussian x "
                                  "RBM-product (H=24, sorted) x Pade-Jastrow, reference "
                                  "importance sampling dt=0.01",
                      "walkers": args.walkers, "steps_per_iter": args.steps_per_iter,
                      "refresh_interval": args.refresh, "burnin_steps": args.burnin,
                      "rbm_contractions": "FP32 matrix-core tiles (V_MFMA_F32_16X16X4_F32)"},
           "forms": res}
    print(json.dumps(out))


def rbmpj_optimisation(args, V, N, H, P):
    """ADAM (lr 0.01, 10 gradient batches) iterations of --steps-per-iter steps,
    FP32 walk: on the device (vmc_optimize, one hipGraph per iteration) and as
    the host loop the reference runs (vmc_run, host ADAM, vmc_set_parameters)."""
    from vmachine_amd import host as HST
    cfg = V.make_config(N, omega=1.0, kinds=[V.SLATER, V.RBM_GAUSSIAN, V.RBM_PRODUCT, V.PADE],
                        sampling=V.IMPORTANCE, step_length=0.01, precision=32,
                        refresh_interval=args.refresh, num_walkers=args.walkers, num_hidden=H,
                        sigma=1.0, sorting=True)
    S, iters = args.steps_per_iter, args.opt_iters
    ws = args.walkers * S * iters
    out = {}
    with V.Engine(cfg) as eng:
        eng.set_parameters(P)
        eng.init_walkers(args.seed)
        burn_in(eng, args.burnin, "rbmpj optimisation")
        Pd, _, _ = eng.optimize(P, 1, S, 0, kind="adam", learning_rate=0.01)   # warm-up
        t0 = time.perf_counter()
        Pd, hist, ms = eng.optimize(Pd, iters, S, 0, kind="adam", learning_rate=0.01)
        wall = time.perf_counter() - t0
        out["device_loop"] = {"iterations_per_s": iters / wall, "walker_steps_per_s": ws / wall,
                              "kernel_walker_steps_per_s": ws / (ms / 1e3),
                              "energy_last": float(hist[-1, 0] / (args.walkers * S))}
        opt = HST.Optimizer("adam", P.size, 0.01)
        Ph = Pd.copy()
        t0 = time.perf_counter()
        for it in range(iters):
            eng.set_parameters(Ph)
            sums, g, ge, _ = eng.run(S, 0, it, opt.batches)
            upd = opt.update(sums["energy"] / sums["samples"], g / sums["gradient_samples"],
                             ge / sums["gradient_samples"])
            Ph = Ph - upd.reshape(Ph.shape)
        wall = time.perf_counter() - t0
        out["host_loop"] = {"iterations_per_s": iters / wall, "walker_steps_per_s": ws / wall}
    print(f"rbmpj optimisation: device loop {out['device_loop']['iterations_per_s']:.3g} it/s, "
          f"host loop {out['host_loop']['iterations_per_s']:.3g} it/s", file=sys.stderr, flush=True)
    return out


BATCH_DOTS = [(2, 1.0), (6, 1.0), (12, 1.0), (20, 0.5), (30, 0.5), (42, 0.28), (56, 0.1)]


def batch_allotment(spec):
    """Walkers per system: 'n2:T' = in proportion to N^2, T in total; an
    integer = that many for every system."""
    if spec.startswith("n2:"):
        total = int(spec[3:])
        w = np.array([N * N for N, _ in BATCH_DOTS], dtype=np.float64)
        return [max(1, int(round(total * x / w.sum()))) for x in w]
    return [int(spec)] * len(BATCH_DOTS)


def batch_workload(args, V):
    """value from the first of --batch-walkers (comma-separated allotments:
    'n2:T' or a count per system); every allotment is measured the same way."""
    lines = [batch_measure(args, V, batch_allotment(c)) for c in str(args.batch_walkers).split(",")]
    out = lines[0]
    out["other_walker_counts"] = lines[1:]
    print(json.dumps(out))


def batch_measure(args, V, Ws):
    """Seven closed-shell dots N = 2..56 (Slater x Gaussian x Pade-Jastrow,
    Coulomb, alpha = gamma = 1, exact Langevin dt --dt with drift limit, FP32
    walk with an FP64 rebuild every --refresh steps), walkers per system Ws.  One optimisation iteration = --batch-steps steps per walker of every
    system, per-system sums and the device ADAM step (learning rate
    --batch-lr).  value = walker-steps/s of all systems together, the
    iteration replayed as a hipGraph, shapes and launch form from the
    autotuner (one kernel per system per refresh segment side by side on
    streams of their own, or one persistent launch per iteration); beside it
    the same iterations with the default shapes, in the other launch form,
    without the graph, and as seven single-system engines run one after
    another (host optimiser)."""
    S, lr = args.batch_steps, args.batch_lr
    systems = [dict(num_particles=N, omega=w, num_walkers=W, step_length=args.dt, alpha=1.0,
                    gamma=1.0) for (N, w), W in zip(BATCH_DOTS, Ws)]
    kinds = [V.SLATER, V.GAUSSIAN, V.PADE]

    def make_batch():
        b = V.Batch(systems, kinds=kinds, sampling=V.IMPORTANCE_EXACT, quirks=0, precision=32,
                    refresh_interval=args.refresh, drift_limit=args.drift_limit)
        b.init_walkers(args.seed)
        b.run(max(args.warmup, 1), S, 0, learning_rate=lr)
        return b

    def time_batch(b, use_graph):
        t0 = time.perf_counter()
        out, ms = b.run(args.steps, S, 0, learning_rate=lr, use_graph=use_graph)
        wall = time.perf_counter() - t0
        return wall, ms / 1000.0, out

    total_ws = sum(Ws) * S * args.steps
    res = {}
    b = make_batch()
    wall, kern, _ = time_batch(b, True)
    res["default_shapes_graph"] = {"walker_steps_per_s": total_ws / wall,
                                   "kernel_walker_steps_per_s": total_ws / kern}
    print(f"batch default shapes: {total_ws / wall:.4g} walker-steps/s", file=sys.stderr, flush=True)
    b.set_launch(V.BATCH_PERSISTENT)
    wall, kern, _ = time_batch(b, True)
    res["default_shapes_persistent_graph"] = {"walker_steps_per_s": total_ws / wall,
                                              "kernel_walker_steps_per_s": total_ws / kern}
    print(f"batch default shapes, persistent: {total_ws / wall:.4g} walker-steps/s",
          file=sys.stderr, flush=True)
    lanes, ns = b.autotune(args.batch_tune_steps)
    launch = {V.BATCH_SIDE_BY_SIDE: "side_by_side", V.BATCH_PERSISTENT: "persistent"}[b.launch()]
    print(f"autotune lanes={lanes.tolist()} ns={ns.tolist()} launch={launch}", file=sys.stderr,
          flush=True)
    tuned = b.launch()
    for other, key in ((V.BATCH_SIDE_BY_SIDE, "side_by_side_graph"),
                       (V.BATCH_PERSISTENT, "persistent_graph")):
        b.set_launch(other)
        wall_o, kern_o, _ = time_batch(b, True)
        res[key] = {"walker_steps_per_s": total_ws / wall_o,
                    "kernel_walker_steps_per_s": total_ws / kern_o}
    b.set_launch(tuned)
    wall, kern, out = time_batch(b, True)
    res["graph"] = {"walker_steps_per_s": total_ws / wall, "kernel_walker_steps_per_s": total_ws / kern,
                    "ms_per_step": 1000.0 * wall / args.steps}
    wall_d, kern_d, _ = time_batch(b, False)
    res["direct"] = {"walker_steps_per_s": total_ws / wall_d,
                     "kernel_walker_steps_per_s": total_ws / kern_d}
    energies = [float(out[-1, s, 0] / (Ws[s] * S)) for s in range(len(systems))]
    params = b.parameters().tolist()
    b.close()
    print(f"batch autotuned graph: {res['graph']['walker_steps_per_s']:.4g}", file=sys.stderr,
          flush=True)

    # the same work as single-system engines, one after another, host optimiser
    from vmachine_amd import host as H
    engines = []
    for s, (N, w) in enumerate(BATCH_DOTS):
        cfg = V.make_config(N, omega=w, kinds=kinds, sampling=V.IMPORTANCE_EXACT,
                            step_length=args.dt, quirks=0, drift_limit=args.drift_limit,
                            precision=32, refresh_interval=args.refresh, num_walkers=Ws[s])
        e = V.Engine(cfg)
        P = np.array([[0.0], [1.0], [1.0]])
        e.set_parameters(P)
        e.init_walkers(args.seed + s)
        engines.append([e, P, H.Optimizer("adam", 3, lr)])

    def seq_iteration(it):
        for rec in engines:
            e, P, opt = rec
            sums, g, ge, _ = e.run(S, 0, it, 1)
            upd = opt.update(sums["energy"] / sums["samples"], g / sums["gradient_samples"],
                             ge / sums["gradient_samples"])
            rec[1] = P - upd.reshape(P.shape)
            e.set_parameters(rec[1])

    for it in range(max(args.warmup, 1)):
        seq_iteration(it)
    t0 = time.perf_counter()
    for it in range(args.steps):
        seq_iteration(max(args.warmup, 1) + it)
    wall_s = time.perf_counter() - t0
    for rec in engines:
        rec[0].close()
    res["sequential_engines"] = {"walker_steps_per_s": total_ws / wall_s}
    value = res["graph"]["walker_steps_per_s"]
    out = {"metric": "walker-steps/s, 7 quantum dots N=2..56 walked side by side",
           "value": value, "unit": "walker-steps/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": res["graph"]["ms_per_step"],
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32 walk, f64 inverse rebuild and accumulation",
           "data": "synthetic (walkers drawn from N(0, 1/omega^2) per system)",
           "config": {"workload": "closed-shell 2-D dots N=2,6,12,20,30,42,56 (omega "
                                  "1,1,1,0.5,0.5,0.28,0.1), Coulomb, Slater x Gaussian x "
                                  "Pade-Jastrow, exact Langevin",
                      "walkers_per_system": Ws, "total_walkers": sum(Ws),
                      "steps_per_iteration": S, "dt": args.dt,
                      "refresh_interval": args.refresh, "learning_rate": lr},
           "autotune": {"lanes_per_walker": lanes.tolist(), "ns_per_walker_step": ns.tolist(),
                        "launch": launch},
           "forms": res,
           "speedup_vs_sequential_engines": value / res["sequential_engines"]["walker_steps_per_s"],
           "energies_last_iteration": energies, "alpha_gamma": params}
    return out


def equilibrate(V, args):
    """Walker positions at equilibrium: from N(0, 1/omega^2) per coordinate (ref
    randomnormal.cpp:17), exact Langevin with a large step (dt=--equil-dt, drift
    limit 1) for --equil-steps steps per walker.  Any exact method samples
    |psi|^2, so the measured forms all start from these positions; at dt=0.005
    Langevin itself would need ~10^6 steps per walker to get there."""
    cfg = V.make_config(args.N, omega=args.omega, kinds=[V.SLATER, V.GAUSSIAN, V.PADE],
                        sampling=V.IMPORTANCE_EXACT, step_length=args.equil_dt, quirks=0,
                        drift_limit=1.0, precision=32, refresh_interval=args.refresh,
                        num_walkers=args.walkers)
    with V.Engine(cfg) as eng:
        eng.set_parameters(np.array([[1.0], [1.0], [1.0]]))
        eng.init_walkers(args.seed)
        burn_in(eng, args.equil_steps, f"equilibration (Langevin dt={args.equil_dt})")
        return eng.positions()


def lp_engine(V, args, pos, walkers=None, refresh=None):
    W = walkers or args.walkers
    cfg = V.make_config(args.N, omega=args.omega, kinds=[V.SLATER, V.GAUSSIAN, V.PADE],
                        sampling=V.LEARNED, step_length=args.dt, quirks=0, precision=32,
                        refresh_interval=refresh or args.refresh, num_walkers=W)
    eng = V.Engine(cfg)
    eng.set_parameters(np.array([[1.0], [1.0], [1.0]]))
    eng.lp_set(lp_initial_weights(args.seed), *lp_hyper(args))
    eng.set_positions(pos[:W])
    eng.set_random_state(args.seed + 14, 0)
    return eng


def lp_hyper(args):
    return args.lp_sigma0, args.lp_mu_max, args.lp_ls_max


def lp_initial_weights(seed):
    """Glorot-uniform hidden layers, zero head: the untrained proposal is a
    symmetric Gaussian random walk of width sigma0/sqrt(omega) (--lp-sigma0)."""
    rng = np.random.default_rng(seed)
    th = np.zeros(V_LP_NPARAM, dtype=np.float32)
    a1, a2 = np.sqrt(6.0 / (18 + 64)), np.sqrt(6.0 / 128)
    th[:64 * 18] = rng.uniform(-a1, a1, 64 * 18)
    w2 = 64 * 18 + 64
    th[w2:w2 + 64 * 64] = rng.uniform(-a2, a2, 64 * 64)
    return th


V_LP_NPARAM = 5636


def copy_workload(args):
    """A device-to-device copy of --copy-mib MiB, 5 times: the reference reading for
    FETCH_SIZE / WRITE_SIZE (each launch reads and writes exactly that many bytes)."""
    import torch
    n = args.copy_mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / 5
    print(json.dumps({"workload": "copy", "mib": args.copy_mib, "bytes_read_per_launch": 4 * n,
                      "bytes_written_per_launch": 4 * n, "GB_per_s": 8.0 * n / dt / 1e9}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed sampling iterations")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--steps-per-iter", type=int, default=1000)
    ap.add_argument("--walkers", type=int, default=None,
                    help="walkers (default 32768; 65536 for the rbmpj workload)")
    ap.add_argument("--N", type=int, default=56)
    ap.add_argument("--omega", type=float, default=0.1)
    ap.add_argument("--dt", type=float, default=0.005)
    ap.add_argument("--refresh", type=int, default=100)
    ap.add_argument("--drift-limit", type=float, default=1.0,
                    help="Umrigar drift limit a of the exact Langevin proposal (0: none)")
    ap.add_argument("--burnin", type=int, default=40000,
                    help="rbmpj workload: burn-in steps per walker")
    ap.add_argument("--equil-steps", type=int, default=60000)
    ap.add_argument("--equil-dt", type=float, default=0.1)
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--iat-steps", type=int, default=1 << 23,
                    help="recorded steps per walker for tau_int of the Langevin forms")
    ap.add_argument("--iat-walkers", type=int, default=2048,
                    help="walkers of the tau_int runs (every one recorded)")
    ap.add_argument("--iat-block", type=int, default=512,
                    help="steps per block mean of the recorded Langevin series")
    ap.add_argument("--lp-iat-steps", type=int, default=1 << 16)
    ap.add_argument("--lp-iat-block", type=int, default=4)
    ap.add_argument("--lp-train-iters", type=int, default=300)
    ap.add_argument("--lp-train-steps", type=int, default=10)
    ap.add_argument("--lp-lr", type=float, default=3e-3)
    ap.add_argument("--lp-sigma0", type=float, default=0.4,
                    help="learned proposal: move scale of the zero-head network, 1/sqrt(omega) "
                         "units (the trained scale stays within sigma0 exp(+-ls_max))")
    ap.add_argument("--lp-mu-max", type=float, default=1.0,
                    help="learned proposal: bound of the mean move, 1/sqrt(omega) units")
    ap.add_argument("--lp-ls-max", type=float, default=2.0,
                    help="learned proposal: bound of |log(scale / sigma0)|")
    ap.add_argument("--fp64-iters", type=int, default=3)
    ap.add_argument("--cpu-steps-per-rank", type=int, default=200000)
    ap.add_argument("--cpu-port", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-iat", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the mixed-size batch figure of the default workload")
    ap.add_argument("--batch-main", default="n2:131072",
                    help="walker allotment of the batch figure in the default workload")
    ap.add_argument("--workload", choices=["dot56", "rbmpj", "batch", "copy"], default="dot56",
                    help="dot56: the headline N=56 workload; rbmpj: RBM x Slater x Pade N=12 H=24; "
                         "batch: seven dots N=2..56 walked side by side; "
                         "copy: a plain device copy of --copy-mib MiB (calibrates HBM byte counters)")
    ap.add_argument("--batch-walkers", default="n2:131072,8192,1024",
                    help="walker allotments, comma-separated, one measurement each: 'n2:T' "
                         "(in proportion to N^2, T in total) or a count per system")
    ap.add_argument("--batch-steps", type=int, default=200)
    ap.add_argument("--batch-lr", type=float, default=0.01)
    ap.add_argument("--batch-tune-steps", type=int, default=50)
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--rbm-forms", default="fp32,fp32_lanes,fp32_mfma_bf16,fp64,opt",
                    help="rbmpj workload: the forms to run (fp32 = matrix-core FP32 tiles; "
                         "opt = ADAM iterations, device loop vs host loop)")
    ap.add_argument("--opt-iters", type=int, default=5,
                    help="rbmpj workload: timed optimisation iterations per loop form")
    args = ap.parse_args()
    if args.walkers is None:
        args.walkers = 65536 if args.workload == "rbmpj" else 32768

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this engine runs independent walkers on one device "
              "and no multi-GPU figure is produced (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "copy":
        copy_workload(args)
        return
    import vmachine_amd as V
    if V.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.workload == "rbmpj":
        rbmpj_workload(args, V)
        return
    if args.workload == "batch":
        batch_workload(args, V)
        return
    n = {2: 1, 6: 3, 12: 6, 20: 10, 30: 15, 42: 21, 56: 28, 72: 36, 90: 45, 110: 55}[args.N]
    pos = equilibrate(V, args)

    # ---- FP32 walk, exact Langevin dt (value) ----
    # every measured form starts from the equilibrated positions with a random
    # stream of its own (reusing the equilibration's stream would replay the
    # noise that produced the starting state)
    eng = make_engine(V, args, 32, args.walkers)
    eng.set_positions(pos)
    eng.set_random_state(args.seed + 11, 0)
    wall, kern, E32, acc32 = timed(eng, args.steps, args.warmup, args.steps_per_iter)
    wsteps = args.walkers * args.steps_per_iter * args.steps
    value = wsteps / wall
    kern_rate = wsteps / kern
    launches = args.steps * ((args.steps_per_iter + args.refresh - 1) // args.refresh)
    launch_s = kern / launches
    steps_per_launch = args.walkers * min(args.refresh, args.steps_per_iter)
    F = flops_per_step(args.N, n, a=acc32)
    B4 = bytes_per_step(args.N, n, 4, a=acc32)
    achieved = F * steps_per_launch / launch_s / 1e12
    eng.close()
    iat = None
    if not args.no_iat:
        with iat_engine(V, args, lambda w: make_engine(V, args, 32, w), pos, 21) as e:
            iat = measure_iat(e, args.iat_steps, args.iat_block)
        print(f"# FP32 tau_int {iat['tau_int']:.0f} steps", file=sys.stderr, flush=True)

    # ---- plain FP64 walk, exact Langevin dt ----
    fp64 = None
    if not args.no_fp64:
        eng = make_engine(V, args, 64, args.walkers)
        eng.set_positions(pos)
        eng.set_random_state(args.seed + 12, 0)
        w64, k64, E64, a64 = timed(eng, args.fp64_iters, 1, args.steps_per_iter)
        r64 = args.walkers * args.steps_per_iter * args.fp64_iters / w64
        fp64 = {"walker_steps_per_s": r64, "energy": E64, "acceptance": a64,
                "kernel_walker_steps_per_s": args.walkers * args.steps_per_iter *
                args.fp64_iters / k64}
        eng.close()
        if not args.no_iat:
            with iat_engine(V, args, lambda w: make_engine(V, args, 64, w), pos, 22) as e:
                it64 = measure_iat(e, args.iat_steps, args.iat_block)
            fp64["iat"] = it64
            fp64["iat_corrected_samples_per_s"] = r64 / (2 * it64["tau_int"])
            print(f"# FP64 tau_int {it64['tau_int']:.0f} steps", file=sys.stderr, flush=True)

    # ---- learned proposal, FP32: online training, then the same measurements ----
    learned = None
    if not args.no_learned:
        eng = lp_engine(V, args, pos)
        t0 = time.perf_counter()
        hist = eng.lp_train(args.lp_train_iters, args.lp_train_steps, args.lp_lr)
        t_train = time.perf_counter() - t0
        theta = eng.lp_get()
        # the rebuild every --refresh accepted moves: every --refresh / a steps
        # at the trained proposal's acceptance a (the Langevin forms accept
        # nearly every move); the walk goes on from the trained engine's
        # walkers and random stream
        acc_train = float(hist[-5:, 1].mean())
        lp_refresh = min(args.steps_per_iter, max(args.refresh, int(args.refresh / max(acc_train, 1e-3))))
        post, (rs, rc) = eng.positions(), eng.random_state()
        eng.close()
        eng = lp_engine(V, args, post, refresh=lp_refresh)
        eng.lp_set(theta, *lp_hyper(args))
        eng.set_random_state(rs, rc)
        wl, kl, El, al = timed(eng, args.steps, args.warmup, args.steps_per_iter)
        rl = wsteps / wl
        learned = {"walker_steps_per_s": rl, "kernel_walker_steps_per_s": wsteps / kl,
                   "energy": El, "acceptance": al,
                   "training": {"iterations": args.lp_train_iters,
                                "walk_steps_between": args.lp_train_steps,
                                "learning_rate": args.lp_lr, "seconds": t_train,
                                "esjd_start": float(hist[:5, 0].mean()),
                                "esjd_end": float(hist[-5:, 0].mean()),
                                "acceptance_end": acc_train},
                   "hyper": {"sigma0": args.lp_sigma0, "mu_max": args.lp_mu_max,
                             "ls_max": args.lp_ls_max},
                   "refresh_interval": lp_refresh}
        eng.close()
        if not args.no_iat:
            def make_lp(w):
                e = lp_engine(V, args, pos, walkers=w, refresh=lp_refresh)
                e.lp_set(theta, *lp_hyper(args))
                return e
            with iat_engine(V, args, make_lp, pos, 23) as e:
                itl = measure_iat(e, args.lp_iat_steps, args.lp_iat_block)
            learned["iat"] = itl
            learned["iat_corrected_samples_per_s"] = rl / (2 * itl["tau_int"])
            print(f"# learned tau_int {itl['tau_int']:.1f} steps", file=sys.stderr, flush=True)

    # ---- the reference's own method (Green-function ratio, quirks) for the CPU's tau ----
    tau_ref = None
    if not args.no_iat and not args.no_cpu:
        def make_compat(w):
            cfg = V.make_config(args.N, omega=args.omega, kinds=[V.SLATER, V.GAUSSIAN, V.PADE],
                                sampling=V.IMPORTANCE, step_length=args.dt, precision=32,
                                refresh_interval=args.refresh, num_walkers=w)
            e = V.Engine(cfg)
            e.set_parameters(np.array([[1.0], [1.0], [1.0]]))
            return e
        with iat_engine(V, args, make_compat, pos, 13) as eng:
            eng.run(5000, 0)
            tau_ref = measure_iat(eng, args.iat_steps, args.iat_block)

    cpu = None if args.no_cpu else cpu_baseline(args)
    if cpu and tau_ref:
        cpu["iat"] = tau_ref
        cpu["iat_corrected_samples_per_s"] = cpu["value"] / (2 * tau_ref["tau_int"])
        cpu["tau_note"] = ("tau of E_L for the reference's own method (its Green-function ratio "
                           "and Pade ratio semantics) measured with the engine's compatible mode "
                           "on the same equilibrated walkers")

    # ---- the mixed-size batch (seven dots, walkers in proportion to N^2) ----
    batch = None
    if not args.no_batch:
        bl = batch_measure(args, V, batch_allotment(args.batch_main))
        batch = {"walker_steps_per_s": bl["value"], "allotment": args.batch_main,
                 "walkers_per_system": bl["config"]["walkers_per_system"],
                 "forms": bl["forms"], "autotune": bl["autotune"],
                 "speedup_vs_sequential_engines": bl["speedup_vs_sequential_engines"],
                 "energies_last_iteration": bl["energies_last_iteration"],
                 "note": "one optimisation iteration = 200 steps per walker of every system + "
                         "device ADAM, replayed as a hipGraph; launch = one kernel per system "
                         "per refresh segment side by side, or one persistent launch per "
                         "iteration (autotuned, both timed); sequential = seven single-system "
                         "engines in turn with the host optimiser"}

    out = {
        "metric": "walker-steps/s and IAT-corrected samples/s at N=56 w=0.1, vs plain FP64",
        "value": value,
        "unit": "walker-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 walk, f64 inverse rebuild and accumulation",
        "data": "synthetic (walkers drawn from N(0, 1/omega^2) per coordinate, equilibrated)",
        "config": {"workload": f"2-D quantum dot N={args.N} omega={args.omega}, Coulomb, "
                               "Slater(Hermite) x Gaussian x Pade-Jastrow (alpha=gamma=1), "
                               f"exact MH Langevin dt={args.dt}, drift limit a={args.drift_limit}",
                   "walkers": args.walkers, "steps_per_iter": args.steps_per_iter,
                   "refresh_interval": args.refresh, "equilibration_steps": args.equil_steps,
                   "equilibration_dt": args.equil_dt,
                   "langevin": ("exact Metropolis-Hastings Langevin (the Green's-function ratio "
                                "in the acceptance) with the Umrigar-Nightingale-Runge drift "
                                f"limit a={args.drift_limit}: F -> 2F/(1+sqrt(1+a F^2 dt/2)) in "
                                "the proposal and its density; still exact, and walkers next "
                                "to a node of the determinant are not frozen by its divergent "
                                "drift" if args.drift_limit > 0 else
                                "exact Metropolis-Hastings Langevin, no drift limit"),
                   "refresh": (f"FP64 Slater-inverse rebuild every {args.refresh} accepted "
                               "moves of every walker on average: every "
                               f"{args.refresh} steps in the FP32 Langevin form, which accepts "
                               "0.9999 of its moves, and every refresh/a steps in the learned "
                               "form, a = its acceptance after training (learned."
                               "refresh_interval)")},
        "kernel_walker_steps_per_s": kern_rate,
        "energy": E32,
        "acceptance": acc32,
        "iat": iat,
        "iat_note": (f"tau_int of E_L per form from {args.iat_walkers} walkers x {args.iat_steps} "
                     f"steps (Langevin forms; learned: x {args.lp_iat_steps}) on fresh engines "
                     "started from the same equilibrated positions: Sokal's windowed sum over "
                     "the autocorrelation about the mean pooled over all walkers, and per-walker "
                     "automated blocking; tau_int = the larger; IAT-corrected samples/s = "
                     "walker-steps/s / (2 tau_int)"),
        "iat_corrected_samples_per_s": (value / (2 * iat["tau_int"])) if iat else None,
        "fp64": fp64,
        "learned": learned,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                     "traffic": PMC_BYTES_PER_WALKER_LAUNCH * args.walkers,
                     "peak_nonpacked": PEAK_FP32_TFLOPS / 2,
                     "frac_nonpacked": achieved / (PEAK_FP32_TFLOPS / 2),
                     "note": "VALU FP32 compute roofline: the walk kernel issues no MFMA and "
                             "keeps walker state in registers, so HBM is far from its bound. "
                             "peak = MI355X's 157.3 TFLOP/s vector FP32, which counts packed "
                             "FP32 (v_pk_fma_f32); the build keeps packed FP32 out of every "
                             "code object, so the reachable VALU ceiling is half of it "
                             "(peak_nonpacked, frac_nonpacked).  Algorithmic flops "
                             f"{F:.0f}/walker-step x {steps_per_launch} walker-steps per launch; "
                             f"streaming-equivalent bytes {B4:.0f}/walker-step; traffic = HBM "
                             "bytes per launch measured by rocprofv3 FETCH_SIZE/WRITE_SIZE "
                             "(scaled by a 1 GiB copy) at N=56, 100 steps per launch"},
        "cpu_baseline": cpu,
    }
    if batch:
        out["batch_c4"] = batch
    if cpu:
        out["gpu_vs_cpu"] = value / cpu["value"]
    if learned and fp64 and "iat_corrected_samples_per_s" in learned and \
            "iat_corrected_samples_per_s" in fp64:
        out["learned_vs_plain_fp64_iat_corrected"] = (learned["iat_corrected_samples_per_s"] /
                                                      fp64["iat_corrected_samples_per_s"])
        if cpu and "iat_corrected_samples_per_s" in cpu:
            out["learned_vs_cpu_iat_corrected"] = (learned["iat_corrected_samples_per_s"] /
                                                   cpu["iat_corrected_samples_per_s"])
    print(json.dumps(out))


if __name__ == "__main__":
    main()
