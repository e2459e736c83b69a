"""Writes the golden fixtures in tests/golden/fixtures/ from the reference itself.

Run in the build container only (it needs the reference tree to build
oracle/_ref):  python tests/golden/make_golden.py

Each scenario writes an input.in, a weights file (read by the reference's
`initialWeights: fromfile`) and a script of random numbers for the harness in
oracle/ref/harness.cpp, which drives the reference's own System, elements,
Metropolis, Sampler and optimiser objects.  The fixture stores the inputs and
the reference's outputs; tests compare the oracle (and through it the HIP
engine) against them.  The random numbers are the engine's own Philox stream,
so the same fixture also pins the engine's random-number protocol.
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402

FIX = os.path.join(HERE, "fixtures")
HARNESS = os.path.join(ROOT, "oracle", "_ref", "vmc_ref_harness")
NAMES = {O.GAUSSIAN: "gaussian", O.SLATER: "slaterDeterminant", O.PADE: "padeJastrow",
         O.RBM_GAUSSIAN: "RBMGaussian", O.RBM_PRODUCT: "RBMProduct"}


def fmt(v):
    return "%.17g" % v


def input_in(sc, nsteps, niter, weights_path):
    L = [f"numParticles: {sc['N']}", f"numDimensions: {sc['D']}",
         "hamiltonian: harmonicOscillator", f"omega: {fmt(sc['omega'])}"]
    if sc["interaction"]:
        L.append("interactionStyle: coulomb")
    for k in sc["kinds"]:
        L.append(f"waveFunctionElement: {NAMES[k]}")
    if O.SLATER in sc["kinds"]:
        L.append("basis: hermite")
    if O.RBM_PRODUCT in sc["kinds"]:
        L.append(f"numHiddenNodes: {sc['num_hidden']}")
    L.append(f"sorting: {'true' if sc.get('sorting') else 'false'}")
    L += [f"numIterations: {niter}", f"numSteps: {nsteps}",
          f"learningRate: {fmt(sc.get('lr', 0.1))}", f"stepLength: {fmt(sc['step_length'])}",
          "sampling: " + ("bruteForce" if sc["sampling"] == O.BRUTE_FORCE else "importanceSampling")]
    opt = sc.get("optimizer", "adam")
    L.append("optimization: " + {"adam": "adam", "sgd": "sgd 0 0", "gd": "gd 0 0"}[opt])
    L.append(f"initialWeights: fromfile {weights_path}")
    return "\n".join(L) + "\n"


def draws(sc, nsteps_total):
    nd = sc["N"] * sc["D"]
    out = {"k": [], "z": [], "u_move": [], "u_acc": []}
    for c in range(nsteps_total):
        k, z, um, ua = O.step_draws(sc["seed"], 0, c, nd)
        out["k"].append(k)
        out["z"].append(z)
        out["u_move"].append(um)
        out["u_acc"].append(ua)
    return out


def run_scenario(name, sc, mode, nsteps, niter=1):
    """mode 'elements': nsteps single-step snapshots; 'iterations': niter x nsteps steps."""
    cfg = O.make_config(**{k: sc[k] for k in ["N", "D", "omega", "interaction", "kinds",
                                              "num_hidden", "sorting", "sampling",
                                              "step_length"]})
    mp = O.max_parameters(cfg)
    E = len(sc["kinds"])
    P = np.array(sc["params"], dtype=np.float64).reshape(E, mp)
    with tempfile.TemporaryDirectory() as td:
        wpath = os.path.join(td, "weights.dat")
        with open(wpath, "w") as f:
            for row in P:
                f.write(" ".join(fmt(v) for v in row) + "\n")
        if mode == "elements":
            text = input_in(sc, 1, 1, wpath)
            total = nsteps
        else:
            text = input_in(sc, nsteps, niter, wpath)
            equil = int(nsteps * 0.001)
            total = niter * (nsteps + equil)
        ipath = os.path.join(td, "input.in")
        with open(ipath, "w") as f:
            f.write(text)
        d = draws(sc, total)
        spath = os.path.join(td, "script.txt")
        with open(spath, "w") as f:
            f.write(f"mode {mode}\n")
            f.write("sampling " + ("brute" if sc["sampling"] == O.BRUTE_FORCE else "importance") + "\n")
            f.write(f"steps {nsteps}\niterations {niter}\n")
            f.write("positions " + " ".join(fmt(v) for v in sc["positions"]) + "\n")
            f.write(f"ints {total} " + " ".join(str(v) for v in d["k"]) + "\n")
            f.write(f"gauss {total} " + " ".join(fmt(v) for v in d["z"]) + "\n")
            if sc["sampling"] == O.BRUTE_FORCE:
                u = []
                for a, b in zip(d["u_move"], d["u_acc"]):
                    u += [a, b]
            else:
                u = d["u_acc"]
            f.write(f"unif {len(u)} " + " ".join(fmt(v) for v in u) + "\n")
        opath = os.path.join(td, "out.json")
        subprocess.run([HARNESS, ipath, spath, opath], check=True, cwd=td,
                       stdout=subprocess.DEVNULL)
        with open(opath) as f:
            ref = json.load(f)
    fixture = {"name": name, "mode": mode, "nsteps": nsteps, "niter": niter,
               "scenario": sc, "input_in": text, "reference": ref}
    with open(os.path.join(FIX, name + ".json"), "w") as f:
        json.dump(fixture, f)
    print("wrote", name)


def scenario(N, kinds, omega=1.0, interaction=1, num_hidden=None, sorting=False,
             sampling=O.IMPORTANCE_COMPAT, step_length=0.01, seed=1, pos_scale=None,
             params=None, **extra):
    rng = np.random.default_rng(seed)
    D = 2
    cfg = O.make_config(N, D, omega, interaction, kinds, num_hidden, None, sorting, sampling,
                        step_length)
    mp = O.max_parameters(cfg)
    if params is None:
        P = np.zeros((len(kinds), mp))
        for e, k in enumerate(kinds):
            npar = O.num_parameters(cfg, e)
            if k == O.GAUSSIAN:
                P[e, 0] = 1.0 + 0.1 * rng.standard_normal()
            elif k == O.PADE:
                P[e, 0] = 0.5 + 0.5 * rng.random()
            elif k in (O.RBM_GAUSSIAN, O.RBM_PRODUCT):
                a = 1.0 / np.sqrt(N * D)
                P[e, :npar] = rng.uniform(-a, a, npar)
        params = P.tolist()
    scale = (1.0 / np.sqrt(omega)) if pos_scale is None else pos_scale
    positions = (scale * rng.standard_normal(N * D)).tolist()
    sc = dict(N=N, D=D, omega=omega, interaction=interaction, kinds=list(kinds),
              num_hidden=N if num_hidden is None else num_hidden, sorting=sorting,
              sampling=sampling, step_length=step_length, seed=seed, params=params,
              positions=positions)
    sc.update(extra)
    return sc


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "all", "ref"], check=True)
    os.makedirs(FIX, exist_ok=True)
    G, S, P, RG, RP = O.GAUSSIAN, O.SLATER, O.PADE, O.RBM_GAUSSIAN, O.RBM_PRODUCT
    # --- element-level snapshots: one Metropolis step per snapshot ---
    run_scenario("el_n2_gauss_pade_importance",
                 scenario(2, [G, P], step_length=0.1, seed=11), "elements", 40)
    run_scenario("el_n2_gauss_pade_brute",
                 scenario(2, [G, P], step_length=0.1, sampling=O.BRUTE_FORCE, seed=12),
                 "elements", 40)
    run_scenario("el_n2_gauss_nointeraction",
                 scenario(2, [G], interaction=0, step_length=0.1, seed=13,
                          params=[[1.0]]), "elements", 20)
    run_scenario("el_n6_slater_gauss_pade",
                 scenario(6, [S, G, P], step_length=0.05, seed=14), "elements", 40)
    run_scenario("el_n12_slater_gauss_pade",
                 scenario(12, [S, G, P], step_length=0.02, seed=15), "elements", 40)
    run_scenario("el_n20_slater_gauss_pade",
                 scenario(20, [S, G, P], omega=0.5, step_length=0.02, seed=16), "elements", 30)
    run_scenario("el_n56_slater_gauss_pade_w0.1",
                 scenario(56, [S, G, P], omega=0.1, step_length=0.005, seed=17), "elements", 30)
    run_scenario("el_n2_rbm",
                 scenario(2, [RG, RP], num_hidden=2, step_length=0.1, sorting=True, seed=18),
                 "elements", 30)
    run_scenario("el_n12_rbmpj_sorted",
                 scenario(12, [S, RG, RP, P], num_hidden=24, step_length=0.01, sorting=True,
                          seed=19), "elements", 12)
    run_scenario("el_n12_rbmpj_unsorted",
                 scenario(12, [S, RG, RP, P], num_hidden=24, step_length=0.01, sorting=False,
                          seed=20), "elements", 12)
    # --- whole iterations: Metropolis loop + Sampler + optimiser ---
    run_scenario("it_n2_gauss_pade_adam",
                 scenario(2, [G, P], step_length=0.1, seed=21, optimizer="adam", lr=0.1),
                 "iterations", 3000, 4)
    run_scenario("it_n2_gauss_pade_brute_sgd",
                 scenario(2, [G, P], step_length=0.1, sampling=O.BRUTE_FORCE, seed=22,
                          optimizer="sgd", lr=0.1), "iterations", 2000, 3)
    run_scenario("it_n6_slater_gauss_pade_adam",
                 scenario(6, [S, G, P], step_length=0.05, seed=23, optimizer="adam", lr=0.1),
                 "iterations", 2000, 3)
    run_scenario("it_n12_rbmpj_adam",
                 scenario(12, [S, RG, RP, P], num_hidden=24, step_length=0.01, sorting=False,
                          seed=24, optimizer="adam", lr=0.01), "iterations", 500, 2)
    run_scenario("it_n2_gauss_gd_nointeraction",
                 scenario(2, [G], interaction=0, step_length=0.1, seed=25, optimizer="gd",
                          lr=0.5, params=[[0.8]]), "iterations", 1000, 3)


if __name__ == "__main__":
    main()
