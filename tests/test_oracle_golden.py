"""Pins the FP64 oracle against golden vectors written by the reference itself.

The fixtures (tests/golden/fixtures/*.json) hold the reference's per-element
ratios, gradients, Laplacians, parameter gradients and energies after each
scripted Metropolis step, and whole optimisation iterations (sampler averages,
optimiser updates, final positions), all produced by oracle/ref/harness.cpp
driving the reference's objects.  The oracle must reproduce them to FP64
rounding.
"""
import glob
import json
import os

import numpy as np
import pytest

from oracle import oracle as O

FIX = os.path.join(os.path.dirname(__file__), "golden", "fixtures")
EL = sorted(glob.glob(os.path.join(FIX, "el_*.json")))
IT = sorted(glob.glob(os.path.join(FIX, "it_*.json")))

RTOL = 1e-9


def _cfg(sc):
    return O.make_config(sc["N"], sc["D"], sc["omega"], sc["interaction"], sc["kinds"],
                         sc["num_hidden"], None, sc["sorting"], sc["sampling"],
                         sc["step_length"], quirks=O.QUIRKS_ALL, basis_scaled=False)


def _close(a, b, scale=None, rtol=RTOL, what=""):
    a = np.asarray(a, dtype=float)
    b = np.asarray(b, dtype=float)
    s = np.max(np.abs(b)) if scale is None else scale
    err = np.max(np.abs(a - b)) if a.size else 0.0
    assert err <= rtol * max(s, 1.0), f"{what}: max abs err {err} (scale {s})"


def _check_snapshot(w, snap, ne, with_ratio):
    _close(w.positions(), snap["positions"], what="positions")
    for e in range(ne):
        ref = snap["elements"][e]
        if with_ratio:
            _close(w.ratio(e), ref["ratio"], what=f"ratio e{e}")
        _close(w.gradient(e), ref["gradient"], what=f"gradient e{e}")
        _close(w.laplacian(e), ref["laplacian"], what=f"laplacian e{e}")
        _close(w.parameter_gradient(e), ref["parameter_gradient"], what=f"pgrad e{e}")
    _close(w.kinetic(), snap["kinetic"], what="kinetic")
    _close(w.external(), snap["external"], what="external")
    _close(w.interaction(), snap["interaction"], what="interaction")


@pytest.mark.parametrize("path", EL, ids=[os.path.basename(p)[:-5] for p in EL])
def test_elements_match_reference(path):
    fx = json.load(open(path))
    sc = fx["scenario"]
    cfg = _cfg(sc)
    w = O.Walker(cfg, sc["params"], sc["positions"])
    snaps = fx["reference"]["snapshots"]
    ne = len(sc["kinds"])
    _check_snapshot(w, snaps[0], ne, False)
    nd = sc["N"] * sc["D"]
    accepted = 0
    for i, snap in enumerate(snaps[1:]):
        k, z, um, ua = O.step_draws(sc["seed"], 0, i, nd)
        acc = w.step(k, z, um, ua)
        ref_acc = not np.array_equal(np.asarray(snap["positions"]), np.asarray(snaps[i]["positions"]))
        assert acc == ref_acc, f"step {i}: accept {acc} vs reference {ref_acc}"
        accepted += acc
        _check_snapshot(w, snap, ne, True)
    assert accepted > 0


@pytest.mark.parametrize("path", IT, ids=[os.path.basename(p)[:-5] for p in IT])
def test_iterations_match_reference(path):
    fx = json.load(open(path))
    sc = fx["scenario"]
    cfg = _cfg(sc)
    ne = len(sc["kinds"])
    mp = O.max_parameters(cfg)
    P = np.array(sc["params"], dtype=float).reshape(ne, mp)
    w = O.Walker(cfg, P, sc["positions"])
    steps = fx["nsteps"]
    equil = int(steps * 0.001)
    opt_kind = {"adam": 0, "sgd": 1, "gd": 2}[sc["optimizer"]]
    batches = 10 if opt_kind == 0 else 1
    opt = O.Optimizer(opt_kind, ne * mp, sc["lr"])
    counter = 0
    for it, ref in enumerate(fx["reference"]["iterations"]):
        counter, acc, g, ge, _ = w.run(sc["seed"], 0, counter, steps, equil, it, batches)
        S = steps // batches
        avgE = acc[0] / steps
        avgG = g.reshape(ne, mp) / S
        avgGE = ge.reshape(ne, mp) / S
        _close(avgE, ref["energy"], what=f"iter {it} energy")
        _close(avgG, ref["avg_gradients"], rtol=1e-8, what=f"iter {it} <G>")
        _close(avgGE, ref["avg_gradients_e"], rtol=1e-8, what=f"iter {it} <GE>")
        P = P - opt.update(avgE, avgG, avgGE).reshape(ne, mp)
        _close(P, ref["parameters"], rtol=1e-8, what=f"iter {it} parameters")
        w.set_parameters(P)
        _close(w.positions(), ref["positions"], rtol=1e-8, what=f"iter {it} positions")
