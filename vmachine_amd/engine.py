"""ctypes binding of the C ABI in include/vmc.h (libvmc_hip.so).

The engine runs only through the HIP library: if libvmc_hip.so is missing,
loading fails with an error that says how to build it; there is no CPU path.
"""
import ctypes
import os
import re

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIB_PATH = os.path.join(PKG, "_lib", "libvmc_hip.so")
HEADER = os.path.join(ROOT, "include", "vmc.h")

GAUSSIAN, SLATER, PADE, RBM_GAUSSIAN, RBM_PRODUCT = 1, 2, 3, 4, 5
DENSITY_RADIAL, DENSITY_SPATIAL, DENSITY_TWO_BODY = 1, 2, 4
# RBM-product contractions (vmc_config.rbm_matrix)
RBM_AUTO, RBM_LANES, RBM_MFMA_F32, RBM_MFMA_BF16 = 0, 1, 2, 3
BRUTE_FORCE, IMPORTANCE, IMPORTANCE_EXACT, LEARNED = 0, 1, 2, 3
LP_NPARAM = 5636
Q2_PADE_RATIO_UPPER, Q3_PADE_GAMMA_HALF = 1 << 2, 1 << 3
Q6_RBM_W_UNSCALED, Q13_RBMG_PSI_RATIO = 1 << 6, 1 << 13
QUIRKS_REFERENCE = Q2_PADE_RATIO_UPPER | Q3_PADE_GAMMA_HALF | Q6_RBM_W_UNSCALED | Q13_RBMG_PSI_RATIO
EVAL_E, EVAL_T, EVAL_VEXT, EVAL_VINT, EVAL_GALPHA, EVAL_GGAMMA = range(6)
EVAL_RATIO_SLATER, EVAL_RATIO_PADE, EVAL_RATIO_GAUSS, EVAL_DRIFT = 6, 7, 8, 9

_D, _I, _U = ctypes.c_double, ctypes.c_int, ctypes.c_uint
_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int)
_FP = ctypes.POINTER(ctypes.c_float)


class VmcConfig(ctypes.Structure):
    _fields_ = [
        ("num_particles", _I), ("num_dimensions", _I), ("omega", _D), ("interaction", _I),
        ("num_elements", _I), ("kinds", _I * 6), ("num_hidden", _I), ("sigma", _D),
        ("sorting", _I), ("sampling", _I), ("step_length", _D), ("quirks", _U),
        ("precision", _I), ("refresh_interval", _I), ("num_walkers", _I), ("drift_limit", _D),
        ("rbm_matrix", _I),
    ]


class VmcSums(ctypes.Structure):
    _fields_ = [(n, _D) for n in [
        "energy", "energy2", "kinetic", "external", "interaction", "kinetic2", "external2",
        "interaction2", "accepted", "samples", "gradient_samples", "kernel_ms", "nonfinite"]]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class VmcError(RuntimeError):
    pass


class VmcBatchSystem(ctypes.Structure):
    _fields_ = [("num_particles", _I), ("omega", _D), ("num_walkers", _I),
                ("step_length", _D), ("alpha", _D), ("gamma", _D)]


class VmcOptimizerConfig(ctypes.Structure):
    _fields_ = [("kind", _I), ("learning_rate", _D), ("gamma", _D), ("monotonic", _D),
                ("batches", _I)]


OPT_ADAM, OPT_SGD, OPT_GD = 0, 1, 2
OPT_OUT = 16   # per iteration: the 14 walk sums (kernel order), alpha, gamma after the step


class VmcBatchConfig(ctypes.Structure):
    _fields_ = [("num_systems", _I), ("systems", ctypes.POINTER(VmcBatchSystem)),
                ("num_elements", _I), ("kinds", _I * 6), ("interaction", _I), ("sampling", _I),
                ("quirks", _U), ("precision", _I), ("refresh_interval", _I),
                ("drift_limit", _D)]


BATCH_OUT = 16   # per system and iteration: the 14 walk sums (VMC_EVAL order), alpha, gamma
BATCH_MAX_SYSTEMS = 8
BATCH_SIDE_BY_SIDE = 0   # vmc_batch_set_launch: one kernel per system per refresh segment
BATCH_PERSISTENT = 1     # one persistent launch per iteration


class LpHyper(ctypes.Structure):
    _fields_ = [("sigma0", _D), ("mu_max", _D), ("ls_max", _D)]


_lib = None


def header_functions():
    """Names of the functions include/vmc.h declares."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(vmc_\w+)\s*\(", text, re.M)))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise VmcError(f"{LIB_PATH} is not built: run `python -m vmachine_amd.build` "
                           "(hipcc --offload-arch=gfx950)")
        L = ctypes.CDLL(LIB_PATH)
        L.vmc_last_error.restype = ctypes.c_char_p
        L.vmc_num_parameters.argtypes = [ctypes.POINTER(VmcConfig), _I]
        L.vmc_max_parameters.argtypes = [ctypes.POINTER(VmcConfig)]
        L.vmc_device_count.restype = _I
        L.vmc_engine_create.argtypes = [ctypes.POINTER(VmcConfig), ctypes.POINTER(ctypes.c_void_p)]
        L.vmc_engine_destroy.argtypes = [ctypes.c_void_p]
        L.vmc_engine_destroy.restype = None
        L.vmc_set_parameters.argtypes = [ctypes.c_void_p, _DP, _I, _I]
        L.vmc_init_walkers.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
        for n in ["vmc_set_positions", "vmc_get_positions", "vmc_set_quantum_force",
                  "vmc_get_quantum_force"]:
            getattr(L, n).argtypes = [ctypes.c_void_p, _DP, _I, _I]
        L.vmc_set_random_state.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64]
        L.vmc_get_random_state.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64),
                                           ctypes.POINTER(ctypes.c_uint64)]
        L.vmc_run.argtypes = [ctypes.c_void_p, _I, _I, _I, _I, ctypes.POINTER(VmcSums), _DP, _DP]
        L.vmc_run_recorded.argtypes = [ctypes.c_void_p, _I, _I, _I, _I, ctypes.POINTER(VmcSums),
                                       _DP, _DP, _I, _FP]
        L.vmc_evaluate.argtypes = [ctypes.c_void_p, _DP, _I, _IP, _DP, _DP, _DP]
        L.vmc_evaluate_elements.argtypes = [ctypes.c_void_p, _DP, _I, _IP, _DP, _DP, _DP, _DP,
                                            _DP, _DP]
        L.vmc_draws.argtypes = [_I, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64, _I, _I,
                                _IP, _DP, _DP, _DP]
        L.vmc_lp_set.argtypes = [ctypes.c_void_p, _FP, _I, ctypes.POINTER(LpHyper)]
        L.vmc_lp_get.argtypes = [ctypes.c_void_p, _FP, _I]
        L.vmc_lp_gradient.argtypes = [ctypes.c_void_p, _DP, _I, _IP, _DP, _DP, _DP]
        L.vmc_lp_train.argtypes = [ctypes.c_void_p, _I, _I, _D, _DP]
        L.vmc_run_instant.argtypes = [ctypes.c_void_p, _I, _I, _I, _I, ctypes.POINTER(VmcSums),
                                      _DP, _DP, _I, _FP]
        L.vmc_run_blocked.argtypes = [ctypes.c_void_p, _I, _I, _I, _I, ctypes.POINTER(VmcSums),
                                      _DP, _DP, _I, _I, _DP]
        L.vmc_set_density.argtypes = [ctypes.c_void_p, _I, _I, _D]
        _I64P = ctypes.POINTER(ctypes.c_int64)
        L.vmc_get_density.argtypes = [ctypes.c_void_p, _I64P, _I64P, _I64P]
        L.vmc_optimize.argtypes = [ctypes.c_void_p, _DP, _I, _I, _I,
                                   ctypes.POINTER(VmcOptimizerConfig), _I, _DP, _DP]
        L.vmc_optimize_static.argtypes = [ctypes.c_void_p, _DP, _I, _I, _I,
                                          ctypes.POINTER(VmcOptimizerConfig), _DP, _DP]
        L.vmc_batch_create.argtypes = [ctypes.POINTER(VmcBatchConfig),
                                       ctypes.POINTER(ctypes.c_void_p)]
        L.vmc_batch_destroy.argtypes = [ctypes.c_void_p]
        L.vmc_batch_destroy.restype = None
        L.vmc_batch_init_walkers.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
        for n in ["vmc_batch_set_positions", "vmc_batch_get_positions"]:
            getattr(L, n).argtypes = [ctypes.c_void_p, _I, _DP, _I, _I]
        L.vmc_batch_get_parameters.argtypes = [ctypes.c_void_p, _DP]
        L.vmc_batch_set_shape.argtypes = [ctypes.c_void_p, _I, _I]
        L.vmc_batch_autotune.argtypes = [ctypes.c_void_p, _I, _IP, _DP]
        L.vmc_batch_set_launch.argtypes = [ctypes.c_void_p, _I]
        L.vmc_batch_get_launch.argtypes = [ctypes.c_void_p, _IP]
        L.vmc_batch_run.argtypes = [ctypes.c_void_p, _I, _I, _I, _I, _D, _I, _DP, _DP]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(_DP)


def make_config(num_particles, omega=1.0, interaction=True, kinds=(SLATER, GAUSSIAN, PADE),
                sampling=IMPORTANCE, step_length=0.01, quirks=QUIRKS_REFERENCE, precision=64,
                refresh_interval=0, num_walkers=1024, num_hidden=None, sigma=None,
                sorting=False, num_dimensions=2, drift_limit=0.0, rbm_matrix=RBM_AUTO):
    c = VmcConfig()
    c.drift_limit = drift_limit
    c.rbm_matrix = rbm_matrix
    c.num_particles = num_particles
    c.num_dimensions = num_dimensions
    c.omega = omega
    c.interaction = int(bool(interaction))
    c.num_elements = len(kinds)
    for i, k in enumerate(kinds):
        c.kinds[i] = k
    c.num_hidden = num_particles if num_hidden is None else num_hidden
    c.sigma = (1.0 / np.sqrt(omega)) if sigma is None else sigma
    c.sorting = int(sorting)
    c.sampling = sampling
    c.step_length = step_length
    c.quirks = quirks
    c.precision = precision
    c.refresh_interval = refresh_interval
    c.num_walkers = num_walkers
    return c


def max_parameters(cfg):
    return lib().vmc_max_parameters(ctypes.byref(cfg))


def device_count():
    return lib().vmc_device_count()


def draws(precision, seed, walker, counter0, n, dof):
    """The walk's random draws (coordinate, normal, move uniform, accept uniform)."""
    k = np.zeros(n, dtype=np.int32)
    z, um, ua = np.zeros(n), np.zeros(n), np.zeros(n)
    L = lib()
    rc = L.vmc_draws(precision, seed, walker, counter0, n, dof, k.ctypes.data_as(_IP), _dp(z),
                     _dp(um), _dp(ua))
    if rc != 0:
        raise VmcError(f"vmc error {rc}: {L.vmc_last_error().decode()}")
    return k, z, um, ua


class Engine:
    """One device engine: W walkers of one system (include/vmc.h)."""

    def __init__(self, cfg):
        self.cfg = cfg
        self.L = lib()
        h = ctypes.c_void_p()
        self._check(self.L.vmc_engine_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h
        self.nd = cfg.num_particles * cfg.num_dimensions
        self.W = cfg.num_walkers
        self.maxpar = max_parameters(cfg)
        self.E = cfg.num_elements

    def _check(self, rc):
        if rc != 0:
            raise VmcError(f"vmc error {rc}: {self.L.vmc_last_error().decode()}")

    def close(self):
        if getattr(self, "h", None):
            self.L.vmc_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_parameters(self, params):
        p = np.ascontiguousarray(params, dtype=np.float64).reshape(self.E, self.maxpar)
        self._check(self.L.vmc_set_parameters(self.h, _dp(p), self.E, self.maxpar))

    def init_walkers(self, seed):
        self._check(self.L.vmc_init_walkers(self.h, seed))

    def set_positions(self, pos):
        p = np.ascontiguousarray(pos, dtype=np.float64).reshape(self.W, self.nd)
        self._check(self.L.vmc_set_positions(self.h, _dp(p), self.W, self.nd))

    def positions(self):
        out = np.zeros((self.W, self.nd))
        self._check(self.L.vmc_get_positions(self.h, _dp(out), self.W, self.nd))
        return out

    def set_quantum_force(self, qf):
        q = np.ascontiguousarray(qf, dtype=np.float64).reshape(self.W, self.nd)
        self._check(self.L.vmc_set_quantum_force(self.h, _dp(q), self.W, self.nd))

    def quantum_force(self):
        out = np.zeros((self.W, self.nd))
        self._check(self.L.vmc_get_quantum_force(self.h, _dp(out), self.W, self.nd))
        return out

    def set_random_state(self, seed, counter=0):
        self._check(self.L.vmc_set_random_state(self.h, seed, counter))

    def random_state(self):
        s, c = ctypes.c_uint64(), ctypes.c_uint64()
        self._check(self.L.vmc_get_random_state(self.h, ctypes.byref(s), ctypes.byref(c)))
        return s.value, c.value

    def optimize(self, params, iterations, steps, equilibration=0, kind="adam",
                 learning_rate=0.1, gamma=0.0, monotonic=0.0, batches=None, use_graph=True):
        """Optimisation iterations on the device (vmc_optimize).  Returns (final
        parameter matrix, history [iterations][OPT_OUT], device milliseconds);
        batches defaults to the reference's (10 for adam, 1 otherwise)."""
        k = {"adam": OPT_ADAM, "sgd": OPT_SGD, "gd": OPT_GD}[kind]
        oc = VmcOptimizerConfig(k, learning_rate, gamma, monotonic,
                                batches if batches else (10 if k == OPT_ADAM else 1))
        p = np.ascontiguousarray(params, dtype=np.float64).reshape(self.E, self.maxpar).copy()
        hist = np.zeros((iterations, OPT_OUT))
        ms = ctypes.c_double(0)
        self._check(self.L.vmc_optimize(self.h, _dp(p), iterations, steps, equilibration,
                                        ctypes.byref(oc), int(bool(use_graph)), _dp(hist),
                                        ctypes.byref(ms)))
        return p, hist, ms.value

    def optimize_static(self, params, iterations, steps, equilibration=0, kind="adam",
                        learning_rate=0.1, gamma=0.0, monotonic=0.0, batches=None):
        """As optimize(use_graph=False) through vmc_optimize_static, which
        allocates no device memory (its loop state lives in a per-GPU array)."""
        k = {"adam": OPT_ADAM, "sgd": OPT_SGD, "gd": OPT_GD}[kind]
        oc = VmcOptimizerConfig(k, learning_rate, gamma, monotonic,
                                batches if batches else (10 if k == OPT_ADAM else 1))
        p = np.ascontiguousarray(params, dtype=np.float64).reshape(self.E, self.maxpar).copy()
        hist = np.zeros((iterations, OPT_OUT))
        ms = ctypes.c_double(0)
        self._check(self.L.vmc_optimize_static(self.h, _dp(p), iterations, steps, equilibration,
                                               ctypes.byref(oc), _dp(hist), ctypes.byref(ms)))
        return p, hist, ms.value

    def run(self, steps, equilibration=0, iteration=0, batches=1, record_walkers=0):
        sums = VmcSums()
        g = np.zeros(max(self.E * self.maxpar, 1))
        ge = np.zeros_like(g)
        if record_walkers:
            en = np.zeros((record_walkers, steps), dtype=np.float32)
            self._check(self.L.vmc_run_recorded(self.h, steps, equilibration, iteration, batches,
                                                ctypes.byref(sums), _dp(g), _dp(ge),
                                                record_walkers,
                                                en.ctypes.data_as(_FP)))
        else:
            en = None
            self._check(self.L.vmc_run(self.h, steps, equilibration, iteration, batches,
                                       ctypes.byref(sums), _dp(g), _dp(ge)))
        n = self.E * self.maxpar
        return (sums.as_dict(), g[:n].reshape(self.E, self.maxpar),
                ge[:n].reshape(self.E, self.maxpar), en)

    def run_blocked(self, steps, block, equilibration=0, iteration=0, batches=1,
                    record_walkers=None):
        """vmc_run_blocked: (sums, g, ge, means [record_walkers][steps // block])
        with means the FP64 averages of `block` consecutive sampled local
        energies of each recorded walker (all walkers by default)."""
        W = self.cfg.num_walkers if record_walkers is None else record_walkers
        sums = VmcSums()
        g = np.zeros(max(self.E * self.maxpar, 1))
        ge = np.zeros_like(g)
        means = np.zeros((W, steps // block), dtype=np.float64)
        self._check(self.L.vmc_run_blocked(self.h, steps, equilibration, iteration, batches,
                                           ctypes.byref(sums), _dp(g), _dp(ge), W, block,
                                           _dp(means)))
        n = self.E * self.maxpar
        return (sums.as_dict(), g[:n].reshape(self.E, self.maxpar),
                ge[:n].reshape(self.E, self.maxpar), means)

    def run_instant(self, steps, equilibration=0, iteration=0, batches=1, record_walkers=1):
        """vmc_run_instant: (sums, g, ge, out [4][record_walkers][steps]) with
        out = local energy, kinetic, external, interaction (FP32)."""
        sums = VmcSums()
        g = np.zeros(max(self.E * self.maxpar, 1))
        ge = np.zeros_like(g)
        out = np.zeros((4, record_walkers, steps), dtype=np.float32)
        self._check(self.L.vmc_run_instant(self.h, steps, equilibration, iteration, batches,
                                           ctypes.byref(sums), _dp(g), _dp(ge), record_walkers,
                                           out.ctypes.data_as(_FP)))
        n = self.E * self.maxpar
        return (sums.as_dict(), g[:n].reshape(self.E, self.maxpar),
                ge[:n].reshape(self.E, self.maxpar), out)

    def set_density(self, flags, bins=1000, max_radius=5.0):
        self._check(self.L.vmc_set_density(self.h, flags, bins, max_radius))
        self._dens_bins = bins

    def density(self):
        """(radial [bins], spatial [bins, bins], two_body [bins, bins]) int64 counts."""
        b = getattr(self, "_dens_bins", 1)
        r = np.zeros(b, dtype=np.int64)
        sp = np.zeros((b, b), dtype=np.int64)
        tb = np.zeros((b, b), dtype=np.int64)
        P = ctypes.POINTER(ctypes.c_int64)
        self._check(self.L.vmc_get_density(self.h, r.ctypes.data_as(P), sp.ctypes.data_as(P),
                                           tb.ctypes.data_as(P)))
        return r, sp, tb

    def evaluate(self, positions, moves, new_values):
        pos = np.ascontiguousarray(positions, dtype=np.float64)
        W = pos.shape[0]
        mv = np.ascontiguousarray(moves, dtype=np.int32)
        nv = np.ascontiguousarray(new_values, dtype=np.float64)
        width = 9 + self.nd
        o0 = np.zeros((W, width))
        o1 = np.zeros((W, width))
        self._check(self.L.vmc_evaluate(self.h, _dp(pos), W, mv.ctypes.data_as(_IP), _dp(nv),
                                        _dp(o0), _dp(o1)))
        return o0, o1

    # ---- learned proposal ----
    def lp_set(self, theta, sigma0=0.2, mu_max=1.0, ls_max=2.0):
        th = np.ascontiguousarray(theta, dtype=np.float32)
        assert th.size == LP_NPARAM
        hp = LpHyper(sigma0, mu_max, ls_max)
        self._check(self.L.vmc_lp_set(self.h, th.ctypes.data_as(_FP), LP_NPARAM, ctypes.byref(hp)))

    def lp_get(self):
        th = np.zeros(LP_NPARAM, dtype=np.float32)
        self._check(self.L.vmc_lp_get(self.h, th.ctypes.data_as(_FP), LP_NPARAM))
        return th

    def lp_gradient(self, positions, electrons, normals):
        """Per-sample (J, A, log|psi'/psi|^2, x', y') and the summed dJ/dtheta."""
        pos = np.ascontiguousarray(positions, dtype=np.float64)
        W = pos.shape[0]
        el = np.ascontiguousarray(electrons, dtype=np.int32)
        nr = np.ascontiguousarray(normals, dtype=np.float64).reshape(W, 2)
        per = np.zeros((W, 5))
        g = np.zeros(LP_NPARAM)
        self._check(self.L.vmc_lp_gradient(self.h, _dp(pos), W, el.ctypes.data_as(_IP), _dp(nr),
                                           _dp(per), _dp(g)))
        return per, g

    def lp_train(self, iterations, walk_steps, learning_rate):
        hist = np.zeros((max(iterations, 1), 3))
        self._check(self.L.vmc_lp_train(self.h, iterations, walk_steps, learning_rate, _dp(hist)))
        return hist[:iterations]

    def evaluate_elements(self, positions, moves, new_values):
        """vmc_evaluate plus per-element ratios [W][E] of the forced move and the
        parameter-gradient matrices [W][E][maxpar] before and after it."""
        pos = np.ascontiguousarray(positions, dtype=np.float64)
        W = pos.shape[0]
        mv = np.ascontiguousarray(moves, dtype=np.int32)
        nv = np.ascontiguousarray(new_values, dtype=np.float64)
        width = 9 + self.nd
        o0, o1 = np.zeros((W, width)), np.zeros((W, width))
        r = np.zeros((W, self.E))
        g0 = np.zeros((W, self.E, max(self.maxpar, 1)))
        g1 = np.zeros_like(g0)
        self._check(self.L.vmc_evaluate_elements(self.h, _dp(pos), W, mv.ctypes.data_as(_IP),
                                                 _dp(nv), _dp(o0), _dp(o1), _dp(r), _dp(g0),
                                                 _dp(g1)))
        mp = self.maxpar
        return o0, o1, r, g0[:, :, :mp], g1[:, :, :mp]


class Batch:
    """Several systems walked side by side, one kernel each (vmc_batch_* in include/vmc.h).

    systems: list of dicts with num_particles, omega, num_walkers, step_length,
    alpha, gamma.  System s draws from seed + s and retraces an Engine of that
    system seeded seed + s."""

    def __init__(self, systems, kinds=(SLATER, GAUSSIAN, PADE), interaction=True,
                 sampling=IMPORTANCE_EXACT, quirks=QUIRKS_REFERENCE, precision=64,
                 refresh_interval=0, drift_limit=0.0):
        self.L = lib()
        self.n = len(systems)
        self._sys = (VmcBatchSystem * self.n)()
        for i, y in enumerate(systems):
            for k in ("num_particles", "omega", "num_walkers", "step_length", "alpha", "gamma"):
                setattr(self._sys[i], k, y[k])
        self.systems = [dict(y) for y in systems]
        c = VmcBatchConfig()
        c.num_systems = self.n
        c.systems = ctypes.cast(self._sys, ctypes.POINTER(VmcBatchSystem))
        c.num_elements = len(kinds)
        for i, k in enumerate(kinds):
            c.kinds[i] = k
        c.interaction = int(bool(interaction))
        c.sampling = sampling
        c.quirks = quirks
        c.precision = precision
        c.refresh_interval = refresh_interval
        c.drift_limit = drift_limit
        self.cfg = c
        h = ctypes.c_void_p()
        self._check(self.L.vmc_batch_create(ctypes.byref(c), ctypes.byref(h)))
        self.h = h

    def _check(self, rc):
        if rc != 0:
            raise VmcError(f"vmc error {rc}: {self.L.vmc_last_error().decode()}")

    def close(self):
        if getattr(self, "h", None):
            self.L.vmc_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def init_walkers(self, seed):
        self._check(self.L.vmc_batch_init_walkers(self.h, seed))

    def _shape(self, s):
        y = self.systems[s]
        return y["num_walkers"], 2 * y["num_particles"]

    def set_positions(self, s, pos):
        W, nd = self._shape(s)
        p = np.ascontiguousarray(pos, dtype=np.float64).reshape(W, nd)
        self._check(self.L.vmc_batch_set_positions(self.h, s, _dp(p), W, nd))

    def get_positions(self, s):
        W, nd = self._shape(s)
        p = np.zeros((W, nd))
        self._check(self.L.vmc_batch_get_positions(self.h, s, _dp(p), W, nd))
        return p

    def parameters(self):
        """[num_systems][2]: alpha, gamma after the last run."""
        ag = np.zeros((self.n, 2))
        self._check(self.L.vmc_batch_get_parameters(self.h, _dp(ag)))
        return ag

    def set_shape(self, s, lanes_per_walker):
        self._check(self.L.vmc_batch_set_shape(self.h, s, lanes_per_walker))

    def set_launch(self, launch):
        """BATCH_SIDE_BY_SIDE (one kernel per system per refresh segment) or
        BATCH_PERSISTENT (one persistent launch per iteration)."""
        self._check(self.L.vmc_batch_set_launch(self.h, int(launch)))

    def launch(self):
        v = ctypes.c_int(0)
        self._check(self.L.vmc_batch_get_launch(self.h, ctypes.byref(v)))
        return v.value

    def autotune(self, steps):
        """(lanes per walker, ns per walker-step) chosen for every system."""
        lanes = np.zeros(self.n, dtype=np.int32)
        ns = np.zeros(self.n)
        self._check(self.L.vmc_batch_autotune(self.h, steps, lanes.ctypes.data_as(_IP), _dp(ns)))
        return lanes, ns

    def run(self, iterations, steps, equilibration=0, batches=1, learning_rate=0.0,
            use_graph=True):
        """(out [iterations][num_systems][BATCH_OUT], device milliseconds)."""
        out = np.zeros((iterations, self.n, BATCH_OUT))
        ms = ctypes.c_double(0)
        self._check(self.L.vmc_batch_run(self.h, iterations, steps, equilibration, batches,
                                         learning_rate, int(bool(use_graph)), _dp(out),
                                         ctypes.byref(ms)))
        return out, ms.value
