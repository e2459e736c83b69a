"""ctypes binding of the FP64 C oracle (oracle/vmc_oracle.c).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by the smoke check in
__graft_entry__.py and by the cpu_baseline leg of bench.py -- always as the
checker, never as the product path.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libvmc_oracle.so")

GAUSSIAN, SLATER, PADE, RBM_GAUSSIAN, RBM_PRODUCT = 1, 2, 3, 4, 5
BRUTE_FORCE, IMPORTANCE_COMPAT, IMPORTANCE_EXACT = 0, 1, 2
Q2, Q3, Q6, Q13 = 1 << 2, 1 << 3, 1 << 6, 1 << 13
QUIRKS_ALL = Q2 | Q3 | Q6 | Q13
NACC = 10
ACC_NAMES = ["E", "E2", "T", "Vext", "Vint", "T2", "Vext2", "Vint2", "accept", "nsamples"]

_D = ctypes.c_double
_I = ctypes.c_int
_P = ctypes.c_void_p
_DP = ctypes.POINTER(ctypes.c_double)
_U64 = ctypes.c_uint64


class Config(ctypes.Structure):
    _fields_ = [
        ("N", _I), ("D", _I), ("omega", _D), ("interaction", _I),
        ("num_elements", _I), ("kinds", _I * 6), ("num_hidden", _I),
        ("sigma", _D), ("sorting", _I), ("sampling", _I), ("step_length", _D),
        ("quirks", ctypes.c_uint), ("basis_scaled", _I), ("drift_limit", _D),
    ]


def build():
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


_lib = None

# learned proposal: network sizes and parameter offsets (vmc_oracle.h)
LP_IN, LP_HID, LP_OUT = 18, 64, 4
LP_W1 = 0
LP_B1 = LP_W1 + LP_HID * LP_IN
LP_W2 = LP_B1 + LP_HID
LP_B2 = LP_W2 + LP_HID * LP_HID
LP_W3 = LP_B2 + LP_HID
LP_B3 = LP_W3 + LP_OUT * LP_HID
LP_NPARAM = LP_B3 + LP_OUT


class LpHyper(ctypes.Structure):
    _fields_ = [("sigma0", ctypes.c_double), ("mu_max", ctypes.c_double),
                ("ls_max", ctypes.c_double)]


def lp_hyper(sigma0=0.2, mu_max=1.0, ls_max=2.0):
    return LpHyper(sigma0, mu_max, ls_max)


def lp_init(rng, zero_head=True):
    """Glorot-uniform layers 1 and 2; the head zero (an untrained proposal is a
    symmetric Gaussian random walk of width sigma0)."""
    th = np.zeros(LP_NPARAM)
    a1 = np.sqrt(6.0 / (LP_IN + LP_HID))
    a2 = np.sqrt(6.0 / (2 * LP_HID))
    th[LP_W1:LP_B1] = rng.uniform(-a1, a1, LP_HID * LP_IN)
    th[LP_W2:LP_B2] = rng.uniform(-a2, a2, LP_HID * LP_HID)
    if not zero_head:
        a3 = np.sqrt(6.0 / (LP_HID + LP_OUT))
        th[LP_W3:LP_B3] = rng.uniform(-a3, a3, LP_OUT * LP_HID)
    return th


def lp_features(cfg, x, i, xi, yi):
    f = np.zeros(LP_IN)
    nn = lib().vmco_lp_features(ctypes.byref(cfg), _dp(np.ascontiguousarray(x, dtype=np.float64)),
                                int(i), float(xi), float(yi), _dp(f))
    return f, nn


def lp_forward(theta, f):
    h1, h2, o = np.zeros(LP_HID), np.zeros(LP_HID), np.zeros(LP_OUT)
    lib().vmco_lp_forward(_dp(np.ascontiguousarray(theta, dtype=np.float64)),
                          _dp(np.ascontiguousarray(f, dtype=np.float64)), _dp(h1), _dp(h2), _dp(o))
    return h1, h2, o


def lp_head(hp, omega, o):
    mu, sig = np.zeros(2), np.zeros(2)
    lib().vmco_lp_head(ctypes.byref(hp), omega, _dp(np.ascontiguousarray(o, dtype=np.float64)),
                       _dp(mu), _dp(sig))
    return mu, sig


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.vmco_walker_new.restype = _P
        L.vmco_walker_new.argtypes = [ctypes.POINTER(Config)]
        L.vmco_walker_free.argtypes = [_P]
        for name in ["vmco_set_parameters", "vmco_init_positions", "vmco_get_positions",
                     "vmco_get_quantum_force", "vmco_set_quantum_force"]:
            getattr(L, name).argtypes = [_P, _DP]
        L.vmco_reset_quantum_force.argtypes = [_P]
        L.vmco_propose.argtypes = [_P, _I, _D]
        L.vmco_reject.argtypes = [_P]
        L.vmco_accept.argtypes = [_P]
        L.vmco_element_ratio.argtypes = [_P, _I]
        L.vmco_element_ratio.restype = _D
        L.vmco_element_gradient.argtypes = [_P, _I, _I]
        L.vmco_element_gradient.restype = _D
        L.vmco_element_laplacian.argtypes = [_P, _I]
        L.vmco_element_laplacian.restype = _D
        L.vmco_element_parameter_gradient.argtypes = [_P, _I, _DP]
        for name in ["vmco_probability_ratio", "vmco_kinetic_energy", "vmco_external_energy",
                     "vmco_interaction_energy"]:
            getattr(L, name).argtypes = [_P]
            getattr(L, name).restype = _D
        L.vmco_step.argtypes = [_P, _I, _D, _D, _D]
        L.vmco_step.restype = _I
        L.vmco_run.argtypes = [_P, _U64, ctypes.c_uint32, _U64, _I, _I, _I, _I, _DP, _DP, _DP, _DP]
        L.vmco_run.restype = _U64
        L.vmco_run_many.argtypes = [ctypes.POINTER(Config), _DP, _I, _DP, _DP, _U64,
                                    ctypes.POINTER(_U64), _I, _I, _I, _I, _I, _DP, _DP, _DP]
        L.vmco_philox4x32_10.argtypes = [ctypes.POINTER(ctypes.c_uint32)] * 3
        L.vmco_step_draws.argtypes = [_U64, ctypes.c_uint32, _U64, _I, ctypes.POINTER(_I),
                                      _DP, _DP, _DP]
        L.vmco_initial_positions.argtypes = [_U64, ctypes.c_uint32, _I, _D, _DP]
        L.vmco_blocking.argtypes = [_DP, _U64, _DP, _DP, ctypes.POINTER(_U64)]
        L.vmco_blocking.restype = _I
        L.vmco_chi2_quantile95.argtypes = [_I]
        L.vmco_chi2_quantile95.restype = _D
        L.vmco_num_orbitals.argtypes = [_I, _I]
        L.vmco_num_orbitals.restype = _I
        L.vmco_orbital_list.argtypes = [_I, _I, ctypes.POINTER(_I)]
        L.vmco_num_parameters.argtypes = [ctypes.POINTER(Config), _I]
        L.vmco_num_parameters.restype = _I
        L.vmco_max_parameters.argtypes = [ctypes.POINTER(Config)]
        L.vmco_max_parameters.restype = _I
        for name in ["vmco_basis_element"]:
            getattr(L, name).argtypes = [_P, _I, _DP]
            getattr(L, name).restype = _D
        for name in ["vmco_basis_element_der", "vmco_basis_element_secder"]:
            getattr(L, name).argtypes = [_P, _I, _I, _DP]
            getattr(L, name).restype = _D
        L.vmco_lp_features.argtypes = [ctypes.POINTER(Config), _DP, _I, _D, _D, _DP]
        L.vmco_lp_features.restype = _I
        L.vmco_lp_forward.argtypes = [_DP, _DP, _DP, _DP, _DP]
        L.vmco_lp_head.argtypes = [ctypes.POINTER(LpHyper), _D, _DP, _DP, _DP]
        L.vmco_lp_step.argtypes = [_P, _DP, ctypes.POINTER(LpHyper), _I, _D, _D, _D]
        L.vmco_lp_step.restype = _I
        L.vmco_lp_sample_grad.argtypes = [_P, _DP, ctypes.POINTER(LpHyper), _I, _D, _D, _DP, _DP]
        L.vmco_lp_sample_grad.restype = _D
        _I64P = ctypes.POINTER(ctypes.c_int64)
        L.vmco_density_add.argtypes = [_DP, _I, _I, _D, _I64P, _I64P, _I64P]
        L.vmco_density_add.restype = None
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(_DP)


def make_config(N, D=2, omega=1.0, interaction=1, kinds=(SLATER, GAUSSIAN, PADE),
                num_hidden=None, sigma=None, sorting=False, sampling=IMPORTANCE_COMPAT,
                step_length=0.01, quirks=QUIRKS_ALL, basis_scaled=False, drift_limit=0.0):
    c = Config()
    c.drift_limit = drift_limit
    c.N, c.D, c.omega, c.interaction = N, D, omega, int(interaction)
    c.num_elements = len(kinds)
    for i, k in enumerate(kinds):
        c.kinds[i] = k
    c.num_hidden = N if num_hidden is None else num_hidden
    c.sigma = (1.0 / np.sqrt(omega)) if sigma is None else sigma
    c.sorting = int(sorting)
    c.sampling = sampling
    c.step_length = step_length
    c.quirks = quirks
    c.basis_scaled = int(basis_scaled)
    return c


def max_parameters(cfg):
    return lib().vmco_max_parameters(ctypes.byref(cfg))


def num_parameters(cfg, e):
    return lib().vmco_num_parameters(ctypes.byref(cfg), e)


class Walker:
    """One walker of the restatement: the reference's elements for one chain."""

    def __init__(self, cfg, params, positions):
        self.cfg = cfg
        self.L = lib()
        self.h = self.L.vmco_walker_new(ctypes.byref(cfg))
        self.nd = cfg.N * cfg.D
        self.maxpar = max_parameters(cfg)
        self.set_parameters(params)
        self.init_positions(positions)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.vmco_walker_free(self.h)
            self.h = None

    def set_parameters(self, params):
        p = np.ascontiguousarray(params, dtype=np.float64).reshape(self.cfg.num_elements, max(self.maxpar, 1))
        self.L.vmco_set_parameters(self.h, _dp(p))

    def init_positions(self, positions):
        x = np.ascontiguousarray(positions, dtype=np.float64)
        assert x.size == self.nd
        self.L.vmco_init_positions(self.h, _dp(x))

    def positions(self):
        out = np.zeros(self.nd)
        self.L.vmco_get_positions(self.h, _dp(out))
        return out

    def quantum_force(self):
        out = np.zeros(self.nd)
        self.L.vmco_get_quantum_force(self.h, _dp(out))
        return out

    def set_quantum_force(self, qf):
        q = np.ascontiguousarray(qf, dtype=np.float64)
        self.L.vmco_set_quantum_force(self.h, _dp(q))

    def propose(self, k, newval):
        self.L.vmco_propose(self.h, int(k), float(newval))

    def reject(self):
        self.L.vmco_reject(self.h)

    def accept(self):
        self.L.vmco_accept(self.h)

    def ratio(self, e=None):
        if e is None:
            return self.L.vmco_probability_ratio(self.h)
        return self.L.vmco_element_ratio(self.h, e)

    def gradient(self, e):
        return np.array([self.L.vmco_element_gradient(self.h, e, k) for k in range(self.nd)])

    def laplacian(self, e):
        return self.L.vmco_element_laplacian(self.h, e)

    def parameter_gradient(self, e):
        out = np.zeros(max(self.maxpar, 1))
        self.L.vmco_element_parameter_gradient(self.h, e, _dp(out))
        return out[: self.maxpar]

    def kinetic(self):
        return self.L.vmco_kinetic_energy(self.h)

    def external(self):
        return self.L.vmco_external_energy(self.h)

    def interaction(self):
        return self.L.vmco_interaction_energy(self.h)

    def local_energy(self):
        return self.kinetic() + self.external() + self.interaction()

    def drift(self):
        g = sum(self.gradient(e) for e in range(self.cfg.num_elements))
        return 2.0 * g

    def step(self, k, z, u_move, u_acc):
        return bool(self.L.vmco_step(self.h, int(k), float(z), float(u_move), float(u_acc)))

    def lp_step(self, theta, hp, i, xi0, xi1, ua):
        th = np.ascontiguousarray(theta, dtype=np.float64)
        return bool(self.L.vmco_lp_step(self.h, _dp(th), ctypes.byref(hp), int(i), float(xi0),
                                        float(xi1), float(ua)))

    def lp_sample_grad(self, theta, hp, i, xi0, xi1, grad=None):
        """J = A |dx|^2 omega of one proposal; adds dJ/dtheta into grad.
        Returns (J, A, log|psi'/psi|^2, x', y', grad)."""
        th = np.ascontiguousarray(theta, dtype=np.float64)
        g = np.zeros(LP_NPARAM) if grad is None else grad
        out = np.zeros(6)
        self.L.vmco_lp_sample_grad(self.h, _dp(th), ctypes.byref(hp), int(i), float(xi0),
                                   float(xi1), _dp(g), _dp(out))
        self.last_L = out[5]
        return out[0], out[1], out[2], out[3], out[4], g

    def run(self, seed, walker, counter, steps, equil, it=0, batches=1, record=False):
        E = self.cfg.num_elements
        acc = np.zeros(NACC)
        g = np.zeros(E * max(self.maxpar, 1))
        ge = np.zeros_like(g)
        en = np.zeros(max(steps, 1)) if record else None
        nxt = self.L.vmco_run(self.h, seed, walker, counter, steps, equil, it, batches,
                              _dp(acc), _dp(g), _dp(ge), _dp(en) if record else None)
        return nxt, acc, g, ge, en


def run_many(cfg, params, positions, seed, counters, steps, equil, it=0, batches=1,
             threads=8, qf=None):
    """Multithreaded restatement over W walkers; positions/counters/qf updated in place."""
    L = lib()
    W = positions.shape[0]
    mp = max(max_parameters(cfg), 1)
    E = cfg.num_elements
    p = np.ascontiguousarray(params, dtype=np.float64).reshape(E, mp)
    acc = np.zeros(NACC)
    g = np.zeros(E * mp)
    ge = np.zeros(E * mp)
    assert positions.dtype == np.float64 and positions.flags.c_contiguous
    assert counters.dtype == np.uint64
    L.vmco_run_many(ctypes.byref(cfg), _dp(p), W, _dp(positions),
                    _dp(qf) if qf is not None else None, seed,
                    counters.ctypes.data_as(ctypes.POINTER(_U64)), steps, equil, it, batches,
                    threads, _dp(acc), _dp(g), _dp(ge))
    return acc, g.reshape(E, mp), ge.reshape(E, mp)


def philox(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().vmco_philox4x32_10(c, k, o)
    return list(o)


def step_draws(seed, walker, counter, nd):
    k = _I()
    z, um, ua = _D(), _D(), _D()
    lib().vmco_step_draws(seed, walker, counter, nd, ctypes.byref(k), ctypes.byref(z),
                          ctypes.byref(um), ctypes.byref(ua))
    return k.value, z.value, um.value, ua.value


def initial_positions(seed, walker, nd, omega):
    out = np.zeros(nd)
    lib().vmco_initial_positions(seed, walker, nd, omega, _dp(out))
    return out


def blocking(x):
    a = np.array(x, dtype=np.float64, copy=True)
    mean, se = _D(), _D()
    nk = _U64()
    rc = lib().vmco_blocking(_dp(a), a.size, ctypes.byref(mean), ctypes.byref(se), ctypes.byref(nk))
    if rc != 0:
        raise ValueError("blocking needs a power-of-two length")
    return mean.value, se.value, nk.value


def density_add(x, N, bins, max_radius, radial=None, spatial=None, two_body=None):
    """Adds one configuration's counts to int64 histograms (radial [bins],
    spatial and two_body [bins, bins]) in place."""
    P = ctypes.POINTER(ctypes.c_int64)
    xx = np.ascontiguousarray(x, dtype=np.float64)
    ptr = [h.ctypes.data_as(P) if h is not None else None for h in (radial, spatial, two_body)]
    for h in (radial, spatial, two_body):
        assert h is None or (h.dtype == np.int64 and h.flags.c_contiguous)
    lib().vmco_density_add(_dp(xx), N, bins, max_radius, *ptr)


def chi2_quantile95(dof):
    return lib().vmco_chi2_quantile95(dof)


def num_orbitals(N, D=2):
    return lib().vmco_num_orbitals(N, D)


def orbital_list(n, D=2):
    out = (_I * (n * D))()
    lib().vmco_orbital_list(n, D, out)
    return np.array(list(out)).reshape(n, D)


class Optimizer:
    """Python mirror of vmco_optimizer for tests (same formulas as the C code)."""

    def __init__(self, kind, size, eta, gamma=0.0, monotonic_exp=0.0):
        self.kind, self.eta, self.gamma, self.mono = kind, eta, gamma, monotonic_exp
        self.m = np.ones(size)
        self.v = np.ones(size)
        self.step = 0

    def update(self, avgE, avgG, avgGE):
        self.step += 1
        g = 2.0 * (np.ravel(avgGE) - avgE * np.ravel(avgG))
        if self.kind == 0:
            b1, b2, eps = 0.89, 0.999, 1e-8
            self.m = b1 * self.m + (1 - b1) * g
            self.v = b2 * self.v + (1 - b2) * g * g
            mh = self.m / (1 - b1 ** self.step)
            vh = self.v / (1 - b2 ** self.step)
            return self.eta * mh / (np.sqrt(vh) + eps)
        self.v = self.gamma * self.v + self.eta * g / (self.step ** self.mono)
        return self.v.copy()
