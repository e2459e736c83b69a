"""ctypes binding of libvmc_host.so (include/vmc_host.h) and the `vmachine`
executable: the input.in parser, optimisers, initial weights and automated
blocking that drive the engine the way the reference's System/Sampler drive
their plugins.  All of it is host code; only `vmachine` runs touch the GPU.
"""
import ctypes
import json
import os
import re
import subprocess

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
HOST_LIB = os.path.join(PKG, "_lib", "libvmc_host.so")
EXE = os.path.join(PKG, "_lib", "vmachine")
HEADER = os.path.join(ROOT, "include", "vmc_host.h")

_D, _I = ctypes.c_double, ctypes.c_int
_PD = ctypes.POINTER(ctypes.c_double)
_lib = None


class HostError(RuntimeError):
    pass


def header_functions():
    """Names of the functions include/vmc_host.h declares."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(vmc_host_\w+)\s*\(", text)))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(HOST_LIB):
            raise HostError(f"{HOST_LIB} is not built: run `python -m vmachine_amd.build`")
        L = ctypes.CDLL(HOST_LIB)
        L.vmc_host_blocking.argtypes = [_PD, ctypes.c_int64, _PD, _PD]
        L.vmc_host_optimizer_create.argtypes = [ctypes.c_char_p, _I, _D, _D, _D]
        L.vmc_host_optimizer_create.restype = ctypes.c_void_p
        L.vmc_host_optimizer_batches.argtypes = [ctypes.c_void_p]
        L.vmc_host_optimizer_update.argtypes = [ctypes.c_void_p, _D, _PD, _PD, _PD]
        L.vmc_host_optimizer_update.restype = None
        L.vmc_host_optimizer_destroy.argtypes = [ctypes.c_void_p]
        L.vmc_host_optimizer_destroy.restype = None
        L.vmc_host_parse.argtypes = [ctypes.c_char_p, ctypes.c_char_p, _I]
        L.vmc_host_initial_weights.argtypes = [ctypes.c_char_p, ctypes.c_uint64, _PD, _I, _I]
        L.vmc_host_main.argtypes = [_I, ctypes.POINTER(ctypes.c_char_p)]
        _lib = L
    return _lib


def _pd(a):
    return a.ctypes.data_as(_PD)


def blocking(x):
    """(mean, standard error) of a 2^d-long series by automated blocking."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    m, s = _D(), _D()
    if lib().vmc_host_blocking(_pd(x), x.size, ctypes.byref(m), ctypes.byref(s)) != 0:
        raise HostError("blocking needs a series of length 2^d, d >= 1")
    return m.value, s.value


class Optimizer:
    """adam / sgd / gd over a flattened elements x max-parameters matrix."""

    def __init__(self, kind, size, learning_rate, gamma=0.0, monotonic=0.0):
        self.size = size
        self._h = lib().vmc_host_optimizer_create(kind.encode(), size, learning_rate, gamma,
                                                  monotonic)
        if not self._h:
            raise HostError(f"optimizer '{kind}' does not exist")

    @property
    def batches(self):
        return lib().vmc_host_optimizer_batches(self._h)

    def update(self, avg_energy, avg_grad, avg_grad_e):
        g = np.ascontiguousarray(avg_grad, dtype=np.float64).ravel()
        ge = np.ascontiguousarray(avg_grad_e, dtype=np.float64).ravel()
        assert g.size == self.size and ge.size == self.size
        out = np.zeros(self.size)
        lib().vmc_host_optimizer_update(self._h, avg_energy, _pd(g), _pd(ge), _pd(out))
        return out

    def __del__(self):
        if getattr(self, "_h", None):
            lib().vmc_host_optimizer_destroy(self._h)
            self._h = None


def parse(path):
    """The resolved configuration of an input.in as a dict; HostError (with the
    reference's message) on a configuration error."""
    buf = ctypes.create_string_buffer(1 << 16)
    rc = lib().vmc_host_parse(str(path).encode(), buf, len(buf))
    if rc != 0:
        raise HostError(buf.value.decode())
    return json.loads(buf.value.decode())


def initial_weights(path, rows, cols, seed=1):
    out = np.zeros((rows, cols))
    rc = lib().vmc_host_initial_weights(str(path).encode(), seed, _pd(out), rows, cols)
    if rc != 0:
        raise HostError(f"initial weights for {path}: status {rc}")
    return out


def run(args, timeout=600, cwd=None):
    """Runs the `vmachine` executable; returns the CompletedProcess."""
    if not os.path.exists(EXE):
        raise HostError(f"{EXE} is not built: run `python -m vmachine_amd.build`")
    return subprocess.run([EXE] + [str(a) for a in args], stdout=subprocess.PIPE,
                          stderr=subprocess.PIPE, text=True, timeout=timeout, cwd=cwd)


def parse_output(text):
    """The iteration table and final block of a `vmachine` run's output."""
    rows, final = [], {}
    names = {"Total": "energy", "Kinetic": "kinetic", "External": "external",
             "Interaction": "interaction"}
    for line in text.splitlines():
        parts = line.split()
        if len(parts) == 11 and parts[0].isdigit():
            rows.append([int(parts[0])] + [float(p) for p in parts[1:]])
            continue
        m = re.match(r"\s*(Total|Kinetic|External|Interaction) energy\s*:\s*(\S+) \(with STD = (\S+)\)",
                     line)
        if m:
            final[names[m.group(1)]] = float(m.group(2))
            final[names[m.group(1)] + "_std"] = float(m.group(3))
        m = re.match(r"\s*Acceptence Ratio\s*:\s*(\S+)", line)
        if m:
            final["acceptance"] = float(m.group(1))
        m = re.match(r"Average CPU time before convergence:\s*(\S+)", line)
        if m:
            final["cpu_time_per_iteration"] = float(m.group(1))
    return rows, final
