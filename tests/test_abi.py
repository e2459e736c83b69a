"""The C ABI library loads without a GPU, exports everything include/vmc.h
declares, and rejects invalid configurations before touching a device."""
import ctypes

import pytest

import vmachine_amd as V
from vmachine_amd import engine as E


def test_library_exports_every_declared_symbol():
    L = E.lib()
    names = E.header_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), f"libvmc_hip.so does not export {n}"


def test_integration_index_lists_every_declared_symbol():
    """The maintainer's index of entry points covers both headers."""
    import os
    import re
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    index = open(os.path.join(root, "INTEGRATION.md")).read().split("## 4. Index of entry points")
    assert len(index) == 2, "INTEGRATION.md has no entry-point index"
    listed = set(re.findall(r"`(vmc_\w+)`", index[1]))
    host = open(os.path.join(root, "include", "vmc_host.h")).read()
    declared = set(E.header_functions()) | set(re.findall(r"\b(vmc_host_\w+)\s*\(", host))
    assert declared - listed == set(), sorted(declared - listed)


def test_parameter_shapes_match_reference():
    c = V.make_config(56, omega=0.1)
    L = E.lib()
    assert [L.vmc_num_parameters(ctypes.byref(c), e) for e in range(3)] == [0, 1, 1]
    assert L.vmc_max_parameters(ctypes.byref(c)) == 1
    c = V.make_config(12, kinds=[V.SLATER, V.RBM_GAUSSIAN, V.RBM_PRODUCT, V.PADE], num_hidden=24)
    assert [L.vmc_num_parameters(ctypes.byref(c), e) for e in range(4)] == [0, 24, 600, 1]
    assert L.vmc_max_parameters(ctypes.byref(c)) == 600


@pytest.mark.parametrize("kw,msg", [
    (dict(num_particles=8), "closed-shell"),
    (dict(num_particles=6, num_dimensions=3), "numDimensions"),
    (dict(num_particles=6, precision=16), "precision"),
    (dict(num_particles=6, num_walkers=0), "num_walkers"),
    (dict(num_particles=6, omega=-1.0), "omega"),
    (dict(num_particles=6, sampling=7), "sampling"),
    (dict(num_particles=6, rbm_matrix=9), "rbm_matrix"),
    # matrix-core RBM contractions: FP32 walks with an RBM-product element, K >= 4
    (dict(num_particles=12, kinds=[V.SLATER, V.RBM_GAUSSIAN, V.RBM_PRODUCT, V.PADE],
          num_hidden=24, precision=64, rbm_matrix=V.RBM_MFMA_F32), "matrix-core"),
    (dict(num_particles=6, kinds=[V.SLATER, V.GAUSSIAN, V.PADE], precision=32,
          rbm_matrix=V.RBM_MFMA_BF16), "matrix-core"),
    (dict(num_particles=2, kinds=[V.RBM_GAUSSIAN, V.RBM_PRODUCT], num_hidden=2, precision=32,
          rbm_matrix=V.RBM_MFMA_F32), "matrix-core"),
])
def test_invalid_configs_are_rejected(kw, msg):
    args = dict(num_particles=6)
    args.update(kw)
    n = args.pop("num_particles")
    c = V.make_config(n, **args)
    with pytest.raises(V.VmcError, match=msg):
        V.Engine(c)


def test_no_device_is_an_error_not_a_fallback():
    if V.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(V.VmcError, match="no HIP device"):
        V.Engine(V.make_config(6))


def _dots(*Ns):
    return [dict(num_particles=N, omega=1.0, num_walkers=64, step_length=0.01, alpha=1.0,
                 gamma=1.0) for N in Ns]


@pytest.mark.parametrize("systems,kw,msg", [
    (_dots(), {}, "systems"),
    (_dots(*([2] * 9)), {}, "systems"),
    (_dots(2, 8), {}, "system 1: slaterDeterminant needs a closed shell"),
    (_dots(2, 6), dict(kinds=(V.SLATER, V.RBM_GAUSSIAN)), "Slater, Gaussian and Pade"),
    (_dots(2, 6), dict(kinds=(V.GAUSSIAN, V.GAUSSIAN)), "twice"),
    (_dots(2, 6), dict(sampling=V.LEARNED), "brute force or Langevin"),
    (_dots(2, 6), dict(precision=16), "precision"),
    (_dots(2, 6), dict(drift_limit=-1.0), "drift_limit"),
    ([dict(num_particles=57, omega=1.0, num_walkers=8, step_length=0.01, alpha=1.0, gamma=1.0)],
     dict(kinds=(V.GAUSSIAN, V.PADE)), "numParticles"),
    ([dict(num_particles=6, omega=0.0, num_walkers=8, step_length=0.01, alpha=1.0, gamma=1.0)],
     {}, "system 0: num_walkers"),
])
def test_invalid_batches_are_rejected(systems, kw, msg):
    with pytest.raises(V.VmcError, match=msg):
        V.Batch(systems, **kw)


def test_batch_without_device_is_an_error():
    if V.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(V.VmcError, match="no HIP device"):
        V.Batch(_dots(2, 6, 12))
