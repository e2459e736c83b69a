"""MI355X-native VMC sampling engine for VMaChine's hot path.

The compute path is libvmc_hip.so (hand-written HIP kernels for gfx950 behind
the C ABI in include/vmc.h); vmachine_amd.engine is its ctypes binding.
"""
from .engine import (  # noqa: F401
    BATCH_PERSISTENT, BATCH_SIDE_BY_SIDE, BRUTE_FORCE, GAUSSIAN, IMPORTANCE, IMPORTANCE_EXACT, LEARNED, LP_NPARAM, PADE, QUIRKS_REFERENCE,
    RBM_AUTO, RBM_GAUSSIAN, RBM_LANES, RBM_MFMA_BF16, RBM_MFMA_F32, RBM_PRODUCT, SLATER, Batch,
    Engine, VmcError, device_count, make_config, max_parameters,
)
