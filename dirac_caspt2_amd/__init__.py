"""MI355X-native CASCI/CASPT2 hot path (DIRAC-CASPT2 compatible), HIP on gfx950.

* ``csrc/``      hand-written HIP kernels + the C ABI (include/dcaspt2_amd.h)
* ``_lib``       ctypes binding of libdcaspt2_amd.so (no CPU fallback)
* ``api``        host-side mirror of the reference driver (Caspt2Context)
* ``problem``    loading integrals (seeded synthetic factors, dense blocks)
* ``synth``      the seeded Kramers-paired workload generator
* ``fortran/``   iso_c_binding interface module for a Fortran host
"""
from .synth import CONFIG_FP64, CONFIG_SMALL, CONFIG_SPLIT, SynthSpec, generate  # noqa: F401

__all__ = ["SynthSpec", "generate", "CONFIG_SMALL", "CONFIG_FP64", "CONFIG_SPLIT"]
