"""dualres_amd: MI355X-native engine for the HMC sampler of the
dual-resolution Gaussian-process image model (asw221/dualres).

The compute path is the C-ABI library ``libdualres_hip.so`` (hand-written
HIP kernels for gfx950) declared in ``include/dualres_hip.h``; this package
is a thin ctypes binding plus input builders.  The C++ host layer
(``dualres_amd/host``) provides the command-line tools.
"""
from ._lib import DrDomainError, DrError, load  # noqa: F401
from .engine import Chain, pow2_embedding  # noqa: F401
