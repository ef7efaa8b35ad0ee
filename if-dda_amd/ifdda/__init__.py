"""Host-side Python mirror of IF-DDA's FFT-accelerated DDA solve on MI355X.

The compute lives in ``libifdda_hip.so`` (HIP for gfx950); this package only
binds its C ABI (``include/ifdda.h``) with ctypes.  Names follow the cdmlib
routines they replace: ``greencalculfft``, ``produitfftmatvect3``,
``inverserig``.  There is no CPU fallback: loading fails loudly if the HIP
library is missing.
"""
from .capi import (  # noqa: F401
    GPBICG1, QMRBICGSTAB1, QMRBICGSTAB2, TFQMR, METHODS, IfddaError, Runner, Solver, lib, lib_path,
    solve_batch, solve_batch_release,
)
