"""MI355X (gfx950) drop-in for PowerSpectra.jl's mode-coupling and covariance
hot path.  The compute path is the HIP library libpsx_hip.so (C ABI in
include/powerspectra_hip.h); this package mirrors the reference's Julia API
on top of it.
"""
from . import _abi
from ._abi import (ArgumentError, AxisError, BudgetError, Context, DeviceError,
                   DomainError, PowerSpectraError, SingularException, default_context)

__all__ = [
    "ArgumentError", "AxisError", "BudgetError", "Context", "DeviceError", "DomainError",
    "PowerSpectraError", "SingularException", "default_context",
]
