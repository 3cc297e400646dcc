"""MI355X (gfx950) implementation of MadSDP's interior-point hot path.

The device library `libmadsdp_hip.so` (C ABI in include/madsdp_hip.h) holds
every kernel and the solver driver; this package mirrors the reference's
Julia API (`MadSDPSolver`, `solve!`, the component types) over it.
"""
from .model import BlockConstraints, Model  # noqa: F401
from .solver import (AdaptiveMu, CholeskyBacktrack, DenseCholesky, DualCentering,  # noqa: F401
                     DualShift, EigenStep, HKM, InitialPoint, MadNLPSchur, MadSDPSolver, Mehrotra, NT,
                     PrimalDual, QRSchur, ScaledIdentity, Status, madsdp)
from .serve import Server, serve  # noqa: F401,E402
