"""tencirpauli_amd: TenCirPauli's matrix-free Pauli-sum MVP on AMD MI355X (gfx950).

The public surface mirrors the reference's operator/plan API for this path:
``PauliOperator.native_mvp_plan()`` / ``mvp()`` / ``restrict_charge()`` with
``U1Sector`` and contiguous-block ``ChargeSector`` restrictions.  All numerics
run in the hand-written HIP kernels of ``libtencirpauli_amd.so`` (C ABI in
include/tencirpauli_amd.h); there is no CPU fallback.
"""

from ._lib import DeviceError, NativeLibraryError, UnsupportedOnDeviceError, device_count
from ._plan import DEFAULT_MAX_BYTES
from .charge import AdditiveCharge, ChargeRestrictedOperator, ChargeSector
from .fermion import FermionOperator, FermionTerm, FermionWord, OperatorSpace
from .lanczos import eigsh
from .pauli import BackendMVPPlan, BackendPlanArrays, ChargeAnalysis, NativeMVPPlan, PauliOperator, PauliTerm
from .symmetry import SectorMvpPlan, U1RestrictedOperator, U1Sector

__version__ = "0.1.0"

__all__ = [
    "DEFAULT_MAX_BYTES",
    "AdditiveCharge",
    "BackendMVPPlan",
    "BackendPlanArrays",
    "ChargeAnalysis",
    "ChargeRestrictedOperator",
    "ChargeSector",
    "DeviceError",
    "FermionOperator",
    "FermionTerm",
    "FermionWord",
    "NativeLibraryError",
    "NativeMVPPlan",
    "OperatorSpace",
    "PauliOperator",
    "PauliTerm",
    "SectorMvpPlan",
    "U1RestrictedOperator",
    "U1Sector",
    "UnsupportedOnDeviceError",
    "device_count",
    "eigsh",
]
