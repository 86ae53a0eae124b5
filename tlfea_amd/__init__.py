"""MI355X-native (gfx950 HIP) TLFEA solve path: TET10/TET4 St. Venant-Kirchhoff force and
stiffness kernels, CSR sparsity / assembly / SpMV, CG and the SyncedNesterov / SyncedAdamW
loops, behind a C-ABI (include/tlfea_hip.h) used by the C++ drop-in classes
(include/tlfea/) and, through ctypes, by this package."""
from . import mesh  # noqa: F401  (numpy-only; no GPU needed)

__all__ = ["mesh"]
