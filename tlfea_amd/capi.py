"""ctypes binding of the C-ABI in ``include/tlfea_hip.h`` (libtlfea_hip.so).

Device buffers are torch tensors on ``cuda:0`` (torch is used only to own device memory);
every call passes raw device pointers and the stream handle, and raises ``TlfeaError`` on
a non-zero status.  There is no CPU fallback: if the library is missing, ``lib()`` raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libtlfea_hip.so")

TLFEA_OK = 0
TLFEA_EINVAL = -1
TLFEA_EHIP = -2
TLFEA_ENOTPD = -3
TLFEA_ECAPACITY = -4
TLFEA_ENOCONV = -5
TILE = 64


class TlfeaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"tlfea error {code}: {msg}")
        self.code = code


class Quad(C.Structure):
    _fields_ = [("n", C.c_int), ("x", C.c_double * 5), ("y", C.c_double * 5), ("z", C.c_double * 5),
                ("w", C.c_double * 5)]


class Material(C.Structure):
    _fields_ = [("lam", C.c_double), ("mu", C.c_double), ("eta_damp", C.c_double), ("lambda_damp", C.c_double)]


vp = C.c_void_p
i32 = C.c_int
i64 = C.c_int64
f64 = C.c_double

# name: (restype, argtypes)
_SIGS = {
    "tlfea_version": (C.c_char_p, []),
    "tlfea_last_error": (C.c_char_p, []),
    "tlfea_device_synchronize": (i32, []),
    "tlfea_padded_elems": (i64, [i32]),
    "tlfea_quadrature": (i32, [i32, C.POINTER(Quad)]),
    "tlfea_precompute_grad": (i32, [i32, i32, vp, vp, vp, vp, C.POINTER(Quad), vp, vp, vp]),
    "tlfea_calc_p": (i32, [i32, i32, vp, vp, vp, vp, vp, C.POINTER(Material), vp, vp, vp, vp, vp, vp]),
    "tlfea_fint_from_p": (i32, [i32, i32, vp, vp, vp, vp, C.POINTER(Quad), vp, vp]),
    "tlfea_internal_force": (i32, [i32, i32, i32, vp, vp, vp, vp, vp, vp, C.POINTER(Quad), C.POINTER(Material), vp,
                                   vp, i32, vp]),
    "tlfea_grad_to_f32": (i32, [i32, i32, vp, vp, C.POINTER(Quad), vp, vp, vp]),
    "tlfea_internal_force_f32": (i32, [i32, i32, i32, vp, vp, vp, vp, vp, vp, C.POINTER(Material), vp, i32, vp]),
    "tlfea_grad_to_ref": (i32, [i32, i32, vp, vp, vp, vp, vp]),
    "tlfea_qp_to_ref": (i32, [i32, i32, vp, vp, vp]),
    "tlfea_pattern_create": (i32, [i32, i32, i32, vp, C.POINTER(vp), vp]),
    "tlfea_pattern_destroy": (i32, [vp]),
    "tlfea_pattern_sizes": (i32, [vp, C.POINTER(i64), C.POINTER(i64), C.POINTER(i64)]),
    "tlfea_pattern_node_csr": (i32, [vp, vp, vp, vp]),
    "tlfea_pattern_dof_csr": (i32, [vp, vp, vp, vp]),
    "tlfea_mass_assemble": (i32, [vp, i32, i32, vp, vp, C.POINTER(Quad), f64, vp, vp]),
    "tlfea_mass_to_csr": (i32, [vp, vp, vp, vp]),
    "tlfea_stiffness_assemble": (i32, [vp, i32, i32, vp, vp, vp, vp, vp, vp, C.POINTER(Quad), C.POINTER(Material),
                                       f64, vp, vp]),
    "tlfea_stiffness_to_csr": (i32, [vp, vp, vp, vp]),
    "tlfea_apply_dirichlet": (i32, [vp, i32, vp, vp, vp, vp]),
    "tlfea_spmv_block": (i32, [vp, vp, vp, vp, vp]),
    "tlfea_spmv_node3": (i32, [vp, vp, vp, vp, vp]),
    "tlfea_spmv_csr": (i32, [i32, vp, vp, vp, vp, vp, vp]),
    "tlfea_reduce_scratch_size": (i64, [i64]),
    "tlfea_dot": (i32, [i64, vp, vp, vp, vp, vp]),
    "tlfea_nrm2": (i32, [i64, vp, vp, vp, vp]),
    "tlfea_axpy": (i32, [i64, f64, vp, vp, vp]),
    "tlfea_scal": (i32, [i64, f64, vp, vp]),
    "tlfea_cg_create": (i32, [vp, C.POINTER(vp)]),
    "tlfea_cg_destroy": (i32, [vp]),
    "tlfea_cg_solve": (i32, [vp, vp, vp, vp, f64, i32, i32, C.POINTER(i32), C.POINTER(f64), vp]),
    "tlfea_add_displacement": (i32, [i32, vp, vp, vp, vp, vp]),
}

_lib = None


def declared_symbols(header: str | None = None):
    """Function names declared in include/*.h (used by the export test)."""
    import glob
    import re
    root = os.path.dirname(HERE)
    names = []
    for h in sorted(glob.glob(os.path.join(root, "include", "*.h"))) if header is None else [header]:
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names += re.findall(r"\b(tlfea_[a-z0-9_]+)\s*\(", txt)
    return sorted(set(names))


def lib():
    """Load libtlfea_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() (hipcc, gfx950) first")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> int:
    if rc != TLFEA_OK:
        raise TlfeaError(rc, lib().tlfea_last_error().decode())
    return rc


def call(name: str, *args) -> int:
    return check(getattr(lib(), name)(*args))


def ptr(t) -> int | None:
    """Raw device pointer of a torch tensor (None -> NULL)."""
    if t is None:
        return None
    return t.data_ptr()


def quadrature(nn: int) -> Quad:
    q = Quad()
    call("tlfea_quadrature", nn, C.byref(q))
    return q


def material(E: float, nu: float, eta: float = 0.0, lambda_d: float = 0.0) -> Material:
    """SVK Lame parameters from E, nu as the reference's SetSVK computes them
    (`src/elements/FEAT10Data.cuh:543-552`)."""
    mu = E / (2 * (1 + nu))
    lam = (E * nu) / ((1 + nu) * (1 - 2 * nu))
    return Material(lam, mu, eta, lambda_d)


def padded_elems(n_elem: int) -> int:
    return int(lib().tlfea_padded_elems(n_elem))
