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
    _fields_ = [("lam", C.c_double), ("mu", C.c_double), ("eta_damp", C.c_double), ("lambda_damp", C.c_double),
                ("model", C.c_int), ("mu10", C.c_double), ("mu01", C.c_double), ("kappa", C.c_double)]


class Body(C.Structure):
    _fields_ = [("nn", C.c_int), ("n_elem", C.c_int), ("n_nodes", C.c_int), ("conn", C.c_void_p),
                ("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p), ("grad_t", C.c_void_p),
                ("detJ_t", C.c_void_p), ("quad", Quad), ("mat", Material), ("pattern", C.c_void_p),
                ("m_sell", C.c_void_p), ("f_ext", C.c_void_p), ("n_fixed", C.c_int), ("fixed_nodes", C.c_void_p),
                ("xref", C.c_void_p), ("yref", C.c_void_p), ("zref", C.c_void_p), ("f_int", C.c_void_p),
                ("constraint", C.c_void_p), ("force_plan", C.c_void_p), ("X0", C.c_void_p), ("Y0", C.c_void_p),
                ("Z0", C.c_void_p)]


class BatchTuning(C.Structure):
    _fields_ = [("chunk_nodes", C.c_int), ("poll", C.c_int), ("compaction", C.c_double),
                ("us_per_iter", C.c_double), ("solve_s", C.c_double)]


class NesterovParams(C.Structure):
    """Same layout as the reference's SyncedNesterovParams (SyncedNesterov.cuh:31-35)."""
    _fields_ = [("alpha", C.c_double), ("rho", C.c_double), ("inner_tol", C.c_double), ("outer_tol", C.c_double),
                ("max_outer", C.c_int), ("max_inner", C.c_int), ("time_step", C.c_double)]


class AdamWParams(C.Structure):
    """Same layout as the reference's SyncedAdamWParams (SyncedAdamW.cuh:32-39)."""
    _fields_ = [("lr", C.c_double), ("beta1", C.c_double), ("beta2", C.c_double), ("eps", C.c_double),
                ("weight_decay", C.c_double), ("lr_decay", C.c_double), ("inner_tol", C.c_double),
                ("outer_tol", C.c_double), ("rho", C.c_double), ("max_outer", C.c_int), ("max_inner", C.c_int),
                ("time_step", C.c_double), ("convergence_check_interval", C.c_int), ("inner_rtol", C.c_double)]


SOLVER_NESTEROV, SOLVER_ADAMW, SOLVER_ADAMW_NOCOOP = 0, 1, 2

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
    "tlfea_force_plan_create": (i32, [i32, i32, i32, vp, C.POINTER(vp), vp]),
    "tlfea_force_plan_destroy": (i32, [vp]),
    "tlfea_force_plan_sizes": (i32, [vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(i64), C.POINTER(i32)]),
    "tlfea_internal_force_planned": (i32, [vp, vp, vp, vp, vp, vp, C.POINTER(Quad), C.POINTER(Material), vp, vp, i32,
                                           vp]),
    "tlfea_internal_force_rest": (i32, [vp, vp, vp, vp, vp, vp, vp, C.POINTER(Quad), C.POINTER(Material), vp, vp, i32,
                                        vp]),
    "tlfea_tangent_apply_planned": (i32, [vp, vp, vp, vp, vp, vp, vp, vp, vp, i32, C.POINTER(Quad), C.POINTER(Material),
                                          vp, vp, vp, vp]),
    "tlfea_precompute_jinv": (i32, [i32, i32, vp, vp, vp, vp, C.POINTER(Quad), vp, vp]),
    "tlfea_jinv_affine": (i32, [i32, i32, vp, vp, C.POINTER(i32), vp]),
    "tlfea_fixed_dof_mask": (i32, [i32, i32, vp, vp, vp]),
    "tlfea_grad_to_f32": (i32, [i32, i32, vp, vp, C.POINTER(Quad), vp, vp, vp]),
    "tlfea_internal_force_f32": (i32, [i32, i32, i32, vp, vp, vp, vp, vp, vp, C.POINTER(Material), vp, i32, vp]),
    "tlfea_grad_to_ref": (i32, [i32, i32, vp, vp, vp, vp, vp]),
    "tlfea_qp_to_ref": (i32, [i32, i32, vp, vp, vp]),
    "tlfea_pattern_create": (i32, [i32, i32, i32, vp, C.POINTER(vp), vp]),
    "tlfea_pattern_create_ex": (i32, [i32, i32, i32, vp, i32, C.POINTER(vp), vp]),
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
    "tlfea_spmv_block_f32": (i32, [vp, vp, vp, vp, vp]),
    "tlfea_spmv_csr": (i32, [i32, vp, vp, vp, vp, vp, vp]),
    "tlfea_reduce_scratch_size": (i64, [i64]),
    "tlfea_dot": (i32, [i64, vp, vp, vp, vp, vp]),
    "tlfea_nrm2": (i32, [i64, vp, vp, vp, vp]),
    "tlfea_axpy": (i32, [i64, f64, vp, vp, vp]),
    "tlfea_scal": (i32, [i64, f64, vp, vp]),
    "tlfea_cg_create": (i32, [vp, C.POINTER(vp)]),
    "tlfea_cg_destroy": (i32, [vp]),
    "tlfea_cg_solve": (i32, [vp, vp, vp, vp, f64, i32, i32, C.POINTER(i32), C.POINTER(f64), vp]),
    "tlfea_cg_solve_matfree": (i32, [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, i32, C.POINTER(Quad), C.POINTER(Material),
                                     vp, vp, vp, f64, i32, i32, C.POINTER(i32), C.POINTER(f64), vp]),
    "tlfea_cg_solve_matfree_noalloc": (i32, [vp, vp, vp, vp, vp, vp, vp, vp, vp, i32, C.POINTER(Quad),
                                             C.POINTER(Material), vp, vp, vp, f64, i32, i32, C.POINTER(i32),
                                             C.POINTER(f64), vp]),
    "tlfea_batch_create": (i32, [i32, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, C.POINTER(vp), vp]),
    "tlfea_batch_destroy": (i32, [vp]),
    "tlfea_batch_sizes": (i32, [vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)]),
    "tlfea_batch_set_chunk": (i32, [vp, i32]),
    "tlfea_batch_tangent_apply": (i32, [vp, vp, vp, vp]),
    "tlfea_batch_tangent_apply_f32": (i32, [vp, vp, vp, vp]),
    "tlfea_batch_cg_solve": (i32, [vp, vp, vp, i32, f64, i32, i32, i32, i32, vp, vp, vp, C.POINTER(i32), vp]),
    "tlfea_batch_set_compaction": (i32, [vp, f64]),
    "tlfea_batch_autotune": (i32, [vp, vp, vp, i32, f64, i32, C.POINTER(BatchTuning), vp]),
    "tlfea_batch_coarse_setup": (i32, [vp, i32, C.POINTER(i32), vp]),
    "tlfea_batch_coarse_apply": (i32, [vp, vp, vp, vp]),
    "tlfea_batch_set_coarse_modes": (i32, [vp, i32]),
    "tlfea_batch_nesterov_solve": (i32, [vp, f64, vp, i32, vp, vp, vp, vp, vp, i32, f64, f64, i32, i32, i32, i32,
                                         vp, vp, vp, vp, vp, C.POINTER(i32), vp]),
    "tlfea_batch_smooth": (i32, [vp, vp, vp, i32, vp]),
    "tlfea_batch_set_mass": (i32, [vp, vp, i32]),
    "tlfea_batch_apply": (i32, [vp, f64, i32, i32, vp, vp, vp]),
    "tlfea_batch_set_displacement": (i32, [vp, vp, vp, vp, vp]),
    "tlfea_batch_internal_force": (i32, [vp, vp, vp]),
    "tlfea_batch_dot": (i32, [vp, vp, vp, vp, vp]),
    "tlfea_batch_precond_setup": (i32, [vp, i32, f64, i32, f64, vp, vp]),
    "tlfea_mpnn_create": (i32, [i32, C.POINTER(vp)]),
    "tlfea_mpnn_destroy": (i32, [vp]),
    "tlfea_mpnn_num_params": (i32, [vp, C.POINTER(i32)]),
    "tlfea_mpnn_set_params": (i32, [vp, vp]),
    "tlfea_mpnn_get_params": (i32, [vp, vp]),
    "tlfea_mpnn_get_grads": (i32, [vp, vp]),
    "tlfea_mpnn_zero_grads": (i32, [vp, vp]),
    "tlfea_mpnn_forward": (i32, [vp, i32, vp, vp, vp, vp, vp]),
    "tlfea_mpnn_loss_grad": (i32, [vp, i32, vp, vp, vp, vp, vp, C.POINTER(f64), vp]),
    "tlfea_mpnn_adam": (i32, [vp, f64, f64, f64, f64, vp]),
    "tlfea_mpnn_backward": (i32, [vp, i32, vp, vp, vp, vp]),
    "tlfea_mpnn_energy_grad": (i32, [i32, vp, vp, vp, vp, vp, vp]),
    "tlfea_mpnn_features": (i32, [i32, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
    "tlfea_mpnn_scale_output": (i32, [i32, vp, vp, vp, vp, vp, vp]),
    "tlfea_add_displacement": (i32, [i32, vp, vp, vp, vp, vp]),
    "tlfea_matrix_to_f32": (i32, [vp, vp, vp, vp]),
    "tlfea_cg_solve_mixed": (i32, [vp, vp, vp, vp, vp, f64, f64, i32, i32, C.POINTER(i32), C.POINTER(f64),
                                   C.POINTER(i32), vp]),
    "tlfea_constraint_eval": (i32, [i32, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
    "tlfea_solver_create": (i32, [i32, C.POINTER(Body), C.POINTER(vp)]),
    "tlfea_solver_destroy": (i32, [vp]),
    "tlfea_solver_update_body": (i32, [vp, C.POINTER(Body)]),
    "tlfea_solver_set_parameters": (i32, [vp, vp]),
    "tlfea_solver_set_restart": (i32, [vp, i32]),
    "tlfea_solver_restarts": (i32, [vp, C.POINTER(i32)]),
    "tlfea_solver_set_matrix_free_mass": (i32, [vp, f64]),
    "tlfea_solver_set_rel_tol": (i32, [vp, f64]),
    "tlfea_solver_set_poll_interval": (i32, [vp, i32]),
    "tlfea_solver_solve": (i32, [vp, vp]),
    "tlfea_solver_state": (i32, [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp)]),
    "tlfea_solver_stats": (i32, [vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(f64), C.POINTER(f64)]),
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
