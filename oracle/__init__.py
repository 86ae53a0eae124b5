"""ctypes front end of the FP64 CPU oracle (``oracle/oracle.cpp``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py``, to check results against; the product never imports it.

Parity status of each function (how it is pinned, see tests/test_oracle.py):

* ``quad_t10`` / ``quad_t4`` -- exact rational constants, sum of weights, exactness on
  monomials of degree <= 3.
* ``grad_t10`` / ``grad_t4`` -- partition of unity, sum_a X_a (x) g_a = I, detJ = 6 vol on
  straight tets, the affine closed form (J^-T dN/dxi), cross-check of the two inverse
  routes.
* ``calc_p`` -- P(I) = 0, P(R) = 0 for rotations, uniaxial closed form, finite difference
  of the SVK energy.
* ``internal_force`` -- f(rest) = 0, sum of forces = 0, finite difference of the total
  energy, patch test on affine displacements.
* ``node_pattern`` / ``dof_pattern`` -- brute force (std::set), compared against an
  independent numpy sort/unique route.
* ``mass_values`` -- symmetry, total mass = rho V, row sums = integral of rho N_i,
  TET4 rank-one rule.
* ``mass_values_exact`` -- the coefficient table against an independent exact rational
  integration of N_a N_b over the tetrahedron (Dirichlet moment formula), positive
  definiteness, total mass = rho V, row sums equal to the 5-point rule's.
* ``stiffness_values`` -- symmetry, rigid modes in the kernel, finite difference of f_int,
  linear-elastic closed form at F = I.
* ``mr_point`` and the Mooney-Rivlin branch of ``calc_p`` / ``internal_force`` /
  ``stiffness_values`` -- P(I) = 0, P = finite difference of W, A = finite difference of P
  (and its major symmetry), frame indifference P(QF) = Q P(F), the small-strain limit
  (linear elasticity with mu = 2 (mu10 + mu01), bulk modulus kappa), f_int = finite
  difference of the total MR energy, K = finite difference of f_int.
* ``tangent_apply`` -- equal to ``spmv`` of the assembled ``stiffness_values`` (SVK and
  Mooney-Rivlin, with damping, deformed states), symmetric; ``true_rel_residual`` uses it.
* ``cg`` / ``band_cholesky_solve`` -- each against dense numpy solves, Euler-Bernoulli
  band on the beam.
* ``nesterov_step`` -- the iterate sequence against an independent pure-Python
  restatement of the loop; one long outer iteration against the direct solve of the same
  implicit-Euler equation.
* ``constraint_jacobian_csr`` -- J J^T = I, J^T equal to the transpose of J, J x = the
  fixed DOFs of x, row sums of J^T equal to the fixed-DOF indicator.
* ``adamw_step`` (coop and nocoop) -- the iterate sequence against torch.optim.AdamW driven
  by ``step_gradient`` (1e-10; a wrong bias-correction step, decay order, weight decay, eps
  or gradient lag fails it); the converged state against the direct solve.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)


class OrcModel(C.Structure):
    _fields_ = [("nn", C.c_int), ("nq", C.c_int), ("n_elem", C.c_int), ("n_nodes", C.c_int),
                ("conn", ip), ("grad", dp), ("detJ", dp), ("qw", dp),
                ("lam", C.c_double), ("mu", C.c_double), ("eta_d", C.c_double), ("lambda_d", C.c_double),
                ("m_off", ip), ("m_col", ip), ("m_val", dp), ("f_ext", dp),
                ("n_fixed", C.c_int), ("fixed", ip), ("xref", dp), ("yref", dp), ("zref", dp)]


class OrcState(C.Structure):
    _fields_ = [("x", dp), ("y", dp), ("z", dp), ("v_guess", dp), ("v_prev", dp), ("lam", dp)]


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (g++ -fopenmp)."""
    import subprocess
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "all"], cwd=_HERE, check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _lib.orc_node_pattern.restype = C.c_int64
        _lib.orc_dof_pattern.restype = C.c_int64
        _lib.orc_energy.restype = C.c_double
        _lib.orc_cg.restype = C.c_int
        _lib.orc_band_cholesky_solve.restype = C.c_int
        _lib.orc_nesterov_step.restype = C.c_int
        _lib.orc_adamw_step.restype = C.c_int
        _lib.orc_num_threads.restype = C.c_int
        for name in ("orc_energy", "orc_cg", "orc_spmv"):
            pass
    return _lib


def _d(a):
    return a.ctypes.data_as(dp) if a is not None else None


def _i(a):
    return a.ctypes.data_as(ip) if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def set_num_threads(n: int) -> None:
    lib().orc_set_num_threads(int(n))


def num_threads() -> int:
    return int(lib().orc_num_threads())


# ---------------------------------------------------------------------------------------
class material:
    """Context manager selecting the material of calc_p / internal_force / energy /
    stiffness_values (and the Nesterov/AdamW restatements) for its body:
    ``with material("mr", mu10, mu01, kappa): ...``; SVK (the lambda, mu arguments)
    outside.  Mirrors the reference's per-element ``d_material_model`` switch."""

    def __init__(self, model="svk", mu10=0.0, mu01=0.0, kappa=0.0):
        self.args = (1 if model == "mr" else 0, float(mu10), float(mu01), float(kappa))

    def __enter__(self):
        lib().orc_set_material(C.c_int(self.args[0]), *[C.c_double(a) for a in self.args[1:]])
        return self

    def __exit__(self, *exc):
        lib().orc_set_material(C.c_int(0), C.c_double(0.0), C.c_double(0.0), C.c_double(0.0))
        return False


def mr_point(F, mu10, mu01, kappa):
    """(W, P, A) of the Mooney-Rivlin material at one 3x3 F (row-major numpy in and out;
    A[i, j, k, l] = dP_ij / dF_kl)."""
    F = np.asarray(F, np.float64)
    F9 = _f64(F.T.reshape(-1))  # column-major
    W = C.c_double(0.0)
    P9 = np.zeros(9)
    A81 = np.zeros(81)
    lib().orc_mr_point(_d(F9), C.c_double(mu10), C.c_double(mu01), C.c_double(kappa), C.byref(W), _d(P9), _d(A81))
    P = P9.reshape(3, 3).T
    A = A81.reshape(3, 3, 3, 3).transpose(3, 2, 1, 0)
    return W.value, P.copy(), A.copy()


def quad_t10():
    q = [np.zeros(5) for _ in range(4)]
    lib().orc_quad_t10(*[_d(a) for a in q])
    return q  # qx, qy, qz, qw


def quad_t4():
    q = [np.zeros(1) for _ in range(4)]
    lib().orc_quad_t4(*[_d(a) for a in q])
    return q


def quad(nn):
    return quad_t10() if nn == 10 else quad_t4()


def grad(mesh):
    """(grad, detJ) in the reference layout for a TET10 or TET4 mesh."""
    conn = _i32(mesh.conn_colmajor())
    x, y, z = _f64(mesh.x), _f64(mesh.y), _f64(mesh.z)
    if mesh.nn == 10:
        qx, qy, qz, _ = quad_t10()
        g = np.zeros(mesh.n_elem * 5 * 30)
        dj = np.zeros(mesh.n_elem * 5)
        lib().orc_t10_grad(mesh.n_elem, _i(conn), _d(x), _d(y), _d(z), _d(qx), _d(qy), _d(qz), _d(g), _d(dj))
    else:
        g = np.zeros(mesh.n_elem * 12)
        dj = np.zeros(mesh.n_elem)
        lib().orc_t4_grad(mesh.n_elem, _i(conn), _d(x), _d(y), _d(z), _d(g), _d(dj))
    return g, dj


def calc_p(mesh, g, lam, mu, x=None, y=None, z=None, v=None, eta=0.0, lambda_d=0.0):
    nq = 5 if mesh.nn == 10 else 1
    conn = _i32(mesh.conn_colmajor())
    x = _f64(mesh.x if x is None else x)
    y = _f64(mesh.y if y is None else y)
    z = _f64(mesh.z if z is None else z)
    n = mesh.n_elem * nq * 9
    F, P, Fd, Pv = (np.zeros(n) for _ in range(4))
    lib().orc_calc_p(mesh.nn, nq, mesh.n_elem, _i(conn), _d(x), _d(y), _d(z), _d(_f64(g)), C.c_double(lam),
                     C.c_double(mu), C.c_double(eta), C.c_double(lambda_d), _d(_f64(v) if v is not None else None),
                     _d(F), _d(P), _d(Fd), _d(Pv))
    return F, P, Fd, Pv


def fint_from_p(mesh, P, g, detJ, f=None):
    nq = 5 if mesh.nn == 10 else 1
    qw = quad(mesh.nn)[3]
    f = np.zeros(3 * mesh.n_nodes) if f is None else f
    lib().orc_fint_from_p(mesh.nn, nq, mesh.n_elem, _i(_i32(mesh.conn_colmajor())), _d(_f64(P)), _d(_f64(g)),
                          _d(_f64(detJ)), _d(qw), _d(f))
    return f


def internal_force(mesh, g, detJ, lam, mu, x=None, y=None, z=None, v=None, eta=0.0, lambda_d=0.0):
    nq = 5 if mesh.nn == 10 else 1
    qw = quad(mesh.nn)[3]
    x = _f64(mesh.x if x is None else x)
    y = _f64(mesh.y if y is None else y)
    z = _f64(mesh.z if z is None else z)
    f = np.zeros(3 * mesh.n_nodes)
    lib().orc_internal_force(mesh.nn, nq, mesh.n_elem, mesh.n_nodes, _i(_i32(mesh.conn_colmajor())), _d(x), _d(y),
                             _d(z), _d(_f64(g)), _d(_f64(detJ)), _d(qw), C.c_double(lam), C.c_double(mu),
                             C.c_double(eta), C.c_double(lambda_d), _d(_f64(v) if v is not None else None), _d(f))
    return f


def energy(mesh, g, detJ, lam, mu, x=None, y=None, z=None):
    nq = 5 if mesh.nn == 10 else 1
    qw = quad(mesh.nn)[3]
    x = _f64(mesh.x if x is None else x)
    y = _f64(mesh.y if y is None else y)
    z = _f64(mesh.z if z is None else z)
    return float(lib().orc_energy(mesh.nn, nq, mesh.n_elem, _i(_i32(mesh.conn_colmajor())), _d(x), _d(y), _d(z),
                                  _d(_f64(g)), _d(_f64(detJ)), _d(qw), C.c_double(lam), C.c_double(mu)))


def node_pattern(mesh):
    conn = _i32(mesh.conn_colmajor())
    nnz = lib().orc_node_pattern(mesh.nn, mesh.n_elem, mesh.n_nodes, _i(conn), None, None)
    off = np.zeros(mesh.n_nodes + 1, np.int32)
    col = np.zeros(nnz, np.int32)
    lib().orc_node_pattern(mesh.nn, mesh.n_elem, mesh.n_nodes, _i(conn), _i(off), _i(col))
    return off, col


def dof_pattern(mesh):
    conn = _i32(mesh.conn_colmajor())
    nnz = lib().orc_dof_pattern(mesh.nn, mesh.n_elem, mesh.n_nodes, _i(conn), None, None)
    off = np.zeros(3 * mesh.n_nodes + 1, np.int32)
    col = np.zeros(nnz, np.int32)
    lib().orc_dof_pattern(mesh.nn, mesh.n_elem, mesh.n_nodes, _i(conn), _i(off), _i(col))
    return off, col


def mass_values(mesh, detJ, rho, off, col):
    nq = 5 if mesh.nn == 10 else 1
    qx, qy, qz, qw = quad(mesh.nn)
    vals = np.zeros(len(col))
    lib().orc_mass_values(mesh.nn, nq, mesh.n_elem, _i(_i32(mesh.conn_colmajor())), _d(_f64(detJ)), _d(qx), _d(qy),
                          _d(qz), _d(qw), C.c_double(rho), _i(_i32(off)), _i(_i32(col)), _d(vals), mesh.n_nodes)
    return vals


# Corner/edge numbering of the TET10 element: edge m joins corners T10_EDGES[m].
T10_EDGES = ((0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3))


def exact_mass_coefficients(nn):
    """C with the exactly integrated element mass M_ab = rho V C_ab of a straight-sided element.

    This has no reference counterpart: mass_matrix_qp_kernel (FEAT10Data.cu:200-271)
    integrates the degree-4 integrand N_a N_b with the 5-point rule, whose negative
    centroid weight makes the TET10 mass indefinite, and the TET4 one
    (FEAT4Data.cu:197-243) with one point, which gives a rank-1 mass.  Exactly
    integrated, TET10: corner-corner 6/420 (diagonal) and 1/420, corner-edge -4/420 when
    the edge has that corner and -6/420 otherwise, edge-edge 32/420 (diagonal), 16/420
    for edges sharing a corner and 8/420 for opposite edges; TET4: (1 + delta_ab)/20.
    """
    if nn == 4:
        return (np.ones((4, 4)) + np.eye(4)) / 20.0
    Cm = np.zeros((10, 10))
    for a in range(4):
        for b in range(4):
            Cm[a, b] = 6.0 if a == b else 1.0
        for m, e in enumerate(T10_EDGES):
            Cm[a, 4 + m] = Cm[4 + m, a] = -4.0 if a in e else -6.0
    for m, e in enumerate(T10_EDGES):
        for k, f in enumerate(T10_EDGES):
            Cm[4 + m, 4 + k] = 32.0 if m == k else (16.0 if set(e) & set(f) else 8.0)
    return Cm / 420.0


def mass_values_exact(mesh, detJ, rho, off, col):
    """Node-level consistent mass on the mass pattern with M_e = rho V_e C
    (``exact_mass_coefficients``), V_e = sum_q w_q detJ_(e,q)."""
    nq = 5 if mesh.nn == 10 else 1
    qw = quad(mesh.nn)[3]
    vol = np.asarray(detJ, np.float64).reshape(mesh.n_elem, nq) @ qw
    Cm = exact_mass_coefficients(mesh.nn)
    conn = np.asarray(mesh.conn_colmajor(), np.int64).reshape(mesh.nn, mesh.n_elem).T
    off = np.asarray(off, np.int64)
    col = np.asarray(col, np.int64)
    vals = np.zeros(len(col))
    for e in range(mesh.n_elem):
        for a in range(mesh.nn):
            i = conn[e, a]
            row = col[off[i]:off[i + 1]]
            k = off[i] + np.searchsorted(row, conn[e])
            np.add.at(vals, k, rho * vol[e] * Cm[a])
    return vals


def stiffness_values(mesh, g, detJ, lam, mu, off, col, x=None, y=None, z=None, h=1.0, eta=0.0, lambda_d=0.0,
                     parallel=False):
    nq = 5 if mesh.nn == 10 else 1
    qw = quad(mesh.nn)[3]
    x = _f64(mesh.x if x is None else x)
    y = _f64(mesh.y if y is None else y)
    z = _f64(mesh.z if z is None else z)
    vals = np.zeros(len(col))
    fn = lib().orc_stiffness_values_omp if parallel else lib().orc_stiffness_values
    fn(mesh.nn, nq, mesh.n_elem, _i(_i32(mesh.conn_colmajor())), _d(x), _d(y), _d(z),
                               _d(_f64(g)), _d(_f64(detJ)), _d(qw), C.c_double(lam), C.c_double(mu), C.c_double(eta),
                               C.c_double(lambda_d), C.c_double(h), _i(_i32(off)), _i(_i32(col)), _d(vals))
    return vals


def tangent_apply(mesh, g, detJ, lam, mu, p, x=None, y=None, z=None, h=1.0, eta=0.0, lambda_d=0.0):
    """(h K(x) + C) p element by element (OpenMP), the operator ``stiffness_values``
    assembles, without the CSR."""
    nq = 5 if mesh.nn == 10 else 1
    qw = quad(mesh.nn)[3]
    x = _f64(mesh.x if x is None else x)
    y = _f64(mesh.y if y is None else y)
    z = _f64(mesh.z if z is None else z)
    out = np.zeros(3 * mesh.n_nodes)
    lib().orc_tangent_apply(mesh.nn, nq, mesh.n_elem, mesh.n_nodes, _i(_i32(mesh.conn_colmajor())), _d(x), _d(y),
                            _d(z), _d(_f64(g)), _d(_f64(detJ)), _d(qw), C.c_double(lam), C.c_double(mu),
                            C.c_double(eta), C.c_double(lambda_d), C.c_double(h), _d(_f64(p)), _d(out))
    return out


def true_rel_residual(mesh, g, detJ, lam, mu, u, f, fixed):
    """||(f - K u)_free|| / ||f_free|| of the Dirichlet-eliminated linear-static system
    (LinearStaticSolver.cu:316-342: fixed DOFs removed, u = 0 there) at the rest positions:
    the true residual behind the CG's stopping rule sqrt(rr/rr0) (:418)."""
    free = np.ones(3 * mesh.n_nodes, bool)
    free.reshape(-1, 3)[np.asarray(fixed, np.int64)] = False
    uu = np.where(free, np.asarray(u, np.float64).reshape(-1), 0.0)
    r = (np.asarray(f, np.float64).reshape(-1) - tangent_apply(mesh, g, detJ, lam, mu, uu))[free]
    return float(np.linalg.norm(r) / np.linalg.norm(np.asarray(f, np.float64).reshape(-1)[free]))


def apply_dirichlet(off, col, vals, f, fixed):
    vals = vals.copy()
    f = f.copy()
    fixed = _i32(fixed)
    lib().orc_apply_dirichlet(len(off) - 1, _i(_i32(off)), _i(_i32(col)), _d(vals), _d(f), len(fixed), _i(fixed))
    return vals, f


def spmv(off, col, vals, xv):
    y = np.zeros(len(off) - 1)
    lib().orc_spmv(len(off) - 1, _i(_i32(off)), _i(_i32(col)), _d(_f64(vals)), _d(_f64(xv)), _d(y))
    return y


def cg(off, col, vals, f, tol=1e-10, max_iter=10000):
    n = len(off) - 1
    u = np.zeros(n)
    res = C.c_double(0.0)
    it = lib().orc_cg(n, _i(_i32(off)), _i(_i32(col)), _d(_f64(vals)), _d(_f64(f)), _d(u), C.c_double(tol),
                      int(max_iter), C.byref(res))
    return u, int(it), float(res.value)


def band_cholesky_solve(off, col, vals, b):
    n = len(off) - 1
    x = np.zeros(n)
    rc = lib().orc_band_cholesky_solve(n, _i(_i32(off)), _i(_i32(col)), _d(_f64(vals)), _d(_f64(b)), _d(x))
    if rc != 0:
        raise np.linalg.LinAlgError(f"matrix not positive definite at row {rc - 1}")
    return x


def constraint_jacobian_csr(fixed, n_nodes, transpose=False):
    """Selection Jacobian of the fixed-node constraints c = x - x_ref at the fixed nodes.

    J (FEAT10Data.cu:420-437,478-506): row 3k + d holds one entry 1.0 at DOF
    3 fixed[k] + d.  J^T (FEAT10Data.cu:439-473,510-562): rows are the 3N DOFs, row
    3 fixed[k] + d holds 1.0 at column 3k + d (columns ascending within a row).
    Returns (offsets, columns, values)."""
    fixed = np.asarray(fixed, np.int64)
    nc = 3 * len(fixed)
    dofs = (3 * fixed[:, None] + np.arange(3)[None, :]).reshape(-1)  # constraint t -> DOF
    if not transpose:
        return np.arange(nc + 1, dtype=np.int32), dofs.astype(np.int32), np.ones(nc)
    rows = 3 * int(n_nodes)
    off = np.zeros(rows + 1, np.int64)
    np.add.at(off, dofs + 1, 1)
    off = np.cumsum(off)
    order = np.lexsort((np.arange(nc), dofs))  # by DOF, then constraint index
    return off.astype(np.int32), order.astype(np.int32), np.ones(nc)


# ---------------------------------------------------------------------------------------
class Problem:
    """Everything the implicit-Euler solvers need, held alive for ctypes."""

    def __init__(self, mesh, E, nu, rho, eta=0.0, lambda_d=0.0, f_ext=None, fixed=None):
        self.mesh = mesh
        self.nn = mesh.nn
        self.nq = 5 if mesh.nn == 10 else 1
        self.mu = E / (2 * (1 + nu))
        self.lam = (E * nu) / ((1 + nu) * (1 - 2 * nu))
        self.eta, self.lambda_d = eta, lambda_d
        self.conn = _i32(mesh.conn_colmajor())
        self.g, self.detJ = grad(mesh)
        self.qw = quad(mesh.nn)[3]
        self.m_off, self.m_col = node_pattern(mesh)
        self.m_val = mass_values(mesh, self.detJ, rho, self.m_off, self.m_col)
        self.f_ext = _f64(mesh.f_ext if f_ext is None else f_ext)
        self.fixed = _i32(mesh.fixed if fixed is None else fixed)
        self.xref, self.yref, self.zref = _f64(mesh.x).copy(), _f64(mesh.y).copy(), _f64(mesh.z).copy()
        self.model = OrcModel(self.nn, self.nq, mesh.n_elem, mesh.n_nodes, _i(self.conn), _d(self.g), _d(self.detJ),
                              _d(self.qw), self.lam, self.mu, eta, lambda_d, _i(self.m_off), _i(self.m_col),
                              _d(self.m_val), _d(self.f_ext), len(self.fixed), _i(self.fixed), _d(self.xref),
                              _d(self.yref), _d(self.zref))

    def set_fixed(self, fixed):
        """Replaces the fixed-node list (FEAT10Data.cu:674-751); the constraint targets stay
        the positions at construction.  The state's multipliers are the caller's: they keep
        their constraint-number indexing, and a new count needs a new array."""
        self.fixed = _i32(fixed)
        self.model.n_fixed = len(self.fixed)
        self.model.fixed = _i(self.fixed)

    def new_state(self):
        n = self.mesh.n_nodes
        return dict(x=_f64(self.mesh.x).copy(), y=_f64(self.mesh.y).copy(), z=_f64(self.mesh.z).copy(),
                    v_guess=np.zeros(3 * n), v_prev=np.zeros(3 * n), lam=np.zeros(3 * len(self.fixed)))

    @staticmethod
    def _cstate(s):
        return OrcState(_d(s["x"]), _d(s["y"]), _d(s["z"]), _d(s["v_guess"]), _d(s["v_prev"]), _d(s["lam"]))

    def nesterov_step(self, s, alpha, rho, inner_tol, outer_tol, max_outer, max_inner, dt):
        return lib().orc_nesterov_step(C.byref(self.model), C.byref(self._cstate(s)), C.c_double(alpha),
                                       C.c_double(rho), C.c_double(inner_tol), C.c_double(outer_tol), int(max_outer),
                                       int(max_inner), C.c_double(dt))

    def adamw_step(self, s, lr, beta1, beta2, eps, weight_decay, lr_decay, inner_tol, outer_tol, rho, max_outer,
                   max_inner, dt, check_interval, inner_rtol, nocoop=False):
        return lib().orc_adamw_step(C.byref(self.model), C.byref(self._cstate(s)), C.c_double(lr), C.c_double(beta1),
                                    C.c_double(beta2), C.c_double(eps), C.c_double(weight_decay),
                                    C.c_double(lr_decay), C.c_double(inner_tol), C.c_double(outer_tol),
                                    C.c_double(rho), int(max_outer), int(max_inner), C.c_double(dt),
                                    int(check_interval), C.c_double(inner_rtol), int(bool(nocoop)))

    def step_gradient(self, xp, yp, zp, v, v_prev, lam, dt, rho):
        g = np.zeros(3 * self.mesh.n_nodes)
        lib().orc_step_gradient(C.byref(self.model), _d(_f64(xp)), _d(_f64(yp)), _d(_f64(zp)), _d(_f64(v)),
                                _d(_f64(v_prev)), _d(_f64(lam)), C.c_double(dt), C.c_double(rho), _d(g))
        return g
