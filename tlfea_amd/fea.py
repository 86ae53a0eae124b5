"""Python host layer over the C-ABI, for tests and the benchmark.

``Body`` holds one mesh's device state the way ``GPU_FEAT10_Data`` / ``GPU_FEAT4_Data``
do in the reference (`src/elements/FEAT10Data.cuh:733-809`), with the device layouts
of ``include/tlfea_hip.h``; ``LinearStatic`` follows ``LinearStaticSolver::Solve``
(`src/solvers/LinearStaticSolver.cu:469-481`).  All compute goes through libtlfea_hip.so.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import capi
from .capi import call, ptr

DEV = "cuda"


def _dev(a, dtype):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).to(DEV)


class Body:
    def __init__(self, mesh, E=0.0, nu=0.0, rho=0.0, eta=0.0, lambda_d=0.0, stream=None, window=None):
        self.mesh = mesh
        self.window = window
        self.nn = mesh.nn
        self.nq = 5 if self.nn == 10 else 1
        self.n_elem = mesh.n_elem
        self.n_nodes = mesh.n_nodes
        self.stream = stream
        self.quad = capi.quadrature(self.nn)
        self.set_material(E, nu, rho, eta, lambda_d)
        ce = np.asarray(mesh.elems)
        if ce.size and (ce.min() < 0 or ce.max() >= self.n_nodes):
            raise ValueError("connectivity references nodes outside [0, n_nodes)")
        if len(mesh.fixed) and (np.min(mesh.fixed) < 0 or np.max(mesh.fixed) >= self.n_nodes):
            raise ValueError("fixed-node list references nodes outside [0, n_nodes)")
        self.conn = _dev(mesh.conn_colmajor(), torch.int32)
        self.X0 = [_dev(a, torch.float64) for a in (mesh.x, mesh.y, mesh.z)]
        self.x = [t.clone() for t in self.X0]
        pe = capi.padded_elems(self.n_elem)
        self.grad_t = torch.zeros(pe * self.nq * self.nn * 3, dtype=torch.float64, device=DEV)
        self.detJ_t = torch.zeros(pe * self.nq, dtype=torch.float64, device=DEV)
        self.f_int = torch.zeros(3 * self.n_nodes, dtype=torch.float64, device=DEV)
        self.f_ext = _dev(mesh.f_ext if mesh.f_ext is not None else np.zeros(3 * self.n_nodes), torch.float64)
        self.fixed = _dev(mesh.fixed, torch.int32)
        self._pattern = None
        self._plan = None
        self._dof_mask = None
        self._jinv = None
        self._affine = None
        self.precompute()

    def set_material(self, E, nu, rho, eta=0.0, lambda_d=0.0):
        self.E, self.nu, self.rho = E, nu, rho
        self.mat = capi.material(E, nu, eta, lambda_d) if (E or nu) else capi.Material(0, 0, eta, lambda_d)

    def set_mooney_rivlin(self, mu10, mu01, kappa):
        """Compressible Mooney-Rivlin in place of SVK, the reference's
        SetMooneyRivlin(mu10, mu01, kappa) (`src/elements/FEAT10Data.cuh:571-587`); damping
        and rho are kept.  set_svk() switches back."""
        self.mat.model, self.mat.mu10, self.mat.mu01, self.mat.kappa = 1, float(mu10), float(mu01), float(kappa)

    def set_svk(self):
        """SVK with the current lambda, mu (`src/elements/FEAT10Data.cuh:533-552`)."""
        self.mat.model, self.mat.mu10, self.mat.mu01, self.mat.kappa = 0, 0.0, 0.0, 0.0

    @property
    def s(self):
        return self.stream

    def precompute(self):
        call("tlfea_precompute_grad", self.nn, self.n_elem, ptr(self.conn), ptr(self.X0[0]), ptr(self.X0[1]),
             ptr(self.X0[2]), C.byref(self.quad), ptr(self.grad_t), ptr(self.detJ_t), self.s)

    def grad_ref(self):
        g = torch.empty(self.n_elem * self.nq * self.nn * 3, dtype=torch.float64, device=DEV)
        d = torch.empty(self.n_elem * self.nq, dtype=torch.float64, device=DEV)
        call("tlfea_grad_to_ref", self.nn, self.n_elem, ptr(self.grad_t), ptr(self.detJ_t), ptr(g), ptr(d), self.s)
        return g, d

    def qp_tensor(self):
        return torch.zeros(capi.padded_elems(self.n_elem) * self.nq * 9, dtype=torch.float64, device=DEV)

    def qp_ref(self, t):
        out = torch.empty(self.n_elem * self.nq * 9, dtype=torch.float64, device=DEV)
        call("tlfea_qp_to_ref", self.nn, self.n_elem, ptr(t), ptr(out), self.s)
        return out

    def calc_p(self, v=None, x=None):
        x = self.x if x is None else x
        F, P, Fd, Pv = (self.qp_tensor() for _ in range(4))
        call("tlfea_calc_p", self.nn, self.n_elem, ptr(self.conn), ptr(x[0]), ptr(x[1]), ptr(x[2]), ptr(self.grad_t),
             C.byref(self.mat), ptr(v), ptr(F), ptr(P), ptr(Fd), ptr(Pv), self.s)
        return F, P, Fd, Pv

    def fint_from_p(self, P, f=None):
        f = self.f_int if f is None else f
        call("tlfea_fint_from_p", self.nn, self.n_elem, ptr(self.conn), ptr(P), ptr(self.grad_t), ptr(self.detJ_t),
             C.byref(self.quad), ptr(f), self.s)
        return f

    def internal_force(self, x=None, v=None, f=None, accumulate=False):
        x = self.x if x is None else x
        f = torch.empty(3 * self.n_nodes, dtype=torch.float64, device=DEV) if f is None else f
        call("tlfea_internal_force", self.nn, self.n_elem, self.n_nodes, ptr(self.conn), ptr(x[0]), ptr(x[1]),
             ptr(x[2]), ptr(self.grad_t), ptr(self.detJ_t), C.byref(self.quad), C.byref(self.mat), ptr(v), ptr(f),
             int(bool(accumulate)), self.s)
        return f

    @property
    def plan(self):
        """Block-local force plan, or None when the mesh order is too scattered for the LDS
        budget (then the atomic-scatter kernel is used)."""
        if self._plan is None:
            h = C.c_void_p()
            rc = capi.lib().tlfea_force_plan_create(self.nn, self.n_elem, self.n_nodes, ptr(self.conn), C.byref(h),
                                                     self.s)
            if rc == capi.TLFEA_ECAPACITY:
                self._plan = False
            else:
                capi.check(rc)
                self._plan = h
        return self._plan or None

    def plan_sizes(self):
        a, b, c, d = C.c_int(), C.c_int(), C.c_int64(), C.c_int()
        call("tlfea_force_plan_sizes", self.plan, C.byref(a), C.byref(b), C.byref(c), C.byref(d))
        return dict(block_elems=a.value, n_blocks=b.value, n_partials=c.value, max_block_nodes=d.value)

    def internal_force_planned(self, x=None, v=None, f=None, accumulate=False):
        x = self.x if x is None else x
        f = torch.empty(3 * self.n_nodes, dtype=torch.float64, device=DEV) if f is None else f
        call("tlfea_internal_force_planned", self.plan, ptr(x[0]), ptr(x[1]), ptr(x[2]), ptr(self.grad_t),
             ptr(self.detJ_t), C.byref(self.quad), C.byref(self.mat), ptr(v), ptr(f), int(bool(accumulate)), self.s)
        return f

    def internal_force_rest(self, x=None, v=None, f=None, accumulate=False):
        """f_int with the shape-function gradients rebuilt from the rest positions (no
        gradient arrays read)."""
        x = self.x if x is None else x
        f = torch.empty(3 * self.n_nodes, dtype=torch.float64, device=DEV) if f is None else f
        X0 = self.X0
        call("tlfea_internal_force_rest", self.plan, ptr(x[0]), ptr(x[1]), ptr(x[2]), ptr(X0[0]), ptr(X0[1]),
             ptr(X0[2]), C.byref(self.quad), C.byref(self.mat), ptr(v), ptr(f), int(bool(accumulate)), self.s)
        return f

    def elastic_material(self):
        """The material without Kelvin-Voigt damping (the tangent of the elastic force)."""
        m = self.mat
        return capi.Material(m.lam, m.mu, 0.0, 0.0, m.model, m.mu10, m.mu01, m.kappa)

    def fixed_dof_mask(self):
        """uint8 [3*n_nodes], 1 on the DOFs of fixed nodes."""
        if self._dof_mask is None:
            m = torch.empty(3 * self.n_nodes, dtype=torch.uint8, device=DEV)
            nf = int(self.fixed.numel())
            call("tlfea_fixed_dof_mask", self.n_nodes, nf, ptr(self.fixed) if nf else None, ptr(m), self.s)
            self._dof_mask = m
        return self._dof_mask

    @property
    def jinv_t(self):
        """Inverse-transposed Jacobians per (element, q), the matrix-free tangent's input."""
        if self._jinv is None:
            pe = capi.padded_elems(self.n_elem)
            self._jinv = torch.zeros(pe * self.nq * 9, dtype=torch.float64, device=DEV)
            call("tlfea_precompute_jinv", self.nn, self.n_elem, ptr(self.conn), ptr(self.X0[0]), ptr(self.X0[1]),
                 ptr(self.X0[2]), C.byref(self.quad), ptr(self._jinv), self.s)
        return self._jinv

    @property
    def affine(self):
        """True when every element's M and det J agree at all quadrature points."""
        if self._affine is None:
            a = C.c_int()
            call("tlfea_jinv_affine", self.nn, self.n_elem, ptr(self.jinv_t), ptr(self.detJ_t), C.byref(a), self.s)
            self._affine = bool(a.value)
        return self._affine

    def tangent_apply(self, p, x=None, out=None, masked=True, affine=True, damping=False):
        """y = K(x) p without assembly; with masked=True fixed rows return p and fixed
        columns are dropped (the Dirichlet-eliminated operator); affine: on a mesh of
        straight-sided elements one quadrature point's M and det J are read; damping: the
        reference's K + C with the body's Kelvin-Voigt coefficients."""
        if self.plan is None:
            raise RuntimeError("matrix-free tangent needs a force plan (mesh too scattered for the LDS budget)")
        x = self.x if x is None else x
        out = torch.empty(3 * self.n_nodes, dtype=torch.float64, device=DEV) if out is None else out
        mat = self.mat if damping else self.elastic_material()
        X0 = self.X0
        call("tlfea_tangent_apply_planned", self.plan, ptr(x[0]), ptr(x[1]), ptr(x[2]), ptr(X0[0]), ptr(X0[1]),
             ptr(X0[2]), ptr(self.jinv_t), ptr(self.detJ_t), int(affine and self.affine),
             C.byref(self.quad), C.byref(mat),
             ptr(self.fixed_dof_mask()) if masked else None, ptr(p), ptr(out), self.s)
        return out

    # ---- sparse ------------------------------------------------------------------
    @property
    def pattern(self):
        if self._pattern is None:
            h = C.c_void_p()
            if self.window is None:
                call("tlfea_pattern_create", self.nn, self.n_elem, self.n_nodes, ptr(self.conn), C.byref(h), self.s)
            else:
                call("tlfea_pattern_create_ex", self.nn, self.n_elem, self.n_nodes, ptr(self.conn), int(self.window),
                     C.byref(h), self.s)
            self._pattern = h
        return self._pattern

    def pattern_sizes(self):
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        call("tlfea_pattern_sizes", self.pattern, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value

    def node_csr(self):
        nnz, _, _ = self.pattern_sizes()
        off = torch.empty(self.n_nodes + 1, dtype=torch.int32, device=DEV)
        col = torch.empty(nnz, dtype=torch.int32, device=DEV)
        call("tlfea_pattern_node_csr", self.pattern, ptr(off), ptr(col), self.s)
        return off, col

    def dof_csr(self):
        _, nnz, _ = self.pattern_sizes()
        off = torch.empty(3 * self.n_nodes + 1, dtype=torch.int32, device=DEV)
        col = torch.empty(nnz, dtype=torch.int32, device=DEV)
        call("tlfea_pattern_dof_csr", self.pattern, ptr(off), ptr(col), self.s)
        return off, col

    def mass_sell(self, rho=None):
        rho = self.rho if rho is None else rho
        _, _, nb = self.pattern_sizes()
        m = torch.empty(nb, dtype=torch.float64, device=DEV)
        call("tlfea_mass_assemble", self.pattern, self.nn, self.n_elem, ptr(self.conn), ptr(self.detJ_t),
             C.byref(self.quad), float(rho), ptr(m), self.s)
        return m

    def mass_csr_values(self, m_sell):
        nnz, _, _ = self.pattern_sizes()
        v = torch.empty(nnz, dtype=torch.float64, device=DEV)
        call("tlfea_mass_to_csr", self.pattern, ptr(m_sell), ptr(v), self.s)
        return v

    def stiffness_sell(self, x=None, h=1.0, out=None):
        x = self.x if x is None else x
        _, _, nb = self.pattern_sizes()
        k = torch.empty(9 * nb, dtype=torch.float64, device=DEV) if out is None else out
        call("tlfea_stiffness_assemble", self.pattern, self.nn, self.n_elem, ptr(self.conn), ptr(x[0]), ptr(x[1]),
             ptr(x[2]), ptr(self.grad_t), ptr(self.detJ_t), C.byref(self.quad), C.byref(self.mat), float(h), ptr(k),
             self.s)
        return k

    def stiffness_csr_values(self, k_sell):
        _, nnz, _ = self.pattern_sizes()
        v = torch.empty(nnz, dtype=torch.float64, device=DEV)
        call("tlfea_stiffness_to_csr", self.pattern, ptr(k_sell), ptr(v), self.s)
        return v

    def apply_dirichlet(self, k_sell, f):
        call("tlfea_apply_dirichlet", self.pattern, int(self.fixed.numel()), ptr(self.fixed), ptr(k_sell), ptr(f),
             self.s)

    def spmv(self, k_sell, xv, y=None):
        y = torch.empty(3 * self.n_nodes, dtype=torch.float64, device=DEV) if y is None else y
        call("tlfea_spmv_block", self.pattern, ptr(k_sell), ptr(xv), ptr(y), self.s)
        return y

    def spmv_mass(self, m_sell, xv, y=None):
        y = torch.empty(3 * self.n_nodes, dtype=torch.float64, device=DEV) if y is None else y
        call("tlfea_spmv_node3", self.pattern, ptr(m_sell), ptr(xv), ptr(y), self.s)
        return y

    def close(self):
        if self._pattern is not None:
            capi.lib().tlfea_pattern_destroy(self._pattern)
            self._pattern = None
        if self._plan:
            capi.lib().tlfea_force_plan_destroy(self._plan)
        self._plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class LinearStatic:
    """K(x) u = f_ext with Dirichlet elimination and device-scalar CG, then x += u
    (`src/solvers/LinearStaticSolver.cu:469-481`)."""

    def __init__(self, body: Body, tol=1e-10, max_iter=10000, check_every=32, mixed=False, refresh=1e-2,
                 matfree=False, two_level=False, cheb_degree=16, cheb_ratio=1000.0):
        """mixed=True: FP32 Krylov recursion with FP64 accumulation and reliable updates
        (tlfea_cg_solve_mixed); tol is then the true FP64 relative residual.
        matfree=True: FP64 CG with K(x) p applied element by element
        (tlfea_cg_solve_matfree); nothing is assembled.  The operator is the assembled
        path's: K + C with Kelvin-Voigt damping.
        two_level=True (not in the reference): the flexible FP64 CG preconditioned by the
        FP32 two-level stage (cheb_degree Chebyshev steps on [lambda_max / cheb_ratio,
        lambda_max] plus the node-plane coarse correction, set up afresh every solve) on
        K(x), through a one-instance batch of the body that follows its positions; the same
        stopping rule.  SVK bodies without Kelvin-Voigt damping."""
        if sum((mixed, matfree, two_level)) > 1:
            raise ValueError("mixed, matfree and two_level are exclusive")
        if two_level and (body.mat.model != 0 or body.mat.eta_damp != 0.0 or body.mat.lambda_damp != 0.0):
            raise ValueError("the two-level path needs an SVK body without damping")
        self.b = body
        self.tol, self.max_iter, self.check_every = tol, max_iter, check_every
        self.mixed, self.refresh, self.matfree = mixed, refresh, matfree
        self.two_level, self.cheb_degree, self.cheb_ratio = two_level, cheb_degree, cheb_ratio
        self.batch, self._batch_key = None, None
        self.k32 = None
        self.refreshes = 0
        h = C.c_void_p()
        call("tlfea_cg_create", body.pattern, C.byref(h))
        self.cg = h
        self.k = None
        self.f = None
        self.u = torch.zeros(3 * body.n_nodes, dtype=torch.float64, device=DEV)
        self.iters = 0
        self.res = 0.0

    def _batch(self):
        """The body as a one-instance batch, rebuilt when its fixed set or material changes
        (the batch copies both at creation)."""
        from .batch import Batch
        b = self.b
        fixed = b.fixed.cpu().numpy()
        key = (fixed.tobytes(), b.mat.lam, b.mat.mu)
        if self.batch is None or key != self._batch_key:
            if self.batch is not None:
                self.batch.close()
            m = b.mesh
            mesh = type(m)(m.x, m.y, m.z, m.elems, fixed=fixed.astype(np.int32), f_ext=None)
            self.batch = Batch([(mesh, dict(E=b.E, nu=b.nu))], stream=b.stream)
            self._batch_key = key
        return self.batch

    def assemble(self):
        if self.two_level:
            self.f = self.b.f_ext  # fixed DOFs are masked inside the solve
            return
        if self.matfree:
            if self.b.plan is None:
                raise RuntimeError("matrix-free CG needs a force plan")
            self.f = self.b.f_ext  # fixed DOFs are zeroed inside the solve
            self.b.fixed_dof_mask()
            self.b.jinv_t
            return
        self.k = self.b.stiffness_sell(out=self.k)
        self.f = self.b.f_ext.clone()
        self.b.apply_dirichlet(self.k, self.f)
        if self.mixed:
            if self.k32 is None:
                self.k32 = torch.empty(self.k.numel(), dtype=torch.float32, device=DEV)
            call("tlfea_matrix_to_f32", self.b.pattern, ptr(self.k), ptr(self.k32), self.b.s)

    def cg_solve(self):
        it = C.c_int()
        res = C.c_double()
        if self.mixed:
            nref = C.c_int()
            rc = capi.lib().tlfea_cg_solve_mixed(self.cg, ptr(self.k), ptr(self.k32), ptr(self.f), ptr(self.u),
                                                 float(self.tol), float(self.refresh), int(self.max_iter),
                                                 int(self.check_every), C.byref(it), C.byref(res), C.byref(nref),
                                                 self.b.s)
            self.refreshes = nref.value
        elif self.two_level:
            b = self.b
            bt = self._batch()
            u = torch.stack([b.x[d] - b.X0[d] for d in range(3)], 1).reshape(-1)
            bt.set_displacement(u if bool(torch.any(u != 0)) else None)
            bt.precond_setup(degree=self.cheb_degree, ratio=self.cheb_ratio)
            bt.coarse_setup()
            bt.solve(tol=float(self.tol), max_iter=int(self.max_iter), f=self.f, out=self.u, precond=True,
                     graph=False)
            self.iters, self.res = int(bt.iters[0]), float(bt.rel[0])
            return int(bt.status[0])
        elif self.matfree:
            b = self.b
            mat = b.mat
            X0 = b.X0
            rc = capi.lib().tlfea_cg_solve_matfree(self.cg, b.plan, ptr(b.x[0]), ptr(b.x[1]), ptr(b.x[2]),
                                                   ptr(X0[0]), ptr(X0[1]), ptr(X0[2]), ptr(b.jinv_t),
                                                   ptr(b.detJ_t), int(b.affine), C.byref(b.quad), C.byref(mat),
                                                   ptr(b.fixed_dof_mask()), ptr(self.f), ptr(self.u),
                                                   float(self.tol), int(self.max_iter), int(self.check_every),
                                                   C.byref(it), C.byref(res), b.s)
        else:
            rc = capi.lib().tlfea_cg_solve(self.cg, ptr(self.k), ptr(self.f), ptr(self.u), float(self.tol),
                                           int(self.max_iter), int(self.check_every), C.byref(it), C.byref(res),
                                           self.b.s)
        self.iters, self.res = it.value, res.value
        if rc not in (capi.TLFEA_OK, capi.TLFEA_ENOCONV, capi.TLFEA_ENOTPD):
            capi.check(rc)
        return rc

    def solve(self):
        self.assemble()
        rc = self.cg_solve()
        x = self.b.x
        call("tlfea_add_displacement", self.b.n_nodes, ptr(self.u), ptr(x[0]), ptr(x[1]), ptr(x[2]), self.b.s)
        return rc

    def close(self):
        if self.cg is not None:
            capi.lib().tlfea_cg_destroy(self.cg)
            self.cg = None
        if getattr(self, "batch", None) is not None:
            self.batch.close()
            self.batch = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_HIP_MEMCPY = None


def _hip_memcpy():
    """hipMemcpy of the HIP runtime, loaded once."""
    global _HIP_MEMCPY
    if _HIP_MEMCPY is None:
        fn = C.CDLL("libamdhip64.so").hipMemcpy
        fn.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _HIP_MEMCPY = fn
    return _HIP_MEMCPY


class Solver:
    """SyncedNesterov / SyncedAdamW / SyncedAdamWNocoop on one Body
    (`src/solvers/SyncedNesterov.cuh:39-179`, `SyncedAdamW.cuh`, `SyncedAdamWNocoop.cuh`)."""

    def __init__(self, body: Body, kind: int = capi.SOLVER_NESTEROV, planned: bool = True, rest: bool = False):
        """planned: the block-local force kernel (else the atomic one); rest (with planned):
        the force kernel that rebuilds the gradients from the rest positions."""
        self.b = body
        self.kind = kind
        self.m_sell = body.mass_sell()
        nf = int(body.fixed.numel())
        self.constraint = torch.zeros(max(1, 3 * nf), dtype=torch.float64, device=DEV)
        self.xref = [t.clone() for t in body.X0]
        d = capi.Body(body.nn, body.n_elem, body.n_nodes, ptr(body.conn), ptr(body.x[0]), ptr(body.x[1]),
                      ptr(body.x[2]), ptr(body.grad_t), ptr(body.detJ_t), body.quad, body.mat, body.pattern,
                      ptr(self.m_sell), ptr(body.f_ext), nf, ptr(body.fixed) if nf else None, ptr(self.xref[0]),
                      ptr(self.xref[1]), ptr(self.xref[2]), ptr(body.f_int), ptr(self.constraint),
                      body.plan if planned else None)
        if planned and rest:
            d.X0, d.Y0, d.Z0 = ptr(body.X0[0]), ptr(body.X0[1]), ptr(body.X0[2])
        self.desc = d
        h = C.c_void_p()
        call("tlfea_solver_create", kind, C.byref(d), C.byref(h))
        self.h = h
        vg, vp_, lam = C.c_void_p(), C.c_void_p(), C.c_void_p()
        call("tlfea_solver_state", h, C.byref(vg), C.byref(vp_), C.byref(lam))
        self._state = (vg.value, vp_.value, lam.value)
        self.params = None

    def set_parameters(self, params):
        self.params = params
        call("tlfea_solver_set_parameters", self.h, C.cast(C.pointer(params), C.c_void_p))

    def set_rel_tol(self, rel_tol):
        call("tlfea_solver_set_rel_tol", self.h, float(rel_tol))

    def set_poll_interval(self, every):
        call("tlfea_solver_set_poll_interval", self.h, int(every))

    def set_restart(self, on=True):
        """Nesterov momentum restart on g.(v_(k+1) - v_k) > 0 (not in the reference)."""
        call("tlfea_solver_set_restart", self.h, int(bool(on)))

    def set_matrix_free_mass(self, rho):
        """Mass term of the gradient applied per element inside the force kernel (density rho,
        the quadrature rule of the assembled mass) instead of from the assembled matrix; 0
        returns to the assembled mass (not in the reference)."""
        call("tlfea_solver_set_matrix_free_mass", self.h, float(rho))

    def restarts(self):
        n = C.c_int()
        call("tlfea_solver_restarts", self.h, C.byref(n))
        return n.value

    def solve(self):
        call("tlfea_solver_solve", self.h, self.b.s)

    def stats(self):
        a, b_, c, d = C.c_int(), C.c_int(), C.c_double(), C.c_double()
        call("tlfea_solver_stats", self.h, C.byref(a), C.byref(b_), C.byref(c), C.byref(d))
        return dict(inner=a.value, outer=b_.value, norm_g=c.value, rel=d.value)

    def _vec(self, addr, n):
        out = torch.empty(n, dtype=torch.float64, device=DEV)
        torch.cuda.synchronize()
        rc = _hip_memcpy()(ptr(out), addr, 8 * n, 3)  # device to device
        if rc != 0:
            raise RuntimeError(f"hipMemcpy failed: {rc}")
        return out

    def v_guess(self):
        return self._vec(self._state[0], 3 * self.b.n_nodes)

    def v_prev(self):
        return self._vec(self._state[1], 3 * self.b.n_nodes)

    def lambda_guess(self):
        return self._vec(self._state[2], 3 * int(self.b.fixed.numel()))

    def close(self):
        if getattr(self, "h", None) is not None:
            capi.lib().tlfea_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
