"""Host layer for batches of independent meshes (``tlfea_batch_*`` in include/tlfea_hip.h).

``Batch`` concatenates a list of meshes with their materials into one device numbering
and solves the linear-static problem K_i u_i = f_i of every instance at once; the
per-instance loop it replaces is LinearStaticSolver::Solve called once per mesh
(`src/solvers/LinearStaticSolver.cu:469-481`).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import capi
from .capi import call, ptr

DEV = "cuda"


def lame(E, nu):
    return E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))


class Batch:
    """items: list of (TetMesh, dict(E=..., nu=...)); every mesh has the same element type."""

    def __init__(self, items, stream=None):
        if not items:
            raise ValueError("empty batch")
        nn = items[0][0].nn
        if any(m.nn != nn for m, _ in items):
            raise ValueError("all instances must have the same element type")
        self.items = items
        self.nn = nn
        self.n_inst = len(items)
        self.stream = stream
        ne = np.array([m.n_elem for m, _ in items], np.int64)
        nv = np.array([m.n_nodes for m, _ in items], np.int64)
        self.elem_off = np.concatenate([[0], np.cumsum(ne)]).astype(np.int32)
        self.node_off = np.concatenate([[0], np.cumsum(nv)]).astype(np.int32)
        self.n_elem = int(self.elem_off[-1])
        self.n_nodes = int(self.node_off[-1])
        elems = np.concatenate([np.asarray(m.elems, np.int64) + int(o) for (m, _), o in zip(items, self.node_off[:-1])])
        if elems.min() < 0 or elems.max() >= self.n_nodes:
            raise ValueError("connectivity references nodes outside the batch")
        conn = np.ascontiguousarray(elems.T.reshape(-1)).astype(np.int32)  # column-major [nn][n_elem]
        xyz = np.concatenate([m.xyz() for m, _ in items])
        mask = np.zeros(3 * self.n_nodes, np.uint8)
        f = np.zeros(3 * self.n_nodes)
        for (m, _), o in zip(items, self.node_off[:-1]):
            fx = np.asarray(m.fixed, np.int64) + int(o)
            mask[(3 * fx[:, None] + np.arange(3)).reshape(-1)] = 1
            if m.f_ext is not None:
                f[3 * int(o): 3 * (int(o) + m.n_nodes)] = m.f_ext
        lam_mu = np.array([lame(mat["E"], mat["nu"]) for _, mat in items])
        self.lam = np.ascontiguousarray(lam_mu[:, 0])
        self.mu = np.ascontiguousarray(lam_mu[:, 1])
        self.conn = torch.as_tensor(conn).to(DEV)
        self.X = [torch.as_tensor(np.ascontiguousarray(xyz[:, d])).to(DEV) for d in range(3)]
        self.mask = torch.as_tensor(mask).to(DEV)
        self.f = torch.as_tensor(f).to(DEV)
        h = C.c_void_p()
        eo, no = self.elem_off, self.node_off
        call("tlfea_batch_create", nn, self.n_inst, eo.ctypes.data, no.ctypes.data, ptr(self.conn), ptr(self.X[0]),
             ptr(self.X[1]), ptr(self.X[2]), self.lam.ctypes.data, self.mu.ctypes.data, ptr(self.mask), C.byref(h),
             self.stream)
        self.h = h
        self.iters = np.zeros(self.n_inst, np.int32)
        self.rel = np.zeros(self.n_inst)
        self.status = np.zeros(self.n_inst, np.int32)
        self.total_iters = 0

    def inst_slice(self, i):
        """DOF slice of instance i."""
        return slice(3 * int(self.node_off[i]), 3 * int(self.node_off[i + 1]))

    def set_chunk(self, nodes):
        call("tlfea_batch_set_chunk", self.h, int(nodes))

    def set_compaction(self, ratio):
        call("tlfea_batch_set_compaction", self.h, float(ratio))

    def autotune(self, f=None, trial_iters=64, tol=1e-8, full_solves=False):
        """Chooses chunk size, graph-replay length and (with full_solves) the compaction
        ratio by timing batched iterations (tlfea_batch_autotune); returns the choice."""
        f = self.f if f is None else f
        u = torch.empty(3 * self.n_nodes, dtype=torch.float64, device=DEV)
        t = capi.BatchTuning()
        call("tlfea_batch_autotune", self.h, ptr(f), ptr(u), int(trial_iters), float(tol), int(bool(full_solves)),
             C.byref(t), self.stream)
        return dict(chunk_nodes=t.chunk_nodes, poll=t.poll, compaction=t.compaction, us_per_iter=t.us_per_iter,
                    solve_s=t.solve_s)

    def n_chunks(self):
        a = C.c_int()
        call("tlfea_batch_sizes", self.h, None, None, None, C.byref(a))
        return a.value

    def tangent_apply(self, p, out=None):
        out = torch.empty(3 * self.n_nodes, dtype=torch.float64, device=DEV) if out is None else out
        call("tlfea_batch_tangent_apply", self.h, ptr(p), ptr(out), self.stream)
        return out

    def tangent_apply_f32(self, p, out=None):
        """(K + c M) p with the FP32 operator of the preconditioner (the mass term when set;
        p and out float32)."""
        out = torch.empty(3 * self.n_nodes, dtype=torch.float32, device=DEV) if out is None else out
        call("tlfea_batch_tangent_apply_f32", self.h, ptr(p), ptr(out), self.stream)
        return out

    def set_mass(self, rho_coef, rule="quadrature"):
        """Operator K + c M per instance (rho_coef[i] = rho_i * c); None removes it.  rule
        "quadrature": the mass of the element's quadrature rule (as the reference's mass
        assembly, indefinite for TET10); "exact": the exactly integrated mass (positive
        definite).  Drops the preconditioner (set it up again for the new operator)."""
        self.precond_degree = 0
        if rho_coef is None:
            call("tlfea_batch_set_mass", self.h, None, 0)
            self.mass_on = False
            return
        rules = {"quadrature": 0, "exact": 1}
        if rule not in rules:
            raise ValueError(f"unknown mass rule {rule!r}")
        rc = np.ascontiguousarray(rho_coef, dtype=np.float64)
        call("tlfea_batch_set_mass", self.h, rc.ctypes.data, rules[rule])
        self.mass_on = True

    def apply(self, p, k_scale=1.0, mass=False, masked=True, out=None):
        """(k_scale K + c M) p (the mass term only when ``mass``); masked=False applies the
        operator without the Dirichlet elimination."""
        out = torch.empty(3 * self.n_nodes, dtype=torch.float64, device=DEV) if out is None else out
        call("tlfea_batch_apply", self.h, float(k_scale), int(bool(mass)), int(bool(masked)), ptr(p), ptr(out),
             self.stream)
        return out

    def set_displacement(self, u=None):
        """u: displacement from the rest positions, AoS [3*n_nodes] on the device -- the
        operator becomes K(X + u) and internal_force() f_int(X + u); None: back to rest."""
        if u is None:
            call("tlfea_batch_set_displacement", self.h, None, None, None, self.stream)
            return
        us = [u[d::3].contiguous() for d in range(3)]
        call("tlfea_batch_set_displacement", self.h, ptr(us[0]), ptr(us[1]), ptr(us[2]), self.stream)

    def internal_force(self, out=None):
        """f_int of every instance at the current positions (all DOFs)."""
        out = torch.empty(3 * self.n_nodes, dtype=torch.float64, device=DEV) if out is None else out
        call("tlfea_batch_internal_force", self.h, ptr(out), self.stream)
        return out

    def dot(self, a, c=None):
        """Per-instance a . c (c = a: squared norms) over all DOFs, a [n_inst] tensor."""
        out = torch.empty(self.n_inst, dtype=torch.float64, device=DEV)
        call("tlfea_batch_dot", self.h, ptr(a), ptr(a if c is None else c), ptr(out), self.stream)
        return out

    def precond_setup(self, degree=6, ratio=30.0, power_iters=20, safety=1.1):
        """FP32 Jacobi-Chebyshev preconditioner of the operator (K, or K + c M when a mass
        term is set, at the current positions); returns the lambda_max(D^-1 A) estimates."""
        lm = np.zeros(self.n_inst)
        call("tlfea_batch_precond_setup", self.h, int(degree), float(ratio), int(power_iters), float(safety),
             lm.ctypes.data, self.stream)
        self.precond_degree = degree
        return lm

    def set_coarse_modes(self, modes):
        """Coarse modes per node plane for the next coarse_setup: 6 (rigid motions, the
        default) or 11 (rigid motions and five in-plane fields; fewer iterations, but a
        set-up about three times longer, which the beams measured here do not win back)."""
        call("tlfea_batch_set_coarse_modes", self.h, int(modes))

    def coarse_setup(self, enable=True):
        """Two-level coarse correction (rigid motions of node planes) inside the FP32 stage;
        returns the coarse dimension (needs precond_setup first)."""
        nc = C.c_int()
        call("tlfea_batch_coarse_setup", self.h, int(bool(enable)), C.byref(nc), self.stream)
        return nc.value

    def coarse_apply(self, r, out=None):
        """z = P A_c^-1 P^T r, the coarse part of the two-level preconditioner (FP64)."""
        out = torch.empty(3 * self.n_nodes, dtype=torch.float64, device=DEV) if out is None else out
        call("tlfea_batch_coarse_apply", self.h, ptr(r), ptr(out), self.stream)
        return out

    def smooth(self, u, sweeps=2, f=None):
        """u += Chebyshev(f - K u), in place (needs precond_setup)."""
        f = self.f if f is None else f
        call("tlfea_batch_smooth", self.h, ptr(f), ptr(u), int(sweeps), self.stream)
        return u

    def solve(self, tol=1e-8, max_iter=100000, u0=None, poll=None, graph=True, f=None, out=None, precond=False):
        """Returns u (device, all instances); per-instance iterations, relative residuals
        and status land in self.iters / self.rel / self.status.  precond=True uses the
        FP32 Chebyshev stage set up by precond_setup() inside a flexible FP64 CG.  poll:
        iterations per graph replay (None: the batch's default or the autotuner's)."""
        f = self.f if f is None else f
        u = out if out is not None else torch.empty(3 * self.n_nodes, dtype=torch.float64, device=DEV)
        if u0 is not None:
            u.copy_(u0)
        tot = C.c_int()
        rc = capi.lib().tlfea_batch_cg_solve(self.h, ptr(f), ptr(u), int(u0 is not None), float(tol), int(max_iter),
                                             int(poll or 0), int(bool(graph)), int(bool(precond)), self.iters.ctypes.data,
                                             self.rel.ctypes.data, self.status.ctypes.data, C.byref(tot),
                                             self.stream)
        self.total_iters = tot.value
        if rc not in (capi.TLFEA_OK, capi.TLFEA_ENOCONV):
            capi.check(rc)
        self.rc = rc
        return u

    def dof_inst(self):
        """Instance index of every DOF (device int64 [3 n_nodes])."""
        if getattr(self, "_dof_inst", None) is None:
            counts = 3 * np.diff(self.node_off)
            self._dof_inst = torch.as_tensor(np.repeat(np.arange(self.n_inst), counts)).to(DEV)
        return self._dof_inst

    def lambda_max(self, iters=60, mass=True, seed=0):
        """Per-instance largest eigenvalue of the batch's operator on the free DOFs (K, plus
        c M when mass and a mass term is set), by FP64 power iteration from a positive start;
        the Rayleigh quotient returned approaches it from below."""
        inst = self.dof_inst()
        fixed = self.mask.bool()
        gen = torch.Generator().manual_seed(seed)
        x = (torch.rand(3 * self.n_nodes, generator=gen, dtype=torch.float64) + 0.5).to(DEV)
        x[fixed] = 0.0
        mass = bool(mass and getattr(self, "mass_on", False))
        for _ in range(iters):
            y = self.apply(x, mass=mass)
            y[fixed] = 0.0
            x = y / torch.sqrt(self.dot(y)).clamp_min(1e-300)[inst]
        y = self.apply(x, mass=mass)
        y[fixed] = 0.0
        return (self.dot(x, y) / self.dot(x).clamp_min(1e-300)).cpu().numpy()

    def nesterov_solve(self, dt, rho, alpha, u_n, v_n, f_ext, v=None, mass_rule="exact", inner_tol=0.0,
                       rel_tol=1e-8, max_inner=200000, restart=True, poll=None, graph=True):
        """One implicit step per instance by the batched SyncedNesterov inner loop
        (tlfea_batch_nesterov_solve): v minimising M (v - v_n)/dt + f_int(X + u_n + dt v) -
        f_ext per instance, from v (when given, updated in place) or v_n.  rho, alpha: per
        instance (density, step size).  Per-instance iterations, |g|/|g_0|, restarts and
        status land in self.iters / self.rel / self.restarts / self.status, step-size halvings of
        the divergence safeguard in self.halvings."""
        rules = {"quadrature": 0, "exact": 1}
        if mass_rule not in rules:
            raise ValueError(f"unknown mass rule {mass_rule!r}")
        use_guess = v is not None
        out = v if use_guess else torch.empty(3 * self.n_nodes, dtype=torch.float64, device=DEV)
        rho = np.ascontiguousarray(np.broadcast_to(np.asarray(rho, np.float64), (self.n_inst,)))
        alpha = np.ascontiguousarray(np.broadcast_to(np.asarray(alpha, np.float64), (self.n_inst,)))
        self.restarts = np.zeros(self.n_inst, np.int32)
        self.halvings = np.zeros(self.n_inst, np.int32)
        tot = C.c_int()
        rc = capi.lib().tlfea_batch_nesterov_solve(
            self.h, float(dt), rho.ctypes.data, rules[mass_rule], alpha.ctypes.data, ptr(u_n), ptr(v_n), ptr(f_ext),
            ptr(out), int(use_guess), float(inner_tol), float(rel_tol), int(max_inner), int(bool(restart)),
            int(poll or 0), int(bool(graph)), self.iters.ctypes.data, self.rel.ctypes.data, self.restarts.ctypes.data,
            self.halvings.ctypes.data, self.status.ctypes.data, C.byref(tot), self.stream)
        self.total_iters = tot.value
        if rc not in (capi.TLFEA_OK, capi.TLFEA_ENOCONV):
            capi.check(rc)
        self.rc = rc
        return out

    def close(self):
        if getattr(self, "h", None) is not None:
            capi.lib().tlfea_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
