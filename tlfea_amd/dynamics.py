"""Implicit time steps for a batch of cantilevers (the dynamic batched configuration).

Each step is the reference's implicit Euler step (SyncedNesterov's g(v) = M (v - v_n)/dt
+ f_int(x_n + dt v) - f_ext = 0, `src/solvers/SyncedNesterov.cu:32-95`) with the internal
force linearised at the rest state, f_int(x) = K (x - X): over the configuration's 200
steps of 1e-4 s the beams move by about a millimetre (strains below 1e-4), where the SVK
tangent at rest is exact to that order.  Per step, for every instance at once:

    (K + M / dt^2) v_{n+1} = M v_n / dt^2 - K u_n / dt + f(t_{n+1}) / dt,
    u_{n+1} = u_n + dt v_{n+1},

solved by the batched FP64 CG to ||rhs - A v|| / ||rhs|| < tol from v_n, with fixed DOFs
eliminated (v = 0 there) instead of the reference's augmented-Lagrangian penalty.  Loads:
gravity as the consistent nodal load M g, plus A sin(2 pi f t) along the instance's
tip-load pattern.

With nonlinear=True each step instead solves the implicit-Euler equation itself,

    g(v) = M (v - v_n) / dt + f_int(X + u_n + dt v) - f(t_{n+1}) = 0,

by inexact Newton iterations from v_n, (M/dt + dt K(x)) dv = -g with K(x) the tangent at the
current positions (tlfea_batch_set_displacement) and f_int from the batched force kernel,
until ||g||/||g(v_n)|| <= tol_nl for every instance (the dynamic-step criterion of the
SyncedNesterov loop, which minimises the same g; solved here with the batched CG, each
solve only to the reduction the worst unconverged instance still needs:
0.5 tol_nl ||g(v_n)|| / ||g||, at least newton_cg_tol, at most 0.1).  An
instance whose ratio has stopped halving below 1e-6 sits at the FP64 floor of evaluating
g (rounding of f_int against the load increment); the step then ends and the instance is
counted in ``floor`` instead of failing.

With nesterov=True each step is instead one SyncedNesterov solve of the same equation for
every instance at once (tlfea_batch_nesterov_solve: the reference's inner Nesterov loop
with momentum restart, per-instance step size 1/(safety dt lambda_max(K + M/dt^2)) from a
power iteration at rest, from v_n) until ||g||/||g(v_n)|| <= tol_nl.

With cheb_degree > 0 every solve is the flexible FP64 CG preconditioned by the FP32
Jacobi-Chebyshev stage of that degree on K + M/dt^2 (and, with coarse=True, the additive
coarse correction P A_c^-1 P^T over rigid motions of node planes), both set up once at the
rest state and kept for every step: the mass term dominates the low modes at these time
steps, so the frozen preconditioner stays close to the current operator.

M is the exactly integrated consistent mass by default (mass_rule="exact").  The
reference's 5-point mass (mass_matrix_qp_kernel, FEAT10Data.cu:200-271) is indefinite,
so K + M/dt^2 is indefinite once M/dt^2 outweighs K -- on soft beams (E of a few MPa)
at dt = 1e-4 it is, and CG stops on a direction of negative curvature.
mass_rule="quadrature" keeps the reference's mass for stiff bodies.
"""
from __future__ import annotations

import numpy as np
import torch

from .batch import Batch

DEV = "cuda"


def dynamic_beams(n, seed):
    """The dynamic configuration's instances: beams of the sampler with a tip-load
    amplitude log-uniform in [10, 1e3] N and a frequency uniform in [1, 20] Hz."""
    from . import mesh as M
    items = M.random_beam_batch(n, seed)
    rng = np.random.default_rng(seed + 7919)
    amp = 10 ** rng.uniform(1, 3, size=n)
    freq = rng.uniform(1, 20, size=n)
    return items, amp, freq


class BatchDynamics:
    def __init__(self, items, amp, freq, dt=1e-4, g=(0.0, 0.0, -9.81), tol=1e-8, batch=None, mass_rule="exact",
                 nonlinear=False, tol_nl=1e-8, newton_cg_tol=1e-10, max_newton=8, cheb_degree=0, coarse=False,
                 nesterov=False, alpha_safety=1.1, max_inner=200000):
        self.b = batch if batch is not None else Batch(items)
        b = self.b
        self.precond = cheb_degree > 0
        self.nesterov, self.max_inner, self.mass_rule = nesterov, max_inner, mass_rule
        self.dt, self.tol = dt, tol
        self.nonlinear, self.tol_nl, self.newton_cg_tol, self.max_newton = nonlinear, tol_nl, newton_cg_tol, max_newton
        b.set_displacement(None)
        self.c = 1.0 / dt ** 2
        rho = np.array([mat.get("rho", 1000.0) for _, mat in items])
        self.rho = rho
        b.set_mass(rho * self.c, rule=mass_rule)
        if nesterov:
            # the step's Hessian is M/dt + dt K = dt (K + M/dt^2)
            self.alpha = 1.0 / (alpha_safety * dt * b.lambda_max())
        if self.precond:
            b.precond_setup(cheb_degree)
            if coarse:
                b.coarse_setup(True)
        n3 = 3 * b.n_nodes
        self.fixed = b.mask.bool()
        counts = 3 * np.diff(b.node_off)
        inst = torch.as_tensor(np.repeat(np.arange(b.n_inst), counts)).to(DEV)
        self.inst = inst
        # tip-load pattern of unit total load per instance
        F = torch.zeros(b.n_inst, 3, dtype=torch.float64, device=DEV)
        F.index_add_(0, inst[0::3], b.f.view(-1, 3))
        fn = torch.linalg.norm(F, dim=1).clamp_min(1e-300)
        self.tip = b.f / fn[inst]
        self.amp = torch.as_tensor(np.asarray(amp, np.float64)).to(DEV)[inst]
        self.omega = torch.as_tensor(2 * np.pi * np.asarray(freq, np.float64)).to(DEV)[inst]
        grep = torch.as_tensor(np.tile(np.asarray(g, np.float64), b.n_nodes)).to(DEV)
        # consistent gravity load g_d sum_j M_ij over all nodes j (fixed ones included)
        self.fg = b.apply(grep, k_scale=0.0, mass=True, masked=False) / self.c
        self.fg[self.fixed] = 0.0
        self.u = torch.zeros(n3, dtype=torch.float64, device=DEV)
        self.v = torch.zeros(n3, dtype=torch.float64, device=DEV)
        self.t = 0.0
        self.iters = []
        self.newton = []
        self.floor = []

    def load(self, t):
        return self.fg + self.tip * (self.amp * torch.sin(self.omega * t))

    def _norms(self, g):
        return torch.sqrt(self.b.dot(g))

    def _residual(self, v, f1):
        """g(v) on the free DOFs, with the batch moved to x = X + u_n + dt v."""
        b, dt = self.b, self.dt
        b.set_displacement(self.u + dt * v)
        g = dt * b.apply(v - self.v, k_scale=0.0, mass=True) + b.internal_force() - f1
        g[self.fixed] = 0.0
        return g

    def step(self):
        if self.nesterov:
            return self._step_nesterov()
        if self.nonlinear:
            return self._step_newton()
        b, dt = self.b, self.dt
        t1 = self.t + dt
        rhs = b.apply(self.v, k_scale=0.0, mass=True) - b.apply(self.u, k_scale=1.0) / dt + self.load(t1) / dt
        rhs[self.fixed] = 0.0
        v1 = b.solve(tol=self.tol, f=rhs, u0=self.v, precond=self.precond)
        if b.rc != 0:
            raise RuntimeError(f"time step {len(self.iters) + 1}: {int((b.status != 0).sum())} instances did not "
                               f"converge")
        self.u = self.u + dt * v1
        self.v = v1
        self.t = t1
        self.iters.append(b.iters.copy())
        return self.u, self.v

    def _step_nesterov(self):
        b, dt = self.b, self.dt
        t1 = self.t + dt
        f1 = self.load(t1)
        v = b.nesterov_solve(dt, self.rho, self.alpha, self.u, self.v, f1, mass_rule=self.mass_rule,
                             rel_tol=self.tol_nl, max_inner=self.max_inner)
        if b.rc != 0:
            raise RuntimeError(f"time step {len(self.iters) + 1}: {int((b.status != 0).sum())} instances did not "
                               f"converge")
        self.u = self.u + dt * v
        self.v = v
        self.t = t1
        self.iters.append(b.iters.copy())
        return self.u, self.v

    def _step_newton(self):
        b, dt = self.b, self.dt
        t1 = self.t + dt
        f1 = self.load(t1)
        v = self.v.clone()
        its = np.zeros(b.n_inst, np.int64)
        g = self._residual(v, f1)
        g0 = self._norms(g)
        k = 0
        prev = None
        at_floor = 0
        while True:
            gn = self._norms(g)
            ok = gn <= self.tol_nl * g0
            if bool(ok.all()):
                break
            ratio = gn / g0.clamp_min(1e-300)
            if prev is not None:
                stalled = (~ok) & (ratio > 0.5 * prev) & (ratio < 1e-6)
                if bool((ok | stalled).all()):
                    at_floor = int(stalled.sum())
                    break
            prev = ratio
            if k == self.max_newton:
                raise RuntimeError(f"time step {len(self.iters) + 1}: Newton did not reach {self.tol_nl:g} in "
                                   f"{k} iterations (worst ratio {float((gn / g0.clamp_min(1e-300)).max()):.2e})")
            # (M/dt + dt K(x)) dv = -g  <=>  (K(x) + M/dt^2) dv = -g/dt; instances already
            # converged get a zero load, which the batched CG finishes at once (and compacts out)
            rhs = -g / dt
            rhs[ok[self.inst]] = 0.0
            worst = float(ratio[~ok].max())
            cg_tol = min(0.1, max(self.newton_cg_tol, 0.5 * self.tol_nl / max(worst, 1e-300)))
            dv = b.solve(tol=cg_tol, f=rhs, precond=self.precond)
            if b.rc != 0:
                raise RuntimeError(f"time step {len(self.iters) + 1}, Newton iteration {k + 1}: "
                                   f"{int((b.status != 0).sum())} instances did not converge")
            its += b.iters
            v = v + dv
            k += 1
            g = self._residual(v, f1)
        b.set_displacement(None)
        self.u = self.u + dt * v
        self.v = v
        self.t = t1
        self.iters.append(its)
        self.newton.append(k)
        self.floor.append(at_floor)
        return self.u, self.v
