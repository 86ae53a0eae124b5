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
    def __init__(self, items, amp, freq, dt=1e-4, g=(0.0, 0.0, -9.81), tol=1e-8, batch=None, mass_rule="exact"):
        self.b = batch if batch is not None else Batch(items)
        b = self.b
        self.dt, self.tol = dt, tol
        self.c = 1.0 / dt ** 2
        rho = np.array([mat.get("rho", 1000.0) for _, mat in items])
        b.set_mass(rho * self.c, rule=mass_rule)
        n3 = 3 * b.n_nodes
        self.fixed = b.mask.bool()
        counts = 3 * np.diff(b.node_off)
        inst = torch.as_tensor(np.repeat(np.arange(b.n_inst), counts)).to(DEV)
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

    def load(self, t):
        return self.fg + self.tip * (self.amp * torch.sin(self.omega * t))

    def step(self):
        b, dt = self.b, self.dt
        t1 = self.t + dt
        rhs = b.apply(self.v, k_scale=0.0, mass=True) - b.apply(self.u, k_scale=1.0) / dt + self.load(t1) / dt
        rhs[self.fixed] = 0.0
        v1 = b.solve(tol=self.tol, f=rhs, u0=self.v)
        if b.rc != 0:
            raise RuntimeError(f"time step {len(self.iters) + 1}: {int((b.status != 0).sum())} instances did not "
                               f"converge")
        self.u = self.u + dt * v1
        self.v = v1
        self.t = t1
        self.iters.append(b.iters.copy())
        return self.u, self.v
