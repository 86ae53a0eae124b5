"""Learned warm start for batched linear-static solves (``tlfea_mpnn_*`` in include/tlfea_hip.h).

The network (4 message-passing layers, width 64, BF16 on the matrix cores, FP32
accumulation) predicts, per node, the displacement scaled by the beam's bending scale
s = |F| Lx^3 / (E Ly Lz min(Ly, Lz)^2) -- the axial component by s min(Ly, Lz) / Lx, the
size of the axial displacement that goes with the cross-sections' rotation, so that the
loss weighs the axial error as much as the beam's energy does -- from: rest coordinates
normalised by the bounding
box, the fixed-node mask, the nodal load over |F|, and per-beam inputs broadcast to every
node (load direction, log10 E, nu, aspect ratios).  The per-beam inputs are there because
four message-passing hops cannot carry a tip load along a beam hundreds of nodes long.

``WarmStart.guess`` returns u0 = s * prediction for every instance of a ``Batch``, scaled
per instance to the energy minimiser along it, u0 <- (f.u0 / u0.K u0) u0, which makes
the energy error ||u0 - u*||_K (the quantity CG minimises) never larger than the zero
guess's; an instance whose guess does not point downhill (f.u0 <= 0) falls back to
u0 = 0.  The residual test ||f - K u0|| > ||f|| is also computed (``policy="residual"``
applies it instead): K amplifies the BF16-level node-to-node noise of any network output
by up to its condition number, so that test rejects guesses that still save CG work.
The CG always finishes from u0 with the plain method's own stopping test, so the guess
changes the iteration count, never the answer's tolerance.

Measured on 256 held-out sampler beams with the weights in ``weights/`` (20,000 training
solves, 8 epochs): the raw guess is 2.5% off the solution in L2 (median), yet its energy
error ||u0 - u*||_K^2 is 156 times the solution's energy ||u*||_K^2.  Chebyshev smoothing
of the guess with the operator itself (degree 16 on [lambda_max / 30, lambda_max], up to
three sweeps) brings that only to 6-28: the error is not node-level noise alone but sits
in cross-section modes 1e3-1e4 times stiffer than the beam's bending, where a 2% error
outweighs the whole solution.  Energy-scaled, every variant stays at 0.7-0.9 of the zero
guess's energy error, and neither the plain CG (3,075 vs 3,006 median iterations) nor the
two-level preconditioned CG (54-55 vs 53) gets shorter from it.
"""
from __future__ import annotations

import ctypes as C
import os
import time

import numpy as np
import torch

from . import capi
from .capi import call, ptr

DEV = "cuda"
HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "weights", "warmstart_beams.npz")
LAYERS = 4
FEAT = 16
OUT = 16


class BatchGraph:
    """Node graph, features, scales and node->instance map of a ``Batch``."""

    def __init__(self, batch):
        self.b = batch
        n = batch.n_nodes
        pat = C.c_void_p()
        call("tlfea_pattern_create", batch.nn, batch.n_elem, n, ptr(batch.conn), C.byref(pat), batch.stream)
        a, b_, c_ = C.c_int64(), C.c_int64(), C.c_int64()
        call("tlfea_pattern_sizes", pat, C.byref(a), C.byref(b_), C.byref(c_))
        self.off = torch.empty(n + 1, dtype=torch.int32, device=DEV)
        self.col = torch.empty(a.value, dtype=torch.int32, device=DEV)
        call("tlfea_pattern_node_csr", pat, ptr(self.off), ptr(self.col), batch.stream)
        capi.lib().tlfea_pattern_destroy(pat)
        counts = np.diff(batch.node_off)
        self.node_inst = torch.as_tensor(np.repeat(np.arange(batch.n_inst, dtype=np.int32), counts)).to(DEV)
        par = np.zeros((batch.n_inst, 12))
        scale = np.zeros((batch.n_inst, 3))
        for i, (m, mat) in enumerate(batch.items):
            X = m.xyz()
            lo, hi = X.min(axis=0), X.max(axis=0)
            ext = np.maximum(hi - lo, 1e-30)
            F = np.zeros(3) if m.f_ext is None else np.asarray(m.f_ext).reshape(-1, 3).sum(axis=0)
            par[i, 0:3], par[i, 3:6], par[i, 6:9] = lo, ext, F
            par[i, 9], par[i, 10] = mat["E"], mat["nu"]
            sb = np.linalg.norm(F) * ext[0] ** 3 / (mat["E"] * ext[1] * ext[2] * min(ext[1], ext[2]) ** 2)
            scale[i] = (sb * min(ext[1], ext[2]) / ext[0], sb, sb)
        self.scale_np = scale
        self.par = torch.as_tensor(par).to(DEV)
        self.scale = torch.as_tensor(scale.reshape(-1)).to(DEV)
        self.feat = torch.empty(n * FEAT, dtype=torch.int16, device=DEV)
        X0 = batch.X
        call("tlfea_mpnn_features", n, ptr(self.node_inst), ptr(X0[0]), ptr(X0[1]), ptr(X0[2]), ptr(batch.mask),
             ptr(batch.f), ptr(self.par), ptr(self.feat), batch.stream)

    def targets(self, u):
        """Scaled displacement targets [n, 3] (FP32) and loss weights (free nodes 1)."""
        s = self.scale.view(-1, 3)[self.node_inst.long()]
        t = (u.view(-1, 3) / s).float()
        w = (self.b.mask.view(-1, 3)[:, 0] == 0).float()
        return t.contiguous(), w.contiguous()


class WarmStart:
    def __init__(self, params=None, layers=LAYERS, seed=0):
        h = C.c_void_p()
        call("tlfea_mpnn_create", layers, C.byref(h))
        self.h = h
        self.layers = layers
        n = C.c_int()
        call("tlfea_mpnn_num_params", h, C.byref(n))
        self.n_params = n.value
        self.set_params(params if params is not None else init_params(layers, seed))

    # ---- parameters -------------------------------------------------------------------
    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.n_params:
            raise ValueError(f"expected {self.n_params} parameters, got {p.size}")
        call("tlfea_mpnn_set_params", self.h, p.ctypes.data)

    def params(self):
        p = np.zeros(self.n_params, np.float32)
        call("tlfea_mpnn_get_params", self.h, p.ctypes.data)
        return p

    def grads(self):
        g = np.zeros(self.n_params, np.float32)
        call("tlfea_mpnn_get_grads", self.h, g.ctypes.data)
        return g

    def save(self, path, **meta):
        os.makedirs(os.path.dirname(path), exist_ok=True)
        np.savez(path, params=self.params(), layers=self.layers, **{k: np.asarray(v) for k, v in meta.items()})

    @classmethod
    def load(cls, path=WEIGHTS):
        with np.load(path, allow_pickle=False) as z:
            return cls(params=z["params"], layers=int(z["layers"]))

    # ---- compute ----------------------------------------------------------------------
    def forward(self, feat, off, col, n, stream=None):
        out = torch.empty(n * OUT, dtype=torch.float32, device=DEV)
        call("tlfea_mpnn_forward", self.h, n, ptr(feat), ptr(off), ptr(col), ptr(out), stream)
        return out

    def loss_grad(self, feat, off, col, n, target, weight, stream=None):
        loss = C.c_double()
        call("tlfea_mpnn_loss_grad", self.h, n, ptr(feat), ptr(off), ptr(col), ptr(target), ptr(weight),
             C.byref(loss), stream)
        return loss.value

    def backward(self, n, off, col, dout, stream=None):
        """Backward of an FP32 [n, 3] output gradient through the last forward."""
        call("tlfea_mpnn_backward", self.h, n, ptr(off), ptr(col), ptr(dout), stream)

    def zero_grads(self, stream=None):
        call("tlfea_mpnn_zero_grads", self.h, stream)

    def adam(self, lr, beta1=0.9, beta2=0.999, eps=1e-8, stream=None):
        call("tlfea_mpnn_adam", self.h, float(lr), float(beta1), float(beta2), float(eps), stream)

    def guess(self, batch, graph=None, policy="energy", smooth=2):
        """u0 for every instance (device FP64).  With smooth > 0 (and the batch's Chebyshev
        stage set up) the raw prediction first gets `smooth` smoothing sweeps on K u = f,
        which damp its node-to-node noise.  policy "energy" (scale to the energy minimiser
        along the guess, fall back when f.u0 <= 0), "residual" (fall back when
        ||f - K u0|| > ||f||) or None (no test).  Per-instance flags and ratios land in
        self.fell_back, self.res_ratio (||f - K u0|| / ||f|| before the test) and
        self.alpha (the energy scaling)."""
        g = graph if graph is not None else BatchGraph(batch)
        n = batch.n_nodes
        out = self.forward(g.feat, g.off, g.col, n, batch.stream)
        u0 = torch.empty(3 * n, dtype=torch.float64, device=DEV)
        call("tlfea_mpnn_scale_output", n, ptr(g.node_inst), ptr(out), ptr(g.scale), ptr(batch.mask), ptr(u0),
             batch.stream)
        ni = batch.n_inst
        self.fell_back = np.zeros(ni, bool)
        self.alpha = np.ones(ni)
        if smooth and getattr(batch, "precond_degree", 0):
            batch.smooth(u0, sweeps=smooth)
        if policy is None:
            return u0
        fixed = batch.mask.bool()
        Ku = batch.tangent_apply(u0)
        f = batch.f.clone()
        f[fixed] = 0.0
        inst = g.node_inst.long().repeat_interleave(3)
        seg = lambda v: torch.zeros(ni, dtype=torch.float64, device=DEV).index_add_(0, inst, v)
        r = f - Ku
        r[fixed] = 0.0
        rr, ff = seg(r * r), seg(f * f)
        self.res_ratio = torch.sqrt(rr / ff.clamp_min(1e-300)).cpu().numpy()
        if policy == "residual":
            bad = (rr > ff).cpu().numpy()
            scale = torch.ones(ni, dtype=torch.float64, device=DEV)
        else:
            fu, uKu = seg(f * u0), seg(torch.where(fixed, torch.zeros_like(u0), u0 * Ku))
            ok = (fu > 0) & (uKu > 0)
            scale = torch.where(ok, fu / uKu.clamp_min(1e-300), torch.zeros_like(fu))
            bad = (~ok).cpu().numpy()
            self.alpha = scale.cpu().numpy()
        self.fell_back = bad
        s = torch.where(torch.as_tensor(bad).to(DEV), torch.zeros_like(scale), scale)
        return u0 * s[inst]

    def close(self):
        if getattr(self, "h", None) is not None:
            capi.lib().tlfea_mpnn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def init_params(layers=LAYERS, seed=0):
    """He-style uniform initialisation (the padded output rows are zero)."""
    rng = np.random.default_rng(seed)
    W, parts = 64, []
    shapes = [(W, FEAT), (W,)] + [(W, 2 * W), (W,)] * layers + [(OUT, W), (OUT,)]
    for k, shp in enumerate(shapes):
        if len(shp) == 2:
            lim = np.sqrt(6.0 / shp[1]) * (1.0 if k == 0 else 0.5)
            a = rng.uniform(-lim, lim, size=shp)
            if k == len(shapes) - 2:
                a[3:] = 0.0
                a[:3] *= 0.1
            parts.append(a.reshape(-1))
        else:
            parts.append(np.zeros(shp))
    return np.concatenate(parts).astype(np.float32)


def _sub_csr(off, col, n0, n1):
    """CSR rows [n0, n1) renumbered to start at 0 (instances never share nodes)."""
    o = off[n0:n1 + 1]
    c = col[o[0].item():o[-1].item()] - n0
    return (o - o[0]).contiguous(), c.contiguous()


def generate(n_inst, seed, group=256, tol=1e-8, log=print, two_level=True):
    """Converged FP64 solves of the beam sampler, returned as training groups.  two_level:
    the flexible FP64 CG with the FP32 two-level preconditioner (the same solutions to the
    same tolerance as the plain CG, about ten times faster on these batches)."""
    from . import mesh as M
    from .batch import Batch
    data = []
    done = 0
    while done < n_inst:
        k = min(group, n_inst - done)
        items = M.random_beam_batch(k, seed + done)
        b = Batch(items)
        if two_level:
            b.precond_setup(degree=6, ratio=250.0, power_iters=20)
            b.coarse_setup()
        u = b.solve(tol=tol, precond=two_level)
        g = BatchGraph(b)
        t, w = g.targets(u)
        ok = b.status == 0
        data.append(dict(feat=g.feat, off=g.off, col=g.col, t=t, w=w, node_off=b.node_off.copy(), ok=ok))
        done += k
        log(f"[learned] generated {done}/{n_inst} solves ({int(ok.sum())} of {k} converged in this group)")
        b.close()
    return data


def train(data, epochs=20, inst_per_step=32, lr=2e-3, lr_final=2e-4, seed=0, net=None, log=print):
    net = net or WarmStart(params=init_params(LAYERS, seed))
    rng = np.random.default_rng(seed)
    steps = []
    for gi, d in enumerate(data):
        ni = len(d["node_off"]) - 1
        for i0 in range(0, ni, inst_per_step):
            steps.append((gi, i0, min(ni, i0 + inst_per_step)))
    total = epochs * len(steps)
    k = 0
    t0 = time.time()
    hist = []
    for ep in range(epochs):
        order = rng.permutation(len(steps))
        el = 0.0
        for j in order:
            gi, i0, i1 = steps[j]
            d = data[gi]
            n0, n1 = int(d["node_off"][i0]), int(d["node_off"][i1])
            off, col = _sub_csr(d["off"], d["col"], n0, n1)
            lr_k = lr * (lr_final / lr) ** (k / max(1, total - 1))
            net.zero_grads()
            el += net.loss_grad(d["feat"][n0 * FEAT:n1 * FEAT], off, col, n1 - n0, d["t"][n0:n1], d["w"][n0:n1])
            net.adam(lr_k)
            k += 1
        hist.append(el / len(steps))
        log(f"[learned] epoch {ep + 1}/{epochs} mean loss {hist[-1]:.4e} ({time.time() - t0:.0f}s)")
    return net, hist


def generate_energy(n_inst, seed, group=256, tol=1e-8, log=print):
    """Like ``generate`` but keeps each group's batch (its operator K) for the energy loss."""
    from . import mesh as M
    from .batch import Batch
    data, done = [], 0
    while done < n_inst:
        k = min(group, n_inst - done)
        b = Batch(M.random_beam_batch(k, seed + done))
        u = b.solve(tol=tol)
        g = BatchGraph(b)
        inst = g.node_inst.long().repeat_interleave(3)
        f = b.f.clone()
        f[b.mask.bool()] = 0.0
        E = torch.zeros(b.n_inst, dtype=torch.float64, device=DEV).index_add_(0, inst, f * u)  # u*.K u* = f.u*
        data.append(dict(batch=b, graph=g, u=u, E=E, inst=inst, node_off=b.node_off.copy()))
        done += k
        log(f"[learned] generated {done}/{n_inst} solves for the energy loss")
    return data


def train_energy(data, net, epochs=4, inst_per_step=32, lr=5e-4, lr_final=5e-5, seed=0, log=print):
    """Minimises the mean over instances of ||scale * out - u*||_K^2 / ||u*||_K^2, the error
    measure CG minimises, with K applied by the batch's matrix-free FP64 operator."""
    rng = np.random.default_rng(seed)
    steps = []
    for gi, d in enumerate(data):
        ni = len(d["node_off"]) - 1
        for i0 in range(0, ni, inst_per_step):
            steps.append((gi, i0, min(ni, i0 + inst_per_step)))
    total, k, hist, t0 = epochs * len(steps), 0, [], time.time()
    for ep in range(epochs):
        el = 0.0
        for j in rng.permutation(len(steps)):
            gi, i0, i1 = steps[j]
            d = data[gi]
            b, g = d["batch"], d["graph"]
            n0, n1 = int(d["node_off"][i0]), int(d["node_off"][i1])
            ns = n1 - n0
            off, col = _sub_csr(g.off, g.col, n0, n1)
            out = net.forward(g.feat[n0 * FEAT:n1 * FEAT], off, col, ns)
            us = torch.empty(3 * ns, dtype=torch.float64, device=DEV)
            ninst = g.node_inst[n0:n1].contiguous()
            call("tlfea_mpnn_scale_output", ns, ptr(ninst), ptr(out), ptr(g.scale), ptr(b.mask[3 * n0:3 * n1]),
                 ptr(us), None)
            e = torch.zeros(3 * b.n_nodes, dtype=torch.float64, device=DEV)
            e[3 * n0:3 * n1] = us - d["u"][3 * n0:3 * n1]
            Ke = b.tangent_apply(e)
            woe = torch.zeros(b.n_inst, dtype=torch.float64, device=DEV)
            woe[i0:i1] = 1.0 / ((i1 - i0) * d["E"][i0:i1])
            dout = torch.empty(3 * ns, dtype=torch.float32, device=DEV)
            call("tlfea_mpnn_energy_grad", ns, ptr(ninst), ptr(g.scale), ptr(Ke[3 * n0:3 * n1].contiguous()),
                 ptr(woe), ptr(dout), None)
            net.zero_grads()
            net.backward(ns, off, col, dout)
            net.adam(lr * (lr_final / lr) ** (k / max(1, total - 1)))
            seg = torch.zeros(b.n_inst, dtype=torch.float64, device=DEV).index_add_(0, d["inst"], e * Ke)
            el += float((seg[i0:i1] / d["E"][i0:i1]).mean())
            k += 1
        hist.append(el / len(steps))
        log(f"[learned] energy epoch {ep + 1}/{epochs} mean relative energy error {hist[-1]:.4e} ({time.time() - t0:.0f}s)")
    return net, hist
