"""GPU checks of the learned warm-start network against its CPU restatement (oracle/mpnn.py).

Bars: forward outputs within 5e-3 (norm-wise) of the restatement that applies the same
BF16 rounding points (the remaining difference is FP32 accumulation order inside the
matrix cores, which can move a BF16 rounding by one unit); gradients within 2e-2 per
parameter block of that restatement and within 5e-2 of the float64 gradient.
"""
import numpy as np
import pytest

import oracle as O
from oracle import mpnn as N

pytestmark = pytest.mark.gpu


def _graph(nx=10, ny=2, nz=2):
    import torch
    from tlfea_amd import mesh as M
    m = M.kuhn_beam(nx, ny, nz)
    off, col = O.node_pattern(m)
    return m, np.asarray(off), np.asarray(col), torch.as_tensor(np.asarray(off, np.int32)).cuda(), \
        torch.as_tensor(np.asarray(col, np.int32)).cuda()


def _feat(n, seed):
    import torch
    x = N.bf16(np.random.default_rng(seed).standard_normal((n, N.FEAT)).astype(np.float32))
    bits = (x.view(np.uint32) >> 16).astype(np.uint16).view(np.int16)
    return x, torch.as_tensor(bits.reshape(-1).copy()).cuda()


def test_forward_matches_restatement(gpu):
    from tlfea_amd.learned import WarmStart, init_params
    m, off, col, doff, dcol = _graph()
    n = m.n_nodes
    x, xd = _feat(n, 1)
    p = init_params(4, seed=3)
    net = WarmStart(params=p)
    out = net.forward(xd, doff, dcol, n).cpu().numpy().reshape(n, N.OUT)
    ref, _ = N.forward(p, x, off, col, emulate=True)
    assert np.linalg.norm(out[:, :3] - ref[:, :3]) <= 5e-3 * np.linalg.norm(ref[:, :3])
    assert np.all(out[:, 3:] == 0.0)


def test_gradients_match_restatement(gpu):
    import torch
    from tlfea_amd.learned import WarmStart, init_params
    m, off, col, doff, dcol = _graph()
    n = m.n_nodes
    x, xd = _feat(n, 2)
    rng = np.random.default_rng(4)
    t = rng.standard_normal((n, 3)).astype(np.float32)
    w = (rng.random(n) > 0.1).astype(np.float32)
    p = init_params(4, seed=5)
    net = WarmStart(params=p)
    net.zero_grads()
    loss = net.loss_grad(xd, doff, dcol, n, torch.as_tensor(t).cuda(), torch.as_tensor(w).cuda())
    g = net.grads().astype(np.float64)
    out16, c16 = N.forward(p, x, off, col, emulate=True)
    g16 = N.backward(c16, out16, t, w)
    out64, c64 = N.forward(p.astype(np.float64), x.astype(np.float64), off, col)
    g64 = N.backward(c64, out64, t.astype(np.float64), w.astype(np.float64))
    assert abs(loss - N.loss(out16, t, w)) <= 1e-3 * abs(loss)
    for name, shp, o in N.layout(4)[0]:
        sz = int(np.prod(shp))
        a, b, c = g[o:o + sz], g16[o:o + sz], g64[o:o + sz]
        assert np.linalg.norm(a - b) <= 2e-2 * np.linalg.norm(b) + 1e-12, name
        assert np.linalg.norm(a - c) <= 5e-2 * np.linalg.norm(c) + 1e-12, name
    # gradients accumulate until zeroed
    net.loss_grad(xd, doff, dcol, n, torch.as_tensor(t).cuda(), torch.as_tensor(w).cuda())
    assert np.allclose(net.grads(), 2 * g, rtol=1e-5, atol=1e-9)


def test_adam_training_reduces_loss(gpu):
    import torch
    from tlfea_amd.learned import WarmStart, init_params
    m, off, col, doff, dcol = _graph(6, 2, 2)
    n = m.n_nodes
    x, xd = _feat(n, 6)
    # a smooth target the network can fit: a cubic in the first feature
    t = np.stack([x[:, 0] ** 3 / 4, 0.5 * x[:, 1], np.zeros(n)], axis=1).astype(np.float32)
    w = np.ones(n, np.float32)
    net = WarmStart(params=init_params(4, seed=7))
    tt, wt = torch.as_tensor(t).cuda(), torch.as_tensor(w).cuda()
    losses = []
    for _ in range(60):
        net.zero_grads()
        losses.append(net.loss_grad(xd, doff, dcol, n, tt, wt))
        net.adam(3e-3)
    assert losses[-1] < 0.3 * losses[0], (losses[0], losses[-1])
    # the padded output rows are never trained
    p = net.params()
    lay = {nme: (shp, o) for nme, shp, o in N.layout(4)[0]}
    shp, o = lay["W_out"]
    assert np.all(p[o + 3 * 64:o + 16 * 64] == 0.0)


def test_warm_start_on_held_out_beams(gpu):
    """The committed weights (trained on 20,000 converged solves of the beam sampler) on
    held-out beams: what the network gains over the zero guess is the prediction itself --
    the raw guess is within 5% (median L2; 2.5% measured on 256 held-out beams) of the FP64
    solution where the zero guess is 100% off; CG from the energy-scaled guess reaches the
    same solution; the zero-weight network falls back on every instance.  The iteration
    counts of both starts are printed: on these slender beams the guess does not shorten CG
    (its error, small in L2, sits in modes 1e3-1e4 times stiffer than the beam's bending
    and carries more energy than the solution, see the learned module's notes), and the
    energy scaling keeps it from costing more than a few per cent."""
    import torch
    from tlfea_amd import mesh as M
    from tlfea_amd.batch import Batch
    from tlfea_amd.learned import BatchGraph, WarmStart, init_params
    items = M.random_beam_batch(16, 9100000)
    b = Batch(items)
    g = BatchGraph(b)
    u_zero = b.solve(tol=1e-8)
    it_zero = b.iters.copy()
    net = WarmStart.load()
    raw = net.guess(b, graph=g, policy=None, smooth=0)
    err = [torch.linalg.norm(raw[b.inst_slice(i)] - u_zero[b.inst_slice(i)]).item()
           / torch.linalg.norm(u_zero[b.inst_slice(i)]).item() for i in range(b.n_inst)]
    assert np.median(err) < 0.05, err
    u0 = net.guess(b, graph=g)
    assert net.fell_back.sum() <= 1
    u_warm = b.solve(tol=1e-8, u0=u0)
    assert b.rc == 0
    it_warm = b.iters.copy()
    for i in range(b.n_inst):
        sl = b.inst_slice(i)
        assert (torch.linalg.norm(u_warm[sl] - u_zero[sl]) / torch.linalg.norm(u_zero[sl])).item() < 1e-6
    print(f"[warm start] iterations zero-start {it_zero.sum()} warm-start {it_warm.sum()} "
          f"ratio {it_warm.sum() / it_zero.sum():.3f}", flush=True)
    assert it_warm.sum() <= 1.1 * it_zero.sum()
    # a network that predicts nothing falls back to the zero guess everywhere
    zero_net = WarmStart(params=np.zeros_like(init_params()))
    u00 = zero_net.guess(b, graph=g)
    assert zero_net.fell_back.all() and torch.count_nonzero(u00).item() == 0
    u_fb = b.solve(tol=1e-8, u0=u00)
    for i in range(b.n_inst):
        sl = b.inst_slice(i)
        assert (torch.linalg.norm(u_fb[sl] - u_zero[sl]) / torch.linalg.norm(u_zero[sl])).item() < 1e-6
    # the same zero-start solve twice: the tangent's in-block LDS atomics change the
    # summation order between runs, and on these beams (kappa up to 1e9) finite-precision
    # CG counts move with it (measured up to 8.6%: 12,869 against 14,074)
    assert np.all(np.abs(b.iters - it_zero) <= np.maximum(3, 0.2 * it_zero))
