"""CPU checks of the network restatement (oracle/mpnn.py) that the GPU kernels are compared with.

The float64 backward is pinned against central finite differences of the float64 forward:
12 random coordinates of every parameter block (relative error 1e-6 above the 1e-9
difference noise) and directional
derivatives along random directions (1e-7), so a dropped term, a transposed weight or a
missing residual path in the restated backward fails here.
"""
import numpy as np

import oracle as O
from oracle import mpnn as N


def _graph():
    from tlfea_amd import mesh as M
    m = M.kuhn_beam(3, 1, 1)
    off, col = O.node_pattern(m)
    return m, np.asarray(off), np.asarray(col)


def test_bf16_rounding():
    a = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -8 + 2 ** -20, 1.0 + 3 * 2 ** -8, -3.14159], np.float32)
    b = N.bf16(a)
    # ties to even: 1 + 2^-8 is halfway between 1 and 1 + 2^-7 -> 1; above halfway -> up
    assert b[0] == 1.0 and b[1] == 1.0 and b[2] == np.float32(1.0 + 2 ** -7) and b[3] == np.float32(1.0 + 2 ** -6)
    assert abs(b[4] - a[4]) <= 2 ** -8 * abs(a[4])


def test_backward_matches_finite_differences():
    m, off, col = _graph()
    n = m.n_nodes
    rng = np.random.default_rng(0)
    x = rng.standard_normal((n, N.FEAT))
    t = rng.standard_normal((n, 3))
    w = (rng.random(n) > 0.2).astype(np.float64)
    p = N.init_params(4, seed=1).astype(np.float64)
    out, cache = N.forward(p, x, off, col)
    g = N.backward(cache, out, t, w)
    f = lambda q: N.loss(N.forward(q, x, off, col)[0], t, w)
    h = 1e-6
    lay, _ = N.layout(4)
    for name, shp, o in lay:
        sz = int(np.prod(shp))
        live = sz if name not in ("W_out", "b_out") else 3 * (sz // 16)
        if name in ("W_out", "b_out"):  # padded rows carry no gradient
            assert np.all(g[o + live:o + sz] == 0.0)
        for k in rng.choice(live, size=min(12, live), replace=False):
            e = np.zeros_like(p)
            e[o + k] = h
            fd = (f(p + e) - f(p - e)) / (2 * h)
            # 1e-6 relative, above the finite-difference noise (~1e-10 for an O(1) loss)
            assert abs(g[o + k] - fd) <= 1e-6 * abs(fd) + 1e-9, (name, k)
    for _ in range(4):
        v = rng.standard_normal(len(p))
        lay_mask = np.ones_like(p)
        for name, shp, o in lay:
            if name in ("W_out", "b_out"):
                sz = int(np.prod(shp))
                lay_mask[o + 3 * (sz // 16):o + sz] = 0.0
        v *= lay_mask
        hv = 1e-6 / np.linalg.norm(v) * np.linalg.norm(p)
        fd = (f(p + hv * v) - f(p - hv * v)) / (2 * hv)
        assert abs(g @ v - fd) <= 1e-7 * np.linalg.norm(g) * np.linalg.norm(v)


def test_emulated_forward_close_to_float64():
    m, off, col = _graph()
    rng = np.random.default_rng(3)
    x = rng.standard_normal((m.n_nodes, N.FEAT))
    p = N.init_params(4, seed=2)
    o64, _ = N.forward(p.astype(np.float64), x, off, col)
    o16, _ = N.forward(p, x, off, col, emulate=True)
    rel = np.linalg.norm(o16[:, :3] - o64[:, :3]) / np.linalg.norm(o64[:, :3])
    assert 1e-5 < rel < 5e-2, rel  # BF16 rounding is visible but small
