"""CPU restatement of the learned warm-start network (test infrastructure only).

The same network as tlfea_amd/csrc/mpnn.hip, written out with numpy:

    h0 = relu(x W_in^T + b_in)
    for l = 1..L:  m = mean_{j in N(i)} h_j     (CSR rows, diagonal included)
                   pre_l = [h, m] W_l^T + b_l
                   h = h + relu(pre_l)
    out = h W_out^T + b_out                    (3 outputs used of 16)
    loss = sum_i w_i |out_i[:3] - t_i|^2 / sum_i w_i

``forward``/``backward`` with ``emulate=False`` run in float64 without rounding: the exact
gradient of the float64 forward, pinned in the tests against central finite differences.
With ``emulate=True`` they apply the kernels' rounding points: BF16 (round to nearest
even) weights, features, activations, aggregated means and, in the backward, the
gradients that enter a GEMM; FP32 accumulation and FP32 biases.  The GPU gradients are
compared with that restatement.

There is no reference counterpart: the network is an addition to the reference's solve path.
Parameter packing (FP32 vector): W_in [64x16], b_in [64], per layer W_l [64x128], b_l [64],
W_out [16x64], b_out [16].
"""
from __future__ import annotations

import numpy as np

W = 64
FEAT = 16
OUT = 16


def layout(layers=4):
    """[(name, shape, offset)] of the packed parameter vector."""
    out, off = [], 0
    shapes = [("W_in", (W, FEAT)), ("b_in", (W,))]
    for l in range(1, layers + 1):
        shapes += [(f"W_{l}", (W, 2 * W)), (f"b_{l}", (W,))]
    shapes += [("W_out", (OUT, W)), ("b_out", (OUT,))]
    for name, shp in shapes:
        out.append((name, shp, off))
        off += int(np.prod(shp))
    return out, off


def unpack(vec, layers=4):
    lay, _ = layout(layers)
    return {n: np.asarray(vec[o:o + int(np.prod(s))]).reshape(s) for n, s, o in lay}


def pack(d, layers=4):
    lay, n = layout(layers)
    v = np.zeros(n, dtype=np.float64)
    for name, shp, o in lay:
        v[o:o + int(np.prod(shp))] = np.asarray(d[name]).reshape(-1)
    return v


def init_params(layers=4, seed=0):
    """He-style uniform initialisation; the padded output rows 3..15 are zero."""
    rng = np.random.default_rng(seed)
    d = {}
    for name, shp, _ in layout(layers)[0]:
        if name.startswith("W"):
            fan_in = shp[1]
            lim = np.sqrt(6.0 / fan_in) * (0.5 if name != "W_in" else 1.0)
            d[name] = rng.uniform(-lim, lim, size=shp)
        else:
            d[name] = np.zeros(shp)
    d["W_out"][3:] = 0.0
    d["W_out"][:3] *= 0.1
    return pack(d, layers).astype(np.float32)


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def _agg(h, off, col, emulate):
    deg = np.diff(off)
    rows = np.repeat(np.arange(len(off) - 1), deg)
    s = np.zeros_like(h)
    np.add.at(s, rows, h[col])
    m = s / deg[:, None].astype(h.dtype)
    return bf16(m) if emulate else m


def _agg_t(dm, off, col):
    deg = np.diff(off).astype(dm.dtype)
    rows = np.repeat(np.arange(len(off) - 1), np.diff(off))
    out = np.zeros_like(dm)
    # symmetric graph: dh_j += sum_{i in row j} dm_i / deg(i)
    np.add.at(out, rows, dm[col] / deg[col][:, None])
    return out


def forward(params, x, off, col, layers=4, emulate=False):
    """Returns (out [n, 16], cache)."""
    dt = np.float32 if emulate else np.float64
    P = {k: v.astype(dt) for k, v in unpack(np.asarray(params, dtype=np.float64), layers).items()}
    Wq = {k: (bf16(v) if emulate else v) for k, v in P.items() if k.startswith("W")}
    x = bf16(np.asarray(x, np.float32)) if emulate else np.asarray(x, np.float64)
    relu = lambda v: np.maximum(v, 0)
    rnd = bf16 if emulate else (lambda v: v)
    pre = [x @ Wq["W_in"].T + P["b_in"]]
    hs = [rnd(relu(pre[0]))]
    ms = []
    for l in range(1, layers + 1):
        h = hs[-1]
        m = _agg(h, off, col, emulate)
        ms.append(m)
        Wl = Wq[f"W_{l}"]
        p = h @ Wl[:, :W].T + m @ Wl[:, W:].T + P[f"b_{l}"]
        pre.append(p)
        hs.append(rnd(h + relu(p)))
    out = hs[-1] @ Wq["W_out"].T + P["b_out"]
    return out, dict(x=x, pre=pre, hs=hs, ms=ms, Wq=Wq, emulate=emulate, layers=layers, off=off, col=col)


def loss(out, target, w):
    d = out[:, :3] - target
    return float(np.sum(w[:, None] * d * d) / np.sum(w))


def backward(cache, out, target, w):
    """Gradient of ``loss`` w.r.t. the packed parameters (float64 vector)."""
    emulate = cache["emulate"]
    dt = np.float32 if emulate else np.float64
    rnd = bf16 if emulate else (lambda v: v)
    L, off, col = cache["layers"], cache["off"], cache["col"]
    Wq, pre, hs, ms, x = cache["Wq"], cache["pre"], cache["hs"], cache["ms"], cache["x"]
    wsum = np.sum(w, dtype=np.float64)
    dout = np.zeros_like(out, dtype=dt)
    dout[:, :3] = (2.0 * w[:, None] * (out[:, :3] - target) / wsum).astype(dt)
    g = {}
    d16 = rnd(dout)
    g["W_out"] = d16.T @ hs[-1]
    g["b_out"] = dout.sum(axis=0)
    dh = d16 @ Wq["W_out"]
    for l in range(L, 0, -1):
        dp = dh * (pre[l] > 0)
        dp16 = rnd(dp)
        Wl = Wq[f"W_{l}"]
        g[f"W_{l}"] = np.concatenate([dp16.T @ hs[l - 1], dp16.T @ ms[l - 1]], axis=1)
        g[f"b_{l}"] = dp.sum(axis=0)
        dm = dp16 @ Wl[:, W:]
        dh = dh + dp16 @ Wl[:, :W] + _agg_t(dm, off, col)
    dp = dh * (pre[0] > 0)
    dp16 = rnd(dp)
    g["W_in"] = dp16.T @ x
    g["b_in"] = dp.sum(axis=0)
    g["W_out"][3:] = 0.0
    g["b_out"][3:] = 0.0
    return pack({k: np.asarray(v, np.float64) for k, v in g.items()}, L)
