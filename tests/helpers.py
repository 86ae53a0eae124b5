"""Shared test fixtures: meshes and deformed states (seeded)."""
import os

import numpy as np

from tlfea_amd import mesh as M

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MESHES = os.path.join(GOLDEN, "meshes")

E_STEEL, NU_STEEL, RHO_STEEL = 2.1e11, 0.3, 7850.0


def lame(E, nu):
    return (E * nu) / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))


def load_mesh(name):
    """Meshes shipped by the reference (data/meshes, BSD-3) and the synthetic beams."""
    if name == "cube":
        m = M.read_tetgen(os.path.join(MESHES, "cube.1.node"), os.path.join(MESHES, "cube.1.ele"))
        m.fixed = np.nonzero(np.abs(m.z) < 1e-12)[0].astype(np.int32)
        f = np.zeros(3 * m.n_nodes)
        f[3 * 6 + 0] = 1000.0  # test_feat10_nesterov.cc: 1000 N in +x on node 6
        m.f_ext = f
        return m
    if name == "beam_vtu":
        return M.clamp_and_tip_load(M.read_vtu_ascii(os.path.join(MESHES, "beam.vtu")), -1000.0)
    if name == "beam_highres_t10":
        return M.clamp_and_tip_load(M.read_vtu_ascii(os.path.join(MESHES, "beam_highres_t10.vtu")), -1000.0)
    if name == "beam_highres_t4":
        return M.clamp_and_tip_load(M.read_vtu_ascii(os.path.join(MESHES, "beam_highres_t4.vtu")), -1000.0)
    if name == "C1":
        return M.beam_config("C1")
    if name == "kuhn_t4":
        return M.clamp_and_tip_load(M.kuhn_beam(12, 3, 2, tet10=False), -500.0)
    if name == "C2":
        return M.beam_config("C2")
    raise KeyError(name)


def deformed(m, seed=0, amp=1e-3, rot=0.2):
    """Positions after a rigid rotation about z plus a seeded random perturbation scaled by
    the mesh size (large enough to make the SVK response nonlinear)."""
    rng = np.random.default_rng(seed)
    X = m.xyz()
    L = np.ptp(X, axis=0).max()
    c, s = np.cos(rot), np.sin(rot)
    R = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1.0]])
    Y = X @ R.T + amp * L * rng.standard_normal(X.shape)
    return Y[:, 0].copy(), Y[:, 1].copy(), Y[:, 2].copy()


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def _inv3(J):
    """3x3 inverses and determinants of a stack, by cofactors (works in longdouble)."""
    c = np.empty_like(J)
    c[:, 0, 0] = J[:, 1, 1] * J[:, 2, 2] - J[:, 1, 2] * J[:, 2, 1]
    c[:, 0, 1] = J[:, 0, 2] * J[:, 2, 1] - J[:, 0, 1] * J[:, 2, 2]
    c[:, 0, 2] = J[:, 0, 1] * J[:, 1, 2] - J[:, 0, 2] * J[:, 1, 1]
    c[:, 1, 0] = J[:, 1, 2] * J[:, 2, 0] - J[:, 1, 0] * J[:, 2, 2]
    c[:, 1, 1] = J[:, 0, 0] * J[:, 2, 2] - J[:, 0, 2] * J[:, 2, 0]
    c[:, 1, 2] = J[:, 0, 2] * J[:, 1, 0] - J[:, 0, 0] * J[:, 1, 2]
    c[:, 2, 0] = J[:, 1, 0] * J[:, 2, 1] - J[:, 1, 1] * J[:, 2, 0]
    c[:, 2, 1] = J[:, 0, 1] * J[:, 2, 0] - J[:, 0, 0] * J[:, 2, 1]
    c[:, 2, 2] = J[:, 0, 0] * J[:, 1, 1] - J[:, 0, 1] * J[:, 1, 0]
    det = J[:, 0, 0] * c[:, 0, 0] + J[:, 0, 1] * c[:, 1, 0] + J[:, 0, 2] * c[:, 2, 0]
    return c / det[:, None, None], det


def fint_extended(m, x, y, z, lam, mu, u=None):
    """SVK internal force (compute_internal_force, FEAT10DataFunc.cuh:404-460, with the
    gradients of FEAT10Data.cu:92-198) in numpy longdouble (64-bit mantissa), with the
    strain from the displacement gradient: a reference for small-strain states, where
    forming F^T F - I in double loses digits.  ``u`` (n_nodes x 3) gives the displacement
    itself, so that no rounding of x = X + u to the coordinates' precision enters."""
    import oracle as O
    ld = np.longdouble
    qx, qy, qz, qw = O.quad(m.nn)
    X = m.xyz()[m.elems].astype(ld)
    if u is None:
        xc = np.stack([x, y, z], axis=1)[m.elems].astype(ld)
        u = xc - X  # exact: differences of doubles fit the longer mantissa here
    else:
        u = np.asarray(u, dtype=ld)[m.elems]
    f = np.zeros((m.n_nodes, 3), dtype=ld)
    lam, mu = ld(lam), ld(mu)
    edges = ((0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3))
    for q in range(len(qw)):
        Lq = [ld(1) - ld(qx[q]) - ld(qy[q]) - ld(qz[q]), ld(qx[q]), ld(qy[q]), ld(qz[q])]
        dL = np.array([[-1, -1, -1], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=ld)
        if m.nn == 4:
            dN = dL
        else:
            dN = np.zeros((10, 3), dtype=ld)
            for a in range(4):
                dN[a] = (4 * Lq[a] - 1) * dL[a]
            for k, (i, j) in enumerate(edges):
                dN[4 + k] = 4 * (Lq[i] * dL[j] + Lq[j] * dL[i])
        J = np.einsum("eai,ak->eik", X, dN)
        Ji, det = _inv3(J)
        g = np.einsum("ak,ekj->eaj", dN, Ji)
        H = np.einsum("eai,eaj->eij", u, g)
        Em = (H + H.transpose(0, 2, 1) + np.einsum("emi,emj->eij", H, H)) / 2
        trE = Em[:, 0, 0] + Em[:, 1, 1] + Em[:, 2, 2]
        S = 2 * mu * Em + lam * trE[:, None, None] * np.eye(3, dtype=ld)
        P = S + np.einsum("eik,ekj->eij", H, S)
        fe = np.einsum("eij,eaj->eai", P, g) * (det * ld(qw[q]))[:, None, None]
        np.add.at(f, m.elems, fe)
    return f.reshape(-1)
