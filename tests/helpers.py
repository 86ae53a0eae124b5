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
