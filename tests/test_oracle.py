"""Pins the FP64 CPU oracle (CPU only).

The reference holds no golden vectors or known-answer tests, so every oracle function is
checked against what the mathematics fixes -- closed forms, invariants, brute force and
independent routes -- chosen so that a dropped term, a wrong sign or index or a transposed
operand fails one of them.
"""
import itertools
import math

import numpy as np
import pytest

import oracle as O
from helpers import E_STEEL, NU_STEEL, RHO_STEEL, deformed, lame, load_mesh, rel_err
from tlfea_amd import mesh as M

SMALL = ["cube", "beam_vtu", "C1", "kuhn_t4"]


def _tet_volume(X):
    return np.linalg.det(np.stack([X[1] - X[0], X[2] - X[0], X[3] - X[0]])) / 6.0


# ---------------------------------------------------------------------------------------
def test_quadrature_constants_and_exactness():
    qx, qy, qz, qw = O.quad_t10()
    assert qw[0] == -4.0 / 5.0 / 6.0 and all(w == 9.0 / 20.0 / 6.0 for w in qw[1:])
    assert abs(qw.sum() - 1.0 / 6.0) < 1e-17
    # exact on x^a y^b z^c with a+b+c <= 3 over the unit tet: a! b! c! / (a+b+c+3)!
    for a, b, c in itertools.product(range(4), repeat=3):
        if a + b + c > 3:
            continue
        exact = math.factorial(a) * math.factorial(b) * math.factorial(c) / math.factorial(a + b + c + 3)
        approx = float(np.sum(qw * qx ** a * qy ** b * qz ** c))
        assert abs(approx - exact) < 1e-16, (a, b, c)
    # degree 4 is NOT exact (the rule is degree 3): guards against a wrong table
    exact4 = math.factorial(4) / math.factorial(7)
    assert abs(float(np.sum(qw * qx ** 4)) - exact4) > 1e-6
    tx, ty, tz, tw = O.quad_t4()
    assert tx[0] == ty[0] == tz[0] == 0.25 and tw[0] == 1.0 / 6.0


# ---------------------------------------------------------------------------------------
@pytest.mark.parametrize("name", SMALL + ["beam_highres_t10", "beam_highres_t4"])
def test_grad_invariants(name):
    m = load_mesh(name)
    g, dj = O.grad(m)
    nn = m.nn
    nq = 5 if nn == 10 else 1
    G = g.reshape(m.n_elem, nq, 3, nn)  # [e, q, d, a]  (grad[a + nn*d])
    # partition of unity: sum_a g_a = 0
    assert np.abs(G.sum(axis=3)).max() < 1e-9 * np.abs(G).max()
    # sum_a X_a (x) g_a = I at every quadrature point
    X = m.xyz()[m.elems]  # [e, a, i]
    ident = np.einsum("eai,eqda->eqid", X, G)
    assert np.abs(ident - np.eye(3)).max() < 1e-10
    # straight-sided tets: det J = 6 * volume at every quadrature point
    V = np.array([_tet_volume(X[e, :4]) for e in range(m.n_elem)])
    assert rel_err(dj.reshape(m.n_elem, nq), np.repeat(6 * V[:, None], nq, 1)) < 1e-12
    assert np.all(dj > 0)


def test_grad_affine_closed_form():
    """For a straight-sided element g_a = J^{-T} dN_a/dxi with J = [X1-X0, X2-X0, X3-X0]
    (an explicit numpy inverse, independent of the pivoting solve)."""
    m = load_mesh("beam_vtu")
    g, dj = O.grad(m)
    qx, qy, qz, _ = O.quad_t10()
    G = g.reshape(m.n_elem, 5, 3, 10)
    for e in range(m.n_elem):
        X = m.xyz()[m.elems[e]]
        J = np.stack([X[1] - X[0], X[2] - X[0], X[3] - X[0]], 1)
        JinvT = np.linalg.inv(J).T
        for q in range(5):
            L = np.array([1 - qx[q] - qy[q] - qz[q], qx[q], qy[q], qz[q]])
            dL = np.array([[-1, -1, -1], [1, 0, 0], [0, 1, 0], [0, 0, 1.0]])
            dN = np.zeros((10, 3))
            for i in range(4):
                dN[i] = (4 * L[i] - 1) * dL[i]
            for k, (i, j) in enumerate(M.T10_EDGES):
                dN[4 + k] = 4 * (L[i] * dL[j] + L[j] * dL[i])
            ref = dN @ JinvT.T  # rows g_a
            assert np.allclose(G[e, q].T, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def test_beam_vtu_detj_value():
    """beam.vtu: 24 straight TET10 of volume 10/24 -> det J = 2.5 everywhere."""
    m = load_mesh("beam_vtu")
    _, dj = O.grad(m)
    assert np.allclose(dj, 2.5, rtol=1e-13)


# ---------------------------------------------------------------------------------------
def _single_elem_mesh(F=None):
    m = M.kuhn_beam(1, 1, 1, 1.0, 1.0, 1.0)
    m.elems = m.elems[:1].copy()
    return m


def _p_at(F, lam, mu):
    """P of the oracle for a homogeneous deformation x = F X on one element."""
    m = _single_elem_mesh()
    g, _ = O.grad(m)
    X = m.xyz()
    x = X @ F.T
    _, P, _, _ = O.calc_p(m, g, lam, mu, x[:, 0].copy(), x[:, 1].copy(), x[:, 2].copy())
    return P.reshape(5, 3, 3).transpose(0, 2, 1)  # stored column-major


def test_svk_closed_forms_and_energy_derivative():
    lam, mu = lame(E_STEEL, NU_STEEL)
    P = _p_at(np.eye(3), lam, mu)
    assert np.abs(P).max() < 1e-3  # rounding of E * 1e-16
    th = 0.7
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    assert np.abs(_p_at(R, lam, mu)).max() < 1e-2
    s = 1.01
    P = _p_at(np.diag([s, 1, 1]), lam, mu)[0]
    e11 = 0.5 * (s * s - 1)
    assert abs(P[0, 0] - s * (lam + 2 * mu) * e11) < 1e-12 * abs(P[0, 0])
    assert abs(P[1, 1] - lam * e11) < 1e-12 * abs(P[1, 1])
    assert abs(P[0, 1]) < 1e-3
    # P = dW/dF with W = lambda/2 (tr E)^2 + mu tr(E^2), central differences
    rng = np.random.default_rng(0)
    F = np.eye(3) + 0.05 * rng.standard_normal((3, 3))

    def W(F):
        E = 0.5 * (F.T @ F - np.eye(3))
        return 0.5 * lam * np.trace(E) ** 2 + mu * np.trace(E @ E)

    P = _p_at(F, lam, mu)[0]
    h = 1e-7
    for i in range(3):
        for j in range(3):
            Fp, Fm = F.copy(), F.copy()
            Fp[i, j] += h
            Fm[i, j] -= h
            assert abs((W(Fp) - W(Fm)) / (2 * h) - P[i, j]) < 1e-6 * np.abs(P).max()


@pytest.mark.parametrize("name", SMALL)
def test_internal_force_invariants(name):
    m = load_mesh(name)
    g, dj = O.grad(m)
    lam, mu = lame(E_STEEL, NU_STEEL)
    f0 = O.internal_force(m, g, dj, lam, mu)
    x, y, z = deformed(m, seed=11)
    f = O.internal_force(m, g, dj, lam, mu, x, y, z)
    assert np.abs(f0).max() < 1e-10 * np.abs(f).max()
    # translation invariance: sum over nodes vanishes
    assert np.abs(f.reshape(-1, 3).sum(0)).max() < 1e-9 * np.abs(f).max()
    # rotation invariance: moments of the internal forces vanish (sum x_a x f_a = 0)
    xa = np.stack([x, y, z], 1)
    mom = np.cross(xa, f.reshape(-1, 3)).sum(0)
    assert np.abs(mom).max() < 1e-8 * np.abs(f).max() * np.abs(xa).max()


def test_internal_force_is_energy_gradient():
    m = load_mesh("cube")
    g, dj = O.grad(m)
    lam, mu = lame(1e9, 0.3)
    x, y, z = deformed(m, seed=12, amp=1e-2)
    f = O.internal_force(m, g, dj, lam, mu, x, y, z)
    h = 1e-7
    for node in (0, 6, 13, 26):
        for d, arr in enumerate((x, y, z)):
            a = arr.copy()
            a[node] += h
            args = [x, y, z]
            args[d] = a
            ep = O.energy(m, g, dj, lam, mu, *args)
            a[node] -= 2 * h
            em = O.energy(m, g, dj, lam, mu, *args)
            fd = (ep - em) / (2 * h)
            assert abs(fd - f[3 * node + d]) < 1e-6 * np.abs(f).max()


@pytest.mark.parametrize("name", ["C1", "kuhn_t4"])
def test_patch_test(name):
    """Affine displacement -> constant P -> zero net force on every interior node."""
    m = load_mesh(name)
    g, dj = O.grad(m)
    lam, mu = lame(E_STEEL, NU_STEEL)
    A = np.array([[1e-3, 2e-4, 0], [0, -5e-4, 1e-4], [3e-4, 0, 2e-4]])
    X = m.xyz()
    x = X + X @ A.T
    f = O.internal_force(m, g, dj, lam, mu, x[:, 0].copy(), x[:, 1].copy(), x[:, 2].copy()).reshape(-1, 3)
    lo, hi = X.min(0), X.max(0)
    tol = 1e-9
    interior = np.all((X > lo + tol) & (X < hi - tol), axis=1)
    assert interior.sum() > 0
    assert np.abs(f[interior]).max() < 1e-9 * np.abs(f).max()
    assert np.abs(f[~interior]).max() > 1e-3 * np.abs(f).max()


# ---------------------------------------------------------------------------------------
@pytest.mark.parametrize("name", SMALL)
def test_patterns_against_numpy_unique(name):
    m = load_mesh(name)
    off, col = O.node_pattern(m)
    E = m.elems.astype(np.int64)
    n = m.n_nodes
    keys = np.unique((E[:, :, None] * n + E[:, None, :]).reshape(-1))
    rows, cols = keys // n, keys % n
    assert np.array_equal(col, cols) and np.array_equal(off, np.searchsorted(rows, np.arange(n + 1)))
    doff, dcol = O.dof_pattern(m)
    assert len(dcol) == 9 * len(col)
    assert doff[-1] == 9 * len(col)
    # DOF row 3i+d = node row i expanded by 3
    for i in (0, n // 2, n - 1):
        r = np.arange(off[i], off[i + 1])
        exp = (3 * col[r][:, None] + np.arange(3)[None, :]).reshape(-1)
        for d in range(3):
            assert np.array_equal(dcol[doff[3 * i + d]:doff[3 * i + d + 1]], exp)


def test_pattern_sizes_of_benchmark_beam():
    m = load_mesh("C1")
    off, col = O.node_pattern(m)
    assert (m.n_elem, m.n_nodes, len(col)) == (480, 1025, 22457)
    doff, dcol = O.dof_pattern(m)
    assert len(dcol) == 202113


@pytest.mark.parametrize("name", SMALL)
def test_mass_matrix(name):
    m = load_mesh(name)
    g, dj = O.grad(m)
    off, col = O.node_pattern(m)
    vals = O.mass_values(m, dj, RHO_STEEL, off, col)
    import scipy.sparse as sp
    Mm = sp.csr_matrix((vals, col, off)).toarray()
    assert np.abs(Mm - Mm.T).max() < 1e-12 * np.abs(Mm).max()
    X = m.xyz()[m.elems]
    V = np.array([_tet_volume(X[e, :4]) for e in range(m.n_elem)])
    assert abs(Mm.sum() - RHO_STEEL * V.sum()) < 1e-12 * RHO_STEEL * V.sum()
    # row sums = rho * integral of N_i: corner -V/20, midside V/5 (TET10); V/4 (TET4)
    rs = np.zeros(m.n_nodes)
    for e in range(m.n_elem):
        for a in range(m.nn):
            w = (-1 / 20 if a < 4 else 1 / 5) if m.nn == 10 else 1 / 4
            rs[m.elems[e, a]] += RHO_STEEL * V[e] * w
    assert np.abs(Mm.sum(1) - rs).max() < 1e-12 * np.abs(rs).max()
    if m.nn == 4:  # 1-point rule: every element entry is rho detJ / 96 = rho V / 16
        e0 = m.elems[0]
        assert abs(Mm[e0[0], e0[1]] - RHO_STEEL * V[0] / 16) < 1e-12 or Mm[e0[0], e0[1]] > RHO_STEEL * V[0] / 16


def _exact_shape_mass(nn):
    """int N_a N_b dV / V by exact rational arithmetic: N as polynomials in the barycentric
    coordinates, int L0^a L1^b L2^c L3^d dV = 3! a! b! c! d! V / (a+b+c+d+3)!."""
    from fractions import Fraction

    def mul(p, q):
        r = {}
        for ea, ca in p.items():
            for eb, cb in q.items():
                k = tuple(x + y for x, y in zip(ea, eb))
                r[k] = r.get(k, 0) + ca * cb
        return r

    def lin(i):
        e = [0, 0, 0, 0]
        e[i] = 1
        return {tuple(e): Fraction(1)}

    if nn == 4:
        N = [lin(i) for i in range(4)]
    else:
        N = []
        for i in range(4):  # L (2L - 1)
            sq = mul(lin(i), lin(i))
            N.append({k: 2 * v for k, v in sq.items()} | {tuple(lin(i))[0]: Fraction(-1)})
        for i, j in O.T10_EDGES:
            N.append({k: 4 * v for k, v in mul(lin(i), lin(j)).items()})

    def integ(p):
        return sum(c * Fraction(6 * math.prod(math.factorial(x) for x in e), math.factorial(sum(e) + 3))
                   for e, c in p.items())

    return np.array([[float(integ(mul(N[a], N[b]))) for b in range(nn)] for a in range(nn)])


@pytest.mark.parametrize("nn", [10, 4])
def test_exact_mass_coefficients(nn):
    Cm = O.exact_mass_coefficients(nn)
    assert np.abs(Cm - _exact_shape_mass(nn)).max() < 1e-15
    assert abs(Cm.sum() - 1.0) < 1e-14 and np.linalg.eigvalsh(Cm).min() > 0


def test_exact_shape_mass_route_matches_quadrature_on_degree_three():
    # the independent route itself: integrating N_a (degree 2) times a linear function is
    # degree 3, which the 5-point rule integrates exactly
    qx, qy, qz, qw = O.quad_t10()
    Cm = _exact_shape_mass(10)
    L = np.stack([1 - qx - qy - qz, qx, qy, qz])
    Nq = np.concatenate([L * (2 * L - 1), np.stack([4 * L[i] * L[j] for i, j in O.T10_EDGES])])
    # sum_b N_b = 1, so row sums of C are int N_a / V
    assert np.abs(Cm.sum(1) - 6 * Nq @ qw).max() < 1e-15


@pytest.mark.parametrize("name", ["cube", "beam_vtu", "kuhn_t4"])
def test_exact_mass_matrix(name):
    m = load_mesh(name)
    g, dj = O.grad(m)
    off, col = O.node_pattern(m)
    import scipy.sparse as sp
    Mx = sp.csr_matrix((O.mass_values_exact(m, dj, RHO_STEEL, off, col), col, off)).toarray()
    Mq = sp.csr_matrix((O.mass_values(m, dj, RHO_STEEL, off, col), col, off)).toarray()
    X = m.xyz()[m.elems]
    V = np.array([_tet_volume(X[e, :4]) for e in range(m.n_elem)])
    assert np.abs(Mx - Mx.T).max() < 1e-12 * np.abs(Mx).max()
    assert abs(Mx.sum() - RHO_STEEL * V.sum()) < 1e-12 * RHO_STEEL * V.sum()
    # the same row sums as the quadrature rule (the integral of N_i is degree 2)
    assert np.abs(Mx.sum(1) - Mq.sum(1)).max() < 1e-12 * np.abs(Mx.sum(1)).max()
    ev = np.linalg.eigvalsh(Mx)
    assert ev.min() > 0
    if m.nn == 10:  # the 5-point rule's mass is indefinite, the exact one is not
        assert np.linalg.eigvalsh(Mq).min() < 0


# ---------------------------------------------------------------------------------------
def _K(m, lam, mu, x=None, y=None, z=None, h=1.0, eta=0.0, ld=0.0):
    g, dj = O.grad(m)
    off, col = O.dof_pattern(m)
    vals = O.stiffness_values(m, g, dj, lam, mu, off, col, x, y, z, h=h, eta=eta, lambda_d=ld)
    import scipy.sparse as sp
    return sp.csr_matrix((vals, col, off)).toarray(), off, col, vals


@pytest.mark.parametrize("name", ["cube", "kuhn_t4"])
def test_stiffness_symmetry_rigid_modes_and_fd(name):
    m = load_mesh(name)
    lam, mu = lame(1e9, 0.3)
    K, *_ = _K(m, lam, mu)
    assert np.abs(K - K.T).max() < 1e-12 * np.abs(K).max()
    X = m.xyz()
    modes = [np.tile(np.eye(3)[d], m.n_nodes) for d in range(3)]
    for d in range(3):  # infinitesimal rotations
        w = np.zeros(3)
        w[d] = 1.0
        modes.append(np.cross(w, X).reshape(-1))
    for r in modes:
        assert np.abs(K @ r).max() < 1e-10 * np.abs(K).max() * np.abs(r).max()
    # tangent = derivative of f_int at a deformed state (central differences)
    g, dj = O.grad(m)
    x, y, z = deformed(m, seed=13, amp=5e-3)
    Kx, *_ = _K(m, lam, mu, x, y, z)
    assert np.abs(Kx - Kx.T).max() < 1e-10 * np.abs(Kx).max()
    u = np.random.default_rng(14).standard_normal(3 * m.n_nodes)
    eps = 1e-7 * np.abs(m.xyz()).max()
    U = u.reshape(-1, 3)
    fp = O.internal_force(m, g, dj, lam, mu, x + eps * U[:, 0], y + eps * U[:, 1], z + eps * U[:, 2])
    fm = O.internal_force(m, g, dj, lam, mu, x - eps * U[:, 0], y - eps * U[:, 1], z - eps * U[:, 2])
    assert rel_err(Kx @ u, (fp - fm) / (2 * eps)) < 1e-6


def test_stiffness_linear_closed_form():
    """At F = I: K_ab = sum_q (lambda g_a (x) g_b + mu (g_a . g_b) I + mu g_b (x) g_a) dV,
    assembled independently in numpy."""
    m = load_mesh("cube")
    lam, mu = lame(1e9, 0.3)
    K, *_ = _K(m, lam, mu, h=2.0)
    g, dj = O.grad(m)
    qw = O.quad_t10()[3]
    G = g.reshape(m.n_elem, 5, 3, 10)
    Kr = np.zeros_like(K)
    for e in range(m.n_elem):
        for q in range(5):
            dV = dj[e * 5 + q] * qw[q]
            for a in range(10):
                for b in range(10):
                    ga, gb = G[e, q, :, a], G[e, q, :, b]
                    blk = lam * np.outer(ga, gb) + mu * (ga @ gb) * np.eye(3) + mu * np.outer(gb, ga)
                    ia, ib = m.elems[e, a], m.elems[e, b]
                    Kr[3 * ia:3 * ia + 3, 3 * ib:3 * ib + 3] += 2.0 * blk * dV
    assert np.abs(K - Kr).max() < 1e-12 * np.abs(Kr).max()


def test_stiffness_damping_block():
    """The Kelvin-Voigt block equals the derivative of the viscous force w.r.t. v at
    F = I: C_ab = (eta (g_a.g_b) I + eta g_b (x) g_a + lambda_d g_a (x) g_b) dV."""
    m = load_mesh("cube")
    lam, mu = lame(1e9, 0.3)
    C, *_ = _K(m, 0.0, 0.0, eta=3.0, ld=5.0)  # elastic part off: the block alone
    g, dj = O.grad(m)
    v = np.random.default_rng(15).standard_normal(3 * m.n_nodes)
    f0 = O.internal_force(m, g, dj, 0.0, 0.0, v=np.zeros_like(v), eta=3.0, lambda_d=5.0)
    f1 = O.internal_force(m, g, dj, 0.0, 0.0, v=v, eta=3.0, lambda_d=5.0)
    assert rel_err(C @ v, f1 - f0) < 1e-10


@pytest.mark.parametrize("name", ["cube", "kuhn_t4"])
def test_damped_stiffness_at_rest_is_shifted_linear_elasticity(name):
    """At the rest positions K + C (the Kelvin-Voigt block added unscaled,
    FEAT10DataFunc.cuh:686-767) equals the undamped tangent with lambda + lambda_d and
    mu + eta: the identity the matrix-free operator relies on for damped bodies.  Away
    from rest it does not hold (the check that would catch a form that ignored x)."""
    m = load_mesh(name)
    lam, mu = lame(2e9, 0.3)
    eta, ld = 4e7, 9e6
    g, dj = O.grad(m)
    off, col = O.dof_pattern(m)
    kc = O.stiffness_values(m, g, dj, lam, mu, off, col, eta=eta, lambda_d=ld)
    ks = O.stiffness_values(m, g, dj, lam + ld, mu + eta, off, col)
    assert rel_err(kc, ks) < 1e-13
    assert rel_err(kc, O.stiffness_values(m, g, dj, lam, mu, off, col)) > 1e-2
    x, y, z = deformed(m, seed=3, amp=2e-2)
    kcx = O.stiffness_values(m, g, dj, lam, mu, off, col, x, y, z, eta=eta, lambda_d=ld)
    ksx = O.stiffness_values(m, g, dj, lam + ld, mu + eta, off, col, x, y, z)
    assert rel_err(kcx, ksx) > 1e-4


# ---------------------------------------------------------------------------------------
def test_dirichlet_against_dense_elimination():
    m = load_mesh("cube")
    lam, mu = lame(1e9, 0.3)
    K, off, col, vals = _K(m, lam, mu)
    v2, f2 = O.apply_dirichlet(off, col, vals, m.f_ext, m.fixed)
    import scipy.sparse as sp
    Kb = sp.csr_matrix((v2, col, off)).toarray()
    fixed = np.zeros(3 * m.n_nodes, bool)
    for n in m.fixed:
        fixed[3 * n:3 * n + 3] = True
    Kd = K.copy()
    Kd[fixed, :] = 0
    Kd[:, fixed] = 0
    Kd[fixed, fixed] = 1
    assert np.array_equal(Kb, Kd)
    fd = m.f_ext.copy()
    fd[fixed] = 0
    assert np.array_equal(f2, fd)


@pytest.mark.parametrize("name", ["cube", "beam_vtu", "C1"])
def test_cg_and_cholesky_against_dense(name):
    m = load_mesh(name)
    lam, mu = lame(E_STEEL, NU_STEEL)
    K, off, col, vals = _K(m, lam, mu)
    v2, f2 = O.apply_dirichlet(off, col, vals, m.f_ext, m.fixed)
    import scipy.sparse as sp
    Kb = sp.csr_matrix((v2, col, off)).toarray()
    ud = np.linalg.solve(Kb, f2)
    uc = O.band_cholesky_solve(off, col, v2, f2)
    assert rel_err(uc, ud) < 1e-10
    u, it, res = O.cg(off, col, v2, f2, tol=1e-12, max_iter=100000)
    assert it > 0 and res < 1e-12
    assert rel_err(u, ud) < 1e-9
    assert np.all(u[np.repeat(np.isin(np.arange(m.n_nodes), m.fixed), 3)] == 0)


def test_euler_bernoulli_band():
    """beam_linear_static.cc:223-234: the C1 tip deflection is within a few % of PL^3/(3EI)."""
    m = load_mesh("C1")
    lam, mu = lame(E_STEEL, NU_STEEL)
    _, off, col, vals = _K(m, lam, mu)
    v2, f2 = O.apply_dirichlet(off, col, vals, m.f_ext, m.fixed)
    u = O.band_cholesky_solve(off, col, v2, f2)
    dz = u[3 * m.meta["tip_nodes"] + 2].min()
    eb = 1000.0 / (3 * E_STEEL * 0.1 ** 4 / 12)
    assert 0.97 < abs(dz) / eb < 1.03


def test_cg_not_positive_definite():
    m = load_mesh("cube")
    lam, mu = lame(-1e9, 0.3)
    _, off, col, vals = _K(m, lam, mu)
    v2, f2 = O.apply_dirichlet(off, col, vals, m.f_ext, m.fixed)
    u, it, res = O.cg(off, col, v2, f2, tol=1e-10, max_iter=100)
    assert it == -1


# ---------------------------------------------------------------------------------------
def _newton(m, prob, dt, mode, rho=0.0, iters=40):
    """Direct (Newton, dense tangent) solutions the solvers must reach:
    mode "step":   M v/dt + f_int(X + dt v) - f_ext + rho dt^2 v|_fixed = 0 on all DOFs
                   (one implicit step with the augmented-Lagrangian penalty at lambda = 0);
    mode "static": f_int(X + dt v) = f_ext on the free DOFs with v = 0 at the fixed nodes
                   (the fixed point of several outer iterations, because the reference
                   sets v_prev = v_guess after every outer iteration, SyncedNesterov.cu:294-296).
    """
    import scipy.sparse as sp
    lam, mu = prob.lam, prob.mu
    M3 = np.kron(sp.csr_matrix((prob.m_val, prob.m_col, prob.m_off)).toarray(), np.eye(3))
    fixed = np.zeros(3 * m.n_nodes, bool)
    for n in m.fixed:
        fixed[3 * n:3 * n + 3] = True
    free = ~fixed
    X = m.xyz()
    v = np.zeros(3 * m.n_nodes)
    for _ in range(iters):
        x = X + dt * v.reshape(-1, 3)
        f = O.internal_force(m, prob.g, prob.detJ, lam, mu, x[:, 0].copy(), x[:, 1].copy(), x[:, 2].copy())
        K, *_ = _K(m, lam, mu, x[:, 0].copy(), x[:, 1].copy(), x[:, 2].copy())
        if mode == "step":
            g = M3 @ v / dt + f - m.f_ext + rho * dt * dt * v * fixed
            J = M3 / dt + dt * K + np.diag(rho * dt * dt * fixed)
            dv = np.linalg.solve(J, -g)
            v += dv
        else:
            g = f - m.f_ext
            dv = np.linalg.solve(dt * K[np.ix_(free, free)], -g[free])
            v[free] += dv
        if np.linalg.norm(dv) < 1e-15 * max(1e-300, np.linalg.norm(v)):
            break
    return v, free


def test_nesterov_single_outer_reaches_implicit_step():
    """One outer iteration of the restated Nesterov loop, run long, lands on the direct
    solution of its implicit step (penalty rho at lambda = 0)."""
    m = load_mesh("cube")
    prob = O.Problem(m, 7e8, 0.33, 2700.0)
    dt, rho = 1e-3, 1e10
    st = prob.new_state()
    prob.nesterov_step(st, 1e-8, rho, 0.0, 0.0, 1, 100000, dt)
    v_ref, free = _newton(m, prob, dt, "step", rho=rho)
    assert rel_err(st["v_guess"], v_ref) < 1e-9


def test_nesterov_outer_loop_approaches_static_equilibrium():
    """With the reference's rho = 1e14 and many outer iterations the step tends to static
    equilibrium (the augmented-Lagrangian fixed point converges slowly: a 5e-3 band)."""
    m = load_mesh("cube")
    prob = O.Problem(m, 7e8, 0.33, 2700.0)
    dt = 1e-3
    st = prob.new_state()
    prob.nesterov_step(st, 1e-8, 1e14, 0.0, 0.0, 60, 5000, dt)
    v_ref, free = _newton(m, prob, dt, "static")
    assert rel_err(st["v_guess"][free], v_ref[free]) < 5e-3
    assert np.abs(st["v_guess"][~free]).max() < 1e-2 * np.abs(v_ref).max()


@pytest.mark.parametrize("nocoop", [False, True])
def test_adamw_single_outer_approaches_implicit_step(nocoop):
    """AdamW's one-outer result approaches the direct implicit-step solution (fixed
    learning rate -> a loose band); its iterate sequence is pinned by the torch.optim.AdamW
    test below."""
    m = load_mesh("cube")
    prob = O.Problem(m, 7e8, 0.33, 2700.0)
    dt, rho = 1e-3, 1e10
    st = prob.new_state()
    prob.adamw_step(st, 1e-4, 0.9, 0.999, 1e-8, 0.0, 0.999, 0.0, 0.0, rho, 1, 4000, dt, 10, 0.0, nocoop=nocoop)
    v_ref, free = _newton(m, prob, dt, "step", rho=rho)
    assert rel_err(st["v_guess"], v_ref) < 5e-2


@pytest.mark.parametrize("nocoop", [False, True])
def test_adamw_trajectory_matches_torch_adamw(nocoop):
    """The oracle's AdamW inner loop against torch.optim.AdamW, an independent
    implementation of the same decoupled-weight-decay Adam (SyncedAdamW.cu:146-167,
    SyncedAdamWNocoop.cu:98-253), driven by the oracle's own step gradient.  The reference's
    conventions map onto it as: the first update uses g = 0 and its bias correction uses
    t = 2 (torch's step counter preset to 1), and the learning rate is decayed before each
    use (lr_k = lr0 * decay^(k+1), set on the parameter group).  Pins the moment updates,
    the bias corrections, eps outside the square root and the weight decay on v_guess."""
    import torch
    m = load_mesh("cube")
    prob = O.Problem(m, 7e8, 0.33, 2700.0)
    dt, rho = 1e-3, 1e10
    lr0, b1, b2, eps, wd, decay = 2e-4, 0.9, 0.999, 1e-8, 1e-4, 0.995
    n_inner = 12
    st = prob.new_state()
    xp, yp, zp = st["x"].copy(), st["y"].copy(), st["z"].copy()
    v_prev, lam = st["v_prev"].copy(), st["lam"].copy()
    prob.adamw_step(st, lr0, b1, b2, eps, wd, decay, 0.0, 0.0, rho, 1, n_inner, dt, 10, 0.0, nocoop=nocoop)

    v = torch.zeros(3 * m.n_nodes, dtype=torch.float64, requires_grad=True)
    opt = torch.optim.AdamW([v], lr=lr0, betas=(b1, b2), eps=eps, weight_decay=wd, foreach=False)
    opt.state[v] = {"step": torch.tensor(1.0), "exp_avg": torch.zeros_like(v),
                    "exp_avg_sq": torch.zeros_like(v)}
    g = np.zeros(3 * m.n_nodes)
    for k in range(n_inner):
        opt.param_groups[0]["lr"] = lr0 * decay ** (k + 1)
        v.grad = torch.from_numpy(g.copy())
        opt.step()
        g = prob.step_gradient(xp, yp, zp, v.detach().numpy(), v_prev, lam, dt, rho)
    ref = v.detach().numpy()
    assert np.abs(ref).max() > 0.0
    assert rel_err(st["v_guess"], ref) < 1e-10


def test_nesterov_trajectory_matches_python_restatement():
    """Independent pure-Python restatement of one Nesterov inner loop (alpha, beta sequence,
    look-ahead, gradient, update) on the cube for a handful of iterations."""
    m = load_mesh("cube")
    prob = O.Problem(m, 7e8, 0.33, 2700.0)
    dt, alpha, rho = 1e-3, 1e-8, 1e10
    st = prob.new_state()
    prob.nesterov_step(st, alpha, rho, 0.0, 0.0, 1, 6, dt)
    import scipy.sparse as sp
    M3 = np.kron(sp.csr_matrix((prob.m_val, prob.m_col, prob.m_off)).toarray(), np.eye(3))
    X = m.xyz()
    vk = np.zeros(3 * m.n_nodes)
    vkm1 = vk.copy()
    t = 1.0
    for _ in range(6):
        tn = 0.5 * (1 + np.sqrt(1 + 4 * t * t))
        y = vk + (t - 1) / tn * (vk - vkm1)
        x = X + dt * y.reshape(-1, 3)
        f = O.internal_force(m, prob.g, prob.detJ, prob.lam, prob.mu, x[:, 0].copy(), x[:, 1].copy(), x[:, 2].copy())
        g = M3 @ y / dt + f - m.f_ext
        for n in m.fixed:
            g[3 * n:3 * n + 3] += dt * rho * (x[n] - X[n])
        vkm1, vk, t = vk, y - alpha * g, tn
    assert rel_err(st["v_guess"], vk) < 1e-12


@pytest.mark.parametrize("name", ["cube", "kuhn_t4"])
def test_extended_precision_force_restatement(name):
    # the longdouble restatement used to assess small-strain accuracy agrees with the oracle
    # where both are accurate (large deformation), and the oracle's own F^T F - I form loses
    # digits at strains near 1e-6 (what the longdouble reference is for)
    from helpers import fint_extended
    m = load_mesh(name)
    lam, mu = lame(E_STEEL, NU_STEEL)
    g, dj = O.grad(m)
    xd = deformed(m, seed=2)
    assert rel_err(fint_extended(m, *xd, lam, mu).astype(np.float64), O.internal_force(m, g, dj, lam, mu, *xd)) < 1e-12
    u = 1e-7 * np.ptp(m.xyz(), axis=0).max() * np.random.default_rng(5).standard_normal((m.n_nodes, 3))
    xs = (m.x + u[:, 0], m.y + u[:, 1], m.z + u[:, 2])
    e = rel_err(O.internal_force(m, g, dj, lam, mu, *xs), fint_extended(m, *xs, lam, mu).astype(np.float64))
    assert 1e-13 < e < 1e-8


# ---------------------------------------------------------------------------------------
# Mooney-Rivlin (no test or example of the reference uses it: pinned by the mathematics)
MR = (0.3e6, 0.1e6, 5e6)  # mu10, mu01, kappa [Pa]


def _rand_F(seed, amp=0.2):
    F = np.eye(3) + amp * np.random.default_rng(seed).standard_normal((3, 3))
    assert np.linalg.det(F) > 0.2
    return F


def test_mr_point_derivatives_and_invariants():
    mu10, mu01, kappa = MR
    W0, P0, A0 = O.mr_point(np.eye(3), *MR)
    assert abs(W0) < 1e-9 * mu10 and np.abs(P0).max() < 1e-9 * mu10
    # small-strain limit: isotropic linear elasticity with shear modulus 2 (mu10 + mu01)
    # and bulk modulus kappa
    mu = 2 * (mu10 + mu01)
    eye = np.eye(3)
    C = (mu * (np.einsum("ik,jl->ijkl", eye, eye) + np.einsum("il,jk->ijkl", eye, eye)) +
         (kappa - 2 * mu / 3) * np.einsum("ij,kl->ijkl", eye, eye))
    assert np.abs(A0 - C).max() < 1e-9 * np.abs(C).max()
    for seed in (1, 2, 3):
        F = _rand_F(seed)
        W, P, A = O.mr_point(F, *MR)
        h = 1e-6
        for k in range(3):
            for l in range(3):
                Fp, Fm = F.copy(), F.copy()
                Fp[k, l] += h
                Fm[k, l] -= h
                Wp, Pp, _ = O.mr_point(Fp, *MR)
                Wm, Pm, _ = O.mr_point(Fm, *MR)
                assert abs((Wp - Wm) / (2 * h) - P[k, l]) < 1e-7 * np.abs(P).max()
                assert np.abs((Pp - Pm) / (2 * h) - A[:, :, k, l]).max() < 1e-7 * np.abs(A).max()
        # hyperelastic: A is a Hessian (major symmetry)
        assert np.abs(A - A.transpose(2, 3, 0, 1)).max() < 1e-12 * np.abs(A).max()
        # frame indifference: W(QF) = W(F), P(QF) = Q P(F)
        Q, _ = np.linalg.qr(np.random.default_rng(seed + 10).standard_normal((3, 3)))
        Q *= np.sign(np.linalg.det(Q))
        Wq, Pq, _ = O.mr_point(Q @ F, *MR)
        assert abs(Wq - W) < 1e-9 * abs(W)
        assert np.abs(Pq - Q @ P).max() < 1e-9 * np.abs(P).max()


def test_mr_uniaxial_closed_form():
    """F = diag(s, 1, 1): J = s, I1 = s^2 + 2, I2 = 2 s^2 + 1 and
    P_11 = t1 (s - I1/(3 s)) + t2 (I1 s - s^3 - 2 I2/(3 s)) + kappa (s - 1)."""
    mu10, mu01, kappa = MR
    for s in (0.9, 1.05, 1.3):
        _, P, _ = O.mr_point(np.diag([s, 1.0, 1.0]), *MR)
        I1, I2 = s * s + 2, 2 * s * s + 1
        t1, t2 = 2 * mu10 * s ** (-2 / 3), 2 * mu01 * s ** (-4 / 3)
        P11 = t1 * (s - I1 / (3 * s)) + t2 * (I1 * s - s ** 3 - 2 * I2 / (3 * s)) + kappa * (s - 1)
        P22 = t1 * (1 - I1 / 3) + t2 * (I1 - 1 - 2 * I2 / 3) + kappa * (s - 1) * s
        assert abs(P[0, 0] - P11) < 1e-12 * abs(P11)
        assert abs(P[1, 1] - P22) < 1e-10 * max(abs(P22), mu10)
        assert np.abs(P - np.diag(np.diag(P))).max() < 1e-9 * mu10


@pytest.mark.parametrize("name", ["cube", "kuhn_t4"])
def test_mr_element_force_and_stiffness(name):
    """With the material switched to Mooney-Rivlin: f_int(rest) = 0, f_int = gradient of the
    total energy, K symmetric with the rigid modes in its kernel at rest and equal to the
    derivative of f_int at a deformed state."""
    m = load_mesh(name)
    g, dj = O.grad(m)
    X = m.xyz()
    scale = np.abs(X).max()
    with O.material("mr", *MR):
        f0 = O.internal_force(m, g, dj, 0.0, 0.0)
        # det F within [0.86, 1.14] on both meshes (the TET4 mesh's elements are small)
        x, y, z = deformed(m, seed=21, amp=1e-2 if m.nn == 10 else 1e-3)
        f = O.internal_force(m, g, dj, 0.0, 0.0, x, y, z)
        assert np.abs(f0).max() < 1e-10 * np.abs(f).max()
        h = 1e-7 * scale
        for node in (0, 3, m.n_nodes // 2):
            for d, arr in enumerate((x, y, z)):
                args = [x, y, z]
                a = arr.copy()
                a[node] += h
                args[d] = a
                ep = O.energy(m, g, dj, 0.0, 0.0, *args)
                a = arr.copy()
                a[node] -= h
                args[d] = a
                em = O.energy(m, g, dj, 0.0, 0.0, *args)
                assert abs((ep - em) / (2 * h) - f[3 * node + d]) < 1e-6 * np.abs(f).max()
        K0, *_ = _K(m, 0.0, 0.0)
        assert np.abs(K0 - K0.T).max() < 1e-12 * np.abs(K0).max()
        modes = [np.tile(np.eye(3)[d], m.n_nodes) for d in range(3)]
        for d in range(3):
            w = np.zeros(3)
            w[d] = 1.0
            modes.append(np.cross(w, X).reshape(-1))
        for r in modes:
            assert np.abs(K0 @ r).max() < 1e-10 * np.abs(K0).max() * np.abs(r).max()
        Kx, *_ = _K(m, 0.0, 0.0, x, y, z)
        assert np.abs(Kx - Kx.T).max() < 1e-10 * np.abs(Kx).max()
        u = np.random.default_rng(22).standard_normal(3 * m.n_nodes)
        eps = 1e-7 * scale
        U = u.reshape(-1, 3)
        fp = O.internal_force(m, g, dj, 0.0, 0.0, x + eps * U[:, 0], y + eps * U[:, 1], z + eps * U[:, 2])
        fm = O.internal_force(m, g, dj, 0.0, 0.0, x - eps * U[:, 0], y - eps * U[:, 1], z - eps * U[:, 2])
        assert rel_err(Kx @ u, (fp - fm) / (2 * eps)) < 1e-6
    # the switch is scoped: SVK again outside
    lam, mu = lame(1e9, 0.3)
    assert np.abs(O.internal_force(m, g, dj, lam, mu, x, y, z) - f).max() > 1e-3 * np.abs(f).max()


# ---------------------------------------------------------------------------------------
def test_constraint_jacobian_csr_is_the_fixed_dof_selection():
    """J selects the fixed DOFs and J^T is its transpose (FEAT10Data.cu:420-562): J J^T = I,
    J^T equals J transposed, J x picks x at 3 fixed[k] + d in constraint order, and J^T's
    row sums mark exactly the fixed DOFs.  An off-by-one in the DOF or a swapped d/k order
    fails these."""
    import scipy.sparse as sp
    rng = np.random.default_rng(3)
    n_nodes = 40
    fixed = rng.choice(n_nodes, 11, replace=False)  # unsorted, as UpdateNodalFixed may pass
    nc = 3 * len(fixed)
    o, c, v = O.constraint_jacobian_csr(fixed, n_nodes)
    J = sp.csr_matrix((v, c, o), shape=(nc, 3 * n_nodes))
    ot, ct, vt = O.constraint_jacobian_csr(fixed, n_nodes, transpose=True)
    JT = sp.csr_matrix((vt, ct, ot), shape=(3 * n_nodes, nc))
    assert abs(J @ JT - sp.eye(nc)).max() == 0.0
    assert abs(J.T - JT).max() == 0.0
    x = rng.standard_normal(3 * n_nodes)
    assert np.array_equal(J @ x, x.reshape(-1, 3)[fixed].reshape(-1))
    mark = np.zeros(3 * n_nodes)
    mark.reshape(-1, 3)[fixed] = 1.0
    assert np.array_equal(np.asarray(JT.sum(axis=1)).ravel(), mark)
    for r in range(3 * n_nodes):  # columns ascending within each J^T row
        assert np.all(np.diff(ct[ot[r]:ot[r + 1]]) > 0)


@pytest.mark.parametrize("name", ["C1", "kuhn_t4", "cube"])
@pytest.mark.parametrize("model", ["svk", "svk_damped", "mr"])
def test_tangent_apply_equals_assembled_spmv(name, model):
    """The element-by-element (h K + C) p equals the assembled values times p at a deformed
    state (a wrong block index or a transposed block fails it), and p.Kq = q.Kp."""
    m = load_mesh(name)
    lam, mu = lame(E_STEEL, NU_STEEL)
    g, dj = O.grad(m)
    x, y, z = deformed(m, seed=4, amp=2e-3)
    eta, ld = (0.3 * mu, 0.2 * lam) if model == "svk_damped" else (0.0, 0.0)
    off, col = O.dof_pattern(m)
    rng = np.random.default_rng(8)
    p, q = rng.standard_normal((2, 3 * m.n_nodes))
    ctx = O.material("mr", 0.3 * mu, 0.1 * mu, lam + 2 * mu / 3) if model == "mr" else O.material()
    with ctx:
        vals = O.stiffness_values(m, g, dj, lam, mu, off, col, x=x, y=y, z=z, h=0.7, eta=eta, lambda_d=ld)
        kp = O.tangent_apply(m, g, dj, lam, mu, p, x=x, y=y, z=z, h=0.7, eta=eta, lambda_d=ld)
        kq = O.tangent_apply(m, g, dj, lam, mu, q, x=x, y=y, z=z, h=0.7, eta=eta, lambda_d=ld)
    assert rel_err(kp, O.spmv(off, col, vals, p)) < 1e-13
    assert abs(q @ kp - p @ kq) < 1e-12 * np.linalg.norm(kp) * np.linalg.norm(q)


def test_true_rel_residual_of_direct_solve_and_of_a_perturbation():
    """The true residual of the banded Cholesky solution on C1 is at the FP64 floor
    eps ||K||_1 ||u|| / ||f|| (1.9e-10 here: the beam is stiff), and a perturbation e of it
    gives exactly ||K e|| / ||f|| on the free DOFs."""
    m = load_mesh("C1")
    lam, mu = lame(E_STEEL, NU_STEEL)
    g, dj = O.grad(m)
    off, col = O.dof_pattern(m)
    vals = O.stiffness_values(m, g, dj, lam, mu, off, col)
    ve, fb = O.apply_dirichlet(off, col, vals, m.f_ext, m.fixed)
    u = O.band_cholesky_solve(off, col, ve, fb)
    k1 = np.bincount(col, weights=np.abs(ve), minlength=u.size).max()  # ||K||_1 (symmetric)
    floor = np.finfo(float).eps * k1 * np.linalg.norm(u) / np.linalg.norm(fb)
    assert 1e-12 < floor < 1e-9
    assert O.true_rel_residual(m, g, dj, lam, mu, u, m.f_ext, m.fixed) < 20 * floor
    e = np.random.default_rng(2).standard_normal(u.size) * 1e-3 * np.abs(u).max()
    e.reshape(-1, 3)[m.fixed] = 0.0
    want = np.linalg.norm(O.spmv(off, col, ve, e)) / np.linalg.norm(fb)
    got = O.true_rel_residual(m, g, dj, lam, mu, u + e, m.f_ext, m.fixed)
    assert abs(got - want) < 1e-6 * want
