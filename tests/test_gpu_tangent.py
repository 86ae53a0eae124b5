"""GPU parity of the matrix-free tangent action y = K(x) p and of the CG built on it.

Bars: K(x) p 1e-12 relative against the oracle's assembled tangent times p (a different
summation order only); symmetry p.Kq = q.Kp to 1e-12; the central difference of the
internal force (an independent route through the force kernel) to 1e-7; matrix-free CG
displacements against the banded Cholesky solve with the same bars as the assembled CG.
"""
import numpy as np
import pytest

import oracle as O
from helpers import E_STEEL, NU_STEEL, RHO_STEEL, deformed, lame, load_mesh, rel_err

pytestmark = pytest.mark.gpu

MESHES = ["cube", "beam_vtu", "C1", "kuhn_t4"]


def _body(m, E=E_STEEL, nu=NU_STEEL, eta=0.0, lambda_d=0.0):
    from tlfea_amd.fea import Body
    return Body(m, E, nu, RHO_STEEL, eta, lambda_d)


def _np(t):
    return t.detach().cpu().numpy()


def _set_positions(b, xyz):
    import torch
    for d in range(3):
        b.x[d].copy_(torch.as_tensor(xyz[d]))


def _dn_table(nn, quad):
    """dN/dxi per (q, node, k) from the shape functions (FEAT10Data.cu:92-198)."""
    dL = np.array([[-1.0, -1, -1], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    out = []
    for q in range(quad.n):
        if nn == 4:
            out.append(dL.copy())
            continue
        L = np.array([1 - quad.x[q] - quad.y[q] - quad.z[q], quad.x[q], quad.y[q], quad.z[q]])
        rows = [(4 * L[a] - 1) * dL[a] for a in range(4)]
        for i, j in ((0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3)):
            rows.append(4 * (L[i] * dL[j] + L[j] * dL[i]))
        out.append(np.array(rows))
    return np.array(out)


@pytest.mark.parametrize("name", ["cube", "kuhn_t4", "beam_highres_t10"])
def test_jinv_rebuilds_gradients(gpu, name):
    """g_a = J^-T dN_a/dxi from the stored inverse Jacobians equals the stored gradients."""
    m = load_mesh(name)
    b = _body(m)
    nq = b.nq
    pe = b.grad_t.numel() // (nq * b.nn * 3)
    M = _np(b.jinv_t).reshape(pe // 64, nq, 9, 64).transpose(0, 3, 1, 2).reshape(pe, nq, 3, 3)[: m.n_elem]
    dn = _dn_table(b.nn, b.quad)                                  # [q, a, k]
    g = np.einsum("eqik,qak->eqai", M, dn)                        # [e, q, a, i]
    gr, _ = b.grad_ref()
    gr = _np(gr).reshape(m.n_elem, nq, 3, b.nn).transpose(0, 1, 3, 2)  # ref layout a + nn*d
    assert rel_err(g, gr) < 1e-14


@pytest.mark.parametrize("name", MESHES)
@pytest.mark.parametrize("masked", [False, True])
def test_tangent_matches_oracle_matrix(gpu, name, masked):
    import torch
    print(f"[tangent] {name} masked={masked}", flush=True)
    m = load_mesh(name)
    b = _body(m)
    xd = deformed(m, seed=3)
    _set_positions(b, xd)
    p = np.random.default_rng(11).standard_normal(3 * m.n_nodes)
    y = _np(b.tangent_apply(torch.as_tensor(p).cuda(), masked=masked))
    go, djo = O.grad(m)
    lam, mu = lame(E_STEEL, NU_STEEL)
    doff, dcol = O.dof_pattern(m)
    vo = O.stiffness_values(m, go, djo, lam, mu, doff, dcol, *xd)
    if masked:
        vo, _ = O.apply_dirichlet(doff, dcol, vo, m.f_ext, m.fixed)
    ref = O.spmv(doff, dcol, vo, p)
    assert rel_err(y, ref) < 1e-12
    if masked and len(m.fixed):
        fd = (3 * np.asarray(m.fixed)[:, None] + np.arange(3)).reshape(-1)
        assert np.array_equal(y[fd], p[fd])


@pytest.mark.parametrize("name", ["C1", "beam_highres_t10", "kuhn_t4"])
def test_tangent_rest_state_form(gpu, name):
    """At the rest positions the kernel's linear-elasticity form (selected when the
    positions are the rest positions' own arrays) equals the general form at u = 0 and the
    oracle's tangent assembled at rest."""
    import torch
    m = load_mesh(name)
    b = _body(m)
    pt = torch.as_tensor(np.random.default_rng(12).standard_normal(3 * m.n_nodes)).cuda()
    y_gen = _np(b.tangent_apply(pt))            # separate x arrays equal to X0: general form
    y_lin = _np(b.tangent_apply(pt, x=b.X0))    # the rest arrays themselves: rest-state form
    assert rel_err(y_lin, y_gen) < 1e-14
    # on an affine TET10 mesh the rest-state form integrates from the corner values instead
    # of the five points (exact for both) with point 0's M: the same operator to rounding
    # (the bar of the affine form, test_affine_mesh)
    y_pts = _np(b.tangent_apply(pt, x=b.X0, affine=False))
    assert rel_err(y_lin, y_pts) < 1e-13
    go, djo = O.grad(m)
    lam, mu = lame(E_STEEL, NU_STEEL)
    doff, dcol = O.dof_pattern(m)
    vo, _ = O.apply_dirichlet(doff, dcol, O.stiffness_values(m, go, djo, lam, mu, doff, dcol), m.f_ext, m.fixed)
    assert rel_err(y_lin, O.spmv(doff, dcol, vo, _np(pt))) < 1e-12


def test_shuffled_mesh_large_blocks(gpu):
    """Elements in random order: each 256-element block touches ~1800 nodes, so every thread
    stages several groups of nodes.  Rest-state, affine and general forms against the
    oracle, and the matrix-free CG (with its fused update) against the oracle's CG."""
    import torch
    from tlfea_amd.fea import LinearStatic
    m = load_mesh("beam_highres_t10")
    m.elems = m.elems[np.random.default_rng(0).permutation(m.n_elem)]
    b = _body(m)
    assert b.plan_sizes()["max_block_nodes"] > 1024
    go, djo = O.grad(m)
    lam, mu = lame(E_STEEL, NU_STEEL)
    doff, dcol = O.dof_pattern(m)
    pt = torch.as_tensor(np.random.default_rng(15).standard_normal(3 * m.n_nodes)).cuda()
    vr, fo = O.apply_dirichlet(doff, dcol, O.stiffness_values(m, go, djo, lam, mu, doff, dcol), m.f_ext, m.fixed)
    assert rel_err(_np(b.tangent_apply(pt, x=b.X0)), O.spmv(doff, dcol, vr, _np(pt))) < 1e-12
    assert rel_err(_np(b.tangent_apply(pt, x=b.X0, affine=False)), O.spmv(doff, dcol, vr, _np(pt))) < 1e-12
    xd = deformed(m, seed=16, amp=1e-4)
    _set_positions(b, xd)
    vd, _ = O.apply_dirichlet(doff, dcol, O.stiffness_values(m, go, djo, lam, mu, doff, dcol, *xd), m.f_ext, m.fixed)
    assert rel_err(_np(b.tangent_apply(pt)), O.spmv(doff, dcol, vd, _np(pt))) < 1e-12
    _set_positions(b, (m.x, m.y, m.z))
    ls = LinearStatic(b, tol=0.0, max_iter=10, matfree=True)
    ls.assemble()
    ls.cg_solve()
    uk, _, _ = O.cg(doff, dcol, vr, fo, tol=0.0, max_iter=10)
    assert ls.iters == 10 and rel_err(_np(ls.u), uk) < 1e-11


def test_affine_mesh(gpu):
    """A mesh of straight-sided elements is reported affine (M and det J equal at every
    quadrature point to 1e-12); the operator read through point 0's M and det J equals the
    per-point operator to rounding and the oracle to the usual bar.  Curving one edge (a
    midside node moved off its edge by 1% of the edge length) makes the mesh non-affine."""
    import torch
    from tlfea_amd import mesh as M
    m = M.clamp_and_tip_load(M.kuhn_beam(20, 2, 2), -1000.0)   # the C1 beam, 480 TET10
    b = _body(m)
    assert b.affine
    xd = deformed(m, seed=6)
    _set_positions(b, xd)
    pt = torch.as_tensor(np.random.default_rng(14).standard_normal(3 * m.n_nodes)).cuda()
    y_aff = _np(b.tangent_apply(pt))
    y_pts = _np(b.tangent_apply(pt, affine=False))
    assert rel_err(y_aff, y_pts) < 1e-13
    go, djo = O.grad(m)
    lam, mu = lame(E_STEEL, NU_STEEL)
    doff, dcol = O.dof_pattern(m)
    vo, _ = O.apply_dirichlet(doff, dcol, O.stiffness_values(m, go, djo, lam, mu, doff, dcol, *xd), m.f_ext, m.fixed)
    assert rel_err(y_aff, O.spmv(doff, dcol, vo, _np(pt))) < 1e-12
    # at rest: the rest-state and affine forms together
    y_rest = _np(b.tangent_apply(pt, x=b.X0))
    vr, _ = O.apply_dirichlet(doff, dcol, O.stiffness_values(m, go, djo, lam, mu, doff, dcol), m.f_ext, m.fixed)
    assert rel_err(y_rest, O.spmv(doff, dcol, vr, _np(pt))) < 1e-12
    mc = M.clamp_and_tip_load(M.kuhn_beam(20, 2, 2), -1000.0)
    nid = int(mc.elems[200, 4])
    a0, a1 = mc.xyz()[mc.elems[200, 0]], mc.xyz()[mc.elems[200, 1]]
    mc.z[nid] += 0.01 * np.linalg.norm(a1 - a0)
    assert not _body(mc).affine
    assert _body(load_mesh("kuhn_t4")).affine  # one quadrature point


@pytest.mark.parametrize("name", ["beam_vtu", "kuhn_t4"])
def test_tangent_symmetric(gpu, name):
    import torch
    m = load_mesh(name)
    b = _body(m)
    _set_positions(b, deformed(m, seed=4))
    rng = np.random.default_rng(12)
    p = torch.as_tensor(rng.standard_normal(3 * m.n_nodes)).cuda()
    q = torch.as_tensor(rng.standard_normal(3 * m.n_nodes)).cuda()
    kp = b.tangent_apply(p, masked=False)
    kq = b.tangent_apply(q, masked=False)
    a, c = torch.dot(q, kp).item(), torch.dot(p, kq).item()
    scale = (torch.linalg.norm(q) * torch.linalg.norm(kp)).item()
    assert abs(a - c) < 1e-12 * scale


@pytest.mark.parametrize("name", ["cube", "kuhn_t4"])
def test_tangent_is_force_derivative(gpu, name):
    """K(x) p against (f_int(x + h p) - f_int(x - h p)) / 2h from the force kernel."""
    import torch
    m = load_mesh(name)
    b = _body(m)
    xd = deformed(m, seed=5, amp=2e-2)
    _set_positions(b, xd)
    p = np.random.default_rng(13).standard_normal(3 * m.n_nodes)
    kp = _np(b.tangent_apply(torch.as_tensor(p).cuda(), masked=False))
    L = np.ptp(m.xyz(), axis=0).max()
    h = 1e-6 * L
    fs = []
    for sgn in (1.0, -1.0):
        xs = [torch.as_tensor(xd[d] + sgn * h * p[d::3]).cuda() for d in range(3)]
        fs.append(_np(b.internal_force(x=xs)))
    fd = (fs[0] - fs[1]) / (2 * h)
    assert rel_err(kp, fd) < 1e-7


# Kelvin-Voigt coefficients of the damped tests: a quarter of mu and half of lambda, so
# that K + C differs from K by tens of percent
ETA_D, LD_D = 0.25 * lame(E_STEEL, NU_STEEL)[1], 0.5 * lame(E_STEEL, NU_STEEL)[0]


def _damped_oracle(m, xyz=(None, None, None)):
    go, djo = O.grad(m)
    lam, mu = lame(E_STEEL, NU_STEEL)
    doff, dcol = O.dof_pattern(m)
    return doff, dcol, O.stiffness_values(m, go, djo, lam, mu, doff, dcol, *xyz, eta=ETA_D, lambda_d=LD_D)


@pytest.mark.parametrize("name", ["cube", "C1", "kuhn_t4"])
def test_damped_tangent(gpu, name):
    """K + C (FEAT10DataFunc.cuh:512-767) applied without assembly against the oracle's
    assembled damped tangent times p, 1e-12: at the rest positions (the rest-state operator
    with lambda + lambda_d, mu + eta) and at deformed positions (the general form, C p =
    F (lambda_d tr(dE) I + 2 eta dE) added to its material term)."""
    import torch
    print(f"[damped tangent] {name}", flush=True)
    m = load_mesh(name)
    b = _body(m, eta=ETA_D, lambda_d=LD_D)
    p = np.random.default_rng(21).standard_normal(3 * m.n_nodes)
    pt = torch.as_tensor(p).cuda()
    kp = _np(b.tangent_apply(pt, x=b.X0, masked=False, damping=True))
    doff, dcol, vo = _damped_oracle(m)
    ref = O.spmv(doff, dcol, vo, p)
    assert rel_err(kp, ref) < 1e-12
    k_only = _np(b.tangent_apply(pt, x=b.X0, masked=False))
    assert rel_err(k_only, ref) > 1e-2  # the damping block is really in it
    xyz = deformed(m, seed=5, amp=2e-2)
    xt = [torch.as_tensor(a).cuda() for a in xyz]
    kpx = _np(b.tangent_apply(pt, x=xt, masked=False, damping=True))
    _, _, vx = _damped_oracle(m, xyz)
    assert rel_err(kpx, O.spmv(doff, dcol, vx, p)) < 1e-12


@pytest.mark.parametrize("name", ["cube", "C1"])
def test_damped_matfree_cg_against_direct(gpu, name):
    """A damped body's linear-static system is (K + C) u = f in the reference (the block is
    added unscaled, h = 1): the matrix-free CG and the assembled CG both against the
    oracle's banded Cholesky solve of it, at the rest positions and at moved ones."""
    import torch
    from tlfea_amd.fea import LinearStatic
    print(f"[damped matfree cg] {name}", flush=True)
    m = load_mesh(name)
    b = _body(m, eta=ETA_D, lambda_d=LD_D)
    for moved in (False, True):
        if moved:
            # strains of ~0.2% (a 1% perturbation of the node positions makes K + C of the
            # slender C1 beam indefinite)
            xyz = deformed(m, seed=9, amp=1e-4)
            for d in range(3):
                b.x[d].copy_(torch.as_tensor(xyz[d]))
        doff, dcol, vo = _damped_oracle(m, xyz if moved else (None, None, None))
        vo, fo = O.apply_dirichlet(doff, dcol, vo, m.f_ext, m.fixed)
        ud = O.band_cholesky_solve(doff, dcol, vo, fo)
        for matfree in (True, False):
            ls = LinearStatic(b, tol=1e-13, max_iter=200000, matfree=matfree)
            ls.assemble()
            assert ls.cg_solve() == 0, (moved, matfree, ls.iters, ls.res)
            assert rel_err(_np(ls.u), ud) < CG_DISP_BAR[name], (moved, matfree)


CG_DISP_BAR = {"cube": 1e-10, "beam_vtu": 1e-10, "C1": 1e-10, "beam_highres_t4": 20 * 3.2e7 * 2.2e-16}
# relative iteration-count bar at tol 1e-8: the assembly and the matrix-free gather sum
# with atomics, so K varies at rounding level between runs and the iteration counts move
# with it (measured 0-8%; C1, kappa 3.7e6, 1,010 against the oracle's 978 in one run)
ITER_BAR = {"cube": 0.03, "beam_vtu": 0.12, "C1": 0.08, "beam_highres_t4": 0.12}


@pytest.mark.parametrize("name", ["cube", "beam_vtu", "C1", "beam_highres_t4"])
def test_matfree_cg_against_direct(gpu, name):
    from tlfea_amd.fea import LinearStatic
    print(f"[matfree cg] {name}", flush=True)
    m = load_mesh(name)
    b = _body(m)
    ls = LinearStatic(b, tol=1e-13, max_iter=200000, matfree=True)
    ls.assemble()
    rc = ls.cg_solve()
    assert rc == 0, (ls.iters, ls.res)
    go, djo = O.grad(m)
    lam, mu = lame(E_STEEL, NU_STEEL)
    doff, dcol = O.dof_pattern(m)
    vo = O.stiffness_values(m, go, djo, lam, mu, doff, dcol)
    vo, fo = O.apply_dirichlet(doff, dcol, vo, m.f_ext, m.fixed)
    ud = O.band_cholesky_solve(doff, dcol, vo, fo)
    assert rel_err(_np(ls.u), ud) < CG_DISP_BAR[name]
    # the same Krylov sequence as the oracle's CG: the iterates after 10 steps agree to
    # rounding (a 1e-15 relative perturbation of K moves them by <= 3e-14 at 10 steps on
    # these meshes; by 30 steps finite-precision CG has amplified it to 1e-3 on the cube
    # and 5e-5 on beam_highres_t4, measured with the oracle), and at a moderate tolerance the iteration counts agree to
    # within ITER_BAR (finite-precision CG counts move with the summation order, the more
    # so the larger kappa)
    lsk = LinearStatic(b, tol=0.0, max_iter=10, matfree=True)
    lsk.assemble()
    lsk.cg_solve()
    uk, _, _ = O.cg(doff, dcol, vo, fo, tol=0.0, max_iter=10)
    assert lsk.iters == 10 and rel_err(_np(lsk.u), uk) < 1e-11
    ls8 = LinearStatic(b, tol=1e-8, max_iter=200000, matfree=True)
    ls8.assemble()
    assert ls8.cg_solve() == 0
    uo, ito, _ = O.cg(doff, dcol, vo, fo, tol=1e-8, max_iter=200000)
    assert abs(ls8.iters - ito) <= max(3, ITER_BAR[name] * ito), (ls8.iters, ito)
    assert rel_err(_np(ls8.u), uo) < 1e-6


@pytest.mark.parametrize("name", ["C1", "kuhn_t4"])
def test_two_level_linear_static(gpu, name):
    """LinearStatic(two_level=True), the Python mirror of the drop-in's
    UseTwoLevelPreconditioner: at tol 1e-12 it matches the oracle's direct solve within
    1e-10 (kappa * eps allows it on these beams) in a fifth of the matrix-free CG's
    iterations or fewer, and a second solve from the moved positions (the operator K(x)
    there) matches the matrix-free second solve."""
    import torch
    from tlfea_amd.fea import LinearStatic
    m = load_mesh(name)
    b = _body(m)
    ls = LinearStatic(b, tol=1e-12, max_iter=200000, two_level=True, cheb_degree=8, cheb_ratio=300.0)
    assert ls.solve() == 0, (ls.iters, ls.res)
    go, djo = O.grad(m)
    lam, mu = lame(E_STEEL, NU_STEEL)
    doff, dcol = O.dof_pattern(m)
    vo, fo = O.apply_dirichlet(doff, dcol, O.stiffness_values(m, go, djo, lam, mu, doff, dcol), m.f_ext, m.fixed)
    assert rel_err(_np(ls.u), O.band_cholesky_solve(doff, dcol, vo, fo)) < 1e-10
    it2 = ls.iters
    assert ls.solve() == 0
    u_tl = _np(ls.u).copy()
    b2 = _body(m)
    lm = LinearStatic(b2, tol=1e-12, max_iter=200000, matfree=True)
    assert lm.solve() == 0
    it_plain = lm.iters
    assert lm.solve() == 0
    print(f"[two-level LinearStatic] {name}: iterations {it2} vs {it_plain}", flush=True)
    assert it2 * 5 <= it_plain
    assert rel_err(u_tl, _np(lm.u)) < 1e-9
    X = m.xyz()
    moved = np.stack([_np(b.x[d]) - X[:, d] for d in range(3)], 1)
    moved_mf = np.stack([_np(b2.x[d]) - X[:, d] for d in range(3)], 1)
    assert rel_err(moved, moved_mf) < 1e-9
    ls.close()


def test_matfree_cg_iteration_limits(gpu):
    """Out of iterations the matrix-free CG reports ENOCONV with the relative residual of the
    returned u (tested after the loop); a tolerance just above that residual with the same
    limit converges (at the last allowed iteration at the latest)."""
    import torch
    from tlfea_amd import capi
    from tlfea_amd.fea import LinearStatic
    m = load_mesh("C1")
    b = _body(m)
    ls = LinearStatic(b, tol=1e-10, max_iter=7, matfree=True)
    ls.assemble()
    assert ls.cg_solve() == capi.TLFEA_ENOCONV and ls.iters == 7
    go, djo = O.grad(m)
    lam, mu = lame(E_STEEL, NU_STEEL)
    doff, dcol = O.dof_pattern(m)
    vo, fo = O.apply_dirichlet(doff, dcol, O.stiffness_values(m, go, djo, lam, mu, doff, dcol), m.f_ext, m.fixed)
    true = np.linalg.norm(fo - O.spmv(doff, dcol, vo, _np(ls.u))) / np.linalg.norm(fo)
    assert abs(ls.res - true) < 1e-8 * true, (ls.res, true)
    edge = LinearStatic(b, tol=1.01 * ls.res, max_iter=7, matfree=True)
    edge.assemble()
    assert edge.cg_solve() == 0 and edge.iters <= 7, (edge.iters, edge.res)


def test_matfree_solve_moves_positions(gpu):
    """LinearStaticSolver's beam scenario (tol 1e-10) through the matrix-free operator: the
    tip deflection sits in the Euler-Bernoulli band (beam_linear_static.cc:223-234)."""
    from tlfea_amd.fea import LinearStatic
    m = load_mesh("C1")
    b = _body(m)
    ls = LinearStatic(b, tol=1e-10, max_iter=50000, matfree=True)
    assert ls.solve() == 0
    z = _np(b.x[2])
    tip = m.meta["tip_nodes"]
    dz = (z - m.z)[tip].min()
    eb = 1000.0 / (3 * E_STEEL * 0.1 ** 4 / 12)
    assert 0.95 * eb < abs(dz) < 1.05 * eb


def test_matfree_cg_not_positive_definite(gpu):
    from tlfea_amd import capi
    from tlfea_amd.fea import LinearStatic
    m = load_mesh("cube")
    b = _body(m, E=-1e9)
    ls = LinearStatic(b, tol=1e-10, max_iter=100, matfree=True)
    ls.assemble()
    assert ls.cg_solve() == capi.TLFEA_ENOTPD
