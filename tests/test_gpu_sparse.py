"""GPU parity of patterns, assembly, Dirichlet elimination, SpMV, BLAS-1 and CG.

Bars: CSR patterns bit-exact (int); mass and stiffness values, SpMV and reductions
1e-12 relative; CG displacements 1e-10 relative against the banded Cholesky solve of the
oracle's matrix (at a CG tolerance of 1e-13).
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from helpers import E_STEEL, NU_STEEL, RHO_STEEL, deformed, lame, load_mesh, rel_err

pytestmark = pytest.mark.gpu

MESHES = ["cube", "beam_vtu", "C1", "beam_highres_t10", "beam_highres_t4", "kuhn_t4"]


def _body(m, E=E_STEEL, nu=NU_STEEL, rho=RHO_STEEL, eta=0.0, lambda_d=0.0):
    from tlfea_amd.fea import Body
    return Body(m, E, nu, rho, eta, lambda_d)


def _np(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("name", MESHES)
def test_patterns_bit_exact(gpu, name):
    print(f"[pattern] {name}", flush=True)
    m = load_mesh(name)
    b = _body(m)
    off, col = b.node_csr()
    ooff, ocol = O.node_pattern(m)
    assert np.array_equal(_np(off), ooff) and np.array_equal(_np(col), ocol)
    doff, dcol = b.dof_csr()
    odoff, odcol = O.dof_pattern(m)
    assert np.array_equal(_np(doff), odoff) and np.array_equal(_np(dcol), odcol)


@pytest.mark.parametrize("name", MESHES)
def test_mass_values(gpu, name):
    m = load_mesh(name)
    b = _body(m)
    ms = b.mass_sell()
    v = _np(b.mass_csr_values(ms))
    go, djo = O.grad(m)
    ooff, ocol = O.node_pattern(m)
    vo = O.mass_values(m, djo, RHO_STEEL, ooff, ocol)
    assert rel_err(v, vo) < 1e-12


@pytest.mark.parametrize("name", MESHES)
@pytest.mark.parametrize("damp", [False, True])
def test_stiffness_values(gpu, name, damp):
    import torch
    print(f"[stiffness] {name} damp={damp}", flush=True)
    m = load_mesh(name)
    eta, ld = (1e5, 3e4) if damp else (0.0, 0.0)
    b = _body(m, eta=eta, lambda_d=ld)
    x, y, z = deformed(m, seed=7, amp=5e-3)
    xt = [torch.as_tensor(a).cuda() for a in (x, y, z)]
    h = 0.37
    k = b.stiffness_sell(x=xt, h=h)
    v = _np(b.stiffness_csr_values(k))
    go, djo = O.grad(m)
    lam, mu = lame(E_STEEL, NU_STEEL)
    doff, dcol = O.dof_pattern(m)
    vo = O.stiffness_values(m, go, djo, lam, mu, doff, dcol, x, y, z, h=h, eta=eta, lambda_d=ld)
    assert rel_err(v, vo) < 1e-12


@pytest.mark.parametrize("name", ["cube", "C1", "beam_highres_t4"])
def test_dirichlet_and_spmv(gpu, name):
    import torch
    m = load_mesh(name)
    b = _body(m)
    k = b.stiffness_sell()
    f = b.f_ext.clone()
    b.apply_dirichlet(k, f)
    v = _np(b.stiffness_csr_values(k))
    go, djo = O.grad(m)
    lam, mu = lame(E_STEEL, NU_STEEL)
    doff, dcol = O.dof_pattern(m)
    vo = O.stiffness_values(m, go, djo, lam, mu, doff, dcol)
    vo, fo = O.apply_dirichlet(doff, dcol, vo, m.f_ext, m.fixed)
    assert rel_err(v, vo) < 1e-12
    assert np.array_equal(_np(f), fo)
    xv = np.random.default_rng(8).standard_normal(3 * m.n_nodes)
    y = _np(b.spmv(k, torch.as_tensor(xv).cuda()))
    assert rel_err(y, O.spmv(doff, dcol, vo, xv)) < 1e-12
    # the plain scalar-CSR SpMV on the reference format agrees as well
    from tlfea_amd import capi
    from tlfea_amd.capi import ptr
    doff_t, dcol_t = b.dof_csr()
    vals_t = b.stiffness_csr_values(k)
    xt = torch.as_tensor(xv).cuda()
    yt = torch.empty_like(xt)
    capi.call("tlfea_spmv_csr", 3 * m.n_nodes, ptr(doff_t), ptr(dcol_t), ptr(vals_t), ptr(xt), ptr(yt), None)
    assert rel_err(_np(yt), y) < 1e-13


def test_mass_spmv(gpu):
    import torch
    m = load_mesh("C1")
    b = _body(m)
    ms = b.mass_sell()
    xv = np.random.default_rng(9).standard_normal(3 * m.n_nodes)
    y = _np(b.spmv_mass(ms, torch.as_tensor(xv).cuda()))
    go, djo = O.grad(m)
    off, col = O.node_pattern(m)
    mv = O.mass_values(m, djo, RHO_STEEL, off, col)
    ref = np.stack([O.spmv(off, col, mv, xv[d::3]) for d in range(3)], axis=1).reshape(-1)
    assert rel_err(y, ref) < 1e-12


def test_blas1(gpu):
    import torch
    from tlfea_amd import capi
    from tlfea_amd.capi import ptr
    rng = np.random.default_rng(10)
    for n in (1, 63, 1000, 3 * 674081):
        a = rng.standard_normal(n)
        bb = rng.standard_normal(n)
        at, bt = torch.as_tensor(a).cuda(), torch.as_tensor(bb).cuda()
        out = torch.zeros(1, dtype=torch.float64, device="cuda")
        scr = torch.zeros(int(capi.lib().tlfea_reduce_scratch_size(n)), dtype=torch.float64, device="cuda")
        capi.call("tlfea_dot", n, ptr(at), ptr(bt), ptr(out), ptr(scr), None)
        assert abs(out.item() - np.dot(a, bb)) <= 1e-12 * np.sqrt(n) * np.linalg.norm(a) * np.linalg.norm(bb)
        capi.call("tlfea_nrm2", n, ptr(at), ptr(out), ptr(scr), None)
        assert abs(out.item() - np.linalg.norm(a)) <= 1e-13 * np.linalg.norm(a) * np.sqrt(n)
        capi.call("tlfea_axpy", n, 0.5, ptr(at), ptr(bt), None)
        assert rel_err(_np(bt), bb + 0.5 * a) < 1e-15
        capi.call("tlfea_scal", n, -2.0, ptr(bt), None)
        assert rel_err(_np(bt), -2.0 * (bb + 0.5 * a)) < 1e-15
    # determinism: the same reduction twice gives identical bits
    n = 3 * 674081
    a = torch.as_tensor(rng.standard_normal(n)).cuda()
    r1 = torch.zeros(1, dtype=torch.float64, device="cuda")
    r2 = torch.zeros(1, dtype=torch.float64, device="cuda")
    scr = torch.zeros(int(capi.lib().tlfea_reduce_scratch_size(n)), dtype=torch.float64, device="cuda")
    capi.call("tlfea_dot", n, ptr(a), ptr(a), ptr(r1), ptr(scr), None)
    capi.call("tlfea_dot", n, ptr(a), ptr(a), ptr(r2), ptr(scr), None)
    assert r1.item() == r2.item()


# Displacement bar per mesh.  1e-10 where the matrix allows it; the high-resolution VTU
# meshes have kappa(K_free) ~ 3e7-5e7 (measured with eigsh), so two FP64 solutions of the
# same system can differ by ~kappa * eps ~ 1e-8 whatever the method (the banded Cholesky's
# own relative residual there is ~3e-9): the bar is 20 * kappa * eps.
CG_DISP_BAR = {"cube": 1e-10, "beam_vtu": 1e-10, "C1": 1e-10, "beam_highres_t4": 20 * 3.2e7 * 2.2e-16}
# relative iteration-count bar at tol 1e-8: the assembly and the matrix-free gather sum
# with atomics, so K varies at rounding level between runs and the iteration counts move
# with it (measured 0-8%; C1, kappa 3.7e6, 1,010 against the oracle's 978 in one run)
ITER_BAR = {"cube": 0.03, "beam_vtu": 0.12, "C1": 0.08, "beam_highres_t4": 0.12}


@pytest.mark.parametrize("name", ["cube", "beam_vtu", "C1", "beam_highres_t4"])
def test_cg_against_direct(gpu, name):
    from tlfea_amd.fea import LinearStatic
    print(f"[cg] {name}", flush=True)
    m = load_mesh(name)
    b = _body(m)
    ls = LinearStatic(b, tol=1e-13, max_iter=200000)
    ls.assemble()
    rc = ls.cg_solve()
    assert rc == 0, (ls.iters, ls.res)
    assert ls.res < 1e-13
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
    lsk = LinearStatic(b, tol=0.0, max_iter=10)
    lsk.assemble()
    lsk.cg_solve()
    uk, _, _ = O.cg(doff, dcol, vo, fo, tol=0.0, max_iter=10)
    assert lsk.iters == 10 and rel_err(_np(lsk.u), uk) < 1e-11
    ls8 = LinearStatic(b, tol=1e-8, max_iter=200000)
    ls8.assemble()
    assert ls8.cg_solve() == 0
    uo, ito, _ = O.cg(doff, dcol, vo, fo, tol=1e-8, max_iter=200000)
    assert abs(ls8.iters - ito) <= max(3, ITER_BAR[name] * ito), (ls8.iters, ito)
    assert rel_err(_np(ls8.u), uo) < 1e-6


def test_cg_reference_tolerance_and_positions(gpu):
    """The reference's beam scenario (tol 1e-10): positions move by u, tip deflection within
    the Euler-Bernoulli band (beam_linear_static.cc:223-234)."""
    from tlfea_amd.fea import LinearStatic
    m = load_mesh("C1")
    b = _body(m)
    ls = LinearStatic(b, tol=1e-10, max_iter=50000)
    assert ls.solve() == 0
    z = _np(b.x[2])
    tip = m.meta["tip_nodes"]
    dz = (z - m.z)[tip].min()
    eb = 1000.0 / (3 * E_STEEL * 0.1 ** 4 / 12)
    assert 0.95 * eb < abs(dz) < 1.05 * eb


def test_cg_not_positive_definite(gpu):
    """Negative E makes K negative definite: CG aborts on p.Kp <= 0
    (LinearStaticSolver.cu:404-407)."""
    from tlfea_amd import capi
    from tlfea_amd.fea import LinearStatic
    m = load_mesh("cube")
    b = _body(m, E=-1e9)
    ls = LinearStatic(b, tol=1e-10, max_iter=100)
    ls.assemble()
    assert ls.cg_solve() == capi.TLFEA_ENOTPD


@pytest.mark.parametrize("matfree", [False, True])
def test_cg_zero_rhs(gpu, matfree):
    from tlfea_amd.fea import LinearStatic
    m = load_mesh("cube")
    m.f_ext = np.zeros(3 * m.n_nodes)
    b = _body(m)
    ls = LinearStatic(b, tol=1e-10, max_iter=100, matfree=matfree)
    ls.assemble()
    assert ls.cg_solve() == 0 and ls.iters == 0
    assert np.all(_np(ls.u) == 0)


# The mixed solve needs kappa(K) * eps_fp32 well below 1: the FP32 recursion stalls on the
# C1 beam (kappa ~ 3.7e6) and on the 200x20x20 beam (kappa ~ 1e8), and on beam_vtu it
# reaches ~7e-10 but not 1e-12, so only the well-conditioned cases are run here.
@pytest.mark.parametrize("name,tol,bar", [("cube", 1e-12, 1e-10), ("beam_vtu", 1e-8, None)])
def test_mixed_cg_against_direct(gpu, name, tol, bar):
    """FP32 Krylov recursion + FP64 reliable updates to a true FP64 residual of tol: the
    reported residual is the FP64 one and, on the cube, the displacements match the FP64
    direct solve."""
    from tlfea_amd.fea import LinearStatic
    m = load_mesh(name)
    b = _body(m)
    ls = LinearStatic(b, tol=tol, max_iter=200000, mixed=True)
    ls.assemble()
    assert ls.cg_solve() == 0, (ls.iters, ls.res)
    assert ls.res < tol and ls.refreshes >= 1
    go, djo = O.grad(m)
    lam, mu = lame(E_STEEL, NU_STEEL)
    doff, dcol = O.dof_pattern(m)
    vo = O.stiffness_values(m, go, djo, lam, mu, doff, dcol)
    vo, fo = O.apply_dirichlet(doff, dcol, vo, m.f_ext, m.fixed)
    if bar is not None:
        ud = O.band_cholesky_solve(doff, dcol, vo, fo)
        assert rel_err(_np(ls.u), ud) < bar
    # the reported residual is the FP64 one: recompute it with the oracle
    r = fo - O.spmv(doff, dcol, vo, _np(ls.u))
    assert np.linalg.norm(r) / np.linalg.norm(fo) < 2 * tol
