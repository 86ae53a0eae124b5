"""GPU parity of the compressible Mooney-Rivlin material (SetMooneyRivlin,
`src/materials/MooneyRivlin.cuh`) against the oracle, with the material switched to
Mooney-Rivlin on both sides.

Bars as for SVK: F, P, internal forces and assembled stiffness 1e-12 relative; the Nesterov
trajectory 1e-8 relative (summation order amplified by the iterations); the drop-in
driver's Nesterov run 1e-6 and its linear-static solve 1e-10, as their SVK tests.
Coefficients: steel-like (shear modulus 2 (mu10 + mu01) = 80 GPa, bulk modulus kappa =
160 GPa), so the step sizes of the SVK tests stay stable.
"""
import os

import numpy as np
import pytest

import oracle as O
from helpers import E_STEEL, MESHES, NU_STEEL, RHO_STEEL, deformed, load_mesh, rel_err

pytestmark = pytest.mark.gpu

MR = (30e9, 10e9, 160e9)


def _body(m, eta=0.0, lambda_d=0.0):
    from tlfea_amd.fea import Body
    b = Body(m, E_STEEL, NU_STEEL, RHO_STEEL, eta, lambda_d)
    b.set_mooney_rivlin(*MR)
    return b


def _np(t):
    return t.detach().cpu().numpy()


# Perturbation amplitudes (fractions of the mesh length) per mesh.  P(F) is ill-conditioned
# at both ends: near inverted elements (J^-2/3, J^-4/3, F^-T) and at small strain under the
# rigid rotation (t1 (F - I1/3 F^-T) cancels to O(strain)).  These keep det F within
# [0.56, 1.49] and the change of P under 2e-16 relative noise on the positions below
# 3e-13, measured on the oracle, so the 1e-12 bar tests the arithmetic.
AMP = {"cube": 1e-2, "C1": 1e-3, "kuhn_t4": 1e-3, "beam_highres_t10": 3e-4}



@pytest.mark.parametrize("name", ["cube", "C1", "kuhn_t4"])
def test_mr_calc_p(gpu, name):
    import torch
    print(f"[mr calc_p] {name}", flush=True)
    m = load_mesh(name)
    b = _body(m)
    x, y, z = deformed(m, seed=31, amp=AMP[name])
    for i, a in enumerate((x, y, z)):
        b.x[i].copy_(torch.as_tensor(a))
    F, P, _, _ = b.calc_p()
    go, _ = O.grad(m)
    with O.material("mr", *MR):
        Fo, Po, _, _ = O.calc_p(m, go, 0.0, 0.0, x, y, z)
    assert rel_err(_np(b.qp_ref(F)), Fo) < 1e-12
    assert rel_err(_np(b.qp_ref(P)), Po) < 1e-12


@pytest.mark.parametrize("name", ["cube", "C1", "beam_highres_t10", "kuhn_t4"])
@pytest.mark.parametrize("damp", [False, True])
@pytest.mark.parametrize("kind", ["atomic", "planned", "rest"])
def test_mr_internal_force(gpu, name, damp, kind):
    import torch
    print(f"[mr force] {name} damp={damp} kind={kind}", flush=True)
    m = load_mesh(name)
    eta, ld = (1e6, 1e5) if damp else (0.0, 0.0)
    b = _body(m, eta, ld)
    x, y, z = deformed(m, seed=32, amp=AMP[name])
    xt = [torch.as_tensor(a).cuda() for a in (x, y, z)]
    v = np.random.default_rng(33).standard_normal(3 * m.n_nodes) * 1e-2
    vt = torch.as_tensor(v).cuda()
    if kind == "rest" and b.plan_sizes()["max_block_nodes"] * (12 if damp else 9) * 8 > 160 * 1024:
        # blocks beyond the rest-form kernel's LDS budget: the entry point refuses the
        # call before any launch (the drop-in then leaves the rest positions out)
        from tlfea_amd.capi import TLFEA_ECAPACITY, TlfeaError
        with pytest.raises(TlfeaError, match="LDS budget") as e:
            b.internal_force_rest(x=xt, v=vt if damp else None)
        assert e.value.code == TLFEA_ECAPACITY
        return
    force = {"atomic": b.internal_force, "planned": b.internal_force_planned, "rest": b.internal_force_rest}[kind]
    f = force(x=xt, v=vt if damp else None)
    go, djo = O.grad(m)
    with O.material("mr", *MR):
        fo = O.internal_force(m, go, djo, 0.0, 0.0, x, y, z, v=v if damp else None, eta=eta, lambda_d=ld)
    assert rel_err(_np(f), fo) < 1e-12
    # the rest state is force-free
    f0 = force(x=[t.clone() for t in b.X0])
    assert np.abs(_np(f0)).max() < 1e-9 * np.abs(fo).max()


@pytest.mark.parametrize("name", ["cube", "C1", "kuhn_t4"])
@pytest.mark.parametrize("damp", [False, True])
def test_mr_stiffness_values(gpu, name, damp):
    import torch
    print(f"[mr stiffness] {name} damp={damp}", flush=True)
    m = load_mesh(name)
    eta, ld = (1e5, 3e4) if damp else (0.0, 0.0)
    b = _body(m, eta, ld)
    x, y, z = deformed(m, seed=34, amp=AMP[name] / 2)
    xt = [torch.as_tensor(a).cuda() for a in (x, y, z)]
    h = 0.37
    v = _np(b.stiffness_csr_values(b.stiffness_sell(x=xt, h=h)))
    go, djo = O.grad(m)
    doff, dcol = O.dof_pattern(m)
    with O.material("mr", *MR):
        vo = O.stiffness_values(m, go, djo, 0.0, 0.0, doff, dcol, x, y, z, h=h, eta=eta, lambda_d=ld)
    assert rel_err(v, vo) < 1e-12


def test_mr_at_rest_equals_matched_linear_elasticity(gpu):
    """At F = I the Mooney-Rivlin tangent is isotropic linear elasticity with mu = 2 (mu10 +
    mu01) and lambda = kappa - 2 mu / 3: the assembled K equals SVK's with those Lame
    parameters."""
    from tlfea_amd.fea import Body
    m = load_mesh("C1")
    b = _body(m)
    k_mr = _np(b.stiffness_csr_values(b.stiffness_sell()))
    mu = 2 * (MR[0] + MR[1])
    lam = MR[2] - 2 * mu / 3
    E = mu * (3 * lam + 2 * mu) / (lam + mu)
    nu = lam / (2 * (lam + mu))
    s = Body(m, E, nu, RHO_STEEL)
    k_svk = _np(s.stiffness_csr_values(s.stiffness_sell()))
    assert rel_err(k_mr, k_svk) < 1e-12


@pytest.mark.parametrize("force", ["planned", "rest"])
def test_mr_nesterov_trajectory(gpu, force):
    from tlfea_amd import capi
    from tlfea_amd.fea import Solver
    print(f"[mr nesterov] C1 force={force}", flush=True)
    m = load_mesh("C1")
    m.f_ext = np.random.default_rng(21).standard_normal(3 * m.n_nodes) * 50.0
    b = _body(m)
    s = Solver(b, capi.SOLVER_NESTEROV, planned=True, rest=force == "rest")
    prob = O.Problem(m, E_STEEL, NU_STEEL, RHO_STEEL)
    p = capi.NesterovParams(alpha=5e-9, rho=1e10, inner_tol=0.0, outer_tol=0.0, max_outer=2, max_inner=25,
                            time_step=1e-3)
    s.set_parameters(p)
    st = prob.new_state()
    import torch
    x, y, z = deformed(m, seed=22, amp=2e-5, rot=0.0)
    for d, a in enumerate((x, y, z)):
        b.x[d].copy_(torch.as_tensor(a))
    st["x"][:], st["y"][:], st["z"][:] = x, y, z
    X0 = m.xyz()
    for _ in range(2):
        s.solve()
        with O.material("mr", *MR):
            n_inner = prob.nesterov_step(st, p.alpha, p.rho, p.inner_tol, p.outer_tol, p.max_outer, p.max_inner,
                                         p.time_step)
        assert s.stats()["inner"] == n_inner == 50
        xg = np.stack([t.cpu().numpy() for t in b.x], 1)
        xo = np.stack([st["x"], st["y"], st["z"]], 1)
        assert rel_err(xg - X0, xo - X0) < 1e-8
        assert rel_err(s.v_guess().cpu().numpy(), st["v_guess"]) < 1e-8


def test_mr_rejected_where_not_covered(gpu):
    """The FP32 force is SVK-only and says so; an unknown model number is refused."""
    import ctypes as C
    import torch
    from tlfea_amd import capi
    from tlfea_amd.capi import TlfeaError
    m = load_mesh("C1")
    b = _body(m)
    u32 = torch.zeros(m.n_nodes, dtype=torch.float32, device="cuda")
    f32 = torch.zeros(3 * m.n_nodes, dtype=torch.float32, device="cuda")
    g32 = torch.zeros(b.grad_t.numel(), dtype=torch.float32, device="cuda")
    d32 = torch.zeros(b.detJ_t.numel(), dtype=torch.float32, device="cuda")
    rc = capi.lib().tlfea_internal_force_f32(b.nn, b.n_elem, b.n_nodes, C.c_void_p(b.conn.data_ptr()),
                                             C.c_void_p(u32.data_ptr()), C.c_void_p(u32.data_ptr()),
                                             C.c_void_p(u32.data_ptr()), C.c_void_p(g32.data_ptr()),
                                             C.c_void_p(d32.data_ptr()), C.byref(b.mat),
                                             C.c_void_p(f32.data_ptr()), 0, None)
    assert rc == capi.TLFEA_EINVAL and b"St. Venant-Kirchhoff" in capi.lib().tlfea_last_error()
    b.mat.model = 7
    with pytest.raises(TlfeaError, match="unknown material model"):
        b.internal_force_planned()


@pytest.mark.parametrize("name", ["C1", "kuhn_t4"])
def test_mr_tangent_apply(gpu, name):
    """The matrix-free Mooney-Rivlin tangent (dP = A : dF without A) against the oracle's
    assembled tangent times p: at a deformed state (per-point and affine geometry) and at
    rest (the linear-elasticity form with the matched Lame parameters)."""
    import torch
    m = load_mesh(name)
    b = _body(m)
    go, djo = O.grad(m)
    doff, dcol = O.dof_pattern(m)
    pt = torch.as_tensor(np.random.default_rng(35).standard_normal(3 * m.n_nodes)).cuda()
    with O.material("mr", *MR):
        vr, _ = O.apply_dirichlet(doff, dcol, O.stiffness_values(m, go, djo, 0.0, 0.0, doff, dcol), m.f_ext, m.fixed)
        x, y, z = deformed(m, seed=36, amp=AMP[name] / 2)
        vd, _ = O.apply_dirichlet(doff, dcol, O.stiffness_values(m, go, djo, 0.0, 0.0, doff, dcol, x, y, z), m.f_ext,
                                  m.fixed)
    assert rel_err(_np(b.tangent_apply(pt, x=b.X0)), O.spmv(doff, dcol, vr, _np(pt))) < 1e-12
    for i, a in enumerate((x, y, z)):
        b.x[i].copy_(torch.as_tensor(a))
    ref = O.spmv(doff, dcol, vd, _np(pt))
    assert rel_err(_np(b.tangent_apply(pt)), ref) < 1e-12
    assert rel_err(_np(b.tangent_apply(pt, affine=False)), ref) < 1e-12


@pytest.mark.parametrize("name", ["cube", "C1"])
def test_mr_damped_tangent_apply(gpu, name):
    """A damped Mooney-Rivlin body's K + C without assembly (C p = F (lambda_d tr(dE) I +
    2 eta dE) added to A : dF) against the oracle's assembled damped Mooney-Rivlin tangent
    times p, at rest and at a deformed state, 1e-12."""
    import torch
    m = load_mesh(name)
    eta, ld = 1e10, 2e10
    b = _body(m, eta=eta, lambda_d=ld)
    go, djo = O.grad(m)
    doff, dcol = O.dof_pattern(m)
    p = np.random.default_rng(37).standard_normal(3 * m.n_nodes)
    pt = torch.as_tensor(p).cuda()
    x, y, z = deformed(m, seed=38, amp=AMP[name] / 2)
    with O.material("mr", *MR):
        vr = O.stiffness_values(m, go, djo, 0.0, 0.0, doff, dcol, eta=eta, lambda_d=ld)
        vd = O.stiffness_values(m, go, djo, 0.0, 0.0, doff, dcol, x, y, z, eta=eta, lambda_d=ld)
    assert rel_err(_np(b.tangent_apply(pt, x=b.X0, masked=False, damping=True)), O.spmv(doff, dcol, vr, p)) < 1e-12
    xt = [torch.as_tensor(a).cuda() for a in (x, y, z)]
    ref = O.spmv(doff, dcol, vd, p)
    assert rel_err(_np(b.tangent_apply(pt, x=xt, masked=False, damping=True)), ref) < 1e-12
    assert rel_err(_np(b.tangent_apply(pt, x=xt, masked=False)), ref) > 1e-3  # the block is in it


def test_mr_matfree_cg_against_direct(gpu):
    """Linear-static solve of a Mooney-Rivlin C1 beam through the matrix-free CG against the
    oracle's banded Cholesky solve of the same tangent, to the assembled CG's bar."""
    from tlfea_amd.fea import LinearStatic
    m = load_mesh("C1")
    b = _body(m)
    ls = LinearStatic(b, tol=1e-13, max_iter=200000, matfree=True)
    ls.assemble()
    assert ls.cg_solve() == 0
    go, djo = O.grad(m)
    doff, dcol = O.dof_pattern(m)
    with O.material("mr", *MR):
        vo = O.stiffness_values(m, go, djo, 0.0, 0.0, doff, dcol)
    vo, fo = O.apply_dirichlet(doff, dcol, vo, m.f_ext, m.fixed)
    ud = O.band_cholesky_solve(doff, dcol, vo, fo)
    assert rel_err(_np(ls.u), ud) < 1e-10


def _run_driver(args, tmp_path):
    import json
    import subprocess
    from tlfea_amd import build as B
    out = str(tmp_path / "pos.bin")
    print("[driver]", " ".join(args), flush=True)
    r = subprocess.run([B.DRIVER, *args, "--out", out], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    info = json.loads(r.stdout.strip().splitlines()[-1])
    return info, np.fromfile(out, dtype=np.float64).reshape(3, -1)


def test_mr_dropin_nesterov_on_cube(gpu, tmp_path):
    """SetMooneyRivlin through the C++ drop-in: the Nesterov example scenario on cube.1 with a
    Mooney-Rivlin material matching E = 7e8, nu = 0.33 at small strain, against the
    oracle's loop with the same material."""
    mr = (1.0e8, 0.3157e8, 6.86e8)
    args = ["nesterov", "--mesh", "tetgen:" + os.path.join(MESHES, "cube.1"), "--fix", "z0", "--node-load", "6,1000",
            "--E", "7e8", "--nu", "0.33", "--rho0", "2700", "--steps", "3", "--mr", ",".join(map(repr, mr))]
    _, pos = _run_driver(args, tmp_path)
    m = load_mesh("cube")
    prob = O.Problem(m, 7e8, 0.33, 2700.0)
    st = prob.new_state()
    with O.material("mr", *mr):
        for _ in range(3):
            prob.nesterov_step(st, 1e-8, 1e14, 1e-6, 1e-6, 5, 300, 1e-3)
    du = pos.T - m.xyz()
    duo = np.stack([st["x"], st["y"], st["z"]], 1) - m.xyz()
    assert rel_err(du, duo) < 1e-6


@pytest.mark.parametrize("assembled", [False, True])
def test_mr_dropin_linear_static(gpu, tmp_path, assembled):
    """LinearStaticSolver on a Mooney-Rivlin element (matrix-free by default, or assembled
    as the reference does); the displacements match the oracle's direct solve of the
    Mooney-Rivlin K within 1e-10."""
    info, pos = _run_driver(["linear_static", "--mesh", "kuhn:20,2,2", "--tol", "1e-12", "--max-iter", "200000",
                             "--load", "-1000", "--mr", ",".join(map(repr, MR)), "--assembled",
                             "1" if assembled else "0"], tmp_path)
    m = load_mesh("C1")
    g, dj = O.grad(m)
    off, col = O.dof_pattern(m)
    with O.material("mr", *MR):
        vals = O.stiffness_values(m, g, dj, 0.0, 0.0, off, col)
    vals, fb = O.apply_dirichlet(off, col, vals, m.f_ext, m.fixed)
    u = O.band_cholesky_solve(off, col, vals, fb).reshape(-1, 3)
    assert rel_err(pos.T - m.xyz(), u) < 1e-10
    assert info["residual"] < 1e-12
    assert info["matrix_free"] is (not assembled)
