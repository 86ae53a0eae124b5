"""GPU parity of the SyncedNesterov / SyncedAdamW / SyncedAdamWNocoop loops against the
oracle's restatement of the same loops.

Trajectories run with zero tolerances (so both sides execute exactly max_outer x max_inner
iterations) from a deformed start under a load on every DOF, and are compared after each
Solve(): displacements, v_guess, v_prev and lambda.  Both sides do the same arithmetic and
differ only in summation order, which the iterations amplify: Nesterov within 1e-8
relative; AdamW within 1e-6, because its normalised step m/(sqrt(v)+eps) turns any
gradient entry near rounding level into a full +-lr step (from the rest state, where
f_int is pure rounding noise, AdamW trajectories of any two FP64 implementations
diverge -- measured on the CPU by perturbing the oracle's inputs at 1e-14).  The stopping
tests are checked separately with the reference's example tolerances.
"""
import numpy as np
import pytest

import oracle as O
from helpers import E_STEEL, NU_STEEL, RHO_STEEL, deformed, load_mesh, rel_err

pytestmark = pytest.mark.gpu


def _setup(name, kind, E=E_STEEL, nu=NU_STEEL, rho=RHO_STEEL, eta=0.0, lambda_d=0.0, planned=True,
           random_load=False, rest=False):
    from tlfea_amd.fea import Body, Solver
    m = load_mesh(name)
    if random_load:  # loads on every DOF keep every gradient entry far above rounding noise
        m.f_ext = np.random.default_rng(21).standard_normal(3 * m.n_nodes) * 50.0
    b = Body(m, E, nu, rho, eta, lambda_d)
    s = Solver(b, kind, planned=planned, rest=rest)
    prob = O.Problem(m, E, nu, rho, eta=eta, lambda_d=lambda_d)
    return m, b, s, prob


def _start_deformed(m, b, st, amp=2e-5):
    """Same deformed start on both sides (x_prev of the first step); the constraint targets
    stay at the rest positions."""
    import torch
    x, y, z = deformed(m, seed=22, amp=amp, rot=0.0)
    for d, a in enumerate((x, y, z)):
        b.x[d].copy_(torch.as_tensor(a))
    st["x"][:], st["y"][:], st["z"][:] = x, y, z


def _compare(m, b, s, prob, st, tol=1e-9):
    x = np.stack([t.cpu().numpy() for t in b.x], 1)
    xo = np.stack([st["x"], st["y"], st["z"]], 1)
    X0 = m.xyz()
    assert rel_err(x - X0, xo - X0) < tol
    assert rel_err(s.v_guess().cpu().numpy(), st["v_guess"]) < tol
    assert rel_err(s.v_prev().cpu().numpy(), st["v_prev"]) < tol
    if len(m.fixed):
        assert rel_err(s.lambda_guess().cpu().numpy(), st["lam"]) < tol


@pytest.mark.parametrize("name,damp,force", [("cube", False, "planned"), ("C1", False, "planned"),
                                             ("C1", True, "planned"), ("C1", False, "atomic"),
                                             ("kuhn_t4", False, "planned"), ("C1", False, "rest"), ("C1", True, "rest"),
                                             ("kuhn_t4", False, "rest"), ("C1", False, "rest_mf"),
                                             ("cube", False, "rest_mf"), ("kuhn_t4", False, "rest_mf")])
def test_nesterov_trajectory(gpu, name, damp, force):
    """force "rest_mf": the mass term applied per element inside the force kernel
    (tlfea_solver_set_matrix_free_mass) against the oracle's assembled mass."""
    from tlfea_amd import capi
    print(f"[nesterov] {name} damp={damp} force={force}", flush=True)
    eta, ld = (2e5, 1e5) if damp else (0.0, 0.0)
    m, b, s, prob = _setup(name, capi.SOLVER_NESTEROV, eta=eta, lambda_d=ld, planned=force != "atomic",
                           random_load=True, rest=force in ("rest", "rest_mf"))
    if force == "rest_mf":
        s.set_matrix_free_mass(RHO_STEEL)
    alpha = 2e-10 if name == "cube" else 5e-9  # stable step for steel on each mesh
    p = capi.NesterovParams(alpha=alpha, rho=1e10, inner_tol=0.0, outer_tol=0.0, max_outer=2, max_inner=25,
                            time_step=1e-3)
    s.set_parameters(p)
    st = prob.new_state()
    _start_deformed(m, b, st)
    for step in range(2):
        s.solve()
        n_inner = prob.nesterov_step(st, p.alpha, p.rho, p.inner_tol, p.outer_tol, p.max_outer, p.max_inner,
                                     p.time_step)
        assert s.stats()["inner"] == n_inner == 50
        _compare(m, b, s, prob, st, tol=1e-8)


def test_nesterov_stopping_rules(gpu):
    """Reference tolerances on the cube (test_feat10_nesterov.cc:189: alpha 1e-8, rho 1e14,
    tolerances 1e-6, 5 outer x 300 inner, dt 1e-3): same stopping decisions as the oracle."""
    from tlfea_amd import capi
    m, b, s, prob = _setup("cube", capi.SOLVER_NESTEROV, E=7e8, nu=0.33, rho=2700.0)
    p = capi.NesterovParams(alpha=1e-8, rho=1e14, inner_tol=1e-6, outer_tol=1e-6, max_outer=5, max_inner=300,
                            time_step=1e-3)
    s.set_parameters(p)
    st = prob.new_state()
    for step in range(3):
        s.solve()
        n = prob.nesterov_step(st, p.alpha, p.rho, p.inner_tol, p.outer_tol, p.max_outer, p.max_inner, p.time_step)
        assert abs(s.stats()["inner"] - n) <= 2, (s.stats(), n)
        _compare(m, b, s, prob, st, tol=1e-6)


@pytest.mark.parametrize("nocoop", [False, True])
@pytest.mark.parametrize("name", ["cube", "C1"])
@pytest.mark.parametrize("rest", [False, True])
def test_adamw_trajectory(gpu, nocoop, name, rest):
    from tlfea_amd import capi
    print(f"[adamw] {name} nocoop={nocoop} rest={rest}", flush=True)
    kind = capi.SOLVER_ADAMW_NOCOOP if nocoop else capi.SOLVER_ADAMW
    m, b, s, prob = _setup(name, kind, eta=1e4, lambda_d=1e3, random_load=True, rest=rest)
    p = capi.AdamWParams(lr=2e-4, beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=1e-4, lr_decay=0.995,
                         inner_tol=0.0, outer_tol=0.0, rho=1e10, max_outer=2, max_inner=30, time_step=1e-3,
                         convergence_check_interval=10, inner_rtol=0.0)
    s.set_parameters(p)
    st = prob.new_state()
    _start_deformed(m, b, st)
    for step in range(2):
        s.solve()
        n = prob.adamw_step(st, p.lr, p.beta1, p.beta2, p.eps, p.weight_decay, p.lr_decay, p.inner_tol, p.outer_tol,
                            p.rho, p.max_outer, p.max_inner, p.time_step, p.convergence_check_interval, p.inner_rtol,
                            nocoop=nocoop)
        assert s.stats()["inner"] == n
        _compare(m, b, s, prob, st, tol=1e-6)


@pytest.mark.parametrize("nocoop", [False, True])
def test_adamw_stopping_rules(gpu, nocoop):
    """beam_simulation.cc:226 parameters (tolerances on) on the cube: same decisions."""
    from tlfea_amd import capi
    kind = capi.SOLVER_ADAMW_NOCOOP if nocoop else capi.SOLVER_ADAMW
    m, b, s, prob = _setup("cube", kind, E=7e8, nu=0.33, rho=2700.0, eta=0.0, lambda_d=0.1)
    p = capi.AdamWParams(lr=2e-4, beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=1e-4, lr_decay=0.995,
                         inner_tol=1e-1, outer_tol=1e-6, rho=1e14, max_outer=5, max_inner=300, time_step=1e-3,
                         convergence_check_interval=10, inner_rtol=0.0)
    s.set_parameters(p)
    st = prob.new_state()
    for step in range(3):
        s.solve()
        n = prob.adamw_step(st, p.lr, p.beta1, p.beta2, p.eps, p.weight_decay, p.lr_decay, p.inner_tol, p.outer_tol,
                            p.rho, p.max_outer, p.max_inner, p.time_step, p.convergence_check_interval, p.inner_rtol,
                            nocoop=nocoop)
        assert abs(s.stats()["inner"] - n) <= 10, (s.stats(), n)


def test_nesterov_rel_tol_reaches_residual(gpu):
    """The optional relative stop: the final step gradient (oracle, fixed DOFs excluded) is
    at most 1e-8 of the initial one."""
    from tlfea_amd import capi
    m, b, s, prob = _setup("cube", capi.SOLVER_NESTEROV, E=7e8, nu=0.33, rho=2700.0)
    p = capi.NesterovParams(alpha=1e-8, rho=1e14, inner_tol=0.0, outer_tol=0.0, max_outer=1, max_inner=200000,
                            time_step=1e-3)
    s.set_parameters(p)
    s.set_rel_tol(1e-8)
    s.set_poll_interval(64)
    s.solve()
    st = s.stats()
    assert st["rel"] <= 1e-8 and st["inner"] < 200000
    X0 = m.xyz()
    free = np.ones(3 * m.n_nodes, bool)
    for n in m.fixed:
        free[3 * n:3 * n + 3] = False
    z = np.zeros(3 * m.n_nodes)
    g0 = prob.step_gradient(X0[:, 0], X0[:, 1], X0[:, 2], z, z, np.zeros(3 * len(m.fixed)), p.time_step, p.rho)
    v = s.v_guess().cpu().numpy()
    # v_prev after Solve() equals v_guess; evaluate g at the converged v with v_prev = 0
    g1 = prob.step_gradient(X0[:, 0], X0[:, 1], X0[:, 2], v, z, np.zeros(3 * len(m.fixed)), p.time_step, p.rho)
    assert np.linalg.norm(g1[free]) <= 2e-8 * np.linalg.norm(g0[free])


@pytest.mark.parametrize("kind_name", ["nesterov", "adamw"])
def test_prescribed_motion_trajectory(gpu, kind_name):
    """Constraint targets moved away from the rest positions (UpdateConstraintTargets,
    FEAT10Data.cuh:635-646): the fixed nodes are pulled by a rigid translation plus a small
    rotation through the augmented-Lagrangian outer loop; trajectories match the oracle's
    restatement with the same targets."""
    import torch
    from tlfea_amd import capi
    kind = capi.SOLVER_NESTEROV if kind_name == "nesterov" else capi.SOLVER_ADAMW
    m, b, s, prob = _setup("cube", kind, random_load=True)
    L = np.ptp(m.xyz(), axis=0).max()
    fx = np.asarray(m.fixed)
    X = m.xyz()[fx]
    c, sn = np.cos(1e-3), np.sin(1e-3)
    R = np.array([[c, -sn, 0.0], [sn, c, 0.0], [0.0, 0.0, 1.0]])
    T = X @ R.T + np.array([1e-4, -2e-4, 5e-5]) * L
    for d, ref in enumerate((prob.xref, prob.yref, prob.zref)):
        ref[fx] = T[:, d]
        tgt = s.xref[d].cpu().numpy()
        tgt[fx] = T[:, d]
        s.xref[d].copy_(torch.as_tensor(tgt))
    st = prob.new_state()
    _start_deformed(m, b, st)
    d0 = np.abs(np.stack([st["x"], st["y"], st["z"]], 1)[fx] - T).max()
    if kind_name == "nesterov":
        p = capi.NesterovParams(alpha=2e-10, rho=1e10, inner_tol=0.0, outer_tol=0.0, max_outer=2, max_inner=25,
                                time_step=1e-3)
        s.set_parameters(p)
        for _ in range(2):
            s.solve()
            prob.nesterov_step(st, p.alpha, p.rho, 0.0, 0.0, p.max_outer, p.max_inner, p.time_step)
            _compare(m, b, s, prob, st, tol=1e-8)
    else:
        p = capi.AdamWParams(lr=2e-4, beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=1e-4, lr_decay=0.995,
                             inner_tol=0.0, outer_tol=0.0, rho=1e10, max_outer=2, max_inner=30, time_step=1e-3,
                             convergence_check_interval=10, inner_rtol=0.0)
        s.set_parameters(p)
        for _ in range(2):
            s.solve()
            prob.adamw_step(st, p.lr, p.beta1, p.beta2, p.eps, p.weight_decay, p.lr_decay, 0.0, 0.0, p.rho,
                            p.max_outer, p.max_inner, p.time_step, p.convergence_check_interval, 0.0)
            _compare(m, b, s, prob, st, tol=1e-6)
    # the penalty pulls the fixed nodes towards the moved targets
    x = np.stack([t.cpu().numpy() for t in b.x], 1)
    assert np.abs(x[fx] - T).max() < d0


def _step_newton_extended(m, prob, dt, rho):
    """v* of one implicit step from rest, M v/dt + f_int(X + dt v) - f_ext + rho dt^2 v|fixed = 0
    (SyncedNesterov.cu:32-95 at lambda = 0, v_prev = 0), by Newton with f_int in longdouble
    from the displacement dt v itself (helpers.fint_extended) and the oracle's tangent and
    mass; returns (v*, g(.) as a function)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    from helpers import fint_extended
    n3 = 3 * m.n_nodes
    fixed = np.zeros(n3, bool)
    fixed.reshape(-1, 3)[m.fixed] = True
    M3 = sp.kron(sp.csr_matrix((prob.m_val, prob.m_col, prob.m_off)), sp.eye(3)).tocsr()
    off, col = O.dof_pattern(m)
    X = m.xyz()

    def grad(v):
        u = dt * v.reshape(-1, 3)
        f = fint_extended(m, None, None, None, prob.lam, prob.mu, u=u).astype(np.float64)
        return M3 @ v / dt + f - m.f_ext + rho * dt * dt * v * fixed

    v = np.zeros(n3)
    for _ in range(8):
        x = X + dt * v.reshape(-1, 3)
        vals = O.stiffness_values(m, prob.g, prob.detJ, prob.lam, prob.mu, off, col, x[:, 0].copy(),
                                  x[:, 1].copy(), x[:, 2].copy())
        J = (M3 / dt + dt * sp.csr_matrix((vals, col, off)) + sp.diags(rho * dt * dt * fixed)).tocsc()
        dv = spl.spsolve(J, -grad(v))
        v += dv
        if np.linalg.norm(dv) < 1e-14 * np.linalg.norm(v):
            break
    return v, grad


@pytest.mark.parametrize("kind", ["nesterov", "adamw", "adamw_nocoop"])
def test_matrix_free_mass_matches_assembled(gpu, kind):
    """The velocity solvers with the mass term applied per element inside the force kernel
    equal the same solvers on the assembled mass, run side by side from the same deformed
    start: positions, v_guess and f_int after each of two Solve() calls (Nesterov 1e-9,
    AdamW 1e-6, see the module note); a damped body is refused."""
    import torch
    from tlfea_amd import capi
    from tlfea_amd.fea import Body, Solver
    kinds = {"nesterov": capi.SOLVER_NESTEROV, "adamw": capi.SOLVER_ADAMW, "adamw_nocoop": capi.SOLVER_ADAMW_NOCOOP}
    m = load_mesh("C1")
    m.f_ext = np.random.default_rng(21).standard_normal(3 * m.n_nodes) * 50.0
    runs = []
    for mf in (False, True):
        b = Body(m, E_STEEL, NU_STEEL, RHO_STEEL)
        s = Solver(b, kinds[kind], planned=True, rest=True)
        if mf:
            s.set_matrix_free_mass(RHO_STEEL)
        if kind == "nesterov":
            p = capi.NesterovParams(alpha=5e-9, rho=1e10, inner_tol=0.0, outer_tol=0.0, max_outer=2, max_inner=25,
                                    time_step=1e-3)
        else:
            p = capi.AdamWParams(lr=2e-4, beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=1e-4, lr_decay=0.995,
                                 inner_tol=0.0, outer_tol=0.0, rho=1e10, max_outer=2, max_inner=30, time_step=1e-3,
                                 convergence_check_interval=10, inner_rtol=0.0)
        s.set_parameters(p)
        x, y, z = deformed(m, seed=22, amp=2e-5, rot=0.0)
        for d, a in enumerate((x, y, z)):
            b.x[d].copy_(torch.as_tensor(a))
        runs.append((b, s))
    tol = 1e-9 if kind == "nesterov" else 1e-6
    X0 = m.xyz()
    for step in range(2):
        out = []
        for b, s in runs:
            s.solve()
            out.append((np.stack([t.cpu().numpy() for t in b.x], 1) - X0, s.v_guess().cpu().numpy(),
                        b.f_int.cpu().numpy()))
        assert runs[0][1].stats()["inner"] == runs[1][1].stats()["inner"]
        for k in range(3):
            assert rel_err(out[1][k], out[0][k]) < tol, (step, k)
    for b, s in runs:
        s.close()
        b.close()
    bd = Body(m, E_STEEL, NU_STEEL, RHO_STEEL, 1e4, 1e3)
    sd = Solver(bd, capi.SOLVER_NESTEROV, planned=True, rest=True)
    with pytest.raises(RuntimeError):
        sd.set_matrix_free_mass(RHO_STEEL)
    sd.close()
    bd.close()


@pytest.mark.parametrize("mf", [False, True])
def test_nesterov_c1_step_to_1e8_with_restart(gpu, mf):
    """The C1 beam (steel, 1000 N tip load): one implicit step of 1e-3 s solved by
    SyncedNesterov to ||g_free|| <= 1e-8 ||g_free(0)||, with the momentum restart (the
    reference's t-sequence alone converges sublinearly: the oracle's restatement is at 1e-3
    after 40,000 iterations).  The force kernel takes the displacement u = (x_prev - X0) +
    dt v: from absolute positions, the coordinates' rounding (1e-16 m against a 2e-4 m
    deflection) puts a floor near 2e-7 under the relative gradient, which the oracle's own
    absolute-position gradient shows at the solution.  Checked against the Newton solution
    of the same step with f_int in longdouble from the displacement: its gradient at the GPU
    velocity and the two velocities."""
    from tlfea_amd import capi
    m, b, s, prob = _setup("C1", capi.SOLVER_NESTEROV, rest=True)
    if mf:  # the mass term inside the force kernel
        s.set_matrix_free_mass(RHO_STEEL)
    dt, rho = 1e-3, 1e10
    p = capi.NesterovParams(alpha=1e-8, rho=rho, inner_tol=0.0, outer_tol=0.0, max_outer=1, max_inner=200000,
                            time_step=dt)
    s.set_parameters(p)
    s.set_rel_tol(1e-8)
    s.set_restart(True)
    s.set_poll_interval(256)
    s.solve()
    st = s.stats()
    print(f"[nesterov C1] {st['inner']} inner iterations, {s.restarts()} restarts, rel {st['rel']:.3e}", flush=True)
    assert st["rel"] <= 1e-8 and st["inner"] < 200000 and s.restarts() > 0
    v = s.v_guess().cpu().numpy()
    v_ref, grad = _step_newton_extended(m, prob, dt, rho)
    free = np.ones(3 * m.n_nodes, bool)
    free.reshape(-1, 3)[m.fixed] = False
    g0 = np.linalg.norm(grad(np.zeros_like(v))[free])
    rel_ext = np.linalg.norm(grad(v)[free]) / g0
    X0 = m.xyz()
    z = np.zeros(3 * m.n_nodes)
    g_abs = prob.step_gradient(X0[:, 0], X0[:, 1], X0[:, 2], v_ref, z, np.zeros(3 * len(m.fixed)), dt, rho)
    floor = np.linalg.norm(g_abs[free]) / g0
    err = rel_err(v, v_ref)
    print(f"[nesterov C1] longdouble rel. gradient at the GPU velocity {rel_ext:.3e}; absolute-position "
          f"gradient at v* {floor:.3e}; |v - v*| / |v*| {err:.3e}", flush=True)
    assert rel_ext < 2e-8
    assert floor > 3e-8  # the floor the displacement form removes
    assert err < 1e-6
