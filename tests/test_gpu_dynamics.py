"""GPU checks of the batched implicit time steps (tlfea_amd/dynamics.py).

Bars: the matrix-free mass term equals the oracle's consistent mass times p to 1e-12, for
the quadrature rule of the reference's assembly (mass_matrix_qp_kernel, 5-point rule) and
for the exactly integrated mass; five steps of a small beam match a dense FP64 solve of the
same implicit Euler equations, built from the oracle's K and M, to 1e-9 (both rules, and a
soft beam whose K + M/dt^2 is indefinite under the 5-point mass with the exact one); a
batch of beams stepped together matches each beam stepped alone (the per-instance plain
path) to 1e-7.
"""
import numpy as np
import pytest

import oracle as O
from helpers import lame, rel_err

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def _beam(nx, ny, nz, load, E=2e9, nu=0.3, rho=2000.0, tet10=True):
    from tlfea_amd import mesh as M
    m = M.clamp_and_tip_load(M.kuhn_beam(nx, ny, nz, 1.0, 0.1, 0.1, tet10=tet10), np.asarray(load, float))
    return m, dict(E=E, nu=nu, rho=rho)


def _dense_KM(m, mat, rule="quadrature"):
    go, djo = O.grad(m)
    lam, mu = lame(mat["E"], mat["nu"])
    doff, dcol = O.dof_pattern(m)
    kv = O.stiffness_values(m, go, djo, lam, mu, doff, dcol)
    n = 3 * m.n_nodes
    K = np.zeros((n, n))
    for r in range(n):
        K[r, dcol[doff[r]:doff[r + 1]]] = kv[doff[r]:doff[r + 1]]
    off, col = O.node_pattern(m)
    mv = (O.mass_values if rule == "quadrature" else O.mass_values_exact)(m, djo, mat["rho"], off, col)
    M = np.zeros((n, n))
    for i in range(m.n_nodes):
        for k in range(off[i], off[i + 1]):
            for d in range(3):
                M[3 * i + d, 3 * col[k] + d] = mv[k]
    return K, M


@pytest.mark.parametrize("rule", ["quadrature", "exact"])
@pytest.mark.parametrize("tet10", [True, False])
def test_mass_term_matches_oracle(gpu, tet10, rule):
    import torch
    from tlfea_amd.batch import Batch
    m, mat = _beam(4, 2, 1, [0, 0, -10.0], tet10=tet10)
    b = Batch([(m, mat)])
    c = 7.0
    b.set_mass(np.array([mat["rho"] * c]), rule=rule)
    p = np.random.default_rng(1).standard_normal(3 * m.n_nodes)
    free = _np(b.mask) == 0
    y = _np(b.apply(torch.as_tensor(p).cuda(), k_scale=0.0, mass=True))
    K, M = _dense_KM(m, mat, rule)
    ref = c * (M @ (p * free))
    assert rel_err(y[free], ref[free]) < 1e-12
    # K + c M in one pass equals the sum of the parts
    y2 = _np(b.apply(torch.as_tensor(p).cuda(), k_scale=1.0, mass=True))
    ref2 = (K + c * M) @ (p * free)
    assert rel_err(y2[free], ref2[free]) < 1e-12


@pytest.mark.parametrize("rule,E,rho", [("quadrature", 2e9, 2000.0), ("exact", 2e9, 2000.0),
                                        ("exact", 1.5e6, 6500.0)])
def test_steps_match_dense_implicit_euler(gpu, rule, E, rho):
    from tlfea_amd.dynamics import BatchDynamics
    m, mat = _beam(4, 1, 1, [0.3, -0.5, -1.0], E=E, rho=rho)
    dt = 1e-4
    dyn = BatchDynamics([(m, mat)], amp=[200.0], freq=[15.0], dt=dt, tol=1e-13, mass_rule=rule)
    K, M = _dense_KM(m, mat, rule)
    free = _np(dyn.b.mask) == 0
    A = (K + M / dt ** 2)[np.ix_(free, free)]
    ev = np.linalg.eigvalsh(A)
    if E < 1e7:  # the soft beam: definite only with the exact mass
        Mq = _dense_KM(m, mat, "quadrature")[1]
        assert ev.min() > 0 and np.linalg.eigvalsh((K + Mq / dt ** 2)[np.ix_(free, free)]).min() < 0
    tip = _np(dyn.tip)
    g = np.tile([0.0, 0.0, -9.81], m.n_nodes)
    fg = (M @ g)
    u = np.zeros(3 * m.n_nodes)
    v = np.zeros(3 * m.n_nodes)
    for k in range(5):
        t1 = (k + 1) * dt
        f = fg + tip * 200.0 * np.sin(2 * np.pi * 15.0 * t1)
        rhs = M @ v / dt ** 2 - K @ u / dt + f / dt
        v = np.zeros_like(v)
        v[free] = np.linalg.solve(A, rhs[free])
        u = u + dt * v
        ug, vg = dyn.step()
        assert rel_err(_np(vg)[free], v[free]) < 1e-9, k
        assert rel_err(_np(ug)[free], u[free]) < 1e-9, k


def test_batch_steps_match_single_beams(gpu):
    from tlfea_amd.dynamics import BatchDynamics, dynamic_beams
    items, amp, freq = dynamic_beams(4, 31)
    # smaller beams than the configuration's, same sampler otherwise
    from tlfea_amd import mesh as M
    small = []
    rng = np.random.default_rng(5)
    for m, mat in items:
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        mm = M.clamp_and_tip_load(M.kuhn_beam(int(rng.integers(4, 12)), 2, 2, 1.0, 0.1, 0.1), d)
        small.append((mm, mat))
    dyn = BatchDynamics(small, amp, freq, tol=1e-12)
    singles = [BatchDynamics([it], [a], [f], tol=1e-12) for it, a, f in zip(small, amp, freq)]
    for k in range(20):
        u, v = dyn.step()
        for i, s in enumerate(singles):
            us, vs = s.step()
            sl = dyn.b.inst_slice(i)
            assert rel_err(_np(u[sl]), _np(us)) < 1e-7, (k, i)
            assert rel_err(_np(v[sl]), _np(vs)) < 1e-7, (k, i)
    its = np.array(dyn.iters)
    assert its.max() < 2000 and its.min() >= 1


def _newton_steps_cpu(m, mat, amp, freq, dt, nsteps, tol=1e-13, g=(0.0, 0.0, -9.81)):
    """Implicit Euler g(v) = M (v - v_n)/dt + f_int(X + u_n + dt v) - f(t1) = 0 solved densely by
    Newton (SyncedNesterov.cu:32-95's equation; exact consistent mass; f_int from the
    longdouble restatement, the tangent from the oracle's assembly at x)."""
    from helpers import fint_extended
    lam, mu = lame(mat["E"], mat["nu"])
    K0, M = _dense_KM(m, mat, "exact")
    n3 = 3 * m.n_nodes
    free = np.ones(n3, bool)
    for nd in np.asarray(m.fixed):
        free[3 * nd:3 * nd + 3] = False
    F = np.asarray(m.f_ext).reshape(-1, 3).sum(0)
    tip = np.asarray(m.f_ext) / np.linalg.norm(F)
    fg = M @ np.tile(np.asarray(g), m.n_nodes)
    fg[~free] = 0.0
    X = m.xyz()
    go, djo = O.grad(m)
    doff, dcol = O.dof_pattern(m)
    u = np.zeros(n3)
    v = np.zeros(n3)
    traj = []
    for k in range(nsteps):
        t1 = (k + 1) * dt
        f1 = fg + tip * amp * np.sin(2 * np.pi * freq * t1)
        vn = v.copy()
        w = vn.copy()

        def resid(w):
            x = X + (u + dt * w).reshape(-1, 3)
            r = M @ (w - vn) / dt + fint_extended(m, x[:, 0], x[:, 1], x[:, 2], lam, mu).astype(np.float64) - f1
            r[~free] = 0.0
            return r, x

        r, x = resid(w)
        r0 = np.linalg.norm(r)
        for _ in range(20):
            if np.linalg.norm(r) <= tol * r0:
                break
            kv = O.stiffness_values(m, go, djo, lam, mu, doff, dcol, x[:, 0].copy(), x[:, 1].copy(), x[:, 2].copy())
            Kt = np.zeros((n3, n3))
            for row in range(n3):
                Kt[row, dcol[doff[row]:doff[row + 1]]] = kv[doff[row]:doff[row + 1]]
            A = (M / dt + dt * Kt)[np.ix_(free, free)]
            dw = np.zeros(n3)
            dw[free] = np.linalg.solve(A, -r[free])
            w = w + dw
            r, x = resid(w)
        v = w
        u = u + dt * v
        traj.append(u.copy())
    return traj


def test_newton_steps_match_dense_nonlinear_euler(gpu):
    """Batched Newton steps (batched f_int at x, K(x) through tlfea_batch_set_displacement) against
    a dense Newton solve of the same nonlinear implicit-Euler equations: 1e-8 over 8 steps;
    the linearised steps' distance from that trajectory is the linearisation error."""
    from tlfea_amd.dynamics import BatchDynamics
    m, mat = _beam(5, 1, 1, [0.3, -0.5, -1.0], E=2e7, rho=1500.0)
    dt, amp, freq, nsteps = 1e-4, 400.0, 12.0, 8
    ref = _newton_steps_cpu(m, mat, amp, freq, dt, nsteps)
    dyn = BatchDynamics([(m, mat)], amp=[amp], freq=[freq], dt=dt, nonlinear=True, tol_nl=1e-10,
                        newton_cg_tol=1e-13)
    lin = BatchDynamics([(m, mat)], amp=[amp], freq=[freq], dt=dt, tol=1e-13)
    free = _np(dyn.b.mask) == 0
    worst_lin = 0.0
    for k in range(nsteps):
        u, _ = dyn.step()
        ul, _ = lin.step()
        assert rel_err(_np(u)[free], ref[k][free]) < 1e-8, k
        worst_lin = max(worst_lin, rel_err(_np(ul)[free], ref[k][free]))
    print(f"[newton] iterations per step {dyn.newton}, at the FP64 floor {dyn.floor}; linearised steps off by "
          f"{worst_lin:.2e}", flush=True)
    assert max(dyn.newton) <= 4 and worst_lin < 1e-2


def test_batch_internal_force_and_moved_operator(gpu):
    """tlfea_batch_internal_force equals each instance's own f_int (oracle) at deformed
    positions; with the positions set, the batched operator equals the oracle's tangent
    assembled there."""
    import torch
    from helpers import deformed
    from tlfea_amd.batch import Batch
    items = [_beam(4, 2, 1, [0, 0, -10.0], E=E, nu=nu) for E, nu in ((2e9, 0.3), (5e7, 0.42))]
    b = Batch(items)
    xs, fo, yo = [], [], []
    p = np.random.default_rng(9).standard_normal(3 * b.n_nodes)
    for i, (m, mat) in enumerate(items):
        xd = deformed(m, seed=20 + i, amp=5e-3)
        xs.append(np.stack(xd, 1))
        lam, mu = lame(mat["E"], mat["nu"])
        go, djo = O.grad(m)
        fo.append(O.internal_force(m, go, djo, lam, mu, *xd))
        doff, dcol = O.dof_pattern(m)
        vo = O.stiffness_values(m, go, djo, lam, mu, doff, dcol, *xd)
        vo, _ = O.apply_dirichlet(doff, dcol, vo, m.f_ext, m.fixed)
        yo.append(O.spmv(doff, dcol, vo, p[b.inst_slice(i)]))
    uall = (np.concatenate(xs) - np.concatenate([m.xyz() for m, _ in items])).reshape(-1)
    b.set_displacement(torch.as_tensor(uall).cuda())
    f = _np(b.internal_force())
    y = _np(b.tangent_apply(torch.as_tensor(p).cuda()))
    for i in range(len(items)):
        sl = b.inst_slice(i)
        assert rel_err(f[sl], fo[i]) < 1e-12, i
        assert rel_err(y[sl], yo[i]) < 1e-12, i
    b.set_displacement(None)
    assert np.abs(_np(b.internal_force())).max() < 1e-6 * np.abs(f).max()


@pytest.mark.parametrize("rule", ["quadrature", "exact"])
def test_fp32_operator_and_diagonal_with_mass(gpu, rule):
    """With a mass term the FP32 operator of the preconditioner is K + c M (against the
    oracle's dense matrices, FP32 bar 1e-5), and the preconditioner's diagonal includes c M:
    its lambda_max(D^-1 A) estimate equals the dense one to 1%, where a diagonal without the
    mass term would be off by more than 10%."""
    import torch
    from tlfea_amd.batch import Batch
    m, mat = _beam(4, 2, 1, [0, 0, -10.0], E=2e7, rho=1500.0)
    b = Batch([(m, mat)])
    c = 1e8
    b.set_mass(np.array([mat["rho"] * c]), rule=rule)
    K, M = _dense_KM(m, mat, rule)
    free = _np(b.mask) == 0
    A = (K + c * M)[np.ix_(free, free)]
    p = np.random.default_rng(3).standard_normal(3 * m.n_nodes)
    y = _np(b.tangent_apply_f32(torch.as_tensor(p, dtype=torch.float32).cuda()))
    assert rel_err(y[free], A @ p[free]) < 1e-5
    lm = b.precond_setup(degree=2, ratio=30.0, power_iters=400, safety=1.0)[0]

    def lmax(d):
        s = 1.0 / np.sqrt(d)
        return np.linalg.eigvalsh(A * s[:, None] * s[None, :]).max()
    ref = lmax(np.diag(A))
    wrong = lmax(np.diag(K[np.ix_(free, free)]))
    assert abs(wrong / ref - 1) > 0.1
    assert abs(lm / ref - 1) < 1e-2, (lm, ref, wrong)


@pytest.mark.parametrize("nonlinear", [False, True])
def test_preconditioned_steps_match_plain(gpu, nonlinear):
    """Time steps whose solves run the flexible FP64 CG with the FP32 two-level preconditioner
    of K + M/dt^2 (set up once at rest) land on the plain FP64 CG's trajectory (1e-7 over 12
    steps: both solve to 1e-10) in a third of its iterations or fewer."""
    from tlfea_amd import mesh as M
    from tlfea_amd.dynamics import BatchDynamics, dynamic_beams
    items, amp, freq = dynamic_beams(4, 41)
    rng = np.random.default_rng(8)
    small = []
    for m, mat in items:
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        small.append((M.clamp_and_tip_load(M.kuhn_beam(int(rng.integers(8, 20)), 2, 2, 1.0, 0.1, 0.1), d), mat))
    kw = dict(tol=1e-10, tol_nl=1e-10, newton_cg_tol=1e-12, nonlinear=nonlinear)
    plain = BatchDynamics(small, amp, freq, **kw)
    pre = BatchDynamics(small, amp, freq, cheb_degree=4, coarse=True, **kw)
    for k in range(12):
        u0, v0 = plain.step()
        u1, v1 = pre.step()
        for i in range(len(small)):
            sl = plain.b.inst_slice(i)
            assert rel_err(_np(u1[sl]), _np(u0[sl])) < 1e-7, (k, i)
            assert rel_err(_np(v1[sl]), _np(v0[sl])) < 1e-7, (k, i)
    ip, iq = np.array(plain.iters), np.array(pre.iters)
    print(f"[precond dynamics nonlinear={nonlinear}] CG iterations per instance-step: plain median "
          f"{np.median(ip):.0f}, preconditioned median {np.median(iq):.0f}", flush=True)
    assert np.median(iq) * 3 <= np.median(ip)


def _step_problem(m, mat, dt, amp=300.0, g=(0.0, 0.0, -9.81)):
    """One implicit step from rest: exact mass M, the load M g + amp * tip pattern, and the
    free-DOF mask."""
    K, M = _dense_KM(m, mat, "exact")
    n3 = 3 * m.n_nodes
    free = np.ones(n3, bool)
    for nd in np.asarray(m.fixed):
        free[3 * nd:3 * nd + 3] = False
    F = np.asarray(m.f_ext).reshape(-1, 3).sum(0)
    f = M @ np.tile(np.asarray(g), m.n_nodes) + np.asarray(m.f_ext) / np.linalg.norm(F) * amp
    f[~free] = 0.0
    return K, M, free, f


@pytest.mark.parametrize("restart", [False, True])
def test_batch_nesterov_iterates_match_restatement(gpu, restart):
    """The batched SyncedNesterov inner loop (tlfea_batch_nesterov_solve) against a numpy
    restatement of SyncedNesterov.cu:97-360 with the fixed DOFs eliminated, driven by the
    oracle's internal force (longdouble) and dense exact mass: the iterate after 3 and after
    40 iterations agrees to 1e-10, and so does the restart count."""
    import torch
    from helpers import fint_extended
    from tlfea_amd.batch import Batch
    m, mat = _beam(4, 1, 1, [0.3, -0.5, -1.0], E=2e7, rho=1500.0)
    dt = 1e-4
    lam, mu = lame(mat["E"], mat["nu"])
    K, M, free, f = _step_problem(m, mat, dt)
    lmax = np.linalg.eigvalsh((K + M / dt ** 2)[np.ix_(free, free)]).max()
    alpha = 1.0 / (1.1 * dt * lmax)
    X = m.xyz()
    n3 = 3 * m.n_nodes
    b = Batch([(m, mat)])
    zero = torch.zeros(n3, dtype=torch.float64, device="cuda")
    fd = torch.as_tensor(f).cuda()
    for K_it in (3, 40):
        # restatement
        vn = np.zeros(n3)
        vk, vkm1 = vn.copy(), vn.copy()
        t, t1 = 1.0, 0.5 * (1 + np.sqrt(5.0))
        rs = 0
        for _ in range(K_it):
            beta = (t - 1.0) / t1
            y = vk + beta * (vk - vkm1)
            y[~free] = 0.0
            vkm1 = vk
            u = (dt * y).reshape(-1, 3)
            fi = fint_extended(m, X[:, 0] + u[:, 0], X[:, 1] + u[:, 1], X[:, 2] + u[:, 2], lam, mu,
                               u=u).astype(np.float64)
            gr = M @ (y - vn) / dt + fi - f
            gr[~free] = 0.0
            vnext = y - alpha * gr
            vnext[~free] = 0.0
            if restart and gr @ (vnext - vkm1) > 0:
                t = 1.0
                rs += 1
            else:
                t = t1
            t1 = 0.5 * (1 + np.sqrt(1 + 4 * t * t))
            vk = vnext
        v = b.nesterov_solve(dt, mat["rho"], alpha, zero, zero, fd, rel_tol=1e-300, max_inner=K_it,
                             restart=restart, graph=False)
        assert int(b.iters[0]) == K_it
        assert rel_err(_np(v), vk) < 1e-10, K_it
        assert int(b.restarts[0]) == rs, (K_it, int(b.restarts[0]), rs)
    b.close()


def test_batch_nesterov_step_against_dense_newton(gpu):
    """Three beams of different sizes and materials in one batch, one implicit step each by
    the batched Nesterov solve to ||g|| <= 1e-11 ||g(v_n)||: every instance's velocity equals
    the dense Newton solution of the same nonlinear step (longdouble f_int) to 1e-7, and the
    same instance solved alone (a batch of one) to 1e-8; instances stop on their own."""
    import torch
    from tlfea_amd.batch import Batch
    specs = [(_beam(4, 1, 1, [0.3, -0.5, -1.0], E=2e7, rho=1500.0), 300.0),
             (_beam(6, 2, 1, [0.0, 1.0, -1.0], E=5e8, rho=4000.0), 900.0),
             (_beam(3, 1, 2, [1.0, 0.0, -0.2], E=3e6, nu=0.4, rho=1100.0), 40.0)]
    dt = 1e-4
    items = [s for s, _ in specs]
    b = Batch(items)
    rho = np.array([mat["rho"] for _, mat in items])
    b.set_mass(rho / dt ** 2, rule="exact")
    alpha = 1.0 / (1.1 * dt * b.lambda_max())
    fs, refs = [], []
    for (m, mat), amp in specs:
        K, M, free, f = _step_problem(m, mat, dt, amp=amp)
        fs.append(f)
        # the dense Newton reference of one step with the same load
        # (a frequency of 1/(4 dt) puts the sine at 1 at t = dt)
        refs.append(_newton_steps_cpu(m, mat, amp, 0.25 / dt, dt, 1, tol=1e-14)[0] / dt)
    f = torch.as_tensor(np.concatenate(fs)).cuda()
    zero = torch.zeros(3 * b.n_nodes, dtype=torch.float64, device="cuda")
    v = _np(b.nesterov_solve(dt, rho, alpha, zero, zero, f, rel_tol=1e-11, restart=True))
    assert b.rc == 0 and (b.rel <= 1e-11).all()
    its = b.iters.copy()
    print(f"[batched nesterov] iterations per instance {its.tolist()}, restarts {b.restarts.tolist()}", flush=True)
    assert len(set(its.tolist())) > 1
    for i, ((m, mat), amp) in enumerate(specs):
        sl = b.inst_slice(i)
        free = _np(b.mask[sl]) == 0
        assert rel_err(v[sl][free], refs[i][free]) < 1e-7, i
        one = Batch([(m, mat)])
        zi = torch.zeros(3 * m.n_nodes, dtype=torch.float64, device="cuda")
        vi = _np(one.nesterov_solve(dt, mat["rho"], alpha[i], zi, zi, torch.as_tensor(fs[i]).cuda(), rel_tol=1e-11))
        assert rel_err(v[sl], vi) < 1e-8, i
        one.close()
    b.close()


def test_nesterov_steps_match_newton_steps(gpu):
    """Time steps of the dynamic batch by batched Nesterov solves land on the Newton
    trajectory of the same nonlinear implicit-Euler equations (1e-7 over 8 steps, both to
    1e-11)."""
    from tlfea_amd import mesh as M
    from tlfea_amd.dynamics import BatchDynamics, dynamic_beams
    items, amp, freq = dynamic_beams(3, 77)
    rng = np.random.default_rng(2)
    small = []
    for m, mat in items:
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        small.append((M.clamp_and_tip_load(M.kuhn_beam(int(rng.integers(4, 10)), 2, 1, 1.0, 0.1, 0.1), d), mat))
    kw = dict(tol_nl=1e-11, newton_cg_tol=1e-13)
    newton = BatchDynamics(small, amp, freq, nonlinear=True, **kw)
    nes = BatchDynamics(small, amp, freq, nesterov=True, **kw)
    for k in range(8):
        u0, v0 = newton.step()
        u1, v1 = nes.step()
        for i in range(len(small)):
            sl = newton.b.inst_slice(i)
            assert rel_err(_np(u1[sl]), _np(u0[sl])) < 1e-7, (k, i)
            assert rel_err(_np(v1[sl]), _np(v0[sl])) < 1e-7, (k, i)
    its = np.array(nes.iters)
    print(f"[nesterov dynamics] inner iterations per instance-step: median {np.median(its):.0f}, max {its.max()}",
          flush=True)


def test_batch_nesterov_step_halving_safeguard(gpu):
    """A step size 8x beyond 1/lambda_max diverges; the batched loop halves it per instance
    and starts again from v_n until it converges, to the same velocity as a run with a
    stable step size (1e-8), while the instance with a stable step is unaffected."""
    import torch
    from tlfea_amd.batch import Batch
    specs = [_beam(4, 1, 1, [0.3, -0.5, -1.0], E=2e7, rho=1500.0), _beam(5, 1, 1, [0.0, 1.0, -1.0], E=8e7, rho=2500.0)]
    dt = 1e-4
    b = Batch(specs)
    rho = np.array([mat["rho"] for _, mat in specs])
    b.set_mass(rho / dt ** 2, rule="exact")
    good = 1.0 / (1.1 * dt * b.lambda_max())
    zero = torch.zeros(3 * b.n_nodes, dtype=torch.float64, device="cuda")
    f = b.f * 300.0
    ref = _np(b.nesterov_solve(dt, rho, good, zero, zero, f, rel_tol=1e-11))
    assert b.rc == 0 and (b.halvings == 0).all()
    bad = good * np.array([8.0, 1.0])
    v = _np(b.nesterov_solve(dt, rho, bad, zero, zero, f, rel_tol=1e-11))
    assert b.rc == 0, b.status
    assert b.halvings[0] >= 3 and b.halvings[1] == 0, b.halvings
    for i in range(2):
        sl = b.inst_slice(i)
        assert rel_err(v[sl], ref[sl]) < 1e-8, i
