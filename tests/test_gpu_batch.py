"""GPU parity of the batched linear-static path against single-mesh solves and the oracle.

Bars: the batched operator equals each instance's own matrix-free tangent to 1e-13 and the
oracle's assembled K p to 1e-12.  CG stops on its recursive residual (the reference's
rule); the true residual ||f - K u|| / ||f|| of the result can only be trusted down to the
FP64 floor eps * ||K|| ||u|| / ||f||, which on these slender beams (kappa 1e6-8e8) is
1e-10..2e-8 -- even the oracle's direct solve sits there.  So the true residual is checked
against tol + 20 * eps * ||K||_1 ||u|| / ||f|| per instance, and displacements against the
banded Cholesky solve to 1e-6.
"""
import numpy as np
import pytest

import oracle as O
from helpers import lame, rel_err

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def _items(seed=0, n=5, big=False):
    """Small uneven cantilevers with random materials and tip loads."""
    from tlfea_amd import mesh as M
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        nx = int(rng.integers(3, 12 if not big else 40))
        ny = int(rng.integers(1, 4))
        nz = int(rng.integers(1, 4))
        L = float(rng.uniform(0.5, 2.0))
        a = float(rng.uniform(0.05, 0.2))
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        m = M.clamp_and_tip_load(M.kuhn_beam(nx, ny, nz, L, a, a), d * float(10 ** rng.uniform(0, 4)))
        out.append((m, dict(E=float(10 ** rng.uniform(6, 10)), nu=float(rng.uniform(0.2, 0.45)))))
    return out


def _oracle_system(m, E, nu):
    go, djo = O.grad(m)
    lam, mu = lame(E, nu)
    doff, dcol = O.dof_pattern(m)
    vo = O.stiffness_values(m, go, djo, lam, mu, doff, dcol)
    vo, fo = O.apply_dirichlet(doff, dcol, vo, m.f_ext, m.fixed)
    return doff, dcol, vo, fo


def _k_norm1(doff, vo):
    """max row sum of |K| (= ||K||_1 for the symmetric K), an upper bound of ||K||_2."""
    return max(np.abs(vo[doff[i]:doff[i + 1]]).sum() for i in range(len(doff) - 1))


def _residual_bar(tol, doff, vo, u, f):
    return tol + 20 * 2.2e-16 * _k_norm1(doff, vo) * np.linalg.norm(u) / np.linalg.norm(f)


def test_fp32_operator_at_rest(gpu):
    """The FP32 operator of the two-level stage at rest (the corner-value form on these
    affine TET10 beams) equals the FP64 operator to FP32 rounding (5e-6 relative per
    instance) and the oracle's assembled K p to the same bar, returns p on the masked rows,
    and two applications give the same bits (each block sums its element outputs in the fixed
    order of the plan's gather tables)."""
    import torch
    from tlfea_amd.batch import Batch
    items = _items(seed=4, n=6, big=True)
    b = Batch(items)
    p = np.random.default_rng(2).standard_normal(3 * b.n_nodes)
    p32 = torch.as_tensor(p, dtype=torch.float32).cuda()
    y32 = _np(b.tangent_apply_f32(p32))
    y32b = _np(b.tangent_apply_f32(p32))
    assert np.array_equal(y32b, y32)
    pr = _np(p32).astype(np.float64)  # the rounded input both sides see
    y64 = _np(b.tangent_apply(torch.as_tensor(pr).cuda()))
    mask = _np(b.mask) != 0
    assert np.array_equal(y32[mask], _np(p32)[mask])
    for i, (m, mat) in enumerate(items):
        sl = b.inst_slice(i)
        assert rel_err(y32[sl], y64[sl]) < 5e-6, i
        doff, dcol, vo, _ = _oracle_system(m, mat["E"], mat["nu"])
        assert rel_err(y32[sl], O.spmv(doff, dcol, vo, pr[sl])) < 5e-6, i


def test_batch_operator_matches_single_and_oracle(gpu):
    import torch
    from tlfea_amd.batch import Batch
    from tlfea_amd.fea import Body
    items = _items(1, 4)
    b = Batch(items)
    rng = np.random.default_rng(2)
    p = rng.standard_normal(3 * b.n_nodes)
    y = _np(b.tangent_apply(torch.as_tensor(p).cuda()))
    for i, (m, mat) in enumerate(items):
        sl = b.inst_slice(i)
        body = Body(m, mat["E"], mat["nu"], 1000.0)
        yi = _np(body.tangent_apply(torch.as_tensor(p[sl]).cuda()))
        assert rel_err(y[sl], yi) < 1e-13
        doff, dcol, vo, _ = _oracle_system(m, mat["E"], mat["nu"])
        assert rel_err(y[sl], O.spmv(doff, dcol, vo, p[sl])) < 1e-12


def test_batch_cg_against_direct(gpu):
    from tlfea_amd.batch import Batch
    items = _items(3, 5)
    b = Batch(items)
    u = _np(b.solve(tol=1e-12, max_iter=100000))
    assert b.rc == 0 and np.all(b.status == 0), (b.status, b.rel)
    assert np.all(b.rel < 1e-12)
    for i, (m, mat) in enumerate(items):
        doff, dcol, vo, fo = _oracle_system(m, mat["E"], mat["nu"])
        ud = O.band_cholesky_solve(doff, dcol, vo, fo)
        ui = u[b.inst_slice(i)]
        r = fo - O.spmv(doff, dcol, vo, ui)
        assert np.linalg.norm(r) / np.linalg.norm(fo) < _residual_bar(1e-12, doff, vo, ui, fo), i
        assert rel_err(ui, ud) < 1e-6, i
        # same Krylov sequence as the oracle's CG: the first 10 iterates agree
        if i == 0:
            uk, _, _ = O.cg(doff, dcol, vo, fo, tol=0.0, max_iter=10)
            b10 = Batch([items[0]])
            u10 = _np(b10.solve(tol=0.0, max_iter=10, graph=False))
            assert b10.iters[0] == 10 and rel_err(u10, uk) < 1e-11


def test_batch_true_residual_and_sequential(gpu):
    """tol 1e-8: the reported residual is the true one (recomputed with the FP64 operator)
    and every instance agrees with its own single-mesh matrix-free solve."""
    import torch
    from tlfea_amd.batch import Batch
    from tlfea_amd.fea import Body, LinearStatic
    items = _items(4, 6, big=True)
    b = Batch(items)
    u = b.solve(tol=1e-8)
    assert b.rc == 0, (b.status, b.rel)
    r = b.f - b.tangent_apply(u)
    r[b.mask.bool()] = 0.0
    f = b.f.clone()
    f[b.mask.bool()] = 0.0
    for i, (m, mat) in enumerate(items):
        sl = b.inst_slice(i)
        rel = (torch.linalg.norm(r[sl]) / torch.linalg.norm(f[sl])).item()
        doff, dcol, vo, fo = _oracle_system(m, mat["E"], mat["nu"])
        assert rel < _residual_bar(1e-8, doff, vo, _np(u[sl]), fo), (i, rel)
        body = Body(m, mat["E"], mat["nu"], 1000.0)
        ls = LinearStatic(body, tol=1e-8, max_iter=200000, matfree=True)
        ls.assemble()
        assert ls.cg_solve() == 0
        # two 1e-8 solutions of the same system differ by at most ~kappa * 1e-8
        assert rel_err(_np(u[sl]), _np(ls.u)) < 1e-4, i
        assert abs(int(b.iters[i]) - ls.iters) <= max(3, 0.25 * ls.iters), (i, b.iters[i], ls.iters)


def test_batch_graph_matches_eager(gpu):
    """The captured iteration graph is replayed many times per solve, with host polls in
    between; three graph solves in a row agree with the eager-launch solve."""
    from tlfea_amd.batch import Batch
    items = _items(5, 5)
    b = Batch(items)
    ue = _np(b.solve(tol=1e-11, graph=False, poll=8))
    b.solve(tol=1e-9, graph=False, poll=8)
    it_e = b.iters.copy()
    for _ in range(3):
        ug = _np(b.solve(tol=1e-11, graph=True, poll=8))
        assert b.rc == 0
        assert rel_err(ug, ue) < 1e-8
        # finite-precision CG counts move with the summation order (the planned kernels'
        # LDS sums follow wave scheduling): on these beams (kappa up to ~1e8) two solves
        # of the same system differ by up to ~15% in count at 1e-11 and ~8% at 1e-9 --
        # the answers agree (above); the counts are a sanity check with a 20% band
        b.solve(tol=1e-9, graph=True, poll=8)
        assert np.all(np.abs(b.iters - it_e) <= np.maximum(3, 0.2 * it_e)), (b.iters, it_e)


@pytest.mark.parametrize("precond", [False, True])
def test_batch_graph_kept_across_solves(gpu, precond):
    """The iteration graph captured by one solve is replayed by the next ones while poll,
    tol and the set-up are unchanged: solves of f, 2 f (another output buffer) and f again
    from the first solution as guess give u, 2 u (CG is homogeneous in f, and doubling is
    exact) and u again, each equal to an eager solve; a new tol re-captures and still
    agrees."""
    import torch
    from tlfea_amd.batch import Batch
    items = _items(6, 4, big=precond)
    b = Batch(items)
    if precond:
        b.precond_setup(degree=6)
        b.coarse_setup()
    f = b.f.clone()
    ue = _np(b.solve(tol=1e-10, graph=False, poll=8, precond=precond))
    u1 = _np(b.solve(tol=1e-10, graph=True, poll=8, precond=precond))
    out = torch.empty_like(b.f)
    u2 = _np(b.solve(tol=1e-10, graph=True, poll=8, precond=precond, f=2 * f, out=out))
    u3 = _np(b.solve(tol=1e-10, graph=True, poll=8, precond=precond, u0=torch.as_tensor(u1).cuda()))
    assert b.rc == 0 and np.all(b.status == 0)
    assert rel_err(u1, ue) < 1e-8
    assert rel_err(u2, 2 * ue) < 1e-8
    assert rel_err(u3, ue) < 1e-8
    u4 = _np(b.solve(tol=1e-12, graph=True, poll=8, precond=precond))
    assert b.rc == 0 and rel_err(u4, ue) < 1e-8
    b.close()


def test_batch_guess_zero_load_and_limits(gpu):
    import torch
    from tlfea_amd import capi
    from tlfea_amd.batch import Batch
    items = _items(6, 4)
    items[2][0].f_ext = np.zeros(3 * items[2][0].n_nodes)  # an unloaded instance
    b = Batch(items)
    u = b.solve(tol=1e-10)
    assert b.rc == 0
    assert b.iters[2] == 0 and np.all(_np(u[b.inst_slice(2)]) == 0)
    # from the converged solution every instance stops at once (tolerance above the true
    # residual of the converged u)
    r = b.f - b.tangent_apply(u)
    r[b.mask.bool()] = 0.0
    f = b.f.clone()
    f[b.mask.bool()] = 0.0
    true = [(torch.linalg.norm(r[b.inst_slice(i)]) / max(torch.linalg.norm(f[b.inst_slice(i)]).item(), 1e-300)).item()
            for i in range(len(items)) if i != 2]
    b.solve(tol=2 * max(true), u0=u.clone())
    assert np.all(b.iters == 0), b.iters
    # an iteration limit below convergence reports the unfinished instances, with the
    # relative residual of the returned u (the last update's residual is tested after the
    # loop)
    u5 = b.solve(tol=1e-12, max_iter=5)
    assert b.rc == capi.TLFEA_ENOCONV
    assert b.status[2] == 0 and np.all(b.status[[0, 1, 3]] == capi.TLFEA_ENOCONV)
    assert np.all(b.iters[[0, 1, 3]] == 5)
    r5 = b.f - b.tangent_apply(u5)
    r5[b.mask.bool()] = 0.0
    for i in (0, 1, 3):
        true5 = (torch.linalg.norm(r5[b.inst_slice(i)]) / torch.linalg.norm(f[b.inst_slice(i)])).item()
        assert abs(b.rel[i] - true5) < 1e-8 * true5, (i, b.rel[i], true5)
    # a tolerance just above the residual after those 5 iterations, with the same limit:
    # converged (at the last allowed iteration at the latest).  Repeated solves agree to
    # rounding only (the tangent's in-block LDS sums follow wave scheduling), which over 5
    # iterations moves the residual by far less than the 1% margin.
    rel5 = b.rel.copy()
    tol5 = 1.01 * float(rel5[[0, 1, 3]].max())
    b.solve(tol=tol5, max_iter=5)
    assert b.rc == 0 and np.all(b.iters <= 5), (b.iters, b.rel, tol5)
    # the chunk size changes the reduction granularity, not the answer
    b.set_chunk(64)
    assert b.n_chunks() > len(items)
    u64 = b.solve(tol=1e-10)
    assert rel_err(_np(u64), _np(u)) < 1e-7


def test_batch_dot(gpu):
    """Per-instance dot products over each instance's DOFs equal numpy's, and repeat
    exactly (fixed summation order)."""
    import torch
    from tlfea_amd.batch import Batch
    items = _items(8, 5)
    b = Batch(items)
    rng = np.random.default_rng(9)
    a = rng.standard_normal(3 * b.n_nodes)
    c = rng.standard_normal(3 * b.n_nodes)
    at, ct = torch.as_tensor(a).cuda(), torch.as_tensor(c).cuda()
    d = _np(b.dot(at, ct))
    ref = np.array([a[3 * b.node_off[i]:3 * b.node_off[i + 1]] @ c[3 * b.node_off[i]:3 * b.node_off[i + 1]]
                    for i in range(len(items))])
    assert np.abs(d - ref).max() < 1e-12 * np.abs(ref).max()
    n2 = _np(b.dot(at))
    assert np.all(n2 > 0) and np.array_equal(n2, _np(b.dot(at, at)))


def test_batch_fp32_chebyshev_preconditioner(gpu):
    """Flexible FP64 CG with the FP32 Jacobi-Chebyshev stage: same solution as the plain
    batched CG (both at tol 1e-10), true FP64 residual within the floor bar, and several
    times fewer outer iterations (2.9-4.9x measured on these beams); compaction of finished instances changes nothing."""
    import torch
    from tlfea_amd.batch import Batch
    items = _items(7, 6, big=True)
    b = Batch(items)
    u0 = _np(b.solve(tol=1e-10))
    it0 = b.iters.copy()
    lm = b.precond_setup(degree=6, ratio=30.0, power_iters=20)
    assert np.all(lm > 0) and np.all(np.isfinite(lm))
    # lambda_max(D^-1 K) of a stiffness matrix lies in (1, 2 * max nodes per element]
    assert np.all(lm > 1.0) and np.all(lm < 20.0), lm
    u1 = _np(b.solve(tol=1e-10, precond=True))
    assert b.rc == 0, (b.status, b.rel)
    assert np.all(b.iters * 2 < it0), (b.iters, it0)
    r = b.f - b.tangent_apply(torch.as_tensor(u1).cuda())
    r[b.mask.bool()] = 0.0
    for i, (m, mat) in enumerate(items):
        sl = b.inst_slice(i)
        doff, dcol, vo, fo = _oracle_system(m, mat["E"], mat["nu"])
        rel = np.linalg.norm(_np(r[sl])) / np.linalg.norm(fo)
        assert rel < _residual_bar(1e-10, doff, vo, u1[sl], fo), (i, rel)
        assert rel_err(u1[sl], u0[sl]) < 1e-6, i
    # without compaction, and without the graph: same answers
    b.set_compaction(0.0)
    u2 = _np(b.solve(tol=1e-10, precond=True, graph=False))
    assert rel_err(u2, u1) < 1e-7


def test_batch_autotune(gpu):
    """The autotuner picks a configuration from its candidate sets, applies it, and a solve
    afterwards (default replay length = the chosen one) gives the solution of an untuned
    batch to rounding; the choice is deterministic in kind (a valid chunk and replay length)
    whatever the timings."""
    from tlfea_amd.batch import Batch
    items = _items(seed=7, n=6, big=True)
    ref = Batch(items)
    u_ref = _np(ref.solve(tol=1e-10))
    b = Batch(items)
    t = b.autotune(trial_iters=16, tol=1e-8, full_solves=True)
    print(f"[autotune] {t}", flush=True)
    assert t["chunk_nodes"] in (256, 512, 1024, 2048, 4096) and t["poll"] in (16, 32, 64)
    assert t["compaction"] in (0.35, 0.5, 0.65, 0.8, 0.9) and t["us_per_iter"] > 0 and t["solve_s"] > 0
    assert b.n_chunks() == sum(-(-m.n_nodes // t["chunk_nodes"]) for m, _ in items)
    u = _np(b.solve(tol=1e-10))
    assert (b.status == 0).all()
    for i in range(len(items)):
        sl = b.inst_slice(i)
        assert rel_err(u[sl], u_ref[sl]) < 1e-8, i


def _plane_basis(m, modes=6):
    """P of the coarse space in numpy: translations and rotations about the centroid of the
    free nodes of each lattice plane x = const (fixed nodes: zero rows); with modes = 11 also
    the in-plane fields of the plane's coordinates (s, t) = (y, z) - centroid: the bending
    warping (s t, (t^2 - s^2)/2) and ((s^2 - t^2)/2, s t), the stretch (s, t) and the shears
    (t, s) and (s, -t) in the (y, z) components."""
    X = m.xyz()
    key = np.unique(np.round(X[:, 0], 9), return_inverse=True)[1]
    free = np.ones(m.n_nodes, bool)
    free[np.asarray(m.fixed)] = False
    cols = []
    for p in range(key.max() + 1):
        nodes = np.nonzero((key == p) & free)[0]
        if len(nodes) == 0:
            continue
        r = X[nodes] - X[nodes].mean(0)
        for j in range(6):
            v = np.zeros((m.n_nodes, 3))
            if j < 3:
                v[nodes, j] = 1.0
            else:
                e = np.eye(3)[j - 3]
                v[nodes] = np.cross(e, r)
            cols.append(v.reshape(-1))
        if modes == 11:
            s_, t_ = r[:, 1], r[:, 2]
            for fy, fz in ((s_ * t_, 0.5 * (t_ * t_ - s_ * s_)), (0.5 * (s_ * s_ - t_ * t_), s_ * t_), (s_, t_),
                           (t_, s_), (s_, -t_)):
                v = np.zeros((m.n_nodes, 3))
                v[nodes, 1], v[nodes, 2] = fy, fz
                cols.append(v.reshape(-1))
    return np.stack(cols, 1)


@pytest.mark.parametrize("modes", [6, 11])
@pytest.mark.parametrize("tet10", [True, False])
def test_coarse_solve_equals_galerkin_projection(gpu, tet10, modes):
    """tlfea_batch_coarse_apply against P (P^T K P)^-1 P^T r built in numpy from the oracle's
    Dirichlet-eliminated K and the lattice planes (any basis of the same space gives the same
    product), with the rigid motions alone and with the five in-plane fields: for a random
    r, and for r = K u_c with u_c in the coarse space, where it returns u_c.  This pins the
    probing assembly of P^T K P, its block factorisation (TET10: two planes per block; TET4:
    one), the restriction and the prolongation of every mode."""
    import torch
    from tlfea_amd import mesh as M
    from tlfea_amd.batch import Batch
    rng = np.random.default_rng(11)
    items = []
    for k in range(4):
        nx, ny, nz = int(rng.integers(4, 15)), int(rng.integers(1, 4)), int(rng.integers(1, 4))
        m = M.clamp_and_tip_load(M.kuhn_beam(nx, ny, nz, float(rng.uniform(0.5, 2)), 0.1, 0.1, tet10=tet10),
                                 -100.0)
        items.append((m, dict(E=float(10 ** rng.uniform(6, 10)), nu=0.3)))
    b = Batch(items)
    b.set_coarse_modes(modes)
    b.precond_setup(degree=4)
    nc = b.coarse_setup()
    r = rng.standard_normal(3 * b.n_nodes)
    z = _np(b.coarse_apply(torch.as_tensor(r).cuda()))
    uc = np.concatenate([_plane_basis(m, modes) @ rng.standard_normal(_plane_basis(m, modes).shape[1])
                         for m, _ in items])
    zc = _np(b.coarse_apply(b.apply(torch.as_tensor(uc).cuda())))
    for i, (m, mat) in enumerate(items):
        sl = b.inst_slice(i)
        doff, dcol, vo, _ = _oracle_system(m, mat["E"], mat["nu"])
        import scipy.sparse as sp
        K = sp.csr_matrix((vo, dcol, doff))
        P = _plane_basis(m, modes)
        want = P @ np.linalg.lstsq(P.T @ (K @ P), P.T @ r[sl], rcond=None)[0]
        print(f"[coarse] tet10={tet10} modes={modes} inst {i}: coarse dim {P.shape[1]}, "
              f"err {rel_err(z[sl], want):.2e}, u_c err {rel_err(zc[sl], uc[sl]):.2e}", flush=True)
        assert rel_err(z[sl], want) < 1e-9
        assert rel_err(zc[sl], uc[sl]) < 1e-9
    assert nc >= sum(_plane_basis(m, modes).shape[1] for m, _ in items)
    b.close()


@pytest.mark.parametrize("seed", [3, 7])
def test_two_level_solve_against_direct_and_plain(gpu, seed):
    """The FP32 two-level preconditioned flexible CG on uneven beams, with the 6 rigid
    coarse modes per plane and with 11: every instance reaches the FP64 recursive residual
    1e-10, agrees with the oracle's banded Cholesky solve within the bar of the plain batched
    CG (1e-6, these beams' kappa allows no tighter), with the plain batched FP64 CG to 1e-6,
    and needs far fewer iterations than it (at least 5x on the batch's slowest instance)."""
    from tlfea_amd.batch import Batch
    items = _items(seed, 6, big=True)
    b = Batch(items)
    u64 = _np(b.solve(tol=1e-10))
    it64 = b.iters.copy()
    b.precond_setup(degree=6)
    b.coarse_setup()
    u2 = _np(b.solve(tol=1e-10, precond=True))
    it2 = b.iters.copy()
    assert np.all(b.status == 0)
    b.set_coarse_modes(11)
    b.coarse_setup()
    u11 = _np(b.solve(tol=1e-10, precond=True))
    it11 = b.iters.copy()
    print(f"[two-level] plain {list(it64)} -> two-level {list(it2)}, 11 modes {list(it11)}",
          flush=True)
    assert np.all(b.status == 0)
    assert it2.max() * 5 <= it64.max() and it11.max() * 5 <= it64.max()
    for i, (m, mat) in enumerate(items):
        sl = b.inst_slice(i)
        doff, dcol, vo, fo = _oracle_system(m, mat["E"], mat["nu"])
        ud = O.band_cholesky_solve(doff, dcol, vo, fo)
        assert rel_err(u2[sl], ud) < 1e-6 and rel_err(u11[sl], ud) < 1e-6
        assert rel_err(u2[sl], u64[sl]) < 1e-6
        assert np.linalg.norm(fo - O.spmv(doff, dcol, vo, u2[sl])) / np.linalg.norm(fo) < _residual_bar(
            1e-10, doff, vo, u2[sl], fo)
    b.close()


def _jittered_beam(nx, ny, nz, tet10, amp, seed):
    """A cantilever (1 x 0.1 x 0.1) whose free vertices are moved by up to amp of the cell
    size along each axis (edge nodes kept at their edge midpoints, so the elements stay
    affine): its node coordinates along the beam are no lattice."""
    from tlfea_amd import mesh as M
    m = M.clamp_and_tip_load(M.kuhn_beam(nx, ny, nz, tet10=tet10), -1.0e3)
    rng = np.random.default_rng(seed)
    X = m.xyz().copy()
    verts = np.unique(m.elems[:, :4])
    move = np.setdiff1d(verts, m.fixed)
    h = np.array([1.0 / nx, 0.1 / ny, 0.1 / nz])
    X[move] += amp * h * rng.uniform(-1.0, 1.0, (len(move), 3))
    if tet10:
        for k, (a, b) in enumerate([(0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3)]):
            X[m.elems[:, 4 + k]] = 0.5 * (X[m.elems[:, a]] + X[m.elems[:, b]])
    jm = M.TetMesh(X[:, 0].copy(), X[:, 1].copy(), X[:, 2].copy(), m.elems)
    return M.clamp_and_tip_load(jm, -1.0e3)


@pytest.mark.parametrize("tet10", [True, False])
def test_two_level_on_unstructured_meshes(gpu, tet10):
    """Meshes whose nodes lie on no lattice planes: the reference's own beam_highres mesh
    (tests/golden/meshes, TET10 or TET4) and a jittered cantilever, one batch.  The coarse
    space bins the nodes along the beam in slabs half as thick as the longest element, so
    every element spans at most three of them; the FP32 two-level solve reaches 1e-10 on
    each instance, agrees with a sparse direct solve of the oracle's assembled system and
    with the plain batched FP64 CG to 1e-6, in at most a fifth of its iterations."""
    import os
    import scipy.sparse as sps
    import scipy.sparse.linalg as spla
    from tlfea_amd import mesh as M
    from tlfea_amd.batch import Batch
    here = os.path.dirname(os.path.abspath(__file__))
    ref = M.read_vtu_ascii(os.path.join(here, "golden", "meshes",
                                        "beam_highres_t10.vtu" if tet10 else "beam_highres_t4.vtu"))
    items = [(M.clamp_and_tip_load(ref, -1.0e3), dict(E=2.1e11, nu=0.3)),
             (_jittered_beam(16, 3, 3, tet10, 0.2, 7), dict(E=7.0e10, nu=0.33))]
    b = Batch(items)
    u64 = _np(b.solve(tol=1e-10))
    it64 = b.iters.copy()
    assert np.all(b.status == 0)
    b.precond_setup(degree=8, ratio=300.0)
    nc = b.coarse_setup()
    u2 = _np(b.solve(tol=1e-10, precond=True))
    it2 = b.iters.copy()
    print(f"[unstructured tet10={tet10}] coarse dim {nc}; plain {list(it64)} -> two-level {list(it2)}",
          flush=True)
    assert np.all(b.status == 0)
    assert it2.max() * 5 <= it64.max()
    for i, (m, mat) in enumerate(items):
        sl = b.inst_slice(i)
        doff, dcol, vo, fo = _oracle_system(m, mat["E"], mat["nu"])
        n = len(fo)
        ud = spla.spsolve(sps.csr_matrix((vo, dcol, doff), shape=(n, n)).tocsc(), fo)
        assert rel_err(u2[sl], ud) < 1e-6
        assert rel_err(u2[sl], u64[sl]) < 1e-6
    b.close()


def test_two_level_on_the_256_beam_sampler_batch(gpu):
    """The batched configuration at full size: the sampler's 256 uneven cantilevers (seed
    2026, 1.68M TET10, 8.2M DOF) solved at once by the FP32 two-level preconditioned FP64
    CG to 1e-8.  Every instance converges; its true residual ||f - K u|| / ||f||, with K
    applied by the FP64 batched operator (itself pinned to the oracle above), is within the
    CG tolerance plus the drift of the recursion; the solutions equal the plain batched
    FP64 CG's to 1e-6; and eight instances spread over the smaller half of the batch match the
    oracle's direct solve to 1e-6."""
    import torch
    from tlfea_amd import mesh as M
    from tlfea_amd.batch import Batch
    items = M.random_beam_batch(256, 2026)
    b = Batch(items)
    print(f"[256 beams] {b.n_elem} elements, {3 * b.n_nodes} DOF", flush=True)
    u64 = b.solve(tol=1e-8)
    assert np.all(b.status == 0)
    b.precond_setup(degree=4)
    b.coarse_setup()
    u = b.solve(tol=1e-8, precond=True)
    assert np.all(b.status == 0)
    print(f"[256 beams] two-level iterations median {np.median(b.iters)} max {b.iters.max()}", flush=True)
    f = b.f.clone()
    f[b.mask.bool()] = 0.0

    def true_rel(v):
        r = f - b.apply(v)
        r[b.mask.bool()] = 0.0
        return np.sqrt(_np(b.dot(r)) / _np(b.dot(f)))

    rel, rel64 = true_rel(u), true_rel(u64)
    worst = int(np.argmax(rel))
    print(f"[256 beams] true rel. residual max {rel.max():.2e} (instance {worst}: plain FP64 CG {rel64[worst]:.2e}); "
          f"plain FP64 CG max {rel64.max():.2e}", flush=True)
    # on the stiffest instances f - K u cannot be formed below eps ||K|| ||u|| / ||f||: both
    # solvers' true residuals sit at that floor there
    assert np.all(rel < np.maximum(2e-8, 3 * rel64))
    un, u64n = _np(u), _np(u64)
    for i in range(256):
        sl = b.inst_slice(i)
        assert rel_err(un[sl], u64n[sl]) < 1e-6
    for i in sorted(range(256), key=lambda k: items[k][0].n_nodes)[:128:16]:
        m, mat = items[i]
        doff, dcol, vo, fo = _oracle_system(m, mat["E"], mat["nu"])
        assert rel_err(un[b.inst_slice(i)], O.band_cholesky_solve(doff, dcol, vo, fo)) < 1e-6
    b.close()
