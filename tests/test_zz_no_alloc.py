"""tlfea_cg_solve_matfree_noalloc (no counterpart in the reference: the matrix-free CG with
its workspace in the library's static device array, for callers that cannot allocate)
against tlfea_cg_solve_matfree on the same inputs, through the C-ABI: two calls in a row,
the second on a smaller body, so that the static workspace is reused with stale contents
from the first call; and the size limit."""
import ctypes as C

import numpy as np
import pytest
import torch

from helpers import E_STEEL, NU_STEEL, RHO_STEEL, load_mesh, rel_err
from tlfea_amd import capi
from tlfea_amd import mesh as M
from tlfea_amd.fea import Body, LinearStatic, ptr

pytestmark = pytest.mark.gpu


def _direct(m):
    """The oracle's banded Cholesky solve of the Dirichlet-eliminated linear-static system."""
    import oracle as O
    from helpers import lame
    lam, mu = lame(E_STEEL, NU_STEEL)
    g, dj = O.grad(m)
    off, col = O.dof_pattern(m)
    vals = O.stiffness_values(m, g, dj, lam, mu, off, col)
    vals, fb = O.apply_dirichlet(off, col, vals, m.f_ext, m.fixed)
    return O.band_cholesky_solve(off, col, vals, fb)


def _noalloc(ls, tol):
    b = ls.b
    ls.assemble()
    u = torch.full_like(ls.u, 7.0)  # garbage: the solve must overwrite every DOF
    it, res = C.c_int(), C.c_double()
    rc = capi.lib().tlfea_cg_solve_matfree_noalloc(
        b.plan, ptr(b.x[0]), ptr(b.x[1]), ptr(b.x[2]), ptr(b.X0[0]), ptr(b.X0[1]), ptr(b.X0[2]), ptr(b.jinv_t),
        ptr(b.detJ_t), int(b.affine), C.byref(b.quad), C.byref(b.mat), ptr(b.fixed_dof_mask()), ptr(ls.f), ptr(u),
        float(tol), 200000, 32, C.byref(it), C.byref(res), b.s)
    torch.cuda.synchronize()
    return rc, u, it.value, res.value


@pytest.mark.parametrize("first,second", [(("kuhn", (30, 3, 3)), ("C1", None)), (("C1", None), ("kuhn_t4", None))])
def test_noalloc_matches_matfree_on_two_calls(gpu, first, second):
    tol = 1e-12
    for name, cells in (first, second):
        print(f"[noalloc] {name} {cells}", flush=True)
        m = M.clamp_and_tip_load(M.kuhn_beam(*cells), -1000.0) if name == "kuhn" else load_mesh(name)
        b = Body(m, E_STEEL, NU_STEEL, RHO_STEEL)
        ls = LinearStatic(b, tol=tol, max_iter=200000, check_every=32, matfree=True)
        ls.assemble()
        assert ls.cg_solve() == capi.TLFEA_OK
        u_ref, it_ref = ls.u.clone(), ls.iters
        rc, u, it, res = _noalloc(ls, tol)
        assert rc == capi.TLFEA_OK
        assert res < tol
        # the same kernels and order of operations; the tangent's in-block LDS sums follow
        # wave scheduling, so the two solves agree to rounding, not bit for bit.  Near the
        # FP64 floor (tol 1e-12) that rounding moves the iteration count a lot -- two runs of
        # the same kernels took 1173 / 1229 iterations on C1 and 562 / 681 on the 468-DOF
        # TET4 beam -- so the outputs are compared: with each other and with the oracle's
        # direct solve, at the drop-in's 1e-10 bar
        un, ur = u.cpu().numpy(), u_ref.cpu().numpy()
        assert rel_err(un, ur) < 1e-10
        ud = _direct(m)
        assert rel_err(un, ud) < 1e-10 and rel_err(ur, ud) < 1e-10
        ls.close()
        b.close()


def test_noalloc_rejects_bodies_beyond_its_limit(gpu):
    """A plan of more than TLFEA_NOALLOC_MAX_NODES nodes is refused with TLFEA_ECAPACITY
    before any launch (the check needs the plan only, so a plan over a huge node count with
    a small connectivity suffices)."""
    n_nodes = 1048576 + 1
    m = M.kuhn_beam(2, 1, 1)
    conn = torch.as_tensor(m.conn_colmajor(), dtype=torch.int32, device="cuda")
    plan = C.c_void_p()
    capi.call("tlfea_force_plan_create", 10, m.n_elem, n_nodes, ptr(conn), C.byref(plan), None)
    d = torch.zeros(3 * n_nodes, dtype=torch.float64, device="cuda")
    q = capi.quadrature(10)
    mat = capi.Material()
    it, res = C.c_int(), C.c_double()
    rc = capi.lib().tlfea_cg_solve_matfree_noalloc(plan, ptr(d), ptr(d), ptr(d), ptr(d), ptr(d), ptr(d), ptr(d),
                                                   ptr(d), 0, C.byref(q), C.byref(mat), None, ptr(d), ptr(d), 1e-8,
                                                   10, 4, C.byref(it), C.byref(res), None)
    capi.call("tlfea_force_plan_destroy", plan)
    assert rc == capi.TLFEA_ECAPACITY
    assert b"exceed" in capi.lib().tlfea_last_error()
