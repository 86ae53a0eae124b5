"""The C++ drop-in API end to end on the GPU (tlfea_driver over FEASolver, GPU_FEAT10_Data /
GPU_FEAT4_Data and the solver classes), on the reference's example scenarios, compared with
the oracle."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from helpers import MESHES, lame, load_mesh, rel_err
from tlfea_amd import build as B
from tlfea_amd import mesh as M

pytestmark = pytest.mark.gpu


def _run(args, tmp_path):
    out = str(tmp_path / "pos.bin")
    print("[driver]", " ".join(args), flush=True)
    r = subprocess.run([B.DRIVER, *args, "--out", out], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    info = json.loads(r.stdout.strip().splitlines()[-1])
    pos = np.fromfile(out, dtype=np.float64).reshape(3, -1)
    return info, pos


@pytest.mark.parametrize("assembled", [False, True])
@pytest.mark.parametrize("spec,name,load", [("kuhn:20,2,2", "C1", -1000.0), ("kuhn4:12,3,2", "kuhn_t4", -500.0),
                                             ("vtu:" + os.path.join(MESHES, "beam.vtu"), "beam_vtu", -1000.0)])
def test_linear_static_dropin(gpu, tmp_path, spec, name, load, assembled):
    """examples/beam_linear_static.cc: clamp x = x_min, the load spread over the x = x_max
    nodes in -z, CG at tol 1e-12 -> displacements against the oracle's banded Cholesky solve
    within 1e-10 (kappa * eps allows it on these meshes), with the matrix-free operator (the
    default once CalcDnDuPre has recorded the rest positions) and with the assembled one."""
    tol = 1e-10
    info, pos = _run(["linear_static", "--mesh", spec, "--tol", "1e-12", "--max-iter", "200000", "--load", str(load),
                      "--assembled", "1" if assembled else "0"], tmp_path)
    assert info["matrix_free"] is (not assembled)
    m = load_mesh(name)
    lam, mu = lame(2.1e11, 0.3)
    g, dj = O.grad(m)
    off, col = O.dof_pattern(m)
    vals = O.stiffness_values(m, g, dj, lam, mu, off, col)
    vals, fb = O.apply_dirichlet(off, col, vals, m.f_ext, m.fixed)
    u = O.band_cholesky_solve(off, col, vals, fb).reshape(-1, 3)
    assert info["n_fixed"] == len(m.fixed)
    du = pos.T - m.xyz()
    assert rel_err(du, u) < tol
    assert info["residual"] < 1e-12


@pytest.mark.parametrize("spec,name,load", [("kuhn:20,2,2", "C1", -1000.0), ("kuhn4:12,3,2", "kuhn_t4", -500.0),
                                             ("vtu:" + os.path.join(MESHES, "beam_highres_t10.vtu"),
                                              "beam_highres_t10", -1000.0)])
def test_linear_static_dropin_two_level(gpu, tmp_path, spec, name, load):
    """LinearStaticSolver::UseTwoLevelPreconditioner (not in the reference): the flexible
    FP64 CG preconditioned by the FP32 two-level stage, at tol 1e-12 on the reference's
    scenario, matches a sparse direct solve of the oracle's system within 1e-10, as the
    plain drop-in does; two Solve calls in a row (the second re-linearised at the moved
    positions, the operator following them through x - X0) end where the plain drop-in's
    two calls end, within 1e-9 of the displacement, in at most a fifth of its iterations."""
    import scipy.sparse as sps
    import scipy.sparse.linalg as spla
    base = ["linear_static", "--mesh", spec, "--tol", "1e-12", "--max-iter", "200000", "--load", str(load)]
    info, pos = _run(base + ["--two-level", "8"], tmp_path)
    assert info["two_level"] is True
    m = load_mesh(name)
    lam, mu = lame(2.1e11, 0.3)
    g, dj = O.grad(m)
    off, col = O.dof_pattern(m)
    vals = O.stiffness_values(m, g, dj, lam, mu, off, col)
    vals, fb = O.apply_dirichlet(off, col, vals, m.f_ext, m.fixed)
    n = len(fb)
    u = spla.spsolve(sps.csr_matrix((vals, col, off), shape=(n, n)).tocsc(), fb).reshape(-1, 3)
    assert rel_err(pos.T - m.xyz(), u) < 1e-10
    assert info["residual"] < 1e-12
    info_p, pos_p = _run(base + ["--solves", "2"], tmp_path)
    info_t, pos_t = _run(base + ["--solves", "2", "--two-level", "8"], tmp_path)
    print(f"[two-level drop-in] {name}: iterations plain {info_p['iters']} two-level {info_t['iters']}", flush=True)
    assert info_t["two_level"] is True and info_p["two_level"] is False
    assert rel_err(pos_t.T - m.xyz(), pos_p.T - m.xyz()) < 1e-9
    assert info_t["iters"] * 5 <= info_p["iters"]


def test_linear_static_dropin_damped(gpu, tmp_path):
    """A damped body through LinearStaticSolver: the reference's system is (K + C) u = f
    (the Kelvin-Voigt block added unscaled).  The first Solve (positions at rest) matches the
    oracle's direct solve of it; a second Solve re-linearises at the moved positions; both
    run matrix-free and match the all-assembled run of the same two solves."""
    lam, mu = lame(2.1e11, 0.3)
    eta, ld = 0.25 * mu, 0.5 * lam
    base = ["linear_static", "--mesh", "kuhn:20,2,2", "--tol", "1e-12", "--max-iter", "200000", "--load", "-1000",
            "--eta", repr(eta), "--lambda-d", repr(ld)]
    info, pos = _run(base, tmp_path)
    assert info["matrix_free_each"] == [True]
    m = load_mesh("C1")
    g, dj = O.grad(m)
    off, col = O.dof_pattern(m)
    vals = O.stiffness_values(m, g, dj, lam, mu, off, col, eta=eta, lambda_d=ld)
    vals, fb = O.apply_dirichlet(off, col, vals, m.f_ext, m.fixed)
    u = O.band_cholesky_solve(off, col, vals, fb).reshape(-1, 3)
    assert rel_err(pos.T - m.xyz(), u) < 1e-10
    info2, pos2 = _run(base + ["--solves", "2"], tmp_path)
    assert info2["matrix_free_each"] == [True, True]
    info3, pos3 = _run(base + ["--solves", "2", "--assembled", "1"], tmp_path)
    assert info3["matrix_free_each"] == [False, False]
    d2, d3 = pos2.T - m.xyz(), pos3.T - m.xyz()
    assert rel_err(d2, d3) < 1e-9


@pytest.mark.parametrize("mode", ["nesterov", "adamw", "adamw_nocoop"])
def test_iterative_dropin_on_cube(gpu, tmp_path, mode):
    """examples/test_feat10_nesterov.cc / test_feat10_adamw.cc: cube.1 (TetGen), z = 0
    fixed, 1000 N in +x on node 6, three Solve() steps with the examples' parameters;
    final positions against the oracle's restatement of the same loop."""
    args = [mode, "--mesh", "tetgen:" + os.path.join(MESHES, "cube.1"), "--fix", "z0", "--node-load", "6,1000",
            "--E", "7e8", "--nu", "0.33", "--rho0", "2700", "--steps", "3"]
    if mode != "nesterov":
        args += ["--lambda-d", "0.1"]
    info, pos = _run(args, tmp_path)
    m = load_mesh("cube")
    prob = O.Problem(m, 7e8, 0.33, 2700.0, lambda_d=0.0 if mode == "nesterov" else 0.1)
    st = prob.new_state()
    counts = []
    for _ in range(3):
        if mode == "nesterov":
            counts.append(prob.nesterov_step(st, 1e-8, 1e14, 1e-6, 1e-6, 5, 300, 1e-3))
        else:
            counts.append(prob.adamw_step(st, 2e-4, 0.9, 0.999, 1e-8, 1e-4, 0.995, 1e-1, 1e-6, 1e14, 5, 300, 1e-3,
                                          10, 0.0, nocoop=(mode == "adamw_nocoop")))
    du = pos.T - m.xyz()
    duo = np.stack([st["x"], st["y"], st["z"]], 1) - m.xyz()
    # AdamW from rest amplifies rounding noise (see test_gpu_solvers); a 5% band there
    assert rel_err(du, duo) < (1e-6 if mode == "nesterov" else 5e-2)
    assert len(info["inner_iters"]) == 3


def _write_raw(path, *arrays):
    with open(path, "wb") as fh:
        for a in arrays:
            np.ascontiguousarray(a, dtype=np.float64).tofile(fh)


@pytest.mark.parametrize("mode", ["nesterov", "adamw", "adamw_nocoop"])
def test_iterative_dropin_trajectory(gpu, tmp_path, mode):
    """The drop-in solver classes against the oracle on a well-conditioned trajectory (the
    C-ABI tests' setting in test_gpu_solvers): steel cube, loads on every DOF, a deformed
    start set by UpdatePositions after CalcDnDuPre, Kelvin-Voigt damping for AdamW, zero
    tolerances so that both sides run exactly 2 Solve() x 2 outer x 25-30 inner iterations.
    The wrapper's parameter plumbing, state reset and body refresh are checked at the
    C-ABI bars (Nesterov 1e-8, AdamW 1e-6)."""
    from helpers import E_STEEL, NU_STEEL, RHO_STEEL, deformed
    m = load_mesh("cube")
    load = np.random.default_rng(21).standard_normal(3 * m.n_nodes) * 50.0
    x0, y0, z0 = deformed(m, seed=22, amp=2e-5, rot=0.0)
    lf, sf = str(tmp_path / "load.bin"), str(tmp_path / "start.bin")
    _write_raw(lf, load)
    _write_raw(sf, x0, y0, z0)
    eta, ld = (0.0, 0.0) if mode == "nesterov" else (1e4, 1e3)
    args = [mode, "--mesh", "tetgen:" + os.path.join(MESHES, "cube.1"), "--fix", "z0", "--load-file", lf,
            "--start-file", sf, "--E", repr(E_STEEL), "--nu", repr(NU_STEEL), "--rho0", repr(RHO_STEEL),
            "--eta", repr(eta), "--lambda-d", repr(ld), "--steps", "2", "--inner-tol", "0", "--outer-tol", "0",
            "--max-outer", "2", "--rho", "1e10", "--dt", "1e-3"]
    args += ["--alpha", "2e-10", "--max-inner", "25"] if mode == "nesterov" else ["--max-inner", "30"]
    info, pos = _run(args, tmp_path)
    prob = O.Problem(m, E_STEEL, NU_STEEL, RHO_STEEL, eta=eta, lambda_d=ld, f_ext=load)
    st = prob.new_state()
    st["x"][:], st["y"][:], st["z"][:] = x0, y0, z0
    counts = []
    for _ in range(2):
        if mode == "nesterov":
            counts.append(prob.nesterov_step(st, 2e-10, 1e10, 0.0, 0.0, 2, 25, 1e-3))
        else:
            counts.append(prob.adamw_step(st, 2e-4, 0.9, 0.999, 1e-8, 1e-4, 0.995, 0.0, 0.0, 1e10, 2, 30, 1e-3, 10,
                                          0.0, nocoop=(mode == "adamw_nocoop")))
    assert info["inner_iters"] == counts
    X0 = m.xyz()
    duo = np.stack([st["x"], st["y"], st["z"]], 1) - X0
    assert rel_err(pos.T - X0, duo) < (1e-8 if mode == "nesterov" else 1e-6)


@pytest.mark.parametrize("fix2", ["x0+z0", "z0:rev"])
def test_update_nodal_fixed_dropin(gpu, tmp_path, fix2):
    """UpdateNodalFixed between Solve calls (FEAT10Data.cu:674-751) through the drop-in, then
    CalcConstraintData and ConvertToCSR_ConstraintJacT as a caller does: the cube is held at
    z = 0 for the first step and then at z = 0 and x = x_min ("x0+z0": another count, the
    multipliers start at zero) or at the same nodes in reverse order ("z0:rev": same count, the
    multipliers keep their constraint numbers, as the reference's solver keeps
    d_lambda_guess).  Three Nesterov steps against the oracle doing the same, with the
    reference example's aluminium-like material (test_feat10_nesterov.cc: E 7e8, nu 0.33,
    rho 2700) so that the displacements stay far above the rounding of the unit-size
    positions they are read from."""
    E, nu, rho0 = 7e8, 0.33, 2700.0
    m = load_mesh("cube")
    load = np.random.default_rng(5).standard_normal(3 * m.n_nodes) * 50.0
    lf = str(tmp_path / "load.bin")
    _write_raw(lf, load)
    args = ["nesterov", "--mesh", "tetgen:" + os.path.join(MESHES, "cube.1"), "--fix", "z0", "--load-file", lf,
            "--E", repr(E), "--nu", repr(nu), "--rho0", repr(rho0), "--steps", "3",
            "--inner-tol", "0", "--outer-tol", "0", "--max-outer", "2", "--max-inner", "25", "--rho", "1e10",
            "--dt", "1e-3", "--alpha", "2e-10", "--update-fixed-after", "1", "--fix2", fix2]
    info, pos = _run(args, tmp_path)
    z0 = np.nonzero(np.abs(m.z) < 1e-8)[0]
    L = np.ptp(m.x)
    new = z0[::-1] if fix2 == "z0:rev" else np.nonzero((np.abs(m.x - m.x.min()) < 5e-4 * L) |
                                                       (np.abs(m.z) < 1e-8))[0]
    assert info["n_fixed"] == len(z0) and info["n_fixed_final"] == len(new)
    assert fix2 == "z0:rev" or len(new) != len(z0)
    prob = O.Problem(m, E, nu, rho0, f_ext=load, fixed=z0)
    st = prob.new_state()
    for step in range(3):
        if step == 1:
            prob.set_fixed(new)
            if len(new) != len(z0):
                st["lam"] = np.zeros(3 * len(new))
        prob.nesterov_step(st, 2e-10, 1e10, 0.0, 0.0, 2, 25, 1e-3)
    X0 = m.xyz()
    duo = np.stack([st["x"], st["y"], st["z"]], 1) - X0
    assert rel_err(pos.T - X0, duo) < 1e-8


def test_constraint_jacobian_csr_dropin(gpu, tmp_path):
    """ConvertToCSR_ConstraintJac and ConvertToCSR_ConstraintJacT (FEAT10Data.cu:478-562)
    through the drop-in: offsets, columns and values equal the oracle's selection matrices
    exactly (J: one 1.0 per constraint row at DOF 3 fixed[k] + d; J^T: rows over the 3N
    DOFs, columns ascending)."""
    jf = str(tmp_path / "jac.txt")
    info, _ = _run(["nesterov", "--mesh", "tetgen:" + os.path.join(MESHES, "cube.1"), "--fix", "z0:rev",
                    "--steps", "0", "--jac-out", jf], tmp_path)
    arrays = {}
    for line in open(jf):
        tok = line.split()
        arrays[tok[0]] = np.array(tok[2:], dtype=np.float64)
        assert int(tok[1]) == len(tok) - 2
    m = load_mesh("cube")
    fixed = np.nonzero(np.abs(m.z) < 1e-8)[0][::-1]
    for t, pre in ((False, "j_"), (True, "jt_")):
        o, c, v = O.constraint_jacobian_csr(fixed, m.n_nodes, transpose=t)
        assert np.array_equal(arrays[pre + "offsets"].astype(np.int64), o)
        assert np.array_equal(arrays[pre + "columns"].astype(np.int64), c)
        assert np.array_equal(arrays[pre + "values"], v)


@pytest.mark.parametrize("spec,name", [("kuhn:20,2,2", "C1"), ("kuhn4:12,3,2", "kuhn_t4")])
def test_vtk_output_after_solve(gpu, tmp_path, spec, name):
    """WriteOutputVTK (FEAT10Data.cu:759-794) and ANCFCPUUtils::WriteFEAT10ToVTK /
    WriteFEAT4ToVTK (mesh_utils.cc:962-1077) after a drop-in linear-static solve: both
    files parse, hold the solved positions to the 6 significant digits of the format, the
    connectivity exactly and the cell type, and the second one the displacement field."""
    vtk, feat = str(tmp_path / "out.vtk"), str(tmp_path / "feat.vtk")
    load = "-500" if name == "kuhn_t4" else "-1000"
    info, pos = _run(["linear_static", "--mesh", spec, "--tol", "1e-10", "--max-iter", "200000", "--load", load,
                      "--vtk", vtk, "--vtk-feat", feat], tmp_path)
    m = load_mesh(name)
    P = pos.T
    nn = m.nn
    for path, title, ptype in ((vtk, f"T{nn} mesh output", "float"),
                               (feat, "FEAT10 Tetrahedral Mesh" if nn == 10 else "FEAT4 Tetrahedral Mesh", "double")):
        v = M.read_vtk_legacy(path)
        assert v["header"] == "# vtk DataFile Version 3.0" and v["title"] == title
        assert v["point_type"] == ptype
        assert np.array_equal(v["cells"], m.elems)
        assert np.all(v["cell_types"] == (24 if nn == 10 else 10))
        assert np.all(np.abs(v["points"] - P) <= 5.01e-6 * np.abs(P) + 1e-300)
    d = M.read_vtk_legacy(feat)["point_vectors"]["displacement"]
    du = P - m.xyz()
    assert np.all(np.abs(d - du) <= 5.01e-6 * np.abs(du) + 1e-12 * np.abs(m.xyz()).max())
