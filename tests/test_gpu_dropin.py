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
