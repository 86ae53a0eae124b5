#!/usr/bin/env python3
"""Benchmark of the TLFEA solve path on one MI355X.

Workload (single large mesh): cantilever 1.0 x 0.1 x 0.1 m, 200x20x20 hex cells x 6 =
480,000 TET10 elements, 674,081 nodes, 2,022,243 DOF; E = 2.1e11 Pa, nu = 0.3,
rho = 7850; nodes at x = 0 fixed; 1e5 N in -z spread over the x = L nodes.

One step = one linear-static solve to a relative residual of 1e-8 (the reference's
LinearStaticSolver::Solve: tangent K at the current positions, Dirichlet elimination,
CG to sqrt(rr/rr0) < tol).  `value` = solves/s of the FP32 + FP64-refinement path: the
flexible FP64 CG (u, r, alpha, beta and the stopping test in FP64) preconditioned by the
FP32 two-level stage -- Jacobi-Chebyshev on the FP32 operator plus a coarse solve over the
rigid motions of the beam's node planes (P^T K P probed with the FP64 operator, block-
tridiagonal factorisation, its dense inverse on a long chain) -- with that setup inside
every step; the Chebyshev degree and interval ([lambda_max / ratio, lambda_max]) are picked
by timing one solve each of (8, 500), (12, 1000) and (16, 1000) before the timed steps.
Beside it,
on the same problem: the plain FP64 path (block assembly + Dirichlet + CG on the assembled
matrix, the reference's arithmetic) and the matrix-free FP64 CG (K(x) p applied element by
element from the inverse Jacobians, in the single-reduction arrangement).  Inputs are
resident in HBM when the timed region starts.

Kernel rooflines are measured live with HIP events on the launch stream (the null stream,
which torch.cuda.Event records on):
  * roofline      -- the step's dominant kernel pair, the FP32 matrix-free tangent at the
                     rest positions (k_tangent<10,5,float,..,LIN,AFF> + k_bgather<float>: the
                     linear-elasticity form of an affine mesh integrated from the corner
                     values), bytes/launch = 80*n_elem + 24*n_nodes (inverse Jacobian 36 B,
                     detJ 4 B of one quadrature point -- the beam's elements are straight-
                     sided, so all five agree -- and connectivity 40 B per element; p read
                     and y written, 24 B per node);
  * tangent_fp64  -- the same in FP64 (k_tangent<10,5,double,..> + k_gather_tangent, the
                     operator of the matrix-free FP64 CG), bytes/launch = 120*n_elem +
                     48*n_nodes;
  * tangent_general -- the same kernel away from the rest positions (x and X0 read too:
                     120*n_elem + 96*n_nodes);
  * spmv_fp64     -- the FP64 block SpMV of the plain path, bytes/launch =
                     76*node_nnz + 52*n_nodes;
  * force_kernel  -- the fused FP64 P + f_int (block-local plan: force + node gather) with
                     the gradients rebuilt from the rest positions (k_force_rest), on the
                     bytes it moves: connectivity 40 B per element, rest and current
                     positions read (48 B) and f_int written (24 B) per node, bytes/launch =
                     40*n_elem + 72*n_nodes;
  * force_kernel_gradients_read -- the same with the gradients read (k_force_planned), on
                     the reference byte model 1280*n_elem + 48*n_nodes (connectivity 40,
                     gradients 1200, detJ 40 per element; x read and f written per node);
  * hbm_copy      -- a 512 MiB device copy and a 512 MiB read-only reduction (torch), the
                     bandwidth a plain stream reaches on this device, beside the 8 TB/s peak.
cpu_baseline: the FP64 oracle on the box's host cores.  Measured end to end on the
small beam (20x2x2 cells, 3,075 DOF: assembly + Dirichlet + CG to the same tolerance) on
all threads of the process's CPU share and on 1 thread; on the 2M-DOF beam a bounded
sample -- the OpenMP assembly of 1/40 of the elements and 10 CG iterations on the full
matrix -- extrapolated to one solve with the plain GPU CG's iteration count (`value`).
config.true_rel_residual: ||(f - K u)_free|| / ||f_free|| of the timed solution, K applied
by the GPU tangent (the CG itself stops on its recursive residual, as the reference does).

`velocity_solvers` (not part of `value`): SyncedNesterov, SyncedAdamW and SyncedAdamWNocoop
on the same beam, one implicit step of 1e-3 s with zero tolerances (one outer iteration of
200 inner iterations), ms per inner iteration after a warm-up solve, with the assembled mass
and (`*_mf`) with the mass term applied inside the force kernel; and `c1_nesterov`
(`c1_nesterov_mf`), the small beam's implicit step of 1e-3 s solved by SyncedNesterov (with
the momentum restart) to a relative gradient of 1e-8, solves/s.

`batched` (not part of `value`): the batch of 256 uneven cantilevers of the beam sampler
(seed 2026) solved to 1e-8 -- all at once by the batched FP64 CG (finished instances
compacted out), from the learned guess, with the FP32 two-level preconditioner (Chebyshev
stage + rigid-plane coarse correction, setup timed and included) inside the FP64 flexible
CG with and without the learned guess, and one after another by the plain FP64 drop-in
path -- in solves/s; and `nesterov_step`, one implicit step of 1e-3 s from rest for every
beam by the batched SyncedNesterov loop to a relative gradient of 1e-8.

`dynamic` (not part of `value`): 64 beams of the sampler under gravity and sinusoidal tip
loads, 200 implicit steps of 1e-4 s for all beams at once -- Newton steps of the nonlinear
implicit-Euler equation to ||g||/||g(v_n)|| <= 1e-8, steps linearised at rest, and the same
steps by the batched SyncedNesterov loop to the same relative gradient -- and 20 Newton
steps of each beam alone (the per-instance path); instance-steps/s.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
# HBM bytes of one matrix-free tangent launch pair (k_tangent + k_gather_tangent) on the
# 200x20x20 beam: rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 (KiB), scales calibrated
# on a 512 MiB copy
PMC_TANGENT_BYTES = (51216 * 2.0 + 29066 + 27367 * 2.0 + 21638) * 1024
# the same for the rest-state (linear-elasticity) form the linear-static step runs
PMC_TANGENT_REST_BYTES = (35011 * 2.0 + 29066 + 27364 * 2.0 + 21803) * 1024
# the FP32 tangent pair of the two-level stage (k_tangent_em<float> + k_bgather<float>)
PMC_TANGENT32_REST_BYTES = (29260 * 2.0 + 14534 + 16766 * 2.0 + 7912) * 1024
# the same for the force kernels with k_gather_partials: gradient-reading (k_force_planned)
# and rest-coordinate (k_force_rest)
PMC_FORCE_BYTES = (306428 * 2.0 + 29085 + 18373 * 2.0 + 15799) * 1024
PMC_FORCE_REST_BYTES = (23357 * 2.0 + 29066 + 18366 * 2.0 + 15799) * 1024


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def batched_beams(n, tol, log, sequential=True):
    """Solves/s on the sampler's batch of uneven beams: batched paths and the plain
    per-instance drop-in path on the same instances."""
    import numpy as np
    import torch
    from tlfea_amd import mesh as M
    from tlfea_amd.batch import Batch
    from tlfea_amd.fea import Body, LinearStatic
    from tlfea_amd.learned import BatchGraph, WarmStart
    items = M.random_beam_batch(n, 2026)
    b = Batch(items)
    res = {"workload": f"{n} cantilevers of the beam sampler (seed 2026): {b.n_elem} TET10 elements, "
                       f"{3 * b.n_nodes} DOF in total; linear-static solves to {tol:g}"}
    # chunk size, graph-replay length and compaction ratio chosen by the autotuner (setup,
    # not part of the timed solves)
    t = time.perf_counter()
    tune = b.autotune(trial_iters=64, tol=tol, full_solves=True)
    res["autotune"] = dict(tune, tuning_s=time.perf_counter() - t)
    log(f"autotune: {res['autotune']}")

    def timed(tag, fn):
        fn()  # warm-up
        torch.cuda.synchronize()
        t = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        ok = int((b.status == 0).sum())
        res[tag] = {"solves_per_s": ok / dt, "s": dt, "converged": ok, "iters_median": float(np.median(b.iters)),
                    "iters_max": int(b.iters.max())}
        log(f"batched {tag}: {dt:.3f} s, {ok}/{n} converged, median {np.median(b.iters):.0f} iterations")
        return out

    u64 = timed("fp64", lambda: b.solve(tol=tol))
    net = WarmStart.load()
    g = BatchGraph(b)
    uw = timed("fp64_learned_start", lambda: b.solve(tol=tol, u0=net.guess(b, graph=g)))
    # FP32 two-level preconditioner inside the FP64 flexible CG: Chebyshev stage (degree 6 on
    # [lambda_max / 250, lambda_max]) and the rigid-plane coarse correction, set up per batch
    # (timed apart and included in the end-to-end rate)
    torch.cuda.synchronize()
    t = time.perf_counter()
    b.precond_setup(degree=6, ratio=250.0, power_iters=20)
    n_coarse = b.coarse_setup()
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t
    # the autotuner's replay length is chosen for the plain CG's thousands of iterations; the
    # two-level solves finish in 50-140, so they poll every 8 (99 ms against 102 at 16 and
    # 108 at 32 on these beams)
    u2 = timed("fp32_twolevel", lambda: b.solve(tol=tol, precond=True, poll=8))
    u2w = timed("fp32_twolevel_learned_start",
                lambda: b.solve(tol=tol, precond=True, poll=8, u0=net.guess(b, graph=g)))
    for k in ("fp32_twolevel", "fp32_twolevel_learned_start"):
        ok = res[k]["converged"]
        res[k]["setup_s"] = setup_s
        res[k]["solves_per_s_with_setup"] = ok / (res[k]["s"] + setup_s)
    res["fp32_twolevel"]["coarse_dim"] = n_coarse
    rd = {}
    for k, u in (("fp64_learned_start", uw), ("fp32_twolevel", u2), ("fp32_twolevel_learned_start", u2w)):
        rd[k] = max(float(torch.linalg.norm(u[b.inst_slice(i)] - u64[b.inst_slice(i)])
                          / torch.linalg.norm(u64[b.inst_slice(i)])) for i in range(n))
    res["max_rel_diff_vs_fp64"] = rd
    res["speedup_fp32_twolevel_vs_fp64_batched"] = res["fp32_twolevel"]["solves_per_s_with_setup"] / res["fp64"][
        "solves_per_s"]
    if not sequential:
        b.close()
        return res
    # the plain drop-in path, one instance after another
    t_seq, ok = 0.0, 0
    for m, mat in items:
        body = Body(m, mat["E"], mat["nu"], mat["rho"])
        ls = LinearStatic(body, tol=tol, max_iter=200000, check_every=32)
        torch.cuda.synchronize()
        t = time.perf_counter()
        rc = ls.solve()
        torch.cuda.synchronize()
        t_seq += time.perf_counter() - t
        ok += rc == 0
        ls.close()
        body.close()
    res["plain_fp64_sequential"] = {"solves_per_s": ok / t_seq, "s": t_seq, "converged": ok}
    log(f"plain sequential: {t_seq:.1f} s for {n} solves")
    for k in ("fp64", "fp64_learned_start", "fp32_twolevel", "fp32_twolevel_learned_start"):
        res[k]["speedup_vs_plain_sequential"] = res[k]["solves_per_s"] / res["plain_fp64_sequential"]["solves_per_s"]
    b.close()
    return res


def batched_nesterov_bench(n, tol, log, dt=1e-3, max_inner=50000):
    """One implicit step of dt from rest for every beam of the sampler's batch by the batched
    SyncedNesterov loop (tlfea_batch_nesterov_solve: per-instance step size from a power
    iteration, restarted momentum) to ||g|| <= tol ||g(0)||."""
    import numpy as np
    import torch
    from tlfea_amd import mesh as M
    from tlfea_amd.batch import Batch
    items = M.random_beam_batch(n, 2026)
    b = Batch(items)
    rho = np.array([mat["rho"] for _, mat in items])
    torch.cuda.synchronize()
    t = time.perf_counter()
    b.set_mass(rho / dt ** 2, rule="exact")
    alpha = 1.0 / (1.1 * dt * b.lambda_max())
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t
    zero = torch.zeros(3 * b.n_nodes, dtype=torch.float64, device="cuda")

    def run():
        return b.nesterov_solve(dt, rho, alpha, zero, zero, b.f, rel_tol=tol, max_inner=max_inner)
    run()
    torch.cuda.synchronize()
    t = time.perf_counter()
    run()
    torch.cuda.synchronize()
    s = time.perf_counter() - t
    ok = int((b.status == 0).sum())
    res = {"workload": f"{n} cantilevers of the beam sampler (seed 2026), {b.n_elem} TET10 elements: one implicit "
                       f"step of {dt:g} s from rest per beam, SyncedNesterov to ||g|| <= {tol:g} ||g(0)|| (restarted "
                       f"momentum, step 1/(1.1 dt lambda_max) per beam)",
           "solves_per_s": ok / s, "s": s, "converged": ok, "setup_s": setup_s,
           "inner_iters_median": float(np.median(b.iters)), "inner_iters_max": int(b.iters.max()),
           "batched_iterations": b.total_iters, "ms_per_batched_iteration": s * 1e3 / max(b.total_iters, 1)}
    log(f"batched Nesterov step: {s:.2f} s, {ok}/{n} converged, median {np.median(b.iters):.0f} / max "
        f"{b.iters.max()} inner iterations")
    b.close()
    return res


def velocity_solvers_bench(m, E, nu, rho, log, inner=200):
    """SyncedNesterov / SyncedAdamW / SyncedAdamWNocoop on the beam (one implicit step of
    1e-3 s each, zero tolerances so that every solve runs one outer iteration of `inner`
    inner iterations): ms per inner iteration, the rest-coordinate force kernel inside; the
    *_mf lines apply the mass term inside that kernel instead of from the assembled mass."""
    import torch
    from tlfea_amd import capi
    from tlfea_amd.fea import Body, Solver
    out = {}
    for name, kind, mf in (("nesterov", capi.SOLVER_NESTEROV, False), ("adamw", capi.SOLVER_ADAMW, False),
                           ("adamw_nocoop", capi.SOLVER_ADAMW_NOCOOP, False),
                           ("nesterov_mf", capi.SOLVER_NESTEROV, True), ("adamw_mf", capi.SOLVER_ADAMW, True)):
        b = Body(m, E, nu, rho)
        s = Solver(b, kind, planned=True, rest=True)
        if mf:
            s.set_matrix_free_mass(rho)
        if kind == capi.SOLVER_NESTEROV:
            p = capi.NesterovParams(alpha=1e-8, rho=1e14, inner_tol=0.0, outer_tol=0.0, max_outer=1,
                                    max_inner=inner, time_step=1e-3)
        else:
            p = capi.AdamWParams(lr=2e-4, beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=1e-4, lr_decay=0.995,
                                 inner_tol=0.0, outer_tol=0.0, rho=1e14, max_outer=1, max_inner=inner,
                                 time_step=1e-3, convergence_check_interval=10, inner_rtol=0.0)
        s.set_parameters(p)
        s.solve()  # warm-up
        torch.cuda.synchronize()
        t = time.perf_counter()
        s.solve()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        n_inner = s.stats()["inner"]
        out[name] = {"inner_iterations": n_inner, "ms_per_inner": dt * 1e3 / max(n_inner, 1)}
        log(f"{name}: {n_inner} inner iterations in {dt * 1e3:.1f} ms ({dt * 1e3 / max(n_inner, 1):.3f} ms each)")
        del s, b
        torch.cuda.empty_cache()
    return out


def c1_nesterov_bench(E, nu, rho, tol, log, reps=3, mf=False):
    """The 20x2x2-cell beam's implicit step (dt 1e-3 s, penalty 1e10 at lambda = 0, step
    1e-8) solved by SyncedNesterov to ||g_free|| <= tol ||g_free(0)||, with the gradient
    restart of the momentum (the reference's sequence alone converges sublinearly)."""
    import torch
    from tlfea_amd import capi, mesh as M
    from tlfea_amd.fea import Body, Solver
    m = M.beam_config("C1")
    b = Body(m, E, nu, rho)
    s = Solver(b, capi.SOLVER_NESTEROV, planned=True, rest=True)
    if mf:  # the mass term inside the force kernel
        s.set_matrix_free_mass(rho)
    p = capi.NesterovParams(alpha=1e-8, rho=1e10, inner_tol=0.0, outer_tol=0.0, max_outer=1, max_inner=200000,
                            time_step=1e-3)
    s.set_rel_tol(tol)
    s.set_restart(True)
    s.set_poll_interval(256)

    def run():
        for d in range(3):
            b.x[d].copy_(b.X0[d])
        s.set_parameters(p)
        s.solve()

    run()  # warm-up
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        run()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / reps
    st = s.stats()
    out = {"workload": f"20x2x2-cell beam ({3 * m.n_nodes} DOF), one implicit step of 1e-3 s to a relative "
                       f"gradient of {tol:g} (restarted momentum)",
           "solves_per_s": 1.0 / dt, "s": dt, "inner_iterations": st["inner"], "restarts": s.restarts(),
           "final_rel_gradient": st["rel"], "ms_per_inner": dt * 1e3 / max(st["inner"], 1)}
    out["matrix_free_mass"] = mf
    log(f"C1 Nesterov to {tol:g}{' (matrix-free mass)' if mf else ''}: {dt * 1e3:.1f} ms, {st['inner']} inner "
        f"iterations, {s.restarts()} restarts")
    s.close()
    b.close()
    return out


def dynamic_beams_bench(n, steps, seq_steps, tol, log):
    """Batched implicit time steps of n beams: Newton steps of the nonlinear equation (the
    headline) and steps linearised at rest, and the Newton steps of the same beams one by
    one."""
    import numpy as np
    import torch
    from tlfea_amd.dynamics import BatchDynamics, dynamic_beams
    items, amp, freq = dynamic_beams(n, 2026)
    res = {"workload": f"{n} beams of the sampler (seed 2026); gravity + sinusoidal tip load; implicit Euler, "
                       f"dt 1e-4 s, exactly integrated consistent mass; Newton steps: ||g||/||g(v_n)|| <= {tol:g} "
                       f"per instance (inexact Newton, batched CG inside); linearised steps: CG to {tol:g}"}
    final = {}
    b = None
    for mode in ("newton", "linearised", "nesterov"):
        kw = dict(tol=tol, tol_nl=tol, nonlinear=mode == "newton", nesterov=mode == "nesterov", batch=b)
        dyn = BatchDynamics(items, amp, freq, **kw)
        b = dyn.b
        dyn.step()  # warm-up (graph capture)
        dyn = BatchDynamics(items, amp, freq, **kw)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            dyn.step()
        torch.cuda.synchronize()
        dt_b = time.perf_counter() - t
        its = np.array(dyn.iters)
        key = "inner_iters" if mode == "nesterov" else "cg_iters"
        r = {"steps": steps, "s": dt_b, "instance_steps_per_s": n * steps / dt_b,
             f"{key}_median": float(np.median(its)), f"{key}_max": int(its.max())}
        if mode == "newton":
            r["newton_iters_mean"] = float(np.mean(dyn.newton))
            r["newton_iters_max"] = int(np.max(dyn.newton))
            r["instance_steps_at_fp64_floor"] = int(np.sum(dyn.floor))
        res[mode] = r
        final[mode] = dyn.u.clone()
        res.setdefault("n_elem", b.n_elem)
        res.setdefault("n_dof", 3 * b.n_nodes)
        log(f"dynamic {mode}: {steps} steps of {n} beams in {dt_b:.2f} s, median {np.median(its):.0f} iterations")
    for other in ("linearised", "nesterov"):
        res[f"{other}_vs_newton_max_rel_diff"] = max(
            float(torch.linalg.norm(final[other][b.inst_slice(i)] - final["newton"][b.inst_slice(i)])
                  / torch.linalg.norm(final["newton"][b.inst_slice(i)])) for i in range(n))
    # the same beams one at a time (Newton), for seq_steps steps, and the agreement after them
    dyn2 = BatchDynamics(items, amp, freq, tol=tol, tol_nl=tol, nonlinear=True, batch=b)
    for _ in range(seq_steps):
        dyn2.step()
    t_seq, diff = 0.0, 0.0
    for i, it in enumerate(items):
        single = BatchDynamics([it], [amp[i]], [freq[i]], tol=tol, tol_nl=tol, nonlinear=True)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(seq_steps):
            single.step()
        torch.cuda.synchronize()
        t_seq += time.perf_counter() - t
        sl = dyn2.b.inst_slice(i)
        nu = float(torch.linalg.norm(single.u))
        diff = max(diff, float(torch.linalg.norm(dyn2.u[sl] - single.u)) / max(nu, 1e-300))
        single.b.close()
    res["one_by_one"] = {"steps": seq_steps, "s": t_seq, "instance_steps_per_s": n * seq_steps / t_seq}
    res["speedup_vs_one_by_one"] = res["newton"]["instance_steps_per_s"] / res["one_by_one"]["instance_steps_per_s"]
    res["max_rel_diff_batched_vs_one_by_one"] = diff
    log(f"dynamic one by one (Newton): {seq_steps} steps each in {t_seq:.2f} s; max rel diff {diff:.2e}")
    b.close()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--kernel-reps", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cells", type=str, default="200,20,20")
    ap.add_argument("--cheb-degree", type=int, default=0,
                    help="Chebyshev degree of the FP32 stage (0: the fastest of (degree, interval ratio) = (8, 500), "
                         "(12, 1000), (16, 1000) on one trial solve each)")
    ap.add_argument("--cheb-ratio", type=float, default=1000.0,
                    help="lambda_max / lower end of the Chebyshev interval, with --cheb-degree")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-beams measurement")
    ap.add_argument("--batch-n", type=int, default=256)
    ap.add_argument("--batch-only", action="store_true", help="only the batched-beams measurement")
    ap.add_argument("--no-seq", action="store_true", help="skip the one-after-another plain solves of the batch")
    ap.add_argument("--no-dynamic", action="store_true", help="skip the dynamic batched measurement")
    ap.add_argument("--no-solvers", action="store_true", help="skip the Nesterov / AdamW inner-iteration timing")
    ap.add_argument("--solvers-only", action="store_true", help="only the Nesterov / AdamW inner-iteration timing")
    ap.add_argument("--dynamic-only", action="store_true", help="only the dynamic batched measurement")
    ap.add_argument("--no-c1", action="store_true", help="with --solvers-only: skip the C1 Nesterov solves")
    ap.add_argument("--batch-nesterov-only", action="store_true",
                    help="only the batched SyncedNesterov step on the sampler's batch")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="one launch each of the planned force kernel, the block SpMV, the matrix-free "
                         "tangent, the rest-coordinate force kernel, the FP32 tangent and a 512 MiB copy "
                         "(the copy calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE), then exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)

    import numpy as np
    import torch
    from tlfea_amd import capi, mesh as M
    from tlfea_amd.fea import Body, LinearStatic

    capi.lib()
    nx, ny, nz = (int(v) for v in args.cells.split(","))
    E, nu, rho = 2.1e11, 0.3, 7850.0
    if args.pmc_probe:
        m = M.clamp_and_tip_load(M.kuhn_beam(nx, ny, nz), -1.0e5)
        b = Body(m, E, nu, rho)
        f = torch.zeros(3 * m.n_nodes, dtype=torch.float64, device="cuda")
        b.plan
        k = b.stiffness_sell()
        p = torch.randn(3 * m.n_nodes, dtype=torch.float64, device="cuda")
        y = torch.empty_like(p)
        torch.cuda.synchronize()
        b.internal_force_planned(f=f)
        b.spmv(k, p, y)
        b.tangent_apply(p, out=y)
        b.internal_force_rest(f=f)
        b.tangent_apply(p, x=b.X0, out=y)
        # the FP32 operator of the two-level stage (a one-instance batch): k_tangent<float> +
        # k_bgather<float>
        from tlfea_amd.batch import Batch
        bc = Batch([(m, dict(E=E, nu=nu))])
        p32 = p.float()
        y32 = torch.empty_like(p32)
        torch.cuda.synchronize()
        bc.tangent_apply_f32(p32, out=y32)
        torch.cuda.synchronize()
        bc.close()
        src = torch.randn(64 * 1024 * 1024, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log("pmc probe done")
        return
    if args.dynamic_only:
        print(json.dumps({"dynamic": dynamic_beams_bench(64, 200, 20, args.tol, log)}))
        return
    if args.solvers_only:
        m = M.clamp_and_tip_load(M.kuhn_beam(nx, ny, nz), -1.0e5)
        vs = velocity_solvers_bench(m, E, nu, rho, log)
        if not args.no_c1:
            vs["c1_nesterov"] = c1_nesterov_bench(E, nu, rho, args.tol, log)
            vs["c1_nesterov_mf"] = c1_nesterov_bench(E, nu, rho, args.tol, log, mf=True)
        print(json.dumps({"velocity_solvers": vs}))
        return
    if args.batch_nesterov_only:
        print(json.dumps({"batched": {"nesterov_step": batched_nesterov_bench(args.batch_n, args.tol, log)}}))
        return
    if args.batch_only:
        bb = batched_beams(args.batch_n, args.tol, log, not args.no_seq)
        bb["nesterov_step"] = batched_nesterov_bench(args.batch_n, args.tol, log)
        print(json.dumps({"batched": bb}))
        return
    t0 = time.time()
    m = M.clamp_and_tip_load(M.kuhn_beam(nx, ny, nz), -1.0e5)
    log(f"mesh {m.n_elem} el / {m.n_nodes} nodes in {time.time() - t0:.1f}s")
    body = Body(m, E, nu, rho)
    node_nnz, dof_nnz, nb = body.pattern_sizes()
    X0 = [t.clone() for t in body.x]

    def timed_solves(ls):
        def step():
            for d in range(3):
                body.x[d].copy_(X0[d])
            rc = ls.solve()
            if rc != 0:
                raise RuntimeError(f"solve did not converge: rc={rc} iters={ls.iters} res={ls.res}")
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        return time.perf_counter() - t

    # plain FP64 path (the drop-in): assembled K, reference CG rule sqrt(rr/rr0) < tol
    ls64 = LinearStatic(body, tol=args.tol, max_iter=200000, check_every=64)
    dt64 = timed_solves(ls64)
    log(f"plain FP64 solve: {dt64 / args.steps * 1e3:.1f} ms, {ls64.iters} CG iterations, rel res {ls64.res:.3e}")
    # matrix-free FP64 path: same CG, K(x) p applied per element
    ls = LinearStatic(body, tol=args.tol, max_iter=200000, check_every=64, matfree=True)
    dt = timed_solves(ls)
    iters = ls.iters
    ms_per_step = dt / args.steps * 1e3
    log(f"matrix-free FP64 solve: {ms_per_step:.1f} ms, {iters} CG iterations, rel res {ls.res:.3e}")
    # the two solutions agree (same Krylov sequence, different summation order)
    u_diff = (torch.linalg.norm(ls.u - ls64.u) / torch.linalg.norm(ls64.u)).item()
    log(f"matrix-free vs assembled displacement rel diff {u_diff:.3e}")
    # true residual of the timed solution: K u by the GPU tangent at the rest positions
    free = body.fixed_dof_mask() == 0
    true_res = (torch.linalg.norm((body.f_ext - body.tangent_apply(ls.u, x=body.X0))[free])
                / torch.linalg.norm(body.f_ext[free])).item()
    log(f"matrix-free solution: true rel. residual {true_res:.3e} (recursive {ls.res:.3e})")

    # FP32 two-level path (the headline): the batch machinery on this one mesh; every step
    # sets up the preconditioner for K and solves in the flexible FP64 CG
    from tlfea_amd.batch import Batch
    bc = Batch([(m, dict(E=E, nu=nu))])

    def step2(cfg):
        deg, ratio = cfg
        bc.precond_setup(degree=deg, ratio=ratio, power_iters=20)
        bc.coarse_setup()
        # eager launches: every set-up invalidates the captured iteration graph, and on this
        # one large instance a replay is no faster than stream launches (50.5 vs 50.6 ms)
        # while its capture costs ~3.7 ms
        u = bc.solve(tol=args.tol, precond=True, max_iter=20000, graph=False)
        if bc.status[0] != 0:
            raise RuntimeError(f"two-level solve did not converge: status {bc.status[0]}")
        return u

    trial = {}
    cfgs = [(args.cheb_degree, args.cheb_ratio)] if args.cheb_degree > 0 else [(8, 500.0), (12, 1000.0), (16, 1000.0)]
    for cfg in cfgs:
        step2(cfg)
        torch.cuda.synchronize()
        t = time.perf_counter()
        step2(cfg)
        torch.cuda.synchronize()
        trial[cfg] = time.perf_counter() - t
    deg2 = min(trial, key=trial.get)
    log(f"FP32 two-level trials (s per solve by Chebyshev degree and interval ratio): {trial}; using {deg2}")
    for _ in range(args.warmup):
        step2(deg2)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        u2 = step2(deg2)
    torch.cuda.synchronize()
    dt2 = time.perf_counter() - t
    iters2, rel2 = int(bc.iters[0]), float(bc.rel[0])
    # where a step's time goes (one more step, synchronised between its phases)
    phase = {}
    for tag, fn in (("chebyshev_setup", lambda: bc.precond_setup(degree=deg2[0], ratio=deg2[1], power_iters=20)),
                    ("coarse_setup", lambda: bc.coarse_setup()),
                    ("solve", lambda: bc.solve(tol=args.tol, precond=True, max_iter=20000, graph=False))):
        torch.cuda.synchronize()
        t = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        phase[tag] = (time.perf_counter() - t) * 1e3
    log(f"FP32 two-level step phases (ms): {phase}")
    u2_diff = (torch.linalg.norm(u2 - ls.u) / torch.linalg.norm(ls.u)).item()
    true_res2 = (torch.linalg.norm((body.f_ext - body.tangent_apply(u2, x=body.X0))[free])
                 / torch.linalg.norm(body.f_ext[free])).item()
    log(f"FP32 two-level solve: {dt2 / args.steps * 1e3:.1f} ms, {iters2} iterations, rel res {rel2:.3e}, true "
        f"{true_res2:.3e}, rel diff vs FP64 {u2_diff:.3e}")

    # ---- kernel rooflines (HIP events on the null stream) ---------------------------
    R = args.kernel_reps
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def event_time(fn):
        # warm up for ~20 ms of GPU work (clocks ramp up after the host-side phases), then
        # the fastest of five back-to-back batches of R calls: neither a cold start nor a
        # host hiccup between two launches enters the per-call time
        t0 = time.perf_counter()
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        while time.perf_counter() - t0 < 0.02:
            for _ in range(R):
                fn()
            torch.cuda.synchronize()
        best = None
        for _ in range(5):
            e0.record()
            for _ in range(R):
                fn()
            e1.record()
            torch.cuda.synchronize()
            t = e0.elapsed_time(e1) / 1e3 / R
            best = t if best is None else min(best, t)
        return best

    for d in range(3):
        body.x[d].copy_(X0[d])
    p = torch.randn(3 * m.n_nodes, dtype=torch.float64, device="cuda")
    y = torch.empty_like(p)
    # FP32 tangent of the two-level stage: inverse Jacobian (36 B) + detJ (4 B) of one point and
    # connectivity (40 B) per element, p read and y written once (24 B) per node
    p32 = p.float()
    y32 = torch.empty_like(p32)
    t32_s = event_time(lambda: bc.tangent_apply_f32(p32, out=y32))
    t32_bytes = (40 if body.affine else 200) * m.n_elem + 40 * m.n_elem + 24 * m.n_nodes
    t32_gbs = t32_bytes / t32_s / 1e9
    # matrix-free tangent as the step runs it (rest positions: the linear-elasticity form;
    # affine mesh: one quadrature point's geometry): inverse Jacobian (72 B) + detJ (8 B)
    # per element and quadrature point read (one point when affine, else 5) + connectivity
    # (40 B) per element, p read and y written once (48 B) per node
    elem_bytes = (80 if body.affine else 400) + 40
    tan_s = event_time(lambda: body.tangent_apply(p, x=body.X0, out=y))
    tan_bytes = elem_bytes * m.n_elem + 48 * m.n_nodes
    tan_gbs = tan_bytes / tan_s / 1e9
    # the general form (positions away from the rest positions; the step's kernel when a
    # solve starts from a deformed state): x and X0 read as well, 96 B per node
    for d in range(3):
        body.x[d].add_(1e-6)
    tgen_s = event_time(lambda: body.tangent_apply(p, out=y))
    tgen_bytes = elem_bytes * m.n_elem + 96 * m.n_nodes
    tgen_gbs = tgen_bytes / tgen_s / 1e9
    for d in range(3):
        body.x[d].copy_(X0[d])
    # FP64 block SpMV: 3x3 block values (72 B) + column index (4 B) per stored block, x read
    # and y written once (48 B per node), row id per node (4 B)
    spmv_s = event_time(lambda: body.spmv(ls64.k, p, y))
    spmv_bytes = 76 * node_nnz + 52 * m.n_nodes
    spmv_gbs = spmv_bytes / spmv_s / 1e9
    f = torch.zeros(3 * m.n_nodes, dtype=torch.float64, device="cuda")
    # fused P + f_int reading the gradients, on the reference byte model: connectivity 40 +
    # gradients 1200 + detJ 40 B per element, x read and f written (48 B) per node
    force_bytes = 1280 * m.n_elem + 48 * m.n_nodes
    force_s = event_time(lambda: body.internal_force_planned(f=f))
    force_gbs = force_bytes / force_s / 1e9
    # the rest-coordinate form reads no gradients: connectivity 40 B per element, rest and
    # current positions read (48 B) and f written (24 B) per node
    rest_bytes = 40 * m.n_elem + 72 * m.n_nodes
    rest_s = event_time(lambda: body.internal_force_rest(f=f))
    rest_gbs = rest_bytes / rest_s / 1e9
    # what a plain stream reaches: a 512 MiB copy (read + write) and a read-only reduction
    src = torch.randn(64 * 1024 * 1024, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    copy_s = event_time(lambda: dst.copy_(src))
    read_s = event_time(lambda: torch.sum(src))
    copy_gbs = 2 * src.numel() * 8 / copy_s / 1e9
    read_gbs = src.numel() * 8 / read_s / 1e9
    del src, dst
    for tag, gbs in (("tangent", tan_gbs), ("tangent fp32", t32_gbs), ("tangent general", tgen_gbs), ("spmv", spmv_gbs),
                     ("force (gradients read)", force_gbs), ("force (rest coordinates)", rest_gbs)):
        # an algorithmic byte model above the peak would mean the kernel skips work it is charged for
        assert gbs <= HBM_PEAK_GBS, f"{tag}: {gbs:.0f} GB/s exceeds the HBM peak: byte model is wrong"
    log(f"tangent FP32 (rest form) {t32_s * 1e6:.1f} us {t32_gbs:.0f} GB/s; "
        f"tangent (rest form) {tan_s * 1e6:.1f} us {tan_gbs:.0f} GB/s; tangent (general) {tgen_s * 1e6:.1f} us "
        f"{tgen_gbs:.0f} GB/s; spmv64 {spmv_s * 1e6:.1f} us {spmv_gbs:.0f} GB/s; "
        f"force (gradients read) {force_s * 1e6:.1f} us {force_gbs:.0f} GB/s; force (rest coordinates) "
        f"{rest_s * 1e6:.1f} us {rest_gbs:.0f} GB/s; copy {copy_gbs:.0f} GB/s, read {read_gbs:.0f} GB/s")

    # ---- CPU baseline: the oracle on the host cores ----------------------------------
    cpu = None
    if not args.no_cpu:
        import oracle as O
        # the process's CPU share (OMP_NUM_THREADS is set to it on the GPU boxes)
        ncores = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
        model = "unknown"
        try:
            with open("/proc/cpuinfo") as fh:
                model = next(l.split(":", 1)[1].strip() for l in fh if l.startswith("model name"))
        except (OSError, StopIteration):
            pass
        lam, mu = E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))
        # (a) the small beam end to end: assembly + Dirichlet + CG to tol
        m1 = M.beam_config("C1")
        g1, dj1 = O.grad(m1)
        off1, col1 = O.dof_pattern(m1)

        def c1_solve(threads, reps=5):
            O.set_num_threads(threads)
            t = time.perf_counter()
            for _ in range(reps):
                vals = O.stiffness_values(m1, g1, dj1, lam, mu, off1, col1, parallel=threads > 1)
                ve, fb = O.apply_dirichlet(off1, col1, vals, m1.f_ext, m1.fixed)
                _, it1, _ = O.cg(off1, col1, ve, fb, tol=args.tol, max_iter=200000)
            return (time.perf_counter() - t) / reps, it1

        c1_all, c1_iters = c1_solve(ncores)
        c1_one, _ = c1_solve(1)
        # the GPU's matrix-free path on the same small beam, for comparison
        b1 = Body(m1, E, nu, rho)
        ls1 = LinearStatic(b1, tol=args.tol, max_iter=200000, check_every=64, matfree=True)
        X1 = [t.clone() for t in b1.x]

        def gpu_c1():
            for d in range(3):
                b1.x[d].copy_(X1[d])
            ls1.solve()
        gpu_c1()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(10):
            gpu_c1()
        torch.cuda.synchronize()
        c1_gpu = (time.perf_counter() - t) / 10
        ls1.close()
        b1.close()
        # (b) the 2M-DOF beam: OpenMP assembly of a slice of the elements on a per-slice
        # pattern + 10 CG iterations on the full matrix, extrapolated
        O.set_num_threads(ncores)
        go, djo = O.grad(m)
        frac_el = max(1, m.n_elem // 40)
        sub = M.TetMesh(m.x, m.y, m.z, m.elems[:frac_el])
        soff, scol = O.dof_pattern(sub)
        sg, sdj = go[: frac_el * 150], djo[: frac_el * 5]
        ta = time.perf_counter()
        O.stiffness_values(sub, sg, sdj, lam, mu, soff, scol, parallel=True)
        t_asm = (time.perf_counter() - ta) * (m.n_elem / frac_el)
        iters64 = ls64.iters
        doff_t, dcol_t = body.dof_csr()
        vals = body.stiffness_csr_values(ls64.k).cpu().numpy()
        doff_n, dcol_n = doff_t.cpu().numpy(), dcol_t.cpu().numpy()
        bvec = np.random.default_rng(0).standard_normal(3 * m.n_nodes)
        tc = time.perf_counter()
        O.cg(doff_n, dcol_n, vals, bvec, tol=0.0, max_iter=10)
        t_it = (time.perf_counter() - tc) / 10
        cpu_solve_s = t_asm + iters64 * t_it
        del vals, doff_n, dcol_n
        cpu = {"value": 1.0 / cpu_solve_s, "unit": "solves/s", "cores": ncores, "kind": "port",
               "cpu_model": model, "nproc": os.cpu_count(),
               "sample": f"oracle FP64 (OpenMP, {ncores} threads): stiffness assembly of {frac_el} of {m.n_elem} "
                         f"elements + 10 CG iterations on the full {dof_nnz}-nnz CSR, extrapolated to the "
                         f"{iters64} iterations of the plain FP64 GPU CG: {t_asm:.1f} s assembly + "
                         f"{t_it * 1e3:.1f} ms/iteration",
               "c1_measured": {"workload": f"20x2x2-cell beam, {3 * m1.n_nodes} DOF: assembly + Dirichlet + CG "
                                           f"to {args.tol:g} ({c1_iters} iterations), end to end",
                               "s_all_threads": c1_all, "threads": ncores, "s_1_thread": c1_one,
                               "gpu_matfree_s": c1_gpu}}
        log(f"cpu baseline ({model}, {ncores} of {os.cpu_count()} threads): C2 {cpu_solve_s:.1f} s per solve "
            f"(extrapolated); C1 measured {c1_all * 1e3:.1f} ms ({ncores} threads), {c1_one * 1e3:.1f} ms "
            f"(1 thread), GPU {c1_gpu * 1e3:.1f} ms")

    velocity = None if args.no_solvers else velocity_solvers_bench(m, E, nu, rho, log)
    if velocity is not None:
        velocity["c1_nesterov"] = c1_nesterov_bench(E, nu, rho, args.tol, log)
        velocity["c1_nesterov_mf"] = c1_nesterov_bench(E, nu, rho, args.tol, log, mf=True)
    batched = None if args.no_batch else batched_beams(args.batch_n, args.tol, log, not args.no_seq)
    if batched is not None:
        batched["nesterov_step"] = batched_nesterov_bench(args.batch_n, args.tol, log)
    dynamic = None if args.no_dynamic else dynamic_beams_bench(64, 200, 20, args.tol, log)

    # HBM bytes per launch on this workload from two rocprofv3 PMC passes (FETCH_SIZE x 2.0
    # + WRITE_SIZE x 1.0, scales calibrated on a 512 MiB copy); known for the default
    # 200x20x20 beam only
    default = (nx, ny, nz) == (200, 20, 20)
    traffic = (890247 * 2.0 + 31606) * 1024 if default else None
    traffic_tan = PMC_TANGENT_REST_BYTES if default else None
    out = {
        "metric": "TET10 solves/s at 1e-8 rel. residual (learned+FP32 vs plain); force-kernel GB/s",
        "value": args.steps / dt2,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt2 / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"linear-static solve to a relative residual of {args.tol:g} (flexible FP64 CG with "
                               f"the FP32 two-level preconditioner, set up per solve), cantilever {nx}x{ny}x{nz} "
                               f"cells x 6 TET10 ({m.n_elem} el, {m.n_nodes} nodes, {3 * m.n_nodes} DOF), 1e5 N "
                               f"tip load",
                   "n_elem": m.n_elem, "n_dof": 3 * m.n_nodes, "iterations": iters2, "chebyshev_degree": deg2[0],
                   "chebyshev_ratio": deg2[1],
                   "degree_trials_s": {f"{k[0]}/{k[1]:g}": v for k, v in trial.items()},
                   "step_phases_ms": phase,
                   "final_rel_residual": rel2, "true_rel_residual": true_res2,
                   "u_rel_diff_vs_matfree_fp64": u2_diff, "K_nnz": dof_nnz,
                   "plain_fp64": {"solves_per_s": args.steps / dt64, "ms_per_solve": dt64 / args.steps * 1e3,
                                  "cg_iterations": ls64.iters, "final_rel_residual": ls64.res,
                                  "what": "block assembly + Dirichlet + CG on the assembled matrix"},
                   "matfree_fp64": {"solves_per_s": args.steps / dt, "ms_per_solve": ms_per_step,
                                    "cg_iterations": iters, "final_rel_residual": ls.res,
                                    "true_rel_residual": true_res, "u_rel_diff_vs_assembled": u_diff},
                   "speedup_vs_plain_fp64": dt64 / dt2, "speedup_vs_matfree_fp64": dt / dt2},
        "roofline": {"bound": "hbm", "achieved": t32_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": t32_gbs / HBM_PEAK_GBS, "traffic": PMC_TANGENT32_REST_BYTES if default else None,
                     "kernel": "k_tangent_em<float> + k_bgather<float> (the FP32 operator of the two-level stage "
                               "at the rest positions: the linear-elasticity form of an affine mesh integrated from "
                               "the corner values, element outputs summed per node in LDS in a fixed order)",
                     "avg_us": t32_s * 1e6, "bytes_per_launch": t32_bytes},
        "tangent_fp64": {"bound": "hbm", "achieved": tan_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": tan_gbs / HBM_PEAK_GBS, "traffic": traffic_tan,
                     "kernel": "k_tangent<10,5,double,true,0,true,true> + k_gather_tangent (matrix-free K p at the "
                               "rest positions, the linear-elasticity form of an affine mesh integrated from the "
                               "corner values)",
                     "avg_us": tan_s * 1e6, "bytes_per_launch": tan_bytes},
        "tangent_general": {"bound": "hbm", "achieved": tgen_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": tgen_gbs / HBM_PEAK_GBS, "traffic": PMC_TANGENT_BYTES if default else None,
                            "kernel": "k_tangent<10,5,double,true,0,false,true> + k_gather_tangent (K(x) p away from "
                                      "rest)",
                            "avg_us": tgen_s * 1e6, "bytes_per_launch": tgen_bytes},
        "spmv_fp64": {"bound": "hbm", "achieved": spmv_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                      "frac": spmv_gbs / HBM_PEAK_GBS, "traffic": traffic, "kernel": "k_cg_spmv/k_spmv_block",
                      "avg_us": spmv_s * 1e6, "bytes_per_launch": spmv_bytes},
        "force_kernel": {"bound": "hbm", "achieved": rest_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": rest_gbs / HBM_PEAK_GBS, "traffic": PMC_FORCE_REST_BYTES if default else None,
                         "kernel": "k_force_rest<10,5,false> + k_gather_partials (gradients rebuilt from the rest "
                                   "positions; on the bytes this kernel moves: connectivity, rest and current "
                                   "positions, f_int)",
                         "avg_us": rest_s * 1e6, "bytes_per_launch": rest_bytes},
        "force_kernel_gradients_read": {"bound": "hbm", "achieved": force_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                        "frac": force_gbs / HBM_PEAK_GBS, "traffic": PMC_FORCE_BYTES if default else None,
                                        "kernel": "k_force_planned<10,5,false> + k_gather_partials",
                                        "avg_us": force_s * 1e6, "bytes_per_launch": force_bytes},
        "hbm_copy": {"copy_gbs": copy_gbs, "read_gbs": read_gbs, "peak": HBM_PEAK_GBS,
                     "what": "torch 512 MiB device copy (read + write bytes) and read-only sum, HIP events"},
        "cpu_baseline": cpu,
        "velocity_solvers": velocity,
        "batched": batched,
        "dynamic": dynamic,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
