"""Trains the learned warm start on converged solves of the beam sampler and evaluates it.

    python -m tlfea_amd.train_warmstart --n-train 20000 --epochs 8 --out tlfea_amd/weights/warmstart_beams.npz

Training data: FP64 batched linear-static solves (tol 1e-8) of ``mesh.random_beam_batch``
instances with seeds from --seed.  Evaluation: --n-eval held-out instances (seed
--eval-seed), batched CG from the zero guess and from the network's guess (with the
per-instance fallback), iterations to 1e-8 and wall time of each, solution agreement.
Prints one JSON line with the results.
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np


def evaluate(net, n_eval, seed, tol=1e-8, log=print):
    import numpy as np
    import torch
    from . import mesh as M
    from .batch import Batch
    from .learned import BatchGraph
    items = M.random_beam_batch(n_eval, seed)
    b = Batch(items)
    g = BatchGraph(b)
    b.solve(tol=tol)  # warm-up (graph capture, caches)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    u_zero = b.solve(tol=tol)
    torch.cuda.synchronize()
    t_zero = time.perf_counter() - t0
    it_zero = b.iters.copy()
    # the raw guess's error and residual against the converged zero-start solution
    raw = net.guess(b, graph=g, policy=None, smooth=0)
    net.guess(b, graph=g, policy="residual", smooth=0)
    res_ratio = net.res_ratio.copy()
    fb_res = int(net.fell_back.sum())
    rel = lambda v: [torch.linalg.norm(v[b.inst_slice(i)] - u_zero[b.inst_slice(i)]).item()
                     / max(torch.linalg.norm(u_zero[b.inst_slice(i)]).item(), 1e-300) for i in range(b.n_inst)]
    raw_err = rel(raw)
    variants = {}
    b.precond_setup(degree=8, ratio=30.0, power_iters=20)
    for name, kw in (("raw", dict(policy=None, smooth=0)), ("smooth2", dict(policy="energy", smooth=2)),
                     ("smooth4", dict(policy="energy", smooth=4))):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ug = net.guess(b, graph=g, **kw)
        b.solve(tol=tol, u0=ug)
        torch.cuda.synchronize()
        variants[name] = dict(t_s=time.perf_counter() - t0, iters_total=int(b.iters.sum()),
                              iters_median=float(np.median(b.iters)), iters_max=int(b.iters.max()),
                              fallbacks=int(net.fell_back.sum()), guess_err_median=float(np.median(rel(ug))),
                              converged=int((b.status == 0).sum()))
        log(f"[learned] variant {name}: {variants[name]}")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    u0 = net.guess(b, graph=g)
    u_warm = b.solve(tol=tol, u0=u0)
    torch.cuda.synchronize()
    t_warm = time.perf_counter() - t0
    it_warm = b.iters.copy()
    # relative error of the guess and agreement of the two converged solutions
    err0, agree = [], []
    for i in range(b.n_inst):
        sl = b.inst_slice(i)
        nz = torch.linalg.norm(u_zero[sl]).item()
        err0.append(torch.linalg.norm(u0[sl] - u_zero[sl]).item() / max(nz, 1e-300))
        agree.append(torch.linalg.norm(u_warm[sl] - u_zero[sl]).item() / max(nz, 1e-300))
    res = dict(n_eval=n_eval, seed=seed, tol=tol, t_zero_s=t_zero, t_warm_s=t_warm, speedup=t_zero / t_warm,
               iters_zero_total=int(it_zero.sum()), iters_warm_total=int(it_warm.sum()),
               iters_zero_median=float(np.median(it_zero)), iters_warm_median=float(np.median(it_warm)),
               iters_zero_max=int(it_zero.max()), iters_warm_max=int(it_warm.max()),
               fallbacks=int(net.fell_back.sum()), guess_rel_err_median=float(np.median(err0)),
               raw_guess_rel_err_median=float(np.median(raw_err)), raw_guess_rel_err_p90=float(np.percentile(raw_err, 90)),
               raw_residual_ratio_median=float(np.median(res_ratio)), residual_policy_fallbacks=fb_res,
               solution_rel_diff_max=float(np.max(agree)), converged_zero=int((b.status == 0).sum()),
               variants=variants)
    log(f"[learned] eval: {json.dumps(res)}")
    b.close()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n-train", type=int, default=20000)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--seed", type=int, default=1000)
    ap.add_argument("--n-eval", type=int, default=256)
    ap.add_argument("--eval-seed", type=int, default=9000000)
    ap.add_argument("--inst-per-step", type=int, default=32)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--out", type=str, default=None)
    ap.add_argument("--load", type=str, default=None, help="evaluate these weights instead of training")
    ap.add_argument("--init", type=str, default=None, help="start from these weights (skips the L2 stage)")
    ap.add_argument("--energy-epochs", type=int, default=0, help="energy-norm fine-tuning epochs")
    ap.add_argument("--n-energy", type=int, default=4096, help="solves kept for the energy loss")
    args = ap.parse_args()
    from . import capi
    from .learned import WEIGHTS, WarmStart, generate, generate_energy, train, train_energy
    capi.lib()
    log = lambda s: print(s, flush=True)
    if args.load:
        res = evaluate(WarmStart.load(args.load), args.n_eval, args.eval_seed, log=log)
        print(json.dumps(res), flush=True)
        return
    t_gen = t_train = 0.0
    hist, ehist = [], []
    if args.init:
        net = WarmStart.load(args.init)
    else:
        t0 = time.time()
        data = generate(args.n_train, args.seed, log=log)
        t_gen = time.time() - t0
        t0 = time.time()
        net, hist = train(data, epochs=args.epochs, inst_per_step=args.inst_per_step, lr=args.lr, seed=args.seed,
                          log=log)
        t_train = time.time() - t0
        del data
    if args.energy_epochs > 0:
        import torch
        torch.cuda.empty_cache()
        t0 = time.time()
        edata = generate_energy(args.n_energy, args.seed + 5000000, log=log)
        t_gen += time.time() - t0
        t0 = time.time()
        net, ehist = train_energy(edata, net, epochs=args.energy_epochs, inst_per_step=args.inst_per_step,
                                  seed=args.seed, log=log)
        t_train += time.time() - t0
        for d in edata:
            d["batch"].close()
        del edata
    out = args.out or WEIGHTS
    net.save(out, n_train=args.n_train, epochs=args.epochs, seed=args.seed, loss_history=np.asarray(hist),
             energy_epochs=args.energy_epochs, n_energy=args.n_energy, energy_history=np.asarray(ehist))
    log(f"[learned] saved {out}")
    res = evaluate(net, args.n_eval, args.eval_seed, log=log)
    res.update(n_train=args.n_train, epochs=args.epochs, t_generate_s=t_gen, t_train_s=t_train,
               final_loss=hist[-1] if hist else None, energy_epochs=args.energy_epochs,
               final_energy_error=ehist[-1] if ehist else None)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
