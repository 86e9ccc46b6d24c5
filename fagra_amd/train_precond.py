"""Train the learned block-Jacobi correction on the GPU.

    python -m fagra_amd.train_precond --seeds 64 --out fagra_amd/learned_precond.f32

Captures every Schur system that LM (FP32, block-Jacobi PCG) attempts on seeded 256-camera
problems (ring256: 60,000 points, tracks U{3..8}, Huber 1 px), trains the network with the
library's forward / backward / Adam kernels, and reports the held-out loss and the PCG
iteration counts with and without the network on a held-out seed.
"""
import argparse
import json
import sys
import time

import numpy as np


def log(*a):
    print(*a, flush=True)


def evaluate(net, seed, workload="ring256", cg_tol=1e-3, lambdas=(1e-4, 1e-3, 1e-2, 1e-1)):
    """PCG iterations of block-Jacobi vs learned on the first linearization of a held-out seed."""
    import fagra_amd as fa
    from fagra_amd import synthetic
    p = synthetic.generate(workload, seed=seed)
    out = {}
    for name, pc in (("block_jacobi", fa.PRECOND_BLOCK_JACOBI), ("learned", fa.PRECOND_LEARNED)):
        ba = fa.BalProblem(p, precision=fa.FP32)
        s = fa.Schur(precision=fa.FP32, relative_tolerance=cg_tol, preconditioner=pc)
        fa.lib().fg_schur_set_mlp(s.h, net.h)
        assert s.prepare(ba) == fa.OK and s.linearize() == fa.OK and s.prepare_damping(1e-6) == fa.OK
        its = []
        for lam in lambdas:
            st, d, pr = s.solve_damped(lam)
            dg = s.diagnostics()
            its.append({"lambda": lam, "status": st, "iterations": dg.iterations, "fallback": dg.used_fallback,
                        "reduced_residual": dg.reduced_relative_residual})
        out[name] = its
        s.close()
        ba.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--seeds", type=int, default=64)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--max-iter", type=int, default=12)
    ap.add_argument("--out", default=None)
    ap.add_argument("--report", default=None)
    args = ap.parse_args()
    from fagra_amd import learned

    t = time.time()
    feats, targs = learned.collect(range(1, args.seeds + 1), max_iter=args.max_iter, log=log)
    log(f"train: {len(feats)} samples from {args.seeds} problems in {time.time() - t:.1f}s")
    hold_f, hold_t = learned.collect([1000], max_iter=args.max_iter)
    net = learned.Network(seed=0)
    base = net.loss_grad(hold_f, hold_t)[0]
    t = time.time()
    losses = net.train(feats, targs, epochs=args.epochs, batch=args.batch, lr=args.lr, seed=0)
    log(f"train: {args.epochs} epochs in {time.time() - t:.1f}s; loss {losses[0]:.5f} -> {losses[-1]:.5f}")
    held = net.loss_grad(hold_f, hold_t)[0]
    log(f"train: held-out loss {base:.5f} (block-Jacobi, G = I) -> {held:.5f}")
    ev = evaluate(net, seed=1000)
    for name, its in ev.items():
        log(f"eval {name}: " + ", ".join(f"lam {r['lambda']:g}: {r['iterations']} it" for r in its))
    if args.out:
        net.save(args.out)
        log(f"train: wrote {args.out}")
    if args.report:
        with open(args.report, "w") as f:
            json.dump({"samples": int(len(feats)), "epoch_loss": losses.tolist(), "heldout_loss_identity": base,
                       "heldout_loss_trained": held, "eval": ev, "args": vars(args)}, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
