"""Train the learned block-Jacobi correction on the GPU.

    python -m fagra_amd.train_precond --out fagra_amd/learned_precond.f32

Captures every Schur system that LM (FP32, block-Jacobi PCG, CG tolerance 1e-3) attempts over
--max-iter accepted steps -- enough for the damping to fall from 1e-3 to the floor 1e-12 -- on
seeded problems of three kinds: ring256 (256 cameras, 60,000 points, tracks U{3..8}), 256-camera
rings (radii 8/10/12, 60,000 points, tracks U{3..10}, 2% outliers) and 400-camera long-tailed
problems (60,000 points, Zipf(1.5) tracks, focal U(400, 800)), all Huber 1 px. Trains the
network with the library's forward / backward / Adam kernels, records the feature range of the
training set (fagra_amd/learned_precond.range.f32: cameras outside it keep block-Jacobi), and
reports the held-out loss and the PCG iteration counts with and without the network on a
held-out seed across the damping range.
"""
import argparse
import json
import sys
import time

import numpy as np


def log(*a):
    print(*a, flush=True)


TRAINING_SETS = [  # (workload, first seed, generator overrides); seeds per set = --seeds / 4, x2 for ring256
    ("ring256", 1, {}),
    ("rings1500", 1, {"n_cameras": 256, "n_points": 60000}),
    ("zipf4000", 1, {"n_cameras": 400, "n_points": 60000}),
]


def evaluate(net, seed, workload="ring256", cg_tol=1e-3,
             lambdas=(1e-12, 1e-10, 1e-8, 1e-6, 1e-4, 1e-3, 1e-2, 1e-1)):
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
    ap.add_argument("--max-iter", type=int, default=40)
    ap.add_argument("--out", default=None)
    ap.add_argument("--report", default=None)
    args = ap.parse_args()
    from fagra_amd import learned

    t = time.time()
    fs, ts = [], []
    for workload, first, overrides in TRAINING_SETS:
        n = args.seeds // 2 if workload == "ring256" else args.seeds // 4
        f, tg = learned.collect(range(first, first + n), workload=workload, max_iter=args.max_iter, log=log,
                                **overrides)
        fs.append(f)
        ts.append(tg)
    feats, targs = np.concatenate(fs), np.concatenate(ts)
    log(f"train: {len(feats)} samples from {args.seeds} problems in {time.time() - t:.1f}s")
    hold_f, hold_t = learned.collect([1000], max_iter=args.max_iter)
    net = learned.Network(seed=0)
    lo, hi = learned.feature_range(feats)
    net.set_input_range(lo, hi)
    log(f"train: damping feature range {lo[-1] * 8:.2f} .. {hi[-1] * 8:.2f} (log10)")
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
        net.save(args.out, range_path=args.out.replace(".f32", ".range.f32"))
        log(f"train: wrote {args.out}")
    if args.report:
        with open(args.report, "w") as f:
            json.dump({"samples": int(len(feats)), "epoch_loss": losses.tolist(), "heldout_loss_identity": base,
                       "heldout_loss_trained": held, "eval": ev, "args": vars(args)}, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
